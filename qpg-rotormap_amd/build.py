"""Build librotormap_hip.so for gfx950 in-tree (hipcc, no JIT cache).

python qpg-rotormap_amd/build.py  ->  qpg-rotormap_amd/rotormap_amd/librotormap_hip.so
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "rotormap_amd")
LIB = os.path.join(OUT_DIR, "librotormap_hip.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off: every a*b+c stays two IEEE roundings (the encoder's and
# the rescoring's bit-exact contracts); packed fp32 ops are disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off"]

SOURCES = [
    ("encode.hip", "hip"),
    ("gemm_topk.hip", "hip"),
    ("fasta.hip", "hip"),
    ("noalloc.hip", "hip"),
    ("capi.cpp", "hipcc-host"),
    ("serve.cpp", "hipcc-host"),
    ("synth.cpp", "host"),
]


def _compile(src, kind):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    if kind == "hip":
        cmd = [HIPCC, "-c", path, "-o", obj] + HIP_FLAGS
    elif kind == "hipcc-host":
        cmd = [HIPCC, "-c", path, "-o", obj, "-O3", "-fPIC", "-std=c++17",
               "-D__HIP_PLATFORM_AMD__", "-x", "c++"]
    else:
        cmd = ["g++", "-c", path, "-o", obj] + HOST_FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(*s), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-fopenmp", "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
