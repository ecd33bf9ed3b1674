"""rm_map_batch_noalloc (no device allocation; its buffers live in a static
device array) against rm_map_batch through the C ABI: equal bit for bit in
every mode, on two calls in a row, the second with a smaller input, and a
clear error for a batch past the scratch array's capacity."""
import ctypes as C

import numpy as np
import pytest

import rotormap_amd as R
from rotormap_amd import _lib as L

pytestmark = pytest.mark.gpu

K = 20_000


@pytest.fixture(scope="module")
def setup(torch_gpu):
    ref = R.synth_reference([1_200_000, 800_000], gc=0.41, seed=91)
    re = R.RopeEncoder(k=K, s=5, m=1, c=5)
    idx = R.build_index(ref, re, kstep=1_000)
    reads = R.sample_reads(ref, 3000, K, err=0.05, seed=92, draw="random", model=1)
    return re, idx, reads


def call(torch, fn, B, re, words, n, mode, k=20, kprime=32):
    dw = torch.from_numpy(np.ascontiguousarray(words[:n]).view(np.int32).reshape(-1).copy()).cuda()
    v = torch.full((k, n), 7.0, dtype=torch.float32, device="cuda").t()
    l = torch.full((k, n), -5, dtype=torch.int32, device="cuda").t()
    nr = torch.empty(n, re.m, dtype=torch.float32, device="cuda").t()
    rc = fn(L.ctx(0), B.handle, L.ptr(dw), n, re.k, re.s, re.m, re.c, k, kprime, R.MODES[mode],
            L.ptr(v), L.ptr(l), n, L.ptr(nr), L.stream_of(0))
    torch.cuda.synchronize()
    return rc, v.cpu().numpy(), l.cpu().numpy(), nr.cpu().numpy()


@pytest.mark.parametrize("mode", ["fp8_rescore", "fp8", "fp16"])
def test_noalloc_equals_map_batch(torch_gpu, setup, mode):
    torch = torch_gpu
    re, idx, reads = setup
    B = R.ResidentIndex(idx.embeds8, fp16=(mode == "fp16"))
    lib = L.lib()
    for n in (2500, 37):  # the second call is smaller: stale scratch must not leak
        rc0, v0, l0, n0 = call(torch, lib.rm_map_batch, B, re, reads.words, n, mode)
        rc1, v1, l1, n1 = call(torch, lib.rm_map_batch_noalloc, B, re, reads.words, n, mode)
        assert rc0 == 0 and rc1 == 0, L.last_error()
        np.testing.assert_array_equal(l1, l0, err_msg=f"{mode} n={n}")
        np.testing.assert_array_equal(v1, v0, err_msg=f"{mode} n={n}")
        np.testing.assert_array_equal(n1, n0, err_msg=f"{mode} n={n}")
        res = R.count_mapped(l1, reads.recs[:n], reads.starts[:n], idx.rec_of_col, idx.starts, K)
        assert res["correct"] >= 0.99 * n


def test_noalloc_rejects_oversized_batch(torch_gpu, setup):
    torch = torch_gpu
    re, idx, reads = setup
    B = R.ResidentIndex(idx.embeds8)
    n = 7000  # 7000 x (2048 + 8192 + partials) bytes > 64 MiB
    words = np.resize(reads.words, (n, reads.words.shape[1]))
    rc, v, l, _ = call(torch, L.lib().rm_map_batch_noalloc, B, re, words, n, "fp8_rescore")
    assert rc == L.RM_ERR_SHAPE and "scratch" in L.last_error()
    assert np.all(l == -5)  # nothing was launched
