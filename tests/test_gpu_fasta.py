"""GPU FASTA parse (rm_fasta_reads) against the oracle, bit for bit: packed
words and headers of randomized files (line widths, CRLF, lowercase, IUPAC
and junk bytes, N runs, short records, a junk prefix, '>' inside lines, a
missing final newline), an unaligned buffer start, and chunked streaming."""
import ctypes as C

import numpy as np
import pytest

import oracle.oracle as O
import rotormap_amd as R
from rotormap_amd import _lib as L

pytestmark = pytest.mark.gpu


def make_fasta(seed, nrec, k, junk=True, final_newline=True):
    rng = np.random.default_rng(seed)
    alpha = np.frombuffer(b"ACGTACGTACGTACGTacgtNNNRYKM ->", np.uint8)
    out = [b"# not a record\n" if junk else b""]
    for r in range(nrec):
        L_ = int(max(1, k + rng.integers(-k // 8, k // 4)))
        seq = alpha[rng.integers(0, len(alpha), L_)]
        width = int(rng.integers(40, 130))
        eol = b"\r\n" if rng.random() < 0.3 else b"\n"
        head = f">read_{r} start={rng.integers(1, 10**8)} src=chr{r % 7}".encode()
        # '>' only inside lines: at a line start it would open a record
        lines = [seq[i:i + width].tobytes() for i in range(0, L_, width)]
        lines = [b"A" + x[1:] if x.startswith(b">") else x for x in lines]
        if rng.random() < 0.1:
            lines.insert(len(lines) // 2, b"")  # an empty line
        out.append(head + eol + eol.join(lines) + eol)
    data = b"".join(out)
    return data if final_newline else data.rstrip(b"\r\n")


def parse_gpu(torch, data, k):
    w, h = R.fasta_reads_gpu(data, k)
    torch.cuda.synchronize()
    return w.cpu().numpy().view(np.uint32), [x.encode("latin-1") for x in h]


@pytest.mark.parametrize("seed,nrec,k", [(1, 300, 2000), (2, 40, 20000), (3, 500, 17), (4, 64, 16)])
def test_fasta_gpu_matches_oracle(torch_gpu, seed, nrec, k):
    data = make_fasta(seed, nrec, k, final_newline=seed % 2 == 0)
    ow, oh = O.fasta_reads(data, k)
    gw, gh = parse_gpu(torch_gpu, data, k)
    assert gh == oh
    np.testing.assert_array_equal(gw, ow)
    assert len(oh) > 0.5 * nrec  # most records are long enough


def test_fasta_gpu_edge_cases(torch_gpu):
    cases = [b">r1 desc\nACGT\n>r2\r\nAC\r\nGT\r\n", b">x\nacgtNRY S->T\n",
             b"leading junk\n>a\nAC\r>b\nGT\n", b">only header", b">a\nACGTACGT",
             b">a\rACGT\n>b\nTTTT\n", b">a\n\nAC\n\n\nGT\n\n", b"ACGT\nACGT\n", b">\n>\nA\n"]
    for data in cases:
        for k in (1, 2, 4, 12):
            ow, oh = O.fasta_reads(data, k)
            gw, gh = parse_gpu(torch_gpu, data, k)
            assert gh == oh, (data, k)
            np.testing.assert_array_equal(gw, ow.reshape(gw.shape), err_msg=repr((data, k)))
    w, h = R.fasta_reads_gpu(b"", 4)
    assert w.shape[0] == 0 and h == []


def test_fasta_gpu_unaligned_buffer(torch_gpu):
    torch = torch_gpu
    data = make_fasta(5, 50, 1000)
    k, W = 1000, (1000 + 15) // 16
    for off in (1, 3, 7, 13):
        buf = torch.from_numpy(np.frombuffer(b"x" * off + data, np.uint8).copy()).cuda()
        cap = data.count(b">") + 1
        words = torch.empty(cap, W, dtype=torch.int32, device="cuda")
        n = C.c_int64()
        L.check(L.lib().rm_fasta_reads(L.ctx(0), C.c_void_p(buf.data_ptr() + off), len(data), k,
                                       L.ptr(words), cap, None, None, C.byref(n), None))
        torch.cuda.synchronize()
        ow, _ = O.fasta_reads(data, k)
        assert n.value == len(ow)
        np.testing.assert_array_equal(words[:n.value].cpu().numpy().view(np.uint32), ow)


def test_fasta_gpu_capacity_error(torch_gpu):
    torch = torch_gpu
    data = make_fasta(6, 30, 100, junk=False)
    raw = torch.from_numpy(np.frombuffer(data, np.uint8).copy()).cuda()
    words = torch.empty(4, 7, dtype=torch.int32, device="cuda")
    n = C.c_int64()
    rc = L.lib().rm_fasta_reads(L.ctx(0), L.ptr(raw), len(data), 100, L.ptr(words), 4, None, None,
                                C.byref(n), None)
    assert rc == L.RM_ERR_SHAPE and n.value == len(O.fasta_reads(data, 100)[1])


def test_iter_fasta_gpu_chunks(torch_gpu, tmp_path):
    data = make_fasta(7, 200, 3000)
    p = tmp_path / "reads.fa"
    p.write_bytes(data)
    ow, oh = O.fasta_reads(data, 3000)
    got_w, got_h, firsts = [], [], []
    for w, h, first in R.iter_fasta_gpu(str(p), 3000, batch_size=37, chunk_bytes=50_000):
        got_w.append(w.cpu().numpy().view(np.uint32))
        got_h.extend(x.encode("latin-1") for x in h)
        firsts.append(first)
    assert got_h == oh
    np.testing.assert_array_equal(np.concatenate(got_w), ow)
    assert firsts[0] == 1 and firsts[1] == 38


def test_fasta_host_entry(torch_gpu):
    data = make_fasta(8, 60, 2500)
    raw = np.frombuffer(data, np.uint8).copy()
    cap, W = data.count(b">") + 1, (2500 + 15) // 16
    words = np.zeros((cap, W), np.uint32)
    off = np.zeros(cap, np.int64)
    ln = np.zeros(cap, np.int32)
    n = C.c_int64()
    L.check(L.lib().rm_fasta_reads_host(L.ctx(0), L.ptr(raw), raw.size, 2500, L.ptr(words), cap,
                                        L.ptr(off), L.ptr(ln), C.byref(n)))
    ow, oh = O.fasta_reads(data, 2500)
    assert n.value == len(oh)
    np.testing.assert_array_equal(words[:n.value], ow)
    assert [raw[o:o + l].tobytes() for o, l in zip(off[:n.value], ln[:n.value])] == oh
