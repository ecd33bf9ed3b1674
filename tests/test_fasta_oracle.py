"""The oracle's FASTA parse against hand-worked cases of the reference
reader's rules (RotorMap/fasta/reader.jl): record starts at line-start '>',
junk before the first record, header capture with CRLF, line terminators
skipped, _LUT3 byte codes (junk -> G), trim to k, skip short records."""
import numpy as np

import oracle.oracle as O


def codes_word(codes):
    w = 0
    for j, c in enumerate(codes):
        w |= c << (2 * j)
    return w


def test_simple_and_crlf():
    w, h = O.fasta_reads(b">r1 desc\nACGT\n>r2\r\nAC\r\nGT\r\n", 4)
    assert h == [b">r1 desc", b">r2"]
    assert w[:, 0].tolist() == [codes_word([0, 1, 2, 3])] * 2


def test_byte_codes_and_junk_bases():
    # lowercase, N, IUPAC, space, '-', and a '>' that is not at a line start
    w, h = O.fasta_reads(b">x\nacgtNRY S->T\n", 12)
    assert w[0, 0] == codes_word([0, 1, 2, 3, 2, 2, 2, 2, 2, 2, 2, 3])


def test_trim_skip_and_order():
    data = b">long\n" + b"T" * 40 + b"\n>short\nACG\n>exact\n" + b"C" * 20 + b"\n"
    w, h = O.fasta_reads(data, 20)
    assert h == [b">long", b">exact"]
    assert w.shape == (2, 2)
    assert w[0].tolist() == [0xFFFFFFFF, 0xFF]          # 20 x T, tail zero
    assert w[1].tolist() == [0x55555555, 0x55]          # 20 x C


def test_junk_prefix_and_record_after_cr():
    data = b"leading junk\n>a\nAC\r>b\nGT\n"
    w, h = O.fasta_reads(data, 2)
    assert h == [b">a", b">b"]  # '>' after a bare '\r' starts a record
    assert w[:, 0].tolist() == [codes_word([0, 1]), codes_word([2, 3])]


def test_header_without_newline_and_eof_cases():
    # no newline after the last header: no read; sequence running to EOF: kept
    assert O.fasta_reads(b">only header", 1)[0].shape[0] == 0
    w, h = O.fasta_reads(b">a\nACGTACGT", 8)
    assert h == [b">a"] and w[0, 0] == codes_word([0, 1, 2, 3] * 2)
    # a header line ended by a bare '\r' before the next record yields nothing
    w, h = O.fasta_reads(b">a\rACGT\n>b\nTTTT\n", 4)
    assert h == [b">b"]


def test_empty_lines_and_empty_input():
    w, h = O.fasta_reads(b">a\n\nAC\n\n\nGT\n\n", 4)
    assert w[0, 0] == codes_word([0, 1, 2, 3])
    assert O.fasta_reads(b"", 4)[0].shape[0] == 0
    assert O.fasta_reads(b"ACGT\nACGT\n", 4)[0].shape[0] == 0  # no record at all


def test_packing_matches_forward_pack():
    rng = np.random.default_rng(0)
    codes = rng.integers(0, 4, 1000).astype(np.uint8)
    seq = "".join("ACGT"[c] for c in codes).encode()
    w, _ = O.fasta_reads(b">r\n" + b"\n".join(seq[i:i + 61] for i in range(0, 1000, 61)) + b"\n", 997)
    np.testing.assert_array_equal(w[0], O.forward_pack(codes[:997]))
