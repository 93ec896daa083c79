"""Checks that pin the oracle without the reference: closed forms, hand-derived
trees, invariants and brute force on tiny inputs.  Each is chosen so that a
plausible slip (wrong rounding form or element rule, a textbook heap instead
of the reference's, swapped children, a wrong bit order) fails."""
import numpy as np
import pytest

from oracle import pyoracle as po


# ------------------------------------------------------------------ prequant rules

def test_tie_rounding_forms_by_element_rule():
    """x * r = 2.5 exactly (eb = 0.5 -> r = 1): the AVX form rounds half to
    even (2), the scalar form half away from zero (3).  1D and 2D use the AVX
    form; 3D uses the scalar form from x = 8*floor(DIM0/8) on; 2D elements
    there are never coded and come back as 0."""
    eb = 0.5
    x1 = np.full((21,), 2.5, np.float32)
    y1 = po.decompress(po.compress(x1, eb)[0], x1.shape)
    assert np.all(y1 == 2.0)
    x2 = np.full((5, 12), 2.5, np.float32)
    y2 = po.decompress(po.compress(x2, eb)[0], x2.shape)
    assert np.all(y2[:, :8] == 2.0) and np.all(y2[:, 8:] == 0.0)
    x3 = np.full((3, 4, 12), 2.5, np.float32)
    y3 = po.decompress(po.compress(x3, eb)[0], x3.shape)
    assert np.all(y3[..., :8] == 2.0) and np.all(y3[..., 8:] == 3.0)
    x3b = np.full((3, 4, 12), -2.5, np.float32)
    y3b = po.decompress(po.compress(x3b, eb)[0], x3b.shape)
    assert np.all(y3b[..., :8] == -2.0) and np.all(y3b[..., 8:] == -3.0)


def _closed_form(x, eb, r2r=False):
    """q by the element rule, then out = fl32(fl64(q) * fl64(2 eb))."""
    rng = po.value_range(x) if r2r else 0.0
    f = po.eb_family(r2r, eb, rng)
    q = np.rint(x * np.float32(f.r_f)).astype(np.float32)
    if x.ndim == 3:
        x8 = (x.shape[-1] // 8) * 8
        qs = np.copysign(np.floor(np.abs(x[..., x8:].astype(np.float64) * f.r_d) + 0.5),
                         x[..., x8:].astype(np.float64) * f.r_d).astype(np.float32)
        q[..., x8:] = qs
    out = (q.astype(np.float64) * f.twoeb_d).astype(np.float32)
    if x.ndim == 2:
        out[:, (x.shape[-1] // 8) * 8:] = 0.0
    return out


@pytest.mark.parametrize("shape", [(1001,), (33, 45), (11, 12, 21), (16, 16, 16)])
@pytest.mark.parametrize("eb,r2r", [(1e-3, False), (1e-4, True)])
def test_round_trip_equals_closed_form(shape, eb, r2r):
    """With every |q| < 2^24 the Lorenzo recurrences are exact, so the
    decompressed field equals the prequantized field scaled back."""
    rng = np.random.default_rng(11)
    x = np.cumsum(rng.normal(0, 0.05, int(np.prod(shape)))).astype(np.float32).reshape(shape)
    a, _, _ = po.compress(x, eb, r2r)
    y = po.decompress(a, shape, r2r, po.value_range(x) if r2r else 0.0)
    assert np.array_equal(y.view(np.uint32), _closed_form(x, eb, r2r).view(np.uint32))


def test_error_bound_with_reference_slack():
    """|x - y| <= eb_final up to the fp32 reciprocal/rounding slack of the
    reference (|x| 2^-23 + ulp/2 per rounding), D2 columns excluded."""
    rng = np.random.default_rng(3)
    x = np.cumsum(rng.normal(0, 0.02, (40, 50, 60)), axis=2).astype(np.float32)
    for eb in (1e-2, 1e-3, 1e-4):
        a, _, ebf = po.compress(x, eb, True)
        y = po.decompress(a, x.shape, True, po.value_range(x))
        err = np.abs(x.astype(np.float64) - y.astype(np.float64))
        slack = np.abs(x).astype(np.float64) * 2.0 ** -22 + 1e-30
        assert np.all(err <= ebf + slack)


# ------------------------------------------------------------------ codes / outliers

def test_codes_range_and_outlier_rule():
    """code in {0} U [1, 2R-1]; a jump of >= R quanta inside a block is an
    outlier whose value is q itself; the next element predicts from it."""
    x = np.zeros(64, np.float32)
    x[10] = 5000.0  # eb 0.5 -> r = 1 -> q = 5000 >= R
    code, outl = po.dualquant(x, 0.5)
    assert code.min() >= 0 and code.max() <= 4095
    assert code[10] == 0 and outl[10] == 5000.0
    assert code[11] == 0  # e = -5000: an outlier too
    assert code[12] == 2048 and code[9] == 2048
    assert int((code == 0).sum()) == 2


def test_block_start_predicts_zero():
    """Every Lorenzo block starts from the zero halo: a constant field gives
    code R + c at block starts and R elsewhere (1D, B=8)."""
    x = np.full(32, 7.0, np.float32)
    code, _ = po.dualquant(x, 0.5)
    assert np.all(code[::8] == 2048 + 7) and np.all(np.delete(code, np.s_[::8]) == 2048)


# ------------------------------------------------------------------ Huffman

def test_sz_heap_tree_by_hand():
    """freq {0:5, 1:1, 2:1}.  Traced by hand through qinsert/qremove of
    src/huffman.hh:169-192: the second qremove returns symbol 0 (freq 5)
    because the first left the heap unsifted, so the merges are (0,1) and
    then (that,2) and the most frequent symbol gets a 2-bit code.  A textbook
    heap would give it 1 bit; swapped children would flip every first bit."""
    freq = np.zeros(8, np.uint64)
    freq[:3] = [5, 1, 1]
    cb, cl, nc, tree = po.codebook(freq, 8)
    assert list(cl[:3]) == [2, 2, 1]
    assert list(cb[:3]) == [0b00, 0b01, 0b1]
    assert nc == 5
    L = [1, 2, 0, 0, 0]
    R = [4, 3, 0, 0, 0]
    C = [0, 0, 0, 1, 2]
    T = [0, 0, 1, 1, 1]
    expect = bytes([0] + L + R) + b"".join(int(c).to_bytes(4, "little") for c in C) + bytes(T)
    assert tree == expect


def test_two_symbols_more_frequent_gets_zero():
    """First removed (less frequent) becomes the right child (bit 1)."""
    freq = np.zeros(8, np.uint64)
    freq[2], freq[5] = 10, 3
    cb, cl, nc, _ = po.codebook(freq, 8)
    assert cl[2] == 1 and cl[5] == 1 and cb[2] == 0 and cb[5] == 1 and nc == 3


def _bitstrings(cb, cl):
    return {s: format(int(cb[s]), f"0{int(cl[s])}b") for s in range(len(cl)) if cl[s]}


@pytest.mark.parametrize("seed", range(5))
def test_codebook_full_and_prefix_free(seed):
    rng = np.random.default_rng(seed)
    nsym = int(rng.integers(2, 300))
    freq = np.zeros(1024, np.uint64)
    idx = rng.choice(1024, nsym, replace=False)
    freq[idx] = rng.integers(1, 10 ** 6, nsym).astype(np.uint64)
    cb, cl, nc, _ = po.codebook(freq, 1024)
    assert nc == 2 * nsym - 1
    assert set(np.nonzero(cl)[0]) == set(idx)
    kraft = sum(2.0 ** -int(cl[s]) for s in idx)
    assert abs(kraft - 1.0) < 1e-12
    bs = sorted(_bitstrings(cb, cl).values())
    for a, b in zip(bs, bs[1:]):
        assert not b.startswith(a)


@pytest.mark.parametrize("seed", range(3))
def test_pack_is_msb_first_concatenation(seed):
    """HuffmanEncode == concatenating the codewords' bit strings, zero-padded
    to a byte (textbook bit writer on Python strings)."""
    rng = np.random.default_rng(seed)
    freq = np.zeros(64, np.uint64)
    freq[:40] = rng.integers(1, 1000, 40).astype(np.uint64)
    cb, cl, _, _ = po.codebook(freq, 64)
    sym = rng.integers(0, 40, 3001).astype(np.int32)
    s = "".join(format(int(cb[v]), f"0{int(cl[v])}b") for v in sym)
    s += "0" * (-len(s) % 8)
    expect = bytes(int(s[i:i + 8], 2) for i in range(0, len(s), 8))
    out = np.zeros(len(sym) * 8 + 8, np.uint8)
    n = po.orc().orc_huffman_pack(sym.ctypes.data, len(sym), cb.ctypes.data, cl.ctypes.data, out.ctypes.data)
    assert out[:n].tobytes() == expect


def test_huffman_section_round_trip():
    rng = np.random.default_rng(9)
    x = np.cumsum(rng.normal(0, 0.3, (30, 40)), axis=1).astype(np.float32)
    a, _, _ = po.compress(x, 1e-3)
    sec = po.archive_sections(a)
    code, _ = po.dualquant(x, 1e-3)
    dec = po.huffman_decode(sec["huff"], x.size)
    assert np.array_equal(dec, code.ravel())


def test_single_symbol_tree_and_empty_stream():
    """One symbol: nodeCount 1, code length 0, no stream bytes; the decoder
    fills the constant (src/huffman.hh:348-352)."""
    x = np.zeros((8, 8, 8), np.float32)
    a, nout, _ = po.compress(x, 1e-3)
    sec = po.archive_sections(a)
    assert int.from_bytes(sec["huff"][:4], "big") == 1
    assert sec["huff_size"] == 8 + 1 + 7
    assert nout == 0


# ------------------------------------------------------------------ archive layout

def test_archive_layout_and_size_formula():
    rng = np.random.default_rng(4)
    x = rng.normal(0, 1, (17, 19, 23)).astype(np.float32)
    a, nout, _ = po.compress(x, 1e-3)
    sec = po.archive_sections(a)
    assert sec["ndim"] == 3 and sec["dims"] == (23, 19, 17)
    assert sec["eb"] == 1e-3 and sec["pad_len"] == 0 and sec["n_outlier"] == nout
    assert len(a) == 4 * (3 + 1) + 8 + 24 + sec["huff_size"] + 4 * nout
    assert int.from_bytes(sec["huff"][4:8], "big") == 4096
