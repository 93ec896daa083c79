"""The library's host codebook builder (the one every compress call uses)
against the oracle's restatement of src/huffman.hh, on histograms chosen to
stress the reference's heap quirks: ties, sparse support, wide dictionaries
and degenerate trees.  Runs without a GPU."""
import numpy as np
import pytest

from oracle import pyoracle as po
from vecsz_amd import VecszError, huffman_codebook


def _hists():
    rng = np.random.default_rng(5)
    yield "two", np.array([3, 9], np.uint64)
    yield "one_symbol", np.array([0, 0, 7, 0], np.uint64)
    yield "all_equal_256", np.full(256, 11, np.uint64)
    yield "powers_of_two", (2 ** np.arange(40, dtype=np.uint64)).astype(np.uint64)
    yield "ties_small", rng.integers(0, 4, 1024).astype(np.uint64)
    d = np.abs(np.arange(4096) - 2048.0)
    yield "laplace_4096", (1e8 * np.exp(-d / 40) / 40 + rng.integers(0, 3, 4096)).astype(np.uint64)
    yield "laplace_sparse", np.where(rng.random(4096) < 0.3, (1e6 * np.exp(-d / 200)).astype(np.uint64), 0).astype(np.uint64)
    yield "heavy_zero_bin", np.concatenate([[10 ** 9], rng.integers(0, 50, 4095)]).astype(np.uint64)
    yield "dict_65536", rng.integers(0, 1000, 65536).astype(np.uint64)
    yield "dict_65536_sparse", np.where(rng.random(65536) < 0.002, rng.integers(1, 10 ** 6, 65536), 0).astype(np.uint64)
    yield "big_counts", rng.integers(2 ** 40, 2 ** 41, 300).astype(np.uint64)


@pytest.mark.parametrize("name,freq", list(_hists()), ids=[n for n, _ in _hists()])
def test_codebook_matches_oracle(name, freq):
    bits, lens, nc, tree = huffman_codebook(freq)
    obits, olens, onc, otree = po.codebook(freq, freq.size)
    assert nc == onc
    assert np.array_equal(lens, olens)
    assert np.array_equal(bits, obits)
    assert tree == otree


def test_codebook_too_long_is_an_error():
    # 85 Fibonacci weights drive the reference's heap to a codeword longer
    # than 64 bits (its >64-bit branch), which this library reports; the
    # oracle refuses the same histogram
    fib = [1, 1]
    while len(fib) < 85:
        fib.append(fib[-1] + fib[-2])
    f = np.array(fib, np.uint64)
    with pytest.raises(VecszError) as e:
        huffman_codebook(f)
    assert e.value.status == 4
    with pytest.raises(po.OracleError):
        po.codebook(f, f.size)


def test_codebook_rejects_empty_histogram():
    with pytest.raises(VecszError):
        huffman_codebook(np.zeros(16, np.uint64))
