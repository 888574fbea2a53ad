"""reduce_into_add_f64 by class folds against the oracle's defined order (row i < n - n%16
into stripe i%16 of its bucket in index order, stripes combined pairwise, tail rows after).

From 17 to 2^20 buckets the chain keys (bucket * 16 + row % 16) are sorted without their top
non-constant byte: rows of one class (the key bits below that byte) come out contiguous and in
row order, and one block per class partitions each chunk of its rows by the top byte in LDS and
folds every chain in row order. The cases reach both block sizes (1024 threads from 16384 rows
per class), the smallest bucket count of the form, 16-bit classes, only the low byte varying,
constant nonzero bits above the top varying byte, skewed buckets, and row counts that are not
multiples of 16 or of a chunk.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from verkehrt_amd import Registry
    return Registry().resolve(0)


def S(t, v):
    from verkehrt_amd import Scalar
    return Scalar(t, v)


def _check(D, k, v, nb):
    got = D.call("reduce_into_add_f64", [D.array(k), D.array(v), S("i64", nb)]).to_numpy()
    want = O.reduce_into(k, v, nb)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (k.size, nb)


def _vals(rng, n):
    # mixed signs and magnitudes, so that the order of the adds shows in the bits
    return rng.standard_normal(n) * np.exp2(rng.integers(-20, 20, n))


@pytest.mark.parametrize("n,nb", [(4_194_304 + 4_099, 1000), (4_500_017, 4096), (3_000_013, 17)])
def test_byte_classes_wide_blocks(D, n, nb):
    # about 16K rows or more per class of the low byte: 1024-thread blocks, chunks of 8192 rows
    rng = np.random.default_rng(nb)
    _check(D, rng.integers(0, nb, n).astype(np.int64), _vals(rng, n), nb)


@pytest.mark.parametrize("n,nb", [(200_003, 300), (1_000_001, 65_536), (2_000_000 + 9, 1 << 20)])
def test_small_and_16_bit_classes(D, n, nb):
    # fewer rows per class: 256-thread blocks; from 4096 buckets on the classes are 16 bits wide
    rng = np.random.default_rng(nb + 1)
    _check(D, rng.integers(0, nb, n).astype(np.int64), _vals(rng, n), nb)


def test_only_low_byte_varies(D):
    # 1000 buckets declared but keys below 16: the chain keys' only varying byte is the low one,
    # so every digit is sorted and each class is one chain
    rng = np.random.default_rng(5)
    n = 1_000_003
    _check(D, rng.integers(0, 16, n).astype(np.int64), _vals(rng, n), 1000)


def test_constant_bits_above_top_digit(D):
    # keys in [2^19, 2^19 + 4096) of 2^20 buckets: chain keys' top byte is a nonzero constant,
    # the byte below it varies over all 256 values
    rng = np.random.default_rng(6)
    n = 1_500_007
    k = (rng.integers(0, 4096, n) + (1 << 19)).astype(np.int64)
    _check(D, k, _vals(rng, n), 1 << 20)


def test_skewed_buckets(D):
    # one bucket holds half the rows, the rest spread over 3000 buckets, a few buckets empty
    rng = np.random.default_rng(7)
    n = 5_000_011
    k = rng.integers(0, 3000, n).astype(np.int64)
    k[rng.random(n) < 0.5] = 1234
    _check(D, k, _vals(rng, n), 3001)
