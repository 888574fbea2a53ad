"""argsort by top-digit buckets (two HBM passes over the two top non-constant digits, then one
LDS sort per 16-bit bucket) against the oracle's LSD radix restatement (radix.rs:187-259).

The bucket path runs for positions-only sorts of at least 2^22 rows with four or more
non-constant digits whose two top digits would give buckets of at most 0.7 x 3072 (or 0.6 x
8192) rows if they were independent. The cases below reach both local capacities, buckets that still
overflow them (correlated top digits: the batched passes over those row ranges), many equal
keys inside a bucket (stability), constant middle digits, i32 keys (4 digits) and f64 bit
patterns, at row counts that are not multiples of any tile.

When the two top digits are too skewed for 16-bit buckets but the three top digits would give
segments that fit (normal floats: sign and exponent in the top bits), the sort takes three HBM
passes, finds the segments where the three top digits change and sorts each segment locally;
the cases below take both local capacities, segments of one key past them, and segments with
several keys past them (the batched passes over their row ranges).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from verkehrt_amd import Registry
    return Registry().resolve(0)


def _check(D, keys, t):
    got = D.call(f"argsort_{t}", [D.array(keys, t)]).to_numpy()
    want = O.argsort(keys)
    assert np.array_equal(got, want), (t, keys.size)


def _rng_u64(rng, n):
    return rng.integers(0, 2**64, n, dtype=np.uint64)


@pytest.mark.parametrize("n", [1 << 22, (1 << 23) + 12345])
def test_uniform_u64(D, n):
    _check(D, _rng_u64(np.random.default_rng(n), n), "u64")


def test_uniform_i64_and_f64_bits(D):
    rng = np.random.default_rng(7)
    n = (1 << 22) + 777
    _check(D, _rng_u64(rng, n).view(np.int64), "i64")
    # every bit pattern, NaNs and infinities included, in the image's total order
    _check(D, _rng_u64(rng, n).view(np.float64), "f64")


def test_uniform_i32(D):
    rng = np.random.default_rng(8)
    n = (1 << 22) + 3
    _check(D, rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32), "i32")


def test_equal_keys_inside_buckets(D):
    # uniform top 16 bits, lower bits from only 4 values: long runs of equal keys per bucket
    rng = np.random.default_rng(9)
    n = (1 << 22) + 99
    top = rng.integers(0, 1 << 16, n, dtype=np.uint64) << np.uint64(48)
    low = rng.integers(0, 4, n, dtype=np.uint64) * np.uint64(0x0101010101)
    _check(D, top | low, "u64")


def test_constant_middle_digits(D):
    # digits 2..5 constant: active digits 0, 1, 6, 7 (two local digits, both in the low bytes)
    rng = np.random.default_rng(10)
    n = 1 << 22
    k = (rng.integers(0, 1 << 16, n, dtype=np.uint64) << np.uint64(48)) | np.uint64(0x0000ABCDEF120000)
    k |= rng.integers(0, 1 << 16, n, dtype=np.uint64)
    _check(D, k, "u64")


def test_large_capacity(D):
    # 32 x 32 top values: about 4096 rows per bucket, the 8192-row local sort
    rng = np.random.default_rng(11)
    n = (1 << 22) + 5
    a = rng.integers(0, 32, n, dtype=np.uint64)
    b = rng.integers(0, 32, n, dtype=np.uint64)
    k = (a << np.uint64(56)) | (b << np.uint64(48)) | rng.integers(0, 1 << 48, n, dtype=np.uint64)
    _check(D, k, "u64")


def test_overflowing_buckets(D):
    # the two top digits equal: each looks uniform alone, but every bucket (j, j) holds n/256
    # rows, above either capacity, so all of them take the batched passes
    rng = np.random.default_rng(12)
    n = (1 << 22) + 17
    j = rng.integers(0, 256, n, dtype=np.uint64)
    k = (j << np.uint64(56)) | (j << np.uint64(48)) | rng.integers(0, 1 << 48, n, dtype=np.uint64)
    _check(D, k, "u64")


def test_some_buckets_overflow(D):
    # a quarter of the rows in the correlated buckets, the rest uniform: both paths in one call
    rng = np.random.default_rng(13)
    n = 1 << 23
    k = _rng_u64(rng, n)
    hot = rng.random(n) < 0.25
    j = rng.integers(0, 64, hot.sum(), dtype=np.uint64)
    k[hot] = (j << np.uint64(56)) | (j << np.uint64(48)) | (k[hot] & np.uint64((1 << 48) - 1))
    _check(D, k, "u64")


def test_i32_overflow_and_three_digits(D):
    # i32 keys with correlated top digits: every bucket overflows and the batched passes sort
    # the two lower digits; with a constant low byte only three digits are active, so the
    # ordinary passes run
    rng = np.random.default_rng(14)
    n = (1 << 22) + 1
    j = rng.integers(0, 128, n, dtype=np.int64)
    k = ((j << 24) | (j << 16) | rng.integers(0, 1 << 16, n, dtype=np.int64)).astype(np.uint32).view(np.int32)
    _check(D, k, "i32")
    k3 = (rng.integers(-2**31, 2**31, n, dtype=np.int64) & ~0xFF).astype(np.int32)
    _check(D, k3, "i32")


def test_large_sub_buckets(D):
    # about 1024 rows per bucket whose next digit takes two values: sub-buckets of about 512
    # rows, too large to count, so those blocks sort by every local digit in LDS
    rng = np.random.default_rng(15)
    n = (1 << 22) + 31
    hi = rng.integers(0, 256, n, dtype=np.uint64)
    lo = rng.integers(0, 16, n, dtype=np.uint64)
    mid = rng.integers(0, 2, n, dtype=np.uint64)
    k = (hi << np.uint64(56)) | (lo << np.uint64(48)) | (mid << np.uint64(40))
    k |= rng.integers(0, 1 << 40, n, dtype=np.uint64)
    _check(D, k, "u64")


@pytest.mark.parametrize("pool", [1024, 4096])
def test_few_distinct_keys(D, pool):
    # keys drawn from a pool of random values: each top-16 bucket holds one key, repeated about
    # n / pool times; 1024 keys overflow every local sort (the positions are copied as they
    # stand, no lower digit varying in any bucket), 4096 keys fit (one-key buckets)
    rng = np.random.default_rng(16 + pool)
    n = (1 << 22) + 3
    vals = _rng_u64(rng, pool)
    _check(D, vals[rng.integers(0, pool, n)], "u64")
    # the same with a second key value in some buckets: lower digits vary there
    k = vals[rng.integers(0, pool, n)]
    k[rng.random(n) < 0.01] ^= np.uint64(0x5A5A)
    _check(D, k, "u64")


@pytest.mark.parametrize("t", ["u64", "i64", "f64", "i32"])
def test_batched_buckets(D, t):
    # batched argsort: items whose keys come from small pools (at least 64 rows per pair of the
    # two top digits) sort by buckets, the others by LSD passes, in one call; pools of 256 keys
    # give one-key segments above the local capacity, pools where pairs share several keys give
    # larger segments (local sorts and range passes)
    rng = np.random.default_rng(31)
    dt = {"u64": np.uint64, "i64": np.int64, "f64": np.float64, "i32": np.int32}[t]

    def pool(g):
        if t == "i32":
            return rng.integers(-2**31, 2**31, g, dtype=np.int64).astype(np.int32)
        return rng.integers(0, 2**64, g, dtype=np.uint64).view(dt)

    xs = []
    specs = [(300000, 256), (5000, 40), (200000, None), (4095, 3), (100000, 64), (1, 1), (250000, "shared"),
             (70000, 1000), (0, 1), (120000, None)]
    for n, g in specs:
        if g is None:
            xs.append(pool(n) if n else np.zeros(0, dt))
        elif g == "shared":
            # pairs of the two top digits each shared by many keys that differ in lower digits
            base = pool(64)
            k = base[rng.integers(0, 64, n)]
            low = rng.integers(0, 1 << 12, n, dtype=np.int64)
            if t == "i32":
                k = (k.astype(np.int64) & ~0xFFF | low).astype(np.int32)
            else:
                k = (k.view(np.uint64) & ~np.uint64(0xFFFFF) | low.astype(np.uint64)).view(dt)
            xs.append(k)
        else:
            xs.append(pool(g)[rng.integers(0, g, n)] if n else np.zeros(0, dt))
    got = D.batch_call(f"argsort_{t}", [[D.array(x, t)] for x in xs])
    for x, p in zip(xs, got):
        assert np.array_equal(p.to_numpy(), O.argsort(x)), (t, x.size)


def _normal_with_specials(rng, n):
    f = rng.standard_normal(n)
    for v in (0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 5e-324, -5e-324):
        f[rng.integers(0, n, 50)] = v
    return f


@pytest.mark.parametrize("n", [(1 << 22) + 7, (1 << 24) + 3])
def test_segments_normal_f64(D, n):
    # three-digit segments: about 140 rows each at 2^22 (3072-row local sorts), more at 2^24
    _check(D, _normal_with_specials(np.random.default_rng(n), n), "f64")


def test_segments_overflow_f64(D):
    # 1% of the rows one key (a segment of one key past the local capacity, written as it
    # stands) and 1% in one segment with random lower bits (past the capacity with many keys:
    # the batched passes over its rows)
    rng = np.random.default_rng(41)
    n = (1 << 22) + 1001
    f = _normal_with_specials(rng, n)
    f[rng.random(n) < 0.01] = 1.0
    hot = rng.random(n) < 0.01
    bits = np.float64(1.5).view(np.uint64) | rng.integers(0, 1 << 40, hot.sum(), dtype=np.uint64)
    f[hot] = bits.view(np.float64)
    _check(D, f, "f64")


def test_segments_skewed_top_byte_u64(D):
    # the top byte one value for 90% of the rows: the two-digit estimate fails, three fit
    rng = np.random.default_rng(42)
    n = (1 << 22) + 77
    top = np.where(rng.random(n) < 0.9, 1, rng.integers(0, 256, n)).astype(np.uint64)
    k = (top << np.uint64(56)) | rng.integers(0, 1 << 56, n, dtype=np.uint64)
    _check(D, k, "u64")
    _check(D, k.view(np.int64), "i64")


def test_segments_three_levels_u64(D):
    # 16 x 16 x 20 values of the three top bytes with weights 0.5 / 1 / 1.6 / 2.4 / 4: segments of
    # about 1.2K and 2.4K rows (3072-row local sorts), 3.8K (5120-row), 5.7K (8192-row) and 9.5K
    # rows (past every capacity: the batched passes over their rows)
    rng = np.random.default_rng(43)
    n = (1 << 24) + 5
    a, b, c = np.meshgrid(np.arange(16), np.arange(16), np.arange(20), indexing="ij")
    combos = ((a.ravel() << 16) | (b.ravel() << 8) | c.ravel()).astype(np.uint64)
    w = rng.choice(np.array([0.5, 1.0, 1.6, 2.4]), combos.size)
    w[rng.integers(0, combos.size, 40)] = 4.0
    top = combos[rng.choice(combos.size, n, p=w / w.sum())]
    k = (top << np.uint64(40)) | rng.integers(0, 1 << 40, n, dtype=np.uint64)
    _check(D, k, "u64")
