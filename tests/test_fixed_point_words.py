"""The mixed P2G's fixed-point words, on the CPU.

The P2G flush turns a float register sum `val` into a 64-bit integer word
q = rint(val * scale) (scale a power of two, ties to even) in two ways:

* the exact split (fx_add_f): q rounded in FP32, hi = floor(q / 2^32) in
  FP32, lo = q - hi 2^32 in FP64, word = hi << 32 | lo;
* the three-instruction word (fx_word_fast), taken when every |val| of the
  wave is below 2^51 / scale: t = val * scale + 1.5 * 2^52 rounded once in
  FP64, word = bits(t) - bits(1.5 * 2^52).

Both must give the same 64-bit word wherever the second is used; numpy's
float64 product of a float32 and a power of two is exact, so the sum with
1.5 * 2^52 is the single rounding of the device's fma.
"""
import numpy as np

MAGIC = 6755399441055744.0                 # 1.5 * 2^52
MAGIC_BITS = np.uint64(0x4338000000000000)
FAST_LIMIT = 2.0 ** 51


def exact_split(val, scale):
    q = np.rint(val.astype(np.float32) * np.float32(scale)).astype(np.float32)
    hi = np.floor(q * np.float32(2.0 ** -32)).astype(np.float32)
    lo = q.astype(np.float64) - hi.astype(np.float64) * 4294967296.0
    return (hi.astype(np.int64).astype(np.uint64) << np.uint64(32)) | \
        lo.astype(np.uint64)


def fast_word(val, scale):
    t = val.astype(np.float32).astype(np.float64) * scale + MAGIC
    return t.view(np.uint64) - MAGIC_BITS


def fast_limit(scale):
    lim = FAST_LIMIT / scale
    return np.float32(2.0 ** 126) if lim > 2.0 ** 126 else np.float32(lim)


def test_fast_word_equals_the_exact_split():
    rng = np.random.default_rng(20261202)
    for e in range(-60, 61, 3):
        scale = 2.0 ** e
        lim = fast_limit(scale)
        # magnitudes over 80 binades around the limit, both signs
        mag = rng.random(20000) * 2.0 ** rng.integers(-40, 40, 20000) * 2.0 ** (51 - e)
        val = np.minimum(mag, 3.0e38).astype(np.float32)
        val *= np.where(rng.random(20000) < 0.5, -1, 1).astype(np.float32)
        # exact halves (ties) of the fixed-point quantum
        ties = ((rng.integers(-1000, 1000, 5000) + 0.5) * 2.0 ** -e).astype(np.float32)
        val = np.concatenate([val, ties])
        use = np.abs(val) < lim
        assert use.sum() > 5000
        a, b = exact_split(val[use], scale), fast_word(val[use], scale)
        assert np.array_equal(a, b), e


def test_words_are_the_rounded_products():
    vals = np.array([0.0, -0.0, 0.5, 1.5, 2.5, -0.5, -1.5, 3.25, -7.0, 1e-30],
                    dtype=np.float32)
    words = fast_word(vals, 1.0).view(np.int64)
    assert words.tolist() == [0, 0, 0, 2, 2, 0, -2, 3, -7, 0]
    big = np.array([2.0 ** 50 + 2.0 ** 27, -(2.0 ** 50)], dtype=np.float32)
    assert fast_word(big, 1.0).view(np.int64).tolist() == [2 ** 50 + 2 ** 27, -(2 ** 50)]


def test_limit_keeps_the_word_in_range():
    for e in range(-120, 121, 7):
        scale = 2.0 ** e
        lim = np.float64(fast_limit(scale))
        # the largest float below the limit still maps below 2^51
        v = np.nextafter(np.float32(lim), np.float32(0))
        assert abs(np.float64(v) * scale) < FAST_LIMIT
