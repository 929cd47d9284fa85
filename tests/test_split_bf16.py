"""The split of FP32 operands into three bf16 parts used by the FP32 LU's
trailing updates (lu.hip, gemm_split_kernel), restated in numpy and checked
against what the arithmetic fixes: the parts sum to x exactly, |m| <= 2^-8 |x|,
|l| <= 2^-16 |x|, and the six kept partial products give a b within
2^-23 (1 + 2^-9) |a b| (the dropped ml + lm + ll: each at most 2^-8 |m_a m_b|),
while dropping any one kept term breaks that bound."""
import numpy as np


def bf16_rne(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest bf16 (ties to even), returned as float32 (the
    rounding v_cvt_pk_bf16_f32 applies to finite values)."""
    b = x.astype(np.float32).view(np.uint32).astype(np.uint64)
    r = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32)


def split3(x: np.ndarray):
    x = x.astype(np.float32)
    h = bf16_rne(x)
    r1 = (x - h).astype(np.float32)
    m = bf16_rne(r1)
    r2 = (r1 - m).astype(np.float32)
    l = bf16_rne(r2)
    return h, m, l, r1, r2


def _samples(n=200000, seed=0):
    rng = np.random.default_rng(seed)
    mant = rng.uniform(1.0, 2.0, n)
    expo = rng.integers(-60, 60, n)
    sign = rng.choice([-1.0, 1.0], n)
    x = (sign * mant * np.exp2(expo)).astype(np.float32)
    # values with all 24 significand bits set, powers of two, bf16 ties
    special = np.array([1.0, -1.0, 2.0 ** -40, np.float32(1.9999999), np.float32(1.00390625),
                        np.float32(1.0 + 2 ** -8 + 2 ** -16), np.float32(-3.0000002)], np.float32)
    return np.concatenate([x, special])


def test_parts_are_exact_and_bounded():
    x = _samples()
    h, m, l, r1, r2 = split3(x)
    # both differences are exact in FP32, and l = r2 exactly (r2 fits a bf16)
    xd = x.astype(np.float64)
    assert np.array_equal(r1.astype(np.float64), xd - h.astype(np.float64))
    assert np.array_equal(r2.astype(np.float64), r1.astype(np.float64) - m.astype(np.float64))
    assert np.array_equal(l, r2)
    assert np.array_equal(h.astype(np.float64) + m.astype(np.float64) + l.astype(np.float64), xd)
    # every part is a bf16 value
    for part in (h, m, l):
        assert not np.any(part.view(np.uint32) & 0xFFFF)
    ax = np.abs(xd)
    assert np.all(np.abs(m.astype(np.float64)) <= 2.0 ** -8 * ax)
    assert np.all(np.abs(l.astype(np.float64)) <= 2.0 ** -16 * ax)


def test_six_products_within_bound():
    a, b = _samples(seed=1), _samples(seed=2)
    ha, ma, la, _, _ = split3(a)
    hb, mb, lb, _, _ = split3(b)
    f = lambda t: t.astype(np.float64)
    terms = {"hh": f(ha) * f(hb), "hm": f(ha) * f(mb), "mh": f(ma) * f(hb),
             "hl": f(ha) * f(lb), "lh": f(la) * f(hb), "mm": f(ma) * f(mb)}
    exact = f(a) * f(b)
    bound = 2.0 ** -23 * (1 + 2.0 ** -9) * np.abs(exact)
    kept = sum(terms.values())
    assert np.all(np.abs(kept - exact) <= bound)
    # each kept term matters: without it the bound fails somewhere
    for name, t in terms.items():
        assert np.any(np.abs(kept - t - exact) > bound), name
