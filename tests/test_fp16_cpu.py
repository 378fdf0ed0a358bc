"""Storage in IEEE fp16 and the 16-bit precision modes (GS_PREC_BF16,
GS_PREC_FP16), CPU side: the fp16 helpers shared by the kernels and the
restatement (splat.js_amd/csrc/gs_math.h) against numpy's float16 (an
independent IEEE implementation), exact unbiasedness of the stochastic
rounding over its random bits, and the restatement's 16-bit SH Adam against
its f32 SH Adam plus those roundings."""
import numpy as np

import oracle as O
from helpers import adam_for_iter


def _bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


def test_fp16_round_to_nearest_even_matches_numpy():
    rng = np.random.default_rng(11)
    u = rng.integers(0, 2 ** 32, 1_000_000, dtype=np.uint64).astype(np.uint32)
    x = u.view(np.float32)
    x = x[np.isfinite(x)]
    # dense around the fp16 range: normal, subnormal, the overflow edge, ties
    mags = np.concatenate([
        rng.uniform(-16, 16, 200_000).astype(np.float32) * np.float32(2.0 ** -10),
        np.exp2(rng.uniform(-27, 17, 200_000)).astype(np.float32),
        np.array([65504, 65519.99, 65520, 65535, 65536, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                  2.0 ** -14, 2.0 ** -14 - 2.0 ** -25, 1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, 0.0], np.float32)])
    x = np.concatenate([x, mags, -mags])
    got = O.f2h_rne(x)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(got, want)
    nf = O.f2h_rne(np.array([np.inf, -np.inf, np.nan], np.float32))
    assert list(nf[:2]) == [0x7C00, 0xFC00] and (nf[2] & 0x7C00) == 0x7C00 and (nf[2] & 0x3FF) != 0


def test_fp16_decode_is_exact():
    b = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16)
    f = O.h2f(b)
    want = b.view(np.float16).astype(np.float32)
    fin = np.isfinite(want)
    np.testing.assert_array_equal(_bits(f)[fin], _bits(want)[fin])
    assert np.all(np.isnan(f[np.isnan(want)])) and np.array_equal(f[np.isinf(want)], want[np.isinf(want)])


def test_fp16_stochastic_rounding_is_unbiased_over_its_random_bits():
    """Over all 2^16 random words a value rounds up with probability equal to
    its distance from the lower fp16 neighbour (in units of that gap): to
    2^-13 for normal results, to 2^-16 for subnormal ones."""
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.standard_normal(8).astype(np.float32) * np.float32(0.05),
                         np.array([1.0, 0.3333, 1e-5, 3e-6, -7.1e-7, 1000.7, 65500.3], np.float32)])
    rnd = np.arange(1 << 16, dtype=np.uint32)
    for x in xs:
        out = O.f2h_sr(np.full(1 << 16, x, np.float32), rnd)
        vals = out.view(np.float16).astype(np.float64)
        lo16 = np.float16(x)
        if abs(float(lo16)) > abs(float(x)):
            lo16 = np.nextafter(lo16, np.float16(0))
        hi16 = np.nextafter(lo16, np.float16(np.copysign(np.inf, float(x))))
        assert set(np.unique(vals)) <= {float(lo16), float(hi16)}, x
        frac = (abs(float(x)) - abs(float(lo16))) / (abs(float(hi16)) - abs(float(lo16)))
        p_up = np.mean(vals == float(hi16))
        tol = 2.0 ** -13 if abs(float(lo16)) >= 2.0 ** -14 else 2.0 ** -16
        assert abs(p_up - frac) <= tol, (x, p_up, frac)


def test_fp16_round_up_is_the_smallest_upper_bound():
    rng = np.random.default_rng(4)
    x = np.exp2(rng.uniform(-20, 15, 100_000)).astype(np.float32)
    up = O.f2h_up(x)
    v = up.view(np.float16).astype(np.float32)
    assert np.all(v >= x)
    below = (up - 1).view(np.float16).astype(np.float32)
    assert np.all(below < x)


def test_sh_adam_16bit_equals_f32_adam_plus_stochastic_rounding():
    """The 16-bit SH Adam keeps f32 moments: they equal the f32 SH Adam's on
    the decoded coefficients, and the stored coefficient is the f32 result
    rounded stochastically with the documented bits (stream 2)."""
    rng = np.random.default_rng(8)
    n, nc, step = 300, 24, 37
    hp = adam_for_iter(step, 2.0)
    for fmt, enc, dec, sr in ((1, O.f2bf_rne, O.bf2f, O.f2bf_sr), (2, O.f2h_rne, O.h2f, O.f2h_sr)):
        sh16 = enc(rng.normal(0, 0.1, (n, nc)).astype(np.float32))
        g = rng.normal(0, 1e-3, (n, nc)).astype(np.float32)
        g[0, 0] = np.nan
        m = rng.normal(0, 1e-4, (n, nc)).astype(np.float32)
        v = np.abs(rng.normal(0, 1e-6, (n, nc))).astype(np.float32)
        sh_f = dec(sh16)
        m_f, v_f = m.copy(), v.copy()
        O.sh_adam(hp, sh_f, g, m_f, v_f)
        O.sh_adam_h16(hp, sh16, g, m, v, step, fmt)
        np.testing.assert_array_equal(_bits(m), _bits(m_f))
        np.testing.assert_array_equal(_bits(v), _bits(v_f))
        bits = np.array([[O.lib().orc_sr_bits(i, c, step, 2) for c in range(nc)] for i in range(n)], np.uint32)
        np.testing.assert_array_equal(sh16, sr(sh_f, bits))
