"""Mixed-precision storage (bf16 SH coefficients and Adam moments), CPU side:
the rounding helpers shared by the kernels and the restatement
(splat.js_amd/csrc/gs_math.h), pinned against independent definitions, and
the restatement's mixed Adam against its f32 Adam plus those roundings."""
import numpy as np

import oracle as O
from helpers import adam_for_iter


def _bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


def _specials():
    return np.array([0.0, -0.0, 1.0, -1.0, 1.00390625, 1.005859375, 1.01171875, 3.3895314e38, 3.4e38,
                     np.inf, -np.inf, np.nan, 1e-40, -1e-45, 1.17549435e-38], np.float32)


def test_bf16_round_to_nearest_even_matches_a_float64_definition():
    rng = np.random.default_rng(5)
    u = rng.integers(0, 2 ** 32, 1_000_000, dtype=np.uint64).astype(np.uint32)
    x = np.concatenate([u.view(np.float32), _specials()])
    x = x[np.isfinite(x)]
    got = O.f2bf_rne(x).astype(np.uint32)
    # independent: the two bf16 neighbours in float64, nearest wins, ties to the even one
    xb = _bits(x)
    lo_b = (xb & 0xFFFF0000).astype(np.uint32)
    lo = lo_b.view(np.float32).astype(np.float64)
    hi_b = lo_b + np.uint32(0x10000)
    # the next bf16 magnitude, exactly in float64 (2^128 past the largest finite one)
    e = np.maximum((xb >> 23) & 0xFF, 1).astype(np.float64)
    hi = lo + np.sign(np.where(x == 0, 1.0, x)) * np.exp2(e - 127.0 - 7.0)
    xd = x.astype(np.float64)
    d_lo, d_hi = np.abs(xd - lo), np.abs(hi - xd)
    up = (d_hi < d_lo) | ((d_hi == d_lo) & (((lo_b >> 16) & 1) == 1))
    want = np.where(up, hi_b >> 16, lo_b >> 16)
    np.testing.assert_array_equal(got, want)
    # non-finite: inf stays inf, NaN stays a NaN
    nf = O.f2bf_rne(np.array([np.inf, -np.inf, np.nan], np.float32))
    assert list(nf[:2]) == [0x7F80, 0xFF80] and (nf[2] & 0x7F80) == 0x7F80 and (nf[2] & 0x7F) != 0


def test_bf16_decode_is_exact():
    b = np.arange(0, 1 << 16, dtype=np.uint32).astype(np.uint16)
    f = O.bf2f(b)
    np.testing.assert_array_equal(_bits(f), b.astype(np.uint32) << 16)


def test_stochastic_rounding_is_exactly_unbiased_over_its_random_bits():
    """For a value x, the 2^16 random words round up exactly low16(bits(x))
    times, so E[rounded] = x exactly (away from the overflow edge)."""
    rng = np.random.default_rng(9)
    xs = np.concatenate([rng.standard_normal(6).astype(np.float32) * np.float32(0.05),
                         np.array([1.0, 0.999, 1e-30, -3.25e5], np.float32)])
    rnd = np.arange(1 << 16, dtype=np.uint32)
    for x in xs:
        out = O.f2bf_sr(np.full(1 << 16, x, np.float32), rnd).astype(np.uint32)
        xb = int(np.float32(x).view(np.uint32))
        ups = int(np.sum(out != (xb >> 16)))
        assert ups == (xb & 0xFFFF), (x, ups)
        mean = O.bf2f(out.astype(np.uint16)).astype(np.float64).mean()
        assert mean == float(x) or abs(mean - float(x)) <= abs(float(x)) * 2.0 ** -40


def test_stochastic_rounding_keeps_non_finite():
    out = O.f2bf_sr(np.array([np.inf, -np.inf, np.nan], np.float32), np.array([0xFFFF] * 3, np.uint32))
    assert list(out[:2]) == [0x7F80, 0xFF80] and (out[2] & 0x7F) != 0


def test_sr_bits_are_uniform_and_uncorrelated():
    """16-bit words over (Gaussian, step, stream, slot): uniform in mean and
    variance, no coincidences between neighbouring slots, streams or steps
    beyond chance (1/65536), and no lag-1 correlation."""
    L = O.lib()
    v = np.array([[[[L.orc_sr_bits(i, s, t, st) for s in range(16)] for st in range(5)] for t in (7, 8)]
                  for i in range(400)], np.float64)       # [gaussian, step, stream, slot]
    assert v.max() < 65536
    u = v / 65536.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1.0 / 12.0) < 0.005
    for a, b in ((v[..., 0::2], v[..., 1::2]), (v[:, :, 0], v[:, :, 1]), (v[:, 0], v[:, 1]), (v[:-1], v[1:])):
        assert np.mean(a == b) < 5e-3
    flat = u.reshape(-1) - 0.5
    assert abs(np.mean(flat[1:] * flat[:-1])) / np.var(flat) < 0.02


def test_sr_bits_of_deep_sh_slots():
    """The SH streams draw bits for up to 48 coefficient slots per (Gaussian,
    step) (degree 3, padded): the 16-bit words stay uniform (mean, variance,
    16-bin chi-square on the top bits) in every third of the slot range, and
    neighbouring slots and streams are uncorrelated."""
    L = O.lib()
    n = 300
    v = np.array([[[L.orc_sr_bits(i, c, 11, st) for c in range(48)] for st in (2, 3, 4)] for i in range(n)],
                 np.float64)                                   # [gaussian, stream, coefficient]
    u = v / 65536.0
    for depth in (slice(0, 16), slice(16, 32), slice(32, 48)):
        w = u[:, :, depth].reshape(-1)
        assert abs(w.mean() - 0.5) < 0.015 and abs(w.var() - 1.0 / 12.0) < 0.006
        hist = np.bincount((w * 16).astype(int), minlength=16)
        e = w.size / 16.0
        assert np.sum((hist - e) ** 2 / e) < 45.0              # chi-square, 15 dof: p ~ 1e-4
    d = u - 0.5
    assert abs(np.mean(d[:, :, 1:] * d[:, :, :-1])) / np.var(d) < 0.02
    assert abs(np.mean(d[:, 1:] * d[:, :-1])) / np.var(d) < 0.02


def test_mixed_adam_is_f32_adam_plus_documented_rounding():
    rng = np.random.default_rng(3)
    n = 257
    P = rng.standard_normal((n, 16)).astype(np.float32)
    P[:, 3:6] = -4.0
    G = (rng.standard_normal((n, 16)) * 0.1).astype(np.float32)
    M16 = O.f2bf_rne(rng.standard_normal((n, 16)).astype(np.float32) * np.float32(0.01))
    V16 = O.f2bf_rne(np.abs(rng.standard_normal((n, 16)).astype(np.float32)) * np.float32(1e-3))
    hp = adam_for_iter(17, 1.0)
    step = 17
    # reference: f32 Adam on the decoded moments, then stochastic rounding of m, v
    Pr, Mr, Vr = P.copy(), O.bf2f(M16), O.bf2f(V16)
    O.adam(hp, Pr, G, Mr, Vr)
    rm = np.array([[O.lib().orc_sr_bits(i, s, step, 0) for s in range(16)] for i in range(n)], np.uint32)
    rv = np.array([[O.lib().orc_sr_bits(i, s, step, 1) for s in range(16)] for i in range(n)], np.uint32)
    Pm, Mm, Vm = P.copy(), M16.copy(), V16.copy()
    O.adam_mixed(hp, Pm, G, Mm, Vm, step)
    np.testing.assert_array_equal(_bits(Pm), _bits(Pr))
    np.testing.assert_array_equal(Mm, O.f2bf_sr(Mr, rm))
    np.testing.assert_array_equal(Vm, O.f2bf_sr(Vr, rv))
    # skipped slots (lr 0: 14, 15) keep their bits
    np.testing.assert_array_equal(Mm[:, 14:], M16[:, 14:])


def test_mixed_sh_adam_is_f32_adam_plus_documented_rounding():
    rng = np.random.default_rng(4)
    n, nc = 131, 24
    S16 = O.f2bf_rne((rng.standard_normal((n, nc)) * 0.05).astype(np.float32))
    G = (rng.standard_normal((n, nc)) * 0.01).astype(np.float32)
    M16 = O.f2bf_rne((rng.standard_normal((n, nc)) * 1e-3).astype(np.float32))
    V16 = O.f2bf_rne((np.abs(rng.standard_normal((n, nc))) * 1e-6).astype(np.float32))
    hp = adam_for_iter(1234, 1.0)
    step = 1234
    Sr, Mr, Vr = O.bf2f(S16), O.bf2f(M16), O.bf2f(V16)
    O.sh_adam(hp, Sr, G, Mr, Vr)

    def r(stream):
        return np.array([[O.lib().orc_sr_bits(i, c, step, stream) for c in range(nc)] for i in range(n)], np.uint32)

    Sm, Mm, Vm = S16.copy(), M16.copy(), V16.copy()
    O.sh_adam_mixed(hp, Sm, G, Mm, Vm, step)
    np.testing.assert_array_equal(Sm, O.f2bf_sr(Sr, r(2)))
    np.testing.assert_array_equal(Mm, O.f2bf_sr(Mr, r(3)))
    np.testing.assert_array_equal(Vm, O.f2bf_sr(Vr, r(4)))
