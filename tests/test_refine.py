"""MCMC-lite refine (trainer.js:666-794).

CPU: the restatement (oracle/refine.py) — its mulberry32 against a literal
JS-semantics emulation (signed ToInt32 steps), and the structural effects of
one call on a random state. GPU: the trainer's gs_refine (16 B per Gaussian
to the host, plan replayed there, spawns applied on the device) against the
restatement on the full state, bit for bit, in f32 and mixed precision, then
two more training steps on both sides (which pins the moment resets)."""
import math

import numpy as np
import pytest

import gsabi
import refine as R
from helpers import OracleTrainer, config1_setup


def _i32(x):
    x &= 0xFFFFFFFF
    return x - (1 << 32) if x & 0x80000000 else x


def _imul(a, b):
    return _i32((a & 0xFFFFFFFF) * (b & 0xFFFFFFFF))


def _js_make_rng(seed):
    """geometry.js:750-758 step by step with JS's signed 32-bit semantics."""
    a = _i32(seed & 0xFFFFFFFF)

    def rng():
        nonlocal a
        a = _i32(a)
        a = _i32(a + 0x6D2B79F5)
        t = _imul(a ^ ((a & 0xFFFFFFFF) >> 15), 1 | a)
        t = _i32((t + _imul(t ^ ((t & 0xFFFFFFFF) >> 7), 61 | t))) ^ t
        return ((t ^ ((t & 0xFFFFFFFF) >> 14)) & 0xFFFFFFFF) / 4294967296
    return rng


def test_make_rng_matches_js_semantics():
    for seed in (0, 1, 7, 0xDEADBEEF, 2 ** 32 - 1):
        a, b = R.make_rng(seed), _js_make_rng(seed)
        for _ in range(2000):
            x, y = a(), b()
            assert x == y and 0.0 <= x < 1.0


def _random_state(n, cap, nc, seed=0):
    rng = np.random.default_rng(seed)
    P = np.zeros((cap, 16), np.float32)
    P[:n, :3] = rng.uniform(-1, 1, (n, 3))
    P[:n, 3:6] = rng.normal(-4, 0.5, (n, 3))
    P[:n, 6:10] = rng.normal(0, 1, (n, 4))
    P[:n, 10:13] = rng.normal(0, 1, (n, 3))
    P[:n, 13] = rng.normal(0, 2.5, n)
    st = {"P": P, "M": np.zeros((cap, 16), np.float32), "V": np.zeros((cap, 16), np.float32),
          "SH": np.zeros((cap, nc), np.float32), "SM": np.zeros((cap, nc), np.float32),
          "SV": np.zeros((cap, nc), np.float32)}
    st["M"][:n] = rng.normal(0, 1, (n, 16))
    st["V"][:n] = 1.0
    st["SH"][:n] = rng.normal(0, 0.05, (n, nc))
    st["SM"][:n] = 1.0
    st["SV"][:n] = 1.0
    return st


def test_refine_restatement_structure():
    n, cap, nc = 4000, 8000, 24
    st = _random_state(n, cap, nc)
    before = {k: v.copy() for k, v in st.items()}
    opa = 1 / (1 + np.exp(-before["P"][:n, 13].astype(np.float64)))
    dead = np.nonzero(opa < 0.02)[0]
    moved, grown, n2 = R.refine(st, n, cap, 100, 60000, R.make_rng(3))
    assert moved == min(len(dead), math.ceil(0.05 * n)) and moved > 0
    assert grown == math.ceil(0.15 * n) and n2 == n + grown
    P, P0 = st["P"], before["P"]
    spawned = np.concatenate([dead[:moved], np.arange(n, n2)])
    untouched = np.setdiff1d(np.arange(n), spawned)
    donors = np.nonzero(opa > 0.4)[0]
    # dead Gaussians past the move cap keep their rows
    np.testing.assert_array_equal(P[dead[moved:]], P0[dead[moved:]])
    # every spawn: rotation and colour of some donor, fresh moments, pad slots zero
    rows = P0[donors, 6:13]
    for t in spawned:
        assert (rows == P[t, 6:13]).all(axis=1).any()
        assert not st["M"][t].any() and not st["V"][t].any() and not st["SM"][t].any()
        assert P[t, 14] == 0 and P[t, 15] == 0
    # relocated dead splats are clones: opacity logit ln(1/3)
    assert np.all(P[dead[:moved], 13] == np.float32(math.log(0.25 / 0.75)))
    # a split shrinks its donor's log-scales by ln(1/1.6) and resets its moments;
    # other live Gaussians keep everything
    changed = [i for i in untouched if not np.array_equal(P[i], P0[i])]
    assert changed, "growth should split some donors"
    for i in changed:
        np.testing.assert_allclose(P[i, 3:6], P0[i, 3:6] + math.log(1 / 1.6), rtol=0, atol=1e-6)
        assert not st["M"][i].any() and not st["SM"][i].any()
        assert opa[i] > 0.4
    # splits keep the donor's opacity on the new half, clones take ln(1/3)
    new = P[n:n2, 13]
    assert np.any(new == np.float32(math.log(0.25 / 0.75))) and np.any(new != np.float32(math.log(0.25 / 0.75)))


def test_refine_needs_sixteen_donors():
    n, cap, nc = 100, 200, 0
    st = _random_state(n, cap, nc)
    st["P"][:n, 13] = -9.0
    st["P"][:10, 13] = 3.0
    assert R.refine(st, n, cap, 0, 60000, R.make_rng(1)) == (0, 0, n)


def _scripted(values):
    """An rng that returns the given values in order (and records how many
    were drawn), standing in for makeRng so the test chooses every random
    decision itself."""
    vals = list(values)
    drawn = []

    def rng():
        assert vals, "the refine drew more random numbers than the script holds"
        drawn.append(vals.pop(0))
        return drawn[-1]
    rng.left = vals
    return rng


def test_refine_hand_worked_donor_choice():
    """A case worked by hand from trainer.js:687-779, with the random stream
    scripted so that each choice is known in advance.

    n = 100, cap = 200. Gaussians 0-79 are donors (opacity logit 3, sigma
    0.95 > 0.4), 80-97 neither (logit -2, sigma 0.12), 98-99 dead (logit -6,
    sigma 0.0025 < 0.02; moveCap = ceil(5) holds both). Donor i has all three
    log-scales -3 - 0.01 * ((37 i) mod 80), so its size rank is (37 i) mod 80
    (0 = largest) and rank r belongs to donor (13 r) mod 80 (13 = 37^-1 mod
    80). Donor 79 (rank 43) is given donor 0's log-scales: a tie for the
    largest, which the stable sort keeps in index order. bigDonors is the top
    max(16, 80 >> 2) = 20, largest first:
        [0, 79, 13, 26, 39, 52, 65, 78, 11, 24, 37, 50, 63, 76, 9, 22, 35, 48, 61, 74]
    (rank 19's donor, 7, falls off behind the tie).
    Relocation (no split draw): dead 98 draws 0.5 -> donors[40] = 40; dead 99
    draws 0.999 -> donors[79] = 79. Growth: ceil(100 * 0.15) = 15 spawns at
    rows 100-114.
      k=0: 0.1 < 0.7 split, 0.0 -> bigDonors[0] = 0; 74 moves into slot 0
      k=1: split, 0.0 -> 74 (the swap-popped donor); 61 moves into slot 0
      k=2: 0.9 no split, 0.0 -> donors[0] = 0, a clone of the shrunk donor 0
      k=3: split, 0.999 -> bigDonors[floor(0.999 * 18)] = bigDonors[17] = 48 (the last)
      k=4: split, 0.05 -> bigDonors[floor(0.05 * 17)] = bigDonors[0] = 61
      k=5..14: 16 left, no split draw; 0.3 -> donors[24] = 24
    Every spawn then draws three Gaussians of two values each: (0.5, 0.0)
    gives sqrt(-2 ln 0.5) cos(0) = sqrt(2 ln 2).
    With the big-donor sort reversed, or the swap-pop replaced by a removal
    in place, the split donors differ."""
    n, cap = 100, 200
    P = np.zeros((cap, 16), np.float32)
    idx = np.arange(n)
    P[:n, 0] = idx * 0.01
    P[:n, 1] = -idx * 0.02
    P[:n, 2] = 4.0 + idx * 0.001
    rank = (37 * idx) % 80
    ls = np.where(idx < 80, -3.0 - 0.01 * rank, -4.0).astype(np.float32)
    ls[79] = ls[0]
    P[:n, 3:6] = ls[:, None]
    P[:n, 6] = 1.0 + idx          # unique rotation and colour per Gaussian: they identify the donor
    P[:n, 7:10] = 0.1
    P[:n, 10] = idx * 0.03
    P[:n, 11:13] = -0.5
    P[:n, 13] = np.where(idx < 80, 3.0, np.where(idx < 98, -2.0, -6.0))
    M = np.ones((cap, 16), np.float32)
    V = np.ones((cap, 16), np.float32)
    st = {"P": P, "M": M, "V": V}
    P0 = P.copy()
    gauss3 = [0.5, 0.0] * 3
    script = ([0.5] + gauss3 + [0.999] + gauss3 +
              [0.1, 0.0] + gauss3 + [0.1, 0.0] + gauss3 + [0.9, 0.0] + gauss3 +
              [0.1, 0.999] + gauss3 + [0.1, 0.05] + gauss3 +
              ([0.3] + gauss3) * 10)
    rng = _scripted(script)
    moved, grown, n2 = R.refine(st, n, cap, 100, 60000, rng)
    assert (moved, grown, n2) == (2, 15, 115)
    assert rng.left == [], "every scripted value is drawn, no more"

    def donor_of(row):
        return int(round(P[row, 6] - 1.0))

    assert donor_of(98) == 40 and donor_of(99) == 79
    spawned = [donor_of(100 + k) for k in range(15)]
    assert spawned == [0, 74, 0, 48, 61] + [24] * 10
    splits = [0, 74, 48, 61]
    shrink, clone = math.log(1 / 1.6), math.log(0.85)
    g = math.sqrt(-2 * math.log(0.5))

    def f32(x):
        return float(np.float32(x))

    # split donors: log-scales shrunk once, moments cleared, opacity kept
    for d in splits:
        assert np.all(P[d, 3:6] == np.float32(f32(float(P0[d, 3]) + shrink)))
        assert not M[d].any() and not V[d].any()
        assert P[d, 13] == P0[d, 13]
    # untouched Gaussians keep their rows and moments
    for i in set(range(98)) - set(splits):
        assert np.array_equal(P[i], P0[i]) and M[i].all()
    for k, d in enumerate([0, 74, 0, 48, 61] + [24] * 10):
        row = 100 + k
        split = k in (0, 1, 3, 4)
        # a clone of donor 0 at k=2 reads its shrunk log-scales
        base = float(P[d, 3]) if not split else f32(float(P0[d, 3]))
        want_ls = f32(base + shrink) if split else f32(base + clone)
        assert np.all(P[row, 3:6] == np.float32(want_ls)), (k, P[row, 3:6], want_ls)
        want_opa = P0[d, 13] if split else np.float32(math.log(0.25 / 0.75))
        assert P[row, 13] == want_opa
        assert np.array_equal(P[row, 6:13], P0[d, 6:13])
        assert P[row, 14] == 0 and P[row, 15] == 0 and not M[row].any() and not V[row].any()
        # jitter: s = exp(mean log-scale of the donor when it spawns), 0.7 g per axis
        # (before a split shrinks the donor; a clone of donor 0 sees it shrunk)
        s = math.exp(float(P0[d, 3])) if split else math.exp(float(P[d, 3]))
        np.testing.assert_allclose(P[row, :3], P0[d, :3] + g * s * 0.7, rtol=1e-6, atol=1e-7)
    for row, d in ((98, 40), (99, 79)):
        assert np.all(P[row, 3:6] == np.float32(f32(float(P0[d, 3]) + clone)))
        assert P[row, 13] == np.float32(math.log(0.25 / 0.75))
    # rows past the new n are untouched
    assert not P[n2:].any()


# ---------------------------------------------------------------- GPU ---

def _pad(a, cap):
    out = np.zeros((cap,) + a.shape[1:], a.dtype)
    out[:a.shape[0]] = a
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f32", "mixed", "bf16", "fp16"])
def test_gpu_refine_matches_restatement(precision):
    p, sh, cams, images, W, H, radius = config1_setup(n=6000)
    p = p.copy()
    rng = np.random.default_rng(11)
    p[rng.random(p.shape[0]) < 0.1, 13] = -5.0          # some dead splats
    n = p.shape[0]
    prec = {"mixed": gsabi.GS_PREC_MIXED, "bf16": gsabi.GS_PREC_BF16,
            "fp16": gsabi.GS_PREC_FP16}.get(precision, gsabi.GS_PREC_F32)
    t = gsabi.GSTrainer(shDeg=2, maxSplats=4 * n, precision=prec)
    t.setup({"data": p, "n": n, "sh": sh}, cams, images, W, H, radius)
    ref = OracleTrainer(p, sh, cams, images, radius, shDeg=2, precision=precision)
    t.iter = ref.iter = 2500
    for _ in range(2):
        t.step_debug(0)
        ref.step(0)
    cap = t.info().cap
    assert cap == 4 * n
    r = t.refine(seed=12345)
    if precision == "mixed":
        st = {"P": _pad(ref.P, cap), "M": _pad(ref.M16, cap), "V": _pad(ref.V16, cap),
              "SH": _pad(ref.SH16, cap), "SM": _pad(ref.SM16, cap), "SV": _pad(ref.SV16, cap)}
    elif precision in ("bf16", "fp16"):
        st = {"P": _pad(ref.P, cap), "M": _pad(ref.M, cap), "V": _pad(ref.V, cap),
              "SH": _pad(ref.SH16, cap), "SM": _pad(ref.SM, cap), "SV": _pad(ref.SV, cap)}
    else:
        st = {"P": _pad(ref.P, cap), "M": _pad(ref.M, cap), "V": _pad(ref.V, cap),
              "SH": _pad(ref.SH, cap), "SM": _pad(ref.SM, cap), "SV": _pad(ref.SV, cap)}
    moved, grown, n2 = R.refine(st, n, cap, ref.iter, 60000, R.make_rng(12345))
    assert (r["moved"], r["grown"], r["n"]) == (moved, grown, n2)
    assert moved > 0 and grown == math.ceil(0.15 * n)
    ref.P = st["P"][:n2].copy()
    if precision == "mixed":
        ref.M16, ref.V16 = st["M"][:n2].copy(), st["V"][:n2].copy()
        ref.SH16, ref.SM16, ref.SV16 = st["SH"][:n2].copy(), st["SM"][:n2].copy(), st["SV"][:n2].copy()
        import oracle as O
        ref.SH = O.bf2f(ref.SH16)
    elif precision in ("bf16", "fp16"):
        ref.M, ref.V = st["M"][:n2].copy(), st["V"][:n2].copy()
        ref.SH16, ref.SM, ref.SV = st["SH"][:n2].copy(), st["SM"][:n2].copy(), st["SV"][:n2].copy()
        ref.SH = ref._dec(ref.SH16)
    else:
        ref.M, ref.V = st["M"][:n2].copy(), st["V"][:n2].copy()
        ref.SH, ref.SM, ref.SV = st["SH"][:n2].copy(), st["SM"][:n2].copy(), st["SV"][:n2].copy()
    out = t.readGaussians()
    np.testing.assert_array_equal(out["data"].view(np.uint32), ref.P.view(np.uint32))
    np.testing.assert_array_equal(out["sh"].view(np.uint32), ref.SH.view(np.uint32))
    # training continues identically (moments were reset the same way)
    for s in range(2):
        t.step_debug(0)
        ref.step(0)
        out = t.readGaussians()
        np.testing.assert_array_equal(out["data"].view(np.uint32), ref.P.view(np.uint32), err_msg=f"step {s}")
        np.testing.assert_array_equal(out["sh"].view(np.uint32), ref.SH.view(np.uint32), err_msg=f"step {s}")
    t.destroy()
