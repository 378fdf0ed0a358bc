"""CPU restatement of the reference's MCMC-lite refine (src/gs/trainer.js:
666-794) on full state arrays, in the reference's arithmetic: JS doubles
for every intermediate, float32 storage (Float32Array stores round to
nearest). Test infrastructure: only tests/ use it, for the expected results
of the trainer's gs_refine (host planning on 16 B per Gaussian + device apply).

The random stream is the reference's makeRng (src/sfm/geometry.js:750-758).
Math.log / Math.exp / Math.cos / Math.sqrt are taken from Python's math
module, which uses the same libm as the trainer's host code; V8's own
implementations may differ from it in the last ulp (parity unpinned there).
"""
import math

import numpy as np

STRIDE = 16


def make_rng(seed):
    """makeRng (geometry.js:750-758): mulberry32, doubles in [0, 1)."""
    a = seed & 0xFFFFFFFF

    def rng():
        nonlocal a
        a = (a + 0x6D2B79F5) & 0xFFFFFFFF                         # a = (a + C) | 0
        t = ((a ^ (a >> 15)) * (a | 1)) & 0xFFFFFFFF              # Math.imul(a ^ a >>> 15, 1 | a)
        t = ((t + (((t ^ (t >> 7)) * (t | 61)) & 0xFFFFFFFF)) & 0xFFFFFFFF) ^ t
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0

    return rng


def _f32(x):
    return np.float32(x)


def refine(state, n, cap, it, horizon, rng, grow_rate=None, grow_until=None, grow_limit=0):
    """state: dict of row arrays P (cap x 16 f32), M, V (cap x 16), and when
    SH is on SH, SM, SV (cap x 3K); moment/SH rows are copied or zeroed as
    whole rows, so any dtype works (f32, or uint16 bf16 bits). Mutates the
    arrays; returns (moved, grown, new n)."""
    P, M, V = state["P"], state["M"], state["V"]
    SH, SM, SV = state.get("SH"), state.get("SM"), state.get("SV")

    def sig(x):
        return 1.0 / (1.0 + math.exp(-x))

    dead, donors = [], []
    for i in range(n):
        o = sig(float(P[i, 13]))
        if o < 0.02:
            dead.append(i)
        elif o > 0.4:
            donors.append(i)
    if len(donors) < 16:
        return 0, 0, n
    move_cap = math.ceil(n * 0.05)
    if len(dead) > move_cap:
        dead = dead[:move_cap]

    def gauss():
        u = max(1e-9, rng())
        w = rng()
        return math.sqrt(-2 * math.log(u)) * math.cos(2 * math.pi * w)

    def mean_log_scale(i):
        return (float(P[i, 3]) + float(P[i, 4]) + float(P[i, 5])) / 3

    # Array.prototype.sort is stable; comparator (a, b) => mls(b) - mls(a)
    big = sorted(donors, key=lambda i: -mean_log_scale(i))
    big = big[:max(16, len(donors) >> 2)]

    def spawn_at(bi, allow_split):
        do_split = allow_split and len(big) > 16 and rng() < 0.7
        if do_split:
            di = int(rng() * len(big))
            don = big[di]
            big[di] = big[-1]
            big.pop()
        else:
            don = donors[int(rng() * len(donors))]
        s = math.exp(mean_log_scale(don))
        for c in range(3):
            P[bi, c] = _f32(float(P[don, c]) + gauss() * s * 0.7)
        P[bi, 3:13] = P[don, 3:13]
        if do_split:
            shrink = math.log(1 / 1.6)
            for k in range(3, 6):
                P[bi, k] = _f32(float(P[bi, k]) + shrink)
                P[don, k] = _f32(float(P[don, k]) + shrink)
            P[bi, 13] = P[don, 13]
            M[don] = 0
            V[don] = 0
        else:
            for k in range(3, 6):
                P[bi, k] = _f32(float(P[bi, k]) + math.log(0.85))
            P[bi, 13] = _f32(math.log(0.25 / 0.75))
        P[bi, 14] = 0
        P[bi, 15] = 0
        M[bi] = 0
        V[bi] = 0
        if SH is not None:
            SH[bi] = SH[don]
            SM[bi] = 0
            SV[bi] = 0
            if do_split:
                SM[don] = 0
                SV[don] = 0

    for i in dead:
        spawn_at(i, False)
    limit = min(cap, grow_limit or cap)
    until = 0.75 * horizon if grow_until is None else grow_until
    rate = 0.15 if grow_rate is None else grow_rate
    grown = max(0, min(math.ceil(n * rate), limit - n)) if it < until else 0
    for k in range(grown):
        spawn_at(n + k, True)
    return len(dead), grown, n + grown
