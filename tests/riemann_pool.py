"""The known-answer pool of the reference's Riemann benchmark (cpp/apps/bench_riemann.cpp:20-117): 65536 face
pairs from its LCG (seed 88172645463325252) with the edge cases it puts first, and its FNV-1a fingerprint of the
fluxes and wave speeds. Restated in plain Python floats (IEEE double, no fused operations)."""
from __future__ import annotations

import math
import struct

import numpy as np

MASK = (1 << 64) - 1
POOL = 1 << 16
# fingerprints printed by the reference's `bench_riemann 65536` (shock_blend = 0), identical with and without SIMD
HASHES = {("ideal", "hllc"): 0x6c20aeb15662ba47, ("ideal", "lowmach"): 0xf854d6c543539129,
          ("stiffened", "hllc"): 0x9be8bdb298df5924, ("stiffened", "lowmach"): 0x025f866d129ea3f5}


class Lcg:
    def __init__(self):
        self.s = 88172645463325252

    def next(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & MASK
        return float(self.s >> 11) * (1.0 / 9007199254740992.0)


def p2c(W, gamma, p_inf, stiff):
    rho, u, v, p, w = W
    ke = 0.5 * rho * (u * u + v * v + w * w)
    e = (p + gamma * p_inf) / (gamma - 1.0) if stiff else p / (gamma - 1.0)
    return [rho, rho * u, rho * v, e + ke, rho * w]


def pool(kind: str):
    stiff = kind == "stiffened"
    gamma, p_inf = (4.4, 6.0e8) if stiff else (1.4, 0.0)
    rng = Lcg()
    UL, UR, axis = [None] * POOL, [None] * POOL, [0] * POOL
    p0 = 1.0e8 if stiff else 1.0
    for k in range(POOL):
        strong = rng.next() < 0.1
        a = [1.0 + 0.2 * rng.next(), 2.0 * rng.next() - 1.0, 2.0 * rng.next() - 1.0, p0 * (1.0 + 0.2 * rng.next()),
             2.0 * rng.next() - 1.0]
        b = list(a)
        jump = 4.0 if strong else 0.02
        b[0] *= 1.0 + jump * rng.next()
        b[1] += jump * (rng.next() - 0.5)
        b[2] += jump * (rng.next() - 0.5)
        b[3] *= 1.0 + jump * rng.next()
        b[4] += 0.1 * jump * (rng.next() - 0.5)
        UL[k] = p2c(a, gamma, p_inf, stiff)
        UR[k] = p2c(b, gamma, p_inf, stiff)
        axis[k] = int(rng.next() * 3.0) % 3
    k = 0

    def put(l, r, ax):
        nonlocal k
        UL[k], UR[k], axis[k] = l, r, ax
        k += 1

    P = lambda W: p2c(W, gamma, p_inf, stiff)
    for ax in range(3):
        a = P([1.0, 0.3, -0.2, p0, 0.1])
        put(a, a, ax)
        put(P([1.0, 0.3, -0.2, p0, 0.1]), P([1.0, -0.3, -0.2, p0, 0.1]), ax)
        put(P([1.0, 0.0, 0.0, p0, 0.0]), P([0.125, -0.0, 0.0, 0.1 * p0, -0.0]), ax)
        put(P([1.0, 0.0, 0.0, p0, 0.0]), P([1.0, 0.0, 0.0, 2.0 * p0, 0.0]), ax)
        put([1e-20, 0.0, 0.0, 1e-20, 0.0], [1.0, 0.0, 0.0, p0, 0.0], ax)
        put([1.0, 5.0, 0.0, 0.1, 0.0], [1.0, 5.0, 0.0, 0.1, 0.0], ax)
        put([-1.0, 1.0, 0.0, 1.0, 0.0], [1.0, 0.0, 0.0, p0, 0.0], ax)
        for ulp in range(-2, 3):
            c = math.sqrt(gamma * (p0 + (p_inf if stiff else 0.0)) / 1.0)
            u = c
            for _ in range(abs(ulp)):
                u = math.nextafter(u, 0.0 if ulp < 0 else 1e300)
            W = [1.0, 0.0, 0.0, p0, 0.0]
            W[(1, 2, 4)[ax]] = u
            Wr = list(W)
            Wr[3] *= 1.0 + 1e-3
            put(P(W), P(Wr), ax)
            put(P(Wr), P(W), ax)
    return np.array(UL), np.array(UR), np.array(axis, dtype=np.int32), gamma, p_inf, stiff


def fingerprint(F, S):
    """FNV-1a over (F[k][0..4], S[k]) in the reference's order (bench_riemann.cpp:108-111), one pass."""
    bits = np.concatenate([np.ascontiguousarray(F, np.float64), np.ascontiguousarray(S, np.float64)[:, None]], axis=1)
    words = bits.view(np.uint64).ravel().tolist()
    h = 1469598103934665603
    for wv in words:
        h = ((h ^ wv) * 1099511628211) & MASK
    return h
