"""Seeded synthetic inputs: counter-based Gaussian streams and U(1) sector profiles.

Everything here is host-side input generation; no device work.  The same
flat buffers feed the device tensors and any CPU checker, so both sides see
bit-identical inputs.

* ``gaussian(seed, stream, n)``: SplitMix64 counter hash -> Box-Muller
  (Steele, Lea & Flood 2014 for SplitMix64; Box & Muller 1958).
* ``gaussian_profile(...)``: bond sector dims ``max(1, round(dmax*exp(-q^2/2s^2)))``
  over consecutive charges, with ``s`` bisected so the total matches chi.
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform_u64(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n 64-bit words: SplitMix64 output at counters offset..offset+n-1 of (seed, stream)."""
    with np.errstate(over="ignore"):
        base = _mix(np.array([(seed * 0x2545F4914F6CDD1D + stream * 0x632BE59BD9B4E019) & (2**64 - 1)],
                             dtype=np.uint64))[0]
        ctr = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        return _mix(base + ctr * _GOLDEN)


def gaussian(seed: int, stream: int, n: int) -> np.ndarray:
    """n standard normals (float64), deterministic in (seed, stream)."""
    m = (n + 1) // 2
    w = uniform_u64(seed, stream, 2 * m)
    u1 = ((w[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (w[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log(u1))
    g = np.empty(2 * m)
    g[0::2] = r * np.cos(2.0 * math.pi * u2)
    g[1::2] = r * np.sin(2.0 * math.pi * u2)
    return g[:n]


def gaussian_profile(nsec: int, chi: int, dmax: int = None, center: float = None) -> List[int]:
    """Sector dims over ``nsec`` consecutive charges with a Gaussian profile summing to ~chi."""
    if nsec <= 0:
        return []
    if chi < nsec:
        raise ValueError("chi must be at least the number of sectors")
    c = (nsec - 1) / 2.0 if center is None else center

    def dims(sig, dm):
        return [max(1, int(round(dm * math.exp(-((q - c) ** 2) / (2 * sig * sig))))) for q in range(nsec)]

    if dmax is None:
        # choose the peak so the profile is a plain Gaussian of total chi
        lo, hi = 1.0, float(chi)
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            s = sum(dims(max(nsec / 6.0, 0.5), mid))
            lo, hi = (mid, hi) if s < chi else (lo, mid)
        return dims(max(nsec / 6.0, 0.5), hi)
    lo, hi = 1e-3, 10.0 * nsec
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        s = sum(dims(mid, dmax))
        lo, hi = (mid, hi) if s < chi else (lo, mid)
    return dims(hi, dmax)


def binom_capped_profile(l: int, L: int, N: int, chi: int, dmax: int = None, nsec: int = None,
                         shift: int = 0) -> List[Tuple[int, int]]:
    """Charge->dim pairs for bond ``l`` of an L-site chain at total charge N.

    Allowed charges q in [max(0, N-(L-l)), min(l, N)] (C9: charge = up count);
    dims follow a Gaussian profile around the middle allowed charge, capped by
    the Hilbert-space dims binomial(l, q)*binomial(L-l, N-q) of each side.
    """
    qmin, qmax = max(0, N - (L - l)), min(l, N)
    qs = list(range(qmin, qmax + 1))
    if nsec is not None and nsec < len(qs):
        mid = (qmin + qmax) // 2 + shift
        lo = max(qmin, min(mid - nsec // 2, qmax - nsec + 1))
        qs = list(range(lo, lo + nsec))
    ds = gaussian_profile(len(qs), max(chi, len(qs)), dmax)
    out = []
    for q, d in zip(qs, ds):
        cap = min(math.comb(l, q), math.comb(L - l, N - q))
        if cap > 0:
            out.append((q, min(d, cap)))
    return out


# ---------------------------------------------------------------------------
# Named H_eff workloads (one two-site bond with its environments)
# ---------------------------------------------------------------------------
HEFF_WORKLOADS = {
    # XXZ chain L=32, chi=64, D=5 (middle bond)
    "xxz_L32_chi64": dict(nsec=13, chi=64, dmax=None, mpo_dims=((-1, 1), (0, 3), (1, 1))),
    # XXZ chain L=100, chi=512, ~30 sectors, dims 1-60, D=5
    "xxz_L100_chi512": dict(nsec=30, chi=512, dmax=60, mpo_dims=((-1, 1), (0, 3), (1, 1))),
    # Heisenberg 6x24 cylinder, chi=2048, ~60 sectors, dims 1-200, D=20 {-1:6, 0:8, +1:6}
    "cyl6x24_chi2048": dict(nsec=60, chi=2048, dmax=200, mpo_dims=((-1, 6), (0, 8), (1, 6))),
    # XXZ chain L=64, chi=256, ~21 sectors, D=5
    "xxz_L64_chi256": dict(nsec=21, chi=256, dmax=None, mpo_dims=((-1, 1), (0, 3), (1, 1))),
    # t-V chain L=128, chi=4096, ~70 sectors dims <=400, D=6
    "tv_L128_chi4096": dict(nsec=70, chi=4096, dmax=400, mpo_dims=((-1, 2), (0, 2), (1, 2))),
    # the reference's own benchmark shape (benchmarkings.ipynb:48-73): chi=1000, D=15
    "refbench_chi1000": dict(explicit=True,
                             left=((2, 6), (3, 52), (4, 167), (5, 275), (6, 275), (7, 167), (8, 52), (9, 6)),
                             right=((3, 8), (4, 62), (5, 178), (6, 254), (7, 252), (8, 176), (9, 62), (10, 8)),
                             mpo_dims=((-1, 4), (0, 7), (1, 4))),
}


def heff_spaces(name: str):
    """Leg maps for a named H_eff workload: (left bond, right bond, mpo bond) charge->dim pairs.

    Layout follows the two-site DMRG conventions: v = (Vl, Vd, Vd, Vr†),
    envL = (Vl†, Vw†, Vl), envR = (Vr, Vw, Vr†), W = (Vw, Vd, Vw†, Vd†).  The
    right bond is the left profile shifted by one charge (two sites add 0..2).
    """
    w = HEFF_WORKLOADS[name]
    if w.get("explicit"):
        return list(w["left"]), list(w["right"]), list(w["mpo_dims"])
    dims = gaussian_profile(w["nsec"], w["chi"], w["dmax"])
    q0 = 0
    left = [(q0 + i, d) for i, d in enumerate(dims)]
    right = [(q0 + 1 + i, d) for i, d in enumerate(dims)]
    return left, right, list(w["mpo_dims"])
