"""Shared test helpers: conversions between the product and the oracle."""
import os

import numpy as np

from oracle import symtensors as ost


def o_space(V):
    """product U1Space -> oracle Space"""
    return ost.Space(list(V.items()), V.isdual)


def p_space(V):
    from mps_amd.symtensors import U1Space
    return U1Space(*V.items(), isdual=V.isdual)


def o_from_flat(charge, spaces, flat):
    """Oracle SymTensor whose blocks are consecutive column-major slices of ``flat``."""
    sects, sizes = ost.allsectorsandsizes(charge, spaces)
    blocks, p = {}, 0
    for s, z in zip(sects, sizes):
        n = int(np.prod(z))
        blocks[s] = np.asarray(flat[p:p + n], dtype=np.float64).reshape(z, order="F")
        p += n
    assert p == len(flat)
    return ost.SymTensor(charge, spaces, blocks, check=True)


def o_flat(T):
    """Oracle SymTensor -> packed flat buffer (sorted sectors, column-major blocks)."""
    parts = [np.asarray(b, dtype=np.float64).reshape(-1, order="F") for b in T.blocks.values()]
    return np.concatenate(parts) if parts else np.zeros(0)


def o_from_product(P):
    """Device SymTensor -> oracle SymTensor (via host copy)."""
    return o_from_flat(P.charge, tuple(o_space(V) for V in P.space), P.data.detach().cpu().double().numpy())


def random_space(rng, nmin=1, nmax=4, cmin=-3, cmax=3, dmax=4, isdual=None):
    n = int(rng.integers(nmin, nmax + 1))
    cs = sorted(rng.choice(np.arange(cmin, cmax + 1), size=n, replace=False).tolist())
    ds = rng.integers(1, dmax + 1, size=n).tolist()
    d = bool(rng.integers(0, 2)) if isdual is None else isdual
    return ost.Space(list(zip(cs, ds)), d)


def rel_block_err(A, B):
    """max over blocks of ||A_b - B_b|| / max(||B_b||, tiny) (norm-wise per block)."""
    worst = 0.0
    for k in B.blocks:
        a, b = np.asarray(A.blocks[k], dtype=np.float64), np.asarray(B.blocks[k], dtype=np.float64)
        nb = np.linalg.norm(b)
        d = np.linalg.norm(a - b)
        worst = max(worst, d / nb if nb > 0 else d)
    return worst


def global_rel_err(A, B):
    fa, fb = o_flat(A), o_flat(B)
    return float(np.linalg.norm(fa - fb) / max(np.linalg.norm(fb), 1e-300))


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_fixture(name):
    """Golden fixture -> ({tensor name: oracle SymTensor}, extra arrays, meta dict).

    npz written by tests/golden/make_golden.py; loaded without pickle."""
    import json
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    tensors = {}
    for k, m in meta.items():
        if not isinstance(m, dict) or "legs" not in m:
            continue
        spaces = tuple(ost.Space([tuple(p) for p in pairs], bool(d)) for d, pairs in m["legs"])
        tensors[k] = o_from_flat(m["charge"], spaces, z[f"flat_{k}"])
    extra = {k: z[k] for k in z.files if k != "meta" and not k.startswith("flat_")}
    return tensors, extra, meta
