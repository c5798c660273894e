"""Shared helpers for the parity tests (structures, comparisons)."""
import functools

import numpy as np

from parphom_amd import synth
from oracle import phonon as O


@functools.lru_cache(maxsize=8)
def structure(m, n, seed):
    st = synth.tbg_structure(m, n)
    fc = synth.force_constants(st, seed)
    return st, fc


@functools.lru_cache(maxsize=8)
def oracle_system(m, n, seed):
    st, fc = structure(m, n, seed)
    return O.System(st.lat, st.pos, st.types, st.masses, fc)


def clusters(lam, rel):
    """Index groups of consecutive eigenvalues closer than rel * max|lam|."""
    lam = np.asarray(lam)
    tol = rel * max(1e-300, np.abs(lam).max())
    groups, cur = [], [0]
    for k in range(1, len(lam)):
        if lam[k] - lam[k - 1] <= tol:
            cur.append(k)
        else:
            groups.append(cur)
            cur = [k]
    groups.append(cur)
    return groups


def vel_projected(Z, dD_list, om_cm, groups):
    """Gauge-invariant velocities: for each cluster, the sorted eigenvalues of
    Z_c^H dD Z_c scaled as write_output.F90:443-446 scales the diagonal."""
    out = np.zeros((Z.shape[1], 3))
    for a, dD in enumerate(dD_list):
        Y = dD @ Z
        for g in groups:
            blk = Z[:, g].conj().T @ Y[:, g]
            ev = np.sort(np.linalg.eigvalsh(0.5 * (blk + blk.conj().T)))
            den = np.maximum(2.0 * np.abs(om_cm[g]), 1e-3)
            out[g, a] = ((((ev / den) * O.C2) * O.C1SQ32) * 100.0) * 2.0 * O.PI32
    return out


def vel_cluster_sums(v, groups):
    return np.array([v[g].sum(axis=0) for g in groups])


def vel_noise(dD_list, om_cm, lam, groups, kappa=4.0, eps=None):
    """Conditioning of the analytic velocity of each mode in FP64.  A backward
    stable eigensolver returns e with errors ~ eps |D| / gap along the other
    eigenvectors (gap = distance to the nearest eigenvalue outside the mode's
    cluster), so <e|dD|e> is determined to ~ 2 eps |D| |dD| / gap; scaled by
    the velocity factor / max(2|omega|, 1e-3).  eps defaults to the FP64 unit
    roundoff; the FP32 path's vectors carry the FP32 one."""
    eps = np.finfo(float).eps if eps is None else eps
    lam = np.asarray(lam)
    n = len(lam)
    dnorm = max(np.linalg.norm(d, 2) for d in dD_list)
    Dnorm = np.abs(lam).max()
    gap = np.full(n, np.inf)
    for g in groups:
        lo, hi = g[0], g[-1]
        d = np.inf
        if lo > 0:
            d = min(d, lam[lo] - lam[lo - 1])
        if hi < n - 1:
            d = min(d, lam[hi + 1] - lam[hi])
        gap[g] = d
    gap = np.maximum(gap, np.finfo(float).eps * Dnorm)
    den = np.maximum(2.0 * np.abs(om_cm), 1e-3)
    const = O.C2 * O.C1SQ32 * 100.0 * 2.0 * O.PI32
    return kappa * 2.0 * n * eps * Dnorm * dnorm / gap * const / den


def assert_vel_close(v, v_ref, groups, noise, rel=1e-6, floor=1e-9):
    """Per-cluster comparison: |a - b| <= rel |b| + floor + noise."""
    for g in groups:
        a = v[g].sum(axis=0)
        b = v_ref[g].sum(axis=0)
        tol = rel * np.abs(v_ref[g]).sum(axis=0) + floor + noise[g].sum()
        assert np.all(np.abs(a - b) <= tol), (g, a, b, tol)
