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
