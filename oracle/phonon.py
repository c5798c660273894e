"""FP64 numpy restatement of parphom.x's phonon path (oracle; test use only).

Reference files cited below are under Phonon_Calculations/SRC/.
"""
from __future__ import annotations

import numpy as np

# f32-rounded literals exactly as the Fortran source promotes them
PI32 = float(np.float32(3.141592653589793))          # dynamical_matrix.F90:211
TOL32 = float(np.float32(1e-10))                     # dynamical_matrix.F90:212
C1 = float(np.float32(15.633302))                    # write_output.F90:135
C2 = float(np.float32(33.35641))                     # write_output.F90:135
K32 = float(np.float32(np.float32(15.633302) * np.float32(33.35641)))  # diagonalize.F90:74
C1SQ32 = float(np.float32(np.float32(15.633302) ** 2))                  # write_output.F90:445
DEL_INC32 = float(np.float32(1e-6))                  # dynamical_matrix.F90:61
DEFAULT_TOL32 = float(np.float32(1e-15))             # read_input.F90:395-396


def geometry(lat: np.ndarray, pos: np.ndarray):
    """Crystal coordinates and 2 pi reciprocal rows.

    lammps_read.F90:198-200 (linsolve: crys = inv(lat)^T r) and :228
    (matinvtransp3: rec = 2 pi inv(lat)^T, rows b_k)."""
    lat = np.asarray(lat, np.float64)
    inv = np.linalg.inv(lat)
    crys = np.asarray(pos, np.float64) @ inv          # row form of inv(lat)^T r
    rec = 2.0 * np.pi * inv.T
    return crys, rec


def fc_to_matrix(fc: np.ndarray) -> np.ndarray:
    """The n x n matrix force_read.F90 builds before symmetrising.

    twod_to_fourd (force_read.F90:223-252) maps matrix element (I, J),
    I = 3i+a, J = 3j+b, to the HDF5 Fortran coordinate (a, b, i, j), i.e. the
    C-order element fc[j][i][b][a]."""
    N = fc.shape[0]
    return np.ascontiguousarray(fc.transpose(1, 3, 0, 2)).reshape(3 * N, 3 * N)


def symmetrise_fc(fc: np.ndarray) -> np.ndarray:
    """Phi_sym = (Phi + Phi^T)/2 as pdgeadd('T', ..., 0.5, Phi, 0.5, Phi)
    does at force_read.F90:187-189."""
    P = fc_to_matrix(fc)
    return 0.5 * P.T + 0.5 * P


IMAGES = np.array([[l // 3 - 1, l % 3 - 1, 0] for l in range(9)], dtype=np.float64)
"""Image order of dynamical_matrix.F90:223-226: (t1, t2) = (-1,-1), (-1,0), ..."""


def image_sets(lat, pos, crys):
    """Minimum-image tie sets, dynamical_matrix.F90:223-237.

    Returns mask (N, N, 9) bool: image l of atom j is in S_ij, and mult (N, N).
    Distances use r_img = lat^T (crys_j + t) against the file position r_i."""
    lat = np.asarray(lat, np.float64)
    N = pos.shape[0]
    mask = np.zeros((N, N, 9), dtype=bool)
    # r_img[j, l, :] = sum_m (crys_j + t_l)_m lat[m, :]
    c = crys[:, None, :] + IMAGES[None, :, :]                    # (N, 9, 3)
    rimg = c[..., 0, None] * lat[0] + c[..., 1, None] * lat[1] + c[..., 2, None] * lat[2]
    step = max(1, int(2_000_000 // max(1, N * 9)))
    for i0 in range(0, N, step):
        i1 = min(N, i0 + step)
        d = pos[i0:i1, None, None, :] - rimg[None, :, :, :]      # (b, N, 9, 3)
        dist = np.sqrt(d[..., 0] ** 2 + d[..., 1] ** 2 + d[..., 2] ** 2)
        mn = dist.min(axis=2, keepdims=True)
        mask[i0:i1] = dist < (mn + TOL32)
    mult = mask.sum(axis=2)
    return mask, mult


def _pair_phases(q, crys, lat, mask, deriv):
    """Phase sum p(i, j) of dynamical_matrix.F90:241-255 (deriv 0: sum of
    exp(-2 pi32 i q.Delta); deriv a in 1..3: sum of -exp(...)*i*(lat^T Delta)_a)."""
    N = crys.shape[0]
    q = np.asarray(q, np.float64)
    delta = (crys[None, :, None, :] + IMAGES[None, None, :, :]) - crys[:, None, None, :]  # (N,N,9,3)
    pre = delta[..., 0] * q[0] + delta[..., 1] * q[1] + delta[..., 2] * q[2]
    theta = (-2.0 * PI32) * pre
    e = np.cos(theta) + 1j * np.sin(theta)
    if deriv == 0:
        term = e
    else:
        a = deriv - 1
        r = delta[..., 0] * lat[0, a] + delta[..., 1] * lat[1, a] + delta[..., 2] * lat[2, a]
        term = -(e * 1j) * r
    term = np.where(mask, term, 0.0)
    return term.sum(axis=2)


def assemble(q, phi_sym, crys, lat, masses, types, mask, mult, deriv=0):
    """D(q) (deriv=0) or the analytic dD/dq_a (deriv=a, Cartesian, 1/A),
    dynamical_matrix.F90:49-191 with create_dynamical_ij :200-262:
    D_IJ = Phi_sym_IJ * p_ij / |S_ij| / sqrt(m_i m_j)."""
    N = crys.shape[0]
    p = _pair_phases(q, crys, lat, mask, deriv) / mult
    m = np.asarray(masses, np.float64)[np.asarray(types) - 1]
    w = p / np.sqrt(m[:, None] * m[None, :])
    W = np.repeat(np.repeat(w, 3, axis=0), 3, axis=1)
    return phi_sym * W


def cd_delta(rec):
    """Cartesian step and its crystal components, dynamical_matrix.F90:136-142."""
    s = max(DEL_INC32 * np.sqrt(abs(rec[0, 0] * rec[1, 1] - rec[0, 1] * rec[1, 0])), DEL_INC32)
    return s


def assemble_cd(q, deriv, phi_sym, crys, lat, rec, masses, types, mask, mult):
    """Central-difference dD/dq_a, dynamical_matrix.F90:133-149:
    (D(q + dc/2) - D(q - dc/2)) / |delta|, delta = s e_a, dc = inv(rec)^T delta."""
    s = cd_delta(rec)
    delta = np.zeros(3)
    delta[deriv - 1] = s
    dc = np.linalg.inv(rec).T @ delta
    q = np.asarray(q, np.float64)
    Dp = assemble(q + dc / 2, phi_sym, crys, lat, masses, types, mask, mult, 0)
    Dm = assemble(q - dc / 2, phi_sym, crys, lat, masses, types, mask, mult, 0)
    return (Dp - Dm) / np.linalg.norm(delta)


def eig(D, jobz="V", rng="A", vl=0.0, vu=0.0, il=1, iu=1, abstol=DEFAULT_TOL32):
    """pzheevx(JOBZ, RANGE, 'U', ...) semantics, diagonalize.F90:73-96, using
    LAPACK zheevx on the upper triangle.  For RANGE='V' the reference passes
    vl/K32 and vu/K32 (cm^-1 inputs scaled once, compared against lambda):
    eigenvalues in (vl/K32, vu/K32] are returned.  Returns (w[M], Z[n, M] or
    None, info)."""
    from scipy.linalg import lapack
    compute_v = 1 if jobz == "V" else 0
    a = np.array(D, dtype=np.complex128, order="F", copy=True)
    kw = dict(compute_v=compute_v, range=rng, lower=0, abstol=abstol)
    if rng == "V":
        kw.update(vl=vl / K32, vu=vu / K32)
    elif rng == "I":
        kw.update(il=il, iu=iu)
    w, z, m, ifail, info = lapack.zheevx(a, **kw)
    w = np.asarray(w[:m])
    Z = np.asarray(z[:, :m]) if compute_v else None
    return w, Z, int(info)


def signed_sqrt(lam):
    """eval <- sign(lambda) sqrt|lambda|, diagonalize.F90:119."""
    lam = np.asarray(lam, np.float64)
    return np.sqrt(np.abs(lam)) * np.where(lam < 0, -1.0, 1.0)


def to_cm(lam):
    """omega (cm^-1) = (s * 15.633302f) * 33.35641f, write_output.F90:135."""
    return (signed_sqrt(lam) * C1) * C2


def velocity(Z, dD_list, omega_cm):
    """Group velocities (m/s), write_output.F90:410-446: for each direction,
    Re(diag(Z^H dD Z)) / max(2|omega|, 1e-3) * 33.35641f * (15.633302f)^2 * 100 * 2 * pi32.
    Returns (M, 3) in the h5py layout of the `vel` dataset."""
    out = np.zeros((Z.shape[1], 3))
    den = np.maximum(2.0 * np.abs(omega_cm), 1e-3)
    for a, dD in enumerate(dD_list):
        V = dD @ Z
        d = np.einsum("im,im->m", Z.conj(), V).real
        out[:, a] = ((((d / den) * C2) * C1SQ32) * 100.0) * 2.0 * PI32
    return out


class System:
    """Convenience bundle: geometry, symmetrised Phi and image sets."""

    def __init__(self, lat, pos, types, masses, fc):
        self.lat = np.asarray(lat, np.float64)
        self.pos = np.asarray(pos, np.float64)
        self.types = np.asarray(types)
        self.masses = np.asarray(masses, np.float64)
        self.crys, self.rec = geometry(self.lat, self.pos)
        self.phi = symmetrise_fc(np.asarray(fc, np.float64))
        self.mask, self.mult = image_sets(self.lat, self.pos, self.crys)
        self.n = self.phi.shape[0]

    def D(self, q, deriv=0, method="A"):
        if deriv and method == "C":
            return assemble_cd(q, deriv, self.phi, self.crys, self.lat, self.rec,
                               self.masses, self.types, self.mask, self.mult)
        return assemble(q, self.phi, self.crys, self.lat, self.masses, self.types,
                        self.mask, self.mult, deriv)

    def solve(self, q, jobz="V", rng="A", vl=0.0, vu=0.0, il=1, iu=1,
              abstol=DEFAULT_TOL32, vel=False, method="A"):
        """One q-point of the reference loop (calculations.F90:49-53):
        returns dict(eval_cm, evec, vel, info)."""
        if vel:
            jobz = "V"
        w, Z, info = eig(self.D(q), jobz, rng, vl, vu, il, iu, abstol)
        om = to_cm(w)
        out = {"lam": w, "eval_cm": om, "evec": Z, "info": info, "vel": None}
        if vel:
            dD = [self.D(q, a, method) for a in (1, 2, 3)]
            out["vel"] = velocity(Z, dD, om)
        return out
