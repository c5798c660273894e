"""Pins of the FP64 oracle (CPU).  Each check is independent of the
vectorised formula it tests: brute-force loops written from the Fortran
loop structure, invariants fixed by the mathematics, library routines of a
different kind, the reference's own example inputs, and finite differences."""
import os

import numpy as np
import pytest

from oracle import phonon as O
from tests import _util as U

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _tblg21():
    z = np.load(os.path.join(GOLD, "tblg_21.8_input.npz"))
    return z, O.System(z["lat"], z["pos"], z["types"], z["masses"], z["fc"])


# ------------------------------------------------------------ geometry
def test_geometry_reciprocal_and_crystal():
    z, S = _tblg21()
    lat = z["lat"]
    assert np.allclose(lat @ S.rec.T, 2 * np.pi * np.eye(3), atol=1e-12)
    # r = lat^T crys (lammps_read.F90:198-200)
    assert np.allclose(S.crys @ lat, z["pos"], atol=1e-12)


# ------------------------------------------------------------ force constants
def test_fc_transpose_quirk_literal_loop():
    """force_read.F90:223-252 literally: (I, J) 1-based -> Fortran coordinate
    (mod(I,3) or 3, mod(J,3) or 3, ceil(I/3), ceil(J/3)), HDF5 reverses the
    dimension order, so the C element is fc[c4-1][c3-1][c2-1][c1-1]."""
    rng = np.random.default_rng(0)
    N = 4
    fc = rng.standard_normal((N, N, 3, 3))
    P = np.zeros((3 * N, 3 * N))
    for I in range(1, 3 * N + 1):
        for J in range(1, 3 * N + 1):
            c1 = I % 3 if I % 3 != 0 else 3
            c2 = J % 3 if J % 3 != 0 else 3
            c3 = I // 3 + 1 if I % 3 != 0 else I // 3
            c4 = J // 3 + 1 if J % 3 != 0 else J // 3
            P[I - 1, J - 1] = fc[c4 - 1, c3 - 1, c2 - 1, c1 - 1]
    assert np.array_equal(O.fc_to_matrix(fc), P)
    # pdgeadd('T', 0.5, P, 0.5, P): symmetric, and equal to symmetrising the natural layout
    S = O.symmetrise_fc(fc)
    natural = fc.transpose(0, 2, 1, 3).reshape(3 * N, 3 * N)
    assert np.array_equal(S, S.T)
    assert np.allclose(S, 0.5 * (natural + natural.T), atol=1e-15)


def test_shipped_fc_properties():
    z, _ = _tblg21()
    fc = z["fc"]
    N = fc.shape[0]
    assert fc.shape == (28, 28, 3, 3) and fc.dtype == np.float64
    # the shipped dataset satisfies the acoustic sum rule to <= 1e-6 and
    # Phi = Phi^T to <= 2.4e-7 (checked with margin)
    assert np.abs(fc.sum(axis=1)).max() < 1e-5
    P = fc.transpose(0, 2, 1, 3).reshape(3 * N, 3 * N)
    assert np.abs(P - P.T).max() < 1e-6


def test_synthetic_fc_exact_symmetry_and_asr():
    st, fc = U.structure(3, 4, 1003)
    N = st.natom
    P = fc.transpose(0, 2, 1, 3).reshape(3 * N, 3 * N)
    assert np.abs(P - P.T).max() < 1e-12
    assert np.abs(fc.sum(axis=1)).max() < 1e-11


# ------------------------------------------------------------ images
def _brute_images(lat, pos, crys, i, j):
    d = []
    for l in range(1, 10):
        a = (l - 1) // 3 - 1
        b = l - 2 - 3 * (a + 1)
        c = [crys[j, 0] + a, crys[j, 1] + b, crys[j, 2]]
        r = [sum(lat[m][k] * c[m] for m in range(3)) for k in range(3)]
        d.append(np.sqrt((pos[i, 0] - r[0]) ** 2 + (pos[i, 1] - r[1]) ** 2 + (pos[i, 2] - r[2]) ** 2))
    mn = min(d)
    return [l - 1 for l in range(1, 10) if d[l - 1] < mn + O.TOL32]


def test_image_sets_brute_force():
    z, S = _tblg21()
    lat, pos = z["lat"], z["pos"]
    for i in range(0, 28, 3):
        for j in range(28):
            sel = _brute_images(lat, pos, S.crys, i, j)
            assert list(np.nonzero(S.mask[i, j])[0]) == sel


def test_image_ties_by_construction():
    """Two atoms half a lattice vector apart along a1 tie between two images."""
    lat = np.array([[4.0, 0, 0], [2.0, np.sqrt(12.0), 0], [0, 0, 10.0]])
    pos = np.array([[0.0, 0.0, 5.0], [2.0, 0.0, 5.0]])
    crys, _ = O.geometry(lat, pos)
    mask, mult = O.image_sets(lat, pos, crys)
    assert mult[0, 1] == 2 and mult[1, 0] == 2 and mult[0, 0] == 1
    # the self pair keeps only the home image
    assert np.nonzero(mask[0, 0])[0].tolist() == [4]


# ------------------------------------------------------------ assembly
def _brute_D(S, q, deriv=0):
    """create_dynamical_ij (dynamical_matrix.F90:200-262) element by element."""
    n = S.n
    D = np.zeros((n, n), complex)
    for I in range(n):
        for J in range(n):
            ia, ja = I // 3, J // 3
            sel = np.nonzero(S.mask[ia, ja])[0]
            ph = 0j
            for l in sel:
                a, b = l // 3 - 1, l % 3 - 1
                delta = [S.crys[ja, 0] + a - S.crys[ia, 0], S.crys[ja, 1] + b - S.crys[ia, 1],
                         S.crys[ja, 2] - S.crys[ia, 2]]
                pre = q[0] * delta[0] + q[1] * delta[1] + q[2] * delta[2]
                e = np.exp(-2 * O.PI32 * 1j * pre)
                if deriv == 0:
                    ph += e
                else:
                    r = [sum(S.lat[m, k] * delta[m] for m in range(3)) for k in range(3)]
                    ph -= e * 1j * r[deriv - 1]
            m_i = S.masses[S.types[ia] - 1]
            m_j = S.masses[S.types[ja] - 1]
            D[I, J] = S.phi[I, J] * ph / len(sel) / np.sqrt(m_i * m_j)
    return D


@pytest.mark.parametrize("deriv", [0, 1, 3])
def test_assembly_brute_force(deriv):
    _, S = _tblg21()
    q = np.array([0.137, -0.291, 0.05])
    assert np.allclose(S.D(q, deriv), _brute_D(S, q, deriv), rtol=0, atol=1e-13)


def test_assembly_invariants():
    z, S = _tblg21()
    q = np.array([0.31, 0.17, 0.0])
    D = S.D(q)
    assert np.abs(D - D.conj().T).max() < 1e-13 * np.abs(D).max()     # Hermitian
    assert np.abs(S.D(-q) - D.conj()).max() < 1e-13 * np.abs(D).max()  # time reversal
    # Gamma: D = Phi_sym / m (one mass type, no ties among interacting pairs)
    D0 = S.D(np.zeros(3))
    m = z["masses"][0]
    inter = np.repeat(np.repeat(S.mult > 1, 3, 0), 3, 1)
    assert np.allclose(D0[~inter], (S.phi / m)[~inter], atol=1e-14)


def test_symmetry_equivalent_points_reference_file():
    """Examples/tblg_21.8/q_pt_symm.dat + check_symm.py: K, K' and K+G give
    the same frequencies."""
    z, S = _tblg21()
    om = [S.solve(q, jobz="N")["eval_cm"] for q in z["q_symm"]]
    # the file's q have 10 significant digits and the shipped force constants
    # are symmetric to ~1e-7, so the spectra agree to ~1e-5 cm^-1
    assert np.abs(om[0] - om[1]).max() < 1e-4
    assert np.abs(om[0] - om[2]).max() < 1e-4


def test_acoustic_zero_modes_at_gamma():
    S = U.oracle_system(3, 4, 1003)
    w, _, _ = O.eig(S.D(np.zeros(3)), jobz="N")
    assert np.abs(w[:3]).max() < 1e-12 * np.abs(w).max()
    assert w[3] > 1e-4


def test_reference_bands_features():
    """Frequencies from the shipped tblg_21.8 inputs against the features of
    Examples/tblg_21.8/Reference_bands.pdf (read off the figure, +-5 cm^-1)."""
    z = np.load(os.path.join(GOLD, "tblg_21.8_oracle.npz"))
    g, k, m = z["eval_cm"]
    for f in (78.2, 287.7, 587.0, 722.5, 975.5, 1470.0, 1684.0):
        assert np.abs(g - f).min() < 5.0
    for f in (82.2, 102.2, 115.4, 124.5, 1708.5):
        assert np.abs(m - f).min() < 5.0
    assert abs(k.max() - 1723.2) < 5.0 and abs(m.max() - 1708.5) < 5.0
    _, S = _tblg21()
    om = S.solve(np.zeros(3), jobz="N")["eval_cm"]
    assert np.allclose(om, g, atol=1e-9)


def test_analytic_derivative_vs_finite_difference():
    _, S = _tblg21()
    q = np.array([0.21, 0.33, 0.0])
    inv = np.linalg.inv(S.rec)
    h = 1e-4
    for a in (1, 2):
        e = np.zeros(3)
        e[a - 1] = h
        dq = e @ inv                      # Cartesian step in reduced coordinates
        fd = (-S.D(q + 2 * dq) + 8 * S.D(q + dq) - 8 * S.D(q - dq) + S.D(q - 2 * dq)) / (12 * h)
        # the phase uses f32(pi): d/dq_cart of exp(-2 pi32 i q.Delta) carries pi32/pi
        an = S.D(q, a) * (O.PI32 / np.pi)
        assert np.abs(fd - an).max() < 1e-7 * np.abs(an).max()


def _cd_error(S, q, a, cd):
    """max |CD - analytic * pi32/pi| / max |analytic| for direction a."""
    an = S.D(q, a) * (O.PI32 / np.pi)
    return np.abs(cd - an).max() / np.abs(an).max()


@pytest.mark.parametrize("a", [1, 2, 3])
def test_central_difference_matches_analytic(a):
    """The oracle's central difference (dynamical_matrix.F90:133-149) is the
    derivative along Cartesian q_a: with the phase exp(-2 pi32 i q.Delta), the
    exact derivative is the analytic dD/dq_a times pi32/pi, and the central
    difference with step s = max(f32(1e-6) sqrt|b1x b2y - b1y b2x|, f32(1e-6))
    is within O(s^2) truncation plus O(eps |D| / s) rounding of it (both far
    below 1e-6 here).  The hexagonal cell makes inv(rec) non-symmetric, so a
    step mapped with inv(rec) instead of inv(rec)^T moves q along the wrong
    direction; a wrong normalisation (2|delta|) halves the result: the planted
    variants below must fail the same check."""
    _, S = _tblg21()
    # a non-orthogonal reciprocal lattice is what separates inv(rec) from inv(rec)^T
    assert abs(S.rec[0] @ S.rec[1]) > 0.1 * np.linalg.norm(S.rec[0]) ** 2
    for q in (np.array([0.21, 0.33, 0.0]), np.array([-0.37, 0.08, 0.11])):
        cd = O.assemble_cd(q, a, S.phi, S.crys, S.lat, S.rec, S.masses, S.types, S.mask, S.mult)
        assert _cd_error(S, q, a, cd) < 1e-6
        # the step itself: |delta| = s, its crystal components solve rec^T dc = delta
        s = O.cd_delta(S.rec)
        assert s == max(O.DEL_INC32 * np.sqrt(abs(S.rec[0, 0] * S.rec[1, 1]
                                                   - S.rec[0, 1] * S.rec[1, 0])), O.DEL_INC32)
        # planted mistakes (on the CPU, in a copy of the oracle's formula)
        delta = np.zeros(3)
        delta[a - 1] = s
        dc_wrong = np.linalg.inv(S.rec) @ delta                  # transposed inv(rec)
        Dp = S.D(q + dc_wrong / 2)
        Dm = S.D(q - dc_wrong / 2)
        wrong_dir = (Dp - Dm) / s
        dc = np.linalg.inv(S.rec).T @ delta
        wrong_norm = (S.D(q + dc / 2) - S.D(q - dc / 2)) / (2 * s)
        if a != 3:          # rec has no xz / yz coupling: the z step is the same either way
            assert _cd_error(S, q, a, wrong_dir) > 1e-2
        assert _cd_error(S, q, a, wrong_norm) > 0.4


# ------------------------------------------------------------ eigen
def test_zheevx_against_eigh_and_ranges():
    S = U.oracle_system(3, 4, 1003)
    D = S.D(np.array([0.1, 0.2, 0.0]))
    w, Z, info = O.eig(D)
    we = np.linalg.eigvalsh(D)
    assert info == 0 and np.abs(w - we).max() < 1e-12
    assert np.linalg.norm(D @ Z - Z * w, axis=0).max() < 1e-11
    wi, _, _ = O.eig(D, jobz="N", rng="I", il=5, iu=9)
    assert np.allclose(wi, we[4:9], atol=1e-12)
    vl, vu = 100.0, 400.0        # cm^-1 as written in the input; compared with lambda
    wv, _, _ = O.eig(D, jobz="N", rng="V", vl=vl, vu=vu)
    sel = we[(we > vl / O.K32) & (we <= vu / O.K32)]
    assert np.allclose(wv, sel, atol=1e-12)


def test_units_constants():
    assert O.K32 == 521.4708251953125
    assert O.PI32 == 3.1415927410125732
    const = O.C2 * O.C1SQ32 * 100 * 2 * O.PI32
    assert abs(const - 5122248.0309) < 1e-3


# ------------------------------------------------------------ velocity
def test_velocity_vs_frequency_finite_difference():
    """v = d Omega / dk: 2 pi c d(omega~)/dq with c = 2.99792458e10 cm/s and
    q in 1/A (the reference's factor 33.35641 * 100 equals 1/c to 1e-7)."""
    S = U.oracle_system(3, 4, 1003)
    q = np.array([0.137, 0.291, 0.0])
    r = S.solve(q, vel=True)
    inv = np.linalg.inv(S.rec)
    h = 2e-5
    om0 = r["eval_cm"]
    groups = U.clusters(r["lam"], 1e-3)   # bands well away from crossings
    single = [g[0] for g in groups if len(g) == 1]
    for a in (0, 1):
        e = np.zeros(3)
        e[a] = h
        dq = e @ inv
        op = S.solve(q + dq, jobz="N")["eval_cm"]
        om_ = S.solve(q - dq, jobz="N")["eval_cm"]
        fd = (op - om_) / (2 * h) * 2 * np.pi * 2.99792458
        v = r["vel"][:, a]
        idx = np.array(single)
        err = np.abs(fd[idx] - v[idx] * O.PI32 / np.pi)
        assert np.median(err / (np.abs(v[idx]) + 1.0)) < 1e-6
        assert (err / (np.abs(v[idx]) + 50.0)).max() < 1e-3


def test_golden_regression_tbg_1_2():
    z = np.load(os.path.join(GOLD, "tbg_1_2_oracle.npz"))
    S = U.oracle_system(1, 2, 11)
    for iq, q in enumerate(z["q"]):
        r = S.solve(q, vel=True)
        assert np.allclose(r["eval_cm"], z["eval_cm"][iq], atol=1e-9)
        assert np.allclose(S.D(q), z["D"][iq], atol=1e-14)


def test_block_factorisation_of_D_and_its_quadratic_forms():
    """The block-sparse forms rest on two facts about the oracle's D(q) and
    dD/dq (dynamical_matrix.F90:241-258): each 3x3 block is a phase sum times
    the q-independent real block Phi~ = Phi_sym / |S| / sqrt(m m), and the
    (j, i) block is the Hermitian transpose of the (i, j) block (mirrored tie
    sets).  With both, the quadratic form z^H Op z over the nonzero lower
    blocks (weight 2 off the diagonal, the diagonal block whole) equals the
    dense form.  Checked on the (3,4) cell, whose spring model leaves 70% of
    the blocks zero."""
    st, _ = U.structure(3, 4, 1003)
    S = U.oracle_system(3, 4, 1003)
    N = st.natom
    m = np.asarray(S.masses)[np.asarray(S.types) - 1]
    blk = S.phi.reshape(N, 3, N, 3).transpose(0, 2, 1, 3)          # (i, j, alpha, beta)
    nz = np.abs(blk).max(axis=(2, 3)) > 0
    assert nz.mean() < 0.4
    phit = blk / (S.mult * np.sqrt(m[:, None] * m[None, :]))[:, :, None, None]
    rng = np.random.default_rng(7)
    q = rng.uniform(-0.5, 0.5, 3) * np.array([1, 1, 0])
    Z = rng.standard_normal((3 * N, 5)) + 1j * rng.standard_normal((3 * N, 5))
    lo_a, lo_b = np.nonzero(np.tril(nz))
    for deriv in (0, 1, 2, 3):
        D = S.D(q, deriv)
        p = O._pair_phases(q, S.crys, S.lat, S.mask, deriv)
        # block = phase sum times Phi~; Hermitian partner
        Db = D.reshape(N, 3, N, 3).transpose(0, 2, 1, 3)
        scale = np.abs(D).max()
        assert np.abs(Db - p[:, :, None, None] * phit).max() <= 1e-14 * scale
        assert np.abs(p - p.T.conj())[nz].max() <= 1e-12 * max(1.0, np.abs(p).max())
        assert np.abs(D - D.conj().T).max() <= 1e-14 * scale
        # lower-block quadratic forms against the dense ones
        zb = Z.reshape(N, 3, -1)
        forms = np.zeros(Z.shape[1])
        for a, b in zip(lo_a, lo_b):
            c = np.einsum("xm,xy,ym->m", zb[a].conj(), phit[a, b], zb[b])
            forms += (1.0 if a == b else 2.0) * np.real(p[a, b] * c)
        dense = np.real(np.einsum("im,ij,jm->m", Z.conj(), D, Z))
        assert np.abs(forms - dense).max() <= 1e-12 * np.abs(dense).max()
