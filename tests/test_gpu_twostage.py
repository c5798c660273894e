"""The two-stage reduction (dense -> band on the matrix cores, bulge chasing,
and the back-transformation of both stages) against LAPACK and the oracle,
and the 64 x 32 GEMM tiles its panel products use (marked gpu)."""
import numpy as np
import pytest

from tests import _util as U
from oracle import phonon as O

pytestmark = pytest.mark.gpu


def _lib():
    from parphom_amd import lib
    lib.load()
    return lib


@pytest.fixture
def two_stage(gpu_available):
    """Make the two-stage reduction the default for this test's handles."""
    lib = _lib()
    yield lib
    lib.set_default_reduction(-1, 32)


@pytest.mark.parametrize("prec,tol", [("D", 1e-12), ("S", 2e-5)])
@pytest.mark.parametrize("ops", ["NN", "CN", "NC"])
@pytest.mark.parametrize("mnk", [(300, 32, 97), (77, 17, 260), (200, 1, 45)])
def test_gemm_narrow_outputs(gpu_available, prec, tol, ops, mnk):
    """Outputs at most 32 columns wide take 64 x 32 tiles."""
    lib = _lib()
    rng = np.random.default_rng(7)
    m, n, k = mnk
    opA, opB = ops
    A = rng.standard_normal((m, k) if opA == "N" else (k, m)) + 1j * rng.standard_normal((m, k) if opA == "N" else (k, m))
    B = rng.standard_normal((k, n) if opB == "N" else (n, k)) + 1j * rng.standard_normal((k, n) if opB == "N" else (n, k))
    ref = (A if opA == "N" else A.conj().T) @ (B if opB == "N" else B.conj().T)
    C, _ = lib.debug_zgemm(A, B, opA, opB, prec=prec, mode=0)
    assert np.abs(C - ref).max() <= tol * np.abs(ref).max() * 10


def _hermitian(n, seed, planted=True):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    if not planted:
        return X + X.conj().T
    Q, _ = np.linalg.qr(X)
    lam = np.sort(rng.uniform(-3, 3, n))
    if n >= 12:
        lam[5:8] = lam[5]                  # a triple
        lam[n // 2:n // 2 + 2] = lam[n // 2]
    else:
        lam[1:3] = lam[1]                  # a pair
    A = (Q * lam) @ Q.conj().T
    return 0.5 * (A + A.conj().T)


@pytest.mark.parametrize("prec", ["D", "S"])
@pytest.mark.parametrize("band", [32, 16])
@pytest.mark.parametrize("n", [200, 77, 33, 34, 5])
def test_heevx_two_stage(two_stage, prec, band, n):
    """Eigenvalues, residuals and orthogonality of the two-stage path on
    random Hermitian matrices (planted degeneracies, sizes around the band:
    n = bw + 1 has a single panel row, n < bw none)."""
    lib = two_stage
    lib.set_default_reduction(1, band)
    for seed, planted in ((n, True), (n + 1, False)):
        A = _hermitian(n, seed, planted)
        w, Z, info = lib.heevx(A, jobz="V", precision=prec)
        wref = np.linalg.eigvalsh(A)
        nrm = np.abs(wref).max()
        tol = 1e-12 if prec == "D" else 1e-5
        assert info == 0
        assert np.abs(w - wref).max() <= tol * nrm
        assert np.linalg.norm(A @ Z - Z * w, axis=0).max() <= (1e-11 if prec == "D" else 1e-5) * nrm
        assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= (1e-10 if prec == "D" else 1e-5)


def test_heevx_two_stage_batch_and_ranges(two_stage):
    """Several matrices per launch (one bulge-chasing workgroup each) and
    range I / V with vectors on the two-stage path, FP64."""
    lib = two_stage
    lib.set_default_reduction(1, 32)
    mats = np.stack([_hermitian(150, s, planted=(s % 2 == 0)) for s in range(5)])
    ws, Zs, infos = lib.heevx(mats, jobz="V", batch=5)
    for A, w, Z, inf in zip(mats, ws, Zs, infos):
        assert inf == 0
        wref = np.linalg.eigvalsh(A)
        assert np.abs(w - wref).max() <= 1e-12 * np.abs(wref).max()
        assert np.linalg.norm(A @ Z - Z * w, axis=0).max() <= 1e-11 * np.abs(wref).max()
    A = mats[1]
    wref = np.linalg.eigvalsh(A)
    w, Z, info = lib.heevx(A, jobz="V", rng="I", il=5, iu=20)
    assert len(w) == 16 and np.abs(w - wref[4:20]).max() < 1e-11
    assert np.linalg.norm(A @ Z - Z * w, axis=0).max() < 1e-10


@pytest.mark.parametrize("prec", ["D", "S"])
def test_run_tbg_3_4_two_stage(two_stage, prec):
    """The (3,4) cell's full pipeline (n = 444: 13 panels, a ragged last
    one) on the two-stage path: frequencies (1e-3 cm^-1) and velocities
    against the oracle as in the one-stage tests."""
    lib = two_stage
    st, fc = U.structure(3, 4, 1003)
    S = U.oracle_system(3, 4, 1003)
    from parphom_amd import synth
    qs = synth.q_path_gkmg()[[0, 5, 26, 39]]
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                    precision=prec, batch=2, streams=2) as h:
        h.reduction(1, 32)
        res = h.run(qs)
    n = 3 * st.natom
    for iq, q in enumerate(qs):
        ref = S.solve(q, vel=True)
        assert res["info"][iq] == 0
        assert np.abs(res["eval_cm"][iq] - ref["eval_cm"]).max() <= 1e-3
        Z = np.asarray(res["evec"][iq])
        D = S.D(q)
        nrm = np.abs(ref["lam"]).max()
        tol = 1e-10 if prec == "D" else 1e-5
        assert np.linalg.norm(D @ Z - Z * ref["lam"], axis=0).max() <= tol * nrm
        assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= tol
        groups = U.clusters(ref["lam"], 1e-6)
        dD = [S.D(q, a) for a in (1, 2, 3)]
        eps = None if prec == "D" else float(np.finfo(np.float32).eps)
        noise = U.vel_noise(dD, ref["eval_cm"], ref["lam"], groups, eps=eps)
        if prec == "D":
            U.assert_vel_close(res["vel"][iq], ref["vel"], groups, noise)
        else:
            vmax = np.abs(ref["vel"]).max()
            for g in groups:
                a, b = res["vel"][iq][g].sum(axis=0), ref["vel"][g].sum(axis=0)
                assert np.all(np.abs(a - b) <= 1e-5 * vmax * len(g) + noise[g].sum())


@pytest.mark.parametrize("prec,tol", [("D", 1e-12), ("S", 2e-5)])
@pytest.mark.parametrize("n,nx", [(200, 32), (333, 17), (64, 5), (130, 96)])
def test_gemm_hermitian_from_lower(gpu_available, prec, tol, n, nx):
    """C = A X for Hermitian A from its lower triangle alone (the lower
    product with the diagonal plus the transposed strict lower product); the
    upper triangle holds garbage that must not be read."""
    lib = _lib()
    rng = np.random.default_rng(11)
    H = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = H + H.conj().T
    X = rng.standard_normal((n, nx)) + 1j * rng.standard_normal((n, nx))
    ref = H @ X
    A = np.tril(H) + np.triu(np.full((n, n), 3e3 - 5e2j), 1)
    C, _ = lib.debug_zgemm(A, X, prec=prec, mode=5)
    assert np.abs(C - ref).max() <= tol * np.abs(ref).max() * 10
