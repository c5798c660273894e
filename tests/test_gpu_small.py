"""The one-launch reduction of small matrices (one workgroup per matrix, the
matrix in LDS, every column in one kernel) against LAPACK, the oracle and the
blocked one-stage reduction it stands in for below n = 96 (FP64) / 128 (FP32)
(marked gpu)."""
import numpy as np
import pytest

from tests import _util as U

pytestmark = pytest.mark.gpu


def _lib():
    from parphom_amd import lib
    lib.load()
    return lib


@pytest.fixture
def reduction(gpu_available):
    """Set the process default of the reduction; restored to automatic after
    the test."""
    lib = _lib()
    yield lambda mode: lib.set_default_reduction(mode, 32)
    lib.set_default_reduction(-1, 32)


def _matrices(n, rng):
    """Hermitian test matrices of order n: random; planted exact degeneracies;
    block diagonal (reflectors of zero columns, complex alpha); complex
    tridiagonal (one nonzero entry below each diagonal entry)."""
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    out = [X + X.conj().T]
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lam = np.sort(rng.uniform(-3, 3, n))
    if n >= 8:
        lam[2:5] = lam[2]
        lam[n - 3:n - 1] = lam[n - 3]
    A2 = (Q * lam) @ Q.conj().T
    out.append(0.5 * (A2 + A2.conj().T))
    A3 = out[0].copy()
    h = n // 2
    A3[h:, :h] = 0
    A3[:h, h:] = 0
    out.append(A3)
    A4 = np.diag(rng.standard_normal(n)).astype(complex)
    off = rng.standard_normal(n - 1) + 1j * rng.standard_normal(n - 1)
    A4 += np.diag(off, -1) + np.diag(off.conj(), 1)
    out.append(A4)
    return np.stack(out)


@pytest.mark.parametrize("prec", ["D", "S"])
@pytest.mark.parametrize("n", [1, 2, 3, 17, 32, 33, 64, 65, 96, 128])
@pytest.mark.parametrize("mode", [2, 0])
def test_small_heevx_against_lapack(reduction, prec, n, mode):
    """Eigenvalues against numpy (LAPACK zheevd), residuals and orthogonality
    of the vectors, for the small-matrix kernel (mode 2) and the blocked
    one-stage reduction (mode 0) on the same batch; n = 128 is FP32-only for
    the small kernel (the FP64 path takes the one-stage reduction there)."""
    lib = _lib()
    reduction(mode)
    rng = np.random.default_rng(100 + n)
    A = _matrices(n, rng)
    w, Z, info = lib.heevx(A, jobz="V", precision=prec)
    tol_w = 1e-12 if prec == "D" else 1e-5
    tol_r = 1e-11 if prec == "D" else 1e-5
    tol_o = 1e-10 if prec == "D" else 1e-5
    assert np.all(np.asarray(info) == 0)
    for i, M in enumerate(A):
        wref = np.linalg.eigvalsh(M)
        nrm = max(np.abs(wref).max(), 1e-300)
        assert len(w[i]) == n
        assert np.abs(w[i] - wref).max() <= tol_w * nrm, (i, np.abs(w[i] - wref).max())
        R = M @ Z[i] - Z[i] * w[i]
        assert np.linalg.norm(R, axis=0).max() <= tol_r * nrm, i
        assert np.abs(Z[i].conj().T @ Z[i] - np.eye(n)).max() <= tol_o, i


@pytest.mark.parametrize("prec", ["D", "S"])
def test_small_heevx_ranges(reduction, prec):
    """Ranges I and V and values only through the small-matrix kernel."""
    lib = _lib()
    reduction(2)
    rng = np.random.default_rng(7)
    n = 80
    A = _matrices(n, rng)[0]
    wref = np.linalg.eigvalsh(A)
    nrm = np.abs(wref).max()
    tol = 1e-12 if prec == "D" else 1e-5
    w, Z, info = lib.heevx(A, jobz="V", rng="I", il=5, iu=20, precision=prec)
    assert info == 0 and len(w) == 16 and np.abs(w - wref[4:20]).max() <= tol * nrm
    assert np.linalg.norm(A @ Z - Z * w, axis=0).max() <= (1e-11 if prec == "D" else 1e-5) * nrm
    vl, vu = wref[10] - 1e-3 * nrm, wref[50] + 1e-3 * nrm
    w, Z, info = lib.heevx(A, jobz="N", rng="V", vl=vl, vu=vu, precision=prec)
    sel = wref[(wref > vl) & (wref <= vu)]
    assert info == 0 and len(w) == len(sel) and np.abs(w - sel).max() <= tol * nrm


@pytest.mark.parametrize("prec", ["D", "S"])
def test_small_run_tbg_1_2(gpu_available, prec):
    """The (1,2) cell (N = 28, n = 84) through pph_run with the automatic
    choice, which takes the small-matrix kernel: frequencies against the
    oracle (<= 1e-3 cm^-1), velocities by gauge-invariant cluster sums
    against the oracle's; frequencies of the one-stage reduction on the same
    q-points agree.  With profiling on, the small kernel records no hemv
    launches and the one-stage reduction does (the path really switched)."""
    lib = _lib()
    st, fc = U.structure(1, 2, 11)
    S = U.oracle_system(1, 2, 11)
    from parphom_amd import synth
    qs = np.vstack([synth.q_path_gkmg()[[0, 7, 20, 40]],
                    np.random.default_rng(3).uniform(0, 1, (12, 3)) * [1, 1, 0]])
    res = {}
    hemv = {}
    for mode in (-1, 0):
        with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                        precision=prec, batch=8) as h:
            h.reduction(mode)
            h.profile(True, reset=True)
            res[mode] = h.run(qs)
            hemv[mode] = h.profile(False)["hemv_launches"][1]
    assert hemv[-1] == 0 and hemv[0] > 0
    r = res[-1]
    for iq, q in enumerate(qs):
        ref = S.solve(q, vel=True)
        assert r["info"][iq] == 0
        assert np.abs(r["eval_cm"][iq] - ref["eval_cm"]).max() <= 1e-3
        assert np.abs(r["eval_cm"][iq] - res[0]["eval_cm"][iq]).max() <= 1e-3
        groups = U.clusters(ref["lam"], 1e-6)
        dD = [S.D(q, a) for a in (1, 2, 3)]
        if prec == "D":
            noise = U.vel_noise(dD, ref["eval_cm"], ref["lam"], groups)
            U.assert_vel_close(r["vel"][iq], ref["vel"], groups, noise)
        else:
            noise = U.vel_noise(dD, ref["eval_cm"], ref["lam"], groups,
                                eps=float(np.finfo(np.float32).eps))
            vmax = np.abs(ref["vel"]).max()
            for g in groups:
                a = r["vel"][iq][g].sum(axis=0)
                b = ref["vel"][g].sum(axis=0)
                assert np.all(np.abs(a - b) <= 1e-5 * vmax * len(g) + noise[g].sum())


def test_small_fp32_velocities_match_fp64(gpu_available):
    """The FP32 path with its FP64 correction of the vectors (the default)
    against the FP64 path on the (1,2) cell, acoustic modes near Gamma
    included: velocities per mode (sums only over modes closer than 1e-9
    max|lambda|, degenerate to FP64 precision) within 1e-7 of the largest.
    omega divides the velocity form, so the lowest modes magnify any error
    of the form: dense FP32 forms of the corrected vectors' FP32 rounding
    missed this by an order of magnitude (1.1e-6 of the largest velocity);
    the automatic choice takes the FP64 block sums on this path."""
    lib = _lib()
    st, fc = U.structure(1, 2, 11)
    qs = np.vstack([[[0, 0.9844, 0], [0.0156, 0, 0], [0, 0.0312, 0], [0.0156, 0.0156, 0],
                     [0.0312, 0, 0], [0.0625, 0.0312, 0], [0.5, 0, 0], [1 / 3, 1 / 3, 0]],
                    np.random.default_rng(12).uniform(0, 1, (8, 3)) * [1, 1, 0]])
    res = {}
    for prec in ("D", "S"):
        with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                        precision=prec, batch=16) as h:
            nblk = h.sparse(1)
            assert (nblk > 0) == (prec == "S")
            res[prec] = h.run(qs)
    vmax = max(np.abs(res["D"]["vel"][iq]).max() for iq in range(len(qs)))
    for iq in range(len(qs)):
        assert res["S"]["info"][iq] == 0 and res["D"]["info"][iq] == 0
        om = res["D"]["eval_cm"][iq]
        assert np.abs(res["S"]["eval_cm"][iq] - om).max() <= 1e-6
        lam = np.sign(om) * om ** 2
        tol = 1e-9 * np.abs(lam).max()
        cuts = np.flatnonzero(np.diff(lam) > tol) + 1
        for g in np.split(np.arange(len(lam)), cuts):
            a = res["S"]["vel"][iq][g].sum(axis=0)
            b = res["D"]["vel"][iq][g].sum(axis=0)
            assert np.abs(a - b).max() <= 1e-7 * vmax, (iq, g, a, b)


@pytest.mark.parametrize("prec", ["D", "S"])
def test_small_bit_identical(gpu_available, prec):
    """Two runs of the same batch through the small-matrix kernel return the
    same bits (fixed-order sums only)."""
    lib = _lib()
    st, fc = U.structure(1, 2, 11)
    qs = np.random.default_rng(9).uniform(0, 1, (16, 3)) * [1, 1, 0]
    out = []
    for _ in range(2):
        with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                        precision=prec, batch=16) as h:
            h.reduction(2)
            out.append(h.run(qs))
    for key in ("eval_cm", "evec", "vel"):
        assert np.array_equal(np.asarray(out[0][key]), np.asarray(out[1][key])), key
