"""GPU parity: the HIP library against the FP64 oracle (marked gpu)."""
import numpy as np
import pytest

from tests import _util as U
from oracle import phonon as O

pytestmark = pytest.mark.gpu


def _lib():
    from parphom_amd import lib
    lib.load()
    return lib


# the second shape is above one large tile (128 x 128 FP32, 64 x 128 FP64)
# with ragged edges in every dimension
@pytest.mark.parametrize("prec,tol", [("D", 1e-12), ("S", 2e-5)])
@pytest.mark.parametrize("ops", ["NN", "CN", "NC", "CC"])
@pytest.mark.parametrize("mnk", [(77, 130, 45), (300, 261, 97)])
def test_gemm_ops(gpu_available, prec, tol, ops, mnk):
    lib = _lib()
    rng = np.random.default_rng(1)
    m, n, k = mnk
    opA, opB = ops
    A = rng.standard_normal((m, k) if opA == "N" else (k, m)) + 1j * rng.standard_normal((m, k) if opA == "N" else (k, m))
    B = rng.standard_normal((k, n) if opB == "N" else (n, k)) + 1j * rng.standard_normal((k, n) if opB == "N" else (n, k))
    Aop = A if opA == "N" else A.conj().T
    Bop = B if opB == "N" else B.conj().T
    ref = Aop @ Bop
    C, _ = lib.debug_zgemm(A, B, opA, opB, prec=prec, mode=0)
    assert np.abs(C - ref).max() <= tol * np.abs(ref).max() * 10


@pytest.mark.parametrize("prec,tol", [("D", 1e-12), ("S", 2e-5)])
@pytest.mark.parametrize("mnk", [(150, 70, 64), (333, 290, 96)])
def test_gemm_dot_splitk_twoproduct(gpu_available, prec, tol, mnk):
    lib = _lib()
    rng = np.random.default_rng(2)
    m, n, k = mnk
    A = rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
    B = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    X = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    ref = A @ B
    _, dot = lib.debug_zgemm(A, B, prec=prec, mode=1, Cin=X)
    dref = np.einsum("ij,ij->j", X.conj(), ref).real
    assert np.abs(dot - dref).max() <= tol * np.abs(dref).max() * 10
    C2, _ = lib.debug_zgemm(A, B, prec=prec, mode=2)
    assert np.abs(C2 - ref).max() <= tol * np.abs(ref).max() * 10
    C3, _ = lib.debug_zgemm(A, B, prec=prec, mode=3)
    assert np.abs(C3 - ref).max() <= tol * np.abs(ref).max() * 10


@pytest.mark.parametrize("prec,tol", [("D", 1e-12), ("S", 2e-5)])
@pytest.mark.parametrize("n,nx", [(64, 10), (200, 77), (333, 290)])
def test_gemm_hermitian_dot(gpu_available, prec, tol, n, nx):
    """diag(X^H A X) from the lower triangle of a Hermitian A; the upper
    triangle holds garbage that must not be read."""
    lib = _lib()
    rng = np.random.default_rng(3)
    H = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = H + H.conj().T
    X = rng.standard_normal((n, nx)) + 1j * rng.standard_normal((n, nx))
    dref = np.einsum("ij,ij->j", X.conj(), H @ X).real
    A = np.tril(H) + np.triu(np.full((n, n), 1e3 + 7e2j), 1)
    A[np.diag_indices(n)] += 0.25j          # an imaginary diagonal is ignored too
    _, dot = lib.debug_zgemm(A, X, prec=prec, mode=4, Cin=X)
    assert np.abs(dot - dref).max() <= tol * np.abs(H).max() * np.abs(X).max() ** 2 * n


@pytest.mark.parametrize("mn,seed", [((1, 2), 11), ((3, 4), 1003)])
@pytest.mark.parametrize("method", ["A", "C"])
def test_assembly_matches_oracle(gpu_available, mn, seed, method):
    lib = _lib()
    st, fc = U.structure(*mn, seed)
    S = U.oracle_system(*mn, seed)
    qs = np.array([[0.0, 0.0, 0.0], [2 / 3, 1 / 3, 0.0], [0.123, 0.456, 0.0], [-0.31, 0.77, 0.2]])
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, vel_method=method) as h:
        D = h.assemble(qs, 0)
        for iq, q in enumerate(qs):
            ref = S.D(q)
            assert np.abs(D[iq] - ref).max() <= 1e-13 * np.abs(ref).max()
        for a in (1, 2, 3):
            dD = h.assemble(qs, a)
            for iq, q in enumerate(qs):
                ref = S.D(q, a, method)
                scale = max(np.abs(ref).max(), 1e-30)
                tol = 1e-12 if method == "A" else 1e-6
                assert np.abs(dD[iq] - ref).max() <= tol * scale, (a, iq)


@pytest.mark.parametrize("prec", ["D", "S"])
def test_heevx_random_hermitian(gpu_available, prec):
    lib = _lib()
    rng = np.random.default_rng(5)
    n = 200
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = X + X.conj().T
    # planted exact degeneracies
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lam = np.sort(rng.uniform(-3, 3, n))
    lam[10:13] = lam[10]
    lam[50:52] = lam[50]
    A2 = (Q * lam) @ Q.conj().T
    A2 = 0.5 * (A2 + A2.conj().T)
    for M in (A, A2):
        w, Z, info = lib.heevx(M, jobz="V", precision=prec)
        wref = np.linalg.eigvalsh(M)
        nrm = np.abs(wref).max()
        tol = 1e-12 if prec == "D" else 1e-5
        assert info == 0
        assert np.abs(w - wref).max() <= tol * nrm * (1 if prec == "D" else 1)
        R = M @ Z - Z * w
        assert np.linalg.norm(R, axis=0).max() <= (1e-11 if prec == "D" else 1e-5) * nrm
        if prec == "D":
            assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= 1e-10


def test_heevx_ranges(gpu_available):
    lib = _lib()
    rng = np.random.default_rng(6)
    n = 150
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = X + X.conj().T
    wref = np.linalg.eigvalsh(A)
    w, Z, info = lib.heevx(A, jobz="V", rng="I", il=5, iu=20)
    assert len(w) == 16 and np.abs(w - wref[4:20]).max() < 1e-11
    assert np.linalg.norm(A @ Z - Z * w, axis=0).max() < 1e-10
    vl, vu = -3.0, 2.5
    w, Z, info = lib.heevx(A, jobz="N", rng="V", vl=vl, vu=vu)
    sel = wref[(wref > vl) & (wref <= vu)]
    assert len(w) == len(sel) and np.abs(w - sel).max() < 1e-11


@pytest.mark.parametrize("n", [150, 600])
def test_heevx_range_v_fp32_bounds(gpu_available, n):
    """RANGE='V' on the FP32 path with bounds placed within the FP32 error of
    an eigenvalue (1e-7 |A| outside the window): the FP32 Sturm counts may
    select such a mode, and its FP64-refined eigenvalue then lies outside
    (vl, vu], so it must be dropped.  Every returned eigenvalue lies in the
    window, matches numpy's to 1e-7 |A|, and only modes within 1e-5 |A| of a
    bound may be missing.  n = 600 runs the two-stage reduction."""
    lib = _lib()
    rng = np.random.default_rng(16 + n)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = X + X.conj().T
    wref = np.linalg.eigvalsh(A)
    nrm = np.abs(wref).max()
    eps = 1e-7 * nrm
    lo, hi = n // 4, (3 * n) // 4
    vl, vu = wref[lo] + eps, wref[hi] - eps          # wref[lo], wref[hi] just outside
    w, Z, info = lib.heevx(A.copy(), jobz="V", rng="V", vl=vl, vu=vu, precision="S")
    assert info == 0
    assert len(w) > 0 and (w > vl).all() and (w <= vu).all()
    sel = wref[(wref > vl) & (wref <= vu)]
    near = np.minimum(np.abs(sel - vl), np.abs(sel - vu)) <= 1e-5 * nrm
    idx = np.searchsorted(sel, w)
    idx = np.clip(idx, 0, len(sel) - 1)
    idx = np.where(np.abs(sel[np.clip(idx - 1, 0, None)] - w) < np.abs(sel[idx] - w),
                   np.clip(idx - 1, 0, None), idx)
    assert np.abs(sel[idx] - w).max() <= 1e-7 * nrm
    missing = np.setdiff1d(np.arange(len(sel)), idx)
    assert near[missing].all() and len(w) >= len(sel) - near.sum()
    assert np.linalg.norm(A @ Z - Z * w, axis=0).max() <= 1e-5 * nrm


@pytest.mark.parametrize("prec", ["D", "S"])
@pytest.mark.parametrize("sparse", [0, 1, 2])
def test_run_tbg_3_4_path(gpu_available, prec, sparse):
    """Config-1 structure: eigenvalues (<= 1e-3 cm^-1) and velocities, with the
    dense products (forced, or automatic on the FP64 path: 30% of its blocks
    are nonzero) and with the block-sparse quadratic forms (forced, or
    automatic on the FP32 path, whose corrected vectors need FP64 forms)."""
    lib = _lib()
    st, fc = U.structure(3, 4, 1003)
    S = U.oracle_system(3, 4, 1003)
    from parphom_amd import synth
    qs = synth.q_path_gkmg()[[0, 5, 26, 30, 39, 50]]
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                    precision=prec) as h:
        assert (h.sparse(sparse) > 0) == (sparse == 2 or (sparse == 1 and prec == "S"))
        res = h.run(qs)
    for iq, q in enumerate(qs):
        ref = S.solve(q, vel=True)
        om = res["eval_cm"][iq]
        assert res["info"][iq] == 0
        assert np.abs(om - ref["eval_cm"]).max() <= 1e-3, np.abs(om - ref["eval_cm"]).max()
        if prec == "D":
            groups = U.clusters(ref["lam"], 1e-6)
            dD = [S.D(q, a) for a in (1, 2, 3)]
            noise = U.vel_noise(dD, ref["eval_cm"], ref["lam"], groups)
            U.assert_vel_close(res["vel"][iq], ref["vel"], groups, noise)
            vp = U.vel_projected(res["evec"][iq], dD, ref["eval_cm"], groups)
            vp_ref = U.vel_projected(ref["evec"], dD, ref["eval_cm"], groups)
            tol = 1e-6 * np.abs(vp_ref) + 1e-9 + noise[:, None]
            assert np.all(np.abs(vp - vp_ref) <= tol)
        else:
            # FP32 path: the velocities are forms of the FP32 vectors (FP32
            # matrix-core products here, FP64 block sums with the block lists),
            # so the same gauge-invariant cluster sums are held to the noise
            # of an FP32 eigenvector (eps32 in the conditioning model) and to
            # 1e-5 of the largest velocity of the q-point
            groups = U.clusters(ref["lam"], 1e-6)
            dD = [S.D(q, a) for a in (1, 2, 3)]
            noise = U.vel_noise(dD, ref["eval_cm"], ref["lam"], groups,
                                eps=float(np.finfo(np.float32).eps))
            vmax = np.abs(ref["vel"]).max()
            for g in groups:
                a = res["vel"][iq][g].sum(axis=0)
                b = ref["vel"][g].sum(axis=0)
                tol = 1e-5 * vmax * len(g) + noise[g].sum()
                assert np.all(np.abs(a - b) <= tol), (g, a, b, tol)


@pytest.mark.parametrize("prec", ["D", "S"])
@pytest.mark.parametrize("nb_panel,nb_back", [(16, 16), (48, 64), (64, 128), (32, 48)])
def test_run_every_blocking(gpu_available, prec, nb_panel, nb_back):
    """The reduction and back-transformation at every supported blocking
    (panels of 16..64 reflectors, back-transformation blocks of 16..128):
    frequencies against the oracle and orthonormal eigenvectors with small
    residuals, on the config-1 structure (n = 444: several panels, a ragged
    last one)."""
    lib = _lib()
    st, fc = U.structure(3, 4, 1003)
    S = U.oracle_system(3, 4, 1003)
    from parphom_amd import synth
    qs = synth.q_path_gkmg()[[0, 13, 30]]
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                    precision=prec) as h:
        h.blocking(nb_panel, nb_back)
        res = h.run(qs)
    n = 3 * st.natom
    for iq, q in enumerate(qs):
        ref = S.solve(q)
        assert res["info"][iq] == 0
        assert np.abs(res["eval_cm"][iq] - ref["eval_cm"]).max() <= 1e-3
        Z = np.asarray(res["evec"][iq])
        D = S.D(q)
        nrm = np.abs(ref["lam"]).max()
        tol = 1e-10 if prec == "D" else 1e-5
        assert np.linalg.norm(D @ Z - Z * ref["lam"], axis=0).max() <= tol * nrm
        assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= (1e-10 if prec == "D" else 1e-5)


def test_blocking_rejects_unsupported(gpu_available):
    lib = _lib()
    st, fc = U.structure(1, 2, 11)
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc) as h:
        for bad in ((8, 128), (80, 128), (40, 128), (32, 256), (32, 24)):
            with pytest.raises(lib.PphError):
                h.blocking(*bad)


@pytest.mark.parametrize("prec", ["S", "D"])
def test_run_two_streams(gpu_available, prec):
    """Several batches in flight (batch 2, two streams, a ragged last batch):
    the same frequencies as one stream and as the oracle, host and device
    outputs."""
    lib = _lib()
    import torch
    st, fc = U.structure(1, 2, 11)
    S = U.oracle_system(1, 2, 11)
    from parphom_amd import synth
    qs = synth.q_random(7, 21)
    runs = {}
    for streams in (1, 2):
        with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                        precision=prec, batch=2, streams=streams) as h:
            runs[streams] = h.run(qs)
            n = 3 * st.natom
            dev = torch.device("cuda:0")
            out = {"eval_cm": torch.zeros((7, n), dtype=torch.float64, device=dev),
                   "m": torch.zeros(7, dtype=torch.int64, device=dev),
                   "nz": torch.zeros(7, dtype=torch.int64, device=dev),
                   "info": torch.zeros(7, dtype=torch.int32, device=dev),
                   "evec": torch.zeros((7, n, n), dtype=torch.complex128, device=dev),
                   "vel": torch.zeros((7, n, 3), dtype=torch.float64, device=dev)}
            h.run_device(qs, out)
            torch.cuda.synchronize()
            assert out["m"].cpu().tolist() == [n] * 7
            assert np.abs(out["eval_cm"].cpu().numpy() - np.stack(runs[streams]["eval_cm"])).max() <= 1e-6
    for iq, q in enumerate(qs):
        ref = S.solve(q)["eval_cm"]
        for streams in (1, 2):
            assert runs[streams]["info"][iq] == 0
            assert np.abs(runs[streams]["eval_cm"][iq] - ref).max() <= 1e-3
        assert np.abs(runs[1]["eval_cm"][iq] - runs[2]["eval_cm"][iq]).max() <= 1e-6


def test_profile_before_first_two_lane_run(gpu_available):
    """Phase timers switched on before any batch has run: the second lane,
    created by pph_run, must time its batches too (every batch counted once,
    every eighth hemv launch of each batch timed; the one-stage reduction is
    forced, since at n = 84 the library picks the small-matrix kernel, which
    has no hemv launches)."""
    lib = _lib()
    st, fc = U.structure(1, 2, 11)
    from parphom_amd import synth
    qs = synth.q_random(7, 22)
    n = 3 * st.natom
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                    batch=2, streams=2) as h:
        h.reduction(0)
        h.profile(enable=True, reset=True)
        res = h.run(qs)
        prof = h.profile(enable=False)
    assert all(i == 0 for i in res["info"])
    nbatch = 4                                        # 2 + 2 + 2 + 1 q-points
    assert prof["total"][1] == nbatch and prof["tridiag"][1] == nbatch
    assert prof["total"][0] > 0.0
    assert prof["hemv_launches"][1] == nbatch * len(range(0, n - 1, 8))


@pytest.mark.parametrize("prec", ["D", "S"])
@pytest.mark.parametrize("sel", ["I", "V", "N"])
def test_run_selections_in_lanes(gpu_available, prec, sel):
    """pph_run with several batches in flight (batch 2, two lanes, a ragged
    last batch) for the other selections of pzheevx: RANGE='I' (modes 5..40),
    RANGE='V' (the reference's vl/K32, vu/K32 scaling, a window through the
    middle of the spectrum) and JOBZ='N' (values only).  The selected
    frequencies and their count against the oracle; where vectors exist, their
    residuals against the oracle's D(q), and each selected mode's velocity
    against Re(z^H dD z) of the library's own vector, scaled as
    write_output.F90:443-446 scales it (independent of the gauge of degenerate
    modes)."""
    lib = _lib()
    from parphom_amd import synth
    st, fc = U.structure(2, 3, 23)
    S = U.oracle_system(2, 3, 23)
    qs = synth.q_random(5, 31)
    if sel == "I":
        opt = dict(jobz="V", rng="I", il=5, iu=40, vel=True)
    elif sel == "V":
        opt = dict(jobz="V", rng="V", vl=300.0, vu=900.0, vel=True)
    else:
        opt = dict(jobz="N", rng="A")
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, precision=prec, batch=2, streams=2,
                    **opt) as h:
        res = h.run(qs)
    const = O.C2 * O.C1SQ32 * 100.0 * 2.0 * O.PI32
    for iq, q in enumerate(qs):
        ref = S.solve(q, jobz=opt["jobz"], rng=opt["rng"], vl=opt.get("vl", 0.0),
                      vu=opt.get("vu", 0.0), il=opt.get("il", 1), iu=opt.get("iu", 1))
        assert res["info"][iq] == 0
        assert res["m"][iq] == len(ref["lam"]) > 0, (res["m"][iq], len(ref["lam"]))
        assert np.abs(res["eval_cm"][iq] - ref["eval_cm"]).max() <= 1e-3
        if opt["jobz"] == "N":
            assert res["evec"] is None
            continue
        Z = np.asarray(res["evec"][iq])
        D = S.D(q)
        nrm = np.abs(np.linalg.eigvalsh(D)).max()
        tol = 1e-10 if prec == "D" else 1e-5
        assert np.linalg.norm(D @ Z - Z * ref["lam"], axis=0).max() <= tol * nrm
        om = np.asarray(res["eval_cm"][iq])
        v = np.asarray(res["vel"][iq])
        expect = np.zeros_like(v)
        for a in range(3):
            dD = S.D(q, a + 1)
            e = np.einsum("ik,ik->k", Z.conj(), dD @ Z).real
            expect[:, a] = e / np.maximum(2.0 * np.abs(om), 1e-3) * const
        # relative to the largest velocity of the q-point (v_z of the flat
        # layers is rounding noise on its own scale)
        vtol = (1e-7 if prec == "D" else 1e-5) * np.abs(expect).max() + 1e-9
        assert np.abs(v - expect).max() <= vtol, np.abs(v - expect).max()


def _reflect(B, rng):
    """B <- H B H for a random complex Householder H = I - 2 u u^H, in place
    (O(n^2): the spectrum of B is unchanged)."""
    n = B.shape[0]
    u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    u /= np.linalg.norm(u)
    x = B @ u
    c = np.vdot(u, x).real
    y = 2.0 * x - 2.0 * c * u
    B -= np.outer(u, y.conj())
    B -= np.outer(y, u.conj())


@pytest.mark.parametrize("prec", ["D", "S"])
def test_heevx_large_n_known_spectrum(gpu_available, prec):
    """n above the size whose tridiagonal fits the eigensolver's LDS (9216):
    bisection and inverse iteration read it from global memory.  A = H2 H1
    diag(lam) H1 H2 has the spectrum lam exactly, with a planted triple and a
    pair 1e-10 apart among the selected lowest modes."""
    lib = _lib()
    rng = np.random.default_rng(9300)
    n = 9300
    lam = np.sort(rng.uniform(-2.0, 8.0, n))
    lam[:3] = -2.5
    lam[3], lam[4] = -2.4, -2.4 + 1e-10
    A = np.diag(lam).astype(np.complex128)
    _reflect(A, rng)
    _reflect(A, rng)
    A = 0.5 * (A + A.conj().T)
    il, iu = 1, 40
    w, Z, info = lib.heevx(A, jobz="V", rng="I", il=il, iu=iu, precision=prec)
    nrm = np.abs(lam).max()
    assert info == 0 and len(w) == iu - il + 1
    tol = 1e-12 if prec == "D" else 1e-6
    assert np.abs(w - lam[il - 1:iu]).max() <= tol * nrm
    R = A @ Z - Z * w
    assert np.linalg.norm(R, axis=0).max() <= (1e-11 if prec == "D" else 1e-5) * nrm
    assert np.abs(Z.conj().T @ Z - np.eye(len(w))).max() <= (1e-10 if prec == "D" else 1e-5)


def test_range_i_fp32_low_modes_at_gamma(gpu_available):
    """FP32 + refinement with a range-I selection from the bottom: the acoustic
    triplet at Gamma (omega = K sqrt(lambda) near 0, where eigenvalue errors are
    magnified most) and the rest of the selection agree with the FP64 path to
    1e-3 cm^-1, with the modes above the selection never computed."""
    lib = _lib()
    from parphom_amd import synth
    st = synth.tbg_structure(8, 9)
    fc = synth.force_constants(st, 2008)
    q = np.zeros((1, 3))
    out = {}
    for prec, rng in (("S", "I"), ("D", "A")):
        with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="N", rng=rng, il=1, iu=40,
                        precision=prec, batch=1) as h:
            out[prec] = h.run(q)["eval_cm"][0]
    assert len(out["S"]) == 40
    assert np.abs(out["S"] - out["D"][:40]).max() <= 1e-3


@pytest.mark.parametrize("prec", ["D", "S"])
def test_block_sparse_products_match_dense(gpu_available, prec):
    """The (8,9) cell's cutoff model has 5% of its force-constant blocks
    nonzero, so the Rayleigh quotients and velocities take the block-sparse
    path by default.  FP64 path: against the dense products (a second run,
    same vectors), frequencies to 1e-6 cm^-1 and velocities to 1e-8 of the
    largest.  FP32 path: its vectors change from run to run inside
    near-degenerate pairs (float atomics in the reduction), so its velocities
    are checked against FP64 forms Re(z^H dD z) formed on the host from the
    same run's vectors (the oracle's dD), to 1e-8 of the largest, for the 30
    lowest and 200 sampled modes; its frequencies against the dense run's to
    1e-5 cm^-1 (they move by ~1e-6 between runs with the vectors)."""
    lib = _lib()
    from parphom_amd import synth
    st = synth.tbg_structure(8, 9)
    fc = synth.force_constants(st, 2008)
    qs = synth.q_grid_shifted(16, 2)[[3, 77]]
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                    precision=prec, batch=2, streams=1) as h:
        assert h.sparse(1) > 0
        r1 = h.run(qs)
        assert h.sparse(0) == 0
        r0 = h.run(qs)
    for iq in range(len(qs)):
        assert r1["info"][iq] == 0 and r0["info"][iq] == 0
        # FP32: the refined frequencies move by ~1e-6 cm^-1 between runs with the vectors
        assert np.abs(r1["eval_cm"][iq] - r0["eval_cm"][iq]).max() <= (1e-6 if prec == "D" else 1e-5)
        if prec == "D":
            v1, v0 = r1["vel"][iq], r0["vel"][iq]
            assert np.abs(v1 - v0).max() <= 1e-8 * np.abs(v0).max()
    if prec == "S":
        Sys = O.System(st.lat, st.pos, st.types, st.masses, fc)
        q = qs[0]
        dD = [Sys.D(q, a) for a in (1, 2, 3)]
        n = 3 * st.natom
        pick = np.concatenate([np.arange(30),
                               np.random.default_rng(1).choice(np.arange(30, n), 200, replace=False)])
        Z = r1["evec"][0][:, pick]
        om = r1["eval_cm"][0][pick]
        v_host = U.vel_projected(Z, dD, om, [[k] for k in range(len(pick))])
        v1 = r1["vel"][0][pick]
        assert np.abs(v1 - v_host).max() <= 1e-8 * np.abs(v_host).max()


def test_block_sparse_central_difference_velocities(gpu_available):
    """Central-difference velocities (vel_method 'C') from the block-sparse
    forms, whose phase sums take the difference (p(q+) - p(q-)) / |delta|
    before the block multiply, against the dense products of the dense
    central-difference dD on the same vectors (FP64 path, (8,9) cell): to
    1e-6 of the largest velocity (the difference quotient cancels about ten
    digits in both)."""
    lib = _lib()
    from parphom_amd import synth
    st = synth.tbg_structure(8, 9)
    fc = synth.force_constants(st, 2008)
    qs = synth.q_grid_shifted(16, 2)[[5]]
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                    vel_method="C", precision="D", batch=1, streams=1) as h:
        assert h.sparse(1) > 0
        r1 = h.run(qs)
        assert h.sparse(0) == 0
        r0 = h.run(qs)
    assert r1["info"][0] == 0 and r0["info"][0] == 0
    assert np.abs(r1["eval_cm"][0] - r0["eval_cm"][0]).max() <= 1e-9
    v1, v0 = r1["vel"][0], r0["vel"][0]
    assert np.abs(v1 - v0).max() <= 1e-6 * np.abs(v0).max()


@pytest.mark.parametrize("prec", ["D", "S"])
def test_block_sparse_range_i_with_velocities(gpu_available, prec):
    """A range-I selection (60 lowest modes, eigenvectors and velocities) on
    the block-sparse path, where the forms run over the selected modes only
    and the FP32 path adds the complement correction with the block-sparse
    D X: FP64 path against its dense products (frequencies 1e-9 cm^-1,
    velocities 1e-8 of the largest); FP32 path against the FP64 path's
    frequencies (1e-3 cm^-1)."""
    lib = _lib()
    from parphom_amd import synth
    st = synth.tbg_structure(8, 9)
    fc = synth.force_constants(st, 2008)
    q = synth.q_grid_shifted(16, 2)[[9]]
    res = {}
    for p, modes in (("D", (1, 0)), (prec, (1,))):
        for mode in modes:
            with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True, rng="I",
                            il=1, iu=60, precision=p, batch=1, streams=1) as h:
                assert (h.sparse(mode) > 0) == (mode == 1)
                r = h.run(q)
            assert r["info"][0] == 0 and len(r["eval_cm"][0]) == 60
            res[(p, mode)] = r
    d1, d0 = res[("D", 1)], res[("D", 0)]
    assert np.abs(d1["eval_cm"][0] - d0["eval_cm"][0]).max() <= 1e-9
    assert np.abs(d1["vel"][0] - d0["vel"][0]).max() <= 1e-8 * np.abs(d0["vel"][0]).max()
    if prec == "S":
        assert np.abs(res[("S", 1)]["eval_cm"][0] - d0["eval_cm"][0]).max() <= 1e-3
