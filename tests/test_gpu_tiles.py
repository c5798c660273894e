"""The GEMM tile shapes the autotuner chooses among (64x64, 128x64, 64x128,
128x128): every op combination, split-K, two-product, triangular and
lower-only forms on ragged shapes against numpy, and whole eigensolves with
each shape against the oracle (marked gpu)."""
import numpy as np
import pytest

from tests import _util as U

pytestmark = pytest.mark.gpu

TILES = ["64x64", "128x64", "64x128", "128x128"]


def _lib():
    from parphom_amd import lib
    lib.load()
    return lib


@pytest.fixture
def tile_default(gpu_available):
    lib = _lib()
    yield lambda t: lib.set_default_gemm_tile(t)
    lib.set_default_gemm_tile("64x64")


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("prec,tol", [("D", 1e-12), ("S", 2e-5)])
@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("mnk", [(77, 130, 45), (300, 261, 97), (129, 65, 33)])
def test_gemm_tile_shapes(tile_default, prec, tol, tile, mnk):
    """C = op(A) op(B) for NN, CN, NC, CC, split-K (mode 2) and the
    two-product form (mode 3) with the given tile shape; the 128x128 tile is
    FP32-only and FP64 rejects it."""
    lib = _lib()
    tile_default(tile)
    rng = np.random.default_rng(sum(mnk))
    m, n, k = mnk
    for opA in "NC":
        for opB in "NC":
            A = _rand(rng, (m, k) if opA == "N" else (k, m))
            B = _rand(rng, (k, n) if opB == "N" else (n, k))
            ref = (A if opA == "N" else A.conj().T) @ (B if opB == "N" else B.conj().T)
            if prec == "D" and tile == "128x128":
                with pytest.raises(lib.PphError):
                    lib.debug_zgemm(A, B, opA, opB, prec=prec, mode=0)
                return
            C, _ = lib.debug_zgemm(A, B, opA, opB, prec=prec, mode=0)
            assert np.abs(C - ref).max() <= tol * np.abs(ref).max() * 10, (opA, opB)
    A = _rand(rng, (m, k))
    B = _rand(rng, (k, n))
    ref = A @ B
    for mode in (2, 3):
        kk = k - k % 32 if mode == 3 else k          # the two-product split is at a multiple of 16
        C, _ = lib.debug_zgemm(A[:, :kk], B[:kk], "N", "N", prec=prec, mode=mode)
        r = A[:, :kk] @ B[:kk]
        assert np.abs(C - r).max() <= tol * np.abs(r).max() * 10, mode


@pytest.mark.parametrize("prec,tol", [("D", 1e-12), ("S", 2e-5)])
@pytest.mark.parametrize("tile", ["128x64", "64x128", "128x128"])
def test_gemm_tile_hermitian_from_lower(tile_default, prec, tol, tile):
    """A B for Hermitian A from its lower triangle (triangular modes 1 and 2
    of the band reduction) with the larger tiles."""
    if prec == "D" and tile == "128x128":
        pytest.skip("FP32-only tile")
    lib = _lib()
    tile_default(tile)
    rng = np.random.default_rng(4)
    for n, k in ((200, 32), (333, 17)):
        X = _rand(rng, (n, n))
        H = X + X.conj().T
        B = _rand(rng, (n, k))
        L = np.tril(H)
        L[np.triu_indices(n, 1)] = 1e30          # the upper triangle must not be read
        C, _ = lib.debug_zgemm(L, B, "N", "N", prec=prec, mode=5)
        ref = H @ B
        assert np.abs(C - ref).max() <= tol * np.abs(ref).max() * 10, n


@pytest.mark.parametrize("prec", ["D", "S"])
@pytest.mark.parametrize("tile", TILES)
def test_run_with_tile(gpu_available, prec, tile):
    """The config-1 structure (n = 444: one-stage reduction with panel
    updates and the blocked back-transformation) with each tile shape on the
    handle: frequencies against the oracle and eigenvector residuals."""
    lib = _lib()
    st, fc = U.structure(3, 4, 1003)
    S = U.oracle_system(3, 4, 1003)
    from parphom_amd import synth
    qs = synth.q_path_gkmg()[[0, 13, 30]]
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                    precision=prec) as h:
        if prec == "D" and tile == "128x128":
            with pytest.raises(lib.PphError):
                h.gemm_tile(tile)
            return
        h.gemm_tile(tile)
        res = h.run(qs)
    for iq, q in enumerate(qs):
        ref = S.solve(q)
        assert res["info"][iq] == 0
        assert np.abs(res["eval_cm"][iq] - ref["eval_cm"]).max() <= 1e-3
        Z = np.asarray(res["evec"][iq])
        nrm = np.abs(ref["lam"]).max()
        tol = 1e-10 if prec == "D" else 1e-5
        assert np.linalg.norm(S.D(q) @ Z - Z * ref["lam"], axis=0).max() <= tol * nrm


@pytest.mark.parametrize("tile", ["128x64", "128x128"])
def test_two_stage_with_tile(tile_default, tile):
    """The two-stage reduction (n = 600, FP32) with the larger tiles."""
    lib = _lib()
    tile_default(tile)
    lib.set_default_reduction(1, 32)
    try:
        rng = np.random.default_rng(21)
        n = 600
        X = _rand(rng, (n, n))
        A = X + X.conj().T
        w, Z, info = lib.heevx(A, jobz="V", precision="S")
    finally:
        lib.set_default_reduction(-1, 32)
    wref = np.linalg.eigvalsh(A)
    nrm = np.abs(wref).max()
    assert info == 0
    assert np.abs(w - wref).max() <= 1e-5 * nrm
    assert np.linalg.norm(A @ Z - Z * w, axis=0).max() <= 1e-5 * nrm
