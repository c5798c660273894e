"""Run-to-run determinism of the HIP path (marked gpu).

No kernel of the phonon path adds into memory with atomics: every sum across
workgroups (the reduction's hemv, the GEMM dot epilogue, split-K, the
block-sparse quadratic forms) stores per-workgroup partials that a later
kernel adds in a fixed order.  So two runs on the same inputs give the same
bits, with the same handle or a fresh one, one lane or two.
"""
import numpy as np
import pytest

from tests import _util as U

pytestmark = pytest.mark.gpu


def _lib():
    from parphom_amd import lib
    lib.load()
    return lib


def _same(a, b, what):
    for key in ("eval_cm", "m", "nz", "info", "evec", "vel"):
        x, y = a.get(key), b.get(key)
        if x is None and y is None:
            continue
        if isinstance(x, list):
            assert len(x) == len(y)
            for i, (u, v) in enumerate(zip(x, y)):
                assert np.array_equal(np.asarray(u), np.asarray(v)), f"{what}: {key}[{i}] differs"
        else:
            assert np.array_equal(np.asarray(x), np.asarray(y)), f"{what}: {key} differs"


@pytest.mark.parametrize("prec,cell,sparse", [("D", (3, 4), 0), ("D", (3, 4), 2),
                                              ("D", (8, 9), 1), ("S", (8, 9), 1),
                                              ("S", (3, 4), 0)])
def test_run_bit_identical(gpu_available, prec, cell, sparse):
    """pph_run twice on one handle and once on a fresh handle with two lanes:
    frequencies, eigenvectors and velocities bit for bit.  The (3,4) cell
    (n = 444) runs the dense products (GEMM dot epilogue) or the block-sparse
    forms; the (8,9) cell (n = 2604) the block-sparse forms, and on the FP32
    path the two-stage reduction."""
    lib = _lib()
    from parphom_amd import synth
    st, fc = U.structure(*cell, 41)
    qs = synth.q_random(6, 43)
    opt = dict(jobz="V", vel=True, precision=prec, batch=3)
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, streams=1, **opt) as h:
        h.sparse(sparse)
        r1 = h.run(qs)
        r2 = h.run(qs)
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, streams=2, **opt) as h:
        h.sparse(sparse)
        r3 = h.run(qs)
    assert all(i == 0 for i in r1["info"])
    _same(r1, r2, "same handle")
    _same(r1, r3, "fresh handle, two lanes")


@pytest.mark.parametrize("prec", ["D", "S"])
def test_heevx_split_k_bit_identical(gpu_available, prec):
    """A single matrix of n = 700: the back-transformation's Wt = V^H Z then
    runs as split-K (3 pieces, summed by a second kernel in order)."""
    lib = _lib()
    rng = np.random.default_rng(71)
    n = 700
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = X + X.conj().T
    w1, Z1, i1 = lib.heevx(A.copy(), jobz="V", precision=prec)
    w2, Z2, i2 = lib.heevx(A.copy(), jobz="V", precision=prec)
    assert i1 == 0 and i2 == 0
    assert np.array_equal(w1, w2) and np.array_equal(Z1, Z2)


@pytest.mark.parametrize("prec", ["D", "S"])
def test_gemm_dot_and_split_k_bit_identical(gpu_available, prec):
    """The GEMM's dot epilogue (mode 1, Hermitian mode 4) and split-K (mode 2)
    through the test hook, twice each."""
    lib = _lib()
    rng = np.random.default_rng(72)
    m, n, k = 333, 150, 333
    A = rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
    H = A + A.conj().T
    B = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    X = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    for mode, AA, Cin in ((1, A, X), (4, H, B), (2, A, None)):
        C1, d1 = lib.debug_zgemm(AA, B, prec=prec, mode=mode, Cin=Cin)
        C2, d2 = lib.debug_zgemm(AA, B, prec=prec, mode=mode, Cin=Cin)
        assert np.array_equal(C1, C2) and np.array_equal(d1, d2), mode
