"""The FP32 path's FP64 refinement of its eigenpairs (marked gpu): the
Rayleigh-Ritz step over clusters of close eigenvalues and the residual
correction of the vectors.  FP32 eigenvectors of eigenvalues closer than the
FP32 error of the reduction mix inside their cluster, and the FP64 Rayleigh
quotients of the mixtures fall anywhere between the true eigenvalues; the
FP64 step over each cluster's span restores the eigenvalues, vectors and
velocity forms.  Every other vector carries a first-order mixture with its
neighbours, which the velocities inherit; each residual-correction step
squares it.

* A Hermitian matrix of known spectrum, A = Q diag(lambda) Q^H with Q from the
  QR factorisation of a seeded complex Gaussian matrix, whose spectrum holds
  pairs split by 1e-6 .. 1e-9 of the largest eigenvalue, a triple and an
  exactly degenerate pair, through pph_heevx (the dense forms): eigenvalues
  against lambda, the split pairs' vectors against the columns of Q, the
  degenerate pair's span against theirs.
* TBG (11,12), n = 4764, four q-points of the 128-point benchmark set, the
  block-sparse forms with velocities: FP32 + refinement against the FP64 path
  with the step on and off.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from parphom_amd import lib
    lib.load()
    return lib


def _log(*a):
    print(*a, flush=True)


def _designed_spectrum(n, lmax, rng):
    """Evenly spaced eigenvalues in [-lmax / 2, lmax] with a jitter of a fifth
    of the spacing, then the designed clusters (index, offsets relative to
    lmax).  The lowest sixteen lie below the low-mode threshold of the
    refinement, so its Rayleigh-Ritz step runs over the exactly degenerate
    triple at index 5 after the cluster step."""
    lo = -0.5 * lmax
    base = lo + (lmax - lo) * np.arange(n) / (n - 1)
    lam = base + rng.uniform(-0.2, 0.2, n) * (lmax - lo) / (n - 1)
    designs = {5: (0.0, 0.0, 0.0), 50: (0.0, 1e-6), 100: (0.0, 1e-7), 150: (0.0, 1e-8),
               200: (0.0, 1e-9), 250: (0.0, 1e-8, 3e-8), 300: (0.0, 0.0)}
    for k, offs in designs.items():
        for j, o in enumerate(offs):
            lam[k + j] = lam[k] + o * lmax
    return lam, designs


@pytest.fixture
def reduction_default(gpu_available):
    lib = _lib()
    yield lib
    lib.set_default_reduction(-1, 32)
    lib.set_default_vector_refinement(2, True)


# per (correction steps, BF16 first step): bounds on 1 - |q^H z| of the
# separated modes, on the residual |A z - w z| / lmax and on |Z^H Z - I|
# (measured at n = 384, both reductions: 2.3e-7 / 1.3e-6 / 4.6e-7 without a
# step; 1e-15 / 3.4e-11 / 6.5e-10 after one FP32 step, 7.4e-10 / 7.7e-9 /
# 1.1e-8 after one split BF16 / FP32-band step; the second step works on a
# band of 16 modes either side, which settles close pairs (the velocities'
# error) but leaves the far pairs' residue of a random spectrum: 1.4e-15 /
# 3.4e-11 / 3.1e-11 after FP32, 1.6e-15 / 7.7e-9 / 1.1e-8 after BF16)
STEP_BOUNDS = {(0, True): (1e-6, 1e-5, 1e-5), (1, False): (1e-12, 1e-10, 1e-8),
               (1, True): (1e-8, 1e-7, 1e-7), (2, False): (1e-12, 1e-10, 1e-9),
               (2, True): (1e-12, 1e-7, 1e-7)}


@pytest.mark.parametrize("steps,bf16", list(STEP_BOUNDS))
@pytest.mark.parametrize("reduction", [0, 1])
@pytest.mark.parametrize("n", [384])
def test_known_spectrum_with_close_pairs(reduction_default, n, reduction, steps, bf16):
    """One-stage and two-stage FP32 reductions (the two-stage one returned
    the exactly degenerate triple inside the low modes' Rayleigh-Ritz
    problem as an exactly diagonal block, on which the device Jacobi solver
    used to lose unitarity), with 0, 1 and 2 correction steps, the first
    with FP32 or split BF16 / FP32-band products."""
    lib = reduction_default
    lib.set_default_reduction(reduction, 32)
    lib.set_default_vector_refinement(steps, bf16)
    rng = np.random.default_rng(2024)
    lmax = 10.0
    lam, designs = _designed_spectrum(n, lmax, rng)
    assert np.all(np.diff(lam) >= 0.0)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = (Q * lam) @ Q.conj().T
    A = 0.5 * (A + A.conj().T)
    w, Z, info = lib.heevx(A, jobz="V", precision="S")
    assert info == 0 and len(w) == n
    err = np.abs(w - lam)
    _log(f"known spectrum n={n}: max |w - lambda| / lmax = {err.max() / lmax:.2e}")
    for k, offs in designs.items():
        g = list(range(k, k + len(offs)))
        _log(f"  cluster at {k} offsets {offs}: |w - lambda| / lmax = {err[g] / lmax}")
    # every eigenvalue to 1e-11 of the largest: a mixed pair would be off by
    # up to its splitting (1e-8 .. 1e-6 of lmax here)
    assert err.max() <= 1e-11 * lmax
    # the split pairs and the triple: each Ritz vector along its eigenvector
    for k, offs in designs.items():
        if k in (5, 300):
            continue
        for j in range(len(offs)):
            if min(abs(offs[j] - o) for i, o in enumerate(offs) if i != j) < 1e-8:
                continue              # 1e-9 split: below what FP64 D resolves to 1e-3
            ov = abs(np.vdot(Q[:, k + j], Z[:, k + j]))
            _log(f"  mode {k + j}: |q^H z| = {ov:.15f}")
            assert ov >= 1.0 - 1e-6, (k, j, ov)
    # the degenerate pair and the 1e-9 pair: the Ritz vectors span the pair
    for k, m in ((5, 3), (200, 2), (300, 2)):
        s = np.linalg.svd(Q[:, k:k + m].conj().T @ Z[:, k:k + m], compute_uv=False)
        assert s.min() >= 1.0 - 1e-6, (k, s)
    # every separated mode's vector, the residuals and the orthogonality
    ovb, resb, orthb = STEP_BOUNDS[(steps, bf16)]
    clustered = {k + j for k, offs in designs.items() for j in range(len(offs))}
    sep = [k for k in range(n) if k not in clustered]
    ov = np.abs(np.einsum("ij,ij->j", Q[:, sep].conj(), Z[:, sep]))
    res = np.linalg.norm(A @ Z - Z * w, axis=0).max() / lmax
    orth = np.abs(Z.conj().T @ Z - np.eye(n)).max()
    _log(f"  steps {steps}: 1 - min |q^H z| = {1 - ov.min():.1e}, residual {res:.1e}, "
         f"orthogonality {orth:.1e}")
    assert 1.0 - ov.min() <= ovb
    assert res <= resb
    assert orth <= orthb


def _resolved_dv(r, rD, nq):
    """Largest velocity difference over the modes the FP64 run resolves (its
    eigenvalues more than 1e-8 max|lambda| from their neighbours), relative
    to the largest velocity of the q-point; and the largest frequency
    difference (cm^-1)."""
    dv = dw = 0.0
    for i in range(nq):
        lam = rD["eval_cm"][i]
        l2 = np.sign(lam) * lam * lam
        gap = np.minimum(np.r_[np.inf, np.diff(l2)], np.r_[np.diff(l2), np.inf])
        ok = gap > 1e-8 * np.abs(l2).max()
        vD = rD["vel"][i]
        dv = max(dv, np.abs(r["vel"][i] - vD)[ok].max() / np.abs(vD).max())
        dw = max(dw, np.abs(r["eval_cm"][i] - lam).max())
    return dw, dv


def test_tbg_8_9_fp32_against_fp64(gpu_available):
    """The (8,9) cell (n = 2604, two-stage reduction with three real products
    per complex step) on the default FP32 path: frequencies within 1e-6
    cm^-1 and velocities of the resolved modes within 1e-6 of the largest,
    against the FP64 path."""
    lib = _lib()
    from parphom_amd import synth
    st, fc, qs = synth.workload("tbg_8_9")
    q = qs[[0, 21, 77]]
    res = {}
    for prec in ("D", "S"):
        with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True,
                        precision=prec, batch=len(q), streams=1) as h:
            res[prec] = h.run(q, evec=False)
    dw, dv = _resolved_dv(res["S"], res["D"], len(q))
    _log(f"(8,9) default FP32 path: max |d omega| = {dw:.2e} cm^-1, max |d v| / max |v| = {dv:.2e}")
    assert dw <= 1e-6 and dv <= 1e-6


def test_tbg_11_12_fp32_against_fp64(gpu_available):
    lib = _lib()
    from parphom_amd import synth
    st, fc, qs = synth.workload("tbg_11_12")
    q = qs[[0, 13, 64, 127]]
    res = {}
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True, precision="D",
                    batch=len(q), streams=1) as h:
        res["D"] = h.run(q, evec=False)
    runs = (("plain", False, 0, True), ("clusters", True, 0, True), ("one step", True, 1, True),
            ("corrected FP32", True, 2, False), ("corrected", True, 2, True))
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True, precision="S",
                    batch=len(q), streams=1) as h:
        for tag, on, steps, bf16 in runs:
            h.cluster_rr(on)
            h.vector_refinement(steps, bf16)
            res[tag] = h.run(q, evec=False)
    dw, dv = {}, {}
    for tag, _, _, _ in runs:
        dw[tag], dv[tag] = _resolved_dv(res[tag], res["D"], len(q))
        _log(f"(11,12) {tag}: max |d omega| against FP64 = {dw[tag]:.2e} cm^-1, "
             f"max |d v| / max |v| = {dv[tag]:.2e}")
    # measured (FP32 reduction and Q2 with three real products per complex
    # step): 4.1e-5 cm^-1 and 9.3e-2 with neither step (pairs closer than the
    # FP32 resolution come out mixed; 3.9e-4 and 0.40 with four products);
    # 4.2e-6 and 4.4e-3 with the cluster step alone (the first-order mixing
    # of FP32 vectors with modes just outside the cluster tolerance, 1e-5
    # max|lambda| away); 4.7e-5 of max|v| after one correction step; after
    # two, 1.7e-7 with FP32 products and 1.6e-7 with the split BF16 /
    # FP32-band first step (the default)
    assert dw["plain"] <= 1e-3
    assert dw["clusters"] <= 2e-5 and dv["clusters"] <= 2e-2
    assert dw["one step"] <= 2e-5 and dv["one step"] <= 2e-4
    assert dw["corrected FP32"] <= 2e-5 and dv["corrected FP32"] <= 1e-6
    assert dw["corrected"] <= 2e-5 and dv["corrected"] <= 1e-6


@pytest.mark.parametrize("mnk", [(77, 130, 45), (300, 261, 97), (129, 65, 33), (256, 256, 512)])
@pytest.mark.parametrize("ops", ["NN", "NC", "CN", "CC"])
def test_bf16_complex_gemm(gpu_available, mnk, ops):
    """The BF16 matrix-core product of the first correction step (complex64
    inputs rounded to BF16, FP32 accumulation) against numpy, on ragged
    shapes: every entry within 4e-3 of the sum of the terms' magnitudes
    (BF16 keeps 8 significant bits; measured at most 2.9e-3)."""
    lib = _lib()
    m, n, k = mnk
    opA, opB = ops
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    A = rng.standard_normal((m, k) if opA == "N" else (k, m)) + \
        1j * rng.standard_normal((m, k) if opA == "N" else (k, m))
    B = rng.standard_normal((k, n) if opB == "N" else (n, k)) + \
        1j * rng.standard_normal((k, n) if opB == "N" else (n, k))
    oA = A if opA == "N" else A.conj().T
    oB = B if opB == "N" else B.conj().T
    C, _ = lib.debug_zgemm(A, B, opA, opB, "S", mode=6)
    scale = np.abs(oA) @ np.abs(oB)
    assert np.max(np.abs(C - oA @ oB) / scale) <= 4e-3
