"""Differentiation through the HIP pipeline (marked gpu): d omega / d q and
d omega / d(spring parameters) from the autograd Function against central
finite differences of the same forward (FP64 path), and the FP32+refine path
against the FP64 path."""
import numpy as np
import pytest
import torch

from parphom_amd import synth

pytestmark = pytest.mark.gpu


def _setup(prec="D"):
    from parphom_amd.autograd import SpringPhonons
    st = synth.tbg_structure(3, 4)
    model = SpringPhonons(st, 1003, precision=prec)
    rng = np.random.default_rng(44)
    q = np.array([[0.137, 0.291, 0.0], [0.61, 0.18, 0.0]])
    w = rng.uniform(0.0, 1.0, size=(len(q), 3 * st.natom))
    return model, q, w


def _loss(model, q, p, w):
    om = model.frequencies(torch.tensor(q), torch.tensor(p))
    return float((torch.tensor(w) * om).sum())


def test_gradients_vs_finite_differences(gpu_available):
    model, q, w = _setup("D")
    p0 = synth.default_params()
    qt = torch.tensor(q, requires_grad=True)
    pt = torch.tensor(p0, requires_grad=True)
    om = model.frequencies(qt, pt)
    (torch.tensor(w) * om).sum().backward()
    gq, gp = qt.grad.numpy(), pt.grad.numpy()
    # d/dq (reduced), central differences with h = 1e-6
    h = 1e-6
    for i in range(len(q)):
        for k in (0, 1):
            dq = np.zeros_like(q)
            dq[i, k] = h
            fd = (_loss(model, q + dq, p0, w) - _loss(model, q - dq, p0, w)) / (2 * h)
            assert abs(fd - gq[i, k]) <= 1e-5 * abs(gq[i, k]) + 1e-3, (i, k, fd, gq[i, k])
    # d/d params, central differences with h = 1e-5
    h = 1e-5
    for k in range(len(p0)):
        dp = np.zeros_like(p0)
        dp[k] = h
        fd = (_loss(model, q, p0 + dp, w) - _loss(model, q, p0 - dp, w)) / (2 * h)
        assert abs(fd - gp[k]) <= 1e-5 * abs(gp[k]) + 1e-3, (k, fd, gp[k])


def test_fp32_path_gradients_close_to_fp64(gpu_available):
    m64, q, w = _setup("D")
    m32, _, _ = _setup("S")
    p0 = synth.default_params()
    res = []
    for m in (m64, m32):
        qt = torch.tensor(q, requires_grad=True)
        pt = torch.tensor(p0, requires_grad=True)
        (torch.tensor(w) * m.frequencies(qt, pt)).sum().backward()
        res.append((qt.grad.numpy(), pt.grad.numpy()))
    assert np.allclose(res[0][0], res[1][0], rtol=1e-3, atol=1e-2)
    assert np.allclose(res[0][1], res[1][1], rtol=1e-4, atol=1e-3)


def test_full_size_tbg_11_12(gpu_available):
    """(11,12), N = 1588, n = 4764.  d loss / d q against the analytic group
    velocities (velocity products and unit factor of write_output.F90:443-446,
    a code path independent of the autograd adjoint) on the FP64 path; the
    FP32+refine gradients against the FP64 ones, relative to the scale of the
    summed terms (the weighted sum of 4764 signed velocities cancels to a small
    remainder; FP32 eigenvectors of near-degenerate moire modes mix inside
    their pairs, to first order in the FP32 vector error, which moves
    individual velocities, not pair sums, so with unequal weights up to
    ~7e-5 of the term scale reach the sum on the two-stage reduction, whose
    FP32 vectors carry twice the one-stage residual; the bound is 1e-4); and
    d loss / d (spring constants) against central differences of the FP64
    forward."""
    from parphom_amd.autograd import SpringPhonons, PI32
    st = synth.tbg_structure(11, 12)
    m64 = SpringPhonons(st, 4011, precision="D")
    m32 = SpringPhonons(st, 4011, precision="S")
    q = synth.q_random(2, 44)
    rng = np.random.default_rng(45)
    n = 3 * st.natom
    w = rng.uniform(0.0, 1.0, size=(len(q), n))
    p0 = synth.default_params()
    grads = {}
    for tag, m in (("D", m64), ("S", m32)):
        qt = torch.tensor(q, requires_grad=True)
        pt = torch.tensor(p0, requires_grad=True)
        (torch.tensor(w) * m.frequencies(qt, pt)).sum().backward()
        grads[tag] = (qt.grad.numpy(), pt.grad.numpy())
    C2 = float(np.float32(33.35641))
    with m64.handle(p0, jobz="V", vel=True, precision="D") as h:
        res = h.run(q)
    for i in range(len(q)):
        v = np.asarray(res["vel"][i])
        conv = C2 / (1e2 * 2 * PI32)
        ref = ((w[i][:, None] * v).sum(axis=0) * conv) @ m64.cart_to_red.T
        scale = ((w[i][:, None] * np.abs(v)).sum(axis=0) * conv) @ np.abs(m64.cart_to_red.T)
        g64, g32 = grads["D"][0][i], grads["S"][0][i]
        assert np.all(np.abs(g64[:2] - ref[:2]) <= 1e-9 * scale[:2] + 1e-6), (i, g64, ref)
        assert np.all(np.abs(g32[:2] - g64[:2]) <= 1e-4 * scale[:2]), (i, g32, g64, scale)
    # two spring constants by central differences of the FP64 forward (one q)
    q1 = q[:1]
    # steps of 1e-6 relative: larger ones let near-degenerate moire modes swap
    # order inside +-h, where the sorted frequencies have kinks
    for k in (0, 6):                       # first-shell radial, interlayer
        hk = 1e-6 * max(1.0, abs(p0[k]))
        dp = np.zeros_like(p0)
        dp[k] = hk
        lp = float((torch.tensor(w[:1]) * m64.frequencies(torch.tensor(q1), torch.tensor(p0 + dp))).sum())
        lm = float((torch.tensor(w[:1]) * m64.frequencies(torch.tensor(q1), torch.tensor(p0 - dp))).sum())
        fd = (lp - lm) / (2 * hk)
        pt1 = torch.tensor(p0, requires_grad=True)
        (torch.tensor(w[:1]) * m64.frequencies(torch.tensor(q1), pt1)).sum().backward()
        g = pt1.grad.numpy()[k]
        assert abs(fd - g) <= 1e-5 * abs(g) + 1e-2, (k, fd, g)
        assert abs(grads["S"][1][k] - grads["D"][1][k]) <= 1e-4 * abs(grads["D"][1][k]) + 1e-2
