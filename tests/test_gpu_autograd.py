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
