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
    remainder; the FP32 path's vectors are corrected in FP64 by default, and
    the two agree to 4e-10 of the term scale; without the correction the FP32
    vectors of near-degenerate moire modes mix inside their pairs and moved
    the sum by up to 7e-5); and
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
        assert np.all(np.abs(g32[:2] - g64[:2]) <= 1e-8 * scale[:2]), (i, g32, g64, scale)
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


def test_config4_tbg_11_12_128q(gpu_available):
    """The differentiation workload at full size: (11,12), n = 4764, the 128
    seeded q of the bench workload, loss = sum_q sum_nu w_nu omega_nu(q) with
    w ~ U(0,1) (seed 44), FP32 + FP64-refinement forward, eigenvectors and
    adjoint kept on the GPU.

    * d loss / d q against the analytic group velocities of a separate
      pph_run_device with velocities (the block-sparse velocity forms and the
      unit factor of write_output.F90:443-446, a code path independent of the
      adjoint's pph_expect products), on the same FP32 vectors: to 1e-7 of
      the scale of the summed terms (measured 6e-9).  Inside a cluster of
      degenerate modes the adjoint differentiates the sorted spectrum of the
      projected block and the velocities are its diagonal, so for this check
      the cluster's weights are equalised (their trace agrees);
    * d loss / d (k_r1, k_t1, k_inter) with the seeded weights against
      central differences of the same forward with h = 1e-5 (values only):
      to 1e-4 relative (measured 3e-6 - 7e-6; the difference of two FP32-path
      losses over 610k modes carries their rounding)."""
    from parphom_amd.autograd import SpringPhonons, PI32, C1, C2, _clusters
    st, _, qs = synth.workload("tbg_11_12")
    q = qs[:128]
    n = 3 * st.natom
    model = SpringPhonons(st, synth.WORKLOADS["tbg_11_12"][2], precision="S")
    w = np.random.default_rng(44).uniform(0.0, 1.0, size=n)
    p0 = synth.default_params()
    qt = torch.tensor(q, requires_grad=True)
    pt = torch.tensor(p0, requires_grad=True)
    om = model.frequencies(qt, pt)
    (gp,) = torch.autograd.grad(om, pt, grad_outputs=torch.tensor(w)[None, :].expand(len(q), n),
                                retain_graph=True)
    gp = gp.numpy()
    # inside a cluster of degenerate modes the adjoint differentiates the
    # sorted spectrum of the projected block, the velocities are the
    # diagonal; with the cluster's weights equalised both give its trace
    omn = om.detach().numpy()
    lam = np.sign(omn) * (omn / (C1 * C2)) ** 2
    w_eff = np.tile(w, (len(q), 1))
    ncl = 0
    for iq in range(len(q)):
        for g in _clusters(lam[iq], model.cluster_rel):
            w_eff[iq, g] = w[g].mean()
            ncl += 1
    (gq,) = torch.autograd.grad(om, qt, grad_outputs=torch.tensor(w_eff))
    gq = gq.numpy()
    del om
    torch.cuda.empty_cache()
    print(f"config 4: adjoint done, |dL/dp| = {np.abs(gp).max():.3e}, {ncl} degenerate clusters",
          flush=True)

    dev = torch.device("cuda", 0)

    def run(p, vel):
        out = {"eval_cm": torch.zeros((len(q), n), dtype=torch.float64, device=dev),
               "m": torch.zeros(len(q), dtype=torch.int64, device=dev),
               "info": torch.zeros(len(q), dtype=torch.int32, device=dev)}
        if vel:
            out["nz"] = torch.zeros(len(q), dtype=torch.int64, device=dev)
            out["evec"] = torch.empty((len(q), n, n), dtype=torch.complex128, device=dev)
            out["vel"] = torch.zeros((len(q), n, 3), dtype=torch.float64, device=dev)
        with model.handle(p, jobz="V" if vel else "N", vel=vel, precision="S") as h:
            h.run_device(q, out)
        torch.cuda.synchronize()
        assert int(out["info"].abs().max()) == 0 and int(out["m"].min()) == n
        res = {k: v.cpu().numpy() for k, v in out.items() if k != "evec"}
        del out
        torch.cuda.empty_cache()
        return res

    r = run(p0, True)
    conv = float(np.float32(33.35641)) / (1e2 * 2 * PI32)
    v = r["vel"]                                            # (nq, n, 3) m/s
    ref = (np.einsum("qk,qka->qa", w_eff, v) * conv) @ model.cart_to_red.T
    scale = (np.einsum("qk,qka->qa", w_eff, np.abs(v)) * conv) @ np.abs(model.cart_to_red.T)
    err = np.abs(gq[:, :2] - ref[:, :2]) / scale[:, :2]
    print(f"config 4: d/dq vs group velocities, max rel {err.max():.2e}", flush=True)
    assert err.max() <= 1e-7

    def loss(p):
        return float((w[None, :] * run(p, False)["eval_cm"]).sum())

    h = 1e-5
    for k in (0, 3, 6):                    # k_r1, k_t1, k_inter
        dp = np.zeros_like(p0)
        dp[k] = h
        fd = (loss(p0 + dp) - loss(p0 - dp)) / (2 * h)
        print(f"config 4: {synth.PARAM_NAMES[k]} adjoint {gp[k]:.6e} fd {fd:.6e}", flush=True)
        assert abs(fd - gp[k]) <= 1e-4 * abs(gp[k]), (k, fd, gp[k])
