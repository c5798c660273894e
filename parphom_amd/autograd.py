"""Differentiable phonon frequencies: a torch.autograd.Function over the HIP
pipeline.

    model = SpringPhonons(structure, seed)              # spring-model force constants
    omega = model.frequencies(q, params)                # (nq, n) cm^-1, torch float64
    loss = (w * omega).sum(); loss.backward()           # q.grad, params.grad

Forward: force constants from the parameters (the model is linear in them,
parphom_amd.synth), then pph_run for frequencies, eigenvectors and the
handle reused for the adjoint.  Backward (Hellmann-Feynman):
d lambda_k / dx = <e_k| dD/dx |e_k> for x = a Cartesian q component
(pph_expect with what = 1..3) or a parameter (pph_expect on a handle built
from that parameter's basis force constants, dD/d theta = D[basis]).
Inside clusters of degenerate eigenvalues the derivative of the sorted
eigenvalues is the sorted spectrum of the projected block Z_c^H dD/dx Z_c
(degenerate-subspace projection).  d omega / d lambda = K^2 / max(2|omega|,
1e-3) with K = f32(15.633302) f32(33.35641), the conversion of
write_output.F90:135 and the floor of write_output.F90:443.
Reduced q: dD/dq_red = 2 pi32 (lat^T)^-1 dD/dq_cart, dq_cart as the analytic
derivative of dynamical_matrix.F90:249-253 defines it.
"""
from __future__ import annotations

import numpy as np
import torch

from . import lib, synth

C1 = float(np.float32(15.633302))
C2 = float(np.float32(33.35641))
K2 = (C1 * C2) ** 2
PI32 = float(np.float32(np.pi))


def _clusters(lam, rel):
    tol = rel * max(1e-300, float(np.abs(lam).max()))
    out, cur = [], [0]
    for k in range(1, len(lam)):
        if lam[k] - lam[k - 1] <= tol:
            cur.append(k)
        else:
            if len(cur) > 1:
                out.append(cur)
            cur = [k]
    if len(cur) > 1:
        out.append(cur)
    return out


class SpringPhonons:
    """Spring-model twisted bilayer graphene with differentiable frequencies."""

    def __init__(self, structure: synth.Structure, seed: int, precision: str = "D",
                 cluster_rel: float = 1e-8):
        self.st = structure
        self.seed = seed
        self.basis = synth.spring_basis(structure, seed)
        self.precision = precision
        self.cluster_rel = cluster_rel
        self._basis_handles = None
        n = 3 * structure.natom
        self.n = n
        lat = np.asarray(structure.lat, np.float64)
        # dD/dq_red = 2 pi32 (lat^T)^-1 dD/dq_cart  (rows: reduced components)
        self.cart_to_red = 2.0 * PI32 * np.linalg.inv(lat.T)

    def fc(self, params):
        return synth.force_constants(self.st, self.seed, params=params, basis=self.basis)

    def handle(self, params, **kw):
        st = self.st
        return lib.Handle(st.lat, st.pos, st.types, st.masses, self.fc(params), **kw)

    def basis_handles(self):
        if self._basis_handles is None:
            st = self.st
            hs = []
            for p in range(len(synth.PARAM_NAMES)):
                e = np.zeros(len(synth.PARAM_NAMES))
                e[p] = 1.0
                hs.append(lib.Handle(st.lat, st.pos, st.types, st.masses, self.fc(e)))
            self._basis_handles = hs
        return self._basis_handles

    def frequencies(self, q, params):
        return _Frequencies.apply(q, params, self)


def _expect(h, q, Z, what):
    """Re diag(Z^H X Z) per q; Z (nq, n, nz) complex128 host array."""
    import ctypes as C
    nq, n, nz = Z.shape
    Zc = np.ascontiguousarray(Z.transpose(0, 2, 1))          # column-major n x nz per q
    qf = np.asfortranarray(np.atleast_2d(np.asarray(q, np.float64)))
    out = np.zeros((nq, nz))
    lib._check(lib.load().pph_expect(h._h, nq, lib._ptr(qf), nq, what, Zc.ctypes.data, nz, nz,
                                     lib._ptr(out), 0), h._h)
    return out


def _dlam(h, q, Z, lam, what, cluster_rel):
    """d lambda / dx for every mode, with projected blocks inside clusters."""
    d = _expect(h, q, Z, what)
    for iq in range(Z.shape[0]):
        groups = _clusters(lam[iq], cluster_rel)
        if not groups:
            continue
        X = h.assemble(q[iq:iq + 1], what)[0]
        for g in groups:
            Zc = Z[iq][:, g]
            blk = Zc.conj().T @ (X @ Zc)
            d[iq, g] = np.sort(np.linalg.eigvalsh(0.5 * (blk + blk.conj().T)))
    return d


class _Frequencies(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, params, model):
        qn = q.detach().cpu().numpy().astype(np.float64)
        pn = params.detach().cpu().numpy().astype(np.float64)
        with model.handle(pn, jobz="V", precision=model.precision) as h:
            res = h.run(qn)
        om = np.stack(res["eval_cm"])                          # (nq, n)
        Z = np.stack(res["evec"])                              # (nq, n, n)
        ctx.model = model
        ctx.qn, ctx.pn, ctx.om, ctx.Z = qn, pn, om, Z
        return torch.from_numpy(om.copy())

    @staticmethod
    def backward(ctx, g):
        model = ctx.model
        qn, pn, om, Z = ctx.qn, ctx.pn, ctx.om, ctx.Z
        gn = g.detach().cpu().numpy()
        lam = np.sign(om) * (om / (C1 * C2)) ** 2
        dw = K2 / np.maximum(2.0 * np.abs(om), 1e-3)           # d omega / d lambda
        coef = gn * dw
        grad_q = grad_p = None
        if ctx.needs_input_grad[0]:
            with model.handle(pn) as h:
                dcart = np.stack([_dlam(h, qn, Z, lam, a, model.cluster_rel) for a in (1, 2, 3)],
                                 axis=-1)                       # (nq, n, 3)
            gc = np.einsum("qk,qka->qa", coef, dcart)
            grad_q = torch.from_numpy(gc @ model.cart_to_red.T)
        if ctx.needs_input_grad[1]:
            gp = np.zeros(len(pn))
            for p, hb in enumerate(model.basis_handles()):
                dp = _dlam(hb, qn, Z, lam, 0, model.cluster_rel)
                gp[p] = float((coef * dp).sum())
            grad_p = torch.from_numpy(gp)
        return grad_q, grad_p, None
