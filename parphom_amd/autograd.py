"""Differentiable phonon frequencies: a torch.autograd.Function over the HIP
pipeline.

    model = SpringPhonons(structure, seed)              # spring-model force constants
    omega = model.frequencies(q, params)                # (nq, n) cm^-1, torch float64
    loss = (w * omega).sum(); loss.backward()           # q.grad, params.grad

Forward: force constants from the parameters (the model is linear in them,
parphom_amd.synth), then pph_run_device: frequencies and eigenvectors stay in
GPU memory (torch tensors the library writes), and only the frequencies come
back to the host.  Backward (Hellmann-Feynman), on the device:
d lambda_k / dx = <e_k| dD/dx |e_k> for x = a Cartesian q component
(pph_expect with what = 1..3) or a parameter (pph_expect on a handle built
from that parameter's basis force constants, dD/d theta = D[basis]), with the
eigenvectors read where the forward left them.
Inside clusters of degenerate eigenvalues the derivative of the sorted
eigenvalues is the sorted spectrum of the projected block Z_c^H dD/dx Z_c
(degenerate-subspace projection).  Its off-diagonal entries come from the
same device product by polarisation, B_ab = ((a + b)^H X (a + b) - B_aa -
B_bb) / 2 + i (B_aa + B_bb - (a + i b)^H X (a + i b)) / 2, on vectors formed
on the GPU; only the k x k eigenvalue problems are solved on the host.
d omega / d lambda = K^2 / max(2|omega|, 1e-3) with K = f32(15.633302)
f32(33.35641), the conversion of write_output.F90:135 and the floor of
write_output.F90:443.
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
    """Re diag(Z^H X Z) per q on the device: Z (nq, cols, n) complex128 CUDA
    tensor (each q's vectors column-major, the layout pph_run_device writes),
    X = D(q) (what 0) or dD/dq_{x,y,z} (what 1..3) of handle h; returns a
    (nq, cols) float64 CUDA tensor."""
    import ctypes as C
    nq, cols, n = Z.shape
    if n != h.n or Z.dtype != torch.complex128 or not Z.is_cuda or not Z.is_contiguous():
        raise ValueError("_expect: Z must be a contiguous complex128 CUDA tensor (nq, cols, n)")
    qf = np.asfortranarray(np.atleast_2d(np.asarray(q, np.float64)))
    out = torch.empty((nq, cols), dtype=torch.float64, device=Z.device)
    lib._check(lib.load().pph_expect(h._h, nq, lib._ptr(qf), nq, what, C.c_void_p(Z.data_ptr()),
                                     cols, cols,
                                     C.cast(C.c_void_p(out.data_ptr()), C.POINTER(C.c_double)), 1),
               h._h)
    return out


def _dlam(h, q, Z, lam, what, cluster_rel):
    """d lambda / dx for every mode (nq, n) on the host, with the projected
    blocks inside clusters of degenerate eigenvalues."""
    d = _expect(h, q, Z, what).cpu().numpy()
    groups = [_clusters(lam[iq], cluster_rel) for iq in range(Z.shape[0])]
    qs = [iq for iq in range(Z.shape[0]) if groups[iq]]
    if not qs:
        return d
    # polarisation vectors of every pair inside every cluster, per q
    pairs = {iq: [(a, b) for g in groups[iq] for ia, a in enumerate(g) for b in g[ia + 1:]]
             for iq in qs}
    npol = 2 * max(len(v) for v in pairs.values())
    P = torch.zeros((len(qs), npol, Z.shape[2]), dtype=torch.complex128, device=Z.device)
    for r, iq in enumerate(qs):
        a = torch.tensor([x for x, _ in pairs[iq]], device=Z.device)
        b = torch.tensor([y for _, y in pairs[iq]], device=Z.device)
        k = len(pairs[iq])
        P[r, 0:2 * k:2] = Z[iq, a] + Z[iq, b]
        P[r, 1:2 * k:2] = Z[iq, a] + 1j * Z[iq, b]
    pol = _expect(h, np.asarray(q)[qs], P, what).cpu().numpy()
    for r, iq in enumerate(qs):
        off = {}
        for t, (a, b) in enumerate(pairs[iq]):
            s = d[iq, a] + d[iq, b]
            off[(a, b)] = 0.5 * (pol[r, 2 * t] - s) + 0.5j * (s - pol[r, 2 * t + 1])
        for g in groups[iq]:
            blk = np.diag(d[iq, g]).astype(np.complex128)
            for ia, a in enumerate(g):
                for ib in range(ia + 1, len(g)):
                    blk[ia, ib] = off[(a, g[ib])]
                    blk[ib, ia] = np.conj(off[(a, g[ib])])
            d[iq, g] = np.sort(np.linalg.eigvalsh(blk))
    return d


class _Frequencies(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, params, model):
        qn = q.detach().cpu().numpy().astype(np.float64)
        pn = params.detach().cpu().numpy().astype(np.float64)
        nq, n = len(qn), model.n
        dev = torch.device("cuda", 0)
        out = {"eval_cm": torch.zeros((nq, n), dtype=torch.float64, device=dev),
               "m": torch.zeros(nq, dtype=torch.int64, device=dev),
               "nz": torch.zeros(nq, dtype=torch.int64, device=dev),
               "info": torch.zeros(nq, dtype=torch.int32, device=dev),
               "evec": torch.empty((nq, n, n), dtype=torch.complex128, device=dev)}
        with model.handle(pn, jobz="V", precision=model.precision) as h:
            h.run_device(qn, out)
        torch.cuda.synchronize()
        if int(out["m"].min()) != n or int(out["info"].abs().max()) != 0:
            raise RuntimeError("frequencies: the eigensolve did not return every mode")
        om = out["eval_cm"].cpu().numpy()                      # (nq, n)
        ctx.model = model
        ctx.qn, ctx.pn, ctx.om, ctx.Z = qn, pn, om, out["evec"]
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
