"""ctypes binding of libparphom_hip.so (the C ABI in include/parphom_hip.h).

The library is built in-tree by `parphom_amd/csrc/Makefile`; importing this
module loads it and raises if it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparphom_hip.so")

PPH_INFO_NOCONV, PPH_INFO_ORTH, PPH_INFO_MEMORY, PPH_INFO_NOTCOMP = 1, 2, 4, 8

_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)


class PphSystem(C.Structure):
    _fields_ = [("natom", C.c_int64), ("ntypes", C.c_int64), ("lat", _dp), ("pos", _dp),
                ("mass", _dp), ("type", _i64p)]


class PphOpts(C.Structure):
    _fields_ = [("jobz", C.c_char), ("range", C.c_char), ("vel_method", C.c_char),
                ("precision", C.c_char), ("vel", C.c_int32), ("device", C.c_int32),
                ("vl", C.c_double), ("vu", C.c_double), ("il", C.c_int64), ("iu", C.c_int64),
                ("abstol", C.c_double), ("orfac", C.c_double), ("batch", C.c_int64),
                ("streams", C.c_int32)]


class PphOut(C.Structure):
    _fields_ = [("eval_cm", _dp), ("ld_eval", C.c_int64), ("m", _i64p), ("nz", _i64p),
                ("evec", C.c_void_p), ("evec_cols", C.c_int64), ("vel", _dp), ("info", _i32p)]


EXPORTED = ("pph_create", "pph_run", "pph_run_device", "pph_assemble", "pph_expect", "pph_heevx",
            "pph_dos_triangulation", "pph_dos_triangulation_noalloc", "pph_dos_smearing",
            "pph_fc_symmetrize", "pph_spring_fc", "pph_destroy",
            "pph_last_error",
            "pph_version", "pph_debug_zgemm", "pph_profile", "pph_debug_blocking",
            "pph_debug_sparse", "pph_debug_reduction", "pph_debug_graphs", "pph_debug_gemm_tile",
            "pph_debug_cluster_rr", "pph_debug_vector_refinement")

# GEMM tile shapes of pph_debug_gemm_tile (0: the process default)
GEMM_TILES = {"64x64": 1, "128x64": 2, "64x128": 3, "128x128": 4}

PHASES = ("assemble", "tridiag", "tri_eig", "back_transform", "refine", "velocity", "total",
          "hemv_launches", "he2hb", "hb2st", "q2")

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library (once).  Raises OSError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is not built; run `make -C parphom_amd/csrc` "
                      "(or __graft_entry__.build())")
    lib = C.CDLL(path)
    lib.pph_create.argtypes = [C.POINTER(PphSystem), _dp, C.POINTER(PphOpts), C.POINTER(C.c_void_p)]
    lib.pph_run.argtypes = [C.c_void_p, C.c_int64, _dp, C.c_int64, C.POINTER(PphOut)]
    lib.pph_run_device.argtypes = [C.c_void_p, C.c_int64, _dp, C.c_int64, C.POINTER(PphOut)]
    lib.pph_assemble.argtypes = [C.c_void_p, C.c_int64, _dp, C.c_int64, C.c_int32, C.c_void_p,
                                 C.c_int32]
    lib.pph_heevx.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.POINTER(PphOpts), _dp,
                              C.c_void_p, _i64p, _i32p, C.c_int32]
    lib.pph_destroy.argtypes = [C.c_void_p]
    lib.pph_last_error.argtypes = [C.c_void_p]
    lib.pph_last_error.restype = C.c_char_p
    lib.pph_version.restype = C.c_char_p
    lib.pph_expect.argtypes = [C.c_void_p, C.c_int64, _dp, C.c_int64, C.c_int32, C.c_void_p,
                               C.c_int64, C.c_int64, _dp, C.c_int32]
    lib.pph_expect.restype = C.c_int
    lib.pph_profile.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _dp, _i64p]
    lib.pph_profile.restype = C.c_int
    lib.pph_debug_blocking.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.pph_debug_blocking.restype = C.c_int
    lib.pph_debug_sparse.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]
    lib.pph_debug_sparse.restype = C.c_int
    lib.pph_debug_reduction.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.pph_debug_reduction.restype = C.c_int
    lib.pph_debug_graphs.argtypes = [C.c_void_p, C.c_int32, _i64p, _i64p]
    lib.pph_debug_graphs.restype = C.c_int
    lib.pph_debug_gemm_tile.argtypes = [C.c_void_p, C.c_int32]
    lib.pph_debug_gemm_tile.restype = C.c_int
    lib.pph_debug_cluster_rr.argtypes = [C.c_void_p, C.c_int32]
    lib.pph_debug_cluster_rr.restype = C.c_int
    lib.pph_debug_vector_refinement.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.pph_debug_vector_refinement.restype = C.c_int
    lib.pph_debug_zgemm.argtypes = [C.c_char, C.c_char, C.c_char, C.c_int64, C.c_int64, C.c_int64,
                                    C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                    C.c_int64, C.c_int32, _dp]
    lib.pph_dos_triangulation.argtypes = [C.c_int64, C.c_int64, _dp, _dp, C.c_int64, _i64p,
                                          C.c_int64, _dp, _dp, _dp, C.c_int32, C.c_int32]
    lib.pph_dos_triangulation_noalloc.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p,
                                                  C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                                  C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    lib.pph_dos_smearing.argtypes = [C.c_int32, C.c_int64, C.c_int64, _dp, _dp, C.c_double,
                                     C.c_int64, C.c_int64, _dp, _dp, C.c_int32, C.c_int32]
    lib.pph_fc_symmetrize.argtypes = [C.c_int64, _dp, C.c_int64, C.c_int32, C.c_int32]
    lib.pph_spring_fc.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, _dp, C.c_uint64, C.c_double, C.c_double,
                                  C.c_double, C.c_void_p, C.c_int32, C.c_int32]
    for name in ("pph_create", "pph_run", "pph_run_device", "pph_assemble", "pph_heevx",
                 "pph_destroy", "pph_debug_zgemm", "pph_dos_triangulation",
                 "pph_dos_triangulation_noalloc", "pph_dos_smearing",
                 "pph_fc_symmetrize", "pph_spring_fc"):
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


class PphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"parphom_hip error {code}: {msg}")
        self.code = code


def _check(rc, handle=None):
    if rc != 0:
        msg = load().pph_last_error(handle)
        raise PphError(rc, msg.decode() if msg else "")


def _ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


def make_opts(jobz="N", rng="A", vel=False, vel_method="A", precision="D", vl=0.0, vu=0.0,
              il=1, iu=1, abstol=float(np.float32(1e-15)), orfac=float(np.float32(1e-15)),
              batch=0, device=0, streams=0) -> PphOpts:
    return PphOpts(jobz.encode(), rng.encode(), vel_method.encode(), precision.encode(),
                   1 if vel else 0, device, vl, vu, il, iu, abstol, orfac, batch, streams)


class Handle:
    """Device state for one structure (pph_create / pph_destroy)."""

    def __init__(self, lat, pos, types, masses, fc, **opt):
        lib = load()
        self.lat = np.asfortranarray(np.asarray(lat, np.float64))
        self.pos = np.asfortranarray(np.asarray(pos, np.float64))
        self.types = np.ascontiguousarray(np.asarray(types, np.int64))
        self.masses = np.ascontiguousarray(np.asarray(masses, np.float64))
        fc = np.ascontiguousarray(np.asarray(fc, np.float64))
        N = self.pos.shape[0]
        if fc.shape != (N, N, 3, 3):
            raise ValueError(f"fc must be (N,N,3,3) with N={N}, got {fc.shape}")
        self.N, self.n = N, 3 * N
        self.opts = make_opts(**opt)
        self._sys = PphSystem(N, len(self.masses), _ptr(self.lat), _ptr(self.pos),
                              _ptr(self.masses), _ptr(self.types, C.c_int64))
        h = C.c_void_p()
        _check(lib.pph_create(C.byref(self._sys), _ptr(fc), C.byref(self.opts), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load().pph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def mcap(self):
        o = self.opts
        if o.range == b"I":
            return max(1, min(self.n, o.iu) - max(1, o.il) + 1)
        return self.n

    def run(self, q, evec=None, vel=None):
        """Host-buffer pipeline.  Returns dict(eval_cm[list of arrays], m, nz,
        info, evec[list], vel[list])."""
        lib = load()
        q = np.asfortranarray(np.atleast_2d(np.asarray(q, np.float64)))
        nq = q.shape[0]
        n, mcap = self.n, self.mcap
        want_vec = self.opts.jobz == b"V" if evec is None else evec
        want_vel = bool(self.opts.vel) if vel is None else vel
        ev = np.zeros((nq, mcap))
        m = np.zeros(nq, np.int64)
        nz = np.zeros(nq, np.int64)
        info = np.zeros(nq, np.int32)
        Z = np.zeros((nq, mcap, n), np.complex128) if want_vec else None
        V = np.zeros((nq, mcap, 3)) if want_vel else None
        out = PphOut(_ptr(ev), mcap, _ptr(m, C.c_int64), _ptr(nz, C.c_int64),
                     Z.ctypes.data if Z is not None else None, mcap,
                     _ptr(V) if V is not None else None, _ptr(info, C.c_int32))
        _check(lib.pph_run(self._h, nq, _ptr(q), nq, C.byref(out)), self._h)
        res = {"eval_cm": [ev[i, :m[i]] for i in range(nq)], "m": m, "nz": nz, "info": info}
        res["evec"] = [Z[i, :nz[i]].T for i in range(nq)] if Z is not None else None
        res["vel"] = [V[i, :nz[i]] for i in range(nq)] if V is not None else None
        return res

    def run_device(self, q, out_tensors):
        """Device-buffer pipeline; out_tensors holds torch tensors on the GPU:
        eval_cm (nq, ld), m (nq,) int64, info (nq,) int32, optional nz, evec
        (nq, cols, n) complex128, vel (nq, cols, 3)."""
        lib = load()
        q = np.asfortranarray(np.atleast_2d(np.asarray(q, np.float64)))
        nq = q.shape[0]
        t = out_tensors
        self._check_device_outputs(nq, t)

        def dp(x, ct=C.c_double):
            return C.cast(C.c_void_p(x.data_ptr()), C.POINTER(ct)) if x is not None else None

        cols = t["evec"].shape[1] if t.get("evec") is not None else (
            t["vel"].shape[1] if t.get("vel") is not None else 0)
        out = PphOut(dp(t["eval_cm"]), t["eval_cm"].shape[1], dp(t["m"], C.c_int64),
                     dp(t.get("nz"), C.c_int64),
                     C.c_void_p(t["evec"].data_ptr()) if t.get("evec") is not None else None,
                     cols, dp(t.get("vel")), dp(t["info"], C.c_int32))
        _check(lib.pph_run_device(self._h, nq, _ptr(q), nq, C.byref(out)), self._h)

    def _check_device_outputs(self, nq, t):
        """The library writes every output at strides derived from its shape
        (eval_cm rows of ld_eval, evec and vel blocks of evec_cols modes, m /
        nz / info for all nq rows), so a tensor of the wrong dtype, layout,
        device or size would be written out of bounds: reject it here."""
        import torch
        want = {"eval_cm": torch.float64, "m": torch.int64, "nz": torch.int64,
                "info": torch.int32, "evec": torch.complex128, "vel": torch.float64}
        dev = None
        for key, dt in want.items():
            x = t.get(key)
            if x is None:
                if key in ("eval_cm", "m", "info"):
                    raise ValueError(f"run_device: output '{key}' is required")
                continue
            if not isinstance(x, torch.Tensor) or x.device.type != "cuda":
                raise ValueError(f"run_device: '{key}' must be a CUDA tensor")
            if x.dtype != dt:
                raise ValueError(f"run_device: '{key}' must be {dt}, got {x.dtype}")
            if not x.is_contiguous():
                raise ValueError(f"run_device: '{key}' must be contiguous")
            if x.device.index != self.opts.device:
                raise ValueError(f"run_device: '{key}' is on cuda:{x.device.index}, the handle "
                                 f"on device {self.opts.device}")
            if x.shape[0] < nq:
                raise ValueError(f"run_device: '{key}' has {x.shape[0]} rows for {nq} q-points")
            dev = x.device
        if t["eval_cm"].dim() != 2 or t["eval_cm"].shape[1] < self.mcap:
            raise ValueError(f"run_device: eval_cm must be (nq, >= {self.mcap})")
        ev, vel = t.get("evec"), t.get("vel")
        if ev is not None and (ev.dim() != 3 or ev.shape[2] != self.n
                               or ev.shape[1] < self.mcap):
            raise ValueError(f"run_device: evec must be (nq, >= {self.mcap}, {self.n})")
        if vel is not None:
            if vel.dim() != 3 or vel.shape[2] != 3 or vel.shape[1] < self.mcap:
                raise ValueError(f"run_device: vel must be (nq, >= {self.mcap}, 3)")
            if ev is not None and vel.shape[1] != ev.shape[1]:
                raise ValueError("run_device: vel and evec must have the same number of columns "
                                 "(both are written evec_cols apart)")
        return dev

    def blocking(self, nb_panel=32, nb_back=128):
        """Reflectors per reduction panel and per back-transformation block
        (a test hook: the defaults are the measured best)."""
        _check(load().pph_debug_blocking(self._h, nb_panel, nb_back), self._h)

    def reduction(self, mode=1, band=0):
        """Reduction to the tridiagonal on this handle: 0 one-stage, 1
        two-stage (dense -> band -> tridiagonal), 2 the one-launch reduction of
        small matrices (n <= 96 FP64, 128 FP32; one-stage above), -1
        automatic; band 16 or 32 (0 keeps the current one)."""
        _check(load().pph_debug_reduction(self._h, mode, band), self._h)

    def gemm_tile(self, tile=0):
        """Tile shape of the eigensolver's matrix-core GEMMs on this handle:
        a GEMM_TILES name or code (0: the process default)."""
        code = GEMM_TILES[tile] if isinstance(tile, str) else int(tile)
        _check(load().pph_debug_gemm_tile(self._h, code), self._h)

    def cluster_rr(self, on=True):
        """The FP32 path's FP64 Rayleigh-Ritz step over clusters of close
        eigenvalues (on by default; range A)."""
        _check(load().pph_debug_cluster_rr(self._h, 1 if on else 0), self._h)

    def vector_refinement(self, steps=2, bf16=True):
        """FP64 residual-correction steps of the FP32 path's eigenvectors
        (0, 1 or 2; 2 by default; range A); bf16: the first step's products
        on the BF16 matrix cores (the default) or the FP32 ones."""
        _check(load().pph_debug_vector_refinement(self._h, int(steps), 1 if bf16 else 0), self._h)

    def graphs(self, mode=-1):
        """Batch graphs (range A): 1 on (the default), 0 off, -1 unchanged.
        Returns (captures, replays) summed over the handle's lanes."""
        c, r = C.c_int64(0), C.c_int64(0)
        _check(load().pph_debug_graphs(self._h, mode, C.byref(c), C.byref(r)), self._h)
        return int(c.value), int(r.value)

    def sparse(self, mode=1):
        """Products with D and dD in the Rayleigh quotients and velocities:
        0 dense, 1 automatic (the default), 2 block-sparse whenever block lists
        exist.  Returns the number of lower blocks now used (0 = dense)."""
        nblk = C.c_int64(0)
        _check(load().pph_debug_sparse(self._h, mode, C.byref(nblk)), self._h)
        return int(nblk.value)

    def profile(self, enable=True, reset=False):
        """Accumulated phase times (ms) and batch counts, then set recording."""
        ms = np.zeros(len(PHASES))
        cnt = np.zeros(len(PHASES), np.int64)
        _check(load().pph_profile(self._h, 1 if enable else 0, 1 if reset else 0, _ptr(ms),
                                  _ptr(cnt, C.c_int64)), self._h)
        return {p: (float(ms[i]), int(cnt[i])) for i, p in enumerate(PHASES)}

    def assemble(self, q, what=0):
        """D(q) (what=0) or dD/dq_{x,y,z} (what=1..3) as (nq, n, n) complex128."""
        lib = load()
        q = np.asfortranarray(np.atleast_2d(np.asarray(q, np.float64)))
        nq = q.shape[0]
        out = np.zeros((nq, self.n, self.n), np.complex128)   # each matrix column-major
        _check(lib.pph_assemble(self._h, nq, _ptr(q), nq, what, out.ctypes.data, 0), self._h)
        return out.transpose(0, 2, 1)


def set_default_gemm_tile(tile="64x64"):
    """Process default of the GEMM tile shape (a GEMM_TILES name or code;
    also used by debug_zgemm and the temporary handles of heevx)."""
    code = GEMM_TILES[tile] if isinstance(tile, str) else int(tile)
    _check(load().pph_debug_gemm_tile(None, code))


def set_default_cluster_rr(on=True):
    """Process default of the FP32 path's cluster Rayleigh-Ritz step (also
    the temporary handles of heevx)."""
    _check(load().pph_debug_cluster_rr(None, 1 if on else 0))


def set_default_vector_refinement(steps=2, bf16=True):
    """Process default of the FP32 path's vector-correction steps and of
    the precision of the first step's products (also the temporary handles of
    heevx)."""
    _check(load().pph_debug_vector_refinement(None, int(steps), 1 if bf16 else 0))


def set_default_reduction(mode=-1, band=0):
    """Process default of the reduction for handles left on automatic (also
    the temporary handles of heevx): 0 one-stage, 1 two-stage, 2 the
    one-launch small-matrix reduction where the order fits, -1 the library's
    choice."""
    _check(load().pph_debug_reduction(None, mode, band))


def heevx(A, jobz="V", rng="A", vl=0.0, vu=0.0, il=1, iu=1, precision="D",
          abstol=float(np.float32(1e-15)), orfac=float(np.float32(1e-15)), batch=0, device=0):
    """pzheevx-semantics eigensolve of Hermitian matrices A (nmat, n, n) or (n, n).
    vl/vu are in lambda units here.  Returns (w list, Z list or None, info)."""
    lib = load()
    A = np.asarray(A, np.complex128)
    single = A.ndim == 2
    if single:
        A = A[None]
    nmat, n, _ = A.shape
    Af = np.ascontiguousarray(A.transpose(0, 2, 1))          # column-major per matrix
    o = make_opts(jobz=jobz, rng=rng, vl=vl, vu=vu, il=il, iu=iu, precision=precision,
                  abstol=abstol, orfac=orfac, batch=batch, device=device)
    w = np.zeros((nmat, n))
    Z = np.zeros((nmat, n, n), np.complex128) if jobz == "V" else None
    m = np.zeros(nmat, np.int64)
    info = np.zeros(nmat, np.int32)
    _check(lib.pph_heevx(n, nmat, Af.ctypes.data, C.byref(o), _ptr(w),
                         Z.ctypes.data if Z is not None else None, _ptr(m, C.c_int64),
                         _ptr(info, C.c_int32), 0))
    ws = [w[i, :m[i]] for i in range(nmat)]
    Zs = [Z[i, :m[i]].T for i in range(nmat)] if Z is not None else None
    if single:
        return ws[0], (Zs[0] if Zs else None), int(info[0])
    return ws, Zs, info


def debug_zgemm(A, B, opA="N", opB="N", prec="D", mode=0, Cin=None):
    """Run the library's MFMA complex GEMM on host arrays (test hook)."""
    lib = load()
    dt = np.complex64 if prec == "S" else np.complex128
    A = np.asfortranarray(A.astype(dt))
    B = np.asfortranarray(B.astype(dt))
    m = A.shape[0] if opA == "N" else A.shape[1]
    k = A.shape[1] if opA == "N" else A.shape[0]
    n = B.shape[1] if opB == "N" else B.shape[0]
    Cm = np.asfortranarray(np.zeros((m, n), dt) if Cin is None else Cin.astype(dt))
    dot = np.zeros(n)
    _check(lib.pph_debug_zgemm(prec.encode(), opA.encode(), opB.encode(), m, n, k, A.ctypes.data,
                               A.shape[0], B.ctypes.data, B.shape[0], Cm.ctypes.data, m, mode,
                               _ptr(dot)))
    return Cm, dot


def dos_triangulation(freq, tri, E, func=None, device=0):
    """Linear-triangulation sums (dos, nd) over all triangles and bands, the
    loops of bz_integration.f90:12-47 on the GPU.  freq, func: (nk, nbands);
    tri: (ntri, 3) 0-based; E ascending."""
    lib = load()
    freq = np.ascontiguousarray(freq, np.float64)
    nk, nb = freq.shape
    fu = None if func is None else np.ascontiguousarray(func, np.float64)
    if fu is not None and fu.shape != freq.shape:
        raise ValueError("func must have the shape of freq")
    tri = np.ascontiguousarray(tri, np.int64).reshape(-1, 3)
    E = np.ascontiguousarray(E, np.float64)
    dos = np.zeros(len(E))
    nd = np.zeros(len(E))
    _check(lib.pph_dos_triangulation(nk, nb, _ptr(freq), None if fu is None else _ptr(fu),
                                     len(tri), _ptr(tri, C.c_int64), len(E), _ptr(E), _ptr(dos),
                                     _ptr(nd), device, 0))
    return dos, nd


def dos_triangulation_noalloc(freq, tri, E, dos, nd, func=None, stream=None):
    """pph_dos_triangulation_noalloc on torch CUDA tensors (no allocation in
    the library): freq, func (nk, nbands) float64, tri (ntri, 3) int64, E
    (nE,) float64 ascending, dos and nd (nE,) float64 outputs, all contiguous
    on one GPU.  Enqueued on `stream` (a torch.cuda.Stream; None = torch's
    current stream); synchronise before reading dos / nd."""
    import torch
    for name, x, dt in (("freq", freq, torch.float64), ("tri", tri, torch.int64),
                        ("E", E, torch.float64), ("dos", dos, torch.float64),
                        ("nd", nd, torch.float64)) + ((("func", func, torch.float64),)
                                                       if func is not None else ()):
        if x.device.type != "cuda" or x.dtype != dt or not x.is_contiguous():
            raise ValueError(f"{name} must be a contiguous {dt} CUDA tensor")
    nk, nb = freq.shape
    if func is not None and func.shape != freq.shape:
        raise ValueError("func must have the shape of freq")
    if tri.dim() != 2 or tri.shape[1] != 3:
        raise ValueError("tri must be (ntri, 3)")
    nE = E.shape[0]
    if dos.shape[0] < nE or nd.shape[0] < nE:
        raise ValueError("dos and nd need nE entries")
    dev = freq.device.index
    st = torch.cuda.current_stream(freq.device) if stream is None else stream
    _check(load().pph_dos_triangulation_noalloc(
        nk, nb, freq.data_ptr(), None if func is None else func.data_ptr(), tri.shape[0],
        tri.data_ptr(), nE, E.data_ptr(), dos.data_ptr(), nd.data_ptr(), dev, st.cuda_stream))


def dos_smearing(method, freq, weight, width, E, natom, device=0):
    """sum_q weight[q] / natom sum_b g(E - freq[q, b]) with a Gaussian
    (method 1 or "gaussian") or Lorentzian (2 or "lorentzian") of `width`,
    bz_integration.f90:140-199 on the GPU."""
    lib = load()
    m = {"gaussian": 1, "lorentzian": 2}.get(method, method)
    freq = np.ascontiguousarray(np.atleast_2d(freq), np.float64)
    nq, nb = freq.shape
    w = np.ascontiguousarray(np.broadcast_to(np.asarray(weight, np.float64), (nq,)))
    E = np.ascontiguousarray(E, np.float64)
    dos = np.zeros(len(E))
    _check(lib.pph_dos_smearing(int(m), nq, nb, _ptr(freq), _ptr(w), float(width), int(natom),
                                len(E), _ptr(E), _ptr(dos), device, 0))
    return dos


def spring_fc(natom, i, j, rvec, kind, params, seed, noise_amp, inter_r0, inter_decay, out=None,
              device=0):
    """Spring-model force constants on the GPU (pph_spring_fc) from a pair
    list.  Returns a (natom, natom, 3, 3) float64 array, or fills `out`, a
    contiguous float64 CUDA tensor of that shape, and returns it."""
    lib = load()
    i = np.ascontiguousarray(i, np.int64)
    j = np.ascontiguousarray(j, np.int64)
    rvec = np.ascontiguousarray(rvec, np.float64).reshape(-1, 3)
    kind = np.ascontiguousarray(kind, np.int32)
    params = np.ascontiguousarray(params, np.float64)
    npairs = len(i)
    if not (len(j) == len(kind) == len(rvec) == npairs) or params.shape != (7,):
        raise ValueError("spring_fc: pair arrays of one length and 7 parameters")
    args = (natom, npairs, i.ctypes.data, j.ctypes.data, rvec.ctypes.data, kind.ctypes.data,
            _ptr(params), int(seed) & (2 ** 64 - 1), float(noise_amp), float(inter_r0),
            float(inter_decay))
    if out is None:
        fc = np.zeros((natom, natom, 3, 3))
        _check(lib.pph_spring_fc(*args, fc.ctypes.data, device, 0))
        return fc
    import torch
    if (not isinstance(out, torch.Tensor) or out.dtype != torch.float64 or not out.is_cuda
            or not out.is_contiguous() or tuple(out.shape) != (natom, natom, 3, 3)):
        raise ValueError("spring_fc: out must be a contiguous float64 CUDA tensor (N, N, 3, 3)")
    # device pointers are not range-checked by the library: check the pairs here
    if npairs and (i.min() < 0 or j.min() < 0 or max(i.max(), j.max()) >= natom or
                   (i == j).any() or kind.min() < 0 or kind.max() > 3):
        raise ValueError("spring_fc: pair or kind out of range")
    dev = out.device
    ti, tj = torch.from_numpy(i).to(dev), torch.from_numpy(j).to(dev)
    tr, tk = torch.from_numpy(rvec).to(dev), torch.from_numpy(kind).to(dev)
    dargs = args[:2] + (ti.data_ptr(), tj.data_ptr(), tr.data_ptr(), tk.data_ptr()) + args[6:]
    _check(lib.pph_spring_fc(*dargs, out.data_ptr(), dev.index or 0, 1))
    return out


def fc_symmetrize(fc, iterations, device=0):
    """symmetrize_fc's iterations (acoustic sum rule, then inversion symmetry)
    on a (N, N, 3, 3) float64 array on the GPU; returns the new array."""
    out = np.array(fc, dtype=np.float64, order="C", copy=True)
    if out.ndim != 4 or out.shape[0] != out.shape[1] or out.shape[2:] != (3, 3):
        raise ValueError("fc must have shape (N, N, 3, 3)")
    _check(load().pph_fc_symmetrize(out.shape[0], _ptr(out), int(iterations), device, 0))
    return out
