"""ctypes binding of libparphom_h5.so (include/parphom_h5.h): the
force-constant dataset reader and the per-q output groups in parphom.x's
layout (write_output.F90:95-464)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libparphom_h5.so")
_lib = None
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built (needs hdf5.h; make -C parphom_amd/csrc h5)")
        L = C.CDLL(LIB_PATH)
        L.pph_h5_read_fc.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, _dp]
        L.pph_h5_write_fc.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, _dp]
        L.pph_h5_update_fc.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, _dp]
        L.pph_h5_open_output.argtypes = [C.c_char_p, C.c_int32, _i64p]
        L.pph_h5_open_readonly.argtypes = [C.c_char_p, _i64p]
        L.pph_h5_write_group.argtypes = [C.c_int64, C.c_int64, _dp, _dp, C.c_int64, C.c_void_p,
                                         C.c_int64, C.c_int64, C.c_void_p]
        L.pph_h5_read_dataset.argtypes = [C.c_int64, C.c_char_p, _dp, C.c_int64, _i64p,
                                          C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.pph_h5_group_count.argtypes = [C.c_int64, _i64p]
        L.pph_h5_group_name.argtypes = [C.c_int64, C.c_int64, C.c_char_p, C.c_int64]
        L.pph_h5_close.argtypes = [C.c_int64]
        L.pph_h5_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class H5Error(RuntimeError):
    pass


def _chk(rc):
    if rc != 0:
        raise H5Error(load().pph_h5_last_error().decode())


def read_fc(path, dset, natom):
    out = np.zeros((natom, natom, 3, 3))
    _chk(load().pph_h5_read_fc(path.encode(), dset.encode(), natom, out.ctypes.data_as(_dp)))
    return out


def write_fc(path, dset, fc):
    fc = np.ascontiguousarray(fc, np.float64)
    _chk(load().pph_h5_write_fc(path.encode(), dset.encode(), fc.shape[0], fc.ctypes.data_as(_dp)))


def update_fc(path, dset, fc):
    """Overwrite an existing force-constant dataset in place."""
    fc = np.ascontiguousarray(fc, np.float64)
    _chk(load().pph_h5_update_fc(path.encode(), dset.encode(), fc.shape[0], fc.ctypes.data_as(_dp)))


class Output:
    """Per-q groups of the output file (create=True truncates)."""

    def __init__(self, path, create=True):
        fid = C.c_int64()
        _chk(load().pph_h5_open_output(path.encode(), 1 if create else 0, C.byref(fid)))
        self.fid = fid.value

    def write(self, qindex, qvec, eval_cm, evec=None, vel=None):
        """qindex 1-based; evec (n, nz) complex; vel (nz, 3)."""
        qv = np.ascontiguousarray(qvec, np.float64)
        ev = np.ascontiguousarray(eval_cm, np.float64)
        n = nz = 0
        ep = vp = None
        if evec is not None:
            Z = np.ascontiguousarray(np.asarray(evec, np.complex128).T)   # (nz, n) rows
            nz, n = Z.shape
            ep = Z.ctypes.data
        if vel is not None:
            V = np.ascontiguousarray(vel, np.float64)
            nz = V.shape[0]
            vp = V.ctypes.data
        _chk(load().pph_h5_write_group(self.fid, qindex, qv.ctypes.data_as(_dp), ev.ctypes.data_as(_dp),
                                       len(ev), ep, n, nz, vp))

    def close(self):
        if self.fid is not None:
            load().pph_h5_close(self.fid)
            self.fid = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def read_output(path):
    """{group name: {dataset: array}} of an output file (any parphom layout)."""
    L = load()
    fid = C.c_int64()
    _chk(L.pph_h5_open_readonly(path.encode(), C.byref(fid)))
    out = {}
    try:
        cnt = C.c_int64()
        _chk(L.pph_h5_group_count(fid.value, C.byref(cnt)))
        buf = C.create_string_buffer(256)
        for i in range(cnt.value):
            _chk(L.pph_h5_group_name(fid.value, i, buf, 256))
            g = buf.value.decode()
            out[g] = {}
            for ds in ("q_vec", "eigenvalues", "evec", "vel"):
                dims = (C.c_int64 * 4)()
                rank, cpx = C.c_int32(), C.c_int32()
                if L.pph_h5_read_dataset(fid.value, f"{g}/{ds}".encode(), None, 0, dims,
                                         C.byref(rank), C.byref(cpx)) != 0:
                    continue
                shape = tuple(dims[k] for k in range(rank.value))
                cnt_d = int(np.prod(shape)) * (2 if cpx.value else 1)
                arr = np.zeros(cnt_d)
                _chk(L.pph_h5_read_dataset(fid.value, f"{g}/{ds}".encode(), arr.ctypes.data_as(_dp),
                                           cnt_d, dims, C.byref(rank), C.byref(cpx)))
                out[g][ds] = arr.view(np.complex128).reshape(shape) if cpx.value else arr.reshape(shape)
    finally:
        L.pph_h5_close(fid.value)
    return out
