"""pph_dos_triangulation_noalloc (marked gpu): the no-allocation variant of
pph_dos_triangulation, called through the C ABI on device tensors, against
the existing entry on the same device inputs, on two calls in a row (the
second smaller, so a stale difference array from the first would show), and
its size limit."""
import ctypes as C

import numpy as np
import pytest

from oracle import dos as OD
from tests.test_dos_oracle import random_case

pytestmark = pytest.mark.gpu


def _lib():
    from parphom_amd import lib
    lib.load()
    return lib


def _existing(lib, torch, freq, tri, E, func):
    """pph_dos_triangulation on the same device inputs (on_device = 1)."""
    nE = E.shape[0]
    dos = torch.zeros(nE, dtype=torch.float64, device="cuda:0")
    nd = torch.zeros(nE, dtype=torch.float64, device="cuda:0")

    def p(x, ct=C.c_double):
        return C.cast(C.c_void_p(x.data_ptr()), C.POINTER(ct))

    torch.cuda.synchronize()
    lib._check(lib.load().pph_dos_triangulation(
        freq.shape[0], freq.shape[1], p(freq), None if func is None else p(func), tri.shape[0],
        p(tri, C.c_int64), nE, p(E), p(dos), p(nd), 0, 1))
    return dos.cpu().numpy(), nd.cpu().numpy()


def _case(torch, seed, nk, nb, ntri, with_func, nE):
    tri, E, freq, func = random_case(seed, nk=nk, nb=nb, ntri=ntri)
    # nE energies, with the first 40 vertex frequencies among them (ties)
    E = np.unique(np.concatenate([np.linspace(-0.5, 10.5, nE - 40), freq.ravel()[:40]]))
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    return (tri, E, freq, func if with_func else None,
            t(freq), t(tri, torch.int64), t(E), t(func) if with_func else None)


def test_noalloc_matches_existing_entry_twice(gpu_available):
    import torch
    lib = _lib()
    stream = torch.cuda.Stream()
    for call, (seed, nk, nb, ntri, with_func, ne) in enumerate(((11, 400, 48, 700, True, 6000),
                                                                (12, 60, 9, 90, False, 300))):
        tri, E, freq, func, df, dt, dE, dfu = _case(torch, seed, nk, nb, ntri, with_func, ne)
        nE = len(E)
        print(f"no-alloc call {call}: nk={nk} nbands={nb} ntri={ntri} nE={nE} "
              f"func={'yes' if with_func else 'none'}", flush=True)
        # outputs start as garbage: the entry clears them itself
        dos = torch.full((nE,), 7.0, dtype=torch.float64, device="cuda:0")
        nd = torch.full((nE,), -3.0, dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        lib.dos_triangulation_noalloc(df, dt, dE, dos, nd, func=dfu, stream=stream)
        stream.synchronize()
        d1, n1 = dos.cpu().numpy(), nd.cpu().numpy()
        d0, n0 = _existing(lib, torch, df, dt, dE, dfu)
        scale_d, scale_n = max(1.0, np.abs(d0).max()), max(1.0, np.abs(n0).max())
        # FP64 atomics sum in a different order from run to run
        assert np.abs(d1 - d0).max() <= 1e-12 * scale_d
        assert np.abs(n1 - n0).max() <= 1e-12 * scale_n
        if call == 1:
            # the smaller case also against the oracle
            do, no = OD.linear_triangulation(tri, E, freq, np.ones_like(freq))
            assert np.abs(d1 - do).max() <= 1e-12 * max(1.0, np.abs(do).max())
            assert np.abs(n1 - no).max() <= 1e-12 * max(1.0, np.abs(no).max())


def test_noalloc_rejects_too_many_energies(gpu_available):
    import torch
    lib = _lib()
    one = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    tri = torch.zeros((1, 3), dtype=torch.int64, device="cuda:0")
    big = (1 << 20) + 1
    rc = lib.load().pph_dos_triangulation_noalloc(1, 1, one.data_ptr(), None, 1, tri.data_ptr(),
                                                 big, one.data_ptr(), one.data_ptr(),
                                                 one.data_ptr(), 0, None)
    assert rc == -1
    assert b"scratch capacity" in lib.load().pph_last_error(None)
