"""CPU oracle for the phonon density of states — TEST INFRASTRUCTURE ONLY.

Restates in numpy FP64 the Brillouin-zone integrals of the reference's
utility layer: linear triangulation (bz_integration.f90:12-138) and
Gaussian / Lorentzian smearing (bz_integration.f90:159-199), with the
normalisation of moire_phonon_utils.density_of_states
(Utilities/pyphutil/pyphutil.py:212-336).  Only tests/ and the cpu_baseline
leg of bench.py may use it; the product path is the HIP library.

Pins (tests/test_dos_oracle.py):
  integrate_triangle  against the reference routine itself, compiled from
                      Utilities/pyphutil/fortran_routines/bz_integration.f90
                      by oracle/Makefile into oracle/_ref (when amdflang is
                      present), on random data with energies placed exactly
                      on vertex frequencies; for f = 1 the number density is
                      the area fraction of the triangle below E (closed form)
                      and dos is its E-derivative (finite difference).
  smearing            against the compiled reference; each Gaussian /
                      Lorentzian integrates to weight / natom per band.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

REF_SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libbzi_ref.so")

# bz_integration.f90:183,192: `double precision, parameter :: PI = 3.141592653589793`
# is a default-real literal, so PI holds the float32 value of pi
PI32 = float(np.float32(3.141592653589793))


def sort_freq(a):
    """The 4-element exchange sort of bz_integration.f90:49-64 on the last two
    axes of a (..., 4, 2) array (column 0 = energy, 1 = function value): for
    i < j, rows i and j are swapped when a[i, 0] > a[j, 0].  Equal energies
    keep the order this sequence of exchanges leaves them in."""
    a = a.copy()
    for i in range(4):
        for j in range(i + 1, 4):
            sw = a[..., i, 0] > a[..., j, 0]
            ti = a[..., i, :].copy()
            a[..., i, :] = np.where(sw[..., None], a[..., j, :], a[..., i, :])
            a[..., j, :] = np.where(sw[..., None], ti, a[..., j, :])
    return a


def integrate_triangle(t, Ef):
    """(dos, nd) of bz_integration.f90:67-134 for sorted (..., 4, 2) arrays t
    and the energy Ef (one of the four energies).  The four position tests are
    applied in the reference's order, each overwriting the previous result."""
    e = t[..., 0]
    f = t[..., 1]
    dos = np.zeros(e.shape[:-1])
    nd = np.zeros(e.shape[:-1])
    with np.errstate(divide="ignore", invalid="ignore"):
        # Ef is the largest: the whole triangle lies below
        m = e[..., 3] == Ef
        dos = np.where(m, 0.0, dos)
        nd = np.where(m, (f[..., 0] + f[..., 1]) + f[..., 2], nd)
        # Ef third: vertices 0 and 1 below, the last above
        m = e[..., 2] == Ef
        e0, f0, e1, f1, e2, f2 = e[..., 0], f[..., 0], e[..., 1], f[..., 1], e[..., 3], f[..., 3]
        del02 = (e2 - Ef) / (e2 - e0)
        del12 = (e2 - Ef) / (e2 - e1)
        ddel02 = 1 / (e0 - e2)
        ddel12 = 1 / (e1 - e2)
        nd3 = ((1 - del12 * del02 ** 2) * f0 + (1 - del02 * del12 ** 2) * f1
               + (1 - 3 * del12 * del02 + del12 * del02 ** 2 + del02 * del12 ** 2) * f2)
        dos3 = ((-2 * del02 * ddel02 * del12 - ddel12 * del02 ** 2) * f0
                + (-ddel02 * del12 ** 2 - 2 * del02 * del12 * ddel12) * f1
                + (-3 * del12 * ddel02 - 3 * del02 * ddel12 + 2 * del02 * del12 * ddel02
                   + ddel12 * del02 ** 2 + 2 * del02 * del12 * ddel12 + ddel02 * del12 ** 2) * f2)
        nd = np.where(m, nd3, nd)
        dos = np.where(m, dos3, dos)
        # Ef second: vertex 0 below, the other two above
        m = e[..., 1] == Ef
        e0, f0, e1, f1, e2, f2 = e[..., 0], f[..., 0], e[..., 2], f[..., 2], e[..., 3], f[..., 3]
        del10 = (Ef - e0) / (e1 - e0)
        del20 = (Ef - e0) / (e2 - e0)
        ddel10 = 1 / (e1 - e0)
        ddel20 = 1 / (e2 - e0)
        ndp = (3 - del10 - del20) * f0 + del10 * f1 + del20 * f2
        dos2 = (ddel10 * del20 * ndp + del10 * ddel20 * ndp
                + del10 * del20 * ((-ddel10 - ddel20) * f0 + ddel10 * f1 + ddel20 * f2))
        nd = np.where(m, ndp * del10 * del20, nd)
        dos = np.where(m, dos2, dos)
        # Ef the smallest: nothing below
        m = e[..., 0] == Ef
        nd = np.where(m, 0.0, nd)
        dos = np.where(m, 0.0, dos)
    return dos, nd


def linear_triangulation(tri, E, freq, func):
    """bz_integration.f90:12-47: for every energy, the sum over triangles and
    bands of integrate_triangle on [E, (f, func) at the three vertices].
    tri (ntri, 3) 0-based indices into the rows of freq / func (nk, nbands)."""
    tri = np.asarray(tri)
    fv = freq[tri]                     # (ntri, 3, nbands)
    gv = func[tri]
    ntri, _, nb = fv.shape
    base = np.zeros((ntri, nb, 4, 2))
    base[:, :, 1:, 0] = fv.transpose(0, 2, 1)
    base[:, :, 1:, 1] = gv.transpose(0, 2, 1)
    dos = np.zeros(len(E))
    nd = np.zeros(len(E))
    for i, Ef in enumerate(E):
        a = base.copy()
        a[:, :, 0, 0] = Ef
        d, n = integrate_triangle(sort_freq(a), Ef)
        dos[i] = d.sum()
        nd[i] = n.sum()
    return dos, nd


def smearing(method, weight, freq, width, E, natom):
    """bz_integration.f90:159-199 for one q-point: sum over bands of a Gaussian
    (method 1) or Lorentzian (method 2) of width `width`, times weight / natom."""
    x = np.asarray(E)[:, None] - np.asarray(freq)[None, :]
    if method == 1:
        d = np.exp(-((x / width) ** 2) / 2) / (np.sqrt(2 * PI32) * width)
    elif method == 2:
        d = ((width / 2) / (x ** 2 + (width / 2) ** 2)) * (1 / PI32)
    else:
        raise ValueError("method must be 1 (Gaussian) or 2 (Lorentzian)")
    return d.sum(axis=1) * weight / natom


def lt_normalisation(lat, ntri, natom, nbands):
    """The factors of pyphutil.py:243-244,301-302: DOS and ND are divided by
    3 ntri and by the moire area in nm^2; ND also gets 3 natom / nbands."""
    area = abs(np.linalg.det(np.asarray(lat)[0:2, 0:2])) * 0.01
    fd = (1 / ntri / 3) / area
    return fd, fd * 3 * natom / nbands


# ---------------------------------------------------------------- the compiled reference
def ref_available():
    return os.path.exists(REF_SO)


def _ref():
    """oracle/_ref/libbzi_ref.so: the reference's bz_integration.f90 built by
    oracle/Makefile with the C entry points of oracle/bzi_ref_wrap.f90."""
    lib = C.CDLL(REF_SO)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    lib.ref_linear_triangulation.argtypes = [C.c_int, C.c_int, C.c_int, ip, C.c_int, dp, dp, dp,
                                             dp, dp]
    lib.ref_standard_integration.argtypes = [C.c_int, C.c_double, C.c_int, dp, C.c_double, C.c_int,
                                             dp, C.c_int, dp]
    return lib


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def ref_linear_triangulation(tri, E, freq, func):
    """The reference's linear_triangulation on (tri 0-based, E, freq, func)."""
    lib = _ref()
    nk, nb = freq.shape
    t1 = np.asfortranarray(np.asarray(tri, np.int32) + 1)
    fr = np.asfortranarray(freq)
    fu = np.asfortranarray(func)
    E = np.ascontiguousarray(E, np.float64)
    dos = np.zeros(len(E))
    nd = np.zeros(len(E))
    lib.ref_linear_triangulation(nk, nb, len(tri), _p(t1, C.c_int), len(E), _p(E), _p(fr), _p(fu),
                                 _p(dos), _p(nd))
    return dos, nd


def ref_smearing(method, weight, freq, width, E, natom):
    """The reference's standard_integration for one q-point."""
    lib = _ref()
    fr = np.ascontiguousarray(freq, np.float64)
    E = np.ascontiguousarray(E, np.float64)
    out = np.zeros(len(E))
    lib.ref_standard_integration(method, weight, len(fr), _p(fr), width, len(E), _p(E), natom,
                                 _p(out))
    return out
