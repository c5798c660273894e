"""Pins of the DOS oracle (CPU): against the reference's own bz_integration.f90
compiled into oracle/_ref (skipped when it is not built), and against closed
forms."""
import numpy as np
import pytest

from oracle import dos as OD


def _need_ref():
    if not OD.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference checkout and amdflang)")


def ref_lt(tri, E, freq, func):
    _need_ref()
    return OD.ref_linear_triangulation(tri, E, freq, func)


def ref_smear(method, weight, freq, width, E, natom):
    _need_ref()
    return OD.ref_smearing(method, weight, freq, width, E, natom)


def random_case(seed, nk=14, nb=6, ntri=20, ties=True):
    rng = np.random.default_rng(seed)
    freq = np.round(rng.uniform(0, 10, (nk, nb)), 2)       # rounded: repeated values
    func = rng.uniform(0.5, 2.0, (nk, nb))
    tri = np.array([rng.choice(nk, 3, replace=False) for _ in range(ntri)])
    E = np.arange(-0.5, 10.5, 0.05)
    if ties:
        E = np.unique(np.concatenate([E, freq.ravel()[:40]]))
    return tri, E, freq, func


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_lt_matches_compiled_reference(seed):
    tri, E, freq, func = random_case(seed)
    d0, n0 = ref_lt(tri, E, freq, func)
    d1, n1 = OD.linear_triangulation(tri, E, freq, func)
    assert np.allclose(d1, d0, rtol=1e-12, atol=1e-12 * np.abs(d0).max())
    assert np.allclose(n1, n0, rtol=1e-12, atol=1e-12 * np.abs(n0).max())


@pytest.mark.parametrize("method", [1, 2])
def test_smearing_matches_compiled_reference(method):
    rng = np.random.default_rng(5)
    freq = np.sort(rng.uniform(0, 100, 40))
    E = np.arange(-5, 105, 0.37)
    d0 = ref_smear(method, 0.7, freq, 1.3, E, 12)
    d1 = OD.smearing(method, 0.7, freq, 1.3, E, 12)
    assert np.allclose(d1, d0, rtol=1e-13, atol=1e-15)


def test_lt_area_fraction_closed_form():
    """f = 1, one band, one triangle: nd = 3 x (area fraction below E) for a
    linear band, and dos = d nd / dE."""
    e0, e1, e2 = 1.0, 2.5, 4.0
    freq = np.array([[e2], [e0], [e1]])       # vertices out of order on purpose
    func = np.ones((3, 1))
    tri = np.array([[0, 1, 2]])
    E = np.linspace(0.0, 5.0, 501)
    dos, nd = OD.linear_triangulation(tri, E, freq, func)
    frac = np.where(E <= e0, 0.0,
                    np.where(E <= e1, (E - e0) ** 2 / ((e1 - e0) * (e2 - e0)),
                             np.where(E <= e2, 1 - (e2 - E) ** 2 / ((e2 - e0) * (e2 - e1)), 1.0)))
    assert np.allclose(nd, 3 * frac, atol=1e-12)
    h = 1e-6
    inner = (E > e0 + 1e-3) & (E < e2 - 1e-3) & (np.abs(E - e1) > 1e-3)
    _, ndp = OD.linear_triangulation(tri, E[inner] + h, freq, func)
    _, ndm = OD.linear_triangulation(tri, E[inner] - h, freq, func)
    assert np.allclose(dos[inner], (ndp - ndm) / (2 * h), rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("method,tol", [(1, 1e-11), (2, 2e-3)])
def test_smearing_normalisation(method, tol):
    """Each band contributes weight / natom in total, times the reference's
    float32 pi: sqrt(pi / pi32) for the Gaussian, pi / pi32 for the
    Lorentzian (which also loses its far tails beyond the energy window)."""
    freq = np.array([40.0, 55.0, 60.0])
    E = np.arange(-2000, 2100, 0.01)
    d = OD.smearing(method, 0.5, freq, 0.8, E, 4)
    total = np.trapezoid(d, E) if hasattr(np, "trapezoid") else np.trapz(d, E)
    expect = 3 * 0.5 / 4 * (np.sqrt(np.pi / OD.PI32) if method == 1 else np.pi / OD.PI32)
    assert abs(total - expect) <= tol * expect


def test_lt_normalisation_factors():
    lat = np.array([[10.0, 0, 0], [5.0, 8.660254037844386, 0], [0, 0, 20.0]])
    fd, fn = OD.lt_normalisation(lat, ntri=50, natom=4, nbands=12)
    area = 10.0 * 8.660254037844386 * 0.01
    assert np.isclose(fd, 1 / 150 / area) and np.isclose(fn, fd * 12 / 12)
