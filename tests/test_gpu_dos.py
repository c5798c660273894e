"""DOS on the GPU (marked gpu): pph_dos_triangulation / pph_dos_smearing
against the oracle (oracle/dos.py, itself pinned to the reference's compiled
bz_integration.f90), and the pyphutil-compatible driver end to end."""

import numpy as np
import pytest

from oracle import dos as OD
from tests.test_dos_oracle import random_case

pytestmark = pytest.mark.gpu


def _lib():
    from parphom_amd import lib
    lib.load()
    return lib


def _close(a, b, rel):
    return np.abs(a - b).max() <= rel * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_triangulation_matches_oracle(gpu_available, seed):
    lib = _lib()
    tri, E, freq, func = random_case(seed, nk=30, nb=9, ntri=50)
    d0, n0 = OD.linear_triangulation(tri, E, freq, func)
    d1, n1 = lib.dos_triangulation(freq, tri, E, func=func)
    assert _close(d1, d0, 1e-12) and _close(n1, n0, 1e-12)


def test_triangulation_default_func(gpu_available):
    lib = _lib()
    tri, E, freq, _ = random_case(7)
    d0, n0 = OD.linear_triangulation(tri, E, freq, np.ones_like(freq))
    d1, n1 = lib.dos_triangulation(freq, tri, E)
    assert _close(d1, d0, 1e-12) and _close(n1, n0, 1e-12)


@pytest.mark.parametrize("method", ["gaussian", "lorentzian"])
def test_smearing_matches_oracle(gpu_available, method):
    lib = _lib()
    rng = np.random.default_rng(11)
    nq, nb = 7, 300
    freq = np.sort(rng.uniform(0, 1600, (nq, nb)), axis=1)
    w = rng.uniform(0.5, 2.0, nq)
    E = np.arange(-10, 1650, 0.5)
    m = 1 if method == "gaussian" else 2
    ref = sum(OD.smearing(m, w[q], freq[q], 2.5, E, 12) for q in range(nq))
    got = lib.dos_smearing(method, freq, w, 2.5, E, 12)
    assert _close(got, ref, 1e-12)


def test_density_of_states_end_to_end(gpu_available, tmp_path):
    """TBG (1,2) on a 6 x 6 grid: frequencies from the HIP path, written in
    parphom.x's HDF5 layout, then the pyphutil-style driver (linear
    triangulation with |v|^2 weights, and Gaussian smearing) against the
    oracle on the same inputs."""
    lib = _lib()
    from parphom_amd import dos, h5, synth
    st = synth.tbg_structure(1, 2)
    fc = synth.force_constants(st, 11)
    n = 3 * st.natom
    mesh = (6, 6, 1)
    g = np.array([[i / 6, j / 6, 0.0] for i in range(6) for j in range(6)])
    with lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True) as hd:
        res = hd.run(g)
    path = str(tmp_path / "bands.hdf5")
    with h5.Output(path) as out:
        for i in range(len(g)):
            out.write(i + 1, g[i], res["eval_cm"][i], vel=res["vel"][i])
    grid_file = tmp_path / "full_grid"
    with open(grid_file, "w") as f:
        for p in g:
            f.write("%.10f %.10f %.10f\n" % tuple(p))
        f.write(" ".join(str(i + 1) for i in range(len(g))) + "\n")
    freq = np.array(res["eval_cm"])
    v2 = np.array([np.sum(np.asarray(v) ** 2, axis=1) for v in res["vel"]])
    E = np.arange(-5.0, 1700.0, 0.5)

    Eo, D, ND = dos.density_of_states(st.lat, st.natom, mesh, path, n, (-5.0, 1700.0), 0.5,
                                      method="linear_triangulation", full_q_grid_file=str(grid_file),
                                      output_file=str(tmp_path / "tdos.dat"), func=v2)
    grid, mapping = dos.read_full_grid(mesh, str(grid_file))
    tri = dos.triangles(st.lat, mesh, grid, mapping)
    d0, n0 = OD.linear_triangulation(tri, E, freq, v2)
    fd, fn = OD.lt_normalisation(st.lat, len(tri), st.natom, n)
    assert np.allclose(Eo, E)
    assert _close(D, d0 * fd, 1e-12) and _close(ND, n0 * fn, 1e-12)
    rows = np.loadtxt(tmp_path / "tdos.dat")
    assert rows.shape == (len(E), 3)

    wfile = tmp_path / "weights"
    with open(wfile, "w") as f:
        for p in g:
            f.write("%.6f %.6f %.6f %.17g\n" % (p[0], p[1], p[2], 1.0 / len(g)))
    _, DG = dos.density_of_states(st.lat, st.natom, mesh, path, n, (-5.0, 1700.0), 0.5,
                                  method="gaussian", q_grid_file=str(wfile), width=3.0,
                                  output_file=str(tmp_path / "gdos.dat"))
    ref = sum(OD.smearing(1, 1.0 / len(g), freq[q], 3.0, E, st.natom) for q in range(len(g)))
    assert _close(DG, ref, 1e-12)
    # the triangle integral of the number density reaches 3 natom x (sum of weights) at the top
    _, Dn, NDn = dos.density_of_states(st.lat, st.natom, mesh, path, n, (-5.0, 1700.0), 0.5,
                                       method="linear_triangulation",
                                       full_q_grid_file=str(grid_file),
                                       output_file=str(tmp_path / "dos.dat"))
    P, _ = dos.pyphutil_lattice(st.lat)
    area = abs(np.linalg.det(P[0:2, 0:2])) * 0.01
    assert np.isclose(NDn[-1] * area, 3 * st.natom, rtol=1e-12)
