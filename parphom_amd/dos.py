"""Phonon density of states from parphom.x output, on the GPU.

Mirrors moire_phonon_utils.density_of_states
(Utilities/pyphutil/pyphutil.py:212-359): the same arguments, input files
(the HDF5 output of parphom.x, the full-grid file, the q-grid weights file)
and output file format.  The integrals themselves (bz_integration.f90) run in
the HIP library (pph_dos_triangulation, pph_dos_smearing); the host keeps the
file handling and the triangulation of the k grid.

    python -m parphom_amd.dos lammps.dat NATOM phbands.hdf5 --mesh 18 18 1 \\
        --full-grid full_grid_18X18 --range -2 1800 --spacing 0.025
"""
from __future__ import annotations

import argparse

import numpy as np

from . import formats, h5, lib


def read_full_grid(mesh, full_grid_file):
    """pyphutil.py:200-209: the first mesh[0]*mesh[1]*mesh[2] rows are the
    reduced k-points of the full grid, the next line maps each of them to its
    irreducible point."""
    nk = int(np.prod(mesh))
    grid = np.loadtxt(full_grid_file, max_rows=nk)
    with open(full_grid_file) as f:
        lines = f.readlines()
    mapping = np.array([int(float(v)) for v in lines[nk].split()])
    return grid, mapping


def pyphutil_lattice(lat_rows):
    """pyphutil keeps the lattice vectors as columns (pyphutil.py:63-65) and
    rec_lat = inv(lat^T) 2 pi (pyphutil.py:73)."""
    P = np.asarray(lat_rows, np.float64).T
    return P, np.linalg.inv(P.T) * 2 * np.pi


def triangles(lat_rows, mesh, grid, mapping):
    """Delaunay triangles of the full grid as pyphutil.py:262-287 builds them:
    points rec_lat^T g (as pyphutil writes it) projected on the plane, vertex
    indices mapped to rows of the irreducible frequency table (sorted unique
    mapping values)."""
    from scipy.spatial import Delaunay
    _, rec = pyphutil_lattice(lat_rows)
    pts = np.array([rec.T @ g for g in grid])[:, :2]
    simp = Delaunay(pts).simplices
    up = np.unique(mapping)
    red = np.searchsorted(up, mapping)          # up.index(mapping[v]) for every v
    return red[simp]


def density_of_states(lat, natom, mesh, data_file, nbands, en_range, en_spacing=1e-3,
                      method="linear_triangulation", q_grid_file=None, width=1e-4,
                      full_q_grid_file=None, output_file="dos.dat", func=None, device=0):
    """Returns (E, DOS) or (E, DOS, ND) and writes output_file in the
    reference's format (pyphutil.py:304-308, 327-331).  lat: 3 x 3, rows are
    the lattice vectors (as parphom.x reads them)."""
    E = np.arange(en_range[0], en_range[1], en_spacing)
    data = h5.read_output(data_file)
    keys = sorted(data)
    if method == "linear_triangulation":
        P, _ = pyphutil_lattice(lat)
        moire_area = abs(np.linalg.det(P[0:2, 0:2])) * 0.01          # nm^2
        grid, mapping = read_full_grid(mesh, full_q_grid_file)
        freq = np.empty((len(np.unique(mapping)), nbands))
        if func is not None and np.shape(func) != freq.shape:
            raise ValueError("func must have shape (irreducible k-points, nbands)")
        for c, k in enumerate(keys[:len(freq)]):
            freq[c] = data[k]["eigenvalues"][:nbands]
        tri = triangles(lat, mesh, grid, mapping)
        dos, nd = lib.dos_triangulation(freq, tri, E, func=func, device=device)
        ntri = len(tri)
        dos *= (1 / ntri / 3) / moire_area
        nd *= (1 / ntri / 3) / moire_area * 3 * natom / nbands
        with open(output_file, "w+") as f:
            for i in range(len(E)):
                f.write("%f\t\t%f\t%f\n" % (E[i], dos[i], nd[i]))
        return E, dos, nd
    if method in ("gaussian", "lorentzian"):
        weights = np.loadtxt(q_grid_file, usecols=(3,))
        freq = np.array([data[k]["eigenvalues"][:nbands] for k in keys])
        w = np.array([weights[int(k) - 1] for k in keys])
        dos = lib.dos_smearing(method, freq, w, width, E, natom, device=device)
        with open(output_file, "w+") as f:
            for i in range(len(E)):
                f.write("%f\t\t%f\n" % (E[i], dos[i]))
        return E, dos
    raise ValueError("method must be linear_triangulation, gaussian or lorentzian")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("lammps_file")
    ap.add_argument("natom", type=int)
    ap.add_argument("data_file")
    ap.add_argument("--atom-style", default="atomic")
    ap.add_argument("--atom-types", type=int, default=1)
    ap.add_argument("--mesh", type=int, nargs=3, default=[1, 1, 1])
    ap.add_argument("--nbands", type=int, default=None)
    ap.add_argument("--range", type=float, nargs=2, required=True)
    ap.add_argument("--spacing", type=float, default=1e-3)
    ap.add_argument("--method", default="linear_triangulation")
    ap.add_argument("--q-grid", default=None)
    ap.add_argument("--full-grid", default=None)
    ap.add_argument("--width", type=float, default=1e-4)
    ap.add_argument("--out", default="dos.dat")
    a = ap.parse_args(argv)
    ld = formats.read_lammps_data(a.lammps_file, a.natom, a.atom_types, a.atom_style)
    density_of_states(ld.lat, a.natom, a.mesh, a.data_file, a.nbands or 3 * a.natom, a.range,
                      a.spacing, a.method, a.q_grid, a.width, a.full_grid, a.out)
    print(f"Written DOS to {a.out}")


if __name__ == "__main__":
    main()
