"""Force-constant symmetrisation on the GPU, mirroring symmetrize_fc of the
reference's producer (Force_Constant_Generation/get_force_constant.py:298-348):
the HDF5 force-constant dataset is read, `no_of_iterations` rounds of the
acoustic sum rule and inversion symmetry run in the HIP library
(pph_fc_symmetrize), and the dataset is overwritten in place.

    python -m parphom_amd.fc FORCE_CONSTANTS NATOM ITERATIONS [--dataset force_constants]
"""
from __future__ import annotations

import argparse

from . import h5, lib


def symmetrize_fc(fc_file, natom, no_of_iterations, dataset="force_constants", device=0):
    fc = h5.read_fc(fc_file, dataset, natom)
    out = lib.fc_symmetrize(fc, no_of_iterations, device=device)
    h5.update_fc(fc_file, dataset, out)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("fc_file")
    ap.add_argument("natom", type=int)
    ap.add_argument("iterations", type=int)
    ap.add_argument("--dataset", default="force_constants")
    a = ap.parse_args(argv)
    symmetrize_fc(a.fc_file, a.natom, a.iterations, a.dataset)
    print(f"{a.iterations} symmetrisation iterations written to {a.fc_file}")


if __name__ == "__main__":
    main()
