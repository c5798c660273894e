"""The force-constant producer's steps on the GPU.

* Spring-model force constants (`spring_force_constants`): the synthetic
  producer that stands in for the reference's finite-displacement LAMMPS
  producer (Force_Constant_Generation/get_force_constant.py:227-295), with
  d fc / d(parameter) from the same kernels (the model is linear in its
  parameters).  The pair list comes from parphom_amd.synth (minimum image,
  neighbour shells); the blocks, the seeded noise and the sum-rule self
  blocks are formed in the HIP library (pph_spring_fc), directly into device
  memory when a CUDA tensor is given.
* Symmetrisation (`symmetrize_fc`), mirroring symmetrize_fc of the same
  producer (get_force_constant.py:298-348): the HDF5 force-constant dataset
  is read, `no_of_iterations` rounds of the acoustic sum rule and inversion
  symmetry run in the library (pph_fc_symmetrize), and the dataset is
  overwritten in place.

    python -m parphom_amd.fc FORCE_CONSTANTS NATOM ITERATIONS [--dataset force_constants]
    python -m parphom_amd.fc --spring M N SEED OUT.h5 [--dataset force_constants]
"""
from __future__ import annotations

import argparse

import numpy as np

from . import h5, lib, synth


def spring_force_constants(st, seed, params=None, out=None, device=0):
    """The spring-model (N, N, 3, 3) force constants of structure `st` on the
    GPU: the values of synth.force_constants(st, seed, params) (to the
    rounding of exp in the interlayer springs).  params = e_k gives
    d fc / d params[k].  Returns a numpy array, or fills and returns `out`
    (a float64 CUDA tensor)."""
    i, j, r, kind = synth.spring_pairs(st)
    p = synth.default_params() if params is None else np.asarray(params, np.float64)
    return lib.spring_fc(st.natom, i, j, r, kind, p, seed, synth.NOISE, synth.INTERLAYER,
                         synth.INTER_DECAY, out=out, device=device)


def spring_param_derivatives(st, seed, device=0):
    """d fc / d params[k] for every parameter k: (7, N, N, 3, 3)."""
    return np.stack([spring_force_constants(st, seed, params=np.eye(7)[k], device=device)
                     for k in range(7)])


def symmetrize_fc(fc_file, natom, no_of_iterations, dataset="force_constants", device=0):
    fc = h5.read_fc(fc_file, dataset, natom)
    out = lib.fc_symmetrize(fc, no_of_iterations, device=device)
    h5.update_fc(fc_file, dataset, out)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--spring", nargs=4, metavar=("M", "N", "SEED", "OUT"),
                    help="write the spring-model force constants of the (M, N) twisted "
                         "bilayer to OUT (HDF5) instead of symmetrising a file")
    ap.add_argument("fc_file", nargs="?")
    ap.add_argument("natom", type=int, nargs="?")
    ap.add_argument("iterations", type=int, nargs="?")
    ap.add_argument("--dataset", default="force_constants")
    a = ap.parse_args(argv)
    if a.spring:
        m, n, seed, path = int(a.spring[0]), int(a.spring[1]), int(a.spring[2]), a.spring[3]
        st = synth.tbg_structure(m, n)
        h5.write_fc(path, a.dataset, spring_force_constants(st, seed))
        print(f"spring-model force constants of the ({m},{n}) cell ({st.natom} atoms) "
              f"written to {path}")
        return
    if a.fc_file is None or a.natom is None or a.iterations is None:
        ap.error("FC_FILE NATOM ITERATIONS are required without --spring")
    symmetrize_fc(a.fc_file, a.natom, a.iterations, a.dataset)
    print(f"{a.iterations} symmetrisation iterations written to {a.fc_file}")


if __name__ == "__main__":
    main()
