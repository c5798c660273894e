"""CPU oracle for the parphom.x phonon path — TEST INFRASTRUCTURE ONLY.

This package restates, in numpy/scipy FP64, what parphom.x computes on its
hot path: the symmetrised force-constant matrix (force_read.F90), the
geometry (lammps_read.F90), the minimum-image dynamical matrix D(q) and its
q-derivatives (dynamical_matrix.F90), the pzheevx eigenpairs with the
reference's range/abstol semantics and unit conversion (diagonalize.F90,
write_output.F90:135), and the analytic group velocity
(write_output.F90:410-446).  Each function cites the reference lines it
follows.

Only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of
`bench.py` may import it, and only as the checker: the product path is the
HIP library under `parphom_amd/`, which never calls into this package.

How each function is pinned (checks that live in tests/test_oracle_*.py):
  geometry       a_i . b_j = 2 pi delta_ij; r = lat^T crys round trip.
  symmetrise_fc  transpose quirk of force_read.F90:223-252 vs the natural
                 layout (equal after symmetrisation); Phi_sym = Phi_sym^T.
  image_sets     brute-force pure-Python loop over the 9 images (tiny cells),
                 tie cases built by hand, unique images for short-range pairs.
  assemble       Hermiticity; D(-q) = conj D(q); spectrum of D(q+G) equals
                 that of D(q) (the reference's own q_pt_symm.dat K/K'/K+G
                 check, Examples/tblg_21.8/check_symm.py); three zero modes
                 at Gamma for an ASR-exact Phi; D(0) = Phi_sym/sqrt(m m);
                 shipped tblg_21.8 FORCE_CONSTANTS reproduces the Gamma/M/K
                 features of Examples/tblg_21.8/Reference_bands.pdf.
  assemble(deriv) analytic dD/dq vs a high-order central difference of D.
  eig            scipy zheevx vs numpy eigh (a different LAPACK driver),
                 residual/orthogonality, range I/V selection on eigh output.
  velocity       v = d omega/dk checked against a finite difference of the
                 frequencies (pins the formula and the m/s unit factor).
"""
