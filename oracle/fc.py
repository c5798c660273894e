"""CPU oracle for the force-constant symmetrisation of the reference's
producer — TEST INFRASTRUCTURE ONLY (tests/ may use it; the product path is
pph_fc_symmetrize in the HIP library).

Restates symmetrize_fc (Force_Constant_Generation/get_force_constant.py:
298-348) for one MPI rank on a (N, N, 3, 3) float64 array: each iteration is
acoustic_sum_rule (:320-331) then inversion_symmetry (:334-348).

Pins (tests/test_fc_oracle.py): after the sum rule every
sum_j fc[i, j, a, b] is zero; after the inversion step
fc[i, j, a, b] == fc[j, i, b, a] exactly; a force-constant array that already
satisfies both is a fixed point; both steps are orthogonal projections, so
the sum-rule residual of the iterates decreases monotonically; and the
vectorised form equals the reference's in-place loops (restated with their
loop order) on small arrays.
"""
import numpy as np


def acoustic_sum_rule(fc):
    """get_force_constant.py:320-331: fc[i, :, a, b] -= mean over the second atom."""
    return fc - fc.mean(axis=1, keepdims=True)


def inversion_symmetry(fc):
    """get_force_constant.py:334-348: fc[i, j, a, b] = fc[j, i, b, a] = their mean
    (every unordered pair once, from the values before the step)."""
    return (fc + fc.transpose(1, 0, 3, 2)) / 2


def symmetrize(fc, iterations):
    out = np.array(fc, dtype=np.float64, copy=True)
    for _ in range(iterations):
        out = inversion_symmetry(acoustic_sum_rule(out))
    return out


def symmetrize_loops(fc, iterations):
    """The reference's loops in their order, in place (small arrays only)."""
    data = np.array(fc, dtype=np.float64, copy=True)
    natom = data.shape[0]
    for _ in range(iterations):
        for i in range(natom):
            for alpha in range(3):
                for beta in range(3):
                    col_j = data[i, :, alpha, beta].copy()
                    col_j -= np.mean(col_j)
                    data[i, :, alpha, beta] = col_j
        for i in range(natom):
            for j in range(i, natom):
                for alpha in range(3):
                    for beta in range(3):
                        mean = (data[i, j, alpha, beta] + data[j, i, beta, alpha]) / 2
                        data[i, j, alpha, beta] = mean
                        data[j, i, beta, alpha] = mean
    return data
