"""Pins of the force-constant symmetrisation oracle (CPU)."""
import numpy as np
import pytest

from oracle import fc as OF


def _random_fc(n, seed):
    return np.random.default_rng(seed).standard_normal((n, n, 3, 3))


def test_sum_rule_zeroes_row_sums():
    fc = OF.acoustic_sum_rule(_random_fc(13, 0))
    assert np.abs(fc.sum(axis=1)).max() <= 1e-13


def test_inversion_is_exact_symmetry():
    fc = OF.inversion_symmetry(_random_fc(11, 1))
    assert np.array_equal(fc, fc.transpose(1, 0, 3, 2))


def test_fixed_point():
    """Symmetric 3x3 pair blocks B_ij shared by (i, j) and (j, i), with self
    blocks -sum_j B_ij, satisfy both rules: the iteration leaves them alone."""
    rng = np.random.default_rng(6)
    n = 9
    fc = np.zeros((n, n, 3, 3))
    for i in range(n):
        for j in range(i + 1, n):
            b = rng.standard_normal((3, 3))
            fc[i, j] = fc[j, i] = b + b.T
    for i in range(n):
        fc[i, i] = -fc[i].sum(axis=0)
    assert np.abs(OF.symmetrize(fc, 3) - fc).max() <= 1e-13 * np.abs(fc).max()


def test_residual_decreases():
    fc = _random_fc(17, 2)
    res = []
    for k in range(1, 7):
        res.append(np.abs(OF.symmetrize(fc, k).sum(axis=1)).max())
    assert all(b <= a * (1 + 1e-12) for a, b in zip(res, res[1:]))
    assert res[-1] < res[0]


@pytest.mark.parametrize("n,it", [(4, 1), (5, 3)])
def test_vectorised_equals_reference_loops(n, it):
    fc = _random_fc(n, 3 + n)
    assert np.allclose(OF.symmetrize(fc, it), OF.symmetrize_loops(fc, it), rtol=0, atol=1e-14)


def test_h5_update_roundtrip(tmp_path):
    """The in-place dataset update used by the symmetriser's file path."""
    from parphom_amd import h5
    p = str(tmp_path / "FC.h5")
    fc = _random_fc(6, 4)
    h5.write_fc(p, "force_constants", fc)
    h5.update_fc(p, "force_constants", 2 * fc)
    assert np.array_equal(h5.read_fc(p, "force_constants", 6), 2 * fc)
    with pytest.raises(h5.H5Error):
        h5.update_fc(p, "force_constants", fc[:5, :5])
