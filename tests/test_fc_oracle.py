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


def _splitmix64_py(x):
    """splitmix64 in Python integers (mod 2^64), the published finaliser."""
    m = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def test_spring_noise_hash_pinned():
    """The spring model's counter-based noise: numpy's wrapping uint64 version
    of splitmix64 against exact integer arithmetic; the first output for
    state 0 is the published 0xE220A8397B1DCDAF; the pair noise is symmetric
    in (i, j) and in its 3x3 components, lies in [-1, 1), and changes with
    the seed."""
    from parphom_amd import synth
    xs = [0, 1, 2 ** 63 + 5, 2 ** 64 - 1, 0x123456789ABCDEF0]
    got = synth.splitmix64(np.array(xs, np.uint64))
    assert [int(v) for v in got] == [_splitmix64_py(x) for x in xs]
    assert _splitmix64_py(0) == 0xE220A8397B1DCDAF
    i = np.array([0, 5, 3, 7])
    j = np.array([5, 0, 7, 3])
    u = synth.pair_noise(11, i, j, 10)
    assert np.array_equal(u[0], u[1]) and np.array_equal(u[2], u[3])
    assert np.array_equal(u, u.transpose(0, 2, 1))
    big = synth.pair_noise(11, np.arange(2000) % 97, np.arange(2000) // 97 + 100, 300)
    assert big.min() >= -1.0 and big.max() < 1.0 and abs(big.mean()) < 0.02
    # component (0, 0) of pair {lo, hi} from the formula itself
    key = 0 * 10 + 5
    h = _splitmix64_py((11 * 0x9E3779B97F4A7C15 + key * 8 + 0) % (1 << 64))
    assert u[0, 0, 0] == 2.0 * ((h >> 11) * 2.0 ** -53) - 1.0
    assert not np.array_equal(u, synth.pair_noise(12, i, j, 10))
