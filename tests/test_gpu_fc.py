"""Force-constant symmetrisation on the GPU (marked gpu) against the oracle."""
import numpy as np
import pytest

from oracle import fc as OF

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,it", [(1, 1), (16, 1), (37, 3), (70, 5)])
def test_symmetrize_matches_oracle(gpu_available, n, it):
    from parphom_amd import lib
    fc = np.random.default_rng(n).standard_normal((n, n, 3, 3))
    got = lib.fc_symmetrize(fc, it)
    ref = OF.symmetrize(fc, it)
    assert np.abs(got - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())
    assert np.array_equal(got, got.transpose(1, 0, 3, 2))


def test_symmetrize_file_in_place(gpu_available, tmp_path):
    from parphom_amd import fc as F, h5
    p = str(tmp_path / "FORCE_CONSTANTS")
    raw = np.random.default_rng(9).standard_normal((23, 23, 3, 3))
    h5.write_fc(p, "force_constants", raw)
    out = F.symmetrize_fc(p, 23, 4)
    back = h5.read_fc(p, "force_constants", 23)
    assert np.array_equal(back, out)
    assert np.abs(back - OF.symmetrize(raw, 4)).max() <= 1e-14 * np.abs(back).max()
