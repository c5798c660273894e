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


@pytest.mark.parametrize("cell", [(1, 2), (3, 4), (11, 12)])
def test_spring_producer_matches_cpu_model(gpu_available, cell):
    """pph_spring_fc against the CPU statement of the spring model
    (parphom_amd/synth.py): the intralayer pair blocks bit for bit (the same
    IEEE operations in the same order), the interlayer ones and the sum-rule
    self blocks to the rounding of exp; the sum rule itself."""
    from parphom_amd import fc as F, synth
    st = synth.tbg_structure(*cell)
    seed = 4011
    ref = synth.force_constants(st, seed)
    got = F.spring_force_constants(st, seed)
    i, j, r, kind = synth.spring_pairs(st)
    intra = kind < 3
    assert np.array_equal(got[i[intra], j[intra]], ref[i[intra], j[intra]])
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    assert np.abs(got.sum(axis=1)).max() <= 1e-12 * np.abs(ref).max()
    zero = np.ones((st.natom, st.natom), bool)
    zero[i, j] = False
    zero[np.arange(st.natom), np.arange(st.natom)] = False
    assert not got[zero].any()


def test_spring_parameter_derivatives_and_device_output(gpu_available):
    """d fc / d params[k] from unit parameters against the CPU model's, the
    model's linearity (fc(p) = sum_k p_k d fc / d p_k) and the CUDA-tensor
    output against the host one, on the (5,6) cell."""
    import torch
    from parphom_amd import fc as F, synth
    st = synth.tbg_structure(5, 6)
    seed = 3005
    d = F.spring_param_derivatives(st, seed)
    basis = synth.spring_basis(st, seed)
    for k in (0, 4, 6):
        ref = synth.force_constants_param_derivative(st, seed, k, basis=basis)
        assert np.abs(d[k] - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())
    p = synth.default_params()
    full = F.spring_force_constants(st, seed, params=p)
    lin = np.tensordot(p, d, axes=1)
    assert np.abs(full - lin).max() <= 1e-12 * np.abs(full).max()
    N = st.natom
    out = torch.zeros((N, N, 3, 3), dtype=torch.float64, device="cuda:0")
    F.spring_force_constants(st, seed, params=p, out=out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), full)
