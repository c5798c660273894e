"""HDF5 layer (CPU): the reader on the reference's own FORCE_CONSTANTS file,
and the output groups in write_output.F90's layout."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")

h5 = pytest.importorskip("parphom_amd.h5")
try:
    h5.load()
except OSError:
    pytest.skip("libparphom_h5.so not built", allow_module_level=True)


def test_read_shipped_force_constants():
    fc = h5.read_fc(os.path.join(GOLD, "tblg_21.8_FORCE_CONSTANTS.h5"), "force_constants", 28)
    ref = np.load(os.path.join(GOLD, "tblg_21.8_input.npz"))["fc"]
    assert np.array_equal(fc, ref)
    with pytest.raises(h5.H5Error):
        h5.read_fc(os.path.join(GOLD, "tblg_21.8_FORCE_CONSTANTS.h5"), "force_constants", 27)
    with pytest.raises(h5.H5Error):
        h5.read_fc(os.path.join(GOLD, "tblg_21.8_FORCE_CONSTANTS.h5"), "nope", 28)


def _write_sample(path):
    rng = np.random.default_rng(3)
    n, nz = 12, 5
    qs = rng.uniform(size=(3, 3))
    evs = [np.sort(rng.uniform(size=n)) for _ in qs]
    Zs = [rng.standard_normal((n, nz)) + 1j * rng.standard_normal((n, nz)) for _ in qs]
    Vs = [rng.standard_normal((nz, 3)) for _ in qs]
    with h5.Output(path) as out:
        for i in range(3):
            out.write(i + 1, qs[i], evs[i], evec=Zs[i], vel=Vs[i])
    return qs, evs, Zs, Vs


def test_output_layout_roundtrip(tmp_path):
    p = str(tmp_path / "out.hdf5")
    qs, evs, Zs, Vs = _write_sample(p)
    data = h5.read_output(p)
    assert sorted(data) == ["000000000001", "000000000002", "000000000003"]
    lay = json.load(open(os.path.join(GOLD, "phbands_layout.json")))
    assert len(lay["first"]) == 12 and lay["datasets"]["eigenvalues"][1] == "float64"
    for i in range(3):
        g = data[f"{i + 1:012d}"]
        assert g["q_vec"].shape == (3,) and np.array_equal(g["q_vec"], qs[i])
        assert np.array_equal(g["eigenvalues"], evs[i])
        assert g["evec"].shape == (5, 12) and np.array_equal(g["evec"], Zs[i].T)
        assert g["vel"].shape == (5, 3) and np.array_equal(g["vel"], Vs[i])


H5PY = "/opt/conda/bin/python3.9"


@pytest.mark.skipif(not os.path.exists(H5PY), reason="no h5py interpreter")
def test_output_read_by_h5py(tmp_path):
    """The way the reference's consumers read it: h5py maps the {r, i}
    compound to complex128 (pyphutil.py:421 indexes f[g]['evec'] as complex;
    check_symm.py reads eigenvalues and vel)."""
    p = str(tmp_path / "out.hdf5")
    qs, evs, Zs, Vs = _write_sample(p)
    code = f"""
import h5py, numpy as np, json
f = h5py.File({p!r}, 'r')
g = f['000000000002']
e = g['evec']
a = e[()]
print(json.dumps({{'dtype': a.dtype.str, 'shape': list(e.shape),
    'r01': float(a[0, 1].real), 'i01': float(a[0, 1].imag),
    'vel': list(g['vel'].shape), 'ev': g['eigenvalues'].dtype.str, 'q': g['q_vec'].dtype.str}}))
"""
    out = json.loads(subprocess.run([H5PY, "-c", code], check=True, capture_output=True,
                                    text=True).stdout)
    assert out["dtype"] == "<c16" and out["shape"] == [5, 12] and out["vel"] == [5, 3]
    assert out["ev"] == "<f8" and out["q"] == "<f8"
    assert out["r01"] == Zs[1][1, 0].real and out["i01"] == Zs[1][1, 0].imag
