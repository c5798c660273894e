"""End to end through the parphom.x-compatible driver (marked gpu): input
file + LAMMPS data + q file + HDF5 force constants -> HDF5 output groups."""
import os

import numpy as np
import pytest

from oracle import phonon as O
from parphom_amd import synth
from tests import _util as U

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["D", "S"])
def test_driver_end_to_end(gpu_available, tmp_path, prec):
    from parphom_amd import driver, h5
    st, fc = U.structure(1, 2, 11)
    synth.write_lammps_data(str(tmp_path / "lammps.dat"), st, "molecular")
    q = np.array([[0.0, 0.0, 0.0], [2 / 3, 1 / 3, 0.0], [0.21, 0.37, 0.0]])
    synth.write_q_file(str(tmp_path / "q.dat"), q)
    h5.write_fc(str(tmp_path / "FC.h5"), "force_constants", fc)
    (tmp_path / "inp.dat").write_text(f"""
lammps file location : {tmp_path}
lammps file name : lammps.dat
natom : {st.natom}
atom types : 1
atom style : molecular
q file location : {tmp_path}
q file name : q.dat
nqpt : 3
force constant file location : {tmp_path}
force constant file name : FC.h5
force constant dataset : force_constants
compute eigvecs : yes
group velocity : yes
range : A
output file name : res
output file location : {tmp_path}
precision : {prec}
print progress : no
""")
    path = driver.run(str(tmp_path / "inp.dat"))
    assert path.endswith("res.hdf5")
    data = h5.read_output(path)
    assert sorted(data) == [f"{i:012d}" for i in (1, 2, 3)]
    S = U.oracle_system(1, 2, 11)
    for i in range(3):
        g = data[f"{i + 1:012d}"]
        ref = S.solve(q[i], vel=True)
        assert np.allclose(g["q_vec"], q[i])
        assert np.abs(g["eigenvalues"] - ref["eval_cm"]).max() <= 1e-3
        assert g["evec"].shape == (S.n, S.n) and g["vel"].shape == (S.n, 3)
        groups = U.clusters(ref["lam"], 1e-6)
        if prec == "D":
            dD = [S.D(q[i], a) for a in (1, 2, 3)]
            U.assert_vel_close(g["vel"], ref["vel"], groups, U.vel_noise(dD, ref["eval_cm"], ref["lam"], groups))
    # a second run renames the output to res_v1.hdf5 (read_input.F90:490-513)
    assert driver.run(str(tmp_path / "inp.dat")).endswith("res_v1.hdf5")
