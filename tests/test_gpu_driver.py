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


@pytest.mark.parametrize("prec", ["D", "S"])
def test_fortran_host_end_to_end(gpu_available, tmp_path, prec):
    """parphom_hip.x (amdflang, ISO_C_BINDING, parphom_amd/fortran) on the (1,2)
    cell: the HDF5 groups against the oracle (frequencies, velocities through
    gauge-invariant cluster sums on the FP64 path) and bit for bit against the
    Python driver's output for the same input (same library, deterministic
    kernels)."""
    import subprocess
    from parphom_amd import driver, h5
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "parphom_amd", "fortran", "parphom_hip.x")
    assert os.path.exists(exe), "parphom_hip.x is not built (make -C parphom_amd/fortran)"
    st, fc = U.structure(1, 2, 11)
    synth.write_lammps_data(str(tmp_path / "lammps.dat"), st, "molecular")
    q = np.array([[0.0, 0.0, 0.0], [2 / 3, 1 / 3, 0.0], [0.21, 0.37, 0.0], [0.5, 0.0, 0.0]])
    synth.write_q_file(str(tmp_path / "q.dat"), q)
    h5.write_fc(str(tmp_path / "FC.h5"), "force_constants", fc)
    text = f"""
lammps file location : {tmp_path}
lammps file name : lammps.dat
natom : {st.natom}
atom types : 1
atom style : molecular
q file location : {tmp_path}
q file name : q.dat
nqpt : {len(q)}
force constant file location : {tmp_path}
force constant file name : FC.h5
force constant dataset : force_constants
compute eigvecs : yes
group velocity : yes
range : A
output file location : {tmp_path}
precision : {prec}
batch : 3
print progress : yes
"""
    (tmp_path / "inp_f.dat").write_text(text + "output file name : fortran\n")
    (tmp_path / "inp_p.dat").write_text(text + "output file name : python\n")
    r = subprocess.run([exe, str(tmp_path / "inp_f.dat")], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "output written to" in r.stdout
    pf = h5.read_output(str(tmp_path / "fortran.hdf5"))
    pp = h5.read_output(driver.run(str(tmp_path / "inp_p.dat"), log=lambda *a: None))
    assert sorted(pf) == sorted(pp) == [f"{i:012d}" for i in range(1, len(q) + 1)]
    S = U.oracle_system(1, 2, 11)
    for i in range(len(q)):
        g, gp = pf[f"{i + 1:012d}"], pp[f"{i + 1:012d}"]
        for key in ("q_vec", "eigenvalues", "evec", "vel"):
            assert np.array_equal(g[key], gp[key]), (i, key)
        ref = S.solve(q[i], vel=True)
        assert np.abs(g["eigenvalues"] - ref["eval_cm"]).max() <= 1e-3
        if prec == "D":
            groups = U.clusters(ref["lam"], 1e-6)
            dD = [S.D(q[i], a) for a in (1, 2, 3)]
            U.assert_vel_close(g["vel"], ref["vel"], groups,
                               U.vel_noise(dD, ref["eval_cm"], ref["lam"], groups))
