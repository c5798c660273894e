"""The Fortran host (parphom_amd/fortran) on the CPU: it compiles with
amdflang, its bind(C) types have the layout of the C structs of
include/parphom_hip.h (checked field by field against the ctypes mirror), and
it reads an inp.dat, the LAMMPS data file, the q file and the HDF5 force
constants before the first library call that needs a GPU."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from parphom_amd import lib, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "parphom_amd", "fortran")
AMDFLANG = os.environ.get("AMDFLANG", "/opt/rocm/bin/amdflang")

pytestmark = pytest.mark.skipif(not os.path.exists(AMDFLANG), reason="amdflang not installed")

LAYOUT_PROGRAM = r"""
program layout
  use iso_c_binding
  use parphom_hip
  implicit none
  type(pph_system), target :: s
  type(pph_opts), target :: o
  type(pph_out), target :: r
  write(*, '(a,i0)') 'pph_system ', c_sizeof(s)
  write(*, '(a,i0)') 'pph_opts ', c_sizeof(o)
  write(*, '(a,i0)') 'pph_out ', c_sizeof(r)
  write(*, '(a,i0)') 'pph_system.lat ', off(c_loc(s%lat), c_loc(s))
  write(*, '(a,i0)') 'pph_system.type ', off(c_loc(s%type), c_loc(s))
  write(*, '(a,i0)') 'pph_opts.vel ', off(c_loc(o%vel), c_loc(o))
  write(*, '(a,i0)') 'pph_opts.vl ', off(c_loc(o%vl), c_loc(o))
  write(*, '(a,i0)') 'pph_opts.il ', off(c_loc(o%il), c_loc(o))
  write(*, '(a,i0)') 'pph_opts.abstol ', off(c_loc(o%abstol), c_loc(o))
  write(*, '(a,i0)') 'pph_opts.batch ', off(c_loc(o%batch), c_loc(o))
  write(*, '(a,i0)') 'pph_opts.streams ', off(c_loc(o%streams), c_loc(o))
  write(*, '(a,i0)') 'pph_out.ld_eval ', off(c_loc(r%ld_eval), c_loc(r))
  write(*, '(a,i0)') 'pph_out.evec_cols ', off(c_loc(r%evec_cols), c_loc(r))
  write(*, '(a,i0)') 'pph_out.info ', off(c_loc(r%info), c_loc(r))
contains
  integer function off(a, b)
    type(c_ptr), intent(in) :: a, b
    off = int(transfer(a, 0_c_intptr_t) - transfer(b, 0_c_intptr_t))
  end function
end program
"""


def test_bind_c_types_match_the_c_structs(tmp_path):
    src = tmp_path / "layout.f90"
    src.write_text(LAYOUT_PROGRAM)
    exe = tmp_path / "layout"
    subprocess.run([AMDFLANG, "-module-dir", str(tmp_path), "-o", str(exe),
                    os.path.join(FDIR, "parphom_hip.f90"), str(src)], check=True,
                   capture_output=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                      text=True).stdout.splitlines())
    got = {k: int(v) for k, v in got.items()}
    types = {"pph_system": lib.PphSystem, "pph_opts": lib.PphOpts, "pph_out": lib.PphOut}
    for key, val in got.items():
        if "." in key:
            t, f = key.split(".")
            assert val == getattr(types[t], f).offset, key
        else:
            assert val == C.sizeof(types[key]), key


def test_host_program_builds_and_reads_its_inputs(tmp_path):
    """make -C parphom_amd/fortran into a scratch copy, then the program on
    the (1,2) cell's inputs: without a GPU it must get through every reader
    (the progress line reports what it read) and stop with pph_create's
    error and exit status 1, not crash."""
    if not os.path.exists(os.path.join(ROOT, "parphom_amd", "libparphom_hip.so")) or \
            not os.path.exists(os.path.join(ROOT, "parphom_amd", "libparphom_h5.so")):
        pytest.skip("libraries not built")
    work = tmp_path / "fortran"
    shutil.copytree(FDIR, work, ignore=shutil.ignore_patterns("parphom_hip.x"))
    # the Makefile links ../lib*.so: give the copy the built libraries
    for so in ("libparphom_hip.so", "libparphom_h5.so"):
        os.symlink(os.path.join(ROOT, "parphom_amd", so), tmp_path / so)
    subprocess.run(["make", "-C", str(work), f"MODDIR={tmp_path / 'mod'}"], check=True,
                   capture_output=True)
    exe = work / "parphom_hip.x"
    assert exe.exists()

    from parphom_amd import h5
    st = synth.tbg_structure(1, 2)
    fc = synth.force_constants(st, 11)
    synth.write_lammps_data(str(tmp_path / "lammps.dat"), st, "molecular")
    synth.write_q_file(str(tmp_path / "q.dat"), np.array([[0.0, 0.0, 0.0], [0.25, 0.5, 0.0]]))
    h5.write_fc(str(tmp_path / "FC.h5"), "force_constants", fc)
    (tmp_path / "inp.dat").write_text(f"""
lammps file location : {tmp_path}
lammps file name : lammps.dat
natom : {st.natom}
atom types : 1
atom style : molecular
q file location : {tmp_path}
q_file_name : q.dat
nqpt : 2
force constant file location : {tmp_path}
force constant file name : FC.h5
compute eigvecs : yes
group velocity : maybe
range : A
precision : S
unknown key : 1
""")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="", ROCR_VISIBLE_DEVICES="")
    r = subprocess.run([str(exe), str(tmp_path / "inp.dat")], capture_output=True, text=True,
                       cwd=tmp_path, env=env, timeout=120)
    out = r.stdout
    assert "Could not interpret command maybe" in out
    assert "Command UNKNOWN KEY not recognised" in out
    assert f"Inputs read: natom={st.natom} nqpt=2 range=A precision=S" in out
    assert "pph_create" in out and r.returncode == 1, out
