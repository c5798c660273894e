"""Writes the small fixtures under tests/golden/ (run in the development
container, where the reference checkout is readable):

  tblg_21.8_input.npz   geometry of Phonon_Calculations/Examples/tblg_21.8/lammps.dat
                        and the shipped force constants
                        Force_Constant_Generation/Examples/tblg_21.8/FORCE_CONSTANTS
                        (dataset `force_constants`, (28,28,3,3) float64), plus the
                        q points of q_pt_symm.dat (K, K', K+G)
  phbands_layout.json   group/dataset names, shapes and dtypes of the shipped output
                        Examples/tblg_21.8/phbands_21.8.hdf5 (layout only: its values
                        were not produced from the shipped force constants)
  tblg_21.8_FORCE_CONSTANTS.h5  a byte copy of that HDF5 file (reader test)
  tblg_21.8_oracle.npz  FP64 oracle frequencies at Gamma, K, M from the shipped inputs
  tbg_1_2_oracle.npz    synthetic (1,2) cell: oracle frequencies, velocities and D(q)
                        at a few q, for regression checks

The HDF5 files are read with the h5py of /opt/conda/bin/python3.9 (the only
interpreter here that has it); everything else runs in this interpreter.
Usage:  python tests/golden/make_golden.py
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

REF = "/root/reference"
EX = os.path.join(REF, "Phonon_Calculations/Examples/tblg_21.8")
FCF = os.path.join(REF, "Force_Constant_Generation/Examples/tblg_21.8/FORCE_CONSTANTS")
H5PY = "/opt/conda/bin/python3.9"


def h5_dump(script: str):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.npz")
        subprocess.run([H5PY, "-c", script.replace("@OUT@", out)], check=True)
        with np.load(out, allow_pickle=False) as z:
            return {k: z[k] for k in z.files}


def main():
    from parphom_amd import formats, synth
    from oracle import phonon as O

    fc = h5_dump(f"""
import h5py, numpy as np
f = h5py.File({FCF!r}, 'r')
d = f['force_constants']
np.savez('@OUT@', fc=d[()], chunks=np.array(d.chunks is None))
""")["fc"]
    ld = formats.read_lammps_data(os.path.join(EX, "lammps.dat"), 28, 1, "molecular")
    qsym = formats.read_q_file(os.path.join(EX, "q_pt_symm.dat"), 3)
    np.savez_compressed(os.path.join(HERE, "tblg_21.8_input.npz"), lat=ld.lat, pos=ld.pos,
                        types=ld.types, masses=ld.masses, fc=fc, q_symm=qsym)
    # the shipped HDF5 file itself (58 KB), to test the HDF5 reader on a file
    # the reference's producer wrote
    shutil.copyfile(FCF, os.path.join(HERE, "tblg_21.8_FORCE_CONSTANTS.h5"))

    lay = h5_dump(f"""
import h5py, numpy as np, json
g = h5py.File({os.path.join(EX, 'phbands_21.8.hdf5')!r}, 'r')
keys = sorted(g.keys())
ds = {{k: [list(g[keys[0]][k].shape), str(g[keys[0]][k].dtype)] for k in g[keys[0]].keys()}}
np.savez('@OUT@', keys=np.array(keys), ds=np.array(json.dumps(ds)),
         q0=g[keys[0]]['q_vec'][()], q1=g[keys[1]]['q_vec'][()])
""")
    layout = {"source": "Phonon_Calculations/Examples/tblg_21.8/phbands_21.8.hdf5",
              "ngroups": int(len(lay["keys"])), "first": str(lay["keys"][0]),
              "last": str(lay["keys"][-1]), "datasets": json.loads(str(lay["ds"])),
              "q_vec_0": lay["q0"].tolist(), "q_vec_1": lay["q1"].tolist()}
    with open(os.path.join(HERE, "phbands_layout.json"), "w") as f:
        json.dump(layout, f, indent=1)

    S = O.System(ld.lat, ld.pos, ld.types, ld.masses, fc)
    qs = np.array([synth.GAMMA, synth.K_POINT, synth.M_POINT])
    om = np.array([S.solve(q, jobz="N")["eval_cm"] for q in qs])
    np.savez_compressed(os.path.join(HERE, "tblg_21.8_oracle.npz"), q=qs, eval_cm=om)

    st = synth.tbg_structure(1, 2)
    fcs = synth.force_constants(st, 11)
    S2 = O.System(st.lat, st.pos, st.types, st.masses, fcs)
    q2 = np.array([[0.0, 0.0, 0.0], [2 / 3, 1 / 3, 0.0], [0.21, 0.37, 0.0], [0.5, 0.25, 0.0]])
    ev, vel, D = [], [], []
    for q in q2:
        r = S2.solve(q, vel=True)
        ev.append(r["eval_cm"])
        vel.append(r["vel"])
        D.append(S2.D(q))
    np.savez_compressed(os.path.join(HERE, "tbg_1_2_oracle.npz"), q=q2, eval_cm=np.array(ev),
                        vel=np.array(vel), D=np.array(D), fc=fcs, lat=st.lat, pos=st.pos)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
