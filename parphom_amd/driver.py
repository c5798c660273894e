"""parphom.x-compatible driver on the MI355X library.

    python -m parphom_amd.driver inp.dat [--precision S|D] [--batch B]

Follows main.F90:40-70: read_input (read_input.F90), read_lammps_data
(lammps_read.F90), read_q_file (q_read.F90), read_force_constants
(force_read.F90, one contiguous HDF5 read), then the q loop of
calculations.F90:49-53 as batched calls of the HIP library, writing one
HDF5 group per q in the layout of write_output.F90 (evec only with
`compute eigvecs : yes`, vel only with `group velocity : yes`).
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import formats, h5, lib


def run(input_file: str, precision: str = None, batch: int = None, log=print) -> str:
    t0 = time.time()
    cfg = formats.read_input(input_file)
    for msg in cfg.messages:
        log(msg)
    prec = precision or cfg.precision
    nb = batch if batch is not None else (cfg.batch or 0)
    ld = formats.read_lammps_data(cfg.lammps_path, cfg.natom, cfg.atom_types, cfg.atom_style)
    q = formats.read_q_file(cfg.q_path, cfg.nqpt)
    fc = h5.read_fc(cfg.fc_path, cfg.fc_dataset, cfg.natom)
    if cfg.print_progress:
        log(f"Inputs read: natom={cfg.natom} nqpt={cfg.nqpt} range={cfg.range_} "
            f"eigvecs={'yes' if cfg.evec_comp else 'no'} velocity={'yes' if cfg.comp_vel else 'no'} "
            f"precision={prec}")
    out_path = cfg.output_location + cfg.output_name
    with lib.Handle(ld.lat, ld.pos, ld.types, ld.masses, fc, jobz=cfg.jobz, rng=cfg.range_,
                    vel=cfg.comp_vel, vel_method=cfg.vel_method, precision=prec, vl=cfg.vl,
                    vu=cfg.vu, il=cfg.il, iu=cfg.iu, abstol=cfg.abstol, orfac=cfg.orfac,
                    batch=nb) as hd, h5.Output(out_path) as out:
        chunk = nb if nb > 0 else 32
        for q0 in range(0, len(q), chunk):
            qs = q[q0:q0 + chunk]
            res = hd.run(qs)
            for k in range(len(qs)):
                iq = q0 + k
                out.write(iq + 1, q[iq], res["eval_cm"][k],
                          evec=res["evec"][k] if (cfg.evec_comp and res["evec"] is not None) else None,
                          vel=res["vel"][k] if (cfg.comp_vel and res["vel"] is not None) else None)
                if res["info"][k]:
                    log(f"q-point {iq + 1}: eigensolver info {int(res['info'][k])}")
            if cfg.print_progress:
                log(f"q-points {q0 + 1}..{q0 + len(qs)} of {len(q)} done "
                    f"({time.time() - t0:.1f} s)")
    return out_path


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("input")
    ap.add_argument("--precision", choices=("S", "D"), default=None)
    ap.add_argument("--batch", type=int, default=None)
    a = ap.parse_args(argv)
    try:
        path = run(a.input, a.precision, a.batch)
    except formats.InputError as e:
        print(str(e), file=sys.stderr)
        return 1
    print(f"output written to {path}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
