"""Readers for parphom.x's input formats: the `key : value` input file
(read_input.F90), the positional LAMMPS data file (lammps_read.F90) and the
q-point file (q_read.F90).  Behaviour, defaults and quirks follow the cited
reference lines; messages the reference prints are collected in `messages`.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import List

import numpy as np

_F32_1EM15 = float(np.float32(1e-15))

_YES = ("yes", "true", "Yes", "True", "YES", "TRUE")
_NO = ("no", "false", "No", "False", "NO", "FALSE")


class InputError(RuntimeError):
    """A condition on which parphom.x prints an error and exits."""


@dataclass
class InputConfig:
    """Parsed input file with read_input.F90:368-404 defaults."""
    lammps_location: str = "./"
    lammps_name: str = "lammps.dat"
    atom_style: str = "A"
    atom_types: int = 1
    natom: int = 1
    q_location: str = "./"
    q_name: str = "q_file.dat"
    nqpt: int = 1
    evec_comp: bool = False
    comp_vel: bool = False
    print_progress: bool = True
    vel_method: str = "A"
    jobz: str = "N"
    range_: str = "A"
    il: int = 1
    iu: int = 1
    vl: float = 0.0
    vu: float = 0.0
    mb: int = -1
    nb: int = -1
    abstol: float = _F32_1EM15
    orfac: float = _F32_1EM15
    fc_location: str = "./"
    fc_name: str = "FC_hdf5"
    fc_dataset: str = "force_constants"
    output_name: str = "results"
    output_location: str = "./"
    messages: List[str] = field(default_factory=list)
    # keys parphom.x does not know (it prints "not recognised" for them),
    # accepted here: precision and batch
    precision: str = "D"
    batch: int = 0

    @property
    def lammps_path(self):
        return self.lammps_location + self.lammps_name

    @property
    def q_path(self):
        return self.q_location + self.q_name

    @property
    def fc_path(self):
        return self.fc_location + self.fc_name

    @property
    def output_path(self):
        return self.output_location + self.output_name


def _fortran_real(s: str) -> float:
    return float(s.strip().split()[0].replace("d", "e").replace("D", "e"))


def _fortran_int(s: str) -> int:
    return int(s.strip().split()[0])


def read_input(path: str, nprocs: int = 1, sanitize: bool = True) -> InputConfig:
    """read_input.F90:35-333: one KEY : value per line, keys case-insensitive
    with spaces or underscores, '#'/'!' comments, unknown keys reported and
    ignored; then sanitize_input (read_input.F90:415-516)."""
    cfg = InputConfig()
    try:
        f = open(path)
    except OSError:
        raise InputError(f"Error reading input file {path}")
    with f:
        for raw in f:
            data = raw.rstrip("\n").strip()
            if not data or data[0] in "#!":
                continue
            p = data.find("!")
            if p >= 0:
                data = data[:p]
            p = data.find("#")
            if p >= 0:
                data = data[:p]
            p = data.find(":")
            key = data[:p] if p >= 0 else ""
            val = data[p + 1:] if p >= 0 else data
            key = key.upper().rstrip()
            v = val.strip()
            if key in ("LAMMPS FILE LOCATION", "LAMMPS_FILE_LOCATION"):
                cfg.lammps_location = v
            elif key in ("LAMMPS_FILE_NAME", "LAMMPS FILE NAME"):
                cfg.lammps_name = v
            elif key == "NATOM":
                cfg.natom = _fortran_int(v)
            elif key in ("ATOM_STYLE", "ATOM STYLE"):
                cfg.atom_style = v
            elif key in ("ATOM_TYPES", "ATOM TYPES"):
                cfg.atom_types = _fortran_int(v)
            elif key in ("FORCE CONSTANT FILE LOCATION", "FORCE_CONSTANT_FILE_LOCATION"):
                cfg.fc_location = v
            elif key in ("FORCE CONSTANT FILE NAME", "FORCE_CONSTANT_FILE_NAME"):
                cfg.fc_name = v
            elif key in ("FORCE CONSTANT DATASET", "FORCE_CONSTANT_DATASET"):
                cfg.fc_dataset = v
            elif key in ("Q_FILE_LOCATION", "Q FILE LOCATION"):
                cfg.q_location = v
            elif key in ("Q_FILE_NAME", "Q FILE NAME"):
                cfg.q_name = v
            elif key == "NQPT":
                cfg.nqpt = _fortran_int(v)
            elif key in ("COMPUTE_EIGVECS", "COMPUTE EIGVECS"):
                if v in _YES:
                    cfg.jobz, cfg.evec_comp = "V", True
                elif v in _NO:
                    cfg.jobz, cfg.evec_comp = "N", False
                else:
                    cfg.messages.append(f"Could not interpret command {v}\nEigenvectors will not be computed.")
                    cfg.jobz, cfg.evec_comp = "N", False
            elif key in ("GROUP VELOCITY", "GROUP_VELOCITY"):
                if v in _YES:
                    cfg.comp_vel = True
                elif v in _NO:
                    cfg.comp_vel = False
                else:
                    cfg.messages.append(f"Could not interpret command {v}\nVelocity will not be computed.")
                    cfg.comp_vel = False
            elif key in ("VELOCITY METHOD", "VELOCITY_METHOD"):
                if v in ("cd", "CD", "central_difference", "CENTRAL_DIFFERENCE"):
                    cfg.vel_method = "C"
                elif v in ("A", "a", "Analytic", "analytic"):
                    cfg.vel_method = "A"
                else:
                    cfg.messages.append(f"Could not interpret command {v}\n"
                                        "Velocity will be computed using analytic derivative.")
                    cfg.vel_method = "A"
            elif key == "RANGE":
                if v in ("A", "V", "I"):
                    cfg.range_ = v
                # anything else leaves the range unchanged (read_input.F90:202-206)
            elif key in ("MAX_EIGVAL", "MAX EIGVAL"):
                cfg.vu = _fortran_real(v)
            elif key in ("MIN_EIGVAL", "MIN EIGVAL"):
                cfg.vl = _fortran_real(v)
            elif key in ("MAX_INDEX", "MAX INDEX"):
                cfg.iu = _fortran_int(v)
            elif key in ("MIN_INDEX", "MIN INDEX"):
                cfg.il = _fortran_int(v)
            elif key == "ABSTOL":
                cfg.abstol = _fortran_real(v)
            elif key == "ORFAC":
                cfg.orfac = _fortran_real(v)
            elif key == "MB":
                cfg.mb = _fortran_int(v)
            elif key == "NB":
                cfg.nb = _fortran_int(v)
            elif key in ("OUTPUT FILE NAME", "OUTPUT_FILE_NAME"):
                cfg.output_name = v
            elif key in ("OUTPUT FILE LOCATION", "OUTPUT_FILE_LOCATION"):
                cfg.output_location = v
            elif key in ("PRINT PROGRESS", "PRINT_PROGRESS"):
                if v in _YES:
                    cfg.print_progress = True
                elif v in _NO:
                    cfg.print_progress = False
                else:
                    cfg.messages.append(f"Could not interpret command {v}\nProgress will be printed.")
                    cfg.print_progress = True
            elif key == "PRECISION":
                cfg.precision = "S" if v.upper() in ("S", "FP32", "SINGLE", "MIXED") else "D"
            elif key == "BATCH":
                cfg.batch = _fortran_int(v)
            else:
                cfg.messages.append(f"Command {key} not recognised\nIgnoring this command")
    if sanitize:
        sanitize_input(cfg, nprocs)
    return cfg


def sanitize_input(cfg: InputConfig, nprocs: int = 1) -> None:
    """read_input.F90:415-516."""
    for attr in ("lammps_location", "q_location", "output_location", "fc_location"):
        loc = getattr(cfg, attr)
        if not loc.endswith("/"):
            setattr(cfg, attr, loc.strip() + "/")
    if cfg.mb == -1:
        cfg.mb = min(32, cfg.natom) // nprocs + 1
    if cfg.nb == -1:
        cfg.nb = min(32, cfg.natom) // nprocs + 1
    if cfg.range_ == "V":
        cfg.il, cfg.iu = 1, cfg.natom
    elif cfg.range_ == "I":
        cfg.vl = cfg.vu = 0.0
    elif cfg.range_ == "A":
        cfg.il, cfg.iu = 1, cfg.natom
        cfg.vl = cfg.vu = 0.0
    if cfg.range_ == "I" and cfg.il > cfg.iu:
        raise InputError("Invalid input for eigenvalue indices to be computed")
    if cfg.range_ == "V" and cfg.vl > cfg.vu:
        raise InputError("Invalid input for the interval in which eigenvalue is to be computed")
    if cfg.comp_vel:
        cfg.jobz = "V"            # vectors computed, not written unless evec_comp
    full = cfg.output_location + cfg.output_name + ".hdf5"
    if os.path.exists(full):
        cfg.messages.append(f"File {cfg.output_name}.hdf5 already exists at {cfg.output_location}")
        k = 0
        while True:
            k += 1
            name = f"{cfg.output_name}_v{k}"
            if not os.path.exists(cfg.output_location + name + ".hdf5"):
                break
        cfg.output_name = name + ".hdf5"
        cfg.messages.append(f"Output file being renamed as {cfg.output_name}")
    else:
        cfg.output_name = cfg.output_name + ".hdf5"


@dataclass
class LammpsData:
    lat: np.ndarray        # rows are lattice vectors
    masses: np.ndarray
    types: np.ndarray      # 1-based
    pos: np.ndarray        # (N, 3)
    image: np.ndarray      # (N, 3) image flags (read, unused)
    layer: np.ndarray = None


def _skip(f, k):
    for _ in range(k):
        if not f.readline():
            raise InputError("unexpected end of LAMMPS file")


def _nums(line, count):
    parts = line.split()
    if len(parts) < count:
        raise InputError(f"malformed LAMMPS line: {line!r}")
    return parts[:count]


def read_lammps_data(path: str, natom: int, ntypes: int, style: str) -> LammpsData:
    """Positional parse of lammps_read.F90:66-192: skip 5 lines; xlo xhi,
    ylo yhi, zlo zhi, xy xz yz; skip 3; `ntypes` mass lines; skip 3; `natom`
    atom lines in the given style.  lat rows: (xhi-xlo,0,0), (xy,yhi-ylo,0),
    (xz,yz,zhi-zlo)."""
    try:
        f = open(path)
    except OSError:
        raise InputError(f"Error reading LAMMPS file {path}")
    with f:
        _skip(f, 5)
        x1, x2 = map(float, _nums(f.readline(), 2))
        y1, y2 = map(float, _nums(f.readline(), 2))
        z1, z2 = map(float, _nums(f.readline(), 2))
        xy, xz, yz = map(float, _nums(f.readline(), 3))
        lat = np.array([[x2 - x1, 0.0, 0.0], [xy, y2 - y1, 0.0], [xz, yz, z2 - z1]])
        _skip(f, 3)
        masses = np.zeros(ntypes)
        for t in range(ntypes):
            p = _nums(f.readline(), 2)
            masses[t] = float(p[1])
        _skip(f, 3)
        types = np.zeros(natom, np.int64)
        pos = np.zeros((natom, 3))
        img = np.zeros((natom, 3), np.int64)
        layer = None
        if style in ("atomic", "Atomic", "ATOMIC"):
            for a in range(natom):
                p = _nums(f.readline(), 8)
                types[a] = int(p[1]); pos[a] = [float(x) for x in p[2:5]]
                img[a] = [int(x) for x in p[5:8]]
        elif style in ("molecular", "Molecular", "MOLECULAR"):
            layer = np.zeros(natom, np.int64)
            for a in range(natom):
                p = _nums(f.readline(), 9)
                layer[a] = int(p[1]); types[a] = int(p[2]); pos[a] = [float(x) for x in p[3:6]]
                img[a] = [int(x) for x in p[6:9]]
        elif style in ("full", "Full", "FULL"):
            layer = np.zeros(natom, np.int64)
            for a in range(natom):
                p = _nums(f.readline(), 10)
                layer[a] = int(p[1]); types[a] = int(p[2]); pos[a] = [float(x) for x in p[4:7]]
                img[a] = [int(x) for x in p[7:10]]
        else:
            raise InputError("Atom style not supported. Supported styles are atomic and molecular")
    return LammpsData(lat=lat, masses=masses, types=types, pos=pos, image=img, layer=layer)


def read_q_file(path: str, nqpt: int) -> np.ndarray:
    """q_read.F90:82-93: nqpt list-directed rows of >= 3 reals (extra columns
    and trailing lines ignored).  Returns (nqpt, 3)."""
    try:
        f = open(path)
    except OSError:
        raise InputError(f"Error reading q-point file {path}")
    q = np.zeros((nqpt, 3))
    with f:
        i = 0
        for line in f:
            parts = line.replace(",", " ").split()
            if not parts:
                continue      # list-directed reads skip blank records
            if len(parts) < 3:
                raise InputError(f"q file line {i + 1}: need 3 values")
            q[i] = [_fortran_real(x) for x in parts[:3]]
            i += 1
            if i == nqpt:
                break
    if i < nqpt:
        raise InputError(f"q file {path} has {i} rows, nqpt = {nqpt}")
    return q
