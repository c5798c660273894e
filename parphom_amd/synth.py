"""Synthetic twisted-bilayer-graphene inputs for the phonon path.

Builds the commensurate (m, n) twisted-bilayer-graphene supercell, a seeded
spring-model force-constant set in the (N, N, 3, 3) layout that parphom.x
reads from HDF5 (`Phonon_Calculations/SRC/force_read.F90:93-158`), the q-point
sets used by the workloads, and writers for the positional LAMMPS data file
(`Phonon_Calculations/SRC/lammps_read.F90:66-192`) and the q file
(`Phonon_Calculations/SRC/q_read.F90:82-93`).

Geometry: graphene lattice a1 = a(1, 0), a2 = a(1/2, sqrt3/2), two carbon
sites per cell at 0 and (a1 + a2)/3.  Layer 1 uses the supercell
t1 = m a1 + n a2, t2 = -n a1 + (m+n) a2; layer 2 is the (n, m) supercell of
the same lattice rotated by theta, cos(theta) = (m^2+4mn+n^2)/(2(m^2+mn+n^2)),
so that both layers share t1, t2.  N = 4 (m^2 + mn + n^2).

Force constants (eV/A^2): intralayer radial / tangential springs on the
1st/2nd/3rd graphene neighbour shells, an interlayer radial spring
k(r) = k0 exp(-(r - 3.35)/0.5) for r < 6 A, the minimum periodic image for
every pair, +-1% seeded multiplicative noise on each pair block (applied
symmetrically, so Phi_ij = Phi_ji^T holds exactly), and self blocks from the
acoustic sum rule, Phi_ii = -sum_{j != i} Phi_ij.

The noise of the pair {i, j} is counter-based, a function of (seed, pair,
component) alone: splitmix64 of seed * 0x9E3779B97F4A7C15 + (min(i,j) N +
max(i,j)) * 8 + c for the six upper-triangle components c, mapped to
[-1, 1).  The library's GPU producer (pph_spring_fc) draws the same numbers.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

A_GRAPHENE = 2.46          # A
INTERLAYER = 3.35          # A
MASS_C = 12.011            # amu
BOX_Z = 20.0               # A, out-of-plane cell height (no z images are used)

# spring constants, eV/A^2, for the 1st/2nd/3rd intralayer shells
K_RADIAL = (25.0, 4.0, 1.0)
K_TANGENTIAL = (10.0, -1.0, 0.5)
K_INTER0 = 0.2
INTER_DECAY = 0.5
INTER_CUTOFF = 6.0
NOISE = 0.01

PARAM_NAMES = ("k_r1", "k_r2", "k_r3", "k_t1", "k_t2", "k_t3", "k_inter")


def default_params() -> np.ndarray:
    return np.array(K_RADIAL + K_TANGENTIAL + (K_INTER0,), dtype=np.float64)


@dataclass
class Structure:
    """A periodic cell: rows of `lat` are the lattice vectors (A), `pos` the
    Cartesian positions (A), `types` 1-based atom types, `masses` per type."""
    lat: np.ndarray
    pos: np.ndarray
    types: np.ndarray
    masses: np.ndarray
    layer: np.ndarray = field(default=None)
    m: int = 0
    n: int = 0

    @property
    def natom(self) -> int:
        return int(self.pos.shape[0])

    @property
    def theta_deg(self) -> float:
        m, n = self.m, self.n
        c = (m * m + 4 * m * n + n * n) / (2.0 * (m * m + m * n + n * n))
        return math.degrees(math.acos(min(1.0, c)))


def _layer_sites(p: int, r: int):
    """Integer graphene lattice points (i, j) inside the (p, r) supercell and
    their exact fractional coordinates (num1/det, num2/det)."""
    det = p * p + p * r + r * r
    pts = []
    for i in range(-r - 1, p + 2):
        for j in range(-1, p + 2 * r + 2):
            n1 = i * (p + r) + j * r
            n2 = -i * r + j * p
            if 0 <= n1 < det and 0 <= n2 < det:
                pts.append((n2, n1, i, j))
    pts.sort()
    assert len(pts) == det, (len(pts), det)
    return [(i, j) for (_, _, i, j) in pts], det


def _rot(theta):
    c, s = math.cos(theta), math.sin(theta)
    return np.array([[c, -s], [s, c]])


def tbg_structure(m: int, n: int, a: float = A_GRAPHENE, d: float = INTERLAYER,
                  box_z: float = BOX_Z) -> Structure:
    """Commensurate (m, n) twisted bilayer graphene, N = 4(m^2+mn+n^2)."""
    if m < 0 or n < 0 or (m == 0 and n == 0):
        raise ValueError("need m, n >= 0, not both zero")
    a1 = np.array([a, 0.0])
    a2 = np.array([0.5 * a, 0.5 * math.sqrt(3.0) * a])
    basis = [np.zeros(2), (a1 + a2) / 3.0]

    t1 = m * a1 + n * a2
    t1p = n * a1 + m * a2
    theta = math.atan2(t1[1], t1[0]) - math.atan2(t1p[1], t1p[0])
    phi = -math.atan2(t1[1], t1[0])          # rotate the whole cell so t1 || x
    L = float(np.linalg.norm(t1))
    lat2 = np.array([[L, 0.0], [0.5 * L, 0.5 * math.sqrt(3.0) * L]])
    inv2 = np.linalg.inv(lat2)

    pos, layer = [], []
    for lay, (p, r, rot) in enumerate(((m, n, _rot(phi)),
                                       (n, m, _rot(phi) @ _rot(theta)))):
        sites, _ = _layer_sites(p, r)
        for (i, j) in sites:
            for b in basis:
                xy = rot @ (i * a1 + j * a2 + b)
                f = xy @ inv2
                f = f - np.floor(f + 1e-12)
                f[np.abs(f) < 1e-12] = 0.0
                xy = f @ lat2
                z = 0.5 * box_z + (lay - 0.5) * d
                pos.append((xy[0], xy[1], z))
                layer.append(lay)
    pos = np.array(pos)
    lat = np.array([[L, 0.0, 0.0], [0.5 * L, 0.5 * math.sqrt(3.0) * L, 0.0],
                    [0.0, 0.0, box_z]])
    N = pos.shape[0]
    assert N == 4 * (m * m + m * n + n * n)
    return Structure(lat=lat, pos=pos, types=np.ones(N, dtype=np.int64),
                     masses=np.array([MASS_C]), layer=np.array(layer, dtype=np.int64),
                     m=m, n=n)


def _min_image_pairs(st: Structure, cutoff: float):
    """All (i, j, rvec) with |rvec| < cutoff, rvec = r_j + T - r_i the minimum
    in-plane periodic image (T over the 3x3 in-plane images), i != j."""
    from scipy.spatial import cKDTree
    lat = st.lat
    pos = st.pos
    shifts = np.array([[t1, t2, 0] for t1 in (-1, 0, 1) for t2 in (-1, 0, 1)], float) @ lat
    img = (pos[None, :, :] + shifts[:, None, :]).reshape(-1, 3)
    tree_img = cKDTree(img)
    tree = cKDTree(pos)
    sm = tree.sparse_distance_matrix(tree_img, cutoff, output_type="ndarray")
    i = sm["i"].astype(np.int64)
    k = sm["j"].astype(np.int64)
    N = pos.shape[0]
    j = k % N
    keep = (i != j)
    i, k, j = i[keep], k[keep], j[keep]
    rvec = img[k] - pos[i]
    # keep the minimum image for each (i, j) pair
    dist = np.linalg.norm(rvec, axis=1)
    order = np.lexsort((dist, j, i))
    i, j, rvec, dist = i[order], j[order], rvec[order], dist[order]
    first = np.ones(len(i), bool)
    first[1:] = (i[1:] != i[:-1]) | (j[1:] != j[:-1])
    return i[first], j[first], rvec[first]


@dataclass
class SpringBasis:
    """Pair list and per-parameter 3x3 blocks so that for parameters p,
    Phi_ij = -(sum_p p_k B_k[pair]) * noise[pair] for i != j."""
    i: np.ndarray
    j: np.ndarray
    blocks: np.ndarray      # (npairs, 7, 3, 3)
    noise: np.ndarray       # (npairs, 3, 3), symmetric, values in [1-NOISE, 1+NOISE]
    natom: int


def spring_pairs(st: Structure):
    """The pairs the spring model couples: (i, j, rvec, kind) with rvec the
    minimum-image separation and kind 0, 1, 2 (intralayer neighbour shell) or
    3 (interlayer, |r| < cutoff); other pairs carry no spring."""
    a = A_GRAPHENE
    shells = (a / math.sqrt(3.0), a, 2.0 * a / math.sqrt(3.0))
    i, j, r = _min_image_pairs(st, INTER_CUTOFF)
    dist = np.linalg.norm(r, axis=1)
    same = st.layer[i] == st.layer[j]
    kind = np.full(len(i), -1, np.int32)
    for s, d0 in enumerate(shells):
        kind[same & (np.abs(dist - d0) < 1e-3)] = s
    kind[(~same) & (dist < INTER_CUTOFF)] = 3
    keep = kind >= 0
    return i[keep], j[keep], r[keep], kind[keep]


def spring_basis(st: Structure, seed: int) -> SpringBasis:
    """Decompose the spring-model force constants into per-parameter blocks."""
    i, j, r, kind = spring_pairs(st)
    dist = np.linalg.norm(r, axis=1)
    rhat = r / dist[:, None]
    rr = rhat[:, :, None] * rhat[:, None, :]
    eye = np.eye(3)[None]
    blocks = np.zeros((len(i), 7, 3, 3))
    for s in range(3):
        sel = kind == s
        blocks[sel, s] = rr[sel]
        blocks[sel, 3 + s] = (eye - rr)[sel]
    inter = kind == 3
    kfac = np.exp(-(dist - INTERLAYER) / INTER_DECAY)
    blocks[inter, 6] = kfac[inter, None, None] * rr[inter]
    # symmetric multiplicative noise, identical for (i, j) and (j, i)
    N = st.natom
    noise = 1.0 + NOISE * pair_noise(seed, i, j, N)
    return SpringBasis(i=i, j=j, blocks=blocks, noise=noise, natom=N)


_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
# upper-triangle component c of a 3x3 block: (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
_TRIU = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def splitmix64(x: np.ndarray) -> np.ndarray:
    """The splitmix64 finaliser on uint64 (wrapping arithmetic)."""
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        z = x + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def pair_noise(seed: int, i, j, N: int) -> np.ndarray:
    """(npairs, 3, 3) symmetric noise in [-1, 1) of the pairs {i, j}: the same
    numbers for (i, j) and (j, i)."""
    i = np.asarray(i, np.int64)
    j = np.asarray(j, np.int64)
    key = (np.minimum(i, j) * N + np.maximum(i, j)).astype(np.uint64)
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * _GOLDEN + key * np.uint64(8)
    out = np.zeros((len(i), 3, 3))
    for c, (a, b) in enumerate(_TRIU):
        h = splitmix64(base + np.uint64(c))
        u = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)   # [0, 1)
        out[:, a, b] = 2.0 * u - 1.0
        out[:, b, a] = out[:, a, b]
    return out


def force_constants(st: Structure, seed: int, params=None, basis: SpringBasis = None,
                    dtype=np.float64) -> np.ndarray:
    """Dense (N, N, 3, 3) force constants, fc[i, j, a, b] = d^2E/du_ia du_jb."""
    if basis is None:
        basis = spring_basis(st, seed)
    p = default_params() if params is None else np.asarray(params, np.float64)
    N = basis.natom
    # the parameter sum in a fixed order (k = 0 .. 6), as the GPU producer adds it
    acc = np.zeros((len(basis.i), 3, 3))
    for k in range(len(p)):
        acc = acc + p[k] * basis.blocks[:, k]
    pair = -acc * basis.noise
    fc = np.zeros((N, N, 3, 3), dtype=np.float64)
    fc[basis.i, basis.j] = pair
    self_blk = np.zeros((N, 3, 3))
    np.add.at(self_blk, basis.i, pair)
    fc[np.arange(N), np.arange(N)] = -self_blk
    return fc.astype(dtype, copy=False)


def force_constants_param_derivative(st: Structure, seed: int, k: int,
                                     basis: SpringBasis = None) -> np.ndarray:
    """d fc / d params[k] (the model is linear in its parameters)."""
    e = np.zeros(7)
    e[k] = 1.0
    return force_constants(st, seed, params=e, basis=basis)


# ---------------------------------------------------------------- q sets
K_POINT = (2.0 / 3.0, 1.0 / 3.0, 0.0)
M_POINT = (0.5, 0.0, 0.0)
GAMMA = (0.0, 0.0, 0.0)


def q_path_gkmg(counts=(26, 13, 22)) -> np.ndarray:
    """Gamma-K-M-Gamma in reduced coordinates of b1, b2 (hexagonal cell with
    60 degrees between a1 and a2); each segment excludes its end point, so the
    closing Gamma is not repeated."""
    nodes = [np.array(GAMMA), np.array(K_POINT), np.array(M_POINT), np.array(GAMMA)]
    out = []
    for s, c in enumerate(counts):
        for t in range(c):
            out.append(nodes[s] + (nodes[s + 1] - nodes[s]) * (t / c))
    return np.array(out)


def q_grid(nk: int, shift=(0.0, 0.0)) -> np.ndarray:
    g = np.arange(nk) / nk
    q1, q2 = np.meshgrid(g, g, indexing="ij")
    q = np.stack([q1.ravel() + shift[0], q2.ravel() + shift[1], np.zeros(nk * nk)], axis=1)
    return q


def q_grid_shifted(nk: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    s = rng.uniform(0.0, 1.0 / nk, size=2)
    return q_grid(nk, (s[0], s[1]))


def q_random(nq: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    q = np.zeros((nq, 3))
    q[:, :2] = rng.uniform(0.0, 1.0, size=(nq, 2))
    return q


# ---------------------------------------------------------------- workloads
WORKLOADS = {
    # name: (m, n, fc seed, q builder)
    "tbg_1_2": (1, 2, 11, lambda: q_grid(64)),
    "tbg_3_4": (3, 4, 1003, lambda: q_path_gkmg()),
    "tbg_8_9": (8, 9, 2008, lambda: q_grid_shifted(16, 2)),
    "tbg_5_6": (5, 6, 3005, lambda: q_grid(64)),
    "tbg_11_12": (11, 12, 4011, lambda: q_random(128, 4)),
    "tbg_30_31": (30, 31, 5030, lambda: np.vstack([np.array([GAMMA, K_POINT, M_POINT]),
                                                    q_random(5, 5)])),
}


def workload(name: str):
    """(Structure, fc (N,N,3,3) f64, q (nq,3)) for a named workload."""
    m, n, seed, qf = WORKLOADS[name]
    st = tbg_structure(m, n)
    return st, force_constants(st, seed), qf()


# ---------------------------------------------------------------- writers
def write_lammps_data(path: str, st: Structure, style: str = "molecular") -> None:
    """Positional LAMMPS data file in the layout lammps_read.F90:66-192 parses:
    5 header lines, xlo/xhi, ylo/yhi, zlo/zhi, xy/xz/yz, 3 lines, masses,
    3 lines, then one atom line per atom with image flags."""
    lat = st.lat
    lines = ["LAMMPS data file (synthetic twisted bilayer graphene)", "",
             f"{st.natom} atoms", f"{len(st.masses)} atom types", "",
             f"{0.0:.16e} {lat[0, 0]:.16e} xlo xhi",
             f"{0.0:.16e} {lat[1, 1]:.16e} ylo yhi",
             f"{0.0:.16e} {lat[2, 2]:.16e} zlo zhi",
             f"{lat[1, 0]:.16e} {lat[2, 0]:.16e} {lat[2, 1]:.16e} xy xz yz",
             "", "Masses", ""]
    for t, mval in enumerate(st.masses):
        lines.append(f"{t + 1} {mval}")
    lines += ["", f"Atoms # {style}", ""]
    for a in range(st.natom):
        x, y, z = st.pos[a]
        t = int(st.types[a])
        mol = int(st.layer[a]) + 1 if st.layer is not None else 1
        if style == "atomic":
            lines.append(f"{a + 1} {t} {x:.16e} {y:.16e} {z:.16e} 0 0 0")
        elif style == "molecular":
            lines.append(f"{a + 1} {mol} {t} {x:.16e} {y:.16e} {z:.16e} 0 0 0")
        elif style == "full":
            lines.append(f"{a + 1} {mol} {t} 0.0 {x:.16e} {y:.16e} {z:.16e} 0 0 0")
        else:
            raise ValueError(style)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_q_file(path: str, q: np.ndarray) -> None:
    with open(path, "w") as f:
        for row in q:
            f.write(f"{row[0]:+.16f}\t{row[1]:+.16f}\t{row[2]:+.16f}\n")
