"""CPU tests of the host side: the C-ABI library loads and exports every
symbol the headers declare (no GPU calls), and the input readers follow
read_input.F90 / lammps_read.F90 / q_read.F90."""
import ctypes
import os
import re

import numpy as np
import pytest

from parphom_amd import formats, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("parphom_hip.h", "parphom_hip_debug.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b(pph_[a-z_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    from parphom_amd import lib
    L = lib.load()
    names = _declared()
    assert {"pph_create", "pph_run", "pph_run_device", "pph_assemble", "pph_heevx",
            "pph_destroy", "pph_last_error", "pph_version"} <= names
    for n in names:
        assert hasattr(L, n), n
    assert set(lib.EXPORTED) >= names
    assert L.pph_version().decode().startswith("parphom_hip")


def test_argument_errors_are_return_codes():
    from parphom_amd import lib
    L = lib.load()
    h = ctypes.c_void_p()
    rc = L.pph_create(None, None, None, ctypes.byref(h))
    assert rc == -1 and b"NULL" in L.pph_last_error(None)
    o = lib.make_opts(rng="I", il=5, iu=2)
    w = np.zeros(4)
    m = np.zeros(1, np.int64)
    info = np.zeros(1, np.int32)
    A = np.zeros((2, 2), np.complex128)
    rc = L.pph_heevx(2, 1, A.ctypes.data, ctypes.byref(o), lib._ptr(w), None,
                     lib._ptr(m, ctypes.c_int64), lib._ptr(info, ctypes.c_int32), 0)
    assert rc == -1 and b"il <= iu" in L.pph_last_error(None)


INP = """# comment line
! another comment
lammps file location         : ./data
LAMMPS_FILE_NAME             : cell.dat   # trailing comment
natom                        : 28
atom types                   : 1
atom style                   : molecular
q file location              : ./
q file name                  : q.dat
nqpt                         : 3
force constant file location : fc
force constant file name     : FORCE_CONSTANTS
force constant dataset       : force_constants
compute eigvecs              : YES
range                        : I
min index                    : 13
max index                    : 16
group velocity               : yes
velocity method              : cd
abstol                       : 1E-14
orfac                        : 1d-14
num qpools                   : 3
output file name             : out
output_file_location         : ./
print progress               : no
"""


def test_read_input_keys_defaults_and_sanitise(tmp_path):
    p = tmp_path / "inp.dat"
    p.write_text(INP)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        cfg = formats.read_input(str(p))
    finally:
        os.chdir(cwd)
    assert cfg.lammps_location == "./data/" and cfg.lammps_name == "cell.dat"
    assert cfg.natom == 28 and cfg.atom_style == "molecular"
    assert cfg.fc_location == "fc/" and cfg.fc_dataset == "force_constants"
    assert cfg.evec_comp and cfg.jobz == "V"
    assert cfg.range_ == "I" and (cfg.il, cfg.iu) == (13, 16) and cfg.vl == cfg.vu == 0.0
    assert cfg.comp_vel and cfg.vel_method == "C"
    assert cfg.abstol == 1e-14 and cfg.orfac == 1e-14
    assert cfg.output_name == "out.hdf5" and not cfg.print_progress
    assert any("NUM QPOOLS" in m and "not recognised" in m for m in cfg.messages)
    # mb = nb = min(32, natom) / nprocs + 1
    assert cfg.mb == cfg.nb == 29


def test_read_input_defaults_and_quirks(tmp_path):
    p = tmp_path / "inp.dat"
    p.write_text("natom : 10\nrange : a\ngroup velocity : yes\ncompute eigvecs : maybe\n")
    cfg = formats.read_input(str(p))
    # lower-case range is not recognised and leaves 'A'; il/iu become 1..natom
    assert cfg.range_ == "A" and (cfg.il, cfg.iu) == (1, 10)
    # velocities force JOBZ='V' but not evec_comp (read_input.F90:483-485)
    assert cfg.jobz == "V" and not cfg.evec_comp
    assert cfg.abstol == float(np.float32(1e-15))
    assert any("Could not interpret" in m for m in cfg.messages)


def test_sanitise_errors_and_output_rename(tmp_path):
    p = tmp_path / "inp.dat"
    p.write_text("range : I\nmin index : 5\nmax index : 2\n")
    with pytest.raises(formats.InputError):
        formats.read_input(str(p))
    p.write_text("range : V\nmin eigval : 5\nmax eigval : 2\n")
    with pytest.raises(formats.InputError):
        formats.read_input(str(p))
    (tmp_path / "res.hdf5").write_text("x")
    (tmp_path / "res_v1.hdf5").write_text("x")
    p.write_text(f"output file name : res\noutput file location : {tmp_path}\n")
    cfg = formats.read_input(str(p))
    assert cfg.output_name == "res_v2.hdf5"


@pytest.mark.parametrize("style", ["atomic", "molecular", "full"])
def test_lammps_roundtrip(tmp_path, style):
    st = synth.tbg_structure(1, 2)
    f = tmp_path / "cell.dat"
    synth.write_lammps_data(str(f), st, style)
    ld = formats.read_lammps_data(str(f), st.natom, 1, style)
    assert np.allclose(ld.lat, st.lat, atol=1e-14)
    assert np.allclose(ld.pos, st.pos, atol=1e-14)
    assert np.array_equal(ld.types, st.types) and ld.masses[0] == st.masses[0]
    with pytest.raises(formats.InputError):
        formats.read_lammps_data(str(f), st.natom, 1, "charge")


def test_q_file_extra_columns_and_trailing_lines(tmp_path):
    f = tmp_path / "q.dat"
    f.write_text("0.1 0.2 0.0 0.5\n\n0.3 0.4 0.0 0.5\n7\n")
    q = formats.read_q_file(str(f), 2)
    assert np.allclose(q, [[0.1, 0.2, 0.0], [0.3, 0.4, 0.0]])
    with pytest.raises(formats.InputError):
        formats.read_q_file(str(f), 5)


def test_synthetic_structures():
    for (m, n) in [(1, 2), (3, 4), (5, 6)]:
        st = synth.tbg_structure(m, n)
        assert st.natom == 4 * (m * m + m * n + n * n)
        # no two atoms closer than the C-C bond within a layer
        from scipy.spatial.distance import pdist
        assert pdist(st.pos).min() > 1.3
    assert abs(synth.tbg_structure(3, 4).theta_deg - 9.43) < 0.01
    assert abs(synth.tbg_structure(11, 12).theta_deg - 2.88) < 0.01
    q = synth.q_path_gkmg()
    assert q.shape == (61, 3) and np.allclose(q[26], synth.K_POINT)


def test_dos_argument_errors():
    """Checked on the host before any device work."""
    from parphom_amd import lib
    freq = np.array([[1.0, 2.0], [1.5, 2.5], [0.5, 3.0]])
    with pytest.raises(lib.PphError, match="out of range"):
        lib.dos_triangulation(freq, [[0, 1, 3]], np.arange(0, 4, 0.5))
    with pytest.raises(lib.PphError, match="ascending"):
        lib.dos_triangulation(freq, [[0, 1, 2]], np.array([1.0, 0.5]))
    with pytest.raises(lib.PphError, match="method"):
        lib.dos_smearing(3, freq, 1.0, 0.1, np.arange(0, 4, 0.5), 1)
    with pytest.raises(lib.PphError, match="ascending"):
        lib.dos_smearing(1, freq[:, ::-1], 1.0, 0.1, np.arange(0, 4, 0.5), 1)


def test_dos_grid_and_triangles(tmp_path):
    """Full-grid file parsing and the Delaunay triangles on the pyphutil
    construction of Cartesian points; every grid cell is split in two."""
    from parphom_amd import dos
    lat = np.array([[2.46, 0.0, 0.0], [1.23, 2.130422493309719, 0.0], [0.0, 0.0, 20.0]])
    m = 4
    g = np.array([[i / m, j / m, 0.0] for i in range(m) for j in range(m)])
    p = tmp_path / "grid"
    with open(p, "w") as f:
        for r in g:
            f.write("%.8f %.8f %.8f\n" % tuple(r))
        f.write(" ".join(str(1 + (k % 5)) for k in range(len(g))) + "\n")
    grid, mapping = dos.read_full_grid((m, m, 1), str(p))
    assert np.allclose(grid, g) and mapping.tolist() == [1 + (k % 5) for k in range(len(g))]
    tri = dos.triangles(lat, (m, m, 1), grid, mapping)
    assert tri.shape == (2 * (m - 1) ** 2, 3)
    assert tri.min() >= 0 and tri.max() < 5


def test_tuner_picks_fastest_candidate_and_records_rejections():
    """The autotuner's selection on a stub handle whose run time depends on
    the configuration (no GPU): the fastest candidate wins, rejected
    blockings stay in the table without a rate, every candidate is warmed up
    once before its timed calls."""
    from parphom_amd import tune

    cost = {(16, 1, 32, 64): 4.0, (16, 1, 32, 128): 3.0, (32, 1, 32, 64): 2.5,
            (32, 1, 32, 128): 2.0, (16, 2, 32, 64): 5.0, (16, 2, 32, 128): 1.0,
            (32, 2, 32, 64): 6.0, (32, 2, 32, 128): 4.5}
    now = [0.0]
    calls = []

    class StubHandle:
        def __init__(self, *a, batch, streams, **k):
            self.cfg = [batch, streams, None, None]

        def __enter__(self):
            return self

        def __exit__(self, *a):
            return False

        def blocking(self, nbp, nbb):
            if nbp == 40:
                raise ValueError("unsupported panel")
            self.cfg[2:] = [nbp, nbb]

        def run(self, q):
            calls.append(tuple(self.cfg))
            now[0] += cost[tuple(self.cfg)] * len(q) / 100.0

    res = tune.tune(None, None, None, None, None, np.zeros((5, 3)), batches=(16, 32),
                    streams=(1, 2), panels=(32, 40), backs=(64, 128), reps=2,
                    handle_factory=StubHandle, clock=lambda: now[0])
    assert res["config"] == {"batch": 16, "streams": 2, "reduction": -1, "tile": "64x64",
                             "nb_panel": 32, "nb_back": 128}
    assert abs(res["qps"] - 100.0) < 1e-9
    rejected = [r for r in res["table"] if r["nb_panel"] == 40]
    assert len(rejected) == 8 and all(r["qps"] is None and "error" in r for r in rejected)
    # one warm-up and two timed calls per accepted candidate
    assert len(calls) == 3 * 8


def test_tuner_records_failing_candidates_and_goes_on():
    """A candidate whose handle cannot be made (a batch that does not fit in
    memory) or whose run raises is recorded with its error; the search goes
    on and picks the best of the others."""
    from parphom_amd import tune
    now = [0.0]

    class StubHandle:
        def __init__(self, *a, batch, streams, **k):
            if batch == 64:
                raise MemoryError("out of device memory")
            self.batch, self.streams = batch, streams

        def __enter__(self):
            return self

        def __exit__(self, *a):
            return False

        def blocking(self, nbp, nbb):
            self.nbb = nbb

        def run(self, q):
            if self.streams == 2 and self.nbb == 64:
                raise RuntimeError("pph_run: out of memory")
            now[0] += (1.0 if self.nbb == 128 else 2.0) * len(q) / 100.0

    res = tune.tune(None, None, None, None, None, np.zeros((5, 3)), batches=(16, 64),
                    streams=(1, 2), panels=(32,), backs=(64, 128), reps=1,
                    handle_factory=StubHandle, clock=lambda: now[0])
    t = res["table"]
    assert len(t) == 8
    assert all(r["qps"] is None and "device memory" in r["error"] for r in t if r["batch"] == 64)
    failed = [r for r in t if r["batch"] == 16 and r["streams"] == 2 and r["nb_back"] == 64]
    assert len(failed) == 1 and failed[0]["qps"] is None and "out of memory" in failed[0]["error"]
    assert res["config"]["nb_back"] == 128 and abs(res["qps"] - 100.0) < 1e-9


def test_tuner_searches_tile_shapes_and_reductions():
    """The tile shape and the reduction are search dimensions: each candidate
    sets them on its handle (tune.apply) before its calls, a shape the
    handle rejects is recorded with its error, and the fastest shape wins."""
    from parphom_amd import tune
    now = [0.0]
    seen = []

    class StubHandle:
        def __init__(self, *a, batch, streams, **k):
            self.tile, self.red = None, -1

        def __enter__(self):
            return self

        def __exit__(self, *a):
            return False

        def reduction(self, mode):
            self.red = mode

        def gemm_tile(self, tile):
            if tile == "128x128":
                raise ValueError("the 128x128 tile is FP32 only")
            self.tile = tile

        def blocking(self, nbp, nbb):
            pass

        def run(self, q):
            seen.append((self.red, self.tile))
            cost = {"64x64": 3.0, "128x64": 2.0, "64x128": 2.5}[self.tile] + (0.5 if self.red == 0 else 0)
            now[0] += cost * len(q) / 100.0

    res = tune.tune(None, None, None, None, None, np.zeros((4, 3)), batches=(8,), streams=(1,),
                    panels=(32,), backs=(128,), reductions=(0, 1),
                    tiles=("64x64", "128x64", "64x128", "128x128"), handle_factory=StubHandle,
                    clock=lambda: now[0])
    assert res["config"]["tile"] == "128x64" and res["config"]["reduction"] == 1
    assert len(res["table"]) == 8
    bad = [r for r in res["table"] if r["tile"] == "128x128"]
    assert len(bad) == 2 and all(r["qps"] is None and "FP32 only" in r["error"] for r in bad)
    assert set(seen) == {(r, t) for r in (0, 1) for t in ("64x64", "128x64", "64x128")}


def test_complement_correction_formula():
    """The refinement's second-order eigenvalue correction for a selection
    that starts at the bottom of the spectrum: with x an approximate lowest
    eigenvector (error spread over all modes, as rounding leaves it), theta its
    Rayleigh quotient and r = A x - theta x, the sum over the computed modes
    -sum_j |x_j^H r|^2 / (rho_j - theta) misses the modes above the
    selection; their share is -r_perp^H (P (A - theta) P)^+ r_perp, r_perp
    the residual outside the computed span.  With it the estimate is third
    order in the error; without it, not.  Conjugate gradients on
    P (A - theta) P (as the library runs them) give the same share."""
    rng = np.random.default_rng(3)
    n, M, eps = 120, 20, 1e-4
    lam = np.sort(np.concatenate([[0.001], rng.uniform(0.5, 10.0, n - 1)]))
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = (Q * lam) @ Q.conj().T
    A = 0.5 * (A + A.conj().T)
    # computed modes: exact ones perturbed by eps in random directions
    X = Q[:, :M] + eps * (rng.standard_normal((n, M)) + 1j * rng.standard_normal((n, M)))
    X /= np.linalg.norm(X, axis=0)
    rho = np.real(np.einsum("im,ij,jm->m", X.conj(), A, X))
    x = X[:, 0]
    th = rho[0]
    r = A @ x - th * x
    inspan = -sum(abs(np.vdot(X[:, j], r)) ** 2 / (rho[j] - th) for j in range(1, M))
    P = np.eye(n) - X @ np.linalg.solve(X.conj().T @ X, X.conj().T)
    rp = P @ r
    Ap = P @ (A - th * np.eye(n)) @ P
    comp = -np.real(np.vdot(rp, np.linalg.lstsq(Ap, rp, rcond=None)[0]))
    # conjugate gradients on the projected operator, from zero
    z = np.zeros(n, complex)
    res = rp.copy()
    pdir = rp.copy()
    rr = np.real(np.vdot(res, res))
    for _ in range(60):
        qv = Ap @ pdir
        al = rr / np.real(np.vdot(pdir, qv))
        z += al * pdir
        res -= al * qv
        rn = np.real(np.vdot(res, res))
        pdir = res + (rn / rr) * pdir
        rr = rn
    comp_cg = -np.real(np.vdot(rp, z))
    assert abs(comp_cg - comp) <= 1e-6 * abs(comp)
    err_th = abs(th - lam[0])
    err_span = abs(th + inspan - lam[0])
    err_full = abs(th + inspan + comp - lam[0])
    assert err_full < 1e-3 * err_th
    assert err_span > 100 * err_full
