"""Throughput of the phonon hot path on one MI355X.

Workload (default): commensurate twisted bilayer graphene (11,12), N = 1588
atoms, n = 3N = 4764 complex Hermitian D(q) from the seeded spring-model
force constants, 128 random q-points; every step is one device batch of q
through assembly -> eigensolve (all eigenpairs) -> group velocities, with
the FP32 matrix-core eigensolver and FP64 refinement (precision 'S').  The
same batch on the FP64 path (precision 'D') is timed alongside, and the
frequencies of both are compared.

Prints one JSON line (see the keys below).  Synthetic data; inputs (force
constants, geometry, tables) are resident on the GPU before timing starts,
outputs stay in device buffers.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12            # B/s
FP32_PEAK = 157.3e12         # FLOP/s (dense, matrix and vector)
FP64_PEAK = 78.6e12
HEMV_TIME_STRIDE = 8         # the library times every 8th hemv launch


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def device_outputs(torch, B, n):
    dev = torch.device("cuda:0")
    return {
        "eval_cm": torch.zeros((B, n), dtype=torch.float64, device=dev),
        "m": torch.zeros(B, dtype=torch.int64, device=dev),
        "nz": torch.zeros(B, dtype=torch.int64, device=dev),
        "info": torch.zeros(B, dtype=torch.int32, device=dev),
        "evec": torch.zeros((B, n, n), dtype=torch.complex128, device=dev),
        "vel": torch.zeros((B, n, 3), dtype=torch.float64, device=dev),
    }


def time_path(lib, torch, st, fc, qs, prec, B, steps, warmup, streams, reduction=-1,
              tile="64x64", prof_steps=3, vec_refine=-1, vec_bf16=-1):
    """Each step is one pph_run_device call over streams x B q-points: the
    library keeps `streams` batches of B in flight (one stream and host thread
    each).  Returns a dict: the timed seconds, the phase profile of the timed
    steps (concurrent batches) and of one batch run alone, the frequencies and
    velocities of the first B q-points, the eigenvectors of the first q, the
    number of force-constant blocks the Rayleigh quotients and velocities use
    (0: dense products), and whether the two-stage reduction ran."""
    n = 3 * st.natom
    h = lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz="V", vel=True, precision=prec,
                   batch=B, streams=streams)
    nblk = h.sparse(1)               # blocks the quadratic forms use (0 = dense products)
    h.reduction(reduction)           # -1: the library's choice
    if not (prec == "D" and tile == "128x128"):     # the 128 x 128 tile is FP32 only
        h.gemm_tile(tile)
    if vec_refine >= 0 or vec_bf16 >= 0:     # FP32 path: FP64 correction steps of the vectors
        h.vector_refinement(vec_refine if vec_refine >= 0 else 2,
                            bf16=bool(vec_bf16) if vec_bf16 >= 0 else True)
    per_step = B * streams
    out = device_outputs(torch, per_step, n)
    nq = len(qs)

    def step_q(i):
        idx = [(i * per_step + k) % nq for k in range(per_step)]
        return qs[idx]

    for i in range(warmup):
        h.run_device(step_q(i), out)
        log(f"[bench] {prec} warmup {i + 1}/{warmup} done")
    torch.cuda.synchronize()
    g0 = h.graphs()
    t0 = time.perf_counter()
    for i in range(steps):
        h.run_device(step_q(warmup + i), out)
        log(f"[bench] {prec} step {i + 1}/{steps} done")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    g1 = h.graphs()
    # the phase profile of the first prof_steps of the same steps, recorded
    # in a second pass (the library's phase timers run every batch eagerly,
    # the timed pass replays the captured batch graphs)
    h.profile(enable=True, reset=True)
    for i in range(min(steps, prof_steps)):
        h.run_device(step_q(warmup + i), out)
    torch.cuda.synchronize()
    prof = h.profile(enable=True, reset=True)
    # one batch alone (untimed; also the accuracy check): the first B q-points
    h.run_device(qs[:B], out)
    prof_iso = h.profile(enable=False)
    res = {"dt": dt, "prof": prof, "iso": prof_iso, "nblk": nblk,
           "ev": out["eval_cm"][:B].cpu().numpy().copy(),
           "vel": out["vel"][:B].cpu().numpy().copy(),
           "z0": out["evec"][0].cpu().numpy().T.copy(),   # eigenvectors of the first q, columns
           "two_stage": prof_iso["q2"][1] > 0,
           # the one-launch small-matrix reduction records no hemv launches
           "small": prof_iso["q2"][1] == 0 and prof_iso["hemv_launches"][1] == 0,
           "graph_replays": g1[1] - g0[1], "graph_captures": g1[0] - g0[0]}
    h.close()
    del out
    torch.cuda.empty_cache()
    return res


def velocity_difference(v, v_ref, om_ref, rel=1e-6):
    """Largest difference of the velocities of two runs on the same q, through
    sums over clusters of eigenvalues closer than rel * max|omega^2| (inside a
    degenerate cluster the individual velocities depend on the basis; their
    sum does not).  Returns (max abs difference m/s, max |v_ref|)."""
    lam = np.sign(om_ref) * om_ref ** 2
    tol = rel * np.abs(lam).max()
    groups, cur = [], [0]
    for k in range(1, len(lam)):
        if lam[k] - lam[k - 1] <= tol:
            cur.append(k)
        else:
            groups.append(cur)
            cur = [k]
    groups.append(cur)
    d = max(float(np.abs(v[g].sum(axis=0) - v_ref[g].sum(axis=0)).max()) for g in groups)
    return d, float(np.abs(v_ref).max())


def split_profile(prof, n):
    """Phase ms per batch, and the mean duration of one hemv launch in ms (the
    library times every HEMV_TIME_STRIDE-th of the n - 1 launches per batch;
    the per-batch hemv figure is that mean times n - 1)."""
    phases = {k: v[0] / max(1, v[1]) for k, v in prof.items() if k != "hemv_launches"}
    ms, cnt = prof["hemv_launches"]
    launch_ms = ms / max(1, cnt)
    phases["hemv"] = launch_ms * (n - 1)
    return phases, launch_ms


def kernel_traffic(kernel, workload, n, B):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of the same workload and batch (counters
    scaled by a device copy of known size), or None when no summary matches."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_b{B}_traffic.md")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        if line.startswith("- ") and ":" in line:
            k, v = line[2:].split(":", 1)
            kv[k.strip()] = v.strip().strip("`")
    if kv.get("n") != str(n) or kv.get("batch") != str(B):
        return None
    return float(kv["traffic_bytes_per_launch"])


def gemm_peak(torch, dtype, size=8192, reps=5):
    """Dense real GEMM rate of the device's library GEMM (torch.matmul) in
    FLOP/s: the measured matrix peak the path roofline is also quoted on."""
    a = torch.randn(size, size, dtype=dtype, device="cuda:0")
    b = torch.randn(size, size, dtype=dtype, device="cuda:0")
    torch.matmul(a, b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        c = torch.matmul(a, b)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    del a, b, c
    torch.cuda.empty_cache()
    return 2.0 * size ** 3 * reps / dt


def path_roofline(N, n, B, qps, peak_spec, peak_meas):
    """Whole-path roofline per q-point: t_ideal = useful flops / matrix peak +
    assembly bytes / HBM peak, with the useful flops of the textbook path
    (16/3 n^3 Hermitian tridiagonalisation + 8 n^3 back-transformation of all
    vectors + 24 n^3 for the three velocity products) and the assembly bytes
    72 N^2 / B (force constants once per batch) + 4 * 16 n^2 (D and three dD in
    c128).  Refinement, the tridiagonal eigensolver and the work the Hermitian
    products save are not counted.  When the velocity forms use only the
    nonzero force-constant blocks, the 24 n^3 of dense velocity products are
    not performed at all; `frac_eigensolve_only` then leaves them out of the
    ideal (16/3 n^3 + 8 n^3 and the assembly bytes)."""
    flops = (16.0 / 3.0 + 8.0 + 24.0) * float(n) ** 3
    flops_eig = (16.0 / 3.0 + 8.0) * float(n) ** 3
    k1 = 72.0 * N * N / B + 4 * 16.0 * n * n
    t_meas = 1.0 / qps
    t_spec = flops / peak_spec + k1 / HBM_PEAK
    t_meas_peak = flops / peak_meas + k1 / HBM_PEAK
    t_eig = flops_eig / peak_spec + k1 / HBM_PEAK
    return {"flops_per_q": flops, "assembly_bytes_per_q": k1,
            "ideal_ms_per_q": 1e3 * t_spec, "measured_ms_per_q": 1e3 * t_meas,
            "frac": t_spec / t_meas, "frac_vs_measured_peak": t_meas_peak / t_meas,
            "frac_eigensolve_only": t_eig / t_meas}


def cpu_baseline(st, fc, q, seconds_cap=60.0):
    """The FP64 oracle (numpy assembly, LAPACK zheevx, numpy GEMMs) on one
    q-point of the same workload, on this host's cores."""
    from oracle import phonon as O
    S = O.System(st.lat, st.pos, st.types, st.masses, fc)
    t0 = time.perf_counter()
    r = S.solve(q, vel=True)
    dt = time.perf_counter() - t0
    return dt, r


def scalapack_baseline(st, fc, q, ranks=16, nb=32, timeout=600):
    """parphom.x's per-q work on the host's cores: the oracle's assembly of
    D(q) and dD/dq (numpy, timed), then MKL ScaLAPACK pzheevx('V', 'A', 'U')
    and the six velocity pzgemm calls with the reference's argument choices
    (oracle/_bin/scalapack_baseline under MPICH, one rank per core, one
    thread each, mb = nb = 32 as in the examples' inp.dat).  Returns None
    when the baseline or mpiexec is not there or the run fails."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_bin", "scalapack_baseline")
    mpiexec = "/opt/conda/bin/mpiexec"
    if not (os.path.exists(exe) and os.path.exists(mpiexec)):
        return None
    from oracle import phonon as O
    S = O.System(st.lat, st.pos, st.types, st.masses, fc)
    t0 = time.perf_counter()
    mats = [S.D(q)] + [S.D(q, a) for a in (1, 2, 3)]
    t_asm = time.perf_counter() - t0
    n = mats[0].shape[0]
    fd, path = tempfile.mkstemp(suffix=".bin")
    try:
        with os.fdopen(fd, "wb") as f:
            for M in mats:
                f.write(np.asfortranarray(M, dtype=np.complex128).tobytes(order="F"))
        del mats
        env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")
        r = subprocess.run([mpiexec, "-n", str(ranks), exe, path, str(n), "1", str(nb)],
                           capture_output=True, text=True, timeout=timeout, env=env)
        line = [l for l in r.stdout.splitlines() if l.startswith("{")]
        if r.returncode != 0 or not line:
            log(f"[bench] scalapack baseline failed (rc {r.returncode}): {r.stderr[-400:]}")
            return None
        out = json.loads(line[-1])
    except (OSError, subprocess.SubprocessError, ValueError) as e:
        log(f"[bench] scalapack baseline failed: {e}")
        return None
    finally:
        os.unlink(path)
    out["asm_s_per_q"] = t_asm
    out["s_per_q"] = t_asm + out["eig_s_per_q"] + out["vel_s_per_q"]
    return out


def dos_workload(nbands=4764, mesh=18, seed=7):
    """Transport-DOS shaped workload (the reference's
    Utilities/Examples/transport_dos.py: 18 x 18 full grid, linear
    triangulation with |v|^2 weights, energies -2 .. 1800 cm^-1 every 0.025):
    synthetic flat moire bands, base frequencies uniform in [0, 1600] cm^-1
    with a nearest-neighbour dispersion of 0.5-3 cm^-1 per band."""
    from parphom_amd import dos as D
    rng = np.random.default_rng(seed)
    base = np.sort(rng.uniform(0.0, 1600.0, nbands))
    amp = rng.uniform(0.5, 3.0, nbands)
    g = np.array([[i / mesh, j / mesh, 0.0] for i in range(mesh) for j in range(mesh)])
    ph = 2 * np.pi * g
    shape = np.cos(ph[:, 0]) + np.cos(ph[:, 1]) + np.cos(ph[:, 0] + ph[:, 1])
    freq = base[None, :] + amp[None, :] * shape[:, None]
    func = 1e6 * (1.0 + 0.5 * np.sin(ph[:, 0])[:, None] * np.cos(base / 50.0)[None, :]) ** 2
    lat = np.array([[100.0, 0, 0], [50.0, 86.60254037844386, 0], [0, 0, 20.0]])
    tri = D.triangles(lat, (mesh, mesh, 1), g, np.arange(1, len(g) + 1))
    E = np.arange(-2.0, 1800.0, 0.025)
    return freq, func, tri, E


def dos_bench(lib, torch, a):
    """One JSON line for the DOS row: linear triangulation on the GPU with
    inputs resident in HBM, against the reference routine compiled from its
    own source (oracle/_ref) on a slice of the energies."""
    freq, func, tri, E = dos_workload()
    nk, nb = freq.shape
    dev = torch.device("cuda:0")
    tf = torch.from_numpy(np.ascontiguousarray(freq)).to(dev)
    tg = torch.from_numpy(np.ascontiguousarray(func)).to(dev)
    tt = torch.from_numpy(np.ascontiguousarray(tri, np.int64)).to(dev)
    tE = torch.from_numpy(E).to(dev)
    td = torch.zeros(len(E), dtype=torch.float64, device=dev)
    tn = torch.zeros(len(E), dtype=torch.float64, device=dev)
    import ctypes as C
    P = lambda t: C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double))
    Pi = lambda t: C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_int64))

    def run():
        lib._check(lib.load().pph_dos_triangulation(nk, nb, P(tf), P(tg), len(tri), Pi(tt), len(E),
                                                    P(tE), P(td), P(tn), 0, 1))
    for _ in range(a.warmup):
        run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    evals = float(len(E)) * len(tri) * nb          # the reference's loop count
    # energies that fall inside a (triangle, band) window [min f, max f]: the
    # ones the kernel evaluates (two FP64 atomics each); the rest are known
    # constants (0 below, the sum of the weights above)
    fv = freq[tri]                                   # (ntri, 3, nb)
    inwin = float(np.sum(np.searchsorted(E, fv.max(axis=1), side="right")
                         - np.searchsorted(E, fv.min(axis=1), side="left")))
    res = {"metric": "linear-triangulation DOS evaluations/s (E x triangles x bands)",
           "value": evals / dt, "unit": "evaluations/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": 1e3 * dt, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic flat bands on an 18 x 18 full grid",
           "config": {"workload": "transport_dos_18x18", "nk": nk, "nbands": nb,
                      "triangles": len(tri), "energies": len(E)},
           "in_window_evaluations": inwin, "in_window_per_s": inwin / dt,
           "fp64_atomics_per_s": 2 * inwin / dt}
    # CPU: the reference routine (oracle/_ref, compiled from bz_integration.f90)
    # on the first 8 energies above 800 cm^-1, all triangles and bands
    from oracle import dos as OD
    if OD.ref_available() and not a.no_cpu:
        i0 = int(np.searchsorted(E, 800.0))
        Es = E[i0:i0 + 8]
        t1 = time.perf_counter()
        d0, n0 = OD.ref_linear_triangulation(tri, Es, freq, func)
        ct = time.perf_counter() - t1
        res["cpu_baseline"] = {"value": len(Es) * len(tri) * nb / ct, "unit": "evaluations/s",
                               "cores": 1, "kind": "reference",
                               "sample": f"{len(Es)} energies x {len(tri)} triangles x {nb} bands"}
        got_d, got_n = td.cpu().numpy()[i0:i0 + 8], tn.cpu().numpy()[i0:i0 + 8]
        res["max_rel_diff_vs_reference"] = float(max(
            np.abs(got_d - d0).max() / max(1e-300, np.abs(d0).max()),
            np.abs(got_n - n0).max() / max(1e-300, np.abs(n0).max())))
    print(json.dumps(res))


def fcsym_bench(lib, torch, a, iterations=10):
    """One JSON line for the force-constant symmetrisation (symmetrize_fc of
    the reference's producer): `iterations` rounds of the acoustic sum rule
    and inversion symmetry on the (N, N, 3, 3) force constants of the
    headline system, resident in HBM."""
    from parphom_amd import synth
    st, fc, _ = synth.workload(a.workload)
    N = st.natom
    dev = torch.device("cuda:0")
    t = torch.from_numpy(np.ascontiguousarray(fc)).to(dev)
    import ctypes as C

    def run():
        lib._check(lib.load().pph_fc_symmetrize(N, C.cast(C.c_void_p(t.data_ptr()),
                                                           C.POINTER(C.c_double)),
                                                iterations, 0, 1))
    for _ in range(a.warmup):
        run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps / iterations      # one iteration
    # per iteration each step reads and writes the 72 N^2-byte array once
    alg = 2 * 2 * 72.0 * N * N
    res = {"metric": "force-constant symmetrisation iterations/s", "value": 1.0 / dt,
           "unit": "iterations/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": 1e3 * dt * iterations, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic spring-model force constants",
           "config": {"workload": a.workload, "natom": N, "iterations_per_step": iterations},
           "roofline": {"bound": "hbm", "achieved": alg / dt / 1e9, "peak": HBM_PEAK / 1e9,
                        "unit": "GB/s", "frac": alg / dt / HBM_PEAK, "traffic": None,
                        "kernel": "k_fc_asr + k_fc_sym (one iteration)"}}
    if not a.no_cpu:
        from oracle import fc as OF
        t1 = time.perf_counter()
        ref = OF.symmetrize(fc, 1)
        ct = time.perf_counter() - t1
        res["cpu_baseline"] = {"value": 1.0 / ct, "unit": "iterations/s", "cores": 1,
                               "kind": "port", "sample": f"1 iteration at N = {N} (numpy)"}
        got = t.cpu().numpy()
        chk = OF.symmetrize(fc, (a.warmup + a.steps) * iterations) if N <= 2000 else None
        if chk is not None:
            res["max_rel_diff_vs_oracle"] = float(np.abs(got - chk).max() / np.abs(chk).max())
        del ref
    print(json.dumps(res))


def large_bench(lib, torch, a, nvec=2000):
    """The large single-GPU workload: (30,31) twisted bilayer graphene, N = 11164
    atoms, n = 33492.  Range I (the lowest `nvec` eigenpairs) on its 8 q-points
    (Gamma, K, M and 5 seeded random), FP32+refine and FP64, then range A
    eigenvalues-only on Gamma, K and M in both precisions.  Every run is
    pph_run with host output buffers, two q-points per call (one per lane, or
    one lane for the range-A FP32 run, whose all-vector refinement workspace
    is ~90 GB per q-point); handle creation (force-constant upload and
    symmetrisation) is not timed.  Reports q-pts/s per run, max |dw| of FP32
    against FP64, and the FP32 eigenvector residual |D z - lam z| / |D|_2 at K
    with D from the library's FP64 assembly."""
    from parphom_amd import synth
    t0 = time.perf_counter()
    st, fc, qs = synth.workload("tbg_30_31")
    n = 3 * st.natom
    log(f"[large] N={st.natom} n={n} nq={len(qs)} (inputs built in {time.perf_counter() - t0:.1f}s)")

    def run(prec, rng, q, streams):
        jobz = "V" if rng == "I" else "N"
        h = lib.Handle(st.lat, st.pos, st.types, st.masses, fc, jobz=jobz, rng=rng, il=1,
                       iu=nvec, vel=False, precision=prec, batch=1, streams=streams)
        ev, info, zk, t = [], [], None, 0.0
        for i in range(0, len(q), 2):
            t1 = time.perf_counter()
            r = h.run(q[i:i + 2])
            t += time.perf_counter() - t1
            ev += r["eval_cm"]
            info += [int(x) for x in r["info"]]
            if r["evec"] is not None and i <= 1 < i + 2:
                zk = r["evec"][1 - i]                      # vectors at K (q index 1)
            log(f"[large] {prec} range {rng}: q {i + 1}-{min(i + 2, len(q))} of {len(q)} "
                f"done ({t:.1f}s)")
        h.close()
        return len(q) / t, ev, info, zk

    qps_i32, ev_i32, inf_i32, zk = run("S", "I", qs, 2)
    qps_i64, ev_i64, inf_i64, _ = run("D", "I", qs, 2)
    q3 = qs[:3]
    qps_a32, ev_a32, inf_a32, _ = run("S", "A", q3, 1)
    qps_a64, ev_a64, inf_a64, _ = run("D", "A", q3, 2)
    dw_i = max(float(np.abs(x - y).max()) for x, y in zip(ev_i32, ev_i64))
    dw_a = max(float(np.abs(x - y).max()) for x, y in zip(ev_a32, ev_a64))
    # residual of the FP32 vectors at K against the FP64 D(K); |D|_2 = max |lambda|
    # from the FP64 range-A eigenvalues at K
    K = float(np.float32(15.633302)) * float(np.float32(33.35641))
    lam_all = np.sign(ev_a64[1]) * (ev_a64[1] / K) ** 2
    dnorm = float(np.abs(lam_all).max())
    h = lib.Handle(st.lat, st.pos, st.types, st.masses, fc, precision="D", batch=1)
    Dk = h.assemble(qs[1:2])[0]                        # column-major buffer, strided view
    h.close()
    dev = torch.device("cuda:0")
    Dg = torch.from_numpy(Dk).to(dev)
    del Dk
    Zg = torch.from_numpy(np.ascontiguousarray(zk)).to(dev)
    lam_k = torch.from_numpy(np.sign(ev_i32[1]) * (ev_i32[1] / K) ** 2).to(dev)
    R = Dg @ Zg - Zg * lam_k[None, :]
    resid = float((torch.linalg.vector_norm(R, dim=0) /
                   torch.linalg.vector_norm(Zg, dim=0)).max().item()) / dnorm
    del Dg, Zg, R
    torch.cuda.empty_cache()
    res = {"metric": f"q-pts/s at 3N={n}: lowest {nvec} eigenpairs, FP32+refine vs FP64",
           "value": qps_i32, "unit": "q-pts/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": 1e3 * len(qs) / qps_i32, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "c64 MFMA (FP32 acc) + f64 refinement",
           "data": "synthetic (seeded spring-model force constants; Gamma, K, M + 5 random q)",
           "config": {"workload": "tbg_30_31", "natom": st.natom, "n": n, "nq": len(qs),
                      "range": "I", "il": 1, "iu": nvec, "eigenvectors": True,
                      "group_velocity": False, "precision": "S", "batch": 1, "streams": 2,
                      "host_buffers": True},
           "fp64_path": {"value": qps_i64, "unit": "q-pts/s"},
           "speedup_vs_fp64": qps_i32 / qps_i64,
           "max_abs_domega_vs_fp64_cm": dw_i,
           "max_eigvec_residual_rel_at_K": resid,
           "range_A_values_only": {"nq": 3, "fp32_qps": qps_a32, "fp64_qps": qps_a64,
                                   "speedup_vs_fp64": qps_a32 / qps_a64,
                                   "max_abs_domega_vs_fp64_cm": dw_a},
           "info": {"I_S": inf_i32, "I_D": inf_i64, "A_S": inf_a32, "A_D": inf_a64}}
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--streams", type=int, default=3,
                    help="batches in flight per pph_run call (one stream each); the default "
                         "and the batch sizes are the measured best of both paths at n = 4764")
    ap.add_argument("--workload", default="tbg_11_12")
    ap.add_argument("--reduction", type=int, default=-1, choices=(-1, 0, 1, 2),
                    help="reduction to the tridiagonal: 0 one-stage, 1 two-stage, 2 the "
                         "one-launch small-matrix kernel (n <= 128 FP32, 96 FP64), "
                         "-1 the library's choice")
    ap.add_argument("--tile", default="64x64", choices=("64x64", "128x64", "64x128", "128x128"),
                    help="GEMM tile shape of the eigensolver (the autotuner's tile dimension)")
    ap.add_argument("--vec-refine", type=int, default=-1, choices=(-1, 0, 1, 2),
                    help="FP64 residual-correction steps of the FP32 path's vectors "
                         "(-1: the library default, 2)")
    ap.add_argument("--vec-bf16", type=int, default=-1, choices=(-1, 0, 1),
                    help="products of the first correction step for pairs more than 16 modes "
                         "apart: 1 BF16 matrix cores, 0 FP32 (-1: the library default, 1)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip timing the FP32 path with the other reduction")
    ap.add_argument("--fcsym", action="store_true",
                    help="measure the force-constant symmetrisation instead (see fcsym_bench)")
    ap.add_argument("--dos", action="store_true",
                    help="measure the linear-triangulation DOS kernel instead (see dos_bench)")
    ap.add_argument("--large", action="store_true",
                    help="run the large single-GPU workload instead (see large_bench)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device 3 times and exit (a copy of known "
                         "size, to scale the FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the phonon path has no multi-GPU mode "
              f"(q-points shard with no exchange); got --gpus {a.gpus}", file=sys.stderr)
        sys.exit(2)

    import torch
    if a.copy_calibration:
        src = torch.ones(a.copy_calibration // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 4, "copies": 3}))
        return
    from parphom_amd import lib, synth
    lib.load()
    if a.dos:
        dos_bench(lib, torch, a)
        return
    if a.fcsym:
        fcsym_bench(lib, torch, a)
        return
    if a.large:
        large_bench(lib, torch, a)
        return
    t0 = time.perf_counter()
    st, fc, qs = synth.workload(a.workload)
    n = 3 * st.natom
    log(f"[bench] workload {a.workload}: N={st.natom} n={n} nq={len(qs)} "
        f"(inputs built in {time.perf_counter() - t0:.1f}s)")

    B, S = a.batch, a.streams
    r32 = time_path(lib, torch, st, fc, qs, "S", B, a.steps, a.warmup, S, a.reduction, a.tile,
                    vec_refine=a.vec_refine, vec_bf16=a.vec_bf16)
    dt32, prof32, iso32, ev32, z32, nblk = (r32["dt"], r32["prof"], r32["iso"], r32["ev"],
                                            r32["z0"], r32["nblk"])
    two, small = r32["two_stage"], r32["small"]
    qps32 = a.steps * B * S / dt32
    res = {"metric": "q-pts/s at 3N=4764 eigpairs+v_g: FP32+refine vs FP64, max|dw|<=1e-3 cm^-1",
           "value": qps32, "unit": "q-pts/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": 1e3 * dt32 / a.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "c64 MFMA (FP32 acc) + f64 refinement",
           "data": "synthetic (seeded spring-model force constants, random q)",
           "config": {"workload": a.workload, "natom": st.natom, "n": n,
                      "q_per_step": B * S, "batch": B, "streams": S,
                      "range": "A", "eigenvectors": True, "group_velocity": True,
                      "precision": "S",
                      "reduction": ("two-stage (band 32)" if two else
                                    "small-matrix kernel" if small else "one-stage"),
                      "gemm_tile": a.tile,
                      "vector_correction_steps": a.vec_refine if a.vec_refine >= 0 else 2,
                      "vector_correction_first_step": (
                          "FP32 products" if a.vec_bf16 == 0 else
                          "BF16 matrix cores for pairs more than 16 modes apart, FP32 for the band"),
                      "batch_graphs_replayed_in_timed_steps": r32["graph_replays"]}}
    phases, hemv_launch_ms = split_profile(prof32, n)
    res["phase_ms_per_batch"] = {k: round(v, 3) for k, v in phases.items()}
    iso, iso_launch_ms = split_profile(iso32, n)
    res["phase_ms_one_batch_alone"] = {k: round(v, 3) for k, v in iso.items()}

    # roofline of the assembly phase (HBM-bound): algorithmic bytes = Phi
    # (8 n^2) + B * D in c64 for the reduction, and either D in c128 and dD in
    # c64 (dense products) or the four phase sums of D and dD on the nblk
    # nonzero blocks (4 c128 each, k_sparse_blocks)
    asm_bytes = 8 * n * n + B * 8 * n * n + (B * 4 * 16.0 * nblk if nblk > 0
                                             else B * (16 + 3 * 8) * n * n)
    # the assembly is one launch per batch; in the timed steps it shares the
    # GPU with the other lane's reduction, so it is rated on the batch run alone
    asm_s = iso["assemble"] / 1e3
    # roofline of the velocity forms (one launch per batch, 3 directions).
    # Dense products: diag(Z^H dD Z) with dD Hermitian is formed from its lower
    # triangle, 2 Re(Z^H A' Z), a triangular n x n times n x n complex product:
    # n^3 / 2 complex multiply-adds = 4 n^3 flops per direction, B * 12 n^3 per
    # launch, on the FP32 matrix cores.  Block-sparse forms (nblk lower 3x3
    # blocks, each a phase sum times a real block Phi~): per block and mode
    # c = z_a^H (Phi~ z_b), a real 3x3 block times a complex 3-vector (36
    # flops) and the dot with the row's 3-vector (24), then per operator
    # Re(p c) weighted and accumulated (4).  On the FP32 path one launch forms
    # the Rayleigh numerators and the three velocity forms (76 flops per block
    # and mode, in FP64 on the vector units): of the FP32 vectors, or, after
    # the FP64 correction of the vectors (the default), of the corrected FP64
    # vectors, in chunks of 8 matrices; its phase also holds the transposes
    # of the vectors to the mode-contiguous layout.
    vel_s = phases["velocity"] / 1e3
    vsteps = a.vec_refine if a.vec_refine >= 0 else 2
    if vel_s <= 0:
        vel_rl = None
    elif nblk > 0:
        vel_flops = B * n * 76.0 * nblk
        vel_rl = {"bound": "fp64 vector", "achieved": vel_flops / vel_s / 1e12,
                  "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                  "frac": vel_flops / vel_s / FP64_PEAK, "traffic": None,
                  "kernel": ("k_sparse_dots<double, 4> (Rayleigh numerators and velocity forms "
                             "of the corrected vectors" if vsteps > 0 else
                             "k_sparse_dots<float, 4> (Rayleigh numerators and velocity forms")
                            + " on %d nonzero force-constant blocks) + k_transpose" % nblk}
    else:
        vel_flops = B * 12.0 * n ** 3
        vel_rl = {"bound": "mfma", "achieved": vel_flops / vel_s / 1e12,
                  "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                  "frac": vel_flops / vel_s / FP32_PEAK, "traffic": None,
                  "kernel": "velocity products (k_zgemm dot epilogue, FP32)"}
    res["roofline_velocity"] = vel_rl
    res["config"]["fc_blocks_used"] = nblk
    res["roofline_assembly"] = {"bound": "hbm", "achieved": asm_bytes / asm_s / 1e9,
                                "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                "frac": asm_bytes / asm_s / HBM_PEAK, "traffic": None,
                                "kernel": "k_assemble" + (" + k_sparse_blocks" if nblk > 0 else "")}
    dominant = max(("assemble", "tridiag", "tri_eig", "back_transform", "refine", "velocity"),
                   key=lambda k: phases[k])
    res["dominant_phase"] = dominant
    if two:
        # dominant kernel of the two-stage path: k_q2_apply, the bulge-chasing
        # reflectors applied to all M = n eigenvectors (MFMA-bound).  Its
        # algorithmic work: every reflector of length l touches l x M entries
        # twice (a dot and an update); the n^2 / 2 reflector entries make
        # n^2 M complex multiply-adds = 8 n^2 M flops per matrix, B matrices
        # per launch (the blocked form's products on the zero part of each
        # block are not counted).  One launch per batch, timed with a HIP-event
        # pair by the library (slot "q2").
        # `roofline` rates the kernel with the GPU to itself (the batch run
        # alone: an event pair around its one launch); in the steps the
        # lanes' launches overlap almost completely (S launches of ~S times
        # the single duration), so the per-launch rate there
        # (`roofline_shared`) is the aggregate divided by S
        q2_flops = B * 8.0 * float(n) ** 3
        q2_ms, q2_cnt = prof32["q2"]
        q2_s = q2_ms / max(1, q2_cnt) / 1e3
        iso_q2 = iso32["q2"][0] / max(1, iso32["q2"][1]) / 1e3
        res["roofline"] = {"bound": "mfma", "achieved": q2_flops / iso_q2 / 1e12,
                           "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                           "frac": q2_flops / iso_q2 / FP32_PEAK,
                           "traffic": kernel_traffic("k_q2_apply", a.workload, n, B),
                           "kernel": "k_q2_apply<float, 32, 8, true>", "flops_per_launch": q2_flops,
                           "avg_launch_us": iso_q2 * 1e6,
                           "note": "one batch alone on the GPU (HIP-event pair on the lane's stream)"}
        res["roofline_shared"] = {
            "bound": "mfma", "achieved": q2_flops / q2_s / 1e12, "peak": FP32_PEAK / 1e12,
            "unit": "TFLOP/s", "frac": q2_flops / q2_s / FP32_PEAK,
            "kernel": "k_q2_apply<float, 32, 8, true>", "avg_launch_us": q2_s * 1e6,
            "aggregate_tflops": S * q2_flops / q2_s / 1e12,
            "note": f"profiled pass over the timed steps' q-points, {S} batches in flight sharing "
                    "the GPU; per-launch rate (aggregate = S x this while the launches overlap)"}
        # the band reduction's matrix-core work: the trailing updates of the
        # lower triangle, about 2 m^2 bw complex multiply-adds per panel (X = A
        # VT and the rank-2 bw update), i.e. 16/3 n^3 flops per matrix in total
        # (the textbook count of a Hermitian reduction); rated on one batch alone
        he_flops = B * 16.0 / 3.0 * float(n) ** 3
        he_s = iso["he2hb"] / 1e3
        res["roofline_band_reduction"] = {
            "bound": "mfma", "achieved": he_flops / he_s / 1e12, "peak": FP32_PEAK / 1e12,
            "unit": "TFLOP/s", "frac": he_flops / he_s / FP32_PEAK,
            "kernel": "he2hb (k_panel_qr + k_zgemm panel products), one batch alone"}
    elif small:
        # the one-launch reduction of small matrices (k_small_hetrd, one
        # workgroup per matrix, the matrix in LDS): 16/3 n^3 flops per matrix
        # (the textbook count of a Hermitian reduction) on the vector units, in
        # the working precision (FP32 here); one launch per batch, rated on the
        # batch run alone
        sm_flops = B * 16.0 / 3.0 * float(n) ** 3
        sm_s = iso["tridiag"] / 1e3
        res["roofline"] = {"bound": "fp32 vector", "achieved": sm_flops / sm_s / 1e12,
                           "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                           "frac": sm_flops / sm_s / FP32_PEAK, "traffic": None,
                           "kernel": "k_small_hetrd<float>", "flops_per_launch": sm_flops,
                           "avg_launch_us": sm_s * 1e6,
                           "note": "one batch alone on the GPU (HIP-event pair on the lane's stream)"}
    else:
        # roofline of the dominant kernel, the hemv of the one-stage reduction
        # (k_hemv_sym, one launch per column j, HBM-bound): algorithmic bytes
        # per launch = the lower triangle of the trailing matrix, m (m + 1) / 2
        # complex64 elements with m = n - 1 - j, plus the vector in and the
        # product out (2 m elements), for each of the B matrices.  The library
        # times the launches of columns j = 0, 8, 16, ... with HIP-event pairs
        # on its stream during the timed steps; bytes are averaged over the
        # same columns.
        m = (n - 1 - np.arange(0, n - 1, HEMV_TIME_STRIDE)).astype(np.float64)
        hemv_bytes = B * 8.0 * float(np.mean(m * (m + 1) / 2 + 2 * m))
        hemv_s = hemv_launch_ms / 1e3
        iso_s = iso_launch_ms / 1e3
        res["roofline"] = {"bound": "hbm", "achieved": hemv_bytes / iso_s / 1e9,
                           "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": hemv_bytes / iso_s / HBM_PEAK,
                           "traffic": kernel_traffic("k_hemv_sym", a.workload, n, B),
                           "kernel": "k_hemv_sym<float>", "bytes_per_launch": hemv_bytes,
                           "avg_launch_us": iso_launch_ms * 1e3,
                           "note": "one batch alone on the GPU (HIP-event pairs on the lane's stream)"}
        res["roofline_shared"] = {
            "bound": "hbm", "achieved": hemv_bytes / hemv_s / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": hemv_bytes / hemv_s / HBM_PEAK, "kernel": "k_hemv_sym<float>",
            "avg_launch_us": hemv_launch_ms * 1e3, "ms_per_batch": round(phases["hemv"], 3),
            "note": f"profiled pass over the timed steps' q-points, {S} batches in flight sharing the GPU"}

    if not a.no_fp64:
        # FP64 workspaces are larger: three quarters of the batch per stream
        # (at n = 4764 and three lanes, 24 per lane is the FP64 path's best:
        # 11.97 q-points/s against 11.09 at 16 x 2 and 11.65 at 32 x 2)
        B64 = max(1, (3 * B) // 4) if S > 1 else B
        # at most 6 timed steps (and 2 warm-up steps) of the FP64 path, so the
        # default run stays within a few minutes whatever --steps is
        steps64 = min(max(1, a.steps - 1), 6)
        r64 = time_path(lib, torch, st, fc, qs, "D", B64, steps64, min(a.warmup, 2), S,
                        a.reduction, a.tile)
        qps64 = steps64 * B64 * S / r64["dt"]
        k = min(B, B64)
        dw = float(np.abs(ev32[:k] - r64["ev"][:k]).max())
        ph64, _ = split_profile(r64["prof"], n)
        res["fp64_path"] = {"value": qps64, "unit": "q-pts/s", "batch": B64, "streams": S,
                            "reduction": ("two-stage (band 32)" if r64["two_stage"] else
                                          "small-matrix kernel" if r64["small"] else "one-stage"),
                            "batch_graphs_replayed_in_timed_steps": r64["graph_replays"],
                            "phase_ms_per_batch": {kk: round(v, 3) for kk, v in ph64.items()}}
        res["speedup_vs_fp64"] = qps32 / qps64
        res["max_abs_domega_vs_fp64_cm"] = dw
        # the FP32 path's velocities against the FP64 path's on the same q,
        # through sums over degenerate clusters (gauge-invariant), all modes
        dv = [velocity_difference(r32["vel"][i], r64["vel"][i], r64["ev"][i]) for i in range(k)]
        res["max_abs_dvel_vs_fp64_m_per_s"] = max(x[0] for x in dv)
        # relative to the largest velocity over the compared q-points (at
        # Gamma every velocity is zero)
        res["max_rel_dvel_vs_fp64"] = max(x[0] for x in dv) / max(1e-300, max(x[1] for x in dv))
        # per mode: only modes closer than 1e-9 max|omega^2| (degenerate to
        # FP64 precision) are summed
        dvm = [velocity_difference(r32["vel"][i], r64["vel"][i], r64["ev"][i], rel=1e-9)
               for i in range(k)]
        res["max_rel_dvel_vs_fp64_per_mode"] = (max(x[0] for x in dvm) /
                                               max(1e-300, max(x[1] for x in dvm)))
        # whole-path roofline of both paths, against the spec-sheet matrix peaks
        # and against what the library GEMM reaches on this device
        p32, p64 = gemm_peak(torch, torch.float32), gemm_peak(torch, torch.float64)
        res["roofline_path"] = {
            "S": path_roofline(st.natom, n, B, qps32, FP32_PEAK, p32),
            "D": path_roofline(st.natom, n, B64, qps64, FP64_PEAK, p64),
            "gemm_peak_measured_tflops": {"fp32": p32 / 1e12, "fp64": p64 / 1e12},
            "gemm_peak_spec_tflops": {"fp32": FP32_PEAK / 1e12, "fp64": FP64_PEAK / 1e12}}
        if not a.no_compare:
            # the FP32 path with the other reduction, same batches and lanes
            other = 0 if (two or small) else 1
            steps_o = min(a.steps, 4)
            ro = time_path(lib, torch, st, fc, qs, "S", B, steps_o, min(a.warmup, 1), S, other,
                           a.tile)
            pho, _ = split_profile(ro["iso"], n)
            dvo = [velocity_difference(ro["vel"][i], r64["vel"][i], r64["ev"][i]) for i in range(k)]
            # the FP32 path without the FP64 correction of its vectors
            steps_v = min(a.steps, 4)
            rv = time_path(lib, torch, st, fc, qs, "S", B, steps_v, min(a.warmup, 1), S,
                           a.reduction, a.tile, vec_refine=0)
            dvv = [velocity_difference(rv["vel"][i], r64["vel"][i], r64["ev"][i], rel=1e-9)
                   for i in range(k)]
            res["fp32_without_vector_correction"] = {
                "value": steps_v * B * S / rv["dt"], "unit": "q-pts/s",
                "max_abs_domega_vs_fp64_cm": float(np.abs(rv["ev"][:k] - r64["ev"][:k]).max()),
                "max_rel_dvel_vs_fp64_per_mode": (max(x[0] for x in dvv) /
                                                  max(1e-300, max(x[1] for x in dvv)))}
            res["fp32_other_reduction"] = {
                "reduction": ("two-stage (band 32)" if ro["two_stage"] else
                              "small-matrix kernel" if ro["small"] else "one-stage"),
                "value": steps_o * B * S / ro["dt"], "unit": "q-pts/s",
                "max_abs_domega_vs_fp64_cm": float(np.abs(ro["ev"][:k] - r64["ev"][:k]).max()),
                "max_rel_dvel_vs_fp64": max(x[0] for x in dvo) / max(1e-300, max(x[1] for x in dvo)),
                "phase_ms_one_batch_alone": {kk: round(v, 3) for kk, v in pho.items()}}

    if not a.no_cpu:
        ranks = max(1, min(16, os.cpu_count() or 1))
        sl = scalapack_baseline(st, fc, qs[0], ranks=ranks)
        if sl is not None:
            # the oracle's eigenvalues for the accuracy checks below (values only)
            from oracle import phonon as O
            r = O.System(st.lat, st.pos, st.types, st.masses, fc).solve(qs[0], jobz="N")
            res["cpu_baseline"] = {
                "value": 1.0 / sl["s_per_q"], "unit": "q-pts/s", "cores": sl["ranks"],
                "kind": "port",
                "sample": f"1 q-point of {a.workload}: numpy assembly of D and dD "
                          f"({sl['asm_s_per_q']:.1f} s), MKL ScaLAPACK pzheevx('V','A','U') "
                          f"({sl['eig_s_per_q']:.1f} s) and 6 pzgemm for the velocities "
                          f"({sl['vel_s_per_q']:.1f} s), {sl['ranks']} MPI ranks x 1 thread "
                          f"on a {sl['nprow']} x {sl['npcol']} grid, blocks of {sl['nb']}"}
            lam0 = float(sl["lowest_lambda"][0])
            res["scalapack_lowest_lambda_vs_oracle"] = abs(lam0 - float(r["lam"][0]))
        else:
            # no MPI / ScaLAPACK: the numpy/LAPACK oracle is the baseline
            cdt, r = cpu_baseline(st, fc, qs[0])
            res["cpu_baseline"] = {"value": 1.0 / cdt, "unit": "q-pts/s", "cores": os.cpu_count(),
                                   "kind": "port",
                                   "sample": f"1 q-point of {a.workload} (assembly + zheevx all "
                                             f"vectors + 3 velocity GEMMs, numpy/OpenBLAS)"}
        res["max_abs_domega_vs_oracle_cm"] = float(np.abs(ev32[0] - r["eval_cm"]).max())
        # eigenvector residuals of the FP32 path against the oracle's FP64 D(q0):
        # |D z - lambda z| / |D|_2 for 256 sampled modes, lambda from the
        # reported frequency (sign(w) (w / K)^2, K = f32(15.633302) f32(33.35641))
        from oracle import phonon as O
        Ssys = O.System(st.lat, st.pos, st.types, st.masses, fc)
        D0 = Ssys.D(qs[0])
        K = float(np.float32(15.633302)) * float(np.float32(33.35641))
        pick = np.random.default_rng(0).choice(n, 256, replace=False)
        Zs = z32[:, pick]
        lam = np.sign(ev32[0][pick]) * (ev32[0][pick] / K) ** 2
        rr = np.linalg.norm(D0 @ Zs - Zs * lam[None, :], axis=0) / np.linalg.norm(Zs, axis=0)
        res["max_eigvec_residual_rel"] = float(rr.max() / np.abs(r["lam"]).max())
    print(json.dumps(res))


if __name__ == "__main__":
    main()
