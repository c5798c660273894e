"""The CPU baseline of bench.py (MKL ScaLAPACK pzheevx and pzgemm under
MPICH, oracle/_bin/scalapack_baseline): its eigenvalues against LAPACK on a
small random Hermitian matrix, on a 2 x 2 process grid with ragged blocks.
Runs on the CPU; skipped where the image has no MPI or ScaLAPACK."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_bin", "scalapack_baseline")
MPIEXEC = "/opt/conda/bin/mpiexec"


@pytest.fixture(scope="module")
def baseline():
    if not (os.path.exists(MPIEXEC) and os.path.exists("/opt/conda/include/mpi.h")):
        pytest.skip("no MPICH / MKL ScaLAPACK in this image")
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "baseline"],
                       capture_output=True, text=True)
    if r.returncode != 0 or not os.path.exists(EXE):
        pytest.skip("the ScaLAPACK baseline does not build here: " + r.stderr[-300:])
    return EXE


def test_scalapack_baseline_eigenvalues(baseline, tmp_path):
    rng = np.random.default_rng(12)
    n, nq = 77, 2                      # 77 = two full blocks of 32 and a ragged 13
    path = tmp_path / "mats.bin"
    lows = []
    with open(path, "wb") as f:
        for _ in range(nq):
            for k in range(4):
                X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
                H = X + X.conj().T
                if k == 0:
                    lows.append(np.linalg.eigvalsh(H)[0])
                f.write(np.asfortranarray(H).tobytes(order="F"))
    env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")
    r = subprocess.run([MPIEXEC, "-n", "4", baseline, str(path), str(n), str(nq), "32"],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr[-500:]
    out = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
    assert out["nprow"] == 2 and out["npcol"] == 2 and out["n"] == n
    assert np.allclose(out["lowest_lambda"], lows, rtol=0, atol=1e-10)
    assert out["eig_s_per_q"] > 0 and out["vel_s_per_q"] > 0
