"""Autotuner for the tile shapes, blocking and batching of the phonon pipeline.

The library's tunables on a given structure and device are the tile shape of
the eigensolver's matrix-core GEMMs (``tile``: 64x64, 128x64, 64x128, and
128x128 on the FP32 path), the reduction (one-stage, two-stage, the one-launch
small-matrix kernel), the reflectors per reduction panel (``nb_panel``: 16,
32, 48, 64), the reflectors per back-transformation block (``nb_back``:
multiples of 16 up to 128), the q-points per device batch (``batch``) and the
batches in flight (``streams``, one stream and host thread each).  ``tune``
times every combination of the given candidates on the caller's own structure
and q-points with the HIP path and returns the fastest, with the whole table;
``apply`` sets a found configuration on a handle.  The measured defaults of
the library (64x64, 32, 128, 32, 2) are the best found this way for twisted
bilayer graphene at n = 4764 without tile search; other sizes can prefer
other values (smaller cells want larger batches).

    from parphom_amd import tune, synth
    st, fc, q = synth.workload("tbg_8_9")
    best = tune.tune(st.lat, st.pos, st.types, st.masses, fc, q)
    # best["config"] -> {"batch": ..., "streams": ..., "reduction": ..., "tile": ...,
    #                    "nb_panel": ..., "nb_back": ...}

or ``python -m parphom_amd.tune --workload tbg_8_9``.
"""
from __future__ import annotations

import argparse
import itertools
import json
import time

import numpy as np


def _default_handle(lat, pos, types, masses, fc, **opts):
    from parphom_amd import lib
    return lib.Handle(lat, pos, types, masses, fc, **opts)


def _device_outputs(nq, n):
    import torch
    dev = torch.device("cuda:0")
    return {"eval_cm": torch.zeros((nq, n), dtype=torch.float64, device=dev),
            "m": torch.zeros(nq, dtype=torch.int64, device=dev),
            "nz": torch.zeros(nq, dtype=torch.int64, device=dev),
            "info": torch.zeros(nq, dtype=torch.int32, device=dev),
            "evec": torch.zeros((nq, n, n), dtype=torch.complex128, device=dev),
            "vel": torch.zeros((nq, n, 3), dtype=torch.float64, device=dev)}


def tune(lat, pos, types, masses, fc, q, precision="S", jobz="V", vel=True,
         batches=(16, 32), streams=(1, 2), panels=(32, 48), backs=(64, 128), reductions=(-1,),
         tiles=("64x64",), reps=1, device_outputs=False, handle_factory=None,
         clock=time.perf_counter, log=None):
    """Time every (batch, streams, reduction, tile, nb_panel, nb_back)
    candidate and return {"config": best, "qps": best rate, "table": [rows]}.
    Each candidate runs one untimed warm-up call and then `reps` timed calls
    of pph_run over batch * streams q-points (taken cyclically from q): with
    host outputs, or with outputs left in device buffers (pph_run_device,
    torch tensors) when device_outputs is set.  reduction: -1 the library's
    choice, 0 one-stage, 1 two-stage, 2 the small-matrix kernel; tile: a
    GEMM tile shape name (lib.GEMM_TILES).  A candidate that fails (a blocking
    or tile the library rejects, a batch that does not fit in memory, any
    library error) is recorded with qps None and its error, and the search
    goes on."""
    make = handle_factory or _default_handle
    q = np.atleast_2d(np.asarray(q, np.float64))
    table = []
    best = None

    def sync():
        if device_outputs:
            import torch
            torch.cuda.synchronize()

    for B, S in itertools.product(batches, streams):
        per_call = B * S
        idx = np.arange(per_call) % len(q)
        qs = q[idx]
        rows = [{"batch": B, "streams": S, "reduction": r, "tile": t, "nb_panel": nbp,
                 "nb_back": nbb, "qps": None}
                for r, t, nbp, nbb in itertools.product(reductions, tiles, panels, backs)]
        try:
            out = _device_outputs(per_call, 3 * len(pos)) if device_outputs else None
            h = make(lat, pos, types, masses, fc, jobz=jobz, vel=vel, precision=precision,
                     batch=B, streams=S)
        except Exception as e:                 # e.g. the batch does not fit in memory
            for row in rows:
                row["error"] = str(e)
            table.extend(rows)
            continue
        call = (lambda h: h.run_device(qs, out)) if device_outputs else (lambda h: h.run(qs))
        with h:
            for row in rows:
                table.append(row)
                try:
                    apply(h, row)
                    call(h)                    # warm-up (allocations, first-use set-up)
                    sync()
                    t0 = clock()
                    for _ in range(reps):
                        call(h)
                    sync()
                    dt = clock() - t0
                except Exception as e:         # keep the row, go on with the next candidate
                    row["error"] = str(e)
                    continue
                row["qps"] = reps * per_call / dt if dt > 0 else float("inf")
                if log:
                    log(f"[tune] batch {B} streams {S} reduction {row['reduction']} "
                        f"tile {row['tile']} panel {row['nb_panel']} back {row['nb_back']}: "
                        f"{row['qps']:.3f} q-pts/s")
                if best is None or row["qps"] > best["qps"]:
                    best = row
        del out
    if best is None:
        raise RuntimeError("tune: no candidate ran")
    cfg = {k: best[k] for k in ("batch", "streams", "reduction", "tile", "nb_panel", "nb_back")}
    return {"config": cfg, "qps": best["qps"], "table": table}


def apply(h, cfg):
    """Set a configuration found by tune (its "config", or a table row) on a
    handle: reduction, GEMM tile shape and blocking.  The batch and streams
    are creation options (lib.Handle(..., batch=, streams=))."""
    if cfg.get("reduction", -1) != -1 and hasattr(h, "reduction"):
        h.reduction(cfg["reduction"])
    if cfg.get("tile") is not None and hasattr(h, "gemm_tile"):
        h.gemm_tile(cfg["tile"])
    h.blocking(cfg["nb_panel"], cfg["nb_back"])


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="tbg_8_9")
    ap.add_argument("--precision", default="S")
    ap.add_argument("--batches", default="16,32")
    ap.add_argument("--streams", default="1,2")
    ap.add_argument("--panels", default="32,48")
    ap.add_argument("--backs", default="64,128")
    ap.add_argument("--reductions", default="-1",
                    help="reductions to try: -1 the library's choice, 0 one-stage, 1 two-stage, "
                         "2 the small-matrix kernel")
    ap.add_argument("--tiles", default="64x64",
                    help="GEMM tile shapes to try (64x64, 128x64, 64x128, 128x128)")
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--device-outputs", action="store_true",
                    help="leave outputs in device buffers (pph_run_device) instead of host arrays")
    a = ap.parse_args(argv)
    from parphom_amd import synth
    st, fc, q = synth.workload(a.workload)
    ints = lambda s: tuple(int(x) for x in s.split(","))   # noqa: E731
    res = tune(st.lat, st.pos, st.types, st.masses, fc, q, precision=a.precision,
               batches=ints(a.batches), streams=ints(a.streams), panels=ints(a.panels),
               backs=ints(a.backs), reductions=ints(a.reductions),
               tiles=tuple(a.tiles.split(",")), reps=a.reps,
               device_outputs=a.device_outputs,
               log=lambda m: print(m, flush=True))
    res["workload"] = a.workload
    print(json.dumps(res))


if __name__ == "__main__":
    main()
