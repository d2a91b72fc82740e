"""Headline benchmark: FP64 neo-Hookean tangent vmult (tensor2 caching,
formulation 1, Newmark mass term) on the Q4 beam [0,4]x[0,1]^2 with
128x32x32 cells (25,610,499 DoFs, 16,384,000 quadrature points).

One "step" is one operator application dst = A src on resident device
vectors.  Prints one JSON line with GDoF/s, the roofline of the dominant
(cell) kernel from HIP events on its own stream, and the CPU baseline (the
oracle, on a bounded sample, timed on this host).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "matrix-free-dealii-precice_amd"))

E_MOD, NU, RHO, BETA, DT = 1.0e6, 0.3, 1000.0, 0.25, 1.0e-3
MU = E_MOD / (2.0 * (1.0 + NU))
ALPHA1 = 1.0 / (BETA * DT * DT)
HBM_PEAK_GBS = 8000.0

BEAM = dict(lo=[0.0, 0.0, 0.0], hi=[4.0, 1.0, 1.0], cells=[128, 32, 32], degree=4)


def canonical_bytes(n_dof, n_cell, p, s=8, words_per_qp=18):
    """s*2*N_dof (src read + dst write) + 4*N_cell*(p+1)^3 (int32 map) + s*18*N_qp."""
    nq = n_cell * (p + 1) ** 3
    return s * 2 * n_dof + 4 * nq + s * words_per_qp * nq


def cpu_baseline(seconds_budget=20.0):
    """Oracle vmult (same operator, same inputs) on a quarter of the beam:
    [0,1]^3 with 32^3 Q4 cells, 6,440,067 DoFs; per-cell work identical."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    m = O.Mesh([0, 0, 0], [1, 1, 1], [32, 32, 32], 4, 3, 1, 7)
    u = m.u0(amp=1e-3, noise=1e-5, seed=0)
    op = O.Solid(m, caching=1)
    op.cache(u)
    x = O.fill_uniform(1, m.n_dofs)
    op.vmult(x)  # warm-up
    times = []
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < seconds_budget and len(times) < 20:
        t0 = time.perf_counter()
        op.vmult(x)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": round(m.n_dofs / t / 1e9, 4), "unit": "GDoF/s", "cores": threads, "kind": "port",
            "sample": f"oracle tensor2 FP64 vmult, [0,1]^3 32^3 Q4 cells ({m.n_dofs} DoFs, same per-cell work "
                      f"as the 128x32x32 beam), median of {len(times)} vmults, {threads} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--degree", type=int, default=BEAM["degree"])
    ap.add_argument("--cells", type=int, nargs=3, default=BEAM["cells"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU decomposition is design-only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import mfgpu

    assert torch.cuda.is_available(), "bench.py needs the GPU"
    dev = torch.device("cuda:0")
    p = args.degree
    mesh = mfgpu.BoxMesh(BEAM["lo"], BEAM["hi"], args.cells, p, 3, 1, 7)
    n_dof, n_cell = mesh.n_dofs, mesh.n_cells
    op = mfgpu.NeoHookOperator(mesh, MU, NU, RHO, ALPHA1, formulation=1, caching="tensor2")
    u0 = torch.from_numpy(mesh.u0(amp=1e-3, noise=1e-5, seed=0)).to(dev)
    op.cache(u0)
    del u0
    src = mfgpu.fill_uniform(1, torch.empty(n_dof, dtype=torch.float64, device=dev))
    dst = torch.empty_like(src)

    for _ in range(args.warmup):
        op.vmult(dst, src)
    torch.cuda.synchronize()
    op.enable_timing(True)
    op.timing()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.vmult(dst, src)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kern_ms, launches = op.timing()
    op.enable_timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    kern_avg_ms = kern_ms / max(launches, 1)
    gdofs = n_dof / (ms_per_step * 1e-3) / 1e9
    B = canonical_bytes(n_dof, n_cell, p)
    achieved = B / (kern_avg_ms * 1e-3) / 1e9

    # per-vmult latency distribution (events around each call)
    lat = []
    for _ in range(min(args.steps, 50)):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        op.vmult(dst, src)
        e1.record()
        e1.synchronize()
        lat.append(e0.elapsed_time(e1))
    out = {
        "metric": "GDoF/s tangent vmult (FP64 neo-Hookean tensor2, Q4 25.6M-DoF beam)",
        "value": round(gdofs, 3),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"tangent vmult, beam [0,4]x[0,1]^2, {args.cells[0]}x{args.cells[1]}x{args.cells[2]} "
                               f"cells, FE_Q({p})^3, QGauss({p + 1}), {n_dof} DoFs, u=0 on x=0, "
                               "linearised at 1e-3 sin sin sin + 1e-5 noise",
                   "n_dofs": n_dof, "n_cells": n_cell, "degree": p, "caching": "tensor2", "formulation": 1},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "apply_cells<4,2,double,QSolidT2F1>", "kernel_avg_ms": round(kern_avg_ms, 4),
                     "canonical_bytes_per_launch": B},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
