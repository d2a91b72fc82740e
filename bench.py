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
# node numbering of every bench mesh (the vector layout in HBM); "cellwise" gives each cell a contiguous block
NUMBERING = "cellwise"

# HBM bytes per vmult of the cell kernel from two rocprofv3 --pmc passes (FETCH_SIZE x 2.00 and WRITE_SIZE
# x 1.00, factors from a copy of known size), with the SHA-256 prefix of the kernel source they were measured
# on.  PMC counters cannot be read from inside this process: when the source differs, the figure no longer
# describes the kernel and the bench reports no traffic until it is re-measured.
TRAFFIC_SOURCE = "profiles/vmult_q4_beam_fp64_pmc_traffic.md"
KERNEL_SOURCE = os.path.join(ROOT, "matrix-free-dealii-precice_amd", "csrc", "cell_kernels.hpp")
MEASURED_TRAFFIC = {((128, 32, 32), 4, "cellwise", "gradient"): (2633974579, "f2c744af6aeb")}


def kernel_source_hash():
    import hashlib
    with open(KERNEL_SOURCE, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:12]


def measured_traffic(key):
    """(bytes, note) for the bench's operator, bytes None when the kernel changed since the measurement."""
    if key not in MEASURED_TRAFFIC:
        return None, "not measured for this configuration"
    b, h = MEASURED_TRAFFIC[key]
    now = kernel_source_hash()
    if h != now:
        return None, f"the kernel source ({now}) differs from the profiled one ({h}): re-measure"
    return b, f"{TRAFFIC_SOURCE}, kernel source sha256 {h}"


def log(msg):
    """Progress on stderr (stdout carries the one JSON line)."""
    print(msg, file=sys.stderr, flush=True)


def canonical_bytes(n_dof, n_cell, p, s=8, words_per_qp=18):
    """s*2*N_dof (src read + dst write) + 4*N_cell*(p+1)^3 (int32 map) + s*18*N_qp."""
    nq = n_cell * (p + 1) ** 3
    return s * 2 * n_dof + 4 * nq + s * words_per_qp * nq


def cpu_baseline(seconds_budget=20.0):
    """CPU restatement of the reference's vmult (same operator, same inputs),
    computed the way the reference's CPU path is: cells in batches of 8 in
    SIMD lanes (VectorizedArray<double>, AVX-512 where the host has it),
    collocation sum factorisation, one OpenMP thread per core over the batches
    of a colour.  Sample: a quarter of the beam's cells, [0,1]^3 with 32^3 Q4
    cells (6,440,067 DoFs), the same per-cell work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    threads = int(os.environ["OMP_NUM_THREADS"])
    m = O.Mesh([0, 0, 0], [1, 1, 1], [32, 32, 32], 4, 3, 1, 7)
    u = m.u0(amp=1e-3, noise=1e-5, seed=0)
    op = O.Solid(m, caching=1)
    op.cache(u)
    fast = O.SimdSolid(op)
    x = O.fill_uniform(1, m.n_dofs)
    y = fast.vmult(x)  # warm-up
    times = []
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < seconds_budget and len(times) < 100:
        t0 = time.perf_counter()
        fast.vmult(x, y)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    simd = f"{fast.isa}, {fast.lanes} doubles per batch"
    return {"value": round(m.n_dofs / t / 1e9, 4), "unit": "GDoF/s", "cores": threads, "kind": "port",
            "label": "CPU restatement of the reference (cross-cell SIMD batches of 8 cells like VectorizedArray, "
                     "collocation sum factorisation, one OpenMP thread per core over coloured batches)",
            "simd": simd, "cpu_model": model,
            "sample": f"tensor2 FP64 vmult, [0,1]^3 32^3 Q4 cells ({m.n_dofs} DoFs, same per-cell work as the "
                      f"128x32x32 beam), median of {len(times)} vmults, {threads} OpenMP threads, {simd}"}


def mgcg_bench(mfgpu, torch, cells, p, level_precision, numbering, x_true_seed=2, storage="folded"):
    """MG-CG on the beam hierarchy (coarse 4x1x1 refined to `cells`): rhs =
    A x_true with x_true from seed 2 (constrained entries zeroed), CG to
    ||r||/||b|| <= 1e-8, Chebyshev degree 5 (range 20, 30 eigenvalue CG steps),
    coarse level FP64 CG to 1e-3.  level_precision FP64: plain FP64;
    FP32: FP32 V-cycle and level caches, FP64 outer CG and coarse solve."""
    n_levels = 1
    while (cells[0] >> n_levels) >= 4 and (cells[1] >> n_levels) >= 1 and (cells[2] >> n_levels) >= 1 and \
            (cells[0] >> n_levels) << n_levels == cells[0]:
        n_levels += 1
    coarse = [c >> (n_levels - 1) for c in cells]
    meshes = [mfgpu.BoxMesh(BEAM["lo"], BEAM["hi"], [c << l for c in coarse], p, 3, 1, 7, numbering=numbering)
              for l in range(n_levels)]
    dev = torch.device("cuda:0")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    precs = [mfgpu.FP64] + [level_precision] * (n_levels - 1)
    ops = [mfgpu.NeoHookOperator(m, MU, NU, RHO, ALPHA1, precision=pr) for m, pr in zip(meshes, precs)]
    fine64 = ops[-1] if level_precision == mfgpu.FP64 else mfgpu.NeoHookOperator(meshes[-1], MU, NU, RHO, ALPHA1)
    for op in ops + [fine64]:  # every level takes the bench's point storage
        op.set_point_storage(mfgpu.T2_GRADIENT if storage == "gradient" else mfgpu.T2_FOLDED)
    trs = [mfgpu.Transfer(ops[l - 1], ops[l], mfgpu.box_children(meshes[l - 1], meshes[l])) for l in range(1, n_levels)]
    u = torch.from_numpy(meshes[-1].u0(amp=1e-3, noise=1e-5, seed=0)).to(dev)
    fine64.cache(u)
    ul = [None] * n_levels
    ul[-1] = u.to(ops[-1].dtype)
    for l in range(n_levels - 1, 0, -1):
        ul[l - 1] = torch.empty(meshes[l - 1].n_dofs, dtype=ops[l - 1].dtype, device=dev)
        trs[l - 1].interpolate(ul[l - 1], ul[l])
    for op, uu in zip(ops, ul):
        if op is not fine64:
            op.cache(uu)
    mg = mfgpu.Multigrid(ops, trs, cheb_degree=5, smoothing_range=20.0, eig_cg_iterations=30, coarse_kind=1,
                         coarse_tolerance=1e-3)
    mg.setup()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    n = meshes[-1].n_dofs
    xt = mfgpu.fill_uniform(x_true_seed, torch.empty(n, dtype=torch.float64, device=dev))
    xt[torch.from_numpy(meshes[-1].constrained()).long().to(dev)] = 0.0
    b = torch.empty_like(xt)
    fine64.vmult(b, xt)
    x = torch.zeros_like(xt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = mfgpu.cg_solve(fine64, b, x, mfgpu.PRECOND_MG, mg, tol_rel=1e-8, max_iter=500)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    err = float(torch.linalg.norm(x - xt) / torch.linalg.norm(xt))
    # operator applications per level in the same solve (counted in a second,
    # untimed solve): sum of their canonical bytes over the solve time
    all_ops = ops if fine64 is ops[-1] else ops + [fine64]
    for op in all_ops:
        op.enable_timing(True)
        op.timing()
    x2 = torch.zeros_like(xt)
    mfgpu.cg_solve(fine64, b, x2, mfgpu.PRECOND_MG, mg, tol_rel=1e-8, max_iter=500)
    counts, bytes_total = [], 0
    for op, m in zip(all_ops, meshes + [meshes[-1]]):
        _, k = op.timing()
        op.enable_timing(False)
        es = 8 if op.dtype == torch.float64 else 4
        nqp = m.n_cells * (p + 1) ** 3
        bytes_total += k * (es * 2 * m.n_dofs + 4 * nqp + 18 * es * nqp)
        counts.append(k)
    del x2
    return {"levels": n_levels, "iterations": info.iterations, "solve_s": round(t_solve, 4),
            "setup_s": round(t_setup, 3), "rel_residual": info.final_residual / float(torch.linalg.norm(b)),
            "err_vs_x_true": err, "vmults_per_level": counts,
            "achieved_gbs_vmults": round(bytes_total / t_solve / 1e9, 1),
            "achieved_note": "canonical bytes of the operator applications only (no vector, transfer or "
                             "coarse-solver bytes) / solve time; vmults_per_level lists levels 0.. then the "
                             "FP64 outer operator when the V-cycle is FP32"}, x


def batched_instances(mfgpu, n_inst=48, seed_base=3, coarse_eig_iterations=100, n_range=(4, 24)):
    """Instance k (seed 3+k): k < n/2 heat (M + dt c K, dt 0.1, c ~ U[0.5, 2], u = 0 on the boundary),
    else solid (neo-Hookean tensor2, u = 0 on x = 0); p ~ U{1..6}, n_q = p+1; unit cube with
    n_x, n_y, n_z ~ U{n_range}.  FP64 fine operator, FP32 MG (levels halve while all counts are even;
    Chebyshev degree 5 on levels > 0, Chebyshev solver on level 0)."""
    import numpy as np
    out = []
    for k in range(n_inst):
        r = np.random.default_rng(seed_base + k)
        kind = "heat" if k < n_inst // 2 else "solid"
        p = int(r.integers(1, 7))
        n = [int(x) for x in r.integers(n_range[0], n_range[1] + 1, size=3)]
        c = float(r.uniform(0.5, 2.0))
        # each instance on its default cell schedule (colours, or the gather for small operators): both sum a
        # node's cell contributions in colour order, as the batched kernels do, so results are bitwise equal
        out.append(mfgpu.Instance(kind, n, p, fine_precision=mfgpu.FP64, level_precision=mfgpu.FP32, dt=0.1,
                                  conductivity=c, coarse_eig_iterations=coarse_eig_iterations, seed=seed_base + k))
    return out


def batched_bench(mfgpu, torch, reps=20, n_range=(4, 24)):
    """Batched vmult (one launch for all 48 FP64 fine operators) and batched
    MG-CG (FP64 outer CG, FP32 V-cycle as a replayed hipGraph), each against
    the same work done instance by instance on the same GPU.  The batched
    solve is timed in steady state (third solve: every V-cycle program of the
    solve already captured)."""
    import numpy as np
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    inst = batched_instances(mfgpu, n_range=n_range)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    log(f"batched: {len(inst)} instances set up in {t_setup:.1f} s")
    n_dofs = sum(i.n_dofs for i in inst)
    srcs = [mfgpu.fill_uniform(500 + k, torch.empty(i.n_dofs, dtype=torch.float64, device=dev))
            for k, i in enumerate(inst)]
    dsts = [torch.empty_like(s) for s in srcs]
    fine = [i.fine for i in inst]

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / n

    t_seq_vmult = timed(lambda: [op.vmult(d, s) for op, d, s in zip(fine, dsts, srcs)], reps)
    batch = mfgpu.Batch(fine)
    t_tune0 = time.perf_counter()
    tuning = batch.tune(srcs, dsts, repetitions=5).as_dict()
    t_tune = time.perf_counter() - t_tune0
    t_batch_vmult = timed(lambda: batch.vmult(dsts, srcs), reps)
    log(f"batched: vmult {t_batch_vmult * 1e3:.3f} ms batched, {t_seq_vmult * 1e3:.3f} ms sequential")
    # canonical bytes: src + dst, int32 map, 18 (solid) / 11 (heat) words per quadrature point
    B = sum(8 * 2 * i.n_dofs + (4 + 8 * (18 if i.kind == "solid" else 11)) * i.fine.n_qp for i in inst)
    # MG-CG: rhs = A x_true (seeds 600+k, constrained entries zeroed), ||r|| <= 1e-8 ||b||
    bs, xts = [], []
    for k, i in enumerate(inst):
        xt = mfgpu.fill_uniform(600 + k, torch.empty(i.n_dofs, dtype=torch.float64, device=dev))
        xt[torch.from_numpy(i.constrained).long().to(dev)] = 0.0
        b = torch.empty_like(xt)
        i.fine.vmult(b, xt)
        bs.append(b)
        xts.append(xt)
    solver = mfgpu.BatchSolver(fine, [i.mg for i in inst])
    fine_t, level_t = solver.tune(repetitions=3)
    log("batched: solver tuned")
    for _ in range(2):  # programs run eagerly on first use and are captured on the second
        xs = [torch.zeros_like(b) for b in bs]
        solver.solve(bs, xs, tol_rel=1e-8)
    log("batched: warm-up solves done")
    xs = [torch.zeros_like(b) for b in bs]
    torch.cuda.synchronize()
    t = time.perf_counter()
    infos, status = solver.solve(bs, xs, tol_rel=1e-8)
    torch.cuda.synchronize()
    t_batch_solve = time.perf_counter() - t
    log(f"batched: solve {t_batch_solve:.3f} s")
    launches, nodes = solver.info()
    xq = [torch.zeros_like(b) for b in bs]
    torch.cuda.synchronize()
    t = time.perf_counter()
    seq_its = []
    for i, b, x in zip(inst, bs, xq):
        seq_its.append(mfgpu.cg_solve(i.fine, b, x, mfgpu.PRECOND_MG, i.mg, tol_rel=1e-8).iterations)
    torch.cuda.synchronize()
    t_seq_solve = time.perf_counter() - t
    same = all(torch.equal(a, b) for a, b in zip(xs, xq))
    err = max(float(torch.linalg.norm(x - xt) / torch.linalg.norm(xt)) for x, xt in zip(xs, xts))
    its = [i.iterations for i in infos]
    return {
        "workload": "48 instances (seed 3+k): 24 heat (M + 0.1 c K, c ~ U[0.5,2]) + 24 neo-Hookean solids, "
                    f"p ~ U{{1..6}}, unit cube n ~ U{{{n_range[0]}..{n_range[1]}}}^3, FP64 fine operators, FP32 MG",
        "n_instances": len(inst), "n_dofs_total": int(n_dofs), "setup_s": round(t_setup, 2),
        "levels": [len(i.meshes) for i in inst],
        "vmult": {"batched_ms": round(t_batch_vmult * 1e3, 4), "sequential_ms": round(t_seq_vmult * 1e3, 4),
                  "speedup": round(t_seq_vmult / t_batch_vmult, 3),
                  "batched_gdofs": round(n_dofs / t_batch_vmult / 1e9, 3),
                  "batched_canonical_gbs": round(B / t_batch_vmult / 1e9, 1), "tuning": tuning,
                  "tune_s": round(t_tune, 2)},
        "mgcg": {"batched_s": round(t_batch_solve, 4), "sequential_s": round(t_seq_solve, 4),
                 "speedup": round(t_seq_solve / t_batch_solve, 3), "iterations_max": max(its),
                 "iterations_mean": round(float(np.mean(its)), 2), "all_converged": all(s == 0 for s in status),
                 "batched_equals_sequential": same, "max_err_vs_x_true": err,
                 "vcycle_launches": launches, "graph_nodes": nodes,
                 "tuning_fine": fine_t.as_dict(), "tuning_levels": level_t.as_dict()},
    }


def serving_bench(mfgpu, n_sessions=32, n_requests=64, rate=40.0, seed=100, max_concurrent=4):
    """32 live solid sessions (p ~ U{2..4}, beam [0,4]x[0,1]^2 with n_x ~ U{8,12,..,48}, n_y = n_z = n_x/4,
    E = 1e6, nu = 0.3, rho = 1000, u0 = 1e-3 sin sin sin + 1e-5 noise); requests are one Newmark step each (Newton to
    ||R||/||R0|| <= 1e-8, MG-CG per linear solve) arriving as a Poisson process (seed 100).  Served with
    continuous batching (mixed-precision MG, up to 4 steps in flight, HBM-budget admission) against the same
    requests run one after another with plain FP64 MG."""
    import numpy as np
    rng = np.random.default_rng(seed)
    shapes = []
    for k in range(n_sessions):
        p = int(rng.integers(2, 5))
        nx = 4 * int(rng.integers(2, 13))
        shapes.append((p, nx))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    sessions = rng.integers(0, n_sessions, n_requests)
    reqs = [(int(s), float(t)) for s, t in zip(sessions, arrivals)]

    def specs(level_precision):
        return [mfgpu.session_spec([0.0, 0.0, 0.0], [4.0, 1.0, 1.0], [nx, nx // 4, nx // 4], p, MU, NU, RHO,
                                   beta=BETA, dt=DT, level_precision=level_precision, tol_f=1e-8,
                                   seed=seed + 1 + k)
                for k, (p, nx) in enumerate(shapes)]

    res, st, _ = mfgpu.serve(specs(mfgpu.FP32), reqs, max_concurrent=max_concurrent, real_time=True)
    log(f"serving: {st.completed} steps, {st.steps_per_s:.1f} steps/s, p50 {st.latency_p50_s * 1e3:.1f} ms")
    # capacity: the same requests all released at t = 0 (saturation)
    sat = [(s, 0.0) for s, _ in reqs]
    res_c, st_c, _ = mfgpu.serve(specs(mfgpu.FP32), sat, max_concurrent=max_concurrent, real_time=False)
    log(f"serving: capacity {st_c.steps_per_s:.1f} steps/s")
    res64, st64, _ = mfgpu.serve(specs(mfgpu.FP64), sat, max_concurrent=1, real_time=False)
    log(f"serving: sequential FP64 {st64.steps_per_s:.1f} steps/s")
    return {
        "workload": f"{n_sessions} solid sessions (p ~ U{{2..4}}, beam n_x ~ U{{8..48}} step 4, n_y = n_z = n_x/4), "
                    f"{n_requests} Newmark-step requests, Poisson rate {rate}/s (seed {seed})",
        "n_dofs": [int(3 * (nx * p + 1) * ((nx // 4) * p + 1) ** 2) for p, nx in shapes],
        "served": {"steps_per_s": round(st.steps_per_s, 2), "latency_p50_ms": round(st.latency_p50_s * 1e3, 2),
                   "latency_p99_ms": round(st.latency_p99_s * 1e3, 2), "makespan_s": round(st.makespan_s, 3),
                   "completed": st.completed, "failed": st.failed, "admissions": st.admissions,
                   "evictions": st.evictions, "max_resident": st.max_resident,
                   "peak_bytes": int(st.peak_bytes), "max_concurrent": max_concurrent,
                   "level_precision": "fp32 (fp64 coarse, outer CG and Newton)"},
        "capacity": {"workload": "the same requests, all released at t = 0", "steps_per_s": round(st_c.steps_per_s, 2),
                     "latency_p50_ms": round(st_c.latency_p50_s * 1e3, 2),
                     "latency_p99_ms": round(st_c.latency_p99_s * 1e3, 2), "makespan_s": round(st_c.makespan_s, 3),
                     "completed": st_c.completed, "failed": st_c.failed, "max_concurrent": max_concurrent},
        "sequential_fp64": {"workload": "the same requests, all released at t = 0, one at a time, FP64 levels",
                            "steps_per_s": round(st64.steps_per_s, 2), "makespan_s": round(st64.makespan_s, 3),
                            "latency_p50_ms": round(st64.latency_p50_s * 1e3, 2), "failed": st64.failed},
        "throughput_vs_sequential_fp64": round(st.steps_per_s / max(st64.steps_per_s, 1e-30), 3),
        "capacity_vs_sequential_fp64": round(st_c.steps_per_s / max(st64.steps_per_s, 1e-30), 3),
        "newton_iterations_mean": round(float(np.mean([r.newton_iterations for r in res])), 2),
        "linear_iterations_mean": round(float(np.mean([r.linear_iterations for r in res])), 2),
    }


def cachings_bench(mfgpu, torch, cells, p, numbering, reps=20):
    """The four MF cachings (same operator) on the beam: FP64 vmult time and
    the words each stores per quadrature point (the memory/arithmetic trade)."""
    dev = torch.device("cuda:0")
    mesh = mfgpu.BoxMesh(BEAM["lo"], BEAM["hi"], cells, p, 3, 1, 7, numbering=numbering)
    u0 = torch.from_numpy(mesh.u0(amp=1e-3, noise=1e-5, seed=0)).to(dev)
    src = mfgpu.fill_uniform(1, torch.empty(mesh.n_dofs, dtype=torch.float64, device=dev))
    out, ref = {}, None
    for name in ("scalar_referential", "tensor2", "tensor2_gradient", "tensor4", "tensor4_ns"):
        op = mfgpu.NeoHookOperator(mesh, MU, NU, RHO, ALPHA1, caching=name.replace("_gradient", ""))
        if name.startswith("tensor2"):
            op.set_point_storage(mfgpu.T2_GRADIENT if name == "tensor2_gradient" else mfgpu.T2_FOLDED)
        op.cache(u0)
        dst = torch.empty_like(src)
        for _ in range(3):
            op.vmult(dst, src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            op.vmult(dst, src)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        if ref is None:
            ref = dst.clone()
        out[name] = {"ms": round(ms, 4), "gdofs": round(mesh.n_dofs / (ms * 1e-3) / 1e9, 3),
                     "cache_words_per_qp": op.n_fields(),
                     "rel_diff_vs_scalar_referential": float(torch.linalg.norm(dst - ref) / torch.linalg.norm(ref))}
        log(f"caching {name}: {ms:.3f} ms")
        del op
        torch.cuda.empty_cache()
    return out


def newton_bench(mfgpu, torch, cells, p, numbering):
    """One Newmark time step (Solid::solve_nonlinear_timestep: Newton with the
    per-iteration cache, diagonal and eigenvalue set-up, MG-CG to 1e-8 per
    linear solve) on the beam hierarchy from u0, body force (0, 0, -9.81):
    FP64 levels, then FP32 levels with FP64 outer CG, residual, Newton update
    and coarse level."""
    n_levels = 1
    while (cells[0] >> n_levels) >= 4 and (cells[0] >> n_levels) << n_levels == cells[0] and \
            (cells[1] >> n_levels) >= 1 and (cells[2] >> n_levels) >= 1:
        n_levels += 1
    coarse = [c >> (n_levels - 1) for c in cells]
    out = {"workload": f"one Newmark step (beta 0.25, dt {DT}) from u0, body force (0, 0, -9.81), Newton to "
                       "||R||/||R0|| <= 1e-8 (at least 3 iterations), MG-CG to 1e-8, Chebyshev degree 5, "
                       "coarse FP64 CG to 1e-3", "levels": n_levels}
    us = {}
    for lp, key in ((mfgpu.FP64, "fp64"), (mfgpu.FP32, "mixed_fp32_levels")):
        meshes = [mfgpu.BoxMesh(BEAM["lo"], BEAM["hi"], [c << l for c in coarse], p, 3, 1, 7, numbering=numbering)
                  for l in range(n_levels)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        prob = mfgpu.SolidProblem(meshes, MU, NU, RHO, beta=BETA, dt=DT, body=(0.0, 0.0, -9.81), level_precision=lp,
                                  tol_f=1e-8, tol_lin=1e-8, coarse_kind=1, coarse_tolerance=1e-3)
        prob.set_state(u=torch.from_numpy(meshes[-1].u0(amp=1e-3, noise=1e-5, seed=0)).cuda())
        torch.cuda.synchronize()
        t_create = time.perf_counter() - t0
        t0 = time.perf_counter()
        info = prob.step()
        torch.cuda.synchronize()
        t_step = time.perf_counter() - t0
        u, _, _ = prob.get_state()
        us[key] = u.clone()
        out[key] = {"newton_iterations": info.newton_iterations, "linear_iterations": info.linear_iterations,
                    "converged": bool(info.converged), "step_s": round(t_step, 4), "create_s": round(t_create, 3),
                    "residual_reduction": info.residual_final / info.residual_0 if info.residual_0 else 0.0}
        log(f"newton {key}: {info.newton_iterations} Newton, {info.linear_iterations} CG, {t_step:.3f} s")
        del prob, u
        torch.cuda.empty_cache()
    out["u_diff_mixed_vs_fp64"] = float(torch.linalg.norm(us["mixed_fp32_levels"] - us["fp64"]) /
                                        torch.linalg.norm(us["fp64"]))
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["step_s"] / out["mixed_fp32_levels"]["step_s"], 3)
    return out


def heat_bench(mfgpu, torch, numbering, n=(96, 96, 96), p=4, reps=30):
    """The heat operator M + dt c K (LaplaceOperator) on the unit cube, all
    faces Dirichlet, variable coefficient c ~ U[0.5, 2] per point: vmult time
    in FP64 and FP32, HIP events on the operator's stream."""
    import numpy as np
    mesh = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], list(n), p, 1, 63, 1, numbering=numbering)
    coef = np.random.default_rng(7).uniform(0.5, 2.0, mesh.n_cells * (p + 1) ** 3)
    out = {"workload": f"heat vmult, [0,1]^3, {n[0]}x{n[1]}x{n[2]} cells, FE_Q({p}), QGauss({p + 1}), "
                       f"{mesh.n_dofs} DoFs, dt = 0.1, c ~ U[0.5,2] per point, {numbering} node numbering"}
    for prec, key in ((mfgpu.FP64, "fp64"), (mfgpu.FP32, "fp32")):
        op = mfgpu.LaplaceOperator(mesh, 0.1, coef, precision=prec)
        src = mfgpu.fill_uniform(1, torch.empty(mesh.n_dofs, dtype=op.dtype, device="cuda"))
        dst = torch.empty_like(src)
        for _ in range(3):
            op.vmult(dst, src)
        torch.cuda.synchronize()
        op.enable_timing(True)
        op.timing()
        for _ in range(reps):
            op.vmult(dst, src)
        ms, k = op.timing()
        ms /= k
        es = 8 if prec == mfgpu.FP64 else 4
        nqp = mesh.n_cells * (p + 1) ** 3
        canonical = es * 2 * mesh.n_dofs + 4 * nqp + 11 * es * nqp  # J^-1, JxW and c per point
        layout = es * 2 * mesh.n_dofs + 4 * nqp + op.n_fields() * es * nqp + (10 * es * mesh.n_cells if op.n_fields() == 1 else 0)
        out[key] = {"ms": round(ms, 4), "gdofs": round(mesh.n_dofs / (ms * 1e-3) / 1e9, 3),
                    "cache_words_per_qp": op.n_fields(), "canonical_bytes": canonical,
                    "canonical_gbs": round(canonical / (ms * 1e-3) / 1e9, 1),
                    "canonical_frac": round(canonical / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "layout_bytes": layout,
                    "layout_frac": round(layout / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "layout_note": "the bytes the layout holds (c per point, J^-1 and det J per cell on affine "
                                   "meshes): the actual-bytes counterpart of canonical_frac"}
        log(f"heat {key}: {ms:.3f} ms")
        del op, src, dst
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mg", action="store_true", help="skip the MG-CG legs")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-instances leg")
    ap.add_argument("--only-batched", action="store_true", help="run the batched-instances leg alone")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-cachings", action="store_true", help="skip the MF-caching comparison leg")
    ap.add_argument("--no-newton", action="store_true", help="skip the Newton time-step leg")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="counter probe: 2 vmults, then 2 copies of a 1 GiB buffer (known bytes), no timing")
    ap.add_argument("--pmc-probe-heat", action="store_true",
                    help="counter probe of the heat leg's FP64 operator: 2 vmults, then 2 copies of a 1 GiB buffer")
    ap.add_argument("--only-serving", action="store_true", help="run the serving leg alone")
    ap.add_argument("--only-newton", action="store_true", help="run the Newton time-step leg alone")
    ap.add_argument("--degree", type=int, default=BEAM["degree"])
    ap.add_argument("--cells", type=int, nargs=3, default=BEAM["cells"])
    ap.add_argument("--numbering", default=NUMBERING, choices=["cellwise", "lexicographic"],
                    help="node numbering of the bench meshes")
    ap.add_argument("--clock-warmup-s", type=float, default=0.5,
                    help="seconds of untimed operator applications before the warm-up steps (clock settling)")
    ap.add_argument("--point-storage", default="gradient", choices=["folded", "gradient"],
                    help="tensor2 point storage of the headline operator (16 or 10 words per point)")
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
    if args.only_batched:
        print(json.dumps({"batched": batched_bench(mfgpu, torch),
                          "batched_small": batched_bench(mfgpu, torch, n_range=(2, 6))}), flush=True)
        return
    if args.only_serving:
        print(json.dumps({"serving": serving_bench(mfgpu)}), flush=True)
        return
    if args.only_newton:
        print(json.dumps({"newton": newton_bench(mfgpu, torch, args.cells, args.degree, args.numbering)}), flush=True)
        return

    def copies_of_known_size():
        a = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        for _ in range(2):
            b.copy_(a)
        torch.cuda.synchronize()

    if args.pmc_probe_heat:
        # the heat leg's FP64 operator: 96^3 Q4 cube, c ~ U[0.5, 2] per point (seed 7)
        hmesh = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [96, 96, 96], 4, 1, 63, 1, numbering=args.numbering)
        coef = np.random.default_rng(7).uniform(0.5, 2.0, hmesh.n_cells * 125)
        hop = mfgpu.LaplaceOperator(hmesh, 0.1, coef, precision=mfgpu.FP64)
        hsrc = mfgpu.fill_uniform(1, torch.empty(hmesh.n_dofs, dtype=torch.float64, device=dev))
        hdst = torch.empty_like(hsrc)
        for _ in range(2):
            hop.vmult(hdst, hsrc)
        copies_of_known_size()
        log("pmc probe done")
        return
    p = args.degree
    mesh = mfgpu.BoxMesh(BEAM["lo"], BEAM["hi"], args.cells, p, 3, 1, 7, numbering=args.numbering)
    n_dof, n_cell = mesh.n_dofs, mesh.n_cells
    op = mfgpu.NeoHookOperator(mesh, MU, NU, RHO, ALPHA1, formulation=1, caching="tensor2")
    gradient_storage = args.point_storage == "gradient"
    op.set_point_storage(mfgpu.T2_GRADIENT if gradient_storage else mfgpu.T2_FOLDED)
    u0 = torch.from_numpy(mesh.u0(amp=1e-3, noise=1e-5, seed=0)).to(dev)
    op.cache(u0)
    del u0
    src = mfgpu.fill_uniform(1, torch.empty(n_dof, dtype=torch.float64, device=dev))
    dst = torch.empty_like(src)

    if args.pmc_probe:
        for _ in range(2):
            op.vmult(dst, src)
        copies_of_known_size()
        log("pmc probe done")
        return

    # From idle, the GPU's clocks take about 40 applications to settle (per-
    # application times measured 0.76 -> 0.90 -> 0.70 ms over the first 40):
    # run the operator for a while before the warm-up steps, so that the timed
    # steps see the steady state rather than the ramp.
    t_w = time.perf_counter()
    while time.perf_counter() - t_w < args.clock_warmup_s:
        for _ in range(10):
            op.vmult(dst, src)
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        op.vmult(dst, src)
    torch.cuda.synchronize()
    # GPU time of the timed region from a HIP event pair at its two ends, on
    # the operator's stream (the default stream, torch's current one): events
    # around every call would add their own packets between the vmults
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        op.vmult(dst, src)
    ev1.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    kern_avg_ms = ev0.elapsed_time(ev1) / args.steps
    gdofs = n_dof / (ms_per_step * 1e-3) / 1e9
    B = canonical_bytes(n_dof, n_cell, p)
    n_qp_all = n_cell * (p + 1) ** 3
    achieved = B / (kern_avg_ms * 1e-3) / 1e9

    traffic = measured_traffic((tuple(args.cells), p, args.numbering, args.point_storage))
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
                               f"linearised at 1e-3 sin sin sin + 1e-5 noise, {args.numbering} node numbering",
                   "n_dofs": n_dof, "n_cells": n_cell, "degree": p, "caching": "tensor2", "formulation": 1,
                   "point_storage": args.point_storage, "words_per_point": op.n_fields()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic[0],
                     "traffic_source": traffic[1],
                     "kernel": f"apply_cells<{p},1,double,{'QSolidT2F1G' if gradient_storage else 'QSolidT2F1'}>",
                     "kernel_avg_ms": round(kern_avg_ms, 4),
                     "kernel_avg_source": "HIP event pair at the two ends of the timed region, on the operator's "
                                          "stream, / steps; one vmult = the 8 colour launches of the kernel",
                     "canonical_bytes_per_launch": B,
                     # src + dst, map, the stored words per point, per-cell geometry (det J_ref; J_ref, J_ref^-1)
                     "layout_bytes_per_launch": 8 * 2 * n_dof + 4 * n_qp_all + op.n_fields() * 8 * n_qp_all
                     + 8 * n_cell * (19 if gradient_storage else 1)},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4)},
        "clock_warmup_s": args.clock_warmup_s,
    }
    del op, src, dst
    torch.cuda.empty_cache()
    if not args.no_mg:
        r64, x64 = mgcg_bench(mfgpu, torch, args.cells, p, mfgpu.FP64, args.numbering, storage=args.point_storage)
        x64 = x64.cpu()
        torch.cuda.empty_cache()
        r32, x32 = mgcg_bench(mfgpu, torch, args.cells, p, mfgpu.FP32, args.numbering, storage=args.point_storage)
        r32["err_vs_fp64_solution"] = float(torch.linalg.norm(x32.cpu() - x64) / torch.linalg.norm(x64))
        out["mgcg"] = {"workload": "MG-CG to ||r||/||b|| <= 1e-8, rhs = A x_true (seed 2), Chebyshev degree 5, "
                                   "coarse FP64 CG to 1e-3", "point_storage": args.point_storage,
                       "fp64": r64, "mixed_fp32_vcycle": r32,
                       "speedup_mixed_vs_fp64": round(r64["solve_s"] / r32["solve_s"], 3)}
    if not args.no_newton:
        torch.cuda.empty_cache()
        out["newton"] = newton_bench(mfgpu, torch, args.cells, p, args.numbering)
    if not args.no_cachings:
        torch.cuda.empty_cache()
        out["cachings"] = cachings_bench(mfgpu, torch, args.cells, p, args.numbering)
        out["heat"] = heat_bench(mfgpu, torch, args.numbering)
    if not args.no_batched:
        torch.cuda.empty_cache()
        out["batched"] = batched_bench(mfgpu, torch)
        torch.cuda.empty_cache()
        out["batched_small"] = batched_bench(mfgpu, torch, n_range=(2, 6))
    if not args.no_serving:
        torch.cuda.empty_cache()
        out["serving"] = serving_bench(mfgpu)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
