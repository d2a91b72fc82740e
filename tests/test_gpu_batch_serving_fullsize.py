"""The batched and serving workloads at their stated sizes.

Batched (48 seeded instances, 24 heat + 24 solid, p ~ U{1..6}, unit cubes
with n ~ U{4..24}^3; the instances the benchmark builds): the batched vmult
and the batched MG-CG equal the per-instance calls bitwise, and a
degree-stratified sample -- the smallest instance of every (kind, degree)
present -- matches the oracle.

Serving (32 seeded sessions up to 48 x 12 x 12 cells, p ~ U{2..4}, 64
requests; the benchmark's sessions): every session's results are the same
alone and in the concurrent mix, and the final states of the three smallest
sessions that received requests match the oracle's Newton-Newmark steps."""
import numpy as np
import pytest
import torch

import bench
import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu


def _lex_dofs(mesh_cellwise, nc):
    """cellwise dof of each lexicographic dof of the same box."""
    lex = mfgpu.BoxMesh(mesh_cellwise._lo, mesh_cellwise._hi, mesh_cellwise.n_cells_dir, mesh_cellwise.degree, nc,
                        1 if nc == 3 else 63, 7 if nc == 3 else 1)
    perm = np.full(lex.n_nodes, -1, dtype=np.int64)
    perm[lex.cell_nodes()] = mesh_cellwise.cell_nodes()
    return (perm[:, None] * nc + np.arange(nc)[None, :]).ravel()


@pytest.fixture(scope="module")
def instances(cuda):
    inst = bench.batched_instances(mfgpu)
    print(f"48-instance batch: {len(inst)} instances, {sum(i.n_dofs for i in inst)} dofs", flush=True)
    return inst


def test_batched_vmult_48_instances(instances):
    inst = instances
    srcs = [mfgpu.fill_uniform(500 + k, torch.empty(i.n_dofs, dtype=torch.float64, device="cuda"))
            for k, i in enumerate(inst)]
    ref = [torch.empty_like(s) for s in srcs]
    for i, s, r in zip(inst, srcs, ref):
        i.fine.vmult(r, s)
    b = mfgpu.Batch([i.fine for i in inst])
    # GATHER twice: the second call with other dst vectors rebuilds its tables
    for sched in (mfgpu.SCHEDULE_QUEUE, mfgpu.SCHEDULE_STREAMS, mfgpu.SCHEDULE_GATHER, mfgpu.SCHEDULE_GATHER):
        b.set_schedule(sched)
        out = [torch.full_like(s, float("nan")) for s in srcs]
        b.vmult(out, srcs)
        torch.cuda.synchronize()
        assert all(torch.equal(o, r) for o, r in zip(out, ref)), sched
    # the smallest instance of every (kind, degree) against the oracle
    strata = {}
    for k, i in enumerate(inst):
        key = (i.kind, i.degree)
        if key not in strata or i.n_dofs < inst[strata[key]].n_dofs:
            strata[key] = k
    print(f"oracle sample: {sorted(strata)}", flush=True)
    for (kind, p), k in sorted(strata.items()):
        i = inst[k]
        m = i.meshes[-1]
        nc = 3 if kind == "solid" else 1
        if kind == "solid":
            om = O.Mesh([0, 0, 0], [1, 1, 1], m.n_cells_dir, p, 3, 1, 7)
            oo = O.Solid(om, caching=1)
            oo.cache(om.u0(amp=1e-3, noise=1e-5, seed=3 + k))
        else:
            om = O.Mesh([0, 0, 0], [1, 1, 1], m.n_cells_dir, p, 1, 63, 1)
            oo = O.Heat(om, dt=0.1, coef=np.full(om.n_cells * (p + 1) ** 3, i.conductivity))
        dof = _lex_dofs(m, nc)
        x = srcs[k][torch.from_numpy(dof).cuda()].cpu().numpy()
        want = oo.vmult(x)
        got = ref[k][torch.from_numpy(dof).cuda()].cpu().numpy()
        r = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"  {kind} p={p} n={m.n_cells_dir}: relative difference to the oracle {r:.2e}", flush=True)
        assert r <= 1e-12, (kind, p)


def test_batched_mgcg_48_instances(instances):
    inst = instances
    bs, xts = [], []
    for k, i in enumerate(inst):
        xt = mfgpu.fill_uniform(600 + k, torch.empty(i.n_dofs, dtype=torch.float64, device="cuda"))
        xt[torch.from_numpy(i.constrained).long().cuda()] = 0.0
        b = torch.empty_like(xt)
        i.fine.vmult(b, xt)
        bs.append(b)
        xts.append(xt)
    solver = mfgpu.BatchSolver([i.fine for i in inst], [i.mg for i in inst])
    for rep in range(2):  # eager programs, then captured graphs
        xs = [torch.zeros_like(b) for b in bs]
        infos, status = solver.solve(bs, xs, tol_rel=1e-8)
        assert all(s == 0 for s in status)
    for i, b, x, xt, inf in zip(inst, bs, xs, xts, infos):
        xq = torch.zeros_like(b)
        info = mfgpu.cg_solve(i.fine, b, xq, mfgpu.PRECOND_MG, i.mg, tol_rel=1e-8)
        assert info.iterations == inf.iterations
        assert torch.equal(xq, x)
        assert float((x - xt).norm() / xt.norm()) <= 1e-5
    print(f"48-instance batch MG-CG: iterations {[inf.iterations for inf in infos]}", flush=True)


def _serving_workload():
    rng = np.random.default_rng(100)
    shapes = []
    for k in range(32):
        p = int(rng.integers(2, 5))
        nx = 4 * int(rng.integers(2, 13))
        shapes.append((p, nx))
    arrivals = np.cumsum(rng.exponential(1.0 / 40.0, 64))
    sessions = rng.integers(0, 32, 64)
    reqs = [(int(s), float(t)) for s, t in zip(sessions, arrivals)]
    specs = [mfgpu.session_spec([0.0, 0.0, 0.0], [4.0, 1.0, 1.0], [nx, nx // 4, nx // 4], p, bench.MU, bench.NU,
                                bench.RHO, beta=bench.BETA, dt=bench.DT, level_precision=mfgpu.FP32, tol_f=1e-8,
                                seed=101 + k)
             for k, (p, nx) in enumerate(shapes)]
    return shapes, reqs, specs


def test_serving_32_sessions(cuda):
    shapes, reqs, specs = _serving_workload()
    alone, st1, fin1 = mfgpu.serve(specs, reqs, max_concurrent=1, real_time=False, final_states=True)
    mix, st4, fin4 = mfgpu.serve(specs, reqs, max_concurrent=4, real_time=False, final_states=True)
    print(f"32-session serving: alone {st1.steps_per_s:.1f} steps/s, mix {st4.steps_per_s:.1f} steps/s "
          f"(p50 {st4.latency_p50_s * 1e3:.1f} ms, p99 {st4.latency_p99_s * 1e3:.1f} ms)", flush=True)
    assert st1.completed == st4.completed == len(reqs) and st4.failed == 0
    for a, b in zip(alone, mix):
        assert (a.status, a.newton_iterations, a.linear_iterations, a.residual_final) == \
               (b.status, b.newton_iterations, b.linear_iterations, b.residual_final)
    for x, y in zip(fin1, fin4):
        assert np.array_equal(x, y)
    # the oracle's Newton-Newmark steps on the three smallest sessions that received requests
    counts = np.bincount([s for s, _ in reqs], minlength=32)
    served = [k for k in range(32) if counts[k] > 0]
    size = lambda k: (shapes[k][1] * shapes[k][0] + 1) * ((shapes[k][1] // 4) * shapes[k][0] + 1) ** 2
    for k in sorted(served, key=size)[:3]:
        p, nx = shapes[k]
        n = [nx, nx // 4, nx // 4]
        levels = 1
        while all((c >> levels) << levels == c for c in n):
            levels += 1
        coarse = [c >> (levels - 1) for c in n]
        meshes = O.box_levels([0, 0, 0], [4.0, 1.0, 1.0], coarse, levels, p, 3)
        ops = [O.Solid(m, alpha1=1.0 / (bench.BETA * bench.DT ** 2), caching=1) for m in meshes]
        fm = meshes[-1]
        u = fm.u0(amp=1e-3, noise=1e-5, seed=101 + k)
        v, a = np.zeros_like(u), np.zeros_like(u)
        for _ in range(counts[k]):
            it, _ = O.newmark_step(meshes, ops, u, v, a, beta=bench.BETA, dt=bench.DT, tol_f=1e-8, tol_u=1e-6,
                                   tol_lin=1e-6)
            assert it > 0
        # the session's mesh is numbered cellwise
        gm = mfgpu.BoxMesh([0, 0, 0], [4.0, 1.0, 1.0], n, p, 3, 1, 7, numbering="cellwise")
        dof = _lex_dofs(gm, 3)
        got = fin4[k][dof]
        d = np.linalg.norm(got - u) / np.linalg.norm(u)
        print(f"  session {k} (p={p}, {n}, {counts[k]} steps): relative difference to the oracle {d:.2e}", flush=True)
        assert d <= 1e-7
