"""Writes small_cube.npz: the oracle's outputs on the small FP64 configuration.

Unit cube [0,1]^3, coarse 2x2x2 refined twice (8^3 cells, 3 levels), vector
FE_Q(2) on Gauss-Lobatto nodes, QGauss(3); neo-Hookean E = 1e6, nu = 0.3,
rho = 1000, Newmark beta = 0.25, dt = 1e-3; u = 0 on x = 0; linearised at
u0 = 1e-3 sin(pi x) sin(pi y) sin(pi z) + 1e-5 U[-1,1] (seed 0).
Stored: layout hashes, dst = A src (src seed 1), and the MG-CG solve of
A x = A x_true (x_true seed 2, constrained entries zero; Chebyshev degree 4,
Chebyshev coarse solver) to ||r||/||b|| <= 1e-8: iterations, solution and
the level eigenvalue estimates.

Run from the repository root: python tests/golden/make_small_cube.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def compute():
    meshes = O.box_levels([0, 0, 0], [1, 1, 1], [2, 2, 2], 3, 2, 3, 1, 7)
    fine = meshes[-1]
    u = fine.u0(amp=1e-3, noise=1e-5, seed=0)
    us = [u]
    for l in range(len(meshes) - 1, 0, -1):
        us.insert(0, O.interpolate(meshes[l - 1], meshes[l], us[0]))
    ops = [O.Solid(m, caching=1) for m in meshes]
    for op, uu in zip(ops, us):
        op.cache(uu)
    src = O.fill_uniform(1, fine.n_dofs)
    dst = ops[-1].vmult(src)
    xt = O.fill_uniform(2, fine.n_dofs)
    xt[fine.constrained()] = 0.0
    b = ops[-1].vmult(xt)
    x, its, rel = O.mg_cg_solve(meshes, ops, b, cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30,
                                coarse_kind=0, tol_rel=1e-8)
    _, eig = O.mg_vmult(meshes, ops, b, cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30, coarse_kind=0)
    return dict(n_dofs=np.int64(fine.n_dofs), n_constrained=np.int64(len(fine.constrained())),
                cell_nodes_sha=np.array(sha(fine.cell_nodes().astype(np.int32))),
                constrained_sha=np.array(sha(fine.constrained().astype(np.int32))),
                u0=u, dst=dst, mg_iterations=np.int64(its), mg_rel_residual=np.float64(rel), mg_x=x,
                x_true_err=np.float64(np.linalg.norm(x - xt) / np.linalg.norm(xt)),
                eig=np.array(eig, dtype=np.float64))


if __name__ == "__main__":
    out = compute()
    np.savez_compressed(os.path.join(HERE, "small_cube.npz"), **out)
    print({k: (v if np.ndim(v) == 0 else np.shape(v)) for k, v in out.items()})
