"""Cellwise node numbering of box meshes (host side, no GPU).

The cellwise numbering must be a pure relabelling of the lexicographic mesh:
the same cells, coordinates, Dirichlet dofs and linearisation point, only
with other node ids.  It must also deliver what it is for: every cell owns one
contiguous block of nodes, and a cell's 3-component nodes fall into fewer
128-byte lines than in the lexicographic layout.
"""
import numpy as np
import pytest

import mfgpu

CASES = [
    ([0, 0, 0], [4, 1, 1], [16, 4, 4], 4, 3, 1),
    ([0, 0, 0], [1, 1, 1], [5, 3, 7], 3, 1, 63),
    ([-1, 0, 2], [1, 0.5, 3], [1, 1, 1], 6, 3, 0),
    ([0, 0, 0], [1, 1, 1], [3, 2, 2], 1, 3, 2 | 16),
    ([0, 0, 0], [1, 1, 1], [4, 4, 4], 2, 3, 1),
    ([0, 0, 0], [2, 1, 1], [6, 3, 3], 5, 3, 1),
]


def pair(lo, hi, n, p, nc, fm):
    cm = 7 if nc == 3 else 1
    a = mfgpu.BoxMesh(lo, hi, n, p, nc, fm, cm)
    b = mfgpu.BoxMesh(lo, hi, n, p, nc, fm, cm, numbering="cellwise")
    perm = np.full(a.n_nodes, -1, dtype=np.int64)
    perm[a.cell_nodes()] = b.cell_nodes()
    return a, b, perm


@pytest.mark.parametrize("lo,hi,n,p,nc,fm", CASES)
def test_cellwise_is_a_relabelling(lo, hi, n, p, nc, fm):
    a, b, perm = pair(lo, hi, n, p, nc, fm)
    assert (a.n_cells, a.n_nodes, a.n_dofs, a.n_constrained) == (b.n_cells, b.n_nodes, b.n_dofs, b.n_constrained)
    assert np.array_equal(np.sort(perm), np.arange(a.n_nodes))  # a bijection on the nodes
    assert np.array_equal(perm[a.cell_nodes()], b.cell_nodes())  # one relabelling for every cell entry
    assert np.array_equal(a.cell_vertices(), b.cell_vertices())
    assert np.array_equal(b.node_coords()[perm], a.node_coords())
    ca = a.constrained().astype(np.int64)
    assert np.array_equal(np.sort(perm[ca // nc] * nc + ca % nc), b.constrained())
    assert np.all(np.diff(b.constrained()) > 0)
    dof = (perm[:, None] * nc + np.arange(nc)[None, :]).ravel()  # cellwise dof of each lexicographic dof
    assert np.array_equal(b.u0(amp=1e-3, noise=1e-5, seed=3)[dof], a.u0(amp=1e-3, noise=1e-5, seed=3))


@pytest.mark.parametrize("lo,hi,n,p,nc,fm", CASES)
def test_cellwise_blocks(lo, hi, n, p, nc, fm):
    """Cell c owns the nodes [c p^3, (c+1) p^3): its high-side half-open box."""
    a, b, perm = pair(lo, hi, n, p, nc, fm)
    N = p + 1
    cn = b.cell_nodes().reshape(-1, N ** 3)
    l = np.arange(N ** 3)
    high = (l % N > 0) & ((l // N) % N > 0) & (l // (N * N) > 0)
    for c in range(b.n_cells):
        assert np.array_equal(np.sort(cn[c][high]), np.arange(c * p ** 3, (c + 1) * p ** 3))
    # the rest are the nodes of the low boundary planes
    x = b.node_coords()[b.n_cells * p ** 3:]
    on_low = np.isclose(x[:, 0], lo[0]) | np.isclose(x[:, 1], lo[1]) | np.isclose(x[:, 2], lo[2])
    assert on_low.all() and len(x) == b.n_nodes - b.n_cells * p ** 3


def test_renumbering_round_trips_and_is_idempotent():
    lo, hi, n, p = [0, 0, 0], [1, 1, 1], [3, 4, 2], 3
    ref = mfgpu.BoxMesh(lo, hi, n, p, 3, 1, 7)
    m = mfgpu.BoxMesh(lo, hi, n, p, 3, 1, 7, numbering="cellwise")
    first = m.cell_nodes()
    mfgpu.check(mfgpu.lib().mfg_box_mesh_renumber(m.h, mfgpu.NUMBERING["cellwise"]))
    assert np.array_equal(m.cell_nodes(), first)
    mfgpu.check(mfgpu.lib().mfg_box_mesh_renumber(m.h, mfgpu.NUMBERING["lexicographic"]))
    assert mfgpu.lib().mfg_box_mesh_numbering(m.h) == mfgpu.NUMBERING["lexicographic"]
    assert np.array_equal(m.cell_nodes(), ref.cell_nodes())
    assert np.array_equal(m.constrained(), ref.constrained())
    assert np.array_equal(m.node_coords(), ref.node_coords())
    assert np.array_equal(m.u0(seed=5), ref.u0(seed=5))
    with pytest.raises(mfgpu.MfgError):
        mfgpu.check(mfgpu.lib().mfg_box_mesh_renumber(m.h, 7))
    with pytest.raises(ValueError):
        mfgpu.BoxMesh(lo, hi, n, p, 3, numbering="random")


def lines_per_cell(mesh, nc=3, line=128):
    """Distinct 128-byte lines of an FP64 nc-component vector that each cell's gather touches."""
    N3 = (mesh.degree + 1) ** 3
    cn = mesh.cell_nodes().reshape(-1, N3).astype(np.int64)
    first = cn * nc * 8 // line
    last = (cn * nc * 8 + nc * 8 - 1) // line
    return np.array([len(np.union1d(f, l)) for f, l in zip(first, last)])


def test_cellwise_gather_touches_fewer_lines():
    # Q4 beam section: lexicographic x-rows of 5 nodes (120 B) straddle two
    # lines; a cellwise block is 64 nodes = 12 whole lines, plus the
    # neighbours' face, edge and vertex pieces (18 lines when ordered well)
    a, b, _ = pair([0, 0, 0], [4, 1, 1], [16, 4, 4], 4, 3, 1)
    la, lb = lines_per_cell(a), lines_per_cell(b)
    interior = [c for c in range(a.n_cells)
                if 0 < c % 16 < 15 and 0 < (c // 16) % 4 < 3 and 0 < c // 64 < 3]
    print("mean lines per interior cell: lexicographic", la[interior].mean(), "cellwise", lb[interior].mean())
    assert lb[interior].max() == lb[interior].min() == 12 + 18
    assert la[interior].mean() > 45
