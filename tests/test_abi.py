"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/mfgpu.h declares, and the host level-layout builder produces index
maps, Dirichlet lists and geometry bit-identical to the oracle's."""
import os
import re

import numpy as np
import pytest

import mfgpu
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(mfg_[a-z0-9_A-Z]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    L = mfgpu.lib()
    syms = declared_symbols()
    assert len(syms) > 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_version_and_error_string():
    L = mfgpu.lib()
    assert b"gfx950" in L.mfg_version()
    with pytest.raises(mfgpu.MfgError) as e:
        mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [1, 1, 1], 9, 3)
    assert e.value.code == 1 and "degree" in str(e.value)


@pytest.mark.parametrize("lo,hi,n,p,nc,fm", [
    ([0, 0, 0], [1, 1, 1], [8, 8, 8], 2, 3, 1),          # small solid case (14,739 dofs)
    ([0, 0, 0], [4, 1, 1], [16, 4, 4], 4, 3, 1),         # beam, Q4
    ([0, 0, 0], [1, 1, 1], [5, 3, 7], 3, 1, 63),         # ragged heat, all faces
    ([-1, 0, 2], [1, 0.5, 3], [1, 1, 1], 6, 3, 0),       # one cell, no constraints
    ([0, 0, 0], [1, 1, 1], [3, 2, 2], 1, 3, 2 | 16),     # two faces
])
def test_layout_bit_identical_to_oracle(lo, hi, n, p, nc, fm):
    cm = 7 if nc == 3 else 1
    a = mfgpu.BoxMesh(lo, hi, n, p, nc, fm, cm)
    b = O.Mesh(lo, hi, n, p, nc, fm, cm)
    assert (a.n_cells, a.n_nodes, a.n_dofs, a.n_constrained) == (b.n_cells, b.n_nodes, b.n_dofs, b.n_constrained)
    assert np.array_equal(a.cell_nodes(), b.cell_nodes())
    assert np.array_equal(a.constrained(), b.constrained())
    assert np.array_equal(a.cell_vertices(), b.cell_vertices())
    assert np.abs(a.node_coords() - b.node_coords()).max() < 1e-15 * max(1.0, np.abs(hi).max())


def test_constrained_count_closed_forms():
    # x = 0 face of an n_x x n_y x n_z Q_p mesh: 3 (n_y p + 1)(n_z p + 1) dofs
    m = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [8, 8, 8], 2, 3, 1, 7)
    assert m.n_constrained == 867 and m.n_dofs == 14739
    m = mfgpu.BoxMesh([0, 0, 0], [4, 1, 1], [128, 32, 32], 4, 3, 1, 7)
    assert m.n_constrained == 3 * 129 * 129 == 49923
    assert m.n_dofs == 513 * 129 * 129 * 3 == 25610499


def test_node_referenced_by_no_cell_is_refused():
    """A level whose node list has a node no cell references would leave that
    node's dst entries unwritten by vmult; the layout builder refuses it before
    any device work."""
    import ctypes
    m = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [2, 2, 2], 1, 3, 1, 7)
    d = mfgpu.LevelDesc()
    ctypes.pointer(d)[0] = m.desc
    d.n_nodes = m.n_nodes + 1
    prm = mfgpu.SolidParams(O.MU, O.NU, O.RHO, O.ALPHA1, 1, 1)
    h = ctypes.c_void_p()
    st = mfgpu.lib().mfg_solid_create(ctypes.byref(d), ctypes.byref(prm), mfgpu.FP64, ctypes.byref(h))
    assert st == 1 and b"no cell" in mfgpu.lib().mfg_last_error()
