"""Python binding of libmfgpu's C ABI (include/mfgpu.h).

The library is the product; this module only loads it with ctypes, checks
status codes (raising MfgError, the counterpart of deal.II's AssertThrow) and
offers thin classes whose method names follow the reference operators
(NeoHookOperator / LaplaceOperator: vmult, vmult_add, Tvmult, cache,
compute_diagonal, el, precondition_Jacobi, memory_consumption, m, n).
Device vectors are torch tensors on the GPU; torch provides memory and
streams, never compute on this path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfgpu.so")

FP64, FP32 = 0, 1
CACHING = {"scalar_referential": 0, "tensor2": 1, "tensor4": 2, "tensor4_ns": 3}
CELLS_COLOURS, CELLS_GATHER = 0, 3
NUMBERING = {"lexicographic": 0, "cellwise": 1}
T2_FOLDED, T2_GRADIENT = 0, 1  # tensor2 formulation-1 point storage

_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class MfgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mfgpu error {code}: {msg}")
        self.code = code


class LevelDesc(ctypes.Structure):
    _fields_ = [("degree", ctypes.c_int32), ("n_q_1d", ctypes.c_int32), ("n_components", ctypes.c_int32),
                ("n_cells", ctypes.c_int32), ("n_nodes", ctypes.c_int64), ("cell_nodes", _c_i32p),
                ("cell_vertices", _c_dp), ("constrained_dofs", _c_i32p), ("n_constrained", ctypes.c_int64),
                ("material_ids", _c_i32p)]


class SolidParams(ctypes.Structure):
    _fields_ = [("mu", ctypes.c_double), ("nu", ctypes.c_double), ("rho", ctypes.c_double),
                ("alpha_1", ctypes.c_double), ("formulation", ctypes.c_int32), ("caching", ctypes.c_int32),
                ("inclusion_mu", ctypes.c_double), ("inclusion_nu", ctypes.c_double),
                ("inclusion_rho", ctypes.c_double)]


def _solid_params(mu, nu, rho, alpha_1, formulation, caching, inclusion=None):
    """SolidParams; inclusion: (mu, nu, rho) of the material on cells whose material id is not 0."""
    imu, inu, irho = inclusion if inclusion is not None else (0.0, 0.0, 0.0)
    return SolidParams(mu, nu, rho, alpha_1, formulation, caching, imu, inu, irho)


_LIB = None

# name -> (restype, argtypes); the exported surface of include/mfgpu.h
SIGNATURES = {
    "mfg_operator_set_cell_schedule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "mfg_operator_cell_schedule": (ctypes.c_int32, [ctypes.c_void_p]),
    "mfg_last_error": (ctypes.c_char_p, []),
    "mfg_version": (ctypes.c_char_p, []),
    "mfg_box_mesh_create_bc": (ctypes.c_int, [_c_dp, _c_dp, _c_i32p, ctypes.c_int32, ctypes.c_int32, _c_i32p,
                                               ctypes.POINTER(_vp)]),
    "mfg_box_mesh_create": (ctypes.c_int, [_c_dp, _c_dp, _c_i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int32, ctypes.POINTER(_vp)]),
    "mfg_box_mesh_destroy": (ctypes.c_int, [_vp]),
    "mfg_box_mesh_set_material_ids": (ctypes.c_int, [_vp, _c_i32p]),
    "mfg_box_mesh_describe": (ctypes.c_int, [_vp, ctypes.POINTER(LevelDesc)]),
    "mfg_box_mesh_node_coords": (ctypes.c_int, [_vp, _c_dp]),
    "mfg_box_mesh_renumber": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mfg_mg_use_graph": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mfg_operator_quadrature_points": (ctypes.c_int, [_vp, _vp]),
    "mfg_heat_assemble_rhs": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "mfg_heat_l2_error": (ctypes.c_int, [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_double)]),
    "mfg_mg_graph_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "mfg_box_mesh_numbering": (ctypes.c_int32, [_vp]),
    "mfg_solid_create": (ctypes.c_int, [ctypes.POINTER(LevelDesc), ctypes.POINTER(SolidParams), ctypes.c_int32,
                                         ctypes.POINTER(_vp)]),
    "mfg_heat_create": (ctypes.c_int, [ctypes.POINTER(LevelDesc), ctypes.c_double, _c_dp, ctypes.c_int32,
                                        ctypes.POINTER(_vp)]),
    "mfg_operator_destroy": (ctypes.c_int, [_vp]),
    "mfg_operator_set_stream": (ctypes.c_int, [_vp, _vp]),
    "mfg_operator_m": (ctypes.c_int64, [_vp]),
    "mfg_operator_n": (ctypes.c_int64, [_vp]),
    "mfg_operator_precision": (ctypes.c_int32, [_vp]),
    "mfg_solid_cache": (ctypes.c_int, [_vp, _vp, ctypes.c_int32]),
    "mfg_operator_vmult": (ctypes.c_int, [_vp, _vp, _vp]),
    "mfg_operator_vmult_add": (ctypes.c_int, [_vp, _vp, _vp]),
    "mfg_operator_Tvmult": (ctypes.c_int, [_vp, _vp, _vp]),
    "mfg_operator_Tvmult_add": (ctypes.c_int, [_vp, _vp, _vp]),
    "mfg_operator_compute_diagonal": (ctypes.c_int, [_vp]),
    "mfg_operator_copy_diagonal": (ctypes.c_int, [_vp, _vp, _vp]),
    "mfg_operator_el": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, _c_dp]),
    "mfg_operator_precondition_jacobi": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double]),
    "mfg_operator_memory_consumption": (ctypes.c_size_t, [_vp]),
    "mfg_operator_qp_data": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_int32)]),
    "mfg_operator_n_colours": (ctypes.c_int32, [_vp]),
    "mfg_operator_set_point_storage": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mfg_operator_point_storage": (ctypes.c_int32, [_vp]),
    "mfg_operator_enable_timing": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mfg_operator_timing": (ctypes.c_int, [_vp, _c_dp, ctypes.POINTER(ctypes.c_int64)]),
    "mfg_box_mesh_cells": (ctypes.c_int, [_vp, _c_i32p]),
    "mfg_box_mesh_boundary_faces": (ctypes.c_int, [_vp, ctypes.c_int32, _c_i32p, ctypes.POINTER(ctypes.c_int64)]),
    "mfg_solid_add_interface_traction": (ctypes.c_int, [_vp, ctypes.c_int64, _c_i32p, _vp, _vp, _vp]),
    "mfg_operator_face_quadrature_points": (ctypes.c_int, [_vp, ctypes.c_int64, _c_i32p, _vp]),
    "mfg_operator_face_values": (ctypes.c_int, [_vp, ctypes.c_int64, _c_i32p, _vp, _vp]),
    "mfg_heat_add_interface_flux": (ctypes.c_int, [_vp, ctypes.c_int64, _c_i32p, _vp, _vp]),
    "mfg_box_mesh_children": (ctypes.c_int, [_vp, _vp, _c_i32p]),
    "mfg_transfer_create": (ctypes.c_int, [_vp, _vp, _c_i32p, ctypes.POINTER(_vp)]),
    "mfg_transfer_destroy": (ctypes.c_int, [_vp]),
    "mfg_transfer_prolongate_add": (ctypes.c_int, [_vp, _vp, _vp]),
    "mfg_transfer_restrict_add": (ctypes.c_int, [_vp, _vp, _vp]),
    "mfg_transfer_interpolate": (ctypes.c_int, [_vp, _vp, _vp]),
    "mfg_mg_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_void_p,
                                      ctypes.POINTER(_vp)]),
    "mfg_mg_destroy": (ctypes.c_int, [_vp]),
    "mfg_mg_setup": (ctypes.c_int, [_vp]),
    "mfg_mg_vmult": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32]),
    "mfg_mg_level_info": (ctypes.c_int, [_vp, ctypes.c_int32, _c_dp, _c_dp, _c_i32p]),
    "mfg_cg_solve": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, ctypes.c_double, _vp, _vp, ctypes.c_double,
                                     ctypes.c_int32, ctypes.c_void_p]),
    "mfg_cg_solve_noalloc": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, ctypes.c_double, _vp, _vp, ctypes.c_double,
                                             ctypes.c_int32, ctypes.c_void_p]),
    "mfg_fill_uniform": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int32, _vp, _vp]),
    "mfg_box_mesh_fill_u0": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _c_dp]),
}


def fill_uniform(seed, t, stream=None):
    """t <- seeded U[-1,1] (splitmix64), generated on the device."""
    import torch
    prec = FP64 if t.dtype == torch.float64 else FP32
    s = None if stream is None else ctypes.c_void_p(int(stream.cuda_stream))
    check(lib().mfg_fill_uniform(seed, t.numel(), prec, _ptr(t), s))
    return t


def lib():
    """Load libmfgpu.so; raises if the HIP library has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(status):
    if status != 0:
        raise MfgError(status, lib().mfg_last_error().decode())


def _ptr(t):
    """Device (or host numpy) pointer of a contiguous tensor/array."""
    if hasattr(t, "data_ptr"):
        assert t.is_contiguous()
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)


def _dtype_of(precision):
    import torch
    return torch.float64 if precision == FP64 else torch.float32


def _check_vec(t, n, dtype, name="vector"):
    """Reject, before any launch, a vector that the kernels would index out of
    bounds or misread: it must be a contiguous device tensor of n entries of
    the operator's precision."""
    if not hasattr(t, "data_ptr") or not t.is_cuda:
        raise ValueError(f"{name}: expected a device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype} does not match the operator's {dtype}")
    if t.numel() != n:
        raise ValueError(f"{name}: {t.numel()} entries, the operator has {n} rows")


def _check_pair(dst, src, n, dtype):
    _check_vec(dst, n, dtype, "dst")
    _check_vec(src, n, dtype, "src")
    if dst.data_ptr() == src.data_ptr():
        raise ValueError("dst and src must be different vectors")


def refined_material_ids(ids, n_cells, times):
    """Material ids of the box mesh refined `times` times from one with n_cells = (n0, n1, n2) cells and
    ids per cell: children inherit their parent's id, as deal.II's refinement does."""
    import numpy as np
    a = np.asarray(ids, dtype=np.int32).reshape(n_cells[2], n_cells[1], n_cells[0])
    for _ in range(times):
        a = a.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)
    return np.ascontiguousarray(a.reshape(-1))


class BoxMesh:
    """Structured box mesh level (cells, nodes, Dirichlet dofs) built on the host."""

    def __init__(self, lo, hi, n_cells, degree, n_components, face_mask=1, comp_mask=7, numbering="lexicographic",
                 face_comp=None):
        """face_comp: optional component mask per boundary face (6 ints: x=lo, x=hi, y=lo, y=hi, z=lo,
        z=hi), which replaces (face_mask, comp_mask)."""
        import numpy as np
        self._lo = np.ascontiguousarray(lo, dtype=np.float64)
        self._hi = np.ascontiguousarray(hi, dtype=np.float64)
        self._n = np.ascontiguousarray(n_cells, dtype=np.int32)
        if face_comp is None:
            face_comp = [comp_mask if (face_mask >> f) & 1 else 0 for f in range(6)]
        self._face_comp = np.ascontiguousarray(face_comp, dtype=np.int32)
        h = _vp()
        check(lib().mfg_box_mesh_create_bc(self._lo.ctypes.data_as(_c_dp), self._hi.ctypes.data_as(_c_dp),
                                            self._n.ctypes.data_as(_c_i32p), degree, n_components,
                                            self._face_comp.ctypes.data_as(_c_i32p), ctypes.byref(h)))
        self.h = h
        self._destroy = lib().mfg_box_mesh_destroy
        if numbering not in NUMBERING:
            raise ValueError(f"numbering must be one of {sorted(NUMBERING)}")
        check(lib().mfg_box_mesh_renumber(self.h, NUMBERING[numbering]))
        self.numbering = numbering
        self.desc = LevelDesc()
        check(lib().mfg_box_mesh_describe(self.h, ctypes.byref(self.desc)))
        self.degree, self.n_components = degree, n_components
        self.n_cells_dir = list(n_cells)
        self.n_cells = self.desc.n_cells
        self.n_nodes = self.desc.n_nodes
        self.n_dofs = self.n_nodes * n_components
        self.n_constrained = self.desc.n_constrained

    def __del__(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def set_material_ids(self, ids):
        """Material id per cell (cell cx + n0 (cy + n1 cz)); None resets all to 0.  Operators built from the
        mesh afterwards use the inclusion material on cells whose id is not 0."""
        import numpy as np
        if ids is None:
            self._material_ids = None
            check(lib().mfg_box_mesh_set_material_ids(self.h, None))
        else:
            self._material_ids = np.ascontiguousarray(ids, dtype=np.int32).reshape(-1)
            if len(self._material_ids) != self.n_cells:
                raise ValueError("one material id per cell")
            check(lib().mfg_box_mesh_set_material_ids(self.h, self._material_ids.ctypes.data_as(_c_i32p)))
        check(lib().mfg_box_mesh_describe(self.h, ctypes.byref(self.desc)))

    def cell_nodes(self):
        import numpy as np
        n = self.n_cells * (self.degree + 1) ** 3
        return np.ctypeslib.as_array(self.desc.cell_nodes, shape=(n,)).copy()

    def boundary_faces(self, boundary_id):
        """(cell, face) pairs on box boundary id 2 d (low side) / 2 d + 1 (high side) of direction d, int32 (n, 2)."""
        import numpy as np
        n = ctypes.c_int64(0)
        check(lib().mfg_box_mesh_boundary_faces(self.h, boundary_id, None, ctypes.byref(n)))
        out = np.zeros(2 * n.value, dtype=np.int32)
        check(lib().mfg_box_mesh_boundary_faces(self.h, boundary_id, out.ctypes.data_as(_c_i32p), ctypes.byref(n)))
        return out.reshape(-1, 2)

    def constrained(self):
        import numpy as np
        if self.n_constrained == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(self.desc.constrained_dofs, shape=(self.n_constrained,)).copy()

    def cell_vertices(self):
        import numpy as np
        return np.ctypeslib.as_array(self.desc.cell_vertices, shape=(self.n_cells * 24,)).copy()

    def node_coords(self):
        import numpy as np
        out = np.zeros(self.n_nodes * 3)
        check(lib().mfg_box_mesh_node_coords(self.h, out.ctypes.data_as(_c_dp)))
        return out.reshape(-1, 3)

    def u0(self, amp=1e-3, noise=1e-5, seed=0):
        """Linearisation point: amp sin(pi x) sin(pi y) sin(pi z) + noise U[-1,1], host array."""
        import numpy as np
        out = np.zeros(self.n_dofs)
        check(lib().mfg_box_mesh_fill_u0(self.h, amp, noise, seed, out.ctypes.data_as(_c_dp)))
        return out


class _Operator:
    def __init__(self, handle, precision, keep):
        self.h = handle
        self.precision = precision
        self._keep = keep
        self._destroy = lib().mfg_operator_destroy

    def __del__(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    @property
    def dtype(self):
        return _dtype_of(self.precision)

    def m(self):
        return lib().mfg_operator_m(self.h)

    def n(self):
        return lib().mfg_operator_n(self.h)

    def set_stream(self, stream):
        check(lib().mfg_operator_set_stream(self.h, ctypes.c_void_p(int(stream.cuda_stream))))

    def vmult(self, dst, src):
        _check_pair(dst, src, self.m(), self.dtype)
        check(lib().mfg_operator_vmult(self.h, _ptr(dst), _ptr(src)))

    def vmult_add(self, dst, src):
        _check_pair(dst, src, self.m(), self.dtype)
        check(lib().mfg_operator_vmult_add(self.h, _ptr(dst), _ptr(src)))

    def Tvmult(self, dst, src):
        _check_pair(dst, src, self.m(), self.dtype)
        check(lib().mfg_operator_Tvmult(self.h, _ptr(dst), _ptr(src)))

    def Tvmult_add(self, dst, src):
        _check_pair(dst, src, self.m(), self.dtype)
        check(lib().mfg_operator_Tvmult_add(self.h, _ptr(dst), _ptr(src)))

    def compute_diagonal(self):
        check(lib().mfg_operator_compute_diagonal(self.h))

    def diagonal(self):
        import torch
        d = torch.empty(self.m(), dtype=self.dtype, device="cuda")
        check(lib().mfg_operator_copy_diagonal(self.h, _ptr(d), None))
        return d

    def get_matrix_diagonal_inverse(self):
        import torch
        d = torch.empty(self.m(), dtype=self.dtype, device="cuda")
        check(lib().mfg_operator_copy_diagonal(self.h, None, _ptr(d)))
        return d

    def el(self, row, col):
        v = ctypes.c_double(0)
        check(lib().mfg_operator_el(self.h, row, col, ctypes.byref(v)))
        return v.value

    def precondition_Jacobi(self, dst, src, omega):
        _check_pair(dst, src, self.m(), self.dtype)
        check(lib().mfg_operator_precondition_jacobi(self.h, _ptr(dst), _ptr(src), omega))

    def memory_consumption(self):
        return lib().mfg_operator_memory_consumption(self.h)

    def n_colours(self):
        return lib().mfg_operator_n_colours(self.h)

    def enable_timing(self, enable=True):
        check(lib().mfg_operator_enable_timing(self.h, 1 if enable else 0))

    def timing(self):
        """(summed cell-kernel ms, launches) since the last call."""
        ms, n = ctypes.c_double(0), ctypes.c_int64(0)
        check(lib().mfg_operator_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def quadrature_points(self):
        """Physical coordinates of the quadrature points, (n_cells * (p+1)^3, 3), input cell order."""
        import torch
        out = torch.empty(self.n_qp * 3, dtype=torch.float64, device="cuda")
        check(lib().mfg_operator_quadrature_points(self.h, _ptr(out)))
        return out.view(self.n_qp, 3)

    def face_quadrature_points(self, faces):
        """Reference coordinates of the face points, (n_faces * (p+1)^2, 3) FP64 device tensor."""
        import numpy as np
        import torch
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        out = torch.empty((len(f) // 2) * (self._keep.degree + 1) ** 2 * 3, dtype=torch.float64, device="cuda")
        check(lib().mfg_operator_face_quadrature_points(self.h, len(f) // 2, f.ctypes.data_as(_c_i32p), _ptr(out)))
        return out.view(-1, 3)

    def face_values(self, faces, u):
        """Values of the FP64 field u at the face points, (n_faces * (p+1)^2, n_components) device tensor."""
        import numpy as np
        import torch
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        nc = self._keep.n_components
        out = torch.empty((len(f) // 2) * (self._keep.degree + 1) ** 2 * nc, dtype=torch.float64, device="cuda")
        check(lib().mfg_operator_face_values(self.h, len(f) // 2, f.ctypes.data_as(_c_i32p), _ptr(u), _ptr(out)))
        return out.view(-1, nc)

    def set_cell_schedule(self, schedule):
        """CELLS_COLOURS (one launch per colour) or CELLS_GATHER (two launches, bitwise the same result)."""
        check(lib().mfg_operator_set_cell_schedule(self.h, schedule))

    def cell_schedule(self):
        return lib().mfg_operator_cell_schedule(self.h)

    def set_point_storage(self, storage):
        """T2_FOLDED (16 words per point) or T2_GRADIENT (10 words, affine meshes); discards the cache."""
        check(lib().mfg_operator_set_point_storage(self.h, storage))

    def point_storage(self):
        return lib().mfg_operator_point_storage(self.h)

    def n_fields(self):
        """Words stored per quadrature point by the operator's cache."""
        nf = ctypes.c_int32(0)
        check(lib().mfg_operator_qp_data(self.h, None, ctypes.byref(nf)))
        return nf.value

    def qp_data(self):
        import torch
        nf = ctypes.c_int32(0)
        check(lib().mfg_operator_qp_data(self.h, None, ctypes.byref(nf)))
        out = torch.empty(nf.value * self.n_qp, dtype=self.dtype, device="cuda")
        check(lib().mfg_operator_qp_data(self.h, _ptr(out), ctypes.byref(nf)))
        return out.view(nf.value, self.n_qp)


class NeoHookOperator(_Operator):
    """Device counterpart of NeoHookOperator<3, Number> (mf_nh_operator.h:99-237)."""

    def __init__(self, mesh, mu, nu, rho, alpha_1, formulation=1, caching="tensor2", precision=FP64, inclusion=None):
        """inclusion: (mu, nu, rho) of set_material's second material, on the mesh's cells whose material id
        is not 0 (mf_nh_operator.h:121-128)."""
        prm = _solid_params(mu, nu, rho, alpha_1, formulation, CACHING[caching], inclusion)
        h = _vp()
        check(lib().mfg_solid_create(ctypes.byref(mesh.desc), ctypes.byref(prm), precision, ctypes.byref(h)))
        super().__init__(h, precision, mesh)
        self.n_qp = mesh.n_cells * (mesh.degree + 1) ** 3

    def cache(self, u):
        import torch
        prec = FP64 if u.dtype == torch.float64 else FP32
        check(lib().mfg_solid_cache(self.h, _ptr(u), prec))

    def add_interface_traction(self, rhs, u, faces, traction):
        """rhs += int t0 . w dA over faces ((n, 2) int32 cell/face pairs), t0 = |det F F^-T N| t at u;
        traction: FP64 device tensor, 3 values per face point ((p+1)^2 points per face)."""
        import numpy as np
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        check(lib().mfg_solid_add_interface_traction(self.h, len(f) // 2, f.ctypes.data_as(_c_i32p), _ptr(u),
                                                     _ptr(traction), _ptr(rhs)))




class LaplaceOperator(_Operator):
    """Device counterpart of LaplaceOperator<3, number> (heat_transfer.h:81-250): M + dt*c*K."""

    def __init__(self, mesh, dt, coefficient=None, precision=FP64):
        import numpy as np
        coef = None if coefficient is None else np.ascontiguousarray(coefficient, dtype=np.float64)
        h = _vp()
        check(lib().mfg_heat_create(ctypes.byref(mesh.desc), dt,
                                    None if coef is None else coef.ctypes.data_as(_c_dp), precision,
                                    ctypes.byref(h)))
        super().__init__(h, precision, mesh)
        self.n_qp = mesh.n_cells * (mesh.degree + 1) ** 3

    def add_interface_flux(self, rhs, faces, flux):
        """rhs -= dt int q phi dA over faces ((n, 2) cell/face pairs), q: FP64 device tensor per face point."""
        import numpy as np
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        check(lib().mfg_heat_add_interface_flux(self.h, len(f) // 2, f.ctypes.data_as(_c_i32p), _ptr(flux),
                                                _ptr(rhs)))

    def assemble_rhs(self, rhs, u_old, u_dirichlet, f_qp=None):
        """LaplaceProblem::assemble_rhs: rhs = int (u_old + dt f) phi - A_unconstrained u_D, constrained rows 0."""
        check(lib().mfg_heat_assemble_rhs(self.h, _ptr(rhs), _ptr(u_old), _ptr(u_dirichlet),
                                          None if f_qp is None else _ptr(f_qp)))

    def l2_error(self, u, exact_qp):
        """LaplaceProblem::compute_error: sqrt(int (exact - u_h)^2)."""
        e = ctypes.c_double(0.0)
        check(lib().mfg_heat_l2_error(self.h, _ptr(u), _ptr(exact_qp), ctypes.byref(e)))
        return e.value


# ------------------------------------------------------------ multigrid ---
PRECOND_NONE, PRECOND_JACOBI, PRECOND_MG = 0, 1, 2


class MGParams(ctypes.Structure):
    _fields_ = [("cheb_degree", ctypes.c_int32), ("smoothing_range", ctypes.c_double),
                ("eig_cg_iterations", ctypes.c_int32), ("coarse_kind", ctypes.c_int32),
                ("coarse_range", ctypes.c_double), ("coarse_tolerance", ctypes.c_double),
                ("coarse_eig_iterations", ctypes.c_int32)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("initial_residual", ctypes.c_double),
                ("final_residual", ctypes.c_double), ("condition_estimate", ctypes.c_double)]


def box_children(coarse, fine):
    import numpy as np
    out = np.zeros(coarse.n_cells * 8, dtype=np.int32)
    check(lib().mfg_box_mesh_children(coarse.h, fine.h, out.ctypes.data_as(_c_i32p)))
    return out


class Transfer:
    """MGTransferMatrixFree between two nested levels (mf_elasticity.h:863-877)."""

    def __init__(self, coarse_op, fine_op, children):
        import numpy as np
        ch = np.ascontiguousarray(children, dtype=np.int32)
        h = _vp()
        check(lib().mfg_transfer_create(coarse_op.h, fine_op.h, ch.ctypes.data_as(_c_i32p), ctypes.byref(h)))
        self.h, self._keep = h, (coarse_op, fine_op)
        self._destroy = lib().mfg_transfer_destroy

    def __del__(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def prolongate_add(self, fine, coarse):
        check(lib().mfg_transfer_prolongate_add(self.h, _ptr(fine), _ptr(coarse)))

    def restrict_add(self, coarse, fine):
        check(lib().mfg_transfer_restrict_add(self.h, _ptr(coarse), _ptr(fine)))

    def interpolate(self, coarse, fine):
        check(lib().mfg_transfer_interpolate(self.h, _ptr(coarse), _ptr(fine)))


class Multigrid:
    """Multigrid + PreconditionMG with Chebyshev smoothing (mf_elasticity.h:1140-1279)."""

    def __init__(self, ops, transfers, cheb_degree=5, smoothing_range=20.0, eig_cg_iterations=30, coarse_kind=0,
                 coarse_range=1e-3, coarse_tolerance=1e-3, coarse_eig_iterations=0):
        L = len(ops)
        oa = (_vp * L)(*[o.h for o in ops])
        ta = (_vp * max(L - 1, 1))(*([t.h for t in transfers] or [None]))
        self.prm = MGParams(cheb_degree, smoothing_range, eig_cg_iterations, coarse_kind, coarse_range,
                            coarse_tolerance, coarse_eig_iterations)
        h = _vp()
        check(lib().mfg_mg_create(L, oa, ta, ctypes.byref(self.prm), ctypes.byref(h)))
        self.h, self._keep = h, (list(ops), list(transfers))
        self._destroy = lib().mfg_mg_destroy

    def __del__(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def setup(self):
        check(lib().mfg_mg_setup(self.h))

    def use_graph(self, enable=True):
        """Replay the V-cycle's sweeps above the coarse level as captured hipGraphs (default off)."""
        check(lib().mfg_mg_use_graph(self.h, 1 if enable else 0))

    def graph_nodes(self):
        n = ctypes.c_int64(0)
        check(lib().mfg_mg_graph_info(self.h, ctypes.byref(n)))
        return n.value

    def vmult(self, dst, src):
        import torch
        if src.dtype not in (torch.float64, torch.float32):
            raise ValueError("src: expected float64 or float32")
        _check_pair(dst, src, self._keep[0][-1].m(), src.dtype)
        prec = FP64 if src.dtype == torch.float64 else FP32
        check(lib().mfg_mg_vmult(self.h, _ptr(dst), _ptr(src), prec))

    def level_info(self, level):
        lmax, lmin, deg = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int32(0)
        check(lib().mfg_mg_level_info(self.h, level, ctypes.byref(lmax), ctypes.byref(lmin), ctypes.byref(deg)))
        return lmax.value, lmin.value, deg.value


def cg_solve(A, b, x, preconditioner=PRECOND_NONE, mg=None, omega=1.0, tol_rel=1e-8, max_iter=10000):
    """SolverCG::solve; returns SolveInfo.  Raises MfgError(4) on no convergence."""
    _check_vec(b, A.m(), A.dtype, "b")
    _check_vec(x, A.m(), A.dtype, "x")
    if b.data_ptr() == x.data_ptr():
        raise ValueError("b and x must be different vectors")
    info = SolveInfo()
    st = lib().mfg_cg_solve(A.h, preconditioner, None if mg is None else mg.h, omega, _ptr(b), _ptr(x), tol_rel,
                            max_iter, ctypes.byref(info))
    check(st)
    return info


# ------------------------------------------------------ Newton-Newmark ---
class NewmarkParams(ctypes.Structure):
    _fields_ = [("beta", ctypes.c_double), ("gamma", ctypes.c_double), ("dt", ctypes.c_double),
                ("body_force", ctypes.c_double * 3), ("max_newton", ctypes.c_int32), ("min_newton", ctypes.c_int32),
                ("tol_f", ctypes.c_double), ("tol_u", ctypes.c_double), ("tol_f_abs", ctypes.c_double),
                ("tol_u_abs", ctypes.c_double), ("tol_lin", ctypes.c_double), ("max_lin", ctypes.c_int32),
                ("cond_cg_iterations", ctypes.c_int32), ("linear_preconditioner", ctypes.c_int32),
                ("jacobi_relaxation", ctypes.c_double)]


# linear preconditioner of the Newton step ("Preconditioner type", parameter_handling.h:181-182)
NEWTON_PRECONDITIONER = {"gmg": 0, "jacobi": 1, "none": 2}


class StepInfo(ctypes.Structure):
    _fields_ = [("newton_iterations", ctypes.c_int32), ("linear_iterations", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("residual_0", ctypes.c_double), ("residual_final", ctypes.c_double),
                ("update_final", ctypes.c_double), ("seconds", ctypes.c_double),
                ("condition_estimate", ctypes.c_double)]


SIGNATURES.update({
    "mfg_solid_residual": (ctypes.c_int, [_vp, _vp, _vp, _c_dp, _vp, _c_dp]),
    "mfg_solid_problem_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(LevelDesc),
                                                 ctypes.POINTER(_c_i32p), ctypes.POINTER(SolidParams),
                                                 ctypes.POINTER(NewmarkParams), ctypes.POINTER(MGParams),
                                                 ctypes.c_int32, ctypes.POINTER(_vp)]),
    "mfg_solid_problem_destroy": (ctypes.c_int, [_vp]),
    "mfg_solid_problem_set_stream": (ctypes.c_int, [_vp, _vp]),
    "mfg_solid_problem_set_state": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "mfg_solid_problem_get_state": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "mfg_solid_problem_step": (ctypes.c_int, [_vp, ctypes.POINTER(StepInfo)]),
    "mfg_solid_problem_memory": (ctypes.c_size_t, [_vp]),
    "mfg_solid_problem_set_interface": (ctypes.c_int, [_vp, ctypes.c_int64, _c_i32p, _vp]),
})


class SessionSpec(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3), ("n_cells", ctypes.c_int32 * 3),
                ("degree", ctypes.c_int32), ("material", SolidParams), ("newmark", NewmarkParams),
                ("mg", MGParams), ("level_precision", ctypes.c_int32), ("u0_amplitude", ctypes.c_double),
                ("u0_noise", ctypes.c_double), ("seed", ctypes.c_uint64)]


class Request(ctypes.Structure):
    _fields_ = [("session", ctypes.c_int32), ("arrival_s", ctypes.c_double)]


class RequestResult(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("start_s", ctypes.c_double), ("end_s", ctypes.c_double),
                ("newton_iterations", ctypes.c_int32), ("linear_iterations", ctypes.c_int32),
                ("status", ctypes.c_int32), ("residual_final", ctypes.c_double)]


class ServerParams(ctypes.Structure):
    _fields_ = [("max_concurrent", ctypes.c_int32), ("hbm_budget_bytes", ctypes.c_size_t),
                ("real_time", ctypes.c_int32)]


class ServeStats(ctypes.Structure):
    _fields_ = [("makespan_s", ctypes.c_double), ("steps_per_s", ctypes.c_double),
                ("latency_p50_s", ctypes.c_double), ("latency_p99_s", ctypes.c_double),
                ("latency_mean_s", ctypes.c_double), ("completed", ctypes.c_int32), ("failed", ctypes.c_int32),
                ("admissions", ctypes.c_int32), ("evictions", ctypes.c_int32), ("max_resident", ctypes.c_int32),
                ("peak_bytes", ctypes.c_size_t)]


SIGNATURES.update({
    "mfg_serve": (ctypes.c_int, [ctypes.POINTER(SessionSpec), ctypes.c_int32, ctypes.POINTER(Request),
                                 ctypes.c_int32, ctypes.POINTER(ServerParams), ctypes.POINTER(RequestResult),
                                 ctypes.POINTER(_c_dp), ctypes.POINTER(ServeStats)]),
    "mfg_session_n_dofs": (ctypes.c_int64, [ctypes.POINTER(SessionSpec)]),
})


def session_spec(lo, hi, n_cells, degree, mu, nu, rho, beta=0.25, gamma=0.5, dt=1e-3, body=(0.0, 0.0, 0.0),
                 level_precision=FP32, tol_f=1e-8, tol_u=1e-6, tol_lin=1e-6, max_newton=10, min_newton=3,
                 cheb_degree=5, smoothing_range=20.0, eig_cg_iterations=30, coarse_kind=0, u0_amplitude=1e-3,
                 u0_noise=1e-5, seed=0, cond_cg_iterations=20):
    s = SessionSpec()
    s.lo[:] = list(lo)
    s.hi[:] = list(hi)
    s.n_cells[:] = list(n_cells)
    s.degree = degree
    s.material = SolidParams(mu, nu, rho, 0.0, 1, CACHING["tensor2"])
    s.newmark = NewmarkParams(beta, gamma, dt, (ctypes.c_double * 3)(*body), max_newton, min_newton, tol_f, tol_u,
                              0.0, 0.0, tol_lin, 1000, cond_cg_iterations)
    s.mg = MGParams(cheb_degree, smoothing_range, eig_cg_iterations, coarse_kind, 1e-3, 1e-3)
    s.level_precision = level_precision
    s.u0_amplitude, s.u0_noise, s.seed = u0_amplitude, u0_noise, seed
    return s


def serve(specs, requests, max_concurrent=4, hbm_budget_bytes=64 << 30, real_time=True, final_states=False):
    """Runs mfg_serve; returns (results, stats, final_u list or None)."""
    import numpy as np
    n_s, n_r = len(specs), len(requests)
    sa = (SessionSpec * n_s)(*specs)
    ra = (Request * n_r)(*[Request(s, t) for s, t in requests])
    res = (RequestResult * max(n_r, 1))()
    prm = ServerParams(max_concurrent, hbm_budget_bytes, 1 if real_time else 0)
    st = ServeStats()
    finals, fptr = None, None
    if final_states:
        finals = [np.zeros(lib().mfg_session_n_dofs(ctypes.byref(sa[i]))) for i in range(n_s)]
        fptr = (_c_dp * n_s)(*[f.ctypes.data_as(_c_dp) for f in finals])
    check(lib().mfg_serve(sa, n_s, ra, n_r, ctypes.byref(prm), res, fptr, ctypes.byref(st)))
    return list(res)[:n_r], st, finals


def solid_residual(op, u, acc, body=(0.0, 0.0, 0.0)):
    """(R, ||R||) of Solid::assemble_residual at (u, acc) on op's layout (FP64)."""
    import torch
    b = (ctypes.c_double * 3)(*body)
    R = torch.empty_like(u)
    nrm = ctypes.c_double(0)
    check(lib().mfg_solid_residual(op.h, _ptr(u), _ptr(acc), b, _ptr(R), ctypes.byref(nrm)))
    return R, nrm.value


class SolidProblem:
    """Newmark + Newton-Raphson + MG-CG on the device (Solid::solve_nonlinear_timestep)."""

    def __init__(self, meshes, mu, nu, rho, beta=0.25, gamma=0.5, dt=1e-3, body=(0.0, 0.0, 0.0), formulation=1,
                 level_precision=FP64, max_newton=10, min_newton=3, tol_f=1e-8, tol_u=1e-6, tol_f_abs=0.0,
                 tol_u_abs=0.0, tol_lin=1e-6, max_lin=1000, cond_cg_iterations=20, cheb_degree=5, smoothing_range=20.0,
                 eig_cg_iterations=30, coarse_kind=0, coarse_range=1e-3, coarse_tolerance=1e-3, caching="tensor2",
                 inclusion=None, preconditioner="gmg", relaxation=0.65):
        """inclusion: (mu, nu, rho) of the material on cells whose material id is not 0 (every level's mesh
        carries its own ids: set_material_ids, refined_material_ids).  preconditioner: "gmg", "jacobi" (with
        relaxation, the reference's default 0.65) or "none", the reference's "Preconditioner type"."""
        L = len(meshes)
        descs = (LevelDesc * L)(*[m.desc for m in meshes])
        self._children = [box_children(meshes[l - 1], meshes[l]) for l in range(1, L)]
        ch = (_c_i32p * max(L - 1, 1))(*([c.ctypes.data_as(_c_i32p) for c in self._children] or [None]))
        sp = _solid_params(mu, nu, rho, 0.0, formulation, CACHING[caching], inclusion)
        nm = NewmarkParams(beta, gamma, dt, (ctypes.c_double * 3)(*body), max_newton, min_newton, tol_f, tol_u,
                           tol_f_abs, tol_u_abs, tol_lin, max_lin, cond_cg_iterations,
                           NEWTON_PRECONDITIONER[preconditioner], relaxation)
        mp = MGParams(cheb_degree, smoothing_range, eig_cg_iterations, coarse_kind, coarse_range, coarse_tolerance)
        h = _vp()
        check(lib().mfg_solid_problem_create(L, descs, ch, ctypes.byref(sp), ctypes.byref(nm), ctypes.byref(mp),
                                             level_precision, ctypes.byref(h)))
        self.h, self._meshes = h, list(meshes)
        self._destroy = lib().mfg_solid_problem_destroy
        self.n_dofs = meshes[-1].n_dofs

    def __del__(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def set_stream(self, stream):
        check(lib().mfg_solid_problem_set_stream(self.h, ctypes.c_void_p(int(stream.cuda_stream))))

    def set_state(self, u=None, v=None, a=None):
        check(lib().mfg_solid_problem_set_state(self.h, None if u is None else _ptr(u),
                                                None if v is None else _ptr(v), None if a is None else _ptr(a)))

    def get_state(self):
        import torch
        u, v, a = (torch.empty(self.n_dofs, dtype=torch.float64, device="cuda") for _ in range(3))
        check(lib().mfg_solid_problem_get_state(self.h, _ptr(u), _ptr(v), _ptr(a)))
        return u, v, a

    def step(self):
        info = StepInfo()
        check(lib().mfg_solid_problem_step(self.h, ctypes.byref(info)))
        return info

    def set_interface(self, faces, traction):
        """Traction on the finest level's faces ((n, 2) cell/face pairs), 3 values per face point in an FP64
        device tensor that the steps read (refresh its contents between steps); faces of length 0 clear it."""
        import numpy as np
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        self._traction = traction if len(f) else None
        check(lib().mfg_solid_problem_set_interface(self.h, len(f) // 2, f.ctypes.data_as(_c_i32p),
                                                    _ptr(traction) if len(f) else None))

    def memory(self):
        return lib().mfg_solid_problem_memory(self.h)


# ---------------------------------------------------- batched instances ---
SCHEDULE_COLOUR, SCHEDULE_QUEUE, SCHEDULE_STREAMS, SCHEDULE_GATHER = 0, 1, 2, 3
BATCH_SOLID_F1, BATCH_SOLID_F0, BATCH_HEAT, BATCH_SOLID_F1_GRADIENT = 0, 1, 2, 3


class BatchTuning(ctypes.Structure):
    _fields_ = [("schedule", ctypes.c_int32), ("cells_per_task", (ctypes.c_int32 * 7) * 4),
                ("waves_per_simd", ctypes.c_int32), ("ms_colour", ctypes.c_double), ("ms_queue", ctypes.c_double),
                ("ms_streams", ctypes.c_double), ("ms_gather", ctypes.c_double)]

    def as_dict(self):
        kinds = ("solid_f1", "solid_f0", "heat", "solid_f1_gradient")
        names = {SCHEDULE_COLOUR: "colour", SCHEDULE_QUEUE: "queue", SCHEDULE_STREAMS: "streams",
                 SCHEDULE_GATHER: "gather"}
        return {"schedule": names[self.schedule],
                "cells_per_task": {k: list(self.cells_per_task[i])[1:] for i, k in enumerate(kinds)},
                "waves_per_simd": self.waves_per_simd, "ms_colour": self.ms_colour, "ms_queue": self.ms_queue,
                "ms_streams": self.ms_streams, "ms_gather": self.ms_gather}


_vpp = ctypes.POINTER(_vp)
SIGNATURES.update({
    "mfg_batch_create": (ctypes.c_int, [ctypes.c_int32, _vpp, _vpp]),
    "mfg_batch_destroy": (ctypes.c_int, [_vp]),
    "mfg_batch_set_stream": (ctypes.c_int, [_vp, _vp]),
    "mfg_batch_set_schedule": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mfg_batch_set_cells_per_task": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "mfg_batch_vmult": (ctypes.c_int, [_vp, _vpp, _vpp]),
    "mfg_batch_vmult_add": (ctypes.c_int, [_vp, _vpp, _vpp]),
    "mfg_batch_synchronize": (ctypes.c_int, [_vp]),
    "mfg_batch_tune": (ctypes.c_int, [_vp, _vpp, _vpp, ctypes.c_int32, ctypes.POINTER(BatchTuning)]),
    "mfg_batch_get_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(BatchTuning)]),
    "mfg_batch_solver_create": (ctypes.c_int, [ctypes.c_int32, _vpp, _vpp, _vpp]),
    "mfg_batch_solver_destroy": (ctypes.c_int, [_vp]),
    "mfg_batch_solver_use_graph": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mfg_batch_solver_set_schedule": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mfg_batch_solver_tune": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(BatchTuning),
                                             ctypes.POINTER(BatchTuning)]),
    "mfg_batch_solver_precondition": (ctypes.c_int, [_vp, _vpp, _vpp]),
    "mfg_batch_solver_solve": (ctypes.c_int, [_vp, _vpp, _vpp, ctypes.c_double, ctypes.c_int32,
                                              ctypes.POINTER(SolveInfo), _c_i32p]),
    "mfg_batch_solver_info": (ctypes.c_int, [_vp, _c_i32p, _c_i32p]),
})


def _ptr_array(ts):
    return (_vp * len(ts))(*[_ptr(t) for t in ts])


class Batch:
    """Many operators (one precision) applied in one launch per vmult."""

    def __init__(self, ops):
        arr = (_vp * len(ops))(*[o.h for o in ops])
        h = _vp()
        check(lib().mfg_batch_create(len(ops), arr, ctypes.byref(h)))
        self.h, self._keep, self.n = h, list(ops), len(ops)
        self._destroy = lib().mfg_batch_destroy

    def __del__(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def set_schedule(self, schedule):
        check(lib().mfg_batch_set_schedule(self.h, schedule))

    def set_cells_per_task(self, kind, degree, cells):
        check(lib().mfg_batch_set_cells_per_task(self.h, kind, degree, cells))

    def vmult(self, dsts, srcs):
        check(lib().mfg_batch_vmult(self.h, _ptr_array(dsts), _ptr_array(srcs)))

    def vmult_add(self, dsts, srcs):
        check(lib().mfg_batch_vmult_add(self.h, _ptr_array(dsts), _ptr_array(srcs)))

    def synchronize(self):
        check(lib().mfg_batch_synchronize(self.h))

    def tune(self, srcs, dsts, repetitions=5):
        t = BatchTuning()
        check(lib().mfg_batch_tune(self.h, _ptr_array(srcs), _ptr_array(dsts), repetitions, ctypes.byref(t)))
        return t

    def tuning(self):
        t = BatchTuning()
        check(lib().mfg_batch_get_tuning(self.h, ctypes.byref(t)))
        return t


class BatchSolver:
    """MG-preconditioned CG on many instances in lockstep; the V-cycle of all
    instances is one program of batched launches replayed as a hipGraph."""

    def __init__(self, fine_ops, mgs):
        n = len(fine_ops)
        fa = (_vp * n)(*[o.h for o in fine_ops])
        ma = (_vp * n)(*[m.h for m in mgs])
        h = _vp()
        check(lib().mfg_batch_solver_create(n, fa, ma, ctypes.byref(h)))
        self.h, self._keep, self.n = h, (list(fine_ops), list(mgs)), n
        self._destroy = lib().mfg_batch_solver_destroy

    def __del__(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def use_graph(self, on=True):
        check(lib().mfg_batch_solver_use_graph(self.h, 1 if on else 0))

    def set_schedule(self, schedule):
        check(lib().mfg_batch_solver_set_schedule(self.h, schedule))

    def tune(self, repetitions=5):
        f, l = BatchTuning(), BatchTuning()
        check(lib().mfg_batch_solver_tune(self.h, repetitions, ctypes.byref(f), ctypes.byref(l)))
        return f, l

    def precondition(self, dsts, srcs):
        check(lib().mfg_batch_solver_precondition(self.h, _ptr_array(dsts), _ptr_array(srcs)))

    def solve(self, bs, xs, tol_rel=1e-8, max_iter=1000):
        """Returns (list of SolveInfo, list of status codes)."""
        infos = (SolveInfo * self.n)()
        status = (ctypes.c_int32 * self.n)()
        check(lib().mfg_batch_solver_solve(self.h, _ptr_array(bs), _ptr_array(xs), tol_rel, max_iter, infos,
                                           status))
        return list(infos), list(status)

    def info(self):
        a, b = ctypes.c_int32(0), ctypes.c_int32(0)
        check(lib().mfg_batch_solver_info(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


class Instance:
    """One problem of a batch on [lo, hi]: the fine operator in fine_precision
    and a geometric MG hierarchy in level_precision (levels halve the cell
    counts while all stay even).  Solid: neo-Hookean tensor2 linearised at
    u0 = amp sin sin sin + noise (levels at the interpolated u), Dirichlet on
    x = lo.  Heat: M + dt c K with constant conductivity c, Dirichlet on all
    faces.  point_storage: the tensor2 point storage of the formulation-1
    solid fine operator; folded by default, since the fused batched kernel
    (one register allocation for all degrees) runs the 10-word storage slower
    at several cells per task."""

    def __init__(self, kind, n_cells, degree, fine_precision=FP64, level_precision=FP32, lo=(0.0, 0.0, 0.0),
                 hi=(1.0, 1.0, 1.0), dt=0.1, conductivity=1.0, formulation=1, cheb_degree=5, smoothing_range=None,
                 eig_cg_iterations=None, coarse_eig_iterations=0, amp=1e-3, noise=1e-5, seed=0,
                 material=(384615.38461538462, 0.3, 1000.0, 4e6), cell_schedule=None, numbering="cellwise",
                 point_storage=T2_FOLDED):
        import numpy as np
        import torch
        self.kind, self.degree, self.conductivity = kind, degree, conductivity
        counts = [list(n_cells)]
        while all(c % 2 == 0 for c in counts[0]):
            counts.insert(0, [c // 2 for c in counts[0]])
        solid = kind == "solid"
        nc, fm, cm = (3, 1, 7) if solid else (1, 63, 1)
        self.meshes = [BoxMesh(lo, hi, c, degree, nc, fm, cm, numbering=numbering) for c in counts]
        fine_mesh = self.meshes[-1]
        mu, nu, rho, alpha_1 = material
        if solid:
            make = lambda m, pr: NeoHookOperator(m, mu, nu, rho, alpha_1, formulation=formulation, precision=pr)
        else:
            def make(m, pr):
                coef = np.full(m.n_cells * (degree + 1) ** 3, conductivity)
                return LaplaceOperator(m, dt, coef, precision=pr)
        self.levels = [make(m, level_precision) for m in self.meshes]
        self.fine = make(fine_mesh, fine_precision)
        if solid and formulation == 1:
            self.fine.set_point_storage(point_storage)
        if cell_schedule is not None:
            for op in self.levels + [self.fine]:
                op.set_cell_schedule(cell_schedule)
        self.transfers = [Transfer(self.levels[l - 1], self.levels[l], box_children(self.meshes[l - 1], self.meshes[l]))
                          for l in range(1, len(self.meshes))]
        if solid:
            u = fine_mesh.u0(amp=amp, noise=noise, seed=seed)
            self.fine.cache(torch.from_numpy(u).to("cuda", _dtype_of(fine_precision)))
            ug = torch.from_numpy(u).to("cuda", _dtype_of(level_precision))
            us = [ug]
            for l in range(len(self.meshes) - 1, 0, -1):
                c = torch.empty(self.meshes[l - 1].n_dofs, dtype=ug.dtype, device="cuda")
                self.transfers[l - 1].interpolate(c, us[0])
                us.insert(0, c)
            for op, uu in zip(self.levels, us):
                op.cache(uu)
        sr = smoothing_range if smoothing_range is not None else (20.0 if solid else 15.0)
        ei = eig_cg_iterations if eig_cg_iterations is not None else (30 if solid else 10)
        self.mg = Multigrid(self.levels, self.transfers, cheb_degree=cheb_degree, smoothing_range=sr,
                            eig_cg_iterations=ei, coarse_kind=0, coarse_eig_iterations=coarse_eig_iterations)
        self.mg.setup()
        self.n_dofs = fine_mesh.n_dofs
        self.constrained = fine_mesh.constrained()
