"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker / CPU baseline.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


class MGParams(ctypes.Structure):
    _fields_ = [("cheb_degree", ctypes.c_int), ("smoothing_range", ctypes.c_double),
                ("eig_cg_iterations", ctypes.c_int), ("coarse_kind", ctypes.c_int),
                ("tol_rel", ctypes.c_double), ("max_iter", ctypes.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        # MFG_ORACLE_LIB: a copy built with a planted mistake (mutation tests only)
        path = os.environ.get("MFG_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        sig = {
            "orc_gll_points": (None, [ctypes.c_int, _dp]),
            "orc_gauss_rule": (None, [ctypes.c_int, _dp, _dp]),
            "orc_lagrange": (None, [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp]),
            "orc_mesh_create": (_vp, [_dp, _dp, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "orc_mesh_create_bc": (_vp, [_dp, _dp, _ip, ctypes.c_int, ctypes.c_int, _ip]),
            "orc_newmark_step_iface": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                      ctypes.POINTER(MGParams), _dp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                                      _dp, ctypes.POINTER(ctypes.c_int), ctypes.c_int64, _ip, _dp]),
            "orc_mesh_destroy": (None, [_vp]),
            "orc_mesh_sizes": (None, [_vp, _lp]),
            "orc_mesh_cell_nodes": (None, [_vp, _ip]),
            "orc_mesh_constrained": (None, [_vp, _ip]),
            "orc_mesh_node_coords": (None, [_vp, _dp]),
            "orc_mesh_cell_vertices": (None, [_vp, _dp]),
            "orc_mesh_set_cell_vertices": (None, [_vp, _dp]),
            "orc_fill_u0": (None, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _dp]),
            "orc_fill_uniform": (None, [ctypes.c_uint64, ctypes.c_int64, _dp]),
            "orc_solid_create": (_vp, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int]),
            "orc_solid_destroy": (None, [_vp]),
            "orc_solid_set_inclusion": (None, [_vp, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "orc_solid_cache": (ctypes.c_int, [_vp, _dp]),
            "orc_solid_vmult": (None, [_vp, _dp, _dp]),
            "orc_solid_vmult_bruteforce": (None, [_vp, _dp, _dp]),
            "orc_solid_diagonal": (None, [_vp, _dp]),
            "orc_solid_tensor2_data": (None, [_vp, _dp]),
            "orc_solid_residual": (None, [_vp, _dp, _dp, _dp, _dp]),
            "orc_heat_create": (_vp, [_vp, ctypes.c_int, ctypes.c_double, _dp]),
            "orc_heat_destroy": (None, [_vp]),
            "orc_heat_vmult": (None, [_vp, _dp, _dp]),
            "orc_heat_vmult_bruteforce": (None, [_vp, _dp, _dp]),
            "orc_heat_vmult_noconstraints": (None, [_vp, _dp, _dp]),
            "orc_heat_diagonal": (None, [_vp, _dp]),
            "orc_heat_load": (None, [_vp, ctypes.c_double, _dp]),
            "orc_operator_mesh": (_vp, [ctypes.c_int, _vp]),
            "orc_prolongate_add": (None, [_vp, _vp, _dp, _dp]),
            "orc_restrict_add": (None, [_vp, _vp, _dp, _dp]),
            "orc_interpolate": (None, [_vp, _vp, _dp, _dp]),
            "orc_mg_cg_solve": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp),
                                               ctypes.POINTER(_vp), ctypes.POINTER(MGParams), _dp, _dp, _dp]),
            "orc_newmark_step": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                ctypes.POINTER(MGParams), _dp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                                _dp, ctypes.POINTER(ctypes.c_int)]),
            "orc_mg_vmult": (None, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                    ctypes.POINTER(MGParams), _dp, _dp, _dp]),
            "orc_cg_solve": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_int, ctypes.c_double, _dp, _dp,
                                            ctypes.c_double, ctypes.c_int, _dp]),
            "orc_cheb_eigenvalues": (None, [ctypes.c_int, _vp, _vp, ctypes.c_int, _dp, _dp]),
            "orc_cheb_dense": (None, [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      _dp, _dp, _dp]),
            "orc_interface_traction": (None, [_vp, ctypes.c_int64, _ip, _dp, _dp, _dp]),
            "orc_face_values": (None, [_vp, ctypes.c_int64, _ip, _dp, _dp]),
            "orc_interface_flux": (None, [_vp, ctypes.c_int64, _ip, _dp, ctypes.c_double, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


# material and Newmark constants of the benchmark configurations
E_MOD, NU, RHO, BETA, DT = 1.0e6, 0.3, 1000.0, 0.25, 1.0e-3
MU = E_MOD / (2.0 * (1.0 + NU))
ALPHA1 = 1.0 / (BETA * DT * DT)


def gll_points(p):
    x = np.zeros(p + 1)
    lib().orc_gll_points(p, dptr(x))
    return x


def gauss_rule(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_gauss_rule(n, dptr(x), dptr(w))
    return x, w


def lagrange(z, x):
    z = np.ascontiguousarray(z, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    val = np.zeros((len(x), len(z)))
    der = np.zeros((len(x), len(z)))
    lib().orc_lagrange(len(z), dptr(z), len(x), dptr(x), dptr(val), dptr(der))
    return val, der


def fill_uniform(seed, n):
    v = np.zeros(n)
    lib().orc_fill_uniform(seed, n, dptr(v))
    return v


class Mesh:
    def __init__(self, lo, hi, n, p, nc, face_mask=1, comp_mask=7, face_comp=None):
        """face_comp: optional component mask per boundary face (6 ints), which
        replaces (face_mask, comp_mask)."""
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        nn = np.ascontiguousarray(n, dtype=np.int32)
        self.lo, self.hi, self.n, self.p, self.nc = lo, hi, list(n), p, nc
        if face_comp is None:
            face_comp = [comp_mask if (face_mask >> f) & 1 else 0 for f in range(6)]
        self.face_comp = np.ascontiguousarray(face_comp, dtype=np.int32)
        self.h = lib().orc_mesh_create_bc(dptr(lo), dptr(hi), iptr(nn), p, nc, iptr(self.face_comp))
        s = np.zeros(6, dtype=np.int64)
        lib().orc_mesh_sizes(self.h, s.ctypes.data_as(_lp))
        self.n_cells, self.n_nodes, self.n_dofs, self.n_constrained, self.nodes_per_cell = (int(v) for v in s[:5])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_destroy(self.h)
            self.h = None

    def cell_nodes(self):
        out = np.zeros(self.n_cells * self.nodes_per_cell, dtype=np.int32)
        lib().orc_mesh_cell_nodes(self.h, iptr(out))
        return out

    def constrained(self):
        out = np.zeros(self.n_constrained, dtype=np.int32)
        lib().orc_mesh_constrained(self.h, iptr(out))
        return out

    def node_coords(self):
        out = np.zeros(self.n_nodes * 3)
        lib().orc_mesh_node_coords(self.h, dptr(out))
        return out.reshape(-1, 3)

    def cell_vertices(self):
        out = np.zeros(self.n_cells * 24)
        lib().orc_mesh_cell_vertices(self.h, dptr(out))
        return out

    def set_cell_vertices(self, v):
        """New cell vertices (n_cells*8*3); nodes move to the trilinear images of their GLL points."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == self.n_cells * 24
        lib().orc_mesh_set_cell_vertices(self.h, dptr(v))

    def boundary_faces(self, boundary_id):
        """(cell, face) pairs on box boundary id 2 d (low side) or 2 d + 1 (high side) of direction d."""
        axis, side = boundary_id // 2, boundary_id % 2
        a, b = (1, 2) if axis == 0 else ((0, 2) if axis == 1 else (0, 1))
        n = self.n
        out = []
        for jb in range(n[b]):
            for ja in range(n[a]):
                cc = [0, 0, 0]
                cc[axis] = n[axis] - 1 if side else 0
                cc[a], cc[b] = ja, jb
                out += [cc[0] + n[0] * (cc[1] + n[1] * cc[2]), boundary_id]
        return np.array(out, dtype=np.int32)

    def interface_traction(self, faces, u, t):
        """rhs of the traction term: int t0 . w dA, t0 = |det F F^-T N| t (3 per face point)."""
        faces = np.ascontiguousarray(faces, dtype=np.int32)
        u = np.ascontiguousarray(u, dtype=np.float64)
        t = np.ascontiguousarray(t, dtype=np.float64)
        rhs = np.zeros(self.n_dofs)
        lib().orc_interface_traction(self.h, len(faces) // 2, iptr(faces), dptr(u), dptr(t), dptr(rhs))
        return rhs

    def face_values(self, faces, u):
        """Field values at the face points, (n_faces * N^2, nc)."""
        faces = np.ascontiguousarray(faces, dtype=np.int32)
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(len(faces) // 2 * (self.p + 1) ** 2 * self.nc)
        lib().orc_face_values(self.h, len(faces) // 2, iptr(faces), dptr(u), dptr(out))
        return out.reshape(-1, self.nc)

    def interface_flux(self, faces, flux, dt):
        """Heat rhs term -dt int q phi dA (q per face point)."""
        faces = np.ascontiguousarray(faces, dtype=np.int32)
        flux = np.ascontiguousarray(flux, dtype=np.float64)
        rhs = np.zeros(self.n_dofs)
        lib().orc_interface_flux(self.h, len(faces) // 2, iptr(faces), dptr(flux), dt, dptr(rhs))
        return rhs

    def u0(self, amp=1e-3, noise=1e-5, seed=0):
        u = np.zeros(self.n_dofs)
        lib().orc_fill_u0(self.h, amp, noise, seed, dptr(u))
        return u


class Solid:
    KIND = 0

    def __init__(self, mesh, n_q=None, mu=MU, nu=NU, rho=RHO, alpha1=ALPHA1, formulation=1, caching=1,
                 inclusion=None, material_ids=None):
        """inclusion: (mu, nu, rho) of the material on the cells whose entry of material_ids (one per cell)
        is not 0."""
        self.mesh = mesh
        n_q = mesh.p + 1 if n_q is None else n_q
        self.mu, self.nu, self.rho, self.alpha1 = mu, nu, rho, alpha1
        self.h = lib().orc_solid_create(mesh.h, n_q, mu, nu, rho, alpha1, formulation, caching)
        if not self.h:
            raise ValueError("unsupported solid operator configuration")
        self.inclusion = None
        if inclusion is not None and material_ids is not None:
            self._ids = np.ascontiguousarray(material_ids, dtype=np.int32).reshape(-1)
            assert len(self._ids) == mesh.n_cells
            self.inclusion = tuple(inclusion)
            lib().orc_solid_set_inclusion(self.h, self._ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          *self.inclusion)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solid_destroy(self.h)
            self.h = None

    def cache(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if lib().orc_solid_cache(self.h, dptr(u)) != 0:
            raise ValueError("det_F is not positive")

    def vmult(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros_like(src)
        lib().orc_solid_vmult(self.h, dptr(dst), dptr(src))
        return dst

    def vmult_bruteforce(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros_like(src)
        lib().orc_solid_vmult_bruteforce(self.h, dptr(dst), dptr(src))
        return dst

    def diagonal(self):
        d = np.zeros(self.mesh.n_dofs)
        lib().orc_solid_diagonal(self.h, dptr(d))
        return d

    def tensor2_data(self):
        nq = self.mesh.n_cells * (self.mesh.p + 1) ** 3
        out = np.zeros(18 * nq)
        lib().orc_solid_tensor2_data(self.h, dptr(out))
        return out.reshape(18, nq)

    def residual(self, u, acc, body=(0.0, 0.0, 0.0)):
        u = np.ascontiguousarray(u, dtype=np.float64)
        acc = np.ascontiguousarray(acc, dtype=np.float64)
        b = np.ascontiguousarray(body, dtype=np.float64)
        R = np.zeros_like(u)
        lib().orc_solid_residual(self.h, dptr(u), dptr(acc), dptr(b), dptr(R))
        return R


class Heat:
    KIND = 1

    def __init__(self, mesh, n_q=None, dt=0.1, coef=None):
        self.mesh = mesh
        n_q = mesh.p + 1 if n_q is None else n_q
        self._coef = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
        self.h = lib().orc_heat_create(mesh.h, n_q, dt, None if self._coef is None else dptr(self._coef))
        if not self.h:
            raise ValueError("unsupported heat operator configuration")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_heat_destroy(self.h)
            self.h = None

    def vmult(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros_like(src)
        lib().orc_heat_vmult(self.h, dptr(dst), dptr(src))
        return dst

    def vmult_noconstraints(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros_like(src)
        lib().orc_heat_vmult_noconstraints(self.h, dptr(dst), dptr(src))
        return dst

    def vmult_bruteforce(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros_like(src)
        lib().orc_heat_vmult_bruteforce(self.h, dptr(dst), dptr(src))
        return dst

    def diagonal(self):
        d = np.zeros(self.mesh.n_dofs)
        lib().orc_heat_diagonal(self.h, dptr(d))
        return d

    def load(self, f):
        b = np.zeros(self.mesh.n_dofs)
        lib().orc_heat_load(self.h, f, dptr(b))
        return b


def prolongate_add(cmesh, fmesh, c, f):
    lib().orc_prolongate_add(cmesh.h, fmesh.h, dptr(c), dptr(f))


def restrict_add(cmesh, fmesh, f, c):
    lib().orc_restrict_add(cmesh.h, fmesh.h, dptr(f), dptr(c))


def interpolate(cmesh, fmesh, f):
    c = np.zeros(cmesh.n_dofs)
    lib().orc_interpolate(cmesh.h, fmesh.h, dptr(np.ascontiguousarray(f)), dptr(c))
    return c


def mg_cg_solve(meshes, ops, b, cheb_degree=5, smoothing_range=20.0, eig_cg_iterations=30, coarse_kind=0,
                tol_rel=1e-8, max_iter=1000, x0=None):
    L = len(meshes)
    mh = (_vp * L)(*[m.h for m in meshes])
    oh = (_vp * L)(*[o.h for o in ops])
    prm = MGParams(cheb_degree, smoothing_range, eig_cg_iterations, coarse_kind, tol_rel, max_iter)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    rel = ctypes.c_double(0.0)
    it = lib().orc_mg_cg_solve(ops[0].KIND, L, mh, oh, ctypes.byref(prm), dptr(np.ascontiguousarray(b)), dptr(x),
                               ctypes.byref(rel))
    return x, it, rel.value


def mg_vmult(meshes, ops, src, cheb_degree=5, smoothing_range=20.0, eig_cg_iterations=30, coarse_kind=0):
    """One V-cycle after set-up; returns (dst, [(lmax, lmin, degree) per level])."""
    L = len(meshes)
    mh = (_vp * L)(*[m.h for m in meshes])
    oh = (_vp * L)(*[o.h for o in ops])
    prm = MGParams(cheb_degree, smoothing_range, eig_cg_iterations, coarse_kind, 0.0, 0)
    dst = np.zeros_like(src)
    eig = np.zeros(3 * L)
    lib().orc_mg_vmult(ops[0].KIND, L, mh, oh, ctypes.byref(prm), dptr(np.ascontiguousarray(src)), dptr(dst),
                       dptr(eig))
    return dst, [tuple(eig[3 * l:3 * l + 3]) for l in range(L)]


def newmark_step(meshes, ops, u, v, a, beta=BETA, gamma=0.5, dt=DT, body=(0.0, 0.0, 0.0), tol_f=1e-8, tol_u=1e-6,
                 tol_f_abs=0.0, tol_u_abs=0.0, tol_lin=1e-6, max_newton=10, min_newton=3, cheb_degree=5,
                 smoothing_range=20.0, eig_cg_iterations=30, coarse_kind=0, faces=None, traction=None):
    """One Newton-Newmark step in place on (u, v, a); returns (newton_its, linear_its).  faces / traction:
    an interface ((cell, face) pairs of the finest mesh, 3 traction values per face point) whose traction
    term every residual adds."""
    L = len(meshes)
    mh = (_vp * L)(*[m.h for m in meshes])
    oh = (_vp * L)(*[o.h for o in ops])
    prm = MGParams(cheb_degree, smoothing_range, eig_cg_iterations, coarse_kind, 1e-6, 1000)
    nm = np.array([beta, gamma, dt, *body, tol_f, tol_u, tol_f_abs, tol_u_abs, tol_lin], dtype=np.float64)
    li = ctypes.c_int(0)
    if faces is None:
        f = np.zeros(2, dtype=np.int32)
        t = np.zeros(3)
        nf = 0
    else:
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        t = np.ascontiguousarray(traction, dtype=np.float64).reshape(-1)
        nf = len(f) // 2
    it = lib().orc_newmark_step_iface(L, mh, oh, ctypes.byref(prm), dptr(nm), max_newton, min_newton, dptr(u),
                                      dptr(v), dptr(a), ctypes.byref(li), nf, iptr(f), dptr(t))
    return it, li.value


def cg_solve(op, b, precond=0, omega=1.0, tol_rel=1e-8, max_iter=10000):
    x = np.zeros_like(b)
    rel = ctypes.c_double(0.0)
    it = lib().orc_cg_solve(op.KIND, op.h, precond, omega, dptr(np.ascontiguousarray(b)), dptr(x), tol_rel,
                            max_iter, ctypes.byref(rel))
    return x, it, rel.value


def cheb_eigenvalues(op, n_it):
    lmax, lmin = ctypes.c_double(0), ctypes.c_double(0)
    lib().orc_cheb_eigenvalues(op.KIND, op.h, op.mesh.h, n_it, ctypes.byref(lmax), ctypes.byref(lmin))
    return lmax.value, lmin.value


def cheb_dense(A, b, degree, smoothing_range, eig_it, x0=None):
    """One Chebyshev application on the dense matrix A: x = p(D^-1 A) D^-1 b
    from a zero guess (x0 None), else x0 + p(D^-1 A) D^-1 (b - A x0).
    Returns (x, dict(lmax, lmin, theta, delta, degree))."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    info = np.zeros(5)
    lib().orc_cheb_dense(n, dptr(A), degree, smoothing_range, eig_it, 1 if x0 is None else 0, dptr(b), dptr(x),
                         dptr(info))
    return x, dict(lmax=info[0], lmin=info[1], theta=info[2], delta=info[3], degree=int(info[4]))


def _simd_lib():
    """The cross-cell SIMD baseline library for this host: AVX-512 if the CPU
    has it, else AVX2."""
    global _SIMD
    if _SIMD is None:
        flags = ""
        try:
            with open("/proc/cpuinfo") as f:
                flags = next((l for l in f if l.startswith("flags")), "")
        except OSError:
            pass
        isa = "avx512" if " avx512f" in flags else "avx2"
        L = ctypes.CDLL(os.path.join(_HERE, f"liboracle_simd_{isa}.so"))
        L.orc_simd_solid_create.restype = _vp
        L.orc_simd_solid_create.argtypes = [ctypes.c_int, _ip, ctypes.c_int64, _ip, ctypes.c_int64, _ip, _dp,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_simd_solid_destroy.argtypes = [_vp]
        L.orc_simd_solid_vmult.argtypes = [_vp, _dp, _dp]
        L.orc_simd_lanes.restype = ctypes.c_int
        _SIMD = (L, isa)
    return _SIMD


_SIMD = None


class SimdSolid:
    """The tangent vmult of a cached tensor2 formulation-1 Solid (degree 2 or
    4) computed the way the reference's CPU path computes it: batches of 8
    cells in SIMD lanes, OpenMP over the batches of a colour (the CPU
    baseline of the benchmark; equals Solid.vmult)."""

    def __init__(self, solid):
        if solid.inclusion is not None:
            raise ValueError("the SIMD restatement takes one material")
        L, self.isa = _simd_lib()
        self.lanes = L.orc_simd_lanes()
        m = solid.mesh
        f = solid.tensor2_data()
        n = np.ascontiguousarray(m.n, dtype=np.int32)
        cn = np.ascontiguousarray(m.cell_nodes(), dtype=np.int32)
        con = np.ascontiguousarray(m.constrained(), dtype=np.int32)
        self.n_dofs = m.n_dofs
        self.h = L.orc_simd_solid_create(m.p, iptr(n), m.n_dofs, iptr(cn), len(con), iptr(con),
                                         dptr(np.ascontiguousarray(f)), solid.mu, solid.nu, solid.rho * solid.alpha1)
        if not self.h:
            raise ValueError("the SIMD baseline covers degrees 2 and 4")
        self._L = L

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_simd_solid_destroy(self.h)
            self.h = None

    def vmult(self, src, dst=None):
        dst = np.empty(self.n_dofs) if dst is None else dst
        self._L.orc_simd_solid_vmult(self.h, dptr(dst), dptr(np.ascontiguousarray(src)))
        return dst


def box_levels(lo, hi, coarse_n, n_levels, p, nc, face_mask=1, comp_mask=7, face_comp=None):
    """Nested structured boxes: level l has coarse_n * 2^l cells per direction."""
    return [Mesh(lo, hi, [c << l for c in coarse_n], p, nc, face_mask, comp_mask, face_comp=face_comp)
            for l in range(n_levels)]
