"""The reference's Turek-Hron flap (turek_gmg_*: turek_hron.h:22-90 with the
.prm of tests/solid/turek_gmg_tensor2) as its 3D variant, shared by the CPU
and GPU tests.

2D case: [0.24899, 0.6] x [0.19, 0.21], 5 x 1 coarse cells refined 3 times,
Q1 / QGauss(2), x = lo clamped, the other three sides (boundary ids 1, 2, 3)
loaded by a constant traction (0, -4) (dummy_tester.cpp:193-197), body force
(0, -0.5), mu = 0.5e6, nu = 0.4, rho = 1000, Newmark beta 0.25 / gamma 0.5,
dt = 0.1, two steps; Newton tolerances 1e-9 / 1e-6 relative, 2e-9 / 5e-15
absolute; linear tolerance 1e-6; GMG with Chebyshev degree 5, range 20, 30
eigenvalue iterations, Chebyshev coarse solver with accurate eigenvalues.

3D variant (turek_hron.h:30-38, 52-58): z in [-0.005, 0.005] with one coarse
cell, the z component held on the two z faces.  Formulation 1 does not depend
on the dimension, and the plane-strain field u(x, y) (u_z = 0), constant in
z, satisfies the 3D discrete equations exactly: every 3D residual entry is the
2D entry times the integral of the z shape function, w_k.  Relative norms are
therefore the 2D ones, and the absolute tolerances scale by the same factors:
||R_3D|| = ||R_2D|| ||w||_2 and ||du_3D|| = ||du_2D|| sqrt(n_z + 1) (one copy
per z node).  The goldens (turek_gmg_tensor2_serial.output:10-17, the same for
tensor4, tensor4_ns and scalar_referential): 4 Newton iterations per step,
watchpoint (0.6, 0.2) displacement (-1.29017e-05, -0.00293199) at t = 0.1 and
(-0.000310726, -0.0140092) at t = 0.2; "Average CG iter = 29"."""
import numpy as np

LO, HI = [0.24899, 0.19, -0.005], [0.6, 0.21, 0.005]
COARSE = [5, 1, 1]
LEVELS = 4                       # coarse mesh + 3 global refinements
FACE_COMP = [7, 0, 0, 0, 4, 4]   # x = lo clamped; z faces: z component
INTERFACE_IDS = (1, 2, 3)        # x = hi, y = lo, y = hi
MU, NU, RHO = 0.5e6, 0.4, 1000.0
BETA, GAMMA, DT = 0.25, 0.5, 0.1
BODY = (0.0, -0.5, 0.0)
TRACTION = (0.0, -4.0, 0.0)
WATCH = (0.6, 0.2)
GOLDEN = {1: (-1.29017e-05, -0.00293199), 2: (-0.000310726, -0.0140092)}
GOLDEN_NEWTON = 4
GOLDEN_AVG_CG = 29


def abs_tolerance_scales(n_z):
    """(||w||_2, sqrt(n_z + 1)) for the z shape functions of n_z Q1 cells."""
    hz = (HI[2] - LO[2]) / n_z
    w = np.full(n_z + 1, hz)
    w[0] = w[-1] = hz / 2
    return float(np.sqrt((w * w).sum())), float(np.sqrt(n_z + 1))


def newton_params(n_z):
    sr, su = abs_tolerance_scales(n_z)
    return dict(tol_f=1e-9, tol_u=1e-6, tol_f_abs=2e-9 * sr, tol_u_abs=5e-15 * su, tol_lin=1e-6, max_newton=10,
                min_newton=3)


def mg_params():
    return dict(cheb_degree=5, smoothing_range=20.0, eig_cg_iterations=30, coarse_kind=0)


def watch_nodes(node_coords):
    X = np.asarray(node_coords)
    return np.where((np.abs(X[:, 0] - WATCH[0]) < 1e-12) & (np.abs(X[:, 1] - WATCH[1]) < 1e-12))[0]


def six_digits(a, b):
    """Equal when printed with 6 significant digits (std::cout's default)."""
    return float(f"{a:.6g}") == float(f"{b:.6g}")
