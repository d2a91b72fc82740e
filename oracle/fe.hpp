// CPU oracle (test infrastructure only; see oracle.h): 1D finite-element
// ingredients, structured box meshes and small dense helpers.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace orc {

using std::vector;

inline uint64_t splitmix64(uint64_t seed, uint64_t index) {
  uint64_t z = seed * 0xD1B54A32D192ED03ull + (index + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// U[-1,1): (x >> 11) * 2^-53 -> U[0,1) -> 2U-1
inline double uniform_pm1(uint64_t seed, uint64_t index) {
  return 2.0 * (double)(splitmix64(seed, index) >> 11) * (1.0 / 9007199254740992.0) - 1.0;
}

// Legendre P_n and P_n' at x in [-1,1] by the three-term recurrence.
inline void legendre(int n, double x, double &p, double &dp) {
  double p0 = 1.0, p1 = x;
  if (n == 0) { p = 1.0; dp = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    double pk = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
    p0 = p1; p1 = pk;
  }
  p = p1;
  dp = (std::abs(1 - x * x) < 1e-300) ? 0.0 : n * (p0 - x * p1) / (1 - x * x);
}

// Gauss-Legendre rule with n points on [0,1].
inline void gauss_rule(int n, double *x, double *w) {
  for (int i = 0; i < n; ++i) {
    double t = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre(n, t, p, dp);
      double dt = p / dp;
      t -= dt;
      if (std::abs(dt) < 1e-16) break;
    }
    double p, dp;
    legendre(n, t, p, dp);
    x[n - 1 - i] = 0.5 * (t + 1.0);
    w[n - 1 - i] = 1.0 / ((1 - t * t) * dp * dp);  // 2/((1-t^2)P'^2) scaled by 1/2
  }
}

// Gauss-Lobatto points (p+1 of them) on [0,1]: endpoints and roots of P_p'.
inline void gll_points(int p, double *x) {
  if (p == 0) { x[0] = 0.5; return; }
  const int N = p;
  vector<double> t(N + 1), told(N + 1);
  for (int i = 0; i <= N; ++i) t[i] = std::cos(M_PI * i / N);
  for (int it = 0; it < 200; ++it) {
    double maxd = 0;
    for (int i = 0; i <= N; ++i) {
      // P_N and P_{N-1}
      double p0 = 1.0, p1 = t[i];
      for (int k = 2; k <= N; ++k) {
        double pk = ((2 * k - 1) * t[i] * p1 - (k - 1) * p0) / k;
        p0 = p1; p1 = pk;
      }
      double pN = (N == 0) ? 1.0 : p1, pNm1 = (N == 1) ? 1.0 : p0;
      if (N == 1) { pN = t[i]; pNm1 = 1.0; }
      double tn = t[i] - (t[i] * pN - pNm1) / ((N + 1) * pN);
      maxd = std::max(maxd, std::abs(tn - t[i]));
      t[i] = tn;
    }
    if (maxd < 1e-16) break;
  }
  for (int i = 0; i <= N; ++i) x[N - i] = 0.5 * (t[i] + 1.0);
  x[0] = 0.0; x[N] = 1.0;
  // exact symmetry
  for (int i = 0; i <= N; ++i) {
    double s = 0.5 * (x[i] + 1.0 - x[N - i]);
    x[i] = s;
  }
  if (N % 2 == 0) x[N / 2] = 0.5;
}

// Lagrange basis on nodes z (m of them): value and derivative at point x.
inline void lagrange_eval(int m, const double *z, double x, double *val, double *der) {
  for (int a = 0; a < m; ++a) {
    double v = 1.0;
    for (int b = 0; b < m; ++b)
      if (b != a) v *= (x - z[b]) / (z[a] - z[b]);
    val[a] = v;
    double d = 0.0;
    for (int c = 0; c < m; ++c) {
      if (c == a) continue;
      double t = 1.0 / (z[a] - z[c]);
      for (int b = 0; b < m; ++b)
        if (b != a && b != c) t *= (x - z[b]) / (z[a] - z[b]);
      d += t;
    }
    der[a] = d;
  }
}

// ------------------------------------------------------------------ mesh ---
struct Mesh {
  int p = 1, N = 2, nc = 1;
  int n[3] = {1, 1, 1};
  int NN[3] = {2, 2, 2};  // nodes per direction
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1}, h[3] = {1, 1, 1};
  int n_cells = 0;
  int64_t n_nodes = 0, n_dofs = 0;
  vector<double> gll;            // p+1 support points on [0,1]
  vector<int32_t> cell_nodes;    // n_cells * N^3
  vector<double> vert;           // n_cells * 8 * 3
  vector<int32_t> constrained;   // sorted dofs
  vector<char> is_con;           // per dof
  vector<double> xnode;          // n_nodes * 3

  int64_t node_index(int I, int J, int K) const {
    return (int64_t)I + (int64_t)NN[0] * ((int64_t)J + (int64_t)NN[1] * K);
  }
  // position of node index along direction d
  double node_coord(int d, int I) const {
    int c = std::min(I / p, n[d] - 1);
    int a = I - c * p;
    return lo[d] + (c + gll[a]) * h[d];
  }
};

// face_comp[f]: components constrained on boundary face f (0:x=lo, 1:x=hi,
// 2:y=lo, 3:y=hi, 4:z=lo, 5:z=hi); a node on several faces takes the union
inline void build_mesh(Mesh &m, const double *lo, const double *hi, const int *n, int p, int nc,
                       const int *face_comp) {
  if (p < 1 || p > 8) throw std::runtime_error("degree out of range");
  m.p = p; m.N = p + 1; m.nc = nc;
  for (int d = 0; d < 3; ++d) {
    m.n[d] = n[d]; m.lo[d] = lo[d]; m.hi[d] = hi[d];
    m.h[d] = (hi[d] - lo[d]) / n[d];
    m.NN[d] = n[d] * p + 1;
  }
  m.gll.resize(p + 1);
  gll_points(p, m.gll.data());
  m.n_cells = n[0] * n[1] * n[2];
  m.n_nodes = (int64_t)m.NN[0] * m.NN[1] * m.NN[2];
  m.n_dofs = m.n_nodes * nc;
  const int N = m.N, N3 = N * N * N;
  m.cell_nodes.resize((size_t)m.n_cells * N3);
  m.vert.resize((size_t)m.n_cells * 24);
  for (int cz = 0; cz < n[2]; ++cz)
    for (int cy = 0; cy < n[1]; ++cy)
      for (int cx = 0; cx < n[0]; ++cx) {
        int cell = cx + n[0] * (cy + n[1] * cz);
        for (int c = 0; c < N; ++c)
          for (int b = 0; b < N; ++b)
            for (int a = 0; a < N; ++a)
              m.cell_nodes[(size_t)cell * N3 + a + N * (b + N * c)] =
                  (int32_t)m.node_index(cx * p + a, cy * p + b, cz * p + c);
        for (int v = 0; v < 8; ++v) {
          int vx = v & 1, vy = (v >> 1) & 1, vz = (v >> 2) & 1;
          m.vert[(size_t)cell * 24 + v * 3 + 0] = lo[0] + (cx + vx) * m.h[0];
          m.vert[(size_t)cell * 24 + v * 3 + 1] = lo[1] + (cy + vy) * m.h[1];
          m.vert[(size_t)cell * 24 + v * 3 + 2] = lo[2] + (cz + vz) * m.h[2];
        }
      }
  m.xnode.resize((size_t)m.n_nodes * 3);
  for (int K = 0; K < m.NN[2]; ++K)
    for (int J = 0; J < m.NN[1]; ++J)
      for (int I = 0; I < m.NN[0]; ++I) {
        int64_t id = m.node_index(I, J, K);
        m.xnode[id * 3 + 0] = m.node_coord(0, I);
        m.xnode[id * 3 + 1] = m.node_coord(1, J);
        m.xnode[id * 3 + 2] = m.node_coord(2, K);
      }
  m.is_con.assign((size_t)m.n_dofs, 0);
  for (int K = 0; K < m.NN[2]; ++K)
    for (int J = 0; J < m.NN[1]; ++J)
      for (int I = 0; I < m.NN[0]; ++I) {
        int comps = 0;
        if (I == 0) comps |= face_comp[0];
        if (I == m.NN[0] - 1) comps |= face_comp[1];
        if (J == 0) comps |= face_comp[2];
        if (J == m.NN[1] - 1) comps |= face_comp[3];
        if (K == 0) comps |= face_comp[4];
        if (K == m.NN[2] - 1) comps |= face_comp[5];
        if (!comps) continue;
        int64_t id = m.node_index(I, J, K);
        for (int c = 0; c < nc; ++c)
          if (comps & (1 << c)) m.is_con[id * nc + c] = 1;
      }
  m.constrained.clear();
  for (int64_t i = 0; i < m.n_dofs; ++i)
    if (m.is_con[i]) m.constrained.push_back((int32_t)i);
}

// ----------------------------------------------------- small 3x3 helpers ---
struct M3 {
  double a[3][3];
};
inline double det3(const double A[3][3]) {
  return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
         A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}
inline void inv3(const double A[3][3], double R[3][3]) {
  double d = det3(A);
  double id = 1.0 / d;
  R[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * id;
  R[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  R[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  R[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * id;
  R[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  R[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  R[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * id;
  R[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  R[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
}

// Trilinear (MappingQ1) Jacobian J[e][f] = dX_e/dxi_f of a hex with
// lexicographic vertices v = vx + 2 vy + 4 vz, at reference point xi.
inline void q1_jacobian(const double *vert24, const double xi[3], double J[3][3]) {
  for (int e = 0; e < 3; ++e)
    for (int f = 0; f < 3; ++f) J[e][f] = 0;
  for (int v = 0; v < 8; ++v) {
    int b[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
    double N1[3], dN1[3];
    for (int d = 0; d < 3; ++d) {
      N1[d] = b[d] ? xi[d] : 1.0 - xi[d];
      dN1[d] = b[d] ? 1.0 : -1.0;
    }
    double g[3] = {dN1[0] * N1[1] * N1[2], N1[0] * dN1[1] * N1[2], N1[0] * N1[1] * dN1[2]};
    for (int e = 0; e < 3; ++e)
      for (int f = 0; f < 3; ++f) J[e][f] += vert24[v * 3 + e] * g[f];
  }
}
inline void q1_position(const double *vert24, const double xi[3], double X[3]) {
  X[0] = X[1] = X[2] = 0;
  for (int v = 0; v < 8; ++v) {
    int b[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
    double w = 1;
    for (int d = 0; d < 3; ++d) w *= b[d] ? xi[d] : 1.0 - xi[d];
    for (int e = 0; e < 3; ++e) X[e] += w * vert24[v * 3 + e];
  }
}

}  // namespace orc
