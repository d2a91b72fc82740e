// Host-side set-up of libmfgpu: last-error storage, 1D finite-element tables
// and the structured box mesh builder that produces the level layout
// (cell->node maps, Dirichlet lists, cell vertices) the operators consume.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"

namespace mfg {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

// Legendre P_n, P_n' and P_n'' at t in (-1,1).
static void legendre3(int n, double t, double &p, double &dp, double &ddp) {
  double pm = 1.0, pc = t;
  if (n == 0) { p = 1; dp = 0; ddp = 0; return; }
  for (int k = 2; k <= n; ++k) {
    double pn = ((2.0 * k - 1.0) * t * pc - (k - 1.0) * pm) / k;
    pm = pc;
    pc = pn;
  }
  p = pc;
  const double s = 1.0 - t * t;
  dp = n * (pm - t * pc) / s;
  // Legendre equation: (1-t^2) P'' - 2 t P' + n(n+1) P = 0
  ddp = (2.0 * t * dp - n * (n + 1.0) * p) / s;
}

std::vector<double> gauss_lobatto_points(int p) {
  std::vector<double> x(p + 1);
  x[0] = 0.0;
  x[p] = 1.0;
  // interior points: roots of P_p' by Newton, started from Chebyshev-Lobatto points
  for (int i = 1; i < p; ++i) {
    double t = -std::cos(M_PI * i / p);
    for (int it = 0; it < 100; ++it) {
      double P, dP, ddP;
      legendre3(p, t, P, dP, ddP);
      const double step = dP / ddP;
      t -= step;
      if (std::abs(step) < 1e-17) break;
    }
    x[i] = 0.5 * (1.0 + t);
  }
  for (int i = 0; i <= p / 2; ++i) {  // enforce the mirror symmetry exactly
    const double a = 0.5 * (x[i] + (1.0 - x[p - i]));
    x[i] = a;
    x[p - i] = 1.0 - a;
  }
  if (p % 2 == 0) x[p / 2] = 0.5;
  return x;
}

void gauss_points(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double t = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double P = 0, dP = 0, ddP = 0;
    for (int it = 0; it < 100; ++it) {
      legendre3(n, t, P, dP, ddP);
      const double step = P / dP;
      t -= step;
      if (std::abs(step) < 1e-17) break;
    }
    legendre3(n, t, P, dP, ddP);
    x[i] = 0.5 * (1.0 + t);
    w[i] = 1.0 / ((1.0 - t * t) * dP * dP);
  }
}

void lagrange_basis(const std::vector<double> &z, double x, double *val, double *der) {
  const int m = (int)z.size();
  // barycentric weights
  std::vector<double> bw(m, 1.0);
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b)
      if (a != b) bw[a] /= (z[a] - z[b]);
  for (int a = 0; a < m; ++a) {
    if (std::abs(x - z[a]) < 1e-15) {
      // at a node: value is the Kronecker delta; derivative from the
      // differentiation-matrix formula D_ab = (w_b/w_a)/(z_a - z_b)
      for (int b = 0; b < m; ++b) val[b] = (a == b) ? 1.0 : 0.0;
      double diag = 0;
      for (int b = 0; b < m; ++b)
        if (b != a) {
          der[b] = (bw[b] / bw[a]) / (z[a] - z[b]);
          diag -= der[b];
        }
      der[a] = diag;
      return;
    }
  }
  // l(x) = prod (x - z_b); l_a(x) = l(x) w_a / (x - z_a)
  double l = 1.0;
  for (int b = 0; b < m; ++b) l *= (x - z[b]);
  double s = 0.0;
  for (int a = 0; a < m; ++a) s += 1.0 / (x - z[a]);
  for (int a = 0; a < m; ++a) {
    val[a] = l * bw[a] / (x - z[a]);
    // l_a'(x) = l_a(x) * (sum_{b != a} 1/(x - z_b))
    der[a] = val[a] * (s - 1.0 / (x - z[a]));
  }
}

double uniform_pm1_host(uint64_t seed, uint64_t index);

Tables1D make_tables(int p) {
  Tables1D t;
  t.N = p + 1;
  const int N = t.N;
  std::vector<double> z = gauss_lobatto_points(p);
  std::vector<double> q, w;
  gauss_points(N, q, w);
  t.points = q;
  t.weight = w;
  t.shape.assign(N * N, 0.0);
  t.colloc.assign(N * N, 0.0);
  std::vector<double> val(N), der(N);
  for (int i = 0; i < N; ++i) {
    lagrange_basis(z, q[i], val.data(), der.data());
    for (int a = 0; a < N; ++a) t.shape[i * N + a] = val[a];
    lagrange_basis(q, q[i], val.data(), der.data());
    for (int r = 0; r < N; ++r) t.colloc[i * N + r] = der[r];
  }
  return t;
}

}  // namespace mfg

// ------------------------------------------------------------ box meshes ---
struct mfg_mesh_t {
  int p = 1, nc = 1;
  int n[3] = {1, 1, 1}, NN[3] = {2, 2, 2};
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  int64_t n_nodes = 0;
  std::vector<int32_t> cell_nodes;
  std::vector<double> vertices;
  std::vector<int32_t> constrained;
  std::vector<double> node_x;
  std::vector<int32_t> lex_id;  // node -> its lexicographic number (identity until renumbered)
  std::vector<int32_t> material_ids;  // per cell; empty: all 0
  int numbering = MFG_NUMBERING_LEXICOGRAPHIC;
};

using namespace mfg;

extern "C" {

const char *mfg_last_error(void) { return g_last_error.c_str(); }
const char *mfg_version(void) { return "mfgpu 0.1 (gfx950)"; }

int mfg_box_mesh_create(const double lo[3], const double hi[3], const int32_t n[3], int32_t degree,
                        int32_t n_components, int32_t face_mask, int32_t comp_mask, mfg_mesh *out) {
  int32_t face_comp[6];
  for (int f = 0; f < 6; ++f) face_comp[f] = ((face_mask >> f) & 1) ? comp_mask : 0;
  return mfg_box_mesh_create_bc(lo, hi, n, degree, n_components, face_comp, out);
}

int mfg_box_mesh_create_bc(const double lo[3], const double hi[3], const int32_t n[3], int32_t degree,
                           int32_t n_components, const int32_t face_comp_mask[6], mfg_mesh *out) {
  return guarded([&] {
    MFG_REQUIRE(out && lo && hi && n && face_comp_mask, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(degree >= 1 && degree <= kMaxDegree, MFG_ERR_ARGUMENT, "degree must be in 1..6");
    MFG_REQUIRE(n_components == 1 || n_components == 3, MFG_ERR_ARGUMENT, "n_components must be 1 or 3");
    auto m = new mfg_mesh_t();
    m->p = degree;
    m->nc = n_components;
    const int p = degree, N = p + 1, N3 = N * N * N;
    for (int d = 0; d < 3; ++d) {
      if (n[d] < 1 || !(hi[d] > lo[d])) { delete m; throw Error(MFG_ERR_ARGUMENT, "bad box extents"); }
      m->n[d] = n[d];
      m->lo[d] = lo[d];
      m->hi[d] = hi[d];
      m->NN[d] = n[d] * p + 1;
    }
    m->n_nodes = (int64_t)m->NN[0] * m->NN[1] * m->NN[2];
    if (m->n_nodes >= ((int64_t)1 << kNodeBits)) { delete m; throw Error(MFG_ERR_ARGUMENT, "mesh too large"); }
    const int64_t n_cells = (int64_t)n[0] * n[1] * n[2];
    m->cell_nodes.resize(n_cells * N3);
    m->vertices.resize(n_cells * 24);
    const std::vector<double> z = gauss_lobatto_points(p);
    double h[3];
    for (int d = 0; d < 3; ++d) h[d] = (hi[d] - lo[d]) / n[d];
    for (int64_t c = 0; c < n_cells; ++c) {
      const int cx = (int)(c % n[0]), cy = (int)((c / n[0]) % n[1]), cz = (int)(c / ((int64_t)n[0] * n[1]));
      int32_t *cn = &m->cell_nodes[c * N3];
      for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j) {
          const int64_t row = (int64_t)(cx * p) + (int64_t)m->NN[0] * ((cy * p + j) + (int64_t)m->NN[1] * (cz * p + k));
          for (int i = 0; i < N; ++i) cn[i + N * (j + N * k)] = (int32_t)(row + i);
        }
      for (int v = 0; v < 8; ++v) {
        const int o[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
        const int cc[3] = {cx, cy, cz};
        for (int d = 0; d < 3; ++d) m->vertices[c * 24 + v * 3 + d] = lo[d] + (cc[d] + o[d]) * h[d];
      }
    }
    m->lex_id.resize(m->n_nodes);
    for (int64_t v = 0; v < m->n_nodes; ++v) m->lex_id[v] = (int32_t)v;
    // node coordinates from the cell-local GLL support points
    m->node_x.resize(m->n_nodes * 3);
    std::vector<double> line[3];
    for (int d = 0; d < 3; ++d) {
      line[d].resize(m->NN[d]);
      for (int I = 0; I < m->NN[d]; ++I) {
        const int c = std::min(I / p, n[d] - 1);
        line[d][I] = lo[d] + (c + z[I - c * p]) * h[d];
      }
    }
    for (int64_t id = 0; id < m->n_nodes; ++id) {
      const int I = (int)(id % m->NN[0]), J = (int)((id / m->NN[0]) % m->NN[1]),
                K = (int)(id / ((int64_t)m->NN[0] * m->NN[1]));
      m->node_x[id * 3 + 0] = line[0][I];
      m->node_x[id * 3 + 1] = line[1][J];
      m->node_x[id * 3 + 2] = line[2][K];
    }
    // homogeneous Dirichlet dofs: a node's component is constrained when any
    // boundary face the node lies on constrains it (the union of the
    // per-boundary-id masks, as make_constraints merges them); sorted by
    // construction
    for (int64_t id = 0; id < m->n_nodes; ++id) {
      const int I = (int)(id % m->NN[0]), J = (int)((id / m->NN[0]) % m->NN[1]),
                K = (int)(id / ((int64_t)m->NN[0] * m->NN[1]));
      const int32_t comps = (I == 0 ? face_comp_mask[0] : 0) | (I == m->NN[0] - 1 ? face_comp_mask[1] : 0) |
                            (J == 0 ? face_comp_mask[2] : 0) | (J == m->NN[1] - 1 ? face_comp_mask[3] : 0) |
                            (K == 0 ? face_comp_mask[4] : 0) | (K == m->NN[2] - 1 ? face_comp_mask[5] : 0);
      for (int c = 0; c < n_components; ++c)
        if (comps & (1 << c)) m->constrained.push_back((int32_t)(id * n_components + c));
    }
    *out = m;
  });
}

int mfg_box_mesh_destroy(mfg_mesh mesh) {
  delete mesh;
  return MFG_OK;
}

int mfg_box_mesh_set_material_ids(mfg_mesh m, const int32_t *ids) {
  return guarded([&] {
    MFG_REQUIRE(m, MFG_ERR_ARGUMENT, "null mesh");
    const size_t n_cells = m->vertices.size() / 24;
    if (ids) m->material_ids.assign(ids, ids + n_cells);
    else m->material_ids.clear();
  });
}

int mfg_box_mesh_describe(mfg_mesh m, mfg_level_desc *d) {
  return guarded([&] {
    MFG_REQUIRE(m && d, MFG_ERR_ARGUMENT, "null argument");
    d->material_ids = m->material_ids.empty() ? nullptr : m->material_ids.data();
    d->degree = m->p;
    d->n_q_1d = m->p + 1;
    d->n_components = m->nc;
    d->n_cells = (int32_t)(m->vertices.size() / 24);
    d->n_nodes = m->n_nodes;
    d->cell_nodes = m->cell_nodes.data();
    d->cell_vertices = m->vertices.data();
    d->constrained_dofs = m->constrained.data();
    d->n_constrained = (int64_t)m->constrained.size();
  });
}

int mfg_box_mesh_fill_u0(mfg_mesh m, double amp, double noise, uint64_t seed, double *out) {
  return guarded([&] {
    MFG_REQUIRE(m && out, MFG_ERR_ARGUMENT, "null argument");
    std::vector<char> con((size_t)m->n_nodes * m->nc, 0);
    for (int32_t c : m->constrained) con[c] = 1;
    for (int64_t node = 0; node < m->n_nodes; ++node) {
      const double *x = &m->node_x[node * 3];
      const double base = amp * std::sin(M_PI * x[0]) * std::sin(M_PI * x[1]) * std::sin(M_PI * x[2]);
      for (int c = 0; c < m->nc; ++c) {
        // the noise is drawn per lexicographic dof: the same field under any numbering
        const int64_t dof = node * m->nc + c, lex_dof = (int64_t)m->lex_id[node] * m->nc + c;
        out[dof] = con[dof] ? 0.0 : base + noise * uniform_pm1_host(seed, (uint64_t)lex_dof);
      }
    }
  });
}

int mfg_box_mesh_boundary_faces(mfg_mesh m, int32_t boundary_id, int32_t *faces, int64_t *n_faces) {
  return guarded([&] {
    MFG_REQUIRE(m && n_faces, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(boundary_id >= 0 && boundary_id < 6, MFG_ERR_ARGUMENT, "box boundary ids are 0..5");
    const int axis = boundary_id / 2, side = boundary_id % 2;
    const int a = axis == 0 ? 1 : 0, b = axis == 2 ? 1 : 2;
    const int64_t count = (int64_t)m->n[a] * m->n[b];
    if (faces) {
      int64_t f = 0;
      for (int jb = 0; jb < m->n[b]; ++jb)
        for (int ja = 0; ja < m->n[a]; ++ja) {
          int cc[3];
          cc[axis] = side ? m->n[axis] - 1 : 0;
          cc[a] = ja;
          cc[b] = jb;
          faces[2 * f] = (int32_t)(cc[0] + (int64_t)m->n[0] * (cc[1] + (int64_t)m->n[1] * cc[2]));
          faces[2 * f + 1] = boundary_id;  // local face 2 axis + side: the same number on a box
          ++f;
        }
    }
    *n_faces = count;
  });
}

int mfg_box_mesh_cells(mfg_mesh m, int32_t n[3]) {
  return guarded([&] {
    MFG_REQUIRE(m && n, MFG_ERR_ARGUMENT, "null argument");
    for (int d = 0; d < 3; ++d) n[d] = m->n[d];
  });
}

// Renumbers the nodes of a box mesh in place: cell_nodes, node coordinates and
// the Dirichlet list follow the permutation; cells keep their order.
//  - LEXICOGRAPHIC: node (I, J, K) of the global lattice is I + NN0 (J + NN1 K).
//  - CELLWISE: nodes are numbered cell by cell, as DoFHandler::distribute_dofs
//    does (mf_elasticity.h:667, heat_transfer.h:437), but every cell (i, j, k)
//    owns exactly the p^3 nodes (i p + a, j p + b, k p + c), a, b, c in 1..p,
//    so block c holds nodes [c p^3, (c + 1) p^3) and, for p = 4, starts on a
//    128-byte line of a 3-component FP64 vector.  The nodes of the three low
//    boundary planes follow all blocks.  Inside a block the nodes go entity by
//    entity -- x edge, y face, y edge, vertex, z edge, x face, z face, interior
//    (an entity named by the directions in which it is not at the cell's high
//    side) -- and lexicographically inside an entity: the pieces a neighbouring
//    cell reads (a face with its edges and vertex) then fall into few lines.
int mfg_box_mesh_renumber(mfg_mesh m, int32_t numbering) {
  return guarded([&] {
    MFG_REQUIRE(m, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(numbering == MFG_NUMBERING_LEXICOGRAPHIC || numbering == MFG_NUMBERING_CELLWISE, MFG_ERR_ARGUMENT,
                "unknown numbering");
    const int p = m->p, N = p + 1, N3 = N * N * N;
    const int64_t n_cells = (int64_t)m->n[0] * m->n[1] * m->n[2];
    std::vector<int32_t> newid((size_t)m->n_nodes, -1);
    if (numbering == MFG_NUMBERING_LEXICOGRAPHIC) {
      for (int64_t c = 0; c < n_cells; ++c) {
        const int cx = (int)(c % m->n[0]), cy = (int)((c / m->n[0]) % m->n[1]), cz = (int)(c / ((int64_t)m->n[0] * m->n[1]));
        for (int l = 0; l < N3; ++l) {
          const int i = l % N, j = (l / N) % N, k = l / (N * N);
          newid[m->cell_nodes[c * N3 + l]] =
              (int32_t)((cx * p + i) + (int64_t)m->NN[0] * ((cy * p + j) + (int64_t)m->NN[1] * (cz * p + k)));
        }
      }
    } else {
      // entity of a block node: bit d set where coordinate d is below the
      // cell's high side; rank of each entity in the numbering order
      static const int kRank[8] = {/*vertex*/ 3, /*x edge*/ 0, /*y edge*/ 2, /*z face*/ 6,
                                   /*z edge*/ 4, /*y face*/ 1, /*x face*/ 5, /*interior*/ 7};
      std::vector<int> block;
      for (int l = 0; l < N3; ++l)
        if (l % N > 0 && (l / N) % N > 0 && l / (N * N) > 0) block.push_back(l);
      auto entity = [&](int l) { return (l % N < p ? 1 : 0) | ((l / N) % N < p ? 2 : 0) | (l / (N * N) < p ? 4 : 0); };
      std::stable_sort(block.begin(), block.end(), [&](int a, int b) { return kRank[entity(a)] < kRank[entity(b)]; });
      int32_t next = 0;
      for (int64_t c = 0; c < n_cells; ++c)
        for (int l : block) newid[m->cell_nodes[c * N3 + l]] = next++;
      for (int64_t c = 0; c < n_cells; ++c)  // the low boundary planes
        for (int l = 0; l < N3; ++l) {
          const int32_t old = m->cell_nodes[c * N3 + l];
          if (newid[old] < 0) newid[old] = next++;
        }
      MFG_REQUIRE(next == m->n_nodes, MFG_ERR_ARGUMENT, "mesh has nodes in no cell");
    }
    for (auto &x : m->cell_nodes) x = newid[x];
    std::vector<double> nx(m->node_x.size());
    for (int64_t v = 0; v < m->n_nodes; ++v)
      for (int d = 0; d < 3; ++d) nx[(size_t)newid[v] * 3 + d] = m->node_x[(size_t)v * 3 + d];
    m->node_x.swap(nx);
    std::vector<int32_t> lex(m->lex_id.size());
    for (int64_t v = 0; v < m->n_nodes; ++v) lex[newid[v]] = m->lex_id[v];
    m->lex_id.swap(lex);
    for (auto &dof : m->constrained) dof = (int32_t)(newid[dof / m->nc] * m->nc + dof % m->nc);
    std::sort(m->constrained.begin(), m->constrained.end());
    m->numbering = numbering;
  });
}

int32_t mfg_box_mesh_numbering(mfg_mesh m) { return m ? m->numbering : -1; }

int mfg_box_mesh_node_coords(mfg_mesh m, double *out) {
  return guarded([&] {
    MFG_REQUIRE(m && out, MFG_ERR_ARGUMENT, "null argument");
    std::memcpy(out, m->node_x.data(), m->node_x.size() * sizeof(double));
  });
}

}  // extern "C"
