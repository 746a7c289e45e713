// ORACLE - test infrastructure only.
//
// CPU restatement of the reference's Galerkin GEP assembly
// (galerkin_sample_gep_hcurl, reference src/fem_problem/galerkin.rs:33-187),
// used to check the HIP path and to time the reference algorithm on the host.
// Only tests, the smoke check and the benchmark's CPU-baseline leg may load
// it; the product library never does.
//
// It follows the reference literally:
//  * Gauss-Legendre nodes by Golub-Welsch: eigenvalues of the Jacobi matrix,
//    weights 2*v0^2, sorted ascending (glq.rs:179-222);
//  * basis functions sampled per (elem, descendant) with the owning elem's
//    Jacobian, points scaled into the descendant's relative range
//    (basis.rs:365-423), HierPoly tables (hierarchical_basis_fns.rs:102-162)
//    and the Legendre-based space (hierarchical_basis_fns.rs:296-646, with the
//    order-10 Q row summing to zero);
//  * per-point integrands of CurlCurl and L2Inner in the reference's operation
//    order with the 2D tensor quadrature sum (integrals.rs:26-92, 292-354,
//    glq.rs:19-32);
//  * per-elem ordered maps keyed [min(dof), max(dof)] with "+=" on collision
//    (sparse_matrix.rs:68-98), merged elem by elem (sparse_matrix.rs:106-120).
//    The reference merges in channel-arrival order (linalg.rs:64-80); the
//    oracle merges in ascending elem id, which is one of those orders.
//
// Parity status of each part is recorded in the tests that pin it (GLQ tables,
// closed-form integrals, golden eigenpair residual, analytic cutoffs).

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <thread>
#include <utility>
#include <vector>

#include "../include/fem2d.h"

namespace oracle {

// ---------------------------------------------------------------------------
// Golub-Welsch (glq.rs:179-222)
// ---------------------------------------------------------------------------

// Symmetric tridiagonal QL with implicit Wilkinson shifts; z tracks the first
// row of the eigenvector matrix only (enough for the weights).
static void tridiag_ql(int n, std::vector<double>& d, std::vector<double>& e, std::vector<double>& z) {
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  if (n > 0) e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-300 + 2.220446049250313e-16 * dd) break;
      }
      if (m != l) {
        if (++iter > 200) break;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i];
          double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          // rotate the tracked eigenvector row
          f = z[i + 1];
          z[i + 1] = s * z[i] + c * f;
          z[i] = c * z[i] - s * f;
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
}

void gauss_quadrature_points(int n, bool include_endpoints, std::vector<double>& pts, std::vector<double>& wts) {
  std::vector<double> d(n, 0.0), e(n, 0.0), z(n, 0.0);
  for (int i = 1; i < n; ++i) {
    double t = 2.0 * i;
    e[i] = 0.5 / std::sqrt(1.0 - 1.0 / (t * t));  // beta_i = 0.5 / sqrt(1 - (2i)^-2)
  }
  if (n > 0) z[0] = 1.0;
  tridiag_ql(n, d, e, z);
  std::vector<std::pair<double, double>> xw(n);
  for (int i = 0; i < n; ++i) xw[i] = {d[i], z[i] * z[i] * 2.0};
  std::sort(xw.begin(), xw.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  pts.clear();
  wts.clear();
  if (include_endpoints) {
    pts.push_back(-1.0);
    wts.push_back(1.0);
  }
  for (auto& p : xw) {
    pts.push_back(p.first);
    wts.push_back(p.second);
  }
  if (include_endpoints) {
    pts.push_back(1.0);
    wts.push_back(1.0);
  }
}

// default_ngq (basis.rs:172-177)
int default_ngq(int max_order) {
  float conv = float(max_order * 4);
  float p2 = std::ceil(std::log2(conv));
  if (!std::isfinite(p2)) return 0;
  return int(std::lround(std::pow(2.0f, p2)));
}

// ---------------------------------------------------------------------------
// Shape-function spaces
// ---------------------------------------------------------------------------

// Tables [order][point]: norm (N), norm_d1 (N'), tang (T), tang_d1 (T').
struct Space {
  std::vector<std::vector<double>> n, nd, t, td;
};

// HierPoly::new_without_d2 (hierarchical_basis_fns.rs:102-162)
static Space hier_poly(int n_max, const std::vector<double>& x) {
  Space s;
  const size_t np = x.size();
  for (int n = 0; n <= n_max; ++n) {
    const double nf = double(n);
    if (n == 0) {
      std::vector<double> t(np), ones(np, 1.0);
      for (size_t p = 0; p < np; ++p) t[p] = 1.0 - x[p];
      s.t.push_back(t);
      s.td.push_back(std::vector<double>(np, -1.0));
      s.n.push_back(ones);
      s.nd.push_back(std::vector<double>(np, 0.0));
    } else if (n == 1) {
      std::vector<double> t(np);
      for (size_t p = 0; p < np; ++p) t[p] = 1.0 + x[p];
      s.t.push_back(t);
      s.td.push_back(std::vector<double>(np, 1.0));
      s.n.push_back(x);
      s.nd.push_back(std::vector<double>(np, 1.0));
    } else {
      std::vector<double> pw(np), pd(np), t(np), td(np);
      for (size_t p = 0; p < np; ++p) {
        pw[p] = s.n[n - 1][p] * x[p];
        pd[p] = nf * s.n[n - 1][p];
      }
      for (size_t p = 0; p < np; ++p) {
        if (n % 2 == 0) {
          t[p] = pw[p] - 1.0;
          td[p] = pd[p];
        } else {
          t[p] = pw[p] - x[p];
          td[p] = pd[p] - 1.0;
        }
      }
      s.n.push_back(pw);
      s.nd.push_back(pd);
      s.t.push_back(t);
      s.td.push_back(td);
    }
  }
  return s;
}

// Euclidean normalisation of T_2..T_8 (hierarchical_basis_fns.rs:206-209); 1 beyond.
static const double kEucNorm[7] = {0.968246, 2.561738, 0.838525, 4.248161, 0.816397, 5.882766, 0.808509};

// Legendre-based space (LegendrePoly + QFunction, hierarchical_basis_fns.rs:296-646):
// N_n = L_n, T_n = c_n * sum_k q_{n,k} L_k with q_{n,n} = 1 and
// q_{n,k} = -(2k+1) / (n(n-1)/2) for k < n, k = n (mod 2).
static Space kol(int n_max, const std::vector<double>& x) {
  const size_t np = x.size();
  std::vector<std::vector<double>> L(n_max + 1, std::vector<double>(np)), Ld(n_max + 1, std::vector<double>(np));
  for (int i = 0; i <= n_max; ++i) {
    const double fi = double(i);
    for (size_t p = 0; p < np; ++p) {
      if (i == 0) {
        L[i][p] = 1.0;
        Ld[i][p] = 0.0;
      } else if (i == 1) {
        L[i][p] = x[p];
        Ld[i][p] = 1.0;
      } else {
        L[i][p] = ((2.0 * fi - 1.0) * x[p] * L[i - 1][p] - (fi - 1.0) * L[i - 2][p]) / fi;
        Ld[i][p] = fi * L[i - 1][p] + x[p] * Ld[i - 1][p];
      }
    }
  }
  Space s;
  s.n = L;
  s.nd = Ld;
  for (int n = 0; n <= n_max; ++n) {
    std::vector<double> t(np), td(np);
    if (n == 0) {
      for (size_t p = 0; p < np; ++p) {
        t[p] = 1.0 - x[p];
        td[p] = -1.0;
      }
    } else if (n == 1) {
      for (size_t p = 0; p < np; ++p) {
        t[p] = 1.0 + x[p];
        td[p] = 1.0;
      }
    } else {
      const int den = n * (n - 1) / 2;
      const double c = n - 2 < 7 ? kEucNorm[n - 2] : 1.0;
      std::vector<double> sv(np, 0.0), sd(np, 0.0);
      for (int k = 0; k <= n; ++k) {
        int num = 0;
        if (k == n) num = den;
        else if ((n - k) % 2 == 0) num = -(2 * k + 1);
        const double w = double(num) / double(den);
        for (size_t p = 0; p < np; ++p) {
          sv[p] += w * L[k][p];
          sd[p] += w * Ld[k][p];
        }
      }
      for (size_t p = 0; p < np; ++p) {
        t[p] = sv[p] * c;
        td[p] = sd[p] * c;
      }
    }
    s.t.push_back(t);
    s.td.push_back(td);
  }
  return s;
}

static Space make_space(int basis, int n_max, const std::vector<double>& x) {
  return basis == FEM2D_BASIS_KOL ? kol(n_max, x) : hier_poly(n_max, x);
}

// ---------------------------------------------------------------------------
// Domain view helpers
// ---------------------------------------------------------------------------

using Range2 = std::array<std::array<double, 2>, 2>;

static Range2 sub_range(int loc, const Range2& r) {
  double mu = (r[0][0] + r[0][1]) / 2.0;
  double mv = (r[1][0] + r[1][1]) / 2.0;
  std::array<double, 2> u = r[0], v = r[1];
  switch (loc) {
    case FEM2D_LOC_SW: u = {r[0][0], mu}; v = {r[1][0], mv}; break;
    case FEM2D_LOC_SE: u = {mu, r[0][1]}; v = {r[1][0], mv}; break;
    case FEM2D_LOC_NW: u = {r[0][0], mu}; v = {mv, r[1][1]}; break;
    case FEM2D_LOC_NE: u = {mu, r[0][1]}; v = {mv, r[1][1]}; break;
    case FEM2D_LOC_W: u = {r[0][0], mu}; break;
    case FEM2D_LOC_E: u = {mu, r[0][1]}; break;
    case FEM2D_LOC_S: v = {r[1][0], mv}; break;
    case FEM2D_LOC_N: v = {mv, r[1][1]}; break;
  }
  return {u, v};
}

struct Tree {
  const fem2d_domain_view* v;
  std::vector<std::vector<int>> children;

  explicit Tree(const fem2d_domain_view* view) : v(view), children(view->n_elems) {
    for (int e = 0; e < v->n_elems; ++e)
      if (v->parent[e] >= 0) children[v->parent[e]].push_back(e);
  }
  // path root..e (inclusive)
  std::vector<int> path(int e) const {
    std::vector<int> p;
    for (int x = e; x >= 0; x = v->parent[x]) p.push_back(x);
    std::reverse(p.begin(), p.end());
    return p;
  }
  // Elem::parametric_range (elem.rs:191-197)
  Range2 parametric_range(int e) const {
    Range2 acc{{{-1.0, 1.0}, {-1.0, 1.0}}};
    auto p = path(e);
    for (size_t k = 1; k < p.size(); ++k) acc = sub_range(v->loc[p[k]], acc);
    return acc;
  }
  // Elem::relative_parametric_range (elem.rs:170-188)
  Range2 relative_range(int e, int ancestor) const {
    Range2 acc{{{-1.0, 1.0}, {-1.0, 1.0}}};
    auto p = path(e);
    size_t k = 0;
    while (k < p.size() && p[k] != ancestor) ++k;
    for (++k; k < p.size(); ++k) acc = sub_range(v->loc[p[k]], acc);
    return acc;
  }
  // Mesh::descendant_elems(e, false): pre-order DFS (mesh.rs:470-493)
  void descendants(int e, std::vector<int>& out) const {
    for (int c : children[e]) {
      out.push_back(c);
      descendants(c, out);
    }
  }
};

// ---------------------------------------------------------------------------
// HierCurlBasisFn (basis.rs:209-423) and the integrals (integrals.rs)
// ---------------------------------------------------------------------------

struct V2 {
  double x, y;
};
static inline V2 vmul(V2 a, double s) { return {a.x * s, a.y * s}; }
static inline V2 vmul(V2 a, V2 b) { return {a.x * b.x, a.y * b.y}; }
static inline double vdot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }

struct BasisFn {
  // constant over the sample points for the axis-aligned affine elements
  V2 jac_u, jac_v, inv_u, inv_v;
  double det;
  V2 para_scale;
  Space us, vs;

  V2 f_u(int i, int j, int m, int n) const { return vmul(vmul(inv_u, us.n[i][m]), vs.t[j][n]); }
  V2 f_v(int i, int j, int m, int n) const { return vmul(vmul(inv_v, us.t[i][m]), vs.n[j][n]); }
  V2 f_u_d1(int i, int j, int m, int n, V2 ps) const {
    return vmul(vmul(inv_u, V2{us.n[i][m] * vs.td[j][n], us.nd[i][m] * vs.t[j][n]}), ps);
  }
  V2 f_v_d1(int i, int j, int m, int n, V2 ps) const {
    return vmul(vmul(inv_v, V2{us.t[i][m] * vs.nd[j][n], us.td[i][m] * vs.n[j][n]}), ps);
  }
  double glq_scale() const { return para_scale.x * para_scale.y; }
  double uv_ratio() const { return jac_u.x / jac_v.y; }
  double vu_ratio() const { return jac_v.y / jac_u.x; }
};

struct Setup {
  const fem2d_domain_view* v;
  Tree tree;
  int basis;
  int i_max, j_max;
  std::vector<double> up, uw, vp, vw;
  explicit Setup(const fem2d_domain_view* view) : v(view), tree(view) {}
};

static double map_range(double val, double in_min, double in_max, double out_min, double out_max) {
  return (val - in_min) * (out_max - out_min) / (in_max - in_min) + out_min;
}

// HierCurlBasisFn::defined_over (basis.rs:365-423)
static BasisFn sample_basis_fn(const Setup& s, int elem, int desc) {
  std::vector<double> u = s.up, v = s.vp;
  double su = 1.0, sv = 1.0;
  if (desc >= 0 && desc != elem) {
    Range2 r = s.tree.relative_range(desc, elem);
    // scale_gauss_quad_points (glq.rs:238-249)
    su = (r[0][1] - r[0][0]) / 2.0;
    double cu = (r[0][1] + r[0][0]) / 2.0;
    sv = (r[1][1] - r[1][0]) / 2.0;
    double cv = (r[1][1] + r[1][0]) / 2.0;
    for (auto& x : u) x = x * su + cu;
    for (auto& y : v) y = y * sv + cv;
  }
  // Element::parametric_mapping over the owning elem's range (element.rs:33-50)
  Range2 pr = s.tree.parametric_range(elem);
  const double* box = s.v->element_box + 4 * elem;
  double x_min = map_range(pr[0][0], -1.0, 1.0, box[0], box[2]);
  double x_max = map_range(pr[0][1], -1.0, 1.0, box[0], box[2]);
  double y_min = map_range(pr[1][0], -1.0, 1.0, box[1], box[3]);
  double y_max = map_range(pr[1][1], -1.0, 1.0, box[1], box[3]);
  double a = (x_max - x_min) / 2.0, b = (y_max - y_min) / 2.0;
  BasisFn f;
  f.jac_u = {a, 0.0};
  f.jac_v = {0.0, b};
  // M2D::det / inverse (space.rs:138-147)
  f.det = f.jac_u.x * f.jac_v.y - f.jac_u.y * f.jac_v.x;
  f.inv_u = {f.jac_v.y / f.det, (-1.0 * f.jac_u.y) / f.det};
  f.inv_v = {(-1.0 * f.jac_v.x) / f.det, f.jac_u.x / f.det};
  f.para_scale = {su, sv};
  f.us = make_space(s.basis, s.i_max, u);
  f.vs = make_space(s.basis, s.j_max, v);
  return f;
}

// real_gauss_quad (glq.rs:19-32)
template <class F>
static double real_gauss_quad(const std::vector<double>& uw, const std::vector<double>& vw, F&& f) {
  double solution = 0.0;
  for (size_t m = 0; m < uw.size(); ++m) {
    double inner = 0.0;
    for (size_t n = 0; n < vw.size(); ++n) inner += f(int(m), int(n)) * vw[n];
    solution += inner * uw[m];
  }
  return solution;
}

static const V2 kCurlOp = {-1.0, 1.0};

static double max_uv_ratios(const BasisFn& p, const BasisFn& q) {
  return double(p.det >= q.det) * p.uv_ratio() + double(p.det < q.det) * q.uv_ratio();
}
static double max_vu_ratios(const BasisFn& p, const BasisFn& q) {
  return double(p.det >= q.det) * p.vu_ratio() + double(p.det < q.det) * q.vu_ratio();
}

// CurlCurl::integrate (integrals.rs:26-92)
static double curl_curl(const Setup& s, int pd, int qd, int pi, int pj, int qi, int qj, const BasisFn& p,
                        const BasisFn& q, double mu) {
  double integral = 0.0;
  if (pd == FEM2D_DIR_U && qd == FEM2D_DIR_U) {
    integral = real_gauss_quad(s.uw, s.vw, [&](int m, int n) {
      double pc = vdot(p.f_u_d1(pi, pj, m, n, q.para_scale), kCurlOp);
      double qc = vdot(q.f_u_d1(qi, qj, m, n, p.para_scale), kCurlOp);
      return pc * qc * max_uv_ratios(p, q);
    });
  } else if (pd == FEM2D_DIR_U && qd == FEM2D_DIR_V) {
    integral = real_gauss_quad(s.uw, s.vw, [&](int m, int n) {
      double pc = vdot(p.f_u_d1(pi, pj, m, n, q.para_scale), kCurlOp);
      double qc = vdot(q.f_v_d1(qi, qj, m, n, p.para_scale), kCurlOp);
      return pc * qc;
    });
  } else if (pd == FEM2D_DIR_V && qd == FEM2D_DIR_U) {
    integral = real_gauss_quad(s.uw, s.vw, [&](int m, int n) {
      double pc = vdot(p.f_v_d1(pi, pj, m, n, q.para_scale), kCurlOp);
      double qc = vdot(q.f_u_d1(qi, qj, m, n, p.para_scale), kCurlOp);
      return pc * qc;
    });
  } else if (pd == FEM2D_DIR_V && qd == FEM2D_DIR_V) {
    integral = real_gauss_quad(s.uw, s.vw, [&](int m, int n) {
      double pc = vdot(p.f_v_d1(pi, pj, m, n, q.para_scale), kCurlOp);
      double qc = vdot(q.f_v_d1(qi, qj, m, n, p.para_scale), kCurlOp);
      return pc * qc * max_vu_ratios(p, q);
    });
  }
  return (1.0 / mu) * integral;
}

// L2Inner::integrate (integrals.rs:292-354)
static double l2_inner(const Setup& s, int pd, int qd, int pi, int pj, int qi, int qj, const BasisFn& p,
                       const BasisFn& q, double eps) {
  const double dmax = std::max(p.det, q.det);
  double integral = 0.0;
  if (pd == FEM2D_DIR_U && qd == FEM2D_DIR_U) {
    integral = real_gauss_quad(s.uw, s.vw, [&](int m, int n) { return vdot(p.f_u(pi, pj, m, n), q.f_u(qi, qj, m, n)) * dmax; });
  } else if (pd == FEM2D_DIR_U && qd == FEM2D_DIR_V) {
    integral = real_gauss_quad(s.uw, s.vw, [&](int m, int n) { return vdot(p.f_u(pi, pj, m, n), q.f_v(qi, qj, m, n)) * dmax; });
  } else if (pd == FEM2D_DIR_V && qd == FEM2D_DIR_U) {
    integral = real_gauss_quad(s.uw, s.vw, [&](int m, int n) { return vdot(p.f_v(pi, pj, m, n), q.f_u(qi, qj, m, n)) * dmax; });
  } else if (pd == FEM2D_DIR_V && qd == FEM2D_DIR_V) {
    integral = real_gauss_quad(s.uw, s.vw, [&](int m, int n) { return vdot(p.f_v(pi, pj, m, n), q.f_v(qi, qj, m, n)) * dmax; });
  }
  return eps * p.glq_scale() * q.glq_scale() * integral;
}

static double integrate(const Setup& s, int kind, int pd, int qd, int pi, int pj, int qi, int qj, const BasisFn& p,
                        const BasisFn& q, double eps, double mu) {
  return kind == FEM2D_INTEGRAL_CURLCURL ? curl_curl(s, pd, qd, pi, pj, qi, qj, p, q, mu)
                                         : l2_inner(s, pd, qd, pi, pj, qi, qj, p, q, eps);
}

using Key = std::pair<uint32_t, uint32_t>;
using SparseMap = std::map<Key, double>;

// SparseMatrix::insert_group (sparse_matrix.rs:68-98)
static void insert_group(SparseMap& m, const std::vector<std::pair<Key, double>>& entries) {
  for (const auto& [rc, val] : entries) {
    Key k = rc.first <= rc.second ? rc : Key{rc.second, rc.first};
    auto it = m.find(k);
    if (it != m.end())
      it->second += val;
    else
      m.emplace(k, val);
  }
}

// SparseMatrix::consume_matrix (sparse_matrix.rs:106-120)
static void consume(SparseMap& dst, const SparseMap& src) {
  for (const auto& [k, val] : src) {
    auto it = dst.find(k);
    if (it != dst.end())
      it->second += val;
    else
      dst.emplace(k, val);
  }
}

struct ElemResult {
  SparseMap a, b;
  int64_t pairs = 0;
};

// One task of the parallel loop over elems (galerkin.rs:73-184)
// row_mask (nullable): only pairs whose entry row min(dof_p, dof_q) is set are
// integrated, so masked rows come out complete and exact at any domain size.
static void assemble_elem(const Setup& s, int e, int ai, int bi, const uint8_t* row_mask, ElemResult& out) {
  const fem2d_domain_view* v = s.v;
  const double eps = v->materials[2 * e], mu = v->materials[2 * e + 1];
  BasisFn local = sample_basis_fn(s, e, -1);
  const int64_t b0 = v->bs_offsets[e], b1 = v->bs_offsets[e + 1];
  std::vector<std::pair<Key, double>> ea, eb;
  for (int64_t p = b0; p < b1; ++p)
    for (int64_t q = p; q < b1; ++q) {
      if (row_mask && !row_mask[std::min(v->bs_dof[p], v->bs_dof[q])]) continue;
      const int pd = v->bs_dir[p], qd = v->bs_dir[q];
      double a = integrate(s, ai, pd, qd, v->bs_i[p], v->bs_j[p], v->bs_i[q], v->bs_j[q], local, local, eps, mu);
      double b = integrate(s, bi, pd, qd, v->bs_i[p], v->bs_j[p], v->bs_i[q], v->bs_j[q], local, local, eps, mu);
      Key k{uint32_t(v->bs_dof[p]), uint32_t(v->bs_dof[q])};
      ea.push_back({k, a});
      eb.push_back({k, b});
    }
  insert_group(out.a, ea);
  insert_group(out.b, eb);
  out.pairs += int64_t(ea.size());

  std::vector<int> desc;
  s.tree.descendants(e, desc);
  ea.clear();
  eb.clear();
  std::vector<std::pair<int, BasisFn>> sampled;  // the reference's sampler cache
  for (int d : desc) sampled.push_back({d, sample_basis_fn(s, e, d)});
  std::vector<BasisFn> desc_local;
  for (int d : desc) desc_local.push_back(sample_basis_fn(s, d, -1));
  for (int64_t p = b0; p < b1; ++p) {
    for (size_t k = 0; k < desc.size(); ++k) {
      const int d = desc[k];
      const BasisFn& ps = sampled[k].second;
      const BasisFn& ql = desc_local[k];
      for (int64_t q = v->bs_offsets[d]; q < v->bs_offsets[d + 1]; ++q) {
        if (row_mask && !row_mask[std::min(v->bs_dof[p], v->bs_dof[q])]) continue;
        const int pd = v->bs_dir[p], qd = v->bs_dir[q];
        double a = integrate(s, ai, pd, qd, v->bs_i[p], v->bs_j[p], v->bs_i[q], v->bs_j[q], ps, ql, eps, mu);
        double b = integrate(s, bi, pd, qd, v->bs_i[p], v->bs_j[p], v->bs_i[q], v->bs_j[q], ps, ql, eps, mu);
        Key key{uint32_t(v->bs_dof[p]), uint32_t(v->bs_dof[q])};
        ea.push_back({key, a});
        eb.push_back({key, b});
      }
    }
  }
  insert_group(out.a, ea);
  insert_group(out.b, eb);
  out.pairs += int64_t(ea.size());
}

}  // namespace oracle

// ---------------------------------------------------------------------------
// C interface (ctypes)
// ---------------------------------------------------------------------------

extern "C" {

typedef struct oracle_gep {
  int64_t n, nnz, pairs;
  int64_t* row_ptr;
  int32_t* col;
  double* a_val;
  double* b_val;
} oracle_gep;

// Gauss-Legendre points by Golub-Welsch; arrays hold n (+2 with endpoints) values.
int oracle_gauss_quadrature_points(int n, int include_endpoints, double* pts, double* wts) {
  std::vector<double> p, w;
  oracle::gauss_quadrature_points(n, include_endpoints != 0, p, w);
  for (size_t k = 0; k < p.size(); ++k) {
    pts[k] = p[k];
    wts[k] = w[k];
  }
  return 0;
}

// Shape tables [order][point] of N, N', T, T' (each (n_max+1)*np doubles).
int oracle_shape_tables(int basis, int n_max, int np, const double* x, double* n, double* nd, double* t, double* td) {
  std::vector<double> xs(x, x + np);
  oracle::Space s = oracle::make_space(basis, n_max, xs);
  for (int o = 0; o <= n_max; ++o)
    for (int p = 0; p < np; ++p) {
      n[o * np + p] = s.n[o][p];
      nd[o * np + p] = s.nd[o][p];
      t[o * np + p] = s.t[o][p];
      td[o * np + p] = s.td[o][p];
    }
  return 0;
}

int oracle_default_ngq(int max_order) { return oracle::default_ngq(max_order); }

// UniformFieldSpace::xy_fields (fields.rs:63-127): per leaf Elem (ascending
// id) the grids x_values / y_values, sized [densities[1]][densities[0]] and
// indexed [m][n] as the reference does, written to out_x / out_y as
// [leaf][m][n]. Returns 1 for MismatchedSolutionSize (fields.rs:68-72) and 2
// where the reference's grid indexing panics (out of range), else 0.
int oracle_xy_fields(const fem2d_domain_view* v, int basis, int du, int dv, const double* sol, int64_t n_sol,
                     double* out_x, double* out_y) {
  if (n_sol != v->n_dofs) return 1;
  oracle::Setup s(v);
  s.basis = basis;
  s.i_max = 0;
  s.j_max = 0;
  for (int e = 0; e < v->n_elems; ++e) {  // Mesh::max_expansion_orders (mesh.rs:626-630)
    s.i_max = std::max<int>(s.i_max, v->orders[2 * e]);
    s.j_max = std::max<int>(s.j_max, v->orders[2 * e + 1]);
  }
  auto uniform_range = [](double lo, double hi, int n) {  // fields.rs:402-405
    const double step = (hi - lo) / double(n - 1);
    std::vector<double> r;
    for (int i = 0; i < n; ++i) r.push_back(double(i) * step + lo);
    return r;
  };
  s.up = uniform_range(-1.0, 1.0, du);
  s.vp = uniform_range(-1.0, 1.0, dv);
  size_t out = 0;
  for (int e = 0; e < v->n_elems; ++e) {
    if (!s.tree.children[e].empty()) continue;
    std::vector<std::vector<double>> xv(size_t(dv), std::vector<double>(size_t(du), 0.0)), yv = xv;
    std::vector<int> anc = s.tree.path(e);  // root .. e
    std::reverse(anc.begin(), anc.end());   // Mesh::ancestor_elems(e, true): e first
    for (int a : anc) {
      const oracle::BasisFn bf = oracle::sample_basis_fn(s, a, e);
      for (int64_t t = v->bs_offsets[a]; t < v->bs_offsets[a + 1]; ++t)
        for (int m = 0; m < du; ++m)
          for (int n = 0; n < dv; ++n) {
            const int i = v->bs_i[t], j = v->bs_j[t];
            oracle::V2 value = v->bs_dir[t] == 0 ? bf.f_u(i, j, m, n)
                               : v->bs_dir[t] == 1 ? bf.f_v(i, j, m, n)
                                                   : oracle::V2{0.0, 0.0};
            value = oracle::vmul(value, sol[v->bs_dof[t]]);
            if (size_t(m) >= xv.size() || size_t(n) >= xv[size_t(m)].size()) return 2;
            xv[size_t(m)][size_t(n)] += value.x;
            yv[size_t(m)][size_t(n)] += value.y;
          }
    }
    for (int m = 0; m < du; ++m)
      for (int n = 0; n < dv; ++n) {
        out_x[out] = xv[size_t(m)][size_t(n)];
        out_y[out] = yv[size_t(m)][size_t(n)];
        ++out;
      }
  }
  return 0;
}

// Galerkin GEP assembly on the CPU. Returns the galerkin.rs error codes
// (1 wrong cc, 2 empty DoF set, 3 invalid GLQ settings) before any work.
// n_threads > 1 runs elem tasks in parallel; results are merged in elem order.
// elem_limit >= 0 restricts the loop to elems [0, elem_limit) (a bounded sample).
// row_mask (nullable, n_dofs bytes) restricts the result to the rows it sets.
int oracle_assemble(const fem2d_domain_view* v, int basis, int ai, int bi, int glq_u, int glq_v, int n_threads,
                    int64_t elem_limit, const uint8_t* row_mask, oracle_gep* out) {
  if (v->cc != FEM2D_CC_HCURL) return FEM2D_ERR_WRONG_CC;
  if (v->n_dofs == 0) return FEM2D_ERR_EMPTY_DOFS;
  const bool given = glq_u > 0 || glq_v > 0;
  if (given && (glq_u < FEM2D_MIN_GLQ_ORDER || glq_v < FEM2D_MIN_GLQ_ORDER)) return FEM2D_ERR_INVALID_GLQ;

  oracle::Setup s(v);
  s.basis = basis;
  s.i_max = 0;
  s.j_max = 0;
  for (int e = 0; e < v->n_elems; ++e) {
    s.i_max = std::max<int>(s.i_max, v->orders[2 * e]);
    s.j_max = std::max<int>(s.j_max, v->orders[2 * e + 1]);
  }
  const int nu = given ? glq_u : oracle::default_ngq(s.i_max);
  const int nv = given ? glq_v : oracle::default_ngq(s.j_max);
  oracle::gauss_quadrature_points(nu, false, s.up, s.uw);
  oracle::gauss_quadrature_points(nv, false, s.vp, s.vw);

  const int ne = elem_limit >= 0 ? int(std::min<int64_t>(elem_limit, v->n_elems)) : v->n_elems;
  std::vector<oracle::ElemResult> res(ne);
  const int nt = std::max(1, n_threads);
  if (nt == 1) {
    for (int e = 0; e < ne; ++e) oracle::assemble_elem(s, e, ai, bi, row_mask, res[e]);
  } else {
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
      pool.emplace_back([&]() {
        for (int e = next++; e < ne; e = next++) oracle::assemble_elem(s, e, ai, bi, row_mask, res[e]);
      });
    for (auto& th : pool) th.join();
  }
  // serial merge (linalg.rs:74-80)
  oracle::SparseMap A, B;
  int64_t pairs = 0;
  for (int e = 0; e < ne; ++e) {
    oracle::consume(A, res[e].a);
    oracle::consume(B, res[e].b);
    pairs += res[e].pairs;
    res[e] = oracle::ElemResult();
  }

  out->n = v->n_dofs;
  out->nnz = int64_t(A.size());
  out->pairs = pairs;
  out->row_ptr = static_cast<int64_t*>(std::calloc(size_t(out->n + 1), sizeof(int64_t)));
  out->col = static_cast<int32_t*>(std::malloc(std::max<size_t>(1, A.size()) * sizeof(int32_t)));
  out->a_val = static_cast<double*>(std::malloc(std::max<size_t>(1, A.size()) * sizeof(double)));
  out->b_val = static_cast<double*>(std::malloc(std::max<size_t>(1, A.size()) * sizeof(double)));
  if (!out->row_ptr || !out->col || !out->a_val || !out->b_val) return FEM2D_ERR_OOM;
  int64_t k = 0;
  auto ib = B.begin();
  for (const auto& [key, val] : A) {
    out->row_ptr[key.first + 1] += 1;
    out->col[k] = int32_t(key.second);
    out->a_val[k] = val;
    out->b_val[k] = ib->second;  // A and B share their key sets (galerkin.rs:124-125, 174-175)
    ++ib;
    ++k;
  }
  for (int64_t r = 0; r < out->n; ++r) out->row_ptr[r + 1] += out->row_ptr[r];
  return 0;
}

void oracle_gep_free(oracle_gep* g) {
  std::free(g->row_ptr);
  std::free(g->col);
  std::free(g->a_val);
  std::free(g->b_val);
  g->row_ptr = nullptr;
  g->col = nullptr;
  g->a_val = g->b_val = nullptr;
}

}  // extern "C"
