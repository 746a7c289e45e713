// MI355X (gfx950) Galerkin GEP assembly: flattened Domain in HBM -> upper-
// triangular CSR of A (CurlCurl) and B (L2Inner) on one shared pattern.
//
// What the reference computes (galerkin.rs:33-187): for every Elem E, every
// pair of E's own basis specs (upper triangle) and every pair (spec of E, spec
// of a descendant D of E), a 2D Gauss-Legendre quadrature of the CurlCurl and
// L2Inner integrands (integrals.rs:26-92, 292-354), accumulated by the key
// [min(dof), max(dof)] (sparse_matrix.rs:68-120).
//
// How it is computed here:
//  * The elements are axis-aligned and affine, so the Jacobian is diagonal and
//    constant per Elem, and each integrand factors into f(x_m) g(y_n). The 2D
//    quadrature is therefore exactly the product of two 1D sums. A 1D sum only
//    depends on the two orders, on which shape functions enter (N, N', T, T')
//    and on the dyadic sub-range the descendant occupies inside the ancestor
//    (basis.rs:372-393). All such 1D sums form a small "Gram library" built
//    once per call; every matrix entry is a scalar prefactor times two library
//    values.
//  * Rows are processed by "row groups": maximal runs of DoFs whose basis specs
//    live on the same Elem(s) (the elem-type DoFs of one leaf, or the DoFs of
//    one edge). Every row of a group has the same candidate column set: the
//    DoFs on the Elems nested with (equal to, ancestor of or descendant of) the
//    group's Elem(s). That set is a union of whole row groups, built per group
//    in LDS (sources -> memberships -> bitonic sort -> dedupe). A row's columns
//    are the suffix of that set starting at the row itself.
//  * Each stored entry (r, c) receives one contribution per nested pair of
//    basis specs of r and c. Two pairs of interior-disjoint Elems can have at
//    most two nested pairs, so an entry has one or two contributions; they are
//    summed in a fixed order and written once. Kernels check this bound and
//    report a violation instead of writing wrong values.
//  * Pattern (count pass + scan) and values (fill pass) are computed on the
//    device in the same call; the fill pass writes col/A/B with consecutive
//    threads on consecutive CSR slots.

#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fem2d.h"
#include "host/capi_internal.hpp"

namespace {

// ---------------------------------------------------------------------------
// Limits of this build (checked; exceeded limits return FEM2D_ERR_UNSUPPORTED)
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;        // threads per row-group block
constexpr int kMaxSrc = 128;       // nested Elems per row group
constexpr int kMaxRaw = 1024;      // (Elem, row group) memberships gathered per row group
constexpr int kMaxDist = 1024;     // distinct row groups in one candidate set
constexpr int kMaxRows = 1024;     // rows per row group
constexpr int kMaxDepth = 24;      // h-refinement depth below a base Elem
constexpr int kMaxRelDepth = 12;   // dyadic halvings per axis covered by the Gram library
constexpr int kMaxOrder = 20;      // MAX_POLYNOMIAL_ORDER (mesh.rs:42)
constexpr int kMaxGlq = 64;        // Gauss points per axis
constexpr int kKinds = 5;          // Gram kinds

enum GramKind { kNN = 0, kNTd = 1, kTdN = 2, kTdTd = 3, kTT = 4 };
enum Rel { kLocal = 0, kAnc = 1, kDesc = 2 };

// device error flags
enum ErrBits : int {
  kErrView = 1,        // view violates Domain invariants
  kErrSize = 2,        // a per-group limit of this build was exceeded
  kErrContrib = 4,     // an entry with more than two contributions
  kErrDepth = 8,       // refinement deeper than kMaxDepth / kMaxRelDepth
};

struct ElemGeo {
  double ju0;     // jac_inv.u[0] = b / det
  double jv1;     // jac_inv.v[1] = a / det
  double det;     // a * b
  double uv;      // a / b
  double vu;      // b / a
  double eps;     // eps_rel.re
  double inv_mu;  // 1 / mu_rel.re
  double pad;
};

struct DevView {
  int n_elems;
  long long n_dofs;
  long long n_bs;
  const int* parent;
  const unsigned char* loc;
  const double* box;
  const double* mat;
  const long long* bs_off;
  const unsigned char* bs_dir;
  const unsigned char* bs_i;
  const unsigned char* bs_j;
  const int* bs_dof;
};

struct Work {
  ElemGeo* geo;
  int* first_child;
  int* max_child;
  int* n_child;
  int4* dyad;  // (du, iu, dv, iv): position of the Elem inside its base Elem
  int* dof_e1;
  int* dof_e2;
  int* dof_k1;
  int* dof_k2;
  int* dof_cnt;
  int* dof_flag;
  int* dof_grp;  // inclusive scan of dof_flag
  int* g_start;  // [G+1]
  int* g_e1;
  int* g_e2;
  int* mem_cnt;  // [E+1] membership counts per Elem
  int* mem_ptr;  // [E+1] exclusive offsets of mem_cnt
  int* mem_group;
  int* mem_bsoff;
  int64_t* row_len;  // [N+1] stored entries per row
  int64_t* row_ptr;  // [N+1] exclusive offsets of row_len
  double* lib;
  const double* glq;  // u points, u weights, v points, v weights (kMaxGlq each)
  int* err;
};

struct LibShape {
  int p1;     // orders 0..p1-1 per axis
  int nr;     // ranges per axis
  int nu, nv; // Gauss points per axis
  int basis;
};

__device__ __forceinline__ void set_err(int* err, int bits) { atomicOr(err, bits); }

// ---------------------------------------------------------------------------
// Shape-function spaces (point evaluation, same recurrences and operation
// order as the reference tables)
// ---------------------------------------------------------------------------

// HierPoly (hierarchical_basis_fns.rs:102-162): N_n = x^n by repeated
// multiplication, N'_n = n x^(n-1); T_0 = 1-x, T_1 = 1+x, T_n = x^n - 1 (even)
// or x^n - x (odd).
__device__ void hier_poly(int n, double x, double& N, double& Nd, double& T, double& Td) {
  if (n == 0) {
    N = 1.0; Nd = 0.0; T = 1.0 - x; Td = -1.0;
    return;
  }
  if (n == 1) {
    N = x; Nd = 1.0; T = 1.0 + x; Td = 1.0;
    return;
  }
  double prev = x;  // x^1
  double pw = x;
  for (int k = 2; k <= n; ++k) {
    prev = pw;
    pw = prev * x;
  }
  const double pd = double(n) * prev;
  N = pw;
  Nd = pd;
  if (n % 2 == 0) {
    T = pw - 1.0;
    Td = pd;
  } else {
    T = pw - x;
    Td = pd - 1.0;
  }
}

// KOLShapeFn: Legendre N_n = L_n (three-term recurrence), T_0/T_1 as HierPoly,
// T_n = c_n * sum_{k<=n} q_{n,k} L_k with q_{n,n} = 1 and
// q_{n,k} = -(2k+1) / (n(n-1)/2) for k < n, k = n (mod 2); c_2..c_8 are the
// Euclidean normalisation constants of hierarchical_basis_fns.rs:206-209.
__device__ void kol(int n, double x, double& N, double& Nd, double& T, double& Td) {
  double L[kMaxOrder + 1], Ld[kMaxOrder + 1];
  for (int i = 0; i <= n; ++i) {
    const double fi = double(i);
    if (i == 0) {
      L[i] = 1.0; Ld[i] = 0.0;
    } else if (i == 1) {
      L[i] = x; Ld[i] = 1.0;
    } else {
      L[i] = ((2.0 * fi - 1.0) * x * L[i - 1] - (fi - 1.0) * L[i - 2]) / fi;
      Ld[i] = fi * L[i - 1] + x * Ld[i - 1];
    }
  }
  N = L[n];
  Nd = Ld[n];
  if (n == 0) {
    T = 1.0 - x; Td = -1.0;
    return;
  }
  if (n == 1) {
    T = 1.0 + x; Td = 1.0;
    return;
  }
  const double kNorm[7] = {0.968246, 2.561738, 0.838525, 4.248161, 0.816397, 5.882766, 0.808509};
  const int den = n * (n - 1) / 2;
  const double c = n - 2 < 7 ? kNorm[n - 2] : 1.0;
  double sv = 0.0, sd = 0.0;
  for (int k = 0; k <= n; ++k) {
    int num = 0;
    if (k == n) num = den;
    else if ((n - k) % 2 == 0) num = -(2 * k + 1);
    const double w = double(num) / double(den);
    sv += w * L[k];
    sd += w * Ld[k];
  }
  T = sv * c;
  Td = sd * c;
}

__device__ __forceinline__ double shape_pick(int basis, int which, int n, double x) {
  double N, Nd, T, Td;
  if (basis == FEM2D_BASIS_KOL) kol(n, x, N, Nd, T, Td);
  else hier_poly(n, x, N, Nd, T, Td);
  switch (which) {
    case 0: return N;
    case 1: return Nd;
    case 2: return T;
    default: return Td;
  }
}

// dyadic sub-range id -> (scale, offset) of the mapped Gauss points
__device__ __forceinline__ void range_of(int rid, double& s, double& c) {
  int du = 31 - __clz(rid + 1);
  int iu = rid + 1 - (1 << du);
  s = ldexp(1.0, -du);
  double lo = -1.0 + double(iu) * ldexp(1.0, 1 - du);
  c = lo + s;
}

// ---------------------------------------------------------------------------
// Prep kernels
// ---------------------------------------------------------------------------

__global__ void k_init(Work w, int n_elems, long long n_dofs) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t < n_elems) {
    w.first_child[t] = 0x7fffffff;
    w.max_child[t] = -1;
    w.n_child[t] = 0;
  }
  if (t < n_dofs) {
    w.dof_e1[t] = 0x7fffffff;
    w.dof_e2[t] = -1;
    w.dof_k1[t] = -1;
    w.dof_k2[t] = -1;
    w.dof_cnt[t] = 0;
  }
  if (t == 0) {
    *w.err = 0;
    w.mem_cnt[n_elems] = 0;
    w.row_len[n_dofs] = 0;
  }
}

__device__ __forceinline__ double map_range(double val, double in_min, double in_max, double out_min,
                                            double out_max) {
  return (val - in_min) * (out_max - out_min) / (in_max - in_min) + out_min;
}

// Elem geometry: parametric range from the chain of HRefLocs (elem.rs:191-197,
// h_refinement.rs:247-279), Jacobian over that range (element.rs:33-50),
// inverse/determinant (space.rs:138-147) and the Elem's dyadic position.
__global__ void k_elem_prep(DevView v, Work w) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= v.n_elems) return;
  unsigned char locs[kMaxDepth];
  int depth = 0;
  int x = e;
  while (v.parent[x] >= 0) {
    const int p = v.parent[x];
    if (p >= v.n_elems || p == x || depth >= kMaxDepth) {
      set_err(w.err, p >= v.n_elems || p == x ? kErrView : kErrDepth);
      return;
    }
    locs[depth++] = v.loc[x];
    x = p;
  }
  if (v.parent[e] >= 0) {
    const int p = v.parent[e];
    atomicMin(&w.first_child[p], e);
    atomicMax(&w.max_child[p], e);
    atomicAdd(&w.n_child[p], 1);
  }
  double u0 = -1.0, u1 = 1.0, v0 = -1.0, v1 = 1.0;
  int du = 0, iu = 0, dv = 0, iv = 0;
  for (int t = depth - 1; t >= 0; --t) {
    const double mu = (u0 + u1) / 2.0, mv = (v0 + v1) / 2.0;
    const int l = locs[t];
    const bool split_u = l <= FEM2D_LOC_E;  // SW, SE, NW, NE, W, E
    const bool split_v = l <= FEM2D_LOC_NE || l >= FEM2D_LOC_S;
    const bool hi_u = l == FEM2D_LOC_SE || l == FEM2D_LOC_NE || l == FEM2D_LOC_E;
    const bool hi_v = l == FEM2D_LOC_NW || l == FEM2D_LOC_NE || l == FEM2D_LOC_N;
    if (l > FEM2D_LOC_N) {
      set_err(w.err, kErrView);
      return;
    }
    if (split_u) {
      if (hi_u) u0 = mu; else u1 = mu;
      iu = 2 * iu + (hi_u ? 1 : 0);
      ++du;
    }
    if (split_v) {
      if (hi_v) v0 = mv; else v1 = mv;
      iv = 2 * iv + (hi_v ? 1 : 0);
      ++dv;
    }
  }
  if (du > kMaxRelDepth || dv > kMaxRelDepth) set_err(w.err, kErrDepth);
  w.dyad[e] = make_int4(du, iu, dv, iv);

  const double* b = v.box + 4 * e;
  const double x_min = map_range(u0, -1.0, 1.0, b[0], b[2]);
  const double x_max = map_range(u1, -1.0, 1.0, b[0], b[2]);
  const double y_min = map_range(v0, -1.0, 1.0, b[1], b[3]);
  const double y_max = map_range(v1, -1.0, 1.0, b[1], b[3]);
  const double a = (x_max - x_min) / 2.0;
  const double bb = (y_max - y_min) / 2.0;
  ElemGeo g;
  g.det = a * bb - 0.0 * 0.0;
  g.ju0 = bb / g.det;
  g.jv1 = a / g.det;
  g.uv = a / bb;
  g.vu = bb / a;
  g.eps = v.mat[2 * e];
  g.inv_mu = 1.0 / v.mat[2 * e + 1];
  g.pad = 0.0;
  w.geo[e] = g;
}

// children of an Elem must have consecutive ids (h_refinement.rs:119-134)
__global__ void k_child_check(DevView v, Work w) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= v.n_elems) return;
  const int n = w.n_child[e];
  if (n == 0) return;
  if (!(n == 2 || n == 4) || w.max_child[e] - w.first_child[e] + 1 != n) set_err(w.err, kErrView);
}

// DoF -> the (at most two) Elems holding its basis specs
__global__ void k_dof_sig(DevView v, Work w) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= v.n_elems) return;
  for (long long t = v.bs_off[e]; t < v.bs_off[e + 1]; ++t) {
    const int d = v.bs_dof[t];
    if (d < 0 || d >= v.n_dofs) {
      set_err(w.err, kErrView);
      continue;
    }
    atomicMin(&w.dof_e1[d], e);
    atomicMax(&w.dof_e2[d], e);
    atomicAdd(&w.dof_cnt[d], 1);
  }
}

__global__ void k_dof_idx(DevView v, Work w) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= v.n_elems) return;
  const long long b0 = v.bs_off[e];
  for (long long t = b0; t < v.bs_off[e + 1]; ++t) {
    const int d = v.bs_dof[t];
    if (d < 0 || d >= v.n_dofs) continue;
    const int cnt = w.dof_cnt[d];
    const int e1 = w.dof_e1[d], e2 = w.dof_e2[d];
    if (cnt == 2 && e1 == e2) {
      set_err(w.err, kErrView);  // two specs of one DoF on the same Elem
      continue;
    }
    if (e == e1) w.dof_k1[d] = int(t - b0);
    else if (e == e2) w.dof_k2[d] = int(t - b0);
  }
}

__global__ void k_dof_flags(Work w, long long n_dofs) {
  long long d = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (d >= n_dofs) return;
  const int cnt = w.dof_cnt[d];
  if (cnt < 1 || cnt > 2) set_err(w.err, cnt > 2 ? kErrContrib : kErrView);
  const int e1 = w.dof_e1[d];
  const int e2 = cnt == 2 ? w.dof_e2[d] : -1;
  int flag = 1;
  if (d > 0) {
    const int pc = w.dof_cnt[d - 1];
    const int p1 = w.dof_e1[d - 1];
    const int p2 = pc == 2 ? w.dof_e2[d - 1] : -1;
    flag = (p1 != e1 || p2 != e2) ? 1 : 0;
  }
  w.dof_flag[d] = flag;
}

__global__ void k_group_starts(Work w, long long n_dofs) {
  long long d = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (d >= n_dofs) return;
  if (w.dof_flag[d]) {
    const int g = w.dof_grp[d] - 1;
    w.g_start[g] = int(d);
    w.g_e1[g] = w.dof_e1[d];
    w.g_e2[g] = w.dof_cnt[d] == 2 ? w.dof_e2[d] : -1;
  }
  if (d == n_dofs - 1) w.g_start[w.dof_grp[d]] = int(n_dofs);
}

// Each Elem's basis-spec list as runs of whole row groups (memberships).
__global__ void k_mem_count(DevView v, Work w) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= v.n_elems) return;
  const long long b0 = v.bs_off[e], b1 = v.bs_off[e + 1];
  int runs = 0;
  long long t = b0;
  while (t < b1) {
    const int d0 = v.bs_dof[t];
    if (d0 < 0 || d0 >= v.n_dofs) {
      set_err(w.err, kErrView);
      break;
    }
    const int g = w.dof_grp[d0] - 1;
    const int gs = w.g_start[g], gn = w.g_start[g + 1] - gs;
    if (d0 != gs || t + gn > b1) {
      set_err(w.err, kErrView);
      break;
    }
    for (int k = 1; k < gn; ++k)
      if (v.bs_dof[t + k] != d0 + k) set_err(w.err, kErrView);
    ++runs;
    t += gn;
  }
  w.mem_cnt[e] = runs;
}

__global__ void k_mem_fill(DevView v, Work w) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= v.n_elems) return;
  const long long b0 = v.bs_off[e], b1 = v.bs_off[e + 1];
  int slot = w.mem_ptr[e];
  const int end = w.mem_ptr[e + 1];
  long long t = b0;
  while (t < b1 && slot < end) {
    const int d0 = v.bs_dof[t];
    if (d0 < 0 || d0 >= v.n_dofs) break;
    const int g = w.dof_grp[d0] - 1;
    const int gn = w.g_start[g + 1] - w.g_start[g];
    w.mem_group[slot] = g;
    w.mem_bsoff[slot] = int(t - b0);
    ++slot;
    t += gn;
  }
}

// 1D Gram library: lib[axis][range][kind][a][b] = sum_m w_m f_a(x'_m) g_b(x_m)
// with x'_m = x_m * s + c mapped into the dyadic sub-range (glq.rs:238-249).
__global__ void k_gram_lib(Work w, LibShape L) {
  const long long per_axis = (long long)L.nr * kKinds * L.p1 * L.p1;
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= 2 * per_axis) return;
  const int axis = int(t / per_axis);
  long long r = t % per_axis;
  const int b = int(r % L.p1);
  r /= L.p1;
  const int a = int(r % L.p1);
  r /= L.p1;
  const int kind = int(r % kKinds);
  const int rid = int(r / kKinds);
  double s, c;
  range_of(rid, s, c);
  const int np = axis == 0 ? L.nu : L.nv;
  const double* pts = w.glq + (axis == 0 ? 0 : 2 * kMaxGlq);
  const double* wts = pts + kMaxGlq;
  // which shape of p (first) and q (second): 0 N, 1 N', 2 T, 3 T'
  int fp = 0, fq = 0;
  switch (kind) {
    case kNN: fp = 0; fq = 0; break;
    case kNTd: fp = 0; fq = 3; break;
    case kTdN: fp = 3; fq = 0; break;
    case kTdTd: fp = 3; fq = 3; break;
    default: fp = 2; fq = 2; break;
  }
  double acc = 0.0;
  for (int m = 0; m < np; ++m) {
    const double xq = pts[m];
    const double xp = rid == 0 ? xq : xq * s + c;
    acc += wts[m] * shape_pick(L.basis, fp, a, xp) * shape_pick(L.basis, fq, b, xq);
  }
  w.lib[t] = acc;
}

// ---------------------------------------------------------------------------
// Row-group candidate sets (shared by the count and fill passes)
// ---------------------------------------------------------------------------

struct SrcRec {
  double cA[4];  // CurlCurl prefactors UU, UV, VU, VV
  double cB[2];  // L2Inner prefactors UU, VV
  int x;         // nested Elem
  int k;         // which Elem of the row group (0, 1)
  int rel;       // kLocal, kAnc, kDesc
  int ru, rv;    // Gram-library range ids
  int pad;
};

struct GroupShared {
  SrcRec src[kMaxSrc];
  unsigned long long key[kMaxRaw];
  short raw_src[kMaxRaw];
  int raw_off[kMaxRaw];
  int dstart[kMaxDist];
  int cum[kMaxDist + 1];
  short ds0[kMaxDist], ds1[kMaxDist];
  int doff0[kMaxDist], doff1[kMaxDist];
  int row_rel[kMaxRows + 1];
  unsigned int row_bs[2][kMaxRows];
  int n_src, n_raw, n_dist, total, u_self, fail;
  int raw_cnt[kMaxSrc + 1];
};

static_assert(sizeof(GroupShared) <= 64 * 1024, "row-group LDS footprint must fit 64 KiB");

__device__ __forceinline__ int rel_range_id(int4 P, int4 Q, int axis) {
  const int dp = axis == 0 ? P.x : P.z;
  const int ip = axis == 0 ? P.y : P.w;
  const int dq = axis == 0 ? Q.x : Q.z;
  const int iq = axis == 0 ? Q.y : Q.w;
  const int dd = dq - dp;
  const int ii = iq - (ip << dd);
  return (1 << dd) - 1 + ii;
}

// Builds the candidate set of row group g in LDS. Returns false (block-uniform)
// if a limit was exceeded; the error flag is set then.
template <bool kWithCoef>
__device__ bool build_group(const DevView& v, const Work& w, int g, GroupShared& S) {
  const int tid = threadIdx.x;
  if (tid == 0) {
    S.fail = 0;
    int n = 0;
    const int e_k[2] = {w.g_e1[g], w.g_e2[g]};
    for (int k = 0; k < 2 && !S.fail; ++k) {
      const int E = e_k[k];
      if (E < 0) break;
      // E itself
      if (n >= kMaxSrc) { S.fail = 1; break; }
      S.src[n].x = E; S.src[n].k = k; S.src[n].rel = kLocal; ++n;
      // ancestors
      for (int a = v.parent[E]; a >= 0; a = v.parent[a]) {
        if (n >= kMaxSrc) { S.fail = 1; break; }
        S.src[n].x = a; S.src[n].k = k; S.src[n].rel = kAnc; ++n;
      }
      // descendants, breadth first, using the source list as the queue
      int q = n;
      int head = -1;  // start with E's children
      while (!S.fail) {
        const int parent_e = head < 0 ? E : S.src[head].x;
        const int nc = w.n_child[parent_e];
        const int fc = w.first_child[parent_e];
        for (int c = 0; c < nc; ++c) {
          if (n >= kMaxSrc) { S.fail = 1; break; }
          S.src[n].x = fc + c; S.src[n].k = k; S.src[n].rel = kDesc; ++n;
        }
        head = head < 0 ? q : head + 1;
        if (head >= n) break;
      }
    }
    S.n_src = n;
    // membership counts -> raw offsets
    int tot = 0;
    for (int s = 0; s < n && !S.fail; ++s) {
      S.raw_cnt[s] = tot;
      const int X = S.src[s].x;
      tot += w.mem_ptr[X + 1] - w.mem_ptr[X];
    }
    S.raw_cnt[n] = tot;
    if (tot > kMaxRaw) S.fail = 1;
    S.n_raw = tot;
  }
  __syncthreads();
  {
    const int fail = S.fail;
    __syncthreads();  // every thread has read the flag before it can be reset
    if (fail) {
      if (tid == 0) set_err(w.err, kErrSize);
      return false;
    }
  }
  const int n_src = S.n_src, n_raw = S.n_raw;
  // gather memberships
  for (int s = tid; s < n_src; s += blockDim.x) {
    const int X = S.src[s].x;
    const int m0 = w.mem_ptr[X], m1 = w.mem_ptr[X + 1];
    int r = S.raw_cnt[s];
    for (int m = m0; m < m1; ++m, ++r) {
      S.key[r] = ((unsigned long long)(unsigned)w.mem_group[m] << 20) | (unsigned long long)r;
      S.raw_src[r] = short(s);
      S.raw_off[r] = w.mem_bsoff[m];
    }
    if (kWithCoef) {
      // block prefactors (see file header): P = p-side Elem, Q = q-side Elem
      const int E = S.src[s].k == 0 ? w.g_e1[g] : w.g_e2[g];
      const int rel = S.src[s].rel;
      const int P = rel == kAnc ? X : E;
      const int Q = rel == kAnc ? E : X;
      const ElemGeo gp = w.geo[P], gq = w.geo[Q];
      int ru = 0, rv = 0;
      double su = 1.0, sv = 1.0;
      if (rel != kLocal) {
        const int4 dp = w.dyad[P], dq = w.dyad[Q];
        ru = rel_range_id(dp, dq, 0);
        rv = rel_range_id(dp, dq, 1);
        su = ldexp(1.0, -(dq.x - dp.x));
        sv = ldexp(1.0, -(dq.z - dp.z));
      }
      const bool p_big = gp.det >= gq.det;
      const double r_uv = p_big ? gp.uv : gq.uv;
      const double r_vu = p_big ? gp.vu : gq.vu;
      const double dmax = gp.det >= gq.det ? gp.det : gq.det;
      const double im = gp.inv_mu;
      S.src[s].cA[0] = im * ((gp.ju0 * 1.0) * (gq.ju0 * su)) * r_uv;
      S.src[s].cA[1] = -(im * (gp.ju0 * (gq.jv1 * sv)));
      S.src[s].cA[2] = -(im * (gp.jv1 * (gq.ju0 * su)));
      S.src[s].cA[3] = im * ((gp.jv1 * 1.0) * (gq.jv1 * sv)) * r_vu;
      const double gs = su * sv;
      S.src[s].cB[0] = gp.eps * gs * dmax * gp.ju0 * gq.ju0;
      S.src[s].cB[1] = gp.eps * gs * dmax * gp.jv1 * gq.jv1;
      S.src[s].ru = ru;
      S.src[s].rv = rv;
    }
  }
  // pad and sort (bitonic) by (row group, gather order)
  int np2 = 1;
  while (np2 < n_raw) np2 <<= 1;
  for (int r = n_raw + tid; r < np2; r += blockDim.x) S.key[r] = ~0ull;
  __syncthreads();
  for (int kk = 2; kk <= np2; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < np2; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const unsigned long long a = S.key[i], b = S.key[ixj];
          const bool up = (i & kk) == 0;
          if ((a > b) == up) {
            S.key[i] = b;
            S.key[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  // dedupe into distinct row groups (one or two sources each) and prefix sums
  if (tid == 0) {
    int nd = 0, run = 0, self = -1, prev = -1;
    for (int t = 0; t < n_raw; ++t) {
      const int h = int(S.key[t] >> 20);
      const int r = int(S.key[t] & 0xFFFFF);
      if (h != prev) {
        if (nd >= kMaxDist) { S.fail = 1; break; }
        const int gs = w.g_start[h];
        S.dstart[nd] = gs;
        S.cum[nd] = run;
        S.ds0[nd] = S.raw_src[r];
        S.doff0[nd] = S.raw_off[r];
        S.ds1[nd] = -1;
        S.doff1[nd] = 0;
        if (h == g) self = run;
        run += w.g_start[h + 1] - gs;
        ++nd;
        prev = h;
      } else {
        if (S.ds1[nd - 1] >= 0) { S.fail = 2; break; }
        S.ds1[nd - 1] = S.raw_src[r];
        S.doff1[nd - 1] = S.raw_off[r];
      }
    }
    S.cum[nd] = run;
    S.n_dist = nd;
    S.total = run;
    S.u_self = self;
    if (self < 0 && !S.fail) S.fail = 3;
  }
  __syncthreads();
  const int fail = S.fail;
  __syncthreads();
  if (fail) {
    if (tid == 0) set_err(w.err, fail == 2 ? kErrContrib : (fail == 3 ? kErrView : kErrSize));
    return false;
  }
  return true;
}

__global__ void __launch_bounds__(kBlock) k_group_count(DevView v, Work w, int n_groups) {
  extern __shared__ unsigned char smem[];
  GroupShared& S = *reinterpret_cast<GroupShared*>(smem);
  for (int g = blockIdx.x; g < n_groups; g += gridDim.x) {
    const int r0 = w.g_start[g], r1 = w.g_start[g + 1];
    if (r1 - r0 > kMaxRows) {
      if (threadIdx.x == 0) set_err(w.err, kErrSize);
      continue;
    }
    if (!build_group<false>(v, w, g, S)) continue;
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x)
      w.row_len[r] = (long long)(S.total - (S.u_self + (r - r0)));
    __syncthreads();
  }
}

__device__ __forceinline__ unsigned pack_bs(const DevView& v, long long t) {
  return unsigned(v.bs_dir[t]) | (unsigned(v.bs_i[t]) << 8) | (unsigned(v.bs_j[t]) << 16);
}

// CurlCurl and L2Inner contributions of one nested pair (p, q)
__device__ __forceinline__ void pair_values(const double* __restrict__ lib, const LibShape& L, const SrcRec& s,
                                            unsigned pbs, unsigned qbs, double& cc, double& l2) {
  const int pd = pbs & 0xff, pi = (pbs >> 8) & 0xff, pj = (pbs >> 16) & 0xff;
  const int qd = qbs & 0xff, qi = (qbs >> 8) & 0xff, qj = (qbs >> 16) & 0xff;
  const long long tab = (long long)L.p1 * L.p1;
  const double* gx = lib + (long long)s.ru * kKinds * tab;
  const double* gy = lib + ((long long)L.nr + s.rv) * kKinds * tab;
  const int ix = pi * L.p1 + qi, iy = pj * L.p1 + qj;
  cc = 0.0;
  l2 = 0.0;
  if (pd == FEM2D_DIR_U && qd == FEM2D_DIR_U) {
    const double xnn = gx[kNN * tab + ix];
    cc = s.cA[0] * xnn * gy[kTdTd * tab + iy];
    l2 = s.cB[0] * xnn * gy[kTT * tab + iy];
  } else if (pd == FEM2D_DIR_U && qd == FEM2D_DIR_V) {
    cc = s.cA[1] * gx[kNTd * tab + ix] * gy[kTdN * tab + iy];
  } else if (pd == FEM2D_DIR_V && qd == FEM2D_DIR_U) {
    cc = s.cA[2] * gx[kTdN * tab + ix] * gy[kNTd * tab + iy];
  } else if (pd == FEM2D_DIR_V && qd == FEM2D_DIR_V) {
    const double ynn = gy[kNN * tab + iy];
    cc = s.cA[3] * gx[kTdTd * tab + ix] * ynn;
    l2 = s.cB[1] * gx[kTT * tab + ix] * ynn;
  }
}

__global__ void __launch_bounds__(kBlock) k_group_fill(DevView v, Work w, int n_groups, LibShape L, int a_is_cc,
                                                      int b_is_cc, int* __restrict__ col, double* __restrict__ aval,
                                                      double* __restrict__ bval) {
  extern __shared__ unsigned char smem[];
  GroupShared& S = *reinterpret_cast<GroupShared*>(smem);
  const int64_t* __restrict__ row_ptr = w.row_ptr;
  for (int g = blockIdx.x; g < n_groups; g += gridDim.x) {
    const int r0 = w.g_start[g], r1 = w.g_start[g + 1];
    const int nrows = r1 - r0;
    if (nrows > kMaxRows) continue;
    if (!build_group<true>(v, w, g, S)) continue;
    const long long seg0 = row_ptr[r0];
    const int e_k[2] = {w.g_e1[g], w.g_e2[g]};
    for (int i = threadIdx.x; i <= nrows; i += blockDim.x) {
      S.row_rel[i] = int(row_ptr[r0 + i] - seg0);
      if (i < nrows) {
        const int r = r0 + i;
        S.row_bs[0][i] = pack_bs(v, v.bs_off[e_k[0]] + w.dof_k1[r]);
        S.row_bs[1][i] = e_k[1] >= 0 ? pack_bs(v, v.bs_off[e_k[1]] + w.dof_k2[r]) : 0u;
      }
    }
    __syncthreads();
    const int seg_n = S.row_rel[nrows];
    const int n_dist = S.n_dist;
    for (int p = threadIdx.x; p < seg_n; p += blockDim.x) {
      // row of this slot
      int lo = 0, hi = nrows;  // row_rel[lo] <= p < row_rel[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (S.row_rel[mid] <= p) lo = mid; else hi = mid;
      }
      const int u = S.u_self + lo + (p - S.row_rel[lo]);
      // distinct row group holding union index u
      int a = 0, b = n_dist;  // cum[a] <= u < cum[b]
      while (b - a > 1) {
        const int mid = (a + b) >> 1;
        if (S.cum[mid] <= u) a = mid; else b = mid;
      }
      const int within = u - S.cum[a];
      const int c = S.dstart[a] + within;
      double cc_sum = 0.0, l2_sum = 0.0;
      for (int side = 0; side < 2; ++side) {
        const int sidx = side == 0 ? S.ds0[a] : S.ds1[a];
        if (sidx < 0) break;
        const SrcRec& s = S.src[sidx];
        const int off = side == 0 ? S.doff0[a] : S.doff1[a];
        const unsigned cbs = pack_bs(v, v.bs_off[s.x] + off + within);
        const unsigned rbs = S.row_bs[s.k][lo];
        double cc, l2;
        if (s.rel == kAnc) pair_values(w.lib, L, s, cbs, rbs, cc, l2);
        else pair_values(w.lib, L, s, rbs, cbs, cc, l2);
        cc_sum += cc;
        l2_sum += l2;
      }
      const long long o = seg0 + p;
      col[o] = c;
      aval[o] = a_is_cc ? cc_sum : l2_sum;
      bval[o] = b_is_cc ? cc_sum : l2_sum;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct fem2d_device_domain {
  int device = 0;
  int cc = 0;
  int n_elems = 0;
  long long n_dofs = 0;
  long long n_bs = 0;
  int max_i = 0, max_j = 0;
  int max_du = 0, max_dv = 0;
  // device arrays
  int* parent = nullptr;
  unsigned char* loc = nullptr;
  double* box = nullptr;
  double* mat = nullptr;
  long long* bs_off = nullptr;
  unsigned char* bs_dir = nullptr;
  unsigned char* bs_i = nullptr;
  unsigned char* bs_j = nullptr;
  int* bs_dof = nullptr;
};

struct fem2d_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  std::vector<DevBuf> bufs;  // grow-only work buffers, by slot
  DevBuf scan_tmp;
  DevBuf out_col, out_a, out_b;
  int* h_pinned = nullptr;  // err, n_groups, nnz (as two ints) staging
  fem2d_timing timing{};
  int max_grid = 0;
};

namespace {

using fem2d::capi::set_error;

int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e), 0);
  return e == hipErrorOutOfMemory ? FEM2D_ERR_OOM : FEM2D_ERR_HIP;
}

#define HIP_TRY(expr)                                  \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_fail(_e, #expr);  \
  } while (0)

int ensure(DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (b.bytes >= bytes) return FEM2D_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  size_t want = bytes + bytes / 8;
  HIP_TRY(hipMalloc(&b.p, want));
  b.bytes = want;
  return FEM2D_OK;
}

template <class T>
int slot(fem2d_ctx* ctx, int k, size_t count, T** out) {
  if (int(ctx->bufs.size()) <= k) ctx->bufs.resize(k + 1);
  if (int rc = ensure(ctx->bufs[k], count * sizeof(T))) return rc;
  *out = static_cast<T*>(ctx->bufs[k].p);
  return FEM2D_OK;
}

template <class T>
int upload(const T* src, size_t count, T** dst) {
  *dst = nullptr;
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(dst), std::max<size_t>(1, count) * sizeof(T)));
  if (count) HIP_TRY(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return FEM2D_OK;
}

int default_ngq(int max_order) {
  // basis.rs:172-177 (f32 arithmetic as in the reference)
  float conv = float(max_order * 4);
  float p2 = std::ceil(std::log2(conv));
  if (!std::isfinite(p2)) return 0;
  return int(std::lround(std::pow(2.0f, p2)));
}

unsigned grid_for(long long n, int block) { return unsigned(std::max<long long>(1, (n + block - 1) / block)); }

int check_common(int cc, long long n_dofs, int glq_u, int glq_v, int basis, int ai, int bi, int precision) {
  if (cc != FEM2D_CC_HCURL) {
    static const char* names[] = {"H(Curl)", "H(Div)", "Discontinuous"};
    set_error(std::string("Wrong Continuity Condition on Domain (required: H(Curl), received: ") +
                  (cc >= 0 && cc <= 2 ? names[cc] : "?") + "); Cannot execute Galerkin Sampling!",
              0);
    return FEM2D_ERR_WRONG_CC;
  }
  if (n_dofs == 0) {
    set_error("No Degrees-of-Freedom Defined over Domain; Cannot execute Galerkin Sampling!", 0);
    return FEM2D_ERR_EMPTY_DOFS;
  }
  if ((glq_u != 0 || glq_v != 0) && (glq_u < FEM2D_MIN_GLQ_ORDER || glq_v < FEM2D_MIN_GLQ_ORDER)) {
    set_error("Invalid GLQ Settings (the number of GLQ points must be at least 4); Cannot execute Galerkin Sampling!",
              0);
    return FEM2D_ERR_INVALID_GLQ;
  }
  if (basis != FEM2D_BASIS_HIERPOLY && basis != FEM2D_BASIS_KOL) {
    set_error("unknown basis space", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if ((ai != FEM2D_INTEGRAL_CURLCURL && ai != FEM2D_INTEGRAL_L2INNER) ||
      (bi != FEM2D_INTEGRAL_CURLCURL && bi != FEM2D_INTEGRAL_L2INNER)) {
    set_error("unknown integral", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if (precision != FEM2D_PRECISION_FP64) {
    set_error("only FEM2D_PRECISION_FP64 is provided", 0);
    return FEM2D_ERR_UNSUPPORTED;
  }
  return FEM2D_OK;
}

int err_to_status(int bits) {
  if (bits & kErrView) {
    set_error("domain view violates the Domain invariants (dof ids, parent links or basis-spec layout)", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  if (bits & kErrContrib) {
    set_error("an entry receives more than two contributions or a DoF has more than two basis specs", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  if (bits & kErrDepth) {
    set_error("refinement deeper than this build supports (24 levels, 12 halvings per axis)", 0);
    return FEM2D_ERR_UNSUPPORTED;
  }
  set_error("a row group exceeds this build's limits (256 nested Elems, 1024 memberships or rows)", 0);
  return FEM2D_ERR_UNSUPPORTED;
}

}  // namespace

extern "C" {

int fem2d_ctx_create(int device, fem2d_ctx** out) {
  if (!out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  *out = nullptr;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    set_error("no such HIP device", 0);
    return FEM2D_ERR_HIP;
  }
  HIP_TRY(hipSetDevice(device));
  auto* ctx = new fem2d_ctx();
  ctx->device = device;
  HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  for (auto& e : ctx->ev) HIP_TRY(hipEventCreate(&e));
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&ctx->h_pinned), 64, hipHostMallocDefault));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  ctx->max_grid = prop.multiProcessorCount * 8;
  *out = ctx;
  return FEM2D_OK;
}

void fem2d_ctx_destroy(fem2d_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& b : ctx->bufs)
    if (b.p) (void)hipFree(b.p);
  for (DevBuf* b : {&ctx->scan_tmp, &ctx->out_col, &ctx->out_a, &ctx->out_b})
    if (b->p) (void)hipFree(b->p);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

void* fem2d_ctx_stream(fem2d_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int fem2d_ctx_synchronize(fem2d_ctx* ctx) {
  if (!ctx) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipDeviceSynchronize());
  return FEM2D_OK;
}

int fem2d_ctx_last_timing(fem2d_ctx* ctx, fem2d_timing* out) {
  if (!ctx || !out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  *out = ctx->timing;
  return FEM2D_OK;
}

int fem2d_device_domain_upload(fem2d_ctx* ctx, const fem2d_domain_view* v, fem2d_device_domain** out) {
  if (!ctx || !v || !out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  *out = nullptr;
  if (v->n_elems < 0 || v->n_dofs < 0 || v->n_dofs > 0x7fffffffLL) {
    set_error("domain view sizes out of range", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  if (v->n_elems > 0 && (!v->parent || !v->loc || !v->element_box || !v->materials || !v->orders || !v->bs_offsets)) {
    set_error("domain view has null arrays", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  const int ne = v->n_elems;
  const long long nbs = ne > 0 ? v->bs_offsets[ne] : 0;
  if (nbs > 0 && (!v->bs_dir || !v->bs_i || !v->bs_j || !v->bs_dof)) {
    set_error("domain view has null basis-spec arrays", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  auto* dd = new fem2d_device_domain();
  dd->device = ctx->device;
  dd->cc = v->cc;
  dd->n_elems = ne;
  dd->n_dofs = v->n_dofs;
  dd->n_bs = nbs;
  // host metadata: maximum orders (Mesh::max_expansion_orders, mesh.rs:626-630)
  // and the deepest dyadic subdivision per axis (sizes the Gram library)
  for (int e = 0; e < ne; ++e) {
    dd->max_i = std::max<int>(dd->max_i, v->orders[2 * e]);
    dd->max_j = std::max<int>(dd->max_j, v->orders[2 * e + 1]);
    int du = 0, dv = 0, steps = 0;
    for (int x = e; x >= 0 && v->parent[x] >= 0 && v->parent[x] < ne && steps <= kMaxDepth;
         x = v->parent[x], ++steps) {
      const int l = v->loc[x];
      if (l <= FEM2D_LOC_E) ++du;
      if (l <= FEM2D_LOC_NE || l >= FEM2D_LOC_S) ++dv;
    }
    dd->max_du = std::max(dd->max_du, du);
    dd->max_dv = std::max(dd->max_dv, dv);
  }
  int rc = FEM2D_OK;
  std::vector<long long> off(v->bs_offsets, v->bs_offsets + ne + 1);
  if ((rc = upload(v->parent, size_t(ne), &dd->parent)) || (rc = upload(v->loc, size_t(ne), &dd->loc)) ||
      (rc = upload(v->element_box, size_t(4 * ne), &dd->box)) ||
      (rc = upload(v->materials, size_t(2 * ne), &dd->mat)) ||
      (rc = upload(off.data(), size_t(ne + 1), &dd->bs_off)) ||
      (rc = upload(v->bs_dir, size_t(nbs), &dd->bs_dir)) || (rc = upload(v->bs_i, size_t(nbs), &dd->bs_i)) ||
      (rc = upload(v->bs_j, size_t(nbs), &dd->bs_j)) || (rc = upload(v->bs_dof, size_t(nbs), &dd->bs_dof))) {
    fem2d_device_domain_free(dd);
    return rc;
  }
  *out = dd;
  return FEM2D_OK;
}

void fem2d_device_domain_free(fem2d_device_domain* dd) {
  if (!dd) return;
  (void)hipSetDevice(dd->device);
  for (void* p : {(void*)dd->parent, (void*)dd->loc, (void*)dd->box, (void*)dd->mat, (void*)dd->bs_off,
                  (void*)dd->bs_dir, (void*)dd->bs_i, (void*)dd->bs_j, (void*)dd->bs_dof})
    if (p) (void)hipFree(p);
  delete dd;
}

int fem2d_assemble_device(fem2d_ctx* ctx, const fem2d_device_domain* dd, int basis, int ai, int bi, int glq_u,
                          int glq_v, int precision, fem2d_gep_device* out) {
  if (!ctx || !dd || !out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if (int rc = check_common(dd->cc, dd->n_dofs, glq_u, glq_v, basis, ai, bi, precision)) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  const int ne = dd->n_elems;
  const long long nd = dd->n_dofs;
  const int nu = glq_u != 0 ? glq_u : default_ngq(dd->max_i);
  const int nv = glq_v != 0 ? glq_v : default_ngq(dd->max_j);
  if (nu > kMaxGlq || nv > kMaxGlq || nu < 1 || nv < 1) {
    set_error("more Gauss points per axis than this build supports (64)", 0);
    return FEM2D_ERR_UNSUPPORTED;
  }
  if (dd->max_i > kMaxOrder || dd->max_j > kMaxOrder) {
    set_error("expansion order above MAX_POLYNOMIAL_ORDER", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  if (dd->max_du > kMaxRelDepth || dd->max_dv > kMaxRelDepth) return err_to_status(kErrDepth);

  // work buffers
  Work w{};
  int rc = FEM2D_OK;
  int k = 0;
  double* glq_dev = nullptr;
  if ((rc = slot(ctx, k++, size_t(std::max(ne, 1)), &w.geo)) ||
      (rc = slot(ctx, k++, size_t(std::max(ne, 1)), &w.first_child)) ||
      (rc = slot(ctx, k++, size_t(std::max(ne, 1)), &w.max_child)) ||
      (rc = slot(ctx, k++, size_t(std::max(ne, 1)), &w.n_child)) ||
      (rc = slot(ctx, k++, size_t(std::max(ne, 1)), &w.dyad)) ||
      (rc = slot(ctx, k++, size_t(nd), &w.dof_e1)) || (rc = slot(ctx, k++, size_t(nd), &w.dof_e2)) ||
      (rc = slot(ctx, k++, size_t(nd), &w.dof_k1)) || (rc = slot(ctx, k++, size_t(nd), &w.dof_k2)) ||
      (rc = slot(ctx, k++, size_t(nd), &w.dof_cnt)) || (rc = slot(ctx, k++, size_t(nd), &w.dof_flag)) ||
      (rc = slot(ctx, k++, size_t(nd), &w.dof_grp)) || (rc = slot(ctx, k++, size_t(nd + 1), &w.g_start)) ||
      (rc = slot(ctx, k++, size_t(nd), &w.g_e1)) || (rc = slot(ctx, k++, size_t(nd), &w.g_e2)) ||
      (rc = slot(ctx, k++, size_t(ne + 1), &w.mem_cnt)) || (rc = slot(ctx, k++, size_t(ne + 1), &w.mem_ptr)) ||
      (rc = slot(ctx, k++, size_t(nd + 1), &w.row_len)) || (rc = slot(ctx, k++, size_t(nd + 1), &w.row_ptr)) || (rc = slot(ctx, k++, size_t(4 * kMaxGlq), &glq_dev)) ||
      (rc = slot(ctx, k++, size_t(1), &w.err)))
    return rc;
  w.glq = glq_dev;

  LibShape L;
  L.p1 = std::max(dd->max_i, dd->max_j) + 1;
  L.nr = (1 << (std::max(dd->max_du, dd->max_dv) + 1)) - 1;
  L.nu = nu;
  L.nv = nv;
  L.basis = basis;
  const long long lib_n = 2LL * L.nr * kKinds * L.p1 * L.p1;
  if ((rc = slot(ctx, k++, size_t(lib_n), &w.lib))) return rc;
  const int k_mem = k++;
  const int k_out0 = k;  // spare

  // Gauss-Legendre nodes and weights (host Newton iteration, FP64)
  std::vector<double> glq(4 * kMaxGlq, 0.0);
  fem2d::gauss_legendre_nodes(nu, glq.data(), glq.data() + kMaxGlq);
  fem2d::gauss_legendre_nodes(nv, glq.data() + 2 * kMaxGlq, glq.data() + 3 * kMaxGlq);

  DevView v;
  v.n_elems = ne;
  v.n_dofs = nd;
  v.n_bs = dd->n_bs;
  v.parent = dd->parent;
  v.loc = dd->loc;
  v.box = dd->box;
  v.mat = dd->mat;
  v.bs_off = dd->bs_off;
  v.bs_dir = dd->bs_dir;
  v.bs_i = dd->bs_i;
  v.bs_j = dd->bs_j;
  v.bs_dof = dd->bs_dof;

  hipStream_t st = ctx->stream;
  HIP_TRY(hipEventRecord(ctx->ev[0], st));
  HIP_TRY(hipMemcpyAsync(glq_dev, glq.data(), glq.size() * sizeof(double), hipMemcpyHostToDevice, st));
  const int B = 256;
  k_init<<<grid_for(std::max<long long>(ne, nd), B), B, 0, st>>>(w, ne, nd);
  k_elem_prep<<<grid_for(ne, B), B, 0, st>>>(v, w);
  k_child_check<<<grid_for(ne, B), B, 0, st>>>(v, w);
  k_dof_sig<<<grid_for(ne, B), B, 0, st>>>(v, w);
  k_dof_idx<<<grid_for(ne, B), B, 0, st>>>(v, w);
  k_dof_flags<<<grid_for(nd, B), B, 0, st>>>(w, nd);
  HIP_TRY(hipGetLastError());

  // dof_flag -> inclusive scan -> row-group ids
  size_t tmp_bytes = 0, t2 = 0;
  HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, w.dof_flag, w.dof_grp, int(nd), st));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, w.row_len, w.row_ptr, int(nd + 1), st));
  tmp_bytes = std::max(tmp_bytes, t2);
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, w.mem_cnt, w.mem_ptr, ne + 1, st));
  tmp_bytes = std::max(tmp_bytes, t2);
  if ((rc = ensure(ctx->scan_tmp, tmp_bytes))) return rc;
  HIP_TRY(hipcub::DeviceScan::InclusiveSum(ctx->scan_tmp.p, tmp_bytes, w.dof_flag, w.dof_grp, int(nd), st));
  k_group_starts<<<grid_for(nd, B), B, 0, st>>>(w, nd);
  k_mem_count<<<grid_for(ne, B), B, 0, st>>>(v, w);
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, tmp_bytes, w.mem_cnt, w.mem_ptr, ne + 1, st));
  // total memberships and number of row groups
  HIP_TRY(hipMemcpyAsync(ctx->h_pinned, w.mem_ptr + ne, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(ctx->h_pinned + 1, w.dof_grp + (nd - 1), sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(ctx->h_pinned + 2, w.err, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (ctx->h_pinned[2]) return err_to_status(ctx->h_pinned[2]);
  const int n_mem = ctx->h_pinned[0];
  const int n_groups = ctx->h_pinned[1];
  if ((rc = slot(ctx, k_mem, size_t(2 * std::max(n_mem, 1)), &w.mem_group))) return rc;
  w.mem_bsoff = w.mem_group + std::max(n_mem, 1);
  (void)k_out0;
  k_mem_fill<<<grid_for(ne, B), B, 0, st>>>(v, w);
  k_gram_lib<<<grid_for(lib_n, B), B, 0, st>>>(w, L);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ctx->ev[1], st));

  // symbolic: row lengths per row group, then row_ptr
  const size_t shmem = sizeof(GroupShared);
  const unsigned ggrid = unsigned(std::min<long long>(n_groups, (long long)ctx->max_grid * 4));
  HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_group_count),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)));
  HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_group_fill),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)));
  HIP_TRY(hipEventRecord(ctx->ev[4], st));
  k_group_count<<<ggrid, kBlock, shmem, st>>>(v, w, n_groups);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ctx->ev[5], st));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, tmp_bytes, w.row_len, w.row_ptr, int(nd + 1), st));
  HIP_TRY(hipMemcpyAsync(ctx->h_pinned + 4, w.row_ptr + nd, sizeof(long long), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(ctx->h_pinned + 2, w.err, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipEventRecord(ctx->ev[2], st));
  HIP_TRY(hipStreamSynchronize(st));
  if (ctx->h_pinned[2]) return err_to_status(ctx->h_pinned[2]);
  long long nnz = 0;
  std::memcpy(&nnz, ctx->h_pinned + 4, sizeof(long long));

  // numeric: columns and values
  if ((rc = ensure(ctx->out_col, size_t(nnz) * sizeof(int))) || (rc = ensure(ctx->out_a, size_t(nnz) * 8)) ||
      (rc = ensure(ctx->out_b, size_t(nnz) * 8)))
    return rc;
  HIP_TRY(hipEventRecord(ctx->ev[6], st));
  k_group_fill<<<ggrid, kBlock, shmem, st>>>(v, w, n_groups, L, ai == FEM2D_INTEGRAL_CURLCURL ? 1 : 0,
                                              bi == FEM2D_INTEGRAL_CURLCURL ? 1 : 0,
                                              static_cast<int*>(ctx->out_col.p), static_cast<double*>(ctx->out_a.p),
                                              static_cast<double*>(ctx->out_b.p));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ctx->ev[7], st));
  HIP_TRY(hipMemcpyAsync(ctx->h_pinned + 2, w.err, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipEventRecord(ctx->ev[3], st));
  HIP_TRY(hipStreamSynchronize(st));
  if (ctx->h_pinned[2]) return err_to_status(ctx->h_pinned[2]);

  float t01 = 0, t12 = 0, t23 = 0, t03 = 0, tc = 0, tf = 0;
  (void)hipEventElapsedTime(&tc, ctx->ev[4], ctx->ev[5]);
  (void)hipEventElapsedTime(&tf, ctx->ev[6], ctx->ev[7]);
  ctx->timing.count_kernel_ms = tc;
  ctx->timing.fill_kernel_ms = tf;
  ctx->timing.nnz = nnz;
  (void)hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]);
  (void)hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]);
  (void)hipEventElapsedTime(&t23, ctx->ev[2], ctx->ev[3]);
  (void)hipEventElapsedTime(&t03, ctx->ev[0], ctx->ev[3]);
  ctx->timing.prep_ms = t01;
  ctx->timing.symbolic_ms = t12;
  ctx->timing.fill_ms = t23;
  ctx->timing.total_ms = t03;
  ctx->timing.n_groups = n_groups;

  out->n = nd;
  out->nnz = nnz;
  out->row_ptr = w.row_ptr;
  out->col = static_cast<int32_t*>(ctx->out_col.p);
  out->a_val = static_cast<double*>(ctx->out_a.p);
  out->b_val = static_cast<double*>(ctx->out_b.p);
  return FEM2D_OK;
}

int fem2d_gep_device_download(fem2d_ctx* ctx, const fem2d_gep_device* src, fem2d_gep_csr* out) {
  if (!ctx || !src || !out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  std::memset(out, 0, sizeof(*out));
  out->n = src->n;
  out->nnz = src->nnz;
  out->row_ptr = static_cast<int64_t*>(std::malloc(size_t(src->n + 1) * sizeof(int64_t)));
  out->col = static_cast<int32_t*>(std::malloc(std::max<size_t>(1, size_t(src->nnz)) * sizeof(int32_t)));
  out->a_val = static_cast<double*>(std::malloc(std::max<size_t>(1, size_t(src->nnz)) * sizeof(double)));
  out->b_val = static_cast<double*>(std::malloc(std::max<size_t>(1, size_t(src->nnz)) * sizeof(double)));
  if (!out->row_ptr || !out->col || !out->a_val || !out->b_val) {
    fem2d_gep_csr_free(out);
    set_error("host allocation failed", 0);
    return FEM2D_ERR_OOM;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpyAsync(out->row_ptr, src->row_ptr, size_t(src->n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                         ctx->stream));
  if (src->nnz > 0) {
    HIP_TRY(hipMemcpyAsync(out->col, src->col, size_t(src->nnz) * sizeof(int32_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(out->a_val, src->a_val, size_t(src->nnz) * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(out->b_val, src->b_val, size_t(src->nnz) * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return FEM2D_OK;
}

void fem2d_gep_csr_free(fem2d_gep_csr* g) {
  if (!g) return;
  std::free(g->row_ptr);
  std::free(g->col);
  std::free(g->a_val);
  std::free(g->b_val);
  g->row_ptr = nullptr;
  g->col = nullptr;
  g->a_val = nullptr;
  g->b_val = nullptr;
}

int fem2d_galerkin_sample_gep_hcurl(fem2d_ctx* ctx, const fem2d_domain_view* domain, int basis, int ai, int bi,
                                    int glq_u, int glq_v, int precision, fem2d_gep_csr* out) {
  if (!domain || !out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  // errors in the reference's order before any device work (galerkin.rs:41-59)
  if (int rc = check_common(domain->cc, domain->n_dofs, glq_u, glq_v, basis, ai, bi, precision)) return rc;
  if (!ctx) {
    set_error("null context", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  fem2d_device_domain* dd = nullptr;
  if (int rc = fem2d_device_domain_upload(ctx, domain, &dd)) return rc;
  fem2d_gep_device dev{};
  int rc = fem2d_assemble_device(ctx, dd, basis, ai, bi, glq_u, glq_v, precision, &dev);
  if (rc == FEM2D_OK) rc = fem2d_gep_device_download(ctx, &dev, out);
  fem2d_device_domain_free(dd);
  return rc;
}

}  // extern "C"
