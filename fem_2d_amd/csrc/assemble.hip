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
//    group's Elem(s). That set is a union of whole row groups. A row's columns
//    are the suffix of that set starting at the row itself.
//  * Each stored entry (r, c) receives one contribution per nested pair of
//    basis specs of r and c. Two pairs of interior-disjoint Elems can have at
//    most two nested pairs, so an entry has one or two contributions; they are
//    summed in a fixed order and written once. Kernels check this bound and
//    report a violation instead of writing wrong values.
//  * One wavefront owns one row group: it builds the candidate set in its own
//    slice of LDS with wave-level scans, a bitonic sort and ballots (no block
//    barriers), then streams the group's rows with consecutive lanes on
//    consecutive CSR slots. Row groups whose candidate set exceeds the wave
//    slice go to a block-per-group path with a larger LDS budget.
//  * Pattern (count pass + scan) and values (fill pass) are computed on the
//    device in the same call. The only host synchronisation is at the end (or
//    when the output buffers must grow).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fem2d.h"
#include "host/capi_internal.hpp"
#include "host/parallel.hpp"
#include "fp32r.hpp"
#include "domain_build.hpp"
#include "pinned_staging.hpp"

namespace {

// ---------------------------------------------------------------------------
// Limits of this build (checked; exceeded limits return FEM2D_ERR_UNSUPPORTED)
// ---------------------------------------------------------------------------
constexpr int kWave = 64;
constexpr int kBlock = 256;        // threads per block (4 wavefronts)
constexpr int kWavesPerBlock = kBlock / kWave;
// wave path (one wavefront per row group)
constexpr int kWS = 32;            // nested Elems per row group
constexpr int kWR = 128;           // memberships gathered per row group
constexpr int kWD = 128;           // distinct row groups in one candidate set
constexpr int kWC = 256;           // candidate columns from the group's first row on
constexpr int kWRows = 128;        // rows per row group
constexpr int kCoef = 8;           // doubles per source record: CurlCurl prefactors (4), L2Inner (2), padding
// block path (one block per row group)
constexpr int kMaxSrc = 128;
constexpr int kMaxRaw = 1024;
constexpr int kMaxDist = 1024;
constexpr int kMaxRows = 1024;
// Stored precision of the fill's output (fem2d_precision): the value types of
// A and B and the streaming fill's launch shape. The wavefronts of a block
// share one LDS copy of the Gram library (up to kLibLds values) beside their
// staged source records. FP64 output: 16 wavefronts and up to 96 KB of
// library (orders to 8 over three levels of refinement), one block per CU,
// measured faster than three 8-wavefront blocks with a 32 KB library; FP32
// output (12 B per entry, more latency-bound) keeps the three smaller blocks
// per CU. kChunk: CSR entries per work item (256 for all; 512 made the FP32
// fill faster on Domains of large row groups and slower on Domains of small
// ones, and the FP64 fill slower on all that were measured).
constexpr int kPrecisions = 3;
template <int kPrecision>
struct Store;
template <>
struct Store<FEM2D_PRECISION_FP64> {
  using A = double;
  using B = double;
  static constexpr bool kRes = false;  // 16-bit residuals stored beside the values (FP32R)
  static constexpr int kBlock = 1024;
  static constexpr int kLibLds = 12288;
  static constexpr int kChunk = 256;
};
template <>
struct Store<FEM2D_PRECISION_FP32> {
  using A = float;
  using B = float;
  static constexpr bool kRes = false;
  static constexpr int kBlock = 512;
  static constexpr int kLibLds = 4096;
  static constexpr int kChunk = 256;
};
template <>
struct Store<FEM2D_PRECISION_FP32R> {  // 16 B per entry: the FP64 shape
  using A = float;
  using B = float;
  static constexpr bool kRes = true;
  static constexpr int kBlock = 1024;
  static constexpr int kLibLds = 12288;
  static constexpr int kChunk = 256;
};
// Launch shape of the streaming fill: the precision's own (kSmall 0), or 512
// threads a block with a 4096-value library copy (kSmall 1), which lets three
// blocks share a CU when the Domain's library is that small
template <int P, int kSmall>
struct FillShape {
  static constexpr int kBlock = kSmall ? 512 : Store<P>::kBlock;
  static constexpr int kLibLds = kSmall ? 4096 : Store<P>::kLibLds;
};
// placements of the fill's Gram library: global memory, an LDS copy in the
// precision's shape, an LDS copy in the small shape
enum FillPlace { kPlaceGlobal = 0, kPlaceLds = 1, kPlaceLdsSmall = 2 };
constexpr int kFillWavesPerSimd = 6;
// latency-bound kernels: at most 64 VGPRs, so that 8 wavefronts per SIMD are resident
constexpr int kLatencyWavesPerSimd = 8;
constexpr int kMaxDepth = 24;      // h-refinement depth below a base Elem
constexpr int kMaxRanges = 16383;  // distinct relative sub-ranges per axis (14-bit ids in the block path)
// spec_pos bits: a spec's index in its Elem's list, and flags
constexpr unsigned kPosIndex = (1u << 29) - 1;
constexpr unsigned kPosNextAdj = 1u << 29;  // the list's next spec belongs to this spec's DoF + 1
constexpr unsigned kPosPrevAdj = 1u << 30;  // the list's previous spec belongs to this spec's DoF - 1
constexpr unsigned kPosLast = 1u << 31;     // the last spec of the list
constexpr int kDenseRangeDepth = 12;  // deepest sub-range of the direct range map
constexpr int kMaxOrder = 20;      // MAX_POLYNOMIAL_ORDER (mesh.rs:42)
constexpr int kMaxGlq = 64;        // Gauss points per axis
constexpr int kKinds = 5;          // Gram kinds
static_assert(kMaxOrder < 32 && kMaxOrder * (kMaxOrder + 1) < 512, "packed spec fields");

enum GramKind { kNN = 0, kNTd = 1, kTdN = 2, kTdTd = 3, kTT = 4 };
enum Rel { kLocal = 0, kAnc = 1, kDesc = 2 };

// device error flags
enum ErrBits : int {
  kErrView = 1,     // view violates Domain invariants
  kErrSize = 2,     // a per-group limit of this build was exceeded
  kErrContrib = 4,  // an entry with more than two contributions / a DoF with more than two specs
  kErrDepth = 8,    // refinement deeper than kMaxDepth
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
  const int* dof_bs;  // [2 n_dofs] each DoF's first and second basis spec (flat index), -1 for none
  int p1;  // orders per axis in the Gram library (max expansion order + 1)
};

struct Work {
  ElemGeo* geo;
  int* depth;        // ancestors of each Elem
  int* n_desc;       // descendants of each Elem
  int* desc_cursor;  // fill cursor of each Elem's descendant part
  int* nested_ptr;   // [E+1] exclusive scan of the lengths 1 + depth + n_desc
  int* nested;       // per Elem: its descendants (any order), itself, its ancestors (parent first); X << 2 | rel
  int4* dyad;  // (du, iu, dv, iv): position of the Elem inside its base Elem
  int* bs_elem;       // per basis spec: owning Elem
  unsigned* bs_pack;  // per basis spec, packed by pack_spec
  int4* dof_sig;  // per DoF (e1, t1, e2, t2): its basis specs t1, t2 (flat index) on its first and second
                  // Elem e1 < e2, in Elem order (-1 for none)
  int* dof_grp;  // row group of each DoF, plus one (inclusive scan of the group starts)
  int* g_start;  // [G+1]
  int* g_e1;
  int* g_e2;
  int2* spec_run;  // [n_bs] per basis spec: row group if a membership starts there (else -1), index in the Elem's list
  unsigned* spec_pos;  // [n_bs] index k in the Elem's list | last in the list << 31 | the list's spec k - 1
                       // belongs to this spec's DoF - 1 << 30 | spec k + 1 to its DoF + 1 << 29
  int* mem_idx;   // [n_bs+1] memberships before each basis spec; Elem X's are [mem_idx[bs_off[X]], mem_idx[bs_off[X+1]])
  long long* tile_sum;  // per-tile sums of the prefix sum being computed
  int* mem_group;
  int* mem_bsoff;
  int64_t* row_len;  // [N+1]
  int64_t* row_ptr;  // [N+1]
  double* lib;
  int lib_range_stride;  // library values per range: p1 * p1 * kKinds
  int lib_nr;            // ranges of the x axis (the y axis's start at lib_nr)
  const long long* ranges;  // [nr_u + nr_v] sorted keys depth << 32 | index of the sub-ranges, per axis
  int lib_nr_v;             // ranges of the y axis
  // for Domains at most kDenseRangeDepth halvings deep: each axis's library
  // range by dyadic id 2^d - 1 + i (-1 where absent), the y axis's from
  // range_map_v; null otherwise (the keys are searched)
  const int* range_map;
  int range_map_v;
  int* err;
  int* n_large;       // row groups handled by the block path
  int* large_list;
  unsigned char* is_large;
  // per row group, fixed-stride records written by the describe kernel and
  // streamed by the fill kernel (groups of the block path have T = -1)
  int4* meta_hdr;      // [G] T (first row's length), rows, sources, first row
  double* meta_coef;   // [G * kWS * kCoef] per source: CurlCurl (4) and L2Inner (2) prefactors, then
                       // int4 {x range base, y range base, k | rel << 1, 0} of the library
  uint4* meta_col;     // [G * kWC] per candidate column t: column, packed spec of the first and second
                       // contribution (~0 if none), their sources (bits 0-7 and 8-15)
  uint2* meta_row;     // [N] per DoF: packed specs on its row group's first / second Elem (written
                       // by the DoF-group scan; a group's rows start at its header's w)
  int meta_cap;        // row groups the records have room for
  int64_t* seg_start;  // [G+1] first CSR entry of each row group
  int* chunk_first;    // [chunk_cap] row group holding the first entry of each chunk
  long long chunk_cap;
  int chunk;           // CSR entries per chunk (the stored precision's Store::kChunk)
  long long n_dofs;
};

// Gauss-Legendre nodes and weights per axis, passed by value to the library kernel
struct GlqArgs {
  double u_pts[kMaxGlq], u_wts[kMaxGlq], v_pts[kMaxGlq], v_wts[kMaxGlq];
};

struct LibShape {
  int p1;      // orders 0..p1-1 per axis
  int tab;     // p1 * p1: values per kind and range
  int range_stride;  // kKinds * tab: values per range
  int nr;      // ranges of the x axis
  int nr_v;    // ranges of the y axis
  int nu, nv;  // Gauss points per axis
  int basis;
};

struct FillOut {
  int* col;
  void* aval;  // Store<precision>::A values
  void* bval;  // Store<precision>::B values
  unsigned* res;  // FP32R: the two 16-bit residuals of each entry (fp32r_pack)
  long long capacity;  // entries the output buffers hold
  int a_is_cc, b_is_cc;
  int precision;  // fem2d_precision
};

__device__ __forceinline__ void set_err(int* err, int bits) { atomicOr(err, bits); }

__device__ __forceinline__ int n_groups_of(const Work& w) { return w.n_dofs > 0 ? w.dof_grp[w.n_dofs - 1] : 0; }

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
__device__ __noinline__ void kol(int n, double x, double& N, double& Nd, double& T, double& Td) {
  // L_k and L'_k are produced by the recurrence and folded into the T sum in
  // increasing k (no per-thread arrays)
  const bool bubble = n >= 2;
  const int den = bubble ? n * (n - 1) / 2 : 1;
  double lk = 1.0, dk = 0.0;  // L_0, L'_0
  double lm1 = 0.0;           // L_{k-1}
  double sv = 0.0, sd = 0.0;
  for (int k = 0; k <= n; ++k) {
    if (k == 1) {
      lm1 = lk;
      lk = x;
      dk = 1.0;
    } else if (k >= 2) {
      const double fk = double(k);
      const double ln = ((2.0 * fk - 1.0) * x * lk - (fk - 1.0) * lm1) / fk;
      const double dn = fk * lk + x * dk;
      lm1 = lk;
      lk = ln;
      dk = dn;
    }
    if (bubble) {
      int num = 0;
      if (k == n) num = den;
      else if ((n - k) % 2 == 0) num = -(2 * k + 1);
      const double wgt = double(num) / double(den);
      sv += wgt * lk;
      sd += wgt * dk;
    }
  }
  N = lk;
  Nd = dk;
  if (n == 0) {
    T = 1.0 - x; Td = -1.0;
    return;
  }
  if (n == 1) {
    T = 1.0 + x; Td = 1.0;
    return;
  }
  const double c = n == 2 ? 0.968246 : n == 3 ? 2.561738 : n == 4 ? 0.838525 : n == 5 ? 4.248161
                 : n == 6 ? 0.816397 : n == 7 ? 5.882766 : n == 8 ? 0.808509 : 1.0;
  T = sv * c;
  Td = sd * c;
}

template <int kBasis>
__device__ __forceinline__ double shape_pick(int which, int n, double x) {
  double N, Nd, T, Td;
  if (kBasis == FEM2D_BASIS_KOL) kol(n, x, N, Nd, T, Td);
  else hier_poly(n, x, N, Nd, T, Td);
  switch (which) {
    case 0: return N;
    case 1: return Nd;
    case 2: return T;
    default: return Td;
  }
}

// A dyadic sub-range of [-1, 1]: index i of the 2^d equal parts at depth d,
// as the key d << 32 | i. The Gram library holds the sub-ranges that occur in
// the Domain (the position of each Elem inside each of its ancestors, and the
// whole range), sorted by key per axis.
__host__ __device__ __forceinline__ long long range_key(int d, int i) { return ((long long)d << 32) | (unsigned)i; }

// sub-range key -> (scale, offset) of the mapped Gauss points
__device__ __forceinline__ void range_of(long long key, double& s, double& c) {
  const int d = int(key >> 32);
  const int i = int(key & 0xffffffffLL);
  s = ldexp(1.0, -d);
  const double lo = -1.0 + double(i) * ldexp(1.0, 1 - d);
  c = lo + s;
}

// library range of Q's position inside its ancestor (or itself) P on one axis:
// binary search of the axis's sorted keys; -1 if absent (an invalid view)
__device__ __forceinline__ int rel_range_id(const Work& w, int4 P, int4 Q, int axis) {
  const int dp = axis == 0 ? P.x : P.z;
  const int ip = axis == 0 ? P.y : P.w;
  const int dq = axis == 0 ? Q.x : Q.z;
  const int iq = axis == 0 ? Q.y : Q.w;
  const int dd = dq - dp;
  if (w.range_map) {
    const int rid = (1 << dd) - 1 + (iq - (ip << dd));
    const int size = axis == 0 ? w.range_map_v : (1 << (kDenseRangeDepth + 1)) - 1;
    return rid >= 0 && rid < size ? w.range_map[(axis == 0 ? 0 : w.range_map_v) + rid] : -1;
  }
  const long long key = range_key(dd, iq - (ip << dd));
  const long long* keys = w.ranges + (axis == 0 ? 0 : w.lib_nr);
  int lo = 0, hi = axis == 0 ? w.lib_nr : w.lib_nr_v;  // keys[lo] <= key < keys[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] <= key) lo = mid; else hi = mid;
  }
  return keys[lo] == key ? lo : -1;
}

// Block prefactors of one nested pair of Elems: P holds the p-side spec (the
// ancestor, or the Elem itself), Q the q-side spec (CurlCurl: the curl of p is
// scaled by Q's para_scale (1,1), the curl of q by P's sub-range scale, the
// ratio comes from the spec with the larger det_jac; L2Inner: both glq_scales,
// the larger det_jac; see integrals.rs:26-92, 292-354, basis.rs:225-252).
__device__ __forceinline__ void block_coefs(const Work& w, int P, int Q, bool local, double* cA, double* cB,
                                            int& ru, int& rv) {
  const ElemGeo gp = w.geo[P], gq = w.geo[Q];
  double su = 1.0, sv = 1.0;
  ru = 0;
  rv = 0;
  if (!local) {
    const int4 dp = w.dyad[P], dq = w.dyad[Q];
    ru = rel_range_id(w, dp, dq, 0);
    rv = rel_range_id(w, dp, dq, 1);
    if (ru < 0 || rv < 0) {  // a position the upload did not see: the view is inconsistent
      set_err(w.err, kErrView);
      ru = rv = 0;
    }
    su = ldexp(1.0, -(dq.x - dp.x));
    sv = ldexp(1.0, -(dq.z - dp.z));
  }
  const bool p_big = gp.det >= gq.det;
  const double r_uv = p_big ? gp.uv : gq.uv;
  const double r_vu = p_big ? gp.vu : gq.vu;
  const double dmax = p_big ? gp.det : gq.det;
  const double im = gp.inv_mu;
  cA[0] = im * ((gp.ju0 * 1.0) * (gq.ju0 * su)) * r_uv;
  cA[1] = -(im * (gp.ju0 * (gq.jv1 * sv)));
  cA[2] = -(im * (gp.jv1 * (gq.ju0 * su)));
  cA[3] = im * ((gp.jv1 * 1.0) * (gq.jv1 * sv)) * r_vu;
  const double gs = su * sv;
  cB[0] = gp.eps * gs * dmax * gp.ju0 * gq.ju0;
  cB[1] = gp.eps * gs * dmax * gp.jv1 * gq.jv1;
}

// CurlCurl and L2Inner contributions of one nested pair (p, q). Directions are
// U = 0 / V = 1 (W never reaches the device: the view check rejects it). With
// pair = 2 dir_p + dir_q the CurlCurl x-kind is NN, NT', T'N, T'T' = pair and
// the y-kind is 3 - pair; L2Inner is non-zero for UU (x NN, y TT) and VV
// (x TT, y NN) only, and shares one factor with CurlCurl.
// Packed basis spec: dir (1 bit), i (5), j (5), i * p1 (9), j * p1 (9). The
// products are the row offsets of i and j in a p1 x p1 library table, so a
// spec pair addresses its table entry without multiplications.
__host__ __device__ __forceinline__ unsigned pack_spec(unsigned d, unsigned i, unsigned j, int p1) {
  return (d & 1u) | (i << 1) | (j << 6) | ((i * unsigned(p1)) << 11) | ((j * unsigned(p1)) << 20);
}
__device__ __forceinline__ unsigned spec_dir(unsigned x) { return x & 1u; }
__device__ __forceinline__ unsigned spec_i(unsigned x) { return (x >> 1) & 31u; }
__device__ __forceinline__ unsigned spec_j(unsigned x) { return (x >> 6) & 31u; }
__device__ __forceinline__ unsigned spec_ip1(unsigned x) { return (x >> 11) & 511u; }
__device__ __forceinline__ unsigned spec_jp1(unsigned x) { return (x >> 20) & 511u; }

// element `idx` of a uniform base with a 32-bit byte offset (lets the compiler
// use the scalar-base + 32-bit vector-offset addressing of global loads)
// streaming store of the output (written once, never read back by the kernel)
template <class T>
__device__ __forceinline__ void st32(T* __restrict__ base, unsigned idx, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * unsigned(sizeof(T))));
}

// A wave-uniform load of data no kernel in flight writes, through the scalar
// cache: with a uniform index it compiles to a scalar load, which the lgkm
// counter tracks, so it does not wait for the wavefront's earlier vector stores
// the way a vector load does (the vector memory counter retires in issue order)
template <class T>
__device__ __forceinline__ T ld_const(const T* base, long long idx) {
  using ConstT = const __attribute__((address_space(4))) T;
  return ((ConstT*)base)[idx];  // only a C-style cast converts to the constant address space
}
__device__ __forceinline__ int4 ld_const(const int4* base, long long idx) {
  const int* p = reinterpret_cast<const int*>(base + idx);
  return make_int4(ld_const(p, 0), ld_const(p, 1), ld_const(p, 2), ld_const(p, 3));
}

template <class T>
__device__ __forceinline__ T ld32(const T* __restrict__ base, unsigned idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * unsigned(sizeof(T)));
}

// CurlCurl and L2Inner contributions of one nested pair (p, q). bx and by are
// the library's range bases on the x and y axes; values are stored as
// [axis][range][kind][a][b] with a = p's order and b = q's order. With
// pair = 2 dir_p + dir_q the CurlCurl x-kind is NN, NT', T'N, T'T' = pair and
// the y-kind is 3 - pair; L2Inner is non-zero for UU (x NN, y TT) and VV
// (x TT, y NN) only and shares one factor with CurlCurl.
__device__ __forceinline__ void pair_values(const double* __restrict__ lib, unsigned tab, const double* cA,
                                            const double* cB, unsigned bx, unsigned by, unsigned pbs, unsigned qbs,
                                            double& cc, double& l2) {
  const unsigned pd = spec_dir(pbs), qd = spec_dir(qbs);
  const unsigned ix = bx + spec_ip1(pbs) + spec_i(qbs);
  const unsigned iy = by + spec_jp1(pbs) + spec_j(qbs);
  const unsigned pair = 2u * pd + qd;
  const double fx = ld32(lib, ix + __umul24(pair, tab));
  const double fy = ld32(lib, iy + __umul24(3u - pair, tab));
  cc = cA[pair] * fx * fy;
  l2 = 0.0;
  if (pd == qd) {
    if (pd == 0) l2 = cB[0] * fx * ld32(lib, iy + __umul24(unsigned(kTT), tab));
    else l2 = cB[1] * ld32(lib, ix + __umul24(unsigned(kTT), tab)) * fy;
  }
}

// ---------------------------------------------------------------------------
// Prep kernels
// ---------------------------------------------------------------------------

__device__ __forceinline__ void init_body(const Work& w, int n_elems, long long n_dofs, long long t) {
  if (t < n_elems) w.desc_cursor[t] = 0;
  if (t == 0) {
    *w.err = 0;
    *w.n_large = 0;
  }
}

__device__ __forceinline__ double map_range(double val, double in_min, double in_max, double out_min,
                                            double out_max) {
  return (val - in_min) * (out_max - out_min) / (in_max - in_min) + out_min;
}

// Elem geometry: parametric range from the chain of HRefLocs (elem.rs:191-197,
// h_refinement.rs:247-279), Jacobian over that range (element.rs:33-50),
// inverse/determinant (space.rs:138-147) and the Elem's dyadic position.
__device__ __forceinline__ void elem_prep_body(const DevView& v, const Work& w, int e) {
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
  // the chain is valid: count this Elem as a descendant of each ancestor
  w.depth[e] = depth;
  for (int a = v.parent[e], i = 0; i < depth; ++i, a = v.parent[a]) atomicAdd(&w.n_desc[a], 1);
  double u0 = -1.0, u1 = 1.0, v0 = -1.0, v1 = 1.0;
  int du = 0, iu = 0, dv = 0, iv = 0;
  for (int t = depth - 1; t >= 0; --t) {
    const double mu = (u0 + u1) / 2.0, mv = (v0 + v1) / 2.0;
    const int l = locs[t];
    if (l > FEM2D_LOC_N) {
      set_err(w.err, kErrView);
      return;
    }
    const bool split_u = l <= FEM2D_LOC_E;  // SW, SE, NW, NE, W, E
    const bool split_v = l <= FEM2D_LOC_NE || l >= FEM2D_LOC_S;
    const bool hi_u = l == FEM2D_LOC_SE || l == FEM2D_LOC_NE || l == FEM2D_LOC_E;
    const bool hi_v = l == FEM2D_LOC_NW || l == FEM2D_LOC_NE || l == FEM2D_LOC_N;
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

// Each Elem writes itself and its ancestors into its own list (after the
// descendant part) and itself into each ancestor's descendant part.
__device__ __forceinline__ void nested_fill_body(const DevView& v, const Work& w, int e) {
  if (e >= v.n_elems) return;
  const int depth = w.depth[e];
  const int own = w.nested_ptr[e] + w.n_desc[e];
  w.nested[own] = (e << 2) | kLocal;
  for (int a = v.parent[e], i = 0; i < depth; ++i, a = v.parent[a]) {
    w.nested[own + 1 + i] = (a << 2) | kAnc;
    const int slot = atomicAdd(&w.desc_cursor[a], 1);
    w.nested[w.nested_ptr[a] + slot] = (e << 2) | kDesc;
  }
}

// Per basis spec: owning Elem and packed spec (one wavefront per Elem, lanes
// over its spec list); the spec must be one of its DoF's listed specs.
__device__ __forceinline__ void spec_body(const DevView& v, const Work& w, long long wave, long long waves) {
  const int lane = threadIdx.x % kWave;
  for (long long e = wave; e < v.n_elems; e += waves) {
    const long long t0 = v.bs_off[e], t1 = v.bs_off[e + 1];
    for (long long t = t0 + lane; t < t1; t += kWave) {
      w.bs_elem[t] = int(e);
      w.bs_pack[t] = pack_spec(v.bs_dir[t], v.bs_i[t], v.bs_j[t], v.p1);
      const int d = v.bs_dof[t];
      const bool prev_adj = t > t0 && v.bs_dof[t - 1] == d - 1;
      const bool next_adj = t + 1 < t1 && v.bs_dof[t + 1] == d + 1;
      w.spec_pos[t] = unsigned(t - t0) | (t + 1 == t1 ? kPosLast : 0u) | (prev_adj ? kPosPrevAdj : 0u) |
                      (next_adj ? kPosNextAdj : 0u);
      // H(curl) specs are U or V directed with orders inside the Gram library
      if (d < 0 || d >= v.n_dofs || v.bs_dir[t] > FEM2D_DIR_V || v.bs_i[t] >= v.p1 || v.bs_j[t] >= v.p1 ||
          (v.dof_bs[2LL * d] != t && v.dof_bs[2LL * d + 1] != t))
        set_err(w.err, kErrView);
    }
  }
}

// Per DoF: its first listed spec exists, the second is absent or another
// spec, and both name this DoF (with the per-spec check above, the DoFs'
// spec lists and the specs' DoF ids describe the same relation).
__device__ __forceinline__ void dof_check_body(const DevView& v, const Work& w, long long d) {
  if (d >= v.n_dofs) return;
  const int ta = v.dof_bs[2 * d], tb = v.dof_bs[2 * d + 1];
  const bool ok_a = ta >= 0 && ta < v.n_bs && v.bs_dof[ta] == d;
  const bool ok_b = tb == -1 || (tb >= 0 && tb < v.n_bs && tb != ta && v.bs_dof[tb] == d);
  if (!ok_a || !ok_b) set_err(w.err, kErrView);
}

// A DoF's signature: the Elems of its specs in increasing order and the specs
// (flat indices); -1 where absent. Out-of-range spec indices (an invalid view,
// rejected by dof_check_body) read as absent.
__device__ __forceinline__ void dof_sig_body(const DevView& v, const Work& w, long long d) {
  if (d >= v.n_dofs) return;
  int ta = v.dof_bs[2 * d], tb = v.dof_bs[2 * d + 1];
  if (ta < 0 || ta >= v.n_bs) ta = -1;
  if (tb < 0 || tb >= v.n_bs) tb = -1;
  int ea = ta >= 0 ? w.bs_elem[ta] : -1, eb = tb >= 0 ? w.bs_elem[tb] : -1;
  if (ea < 0 || (eb >= 0 && eb < ea)) {
    const int te = ta, ee = ea;
    ta = tb;
    ea = eb;
    tb = te;
    eb = ee;
  }
  w.dof_sig[d] = make_int4(ea, ta, eb, tb);
}

// ---------------------------------------------------------------------------
// Prefix sums with fused producers and consumers (reduce, then scan):
// Op::load(i) yields item i (no side effects: it runs in both kernels), a
// count or a struct whose field n is the count and whose other fields carry
// what the store needs, and Op::store(i, exclusive, item) consumes the
// result. k_scan_reduce writes the sum of every tile of kBlock * R items (R
// rounds, ScanRounds); k_scan_apply derives each tile's prefix from those
// sums and scans the tile in R striped rounds of kBlock consecutive items
// (coalesced loads and stores). No block waits on another.
// ---------------------------------------------------------------------------
constexpr int kScanRounds = 4;  // rounds per tile unless the operation sets its own kRounds

// rounds per tile of an operation: Op::kRounds where declared
template <class Op, class = void>
struct ScanRounds {
  static constexpr int value = kScanRounds;
};
template <class Op>
struct ScanRounds<Op, decltype(void(Op::kRounds))> {
  static constexpr int value = Op::kRounds;
};

__host__ __device__ constexpr long long scan_tiles(long long n, int rounds = kScanRounds) {
  return (n + (long long)kBlock * rounds - 1) / ((long long)kBlock * rounds);
}

__device__ __forceinline__ long long item_count(long long x) { return x; }
template <class I>
__device__ __forceinline__ long long item_count(const I& x) {
  return x.n;
}

// block-wide sum; every thread gets the total (kBlock threads, shared words sw[kWavesPerBlock])
__device__ __forceinline__ long long block_sum(long long x, long long* sw) {
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
  if (lane == 0) sw[wid] = x;
  __syncthreads();
  long long t = 0;
#pragma unroll
  for (int i = 0; i < kWavesPerBlock; ++i) t += sw[i];
  __syncthreads();
  return t;
}

// tile sums, by blocks bid = 0..nblk-1 of the launch's share for this scan
template <class Op>
__device__ __forceinline__ void scan_reduce_body(long long n, const Op& op, long long* tile_sum, long long bid,
                                                 long long nblk) {
  __shared__ long long sw[kWavesPerBlock];
  constexpr int R = ScanRounds<Op>::value;
  const long long n_tiles = scan_tiles(n, R);
  for (long long tile = bid; tile < n_tiles; tile += nblk) {
    long long sum = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const long long i = tile * (long long)(kBlock * R) + (long long)k * kBlock + threadIdx.x;
      if (i < n) sum += item_count(op.load(i));
    }
    sum = block_sum(sum, sw);
    if (threadIdx.x == 0) tile_sum[tile] = sum;
  }
}

// All loads of a tile are issued before any result is needed; the rounds are
// scanned within each wave, and one exchange of wave totals through LDS gives
// every round its offset.
template <class Op>
__device__ __forceinline__ void scan_apply_body(long long n, const Op& op, const long long* tile_sum, long long bid,
                                                long long nblk) {
  constexpr int R = ScanRounds<Op>::value;
  __shared__ long long sw[R][kWavesPerBlock];
  __shared__ long long sp[kWavesPerBlock];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const long long n_tiles = scan_tiles(n, R);
  for (long long tile = bid; tile < n_tiles; tile += nblk) {
    const long long base = tile * (long long)(kBlock * R) + threadIdx.x;
    using Item = decltype(op.load(0));
    Item item[R];
    long long incl[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const long long i = base + (long long)k * kBlock;
      item[k] = i < n ? op.load(i) : Item{};
      incl[k] = item_count(item[k]);
    }
    long long pre = 0;  // sum of the tiles before this one
    for (long long t = threadIdx.x; t < tile; t += kBlock) pre += tile_sum[t];
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) pre += __shfl_xor(pre, d, kWave);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const long long y = __shfl_up(incl[k], d, kWave);
        if (lane >= d) incl[k] += y;
      }
    }
    if (lane == kWave - 1) {
#pragma unroll
      for (int k = 0; k < R; ++k) sw[k][wid] = incl[k];
    }
    if (lane == 0) sp[wid] = pre;
    __syncthreads();
    long long carry = 0;
#pragma unroll
    for (int v = 0; v < kWavesPerBlock; ++v) carry += sp[v];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      long long off = carry;
#pragma unroll
      for (int v = 0; v < kWavesPerBlock; ++v) {
        off += v < wid ? sw[k][v] : 0;
        carry += sw[k][v];
      }
      const long long i = base + (long long)k * kBlock;
      if (i < n) op.store(i, off + incl[k] - item_count(item[k]), item[k]);
    }
    __syncthreads();  // sw and sp are rewritten by the next tile
  }
}

// DoF signatures -> row groups. Item d is 1 where a row group starts (its
// pair of Elems differs from DoF d - 1's); the inclusive sum is the group id.
// The store also marks, for each spec of DoF d, whether a membership starts
// there: a membership is a run of an Elem's spec list that is one whole row
// group (it starts at the group's first DoF and ends at its last).
struct DofGroupScan {
  static constexpr int kRounds = 1;  // the items are large: one per thread keeps 8 wavefronts per SIMD
  DevView v;
  Work w;
  struct Item {
    long long n;  // 1 where a row group starts
    int4 sig;     // the DoF's (e1, t1, e2, t2)
    bool last;    // the next DoF starts another group (or there is none)
  };
  __device__ Item load(long long d) const {
    Item it;
    it.sig = w.dof_sig[d];
    it.n = 1;
    if (d > 0) {
      const int4 p = w.dof_sig[d - 1];
      it.n = (p.x != it.sig.x || p.z != it.sig.z) ? 1 : 0;
    }
    it.last = true;
    if (d + 1 < v.n_dofs) {
      const int4 q = w.dof_sig[d + 1];
      it.last = q.x != it.sig.x || q.z != it.sig.z;
    }
    return it;
  }
  // spec t of DoF d in group g; first: d is the group's first DoF. The run
  // continues from the previous spec of the Elem's list when that spec is DoF
  // d - 1's and d is not a group's first, and on to the next one likewise.
  __device__ void mark_spec(int t, int g, bool first, bool last) const {
    const unsigned pos = w.spec_pos[t];
    const int k = int(pos & kPosIndex);
    const bool start = (pos & kPosPrevAdj) == 0u || first;
    const bool end = (pos & kPosNextAdj) == 0u || last;
    if ((start && !first) || (end && !last)) set_err(w.err, kErrView);  // a run that is not a whole group
    w.spec_run[t] = make_int2(start ? g : -1, k);
  }
  __device__ void store(long long d, long long excl, const Item& it) const {
    const int e1 = it.sig.x, e2 = it.sig.z;
    if (e1 < 0 || (e2 >= 0 && e2 == e1)) set_err(w.err, kErrView);  // no spec, or two on one Elem
    const int g1 = int(excl + it.n);  // group id + 1
    w.dof_grp[d] = g1;
    if (it.n) {
      w.g_start[g1 - 1] = int(d);
      w.g_e1[g1 - 1] = e1;
      w.g_e2[g1 - 1] = e2;
    }
    if (d == w.n_dofs - 1) w.g_start[g1] = int(w.n_dofs);
    // the fill's row record of this DoF: its specs packed
    w.meta_row[d] = make_uint2(it.sig.y >= 0 ? w.bs_pack[it.sig.y] : 0u, it.sig.w >= 0 ? w.bs_pack[it.sig.w] : 0u);
    if (e1 >= 0) mark_spec(it.sig.y, g1 - 1, it.n != 0, it.last);
    if (e2 >= 0) mark_spec(it.sig.w, g1 - 1, it.n != 0, it.last);
  }
};

// Each Elem's basis-spec list as memberships (runs marked by DofGroupScan).
// Item t is 1 at the first spec of a membership; the exclusive sum numbers
// the memberships; item n_bs (0) gives their total.
struct MembershipScan {
  DevView v;
  Work w;
  struct Item {
    long long n;
    int2 run;
  };
  __device__ Item load(long long t) const {
    Item it;
    it.run = t < v.n_bs ? w.spec_run[t] : make_int2(-1, 0);
    it.n = it.run.x >= 0 ? 1 : 0;
    return it;
  }
  __device__ void store(long long t, long long excl, const Item& it) const {
    w.mem_idx[t] = int(excl);
    if (it.n) {
      w.mem_group[excl] = it.run.x;
      w.mem_bsoff[excl] = it.run.y;
    }
  }
};

// nested-list offsets: item e is the length 1 + depth + descendants of Elem e
struct NestedScan {
  Work w;
  int n_elems;
  __device__ long long load(long long e) const { return e < n_elems ? 1 + w.depth[e] + w.n_desc[e] : 0; }
  __device__ void store(long long e, long long excl, long long) const { w.nested_ptr[e] = int(excl); }
};

// CSR layout by row groups. Item g is the length of group g's segment: row q
// of a streaming-fill group holds T - q entries, so the segment is
// nrows T - nrows (nrows - 1) / 2 long; a block-path group sums the row
// lengths its count kernel wrote. The exclusive sum is the segment's first
// entry. The store writes the group id for the chunks of w.chunk entries that
// start inside the segment and, for a block-path group, row_ptr of its rows
// (the streaming fill writes those of the other groups). The item past the
// last group carries nnz and writes the call's status block (nnz, error bits,
// block-path groups, row groups).
struct GroupScan {
  static constexpr int kRounds = 1;  // one group per thread: the store's chunk loop is the long part
  Work w;
  long long n_items;  // row groups the host sized the buffers for, plus one
  long long* status;
  __device__ long long seg_len(int g, int r0, int nrows, int4 hdr) const {
    if (hdr.x > 0) return (long long)nrows * hdr.x - (long long)nrows * (nrows - 1) / 2;
    long long len = 0;
    for (int q = 0; q < nrows; ++q) len += w.row_len[r0 + q];
    return len;
  }
  __device__ long long load(long long g) const {
    if (*w.err || g >= n_groups_of(w)) return 0;
    const int r0 = w.g_start[g];
    return seg_len(int(g), r0, w.g_start[g + 1] - r0, w.meta_hdr[g]);
  }
  __device__ void store(long long g, long long excl, long long len) const {
    w.seg_start[g] = excl;
    const int n_groups = n_groups_of(w);
    if (g == n_items - 1) {
      w.row_ptr[w.n_dofs] = excl;
      status[0] = excl;
      status[1] = *w.err;
      status[2] = *w.n_large;
      status[3] = n_groups;
    }
    if (*w.err || g >= n_groups) return;  // an invalid view: the group records are not trustworthy
    const int4 hdr = w.meta_hdr[g];
    if (hdr.x <= 0) {  // a block-path group (the streaming fill writes the row pointers of the others)
      const int r0 = w.g_start[g], nrows = w.g_start[g + 1] - r0;
      long long at = excl;
      for (int q = 0; q < nrows; ++q) {
        w.row_ptr[r0 + q] = at;
        at += w.row_len[r0 + q];
      }
    }
    const long long C = w.chunk;
    for (long long c = (excl + C - 1) / C; c * C < excl + len && c < w.chunk_cap; ++c) w.chunk_first[c] = int(g);
  }
};

// 1D Gram library: lib[axis][range][kind][a][b] = sum_m w_m f_a(x'_m) g_b(x_m)
// with x'_m = x_m * s + c mapped into the dyadic sub-range (glq.rs:238-249).
// The block's values share few (range, order, point) shape evaluations: with
// tab (dynamic LDS of gram_table_bytes) the block first evaluates each of them
// once (N, T and T' of order n at point m, on its sub-ranges and at the
// unmapped points), then every value is a dot product of two table rows; the
// values are bit for bit those of evaluating the shapes per term (same
// functions, same points, same products in the same order).
template <int kBasis>
__device__ __forceinline__ void gram_body(const Work& w, const LibShape& L, const GlqArgs& glq, long long block,
                                          double* tab) {
  __shared__ double s_pts[2][kMaxGlq], s_wts[2][kMaxGlq];
  for (int i = threadIdx.x; i < kMaxGlq; i += blockDim.x) {
    s_pts[0][i] = glq.u_pts[i];
    s_wts[0][i] = glq.u_wts[i];
    s_pts[1][i] = glq.v_pts[i];
    s_wts[1][i] = glq.v_wts[i];
  }
  __syncthreads();
  // layout [axis][range][kind][a][b] (nr_u ranges on the x axis, nr_v on the
  // y axis); the library holds fewer than 2^31 values
  const unsigned stride = unsigned(kKinds * L.p1 * L.p1);
  const unsigned x_part = unsigned(L.nr) * stride, total = unsigned(L.nr + L.nr_v) * stride;
  if (block >= ((long long)total + blockDim.x - 1) / blockDim.x) return;  // block-uniform
  const unsigned p1 = unsigned(L.p1);
  const unsigned t0 = unsigned(block) * blockDim.x;
  const unsigned R0 = t0 / stride;  // the block's first range (over both axes)
  const int npm = max(L.nu, L.nv);
  // shape index of the table: N, T, T' (N' is not used by the library)
  auto slot = [](int f) { return f == 0 ? 0 : f == 2 ? 1 : 2; };
  if (tab) {
    const unsigned t1 = min(t0 + blockDim.x, total);
    const int nR = int((t1 - 1) / stride - R0) + 1;
    // rows: nR sub-ranges, then the x and y axes' unmapped points
    const int items = (nR + 2) * int(p1) * npm;
    for (int i = threadIdx.x; i < items; i += blockDim.x) {
      const int m = i % npm;
      const int r = i / npm;
      const int n = r % int(p1);
      const int row = r / int(p1);
      int axis;
      double x;
      if (row < nR) {
        const unsigned R = R0 + unsigned(row);
        axis = R >= unsigned(L.nr) ? 1 : 0;
        if (m >= (axis == 0 ? L.nu : L.nv)) continue;
        const long long key = w.ranges[R];
        double sc, c;
        range_of(key, sc, c);
        const double xq = s_pts[axis][m];
        x = key == 0 ? xq : xq * sc + c;
      } else {
        axis = row - nR;
        if (m >= (axis == 0 ? L.nu : L.nv)) continue;
        x = s_pts[axis][m];
      }
      double N, Nd, T, Td;
      if (kBasis == FEM2D_BASIS_KOL) kol(n, x, N, Nd, T, Td);
      else hier_poly(n, x, N, Nd, T, Td);
      double* out = tab + (size_t(row) * 3 * p1 + n) * npm + m;
      out[0] = N;
      out[size_t(p1) * npm] = T;
      out[2 * size_t(p1) * npm] = Td;
    }
    __syncthreads();
  }
  const unsigned t = t0 + threadIdx.x;
  if (t >= total) return;
  const int axis = t >= x_part ? 1 : 0;
  unsigned r = t - (axis ? x_part : 0u);
  const int b = int(r % p1);
  r /= p1;
  const int a = int(r % p1);
  r /= p1;
  const int kind = int(r % unsigned(kKinds));
  const int rid = int(r / unsigned(kKinds));
  const long long key = w.ranges[(axis ? L.nr : 0) + rid];
  double s, c;
  range_of(key, s, c);
  const int np = axis == 0 ? L.nu : L.nv;
  const double* pts = s_pts[axis];
  const double* wts = s_wts[axis];
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
  if (tab) {
    const unsigned R = t / stride;
    const int nR = int((min(t0 + blockDim.x, total) - 1) / stride - R0) + 1;
    const double* P = tab + ((size_t(R - R0) * 3 + slot(fp)) * p1 + a) * npm;
    const double* Q = tab + ((size_t(nR + axis) * 3 + slot(fq)) * p1 + b) * npm;
#pragma unroll 1
    for (int m = 0; m < np; ++m) acc += wts[m] * P[m] * Q[m];
  } else {
#pragma unroll 1
    for (int m = 0; m < np; ++m) {
      const double xq = pts[m];
      const double xp = key == 0 ? xq : xq * s + c;
      acc += wts[m] * shape_pick<kBasis>(fp, a, xp) * shape_pick<kBasis>(fq, b, xq);
    }
  }
  w.lib[t] = acc;
}

// LDS bytes of gram_body's shape table for 256-thread blocks (0: too large,
// the values evaluate their shapes per term)
size_t gram_table_bytes(const LibShape& L) {
  const long long stride = (long long)kKinds * L.p1 * L.p1;
  const long long ranges = (kBlock + stride - 1) / stride + 1;  // most sub-ranges one block meets
  const size_t bytes = size_t(ranges + 2) * 3 * size_t(L.p1) * size_t(std::max(L.nu, L.nv)) * sizeof(double);
  return bytes <= 48 * 1024 ? bytes : 0;
}

// ---------------------------------------------------------------------------
// Preparation launches. Work with no dependence between its parts shares one
// launch, split by block ranges (each launch costs a few microseconds on its
// own, more than most of these parts).
// ---------------------------------------------------------------------------

// The first launch of an assembly, by block ranges of three independent
// parts: the Gram library (a function of the Domain's dyadic sub-ranges, the
// shape-function space and the Gauss points; gram_blocks blocks), the clearing
// of the counters (init_blocks), and the first pass of the nested-list and
// row-group prefix sums (gn and gd blocks, from the Domain alone). One launch
// saves the others' ramps.
__global__ void __launch_bounds__(kBlock) k_gram(DevView v, Work w, LibShape L, GlqArgs glq, int use_table,
                                                 unsigned gram_blocks, unsigned init_blocks, unsigned gn,
                                                 long long* ts_nested, long long* ts_dof) {
  if (blockIdx.x >= gram_blocks + init_blocks) {
    const unsigned b = blockIdx.x - gram_blocks - init_blocks;
    if (b < gn) scan_reduce_body(v.n_elems + 1, NestedScan{w, v.n_elems}, ts_nested, b, gn);
    else scan_reduce_body(v.n_dofs, DofGroupScan{v, w}, ts_dof, b - gn, gridDim.x - gram_blocks - init_blocks - gn);
    return;
  }
  if (blockIdx.x >= gram_blocks) {
    init_body(w, v.n_elems, v.n_dofs, (blockIdx.x - gram_blocks) * (long long)kBlock + threadIdx.x);
    return;
  }
  extern __shared__ double gram_tab[];
  double* tab = use_table ? gram_tab : nullptr;
  // one instance per shape-function space (no evaluation of the other's recurrence)
  if (L.basis == FEM2D_BASIS_KOL) gram_body<FEM2D_BASIS_KOL>(w, L, glq, blockIdx.x, tab);
  else gram_body<FEM2D_BASIS_HIERPOLY>(w, L, glq, blockIdx.x, tab);
}

// per DoF the Elems and specs of its two basis specs (after k_domain_prep)
__global__ void __launch_bounds__(kBlock) k_dof_sig(DevView v, Work w) {
  dof_sig_body(v, w, blockIdx.x * (long long)kBlock + threadIdx.x);
}

// Domain data derived once, when the Domain is uploaded (the device Domain is
// immutable afterwards): per Elem its depth, descendant count, position inside
// its base Elem and geometry; per basis spec its owning Elem, packed spec and
// position in the Elem's list (the reference's BasisSpec carries these); and
// the checks that the view's DoF and spec relations agree.
__global__ void __launch_bounds__(kBlock) k_domain_prep(DevView v, Work w, unsigned prep_blocks,
                                                        unsigned spec_blocks) {
  if (blockIdx.x < prep_blocks) {
    elem_prep_body(v, w, int(blockIdx.x * kBlock + threadIdx.x));
  } else if (blockIdx.x < prep_blocks + spec_blocks) {
    spec_body(v, w, (long long)(blockIdx.x - prep_blocks) * kWavesPerBlock + threadIdx.x / kWave,
              (long long)spec_blocks * kWavesPerBlock);
  } else {
    dof_check_body(v, w, (long long)(blockIdx.x - prep_blocks - spec_blocks) * kBlock + threadIdx.x);
  }
}

// second pass of two independent prefix sums in one launch (the first pass
// shares k_gram's launch)
template <class OpA, class OpB>
__global__ void __launch_bounds__(kBlock) k_scan_apply2(long long na, OpA a, const long long* ts_a, unsigned nba,
                                                        long long nb, OpB b, const long long* ts_b) {
  if (blockIdx.x < nba)
    scan_apply_body(na, a, ts_a, blockIdx.x, nba);
  else
    scan_apply_body(nb, b, ts_b, blockIdx.x - nba, gridDim.x - nba);
}

template <class Op>
__global__ void __launch_bounds__(kBlock) k_scan_reduce(long long n, Op op, long long* tile_sum) {
  scan_reduce_body(n, op, tile_sum, blockIdx.x, gridDim.x);
}

template <class Op>
__global__ void __launch_bounds__(kBlock) k_scan_apply(long long n, Op op, const long long* tile_sum) {
  scan_apply_body(n, op, tile_sum, blockIdx.x, gridDim.x);
}

// the nested lists; the first pass of the membership numbering
__global__ void __launch_bounds__(kBlock) k_nested_members(DevView v, Work w, unsigned fill_blocks, long long n_mem,
                                                           MembershipScan mem, long long* ts_mem) {
  if (blockIdx.x < fill_blocks)
    nested_fill_body(v, w, int(blockIdx.x * kBlock + threadIdx.x));
  else
    scan_reduce_body(n_mem, mem, ts_mem, blockIdx.x - fill_blocks, gridDim.x - fill_blocks);
}

// ---------------------------------------------------------------------------
// Nested-Elem sources of a row group (shared by both paths)
// ---------------------------------------------------------------------------

// The sources of row group g: the Elems nested with its first Elem (k = 0)
// and with its second Elem (k = 1), read from the nested lists. Returns the
// count; thread t < count of the calling group (stride nthreads) stores source t.
// g is uniform over the caller's wavefront: the group's Elems and list bounds
// come through scalar loads.
__device__ __forceinline__ int gather_sources(const Work& w, int g, int cap, int* xs, int* info, int t,
                                              int nthreads) {
  const int e1 = ld_const(w.g_e1, g), e2 = ld_const(w.g_e2, g);
  if (e1 < 0) return 0;  // a DoF without basis specs (flagged as an invalid view)
  const int a1 = ld_const(w.nested_ptr, e1), n1 = ld_const(w.nested_ptr, e1 + 1) - a1;
  const int a2 = e2 >= 0 ? ld_const(w.nested_ptr, e2) : 0;
  const int n2 = e2 >= 0 ? ld_const(w.nested_ptr, e2 + 1) - a2 : 0;
  const int n = n1 + n2;
  if (n > cap) return n;
  for (int s = t; s < n; s += nthreads) {
    const int k = s < n1 ? 0 : 1;
    const int x = w.nested[k == 0 ? a1 + s : a2 + (s - n1)];
    xs[s] = x >> 2;
    info[s] = k | ((x & 3) << 1);
  }
  return n;
}

// ---------------------------------------------------------------------------
// Wave path: one wavefront per row group
// ---------------------------------------------------------------------------

// a wave-uniform value loaded through the vector path, moved to a scalar register
__device__ __forceinline__ long long uniform(long long x) {
  const unsigned long long u = (unsigned long long)x;
  return (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(int(u >> 32)) << 32) |
                     (unsigned)__builtin_amdgcn_readfirstlane(int(u & 0xffffffffu)));
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_incl_scan(int x, int lane) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int y = __shfl_up(x, d, kWave);
    if (lane >= d) x += y;
  }
  return x;
}

// Per-wave LDS of the candidate-set build (both wave kernels)
struct WaveShared {
  int src_x[kWS];
  int src_info[kWS];  // k | rel << 1 | ru << 3 | rv << 17
  int src_bs[kWS];    // first basis spec of each source (bs_offsets)
  int src_m0[kWS];    // first membership of each source
  int raw_base[kWS];  // first raw candidate of each source
  union {
    struct {
      unsigned long long key[kWR];  // row group << 20 | raw index
      int raw_meta[kWR];            // source | bs offset << 8
    } sort;
    uint2 row_pk[kWRows];  // fill: packed specs of the group's rows on its first / second Elem
  } u;
  int dstart[kWD];
  int cum[kWD + 1];
  int dmeta0[kWD];  // source | bs offset << 8 of the first contribution to a candidate group
  int dmeta1[kWD];  // the same for the second contribution, -1 if none
  int n_src, n_dist, total, u_self;
};

// Builds row group g's candidate set in this wave's LDS slice. Returns 0 when
// built and inside the wave path's limits, 1 when the group needs the block
// path, 2 on an invalid domain (error flag set). With coef non-null the
// per-source prefactors are written there (kCoef doubles per source) and the
// source infos gain their dyadic range ids.
__device__ int wave_build(const DevView& v, const Work& w, int g, WaveShared& S, int lane, double* coef) {
  const int n_src = gather_sources(w, g, kWS, S.src_x, S.src_info, lane, kWave);
  if (n_src > kWS) return 1;
  if (lane == 0) {
    S.n_src = n_src;
    S.u_self = -1;
  }
  wave_sync();
  // membership counts -> gather offsets (n_src <= 32)
  int cnt = 0, X = -1;
  if (lane < n_src) {
    X = S.src_x[lane];
    const long long b0 = v.bs_off[X];
    const int m0 = w.mem_idx[b0];
    cnt = w.mem_idx[v.bs_off[X + 1]] - m0;
    S.src_bs[lane] = int(b0);
    S.src_m0[lane] = m0;
  }
  const int incl = wave_incl_scan(cnt, lane);
  const int n_raw = __shfl(incl, kWave - 1, kWave);
  if (n_raw > kWR) return 1;
  if (lane < n_src) S.raw_base[lane] = incl - cnt;
  wave_sync();
  // raw candidates (one membership of one source each) gathered across lanes;
  // the source of item r is the last one whose first item is at most r
  for (int r = lane; r < n_raw; r += kWave) {
    int a = 0, b = n_src;  // raw_base[a] <= r < raw_base[b] (raw_base[n_src] = n_raw)
    while (b - a > 1) {
      const int mid = (a + b) >> 1;
      if (S.raw_base[mid] <= r) a = mid; else b = mid;
    }
    const int m = S.src_m0[a] + (r - S.raw_base[a]);
    S.u.sort.key[r] = ((unsigned long long)(unsigned)w.mem_group[m] << 20) | (unsigned long long)r;
    S.u.sort.raw_meta[r] = a | (w.mem_bsoff[m] << 8);
  }
  if (lane < n_src) {
    if (coef) {
      const int info = S.src_info[lane];
      const int k = info & 1, rel = (info >> 1) & 3;
      const int E = k == 0 ? w.g_e1[g] : w.g_e2[g];
      const int P = rel == kAnc ? X : E;
      const int Q = rel == kAnc ? E : X;
      int ru, rv;
      double cA[4], cB[2];
      block_coefs(w, P, Q, rel == kLocal, cA, cB, ru, rv);
      double* out = coef + lane * kCoef;
      out[0] = cA[0];
      out[1] = cA[1];
      out[2] = cA[2];
      out[3] = cA[3];
      out[4] = cB[0];
      out[5] = cB[1];
      // library range bases of the pair and which of the group's Elems it uses
      int* tail = reinterpret_cast<int*>(out + 6);
      tail[0] = ru * w.lib_range_stride;
      tail[1] = (w.lib_nr + rv) * w.lib_range_stride;
      tail[2] = k | (rel << 1);
      tail[3] = 0;
      S.src_info[lane] = k | (rel << 1) | (ru << 3) | (rv << 17);
    }
  }
  wave_sync();  // every raw key is in LDS
  // rank sort by (row group, gather order): the keys are distinct (they carry
  // the raw index), so key i belongs at #{k : key_k < key_i}. Each lane ranks
  // its two keys against every key (LDS broadcasts), then all write in place.
  static_assert(kWR == 2 * kWave, "two raw candidates per lane");
  unsigned long long mine[2];
  int rank[2] = {0, 0};
  for (int h = 0; h < 2; ++h) mine[h] = lane + h * kWave < n_raw ? S.u.sort.key[lane + h * kWave] : ~0ull;
  for (int k = 0; k < n_raw; ++k) {
    const unsigned long long kk = S.u.sort.key[k];
    rank[0] += kk < mine[0] ? 1 : 0;
    rank[1] += kk < mine[1] ? 1 : 0;
  }
  wave_sync();  // every lane has read every key
  for (int h = 0; h < 2; ++h)
    if (lane + h * kWave < n_raw) S.u.sort.key[rank[h]] = mine[h];
  wave_sync();
  // dedupe: first pass writes every new row group, second the second sources
  int running = 0;
  int bad = 0;
  for (int base = 0; base < n_raw; base += kWave) {
    const int t = base + lane;
    const bool valid = t < n_raw;
    const unsigned long long kt = valid ? S.u.sort.key[t] : 0ull;
    const int h = valid ? int(kt >> 20) : -1;
    const int hp = (valid && t > 0) ? int(S.u.sort.key[t - 1] >> 20) : -2;
    const bool is_new = valid && h != hp;
    const unsigned long long mask = __ballot(is_new);
    const unsigned long long upto = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1ull);
    const int didx = running + __popcll(mask & upto) - 1;
    if (valid && didx >= kWD) bad |= 1;
    if (valid && is_new && didx < kWD) {
      const int gs = w.g_start[h];
      S.dstart[didx] = gs;
      S.cum[didx] = w.g_start[h + 1] - gs;  // size, prefix-summed below
      S.dmeta0[didx] = S.u.sort.raw_meta[int(kt & 0xFFFFF)];
      S.dmeta1[didx] = -1;
    }
    wave_sync();
    if (valid && !is_new && didx < kWD) {
      const int hpp = t >= 2 ? int(S.u.sort.key[t - 2] >> 20) : -3;
      if (hpp == h) bad |= 2;  // a third source for one column
      S.dmeta1[didx] = S.u.sort.raw_meta[int(kt & 0xFFFFF)];
    }
    running += __popcll(mask);
    wave_sync();
  }
  if (__ballot(bad != 0)) {
    if (__ballot((bad & 2) != 0)) {
      if (lane == 0) set_err(w.err, kErrContrib);
      return 2;
    }
    return 1;
  }
  const int n_dist = running;
  // prefix sums of the distinct group sizes; locate the group itself
  int carry = 0;
  const int gs_self = ld_const(w.g_start, g);
  for (int base = 0; base < n_dist; base += kWave) {
    const int t = base + lane;
    const int sz = t < n_dist ? S.cum[t] : 0;
    const int inc = wave_incl_scan(sz, lane);
    if (t < n_dist) {
      S.cum[t] = carry + inc - sz;
      if (S.dstart[t] == gs_self) S.u_self = carry + inc - sz;
    }
    carry += __shfl(inc, kWave - 1, kWave);
  }
  if (lane == 0) {
    S.cum[n_dist] = carry;
    S.n_dist = n_dist;
    S.total = carry;
  }
  wave_sync();
  if (S.u_self < 0) {
    if (lane == 0) set_err(w.err, kErrView);
    return 2;
  }
  if (S.total - S.u_self > kWC || ld_const(w.g_start, g + 1) - gs_self > kWRows) return 1;
  return 0;
}

// One wavefront per row group: build the candidate set, then write the row
// lengths and the group's fixed-stride records for the streaming fill:
// source prefactors and infos, the packed specs of its rows, and for each
// candidate column t = 0..T-1 of its first row (candidate index u_self + t)
// the column and the packed specs of its one or two contributions.
__global__ void __launch_bounds__(kBlock, kLatencyWavesPerSimd) k_describe_wave(DevView v, Work w) {
  if (*w.err) return;  // an invalid view was detected: nothing below may trust it
  __shared__ WaveShared shared[kWavesPerBlock];
  const int lane = threadIdx.x % kWave;
  const int wid = __builtin_amdgcn_readfirstlane(int(threadIdx.x / kWave));
  WaveShared& S = shared[wid];
  const int n_groups = n_groups_of(w);
  // static assignment: a shared work counter would serialise on device-scope
  // atomics, which span the XCDs' separate L2 caches
  const int stride = gridDim.x * kWavesPerBlock;
  for (int g = blockIdx.x * kWavesPerBlock + wid; g < n_groups; g += stride) {
    if (g >= w.meta_cap) {  // more row groups than the view's DoF signatures give
      if (lane == 0) set_err(w.err, kErrView);
      break;
    }
    const long long gs = (long long)g * kWS;
    const int st = wave_build(v, w, g, S, lane, w.meta_coef + gs * kCoef);
    const int r0 = ld_const(w.g_start, g), r1 = ld_const(w.g_start, g + 1);
    const int nrows = r1 - r0;
    if (st == 1) {
      if (lane == 0) {
        w.is_large[g] = 1;
        w.large_list[atomicAdd(w.n_large, 1)] = g;
        w.meta_hdr[g] = make_int4(-1, nrows, 0, 0);
      }
    } else if (st == 0) {
      const int u_self = S.u_self, n_dist = S.n_dist, n_src = S.n_src;  // set by wave_build
      const int T = S.total - u_self;
      if (lane == 0) {
        w.is_large[g] = 0;
        w.meta_hdr[g] = make_int4(T, nrows, n_src, r0);
      }
      for (int t = lane; t < T; t += kWave) {
        const int u = u_self + t;
        int a = 0, b = n_dist;  // cum[a] <= u < cum[b]
        while (b - a > 1) {
          const int mid = (a + b) >> 1;
          if (S.cum[mid] <= u) a = mid; else b = mid;
        }
        const int within = u - S.cum[a];
        const int m0 = S.dmeta0[a], m1 = S.dmeta1[a];
        uint4 rec;
        rec.x = unsigned(S.dstart[a] + within);
        rec.y = w.bs_pack[S.src_bs[m0 & 0xff] + (m0 >> 8) + within];
        rec.z = m1 < 0 ? ~0u : w.bs_pack[S.src_bs[m1 & 0xff] + (m1 >> 8) + within];
        rec.w = unsigned(m0 & 0xff) | (m1 < 0 ? 0u : unsigned(m1 & 0xff) << 8);
        w.meta_col[(long long)g * kWC + t] = rec;
      }
    }
    wave_sync();
  }
}

__device__ __forceinline__ void write_entry(const FillOut& o, long long pos, int c, double cc, double l2) {
  if (pos >= 0 && pos < o.capacity) {
    o.col[pos] = c;
    const double a = o.a_is_cc ? cc : l2, b = o.b_is_cc ? cc : l2;
    switch (o.precision) {
      case FEM2D_PRECISION_FP64:
        static_cast<double*>(o.aval)[pos] = a;
        static_cast<double*>(o.bval)[pos] = b;
        break;
      case FEM2D_PRECISION_FP32:
        static_cast<float*>(o.aval)[pos] = float(a);
        static_cast<float*>(o.bval)[pos] = float(b);
        break;
      default: {  // FEM2D_PRECISION_FP32R
        float ha, hb;
        int qa, qb;
        fem2d::fp32r_split(a, ha, qa);
        fem2d::fp32r_split(b, hb, qb);
        static_cast<float*>(o.aval)[pos] = ha;
        static_cast<float*>(o.bval)[pos] = hb;
        o.res[pos] = fem2d::fp32r_pack(qa, qb);
      }
    }
  }
}

// contribution of one (row spec, column spec) pair through source s of the
// group; the source record (staged in LDS) holds the prefactors, the library
// range bases and which of the group's Elems (k) and which relation (rel) the
// pair has. lib is the Gram library in global memory or its copy in LDS.
__device__ __forceinline__ void flat_contrib(const double* __restrict__ lib, unsigned tab,
                                             const double* __restrict__ coef_g, unsigned s, unsigned cbs, uint2 rp,
                                             double& cc, double& l2) {
  const unsigned ci = s * unsigned(kCoef);
  const int4 tail = reinterpret_cast<const int4*>(coef_g)[ci / 2u + 3u];
  const unsigned rpk = (tail.z & 1) ? rp.y : rp.x;
  const bool anc = (tail.z >> 1) == kAnc;
  // the p side of the reference's pair is the ancestor's spec (selected as
  // whole spec words, so that the field extraction below is branch-free)
  const unsigned ps = anc ? cbs : rpk, qs = anc ? rpk : cbs;
  const unsigned pd = spec_dir(ps), qd = spec_dir(qs);
  const unsigned pair = 2u * pd + qd;
  // table row a = p's order, entry b = q's order
  const unsigned ix = unsigned(tail.x) + spec_ip1(ps) + spec_i(qs);
  const unsigned iy = unsigned(tail.y) + spec_jp1(ps) + spec_j(qs);
  const double fx = ld32(lib, ix + __umul24(pair, tab));
  const double fy = ld32(lib, iy + __umul24(3u - pair, tab));
  cc = coef_g[ci + pair] * fx * fy;
  // L2Inner (UU, VV only): the TT factor replaces the N N factor's partner on
  // the axis where the pair has its T functions; read unconditionally
  const double ftt = ld32(lib, (pd ? ix : iy) + __umul24(unsigned(kTT), tab));
  const double l2v = (coef_g[ci + 4u + pd] * (pd ? ftt : fx)) * (pd ? fy : ftt);
  l2 = pd == qd ? l2v : 0.0;
}

// Entry p of a row group's segment: row q = max{q : off(q) <= p} with
// off(q) = q T - q (q - 1) / 2, candidate column t = q + p - off(q). The float
// root (all operands below 2^24, hardware square root) lands within one row of
// q; one branch-free step each way fixes it, and the clamp keeps q inside the
// group's row records whatever the rounding. Returns q and oq = off(q).
__device__ __forceinline__ void segment_row(int p, int T, int nrows, float bq, int& q, int& oq) {
  q = int((bq - __builtin_amdgcn_sqrtf(fmaxf(bq * bq - 8.0f * float(p), 0.0f))) * 0.5f);
  q = min(max(q, 0), nrows - 1);
  // off(q) = q (2T - q + 1) / 2; both factors stay below 2^24
  oq = int(__umul24(unsigned(q), unsigned(2 * T - q + 1)) >> 1);
  const bool down = oq > p && q > 0;
  q -= down ? 1 : 0;
  oq -= down ? T - q : 0;
  const bool up = q + 1 < nrows && oq + (T - q) <= p;
  oq += up ? T - q : 0;
  q += up ? 1 : 0;
}

// Moves (q, oq) forward to the row of a later entry p (row q has T - q
// entries; q stays below nrows, so the loop ends).
__device__ __forceinline__ void segment_advance(int p, int T, int nrows, int& q, int& oq) {
  while (q + 1 < nrows && oq + (T - q) <= p) {
    oq += T - q;
    ++q;
  }
}

// candidate column of entry p in row q (row start oq), clamped to the group's
// column records
__device__ __forceinline__ int segment_col(int p, int T, int q, int oq) { return min(max(q + (p - oq), q), T - 1); }

// Two entries (rows qa, qb; candidate columns ta, tb) of a row group's segment
// from the group's records (bases are wave-uniform; offsets are 32-bit). The two are computed side by side with
// no branch between them, so that their loads are in flight together; the
// second contribution (both specs of the row and of the column on the same
// two Elems, a few entries in a hundred) follows.
struct EntryRecs {  // the column and row records of two entries
  uint4 ra, rb;
  uint2 rpa, rpb;
};

__device__ __forceinline__ EntryRecs load_recs2(const uint4* __restrict__ col_g, const uint2* __restrict__ row_g,
                                                int qa, int ta, int qb, int tb) {
  EntryRecs r;
  r.ra = ld32(col_g, unsigned(ta));
  r.rb = ld32(col_g, unsigned(tb));
  r.rpa = ld32(row_g, unsigned(qa));
  r.rpb = ld32(row_g, unsigned(qb));
  return r;
}

__device__ __forceinline__ void flat_entry2(const EntryRecs& rec, const double* __restrict__ coef_g,
                                            const double* __restrict__ lib, unsigned tab, int& ca, double& cca,
                                            double& l2a, int& cb, double& ccb, double& l2b) {
  const uint4 ra = rec.ra, rb = rec.rb;
  const uint2 rpa = rec.rpa, rpb = rec.rpb;
  ca = int(ra.x);
  cb = int(rb.x);
  flat_contrib(lib, tab, coef_g, ra.w & 0xffu, ra.y, rpa, cca, l2a);
  flat_contrib(lib, tab, coef_g, rb.w & 0xffu, rb.y, rpb, ccb, l2b);
  if (ra.z != ~0u) {
    double cc1, l21;
    flat_contrib(lib, tab, coef_g, (ra.w >> 8) & 0xffu, ra.z, rpa, cc1, l21);
    cca += cc1;
    l2a += l21;
  }
  if (rb.z != ~0u) {
    double cc1, l21;
    flat_contrib(lib, tab, coef_g, (rb.w >> 8) & 0xffu, rb.z, rpb, cc1, l21);
    ccb += cc1;
    l2b += l21;
  }
}

// Entry j of a chunk in the stored precision P, as three (four for FP32R)
// coalesced, non-temporal stores; FP32R splits each value into its FP32 value
// and 16-bit residual (fp32r.hpp)
template <int P>
__device__ __forceinline__ void store_entry(int* __restrict__ col_out, typename Store<P>::A* __restrict__ a_out,
                                            typename Store<P>::B* __restrict__ b_out, unsigned* __restrict__ r_out,
                                            unsigned j, int c, double a, double b) {
  st32(col_out, j, c);
  if constexpr (Store<P>::kRes) {
    float ha, hb;
    int qa, qb;
    fem2d::fp32r_split(a, ha, qa);
    fem2d::fp32r_split(b, hb, qb);
    st32(a_out, j, ha);
    st32(b_out, j, hb);
    st32(r_out, j, fem2d::fp32r_pack(qa, qb));
  } else {
    st32(a_out, j, typename Store<P>::A(a));
    st32(b_out, j, typename Store<P>::B(b));
  }
}

// Streaming fill: the CSR is cut into chunks of w.chunk consecutive entries
// (the size the layout scan used: Store<P>::kChunk, or the one tuned for the
// Domain); wavefront v takes chunks v, v + waves, ..., so that the wavefronts
// work on neighbouring chunks at any time and share each row group's records
// in L2 (one contiguous range per wavefront, with no such sharing, ran 2.5-4x
// slower). For each row group a chunk meets, the
// wavefront stages the group's source records in LDS and writes the entries
// with consecutive lanes on consecutive entries (coalesced, non-temporal
// stores), reading the group's column and row records and the Gram library (a
// per-block LDS copy when it fits). Entries of block-path groups are left to
// k_fill_large. kACurlCurl / kBCurlCurl: A and B take the CurlCurl values
// (else L2Inner); P: the stored precision (Store<P>; the arithmetic is FP64
// in every precision, values are rounded once at the store).
template <bool kLibInLds, bool kACurlCurl, bool kBCurlCurl, int P, int kSmall>
__global__ void __launch_bounds__((FillShape<P, kSmall>::kBlock), kFillWavesPerSimd)
    k_fill_flat(Work w, LibShape L, FillOut o) {
  using VA = typename Store<P>::A;
  using VB = typename Store<P>::B;
  if (*w.err) return;
  // the source records of the row group being written, per wavefront
  __shared__ double4 src_lds[FillShape<P, kSmall>::kBlock / kWave][kWS * kCoef / 4];
  // the block's copy of the Gram library (when it fits)
  __shared__ double lib_lds[kLibInLds ? FillShape<P, kSmall>::kLibLds : 1];
  const long long nnz = w.row_ptr[w.n_dofs];
  if (nnz > o.capacity) return;  // the host grows the buffers and runs the fill again
  if (kLibInLds) {
    const int n_lib = (L.nr + L.nr_v) * L.range_stride;
    for (int i = threadIdx.x; i < n_lib; i += blockDim.x) lib_lds[i] = w.lib[i];
    __syncthreads();
  }
  const int lane = threadIdx.x % kWave;
  const int wave = __builtin_amdgcn_readfirstlane(int((blockIdx.x * blockDim.x + threadIdx.x) / kWave));
  const int waves = __builtin_amdgcn_readfirstlane(int(gridDim.x) * (FillShape<P, kSmall>::kBlock / kWave));
  const long long chunk = w.chunk;  // the layout scan's chunk size (default or tuned per Domain)
  const long long n_chunks = min((nnz + chunk - 1) / chunk, w.chunk_cap);
  const double* __restrict__ lib = kLibInLds ? lib_lds : w.lib;
  const unsigned tab = unsigned(L.tab);
  // per-group values are wave-uniform: they come through scalar loads and stay
  // in scalar registers (with the record bases). Loading each chunk's first row
  // group one chunk ahead was measured 2-3 % slower.
  for (long long ch = wave; ch < n_chunks; ch += waves) {
    const long long pbase = ch * chunk;
    const int len = int(min(chunk, nnz - pbase));
    const long long pb = uniform(pbase);
    int* __restrict__ col_out = o.col + pb;
    VA* __restrict__ a_out = static_cast<VA*>(o.aval) + pb;
    VB* __restrict__ b_out = static_cast<VB*>(o.bval) + pb;
    unsigned* __restrict__ r_out = Store<P>::kRes ? o.res + pb : nullptr;
    int g = ld_const(w.chunk_first, ch);
    int j0 = 0;
    while (j0 < len) {
      const long long S = ld_const(w.seg_start, g), S1 = ld_const(w.seg_start, g + 1);
      const int j1 = int(min((long long)len, S1 - pbase));
      const int4 hdr = ld_const(w.meta_hdr, g);
      if (hdr.x > 0) {
        const int T = hdr.x, nrows = hdr.y;
        const int off = int(pbase - S);  // segment position of chunk entry 0
        const float bq = 2.0f * float(T) + 1.0f;
        const uint4* col_g = w.meta_col + (size_t)g * kWC;
        const uint2* row_g = w.meta_row + hdr.w;  // the group's rows are DoFs hdr.w, hdr.w + 1, ...
        // two entries per lane and iteration; past the end of the group's
        // part of the chunk the second one repeats the first and is not stored.
        // Each lane finds the row of its first entry from the root and then
        // follows its entries (64 apart) row by row. The rows and columns are
        // clamped into the group's records, so a lane past the part's end
        // loads in bounds.
        int qa, oqa;
        segment_row(j0 + lane + off, T, nrows, bq, qa, oqa);
        int j = j0 + lane;
        bool second = j + kWave < j1;
        int qb = qa, oqb = oqa;
        {
          const int pa = j + off, pb = (second ? j + kWave : j) + off;
          segment_advance(pa, T, nrows, qa, oqa);
          qb = qa;
          oqb = oqa;
          segment_advance(pb, T, nrows, qb, oqb);
        }
        // the first iteration's records are loaded before the source records
        // are staged, so that the two loads are in flight together
        EntryRecs rec = load_recs2(col_g, row_g, qa, segment_col(j + off, T, qa, oqa), qb,
                                   segment_col((second ? j + kWave : j) + off, T, qb, oqb));
        double4* stage = src_lds[threadIdx.x / kWave];
        const double4* src_g = reinterpret_cast<const double4*>(w.meta_coef + (size_t)g * kWS * kCoef);
        for (int i = lane; i < hdr.z * (kCoef / 4); i += kWave) stage[i] = src_g[i];
        wave_sync();
        const double* coef_g = reinterpret_cast<const double*>(stage);
        while (j < j1) {
          int c0, c1;
          double cc0, l20, cc1, l21;
          flat_entry2(rec, coef_g, lib, tab, c0, cc0, l20, c1, cc1, l21);
          store_entry<P>(col_out, a_out, b_out, r_out, unsigned(j), c0, kACurlCurl ? cc0 : l20,
                         kBCurlCurl ? cc0 : l20);
          if (second)
            store_entry<P>(col_out, a_out, b_out, r_out, unsigned(j + kWave), c1, kACurlCurl ? cc1 : l21,
                           kBCurlCurl ? cc1 : l21);
          j += 2 * kWave;
          if (j >= j1) break;
          qa = qb;
          oqa = oqb;
          second = j + kWave < j1;
          const int pa = j + off, pb = (second ? j + kWave : j) + off;
          segment_advance(pa, T, nrows, qa, oqa);
          qb = qa;
          oqb = oqa;
          segment_advance(pb, T, nrows, qb, oqb);
          rec = load_recs2(col_g, row_g, qa, segment_col(pa, T, qa, oqa), qb, segment_col(pb, T, qb, oqb));
        }
        wave_sync();  // every lane is done with the staged records
      }
      j0 = j1;
      ++g;
    }
  }
  // row pointers of the streaming-fill groups' rows, after the entries (the
  // fill reads none of them; written last they overlap other wavefronts'
  // final chunks): row q of group g starts q T - q (q - 1) / 2 entries into
  // the group's segment (the layout scan writes those of block-path groups)
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < w.n_dofs;
       r += (long long)gridDim.x * blockDim.x) {
    const int g = w.dof_grp[r] - 1;
    const int T = w.meta_hdr[g].x;
    if (T > 0) {
      const long long q = r - w.g_start[g];
      w.row_ptr[r] = w.seg_start[g] + q * T - q * (q - 1) / 2;
    }
  }
}

// one instance per library placement, integral pair, stored precision and shape
template <bool kLibInLds, int P, int kSmall>
void launch_fill_lib(bool a_cc, bool b_cc, unsigned grid, hipStream_t st, const Work& w, const LibShape& L,
                     const FillOut& o) {
  constexpr int B = FillShape<P, kSmall>::kBlock;
  if (a_cc && !b_cc) k_fill_flat<kLibInLds, true, false, P, kSmall><<<grid, B, 0, st>>>(w, L, o);
  else if (!a_cc && b_cc) k_fill_flat<kLibInLds, false, true, P, kSmall><<<grid, B, 0, st>>>(w, L, o);
  else if (a_cc) k_fill_flat<kLibInLds, true, true, P, kSmall><<<grid, B, 0, st>>>(w, L, o);
  else k_fill_flat<kLibInLds, false, false, P, kSmall><<<grid, B, 0, st>>>(w, L, o);
}

// the small shape exists for the precisions whose own shape is larger
template <int P>
constexpr bool kHasSmall = Store<P>::kBlock > 512;

template <int P>
void launch_fill_prec(int place, bool a_cc, bool b_cc, unsigned grid, hipStream_t st, const Work& w,
                      const LibShape& L, const FillOut& o) {
  if constexpr (kHasSmall<P>) {
    if (place == kPlaceLdsSmall) return launch_fill_lib<true, P, 1>(a_cc, b_cc, grid, st, w, L, o);
  }
  if (place != kPlaceGlobal) launch_fill_lib<true, P, 0>(a_cc, b_cc, grid, st, w, L, o);
  else launch_fill_lib<false, P, 0>(a_cc, b_cc, grid, st, w, L, o);
}

void launch_fill(int place, bool a_cc, bool b_cc, unsigned grid, hipStream_t st, const Work& w, const LibShape& L,
                 const FillOut& o) {
  switch (o.precision) {
    case FEM2D_PRECISION_FP32: launch_fill_prec<FEM2D_PRECISION_FP32>(place, a_cc, b_cc, grid, st, w, L, o); break;
    case FEM2D_PRECISION_FP32R:
      launch_fill_prec<FEM2D_PRECISION_FP32R>(place, a_cc, b_cc, grid, st, w, L, o);
      break;
    default: launch_fill_prec<FEM2D_PRECISION_FP64>(place, a_cc, b_cc, grid, st, w, L, o);
  }
}

template <int P>
long long lib_lds_cap(int place) {
  if constexpr (kHasSmall<P>) {
    if (place == kPlaceLdsSmall) return FillShape<P, 1>::kLibLds;
  }
  return place == kPlaceLds ? Store<P>::kLibLds : 0;
}

// the library fits the fill's LDS copy of a placement (kPlaceGlobal always
// fits; kPlaceLdsSmall only for the precisions that have the small shape)
bool place_fits(int precision, int place, long long lib_n) {
  if (place == kPlaceGlobal) return true;
  switch (precision) {
    case FEM2D_PRECISION_FP32: return lib_n <= lib_lds_cap<FEM2D_PRECISION_FP32>(place);
    case FEM2D_PRECISION_FP32R: return lib_n <= lib_lds_cap<FEM2D_PRECISION_FP32R>(place);
    default: return lib_n <= lib_lds_cap<FEM2D_PRECISION_FP64>(place);
  }
}

// the placement an assembly uses: the tuned one, else the LDS copy when the
// library fits it
int fill_place(int precision, int tuned, long long lib_n) {
  if (tuned >= 0 && place_fits(precision, tuned, lib_n)) return tuned;
  return place_fits(precision, kPlaceLds, lib_n) ? kPlaceLds : kPlaceGlobal;
}

// ---------------------------------------------------------------------------
// Block path: one block per large row group
// ---------------------------------------------------------------------------

struct GroupShared {
  double cA[kMaxSrc][4];
  double cB[kMaxSrc][2];
  int src_x[kMaxSrc];
  int src_info[kMaxSrc];
  int raw_cnt[kMaxSrc + 1];  // first raw candidate of each source
  int raw_m0[kMaxSrc];       // first membership of each source
  unsigned long long key[kMaxRaw];
  short raw_src[kMaxRaw];
  int raw_off[kMaxRaw];
  int dstart[kMaxDist];
  int cum[kMaxDist + 1];
  short ds0[kMaxDist], ds1[kMaxDist];
  int doff0[kMaxDist], doff1[kMaxDist];
  int row_rel[kMaxRows + 1];
  unsigned int row_bs[2][kMaxRows];
  int wave_sum[kWavesPerBlock];  // block-wide scans
  int n_src, n_raw, n_dist, total, u_self, fail;
};

// Exclusive block-wide prefix sum of one int per thread (kBlock threads);
// *total receives the sum. Two barriers; every thread must call it.
__device__ __forceinline__ int block_excl_scan(int x, int* wave_sum, int* total) {
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const int incl = wave_incl_scan(x, lane);
  if (lane == kWave - 1) wave_sum[wid] = incl;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int i = 0; i < kWavesPerBlock; ++i) {
    before += i < wid ? wave_sum[i] : 0;
    all += wave_sum[i];
  }
  __syncthreads();  // wave_sum is reused by the next scan
  *total = all;
  return before + incl - x;
}

static_assert(sizeof(GroupShared) <= 64 * 1024, "row-group LDS footprint must fit 64 KiB");

// Builds the candidate set of row group g in LDS. Returns false (block-uniform)
// if a limit was exceeded; the error flag is set then.
template <bool kWithCoef>
__device__ bool build_group(const DevView& v, const Work& w, int g, GroupShared& S) {
  const int tid = threadIdx.x;
  const int n = gather_sources(w, g, kMaxSrc, S.src_x, S.src_info, tid, blockDim.x);
  if (n > kMaxSrc) {  // block-uniform
    if (tid == 0) set_err(w.err, kErrSize);
    return false;
  }
  __syncthreads();
  // membership counts of the sources -> their first raw candidates (one source
  // per thread: kMaxSrc <= kBlock)
  static_assert(kMaxSrc <= kBlock, "one source per thread");
  int cnt = 0, m0 = 0;
  if (tid < n) {
    const int X = S.src_x[tid];
    m0 = w.mem_idx[v.bs_off[X]];
    cnt = w.mem_idx[v.bs_off[X + 1]] - m0;
  }
  int n_raw = 0;
  const int first_raw = block_excl_scan(cnt, S.wave_sum, &n_raw);
  if (n_raw > kMaxRaw) {  // block-uniform
    if (tid == 0) set_err(w.err, kErrSize);
    return false;
  }
  if (tid < n) {
    S.raw_cnt[tid] = first_raw;
    S.raw_m0[tid] = m0;
  }
  if (tid == 0) {
    S.raw_cnt[n] = n_raw;
    S.n_src = n;
    S.n_raw = n_raw;
  }
  __syncthreads();
  const int n_src = n;
  // raw candidates (one membership of one source each), one per thread; the
  // source of item r is the last one whose first item is at most r
  for (int r = tid; r < n_raw; r += blockDim.x) {
    int a = 0, b = n_src;  // raw_cnt[a] <= r < raw_cnt[b]
    while (b - a > 1) {
      const int mid = (a + b) >> 1;
      if (S.raw_cnt[mid] <= r) a = mid; else b = mid;
    }
    const int m = S.raw_m0[a] + (r - S.raw_cnt[a]);
    S.key[r] = ((unsigned long long)(unsigned)w.mem_group[m] << 20) | (unsigned long long)r;
    S.raw_src[r] = short(a);
    S.raw_off[r] = w.mem_bsoff[m];
  }
  for (int s = tid; s < n_src; s += blockDim.x) {
    const int X = S.src_x[s];
    if (kWithCoef) {
      const int info = S.src_info[s];
      const int k = info & 1, rel = (info >> 1) & 3;
      const int E = k == 0 ? w.g_e1[g] : w.g_e2[g];
      const int P = rel == kAnc ? X : E;
      const int Q = rel == kAnc ? E : X;
      int ru, rv;
      block_coefs(w, P, Q, rel == kLocal, S.cA[s], S.cB[s], ru, rv);
      S.src_info[s] = k | (rel << 1) | (ru << 3) | (rv << 17);
    }
  }
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
          if ((a > b) == ((i & kk) == 0)) {
            S.key[i] = b;
            S.key[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  // dedupe, all threads: each takes 4 consecutive sorted keys; a key whose row
  // group differs from its predecessor's starts a distinct group (first
  // contribution), a repeat is the group's second contribution (a third is
  // an error). Distinct groups are numbered by a block-wide prefix sum.
  static_assert(kMaxRaw == 4 * kBlock && kMaxDist == 4 * kBlock, "four keys / groups per thread");
  int h4[4], r4[4];
  bool new4[4], valid4[4];
  int n_new = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int t = 4 * tid + k;
    valid4[k] = t < n_raw;
    const unsigned long long kt = valid4[k] ? S.key[t] : 0ull;
    h4[k] = int(kt >> 20);
    r4[k] = int(kt & 0xFFFFF);
    new4[k] = valid4[k] && (t == 0 || int(S.key[t - 1] >> 20) != h4[k]);
    n_new += new4[k] ? 1 : 0;
  }
  int nd = 0;
  const int first = block_excl_scan(n_new, S.wave_sum, &nd);
  if (nd > kMaxDist) {  // block-uniform
    if (tid == 0) set_err(w.err, kErrSize);
    return false;
  }
  {
    int d = first;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (new4[k]) {
        const int gs = w.g_start[h4[k]];
        S.dstart[d] = gs;
        S.cum[d] = w.g_start[h4[k] + 1] - gs;  // the group's size, prefix-summed below
        S.ds0[d] = S.raw_src[r4[k]];
        S.doff0[d] = S.raw_off[r4[k]];
        S.ds1[d] = -1;
        S.doff1[d] = 0;
        ++d;
      }
  }
  if (tid == 0) {
    S.fail = 0;
    S.u_self = -1;
  }
  __syncthreads();
  {
    int d = first;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (new4[k]) {
        ++d;
      } else if (valid4[k]) {
        const int t = 4 * tid + k;
        if (t >= 2 && int(S.key[t - 2] >> 20) == h4[k]) {
          S.fail = 2;  // a third contribution to one candidate group
        } else {
          S.ds1[d - 1] = S.raw_src[r4[k]];
          S.doff1[d - 1] = S.raw_off[r4[k]];
        }
      }
    }
  }
  __syncthreads();
  // the candidate columns before each distinct group; the group itself
  int sz[4], n_cols = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    sz[k] = 4 * tid + k < nd ? S.cum[4 * tid + k] : 0;
    n_cols += sz[k];
  }
  int total = 0;
  int run = block_excl_scan(n_cols, S.wave_sum, &total);
  const int gs_self = w.g_start[g];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int d = 4 * tid + k;
    if (d < nd) {
      S.cum[d] = run;
      if (S.dstart[d] == gs_self) S.u_self = run;
    }
    run += sz[k];
  }
  if (tid == 0) {
    S.cum[nd] = total;
    S.n_dist = nd;
    S.total = total;
  }
  __syncthreads();
  if (tid == 0 && !S.fail && S.u_self < 0) S.fail = 3;  // the group is not among its own candidates
  __syncthreads();
  const int fail = S.fail;
  __syncthreads();
  if (fail) {
    if (tid == 0) set_err(w.err, fail == 2 ? kErrContrib : (fail == 3 ? kErrView : kErrSize));
    return false;
  }
  return true;
}

__global__ void __launch_bounds__(kBlock) k_count_large(DevView v, Work w) {
  if (*w.err) return;  // an invalid view was detected: nothing below may trust it
  extern __shared__ unsigned char smem[];
  GroupShared& S = *reinterpret_cast<GroupShared*>(smem);
  const int n_large = *w.n_large;
  for (int i = blockIdx.x; i < n_large; i += gridDim.x) {
    const int g = w.large_list[i];
    const int r0 = w.g_start[g], r1 = w.g_start[g + 1];
    if (r1 - r0 > kMaxRows) {
      if (threadIdx.x == 0) set_err(w.err, kErrSize);
      continue;
    }
    if (!build_group<false>(v, w, g, S)) continue;
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x)
      w.row_len[r] = (int64_t)(S.total - (S.u_self + (r - r0)));
    __syncthreads();
  }
}

// Entries of the block-path row groups. Everything this kernel reads about
// the CSR layout was written by the layout scan of the same assembly: the
// group's segment [seg_start[g], seg_start[g + 1]) and the row pointers of its
// own rows (the layout scan writes those of block-path groups; the row
// pointers of the other groups are written by the streaming fill, which skips
// that work on an attempt whose output buffers are too small). Every candidate
// index is checked against the group's candidate set and the source Elem's
// spec list before it is used; a violation sets kErrView and the entry is
// skipped.
__global__ void __launch_bounds__(kBlock) k_fill_large(DevView v, Work w, LibShape L, FillOut o) {
  if (*w.err) return;  // an invalid view was detected: nothing below may trust it
  // the host grows the buffers and runs the assembly again (as k_fill_flat)
  if (w.row_ptr[w.n_dofs] > o.capacity) return;
  extern __shared__ unsigned char smem[];
  GroupShared& S = *reinterpret_cast<GroupShared*>(smem);
  const int64_t* __restrict__ row_ptr = w.row_ptr;
  const int n_large = *w.n_large;
  for (int i = blockIdx.x; i < n_large; i += gridDim.x) {
    const int g = w.large_list[i];
    const int r0 = w.g_start[g], r1 = w.g_start[g + 1];
    const int nrows = r1 - r0;
    if (nrows > kMaxRows) continue;
    if (!build_group<true>(v, w, g, S)) continue;
    const long long seg0 = w.seg_start[g];
    const long long seg_len = w.seg_start[g + 1] - seg0;
    const int T = S.total - S.u_self;  // candidate columns of the group's first row
    // the segment holds rows of T, T - 1, ... entries (block-uniform check)
    if (seg_len != (long long)nrows * T - (long long)nrows * (nrows - 1) / 2 || nrows > T) {
      if (threadIdx.x == 0) set_err(w.err, kErrView);
      __syncthreads();
      continue;
    }
    const int e_k[2] = {w.g_e1[g], w.g_e2[g]};
    for (int q = threadIdx.x; q <= nrows; q += blockDim.x) {
      S.row_rel[q] = q < nrows ? int(row_ptr[r0 + q] - seg0) : int(seg_len);
      if (q < nrows) {
        const int r = r0 + q;
        const int4 sg = w.dof_sig[r];
        S.row_bs[0][q] = w.bs_pack[sg.y];
        S.row_bs[1][q] = e_k[1] >= 0 ? w.bs_pack[sg.w] : 0u;
      }
    }
    __syncthreads();
    const int seg_n = int(seg_len);
    const int n_dist = S.n_dist;
    for (int p = threadIdx.x; p < seg_n; p += blockDim.x) {
      int lo = 0, hi = nrows;  // row_rel[lo] <= p < row_rel[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (S.row_rel[mid] <= p) lo = mid; else hi = mid;
      }
      const int u = S.u_self + lo + (p - S.row_rel[lo]);
      if (u < S.u_self || u >= S.total) {  // outside the candidate set: the layout disagrees with it
        set_err(w.err, kErrView);
        continue;
      }
      int a = 0, b = n_dist;
      while (b - a > 1) {
        const int mid = (a + b) >> 1;
        if (S.cum[mid] <= u) a = mid; else b = mid;
      }
      const int within = u - S.cum[a];
      const int c = S.dstart[a] + within;
      double cc_sum = 0.0, l2_sum = 0.0;
      bool ok = true;
      for (int side = 0; side < 2; ++side) {
        const int sidx = side == 0 ? S.ds0[a] : S.ds1[a];
        if (sidx < 0) break;
        const int info = S.src_info[sidx];
        const int k = info & 1, rel = (info >> 1) & 3;
        const int ru = (info >> 3) & 0x3fff, rv = (info >> 17) & 0x3fff;
        const int off = side == 0 ? S.doff0[a] : S.doff1[a];
        const int X = S.src_x[sidx];
        const long long t = v.bs_off[X] + off + within;
        if (t >= v.bs_off[X + 1]) {  // past the source Elem's spec list
          ok = false;
          break;
        }
        const unsigned cbs = w.bs_pack[t];
        const unsigned rbs = S.row_bs[k][lo];
        double cc, l2;
        const unsigned bx = unsigned(ru * L.range_stride), by = unsigned((L.nr + rv) * L.range_stride);
        if (rel == kAnc) pair_values(w.lib, unsigned(L.tab), S.cA[sidx], S.cB[sidx], bx, by, cbs, rbs, cc, l2);
        else pair_values(w.lib, unsigned(L.tab), S.cA[sidx], S.cB[sidx], bx, by, rbs, cbs, cc, l2);
        cc_sum += cc;
        l2_sum += l2;
      }
      if (!ok) {
        set_err(w.err, kErrView);
        continue;
      }
      write_entry(o, seg0 + p, c, cc_sum, l2_sum);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Field evaluation: UniformFieldSpace::xy_fields (fields.rs:63-127)
// ---------------------------------------------------------------------------

// Halves [lo, hi] per axis down a chain of HRefLocs (h_refinement.rs:247-279),
// locs[from - 1] first (the step nearest the top), locs[to] last.
__device__ __forceinline__ void fold_range(const unsigned char* locs, int from, int to, double& u0, double& u1,
                                           double& v0, double& v1) {
  u0 = -1.0; u1 = 1.0; v0 = -1.0; v1 = 1.0;
  for (int t = from - 1; t >= to; --t) {
    const double mu = (u0 + u1) / 2.0, mv = (v0 + v1) / 2.0;
    const int l = locs[t];
    const bool split_u = l <= FEM2D_LOC_E;
    const bool split_v = l <= FEM2D_LOC_NE || l >= FEM2D_LOC_S;
    const bool hi_u = l == FEM2D_LOC_SE || l == FEM2D_LOC_NE || l == FEM2D_LOC_E;
    const bool hi_v = l == FEM2D_LOC_NW || l == FEM2D_LOC_NE || l == FEM2D_LOC_N;
    if (split_u) {
      if (hi_u) u0 = mu; else u1 = mu;
    }
    if (split_v) {
      if (hi_v) v0 = mv; else v1 = mv;
    }
  }
}

// One thread per grid point (leaf k, u point m, v point n) of every leaf:
// the sum over the leaf and its ancestors (leaf first, as
// Mesh::ancestor_elems lists them), over each one's basis specs in list
// order, of f_u or f_v (basis.rs:225-233) at the point mapped into that
// Elem's parametric space (defined_over with the leaf as descendant,
// basis.rs:365-423), times the spec's DoF value. Every product is formed in
// the reference's order: (jac_inv row * shape) * shape, then * solution, with
// jac_inv = [[b, -0], [-0, a]] / (a b) (space.rs:138-147).
template <int kBasis>
__global__ void __launch_bounds__(kBlock) k_xy_fields(DevView v, const int* __restrict__ leaves, long long n_leaves,
                                                      int d, const double* __restrict__ pts,
                                                      const double* __restrict__ sol, double* __restrict__ fx,
                                                      double* __restrict__ fy, int* __restrict__ err) {
  const long long per_leaf = (long long)d * d;
  const long long total = n_leaves * per_leaf;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const int leaf = leaves[idx / per_leaf];
    const int m = int((idx / d) % d), n = int(idx % d);
    unsigned char locs[kMaxDepth];
    int depth = 0;
    bool ok = true;
    for (int x = leaf; v.parent[x] >= 0;) {
      const int p = v.parent[x];
      if (p >= v.n_elems || p == x || depth >= kMaxDepth || v.loc[x] > FEM2D_LOC_N) {
        ok = false;
        break;
      }
      locs[depth++] = v.loc[x];
      x = p;
    }
    double sx = 0.0, sy = 0.0;
    if (!ok) {
      set_err(err, kErrView);
      fx[idx] = sx;
      fy[idx] = sy;
      continue;
    }
    int A = leaf;
    for (int t = 0; t <= depth; ++t) {
      // the Elem t steps above the leaf: its parametric range and Jacobian
      // (element.rs:33-50) ...
      double u0, u1, v0, v1;
      fold_range(locs, depth, t, u0, u1, v0, v1);
      const double* bx = v.box + 4 * A;
      const double a = (map_range(u1, -1.0, 1.0, bx[0], bx[2]) - map_range(u0, -1.0, 1.0, bx[0], bx[2])) / 2.0;
      const double b = (map_range(v1, -1.0, 1.0, bx[1], bx[3]) - map_range(v0, -1.0, 1.0, bx[1], bx[3])) / 2.0;
      const double det = a * b - 0.0 * 0.0;
      const double ju0 = b / det, jz = (-1.0 * 0.0) / det, jv1 = a / det;
      // ... and the leaf's grid point in it (glq.rs:238-249 over the leaf's
      // range relative to this Elem; the leaf itself takes the points as given)
      double x = pts[m], y = pts[n];
      if (t > 0) {
        double r0, r1, q0, q1;
        fold_range(locs, t, 0, r0, r1, q0, q1);
        x = x * ((r1 - r0) / 2.0) + (r1 + r0) / 2.0;
        y = y * ((q1 - q0) / 2.0) + (q1 + q0) / 2.0;
      }
      for (long long k = v.bs_off[A]; k < v.bs_off[A + 1]; ++k) {
        const int dof = v.bs_dof[k];
        if (dof < 0 || dof >= v.n_dofs) {
          set_err(err, kErrView);
          continue;
        }
        const int dir = v.bs_dir[k], i = v.bs_i[k], j = v.bs_j[k];
        const double s = sol[dof];
        double Nu, Ndu, Tu, Tdu, Nv, Ndv, Tv, Tdv;
        if (kBasis == FEM2D_BASIS_KOL) {
          kol(i, x, Nu, Ndu, Tu, Tdu);
          kol(j, y, Nv, Ndv, Tv, Tdv);
        } else {
          hier_poly(i, x, Nu, Ndu, Tu, Tdu);
          hier_poly(j, y, Nv, Ndv, Tv, Tdv);
        }
        if (dir == 0) {  // f_u = jac_inv.u * N_i(x) * T_j(y)
          sx += ((ju0 * Nu) * Tv) * s;
          sy += ((jz * Nu) * Tv) * s;
        } else if (dir == 1) {  // f_v = jac_inv.v * T_i(x) * N_j(y)
          sx += ((jz * Tu) * Nv) * s;
          sy += ((jv1 * Tu) * Nv) * s;
        } else {  // any other direction contributes the zero vector times the DoF value
          sx += 0.0 * s;
          sy += 0.0 * s;
        }
      }
      A = v.parent[A];
    }
    fx[idx] = sx;
    fy[idx] = sy;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// what a recorded assembly graph depends on besides its buffers
struct GraphKey {
  unsigned long long domain = 0;
  int basis = -1, ai = -1, bi = -1, nu = 0, nv = 0;
  long long capacity = -1;
  unsigned long long alloc_generation = 0;
  bool run_large = true;
  int precision = -1;
  int chunk = 0;
  int lib_lds = -1;
  bool operator==(const GraphKey& o) const {
    return domain == o.domain && basis == o.basis && ai == o.ai && bi == o.bi && nu == o.nu && nv == o.nv &&
           capacity == o.capacity && alloc_generation == o.alloc_generation && run_large == o.run_large &&
           precision == o.precision && chunk == o.chunk && lib_lds == o.lib_lds;
  }
};

std::atomic<unsigned long long> g_domain_serial{0};

}  // namespace

struct fem2d_device_domain {
  unsigned long long serial = 0;  // unique per upload (keys recorded graphs)
  int device = 0;
  int cc = 0;
  int n_elems = 0;
  long long n_dofs = 0;
  long long n_bs = 0;
  int max_i = 0, max_j = 0;
  int nr_u = 0, nr_v = 0;      // distinct sub-ranges (Gram library ranges) per axis
  long long* ranges = nullptr;  // [nr_u + nr_v] their keys, sorted per axis (device)
  int* range_map = nullptr;     // direct map by dyadic id for shallow Domains (device), else null
  int range_map_v = 0;          // offset of the y axis's map (the x axis's size)
  long long nested_cap = 0;  // sum over Elems of 1 + 2 * depth: entries of the nested lists
  long long n_groups = 0;    // row groups of the view's DoF signatures (sizes the per-group records)
  // whether any row group needs the block path: a property of the Domain,
  // learned by the first assembly (-1 unknown, 0 none, 1 some)
  mutable int large_state = -1;
  // the streaming fill's chunk size and library placement (1 LDS, 0 global
  // memory) per stored precision, chosen by timing the candidates on this
  // Domain (chunk 0: not tuned yet)
  mutable int tuned_chunk[kPrecisions] = {};
  mutable int tuned_lds[kPrecisions] = {-1, -1, -1};
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
  int* dof_bs = nullptr;
  int* leaves = nullptr;  // Elems without children, ascending (the field grids)
  long long n_leaves = 0;
  // derived on the device at upload (k_domain_prep)
  ElemGeo* geo = nullptr;
  int* depth = nullptr;
  int* n_desc = nullptr;
  int4* dyad = nullptr;
  int* bs_elem = nullptr;
  unsigned* bs_pack = nullptr;
  unsigned* spec_pos = nullptr;
  int4* dof_sig = nullptr;
  // one allocation holding every array above (Domains built on the device), else null
  void* arena = nullptr;
};

struct fem2d_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // the block-path fill's branch of a recorded assembly (beside the streaming
  // fill), forked from and joined back into `stream` by the two events
  hipStream_t side_stream = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  hipEvent_t ev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  std::vector<DevBuf> bufs;  // grow-only work buffers, by slot
  DevBuf out_col, out_a, out_b;
  DevBuf out_r;  // FP32R residuals (grown on the first FP32R assembly that needs them)
  DevBuf field_scratch;  // grid points and status word of the field evaluation
  long long out_capacity = 0;
  long long* h_status = nullptr;  // mapped pinned status block: nnz, err, n_large, n_groups
  void* h_stage = nullptr;        // pinned upload staging of the device Domain builder (grow-only)
  size_t h_stage_bytes = 0;
  long long* d_status = nullptr;  // its device address
  GlqArgs glq{};
  int glq_n[2] = {0, 0};  // point counts glq holds
  fem2d_timing timing{};
  int timing_events = 0;  // event times of the latest assembly not yet read: 1 graph, 2 kernel by kernel
  int n_cu = 0;
  int count_grid = 0;  // resident blocks of the wave kernels on this device
  // resident blocks of the streaming fill per stored precision, with the library in
  // global memory / in LDS
  int fill_grid[kPrecisions][3] = {};  // resident blocks of the fill per precision and FillPlace
  bool use_graphs = true;
  bool autotune = true;  // time the fill's chunk sizes on a Domain's first assembly
  int fill_chunk = 0;     // fill chunk size forced for every assembly (0: tuned or default)
  hipGraphExec_t graph_exec[1] = {nullptr};  // the last recorded assembly
  GraphKey graph_key;
  void* solver = nullptr;  // eigensolver handle, made by the first solve
  // generation of the latest assembly whose outputs the buffers hold (0: none)
  unsigned long long out_generation = 0;
  fem2d_eigen_stats eigen_stats{};  // of the last sparse eigensolve
};

namespace {
// generations of device GEP handles, unique in the process
std::atomic<unsigned long long> g_gep_generation{0};

// Work memory of fem2d_assemble_device_noalloc: one static array per device,
// held until the process ends. Host code never passes `scratch` itself to a
// kernel (in host code the name is a stand-in, not a device address): the
// device address of each device's copy is looked up once, with that device
// current, by fem2d_ctx_create, and only that address is used.
constexpr size_t kScratchBytes = size_t(FEM2D_NOALLOC_SCRATCH_BYTES);
constexpr int kMaxScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// device address of this device's scratch array (the device is current)
int register_scratch(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return FEM2D_OK;  // the no-allocation entry point rejects it
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return FEM2D_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    fem2d::capi::set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e), 0);
    return e == hipErrorOutOfMemory ? FEM2D_ERR_OOM : FEM2D_ERR_HIP;
  }
  g_scratch_ptr[device] = p;
  return FEM2D_OK;
}

void* scratch_of(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}
}  // namespace

int fem2d::capi::ctx_device(fem2d_ctx* ctx) { return ctx->device; }
fem2d_eigen_stats* fem2d::capi::ctx_eigen_stats(fem2d_ctx* ctx) { return &ctx->eigen_stats; }

int fem2d::capi::check_gep_handle(fem2d_ctx* ctx, const fem2d_gep_device* gep) {
  if (gep->generation != 0 && gep->generation != ctx->out_generation) {
    fem2d::capi::set_error(
        "stale device GEP handle: its buffers belong to an earlier assembly (or another context) and may have "
        "been overwritten or freed",
        0);
    return FEM2D_ERR_INVALID_ARG;
  }
  return FEM2D_OK;
}
void** fem2d::capi::ctx_solver_slot(fem2d_ctx* ctx) { return &ctx->solver; }

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

#define HIP_TRY_AT(expr, where)                                                   \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) return hip_fail(_e, (std::string(#expr) + " " + where).c_str()); \
  } while (0)

// context slots of the assembly's row pointers and fill chunk table (after
// the work arrays that bind_work numbers)
constexpr int kRowPtrSlot = 48;
constexpr int kChunkSlot = 49;

// bumped whenever a device buffer is (re)allocated: recorded graphs hold
// buffer addresses and are recorded again after any change
std::atomic<unsigned long long> g_alloc_generation{0};

int ensure(DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (b.bytes >= bytes) return FEM2D_OK;
  ++g_alloc_generation;
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
int alloc_device(size_t count, T** dst) {
  *dst = nullptr;
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(dst), std::max<size_t>(1, count) * sizeof(T)));
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
  if (precision != FEM2D_PRECISION_FP64 && precision != FEM2D_PRECISION_FP32 &&
      precision != FEM2D_PRECISION_FP32R) {
    set_error("unknown precision", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  return FEM2D_OK;
}

int err_to_status(long long bits) {
  if (bits & kErrView) {
    set_error("domain view violates the Domain invariants (dof ids, parent links or basis-spec layout)", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  if (bits & kErrContrib) {
    set_error("an entry receives more than two contributions or a DoF has more than two basis specs", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  if (bits & kErrDepth) {
    set_error("refinement deeper than this build supports (24 levels below a base Elem)", 0);
    return FEM2D_ERR_UNSUPPORTED;
  }
  set_error("a row group exceeds this build's limits (128 nested Elems, 1024 memberships or rows)", 0);
  return FEM2D_ERR_UNSUPPORTED;
}

// ---------------------------------------------------------------------------
// One assembly: its device arrays, launch shapes and kernel sequence, shared
// by fem2d_assemble_device (arrays in the context's grow-only slots) and
// fem2d_assemble_device_noalloc (arrays carved from the static scratch array)
// ---------------------------------------------------------------------------

struct Assembly {
  DevView v{};
  Work w{};
  LibShape L{};
  long long lib_n = 0;
  long long n_grp_items = 0;  // row groups the records are sized for, plus one
  long long tiles_nested = 0, tiles_dof = 0, tiles_mem = 0, tiles_grp = 0;
  long long *ts_nested = nullptr, *ts_dof = nullptr, *ts_mem = nullptr, *ts_grp = nullptr;
  bool run_large = true;  // launch the block path (unless the Domain is known to need none)
  int lib_lds = -1;       // the fill's FillPlace (-1: untuned, the LDS copy when it fits)
};

// Option and size checks after check_common; *nu, *nv receive the Gauss point counts.
int check_assembly(const fem2d_device_domain* dd, int glq_u, int glq_v, int* nu, int* nv) {
  if (dd->n_bs == 0 || dd->n_elems == 0) {
    set_error("domain view has DoFs but no basis specs", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  *nu = glq_u != 0 ? glq_u : default_ngq(dd->max_i);
  *nv = glq_v != 0 ? glq_v : default_ngq(dd->max_j);
  if (*nu > kMaxGlq || *nv > kMaxGlq || *nu < 1 || *nv < 1) {
    set_error("more Gauss points per axis than this build supports (64)", 0);
    return FEM2D_ERR_UNSUPPORTED;
  }
  if (dd->max_i > kMaxOrder || dd->max_j > kMaxOrder) {
    set_error("expansion order above MAX_POLYNOMIAL_ORDER", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  if (dd->nr_u > kMaxRanges || dd->nr_v > kMaxRanges) {
    set_error("more distinct refinement sub-ranges per axis than this build supports (16383)", 0);
    return FEM2D_ERR_UNSUPPORTED;
  }
  return FEM2D_OK;
}

// Binds every work array of the assembly except row_ptr (an output) and
// chunk_first (sized by the output capacity) through bind(k, count, &ptr),
// k numbering the arrays; fills in the view, the library shape and the
// prefix-sum tile layout.
// The device Domain's input arrays as the kernels take them
DevView dev_view(const fem2d_device_domain* dd) {
  DevView v{};
  v.n_elems = dd->n_elems;
  v.n_dofs = dd->n_dofs;
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
  v.dof_bs = dd->dof_bs;
  v.p1 = std::max(dd->max_i, dd->max_j) + 1;
  return v;
}

// k_domain_prep and k_dof_sig into the device Domain's derived arrays (already
// allocated), with the error word at d_err; no synchronisation. With zeroed the
// caller has already cleared depth, n_desc and the error word.
int domain_prep_launch(fem2d_ctx* ctx, fem2d_device_domain* dd, int* d_err, bool zeroed) {
  const int ne = dd->n_elems;
  const size_t E1 = size_t(std::max(ne, 1));
  hipStream_t st = ctx->stream;
  Work w{};
  w.geo = dd->geo;
  w.depth = dd->depth;
  w.n_desc = dd->n_desc;
  w.dyad = dd->dyad;
  w.bs_elem = dd->bs_elem;
  w.bs_pack = dd->bs_pack;
  w.spec_pos = dd->spec_pos;
  w.dof_sig = dd->dof_sig;
  w.err = d_err;
  if (!zeroed) {
    HIP_TRY(hipMemsetAsync(dd->depth, 0, E1 * sizeof(int), st));
    HIP_TRY(hipMemsetAsync(dd->n_desc, 0, E1 * sizeof(int), st));
    HIP_TRY(hipMemsetAsync(w.err, 0, sizeof(int), st));
  }
  const unsigned prep_blocks = grid_for(ne, kBlock);
  const unsigned spec_blocks =
      unsigned(std::min<long long>(std::max<long long>(1, (ne + kWavesPerBlock - 1) / kWavesPerBlock), 1 << 16));
  k_domain_prep<<<prep_blocks + spec_blocks + grid_for(dd->n_dofs, kBlock), kBlock, 0, st>>>(
      dev_view(dd), w, prep_blocks, spec_blocks);
  // the DoF signatures read the specs' owners written above
  k_dof_sig<<<grid_for(dd->n_dofs, kBlock), kBlock, 0, st>>>(dev_view(dd), w);
  HIP_TRY_AT(hipGetLastError(), "after the Domain preparation kernels");
  return FEM2D_OK;
}

// the device Domain's derived arrays, allocated here, then k_domain_prep; the
// view's relation checks surface as the call's status
int domain_prep(fem2d_ctx* ctx, fem2d_device_domain* dd) {
  const size_t E1 = size_t(std::max(dd->n_elems, 1)), B1 = size_t(std::max(dd->n_bs, 1LL));
  if (int rc = alloc_device(E1, &dd->geo)) return rc;
  if (int rc = alloc_device(E1, &dd->depth)) return rc;
  if (int rc = alloc_device(E1, &dd->n_desc)) return rc;
  if (int rc = alloc_device(E1, &dd->dyad)) return rc;
  if (int rc = alloc_device(B1, &dd->bs_elem)) return rc;
  if (int rc = alloc_device(B1, &dd->bs_pack)) return rc;
  if (int rc = alloc_device(B1, &dd->spec_pos)) return rc;
  if (int rc = alloc_device(size_t(std::max(dd->n_dofs, 1LL)), &dd->dof_sig)) return rc;
  int* d_err = nullptr;
  if (int rc = alloc_device(2, &d_err)) return rc;
  struct FreeErr {
    int* p;
    ~FreeErr() { (void)hipFree(p); }
  } free_err{d_err};
  if (int rc = domain_prep_launch(ctx, dd, d_err, false)) return rc;
  int err = 0;
  HIP_TRY(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return err ? err_to_status(err) : FEM2D_OK;
}

template <class Bind>
int bind_work(const fem2d_device_domain* dd, int nu, int nv, int basis, Bind&& bind, Assembly& a) {
  const int ne = dd->n_elems;
  const long long nd = dd->n_dofs;
  const long long nbs = dd->n_bs;
  Work& w = a.w;
  w = Work{};
  w.n_dofs = nd;
  int rc = FEM2D_OK;
  int k = 0;
  const size_t E1 = size_t(std::max(ne, 1)), N1 = size_t(nd + 1), B1 = size_t(nbs + 1);
  // the Domain's derived arrays (read only by the assembly)
  w.geo = dd->geo;
  w.depth = dd->depth;
  w.n_desc = dd->n_desc;
  w.dyad = dd->dyad;
  w.bs_elem = dd->bs_elem;
  w.bs_pack = dd->bs_pack;
  w.spec_pos = dd->spec_pos;
  w.dof_sig = dd->dof_sig;
  if ((rc = bind(k++, E1, &w.desc_cursor)) || (rc = bind(k++, size_t(ne + 1), &w.nested_ptr)) ||
      (rc = bind(k++, size_t(std::max(1LL, dd->nested_cap)), &w.nested)) ||
      (rc = bind(k++, B1, &w.spec_run)) ||
      (rc = bind(k++, N1, &w.dof_grp)) || (rc = bind(k++, N1, &w.g_start)) || (rc = bind(k++, N1, &w.g_e1)) ||
      (rc = bind(k++, N1, &w.g_e2)) || (rc = bind(k++, B1, &w.mem_idx)) || (rc = bind(k++, B1, &w.mem_group)) ||
      (rc = bind(k++, B1, &w.mem_bsoff)) || (rc = bind(k++, N1, &w.row_len)) || (rc = bind(k++, size_t(8), &w.err)) ||
      (rc = bind(k++, N1, &w.large_list)) || (rc = bind(k++, N1, &w.is_large)))
    return rc;
  w.n_large = w.err + 1;
  const size_t G1 = size_t(std::max(1LL, dd->n_groups));
  w.meta_cap = int(G1);
  if ((rc = bind(k++, G1, &w.meta_hdr)) || (rc = bind(k++, G1 * kWS * kCoef, &w.meta_coef)) ||
      (rc = bind(k++, G1 * kWC, &w.meta_col)) || (rc = bind(k++, N1, &w.meta_row)) ||
      (rc = bind(k++, G1 + 1, &w.seg_start)))
    return rc;

  LibShape& L = a.L;
  L.p1 = std::max(dd->max_i, dd->max_j) + 1;
  L.tab = L.p1 * L.p1;
  L.range_stride = kKinds * L.tab;
  L.nr = dd->nr_u;
  L.nr_v = dd->nr_v;
  L.nu = nu;
  L.nv = nv;
  L.basis = basis;
  a.lib_n = (long long)(L.nr + L.nr_v) * kKinds * L.p1 * L.p1;
  w.lib_range_stride = L.range_stride;
  w.lib_nr = L.nr;
  w.lib_nr_v = L.nr_v;
  w.ranges = dd->ranges;
  w.range_map = dd->range_map;
  w.range_map_v = dd->range_map_v;
  if ((rc = bind(k++, size_t(a.lib_n), &w.lib))) return rc;

  // tile sums of the four prefix sums, one region each (two run side by side)
  a.n_grp_items = (long long)G1 + 1;
  a.tiles_nested = scan_tiles(ne + 1);
  a.tiles_dof = scan_tiles(nd, ScanRounds<DofGroupScan>::value);
  a.tiles_mem = scan_tiles(nbs + 1);
  a.tiles_grp = scan_tiles(a.n_grp_items, ScanRounds<GroupScan>::value);
  if ((rc = bind(k++, size_t(a.tiles_nested + a.tiles_dof + a.tiles_mem + a.tiles_grp), &w.tile_sum))) return rc;
  a.ts_nested = w.tile_sum;
  a.ts_dof = a.ts_nested + a.tiles_nested;
  a.ts_mem = a.ts_dof + a.tiles_dof;
  a.ts_grp = a.ts_mem + a.tiles_mem;

  a.v = dev_view(dd);
  a.run_large = dd->large_state != 0;
  return FEM2D_OK;
}

// Gauss-Legendre nodes and weights per axis in the context (a function of the
// point counts only; host arithmetic)
void ensure_glq(fem2d_ctx* ctx, int nu, int nv) {
  if (ctx->glq_n[0] != nu || ctx->glq_n[1] != nv) {
    ctx->glq = GlqArgs{};
    fem2d::gauss_legendre_nodes(nu, ctx->glq.u_pts, ctx->glq.u_wts);
    fem2d::gauss_legendre_nodes(nv, ctx->glq.v_pts, ctx->glq.v_wts);
    ctx->glq_n[0] = nu;
    ctx->glq_n[1] = nv;
  }
}

// Part (a): preparation and symbolic phase, ending with the CSR layout and the
// status block (nnz, error bits, block-path groups, row groups) at status.
// With fine_events, events between the stages time them and the describe
// kernel.
int launch_prep(fem2d_ctx* ctx, hipStream_t st, const Assembly& a, long long* status, bool fine_events) {
  const Work& w = a.w;
  const DevView& v = a.v;
  const int ne = v.n_elems;
  const long long nd = v.n_dofs, nbs = v.n_bs;
  const int B = 256;
  auto scan_grid = [&](long long tiles) { return unsigned(std::max(1LL, std::min(tiles, 8LL * ctx->n_cu))); };
  // the Gram library, the counters' clearing and the first pass of the
  // nested-list offsets and the row groups, side by side
  const size_t gram_lds = gram_table_bytes(a.L);
  const unsigned gram_blocks = grid_for(a.lib_n, B), init_blocks = grid_for(ne, B);
  const unsigned gn = scan_grid(a.tiles_nested), gd = scan_grid(a.tiles_dof), gm = scan_grid(a.tiles_mem);
  k_gram<<<gram_blocks + init_blocks + gn + gd, B, gram_lds, st>>>(v, w, a.L, ctx->glq, gram_lds > 0 ? 1 : 0,
                                                                   gram_blocks, init_blocks, gn, a.ts_nested,
                                                                   a.ts_dof);
  HIP_TRY_AT(hipGetLastError(), "after the first preparation kernel");
  // second pass of both
  k_scan_apply2<<<gn + gd, kBlock, 0, st>>>(ne + 1, NestedScan{w, ne}, a.ts_nested, gn, nd, DofGroupScan{v, w},
                                            a.ts_dof);
  // nested lists beside the first pass of the membership numbering
  const unsigned fill_blocks = grid_for(ne, B);
  k_nested_members<<<fill_blocks + gm, kBlock, 0, st>>>(v, w, fill_blocks, nbs + 1, MembershipScan{v, w}, a.ts_mem);
  k_scan_apply<<<gm, kBlock, 0, st>>>(nbs + 1, MembershipScan{v, w}, a.ts_mem);
  HIP_TRY_AT(hipGetLastError(), "after the nested-list and row-group kernels");
  if (fine_events) HIP_TRY(hipEventRecord(ctx->ev[1], st));
  // symbolic: per-group candidate sets, row lengths and records, then row_ptr
  if (fine_events) HIP_TRY(hipEventRecord(ctx->ev[4], st));
  k_describe_wave<<<unsigned(ctx->count_grid), kBlock, 0, st>>>(v, w);
  if (fine_events) HIP_TRY(hipEventRecord(ctx->ev[5], st));
  const unsigned large_grid = unsigned(std::max(1, ctx->n_cu * 2));
  if (a.run_large) k_count_large<<<large_grid, kBlock, sizeof(GroupShared), st>>>(v, w);
  HIP_TRY_AT(hipGetLastError(), "after the describe kernels");
  // CSR layout and the status block
  const GroupScan gs{w, a.n_grp_items, status};
  k_scan_reduce<<<scan_grid(a.tiles_grp), kBlock, 0, st>>>(a.n_grp_items, gs, a.ts_grp);
  k_scan_apply<<<scan_grid(a.tiles_grp), kBlock, 0, st>>>(a.n_grp_items, gs, a.ts_grp);
  HIP_TRY_AT(hipGetLastError(), "after the CSR layout scan");
  if (fine_events) HIP_TRY(hipEventRecord(ctx->ev[2], st));
  return FEM2D_OK;
}

// Part (b): the streaming fill (columns and values; with too small output
// buffers it writes nothing)
int launch_stream_fill(fem2d_ctx* ctx, hipStream_t st, const Assembly& a, const FillOut& o) {
  const int pi = o.precision;
  const int place = fill_place(o.precision, a.lib_lds, a.lib_n);
  launch_fill(place, o.a_is_cc != 0, o.b_is_cc != 0, unsigned(ctx->fill_grid[pi][place]), st, a.w, a.L, o);
  HIP_TRY_AT(hipGetLastError(), "after the fill kernel");
  return FEM2D_OK;
}

// Part (c): the block-path fill (Domains with large row groups)
int launch_large_fill(fem2d_ctx* ctx, hipStream_t st, const Assembly& a, const FillOut& o) {
  const unsigned large_grid = unsigned(std::max(1, ctx->n_cu * 2));
  k_fill_large<<<large_grid, kBlock, sizeof(GroupShared), st>>>(a.v, a.w, a.L, o);
  HIP_TRY_AT(hipGetLastError(), "after the block-path fill kernel");
  return FEM2D_OK;
}

// resident blocks of the streaming fill of one stored precision (the fill
// assigns its chunks statically: exactly these are launched)
template <int P>
int fill_grids(fem2d_ctx* ctx) {
  int per_cu = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fill_flat<false, true, false, P, 0>,
                                                       Store<P>::kBlock, 0));
  ctx->fill_grid[P][kPlaceGlobal] = std::max(1, per_cu) * ctx->n_cu;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fill_flat<true, true, false, P, 0>,
                                                       Store<P>::kBlock, 0));
  ctx->fill_grid[P][kPlaceLds] = std::max(1, per_cu) * ctx->n_cu;
  ctx->fill_grid[P][kPlaceLdsSmall] = ctx->fill_grid[P][kPlaceLds];
  if constexpr (kHasSmall<P>) {
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fill_flat<true, true, false, P, 1>,
                                                         FillShape<P, 1>::kBlock, 0));
    ctx->fill_grid[P][kPlaceLdsSmall] = std::max(1, per_cu) * ctx->n_cu;
  }
  return FEM2D_OK;
}

FillOut make_fill_out(int precision, int ai, int bi, long long capacity, int* col, void* a_val, void* b_val,
                      unsigned* res) {
  FillOut o;
  o.col = col;
  o.aval = a_val;
  o.bval = b_val;
  o.res = res;
  o.precision = precision;
  o.capacity = capacity;
  o.a_is_cc = ai == FEM2D_INTEGRAL_CURLCURL ? 1 : 0;
  o.b_is_cc = bi == FEM2D_INTEGRAL_CURLCURL ? 1 : 0;
  return o;
}

// chunk sizes the fill's autotuning times on a Domain's first assembly, each
// placement and size timed kTuneReps times (the fastest launch counts)
constexpr int kChunkCandidates[] = {128, 256, 512, 1024};
constexpr int kTuneReps = 5;

// FEM2D_TUNE_VERBOSE=1 prints every tuning trial's time to stderr
bool tune_verbose() {
  static const bool on = [] {
    const char* e = std::getenv("FEM2D_TUNE_VERBOSE");
    return e != nullptr && e[0] != '\0' && e[0] != '0';
  }();
  return on;
}

// CSR entries per fill chunk for a stored precision (before any tuning)
int chunk_of(int precision) {
  switch (precision) {
    case FEM2D_PRECISION_FP32: return Store<FEM2D_PRECISION_FP32>::kChunk;
    case FEM2D_PRECISION_FP32R: return Store<FEM2D_PRECISION_FP32R>::kChunk;
    default: return Store<FEM2D_PRECISION_FP64>::kChunk;
  }
}

// w.chunk for a chunk size (0: the stored precision's default)
void set_chunk(Work& w, int chunk, int precision) { w.chunk = chunk > 0 ? chunk : chunk_of(precision); }

// Host metadata of a device Domain that depends on the Elem tree only: the
// sub-ranges of the Gram library (the position of each Elem inside each of its
// ancestors, per axis: h_refinement.rs:247-279 halves the range per HRefLoc,
// plus the whole range), their direct map for shallow Domains, the size of the
// nested lists and the leaves.
struct TreeMeta {
  long long nested_cap = 0;     // sum over Elems of 1 + 2 * depth
  std::vector<long long> keys;  // range keys, ascending per axis, x axis then y
  int nr_u = 0, nr_v = 0;
  std::vector<int> range_map;   // empty unless every key is at most kDenseRangeDepth deep
  int range_map_v = 0;
  std::vector<int> leaves;      // Elems without children, ascending
};

void tree_meta(int ne, const int* parent, const unsigned char* loc, TreeMeta& m) {
  // keys up to kBitDepth halvings deep are collected in a bitmap indexed by
  // 2^d - 1 + i, whose ascending order is the keys' order; deeper ones are sorted
  constexpr int kBitDepth = 16;
  std::vector<unsigned long long> bits[2];
  std::vector<long long> deep[2];
  for (int ax = 0; ax < 2; ++ax) bits[ax].assign(((size_t(2) << kBitDepth) + 63) / 64, 0);
  auto add = [&](int ax, int d, int i) {
    if (d <= kBitDepth) {
      const size_t b = (size_t(1) << d) - 1 + size_t(i);
      bits[ax][b >> 6] |= 1ull << (b & 63);
    } else {
      deep[ax].push_back(range_key(d, i));
    }
  };
  add(0, 0, 0);
  add(1, 0, 0);
  m.nested_cap = 0;
  for (int e = 0; e < ne; ++e) {
    unsigned char locs[kMaxDepth + 1];
    int steps = 0;
    for (int x = e; x >= 0 && parent[x] >= 0 && parent[x] < ne && steps <= kMaxDepth; x = parent[x], ++steps)
      locs[steps] = loc[x];
    m.nested_cap += 1 + 2LL * steps;
    if (steps > kMaxDepth) continue;  // rejected by the device checks
    // Elem e inside its ancestor t levels up: the steps locs[t-1] (topmost) .. locs[0]
    for (int t = 1; t <= steps; ++t) {
      int du = 0, iu = 0, dv = 0, iv = 0;
      for (int k = t - 1; k >= 0; --k) {
        const int l = locs[k];
        if (l <= FEM2D_LOC_E) {
          iu = 2 * iu + (l == FEM2D_LOC_SE || l == FEM2D_LOC_NE || l == FEM2D_LOC_E ? 1 : 0);
          ++du;
        }
        if (l <= FEM2D_LOC_NE || l >= FEM2D_LOC_S) {
          iv = 2 * iv + (l == FEM2D_LOC_NW || l == FEM2D_LOC_NE || l == FEM2D_LOC_N ? 1 : 0);
          ++dv;
        }
      }
      add(0, du, iu);
      add(1, dv, iv);
    }
  }
  std::vector<long long> keys[2];
  for (int ax = 0; ax < 2; ++ax) {
    for (size_t w = 0; w < bits[ax].size(); ++w)
      for (unsigned long long v = bits[ax][w]; v; v &= v - 1) {
        const size_t b = w * 64 + size_t(__builtin_ctzll(v));
        const int d = 63 - __builtin_clzll((unsigned long long)(b + 1));
        keys[ax].push_back(range_key(d, int(b + 1 - (size_t(1) << d))));
      }
    std::sort(deep[ax].begin(), deep[ax].end());
    deep[ax].erase(std::unique(deep[ax].begin(), deep[ax].end()), deep[ax].end());
    keys[ax].insert(keys[ax].end(), deep[ax].begin(), deep[ax].end());
  }
  m.nr_u = int(std::min<size_t>(keys[0].size(), size_t(1) << 30));
  m.nr_v = int(std::min<size_t>(keys[1].size(), size_t(1) << 30));
  // shallow Domains (every sub-range at most kDenseRangeDepth halvings deep):
  // a direct map by dyadic id, one load instead of a search per lookup
  m.range_map.clear();
  m.range_map_v = 0;
  if (int(keys[0].back() >> 32) <= kDenseRangeDepth && int(keys[1].back() >> 32) <= kDenseRangeDepth) {
    const int su = (1 << (int(keys[0].back() >> 32) + 1)) - 1;
    const int sv = (1 << (kDenseRangeDepth + 1)) - 1;
    m.range_map.assign(size_t(su + sv), -1);
    for (size_t k = 0; k < keys[0].size(); ++k)
      m.range_map[size_t((1 << int(keys[0][k] >> 32)) - 1 + int(keys[0][k] & 0xffffffffLL))] = int(k);
    for (size_t k = 0; k < keys[1].size(); ++k)
      m.range_map[size_t(su + (1 << int(keys[1][k] >> 32)) - 1 + int(keys[1][k] & 0xffffffffLL))] = int(k);
    m.range_map_v = su;
  }
  m.keys = std::move(keys[0]);
  m.keys.insert(m.keys.end(), keys[1].begin(), keys[1].end());
  std::vector<char> has_child(size_t(ne), 0);
  for (int e = 0; e < ne; ++e)
    if (parent[e] >= 0 && parent[e] < ne) has_child[size_t(parent[e])] = 1;
  m.leaves.clear();
  for (int e = 0; e < ne; ++e)
    if (!has_child[size_t(e)]) m.leaves.push_back(e);
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
  HIP_TRY(hipStreamCreateWithFlags(&ctx->side_stream, hipStreamNonBlocking));
  for (auto& e : ctx->ev) HIP_TRY(hipEventCreate(&e));
  for (hipEvent_t* e : {&ctx->fork_ev, &ctx->join_ev}) HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&ctx->h_status), 64, hipHostMallocMapped));
  HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx->d_status), ctx->h_status, 0));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  ctx->n_cu = prop.multiProcessorCount;
  // the wave kernels assign row groups statically: launch exactly the blocks
  // that are resident at once
  int per_cu = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_describe_wave, kBlock, 0));
  ctx->count_grid = std::max(1, per_cu) * ctx->n_cu;
  if (int rc = fill_grids<FEM2D_PRECISION_FP64>(ctx)) return rc;
  if (int rc = fill_grids<FEM2D_PRECISION_FP32>(ctx)) return rc;
  if (int rc = fill_grids<FEM2D_PRECISION_FP32R>(ctx)) return rc;
  HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_count_large),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(GroupShared))));
  HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fill_large),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(GroupShared))));
  if (int rc = register_scratch(device)) {
    fem2d_ctx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return FEM2D_OK;
}

void fem2d_ctx_destroy(fem2d_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& b : ctx->bufs)
    if (b.p) (void)hipFree(b.p);
  for (DevBuf* b : {&ctx->out_col, &ctx->out_a, &ctx->out_b, &ctx->out_r, &ctx->field_scratch})
    if (b->p) (void)hipFree(b->p);
  for (auto& g : ctx->graph_exec)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : {ctx->fork_ev, ctx->join_ev})
    if (e) (void)hipEventDestroy(e);
  if (ctx->h_status) (void)hipHostFree(ctx->h_status);
  if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
  fem2d::capi::release_solver(ctx->solver);
  if (ctx->side_stream) (void)hipStreamDestroy(ctx->side_stream);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

void* fem2d_ctx_stream(fem2d_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int fem2d_ctx_set_graphs(fem2d_ctx* ctx, int enable) {
  if (!ctx) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  ctx->use_graphs = enable != 0;
  return FEM2D_OK;
}

int fem2d_ctx_last_eigen_stats(fem2d_ctx* ctx, fem2d_eigen_stats* out) {
  if (!ctx || !out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  *out = ctx->eigen_stats;
  return FEM2D_OK;
}

int fem2d_ctx_set_autotune(fem2d_ctx* ctx, int enable) {
  if (!ctx) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  ctx->autotune = enable != 0;
  return FEM2D_OK;
}

int fem2d_ctx_set_fill_chunk(fem2d_ctx* ctx, int chunk) {
  if (!ctx) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if (chunk != 0 && (chunk < kWave || chunk > 4096 || (chunk & (chunk - 1)) != 0)) {
    set_error("fill chunk: 0, or a power of two from 64 to 4096", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  ctx->fill_chunk = chunk;
  return FEM2D_OK;
}

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
  if (ctx->timing_events) {  // the latest assembly's events (it has synchronised)
    float t01 = 0, t12 = 0, t23 = 0, t03 = 0, tc = 0, tf = 0;
    (void)hipSetDevice(ctx->device);
    (void)hipEventElapsedTime(&t03, ctx->ev[0], ctx->ev[3]);
    if (ctx->timing_events == 2) {  // per-stage events exist in kernel-by-kernel mode only
      (void)hipEventElapsedTime(&tf, ctx->ev[6], ctx->ev[7]);
      (void)hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]);
      (void)hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]);
      (void)hipEventElapsedTime(&t23, ctx->ev[2], ctx->ev[3]);
      (void)hipEventElapsedTime(&tc, ctx->ev[4], ctx->ev[5]);
    }
    (void)hipGetLastError();  // a timing query failure is not an assembly failure
    ctx->timing.prep_ms = t01;
    ctx->timing.symbolic_ms = t12;
    ctx->timing.fill_ms = t23;
    ctx->timing.total_ms = t03;
    ctx->timing.count_kernel_ms = tc;
    ctx->timing.fill_kernel_ms = tf;
    ctx->timing_events = 0;
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
  if ((nbs > 0 && (!v->bs_dir || !v->bs_i || !v->bs_j || !v->bs_dof)) || (v->n_dofs > 0 && !v->dof_bs)) {
    set_error("domain view has null basis-spec arrays", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  if (nbs > 0x7fffffffLL) {
    set_error("more basis specs than this build supports (2^31)", 0);
    return FEM2D_ERR_UNSUPPORTED;
  }
  if (ne > 0 && v->bs_offsets[0] != 0) {
    set_error("domain view basis-spec offsets must start at 0", 0);
    return FEM2D_ERR_INVALID_VIEW;
  }
  for (int e = 0; e < ne; ++e)
    if (v->bs_offsets[e + 1] < v->bs_offsets[e]) {
      set_error("domain view basis-spec offsets decrease", 0);
      return FEM2D_ERR_INVALID_VIEW;
    }
  auto* dd = new fem2d_device_domain();
  dd->serial = ++g_domain_serial;
  dd->device = ctx->device;
  dd->cc = v->cc;
  dd->n_elems = ne;
  dd->n_dofs = v->n_dofs;
  dd->n_bs = nbs;
  // host metadata: maximum orders (Mesh::max_expansion_orders, mesh.rs:626-630)
  // and the Elem tree's
  for (int e = 0; e < ne; ++e) {
    dd->max_i = std::max<int>(dd->max_i, v->orders[2 * e]);
    dd->max_j = std::max<int>(dd->max_j, v->orders[2 * e + 1]);
  }
  TreeMeta tm;
  tree_meta(ne, v->parent, v->loc, tm);
  dd->nested_cap = tm.nested_cap;
  dd->nr_u = tm.nr_u;
  dd->nr_v = tm.nr_v;
  dd->range_map_v = tm.range_map_v;
  {
    // row groups: runs of DoFs whose basis specs lie on the same Elems (spec
    // lists are visited in Elem order, so the first Elem is the smaller one)
    std::vector<int> s0(size_t(v->n_dofs), -1), s1(size_t(v->n_dofs), -1);
    for (int e = 0; e < ne; ++e)
      for (long long t = v->bs_offsets[e]; t < v->bs_offsets[e + 1]; ++t) {
        const int d = v->bs_dof[t];
        if (d < 0 || d >= v->n_dofs) continue;  // rejected by the device checks
        if (s0[d] < 0) s0[d] = e;
        else if (s1[d] < 0) s1[d] = e;
      }
    for (long long d = 0; d < v->n_dofs; ++d)
      if (d == 0 || s0[d] != s0[d - 1] || s1[d] != s1[d - 1]) ++dd->n_groups;
  }
  int rc = FEM2D_OK;
  std::vector<long long> off(v->bs_offsets, v->bs_offsets + ne + 1);
  if (ne == 0) off.assign(1, 0);
  if ((rc = upload(v->parent, size_t(ne), &dd->parent)) || (rc = upload(v->loc, size_t(ne), &dd->loc)) ||
      (rc = upload(v->element_box, size_t(4 * ne), &dd->box)) ||
      (rc = upload(v->materials, size_t(2 * ne), &dd->mat)) ||
      (rc = upload(off.data(), off.size(), &dd->bs_off)) ||
      (rc = upload(v->bs_dir, size_t(nbs), &dd->bs_dir)) || (rc = upload(v->bs_i, size_t(nbs), &dd->bs_i)) ||
      (rc = upload(v->bs_j, size_t(nbs), &dd->bs_j)) || (rc = upload(v->bs_dof, size_t(nbs), &dd->bs_dof)) ||
      (rc = upload(v->dof_bs, size_t(2 * v->n_dofs), &dd->dof_bs)) ||
      (rc = upload(tm.keys.data(), tm.keys.size(), &dd->ranges)) ||
      (!tm.range_map.empty() && (rc = upload(tm.range_map.data(), tm.range_map.size(), &dd->range_map))) ||
      (rc = domain_prep(ctx, dd)) || (rc = upload(tm.leaves.data(), tm.leaves.size(), &dd->leaves))) {
    fem2d_device_domain_free(dd);
    return rc;
  }
  dd->n_leaves = (long long)tm.leaves.size();
  *out = dd;
  return FEM2D_OK;
}

void fem2d_device_domain_free(fem2d_device_domain* dd) {
  if (!dd) return;
  (void)hipSetDevice(dd->device);
  if (dd->arena) {
    (void)hipFree(dd->arena);
    delete dd;
    return;
  }
  for (void* p : {(void*)dd->parent, (void*)dd->loc, (void*)dd->box, (void*)dd->mat, (void*)dd->bs_off,
                  (void*)dd->bs_dir, (void*)dd->bs_i, (void*)dd->bs_j, (void*)dd->bs_dof, (void*)dd->dof_bs,
                  (void*)dd->leaves, (void*)dd->ranges, (void*)dd->range_map, (void*)dd->geo, (void*)dd->depth,
                  (void*)dd->n_desc, (void*)dd->dyad, (void*)dd->bs_elem, (void*)dd->bs_pack, (void*)dd->spec_pos,
                  (void*)dd->dof_sig})
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
  int nu = 0, nv = 0;
  if (int rc = check_assembly(dd, glq_u, glq_v, &nu, &nv)) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  // whatever happens below, the output buffers no longer hold the previous
  // assembly's GEP: its handles become stale; nor do the events its times
  ctx->out_generation = 0;
  ctx->timing_events = 0;
  const long long nd = dd->n_dofs;

  // work buffers (grow-only, reused across calls)
  Assembly a;
  int rc = bind_work(dd, nu, nv, basis,
                     [&](int k, size_t count, auto** ptr) { return slot(ctx, k, count, ptr); }, a);
  if (!rc) rc = slot(ctx, kRowPtrSlot, size_t(nd + 1), &a.w.row_ptr);
  if (rc) return rc;
  a.lib_lds = dd->tuned_lds[precision];
  Work& w = a.w;
  const bool run_large = a.run_large;
  hipStream_t st = ctx->stream;
  ensure_glq(ctx, nu, nv);

  // The assembly in three parts: (a) preparation and symbolic phase, ending
  // with the CSR layout and the status block, (b) the streaming fill, (c) the
  // block-path fill (only for Domains with large row groups). With graphs the
  // three are recorded as one graph, replayed while the Domain, the options
  // and the buffers stay the same, and timed end to end; kernel by kernel,
  // events between the stages also time the stages and the describe and fill
  // kernels on their own.
  // record the work of parts() into ctx->graph_exec[i]
  auto record = [&](int i, const std::function<int()>& part) -> int {
    if (ctx->graph_exec[i]) (void)hipGraphExecDestroy(ctx->graph_exec[i]);
    ctx->graph_exec[i] = nullptr;
    (void)hipGetLastError();  // no stale error is reported as a recording failure
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int erc = part();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &graph);  // always leave capture mode
    if (erc != FEM2D_OK || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return erc != FEM2D_OK ? erc : hip_fail(ec, "hipStreamEndCapture");
    }
    const hipError_t ei = hipGraphInstantiate(&ctx->graph_exec[i], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      ctx->graph_exec[i] = nullptr;
      return hip_fail(ei, "hipGraphInstantiate");
    }
    return FEM2D_OK;
  };

  long long nnz = 0;
  int n_large = 0;
  bool fine_timing = false;
  const bool need_res = precision == FEM2D_PRECISION_FP32R;
  for (int attempt = 0; attempt < 2; ++attempt) {
    // FP32R keeps its residuals in a buffer of their own, grown with the others
    if (need_res && (rc = ensure(ctx->out_r, size_t(std::max(1LL, ctx->out_capacity)) * sizeof(uint32_t))))
      return rc;
    const FillOut o = make_fill_out(precision, ai, bi, ctx->out_capacity, static_cast<int*>(ctx->out_col.p),
                                    ctx->out_a.p, ctx->out_b.p, static_cast<unsigned*>(ctx->out_r.p));
    set_chunk(w, ctx->fill_chunk ? ctx->fill_chunk : dd->tuned_chunk[precision], precision);
    w.chunk_cap = o.capacity / w.chunk + 1;
    if ((rc = slot(ctx, kChunkSlot, size_t(w.chunk_cap), &w.chunk_first))) return rc;
    const bool fill = o.capacity > 0;
    const bool tail = fill && run_large;  // part (c) has work
    if (ctx->use_graphs) {
      const GraphKey key{dd->serial, basis,     ai,       bi, nu, nv, o.capacity, g_alloc_generation.load(),
                         run_large,  precision, w.chunk, a.lib_lds};
      if (!ctx->graph_exec[0] || !(key == ctx->graph_key)) {
        ctx->graph_key = GraphKey{};
        if ((rc = record(0, [&] {
               int prc = launch_prep(ctx, st, a, ctx->d_status, false);
               if (prc != FEM2D_OK) return prc;
               // the block-path fill writes segments of its own, from row
               // pointers the layout scan wrote: a branch beside the streaming
               // fill (a branch costs the graph ~20 us, so only this one)
               if (tail) {
                 HIP_TRY(hipEventRecord(ctx->fork_ev, st));
                 HIP_TRY(hipStreamWaitEvent(ctx->side_stream, ctx->fork_ev, 0));
                 if ((prc = launch_large_fill(ctx, ctx->side_stream, a, o))) return prc;
                 HIP_TRY(hipEventRecord(ctx->join_ev, ctx->side_stream));
               }
               if (fill && (prc = launch_stream_fill(ctx, st, a, o))) return prc;
               if (tail) HIP_TRY(hipStreamWaitEvent(st, ctx->join_ev, 0));
               return prc;
             })))
          return rc;
        ctx->graph_key = key;
      }
      // one graph for the whole assembly; only its end points are timed
      HIP_TRY(hipEventRecord(ctx->ev[0], st));
      HIP_TRY(hipGraphLaunch(ctx->graph_exec[0], st));
      HIP_TRY(hipEventRecord(ctx->ev[3], st));
      fine_timing = false;
    } else {
      HIP_TRY(hipEventRecord(ctx->ev[0], st));
      if ((rc = launch_prep(ctx, st, a, ctx->d_status, true))) return rc;
      HIP_TRY(hipEventRecord(ctx->ev[6], st));
      if (fill && (rc = launch_stream_fill(ctx, st, a, o))) return rc;
      HIP_TRY(hipEventRecord(ctx->ev[7], st));
      if (tail && (rc = launch_large_fill(ctx, st, a, o))) return rc;
      HIP_TRY(hipEventRecord(ctx->ev[3], st));
      fine_timing = true;
    }
    HIP_TRY(hipStreamSynchronize(st));
    const volatile long long* stat = ctx->h_status;
    if (stat[1]) return err_to_status(stat[1]);
    nnz = stat[0];
    n_large = int(stat[2]);
    if (nnz <= ctx->out_capacity) break;
    if (attempt == 1) {
      set_error("output capacity mismatch after growing the buffers", 0);
      return FEM2D_ERR_HIP;
    }
    const size_t cap = size_t(nnz + nnz / 16 + 1024);
    if ((rc = ensure(ctx->out_col, cap * sizeof(int))) || (rc = ensure(ctx->out_a, cap * 8)) ||
        (rc = ensure(ctx->out_b, cap * 8)))
      return rc;
    ctx->out_capacity = (long long)cap;
  }

  // the event times are read when the caller asks for them (fem2d_ctx_last_timing)
  ctx->timing_events = fine_timing ? 2 : 1;
  ctx->timing.nnz = nnz;
  ctx->timing.n_groups = ctx->h_status[3];
  ctx->timing.n_large = n_large;
  if (!run_large && n_large > 0) {
    // cannot happen for an unchanged Domain; assemble again with the block path
    dd->large_state = 1;
    return fem2d_assemble_device(ctx, dd, basis, ai, bi, glq_u, glq_v, precision, out);
  }
  dd->large_state = n_large > 0 ? 1 : 0;
  a.run_large = n_large > 0;  // the tuning trials below skip the block path's count when it has no group
  ctx->timing.fill_chunk = w.chunk;
  ctx->timing.fill_lib_lds = fill_place(precision, a.lib_lds, a.lib_n);
  if (ctx->autotune && !ctx->fill_chunk && dd->tuned_chunk[precision] == 0) {
    // the first assembly of this Domain in this precision: time the streaming
    // fill at each candidate chunk size and library placement (best of
    // kTuneReps launches each, after the layout scan for that size), keep the fastest,
    // and assemble again with it
    int best = w.chunk, best_lds = -1;
    float best_ms = HUGE_VALF;
    const FillOut o = make_fill_out(precision, ai, bi, ctx->out_capacity, static_cast<int*>(ctx->out_col.p),
                                    ctx->out_a.p, ctx->out_b.p, static_cast<unsigned*>(ctx->out_r.p));
    // the repetitions are the outer loop, so that a drift of the GPU's clocks
    // over the trials favours no candidate
    for (int rep = 0; rep < kTuneReps; ++rep)
      for (int place : {kPlaceLdsSmall, kPlaceLds, kPlaceGlobal}) {
        if (!place_fits(precision, place, a.lib_n)) continue;
        a.lib_lds = place;
        for (int c : kChunkCandidates) {
          set_chunk(w, c, precision);
          w.chunk_cap = o.capacity / w.chunk + 1;
          if ((rc = slot(ctx, kChunkSlot, size_t(w.chunk_cap), &w.chunk_first))) return rc;
          if ((rc = launch_prep(ctx, st, a, ctx->d_status, false))) return rc;
          HIP_TRY(hipEventRecord(ctx->ev[6], st));
          if ((rc = launch_stream_fill(ctx, st, a, o))) return rc;
          HIP_TRY(hipEventRecord(ctx->ev[7], st));
          HIP_TRY(hipStreamSynchronize(st));
          float ms = 0.0f;
          HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[6], ctx->ev[7]));
          if (ctx->h_status[1]) return err_to_status(ctx->h_status[1]);
          if (tune_verbose())
            std::fprintf(stderr, "[fem2d tune] precision %d placement %d chunk %d: %.4f ms\n", precision, place, c,
                         double(ms));
          if (ms < best_ms) {
            best_ms = ms;
            best = c;
            best_lds = place;
          }
        }
      }
    dd->tuned_chunk[precision] = best;
    dd->tuned_lds[precision] = best_lds;
    return fem2d_assemble_device(ctx, dd, basis, ai, bi, glq_u, glq_v, precision, out);
  }

  out->n = nd;
  out->nnz = nnz;
  out->row_ptr = w.row_ptr;
  out->col = static_cast<int32_t*>(ctx->out_col.p);
  const bool f32 = precision != FEM2D_PRECISION_FP64;  // FP32 and FP32R store FP32 values
  out->precision = precision;
  out->a_val = f32 ? nullptr : static_cast<double*>(ctx->out_a.p);
  out->b_val = f32 ? nullptr : static_cast<double*>(ctx->out_b.p);
  out->a_val32 = f32 ? static_cast<float*>(ctx->out_a.p) : nullptr;
  out->b_val32 = f32 ? static_cast<float*>(ctx->out_b.p) : nullptr;
  out->res16 = precision == FEM2D_PRECISION_FP32R ? static_cast<uint32_t*>(ctx->out_r.p) : nullptr;
  ctx->out_generation = ++g_gep_generation;
  out->generation = ctx->out_generation;
  return FEM2D_OK;
}

int fem2d_assemble_device_noalloc(fem2d_ctx* ctx, const fem2d_device_domain* dd, int basis, int ai, int bi,
                                  int glq_u, int glq_v, int precision, int64_t capacity, int64_t* row_ptr,
                                  int32_t* col, void* a_val, void* b_val, uint32_t* res16, int64_t* nnz_out) {
  if (!ctx || !dd || !nnz_out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  *nnz_out = -1;
  if (int rc = check_common(dd->cc, dd->n_dofs, glq_u, glq_v, basis, ai, bi, precision)) return rc;
  int nu = 0, nv = 0;
  if (int rc = check_assembly(dd, glq_u, glq_v, &nu, &nv)) return rc;
  if (capacity < 0 || !row_ptr ||
      (capacity > 0 && (!col || !a_val || !b_val || (precision == FEM2D_PRECISION_FP32R && !res16)))) {
    set_error("output arrays missing or negative capacity", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  unsigned char* base = static_cast<unsigned char*>(scratch_of(ctx->device));
  if (!base) {
    set_error("no scratch array registered for this device", 0);
    return FEM2D_ERR_UNSUPPORTED;
  }
  // carve the work arrays from the scratch array (256-byte aligned); the
  // total is known only at the end, so overflow is noted and reported after
  size_t used = 0;
  auto carve = [&](size_t bytes) -> void* {
    const size_t at = (used + 255) & ~size_t(255);
    used = at + std::max<size_t>(bytes, 8);
    return used <= kScratchBytes ? base + at : nullptr;
  };
  Assembly a;
  int rc = bind_work(dd, nu, nv, basis,
                     [&](int, size_t count, auto** ptr) {
                       using T = std::remove_pointer_t<std::remove_pointer_t<decltype(ptr)>>;
                       *ptr = static_cast<T*>(carve(count * sizeof(T)));
                       return FEM2D_OK;
                     },
                     a);
  if (rc) return rc;
  a.lib_lds = dd->tuned_lds[precision];
  Work& w = a.w;
  w.row_ptr = row_ptr;
  set_chunk(w, ctx->fill_chunk ? ctx->fill_chunk : dd->tuned_chunk[precision], precision);
  w.chunk_cap = capacity / w.chunk + 1;
  w.chunk_first = static_cast<int*>(carve(size_t(w.chunk_cap) * sizeof(int)));
  if (used > kScratchBytes) {
    set_error("the assembly of this Domain needs " + std::to_string(used) + " bytes of work memory; the " +
                  "no-allocation entry point has " + std::to_string(kScratchBytes) + " (FEM2D_NOALLOC_SCRATCH_BYTES)",
              0);
    return FEM2D_ERR_UNSUPPORTED;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  ensure_glq(ctx, nu, nv);
  const FillOut o = make_fill_out(precision, ai, bi, capacity, col, a_val, b_val, res16);
  // kernel by kernel on the context's stream: k_gram's clearing blocks reset every counter
  // and flag the kernels accumulate into, so the array's earlier contents do
  // not matter
  if ((rc = launch_prep(ctx, st, a, ctx->d_status, false))) return rc;
  if (capacity > 0 && (rc = launch_stream_fill(ctx, st, a, o))) return rc;
  if (capacity > 0 && a.run_large && (rc = launch_large_fill(ctx, st, a, o))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  const volatile long long* stat = ctx->h_status;
  if (stat[1]) return err_to_status(stat[1]);
  const long long nnz = stat[0];
  const int n_large = int(stat[2]);
  if (!a.run_large && n_large > 0) {
    // cannot happen for an unchanged Domain; assemble again with the block path
    dd->large_state = 1;
    return fem2d_assemble_device_noalloc(ctx, dd, basis, ai, bi, glq_u, glq_v, precision, capacity, row_ptr, col,
                                         a_val, b_val, res16, nnz_out);
  }
  dd->large_state = n_large > 0 ? 1 : 0;
  *nnz_out = nnz;
  if (nnz > capacity) {
    set_error("output capacity " + std::to_string(capacity) + " is below the " + std::to_string(nnz) +
                  " stored entries of this GEP (returned in nnz); nothing was written beyond row_ptr",
              0);
    return FEM2D_ERR_INVALID_ARG;
  }
  return FEM2D_OK;
}

int fem2d_device_malloc(fem2d_ctx* ctx, int64_t bytes, void** out) {
  if (!ctx || !out || bytes < 0) {
    set_error("invalid argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMalloc(out, size_t(std::max<int64_t>(bytes, 1))));
  return FEM2D_OK;
}

int fem2d_device_free(fem2d_ctx* ctx, void* p) {
  if (!ctx) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  if (p) HIP_TRY(hipFree(p));
  return FEM2D_OK;
}

int fem2d_memcpy_to_device(fem2d_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx || bytes < 0 || (bytes > 0 && (!dst || !src))) {
    set_error("invalid argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  if (bytes) HIP_TRY(hipMemcpy(dst, src, size_t(bytes), hipMemcpyHostToDevice));
  return FEM2D_OK;
}

int fem2d_memcpy_to_host(fem2d_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx || bytes < 0 || (bytes > 0 && (!dst || !src))) {
    set_error("invalid argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  if (bytes) HIP_TRY(hipMemcpy(dst, src, size_t(bytes), hipMemcpyDeviceToHost));
  return FEM2D_OK;
}

int fem2d_device_domain_num_leaves(const fem2d_device_domain* dd, int64_t* out) {
  if (!dd || !out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  *out = dd->n_leaves;
  return FEM2D_OK;
}

int fem2d_device_xy_fields(fem2d_ctx* ctx, const fem2d_device_domain* dd, int basis, int density_u, int density_v,
                           const double* solution, int64_t n_solution, double* out_x, double* out_y) {
  if (!ctx || !dd) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if (basis != FEM2D_BASIS_HIERPOLY && basis != FEM2D_BASIS_KOL) {
    set_error("unknown basis", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if (density_u < 0 || density_v < 0 || density_u > 4096 || density_v > 4096) {
    set_error("densities outside [0, 4096]", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  // fields.rs:68-72
  if (n_solution != dd->n_dofs) {
    set_error("Domain size (" + std::to_string(dd->n_dofs) + ") does not match solution size (" +
                  std::to_string(n_solution) + ")!",
              0);
    return FEM2D_ERR_MISMATCHED_SOLUTION_SIZE;
  }
  // fields.rs:80-81, 106-107: the value grids are [densities[1]][densities[0]]
  // and are indexed [m][n] with m < densities[0], n < densities[1], which
  // leaves the grid unless the densities are equal (once any spec is visited)
  if (density_u != density_v && density_u > 0 && density_v > 0 && dd->n_dofs > 0) {
    set_error("index out of bounds: the field grid is indexed [u][v] but sized [v][u]", 0);
    return FEM2D_ERR_PANIC;
  }
  const int d = density_u;
  const long long total = dd->n_leaves * (long long)d * (d > 0 && density_v > 0 ? d : 0);
  if (total == 0) return FEM2D_OK;
  if (!solution || !out_x || !out_y) {
    set_error("null array", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  // uniform_range(-1.0, 1.0, d) (fields.rs:402-405), followed by a zeroed
  // status word, staged in one copy to the context's scratch buffer
  std::vector<double> pts(static_cast<size_t>(d) + 1, 0.0);
  const double step = (1.0 - -1.0) / double(d - 1);
  for (int i = 0; i < d; ++i) pts[size_t(i)] = double(i) * step + -1.0;
  int rc = ensure(ctx->field_scratch, pts.size() * sizeof(double));
  double* d_pts = static_cast<double*>(ctx->field_scratch.p);
  int* d_err = reinterpret_cast<int*>(d_pts + d);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(d_pts, pts.data(), pts.size() * sizeof(double), hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // pts is read before any return below
    if (e != hipSuccess) rc = hip_fail(e, "field grid upload");
  }
  if (!rc) {
    DevView v{};
    v.n_elems = dd->n_elems;
    v.n_dofs = dd->n_dofs;
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
    v.dof_bs = dd->dof_bs;
    const unsigned grid =
        unsigned(std::min<long long>((total + kBlock - 1) / kBlock, std::max(1, ctx->n_cu) * 16LL));
    if (basis == FEM2D_BASIS_KOL)
      k_xy_fields<FEM2D_BASIS_KOL><<<grid, kBlock, 0, ctx->stream>>>(v, dd->leaves, dd->n_leaves, d, d_pts, solution,
                                                                    out_x, out_y, d_err);
    else
      k_xy_fields<FEM2D_BASIS_HIERPOLY><<<grid, kBlock, 0, ctx->stream>>>(v, dd->leaves, dd->n_leaves, d, d_pts,
                                                                         solution, out_x, out_y, d_err);
    hipError_t e = hipGetLastError();
    int h_err = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h_err, d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = hip_fail(e, "field evaluation");
    else if (h_err) {
      set_error("domain view violates the Domain invariants (parent chain, location or DoF id)", 0);
      rc = FEM2D_ERR_INVALID_VIEW;
    }
  }
  return rc;
}

int fem2d_xy_fields(fem2d_ctx* ctx, const fem2d_domain_view* view, int basis, int density_u, int density_v,
                    const double* solution, int64_t n_solution, double* out_x, double* out_y) {
  if (!ctx || !view) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if (n_solution != view->n_dofs) {
    set_error("Domain size (" + std::to_string(view->n_dofs) + ") does not match solution size (" +
                  std::to_string(n_solution) + ")!",
              0);
    return FEM2D_ERR_MISMATCHED_SOLUTION_SIZE;
  }
  fem2d_device_domain* dd = nullptr;
  if (int rc = fem2d_device_domain_upload(ctx, view, &dd)) return rc;
  const long long total =
      dd->n_leaves * (long long)std::max(0, density_u) * (long long)std::max(0, density_v);
  double *d_sol = nullptr, *d_x = nullptr, *d_y = nullptr;
  int rc = FEM2D_OK;
  if (density_u == density_v && total > 0) {
    if (!solution || !out_x || !out_y) {
      set_error("null array", 0);
      rc = FEM2D_ERR_INVALID_ARG;
    }
    if (!rc) rc = upload(solution, size_t(n_solution), &d_sol);
    if (!rc) rc = alloc_device(size_t(total), &d_x);
    if (!rc) rc = alloc_device(size_t(total), &d_y);
  }
  if (!rc) rc = fem2d_device_xy_fields(ctx, dd, basis, density_u, density_v, d_sol, n_solution, d_x, d_y);
  if (!rc && d_x) {
    hipError_t e = hipMemcpy(out_x, d_x, size_t(total) * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out_y, d_y, size_t(total) * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "field download");
  }
  for (void* p : {(void*)d_sol, (void*)d_x, (void*)d_y})
    if (p) (void)hipFree(p);
  fem2d_device_domain_free(dd);
  return rc;
}

int fem2d_gep_device_download(fem2d_ctx* ctx, const fem2d_gep_device* src, fem2d_gep_csr* out) {
  if (!ctx || !src || !out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  std::memset(out, 0, sizeof(*out));
  if (int rc = fem2d::capi::check_gep_handle(ctx, src)) return rc;
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
  // through pinned staging buffers; each chunk is spread over the destination
  // (pageable, first touched here) by several host threads while the next
  // chunk is in flight
  fem2d::PinnedStaging ps;
  HIP_TRY(ps.init(ctx->stream));
  const int threads = std::min(8, fem2d::default_threads());
  auto copy_to = [&](void* dst) {
    return [dst, threads](const unsigned char* p, size_t off, size_t len) -> int {
      unsigned char* d = static_cast<unsigned char*>(dst) + off;
      fem2d::parallel_ranges(len, threads, size_t(1) << 22,
                             [&](size_t b, size_t e) { std::memcpy(d + b, p + b, e - b); });
      return 0;
    };
  };
  // FP32 values widened to double on the way
  auto widen_to = [&](double* dst) {
    return [dst, threads](const unsigned char* p, size_t off, size_t len) -> int {
      double* d = dst + off / sizeof(float);
      fem2d::parallel_ranges(len / sizeof(float), threads, size_t(1) << 20, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
          float x;
          std::memcpy(&x, p + i * sizeof(float), sizeof(float));
          d[i] = double(x);
        }
      });
      return 0;
    };
  };
  // FP32R residuals applied to values already widened from their FP32 parts
  auto add_residuals = [&](double* da, double* db) {
    return [da, db, threads](const unsigned char* p, size_t off, size_t len) -> int {
      const size_t k0 = off / sizeof(uint32_t);
      fem2d::parallel_ranges(len / sizeof(uint32_t), threads, size_t(1) << 20, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
          uint32_t r;
          std::memcpy(&r, p + i * sizeof(uint32_t), sizeof(uint32_t));
          da[k0 + i] = fem2d::fp32r_join(float(da[k0 + i]), fem2d::fp32r_res_a(r));
          db[k0 + i] = fem2d::fp32r_join(float(db[k0 + i]), fem2d::fp32r_res_b(r));
        }
      });
      return 0;
    };
  };
  hipError_t e = hipSuccess;
  const size_t nnz = size_t(std::max<int64_t>(0, src->nnz));
  // each value array in the width its precision stores (FP32 widened to double)
  if (src->precision != FEM2D_PRECISION_FP64 && src->precision != FEM2D_PRECISION_FP32 &&
      src->precision != FEM2D_PRECISION_FP32R) {
    fem2d_gep_csr_free(out);
    set_error("unknown precision", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  const bool f32 = src->precision == FEM2D_PRECISION_FP32, f32r = src->precision == FEM2D_PRECISION_FP32R;
  if ((f32 || f32r) && src->nnz > 0 && (!src->a_val32 || !src->b_val32 || (f32r && !src->res16))) {
    fem2d_gep_csr_free(out);
    set_error("GEP arrays missing", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  int rc = ps.run(src->row_ptr, size_t(src->n + 1) * sizeof(int64_t), copy_to(out->row_ptr), &e);
  if (!rc) rc = ps.run(src->col, nnz * sizeof(int32_t), copy_to(out->col), &e);
  if (!rc && (f32 || f32r)) rc = ps.run(src->a_val32, nnz * sizeof(float), widen_to(out->a_val), &e);
  if (!rc && (f32 || f32r)) rc = ps.run(src->b_val32, nnz * sizeof(float), widen_to(out->b_val), &e);
  // FP32R: each widened value gains its residual (fp32r_join), both matrices from one word
  if (!rc && f32r) rc = ps.run(src->res16, nnz * sizeof(uint32_t), add_residuals(out->a_val, out->b_val), &e);
  if (!rc && !f32 && !f32r) rc = ps.run(src->a_val, nnz * sizeof(double), copy_to(out->a_val), &e);
  if (!rc && !f32 && !f32r) rc = ps.run(src->b_val, nnz * sizeof(double), copy_to(out->b_val), &e);
  if (rc) {
    fem2d_gep_csr_free(out);
    return rc == -1 ? hip_fail(e, "GEP download") : rc;
  }
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

int fem2d_device_domain_info(const fem2d_device_domain* dd, int32_t* n_elems, int64_t* n_dofs, int64_t* n_bs) {
  if (!dd) {
    set_error("null device domain", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if (n_elems) *n_elems = dd->n_elems;
  if (n_dofs) *n_dofs = dd->n_dofs;
  if (n_bs) *n_bs = dd->n_bs;
  return FEM2D_OK;
}

int fem2d_device_domain_download_view(fem2d_ctx* ctx, const fem2d_device_domain* dd, int32_t* parent, uint8_t* loc,
                                      double* element_box, double* materials, int64_t* bs_offsets, uint8_t* bs_dir,
                                      uint8_t* bs_i, uint8_t* bs_j, int32_t* bs_dof, int32_t* dof_bs) {
  if (!ctx || !dd) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  HIP_TRY(hipSetDevice(dd->device));
  const size_t ne = size_t(dd->n_elems), nbs = size_t(dd->n_bs), nd = size_t(dd->n_dofs);
  struct Part {
    void* dst;
    const void* src;
    size_t bytes;
  } parts[] = {{parent, dd->parent, ne * sizeof(int32_t)},
               {loc, dd->loc, ne},
               {element_box, dd->box, 4 * ne * sizeof(double)},
               {materials, dd->mat, 2 * ne * sizeof(double)},
               {bs_offsets, dd->bs_off, (ne + 1) * sizeof(int64_t)},
               {bs_dir, dd->bs_dir, nbs},
               {bs_i, dd->bs_i, nbs},
               {bs_j, dd->bs_j, nbs},
               {bs_dof, dd->bs_dof, nbs * sizeof(int32_t)},
               {dof_bs, dd->dof_bs, 2 * nd * sizeof(int32_t)}};
  for (const Part& q : parts)
    if (q.dst && q.bytes) HIP_TRY(hipMemcpyAsync(q.dst, q.src, q.bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return FEM2D_OK;
}

}  // extern "C"

namespace {

// Sub-allocations of one device allocation, each 256-byte aligned and at
// least one element long
struct ArenaPlan {
  size_t bytes = 0;
  template <class T>
  size_t take(size_t count) {
    const size_t off = bytes;
    bytes += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~size_t(255);
    return off;
  }
};

}  // namespace

// Domain::from_mesh on the device (the kernels of domain_build.hip). One
// allocation holds the Domain and the build's work arrays; the flattened Mesh
// and the Elem tree's metadata go up in one copy; the counts' totals come back
// once to size the lists, and the status once at the end.
int fem2d::capi::device_domain_from_flat(fem2d_ctx* ctx, const MeshFlat& m, int32_t cc, fem2d_device_domain** out) {
  namespace B = fem2d::build;
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  const int ne = m.n_elems, ned = m.n_edges;
  const size_t E = size_t(ne), G = size_t(ned);
  const size_t cap = size_t(std::max<int64_t>(m.uv_specs, 1));
  TreeMeta tm;
  tree_meta(ne, m.parent.data(), m.loc.data(), tm);

  // the host-written part first (one copy), then the device-written part
  ArenaPlan plan;
  const size_t o_status = plan.take<B::Status>(1);
  const size_t o_parent = plan.take<int>(E), o_loc = plan.take<unsigned char>(E);
  const size_t o_box = plan.take<double>(4 * E), o_mat = plan.take<double>(2 * E);
  const size_t o_edges = plan.take<int4>(E), o_meta = plan.take<unsigned>(E);
  const size_t o_side = plan.take<int2>(G), o_eparent = plan.take<int>(G);
  const size_t o_child = plan.take<int2>(G), o_open = plan.take<unsigned char>(G);
  const size_t o_ranges = plan.take<long long>(tm.keys.size());
  const size_t o_rmap = plan.take<int>(tm.range_map.size());
  const size_t o_leaves = plan.take<int>(tm.leaves.size());
  // the preparation kernels accumulate into these two: they go up as zeros
  const size_t o_depth = plan.take<int>(E), o_ndesc = plan.take<int>(E);
  const size_t host_bytes = plan.bytes;
  const size_t o_bs_off = plan.take<long long>(E + 1), o_elem_base = plan.take<long long>(E + 1);
  const size_t o_edge_base = plan.take<long long>(G + 1);
  const size_t o_tiles = plan.take<long long>(size_t(B::scan_tiles(ne, ned)));
  const size_t o_act = plan.take<int2>(G), o_npairs = plan.take<int>(G);
  const size_t o_nelem = plan.take<int>(E), o_len = plan.take<int>(E);
  const size_t o_dir = plan.take<unsigned char>(cap), o_bi = plan.take<unsigned char>(cap);
  const size_t o_bj = plan.take<unsigned char>(cap), o_bdof = plan.take<int>(cap);
  const size_t o_dof_bs = plan.take<int>(2 * cap);
  const size_t o_geo = plan.take<ElemGeo>(E);
  const size_t o_dyad = plan.take<int4>(E);
  const size_t o_bs_elem = plan.take<int>(cap), o_bs_pack = plan.take<unsigned>(cap);
  const size_t o_spec_pos = plan.take<unsigned>(cap), o_dof_sig = plan.take<int4>(cap);

  // the context's pinned staging buffer (grow-only; free again once this call
  // has synchronised), so the upload is one DMA transfer
  if (ctx->h_stage_bytes < host_bytes) {
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    ctx->h_stage = nullptr;
    ctx->h_stage_bytes = 0;
    HIP_TRY(hipHostMalloc(&ctx->h_stage, host_bytes + host_bytes / 4, hipHostMallocDefault));
    ctx->h_stage_bytes = host_bytes + host_bytes / 4;
  }
  unsigned char* stage = static_cast<unsigned char*>(ctx->h_stage);
  std::memset(stage + o_status, 0, sizeof(B::Status));
  std::memset(stage + o_depth, 0, host_bytes - o_depth);  // depth and n_desc
  auto put = [&](size_t off, const void* src, size_t bytes) {
    if (bytes) std::memcpy(stage + off, src, bytes);
  };
  put(o_parent, m.parent.data(), E * sizeof(int));
  put(o_loc, m.loc.data(), E);
  put(o_box, m.box.data(), 4 * E * sizeof(double));
  put(o_mat, m.mat.data(), 2 * E * sizeof(double));
  put(o_edges, m.edges.data(), 4 * E * sizeof(int));
  put(o_meta, m.meta.data(), E * sizeof(unsigned));
  put(o_side, m.side.data(), 2 * G * sizeof(int));
  put(o_eparent, m.edge_parent.data(), G * sizeof(int));
  put(o_child, m.child.data(), 2 * G * sizeof(int));
  put(o_open, m.open_base.data(), G);
  put(o_ranges, tm.keys.data(), tm.keys.size() * sizeof(long long));
  put(o_rmap, tm.range_map.data(), tm.range_map.size() * sizeof(int));
  put(o_leaves, tm.leaves.data(), tm.leaves.size() * sizeof(int));

  auto* dd = new fem2d_device_domain();
  struct Guard {
    fem2d_device_domain*& p;
    ~Guard() { fem2d_device_domain_free(p); }
  } guard{dd};
  dd->serial = ++g_domain_serial;
  dd->device = ctx->device;
  dd->cc = cc;
  dd->n_elems = ne;
  dd->max_i = m.max_i;
  dd->max_j = m.max_j;
  dd->nested_cap = tm.nested_cap;
  dd->nr_u = tm.nr_u;
  dd->nr_v = tm.nr_v;
  dd->range_map_v = tm.range_map_v;
  dd->n_leaves = (long long)tm.leaves.size();
  HIP_TRY(hipMalloc(&dd->arena, plan.bytes));
  unsigned char* base = static_cast<unsigned char*>(dd->arena);
  auto at = [&](size_t off) { return static_cast<void*>(base + off); };
  dd->parent = static_cast<int*>(at(o_parent));
  dd->loc = static_cast<unsigned char*>(at(o_loc));
  dd->box = static_cast<double*>(at(o_box));
  dd->mat = static_cast<double*>(at(o_mat));
  dd->ranges = static_cast<long long*>(at(o_ranges));
  dd->range_map = tm.range_map.empty() ? nullptr : static_cast<int*>(at(o_rmap));
  dd->leaves = static_cast<int*>(at(o_leaves));
  dd->bs_off = static_cast<long long*>(at(o_bs_off));
  dd->bs_dir = static_cast<unsigned char*>(at(o_dir));
  dd->bs_i = static_cast<unsigned char*>(at(o_bi));
  dd->bs_j = static_cast<unsigned char*>(at(o_bj));
  dd->bs_dof = static_cast<int*>(at(o_bdof));
  dd->dof_bs = static_cast<int*>(at(o_dof_bs));
  dd->geo = static_cast<ElemGeo*>(at(o_geo));
  dd->depth = static_cast<int*>(at(o_depth));
  dd->n_desc = static_cast<int*>(at(o_ndesc));
  dd->dyad = static_cast<int4*>(at(o_dyad));
  dd->bs_elem = static_cast<int*>(at(o_bs_elem));
  dd->bs_pack = static_cast<unsigned*>(at(o_bs_pack));
  dd->spec_pos = static_cast<unsigned*>(at(o_spec_pos));
  dd->dof_sig = static_cast<int4*>(at(o_dof_sig));
  auto* status = static_cast<B::Status*>(at(o_status));

  B::Args a;
  a.n_elems = ne;
  a.n_edges = ned;
  a.spec_cap = (long long)cap;
  a.elem_edges = static_cast<const int4*>(at(o_edges));
  a.elem_meta = static_cast<const unsigned*>(at(o_meta));
  a.side = static_cast<const int2*>(at(o_side));
  a.edge_parent = static_cast<const int*>(at(o_eparent));
  a.child = static_cast<const int2*>(at(o_child));
  a.open_base = static_cast<const unsigned char*>(at(o_open));
  a.act = static_cast<int2*>(at(o_act));
  a.n_pairs = static_cast<int*>(at(o_npairs));
  a.n_elem = static_cast<int*>(at(o_nelem));
  a.list_len = static_cast<int*>(at(o_len));
  a.elem_base = static_cast<long long*>(at(o_elem_base));
  a.edge_base = static_cast<long long*>(at(o_edge_base));
  a.tile_sums = static_cast<long long*>(at(o_tiles));
  a.status = status;
  a.bs_off = dd->bs_off;
  a.bs_dir = dd->bs_dir;
  a.bs_i = dd->bs_i;
  a.bs_j = dd->bs_j;
  a.bs_dof = dd->bs_dof;
  a.dof_bs = dd->dof_bs;

  hipStream_t st = ctx->stream;
  HIP_TRY(hipMemcpyAsync(dd->arena, stage, host_bytes, hipMemcpyHostToDevice, st));
  B::Status h{};
  if (ne > 0) {
    HIP_TRY_AT(B::launch_counts(a, st), "launching the Domain counts");
    HIP_TRY(hipMemcpyAsync(&h, status, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (h.flags & B::kFlagTooDeep) {
      set_error("an Edge lies more than " + std::to_string(B::kMaxEdgeDepth) +
                    " refinements below its base Edge (the device Domain build's limit)",
                0);
      return FEM2D_ERR_UNSUPPORTED;
    }
    if (h.flags & (B::kFlagNoPair | B::kFlagInconsistent)) return kActivationFailed;
    if (h.elem_dofs + h.edge_dofs > 0x7fffffffLL) {
      set_error("too many DoFs for 32-bit ids", 0);
      return FEM2D_ERR_PANIC;
    }
    if (h.specs > (long long)cap || h.elem_dofs + h.edge_dofs > h.specs) {
      set_error("device Domain build: list sizes beyond their bound", 0);
      return FEM2D_ERR_HIP;
    }
    HIP_TRY_AT(B::launch_lists(a, st), "launching the Domain lists");
  } else {
    HIP_TRY(hipMemsetAsync(dd->bs_off, 0, sizeof(long long), st));
  }
  dd->n_dofs = h.elem_dofs + h.edge_dofs;
  dd->n_bs = h.specs;
  if (int rc = domain_prep_launch(ctx, dd, &status->err, true)) return rc;
  HIP_TRY_AT(B::launch_group_count(dd->dof_bs, dd->bs_elem, dd->n_dofs, dd->n_bs, status, st), "launching the group count");
  HIP_TRY(hipMemcpyAsync(&h, status, sizeof(h), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (h.flags & B::kFlagCapacity) {
    set_error("device Domain build: a list entry beyond the arrays' capacity", 0);
    return FEM2D_ERR_HIP;
  }
  if (h.err) return err_to_status(h.err);
  dd->n_groups = (long long)h.groups;
  *out = dd;
  dd = nullptr;  // owned by the caller now
  return FEM2D_OK;
}
