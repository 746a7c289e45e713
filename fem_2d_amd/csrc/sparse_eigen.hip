// Sparse shift-invert eigensolve of an assembled GEP on the GPU: the eigenpair
// of A x = lambda B x nearest a target, directly from the device-resident
// upper-triangular CSR, for GEPs of any size whose factor fits HBM. This is
// the computation slepc_solve_gep (slepc_solve.rs:14-68) hands to SLEPc with a
// target eigenvalue (shift-and-invert around it).
//
// Ordering (host, from the pattern alone): breadth-first level sets of the
// matrix graph from a pseudo-peripheral DoF (George-Liu). Every neighbour of
// a DoF lies in its own level or in one of the two next to it, so any
// combination of A and B in level order is block tridiagonal. Consecutive small
// levels are merged into blocks (which keeps it block tridiagonal); the
// components of a disconnected graph (independent instances) follow one
// another, and their levels merge the same way.
//
// Factorisation (device, dense blocks on rocSOLVER / rocBLAS): the block
// Cholesky factorisation of the SPD matrix K' = A + tau B (tau =
// max(|target|, 1)). With D_i its diagonal blocks and U_i the blocks coupling
// block i to block i + 1,
//   S_0 = D_0,  S_i = L_i L_i^T (potrf),  W_i = L_i^-1 U_i (trsm),
//   S_{i+1} = D_{i+1} - W_i^T W_i (syrk).
// Once W_i is formed, L_i is replaced by its inverse (trtri), kept beside its
// transpose: 2 sum_i w_i^2 doubles of HBM. Cholesky is backward stable for
// SPD matrices, so the Schur complements stay SPD even though K' reaches
// condition numbers ~1e13 on high-order bases (the mass matrix's smallest
// eigenvalues are ~1e-10). Explicit inverses of the S_i lose accuracy with the
// square of that condition number and broke the elimination down past ~1e5
// DoFs; eliminating the indefinite A - target B the same way is not stable
// either. The inverses of the triangular factors lose only its first power: M
// below solves K' to ~1e-3 relative on the 4 496-DoF precision-study GEP.
//
// Solve K' x = b in level order (the preconditioner M = K'^-1):
//   forward   z_i = S_i^-1 (b_i - U_{i-1}^T z_{i-1})
//   backward  x_last = z_last,  x_i = z_i - S_i^-1 (U_i x_{i+1})
// with S_i^-1 v = L_i^-T (L_i^-1 v): L_i^-1 v as the column dot products of
// the stored transpose, L_i^-T of the result as those of L_i^-1 itself, one
// wavefront per column reading it contiguously (HBM-bound; rocBLAS trmv ran
// these at ~1.4 TB/s), and sparse products with the coupling entries, taken
// from the CSR in place.
//
// Eigenpair: shift-invert inverse iteration in residual-correction form,
// x <- x - d with (A - target B) d = A x - lambda B x solved by GMRES
// preconditioned with M, and lambda the Rayleigh quotient, B-normalised. With
// d exact the step is (lambda - target) (A - target B)^-1 B x, plain inverse
// iteration, which converges to the eigenvector whose eigenvalue is nearest
// the target; written as a correction, its fixed points are exact eigenpairs
// whatever the inner accuracy, so the residual reaches FP64 level.

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fem2d.h"
#include "eigen_common.hpp"
#include "fp32r.hpp"
#include "host/capi_internal.hpp"

namespace {

using namespace fem2d::eig;

// consecutive levels are merged into one block while the block stays within
// this width (levels wider than it are blocks of their own)
constexpr long long kMergeWidth = 2048;
constexpr int kMaxIter = 200;
constexpr double kTol = 1e-12, kTolStall = 1e-10;

// ---------------------------------------------------------------------------
// Host symbolic phase: level-set ordering and the entry lists of the blocks
// ---------------------------------------------------------------------------

struct Plan {
  long long n = 0;
  int nb = 0;                      // blocks
  std::vector<long long> bstart;   // [nb + 1] first level-order position of each block
  std::vector<int> perm;           // level-order position -> DoF
  std::vector<int> block_of;       // level-order position -> block
  // entries (stored index k, local row, local column) of each block's
  // diagonal block D_i and coupling block U_i (rows in block i, columns in i + 1)
  std::vector<long long> d_off, u_off;  // [nb + 1]
  std::vector<int> d_k, d_r, d_c, u_k, u_r, u_c;
  // per level-order position p: the coupling entries of its row of U (columns
  // in the next block) and of U^T (columns in the previous block), with the
  // columns as level-order positions
  std::vector<long long> ur_ptr, ut_ptr;  // [n + 1]
  std::vector<int> ur_k, ur_c, ut_k, ut_c;
  // the strict lower triangle of the full matrix by rows (DoF order): the
  // stored entries (k, row) of column r, rows ascending
  std::vector<long long> lo_ptr;  // [n + 1]
  std::vector<int> lo_k, lo_row;
  double sum_w2 = 0.0;       // doubles of the kept inverses
  long long max_coupling = 0;  // largest w_i w_{i+1}
  long long max_w = 0;
};

// levels of the component of `start` by breadth-first search; returns the
// number of levels; lvl[v] >= 0 for the visited DoFs, order receives them by
// level (ascending DoF id within a level)
int bfs_levels(long long start, const std::vector<long long>& ap, const std::vector<int>& aj, std::vector<int>& lvl,
               std::vector<int>& order, std::vector<int>& touched) {
  for (int v : touched) lvl[size_t(v)] = -1;
  touched.clear();
  order.clear();
  lvl[size_t(start)] = 0;
  touched.push_back(int(start));
  order.push_back(int(start));
  size_t head = 0;
  int depth = 0;
  while (head < order.size()) {
    const size_t level_end = order.size();
    for (; head < level_end; ++head) {
      const int v = order[head];
      for (long long j = ap[size_t(v)]; j < ap[size_t(v) + 1]; ++j) {
        const int u = aj[size_t(j)];
        if (lvl[size_t(u)] < 0) {
          lvl[size_t(u)] = lvl[size_t(v)] + 1;
          touched.push_back(u);
          order.push_back(u);
        }
      }
    }
    if (order.size() > level_end) {
      std::sort(order.begin() + long(level_end), order.end());
      ++depth;
    }
  }
  return depth + 1;
}

int build_plan(long long n, const std::vector<int64_t>& rp, const std::vector<int32_t>& col, Plan& P) {
  P.n = n;
  // full adjacency without the diagonal
  std::vector<long long> deg(size_t(n) + 1, 0);
  for (long long r = 0; r < n; ++r)
    for (long long k = rp[size_t(r)]; k < rp[size_t(r) + 1]; ++k) {
      const int c = col[size_t(k)];
      if (c < r || c >= n) {
        set_error("GEP is not upper-triangular CSR with columns inside the matrix", 0);
        return FEM2D_ERR_INVALID_ARG;
      }
      if (c != r) {
        ++deg[size_t(r) + 1];
        ++deg[size_t(c) + 1];
      }
    }
  std::vector<long long> ap(size_t(n) + 1, 0);
  for (long long v = 0; v < n; ++v) ap[size_t(v) + 1] = ap[size_t(v)] + deg[size_t(v) + 1];
  std::vector<int> aj(static_cast<size_t>(ap[size_t(n)]));
  {
    std::vector<long long> fill(ap.begin(), ap.end() - 1);
    for (long long r = 0; r < n; ++r)
      for (long long k = rp[size_t(r)]; k < rp[size_t(r) + 1]; ++k) {
        const int c = col[size_t(k)];
        if (c == r) continue;
        aj[size_t(fill[size_t(r)]++)] = c;
        aj[size_t(fill[size_t(c)]++)] = int(r);
      }
  }
  // levels, component by component, each from a pseudo-peripheral DoF
  std::vector<int> lvl(size_t(n), -1), order, touched;
  std::vector<char> done(size_t(n), 0);
  std::vector<long long> level_sizes;  // over all components, in sequence
  P.perm.clear();
  P.perm.reserve(size_t(n));
  for (long long seed = 0; seed < n; ++seed) {
    if (done[size_t(seed)]) continue;
    long long s = seed;
    int depth = bfs_levels(s, ap, aj, lvl, order, touched);
    for (int it = 0; it < 8; ++it) {
      // the last level's DoF of smallest degree (lowest id on a tie)
      long long cand = -1;
      for (int v : order)
        if (lvl[size_t(v)] == depth - 1 &&
            (cand < 0 || ap[size_t(v) + 1] - ap[size_t(v)] < ap[size_t(cand) + 1] - ap[size_t(cand)]))
          cand = v;
      std::vector<int> lvl2 = lvl;  // keep the current level structure if the candidate is no better
      std::vector<int> order2 = order;
      const int d2 = bfs_levels(cand, ap, aj, lvl, order, touched);
      if (d2 <= depth) {
        lvl = std::move(lvl2);
        order = std::move(order2);
        break;
      }
      depth = d2;
      s = cand;
    }
    std::vector<long long> sizes(size_t(depth), 0);
    for (int v : order) {
      ++sizes[size_t(lvl[size_t(v)])];
      done[size_t(v)] = 1;
      P.perm.push_back(v);
    }
    level_sizes.insert(level_sizes.end(), sizes.begin(), sizes.end());
    for (int v : order) lvl[size_t(v)] = -1;  // bfs_levels resets only what it touched itself
    touched.clear();
  }
  // blocks: consecutive levels merged while within kMergeWidth
  P.bstart.assign(1, 0);
  long long cur = 0, at = 0;
  for (long long w : level_sizes) {
    if (cur > 0 && cur + w > kMergeWidth) {
      P.bstart.push_back(at);
      cur = 0;
    }
    cur += w;
    at += w;
  }
  if (cur > 0) P.bstart.push_back(at);
  P.nb = int(P.bstart.size()) - 1;
  std::vector<int> pos(static_cast<size_t>(n));
  for (long long p = 0; p < n; ++p) pos[size_t(P.perm[size_t(p)])] = int(p);
  P.block_of.assign(size_t(n), 0);
  for (int b = 0; b < P.nb; ++b) {
    for (long long p = P.bstart[size_t(b)]; p < P.bstart[size_t(b) + 1]; ++p) P.block_of[size_t(p)] = b;
    const long long w = P.bstart[size_t(b) + 1] - P.bstart[size_t(b)];
    P.sum_w2 += double(w) * double(w);
    P.max_w = std::max(P.max_w, w);
    if (b + 1 < P.nb)
      P.max_coupling = std::max(P.max_coupling, w * (P.bstart[size_t(b) + 2] - P.bstart[size_t(b) + 1]));
  }
  // entry lists by block (counting sort on the block index)
  P.d_off.assign(size_t(P.nb) + 1, 0);
  P.u_off.assign(size_t(P.nb) + 1, 0);
  P.ur_ptr.assign(size_t(n) + 1, 0);
  P.ut_ptr.assign(size_t(n) + 1, 0);
  auto classify = [&](long long r, long long c, int& br, int& bc, long long& pr, long long& pc) {
    pr = pos[size_t(r)];
    pc = pos[size_t(c)];
    br = P.block_of[size_t(pr)];
    bc = P.block_of[size_t(pc)];
    if (bc < br) {  // the coupling is stored with the earlier block as rows
      std::swap(br, bc);
      std::swap(pr, pc);
    }
  };
  for (long long r = 0; r < n; ++r)
    for (long long k = rp[size_t(r)]; k < rp[size_t(r) + 1]; ++k) {
      int br, bc;
      long long pr, pc;
      classify(r, col[size_t(k)], br, bc, pr, pc);
      if (br == bc) {
        ++P.d_off[size_t(br) + 1];
      } else if (bc == br + 1) {
        ++P.u_off[size_t(br) + 1];
        ++P.ur_ptr[size_t(pr) + 1];
        ++P.ut_ptr[size_t(pc) + 1];
      } else {
        set_error("internal error: an entry couples blocks that are not adjacent in the level ordering", 0);
        return FEM2D_ERR_HIP;
      }
    }
  for (int b = 0; b < P.nb; ++b) {
    P.d_off[size_t(b) + 1] += P.d_off[size_t(b)];
    P.u_off[size_t(b) + 1] += P.u_off[size_t(b)];
  }
  for (long long p = 0; p < n; ++p) {
    P.ur_ptr[size_t(p) + 1] += P.ur_ptr[size_t(p)];
    P.ut_ptr[size_t(p) + 1] += P.ut_ptr[size_t(p)];
  }
  const long long nd = P.d_off[size_t(P.nb)], nu = P.u_off[size_t(P.nb)];
  P.d_k.resize(size_t(nd));
  P.d_r.resize(size_t(nd));
  P.d_c.resize(size_t(nd));
  P.u_k.resize(size_t(nu));
  P.u_r.resize(size_t(nu));
  P.u_c.resize(size_t(nu));
  P.ur_k.resize(size_t(nu));
  P.ur_c.resize(size_t(nu));
  P.ut_k.resize(size_t(nu));
  P.ut_c.resize(size_t(nu));
  {
    std::vector<long long> fd(P.d_off.begin(), P.d_off.end() - 1), fu(P.u_off.begin(), P.u_off.end() - 1);
    std::vector<long long> fr(P.ur_ptr.begin(), P.ur_ptr.end() - 1), ft(P.ut_ptr.begin(), P.ut_ptr.end() - 1);
    for (long long r = 0; r < n; ++r)
      for (long long k = rp[size_t(r)]; k < rp[size_t(r) + 1]; ++k) {
        int br, bc;
        long long pr, pc;
        classify(r, col[size_t(k)], br, bc, pr, pc);
        if (br == bc) {
          const long long j = fd[size_t(br)]++;
          P.d_k[size_t(j)] = int(k);
          P.d_r[size_t(j)] = int(pr - P.bstart[size_t(br)]);
          P.d_c[size_t(j)] = int(pc - P.bstart[size_t(br)]);
        } else {
          const long long j = fu[size_t(br)]++;
          P.u_k[size_t(j)] = int(k);
          P.u_r[size_t(j)] = int(pr - P.bstart[size_t(br)]);
          P.u_c[size_t(j)] = int(pc - P.bstart[size_t(bc)]);
          const long long a = fr[size_t(pr)]++;
          P.ur_k[size_t(a)] = int(k);
          P.ur_c[size_t(a)] = int(pc);
          const long long t = ft[size_t(pc)]++;
          P.ut_k[size_t(t)] = int(k);
          P.ut_c[size_t(t)] = int(pr);
        }
      }
  }
  // the strict lower triangle by rows: stored entries of each column
  P.lo_ptr.assign(size_t(n) + 1, 0);
  for (long long r = 0; r < n; ++r)
    for (long long k = rp[size_t(r)]; k < rp[size_t(r) + 1]; ++k)
      if (col[size_t(k)] != r) ++P.lo_ptr[size_t(col[size_t(k)]) + 1];
  for (long long v = 0; v < n; ++v) P.lo_ptr[size_t(v) + 1] += P.lo_ptr[size_t(v)];
  P.lo_k.resize(size_t(P.lo_ptr[size_t(n)]));
  P.lo_row.resize(size_t(P.lo_ptr[size_t(n)]));
  {
    std::vector<long long> fl(P.lo_ptr.begin(), P.lo_ptr.end() - 1);
    for (long long r = 0; r < n; ++r)
      for (long long k = rp[size_t(r)]; k < rp[size_t(r) + 1]; ++k) {
        const int c = col[size_t(k)];
        if (c == r) continue;
        const long long j = fl[size_t(c)]++;
        P.lo_k[size_t(j)] = int(k);
        P.lo_row[size_t(j)] = int(r);
      }
  }
  return FEM2D_OK;
}

// ---------------------------------------------------------------------------
// Device kernels
// ---------------------------------------------------------------------------

// FP64 copies of the stored values (FP32 widened; FP32R joined with its residuals)
template <class V>
__global__ void k_values64(long long nnz, const V* __restrict__ a, const V* __restrict__ b,
                           const uint32_t* __restrict__ res, double* __restrict__ a64, double* __restrict__ b64) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nnz; k += stride) {
    double av = double(a[k]), bv = double(b[k]);
    if (res) {
      const uint32_t q = res[k];
      av = fem2d::fp32r_join(float(av), fem2d::fp32r_res_a(q));
      bv = fem2d::fp32r_join(float(bv), fem2d::fp32r_res_b(q));
    }
    a64[k] = av;
    b64[k] = bv;
  }
}

__global__ void k_fill(double* __restrict__ p, long long count, double v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count; i += stride) p[i] = v;
}

// *out = min over j of L[j, j]^2 from the inverse factor's diagonal 1 / L[j, j]
// (as the bits of a non-negative double, whose integer order is the numeric
// order; *out starts at all ones)
__global__ void k_min_pivot(const double* __restrict__ Linv, long long w, unsigned long long* __restrict__ out) {
  const long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (j >= w) return;
  const double v = 1.0 / (Linv[j + j * w] * Linv[j + j * w]);
  atomicMin(out, static_cast<unsigned long long>(__double_as_longlong(v)));
}

// sum over the wavefront (every lane receives it)
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

// u = l^T for the lower triangle of a column-major w x w block: u[j + i w] =
// l[i + j w] for i >= j (u's strict lower triangle is left as it is)
__global__ void k_transpose_lower(const double* __restrict__ l, double* __restrict__ u, long long w) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < w * w; k += stride) {
    const long long i = k % w, j = k / w;
    if (i >= j) u[j + i * w] = l[k];
  }
}

// y[c] = sum_r m[r + c w] x[r] (+ beta y[c]) over the rows r <= c of an
// upper-triangular column-major block (kUpper) or r >= c of a lower one: one
// wavefront per column, reading it contiguously (four loads in flight a lane)
template <bool kUpper>
__global__ void k_tri_cols(const double* __restrict__ m, long long w, const double* __restrict__ x,
                           double* __restrict__ y, double beta) {
  const int lane = threadIdx.x % kWave;
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave;
  const long long waves = ((long long)gridDim.x * blockDim.x) / kWave;
  for (long long c = wave; c < w; c += waves) {
    const long long r0 = kUpper ? 0 : c, r1 = kUpper ? c + 1 : w;
    const double* __restrict__ col = m + c * w;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    long long r = r0 + lane;
    for (; r + 3 * kWave < r1; r += 4 * kWave) {
      s0 += col[r] * x[r];
      s1 += col[r + kWave] * x[r + kWave];
      s2 += col[r + 2 * kWave] * x[r + 2 * kWave];
      s3 += col[r + 3 * kWave] * x[r + 3 * kWave];
    }
    for (; r < r1; r += kWave) s0 += col[r] * x[r];
    const double s = wave_sum((s0 + s1) + (s2 + s3));
    if (lane == 0) y[c] = beta == 0.0 ? s : s + beta * y[c];
  }
}

// entries of K = A - sigma B into a dense column-major block (leading
// dimension ld); sym: also at the transposed position (a diagonal block)
__global__ void k_scatter(long long count, const int* __restrict__ ks, const int* __restrict__ rs,
                          const int* __restrict__ cs, const double* __restrict__ a, const double* __restrict__ b,
                          double sigma, double* __restrict__ M, long long ld, int sym) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < count; j += stride) {
    const long long k = ks[j], r = rs[j], c = cs[j];
    const double v = a[k] - sigma * b[k];
    M[r + c * ld] = v;
    if (sym) M[c + r * ld] = v;
  }
}


// y[p] = beta y[p] - sum over the listed entries (k, c) of p: K_k x[c], for
// the level-order positions p in [p0, p1) (one wavefront per row)
__global__ void k_coupling(long long p0, long long p1, const long long* __restrict__ ptr, const int* __restrict__ ks,
                           const int* __restrict__ cs, const double* __restrict__ a, const double* __restrict__ b,
                           double sigma, const double* __restrict__ x, double* __restrict__ y, double beta) {
  const int lane = threadIdx.x % kWave;
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave;
  const long long waves = (gridDim.x * (long long)blockDim.x) / kWave;
  for (long long p = p0 + wave; p < p1; p += waves) {
    double s = 0.0;
    for (long long j = ptr[p] + lane; j < ptr[p + 1]; j += kWave) {
      const long long k = ks[j];
      s += (a[k] - sigma * b[k]) * x[cs[j]];
    }
    s = wave_sum(s);
    if (lane == 0) y[p] = (beta == 0.0 ? 0.0 : beta * y[p]) - s;
  }
}

// ax = A x and bx = B x of the full symmetric matrices from the stored upper
// triangle and its strict lower triangle by rows (one wavefront per row)
__global__ void k_spmv2(long long n, const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                        const long long* __restrict__ lo_ptr, const int* __restrict__ lo_k,
                        const int* __restrict__ lo_row, const double* __restrict__ a, const double* __restrict__ b,
                        const double* __restrict__ x, double* __restrict__ ax, double* __restrict__ bx) {
  const int lane = threadIdx.x % kWave;
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave;
  const long long waves = (gridDim.x * (long long)blockDim.x) / kWave;
  for (long long r = wave; r < n; r += waves) {
    double sa = 0.0, sb = 0.0;
    for (long long k = rp[r] + lane; k < rp[r + 1]; k += kWave) {
      const double xc = x[col[k]];
      sa += a[k] * xc;
      sb += b[k] * xc;
    }
    for (long long j = lo_ptr[r] + lane; j < lo_ptr[r + 1]; j += kWave) {
      const long long k = lo_k[j];
      const double xr = x[lo_row[j]];
      sa += a[k] * xr;
      sb += b[k] * xr;
    }
    sa = wave_sum(sa);
    sb = wave_sum(sb);
    if (lane == 0) {
      ax[r] = sa;
      bx[r] = sb;
    }
  }
}

// out = ax - sigma bx
__global__ void k_shifted(long long n, const double* __restrict__ ax, const double* __restrict__ bx, double sigma,
                          double* __restrict__ out) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = ax[i] - sigma * bx[i];
}

// dst[p] = src[perm[p]] (to level order) or dst[perm[p]] = src[p] (back)
__global__ void k_permute(long long n, const int* __restrict__ perm, const double* __restrict__ src,
                          double* __restrict__ dst, int back) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += stride) {
    if (back) dst[perm[p]] = src[p];
    else dst[p] = src[perm[p]];
  }
}

unsigned row_grid(long long rows) {
  return unsigned(std::max<long long>(1, std::min<long long>((rows + 3) / 4, 1 << 16)));
}

// ---------------------------------------------------------------------------
// Numeric phase and solves
// ---------------------------------------------------------------------------

struct Factor {
  const Plan* P = nullptr;
  rocblas_handle h = nullptr;
  hipStream_t st = nullptr;
  double sigma = 0.0;  // the factored matrix is A - sigma B (sigma = -tau: A + tau B)
  const double* a = nullptr;  // FP64 values of the stored entries
  const double* b = nullptr;
  double* chol = nullptr;     // the inverses of the blocks' Cholesky factors L_i^-1 (lower, column-major), at off[i]
  double* cholT = nullptr;    // their transposes (upper, column-major), at the same offsets
  std::vector<long long> off;
  // device copies of the plan's lists
  int *d_k, *d_r, *d_c, *u_k, *u_r, *u_c, *ur_k, *ur_c, *ut_k, *ut_c, *perm;
  long long *ur_ptr, *ut_ptr;
  double *y, *z, *t;  // level-order work vectors
  double* tmp = nullptr;  // one block's intermediate L_i^-1 v
  unsigned long long* pivots = nullptr;  // [nb] smallest squared diagonal entry of each factor (as bits)
  double min_pivot = 0.0;                // smallest over all factors
  int bad_block = -1;                    // first block that is not positive definite
};

template <class T>
int upload_vec(DevMem& m, const std::vector<T>& v, T** out) {
  EIG_HIP(m.alloc(out, v.size()));
  if (!v.empty()) EIG_HIP(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return FEM2D_OK;
}

// The block Cholesky factors L_i of A - sigma B (SPD for sigma = -tau).
// F.bad_block receives the first block whose Schur complement is not
// positive definite (potrf's info), -1 when there is none.
int factorize(Factor& F, double* wbuf, rocblas_int* d_info) {
  const Plan& P = *F.P;
  const hipStream_t st = F.st;
  const double one = 1.0, m_one = -1.0;
  auto w_of = [&](int i) { return P.bstart[size_t(i) + 1] - P.bstart[size_t(i)]; };
  auto load_diag = [&](int i) -> int {  // block i of the slot area = D_i (both triangles)
    const long long w = w_of(i);
    double* S = F.chol + F.off[size_t(i)];
    k_fill<<<grid_for(w * w), kBlock, 0, st>>>(S, w * w, 0.0);
    const long long c0 = P.d_off[size_t(i)], c1 = P.d_off[size_t(i) + 1];
    if (c1 > c0)
      k_scatter<<<grid_for(c1 - c0), kBlock, 0, st>>>(c1 - c0, F.d_k + c0, F.d_r + c0, F.d_c + c0, F.a, F.b, F.sigma,
                                                     S, w, 1);
    EIG_HIP(hipGetLastError());
    return FEM2D_OK;
  };
  if (int rc = load_diag(0)) return rc;
  for (int i = 0; i < P.nb; ++i) {
    const long long w = w_of(i);
    const rocblas_int wi = rocblas_int(w);
    double* L = F.chol + F.off[size_t(i)];
    EIG_BLAS(rocsolver_dpotrf(F.h, rocblas_fill_lower, wi, L, wi, d_info + i));
    if (i + 1 == P.nb) {
      EIG_BLAS(rocsolver_dtrtri(F.h, rocblas_fill_lower, rocblas_diagonal_non_unit, wi, L, wi, d_info + P.nb + i));
      k_transpose_lower<<<grid_for(w * w), kBlock, 0, st>>>(L, F.cholT + F.off[size_t(i)], w);
      EIG_HIP(hipGetLastError());
      break;
    }
    const long long w2 = w_of(i + 1);
    const rocblas_int wn = rocblas_int(w2);
    // W_i = L_i^-1 U_i (w x w2, leading dimension w)
    k_fill<<<grid_for(w * w2), kBlock, 0, st>>>(wbuf, w * w2, 0.0);
    const long long u0 = P.u_off[size_t(i)], u1 = P.u_off[size_t(i) + 1];
    if (u1 > u0)
      k_scatter<<<grid_for(u1 - u0), kBlock, 0, st>>>(u1 - u0, F.u_k + u0, F.u_r + u0, F.u_c + u0, F.a, F.b, F.sigma,
                                                     wbuf, w, 0);
    EIG_HIP(hipGetLastError());
    EIG_BLAS(rocblas_dtrsm(F.h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                           rocblas_diagonal_non_unit, wi, wn, &one, L, wi, wbuf, wi));
    // L_i is needed no more: keep its inverse (and that inverse's transpose)
    // for the solves
    EIG_BLAS(rocsolver_dtrtri(F.h, rocblas_fill_lower, rocblas_diagonal_non_unit, wi, L, wi, d_info + P.nb + i));
    k_transpose_lower<<<grid_for(w * w), kBlock, 0, st>>>(L, F.cholT + F.off[size_t(i)], w);
    EIG_HIP(hipGetLastError());
    // S_{i+1} = D_{i+1} - W_i^T W_i (lower triangle, which is all potrf reads)
    if (int rc = load_diag(i + 1)) return rc;
    EIG_BLAS(rocblas_dsyrk(F.h, rocblas_fill_lower, rocblas_operation_transpose, wn, wi, &m_one, wbuf, wi, &one,
                           F.chol + F.off[size_t(i) + 1], wn));
  }
  std::vector<rocblas_int> info(2 * size_t(P.nb));
  EIG_HIP(hipMemcpyAsync(info.data(), d_info, info.size() * sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  // the smallest pivot L_jj^2 of each factor
  EIG_HIP(hipMemsetAsync(F.pivots, 0xff, size_t(P.nb) * sizeof(unsigned long long), st));
  for (int i = 0; i < P.nb; ++i) {
    const long long w = w_of(i);
    k_min_pivot<<<grid_for(w), kBlock, 0, st>>>(F.chol + F.off[size_t(i)], w, F.pivots + i);
  }
  EIG_HIP(hipGetLastError());
  std::vector<unsigned long long> piv(size_t(P.nb));
  EIG_HIP(hipMemcpyAsync(piv.data(), F.pivots, piv.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  EIG_HIP(hipStreamSynchronize(st));
  F.bad_block = -1;
  F.min_pivot = HUGE_VAL;
  for (int i = 0; i < P.nb; ++i) {
    if ((info[size_t(i)] != 0 || info[size_t(P.nb + i)] != 0) && F.bad_block < 0) F.bad_block = i;
    double v;
    std::memcpy(&v, &piv[size_t(i)], sizeof v);
    F.min_pivot = std::min(F.min_pivot, v);
  }
  return FEM2D_OK;
}

// out = K^-1 f (both in DoF order)
int solve(Factor& F, const double* f, double* out) {
  const Plan& P = *F.P;
  const hipStream_t st = F.st;
  const long long n = P.n;
  // out <- S_i^-1 in + beta out = L_i^-T (L_i^-1 in) + beta out: L_i^-1 in as
  // the column dots of its transpose, then L_i^-T of that as the column dots
  // of L_i^-1 itself (both read their blocks contiguously)
  auto chol_solve = [&](int i, const double* in, double* out, double beta) -> int {
    const long long w = P.bstart[size_t(i) + 1] - P.bstart[size_t(i)];
    const unsigned grid = grid_for(w * kWave);
    k_tri_cols<true><<<grid, kBlock, 0, st>>>(F.cholT + F.off[size_t(i)], w, in, F.tmp, 0.0);
    k_tri_cols<false><<<grid, kBlock, 0, st>>>(F.chol + F.off[size_t(i)], w, F.tmp, out, beta);
    EIG_HIP(hipGetLastError());
    return FEM2D_OK;
  };
  k_permute<<<grid_for(n), kBlock, 0, st>>>(n, F.perm, f, F.y, 0);
  EIG_HIP(hipGetLastError());
  for (int i = 0; i < P.nb; ++i) {
    const long long p0 = P.bstart[size_t(i)], p1 = P.bstart[size_t(i) + 1];
    // z_i = S_i^-1 y_i
    if (int rc = chol_solve(i, F.y + p0, F.z + p0, 0.0)) return rc;
    if (i + 1 < P.nb) {  // y_{i+1} -= U_i^T z_i
      const long long q1 = P.bstart[size_t(i) + 2];
      k_coupling<<<row_grid(q1 - p1), kBlock, 0, st>>>(p1, q1, F.ut_ptr, F.ut_k, F.ut_c, F.a, F.b, F.sigma, F.z, F.y,
                                                      1.0);
      EIG_HIP(hipGetLastError());
    }
  }
  for (int i = P.nb - 2; i >= 0; --i) {
    const long long p0 = P.bstart[size_t(i)], p1 = P.bstart[size_t(i) + 1];
    // t_i = -U_i x_{i+1}; x_i = z_i + S_i^-1 t_i
    k_coupling<<<row_grid(p1 - p0), kBlock, 0, st>>>(p0, p1, F.ur_ptr, F.ur_k, F.ur_c, F.a, F.b, F.sigma, F.z, F.t,
                                                    0.0);
    EIG_HIP(hipGetLastError());
    if (int rc = chol_solve(i, F.t + p0, F.z + p0, 1.0)) return rc;
  }
  k_permute<<<grid_for(n), kBlock, 0, st>>>(n, F.perm, F.z, out, 1);
  EIG_HIP(hipGetLastError());
  return FEM2D_OK;
}

// The sparse matrices in place: A x and B x of the full symmetric matrices
struct SymMats {
  long long n = 0;
  const int64_t* rp = nullptr;
  const int32_t* col = nullptr;
  const long long* lo_ptr = nullptr;
  const int* lo_k = nullptr;
  const int* lo_row = nullptr;
  const double* a = nullptr;
  const double* b = nullptr;
  unsigned rows = 1;
  hipStream_t st = nullptr;
  int apply(const double* x, double* ax, double* bx) const {
    k_spmv2<<<rows, kBlock, 0, st>>>(n, rp, col, lo_ptr, lo_k, lo_row, a, b, x, ax, bx);
    EIG_HIP(hipGetLastError());
    return FEM2D_OK;
  }
};

// work vectors of the inner solve
constexpr int kGmresRestart = 30;
constexpr int kGmresCycles = 4;
constexpr double kInnerTol = 1e-4;  // inner relative residual that the outer iteration needs
struct Gmres {
  double* v[kGmresRestart + 1];
  double* z[kGmresRestart];
  double *w, *ax, *bx;
};

// d ~ (A - target B)^-1 r by GMRES on (A - target B) M u = r, right
// preconditioned with M the block inverse of the SPD matrix A + tau B, from
// d = 0, restarted, until the relative residual is below kInnerTol. The
// preconditioned operator's eigenvalues are (lambda - target) / (lambda + tau)
// over the pencil's: -target / tau on the gradient null space, near 1 for high
// modes, and near 0 only for the few modes close to the target.
int inner_solve(Factor& F, const SymMats& S, double target, rocblas_handle h, const double* r, double* d, Gmres& G,
                fem2d_eigen_stats* stats) {
  const rocblas_int ni = rocblas_int(S.n);
  const hipStream_t st = S.st;
  double rn = 0.0;
  EIG_BLAS(rocblas_dnrm2(h, ni, r, 1, &rn));
  const int steps0 = stats->inner_iterations;
  struct Count {  // the steps of this solve, recorded however it returns
    fem2d_eigen_stats* s;
    int from;
    ~Count() { s->max_inner_per_solve = std::max(s->max_inner_per_solve, double(s->inner_iterations - from)); }
  } count{stats, steps0};
  k_fill<<<grid_for(S.n), kBlock, 0, st>>>(d, S.n, 0.0);
  EIG_HIP(hipGetLastError());
  if (rn == 0.0) return FEM2D_OK;
  const double m_one = -1.0;
  for (int cycle = 0; cycle < kGmresCycles; ++cycle) {
    // v0 = r - K d
    EIG_BLAS(rocblas_dcopy(h, ni, r, 1, G.v[0], 1));
    if (cycle > 0) {
      if (int rc = S.apply(d, G.ax, G.bx)) return rc;
      k_shifted<<<grid_for(S.n), kBlock, 0, st>>>(S.n, G.ax, G.bx, target, G.w);
      EIG_HIP(hipGetLastError());
      EIG_BLAS(rocblas_daxpy(h, ni, &m_one, G.w, 1, G.v[0], 1));
    }
    double beta = 0.0;
    EIG_BLAS(rocblas_dnrm2(h, ni, G.v[0], 1, &beta));
    if (beta <= kInnerTol * rn) return FEM2D_OK;
    const double ib = 1.0 / beta;
    EIG_BLAS(rocblas_dscal(h, ni, &ib, G.v[0], 1));
    static thread_local double H[kGmresRestart + 1][kGmresRestart];
    double cs[kGmresRestart] = {}, sn[kGmresRestart] = {}, g[kGmresRestart + 1] = {};
    for (auto& row : H)
      for (double& x : row) x = 0.0;
    g[0] = beta;
    int m = 0;
    for (int j = 0; j < kGmresRestart; ++j) {
      if (int rc = solve(F, G.v[j], G.z[j])) return rc;  // z_j = M v_j
      if (int rc = S.apply(G.z[j], G.ax, G.bx)) return rc;
      k_shifted<<<grid_for(S.n), kBlock, 0, st>>>(S.n, G.ax, G.bx, target, G.w);  // w = K z_j
      EIG_HIP(hipGetLastError());
      for (int i = 0; i <= j; ++i) {  // modified Gram-Schmidt
        EIG_BLAS(rocblas_ddot(h, ni, G.w, 1, G.v[i], 1, &H[i][j]));
        const double neg = -H[i][j];
        EIG_BLAS(rocblas_daxpy(h, ni, &neg, G.v[i], 1, G.w, 1));
      }
      EIG_BLAS(rocblas_dnrm2(h, ni, G.w, 1, &H[j + 1][j]));
      if (H[j + 1][j] > 0.0) {
        const double inv = 1.0 / H[j + 1][j];
        EIG_BLAS(rocblas_dcopy(h, ni, G.w, 1, G.v[j + 1], 1));
        EIG_BLAS(rocblas_dscal(h, ni, &inv, G.v[j + 1], 1));
      }
      for (int i = 0; i < j; ++i) {  // earlier rotations on the new column
        const double t = cs[i] * H[i][j] + sn[i] * H[i + 1][j];
        H[i + 1][j] = -sn[i] * H[i][j] + cs[i] * H[i + 1][j];
        H[i][j] = t;
      }
      const double den = std::hypot(H[j][j], H[j + 1][j]);
      cs[j] = den > 0.0 ? H[j][j] / den : 1.0;
      sn[j] = den > 0.0 ? H[j + 1][j] / den : 0.0;
      H[j][j] = den;
      H[j + 1][j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      m = j + 1;
      ++stats->inner_iterations;
      if (std::fabs(g[j + 1]) <= 0.5 * kInnerTol * rn || !(den > 0.0)) break;
    }
    // y from the triangular system, d += sum_j y_j z_j
    double y[kGmresRestart] = {};
    for (int i = m - 1; i >= 0; --i) {
      double t = g[i];
      for (int k = i + 1; k < m; ++k) t -= H[i][k] * y[k];
      y[i] = H[i][i] != 0.0 ? t / H[i][i] : 0.0;
    }
    for (int i = 0; i < m; ++i) EIG_BLAS(rocblas_daxpy(h, ni, &y[i], G.z[i], 1, d, 1));
  }
  return FEM2D_OK;
}

int sparse_solve(fem2d_ctx* ctx, const fem2d_gep_device* gep, double target, fem2d_eigenpair* out) {
  if (int rc = check_device_args(ctx, gep, out)) return rc;
  const long long n = gep->n, nnz = gep->nnz;
  if (n == 0) {
    set_error("Only spurious modes were found!", 0);
    return FEM2D_ERR_SPURIOUSLY_CONVERGED;
  }
  const int prec = gep->precision;
  const bool f32r = prec == FEM2D_PRECISION_FP32R, f32 = prec == FEM2D_PRECISION_FP32 || f32r;
  if (prec != FEM2D_PRECISION_FP64 && !f32) {
    set_error("unknown precision", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if (nnz < 0 || nnz > 0x7fffffffLL || n > 0x7fffffffLL || !gep->row_ptr ||
      (nnz > 0 && (!gep->col || (f32 ? !gep->a_val32 || !gep->b_val32 : !gep->a_val || !gep->b_val) ||
                   (f32r && !gep->res16)))) {
    set_error("GEP sizes or arrays invalid", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  EIG_HIP(hipSetDevice(fem2d::capi::ctx_device(ctx)));
  hipStream_t st = static_cast<hipStream_t>(fem2d_ctx_stream(ctx));
  rocblas_handle h;
  if (int rc = solver_handle(ctx, st, &h)) return rc;
  fem2d_eigen_stats* stats = fem2d::capi::ctx_eigen_stats(ctx);
  *stats = fem2d_eigen_stats{};
  stats->n = n;
  using clock = std::chrono::steady_clock;
  auto ms_since = [](clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(clock::now() - t0).count();
  };
  const auto t_sym = clock::now();
  // symbolic phase on the host, from the pattern
  std::vector<int64_t> rp(size_t(n) + 1);
  std::vector<int32_t> col(static_cast<size_t>(nnz));
  EIG_HIP(hipMemcpyAsync(rp.data(), gep->row_ptr, rp.size() * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  if (nnz) EIG_HIP(hipMemcpyAsync(col.data(), gep->col, col.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  EIG_HIP(hipStreamSynchronize(st));
  if (rp[0] != 0 || rp[size_t(n)] != nnz) {
    set_error("row_ptr does not span [0, nnz]", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  for (long long r = 0; r < n; ++r)
    if (rp[size_t(r) + 1] < rp[size_t(r)]) {
      set_error("row_ptr decreases", 0);
      return FEM2D_ERR_INVALID_ARG;
    }
  Plan P;
  if (int rc = build_plan(n, rp, col, P)) return rc;
  stats->blocks = P.nb;
  stats->max_block = int(P.max_w);
  stats->factor_bytes = 16.0 * P.sum_w2;  // the inverse factors and their transposes
  stats->symbolic_ms = ms_since(t_sym);
  // memory: the factors, one coupling-size block, values, lists, vectors
  size_t free_b = 0, total_b = 0;
  EIG_HIP(hipMemGetInfo(&free_b, &total_b));
  const double need = 16.0 * P.sum_w2 + 8.0 * double(P.max_coupling) + 16.0 * double(nnz) +
                      4.0 * double(P.d_k.size() * 3 + P.u_k.size() * 7 + P.lo_k.size() * 2) + 120.0 * double(n);
  if (need > 0.92 * double(free_b)) {
    set_error("the block factorisation of this GEP needs " + std::to_string(need / 1e9) + " GB of device memory, " +
                  std::to_string(double(free_b) / 1e9) + " GB are free",
              0);
    return FEM2D_ERR_PROBLEM_TOO_LARGE;
  }
  DevMem m;
  Factor F;
  F.P = &P;
  F.h = h;
  F.st = st;
  // FP64 values
  double *a64 = nullptr, *b64 = nullptr;
  if (f32) {
    EIG_HIP(m.alloc(&a64, size_t(nnz)));
    EIG_HIP(m.alloc(&b64, size_t(nnz)));
    k_values64<<<grid_for(nnz), kBlock, 0, st>>>(nnz, gep->a_val32, gep->b_val32, f32r ? gep->res16 : nullptr, a64,
                                                 b64);
    EIG_HIP(hipGetLastError());
    F.a = a64;
    F.b = b64;
  } else {
    F.a = gep->a_val;
    F.b = gep->b_val;
  }
  int rc = FEM2D_OK;
  long long *lo_ptr;
  int *lo_k, *lo_row;
  if ((rc = upload_vec(m, P.d_k, &F.d_k)) || (rc = upload_vec(m, P.d_r, &F.d_r)) ||
      (rc = upload_vec(m, P.d_c, &F.d_c)) || (rc = upload_vec(m, P.u_k, &F.u_k)) ||
      (rc = upload_vec(m, P.u_r, &F.u_r)) || (rc = upload_vec(m, P.u_c, &F.u_c)) ||
      (rc = upload_vec(m, P.ur_k, &F.ur_k)) || (rc = upload_vec(m, P.ur_c, &F.ur_c)) ||
      (rc = upload_vec(m, P.ut_k, &F.ut_k)) || (rc = upload_vec(m, P.ut_c, &F.ut_c)) ||
      (rc = upload_vec(m, P.perm, &F.perm)) || (rc = upload_vec(m, P.ur_ptr, &F.ur_ptr)) ||
      (rc = upload_vec(m, P.ut_ptr, &F.ut_ptr)) || (rc = upload_vec(m, P.lo_ptr, &lo_ptr)) ||
      (rc = upload_vec(m, P.lo_k, &lo_k)) || (rc = upload_vec(m, P.lo_row, &lo_row)))
    return rc;
  F.off.assign(size_t(P.nb) + 1, 0);
  for (int i = 0; i < P.nb; ++i) {
    const long long w = P.bstart[size_t(i) + 1] - P.bstart[size_t(i)];
    F.off[size_t(i) + 1] = F.off[size_t(i)] + w * w;
  }
  double *wbuf, *x, *bx, *ax, *y, *ay, *by, *r, *dy;
  rocblas_int* d_info;
  EIG_HIP(m.alloc(&F.chol, size_t(F.off[size_t(P.nb)])));
  EIG_HIP(m.alloc(&F.cholT, size_t(F.off[size_t(P.nb)])));
  EIG_HIP(m.alloc(&F.tmp, size_t(std::max(1LL, P.max_w))));
  EIG_HIP(m.alloc(&wbuf, size_t(std::max(1LL, P.max_coupling))));
  EIG_HIP(m.alloc(&d_info, 2 * size_t(P.nb)));
  EIG_HIP(m.alloc(&F.pivots, size_t(P.nb)));
  for (double** v : {&F.y, &F.z, &F.t, &x, &bx, &ax, &y, &ay, &by, &r, &dy}) EIG_HIP(m.alloc(v, size_t(n)));
  Gmres G;
  for (double*& v : G.v) EIG_HIP(m.alloc(&v, size_t(n)));
  for (double*& v : G.z) EIG_HIP(m.alloc(&v, size_t(n)));
  for (double** v : {&G.w, &G.ax, &G.bx}) EIG_HIP(m.alloc(v, size_t(n)));
  const auto t_fac = clock::now();
  // the preconditioner: the block Cholesky factorisation of the SPD matrix A + tau B
  const double tau = std::max(std::fabs(target), 1.0);
  F.sigma = -tau;
  if ((rc = factorize(F, wbuf, d_info))) return rc;
  stats->min_pivot = F.min_pivot;
  stats->bad_block = F.bad_block;
  if (F.bad_block >= 0) {
    set_error("A + tau B is not positive definite (block " + std::to_string(F.bad_block) + " of " +
                  std::to_string(P.nb) + "): B is not positive definite or A not semi-definite",
              0);
    return FEM2D_ERR_FAILED_TO_INVERT_B;
  }
  stats->factor_ms = ms_since(t_fac);  // factorize ends with a stream synchronisation
  const auto t_it = clock::now();
  SymMats S;
  S.n = n;
  S.rp = gep->row_ptr;
  S.col = gep->col;
  S.lo_ptr = lo_ptr;
  S.lo_k = lo_k;
  S.lo_row = lo_row;
  S.a = F.a;
  S.b = F.b;
  S.rows = row_grid(n);
  S.st = st;
  const rocblas_int ni = rocblas_int(n);
  const std::vector<double> x0 = start_vector(n);
  {
    // the preconditioner's accuracy: z = x0, |M (A + tau B) z - z| / |z|
    EIG_HIP(hipMemcpyAsync(x, x0.data(), size_t(n) * sizeof(double), hipMemcpyHostToDevice, st));
    if ((rc = S.apply(x, ax, bx))) return rc;
    k_shifted<<<grid_for(n), kBlock, 0, st>>>(n, ax, bx, F.sigma, r);
    EIG_HIP(hipGetLastError());
    if ((rc = solve(F, r, dy))) return rc;
    const double m_one = -1.0;
    EIG_BLAS(rocblas_daxpy(h, ni, &m_one, x, 1, dy, 1));
    double e = 0.0, z = 0.0;
    EIG_BLAS(rocblas_dnrm2(h, ni, dy, 1, &e));
    EIG_BLAS(rocblas_dnrm2(h, ni, x, 1, &z));
    stats->precond_error = e / z;
  }
  if (verbose())
    std::fprintf(stderr, "[fem2d sparse] n=%lld blocks=%d max_block=%d factor %.1f ms min pivot %.3e precond error %.3e\n",
                 n, P.nb, int(P.max_w), stats->factor_ms, stats->min_pivot, stats->precond_error);
  EIG_HIP(hipMemcpyAsync(y, x0.data(), size_t(n) * sizeof(double), hipMemcpyHostToDevice, st));
  EIG_HIP(hipStreamSynchronize(st));  // x0 is read before any return below
  double lam_prev = HUGE_VAL;
  for (int it = 0; it <= kMaxIter; ++it) {
    // x = y / sqrt(y^T B y), with A x and B x; lambda = x^T A x
    if ((rc = S.apply(y, ay, by))) return rc;
    stats->iterations = it;
    double s = 0.0;
    EIG_BLAS(rocblas_ddot(h, ni, y, 1, by, 1, &s));
    if (!(s > 0.0) || !std::isfinite(s)) {
      set_error("the shift-invert iteration broke down (B is not positive definite on the iterate)", 0);
      return FEM2D_ERR_NOT_CONVERGED;
    }
    const double inv = 1.0 / std::sqrt(s);
    EIG_BLAS(rocblas_dcopy(h, ni, y, 1, x, 1));
    EIG_BLAS(rocblas_dscal(h, ni, &inv, x, 1));  // x^T B x = 1
    EIG_BLAS(rocblas_dcopy(h, ni, by, 1, bx, 1));
    EIG_BLAS(rocblas_dscal(h, ni, &inv, bx, 1));
    EIG_BLAS(rocblas_dcopy(h, ni, ay, 1, ax, 1));
    EIG_BLAS(rocblas_dscal(h, ni, &inv, ax, 1));
    double lam = 0.0, na = 0.0, nb = 0.0, nr = 0.0;
    EIG_BLAS(rocblas_ddot(h, ni, x, 1, ax, 1, &lam));
    EIG_BLAS(rocblas_dnrm2(h, ni, ax, 1, &na));
    EIG_BLAS(rocblas_dnrm2(h, ni, bx, 1, &nb));
    const double neg = -lam;
    EIG_BLAS(rocblas_dcopy(h, ni, ax, 1, r, 1));
    EIG_BLAS(rocblas_daxpy(h, ni, &neg, bx, 1, r, 1));  // r = A x - lambda B x
    EIG_BLAS(rocblas_dnrm2(h, ni, r, 1, &nr));
    if (!std::isfinite(lam) || !std::isfinite(nr)) {
      set_error("the shift-invert iteration produced a non-finite iterate", 0);
      return FEM2D_ERR_NOT_CONVERGED;
    }
    // relative residual |A x - lambda B x| / max(|A x|, |target| |B x|): the second
    // term measures an iterate of the gradient null space (lambda ~ 0, A x ~ 0)
    const double scale = std::max(na, std::fabs(target) * nb);
    const bool stalled = std::fabs(lam - lam_prev) <= 1e-14 * std::fabs(lam);
    lam_prev = lam;
    stats->residual = nr / scale;
    stats->iterate_ms = ms_since(t_it);
    if (verbose())
      std::fprintf(stderr, "[fem2d sparse] outer %d lambda %.15g residual %.3e inner %d (%.0f ms)\n", it, lam,
                   stats->residual, stats->inner_iterations, stats->iterate_ms);
    if (it > 0 && (nr <= kTol * scale || (stalled && nr <= kTolStall * scale) || nr == 0.0)) {
      double* v = static_cast<double*>(std::malloc(size_t(n) * sizeof(double)));
      if (!v) {
        set_error("host allocation of the eigenvector failed", 0);
        return FEM2D_ERR_OOM;
      }
      const hipError_t e1 = hipMemcpyAsync(v, x, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, st);
      const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(st) : e1;
      if (e2 != hipSuccess) {
        std::free(v);
        return hip_fail(e2, "eigenvector download");
      }
      out->value = lam;
      out->n = n;
      out->vector = v;
      return FEM2D_OK;
    }
    if (it == kMaxIter) break;
    // y = x - d with d ~ (A - target B)^-1 r: with d exact this is
    // (lambda - target) (A - target B)^-1 B x, one step of shift-invert inverse
    // iteration; its fixed points are eigenpairs whatever the inner accuracy
    if ((rc = inner_solve(F, S, target, h, r, dy, G, stats))) return rc;
    EIG_BLAS(rocblas_dcopy(h, ni, x, 1, y, 1));
    const double m_one = -1.0;
    EIG_BLAS(rocblas_daxpy(h, ni, &m_one, dy, 1, y, 1));
  }
  set_error("the shift-invert iteration did not converge (a target equidistant from two eigenvalues?)", 0);
  return FEM2D_ERR_NOT_CONVERGED;
}

}  // namespace

extern "C" {

int fem2d_gep_device_solve_sparse(fem2d_ctx* ctx, const fem2d_gep_device* gep, double target_eigenvalue,
                                  fem2d_eigenpair* out) {
  return sparse_solve(ctx, gep, target_eigenvalue, out);
}

int fem2d_solve_gep_sparse(fem2d_ctx* ctx, const fem2d_gep_csr* gep, double target_eigenvalue, fem2d_eigenpair* out) {
  return solve_host(ctx, gep, target_eigenvalue, out, sparse_solve);
}

}  // extern "C"
