// Dense generalized eigensolves of an assembled GEP, on the context's stream.
//
// fem2d_nalgebra_solve_gep replaces nalgebra_solve_gep (nalgebra_solve.rs:14-50)
// with its arithmetic: B = L L^T, C = B^-1 A, then the symmetric eigenproblem
// of C's lower triangle (C is not symmetric; nalgebra's SymmetricEigen reads
// only that triangle), the eigenvalue nearest the target and its unit vector.
// fem2d_solve_gep_dense solves A x = lambda B x itself (symmetric-definite),
// the problem slepc_solve_gep (slepc_solve.rs:14-68) hands to SLEPc with a
// target: shift-invert inverse iteration on the LU factors of A - target B,
// falling back to the whole spectrum (dsygvd) when the iteration is
// inconclusive.
//
// The upper-triangular CSR is expanded to full symmetric column-major matrices
// in HBM by k_expand (one wavefront per row); the factorisations, the product
// and the eigensolves are rocSOLVER / rocBLAS calls on the context's stream.
// Only the n eigenvalues and the chosen vector come back to the host.

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
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

constexpr double kSpurious = 1e-12;  // nalgebra_solve.rs:24

__global__ void k_zero(double* __restrict__ p, long long count) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count; i += stride) p[i] = 0.0;
}

__global__ void k_set_int(int* __restrict__ p, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// Full symmetric A and B (column-major, leading dimension n) from the stored
// upper triangle, one wavefront per row. An entry outside the upper triangle
// is not written and sets *err.
// V: the stored value type; res: FP32R residuals (fp32r.hpp), null otherwise.
template <class V>
__global__ void k_expand(long long n, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                         const V* __restrict__ a, const V* __restrict__ b, const uint32_t* __restrict__ res,
                         double* __restrict__ A, double* __restrict__ B, int* __restrict__ err) {
  const int lane = threadIdx.x % kWave;
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave;
  const long long waves = (gridDim.x * (long long)blockDim.x) / kWave;
  for (long long r = wave; r < n; r += waves) {
    const int64_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
    for (int64_t k = k0 + lane; k < k1; k += kWave) {
      const long long c = col[k];
      if (c < r || c >= n) {
        atomicOr(err, 1);
        continue;
      }
      double av = double(a[k]), bv = double(b[k]);
      if (res) {
        const uint32_t q = res[k];
        av = fem2d::fp32r_join(float(av), fem2d::fp32r_res_a(q));
        bv = fem2d::fp32r_join(float(bv), fem2d::fp32r_res_b(q));
      }
      A[r + c * n] = av;
      A[c + r * n] = av;
      B[r + c * n] = bv;
      B[c + r * n] = bv;
    }
  }
}

// M(j, i) = M(i, j) for i > j: the upper triangle from the lower one.
__global__ void k_lower_to_upper(long long n, double* __restrict__ M) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n * n; idx += stride) {
    const long long i = idx % n, j = idx / n;
    if (i > j) M[j + i * n] = M[i + j * n];
  }
}

// Dense A and B of a device GEP.
int expand(const fem2d_gep_device* g, hipStream_t st, DevMem& m, double** A, double** B) {
  const long long n = g->n;
  const bool f32r = g->precision == FEM2D_PRECISION_FP32R;
  const bool fp32 = g->precision == FEM2D_PRECISION_FP32 || f32r;  // FP32 value arrays
  if (g->precision != FEM2D_PRECISION_FP64 && !fp32) {
    set_error("unknown precision", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if (g->nnz < 0 || !g->row_ptr ||
      (g->nnz > 0 && (!g->col || (fp32 ? !g->a_val32 || !g->b_val32 : !g->a_val || !g->b_val) ||
                      (f32r && !g->res16)))) {
    set_error("GEP arrays missing", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  EIG_HIP(m.alloc(A, size_t(n * n)));
  EIG_HIP(m.alloc(B, size_t(n * n)));
  int* err = nullptr;
  EIG_HIP(m.alloc(&err, 1));
  k_zero<<<grid_for(n * n), kBlock, 0, st>>>(*A, n * n);
  EIG_HIP(hipGetLastError());
  k_zero<<<grid_for(n * n), kBlock, 0, st>>>(*B, n * n);
  EIG_HIP(hipGetLastError());
  k_set_int<<<1, kWave, 0, st>>>(err, 0);
  EIG_HIP(hipGetLastError());
  const unsigned row_blocks = unsigned(std::max<long long>(1, std::min<long long>((n + 3) / 4, 1 << 16)));
  if (fp32)
    k_expand<<<row_blocks, kBlock, 0, st>>>(n, g->row_ptr, g->col, g->a_val32, g->b_val32, f32r ? g->res16 : nullptr,
                                            *A, *B, err);
  else
    k_expand<<<row_blocks, kBlock, 0, st>>>(n, g->row_ptr, g->col, g->a_val, g->b_val,
                                            static_cast<const uint32_t*>(nullptr), *A, *B, err);
  EIG_HIP(hipGetLastError());
  int h_err = 0;
  EIG_HIP(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, st));
  EIG_HIP(hipStreamSynchronize(st));
  if (h_err) {
    set_error("GEP is not upper-triangular CSR with columns inside the matrix", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  return FEM2D_OK;
}

// The eigenpair whose value is nearest target (nalgebra_solve.rs:28-46): D
// holds the n eigenvalues in ascending order, column k of V (on the device,
// leading dimension n) the vector of D[k].
int pick(const std::vector<double>& D, const double* V, long long n, double target, hipStream_t st,
         fem2d_eigenpair* out) {
  if (std::all_of(D.begin(), D.end(), [](double e) { return std::fabs(e) < kSpurious; })) {
    set_error("Only spurious modes were found!", 0);
    return FEM2D_ERR_SPURIOUSLY_CONVERGED;
  }
  size_t best = 0;
  double delta_min = HUGE_VAL;
  for (size_t k = 0; k < D.size(); ++k) {
    const double delta = std::fabs(D[k] - target);
    if (delta < delta_min) {
      delta_min = delta;
      best = k;
    }
  }
  double* v = static_cast<double*>(std::malloc(size_t(n) * sizeof(double)));
  if (!v) {
    set_error("host allocation of the eigenvector failed", 0);
    return FEM2D_ERR_OOM;
  }
  const hipError_t e1 = hipMemcpyAsync(v, V + best * size_t(n), size_t(n) * sizeof(double), hipMemcpyDeviceToHost, st);
  const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(st) : e1;
  if (e2 != hipSuccess) {
    std::free(v);
    return hip_fail(e2, "eigenvector download");
  }
  out->value = D[best];
  out->n = n;
  out->vector = v;
  return FEM2D_OK;
}

int nalgebra_solve(fem2d_ctx* ctx, const fem2d_gep_device* gep, double target, fem2d_eigenpair* out) {
  if (int rc = check_device_args(ctx, gep, out)) return rc;
  const long long n = gep->n;
  if (n > FEM2D_NALGEBRA_MAX_DENSE_SIZE) {  // nalgebra_solve.rs:15-17
    set_error("Matrices Exceeded Maximum Size (1000x1000); Cannot Solve!", 0);
    return FEM2D_ERR_PROBLEM_TOO_LARGE;
  }
  if (n == 0) {  // no eigenvalues: all of none are spurious (nalgebra_solve.rs:24)
    set_error("Only spurious modes were found!", 0);
    return FEM2D_ERR_SPURIOUSLY_CONVERGED;
  }
  EIG_HIP(hipSetDevice(fem2d::capi::ctx_device(ctx)));
  hipStream_t st = static_cast<hipStream_t>(fem2d_ctx_stream(ctx));
  rocblas_handle h;
  if (int rc = solver_handle(ctx, st, &h)) return rc;
  DevMem m;
  double *A, *B, *C, *D, *E;
  if (int rc = expand(gep, st, m, &A, &B)) return rc;
  EIG_HIP(m.alloc(&C, size_t(n * n)));
  EIG_HIP(m.alloc(&D, size_t(n)));
  EIG_HIP(m.alloc(&E, size_t(n)));
  rocblas_int* d_info;
  EIG_HIP(m.alloc(&d_info, 1));
  rocblas_int info = 0;
  const rocblas_int ni = rocblas_int(n);
  // B = L L^T, then B^-1 from the factor (nalgebra_solve.rs:19-20)
  EIG_BLAS(rocsolver_dpotrf(h, rocblas_fill_lower, ni, B, ni, d_info));
  if (int rc = read_info(d_info, st, &info)) return rc;
  if (info != 0) {
    set_error("Failed to invert B-matrix (via cholesky); likely ill-conditioned!", int(info));
    return FEM2D_ERR_FAILED_TO_INVERT_B;
  }
  EIG_BLAS(rocsolver_dpotri(h, rocblas_fill_lower, ni, B, ni, d_info));
  if (int rc = read_info(d_info, st, &info)) return rc;
  if (info != 0) {
    set_error("Failed to invert B-matrix (via cholesky); likely ill-conditioned!", int(info));
    return FEM2D_ERR_FAILED_TO_INVERT_B;
  }
  k_lower_to_upper<<<grid_for(n * n), kBlock, 0, st>>>(n, B);
  EIG_HIP(hipGetLastError());
  // C = B^-1 A (nalgebra_solve.rs:21)
  const double one = 1.0, zero = 0.0;
  EIG_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, ni, ni, ni, &one, B, ni, A, ni, &zero, C,
                         ni));
  // SymmetricEigen::new(C) reads C's lower triangle (nalgebra_solve.rs:22)
  EIG_BLAS(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, ni, C, ni, D, E, d_info));
  if (int rc = read_info(d_info, st, &info)) return rc;
  if (info != 0) {
    set_error("the symmetric eigensolver did not converge", int(info));
    return FEM2D_ERR_NOT_CONVERGED;
  }
  std::vector<double> hD(static_cast<size_t>(n));
  EIG_HIP(hipMemcpyAsync(hD.data(), D, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, st));
  EIG_HIP(hipStreamSynchronize(st));
  return pick(hD, C, n, target, st, out);
}

// X = I (column-major n x n)
__global__ void k_identity(long long n, double* __restrict__ X) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n * n; i += stride)
    X[i] = i % n == i / n ? 1.0 : 0.0;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// The two substitutions of an LU solve K y = b (K = P L U, getrf's row
// interchanges applied to b first) as products with the inverse factors,
// stored transposed so that each output entry is one wavefront's dot product
// with a contiguous column: t = L^-1 P^T b with column c of X1 = (L^-1)^T
// (rows 0..c), then y = U^-1 t with column c of X2 = (U^-1)^T (rows c..n-1).
// src_of[r] is the entry of b that lands in row r after the interchanges.
template <bool kLower>
__global__ void k_lu_apply(long long n, const double* __restrict__ X, const int* __restrict__ src_of,
                           const double* __restrict__ d, const double* __restrict__ x, double* __restrict__ y) {
  const int lane = threadIdx.x % kWave;
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave;
  const long long waves = ((long long)gridDim.x * blockDim.x) / kWave;
  for (long long c = wave; c < n; c += waves) {
    const long long r0 = kLower ? 0 : c, r1 = kLower ? c + 1 : n;
    const double* __restrict__ col = X + c * n;
    double s0 = 0.0, s1 = 0.0;
    long long r = r0 + lane;
    // the lower pass gathers the interchanged, scaled right-hand side d[i] b[i]
    auto in = [&](long long q) {
      if (!kLower) return x[q];
      const int i = src_of[q];
      return d[i] * x[i];
    };
    for (; r + kWave < r1; r += 2 * kWave) {
      s0 += col[r] * in(r);
      s1 += col[r + kWave] * in(r + kWave);
    }
    for (; r < r1; r += kWave) s0 += col[r] * in(r);
    const double s = wave_sum(s0 + s1);
    if (lane == 0) y[c] = kLower ? s : d[c] * s;  // the upper pass scales its result back
  }
}

// d[i] = 1 / sqrt(|K(i, i)|) (1 where that is not a positive finite number)
__global__ void k_diag_scale(long long n, const double* __restrict__ K, double* __restrict__ d) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) {
    const double v = 1.0 / std::sqrt(std::fabs(K[i + i * n]));
    d[i] = v > 0.0 && v < HUGE_VAL ? v : 1.0;
  }
}

// S = D K D, D = diag(d)
__global__ void k_scale_sym(long long n, const double* __restrict__ K, const double* __restrict__ d,
                            double* __restrict__ S) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n * n; idx += stride) {
    const long long i = idx % n, j = idx / n;
    S[idx] = d[i] * K[idx] * d[j];
  }
}

// y[i] = d[i] x[i]
__global__ void k_scale_vec(long long n, const double* __restrict__ d, const double* __restrict__ x,
                            double* __restrict__ y) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) y[i] = d[i] * x[i];
}

// K = A - shift B, elementwise over the dense matrices
__global__ void k_shift(long long count, const double* __restrict__ A, const double* __restrict__ B, double shift,
                        double* __restrict__ K) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count; i += stride) K[i] = A[i] - shift * B[i];
}

// The eigenpair of A x = lambda B x nearest `target` by shift-invert inverse
// iteration: K = A - target B is LU-factored once; x <- K^-1 B x, B-normalised,
// converges to the eigenvector whose eigenvalue is nearest the target (the
// dominant one of K^-1 B), lambda = x^T A x. Returns 1 when the iteration is
// inconclusive (K singular, no convergence, or an eigenvalue so small that the
// spurious-mode test needs the whole spectrum), leaving A and B unchanged.
// The solves with K use its inverse LU factors (two trsm against the identity,
// once), each solve two bandwidth-bound passes over them (k_lu_apply); every
// solve is followed by one step of iterative refinement against K itself.
int inverse_iteration(rocblas_handle h, hipStream_t st, long long n, const double* A, const double* B, double* K,
                      double target, DevMem& m, double* value, std::vector<double>* vec) {
  const rocblas_int ni = rocblas_int(n);
  rocblas_int *ipiv, *d_info;
  double *x, *y, *bx, *ax, *r, *K0;
  EIG_HIP(m.alloc(&K0, size_t(n * n)));
  EIG_HIP(m.alloc(&r, size_t(n)));
  EIG_HIP(m.alloc(&ipiv, size_t(n)));
  EIG_HIP(m.alloc(&d_info, 1));
  EIG_HIP(m.alloc(&x, size_t(n)));
  EIG_HIP(m.alloc(&y, size_t(n)));
  EIG_HIP(m.alloc(&bx, size_t(n)));
  EIG_HIP(m.alloc(&ax, size_t(n)));
  k_shift<<<grid_for(n * n), kBlock, 0, st>>>(n * n, A, B, target, K0);
  EIG_HIP(hipGetLastError());
  // the factored matrix is D K D, D the inverse square roots of K's diagonal:
  // the hierarchical shape functions' scales differ by orders of magnitude
  double* dsc;
  EIG_HIP(m.alloc(&dsc, size_t(n)));
  k_diag_scale<<<grid_for(n), kBlock, 0, st>>>(n, K0, dsc);
  k_scale_sym<<<grid_for(n * n), kBlock, 0, st>>>(n, K0, dsc, K);
  EIG_HIP(hipGetLastError());
  EIG_BLAS(rocsolver_dgetrf(h, ni, ni, K, ni, ipiv, d_info));
  rocblas_int info = 0;
  if (int rc = read_info(d_info, st, &info)) return rc;
  if (info != 0) return 1;  // the target is an eigenvalue to working precision
  // the row interchanges as a gather map, and the transposed inverse factors
  // (up to kInverseMax DoFs: beyond, the two trsm cost more than the
  // substitutions they replace)
  constexpr long long kInverseMax = 2048;
  bool exact = n > kInverseMax;
  double *X1 = nullptr, *X2 = nullptr;
  int* src_of = nullptr;
  if (!exact) {
    EIG_HIP(m.alloc(&X1, size_t(n * n)));
    EIG_HIP(m.alloc(&X2, size_t(n * n)));
    EIG_HIP(m.alloc(&src_of, size_t(n)));
  }
  if (!exact) {
    std::vector<rocblas_int> piv(static_cast<size_t>(n));
    EIG_HIP(hipMemcpyAsync(piv.data(), ipiv, size_t(n) * sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
    EIG_HIP(hipStreamSynchronize(st));
    std::vector<int> src(static_cast<size_t>(n));
    for (long long i = 0; i < n; ++i) src[size_t(i)] = int(i);
    for (long long i = 0; i < n; ++i) {  // getrs applies the interchanges in order (1-based pivots)
      const long long p = piv[size_t(i)] - 1;
      if (p < 0 || p >= n) {
        set_error("getrf returned a pivot outside the matrix", 0);
        return FEM2D_ERR_HIP;
      }
      std::swap(src[size_t(i)], src[size_t(p)]);
    }
    EIG_HIP(hipMemcpyAsync(src_of, src.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice, st));
    EIG_HIP(hipStreamSynchronize(st));  // src is read before it goes out of scope
  }
  if (!exact) {
    k_identity<<<grid_for(n * n), kBlock, 0, st>>>(n, X1);
    EIG_HIP(hipGetLastError());
    k_identity<<<grid_for(n * n), kBlock, 0, st>>>(n, X2);
    EIG_HIP(hipGetLastError());
    const double one = 1.0;
    // L^T X1 = I and U^T X2 = I: X1 = (L^-1)^T, X2 = (U^-1)^T
    EIG_BLAS(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                           rocblas_diagonal_unit, ni, ni, &one, K, ni, X1, ni));
    EIG_BLAS(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                           rocblas_diagonal_non_unit, ni, ni, &one, K, ni, X2, ni));
  }
  double* t;
  EIG_HIP(m.alloc(&t, size_t(n)));
  const unsigned apply_grid = unsigned(std::max<long long>(1, std::min<long long>((n + 3) / 4, 1 << 16)));
  // y = K^-1 b = D (D K D)^-1 D b: the inverse factors, or (exact) getrs's
  // substitutions when those prove too inaccurate for this K
  auto lu_solve = [&](const double* b, double* y) -> int {
    if (!exact) {
      k_lu_apply<true><<<apply_grid, kBlock, 0, st>>>(n, X1, src_of, dsc, b, t);
      k_lu_apply<false><<<apply_grid, kBlock, 0, st>>>(n, X2, nullptr, dsc, t, y);
      EIG_HIP(hipGetLastError());
      return FEM2D_OK;
    }
    k_scale_vec<<<grid_for(n), kBlock, 0, st>>>(n, dsc, b, t);
    EIG_HIP(hipGetLastError());
    EIG_BLAS(rocsolver_dgetrs(h, rocblas_operation_none, ni, 1, K, ni, ipiv, t, ni));
    k_scale_vec<<<grid_for(n), kBlock, 0, st>>>(n, dsc, t, y);
    EIG_HIP(hipGetLastError());
    return FEM2D_OK;
  };
  const std::vector<double> x0 = start_vector(n);
  EIG_HIP(hipMemcpyAsync(x, x0.data(), size_t(n) * sizeof(double), hipMemcpyHostToDevice, st));
  EIG_HIP(hipStreamSynchronize(st));  // x0 is read before any return below
  const double one = 1.0, zero = 0.0;
  EIG_BLAS(rocblas_dgemv(h, rocblas_operation_none, ni, ni, &one, B, ni, x, 1, &zero, bx, 1));
  // converged: a relative residual |A x - lambda B x| / |A x| of kTol, or of
  // kTolStall once the Rayleigh quotient no longer moves (the residual's
  // rounding floor grows with the conditioning of A and B)
  constexpr int kMaxIter = 100, kRefine = 6;
  constexpr double kTol = 1e-12, kTolStall = 1e-10, kSolveTol = 1e-8;
  double lam_prev = HUGE_VAL;
  for (int it = 0; it < kMaxIter; ++it) {
    // y = K^-1 B x, then iterative refinement against the unfactored K while
    // the residual B x - K y keeps shrinking (at most kRefine steps): the
    // products with the inverse factors leave errors of the size of their
    // conditioning. A solve the refinement leaves inaccurate (the first
    // iteration shows it) switches every later solve to getrs.
    const double m_one = -1.0;
    double nb = 0.0;
    if (!exact) EIG_BLAS(rocblas_dnrm2(h, ni, bx, 1, &nb));
    for (int attempt = 0; attempt < 2; ++attempt) {
      if (int rc = lu_solve(bx, y)) return rc;
      double nres = HUGE_VAL, nr_prev = HUGE_VAL;
      // getrs: one step and no residual norms, its substitutions being backward stable
      const int steps = exact ? 1 : kRefine;
      for (int k = 0; k <= steps; ++k) {
        if (exact && k == steps) break;
        EIG_BLAS(rocblas_dcopy(h, ni, bx, 1, r, 1));
        EIG_BLAS(rocblas_dgemv(h, rocblas_operation_none, ni, ni, &m_one, K0, ni, y, 1, &one, r, 1));  // B x - K y
        if (!exact) {
          EIG_BLAS(rocblas_dnrm2(h, ni, r, 1, &nres));
          if (k == steps || !(nres > 1e-15 * nb) || !(nres < 0.5 * nr_prev)) break;
          nr_prev = nres;
        }
        if (int rc = lu_solve(r, ax)) return rc;  // the correction (ax is free until below)
        EIG_BLAS(rocblas_daxpy(h, ni, &one, ax, 1, y, 1));
      }
      if (exact || it > 0 || nres <= kSolveTol * nb) break;
      if (verbose())
        std::fprintf(stderr, "[fem2d dense] n=%lld inverse factors leave a residual %.3e: getrs from here on\n", n,
                     nres / nb);
      exact = true;
    }
    EIG_BLAS(rocblas_dgemv(h, rocblas_operation_none, ni, ni, &one, B, ni, y, 1, &zero, bx, 1));
    double s = 0.0;
    EIG_BLAS(rocblas_ddot(h, ni, y, 1, bx, 1, &s));
    if (!(s > 0.0) || !std::isfinite(s)) return 1;
    const double inv = 1.0 / std::sqrt(s);
    EIG_BLAS(rocblas_dcopy(h, ni, y, 1, x, 1));
    EIG_BLAS(rocblas_dscal(h, ni, &inv, x, 1));   // x^T B x = 1
    EIG_BLAS(rocblas_dscal(h, ni, &inv, bx, 1));  // B x
    EIG_BLAS(rocblas_dgemv(h, rocblas_operation_none, ni, ni, &one, A, ni, x, 1, &zero, ax, 1));
    double lam = 0.0, na = 0.0, nr = 0.0;
    EIG_BLAS(rocblas_ddot(h, ni, x, 1, ax, 1, &lam));
    EIG_BLAS(rocblas_dnrm2(h, ni, ax, 1, &na));
    const double neg = -lam;
    EIG_BLAS(rocblas_dcopy(h, ni, ax, 1, y, 1));
    EIG_BLAS(rocblas_daxpy(h, ni, &neg, bx, 1, y, 1));  // A x - lambda B x
    EIG_BLAS(rocblas_dnrm2(h, ni, y, 1, &nr));
    if (!std::isfinite(lam) || !std::isfinite(nr)) return 1;
    const bool stalled = std::fabs(lam - lam_prev) <= 1e-14 * std::fabs(lam);
    lam_prev = lam;
    if (verbose())
      std::fprintf(stderr, "[fem2d dense] n=%lld it=%d lambda=%.17g residual=%.3e\n", n, it, lam, nr / na);
    if (nr <= kTol * na || (stalled && nr <= kTolStall * na)) {
      if (std::fabs(lam) < kSpurious) return 1;
      vec->resize(static_cast<size_t>(n));
      EIG_HIP(hipMemcpyAsync(vec->data(), x, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, st));
      EIG_HIP(hipStreamSynchronize(st));
      *value = lam;
      return 0;
    }
  }
  return 1;
}

int dense_solve(fem2d_ctx* ctx, const fem2d_gep_device* gep, double target, fem2d_eigenpair* out) {
  if (int rc = check_device_args(ctx, gep, out)) return rc;
  const long long n = gep->n;
  if (n > FEM2D_DENSE_GEP_MAX_SIZE) {
    set_error("Matrices exceed the dense solver's maximum size (32768 DoFs)", 0);
    return FEM2D_ERR_PROBLEM_TOO_LARGE;
  }
  if (n == 0) {
    set_error("Only spurious modes were found!", 0);
    return FEM2D_ERR_SPURIOUSLY_CONVERGED;
  }
  EIG_HIP(hipSetDevice(fem2d::capi::ctx_device(ctx)));
  hipStream_t st = static_cast<hipStream_t>(fem2d_ctx_stream(ctx));
  rocblas_handle h;
  if (int rc = solver_handle(ctx, st, &h)) return rc;
  DevMem m;
  double *A, *B, *C, *D, *E;
  if (int rc = expand(gep, st, m, &A, &B)) return rc;
  EIG_HIP(m.alloc(&C, size_t(n * n)));
  EIG_HIP(m.alloc(&D, size_t(n)));
  EIG_HIP(m.alloc(&E, size_t(n)));
  rocblas_int* d_info;
  EIG_HIP(m.alloc(&d_info, 1));
  rocblas_int info = 0;
  const rocblas_int ni = rocblas_int(n);
  // B positive definite: its Cholesky factorisation (of a copy) succeeds, the
  // condition the generalized solver checks first
  EIG_HIP(hipMemcpyAsync(C, B, size_t(n * n) * sizeof(double), hipMemcpyDeviceToDevice, st));
  EIG_BLAS(rocsolver_dpotrf(h, rocblas_fill_upper, ni, C, ni, d_info));
  if (int rc = read_info(d_info, st, &info)) return rc;
  if (info != 0) {
    set_error("Failed to invert B-matrix (via cholesky); likely ill-conditioned!", int(info));
    return FEM2D_ERR_FAILED_TO_INVERT_B;
  }
  {
    double lam = 0.0;
    std::vector<double> vec;
    const int it = inverse_iteration(h, st, n, A, B, C, target, m, &lam, &vec);
    if (it < 0 || it > 1) return it;  // a HIP or rocBLAS failure (status codes are >= 100)
    if (it == 0) {
      double* v = static_cast<double*>(std::malloc(size_t(n) * sizeof(double)));
      if (!v) {
        set_error("host allocation of the eigenvector failed", 0);
        return FEM2D_ERR_OOM;
      }
      std::memcpy(v, vec.data(), size_t(n) * sizeof(double));
      out->value = lam;
      out->n = n;
      out->vector = v;
      return FEM2D_OK;
    }
  }
  // the whole spectrum (dsygvd): inverse iteration was inconclusive
  EIG_BLAS(rocsolver_dsygvd(h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, ni, A, ni, B, ni, D, E,
                            d_info));
  if (int rc = read_info(d_info, st, &info)) return rc;
  if (info > ni) {
    set_error("Failed to invert B-matrix (via cholesky); likely ill-conditioned!", int(info - ni));
    return FEM2D_ERR_FAILED_TO_INVERT_B;
  }
  if (info != 0) {
    set_error("the generalized eigensolver did not converge", int(info));
    return FEM2D_ERR_NOT_CONVERGED;
  }
  std::vector<double> hD(static_cast<size_t>(n));
  EIG_HIP(hipMemcpyAsync(hD.data(), D, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, st));
  EIG_HIP(hipStreamSynchronize(st));
  return pick(hD, A, n, target, st, out);
}

}  // namespace

void fem2d::capi::release_solver(void* handle) {
  if (handle) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(handle));
}

extern "C" {

void fem2d_eigenpair_free(fem2d_eigenpair* pair) {
  if (!pair) return;
  std::free(pair->vector);
  pair->vector = nullptr;
  pair->n = 0;
}

int fem2d_gep_device_nalgebra_solve(fem2d_ctx* ctx, const fem2d_gep_device* gep, double target_eigenvalue,
                                    fem2d_eigenpair* out) {
  return nalgebra_solve(ctx, gep, target_eigenvalue, out);
}

int fem2d_nalgebra_solve_gep(fem2d_ctx* ctx, const fem2d_gep_csr* gep, double target_eigenvalue,
                             fem2d_eigenpair* out) {
  if (gep && gep->n > FEM2D_NALGEBRA_MAX_DENSE_SIZE && out) {  // before any upload, as the reference
    out->value = 0.0;
    out->n = 0;
    out->vector = nullptr;
    set_error("Matrices Exceeded Maximum Size (1000x1000); Cannot Solve!", 0);
    return FEM2D_ERR_PROBLEM_TOO_LARGE;
  }
  return solve_host(ctx, gep, target_eigenvalue, out, nalgebra_solve);
}

int fem2d_gep_device_solve_dense(fem2d_ctx* ctx, const fem2d_gep_device* gep, double target_eigenvalue,
                                 fem2d_eigenpair* out) {
  return dense_solve(ctx, gep, target_eigenvalue, out);
}

int fem2d_solve_gep_dense(fem2d_ctx* ctx, const fem2d_gep_csr* gep, double target_eigenvalue, fem2d_eigenpair* out) {
  if (gep && gep->n > FEM2D_DENSE_GEP_MAX_SIZE && out) {
    out->value = 0.0;
    out->n = 0;
    out->vector = nullptr;
    set_error("Matrices exceed the dense solver's maximum size (32768 DoFs)", 0);
    return FEM2D_ERR_PROBLEM_TOO_LARGE;
  }
  return solve_host(ctx, gep, target_eigenvalue, out, dense_solve);
}

}  // extern "C"
