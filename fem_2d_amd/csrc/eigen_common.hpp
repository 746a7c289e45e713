// Helpers shared by the dense and sparse eigensolves of a GEP: device
// allocations freed with the call, status mapping of HIP and rocBLAS
// failures, the context's rocBLAS handle, argument checks, host CSR upload.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fem2d.h"
#include "host/capi_internal.hpp"

namespace fem2d {
namespace eig {

using fem2d::capi::set_error;

constexpr int kWave = 64;
constexpr int kBlock = 256;

// progress lines on stderr when FEM2D_EIGEN_VERBOSE is set (to anything but 0)
inline bool verbose() {
  static const bool v = [] {
    const char* e = std::getenv("FEM2D_EIGEN_VERBOSE");
    return e && *e && std::strcmp(e, "0") != 0;
  }();
  return v;
}

inline unsigned grid_for(long long count) {
  return unsigned(std::max<long long>(1, std::min<long long>((count + kBlock - 1) / kBlock, 1 << 16)));
}

// device allocations of one call, freed when it returns
struct DevMem {
  std::vector<void*> ptrs;
  ~DevMem() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t alloc(T** p, size_t count) {
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, count ? count * sizeof(T) : 8);
    if (e == hipSuccess) ptrs.push_back(q);
    *p = static_cast<T*>(q);
    return e;
  }
};

inline int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e), 0);
  return e == hipErrorOutOfMemory ? FEM2D_ERR_OOM : FEM2D_ERR_HIP;
}

inline int blas_fail(rocblas_status s, const char* what) {
  set_error(std::string(what) + ": " + rocblas_status_to_string(s), 0);
  return s == rocblas_status_memory_error ? FEM2D_ERR_OOM : FEM2D_ERR_HIP;
}

#define EIG_HIP(expr)                                 \
  do {                                                \
    hipError_t _e = (expr);                           \
    if (_e != hipSuccess) return hip_fail(_e, #expr); \
  } while (0)

#define EIG_BLAS(expr)                                               \
  do {                                                               \
    rocblas_status _s = (expr);                                      \
    if (_s != rocblas_status_success) return blas_fail(_s, #expr);   \
  } while (0)

// The context's rocBLAS handle (made on first use, bound to the context's stream).
inline int solver_handle(fem2d_ctx* ctx, hipStream_t st, rocblas_handle* out) {
  void** slot = fem2d::capi::ctx_solver_slot(ctx);
  if (!*slot) {
    rocblas_handle h = nullptr;
    EIG_BLAS(rocblas_create_handle(&h));
    *slot = h;
  }
  *out = static_cast<rocblas_handle>(*slot);
  EIG_BLAS(rocblas_set_stream(*out, st));
  return FEM2D_OK;
}

inline int read_info(const rocblas_int* d_info, hipStream_t st, rocblas_int* h_info) {
  EIG_HIP(hipMemcpyAsync(h_info, d_info, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  EIG_HIP(hipStreamSynchronize(st));
  return FEM2D_OK;
}

inline int check_device_args(fem2d_ctx* ctx, const fem2d_gep_device* gep, fem2d_eigenpair* out) {
  if (!ctx || !gep || !out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  out->value = 0.0;
  out->n = 0;
  out->vector = nullptr;
  if (int rc = fem2d::capi::check_gep_handle(ctx, gep)) return rc;
  if (gep->n < 0) {
    set_error("negative GEP size", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  out->value = 0.0;
  out->n = 0;
  out->vector = nullptr;
  return FEM2D_OK;
}

// A host CSR on the device (checked on the host before any device read).
inline int upload(const fem2d_gep_csr* gep, DevMem& m, fem2d_gep_device* d) {
  if (gep->n < 0 || gep->nnz < 0 || !gep->row_ptr || (gep->nnz > 0 && (!gep->col || !gep->a_val || !gep->b_val))) {
    set_error("GEP sizes or arrays invalid", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  for (long long r = 0; r < gep->n; ++r)
    if (gep->row_ptr[r + 1] < gep->row_ptr[r]) {
      set_error("row_ptr decreases", 0);
      return FEM2D_ERR_INVALID_ARG;
    }
  if (gep->row_ptr[0] != 0 || gep->row_ptr[gep->n] != gep->nnz) {
    set_error("row_ptr does not span [0, nnz]", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  *d = fem2d_gep_device{};
  d->n = gep->n;
  d->nnz = gep->nnz;
  d->precision = FEM2D_PRECISION_FP64;
  EIG_HIP(m.alloc(&d->row_ptr, size_t(gep->n + 1)));
  EIG_HIP(m.alloc(&d->col, size_t(gep->nnz)));
  EIG_HIP(m.alloc(&d->a_val, size_t(gep->nnz)));
  EIG_HIP(m.alloc(&d->b_val, size_t(gep->nnz)));
  EIG_HIP(hipMemcpy(d->row_ptr, gep->row_ptr, size_t(gep->n + 1) * 8, hipMemcpyHostToDevice));
  if (gep->nnz > 0) {
    EIG_HIP(hipMemcpy(d->col, gep->col, size_t(gep->nnz) * 4, hipMemcpyHostToDevice));
    EIG_HIP(hipMemcpy(d->a_val, gep->a_val, size_t(gep->nnz) * 8, hipMemcpyHostToDevice));
    EIG_HIP(hipMemcpy(d->b_val, gep->b_val, size_t(gep->nnz) * 8, hipMemcpyHostToDevice));
  }
  return FEM2D_OK;
}

template <class Solve>
int solve_host(fem2d_ctx* ctx, const fem2d_gep_csr* gep, double target, fem2d_eigenpair* out, Solve solve) {
  if (!ctx || !gep || !out) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  EIG_HIP(hipSetDevice(fem2d::capi::ctx_device(ctx)));
  DevMem m;
  fem2d_gep_device d;
  if (int rc = upload(gep, m, &d)) return rc;
  return solve(ctx, &d, target, out);
}


// a fixed start vector of n components in [0.5, 1.5) (SplitMix64)
inline std::vector<double> start_vector(long long n) {
  std::vector<double> x0(static_cast<size_t>(n));
  unsigned long long state = 0x9E3779B97F4A7C15ull;
  for (auto& v : x0) {
    unsigned long long z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    v = 0.5 + double(z >> 11) * 0x1.0p-53;
  }
  return x0;
}

}  // namespace eig
}  // namespace fem2d
