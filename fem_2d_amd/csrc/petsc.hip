// PETSc AIJ binary files of an assembled GEP: the on-disk input of SLEPc.
//
// Replaces GEP::print_to_petsc_binary_files (linalg.rs:44-52), which converts
// each SparseMatrix to an AIJMatrixBinary (sparse_matrix.rs:184-223) and writes
// it with print_to_petsc_binary_file (sparse_matrix.rs:225-265). The file holds
// the full symmetric matrix (both triangles), rows in order and columns
// ascending within a row, as big-endian words:
//   u32 1211216 (PETSc MAT_FILE_CLASSID), u32 rows, u32 columns, u32 nnz,
//   u32 row lengths [rows], u32 column indices [nnz], f64 values [nnz].
//
// Device algorithm (integer/byte work, bound by HBM):
//  1. per stored upper entry: key = its column if it lies strictly above the
//     diagonal, else n; payload = its index;
//  2. a stable radix sort by key lists the strict upper entries by column with
//     their rows ascending, i.e. the lower triangle row by row;
//  3. full row r = its lower part (sorted entries with key r) followed by its
//     upper row, at offset lower_start[r] + row_ptr[r] (no scan needed);
//  4. the big-endian words are produced by the kernels; A and B share the
//     pattern, which is built once.

#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <utility>
#include <string>
#include <vector>

#include "../../include/fem2d.h"
#include "fp32r.hpp"
#include "host/capi_internal.hpp"
#include "pinned_staging.hpp"

namespace {

using fem2d::capi::set_error;

constexpr uint32_t kMatFileClassId = 1211216;
constexpr int kWave = 64;

__device__ __forceinline__ uint32_t be32(uint32_t x) { return __builtin_bswap32(x); }
__device__ __forceinline__ uint64_t be64(double v) { return __builtin_bswap64(__double_as_longlong(v)); }

// Sort keys of the stored entries (one wavefront per row); flags entries below
// the diagonal or outside the matrix.
__global__ void k_petsc_keys(long long n, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                             uint32_t* __restrict__ key, uint32_t* __restrict__ idx, uint32_t* __restrict__ row_of,
                             int* __restrict__ err) {
  const int lane = threadIdx.x % kWave;
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave;
  const long long waves = (gridDim.x * (long long)blockDim.x) / kWave;
  for (long long r = wave; r < n; r += waves) {
    const int64_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
    for (int64_t k = k0 + lane; k < k1; k += kWave) {
      const int32_t c = col[k];
      if (c < r || c >= n) atomicOr(err, 1);
      key[k] = c > r && c < n ? uint32_t(c) : uint32_t(n);
      idx[k] = uint32_t(k);
      row_of[k] = uint32_t(r);
    }
  }
}

// Per row: start of its lower part among the sorted strict entries, full row
// length (big-endian, for the file) and full row offset.
__global__ void k_petsc_rows(long long n, long long nnz, const int64_t* __restrict__ row_ptr,
                             const uint32_t* __restrict__ skey, int64_t* __restrict__ lower_start,
                             uint32_t* __restrict__ counts_be) {
  const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (r > n) return;
  // first sorted position with key >= r
  long long lo = 0, hi = nnz;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (skey[mid] < uint32_t(r)) lo = mid + 1; else hi = mid;
  }
  lower_start[r] = lo;
  if (r == n) return;
  long long lo2 = lo, hi2 = nnz;  // first position with key >= r + 1
  while (lo2 < hi2) {
    const long long mid = (lo2 + hi2) >> 1;
    if (skey[mid] < uint32_t(r + 1)) lo2 = mid + 1; else hi2 = mid;
  }
  counts_be[r] = be32(uint32_t((lo2 - lo) + (row_ptr[r + 1] - row_ptr[r])));
}

// Full rows in big-endian words: lower part from the sorted strict entries,
// then the stored upper row (one wavefront per row).
// V: the GEP's stored value type (FP32 values are widened to the format's f64);
// res: FP32R residuals (fp32r.hpp), null otherwise
template <class V>
__global__ void k_petsc_fill(long long n, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                             const V* __restrict__ a, const V* __restrict__ b, const uint32_t* __restrict__ res,
                             const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ row_of,
                             const int64_t* __restrict__ lower_start, uint32_t* __restrict__ col_be,
                             uint64_t* __restrict__ a_be, uint64_t* __restrict__ b_be) {
  const int lane = threadIdx.x % kWave;
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave;
  const long long waves = (gridDim.x * (long long)blockDim.x) / kWave;
  for (long long r = wave; r < n; r += waves) {
    const int64_t ls = lower_start[r], nl = lower_start[r + 1] - ls;
    const int64_t k0 = row_ptr[r], nu = row_ptr[r + 1] - k0;
    const int64_t out0 = ls + k0;
    for (int64_t j = lane; j < nl + nu; j += kWave) {
      uint32_t c;
      int64_t k;
      if (j < nl) {
        k = sidx[ls + j];
        c = row_of[k];
      } else {
        k = k0 + (j - nl);
        c = uint32_t(col[k]);
      }
      col_be[out0 + j] = be32(c);
      double av = double(a[k]), bv = double(b[k]);
      if (res) {
        const uint32_t rr = res[k];
        av = fem2d::fp32r_join(float(av), fem2d::fp32r_res_a(rr));
        bv = fem2d::fp32r_join(float(bv), fem2d::fp32r_res_b(rr));
      }
      a_be[out0 + j] = be64(av);
      b_be[out0 + j] = be64(bv);
    }
  }
}

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

int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e), 0);
  return e == hipErrorOutOfMemory ? FEM2D_ERR_OOM : FEM2D_ERR_HIP;
}

#define PETSC_TRY(expr)                                \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_fail(_e, #expr);  \
  } while (0)

// One AIJ file: the header, then the row lengths, column indices and values,
// each streamed from the device arrays the kernels produced.
int write_file(const std::string& path, uint32_t n, uint32_t nnz, const uint32_t* counts_be, const uint32_t* col_be,
               const uint64_t* val_be, fem2d::PinnedStaging& ps) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    set_error("cannot create " + path, 0);
    return FEM2D_ERR_IO;
  }
  auto to_file = [&](const unsigned char* p, size_t, size_t len) -> int {
    if (std::fwrite(p, 1, len, f) == len) return FEM2D_OK;
    set_error("short write to " + path, 0);
    return FEM2D_ERR_IO;
  };
  const uint32_t header[4] = {__builtin_bswap32(kMatFileClassId), __builtin_bswap32(n), __builtin_bswap32(n),
                              __builtin_bswap32(nnz)};
  int rc = to_file(reinterpret_cast<const unsigned char*>(header), 0, sizeof(header));
  const std::pair<const void*, size_t> sections[3] = {
      {counts_be, size_t(n) * 4}, {col_be, size_t(nnz) * 4}, {val_be, size_t(nnz) * 8}};
  for (const auto& sec : sections) {
    if (rc) break;
    hipError_t e = hipSuccess;
    rc = ps.run(sec.first, sec.second, to_file, &e);
    if (rc == -1) rc = hip_fail(e, "PETSc file staging copy");
  }
  if (std::fclose(f) != 0 && !rc) {
    set_error("short write to " + path, 0);
    rc = FEM2D_ERR_IO;
  }
  return rc;
}

}  // namespace

extern "C" {

int fem2d_gep_device_print_to_petsc_binary_files(fem2d_ctx* ctx, const fem2d_gep_device* gep, const char* dir,
                                                 const char* prefix) {
  if (!ctx || !gep || !dir || !prefix) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if (int rc = fem2d::capi::check_gep_handle(ctx, gep)) return rc;
  const long long n = gep->n, nnz = gep->nnz;
  if (n < 0 || nnz < 0 || n >= 0xffffffffLL || 2 * nnz >= 0xffffffffLL) {
    set_error("matrix too large for the PETSc binary format's 32-bit sizes", 0);
    return FEM2D_ERR_UNSUPPORTED;
  }
  {
    const int p = gep->precision;
    const bool f32r = p == FEM2D_PRECISION_FP32R, f32 = p == FEM2D_PRECISION_FP32 || f32r;
    if (p != FEM2D_PRECISION_FP64 && !f32) {
      set_error("unknown precision", 0);
      return FEM2D_ERR_INVALID_ARG;
    }
    if (!gep->row_ptr || (nnz > 0 && (!gep->col || (f32 ? !gep->a_val32 || !gep->b_val32 : !gep->a_val || !gep->b_val) ||
                                      (f32r && !gep->res16)))) {
      set_error("GEP arrays missing", 0);
      return FEM2D_ERR_INVALID_ARG;
    }
  }
  PETSC_TRY(hipSetDevice(fem2d::capi::ctx_device(ctx)));
  hipStream_t st = static_cast<hipStream_t>(fem2d_ctx_stream(ctx));
  DevMem m;
  uint32_t *key, *skey, *idx, *sidx, *row_of, *counts_be, *col_be;
  uint64_t *a_be, *b_be;
  int64_t* lower_start;
  int* err;
  PETSC_TRY(m.alloc(&key, size_t(nnz)));
  PETSC_TRY(m.alloc(&skey, size_t(nnz)));
  PETSC_TRY(m.alloc(&idx, size_t(nnz)));
  PETSC_TRY(m.alloc(&sidx, size_t(nnz)));
  PETSC_TRY(m.alloc(&row_of, size_t(nnz)));
  PETSC_TRY(m.alloc(&lower_start, size_t(n + 1)));
  PETSC_TRY(m.alloc(&counts_be, size_t(n)));
  PETSC_TRY(m.alloc(&err, 1));
  const int zero = 0;
  PETSC_TRY(hipMemcpy(err, &zero, sizeof(int), hipMemcpyHostToDevice));

  const int block = 256;
  const unsigned row_waves = unsigned(std::min<long long>((n + 3) / 4, 1 << 16));
  if (n > 0) {
    k_petsc_keys<<<std::max(1u, row_waves), block, 0, st>>>(n, gep->row_ptr, gep->col, key, idx, row_of, err);
    PETSC_TRY(hipGetLastError());
  }
  int end_bit = 1;
  while (end_bit < 32 && (1LL << end_bit) <= n) ++end_bit;
  size_t tmp_bytes = 0;
  PETSC_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, skey, idx, sidx, int(nnz), 0, end_bit, st));
  void* tmp = nullptr;
  PETSC_TRY(m.alloc(reinterpret_cast<unsigned char**>(&tmp), tmp_bytes));
  PETSC_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, skey, idx, sidx, int(nnz), 0, end_bit, st));
  k_petsc_rows<<<unsigned((n + 1 + block - 1) / block), block, 0, st>>>(n, nnz, gep->row_ptr, skey, lower_start,
                                                                        counts_be);
  PETSC_TRY(hipGetLastError());
  int h_err = 0;
  long long n_strict = 0;
  PETSC_TRY(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, st));
  PETSC_TRY(hipMemcpyAsync(&n_strict, lower_start + n, sizeof(long long), hipMemcpyDeviceToHost, st));
  PETSC_TRY(hipStreamSynchronize(st));
  if (h_err) {
    set_error("GEP is not upper-triangular CSR with columns inside the matrix", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  const long long full = nnz + n_strict;
  PETSC_TRY(m.alloc(&col_be, size_t(full)));
  PETSC_TRY(m.alloc(&a_be, size_t(full)));
  PETSC_TRY(m.alloc(&b_be, size_t(full)));
  if (n > 0) {
    if (gep->precision != FEM2D_PRECISION_FP64)
      k_petsc_fill<<<std::max(1u, row_waves), block, 0, st>>>(
          n, gep->row_ptr, gep->col, gep->a_val32, gep->b_val32,
          gep->precision == FEM2D_PRECISION_FP32R ? gep->res16 : nullptr, sidx, row_of, lower_start, col_be, a_be, b_be);
    else
      k_petsc_fill<<<std::max(1u, row_waves), block, 0, st>>>(n, gep->row_ptr, gep->col, gep->a_val, gep->b_val,
                                                              static_cast<const uint32_t*>(nullptr), sidx, row_of,
                                                              lower_start, col_be, a_be, b_be);
    PETSC_TRY(hipGetLastError());
  }
  PETSC_TRY(hipStreamSynchronize(st));
  fem2d::PinnedStaging ps;
  PETSC_TRY(ps.init(st));
  // linalg.rs:49-50: "{dir}/tmp/{prefix}_a.dat" and "{dir}/tmp/{prefix}_b.dat"
  const std::string base = std::string(dir) + "/tmp/" + prefix;
  if (int rc = write_file(base + "_a.dat", uint32_t(n), uint32_t(full), counts_be, col_be, a_be, ps)) return rc;
  return write_file(base + "_b.dat", uint32_t(n), uint32_t(full), counts_be, col_be, b_be, ps);
}

int fem2d_gep_csr_print_to_petsc_binary_files(fem2d_ctx* ctx, const fem2d_gep_csr* gep, const char* dir,
                                              const char* prefix) {
  if (!ctx || !gep || !dir || !prefix) {
    set_error("null argument", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  if (gep->n < 0 || gep->nnz < 0) {
    set_error("negative GEP sizes", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  PETSC_TRY(hipSetDevice(fem2d::capi::ctx_device(ctx)));
  DevMem m;
  fem2d_gep_device d{};
  d.n = gep->n;
  d.nnz = gep->nnz;
  PETSC_TRY(m.alloc(&d.row_ptr, size_t(gep->n + 1)));
  PETSC_TRY(m.alloc(&d.col, size_t(gep->nnz)));
  PETSC_TRY(m.alloc(&d.a_val, size_t(gep->nnz)));
  PETSC_TRY(m.alloc(&d.b_val, size_t(gep->nnz)));
  PETSC_TRY(hipMemcpy(d.row_ptr, gep->row_ptr, size_t(gep->n + 1) * 8, hipMemcpyHostToDevice));
  if (gep->nnz > 0) {
    PETSC_TRY(hipMemcpy(d.col, gep->col, size_t(gep->nnz) * 4, hipMemcpyHostToDevice));
    PETSC_TRY(hipMemcpy(d.a_val, gep->a_val, size_t(gep->nnz) * 8, hipMemcpyHostToDevice));
    PETSC_TRY(hipMemcpy(d.b_val, gep->b_val, size_t(gep->nnz) * 8, hipMemcpyHostToDevice));
  }
  // the host CSR must be consistent before the device reads rows by row_ptr
  for (long long r = 0; r < gep->n; ++r)
    if (gep->row_ptr[r + 1] < gep->row_ptr[r]) {
      set_error("row_ptr decreases", 0);
      return FEM2D_ERR_INVALID_ARG;
    }
  if (gep->row_ptr[0] != 0 || gep->row_ptr[gep->n] != gep->nnz) {
    set_error("row_ptr does not span [0, nnz]", 0);
    return FEM2D_ERR_INVALID_ARG;
  }
  return fem2d_gep_device_print_to_petsc_binary_files(ctx, &d, dir, prefix);
}

}  // extern "C"
