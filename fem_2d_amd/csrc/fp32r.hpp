// FP32R storage of an FP64 value: an FP32 value plus a 16-bit residual.
//
//   hi = fl32(x)                      (round to nearest)
//   q  = rint((x - hi) * 2^(38 - e))  with e the binary exponent of hi
//   x' = hi + q * 2^(e - 38)
//
// x - hi is exact in FP64 and at most half an FP32 ulp (2^(e - 24)), so
// |q| <= 2^14 fits a signed 16-bit field, and |x' - x| <= 2^(e - 39) <=
// 2^-39 |x|: about 39 significant bits in 6 bytes. hi alone is the value
// rounded to FP32, so an FP32 consumer reads it as it is.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

namespace fem2d {

// binary exponent e of an FP32 value (hi = m 2^e, 1 <= m < 2); subnormals and
// zero take the exponent of the smallest normal, whose ulp they share
__host__ __device__ inline int fp32r_exponent(float hi) {
  uint32_t b;
  std::memcpy(&b, &hi, sizeof b);
  const int eb = int((b >> 23) & 0xffu);
  return (eb == 0 ? 1 : eb) - 127;
}

__host__ __device__ inline void fp32r_split(double x, float& hi, int& q) {
  hi = float(x);
  const double d = x - double(hi);
  const double s = ldexp(d, 38 - fp32r_exponent(hi));
  q = int(rint(s));
  q = q > 32767 ? 32767 : (q < -32768 ? -32768 : q);
}

__host__ __device__ inline double fp32r_join(float hi, int q) {
  return double(hi) + ldexp(double(q), fp32r_exponent(hi) - 38);
}

// the two residuals of one stored entry in one word: A's in the low half, B's
// in the high half
__host__ __device__ inline uint32_t fp32r_pack(int qa, int qb) {
  return uint32_t(uint16_t(int16_t(qa))) | (uint32_t(uint16_t(int16_t(qb))) << 16);
}
__host__ __device__ inline int fp32r_res_a(uint32_t r) { return int(int16_t(uint16_t(r & 0xffffu))); }
__host__ __device__ inline int fp32r_res_b(uint32_t r) { return int(int16_t(uint16_t(r >> 16))); }

}  // namespace fem2d
