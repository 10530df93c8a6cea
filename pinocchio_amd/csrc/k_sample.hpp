// k_sample.hpp -- on-device generation of seeded batch inputs (q, v, a, tau) with the
// same counter-based stream as the host sampler (pinocchio_amd/models.py): draw d of a
// stream with seed s is splitmix64's output for the counter s + (d + 1) * 0x9E3779B97F4A7C15,
// so every element is computed independently from its draw index.  Values are formed in
// FP64 and rounded to FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbd {
namespace sample {

__device__ __forceinline__ uint64_t draw(uint64_t seed, uint64_t d) {
  uint64_t z = seed + (d + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unit(uint64_t seed, uint64_t d) {
  return (double)(draw(seed, d) >> 11) * (1.0 / 9007199254740992.0);
}

// rows [0, R) of dst (row r at dst[r * ld + b]) ~ U(lo, hi); row r of sample b is draw
// base + r * B + b
__global__ void k_uniform(uint64_t seed, uint64_t base, int R, int64_t B, double lo, double hi, float* dst,
                          int64_t ld) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)R * B) return;
  const int64_t r = e / B, b = e - r * B;
  dst[r * ld + b] = (float)(lo + (hi - lo) * unit(seed, base + (uint64_t)e));
}

// unit quaternions [x, y, z, w] (Shoemake) into rows 0..3 of dst; uses draws
// base + k * B + b, k = 0, 1, 2
__global__ void k_quaternion(uint64_t seed, uint64_t base, int64_t B, float* dst, int64_t ld) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double u0 = unit(seed, base + b), u1 = unit(seed, base + B + b), u2 = unit(seed, base + 2 * B + b);
  const double s1 = sqrt(1.0 - u0), s2 = sqrt(u0);
  const double twoPi = 6.283185307179586;
  const double t1 = twoPi * u1, t2 = twoPi * u2;
  dst[0 * ld + b] = (float)(s1 * sin(t1));
  dst[1 * ld + b] = (float)(s1 * cos(t1));
  dst[2 * ld + b] = (float)(s2 * sin(t2));
  dst[3 * ld + b] = (float)(s2 * cos(t2));
}

// rows [0, R) ~ sigma N(0, 1) by Box-Muller over the flattened index i = r * B + b:
// with m = ceil(R B / 2) pairs (u1 = draws base .. base+m-1, u2 = the next m),
// z_i = sqrt(-2 ln(1 - u1_i)) cos(2 pi u2_i) for i < m, else the sine of pair i - m
__global__ void k_normal(uint64_t seed, uint64_t base, int R, int64_t B, double sigma, float* dst, int64_t ld) {
  const int64_t n = (int64_t)R * B;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t m = (n + 1) / 2;
  const int64_t j = e < m ? e : e - m;
  const double u1 = unit(seed, base + j), u2 = unit(seed, base + m + j);
  const double rad = sqrt(-2.0 * log(1.0 - u1));
  const double ang = 6.283185307179586 * u2;
  const double z = e < m ? rad * cos(ang) : rad * sin(ang);
  const int64_t r = e / B, b = e - r * B;
  dst[r * ld + b] = (float)(sigma * z);
}

}  // namespace sample
}  // namespace rbd
