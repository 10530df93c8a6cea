// dpp.hpp -- cross-lane moves within 16-lane rows and quads through DPP (data-parallel
// primitives: operand modifiers of ordinary VALU moves, no LDS round trip), for kernels
// that give each problem a row of 16 lanes or a quad of 4.
//
//   bcast16<E>(x)   every lane of a 16-lane row receives lane E's x (row_newbcast, CDNA2+)
//   bcast4<E>(x)    every lane of a quad receives lane E's x (quad_perm [E E E E])
//   sum16 / max16   all-reduce over a 16-lane row: quad_perm xor 1, xor 2, then
//                   row_half_mirror (quads 0 <-> 1, 2 <-> 3) and row_mirror (halves);
//                   each step combines a lane with a partner holding a disjoint set, and
//                   the combinations are commutative, so every lane ends with the same bits
//   sum4 / max4     the same over a quad (the first two steps)
//   triple3(x, k)   the three values of lanes i - k .. i - k + 2 (k = 0, 1, 2: the lane's
//                   place in a group of three within its row), by row shifts
//
// The whole row (quad) must be active: callers keep control flow uniform per problem.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbd {

namespace dpp {
constexpr int kQuadXor1 = 0xB1;       // quad_perm [1, 0, 3, 2]
constexpr int kQuadXor2 = 0x4E;       // quad_perm [2, 3, 0, 1]
constexpr int kRowMirror = 0x140;
constexpr int kRowHalfMirror = 0x141;
constexpr int kRowNewBcast = 0x150;   // + lane within the row
constexpr int kRowShl1 = 0x101, kRowShl2 = 0x102;  // lane i reads lane i + 1, i + 2 of its row
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112;  // lane i reads lane i - 1, i - 2
constexpr int quadBcast(int e) { return e | (e << 2) | (e << 4) | (e << 6); }

template<int CTRL> __device__ __forceinline__ double mov(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32));
}
}  // namespace dpp

template<int E> __device__ __forceinline__ double bcast16(double x) {
  static_assert(E >= 0 && E < 16, "lane within the row");
  return dpp::mov<dpp::kRowNewBcast + E>(x);
}
template<int E> __device__ __forceinline__ double bcast4(double x) {
  static_assert(E >= 0 && E < 4, "lane within the quad");
  return dpp::mov<dpp::quadBcast(E)>(x);
}
__device__ __forceinline__ double sum4(double v) {
  v += dpp::mov<dpp::kQuadXor1>(v);
  return v + dpp::mov<dpp::kQuadXor2>(v);
}
__device__ __forceinline__ double max4(double v) {
  v = fmax(v, dpp::mov<dpp::kQuadXor1>(v));
  return fmax(v, dpp::mov<dpp::kQuadXor2>(v));
}
// o = (x of lane i - k, i - k + 1, i - k + 2): the lane's own triple of consecutive lanes
// whose first lane sits k before it (the groups must not cross a row boundary)
__device__ __forceinline__ void triple3(double x, int k, double (&o)[3]) {
  const double l1 = dpp::mov<dpp::kRowShl1>(x), l2 = dpp::mov<dpp::kRowShl2>(x);
  const double r1 = dpp::mov<dpp::kRowShr1>(x), r2 = dpp::mov<dpp::kRowShr2>(x);
  o[0] = k == 0 ? x : (k == 1 ? r1 : r2);
  o[1] = k == 0 ? l1 : (k == 1 ? x : r1);
  o[2] = k == 0 ? l2 : (k == 1 ? l1 : x);
}
__device__ __forceinline__ double sum16(double v) {
  v = sum4(v);
  v += dpp::mov<dpp::kRowHalfMirror>(v);
  return v + dpp::mov<dpp::kRowMirror>(v);
}
__device__ __forceinline__ double max16(double v) {
  v = max4(v);
  v = fmax(v, dpp::mov<dpp::kRowHalfMirror>(v));
  return fmax(v, dpp::mov<dpp::kRowMirror>(v));
}

}  // namespace rbd
