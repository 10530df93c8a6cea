// kernels_common.hpp -- shared pieces of the per-lane sweep kernels: batch I/O staging
// through LDS, model-constant access, joint transforms.
//
// One lane evaluates one configuration.  A workgroup is one 64-lane wave that owns a
// tile of 64 consecutive configurations; its inputs are copied into LDS with coalesced
// loads (whatever the caller's layout) and read back per lane as the sweep needs them,
// so no input occupies a register longer than its use.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spatial.hpp"
#include "topo.hpp"

namespace rbd {

constexpr int kTile = 64;   // configurations per wave
constexpr int kPitch = 65;  // LDS row pitch (odd: conflict-free column access)

// Strided batch view: element (k, b) at p[k * rs + b * bs].
template<class S> struct View {
  S* p;
  int64_t rs, bs;
};
template<class S> struct CView {
  const S* p;
  int64_t rs, bs;
};

__device__ __forceinline__ int laneId() { return threadIdx.x & 63; }

// Scheduling fence between per-joint steps of an unrolled sweep: keeps the machine
// scheduler from hoisting the next joint's constant loads across the whole tree, which
// would stretch live ranges far beyond the open path.
#ifndef RBD_FENCE_MASK
#define RBD_FENCE_MASK 0  // instruction classes allowed across (sched_barrier mask); 0: none
#endif
__device__ __forceinline__ void stepFence() {
#ifndef RBD_NO_FENCE
  __builtin_amdgcn_sched_barrier(RBD_FENCE_MASK);
#endif
}

// Model-constant pointer made opaque to the optimiser at the point of use: uniform loads
// through it cannot be merged with, or hoisted above, those of earlier steps, so a
// joint's constants occupy scalar registers only during that joint's step.
template<class S> __device__ __forceinline__ const S* stepConsts(const S* mc) {
  // the pointer is wave-uniform; make that explicit so it always lives in SGPRs
  const uint64_t u = reinterpret_cast<uint64_t>(mc);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  mc = reinterpret_cast<const S*>(((uint64_t)hi << 32) | lo);
  asm volatile("" : "+s"(mc));
  return mc;
}

// Pins freshly computed values at this point of the instruction stream: the empty asm
// consumes and redefines them, so their arithmetic cannot float past later scheduling
// fences (a fence orders only side-effecting instructions, not pure arithmetic).
#ifndef RBD_NO_PIN
template<class S> __device__ __forceinline__ void pin(S& a, S& b) { asm volatile("" : "+v"(a), "+v"(b)); }
template<class S> __device__ __forceinline__ void pin(S& a) { asm volatile("" : "+v"(a)); }
// every scalar member of a struct of scalars (V3, Mot, Frc, Sym6, Xform, CInertia, ...)
template<class S, class T> __device__ __forceinline__ void pinAll(T& o) {
  static_assert(sizeof(T) % sizeof(S) == 0, "struct of scalars expected");
  S* f = reinterpret_cast<S*>(&o);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / sizeof(S)); ++k) asm volatile("" : "+v"(f[k]));
}
#else
template<class S> __device__ __forceinline__ void pin(S&, S&) {}
template<class S> __device__ __forceinline__ void pin(S&) {}
template<class S, class T> __device__ __forceinline__ void pinAll(T&) {}
#endif

// Per-lane pointer made opaque at the point of use, so loads through it are neither
// merged with nor reused from (CSE) earlier steps: values re-read later stay out of the
// register file in between.
template<class S> __device__ __forceinline__ S* laneLaunder(S* p) {
  asm volatile("" : "+v"(p));
  return p;
}

// Copy rows [0, R) of the tile starting at configuration b0 into lds[k * kPitch + b].
// Lanes past the batch end get zeros (keeps every lane on the fast math paths).  kStageU
// loads per lane are issued before the first is waited on (the addresses of lanes past
// the batch end are clamped to the last configuration, so no load needs a branch): with
// one wave per SIMD, a load-wait-store loop would expose the full memory latency once per
// row.
constexpr int kStageU = 8;
// (the per-batch loops below use static_for, so the kStageU values always live in
// registers: a plain loop over a local array may be left rolled in very large kernels,
// which puts the array on the stack)
template<int R, class S, int kU = kStageU>
__device__ __forceinline__ void stageIn(S* __restrict__ lds, CView<S> g, int64_t b0, int nb) {
  const int t = laneId();
  if constexpr (kU == 1) {  // compact loop: one row in flight
    if (g.rs == 1) {
      for (int e = t; e < R * kTile; e += kTile) {
        const int b = e / R, k = e - b * R;
        lds[k * kPitch + b] = b < nb ? g.p[(b0 + b) * g.bs + k] : S(0);
      }
    } else {
      for (int e = t; e < R * kTile; e += kTile) {
        const int k = e >> 6, b = e & 63;
        lds[k * kPitch + b] = b < nb ? g.p[k * g.rs + (b0 + b) * g.bs] : S(0);
      }
    }
    return;
  }
  if (g.rs == 1) {  // configuration-contiguous (column major, Eigen default)
    static_for<0, (R + kU - 1) / kU>([&](auto I) {
      constexpr int e0 = int(I) * kU * kTile;
      S x[kU];
      static_for<0, kU>([&](auto U) {
        const int e = e0 + int(U) * kTile + t;
        const int ec = e < R * kTile ? e : R * kTile - 1;
        const int b = ec / R, k = ec - b * R;
        const int bc = b < nb ? b : nb - 1;
        const S v = g.p[(b0 + bc) * g.bs + k];
        x[U] = b < nb ? v : S(0);
      });
      static_for<0, kU>([&](auto U) {
        const int e = e0 + int(U) * kTile + t;
        if (e < R * kTile) {
          const int b = e / R, k = e - b * R;
          lds[k * kPitch + b] = x[U];
        }
      });
    });
  } else {  // batch-fastest or general strides: row k of the tile is one strided segment
    const int bc = t < nb ? t : nb - 1;
    static_for<0, (R + kU - 1) / kU>([&](auto I) {
      constexpr int k0 = int(I) * kU;
      S x[kU];
      static_for<0, kU>([&](auto U) {
        if constexpr (k0 + int(U) < R) {
          const S v = g.p[(k0 + int(U)) * g.rs + (b0 + bc) * g.bs];
          x[U] = t < nb ? v : S(0);
        }
      });
      static_for<0, kU>([&](auto U) {
        if constexpr (k0 + int(U) < R) lds[(k0 + int(U)) * kPitch + t] = x[U];
      });
    });
  }
}
template<int R, class S, int kU = kStageU>
__device__ __forceinline__ void stageOut(const S* __restrict__ lds, View<S> g, int64_t b0, int nb) {
  const int t = laneId();
  if constexpr (kU == 1) {  // compact loop: one row in flight
    if (g.rs == 1) {
      for (int e = t; e < R * kTile; e += kTile) {
        const int b = e / R, k = e - b * R;
        if (b < nb) g.p[(b0 + b) * g.bs + k] = lds[k * kPitch + b];
      }
    } else {
      for (int e = t; e < R * kTile; e += kTile) {
        const int k = e >> 6, b = e & 63;
        if (b < nb) g.p[k * g.rs + (b0 + b) * g.bs] = lds[k * kPitch + b];
      }
    }
    return;
  }
  if (g.rs == 1) {
    static_for<0, (R + kStageU - 1) / kStageU>([&](auto I) {
      constexpr int e0 = int(I) * kStageU * kTile;
      S x[kStageU];
      static_for<0, kStageU>([&](auto U) {
        const int e = e0 + int(U) * kTile + t;
        const int ec = e < R * kTile ? e : R * kTile - 1;
        const int b = ec / R, k = ec - b * R;
        x[U] = lds[k * kPitch + b];
      });
      static_for<0, kStageU>([&](auto U) {
        const int e = e0 + int(U) * kTile + t;
        const int b = e / R, k = e - b * R;
        if (e < R * kTile && b < nb) g.p[(b0 + b) * g.bs + k] = x[U];
      });
    });
  } else {
    static_for<0, (R + kStageU - 1) / kStageU>([&](auto I) {
      constexpr int k0 = int(I) * kStageU;
      S x[kStageU];
      static_for<0, kStageU>([&](auto U) {
        if constexpr (k0 + int(U) < R) x[U] = lds[(k0 + int(U)) * kPitch + t];
      });
      static_for<0, kStageU>([&](auto U) {
        if constexpr (k0 + int(U) < R)
          if (t < nb) g.p[(k0 + int(U)) * g.rs + (b0 + t) * g.bs] = x[U];
      });
    });
  }
}

// ---- model constants ----------------------------------------------------------------
template<class S> __device__ __forceinline__ R3<S> mcRot(const S* __restrict__ mc, int j) {
  const S* r = mc + j * MC_REC + MC_R;
  return {r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8]};
}
template<class S> __device__ __forceinline__ V3<S> mcPos(const S* __restrict__ mc, int j) {
  const S* r = mc + j * MC_REC + MC_P;
  return {r[0], r[1], r[2]};
}
template<class S> __device__ __forceinline__ Body<S> mcBody(const S* __restrict__ mc, int j) {
  const S* r = mc + j * MC_REC;
  return {r[MC_M], {r[MC_C], r[MC_C + 1], r[MC_C + 2]}, r[MC_I], r[MC_I + 1], r[MC_I + 2], r[MC_I + 3], r[MC_I + 4],
          r[MC_I + 5]};
}
template<class S> __device__ __forceinline__ Sym6<S> mcSym6(const S* __restrict__ mc, int j) {
  const S* r = mc + j * MC_REC + MC_Y6;
  Sym6<S> o;
#pragma unroll
  for (int k = 0; k < 21; ++k) o.a[k] = r[k];
  return o;
}
template<class S> __device__ __forceinline__ CInertia<S> mcComposite(const S* __restrict__ mc, int j) {
  const S* r = mc + j * MC_REC + MC_CI;
  return {r[0], {r[1], r[2], r[3]}, r[4], r[5], r[6], r[7], r[8], r[9]};
}

// ---- joint transforms ---------------------------------------------------------------
// Branch-free sine and cosine of a joint angle.  The library routine carries a
// large-argument reduction path whose registers, inlined once per joint, would be live
// across the whole sweep; this one is straight-line code:
//   k = rint(x * 2/pi), r = x - k pi/2 (Cody-Waite, pi/2 split in three parts, FMA),
//   sin r, cos r by their Taylor series on |r| <= pi/4 (truncation below 3e-9 in FP32
//   with terms to r^9 / r^10, below 1e-19 in FP64 with terms to r^17 / r^18),
//   then the quadrant k mod 4 selects and signs the pair.
// Accurate to a few ulp for |x| up to ~1e4 (FP32) / ~1e8 (FP64), far beyond joint ranges.
template<class S> struct TrigConst;
template<> struct TrigConst<float> {
  static constexpr float twoOverPi = 0.636619772367581343f;
  static constexpr float c1 = 1.5707963705062866f, c2 = -4.371138828673793e-08f, c3 = -1.7151245100058819e-15f;
  static constexpr int nsin = 4, ncos = 5;
};
template<> struct TrigConst<double> {
  static constexpr double twoOverPi = 0.6366197723675814;
  static constexpr double c1 = 1.5707963267948966, c2 = 6.123233995736766e-17, c3 = -1.4973849048591698e-33;
  static constexpr int nsin = 8, ncos = 9;
};
template<class S> RBD_HD void sincosT(S x, S* s, S* c) {
  using K = TrigConst<S>;
  const S k = rint(x * K::twoOverPi);
  S r = fma(-k, K::c1, x);
  r = fma(-k, K::c2, r);
  r = fma(-k, K::c3, r);
  const S r2 = r * r;
  // sin r = r (1 - r^2/3! + r^4/5! - ...), cos r = 1 - r^2/2! + r^4/4! - ...  (Horner)
  S ps = S(0), pc = S(0);
  {
    // coefficients (-1)^n / (2n+1)!  and  (-1)^n / (2n)!  for n = 1..N, built exactly
    S fs[10], fc[10];
    S f = S(1);
#pragma unroll
    for (int n = 1; n <= 9; ++n) {
      fc[n] = ((n & 1) ? S(-1) : S(1)) / (f * S(2 * n - 1) * S(2 * n));
      f = f * S(2 * n - 1) * S(2 * n);
      fs[n] = ((n & 1) ? S(-1) : S(1)) / (f * S(2 * n + 1));
    }
#pragma unroll
    for (int n = K::nsin; n >= 1; --n) ps = fma(ps, r2, fs[n]);
#pragma unroll
    for (int n = K::ncos; n >= 1; --n) pc = fma(pc, r2, fc[n]);
  }
  const S sr = fma(r * r2, ps, r);
  const S cr = fma(r2, pc, S(1));
  const int q = int(k) & 3;
  const S s0 = (q & 1) ? cr : sr;
  const S c0 = (q & 1) ? sr : cr;
  *s = (q & 2) ? -s0 : s0;
  *c = ((q + 1) & 2) ? -c0 : c0;
}

// liMi = jointPlacement * jointTransform(q) for a revolute joint (joint-revolute.hpp:793-820)
template<class S> __device__ __forceinline__ Xform<S> revoluteX(const S* __restrict__ mc, int j, int ax, S sn, S cs) {
  return {rotAbout(mcRot(mc, j), ax, sn, cs), mcPos(mc, j)};
}
// liMi for a free-flyer joint (joint-free-flyer.hpp:333-375); q7 = [x, y, z, qx, qy, qz, qw]
template<class S>
__device__ __forceinline__ Xform<S> freeflyerX(const S* __restrict__ mc, int j, const S* q7, bool identityPlacement) {
  const Xform<S> J{quatToRot(q7[3], q7[4], q7[5], q7[6]), {q7[0], q7[1], q7[2]}};
  if (identityPlacement) return J;
  return compose(Xform<S>{mcRot(mc, j), mcPos(mc, j)}, J);
}

template<class S> __device__ __forceinline__ Mot<S> axisMot(int ax, S w) {
  Mot<S> m = zeroMot<S>();
  if (ax == 0) m.a.x = w;
  else if (ax == 1) m.a.y = w;
  else m.a.z = w;
  return m;
}

// 6x6 SPD solve in double through an LLT (PerformStYSInversion,
// joint-common-operations.hpp:22-45, for the free-flyer D = S^T Ia S block), in place on
// the packed lower triangle a[r (r + 1) / 2 + c] (r >= c): a becomes L, b becomes the
// solution.  Packed and in place so that the factor costs 21 doubles of registers at the
// point of the sweep where every joint's backward-pass state is still live.
__device__ __forceinline__ constexpr int lowIdx(int r, int c) { return r * (r + 1) / 2 + c; }
__device__ __forceinline__ void spdSolve6(double (&a)[21], double (&b)[6]) {
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double s = a[lowIdx(j, j)];
#pragma unroll
    for (int k = 0; k < j; ++k) s -= a[lowIdx(j, k)] * a[lowIdx(j, k)];
    const double ljj = sqrt(s);
    const double inv = 1.0 / ljj;
    a[lowIdx(j, j)] = ljj;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      double t = a[lowIdx(i, j)];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= a[lowIdx(i, k)] * a[lowIdx(j, k)];
      a[lowIdx(i, j)] = t * inv;
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double t = b[i];
#pragma unroll
    for (int k = 0; k < i; ++k) t -= a[lowIdx(i, k)] * b[k];
    b[i] = t / a[lowIdx(i, i)];
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double t = b[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) t -= a[lowIdx(k, i)] * b[k];
    b[i] = t / a[lowIdx(i, i)];
  }
}

}  // namespace rbd
