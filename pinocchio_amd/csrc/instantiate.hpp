// instantiate.hpp -- per-topology kernel instantiation and launcher tables.
#pragma once

#include "ops.hpp"
#include "k_dyn.hpp"
#include "k_crba.hpp"
#include "k_deriv.hpp"
#include "k_kin.hpp"
#include "k_kkt.hpp"
#include "_gen/topologies.inc"

namespace rbd {

inline unsigned gridFor(int64_t B) { return (unsigned)((B + kTile - 1) / kTile); }

template<class TP, class S>
void launchRnea(hipStream_t st, GM, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> a, View<S> out) {
  hipLaunchKernelGGL((k_rnea<TP, S>), dim3(gridFor(B)), dim3(kTile), 0, st, mc, B, q, v, a, out);
}
// Below about four tiles per CU the sweeps are latency-bound with one wave per tile.
inline int wavesForBatch(int64_t B) {
  const int64_t tiles = (B + kTile - 1) / kTile;
  return tiles >= 1024 ? 1 : tiles >= 512 ? 2 : 4;
}
// a subtree split pays when the busiest wave (trunk included) has well under the whole tree
template<class TP, int G> constexpr bool abaSplitPays() {
  constexpr WaveSplit<TP, G> WS{};
  int trunk = 0, busiest = 0;
  for (int j = 1; j < TP::NJ; ++j) trunk += WS.owner[j] < 0 ? 1 : 0;
  for (int k = 0; k < G; ++k) busiest = WS.load[k] > busiest ? WS.load[k] : busiest;
  return 5 * (trunk + busiest) <= 3 * (TP::NJ - 1) && abaWavesLdsBytes<TP, G>() <= 160 * 1024;
}
// ... and its LDS (sized for 4-byte scalars) fits at the scalar's width
template<class TP, class S, int G> constexpr bool abaSplitPaysFor() {
  return abaSplitPays<TP, G>() && abaWavesLdsBytes<TP, G>() / 4 * (int)sizeof(S) <= 160 * 1024;
}
// RNEA derivatives: the same test, with the LDS of the staged inputs and the 43-float
// subtree slots
template<class TP, int G> constexpr bool rneadSplitPays() {
  constexpr WaveSplit<TP, G> WS{};
  int trunk = 0, busiest = 0;
  for (int j = 1; j < TP::NJ; ++j) trunk += WS.owner[j] < 0 ? 1 : 0;
  for (int k = 0; k < G; ++k) busiest = WS.load[k] > busiest ? WS.load[k] : busiest;
  constexpr int lds = ((TP::T.nq + 2 * TP::T.nv) * kPitch + WS.nslots * kRneadSlot * kTile) * 4;
  return 5 * (trunk + busiest) <= 3 * (TP::NJ - 1) && lds <= 160 * 1024;
}
// waves of one kernel resident per CU with `threads` threads per workgroup (0 if unknown)
inline int residentWaves(const void* kernel, int threads) {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, threads, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return blocks * (threads / 64);
}
// Waves per tile: chosen from the batch size, and four waves at every size when the four-wave kernel keeps at least twice as many waves
// resident per CU as the one-wave kernel (humanoid: 389 -> 243 registers per lane, one ->
// two waves per SIMD; ABA at 65,536 configurations 67.7 -> 61.7 us).
template<class TP, class S>
void launchAba(hipStream_t st, GM, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> tau, View<S> out) {
  if constexpr (std::is_same_v<S, double>) {
    // FP64: the one-wave kernel spills its per-joint state (twice the registers of FP32);
    // the four-wave split holds it, where the tree splits
    if constexpr (abaSplitPaysFor<TP, S, 4>())
      hipLaunchKernelGGL((k_aba_waves<TP, S, 4>), dim3(gridFor(B)), dim3(kTile * 4), 0, st, mc, B, q, v, tau, out);
    else
      hipLaunchKernelGGL((k_aba<TP, S>), dim3(gridFor(B)), dim3(kTile), 0, st, mc, B, q, v, tau, out);
  } else {
    bool denser = false;
    if constexpr (abaSplitPays<TP, 4>()) {
      static const bool d = [] {
        const int one = residentWaves((const void*)&k_aba<TP, S>, kTile);
        return one > 0 && residentWaves((const void*)&k_aba_waves<TP, S, 4>, kTile * 4) >= 2 * one;
      }();
      denser = d;
    }
    const int waves = denser ? 4 : wavesForBatch(B);
    if constexpr (abaSplitPays<TP, 4>()) {
      if (waves >= 4) {
        hipLaunchKernelGGL((k_aba_waves<TP, S, 4>), dim3(gridFor(B)), dim3(kTile * 4), 0, st, mc, B, q, v, tau, out);
        return;
      }
    }
    if constexpr (abaSplitPays<TP, 2>()) {
      if (waves >= 2) {
        hipLaunchKernelGGL((k_aba_waves<TP, S, 2>), dim3(gridFor(B)), dim3(kTile * 2), 0, st, mc, B, q, v, tau, out);
        return;
      }
    }
    hipLaunchKernelGGL((k_aba<TP, S>), dim3(gridFor(B)), dim3(kTile), 0, st, mc, B, q, v, tau, out);
  }
}
// CRBA: two waves per tile (output emission split) keep two waves resident per SIMD at
// large batches
template<class TP, class S>
void launchCrba(hipStream_t st, GM, const S* mc, int64_t B, CView<S> q, View<S> M, int upper) {
  if (TP::NJ <= 3)
    hipLaunchKernelGGL((k_crba<TP, S, 1>), dim3(gridFor(B)), dim3(kTile), 0, st, mc, B, q, M, upper);
  else
    hipLaunchKernelGGL((k_crba<TP, S, 2>), dim3(gridFor(B)), dim3(kTile * 2), 0, st, mc, B, q, M, upper);
}
template<class TP, class S> void launchChol(hipStream_t st, GM, int64_t B, CView<S> M, View<double> U, View<double> D) {
  hipLaunchKernelGGL((k_chol<TP, S, double>), dim3(gridFor(B)), dim3(kTile), 0, st, B, M, U, D);
}
// FP32 factor (ABA derivatives, computeMinverse): M staged into LDS first when the batch
// leaves about two tiles per CU or fewer
template<class TP> void launchChol32(hipStream_t st, GM, int64_t B, CView<float> M, View<float> U, View<float> D) {
  const bool staged = (B + kTile - 1) / kTile <= 512;
  if (staged && cholStagedLdsBytes<TP, float>() <= 160 * 1024)
    hipLaunchKernelGGL((k_chol_staged<TP, float, float>), dim3(gridFor(B)), dim3(kTile), 0, st, B, M, U, D);
  else
    hipLaunchKernelGGL((k_chol<TP, float, float>), dim3(gridFor(B)), dim3(kTile), 0, st, B, M, U, D);
}
template<class TP, class S>
void launchCholSolve(hipStream_t st, GM, int64_t B, CView<double> U, CView<double> D, View<S> x) {
  hipLaunchKernelGGL((k_chol64_solve<TP, S>), dim3(gridFor(B)), dim3(kTile), 0, st, B, U, D, x);
}
// RNEA derivatives: with few tiles, 4 waves per tile, each running a share of the subtrees
// (WaveSplit), shorten the critical path
template<class TP, class S>
void launchRneaD(hipStream_t st, GM, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> a, View<S> dq, View<S> dv,
                 View<S> M, int mask, int waves) {
  if (waves <= 0) waves = wavesForBatch(B);
  // (FP64: one wave per tile; the split's LDS, sized for 4-byte scalars, does not fit)
  if constexpr (std::is_same_v<S, float> && rneadSplitPays<TP, 4>()) {
    if (waves >= 2) {
      hipLaunchKernelGGL((k_rnead<TP, S, 4>), dim3(gridFor(B)), dim3(kTile * 4), 0, st, mc, B, q, v, a, dq, dv, M,
                         mask);
      return;
    }
  }
  hipLaunchKernelGGL((k_rnead<TP, S, 1>), dim3(gridFor(B)), dim3(kTile), 0, st, mc, B, q, v, a, dq, dv, M, mask);
}
template<class TP, class S>
void launchAbaDSolve(hipStream_t st, GM, int64_t B, CView<S> U, CView<S> D, View<S> Xq, View<S> Xv, View<S> Minv,
                     int jobBegin) {
  constexpr int CT = SolveTile<S>::CT;
  hipLaunchKernelGGL((k_abad_solve<TP, S>), dim3((unsigned)((B + CT - 1) / CT)), dim3(kTile * kSolveWaves), 0, st, B, U,
                     D, Xq, Xv, Minv, jobBegin);
}


template<class TP, class S>
void launchKin(hipStream_t st, GM, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> a, View<S> oM, View<S> vo,
               View<S> ao, View<S> J) {
  hipLaunchKernelGGL((k_kin<TP, S>), dim3(gridFor(B)), dim3(kTile), 0, st, mc, B, q, v, a, oM, vo, ao, J);
}

template<class TP, class S>
void launchKkt(hipStream_t st, GM, int64_t B, CView<double> U, CView<double> D, int nc, CView<S> J, CView<S> tau,
               CView<S> nle, CView<S> gamma, double damping, int mode, double rcoef, const KktProx& prox,
               View<S> ddq, View<S> lam) {
  using K = KktShape<S>;
  hipLaunchKernelGGL((k_kkt<TP, S>), dim3((unsigned)((B + K::T - 1) / K::T)), dim3(64 * K::W), 0, st, B, U, D, nc, J,
                     tau, nle, gamma, damping, mode, rcoef, prox, ddq, lam);
}

template<class TP> TopoOps makeOpsBase() {
  TopoOps o{};
  o.name = TP::name;
  o.njoints = TP::NJ;
  o.parent = TP::parent;
  o.type = TP::type;
  o.nq = TP::T.nq;
  o.nv = TP::T.nv;
  o.compact = TP::T.compact ? 1 : 0;
  return o;
}

// Kernel families, one translation unit each so that a topology's kernels compile in
// parallel.  "dyn" (RNEA, ABA) exists for every topology; the others need a compact
// (DFS-ordered) tree.
template<class TP> void add_dyn(TopoOps& o) {
  o.rnea_f32 = &launchRnea<TP, float>;
  o.rnea_f64 = &launchRnea<TP, double>;
  o.aba_f32 = &launchAba<TP, float>;
  o.aba_f64 = &launchAba<TP, double>;
}
template<class TP> void add_crba(TopoOps& o) {
  o.crba_f32 = &launchCrba<TP, float>;
  o.crba_f64 = &launchCrba<TP, double>;
}
template<class TP> void add_chol(TopoOps& o) {
  o.chol_f32 = &launchChol<TP, float>;
  o.chol_f64 = &launchChol<TP, double>;
  o.cholsolve_f32 = &launchCholSolve<TP, float>;
  o.cholsolve_f64 = &launchCholSolve<TP, double>;
  o.chol32_f32 = &launchChol32<TP>;
  o.nnzU = nnzOf<TP>();
}
template<class TP> void add_rnead(TopoOps& o) {
  o.rnead_f32 = &launchRneaD<TP, float>;
  o.rnead_f64 = &launchRneaD<TP, double>;
}
template<class TP> void add_abads(TopoOps& o) {
  o.abad_solve_f32 = &launchAbaDSolve<TP, float>;
  if constexpr (abadSolveLdsBytes<TP, double>() <= 160 * 1024) o.abad_solve_f64 = &launchAbaDSolve<TP, double>;
}
// (topologies whose staged FP64 factor fits the LDS)
template<class TP> void add_kkt(TopoOps& o) {
  if constexpr (kktLdsBytes<TP, float>() <= 160 * 1024 && kktLdsBytes<TP, double>() <= 160 * 1024) {
    o.kkt_f32 = &launchKkt<TP, float>;
    o.kkt_f64 = &launchKkt<TP, double>;
  }
}
template<class TP> void add_kin(TopoOps& o) {
  o.kin_f32 = &launchKin<TP, float>;
  o.kin_f64 = &launchKin<TP, double>;
}

}  // namespace rbd

// one kernel family of one topology
#define RBD_FILL(NAME, FAM) \
  namespace rbd {           \
  void fill_##NAME##_##FAM(TopoOps& o) { add_##FAM<Topo_##NAME>(o); } \
  }
