// k_dyn.hpp -- batched inverse dynamics (RNEA) and forward dynamics (ABA) kernels.
//
// One lane = one configuration; the tree walk is fully unrolled for the model's
// topology (TP) and follows the PRE/CLOSE schedule of topo.hpp, so only the open path
// of the tree is live.  Both sweeps run in local joint frames: in FP32 the local-frame
// ABA is about 20x more accurate than the world-frame one on the 36-DoF humanoid
// (max rel. err 3.3e-6 vs 6.0e-5 over 8192 configurations), because world-frame
// articulated inertias carry large parallel-axis terms.
//
// RNEA follows rnea.hxx:17-161 (RneaForwardStep / RneaBackwardStep, armature at 158).
// ABA computes the same function as aba.hxx (WORLD 79-293 / LOCAL 354-540).  Its
// acceleration pass is re-associated: the velocity-product bias c_i of each joint is
// accumulated down the tree during the forward pass (chat_i = X_i^-1 chat_p + c_i), so the
// third pass only needs, per joint, d_i = U_i D_i^-1 and the scalar
// s_i = D_i^-1 u_i - d_i . chat_i:   qdd_i = s_i - d_i . (X_i^-1 ahat_p),
// ahat_i = X_i^-1 ahat_p + S_i qdd_i  (ahat = a - chat, ahat_0 = -g).
#pragma once

#include "kernels_common.hpp"

namespace rbd {

template<class TP, class S> struct DynCfg {
  static constexpr auto& T = TP::T;
  static constexpr int NJ = TP::NJ;
  static constexpr int NQ = T.nq;
  static constexpr int NV = T.nv;
};

template<class S> __device__ __forceinline__ Mot<S> mcGravityNeg(const S* __restrict__ mc, int off) {
  return {{-mc[off], -mc[off + 1], -mc[off + 2]}, {-mc[off + 3], -mc[off + 4], -mc[off + 5]}};
}

// ---------------------------------------------------------------------------------------
// RNEA: tau = M(q) a + C(q, v) v + g(q) + armature * a
// ---------------------------------------------------------------------------------------
// LDS rows a tile of this topology needs (q, v and one nv-vector in/out)
template<class TP> constexpr int dynLdsRows() { return TP::T.nq + 2 * TP::T.nv; }

// One 64-configuration tile starting at b0; `lds` holds dynLdsRows<TP>() * kPitch scalars.
// (kU: loads in flight per lane while staging; the fused mixed-batch kernel, already at
// the register limit, stages one row at a time)
template<class TP, class S, int kU = kStageU>
__device__ __forceinline__ void rneaTile(const S* __restrict__ mc_in, int64_t B, int64_t b0, CView<S> gq,
                                         CView<S> gv, CView<S> ga, View<S> gtau, S* __restrict__ lds) {
  using C = DynCfg<TP, S>;
  constexpr auto& T = TP::T;
  constexpr int NJ = C::NJ, NQ = C::NQ, NV = C::NV;
  S* sq = lds;
  S* sv = lds + NQ * kPitch;
  S* sa = lds + (NQ + NV) * kPitch;  // a in, tau out
  const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
  stageIn<NQ, S, kU>(sq, gq, b0, nb);
  stageIn<NV, S, kU>(sv, gv, b0, nb);
  stageIn<NV, S, kU>(sa, ga, b0, nb);
  __syncthreads();
  const int L = laneId();
  const Mot<S> a0 = mcGravityNeg(mc_in, mcGravity<NJ, NV>());

  Mot<S> vj[NJ], aj[NJ];
  Frc<S> fj[NJ];
  S sn[NJ], cs[NJ];

  static_for<0, T.nev>([&](auto E) {
    constexpr int e = E;
    stepFence();
    const S* __restrict__ mc = stepConsts(mc_in);
    constexpr int i = TP::T.evJoint[e];
    constexpr int p = TP::T.parent[i];
    constexpr int ty = TP::T.type[i];
    constexpr int iq = TP::T.iq[i], iv = TP::T.iv[i];
    if constexpr (TP::T.evKind[e] == EV_PRE) {
      Xform<S> X;
      Mot<S> jv, jacc;
      if constexpr (ty == JT_FF) {
        S q7[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) q7[k] = sq[(iq + k) * kPitch + L];
        X = freeflyerX(mc, i, q7, false);
        jv = {{sv[(iv + 0) * kPitch + L], sv[(iv + 1) * kPitch + L], sv[(iv + 2) * kPitch + L]},
              {sv[(iv + 3) * kPitch + L], sv[(iv + 4) * kPitch + L], sv[(iv + 5) * kPitch + L]}};
        jacc = {{sa[(iv + 0) * kPitch + L], sa[(iv + 1) * kPitch + L], sa[(iv + 2) * kPitch + L]},
                {sa[(iv + 3) * kPitch + L], sa[(iv + 4) * kPitch + L], sa[(iv + 5) * kPitch + L]}};
      } else {
        constexpr int ax = ty - JT_RX;
        sincosT(sq[iq * kPitch + L], &sn[i], &cs[i]);
        X = revoluteX(mc, i, ax, sn[i], cs[i]);
        jv = axisMot(ax, sv[iv * kPitch + L]);
        jacc = axisMot(ax, sa[iv * kPitch + L]);
      }
      Mot<S> vi = jv;
      if constexpr (p > 0) vi = vi + actInv(X, vj[p]);
      Mot<S> ai = crossMM(vi, jv) + jacc;
      if constexpr (p > 0) ai = ai + actInv(X, aj[p]);
      else ai = ai + actInv(X, a0);
      vj[i] = vi;
      aj[i] = ai;
      const Body<S> Y = mcBody(mc, i);
      fj[i] = inertiaMul(Y, ai) + crossMF(vi, inertiaMul(Y, vi));
      pinAll<S>(vj[i]);
      pinAll<S>(aj[i]);
      pinAll<S>(fj[i]);
      if constexpr (ty != JT_FF) pin(sn[i], cs[i]);
    } else {
      if constexpr (ty == JT_FF) {
        const S f6[6] = {fj[i].l.x, fj[i].l.y, fj[i].l.z, fj[i].a.x, fj[i].a.y, fj[i].a.z};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const S arm = mc[mcArmature<NJ>() + iv + k];
          sa[(iv + k) * kPitch + L] = f6[k] + arm * sa[(iv + k) * kPitch + L];
        }
        if constexpr (p > 0) {
          S q7[7];
#pragma unroll
          for (int k = 0; k < 7; ++k) q7[k] = sq[(iq + k) * kPitch + L];
          const Xform<S> X = freeflyerX(mc, i, q7, false);
          fj[p] = fj[p] + act(X, fj[i]);
        }
      } else {
        constexpr int ax = ty - JT_RX;
        const S arm = mc[mcArmature<NJ>() + iv];
        sa[iv * kPitch + L] = comp(fj[i], 3 + ax) + arm * sa[iv * kPitch + L];
        if constexpr (p > 0) {
          const Xform<S> X = revoluteX(mc, i, ax, sn[i], cs[i]);
          fj[p] = fj[p] + act(X, fj[i]);
          pinAll<S>(fj[p]);
        }
      }
    }
  });
  __syncthreads();
  stageOut<NV, S, kU>(sa, gtau, b0, nb);
}

template<class TP, class S>
__global__ __launch_bounds__(64) void k_rnea(const S* __restrict__ mc, int64_t B, CView<S> gq, CView<S> gv,
                                             CView<S> ga, View<S> gtau) {
  __shared__ S lds[dynLdsRows<TP>() * kPitch];
  // a single wave stages the whole tile: 16 loads in flight per lane
  rneaTile<TP, S, 2 * kStageU>(mc, B, int64_t(blockIdx.x) * kTile, gq, gv, ga, gtau, lds);
}

// ---------------------------------------------------------------------------------------
// ABA: ddq = FD(q, v, tau)
//
// G = 1: one wave runs the whole tree for its 64 configurations.  G > 1 (small batches,
// where one wave per tile leaves most SIMDs idle): G waves share the tile and split the
// tree by subtrees (WaveSplit); this function is then the body of wave W.  Trunk joints
// are computed by every wave; a subtree hands its articulated inertia and bias force to
// its trunk parent through an LDS slot (xs), read after a barrier.  ddq goes to a
// separate LDS block (so) because trunk joints are evaluated by several waves.
// ---------------------------------------------------------------------------------------
constexpr int kAbaSlot = 27;  // floats per slot: 21 (articulated inertia) + 6 (bias force)
template<class TP> constexpr int abaLdsRows(int G) { return TP::T.nq + (G > 1 ? 3 : 2) * TP::T.nv; }

template<class TP, class S, int G = 1, int W = 0, int kU = kStageU>
__device__ __forceinline__ void abaTile(const S* __restrict__ mc_in, int64_t B, int64_t b0, CView<S> gq,
                                        CView<S> gv, CView<S> gt, View<S> gddq, S* __restrict__ lds,
                                        S* __restrict__ xs = nullptr) {
  using C = DynCfg<TP, S>;
  constexpr auto& T = TP::T;
  constexpr int NJ = C::NJ, NQ = C::NQ, NV = C::NV;
  constexpr WaveSplit<TP, G> WS{};
  S* sq = lds;
  S* sv = lds + NQ * kPitch;
  S* st = lds + (NQ + NV) * kPitch;                    // tau in
  S* so = G > 1 ? lds + (NQ + 2 * NV) * kPitch : st;  // ddq out
  const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
  // the three inputs are staged by different waves when there are several (more loads
  // in flight per tile)
  if constexpr (W == 0) stageIn<NQ, S, kU>(sq, gq, b0, nb);
  if constexpr (W == 1 % G) stageIn<NV, S, kU>(sv, gv, b0, nb);
  if constexpr (W == 2 % G) stageIn<NV, S, kU>(st, gt, b0, nb);
  __syncthreads();
  const int L = laneId();
  const Mot<S> a0 = mcGravityNeg(mc_in, mcGravity<NJ, NV>());

  Mot<S> vj[NJ], chat[NJ], ahat[NJ];
  Sym6<S> Iacc[NJ];
  Frc<S> pacc[NJ], dd[NJ];
  S ss[NJ], sn[NJ], cs[NJ];

  auto ldV = [&](int k) { return sv[k * kPitch + L]; };

  // ---- passes 1 + 2, interleaved in DFS order ----
  static_for<0, T.nev>([&](auto E) {
    constexpr int e = E;
    constexpr int i = TP::T.evJoint[e];
    if constexpr (WS.runs(i, W)) {
    stepFence();
    const S* __restrict__ mc = stepConsts(mc_in);
    constexpr int p = TP::T.parent[i];
    constexpr int ty = TP::T.type[i];
    constexpr int iq = TP::T.iq[i], iv = TP::T.iv[i];
    if constexpr (TP::T.evKind[e] == EV_PRE) {
      Xform<S> X;
      Mot<S> jv;
      if constexpr (ty == JT_FF) {
        static_assert(p == 0, "free-flyer joints are supported at the root only");
        S q7[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) q7[k] = sq[(iq + k) * kPitch + L];
        X = freeflyerX(mc, i, q7, false);
        jv = {{ldV(iv), ldV(iv + 1), ldV(iv + 2)}, {ldV(iv + 3), ldV(iv + 4), ldV(iv + 5)}};
      } else {
        constexpr int ax = ty - JT_RX;
        sincosT(sq[iq * kPitch + L], &sn[i], &cs[i]);
        X = revoluteX(mc, i, ax, sn[i], cs[i]);
        jv = axisMot(ax, ldV(iv));
      }
      Mot<S> vi = jv;
      if constexpr (p > 0) vi = vi + actInv(X, vj[p]);
      vj[i] = vi;
      Mot<S> ci = crossMM(vi, jv);
      if constexpr (p > 0) chat[i] = actInv(X, chat[p]) + ci;
      else chat[i] = ci;
      pinAll<S>(vj[i]);
      pinAll<S>(chat[i]);
      if constexpr (ty != JT_FF) pin(sn[i], cs[i]);
    } else {
      // ---- CLOSE(i): articulated inertia and bias force of the subtree ----
      const Body<S> Y = mcBody(mc, i);
      Sym6<S> Ia = mcSym6(mc, i);
      Frc<S> pA = crossMF(vj[i], inertiaMul(Y, vj[i]));
      if constexpr (WS.nReg[i] > 0) {
        symAdd(Ia, Iacc[i]);
        pA = pA + pacc[i];
      }
      if constexpr (G > 1 && WS.ldsIn[i]) {
        // contributions of subtrees owned by other waves (fixed order: identical in every wave)
        __syncthreads();
        static_for<1, NJ>([&](auto Cc) {
          constexpr int c = Cc;
          if constexpr (TP::T.parent[c] == i && WS.slot[c] >= 0) {
            const S* x = xs + WS.slot[c] * kAbaSlot * kTile + L;
            Sym6<S> Ic;
#pragma unroll
            for (int k = 0; k < 21; ++k) Ic.a[k] = x[k * kTile];
            symAdd(Ia, Ic);
            pA = pA + Frc<S>{{x[21 * kTile], x[22 * kTile], x[23 * kTile]},
                             {x[24 * kTile], x[25 * kTile], x[26 * kTile]}};
          }
        });
      }
      if constexpr (ty == JT_FF) {
        // root free-flyer: D = Ia + armature (6x6, FP64 LLT), qdd = D^-1 (u - Ia a')
        S q7[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) q7[k] = sq[(iq + k) * kPitch + L];
        const Xform<S> X = freeflyerX(mc, i, q7, false);
        const Mot<S> ap = actInv(X, a0);  // c_root = 0, chat_root = 0
        const Frc<S> Iap = symMul6(Ia, ap);
        double A[21], x[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
#pragma unroll
          for (int c = 0; c <= r; ++c) A[lowIdx(r, c)] = double(sget(Ia, r, c));
          A[lowIdx(r, r)] += double(mc[mcArmature<NJ>() + iv + r]);
          x[r] = double(st[(iv + r) * kPitch + L]) - double(comp(pA, r)) - double(comp(Iap, r));
        }
        spdSolve6(A, x);
        if constexpr (WS.writer(i) == W) {
#pragma unroll
          for (int r = 0; r < 6; ++r) so[(iv + r) * kPitch + L] = S(x[r]);
        }
        ahat[i] = {{ap.l.x + S(x[0]), ap.l.y + S(x[1]), ap.l.z + S(x[2])},
                   {ap.a.x + S(x[3]), ap.a.y + S(x[4]), ap.a.z + S(x[5])}};
        pinAll<S>(ahat[i]);
      } else {
        constexpr int ax = ty - JT_RX;
        const Frc<S> U = symCol(Ia, 3 + ax);
        const S D = comp(U, 3 + ax) + mc[mcArmature<NJ>() + iv];
        const S Dinv = S(1) / D;
        const S u = st[iv * kPitch + L] - comp(pA, 3 + ax);
        dd[i] = scale(Dinv, U);
        ss[i] = Dinv * u - dot(chat[i], dd[i]);
        pinAll<S>(dd[i]);
        pin(ss[i]);
        if constexpr (p > 0) {
          symRank1Sub(Ia, U, Dinv);
          const Mot<S> ci = crossMM(vj[i], axisMot(ax, ldV(iv)));
          const Frc<S> pa = pA + symMul6(Ia, ci) + scale(Dinv * u, U);
          const Xform<S> X = revoluteX(mc, i, ax, sn[i], cs[i]);
          const Sym6<S> Iup = transportUp(X, Ia);
          const Frc<S> pup = act(X, pa);
          if constexpr (G > 1 && WS.slot[i] >= 0) {
            S* x = xs + WS.slot[i] * kAbaSlot * kTile + L;
#pragma unroll
            for (int k = 0; k < 21; ++k) x[k * kTile] = Iup.a[k];
            x[21 * kTile] = pup.l.x;
            x[22 * kTile] = pup.l.y;
            x[23 * kTile] = pup.l.z;
            x[24 * kTile] = pup.a.x;
            x[25 * kTile] = pup.a.y;
            x[26 * kTile] = pup.a.z;
          } else {
            if constexpr (WS.regIdx[i] == 0) {
              Iacc[p] = Iup;
              pacc[p] = pup;
            } else {
              symAdd(Iacc[p], Iup);
              pacc[p] = pacc[p] + pup;
            }
            pinAll<S>(Iacc[p]);
            pinAll<S>(pacc[p]);
          }
        }
      }
    }
    }
  });

  // ---- pass 3: accelerations, DFS pre-order ----
  static_for<0, NJ - 1>([&](auto K) {
    constexpr int i = TP::T.order[K];
    if constexpr (WS.runs(i, W)) {
    stepFence();
    const S* __restrict__ mc = stepConsts(mc_in);
    constexpr int p = TP::T.parent[i];
    constexpr int ty = TP::T.type[i];
    constexpr int iv = TP::T.iv[i];
    if constexpr (ty != JT_FF) {
      constexpr int ax = ty - JT_RX;
      // the joint's sine and cosine are recomputed rather than kept from pass 1: two
      // registers per joint across the whole backward pass cost more than the polynomial
      S sn3, cs3;
      sincosT(sq[TP::T.iq[i] * kPitch + L], &sn3, &cs3);
      const Xform<S> X = revoluteX(mc, i, ax, sn3, cs3);
      Mot<S> t;
      if constexpr (p > 0) t = actInv(X, ahat[p]);
      else t = actInv(X, a0);
      const S qdd = ss[i] - dot(t, dd[i]);
      if constexpr (WS.writer(i) == W) so[iv * kPitch + L] = qdd;
      if constexpr (TP::T.nchild[i] > 0) {
        ahat[i] = t + axisMot(ax, qdd);
        pinAll<S>(ahat[i]);
      }
    }
    }
  });
  __syncthreads();
  if (W == 0) stageOut<NV, S, kU>(so, gddq, b0, nb);
}

template<class TP, class S>
__global__ __launch_bounds__(64) void k_aba(const S* __restrict__ mc, int64_t B, CView<S> gq, CView<S> gv,
                                            CView<S> gt, View<S> gddq) {
  __shared__ S lds[dynLdsRows<TP>() * kPitch];
  abaTile<TP, S, 1, 0, 2 * kStageU>(mc, B, int64_t(blockIdx.x) * kTile, gq, gv, gt, gddq, lds);
}

// G waves per 64-configuration tile, split by subtrees (small batches)
template<class TP, int G> constexpr int abaWavesLdsBytes() {
  constexpr WaveSplit<TP, G> WS{};
  return (abaLdsRows<TP>(G) * kPitch + (WS.nslots > 0 ? WS.nslots : 1) * kAbaSlot * kTile) * 4;
}
template<class TP, class S, int G>
__global__ __launch_bounds__(64 * G) void k_aba_waves(const S* __restrict__ mc, int64_t B, CView<S> gq, CView<S> gv,
                                                      CView<S> gt, View<S> gddq) {
  constexpr WaveSplit<TP, G> WS{};
  __shared__ S lds[abaLdsRows<TP>(G) * kPitch];
  __shared__ S xs[(WS.nslots > 0 ? WS.nslots : 1) * kAbaSlot * kTile];
  const int w = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  static_for<0, G>([&](auto Wc) {
    constexpr int W = Wc;
    if (w == W) abaTile<TP, S, G, W>(mc, B, int64_t(blockIdx.x) * kTile, gq, gv, gt, gddq, lds, xs);
  });
}

// ---------------------------------------------------------------------------------------
// Mixed-robot batch: one launch over instances of several models.  Each workgroup takes
// one entry of a tile table (group, first 64-instance chunk, number of chunks) and runs
// RNEA then ABA on those chunks with the group's compile-time topology; the topology
// branch is uniform per workgroup, so waves never diverge.  Tiles are ordered by
// decreasing cost, so the hardware dispatcher balances the uneven mix.
// ---------------------------------------------------------------------------------------
struct MultiGroup {
  const float* mc;
  int64_t count;
  const float* q;
  const float* v;
  const float* a;
  const float* tau;
  float* tau_out;
  float* ddq_out;
  int32_t topo;
  int32_t pad;
};
struct MultiTile {
  int32_t group, chunk0, nchunks, pad;
};

template<class TP0, class... TPs> constexpr int maxDynRows() {
  if constexpr (sizeof...(TPs) == 0) return dynLdsRows<TP0>();
  else return dynLdsRows<TP0>() > maxDynRows<TPs...>() ? dynLdsRows<TP0>() : maxDynRows<TPs...>();
}
template<class F, class TP0, class... TPs> __device__ __forceinline__ void dispatchTopo(int id, F& f) {
  if (id == 0) f(TP0{});
  else if constexpr (sizeof...(TPs) > 0) dispatchTopo<F, TPs...>(id - 1, f);
}

template<class... TPs>
__global__ __launch_bounds__(64) void k_multi(const MultiGroup* __restrict__ groups,
                                              const MultiTile* __restrict__ tiles) {
  __shared__ float lds[maxDynRows<TPs...>() * kPitch];
  const MultiTile t = tiles[blockIdx.x];
  const MultiGroup g = groups[t.group];
  auto run = [&](auto tp) {
    using TP = decltype(tp);
    const int64_t n = g.count;
#pragma unroll 1
    for (int c = 0; c < t.nchunks; ++c) {
      const int64_t b0 = int64_t(t.chunk0 + c) * kTile;
      rneaTile<TP, float, 1>(g.mc, n, b0, CView<float>{g.q, n, 1}, CView<float>{g.v, n, 1}, CView<float>{g.a, n, 1},
                          View<float>{g.tau_out, n, 1}, lds);
      __syncthreads();
      abaTile<TP, float, 1, 0, 1>(g.mc, n, b0, CView<float>{g.q, n, 1}, CView<float>{g.v, n, 1}, CView<float>{g.tau, n, 1},
                         View<float>{g.ddq_out, n, 1}, lds);
      __syncthreads();
    }
  };
  dispatchTopo<decltype(run), TPs...>(g.topo, run);
}

}  // namespace rbd
