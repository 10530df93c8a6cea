// k_kin.hpp -- batched forward kinematics and joint Jacobians (kinematics.hxx:121-145,
// 209-238; jacobian.hxx:23-62): the pass that RNEA, CRBA and the derivatives open with,
// with its results written out instead of consumed.
//
// One lane per configuration, DFS pre-order over the compile-time tree:
//   oMi = oMi[parent] liMi                                   -> gM (12 per joint: R row major, p)
//   v_i = vJ + liMi^-1 v_parent                              -> gV (6 per joint, local frame)
//   a_i = S qdd + v_i x vJ + liMi^-1 a_parent                -> gA (6 per joint, local frame)
//   J[:, dof] = oMi.act(S_dof)  (WORLD frame)                -> gJ (element r + 6 dof)
// Any output may be absent (null pointer); all outputs leave with nontemporal stores.
#pragma once

#include "k_crba.hpp"
#include "kernels_common.hpp"

namespace rbd {

template<class TP> constexpr int kinLdsRows() { return TP::T.nq + 2 * TP::T.nv; }

template<class TP, class S>
__global__ __launch_bounds__(64) void k_kin(const S* __restrict__ mc_in, int64_t B, CView<S> gq, CView<S> gv,
                                            CView<S> ga, View<S> gM, View<S> gV, View<S> gA, View<S> gJ) {
  constexpr auto& T = TP::T;
  constexpr int NJ = TP::NJ, NQ = T.nq, NV = T.nv;
  __shared__ S lds[kinLdsRows<TP>() * kPitch];
  S* sq = lds;
  S* sv = lds + NQ * kPitch;
  S* sa = lds + (NQ + NV) * kPitch;
  const bool wantV = gV.p != nullptr, wantA = gA.p != nullptr;
  const int64_t b0 = int64_t(blockIdx.x) * kTile;
  const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
  stageIn<NQ>(sq, gq, b0, nb);
  if (wantV) stageIn<NV>(sv, gv, b0, nb);
  if (wantA) stageIn<NV>(sa, ga, b0, nb);
  __syncthreads();
  const int L = laneId();
  const int64_t b = b0 + L;
  const bool live = L < nb;

  Xform<S> oM[NJ];
  Mot<S> vj[NJ], aj[NJ];
  auto put = [&](View<S> g, int row, S x) {
    if (live) stOut(&g.p[(int64_t)row * g.rs + b * g.bs], x);
  };

  static_for<0, NJ - 1>([&](auto K) {
    constexpr int i = TP::T.order[K];
    stepFence();
    const S* __restrict__ mc = stepConsts(mc_in);
    constexpr int p = TP::T.parent[i];
    constexpr int ty = TP::T.type[i];
    constexpr int iq = TP::T.iq[i], iv = TP::T.iv[i], nvi = TP::T.nvj[i];
    Xform<S> X;
    Mot<S> vJ;
    if constexpr (ty == JT_FF) {
      S q7[7];
#pragma unroll
      for (int k = 0; k < 7; ++k) q7[k] = sq[(iq + k) * kPitch + L];
      X = freeflyerX(mc, i, q7, false);
      if (wantV)
        vJ = {{sv[iv * kPitch + L], sv[(iv + 1) * kPitch + L], sv[(iv + 2) * kPitch + L]},
              {sv[(iv + 3) * kPitch + L], sv[(iv + 4) * kPitch + L], sv[(iv + 5) * kPitch + L]}};
    } else {
      S sn, cs;
      sincosT(sq[iq * kPitch + L], &sn, &cs);
      X = revoluteX(mc, i, ty - JT_RX, sn, cs);
      if (wantV) vJ = axisMot(ty - JT_RX, sv[iv * kPitch + L]);
    }
    if constexpr (p > 0) oM[i] = compose(oM[p], X);
    else oM[i] = X;
    if (gM.p) {
      const R3<S>& R = oM[i].R;
      const S r9[9] = {R.m00, R.m01, R.m02, R.m10, R.m11, R.m12, R.m20, R.m21, R.m22};
#pragma unroll
      for (int k = 0; k < 9; ++k) put(gM, (i - 1) * 12 + k, r9[k]);
      put(gM, (i - 1) * 12 + 9, oM[i].p.x);
      put(gM, (i - 1) * 12 + 10, oM[i].p.y);
      put(gM, (i - 1) * 12 + 11, oM[i].p.z);
    }
    if (wantV) {
      Mot<S> vi = vJ;
      if constexpr (p > 0) vi = vi + actInv(X, vj[p]);
      vj[i] = vi;
      const S o[6] = {vi.l.x, vi.l.y, vi.l.z, vi.a.x, vi.a.y, vi.a.z};
#pragma unroll
      for (int k = 0; k < 6; ++k) put(gV, (i - 1) * 6 + k, o[k]);
      if (wantA) {
        Mot<S> ai = crossMM(vi, vJ);
        if constexpr (ty == JT_FF) {
          ai = ai + Mot<S>{{sa[iv * kPitch + L], sa[(iv + 1) * kPitch + L], sa[(iv + 2) * kPitch + L]},
                           {sa[(iv + 3) * kPitch + L], sa[(iv + 4) * kPitch + L], sa[(iv + 5) * kPitch + L]}};
        } else {
          ai = ai + axisMot(ty - JT_RX, sa[iv * kPitch + L]);
        }
        if constexpr (p > 0) ai = ai + actInv(X, aj[p]);
        aj[i] = ai;
        const S oa[6] = {ai.l.x, ai.l.y, ai.l.z, ai.a.x, ai.a.y, ai.a.z};
#pragma unroll
        for (int k = 0; k < 6; ++k) put(gA, (i - 1) * 6 + k, oa[k]);
      }
    }
    if (gJ.p) {
#pragma unroll
      for (int k = 0; k < nvi; ++k) {
        const Mot<S> c = worldAxis<ty>(oM[i], k);
        const S o[6] = {c.l.x, c.l.y, c.l.z, c.a.x, c.a.y, c.a.z};
#pragma unroll
        for (int r = 0; r < 6; ++r) put(gJ, (iv + k) * 6 + r, o[r]);
      }
    }
  });
}

}  // namespace rbd
