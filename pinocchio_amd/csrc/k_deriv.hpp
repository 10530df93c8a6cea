// k_deriv.hpp -- batched analytical derivatives of RNEA and ABA.
//
// computeRNEADerivatives (rnea-derivatives.hxx:242-541) works in the world frame with
// per-dof column data J, dV/dq, dA/dq, dA/dv and per-body composite quantities Ycrb,
// doYcrb (= Ycrb.variation(ov) + [oh x*], inertia.hpp:750-787 and
// rnea-derivatives.hxx:324-333) and of.  The reference fills row blocks
// (idx_v, subtree) and column blocks (subtree+, idx_v) when it reaches the ancestor; here
// every entry pairing an ancestor dof r with a dof c is produced when c's subtree closes,
// from c's column forces (dFda_c, dFdq_c, dFdv_c, dYtJ_c) and r's column data, so the
// column forces never have to be stored:
//   dtau_dq[r][c] = J_r . (dFdq_c + J_c x* of_c)   dtau_dq[c][r] = dFda_c . dAdq_r + dYtJ_c . dVdq_r
//   dtau_dv[r][c] = J_r . dFdv_c                    dtau_dv[c][r] = dFda_c . dAdv_r + dYtJ_c . J_r
//   M[r][c] = M[c][r] = J_r . dFda_c
// and the entries inside one joint's block use dFdq_c before the J_c x* of_c term, as the
// reference does.
//
// As in CRBA (k_crba.hpp), the world frame's origin is put at the floating base: the
// outputs are joint-space quantities, and composite inertias about a nearby origin keep
// more FP32 bits (humanoid, 16384 configurations: ddq_dq max error 1.5e-4 -> 3.6e-5).
//
// computeABADerivatives (aba-derivatives.hxx:380-453) returns ddq_dq = -Minv dtau_dq,
// ddq_dv = -Minv dtau_dv, ddq_dtau = Minv with dtau evaluated at a = ddq (the identity
// its own Bwd2 pass implements).  The engine runs ABA (k_aba), this RNEA-derivative
// kernel at a = ddq, then k_abad_solve, which factors M = U D U^T over the tree pattern in
// registers and applies M^-1 to the 2 nv derivative columns and the nv unit columns.
// FP32 throughout: an FP32 tree factor of the FP32 M gives Minv within 3.2e-6 of FP64 on
// the humanoid, against 2.8e-5 for the reference's FP32 Fcrb sweep.
#pragma once

#include "k_crba.hpp"

namespace rbd {

// time variation of a world-frame body inertia (origin form m, h, Io) along ov, plus the
// [oh x*] terms: the blocks (LIN,ANG), (ANG,LIN), (ANG,ANG); (LIN,LIN) is zero.
template<class S> struct DY {
  R3<S> LA, AL, AA;
};
template<class S> RBD_HD R3<S> skewM(V3<S> v) { return {S(0), -v.z, v.y, v.z, S(0), -v.x, -v.y, v.x, S(0)}; }
// [u]x [w]x = w u^T - (u.w) I
template<class S> RBD_HD R3<S> skewSq(V3<S> u, V3<S> w) {
  const S d = dot(u, w);
  return {w.x * u.x - d, w.x * u.y, w.x * u.z, w.y * u.x, w.y * u.y - d, w.y * u.z, w.z * u.x, w.z * u.y, w.z * u.z - d};
}
template<class S> RBD_HD R3<S> add(const R3<S>& a, const R3<S>& b) {
  return {a.m00 + b.m00, a.m01 + b.m01, a.m02 + b.m02, a.m10 + b.m10, a.m11 + b.m11,
          a.m12 + b.m12, a.m20 + b.m20, a.m21 + b.m21, a.m22 + b.m22};
}
template<class S> RBD_HD R3<S> sub(const R3<S>& a, const R3<S>& b) {
  return {a.m00 - b.m00, a.m01 - b.m01, a.m02 - b.m02, a.m10 - b.m10, a.m11 - b.m11,
          a.m12 - b.m12, a.m20 - b.m20, a.m21 - b.m21, a.m22 - b.m22};
}
template<class S> RBD_HD R3<S> transpose(const R3<S>& a) {
  return {a.m00, a.m10, a.m20, a.m01, a.m11, a.m21, a.m02, a.m12, a.m22};
}
template<class S> RBD_HD R3<S> scaleM(S s, const R3<S>& a) {
  return {s * a.m00, s * a.m01, s * a.m02, s * a.m10, s * a.m11, s * a.m12, s * a.m20, s * a.m21, s * a.m22};
}
template<class S> RBD_HD DY<S> inertiaVariation(const CInertia<S>& Y, const Mot<S>& v, const Frc<S>& oh) {
  const R3<S> Io = {Y.xx, Y.xy, Y.xz, Y.xy, Y.yy, Y.yz, Y.xz, Y.yz, Y.zz};
  const R3<S> W = skewM(v.a);
  // inertia.hpp:750-787 in origin form: m c = h
  R3<S> LA = sub(add(scaleM(-Y.m, skewM(v.l)), skewSq(Y.h, v.a)), skewSq(v.a, Y.h));
  R3<S> AA = sub(sub(scaleM(S(-1), skewSq(v.l, Y.h)), skewSq(Y.h, v.l)), mul(Io, W));
  AA = add(AA, mul(W, Io));
  const R3<S> AL = transpose(LA);
  const R3<S> Fl = skewM(-oh.l), Fa = skewM(-oh.a);
  return {add(LA, Fl), add(AL, Fl), add(AA, Fa)};
}
template<class S> RBD_HD void dyAdd(DY<S>& a, const DY<S>& b) {
  a.LA = add(a.LA, b.LA);
  a.AL = add(a.AL, b.AL);
  a.AA = add(a.AA, b.AA);
}
// D x with D = [[0, LA], [AL, AA]]
template<class S> RBD_HD Frc<S> dyMul(const DY<S>& D, const Mot<S>& x) {
  return {mul(D.LA, x.a), mul(D.AL, x.l) + mul(D.AA, x.a)};
}
// D^T x
template<class S> RBD_HD Frc<S> dyMulT(const DY<S>& D, const Mot<S>& x) {
  return {mulT(D.AL, x.a), mulT(D.LA, x.l) + mulT(D.AA, x.a)};
}
template<class S> RBD_HD S dotFF(const Frc<S>& a, const Mot<S>& b) { return dot(a.l, b.l) + dot(a.a, b.a); }

template<class S> struct ColData {
  Mot<S> J, dVdq, dAdq, dAdv;
};

// ---------------------------------------------------------------------------------------
// computeRNEADerivatives: dtau_dq, dtau_dv, M (any may be skipped) and optionally tau
// ---------------------------------------------------------------------------------------
// G > 1 (small batches): G waves share the tile and split the tree by subtrees
// (WaveSplit, as in ABA): a wave runs the events of its subtrees and of the replicated
// trunk, emits the entries of the joints it owns (wave 0 those of the trunk), and hands
// a subtree's composite inertia, its variation and its force (43 floats) to the trunk
// parent through an LDS slot read after a barrier.  This is the body of wave W.
constexpr int kRneadSlot = 43;
template<class TP, class S, int G, int W>
__device__ __forceinline__ void rneadTile(const S* __restrict__ mc_in, int64_t B, int64_t b0, CView<S> gq,
                                          CView<S> gv, CView<S> ga, View<S> gdq, View<S> gdv, View<S> gM,
                                          int writeMask, S* __restrict__ sq, S* __restrict__ sv, S* __restrict__ sa,
                                          S* __restrict__ xs) {
  constexpr auto& T = TP::T;
  constexpr int NJ = TP::NJ, NQ = T.nq, NV = T.nv;
  constexpr WaveSplit<TP, G> WS{};
  const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
  // the three inputs are staged by different waves when there are several
  if constexpr (W == 0) stageIn<NQ>(sq, gq, b0, nb);
  if constexpr (W == 1 % G) stageIn<NV>(sv, gv, b0, nb);
  if constexpr (W == 2 % G) stageIn<NV>(sa, ga, b0, nb);
  __syncthreads();
  const int L = laneId();
  const int64_t b = b0 + L;
  const bool live = L < nb;
#ifdef RBD_RNEAD_FORCE_MASK
  writeMask = RBD_RNEAD_FORCE_MASK;
#endif
  const bool wq = live && (writeMask & 1), wv = live && (writeMask & 2), wM = live && (writeMask & 4);
  const S* gr = mc_in + mcGravity<NJ, NV>();
  const Mot<S> g = {{gr[0], gr[1], gr[2]}, {gr[3], gr[4], gr[5]}};
  const Mot<S> a0 = {{-g.l.x, -g.l.y, -g.l.z}, {-g.a.x, -g.a.y, -g.a.z}};

  Xform<S> oM[NJ];
  Mot<S> ov[NJ], oa[NJ];  // world velocity and gravity-shifted acceleration
  ColData<S> cd[NJ];      // revolute joints only
  CInertia<S> Yacc[NJ];
  DY<S> DYacc[NJ];
  Frc<S> Facc[NJ];

  // column data of dof k of joint a (the root free-flyer's is derived on the fly)
  auto colData = [&](auto A, int k) -> ColData<S> {
    constexpr int a = decltype(A)::value;
    if constexpr (TP::T.type[a] == JT_FF) {
      const Mot<S> J = worldAxis<JT_FF>(oM[a], k);
      return {J, zeroMot<S>(), crossMM(a0, J), crossMM(ov[a], J)};
    } else {
      return cd[a];
    }
  };

  static_for<0, T.nev>([&](auto E) {
    constexpr int e = E;
    constexpr int i = TP::T.evJoint[e];
    if constexpr (WS.runs(i, W)) {
    stepFence();
    const S* __restrict__ mc = stepConsts(mc_in);
    constexpr int p = TP::T.parent[i];
    constexpr int ty = TP::T.type[i];
    constexpr int iq = TP::T.iq[i], iv = TP::T.iv[i], nvi = TP::T.nvj[i];
    if constexpr (TP::T.evKind[e] == EV_PRE) {
      if constexpr (ty == JT_FF) {
        static_assert(p == 0, "free-flyer joints are supported at the root only");
        S q7[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) q7[k] = sq[(iq + k) * kPitch + L];
        oM[i] = freeflyerX(mc, i, q7, false);
        oM[i].p = {S(0), S(0), S(0)};  // frame origin at the floating base
        Mot<S> vJ = zeroMot<S>(), aJ = zeroMot<S>();
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const Mot<S> J = worldAxis<JT_FF>(oM[i], k);
          vJ = vJ + scale(sv[(iv + k) * kPitch + L], J);
          aJ = aJ + scale(sa[(iv + k) * kPitch + L], J);
        }
        ov[i] = vJ;
        oa[i] = a0 + aJ;
      } else {
        S sn, cs;
        sincosT(sq[iq * kPitch + L], &sn, &cs);
        const Xform<S> X = revoluteX(mc, i, ty - JT_RX, sn, cs);
        if constexpr (p > 0) oM[i] = compose(oM[p], X);
        else oM[i] = X;
        const Mot<S> J = worldAxis<ty>(oM[i], 0);
        const S qd = sv[iv * kPitch + L], qdd = sa[iv * kPitch + L];
        const Mot<S> vJ = scale(qd, J);
        if constexpr (p > 0) {
          ov[i] = ov[p] + vJ;
          oa[i] = oa[p] + crossMM(ov[p], vJ) + scale(qdd, J);
          const Mot<S> dVdq = crossMM(ov[p], J);
          cd[i] = {J, dVdq, crossMM(oa[p], J) + crossMM(ov[p], dVdq), crossMM(ov[i], J) + dVdq};
        } else {
          ov[i] = vJ;
          oa[i] = a0 + scale(qdd, J);
          cd[i] = {J, zeroMot<S>(), crossMM(a0, J), crossMM(ov[i], J)};
        }
      }
    } else {
      // ---- CLOSE(i): composite quantities of the subtree ----
      CInertia<S> Y = worldInertia(mc, i, oM[i]);
      const Frc<S> oh = cmul(Y, ov[i]);
      Frc<S> of = cmul(Y, oa[i]) + crossMF(ov[i], oh);
      DY<S> DYc = inertiaVariation(Y, ov[i], oh);
      if constexpr (WS.nReg[i] > 0) {
        cadd(Y, Yacc[i]);
        dyAdd(DYc, DYacc[i]);
        of = of + Facc[i];
      }
      if constexpr (G > 1 && WS.ldsIn[i]) {
        // subtrees of other waves (fixed order, so every wave sums identically)
        __syncthreads();
        static_for<1, NJ>([&](auto Cc) {
          constexpr int c = Cc;
          if constexpr (TP::T.parent[c] == i && WS.slot[c] >= 0) {
            const S* x = xs + WS.slot[c] * kRneadSlot * kTile + L;
            CInertia<S> Yc;
            DY<S> Dc;
            Frc<S> Fc;
            S* yc = reinterpret_cast<S*>(&Yc);
            S* dc = reinterpret_cast<S*>(&Dc);
            S* fc = reinterpret_cast<S*>(&Fc);
#pragma unroll
            for (int k = 0; k < 10; ++k) yc[k] = x[k * kTile];
#pragma unroll
            for (int k = 0; k < 27; ++k) dc[k] = x[(10 + k) * kTile];
#pragma unroll
            for (int k = 0; k < 6; ++k) fc[k] = x[(37 + k) * kTile];
            cadd(Y, Yc);
            dyAdd(DYc, Dc);
            of = of + Fc;
          }
        });
      }
      if constexpr (WS.writer(i) == W) {
#pragma unroll
      for (int k = 0; k < nvi; ++k) {
        const int c = iv + k;
        const ColData<S> C = colData(IC<i>{}, k);
        const Frc<S> dFda = cmul(Y, C.J);
        Frc<S> dFdq = cmul(Y, C.dAdq);
        if constexpr (p > 0) dFdq = dFdq + dyMul(DYc, C.dVdq);
        const Frc<S> dFdv = dyMul(DYc, C.J) + cmul(Y, C.dAdv);
        const Frc<S> dYtJ = dyMulT(DYc, C.J);
        // entries inside this joint's block
#pragma unroll
        for (int k2 = 0; k2 < nvi; ++k2) {
          const int r = iv + k2;
          const Mot<S> Jr = colData(IC<i>{}, k2).J;
          if (wq) stMat(gdq, b, NV, r, c, dot(Jr, dFdq));
          if (wv) stMat(gdv, b, NV, r, c, dot(Jr, dFdv));
          if (wM) stMat(gM, b, NV, r, c, dot(Jr, dFda) + (r == c ? mc[mcArmature<NJ>() + c] : S(0)));
        }
        const Frc<S> dFdqA = dFdq + crossMF(C.J, of);
        // pairs with every ancestor dof
        static_for<1, NJ>([&](auto A) {
          constexpr int a = A;
          if constexpr (a != i && TP::T.pos[a] < TP::T.pos[i] && TP::T.pos[a] + TP::T.subsize[a] > TP::T.pos[i]) {
#pragma unroll
            for (int k2 = 0; k2 < TP::T.nvj[a]; ++k2) {
              const int r = TP::T.iv[a] + k2;
              const ColData<S> R = colData(A, k2);
              if (wq) {
                stMat(gdq, b, NV, r, c, dot(R.J, dFdqA));
                stMat(gdq, b, NV, c, r, dot(R.dAdq, dFda) + dot(R.dVdq, dYtJ));
              }
              if (wv) {
                stMat(gdv, b, NV, r, c, dot(R.J, dFdv));
                stMat(gdv, b, NV, c, r, dot(R.dAdv, dFda) + dot(R.J, dYtJ));
              }
              if (wM) {
                const S m = dot(R.J, dFda);
                stMat(gM, b, NV, r, c, m);
                stMat(gM, b, NV, c, r, m);
              }
            }
          }
        });
      }
      }
      if constexpr (p > 0) {
        if constexpr (G > 1 && WS.slot[i] >= 0) {
          S* x = xs + WS.slot[i] * kRneadSlot * kTile + L;
          const S* yc = reinterpret_cast<const S*>(&Y);
          const S* dc = reinterpret_cast<const S*>(&DYc);
          const S* fc = reinterpret_cast<const S*>(&of);
#pragma unroll
          for (int k = 0; k < 10; ++k) x[k * kTile] = yc[k];
#pragma unroll
          for (int k = 0; k < 27; ++k) x[(10 + k) * kTile] = dc[k];
#pragma unroll
          for (int k = 0; k < 6; ++k) x[(37 + k) * kTile] = fc[k];
        } else if constexpr (WS.regIdx[i] == 0) {
          Yacc[p] = Y;
          DYacc[p] = DYc;
          Facc[p] = of;
        } else {
          cadd(Yacc[p], Y);
          dyAdd(DYacc[p], DYc);
          Facc[p] = Facc[p] + of;
        }
      }
    }
    }
  });
  if (live) {
    static_for<0, NV>([&](auto C) {
      static_for<0, NV>([&](auto R) {
        if constexpr (!CrbaPattern<TP>::nz(R, C) && int(C) % G == W) {
          if (wq) stMat(gdq, b, NV, R, C, S(0));
          if (wv) stMat(gdv, b, NV, R, C, S(0));
          if (wM) stMat(gM, b, NV, R, C, S(0));
        }
      });
    });
  }
}

template<class TP, class S, int G = 1>
__global__ __launch_bounds__(64 * G) void k_rnead(const S* __restrict__ mc_in, int64_t B, CView<S> gq, CView<S> gv,
                                                  CView<S> ga, View<S> gdq, View<S> gdv, View<S> gM, int writeMask) {
  static_assert(G >= 1 && G <= 8, "1 to 8 waves per tile");
  constexpr int NQ = TP::T.nq, NV = TP::T.nv;
  constexpr WaveSplit<TP, G> WS{};
  __shared__ S sq[NQ * kPitch];
  __shared__ S sv[NV * kPitch];
  __shared__ S sa[NV * kPitch];
  __shared__ S xs[G > 1 && WS.nslots > 0 ? WS.nslots * kRneadSlot * kTile : 1];
  const int64_t b0 = int64_t(blockIdx.x) * kTile;
  if constexpr (G == 1) {
    rneadTile<TP, S, 1, 0>(mc_in, B, b0, gq, gv, ga, gdq, gdv, gM, writeMask, sq, sv, sa, xs);
  } else {
    const int w = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    static_for<0, G>([&](auto Wc) {
      constexpr int W = Wc;
      if (w == W) rneadTile<TP, S, G, W>(mc_in, B, b0, gq, gv, ga, gdq, gdv, gM, writeMask, sq, sv, sa, xs);
    });
  }
}

// ---------------------------------------------------------------------------------------
// ABA-derivative solve.  Input: the packed tree factor of M (k_chol, batch-fastest
// workspace, in the scalar of the outputs), dtau_dq in gXq and dtau_dv in gXv.  Output, in
// place: -Minv dtau_dq, -Minv dtau_dv, and Minv in gMinv.  Jobs (right-hand sides) before
// jobBegin are skipped: jobBegin = 2 nv computes Minv alone (computeMinverse).
//
// A workgroup owns a tile of consecutive configurations: their factors (nnz + nv scalars
// each) are staged in LDS as [entry][configuration], and its waves share out the 3 nv
// right-hand sides; every column load and store is a coalesced run of the batch-fastest
// layout and every factor read a conflict-free LDS row.  FP32: 64 configurations per
// tile, lane b of a wave solves configuration b0 + b, two right-hand sides at a time.
// FP64: 32 configurations per tile (the factor of 64 would not fit the LDS), the two
// halves of a wave on two right-hand sides of the same 32 configurations, one each.
// ---------------------------------------------------------------------------------------
constexpr int kSolveWaves = 12;  // 168 VGPRs: three waves per SIMD at one workgroup per CU

template<class S> struct SolveTile {
  static constexpr int CT = sizeof(S) == 4 ? kTile : kTile / 2;  // configurations per tile
  static constexpr int NR = sizeof(S) == 4 ? 2 : 1;              // right-hand sides per lane
};
template<class TP, class S> constexpr int abadSolveLdsBytes() {
  return (nnzOf<TP>() + TP::T.nv) * SolveTile<S>::CT * (int)sizeof(S);
}

template<class TP, class S>
__global__ __launch_bounds__(64 * kSolveWaves) void k_abad_solve(int64_t B, CView<S> gU, CView<S> gD, View<S> gXq,
                                                                 View<S> gXv, View<S> gMinv, int jobBegin) {
  constexpr auto& T = TP::T;
  constexpr int NV = T.nv;
  constexpr int NNZ = nnzOf<TP>();
  constexpr int CT = SolveTile<S>::CT, NR = SolveTile<S>::NR, HALVES = kTile / CT;
  using P = CholPattern<TP>;
  __shared__ S sU[(NNZ + NV) * CT];  // U entries then D^-1, [entry][configuration]
  const int64_t b0 = int64_t(blockIdx.x) * CT;
  const int nb = (int)(B - b0 < CT ? B - b0 : CT);
  const int L = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int cl = L % CT, h = L / CT;
  // staged kStageU entries per thread at a time, all loads issued before the first is
  // used (lanes past the batch end read the last configuration's entry and store 1)
  {
    constexpr int kTot = (NNZ + NV) * CT, kStep = 64 * kSolveWaves;
    static_assert(kStep % CT == 0, "a thread's entries share its configuration");
    const int b = threadIdx.x % CT;
    const int bc = b < nb ? b : nb - 1;
    static_for<0, (kTot + kStep * kStageU - 1) / (kStep * kStageU)>([&](auto I) {
      S x[kStageU];
      static_for<0, kStageU>([&](auto U) {
        const int e = (int(I) * kStageU + int(U)) * kStep + (int)threadIdx.x;
        const int k = (e < kTot ? e : kTot - 1) / CT;
        x[U] = k < NNZ ? gU.p[(int64_t)k * gU.rs + (b0 + bc) * gU.bs] : gD.p[(int64_t)(k - NNZ) * gD.rs + (b0 + bc) * gD.bs];
      });
      static_for<0, kStageU>([&](auto U) {
        const int e = (int(I) * kStageU + int(U)) * kStep + (int)threadIdx.x;
        if (e < kTot) {
          const int k = e / CT;
          sU[e] = b >= nb ? S(1) : (k < NNZ ? x[U] : S(1) / x[U]);
        }
      });
    });
  }
  __syncthreads();
  const bool live = cl < nb;
  const int64_t b = b0 + cl;

  // right-hand sides two (FP32) or one (FP64) at a time: x <- M^-1 x (Uiv, D^-1, Utiv;
  // cholesky.hxx:466-520).  A compiler memory clobber at every elimination step keeps the
  // factor reads where they are: without it they are hoisted out of the job loop and all
  // 337 values stay live.
  const S* su = sU + cl;
  // U^-1 runs by columns: once x[jc] is final it updates its ancestors' entries
  // (independent multiply-adds, where the reference's row dot products chain).
  auto solve2 = [&](S (&x)[NV], S (&y)[NV]) {
    static_for<0, NV>([&](auto JJ) {
      constexpr int jc = NV - 1 - JJ;
      stepFence();
      asm volatile("" ::: "memory");
      static_for<0, jc>([&](auto II) {
        constexpr int i = II;
        if constexpr (P::anc(i, jc)) {
          constexpr int pk = P::packed(i, jc);
          const S u = su[pk * CT];
          x[i] -= u * x[jc];
          y[i] -= u * y[jc];
        }
      });
      pin(x[jc], y[jc]);
    });
    asm volatile("" ::: "memory");
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const S di = su[(NNZ + k) * CT];
      x[k] *= di;
      y[k] *= di;
    }
    static_for<0, NV>([&](auto KK) {
      constexpr int k = KK;
      stepFence();
      asm volatile("" ::: "memory");
      static_for<k + 1, NV>([&](auto EE) {
        constexpr int e2 = EE;
        if constexpr (P::anc(k, e2)) {
          constexpr int pk = P::packed(k, e2);
          const S u = su[pk * CT];
          x[e2] -= u * x[k];
          y[e2] -= u * y[k];
          pin(x[e2], y[e2]);
        }
      });
    });
  };
  auto solve1 = [&](S (&x)[NV]) {
    static_for<0, NV>([&](auto JJ) {
      constexpr int jc = NV - 1 - JJ;
      stepFence();
      asm volatile("" ::: "memory");
      static_for<0, jc>([&](auto II) {
        constexpr int i = II;
        if constexpr (P::anc(i, jc)) {
          constexpr int pk = P::packed(i, jc);
          x[i] -= su[pk * CT] * x[jc];
        }
      });
      pin(x[jc]);
    });
    asm volatile("" ::: "memory");
#pragma unroll
    for (int k = 0; k < NV; ++k) x[k] *= su[(NNZ + k) * CT];
    static_for<0, NV>([&](auto KK) {
      constexpr int k = KK;
      stepFence();
      asm volatile("" ::: "memory");
      static_for<k + 1, NV>([&](auto EE) {
        constexpr int e2 = EE;
        if constexpr (P::anc(k, e2)) {
          constexpr int pk = P::packed(k, e2);
          x[e2] -= su[pk * CT] * x[k];
          pin(x[e2]);
        }
      });
    });
  };
  // job j < NV: column j of dtau_dq; < 2 NV: of dtau_dv; else unit column j - 2 NV
  // the three outputs share one layout (same rs, bs): select the base pointer only
  const int64_t rs = gXq.rs, lane = b * gXq.bs;
  auto colPtr = [&](int j) -> S* {
    S* base = j < NV ? gXq.p : (j < 2 * NV ? gXv.p : gMinv.p);
    const int c = j < NV ? j : (j < 2 * NV ? j - NV : j - 2 * NV);
    return base + (int64_t)(c * NV) * rs + lane;
  };
  auto load = [&](int j, S (&x)[NV]) {
    if (j < 2 * NV && live) {
      const S* p = colPtr(j);
#pragma unroll
      for (int r = 0; r < NV; ++r) x[r] = p[r * rs];
    } else {
      const int u = j - 2 * NV;
#pragma unroll
      for (int r = 0; r < NV; ++r) x[r] = (r == u) ? S(1) : S(0);
    }
  };
  auto store = [&](int j, const S (&x)[NV]) {
    if (!live || j >= 3 * NV) return;
    S* p = colPtr(j);
    const S sgn = j < 2 * NV ? S(-1) : S(1);
#pragma unroll
    for (int r = 0; r < NV; ++r) stOut(&p[r * rs], sgn * x[r]);
  };
  if constexpr (NR == 2) {
#pragma unroll 1
    for (int j = jobBegin + 2 * w; j < 3 * NV; j += 2 * kSolveWaves) {
      S x[NV], y[NV];
      load(j, x);
      load(j + 1 < 3 * NV ? j + 1 : j, y);
      solve2(x, y);
      store(j, x);
      if (j + 1 < 3 * NV) store(j + 1, y);
    }
  } else {
#pragma unroll 1
    for (int j = jobBegin + HALVES * w + h; j < 3 * NV; j += HALVES * kSolveWaves) {
      S x[NV];
      load(j, x);
      solve1(x);
      store(j, x);
    }
  }
}

}  // namespace rbd
