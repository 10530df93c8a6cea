// k_crba.hpp -- batched joint-space inertia matrix (CRBA) and tree-sparse Cholesky.
//
// CRBA follows crbaWorldConvention (crba.hxx:20-99, 498-548): world-frame joint axes
// J_c = oMi.act(S_c) and composite inertias Ycrb (additive in the world frame).  Instead of
// the reference's row blocks M[i, subtree(i)] = J_i^T Ycrb_i J[:, subtree] (which keep
// a 6 x nv Ag matrix), the column of each dof c is produced when its subtree closes:
// F_c = Ycrb_j J_c and M[r, c] = J_r . F_c for every ancestor-or-self dof r -- the same
// products, grouped so that only the open path (oMi per joint) is live.
//
// The "world" frame of the sweep has its origin at the floating base of the subtree (the
// free flyer's translation is dropped; the rotation is kept): M is a joint-space quantity
// and does not depend on that choice, but composite inertias taken about a far origin
// carry parallel-axis terms m |p|^2 that cost FP32 bits (humanoid, base within 1 m of the
// origin: max error of M 1.0e-6 -> 2.9e-7; the ABA-derivative pipeline's M^-1 7.1e-5 ->
// 1.9e-5 over 16384 configurations).
//
// cholesky::decompose (cholesky.hxx:17-62) runs in FP64 as a right-looking elimination
// over the tree pattern: column j (from nv-1 down to 0) yields D_j = A_jj and
// U_rj = A_rj / D_j for its ancestor rows, then updates A_rr' -= U_rj A_r'j for ancestor
// pairs.  A entries are loaded lazily on first touch, so the live set is the pattern of
// the open path, not the whole matrix.
#pragma once

#include "kernels_common.hpp"

namespace rbd {

// world-frame axis of dof k (0..nvj-1) of joint i with world transform oM
template<int TY, class S> __device__ __forceinline__ Mot<S> worldAxis(const Xform<S>& oM, int k) {
  if constexpr (TY == JT_FF) {
    if (k < 3) return {col(oM.R, k), {S(0), S(0), S(0)}};
    const V3<S> w = col(oM.R, k - 3);
    return {cross(oM.p, w), w};
  } else {
    const V3<S> w = col(oM.R, TY - JT_RX);
    return {cross(oM.p, w), w};
  }
}

// composite inertia of body j expressed in the world (origin form)
template<class S> __device__ __forceinline__ CInertia<S> worldInertia(const S* __restrict__ mc, int j, const Xform<S>& oM) {
  return transportUp(oM, mcComposite(mc, j));
}

template<class TP> struct CrbaPattern {
  static constexpr auto& T = TP::T;
  static constexpr int NV = T.nv;
  // structural non-zero of M: r ancestor-or-self of c (dof tree)
  static constexpr bool nz(int r, int c) {
    if (r > c) return nz(c, r);
    for (int x = c; x >= 0; x = T.dofParent[x])
      if (x == r) return true;
    return false;
  }
};

// Matrix outputs are written once and not read again by the kernel: nontemporal stores
// stream them to HBM without displacing cached inputs (CRBA at 65536 configurations:
// 114 -> 85 us on MI355X).
template<class S> __device__ __forceinline__ void stOut(S* p, S x) { __builtin_nontemporal_store(x, p); }
template<class S> __device__ __forceinline__ void stMat(View<S> g, int64_t b, int nv, int r, int c, S x) {
  stOut(&g.p[(int64_t)(r + c * nv) * g.rs + b * g.bs], x);
}

// ---------------------------------------------------------------------------------------
// CRBA: M(q), symmetric (or upper row blocks only), armature on the diagonal
// ---------------------------------------------------------------------------------------
// G > 1: G waves share the tile; all run the forward pass and the composite-inertia
// accumulation (cheap), and joint i's entries are computed and stored by wave
// EmitGroups<TP, G>::grp[i] only, so the store-heavy emission is spread over more waves.
// One 64-configuration tile starting at b0, run by G waves (w: this wave); `sq` holds
// NQ * kPitch scalars of LDS.
template<class TP, class S, int G>
__device__ __forceinline__ void crbaTile(const S* __restrict__ mc_in, int64_t B, int64_t b0, CView<S> gq, View<S> gM,
                                         int upperOnly, S* __restrict__ sq, int w) {
  constexpr auto& T = TP::T;
  constexpr int NJ = TP::NJ, NQ = T.nq, NV = T.nv;
  constexpr EmitGroups<TP, G> EG{};
  const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
  if constexpr (G == 1) {
    stageIn<NQ>(sq, gq, b0, nb);
  } else {  // two waves stage half of q each
    constexpr int NQh = (NQ + 1) / 2;
    if (w == 0) stageIn<NQh>(sq, gq, b0, nb);
    else if (w == 1) stageIn<NQ - NQh>(sq + NQh * kPitch, CView<S>{gq.p + (int64_t)NQh * gq.rs, gq.rs, gq.bs}, b0, nb);
  }
  __syncthreads();
  const int L = laneId();
  const int64_t b = b0 + L;
  const bool live = L < nb;

  Xform<S> oM[NJ];
  CInertia<S> Yacc[NJ];

  static_for<0, T.nev>([&](auto E) {
    constexpr int e = E;
    stepFence();
    const S* __restrict__ mc = stepConsts(mc_in);
    constexpr int i = TP::T.evJoint[e];
    constexpr int p = TP::T.parent[i];
    constexpr int ty = TP::T.type[i];
    constexpr int iq = TP::T.iq[i], iv = TP::T.iv[i], nvi = TP::T.nvj[i];
    if constexpr (TP::T.evKind[e] == EV_PRE) {
      Xform<S> X;
      if constexpr (ty == JT_FF) {
        S q7[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) q7[k] = sq[(iq + k) * kPitch + L];
        X = freeflyerX(mc, i, q7, false);
        X.p = {S(0), S(0), S(0)};  // frame origin at the floating base (see the header)
      } else {
        S sn, cs;
        sincosT(sq[iq * kPitch + L], &sn, &cs);
        X = revoluteX(mc, i, ty - JT_RX, sn, cs);
      }
      if constexpr (p > 0) oM[i] = compose(oM[p], X);
      else oM[i] = X;
    } else {
      CInertia<S> Y = worldInertia(mc, i, oM[i]);
      if constexpr (TP::T.nchild[i] > 0) cadd(Y, Yacc[i]);
      if (G == 1 || w == EG.grp[i]) {
#pragma unroll
      for (int k = 0; k < nvi; ++k) {
        const int c = iv + k;
        const Frc<S> F = cmul(Y, worldAxis<ty>(oM[i], k));
        // own block (all dofs of this joint): one product per pair, written to both
        // triangles (the reference writes the whole row block, so a free-flyer's 6x6
        // block is full even in upper-only mode)
#pragma unroll
        for (int k2 = 0; k2 <= k; ++k2) {
          const int r = iv + k2;
          S m = dot(worldAxis<ty>(oM[i], k2), F);
          if (r == c) m += mc[mcArmature<NJ>() + c];
          if (live) {
            stMat(gM, b, NV, r, c, m);
            if (r != c) stMat(gM, b, NV, c, r, m);
          }
        }
        // ancestors
        static_for<1, NJ>([&](auto A) {
          constexpr int a = A;
          if constexpr (a != i && TP::T.pos[a] < TP::T.pos[i] && TP::T.pos[a] + TP::T.subsize[a] > TP::T.pos[i]) {
            constexpr int tya = TP::T.type[a];
#pragma unroll
            for (int k2 = 0; k2 < TP::T.nvj[a]; ++k2) {
              const int r = TP::T.iv[a] + k2;
              const S m = dot(worldAxis<tya>(oM[a], k2), F);
              if (live) {
                stMat(gM, b, NV, r, c, m);
                if (!upperOnly) stMat(gM, b, NV, c, r, m);
              }
            }
          }
        });
      }
      }
      if constexpr (p > 0) {
        if constexpr (TP::T.childIdx[i] == 0) Yacc[p] = Y;
        else cadd(Yacc[p], Y);
      }
    }
  });
  // structural zeros (upper triangle only in upper-only mode)
  if (live) {
    static_for<0, NV>([&](auto C) {
      static_for<0, NV>([&](auto R) {
        if constexpr (!CrbaPattern<TP>::nz(R, C)) {
          if ((G == 1 || w == (int)C % G) && (!upperOnly || R <= C)) stMat(gM, b, NV, R, C, S(0));
        }
      });
    });
  }
}

template<class TP, class S, int G = 1>
__global__ __launch_bounds__(64 * G) void k_crba(const S* __restrict__ mc_in, int64_t B, CView<S> gq, View<S> gM,
                                                 int upperOnly) {
  __shared__ S sq[TP::T.nq * kPitch];
  const int w = G > 1 ? (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) : 0;
  crbaTile<TP, S, G>(mc_in, B, int64_t(blockIdx.x) * kTile, gq, gM, upperOnly, sq, w);
}

// ---------------------------------------------------------------------------------------
// FP64 tree-sparse Cholesky M = U D U^T (cholesky::decompose), M read from its upper
// structural entries (any dtype), U written packed (strictly-upper pattern, column by
// column, rows from the root down) and D as nv doubles.
// ---------------------------------------------------------------------------------------
// number of strictly-upper structural non-zeros of M (the packed factor's length)
template<class TP> constexpr int nnzOf() {
  int n = 0;
  for (int c = 0; c < TP::T.nv; ++c)
    for (int x = TP::T.dofParent[c]; x >= 0; x = TP::T.dofParent[x]) ++n;
  return n;
}

template<class TP> struct CholPattern {
  static constexpr auto& T = TP::T;
  static constexpr int NV = T.nv;
  static constexpr int depthOf(int c) {
    int d = 0;
    for (int x = T.dofParent[c]; x >= 0; x = T.dofParent[x]) ++d;
    return d;
  }
  // packed index of (r, c), r a strict ancestor of c; rows listed from the root down
  static constexpr int packed(int r, int c) {
    int base = 0;
    for (int k = 0; k < c; ++k) base += depthOf(k);
    int pos = 0;
    for (int x = T.dofParent[c]; x >= 0; x = T.dofParent[x])
      if (x > r) ++pos;
    return base + (depthOf(c) - 1 - pos);
  }
  static constexpr bool anc(int r, int c) {  // r strict ancestor of c
    for (int x = T.dofParent[c]; x >= 0; x = T.dofParent[x])
      if (x == r) return true;
    return false;
  }
  // first column (in elimination order nv-1 .. 0) that touches A[r][c] (r <= c):
  // the largest k >= c in the subtree of c (for r < c, k must also descend from r)
  static constexpr int firstTouch(int r, int c) {
    int best = c;
    for (int k = NV - 1; k > c; --k)
      if (anc(c, k)) {
        best = k;
        break;
      }
    (void)r;
    return best;
  }
};

// SC is the arithmetic (and output) type of the factor: double for cholesky::decompose,
// float inside the ABA-derivative pipeline.
template<class TP, class SIN, class SC>
__global__ __launch_bounds__(64) void k_chol(int64_t B, CView<SIN> gM, View<SC> gU, View<SC> gD) {
  constexpr auto& T = TP::T;
  constexpr int NV = T.nv;
  using P = CholPattern<TP>;
  const int L = laneId();
  const int64_t b = int64_t(blockIdx.x) * kTile + L;
  if (b >= B) return;
  SC A[NV][NV];  // only pattern entries (r <= c, r ancestor-or-self of c) are materialised
  auto ldM = [&](int r, int c) { return SC(gM.p[(int64_t)(r + c * NV) * gM.rs + b * gM.bs]); };
  static_for<0, NV>([&](auto JJ) {
    constexpr int j = NV - 1 - JJ;
    stepFence();
    // lazily load the column-j entries and the pairs this column updates
    if constexpr (P::firstTouch(j, j) == j) A[j][j] = ldM(j, j);
    static_for<0, NV>([&](auto RR) {
      constexpr int r = RR;
      if constexpr (P::anc(r, j) && P::firstTouch(r, j) == j) A[r][j] = ldM(r, j);
    });
    const SC Dj = A[j][j];
    const SC Dinv = SC(1) / Dj;
    stOut(&gD.p[(int64_t)j * gD.rs + b * gD.bs], Dj);
    SC Uc[NV];
    static_for<0, NV>([&](auto RR) {
      constexpr int r = RR;
      if constexpr (P::anc(r, j)) {
        Uc[r] = A[r][j] * Dinv;
        constexpr int pk = P::packed(r, j);
        stOut(&gU.p[(int64_t)pk * gU.rs + b * gU.bs], Uc[r]);
      }
    });
    // rank-1 update of the ancestor block: A[r][r2] -= U[r] * A[r2][j]  (r <= r2)
    static_for<0, NV>([&](auto R2) {
      constexpr int r2 = R2;
      if constexpr (P::anc(r2, j)) {
        // A[r2][r2] and A[r][r2] for r ancestor of r2 (both ancestors of j)
        if constexpr (P::firstTouch(r2, r2) == j) A[r2][r2] = ldM(r2, r2);
        A[r2][r2] -= Uc[r2] * A[r2][j];
        static_for<0, NV>([&](auto R1) {
          constexpr int r1 = R1;
          if constexpr (P::anc(r1, r2)) {
            if constexpr (P::firstTouch(r1, r2) == j) A[r1][r2] = ldM(r1, r2);
            A[r1][r2] -= Uc[r1] * A[r2][j];
          }
        });
      }
    });
  });
}

// The same factorisation with M's pattern entries staged into LDS first, kStageU loads in
// flight per lane, instead of lazily per column: for small batches (one tile per CU) the
// per-column loads otherwise expose the memory latency once per column.  LDS holds the
// strictly-upper pattern entries in packed order, then the diagonal; [entry][lane].
template<class TP> struct CholEntries {
  int row[nnzOf<TP>() + TP::T.nv > 0 ? nnzOf<TP>() + TP::T.nv : 1] = {};
  int col[nnzOf<TP>() + TP::T.nv > 0 ? nnzOf<TP>() + TP::T.nv : 1] = {};
  constexpr CholEntries() {
    using P = CholPattern<TP>;
    constexpr int NV = TP::T.nv;
    for (int c = 0; c < NV; ++c) {
      for (int r = 0; r < c; ++r)
        if (P::anc(r, c)) {
          row[P::packed(r, c)] = r;
          col[P::packed(r, c)] = c;
        }
      row[nnzOf<TP>() + c] = c;
      col[nnzOf<TP>() + c] = c;
    }
  }
};
template<class TP, class SIN> constexpr int cholStagedLdsBytes() {
  return (nnzOf<TP>() + TP::T.nv) * kTile * (int)sizeof(SIN);
}
template<class TP, class SIN, class SC>
__global__ __launch_bounds__(64) void k_chol_staged(int64_t B, CView<SIN> gM, View<SC> gU, View<SC> gD) {
  constexpr auto& T = TP::T;
  constexpr int NV = T.nv, NNZ = nnzOf<TP>(), NE = NNZ + NV;
  using P = CholPattern<TP>;
  constexpr CholEntries<TP> CE{};
  __shared__ SIN sM[NE * kTile];
  const int L = laneId();
  const int64_t b0 = int64_t(blockIdx.x) * kTile;
  const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
  {
    const int64_t bc = b0 + (L < nb ? L : nb - 1);
    static_for<0, (NE + kStageU - 1) / kStageU>([&](auto I) {
      SIN x[kStageU];
      static_for<0, kStageU>([&](auto U) {
        constexpr int k = int(I) * kStageU + int(U);
        if constexpr (k < NE) x[U] = gM.p[(int64_t)(CE.row[k] + CE.col[k] * NV) * gM.rs + bc * gM.bs];
      });
      static_for<0, kStageU>([&](auto U) {
        constexpr int k = int(I) * kStageU + int(U);
        if constexpr (k < NE) sM[k * kTile + L] = x[U];
      });
    });
  }
  // every lane reads only its own LDS column: no barrier needed between the two phases
  if (L >= nb) return;
  const int64_t b = b0 + L;
  SC A[NV][NV];
  auto ldM = [&](auto R, auto C) {
    constexpr int r = decltype(R)::value, c = decltype(C)::value;
    constexpr int k = r == c ? NNZ + c : P::packed(r, c);
    return SC(sM[k * kTile + L]);
  };
  static_for<0, NV>([&](auto JJ) {
    constexpr int j = NV - 1 - JJ;
    stepFence();
    if constexpr (P::firstTouch(j, j) == j) A[j][j] = ldM(IC<j>{}, IC<j>{});
    static_for<0, NV>([&](auto RR) {
      constexpr int r = RR;
      if constexpr (P::anc(r, j) && P::firstTouch(r, j) == j) A[r][j] = ldM(IC<r>{}, IC<j>{});
    });
    const SC Dj = A[j][j];
    const SC Dinv = SC(1) / Dj;
    stOut(&gD.p[(int64_t)j * gD.rs + b * gD.bs], Dj);
    SC Uc[NV];
    static_for<0, NV>([&](auto RR) {
      constexpr int r = RR;
      if constexpr (P::anc(r, j)) {
        Uc[r] = A[r][j] * Dinv;
        constexpr int pk = P::packed(r, j);
        stOut(&gU.p[(int64_t)pk * gU.rs + b * gU.bs], Uc[r]);
      }
    });
    static_for<0, NV>([&](auto R2) {
      constexpr int r2 = R2;
      if constexpr (P::anc(r2, j)) {
        if constexpr (P::firstTouch(r2, r2) == j) A[r2][r2] = ldM(IC<r2>{}, IC<r2>{});
        A[r2][r2] -= Uc[r2] * A[r2][j];
        static_for<0, NV>([&](auto R1) {
          constexpr int r1 = R1;
          if constexpr (P::anc(r1, r2)) {
            if constexpr (P::firstTouch(r1, r2) == j) A[r1][r2] = ldM(IC<r1>{}, IC<r2>{});
            A[r1][r2] -= Uc[r1] * A[r2][j];
          }
        });
      }
    });
  });
}

// ---------------------------------------------------------------------------------------
// x <- M^-1 x with a packed FP64 factor (cholesky::solve = Uiv, D^-1, Utiv,
// cholesky.hxx:193-317, 466-520); arithmetic in FP64, x in S.
// ---------------------------------------------------------------------------------------
template<class TP, class S>
__global__ __launch_bounds__(64) void k_chol64_solve(int64_t B, CView<double> gU, CView<double> gD, View<S> gx) {
  constexpr auto& T = TP::T;
  constexpr int NV = T.nv;
  using P = CholPattern<TP>;
  const int L = laneId();
  const int64_t b = int64_t(blockIdx.x) * kTile + L;
  if (b >= B) return;
  double x[NV];
  S* __restrict__ px = gx.p + b * gx.bs;
  const double* __restrict__ pU = gU.p + b * gU.bs;
  const double* __restrict__ pD = gD.p + b * gD.bs;
#pragma unroll
  for (int k = 0; k < NV; ++k) x[k] = double(px[k * gx.rs]);
  static_for<0, NV>([&](auto KK) {
    constexpr int k = NV - 1 - KK;
    stepFence();
    const double* __restrict__ pUk = laneLaunder(pU);
    static_for<k + 1, NV>([&](auto EE) {
      constexpr int e2 = EE;
      if constexpr (P::anc(k, e2)) {
        constexpr int pk = P::packed(k, e2);
        x[k] -= pUk[pk * gU.rs] * x[e2];
      }
    });
  });
#pragma unroll
  for (int k = 0; k < NV; ++k) x[k] /= pD[k * gD.rs];
  static_for<0, NV>([&](auto KK) {
    constexpr int k = KK;
    stepFence();
    const double* __restrict__ pUk = laneLaunder(pU);
    static_for<k + 1, NV>([&](auto EE) {
      constexpr int e2 = EE;
      if constexpr (P::anc(k, e2)) {
        constexpr int pk = P::packed(k, e2);
        x[e2] -= pUk[pk * gU.rs] * x[k];
      }
    });
  });
#pragma unroll
  for (int k = 0; k < NV; ++k) px[k * gx.rs] = S(x[k]);
}

}  // namespace rbd
