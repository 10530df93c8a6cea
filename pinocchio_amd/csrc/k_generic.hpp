// k_generic.hpp -- run-time-topology kernels: RNEA, ABA, CRBA, the tree Cholesky and its
// solves, RNEA derivatives, the ABA-derivative solve and forward kinematics for ANY tree of
// revolute X/Y/Z joints with free-flyer roots, in FP32 or FP64.
//
// The prebuilt kernels (k_dyn.hpp, k_crba.hpp, k_deriv.hpp) unroll a tree that is known at
// compile time, so every per-joint quantity is a named register.  These kernels read the
// tree as data instead -- the same model a caller builds with ModelTpl::addJoint
// (multibody/model.hxx:61-170) -- and follow the same depth-first PRE / CLOSE event
// schedule:
//
//   PRE(i)   forward-pass work of joint i (DFS pre-order);
//   CLOSE(i) backward-pass work of joint i once its whole subtree has been visited.
//
// All lanes of a wave walk the same tree, so the event index, the joint record and every
// table entry are wave-uniform (scalar loads, uniform branches).  Only the open path of
// the tree is live at any time; its per-joint state is kept in "lane memory" indexed by
// the joint's depth ("level"; a child's level is its parent's plus one), and per-dof
// column data by dof depth.  Lane memory is [slot][lane] for the 64 lanes of a wave, so a
// wave's access to one slot is one contiguous 256- (FP32) or 512-byte (FP64) row: either
// LDS (when the model's state fits) or a per-workgroup region of a device buffer.
//
// CRBA and the RNEA derivatives put the origin of their world frame at each floating base,
// as the prebuilt kernels do (k_crba.hpp): their outputs are joint-space quantities.
//
// Algorithms and formulations are those of the prebuilt kernels: local-frame RNEA
// (rnea.hxx:17-161) and ABA with the bias re-associated down the tree (aba.hxx:79-293),
// world-frame CRBA (crba.hxx:20-99, 498-548) and RNEA derivatives
// (rnea-derivatives.hxx:242-541) emitting each (ancestor, dof) entry at the dof's CLOSE,
// the right-looking tree factorisation M = U D U^T (cholesky.hxx:17-62) and its solves
// (cholesky.hxx:193-317, 466-520).
#pragma once

#include "k_deriv.hpp"
#include "k_dyn.hpp"

namespace rbd {

// Device tables of one model's topology (built on the host at model upload).
struct GProg {
  const int32_t* ev;       // DFS event schedule: (joint << 1) | kind (EV_PRE / EV_CLOSE)
  const int32_t* jr;       // GJ_REC ints per joint
  const int32_t* colPtr;   // nv + 1: column c of the packed strictly-upper pattern is [colPtr[c], colPtr[c+1])
  const int32_t* patRow;   // nnz: row of each packed entry (ancestor dofs of the column, root first)
  const uint32_t* nzMask;  // nzW words per column: bit r set where (r, c) is a structural non-zero of M
  int nev, nj, nq, nv, nlev, nvlev, nnz, nzW;
};
// joint record: parent, type, first q / v index, dof count, level (depth - 1), dof level
// (number of strict-ancestor dofs), number of children
enum : int { GJ_PARENT = 0, GJ_TYPE, GJ_IQ, GJ_IV, GJ_NV, GJ_LEVEL, GJ_DOFLEVEL, GJ_NCHILD, GJ_REC };

// slots per level / per dof level of each kernel
constexpr int kGRneaLvl = 18;   // v, a (local), f
constexpr int kGAbaLvl = 39;    // v, chat, articulated inertia (21), bias force (6)
constexpr int kGAbaJnt = 7;     // per joint: U D^-1 (6), D^-1 u - chat . U D^-1 (FF root: its ahat)
constexpr int kGCrbaLvl = 22;   // oMi (12), composite inertia (10)
constexpr int kGCrbaDof = 6;    // world axis of the dof
constexpr int kGRneadLvl = 67;  // oMi, ov, oa, Y (10), dY (27), of
constexpr int kGRneadDof = 24;  // J, dV/dq, dA/dq, dA/dv of the dof
constexpr int kGKinLvl = 24;    // oMi, v, a

// This lane's view of lane memory: slot k at p[k * kTile].
template<class S> struct Lane {
  S* p;
  __device__ __forceinline__ S ld(int k) const { return p[k * kTile]; }
  __device__ __forceinline__ void st(int k, S x) const { p[k * kTile] = x; }
  template<class T> __device__ __forceinline__ T get(int k0) const {
    static_assert(sizeof(T) % sizeof(S) == 0, "struct of scalars expected");
    T o;
    S* f = reinterpret_cast<S*>(&o);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / sizeof(S)); ++k) f[k] = p[(k0 + k) * kTile];
    return o;
  }
  template<class T> __device__ __forceinline__ void put(int k0, const T& o) const {
    const S* f = reinterpret_cast<const S*>(&o);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / sizeof(S)); ++k) p[(k0 + k) * kTile] = f[k];
  }
  template<class T> __device__ __forceinline__ void add(int k0, const T& o) const {
    const S* f = reinterpret_cast<const S*>(&o);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / sizeof(S)); ++k) p[(k0 + k) * kTile] += f[k];
  }
};

// Lane memory of this workgroup's wave `w`: dynamic LDS (kLds) or the workgroup's region
// of `ws` (`stride` scalars per wave).
template<class S, bool kLds>
__device__ __forceinline__ Lane<S> laneMem(S* __restrict__ ws, int64_t stride, int w = 0) {
  if constexpr (kLds) {
    extern __shared__ __align__(16) unsigned char g_lds[];
    return {reinterpret_cast<S*>(g_lds) + stride * w + laneId()};
  } else {
    return {ws + ((int64_t)blockIdx.x * (blockDim.x >> 6) + w) * stride + laneId()};
  }
}

template<class S> __device__ __forceinline__ S gin(CView<S> g, int k, int64_t b) {
  return g.p[(int64_t)k * g.rs + b * g.bs];
}
template<class S> __device__ __forceinline__ S& gat(View<S> g, int k, int64_t b) {
  return g.p[(int64_t)k * g.rs + b * g.bs];
}

// liMi of joint i (placement * joint transform) from this lane's q
template<class S>
__device__ __forceinline__ Xform<S> gJointX(const S* __restrict__ mc, int i, int ty, CView<S> gq, int iq, int64_t b) {
  if (ty == JT_FF) {
    S q7[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) q7[k] = gin(gq, iq + k, b);
    return freeflyerX(mc, i, q7, false);
  }
  S sn, cs;
  sincosT(gin(gq, iq, b), &sn, &cs);
  return revoluteX(mc, i, ty - JT_RX, sn, cs);
}
// S x for the joint's slice x of a tangent vector (joint velocity / acceleration)
template<class S> __device__ __forceinline__ Mot<S> gJointMot(int ty, CView<S> g, int iv, int64_t b) {
  if (ty == JT_FF)
    return {{gin(g, iv, b), gin(g, iv + 1, b), gin(g, iv + 2, b)},
            {gin(g, iv + 3, b), gin(g, iv + 4, b), gin(g, iv + 5, b)}};
  return axisMot(ty - JT_RX, gin(g, iv, b));
}
// world-frame axis of dof k of a joint of type ty with world placement oM
template<class S> __device__ __forceinline__ Mot<S> gWorldAxis(int ty, const Xform<S>& oM, int k) {
  if (ty == JT_FF) {
    if (k < 3) return {col(oM.R, k), {S(0), S(0), S(0)}};
    const V3<S> w = col(oM.R, k - 3);
    return {cross(oM.p, w), w};
  }
  const V3<S> w = col(oM.R, ty - JT_RX);
  return {cross(oM.p, w), w};
}
template<class S> __device__ __forceinline__ S f6(const Frc<S>& f, int k) { return comp(f, k); }
// column 3 + ax of a packed symmetric 6x6 (compile-time indices, so Ia stays in registers)
template<class S> __device__ __forceinline__ Frc<S> symColAng(const Sym6<S>& I, int ax) {
  return ax == 0 ? symCol(I, 3) : (ax == 1 ? symCol(I, 4) : symCol(I, 5));
}

__device__ __forceinline__ bool gNz(const GProg& g, int r, int c) {
  return (g.nzMask[c * g.nzW + (r >> 5)] >> (r & 31)) & 1u;
}

// Persistent tile loop: workgroup x handles tiles x, x + grid, ...; a tile is kTile
// configurations.  Lanes past the batch end evaluate the last configuration and store
// nothing.
#define RBD_TILE_LOOP(B)                                                          \
  const int64_t ntiles_ = ((B) + kTile - 1) / kTile;                              \
  for (int64_t tile_ = blockIdx.x; tile_ < ntiles_; tile_ += gridDim.x)

// ---------------------------------------------------------------------------------------
// RNEA (rnea.hxx:17-161, local frames): tau = M(q) a + C(q, v) v + g(q) + armature * a,
// minus the external forces' generalized forces when gfext.p is set (rnea.hpp:60-79:
// 6 rows per joint 1.., [linear; angular] in the joint's frame)
// ---------------------------------------------------------------------------------------
template<class S> __device__ __forceinline__ Frc<S> gFext(const CView<S>& gf, int i, int64_t b) {
  const int r = (i - 1) * 6;
  return {{gin(gf, r, b), gin(gf, r + 1, b), gin(gf, r + 2, b)}, {gin(gf, r + 3, b), gin(gf, r + 4, b), gin(gf, r + 5, b)}};
}
template<class S, bool kLds>
__global__ __launch_bounds__(64) void g_rnea(GProg g, const S* __restrict__ mc, int64_t B, CView<S> gq, CView<S> gv,
                                             CView<S> ga, CView<S> gfext, View<S> gtau, S* __restrict__ ws) {
  const int L = laneId();
  const Lane<S> m = laneMem<S, kLds>(ws, (int64_t)g.nlev * kGRneaLvl * kTile);
  const Mot<S> a0 = mcGravityNeg(mc, g.nj * MC_REC + g.nv);
  const S* __restrict__ arm = mc + g.nj * MC_REC;
  RBD_TILE_LOOP(B) {
    const int64_t b0 = tile_ * kTile;
    const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
    const bool live = L < nb;
    const int64_t b = b0 + (live ? L : nb - 1);
    for (int e = 0; e < g.nev; ++e) {
      const int ev = g.ev[e];
      const int i = ev >> 1;
      const int* __restrict__ jr = g.jr + i * GJ_REC;
      const int p = jr[GJ_PARENT], ty = jr[GJ_TYPE], iq = jr[GJ_IQ], iv = jr[GJ_IV];
      const int s0 = jr[GJ_LEVEL] * kGRneaLvl, sp = s0 - kGRneaLvl;
      if ((ev & 1) == EV_PRE) {
        const Xform<S> X = gJointX(mc, i, ty, gq, iq, b);
        const Mot<S> vJ = gJointMot(ty, gv, iv, b);
        Mot<S> vi = vJ, ai = gJointMot(ty, ga, iv, b);
        if (p > 0) {
          vi = vi + actInv(X, m.template get<Mot<S>>(sp));
          ai = ai + actInv(X, m.template get<Mot<S>>(sp + 6));
        } else {
          ai = ai + actInv(X, a0);
        }
        ai = ai + crossMM(vi, vJ);
        const Body<S> Y = mcBody(mc, i);
        m.put(s0, vi);
        m.put(s0 + 6, ai);
        Frc<S> fi = inertiaMul(Y, ai) + crossMF(vi, inertiaMul(Y, vi));
        if (gfext.p) fi = fi - gFext(gfext, i, b);
        m.put(s0 + 12, fi);
      } else {
        const Frc<S> f = m.template get<Frc<S>>(s0 + 12);
        if (live) {
          if (ty == JT_FF) {
#pragma unroll
            for (int k = 0; k < 6; ++k) gat(gtau, iv + k, b) = f6(f, k) + arm[iv + k] * gin(ga, iv + k, b);
          } else {
            gat(gtau, iv, b) = f6(f, 3 + ty - JT_RX) + arm[iv] * gin(ga, iv, b);
          }
        }
        if (p > 0) m.add(sp + 12, act(gJointX(mc, i, ty, gq, iq, b), f));
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// ABA (aba.hxx:79-293 computed in local frames, bias re-associated as in k_dyn.hpp):
//   passes 1 + 2 over the events; pass 3 (accelerations) over the PRE events.  External
//   forces (aba.hpp:63-85) enter the bias force, p_i = v_i x* I_i v_i - fext_i.
// ---------------------------------------------------------------------------------------
template<class S, bool kLds>
__global__ __launch_bounds__(64) void g_aba(GProg g, const S* __restrict__ mc, int64_t B, CView<S> gq, CView<S> gv,
                                            CView<S> gt, CView<S> gfext, View<S> gddq, S* __restrict__ ws) {
  const int L = laneId();
  const int jbase = g.nlev * kGAbaLvl;
  const Lane<S> m = laneMem<S, kLds>(ws, (int64_t)(jbase + g.nj * kGAbaJnt) * kTile);
  const Mot<S> a0 = mcGravityNeg(mc, g.nj * MC_REC + g.nv);
  const S* __restrict__ arm = mc + g.nj * MC_REC;
  RBD_TILE_LOOP(B) {
    const int64_t b0 = tile_ * kTile;
    const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
    const bool live = L < nb;
    const int64_t b = b0 + (live ? L : nb - 1);
    for (int e = 0; e < g.nev; ++e) {
      const int ev = g.ev[e];
      const int i = ev >> 1;
      const int* __restrict__ jr = g.jr + i * GJ_REC;
      const int p = jr[GJ_PARENT], ty = jr[GJ_TYPE], iq = jr[GJ_IQ], iv = jr[GJ_IV];
      const int s0 = jr[GJ_LEVEL] * kGAbaLvl, sp = s0 - kGAbaLvl;
      const int sj = jbase + i * kGAbaJnt;
      if ((ev & 1) == EV_PRE) {
        const Xform<S> X = gJointX(mc, i, ty, gq, iq, b);
        const Mot<S> vJ = gJointMot(ty, gv, iv, b);
        Mot<S> vi = vJ;
        if (p > 0) vi = vi + actInv(X, m.template get<Mot<S>>(sp));
        Mot<S> chat = crossMM(vi, vJ);
        if (p > 0) chat = chat + actInv(X, m.template get<Mot<S>>(sp + 6));
        m.put(s0, vi);
        m.put(s0 + 6, chat);
        m.put(s0 + 12, mcSym6(mc, i));
        Frc<S> pi = crossMF(vi, inertiaMul(mcBody(mc, i), vi));
        if (gfext.p) pi = pi - gFext(gfext, i, b);
        m.put(s0 + 33, pi);
      } else {
        Sym6<S> Ia = m.template get<Sym6<S>>(s0 + 12);
        const Frc<S> pA = m.template get<Frc<S>>(s0 + 33);
        if (ty == JT_FF) {
          // free-flyer root: D = Ia + armature (6x6, FP64 LLT), qdd = D^-1 (tau - pA - Ia a')
          const Xform<S> X = gJointX(mc, i, ty, gq, iq, b);
          const Mot<S> ap = actInv(X, a0);
          const Frc<S> Iap = symMul6(Ia, ap);
          double A[21], x[6];
#pragma unroll
          for (int r = 0; r < 6; ++r) {
#pragma unroll
            for (int c = 0; c <= r; ++c) A[lowIdx(r, c)] = double(sget(Ia, r, c));
            A[lowIdx(r, r)] += double(arm[iv + r]);
            x[r] = double(gin(gt, iv + r, b)) - double(f6(pA, r)) - double(f6(Iap, r));
          }
          spdSolve6(A, x);
          if (live) {
#pragma unroll
            for (int r = 0; r < 6; ++r) gat(gddq, iv + r, b) = S(x[r]);
          }
          m.put(sj, Mot<S>{{ap.l.x + S(x[0]), ap.l.y + S(x[1]), ap.l.z + S(x[2])},
                           {ap.a.x + S(x[3]), ap.a.y + S(x[4]), ap.a.z + S(x[5])}});
        } else {
          const int ax = ty - JT_RX;
          const Frc<S> U = symColAng(Ia, ax);
          const S Dinv = S(1) / (f6(U, 3 + ax) + arm[iv]);
          const S u = gin(gt, iv, b) - f6(pA, 3 + ax);
          const Frc<S> dd = scale(Dinv, U);
          m.put(sj, dd);
          m.st(sj + 6, Dinv * u - dot(m.template get<Mot<S>>(s0 + 6), dd));
          if (p > 0) {
            symRank1Sub(Ia, U, Dinv);
            const Mot<S> ci = crossMM(m.template get<Mot<S>>(s0), axisMot(ax, gin(gv, iv, b)));
            const Frc<S> pa = pA + symMul6(Ia, ci) + scale(Dinv * u, U);
            const Xform<S> X = gJointX(mc, i, ty, gq, iq, b);
            m.add(sp + 12, transportUp(X, Ia));
            m.add(sp + 33, act(X, pa));
          }
        }
      }
    }
    // pass 3: accelerations in DFS pre-order (ahat kept per level in the v slots)
    for (int e = 0; e < g.nev; ++e) {
      const int ev = g.ev[e];
      if ((ev & 1) != EV_PRE) continue;
      const int i = ev >> 1;
      const int* __restrict__ jr = g.jr + i * GJ_REC;
      const int p = jr[GJ_PARENT], ty = jr[GJ_TYPE], iq = jr[GJ_IQ], iv = jr[GJ_IV];
      const int s0 = jr[GJ_LEVEL] * kGAbaLvl, sp = s0 - kGAbaLvl;
      const int sj = jbase + i * kGAbaJnt;
      if (ty == JT_FF) {
        m.put(s0, m.template get<Mot<S>>(sj));
      } else {
        const Xform<S> X = gJointX(mc, i, ty, gq, iq, b);
        const Mot<S> t = actInv(X, p > 0 ? m.template get<Mot<S>>(sp) : a0);
        const S qdd = m.ld(sj + 6) - dot(t, m.template get<Frc<S>>(sj));
        if (live) gat(gddq, iv, b) = qdd;
        m.put(s0, t + axisMot(ty - JT_RX, qdd));
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// CRBA (crba.hxx:20-99, 498-548; world frame): M(q), armature on the diagonal; upperOnly
// writes the reference's row blocks only (plus full own blocks, as the reference).
// ---------------------------------------------------------------------------------------
template<class S> __device__ __forceinline__ void gStMat(View<S> g, int64_t b, int nv, int r, int c, S x) {
  stOut(&g.p[(int64_t)(r + c * nv) * g.rs + b * g.bs], x);
}

template<class S, bool kLds>
__global__ __launch_bounds__(64) void g_crba(GProg g, const S* __restrict__ mc, int64_t B, CView<S> gq, View<S> gM,
                                             int upperOnly, S* __restrict__ ws) {
  const int L = laneId();
  const int dbase = g.nlev * kGCrbaLvl;
  const Lane<S> m = laneMem<S, kLds>(ws, (int64_t)(dbase + g.nvlev * kGCrbaDof) * kTile);
  const S* __restrict__ arm = mc + g.nj * MC_REC;
  const int nv = g.nv;
  RBD_TILE_LOOP(B) {
    const int64_t b0 = tile_ * kTile;
    const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
    const bool live = L < nb;
    const int64_t b = b0 + (live ? L : nb - 1);
    for (int e = 0; e < g.nev; ++e) {
      const int ev = g.ev[e];
      const int i = ev >> 1;
      const int* __restrict__ jr = g.jr + i * GJ_REC;
      const int p = jr[GJ_PARENT], ty = jr[GJ_TYPE], iv = jr[GJ_IV], nvi = jr[GJ_NV];
      const int s0 = jr[GJ_LEVEL] * kGCrbaLvl, sp = s0 - kGCrbaLvl;
      const int sd = dbase + jr[GJ_DOFLEVEL] * kGCrbaDof;
      if ((ev & 1) == EV_PRE) {
        Xform<S> X = gJointX(mc, i, ty, gq, jr[GJ_IQ], b);
        if (ty == JT_FF) X.p = {S(0), S(0), S(0)};  // frame origin at the floating base
        const Xform<S> oM = p > 0 ? compose(m.template get<Xform<S>>(sp), X) : X;
        m.put(s0, oM);
        m.put(s0 + 12, worldInertia(mc, i, oM));
        for (int k = 0; k < nvi; ++k) m.put(sd + k * kGCrbaDof, gWorldAxis(ty, oM, k));
      } else {
        const CInertia<S> Y = m.template get<CInertia<S>>(s0 + 12);
        const int cs = g.colPtr[iv], na = jr[GJ_DOFLEVEL];
        for (int k = 0; k < nvi; ++k) {
          const int c = iv + k;
          const Frc<S> F = cmul(Y, m.template get<Mot<S>>(sd + k * kGCrbaDof));
          for (int k2 = 0; k2 <= k; ++k2) {
            const int r = iv + k2;
            S x = dot(m.template get<Mot<S>>(sd + k2 * kGCrbaDof), F);
            if (r == c) x += arm[c];
            if (live) {
              gStMat(gM, b, nv, r, c, x);
              if (r != c) gStMat(gM, b, nv, c, r, x);
            }
          }
          for (int s = 0; s < na; ++s) {
            const int r = g.patRow[cs + s];
            const S x = dot(m.template get<Mot<S>>(dbase + s * kGCrbaDof), F);
            if (live) {
              gStMat(gM, b, nv, r, c, x);
              if (!upperOnly) gStMat(gM, b, nv, c, r, x);
            }
          }
        }
        if (p > 0) m.add(sp + 12, Y);
      }
    }
    if (live) {
      for (int c = 0; c < nv; ++c)
        for (int r = 0; r < nv; ++r)
          if (!gNz(g, r, c) && (!upperOnly || r <= c)) gStMat(gM, b, nv, r, c, S(0));
    }
  }
}

// ---------------------------------------------------------------------------------------
// Tree factorisation M = U D U^T (cholesky::decompose, cholesky.hxx:17-62), right-looking
// over the packed pattern, in place in the output: U holds the pattern entries of M and
// D its diagonal, then column j = nv-1 .. 0 is eliminated:
//   D[r2] -= A[r2][j]^2 / D[j],  A[r1][r2] -= A[r1][j] A[r2][j] / D[j]   (r1 < r2 ancestors of j)
//   U[r][j] = A[r][j] / D[j].
// The ancestors of an ancestor r2 of j are the ancestors of j above it, so with both
// columns listed root first, A[r1][r2] for the s1-th ancestor r1 of j sits at
// colPtr[r2] + s1.  One thread per configuration; SIN is M's type, SC the factor's.
// ---------------------------------------------------------------------------------------
template<class SIN, class SC>
__global__ __launch_bounds__(64) void g_chol(GProg g, int64_t B, CView<SIN> gM, View<SC> gU, View<SC> gD) {
  const int64_t b = (int64_t)blockIdx.x * kTile + laneId();
  if (b >= B) return;
  const int nv = g.nv;
  SC* __restrict__ U = gU.p + b * gU.bs;
  SC* __restrict__ D = gD.p + b * gD.bs;
  const int64_t us = gU.rs, ds = gD.rs;
  for (int c = 0; c < nv; ++c) {
    D[c * ds] = SC(gin(gM, c + c * nv, b));
    for (int k = g.colPtr[c]; k < g.colPtr[c + 1]; ++k) U[k * us] = SC(gin(gM, g.patRow[k] + c * nv, b));
  }
  for (int j = nv - 1; j >= 0; --j) {
    const int cj = g.colPtr[j], na = g.colPtr[j + 1] - cj;
    const SC Dinv = SC(1) / D[j * ds];
    for (int s2 = 0; s2 < na; ++s2) {
      const int r2 = g.patRow[cj + s2];
      const SC w2 = U[(cj + s2) * us];
      const SC t = Dinv * w2;
      D[r2 * ds] -= t * w2;
      const int c2 = g.colPtr[r2];
      for (int s1 = 0; s1 < s2; ++s1) U[(c2 + s1) * us] -= U[(cj + s1) * us] * t;
    }
    for (int s = 0; s < na; ++s) U[(cj + s) * us] *= Dinv;
  }
}

// x <- M^-1 x with a packed FP64 factor: U^-1 by columns, D^-1, U^-T by rows
// (cholesky.hxx:193-317, 466-520); FP64 arithmetic in lane memory, x in S.
template<class S, bool kLds>
__global__ __launch_bounds__(64) void g_chol_solve(GProg g, int64_t B, CView<double> gU, CView<double> gD, View<S> gx,
                                                   double* __restrict__ ws) {
  const int L = laneId();
  const int nv = g.nv;
  const Lane<double> w = laneMem<double, kLds>(ws, (int64_t)nv * kTile);
  RBD_TILE_LOOP(B) {
    const int64_t b = tile_ * kTile + L;
    if (b >= B) continue;
    for (int k = 0; k < nv; ++k) w.st(k, double(gat(gx, k, b)));
    for (int j = nv - 1; j >= 0; --j) {
      const double xj = w.ld(j);
      for (int k = g.colPtr[j]; k < g.colPtr[j + 1]; ++k) {
        const int r = g.patRow[k];
        w.st(r, w.ld(r) - gin(gU, k, b) * xj);
      }
    }
    for (int k = 0; k < nv; ++k) w.st(k, w.ld(k) / gin(gD, k, b));
    for (int j = 0; j < nv; ++j) {
      double acc = w.ld(j);
      for (int k = g.colPtr[j]; k < g.colPtr[j + 1]; ++k) acc -= gin(gU, k, b) * w.ld(g.patRow[k]);
      w.st(j, acc);
    }
    for (int k = 0; k < nv; ++k) gat(gx, k, b) = S(w.ld(k));
  }
}

// ---------------------------------------------------------------------------------------
// computeRNEADerivatives (rnea-derivatives.hxx:242-541, world frame): dtau_dq, dtau_dv, M
// (mask bits 1, 2, 4), each (ancestor r, dof c) pair produced at c's CLOSE as in k_deriv.hpp.
// ---------------------------------------------------------------------------------------
template<class S, bool kLds>
__global__ __launch_bounds__(64) void g_rnead(GProg g, const S* __restrict__ mc, int64_t B, CView<S> gq, CView<S> gv,
                                              CView<S> ga, View<S> gdq, View<S> gdv, View<S> gM, int writeMask,
                                              S* __restrict__ ws) {
  const int L = laneId();
  const int dbase = g.nlev * kGRneadLvl;
  const Lane<S> m = laneMem<S, kLds>(ws, (int64_t)(dbase + g.nvlev * kGRneadDof) * kTile);
  const S* __restrict__ arm = mc + g.nj * MC_REC;
  const S* __restrict__ gr = mc + g.nj * MC_REC + g.nv;
  const Mot<S> a0 = {{-gr[0], -gr[1], -gr[2]}, {-gr[3], -gr[4], -gr[5]}};
  const int nv = g.nv;
  RBD_TILE_LOOP(B) {
    const int64_t b0 = tile_ * kTile;
    const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
    const bool live = L < nb;
    const int64_t b = b0 + (live ? L : nb - 1);
    const bool wq = live && (writeMask & 1), wv = live && (writeMask & 2), wM = live && (writeMask & 4);
    for (int e = 0; e < g.nev; ++e) {
      const int ev = g.ev[e];
      const int i = ev >> 1;
      const int* __restrict__ jr = g.jr + i * GJ_REC;
      const int p = jr[GJ_PARENT], ty = jr[GJ_TYPE], iv = jr[GJ_IV], nvi = jr[GJ_NV];
      const int s0 = jr[GJ_LEVEL] * kGRneadLvl, sp = s0 - kGRneadLvl;
      const int sd = dbase + jr[GJ_DOFLEVEL] * kGRneadDof;
      if ((ev & 1) == EV_PRE) {
        const Xform<S> X = gJointX(mc, i, ty, gq, jr[GJ_IQ], b);
        Xform<S> oM;
        Mot<S> ov, oa;
        if (ty == JT_FF) {
          oM = X;
          oM.p = {S(0), S(0), S(0)};  // frame origin at the floating base
          Mot<S> vJ = zeroMot<S>(), aJ = zeroMot<S>();
          for (int k = 0; k < 6; ++k) {
            const Mot<S> J = gWorldAxis(JT_FF, oM, k);
            vJ = vJ + scale(gin(gv, iv + k, b), J);
            aJ = aJ + scale(gin(ga, iv + k, b), J);
          }
          ov = vJ;
          oa = a0 + aJ;
          for (int k = 0; k < 6; ++k) {
            const Mot<S> J = gWorldAxis(JT_FF, oM, k);
            m.put(sd + k * kGRneadDof, ColData<S>{J, zeroMot<S>(), crossMM(a0, J), crossMM(ov, J)});
          }
        } else {
          oM = p > 0 ? compose(m.template get<Xform<S>>(sp), X) : X;
          const Mot<S> J = gWorldAxis(ty, oM, 0);
          const S qd = gin(gv, iv, b), qdd = gin(ga, iv, b);
          const Mot<S> vJ = scale(qd, J);
          if (p > 0) {
            const Mot<S> ovp = m.template get<Mot<S>>(sp + 12), oap = m.template get<Mot<S>>(sp + 18);
            ov = ovp + vJ;
            oa = oap + crossMM(ovp, vJ) + scale(qdd, J);
            const Mot<S> dVdq = crossMM(ovp, J);
            m.put(sd, ColData<S>{J, dVdq, crossMM(oap, J) + crossMM(ovp, dVdq), crossMM(ov, J) + dVdq});
          } else {
            ov = vJ;
            oa = a0 + scale(qdd, J);
            m.put(sd, ColData<S>{J, zeroMot<S>(), crossMM(a0, J), crossMM(ov, J)});
          }
        }
        m.put(s0, oM);
        m.put(s0 + 12, ov);
        m.put(s0 + 18, oa);
        const CInertia<S> Y = worldInertia(mc, i, oM);
        const Frc<S> oh = cmul(Y, ov);
        m.put(s0 + 24, Y);
        m.put(s0 + 34, inertiaVariation(Y, ov, oh));
        m.put(s0 + 61, cmul(Y, oa) + crossMF(ov, oh));
      } else {
        const CInertia<S> Y = m.template get<CInertia<S>>(s0 + 24);
        const DY<S> DYc = m.template get<DY<S>>(s0 + 34);
        const Frc<S> of = m.template get<Frc<S>>(s0 + 61);
        const int cs = g.colPtr[iv], na = jr[GJ_DOFLEVEL];
        for (int k = 0; k < nvi; ++k) {
          const int c = iv + k;
          const ColData<S> C = m.template get<ColData<S>>(sd + k * kGRneadDof);
          const Frc<S> dFda = cmul(Y, C.J);
          const Frc<S> dFdq = cmul(Y, C.dAdq) + dyMul(DYc, C.dVdq);
          const Frc<S> dFdv = dyMul(DYc, C.J) + cmul(Y, C.dAdv);
          const Frc<S> dYtJ = dyMulT(DYc, C.J);
          for (int k2 = 0; k2 < nvi; ++k2) {
            const int r = iv + k2;
            const Mot<S> Jr = m.template get<Mot<S>>(sd + k2 * kGRneadDof);
            if (wq) gStMat(gdq, b, nv, r, c, dot(Jr, dFdq));
            if (wv) gStMat(gdv, b, nv, r, c, dot(Jr, dFdv));
            if (wM) gStMat(gM, b, nv, r, c, dot(Jr, dFda) + (r == c ? arm[c] : S(0)));
          }
          const Frc<S> dFdqA = dFdq + crossMF(C.J, of);
          for (int s = 0; s < na; ++s) {
            const int r = g.patRow[cs + s];
            const ColData<S> R = m.template get<ColData<S>>(dbase + s * kGRneadDof);
            if (wq) {
              gStMat(gdq, b, nv, r, c, dot(R.J, dFdqA));
              gStMat(gdq, b, nv, c, r, dot(R.dAdq, dFda) + dot(R.dVdq, dYtJ));
            }
            if (wv) {
              gStMat(gdv, b, nv, r, c, dot(R.J, dFdv));
              gStMat(gdv, b, nv, c, r, dot(R.dAdv, dFda) + dot(R.J, dYtJ));
            }
            if (wM) {
              const S x = dot(R.J, dFda);
              gStMat(gM, b, nv, r, c, x);
              gStMat(gM, b, nv, c, r, x);
            }
          }
        }
        if (p > 0) {
          m.add(sp + 24, Y);
          m.add(sp + 34, DYc);
          m.add(sp + 61, of);
        }
      }
    }
    if (wq || wv || wM) {
      for (int c = 0; c < nv; ++c)
        for (int r = 0; r < nv; ++r)
          if (!gNz(g, r, c)) {
            if (wq) gStMat(gdq, b, nv, r, c, S(0));
            if (wv) gStMat(gdv, b, nv, r, c, S(0));
            if (wM) gStMat(gM, b, nv, r, c, S(0));
          }
    }
  }
}

// ---------------------------------------------------------------------------------------
// ABA-derivative solve (aba-derivatives.hxx:448-452 as sparse solves): in place
// Xq <- -M^-1 Xq, Xv <- -M^-1 Xv, Minv <- M^-1 from the packed factor (batch-fastest
// workspace, any S).  A workgroup of kGSolveWaves waves owns a tile; its waves share out
// the 3 nv right-hand sides (jobs before jobBegin are skipped) and each keeps its
// right-hand side in LDS as [dof][lane].  kLdsF: the tile's factor (U, then D^-1) is
// staged in LDS too; otherwise it is read from the workspace.
// ---------------------------------------------------------------------------------------
constexpr int kGSolveWaves = 4;

template<class S, bool kLdsF>
__global__ __launch_bounds__(64 * kGSolveWaves) void g_abad_solve(GProg g, int64_t B, CView<S> gU, CView<S> gD,
                                                                  View<S> gXq, View<S> gXv, View<S> gMinv,
                                                                  int jobBegin) {
  extern __shared__ __align__(16) unsigned char g_lds[];
  const int nv = g.nv, nnz = g.nnz;
  S* sx = reinterpret_cast<S*>(g_lds);                  // kGSolveWaves x nv x kTile
  S* sF = sx + (int64_t)kGSolveWaves * nv * kTile;      // (nnz + nv) x kTile when kLdsF
  const int L = threadIdx.x & 63, w = threadIdx.x >> 6;
  const Lane<S> x{sx + (int64_t)w * nv * kTile + L};
  RBD_TILE_LOOP(B) {
    const int64_t b0 = tile_ * kTile;
    const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
    const bool live = L < nb;
    const int64_t b = b0 + (live ? L : nb - 1);
    if constexpr (kLdsF) {
      __syncthreads();  // the previous tile's solves are done with sF
      for (int e = threadIdx.x; e < (nnz + nv) * kTile; e += 64 * kGSolveWaves) {
        const int k = e >> 6, lb = e & 63;
        const int64_t bb = b0 + (lb < nb ? lb : nb - 1);
        sF[e] = k < nnz ? gin(gU, k, bb) : S(1) / gin(gD, k - nnz, bb);
      }
      __syncthreads();
    }
    auto U = [&](int k) -> S {
      if constexpr (kLdsF) return sF[k * kTile + L];
      else return gin(gU, k, b);
    };
    auto Dinv = [&](int k) -> S {
      if constexpr (kLdsF) return sF[(nnz + k) * kTile + L];
      else return S(1) / gin(gD, k, b);
    };
    for (int j = jobBegin + w; j < 3 * nv; j += kGSolveWaves) {
      View<S> out = j < nv ? gXq : (j < 2 * nv ? gXv : gMinv);
      const int c = j < nv ? j : (j < 2 * nv ? j - nv : j - 2 * nv);
      if (j < 2 * nv) {
        for (int r = 0; r < nv; ++r) x.st(r, gat(out, r + c * nv, b));
      } else {
        for (int r = 0; r < nv; ++r) x.st(r, r == c ? S(1) : S(0));
      }
      for (int jc = nv - 1; jc >= 0; --jc) {
        const S xj = x.ld(jc);
        for (int k = g.colPtr[jc]; k < g.colPtr[jc + 1]; ++k) {
          const int r = g.patRow[k];
          x.st(r, x.ld(r) - U(k) * xj);
        }
      }
      for (int k = 0; k < nv; ++k) x.st(k, x.ld(k) * Dinv(k));
      for (int jc = 0; jc < nv; ++jc) {
        S acc = x.ld(jc);
        for (int k = g.colPtr[jc]; k < g.colPtr[jc + 1]; ++k) acc -= U(k) * x.ld(g.patRow[k]);
        x.st(jc, acc);
      }
      if (live) {
        const S sgn = j < 2 * nv ? S(-1) : S(1);
        for (int r = 0; r < nv; ++r) stOut(&gat(out, r + c * nv, b), sgn * x.ld(r));
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// Forward kinematics and WORLD joint Jacobians (kinematics.hxx:121-145, 209-238;
// jacobian.hxx:23-62); outputs as k_kin.hpp, any may be absent.
// ---------------------------------------------------------------------------------------
template<class S, bool kLds>
__global__ __launch_bounds__(64) void g_kin(GProg g, const S* __restrict__ mc, int64_t B, CView<S> gq, CView<S> gv,
                                            CView<S> ga, View<S> gM, View<S> gV, View<S> gA, View<S> gJ,
                                            S* __restrict__ ws) {
  const int L = laneId();
  const Lane<S> m = laneMem<S, kLds>(ws, (int64_t)g.nlev * kGKinLvl * kTile);
  const bool wantV = gV.p != nullptr, wantA = gA.p != nullptr;
  RBD_TILE_LOOP(B) {
    const int64_t b0 = tile_ * kTile;
    const int nb = (int)(B - b0 < kTile ? B - b0 : kTile);
    const bool live = L < nb;
    const int64_t b = b0 + (live ? L : nb - 1);
    auto put = [&](View<S> o, int row, S x) {
      if (live) stOut(&o.p[(int64_t)row * o.rs + b * o.bs], x);
    };
    for (int e = 0; e < g.nev; ++e) {
      const int ev = g.ev[e];
      if ((ev & 1) != EV_PRE) continue;
      const int i = ev >> 1;
      const int* __restrict__ jr = g.jr + i * GJ_REC;
      const int p = jr[GJ_PARENT], ty = jr[GJ_TYPE], iv = jr[GJ_IV], nvi = jr[GJ_NV];
      const int s0 = jr[GJ_LEVEL] * kGKinLvl, sp = s0 - kGKinLvl;
      const Xform<S> X = gJointX(mc, i, ty, gq, jr[GJ_IQ], b);
      const Xform<S> oM = p > 0 ? compose(m.template get<Xform<S>>(sp), X) : X;
      m.put(s0, oM);
      if (gM.p) {
        const R3<S>& R = oM.R;
        const S r12[12] = {R.m00, R.m01, R.m02, R.m10, R.m11, R.m12, R.m20, R.m21, R.m22, oM.p.x, oM.p.y, oM.p.z};
#pragma unroll
        for (int k = 0; k < 12; ++k) put(gM, (i - 1) * 12 + k, r12[k]);
      }
      if (wantV) {
        const Mot<S> vJ = gJointMot(ty, gv, iv, b);
        Mot<S> vi = vJ;
        if (p > 0) vi = vi + actInv(X, m.template get<Mot<S>>(sp + 12));
        m.put(s0 + 12, vi);
        const S o[6] = {vi.l.x, vi.l.y, vi.l.z, vi.a.x, vi.a.y, vi.a.z};
#pragma unroll
        for (int k = 0; k < 6; ++k) put(gV, (i - 1) * 6 + k, o[k]);
        if (wantA) {
          Mot<S> ai = crossMM(vi, vJ) + gJointMot(ty, ga, iv, b);
          if (p > 0) ai = ai + actInv(X, m.template get<Mot<S>>(sp + 18));
          m.put(s0 + 18, ai);
          const S oa[6] = {ai.l.x, ai.l.y, ai.l.z, ai.a.x, ai.a.y, ai.a.z};
#pragma unroll
          for (int k = 0; k < 6; ++k) put(gA, (i - 1) * 6 + k, oa[k]);
        }
      }
      if (gJ.p) {
        for (int k = 0; k < nvi; ++k) {
          const Mot<S> c = gWorldAxis(ty, oM, k);
          const S o[6] = {c.l.x, c.l.y, c.l.z, c.a.x, c.a.y, c.a.z};
#pragma unroll
          for (int r = 0; r < 6; ++r) put(gJ, (iv + k) * 6 + r, o[r]);
        }
      }
    }
  }
}

}  // namespace rbd
