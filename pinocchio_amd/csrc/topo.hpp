// topo.hpp -- compile-time kinematic-tree topology and sweep schedule.
//
// A model's topology (parents, joint types) is baked into each kernel instantiation so
// that every per-joint quantity lives in registers with a compile-time index.  From the
// parents array this header derives, at compile time, what the reference's Model/Data
// hold at run time (model.hxx:61-170 idx_q/idx_v, data.hxx:196-315 nvSubtree,
// parents_fromRow) plus the depth-first sweep schedule the kernels follow:
//
//   PRE(i)   forward-pass work of joint i (parents first, DFS pre-order);
//   CLOSE(i) backward-pass work of joint i, issued as soon as the last joint of its
//            subtree has been visited (children first).
//
// Interleaving the two passes this way keeps only the open path of the tree live, which
// is what lets a whole configuration's sweep stay in one lane's registers.
#pragma once

namespace rbd {

enum : int { JT_U = 0, JT_FF = 1, JT_RX = 2, JT_RY = 3, JT_RZ = 4 };
enum : int { EV_PRE = 0, EV_CLOSE = 1 };

constexpr int kMaxJ = 64;
constexpr int kMaxV = 72;

template<int NJ> struct TopoData {
  int parent[NJ] = {};
  int type[NJ] = {};
  int nqj[NJ] = {}, nvj[NJ] = {}, iq[NJ] = {}, iv[NJ] = {};
  int nq = 0, nv = 0;
  int order[NJ] = {};    // DFS pre-order of joints 1..NJ-1 (entries 0..NJ-2)
  int pos[NJ] = {};      // position of joint j in `order`
  int subsize[NJ] = {};  // joints in the subtree of j, j included
  int nvSub[NJ] = {};    // dofs in the subtree of j (nvSubtree when compact)
  int nchild[NJ] = {};
  int depth[NJ] = {};    // root joint has depth 1
  int childIdx[NJ] = {}; // rank of j among its parent's children (in DFS order)
  int nev = 0;
  int evKind[2 * NJ] = {};
  int evJoint[2 * NJ] = {};
  bool compact = true;   // DFS pre-order is the index order (crba.hxx:573-595)
  bool hasFF = false;
  // dof-level tree: parent dof of each dof (parents_fromRow, data.hxx:244-315),
  // joint owning each dof
  int dofParent[kMaxV] = {};
  int dofJoint[kMaxV] = {};
  int maxDepth = 0;
};

template<int NJ>
constexpr TopoData<NJ> makeTopo(const int (&parent)[NJ], const int (&type)[NJ]) {
  TopoData<NJ> t{};
  for (int j = 0; j < NJ; ++j) {
    t.parent[j] = parent[j];
    t.type[j] = type[j];
  }
  for (int j = 1; j < NJ; ++j) {
    t.nqj[j] = type[j] == JT_FF ? 7 : 1;
    t.nvj[j] = type[j] == JT_FF ? 6 : 1;
    t.iq[j] = t.nq;
    t.iv[j] = t.nv;
    t.nq += t.nqj[j];
    t.nv += t.nvj[j];
    if (type[j] == JT_FF) t.hasFF = true;
  }
  for (int j = 1; j < NJ; ++j) {
    t.depth[j] = parent[j] == 0 ? 1 : t.depth[parent[j]] + 1;
    if (t.depth[j] > t.maxDepth) t.maxDepth = t.depth[j];
    if (parent[j] > 0) t.nchild[parent[j]] += 1;
  }
  // DFS pre-order with children visited in increasing index order
  int stack[NJ + 1] = {};
  int sp = 0, n = 0;
  for (int r = NJ - 1; r >= 1; --r)
    if (parent[r] == 0) stack[sp++] = r;
  while (sp > 0) {
    const int j = stack[--sp];
    t.order[n] = j;
    t.pos[j] = n;
    ++n;
    for (int c = NJ - 1; c > j; --c)
      if (parent[c] == j) stack[sp++] = c;
  }
  for (int k = 0; k < NJ - 1; ++k)
    if (t.order[k] != k + 1) t.compact = false;
  for (int j = NJ - 1; j >= 1; --j) {
    t.subsize[j] += 1;
    t.nvSub[j] += t.nvj[j];
    if (parent[j] > 0) {
      t.subsize[parent[j]] += t.subsize[j];
      t.nvSub[parent[j]] += t.nvSub[j];
    }
  }
  {
    int seen[NJ] = {};
    for (int k = 0; k < NJ - 1; ++k) {
      const int j = t.order[k];
      const int p = parent[j];
      t.childIdx[j] = seen[p];
      seen[p] += 1;
    }
  }
  // event schedule: PRE(order[k]) then CLOSE of every joint whose subtree ends at k,
  // deepest first.
  for (int k = 0; k < NJ - 1; ++k) {
    const int j = t.order[k];
    t.evKind[t.nev] = EV_PRE;
    t.evJoint[t.nev] = j;
    ++t.nev;
    for (int kk = k; kk >= 0; --kk) {
      const int a = t.order[kk];
      if (t.pos[a] + t.subsize[a] - 1 == k) {
        t.evKind[t.nev] = EV_CLOSE;
        t.evJoint[t.nev] = a;
        ++t.nev;
      }
    }
  }
  for (int j = 1; j < NJ; ++j) {
    const int p = parent[j];
    t.dofParent[t.iv[j]] = p > 0 ? t.iv[p] + t.nvj[p] - 1 : -1;
    t.dofJoint[t.iv[j]] = j;
    for (int r = 1; r < t.nvj[j]; ++r) {
      t.dofParent[t.iv[j] + r] = t.iv[j] + r - 1;
      t.dofJoint[t.iv[j] + r] = j;
    }
  }
  return t;
}

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template<int I> struct IC {
  static constexpr int value = I;
  constexpr operator int() const { return I; }
};
template<int B, int E, class F> __host__ __device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(IC<B>{});
    static_for<B + 1, E>(f);
  }
}

// Subtree split of one sweep over G waves that share a tile of configurations.
// Subtrees larger than total / G are opened (their root joins the "trunk", which every
// wave computes redundantly) until the remaining subtrees fit; those are dealt to the
// waves largest first.  A wave owns whole subtrees, so the forward and backward passes of
// a subtree never leave its wave; what a subtree hands to its trunk parent goes through
// an LDS slot (one per subtree root below the trunk), read after a workgroup barrier.
//   owner[j]   -1 for trunk joints, else the wave owning j
//   slot[j]    LDS slot of subtree root j whose parent is a trunk joint, else -1
//   ldsIn[j]   trunk joint j receives at least one contribution through a slot
//   nReg[j]    children of j whose contributions stay in registers (same owner)
//   regIdx[j]  rank of j among its parent's register children
template<class TP, int G> struct WaveSplit {
  static constexpr int NJ = TP::NJ;
  int owner[NJ] = {}, slot[NJ] = {}, nReg[NJ] = {}, regIdx[NJ] = {};
  bool ldsIn[NJ] = {};
  int nslots = 0;
  int load[8] = {};
  constexpr WaveSplit() {
    constexpr auto& T = TP::T;
    bool trunk[NJ] = {}, front[NJ] = {};
    for (int j = 1; j < NJ; ++j) front[j] = T.parent[j] == 0;
    if (G > 1) {
      const int target = (NJ - 1 + G - 1) / G;
      for (int it = 0; it < NJ; ++it) {
        int big = -1;
        for (int j = 1; j < NJ; ++j)
          if (front[j] && (big < 0 || T.subsize[j] > T.subsize[big])) big = j;
        if (big < 0 || T.subsize[big] <= target || T.nchild[big] == 0) break;
        front[big] = false;
        trunk[big] = true;
        for (int c = 1; c < NJ; ++c)
          if (T.parent[c] == big) front[c] = true;
      }
    }
    // deal frontier subtrees to waves, largest first
    int rootOwner[NJ] = {};
    bool done[NJ] = {};
    for (int it = 0; it < NJ; ++it) {
      int big = -1;
      for (int j = 1; j < NJ; ++j)
        if (front[j] && !done[j] && (big < 0 || T.subsize[j] > T.subsize[big])) big = j;
      if (big < 0) break;
      done[big] = true;
      int w = 0;
      for (int k = 1; k < G; ++k)
        if (load[k] < load[w]) w = k;
      rootOwner[big] = w;
      load[w] += T.subsize[big];
    }
    for (int j = 1; j < NJ; ++j) {
      if (trunk[j]) {
        owner[j] = -1;
        continue;
      }
      int r = j;
      while (!front[r]) r = T.parent[r];
      owner[j] = rootOwner[r];
    }
    owner[0] = -1;
    for (int j = 1; j < NJ; ++j) {
      const int p = T.parent[j];
      slot[j] = -1;
      if (p > 0 && owner[j] != owner[p]) {
        slot[j] = nslots++;
        ldsIn[p] = true;
      }
    }
    // register children in DFS order (the order their CLOSE events come in)
    for (int k = 0; k < NJ - 1; ++k) {
      const int j = T.order[k];
      const int p = T.parent[j];
      if (p > 0 && slot[j] < 0) regIdx[j] = nReg[p]++;
    }
  }
  constexpr bool runs(int j, int w) const { return owner[j] < 0 || owner[j] == w; }
  // the wave that stores joint j's outputs (wave 0 for trunk joints)
  constexpr int writer(int j) const { return owner[j] < 0 ? 0 : owner[j]; }
};

// Output-emission groups for G waves sharing one tile of configurations: every wave runs
// the forward pass and the subtree accumulations, and joint i's output entries are
// computed and stored by wave grp[i] only.  Joints are dealt to the least-loaded wave in
// decreasing order of their emission cost nv_i (nv_i + ancestor dofs).
template<class TP, int G> struct EmitGroups {
  int grp[TP::NJ];
  constexpr EmitGroups() : grp{} {
    constexpr auto& T = TP::T;
    long load[8] = {};
    bool done[TP::NJ] = {};
    for (int it = 1; it < TP::NJ; ++it) {
      int best = 1;
      long bc = -1;
      for (int j = 1; j < TP::NJ; ++j) {
        if (done[j]) continue;
        int anc = 0;
        for (int a = T.parent[j]; a > 0; a = T.parent[a]) anc += T.nvj[a];
        const long c = (long)T.nvj[j] * (anc + T.nvj[j]);
        if (c > bc) {
          bc = c;
          best = j;
        }
      }
      done[best] = true;
      int g = 0;
      for (int k = 1; k < G; ++k)
        if (load[k] < load[g]) g = k;
      grp[best] = g;
      load[g] += bc;
    }
  }
};

// Per-joint record of model constants (device buffer, read with uniform loads).
// Offsets in scalars; one record per joint, joint 0 included.
enum : int {
  MC_R = 0,      // placement rotation, 9, row major
  MC_P = 9,      // placement translation, 3
  MC_M = 12,     // mass
  MC_C = 13,     // lever, 3
  MC_I = 16,     // rotational inertia about CoM (xx, xy, yy, xz, yz, zz), 6
  MC_Y6 = 22,    // body inertia as packed symmetric 6x6, 21
  MC_CI = 43,    // composite-origin form (m, h(3), Io(6)) = 10
  MC_REC = 53
};
// after NJ records: armature[nv], then gravity[6]
template<int NJ> constexpr int mcArmature() { return NJ * MC_REC; }
template<int NJ, int NV> constexpr int mcGravity() { return NJ * MC_REC + NV; }
template<int NJ, int NV> constexpr int mcSize() { return NJ * MC_REC + NV + 6; }

}  // namespace rbd
