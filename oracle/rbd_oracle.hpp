// rbd_oracle.hpp -- CPU restatement of the reference's batched rigid-body-dynamics path.
//
// TEST INFRASTRUCTURE ONLY.  This header is the oracle that the GPU engine is checked
// against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it (through oracle/_build/liboracle.so).  The product library never links it.
//
// It restates, in dependency-free C++ templated on the scalar type, the algorithms of
// pinocchio 3.7.0 that rneaInParallel / abaInParallel batch on the CPU:
//   spatial algebra          include/pinocchio/spatial/{se3-tpl,motion-dense,force-dense,
//                            inertia,symmetric3,skew,act-on-set}.hpp
//   joints                   multibody/joint/joint-free-flyer.hpp:333-391,
//                            multibody/joint/joint-revolute.hpp:793-836
//   topology helpers         multibody/data.hxx:196-315
//   forward kinematics       algorithm/kinematics.hxx:121-145, 209-238
//   joint Jacobians          algorithm/jacobian.hxx:23-62
//   rnea                     algorithm/rnea.hxx:17-161
//   aba (WORLD, LOCAL)       algorithm/aba.hxx:79-293, 354-540
//   crba (WORLD)             algorithm/crba.hxx:20-99, 498-548
//   cholesky                 algorithm/cholesky.hxx:17-62, 193-317, 466-597
//   rnea derivatives         algorithm/rnea-derivatives.hxx:235-541
//   aba derivatives          algorithm/aba-derivatives.hxx:16-453
//   integrate (Lie group)    multibody/liegroup/special-euclidean.hpp:659-698,
//                            spatial/explog-quaternion.hpp:92-147
//   batch wrappers           algorithm/parallel/rnea.hpp:39-83, algorithm/parallel/aba.hpp:40-84
//   constraint dynamics      algorithm/constrained-dynamics.hxx:71-500, contact-info.hpp:426-700,
//                            spatial/log.hxx:58-230
//
// Only the joint types the engine supports are restated: FreeFlyer and Revolute{X,Y,Z}.
// Conventions follow the reference: 6-vectors are [linear; angular]; Inertia = (m, c, I_c)
// with I_c a Symmetric3 stored (xx, xy, yy, xz, yz, zz) about the centre of mass; the
// free-flyer configuration is [x, y, z, qx, qy, qz, qw]; gravity is a motion whose
// linear part is (0, 0, -9.81).
//
// Parity pinning: every algorithm here is checked in tests/test_oracle.py against closed
// forms, energy identities, brute force and finite differences (see that file's header).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <vector>

namespace oracle {

// ----------------------------------------------------------------------------------------
// Small fixed-size algebra
// ----------------------------------------------------------------------------------------
template<class S> struct V3 {
  S x[3];
  S& operator[](int i) { return x[i]; }
  const S& operator[](int i) const { return x[i]; }
};
template<class S> inline V3<S> v3(S a, S b, S c) { return V3<S>{{a, b, c}}; }
template<class S> inline V3<S> zero3() { return V3<S>{{S(0), S(0), S(0)}}; }
template<class S> inline V3<S> operator+(const V3<S>& a, const V3<S>& b) { return v3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
template<class S> inline V3<S> operator-(const V3<S>& a, const V3<S>& b) { return v3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
template<class S> inline V3<S> operator-(const V3<S>& a) { return v3(-a[0], -a[1], -a[2]); }
template<class S> inline V3<S> operator*(S s, const V3<S>& a) { return v3(s * a[0], s * a[1], s * a[2]); }
template<class S> inline V3<S>& operator+=(V3<S>& a, const V3<S>& b) { a = a + b; return a; }
template<class S> inline V3<S>& operator-=(V3<S>& a, const V3<S>& b) { a = a - b; return a; }
template<class S> inline S dot(const V3<S>& a, const V3<S>& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template<class S> inline V3<S> cross(const V3<S>& a, const V3<S>& b) {
  return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}

// 3x3 matrix, row major.
template<class S> struct M3 {
  S m[9];
  S& operator()(int r, int c) { return m[r * 3 + c]; }
  const S& operator()(int r, int c) const { return m[r * 3 + c]; }
};
template<class S> inline M3<S> eye3() { M3<S> r{}; r(0, 0) = r(1, 1) = r(2, 2) = S(1); return r; }
template<class S> inline M3<S> zeroM3() { M3<S> r{}; for (auto& e : r.m) e = S(0); return r; }
template<class S> inline V3<S> operator*(const M3<S>& A, const V3<S>& x) {
  return v3(A(0, 0) * x[0] + A(0, 1) * x[1] + A(0, 2) * x[2], A(1, 0) * x[0] + A(1, 1) * x[1] + A(1, 2) * x[2],
            A(2, 0) * x[0] + A(2, 1) * x[1] + A(2, 2) * x[2]);
}
template<class S> inline V3<S> mulT(const M3<S>& A, const V3<S>& x) {  // A^T x
  return v3(A(0, 0) * x[0] + A(1, 0) * x[1] + A(2, 0) * x[2], A(0, 1) * x[0] + A(1, 1) * x[1] + A(2, 1) * x[2],
            A(0, 2) * x[0] + A(1, 2) * x[1] + A(2, 2) * x[2]);
}
template<class S> inline M3<S> operator*(const M3<S>& A, const M3<S>& B) {
  M3<S> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r(i, j) = A(i, 0) * B(0, j) + A(i, 1) * B(1, j) + A(i, 2) * B(2, j);
  return r;
}
template<class S> inline M3<S> transpose(const M3<S>& A) {
  M3<S> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r(i, j) = A(j, i);
  return r;
}
template<class S> inline M3<S> operator+(const M3<S>& A, const M3<S>& B) { M3<S> r; for (int k = 0; k < 9; ++k) r.m[k] = A.m[k] + B.m[k]; return r; }
template<class S> inline M3<S> operator-(const M3<S>& A, const M3<S>& B) { M3<S> r; for (int k = 0; k < 9; ++k) r.m[k] = A.m[k] - B.m[k]; return r; }
template<class S> inline M3<S> operator-(const M3<S>& A) { M3<S> r; for (int k = 0; k < 9; ++k) r.m[k] = -A.m[k]; return r; }
template<class S> inline M3<S> skew(const V3<S>& v) {  // spatial/skew.hpp: skew(v) w = v x w
  M3<S> r;
  r(0, 0) = S(0);  r(0, 1) = -v[2]; r(0, 2) = v[1];
  r(1, 0) = v[2];  r(1, 1) = S(0);  r(1, 2) = -v[0];
  r(2, 0) = -v[1]; r(2, 1) = v[0];  r(2, 2) = S(0);
  return r;
}
// spatial/skew.hpp:182-197: C = [u]x [v]x = v u^T - (u.v) I
template<class S> inline M3<S> skewSquare(const V3<S>& u, const V3<S>& v) {
  M3<S> r;
  const S d = dot(u, v);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r(i, j) = v[i] * u[j] - (i == j ? d : S(0));
  return r;
}

// Symmetric 3x3 in the reference's storage order (xx, xy, yy, xz, yz, zz), symmetric3.hpp.
template<class S> struct Sym3 {
  S d[6];
  M3<S> matrix() const {
    M3<S> r;
    r(0, 0) = d[0]; r(0, 1) = d[1]; r(0, 2) = d[3];
    r(1, 0) = d[1]; r(1, 1) = d[2]; r(1, 2) = d[4];
    r(2, 0) = d[3]; r(2, 1) = d[4]; r(2, 2) = d[5];
    return r;
  }
  static Sym3 fromMatrix(const M3<S>& A) {  // symmetric part
    Sym3 s;
    s.d[0] = A(0, 0); s.d[1] = S(0.5) * (A(0, 1) + A(1, 0)); s.d[2] = A(1, 1);
    s.d[3] = S(0.5) * (A(0, 2) + A(2, 0)); s.d[4] = S(0.5) * (A(1, 2) + A(2, 1)); s.d[5] = A(2, 2);
    return s;
  }
  V3<S> operator*(const V3<S>& v) const {
    return v3(d[0] * v[0] + d[1] * v[1] + d[3] * v[2], d[1] * v[0] + d[2] * v[1] + d[4] * v[2],
              d[3] * v[0] + d[4] * v[1] + d[5] * v[2]);
  }
};

// ----------------------------------------------------------------------------------------
// Spatial algebra: Motion, Force, SE3, Inertia, 6x6
// ----------------------------------------------------------------------------------------
template<class S> struct Motion { V3<S> lin, ang; };
template<class S> struct Force { V3<S> lin, ang; };
template<class S> inline Motion<S> zeroMotion() { return Motion<S>{zero3<S>(), zero3<S>()}; }
template<class S> inline Force<S> zeroForce() { return Force<S>{zero3<S>(), zero3<S>()}; }
template<class S> inline Motion<S> operator+(const Motion<S>& a, const Motion<S>& b) { return {a.lin + b.lin, a.ang + b.ang}; }
template<class S> inline Motion<S> operator-(const Motion<S>& a, const Motion<S>& b) { return {a.lin - b.lin, a.ang - b.ang}; }
template<class S> inline Motion<S> operator*(S s, const Motion<S>& a) { return {s * a.lin, s * a.ang}; }
template<class S> inline Force<S> operator+(const Force<S>& a, const Force<S>& b) { return {a.lin + b.lin, a.ang + b.ang}; }
template<class S> inline Force<S> operator*(S s, const Force<S>& a) { return {s * a.lin, s * a.ang}; }
template<class S> inline S dot(const Motion<S>& m, const Force<S>& f) { return dot(m.lin, f.lin) + dot(m.ang, f.ang); }

// motion-dense.hpp:217-245 : v1 ^ v2  (motion cross motion)
template<class S> inline Motion<S> crossMM(const Motion<S>& a, const Motion<S>& b) {
  return {cross(a.ang, b.lin) + cross(a.lin, b.ang), cross(a.ang, b.ang)};
}
// motion-dense.hpp:247-262 : v.cross(f)  (motion cross force, dual action)
template<class S> inline Force<S> crossMF(const Motion<S>& v, const Force<S>& f) {
  return {cross(v.ang, f.lin), cross(v.ang, f.ang) + cross(v.lin, f.lin)};
}

template<class S> struct SE3 {
  M3<S> R;
  V3<S> p;
  static SE3 identity() { return SE3{eye3<S>(), zero3<S>()}; }
};
// se3-tpl.hpp:327 : composition
template<class S> inline SE3<S> operator*(const SE3<S>& a, const SE3<S>& b) { return SE3<S>{a.R * b.R, a.p + a.R * b.p}; }
// motion-dense.hpp:264-288 : se3Action / se3ActionInverse on motions
template<class S> inline Motion<S> act(const SE3<S>& M, const Motion<S>& m) {
  const V3<S> w = M.R * m.ang;
  return {M.R * m.lin + cross(M.p, w), w};
}
template<class S> inline Motion<S> actInv(const SE3<S>& M, const Motion<S>& m) {
  return {mulT(M.R, m.lin - cross(M.p, m.ang)), mulT(M.R, m.ang)};
}
// force-dense.hpp:214-243 : se3Action / se3ActionInverse on forces
template<class S> inline Force<S> act(const SE3<S>& M, const Force<S>& f) {
  const V3<S> l = M.R * f.lin;
  return {l, M.R * f.ang + cross(M.p, l)};
}
template<class S> inline Force<S> actInv(const SE3<S>& M, const Force<S>& f) {
  return {mulT(M.R, f.lin), mulT(M.R, f.ang - cross(M.p, f.lin))};
}

// Row-major 6x6 with [linear; angular] blocks.
template<class S> struct M6 {
  S m[36];
  S& operator()(int r, int c) { return m[r * 6 + c]; }
  const S& operator()(int r, int c) const { return m[r * 6 + c]; }
  static M6 zero() { M6 r; for (auto& e : r.m) e = S(0); return r; }
  M3<S> block(int r0, int c0) const {
    M3<S> b;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) b(i, j) = (*this)(r0 + i, c0 + j);
    return b;
  }
  void setBlock(int r0, int c0, const M3<S>& b) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) (*this)(r0 + i, c0 + j) = b(i, j);
  }
  void addBlock(int r0, int c0, const M3<S>& b) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) (*this)(r0 + i, c0 + j) += b(i, j);
  }
};
template<class S> inline void toVec(const Motion<S>& m, S* x) { for (int k = 0; k < 3; ++k) { x[k] = m.lin[k]; x[3 + k] = m.ang[k]; } }
template<class S> inline void toVec(const Force<S>& m, S* x) { for (int k = 0; k < 3; ++k) { x[k] = m.lin[k]; x[3 + k] = m.ang[k]; } }
template<class S> inline Motion<S> motionFrom(const S* x) { return {v3(x[0], x[1], x[2]), v3(x[3], x[4], x[5])}; }
template<class S> inline Force<S> forceFrom(const S* x) { return {v3(x[0], x[1], x[2]), v3(x[3], x[4], x[5])}; }
template<class S> inline void matvec6(const M6<S>& A, const S* x, S* y) {
  for (int i = 0; i < 6; ++i) {
    S s = S(0);
    for (int j = 0; j < 6; ++j) s += A(i, j) * x[j];
    y[i] = s;
  }
}
template<class S> inline void matTvec6(const M6<S>& A, const S* x, S* y) {
  for (int i = 0; i < 6; ++i) {
    S s = S(0);
    for (int j = 0; j < 6; ++j) s += A(j, i) * x[j];
    y[i] = s;
  }
}

// inertia.hpp: InertiaTpl (m, c, I_c)
template<class S> struct Inertia {
  S m;
  V3<S> c;
  Sym3<S> I;
  static Inertia zero() { Inertia r; r.m = S(0); r.c = zero3<S>(); for (auto& e : r.I.d) e = S(0); return r; }
  // inertia.hpp:729-735 : f = Y v
  Force<S> operator*(const Motion<S>& v) const {
    const V3<S> fl = m * (v.lin - cross(c, v.ang));
    return {fl, I * v.ang + cross(c, fl)};
  }
  // inertia.hpp:481-497 : 6x6 matrix
  M6<S> matrix() const {
    M6<S> r = M6<S>::zero();
    for (int k = 0; k < 3; ++k) r(k, k) = m;
    const M3<S> mcx = skew(m * c);
    r.setBlock(0, 3, -mcx);
    r.setBlock(3, 0, mcx);
    // I_c - m [c]x [c]x  (AlphaSkewSquare, symmetric3.hpp:222-270)
    const M3<S> cc = skewSquare(c, c);
    M3<S> b = I.matrix();
    for (int k = 0; k < 9; ++k) b.m[k] -= m * cc.m[k];
    r.setBlock(3, 3, b);
    return r;
  }
  // inertia.hpp:641-657 : composite of two inertias
  Inertia operator+(const Inertia& o) const {
    const S eps = std::numeric_limits<S>::epsilon();
    const S mab = m + o.m;
    const S mab_inv = S(1) / std::max(mab, eps);
    const V3<S> AB = c - o.c;
    Inertia r;
    r.m = mab;
    r.c = mab_inv * (m * c + o.m * o.c);
    // I = I_a + I_b - (m_a m_b / (m_a + m_b)) [AB]x^2
    const M3<S> sq = skewSquare(AB, AB);
    const S k = m * o.m * mab_inv;
    M3<S> Im = I.matrix() + o.I.matrix();
    for (int e = 0; e < 9; ++e) Im.m[e] -= k * sq.m[e];
    r.I = Sym3<S>::fromMatrix(Im);
    return r;
  }
  // inertia.hpp:750-787 : time variation of the inertia along v (matrix form)
  M6<S> variation(const Motion<S>& v) const {
    M6<S> res = M6<S>::zero();
    const Motion<S> mv = m * v;
    M3<S> la = -skew(mv.lin) - skewSquare(mv.ang, c) + skewSquare(c, mv.ang);
    res.setBlock(0, 3, la);
    res.setBlock(3, 0, transpose(la));
    M3<S> aa = -skewSquare(mv.lin, c) - skewSquare(c, mv.lin);
    // I - AlphaSkewSquare(m, c) : rotational inertia about the frame origin
    const M3<S> cc = skewSquare(c, c);
    M3<S> ll = I.matrix();
    for (int e = 0; e < 9; ++e) ll.m[e] -= m * cc.m[e];
    aa = aa - ll * skew(v.ang);
    aa = aa + skew(v.ang) * ll;
    res.setBlock(3, 3, aa);
    return res;
  }
};
// inertia.hpp:873-880 : se3Action on an inertia
template<class S> inline Inertia<S> act(const SE3<S>& M, const Inertia<S>& Y) {
  Inertia<S> r;
  r.m = Y.m;
  r.c = M.R * Y.c + M.p;
  r.I = Sym3<S>::fromMatrix(M.R * Y.I.matrix() * transpose(M.R));
  return r;
}
// rnea-derivatives.hxx:324-333 : add the matrix of f x* (.) to a 6x6 block
template<class S> inline void addForceCrossMatrix(const Force<S>& f, M6<S>& M) {
  M.addBlock(0, 3, skew(-f.lin));
  M.addBlock(3, 0, skew(-f.lin));
  M.addBlock(3, 3, skew(-f.ang));
}
// aba.hxx:21-76 : SE3 action on a 6x6 (articulated) inertia, X^* I X^{-1}
template<class S> inline M6<S> actOn6(const SE3<S>& M, const M6<S>& I) {
  // Build the 6x6 action matrix of M on motions: X = [[R, [p]x R], [0, R]]
  M6<S> X = M6<S>::zero();
  X.setBlock(0, 0, M.R);
  X.setBlock(0, 3, skew(M.p) * M.R);
  X.setBlock(3, 3, M.R);
  // Force action is X^{-T}: [[R, 0], [[p]x R, R]]; result = Xf I X^{-1} with X^{-1} = Xf^T
  M6<S> Xf = M6<S>::zero();
  Xf.setBlock(0, 0, M.R);
  Xf.setBlock(3, 0, skew(M.p) * M.R);
  Xf.setBlock(3, 3, M.R);
  M6<S> T = M6<S>::zero(), r = M6<S>::zero();
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      S s = S(0);
      for (int k = 0; k < 6; ++k) s += Xf(i, k) * I(k, j);
      T(i, j) = s;
    }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      S s = S(0);
      for (int k = 0; k < 6; ++k) s += T(i, k) * Xf(j, k);
      r(i, j) = s;
    }
  return r;
}

// ----------------------------------------------------------------------------------------
// Dense helpers (column-major, like Eigen's default)
// ----------------------------------------------------------------------------------------
template<class S> struct Mat {
  int rows = 0, cols = 0;
  std::vector<S> a;
  Mat() = default;
  Mat(int r, int c) : rows(r), cols(c), a((size_t)r * c, S(0)) {}
  S& operator()(int r, int c) { return a[(size_t)c * rows + r]; }
  const S& operator()(int r, int c) const { return a[(size_t)c * rows + r]; }
  void setZero() { std::fill(a.begin(), a.end(), S(0)); }
};
// 6 x n column set: column k is a Motion or a Force.
template<class S> inline Motion<S> colM(const Mat<S>& A, int k) { return motionFrom(&A.a[(size_t)k * 6]); }
template<class S> inline Force<S> colF(const Mat<S>& A, int k) { return forceFrom(&A.a[(size_t)k * 6]); }
template<class S> inline void setCol(Mat<S>& A, int k, const Motion<S>& m) { toVec(m, &A.a[(size_t)k * 6]); }
template<class S> inline void setCol(Mat<S>& A, int k, const Force<S>& f) { toVec(f, &A.a[(size_t)k * 6]); }
template<class S> inline S colDot(const Mat<S>& A, int i, const Mat<S>& B, int j) {
  S s = S(0);
  for (int k = 0; k < 6; ++k) s += A.a[(size_t)i * 6 + k] * B.a[(size_t)j * 6 + k];
  return s;
}

// ----------------------------------------------------------------------------------------
// Model description and model
// ----------------------------------------------------------------------------------------
enum JointType : int32_t { JT_UNIVERSE = 0, JT_FREEFLYER = 1, JT_RX = 2, JT_RY = 3, JT_RZ = 4 };

// Flattened model descriptor shared with the C API (all arrays have njoints entries;
// entry 0 is the universe and is ignored).
struct ModelDesc {
  int32_t njoints;
  const int32_t* parents;
  const int32_t* jtype;
  const double* placement;  // [njoints][12]: R row-major (9) then translation (3)
  const double* inertia;    // [njoints][10]: mass, lever (3), I_c (xx, xy, yy, xz, yz, zz)
  const double* armature;   // [nv] or null
  const double* gravity;    // [6] linear then angular, or null for (0,0,-9.81, 0,0,0)
};

template<class S> struct Model {
  int njoints = 0, nq = 0, nv = 0;
  std::vector<int> parents, jtype, idx_q, idx_v, nqs, nvs;
  std::vector<SE3<S>> placements;
  std::vector<Inertia<S>> inertias;
  std::vector<S> armature;
  Motion<S> gravity;
  // topology helpers, data.hxx:196-315
  std::vector<int> lastChild, nvSubtree, parents_fromRow, nvSubtree_fromRow;

  explicit Model(const ModelDesc& d) {
    njoints = d.njoints;
    if (njoints < 1) throw std::invalid_argument("model needs at least the universe joint");
    parents.assign(d.parents, d.parents + njoints);
    jtype.assign(d.jtype, d.jtype + njoints);
    idx_q.assign(njoints, 0); idx_v.assign(njoints, 0); nqs.assign(njoints, 0); nvs.assign(njoints, 0);
    placements.resize(njoints);
    inertias.resize(njoints);
    // model.hxx:61-170 (addJoint): indices are running sums in insertion order.
    for (int i = 1; i < njoints; ++i) {
      if (parents[i] < 0 || parents[i] >= i) throw std::invalid_argument("parent index must precede child");
      switch (jtype[i]) {
        case JT_FREEFLYER: nqs[i] = 7; nvs[i] = 6; break;
        case JT_RX: case JT_RY: case JT_RZ: nqs[i] = 1; nvs[i] = 1; break;
        default: throw std::invalid_argument("unsupported joint type");
      }
      idx_q[i] = nq; idx_v[i] = nv;
      nq += nqs[i]; nv += nvs[i];
    }
    for (int i = 0; i < njoints; ++i) {
      const double* P = d.placement + 12 * i;
      for (int k = 0; k < 9; ++k) placements[i].R.m[k] = S(P[k]);
      for (int k = 0; k < 3; ++k) placements[i].p[k] = S(P[9 + k]);
      const double* Y = d.inertia + 10 * i;
      inertias[i].m = S(Y[0]);
      for (int k = 0; k < 3; ++k) inertias[i].c[k] = S(Y[1 + k]);
      for (int k = 0; k < 6; ++k) inertias[i].I.d[k] = S(Y[4 + k]);
    }
    armature.assign(nv, S(0));
    if (d.armature) for (int k = 0; k < nv; ++k) armature[k] = S(d.armature[k]);
    if (d.gravity) gravity = Motion<S>{v3(S(d.gravity[0]), S(d.gravity[1]), S(d.gravity[2])),
                                       v3(S(d.gravity[3]), S(d.gravity[4]), S(d.gravity[5]))};
    else gravity = Motion<S>{v3(S(0), S(0), S(-9.81)), zero3<S>()};
    computeTopology();
  }

  void computeTopology() {
    // data.hxx:196-242 (computeLastChild) and 244-315 (computeParents_fromRow)
    lastChild.assign(njoints, -1);
    nvSubtree.assign(njoints, 0);
    for (int i = njoints - 1; i >= 0; --i) {
      if (lastChild[i] == -1) lastChild[i] = i;
      const int p = parents[i];
      if (i > 0) lastChild[p] = std::max(lastChild[i], lastChild[p]);
      nvSubtree[i] = idx_v[lastChild[i]] + nvs[lastChild[i]] - idx_v[i];
    }
    parents_fromRow.assign(nv, -1);
    nvSubtree_fromRow.assign(nv, 0);
    for (int j = 1; j < njoints; ++j) {
      const int p = parents[j];
      parents_fromRow[idx_v[j]] = p > 0 ? idx_v[p] + nvs[p] - 1 : -1;
      nvSubtree_fromRow[idx_v[j]] = nvSubtree[j];
      for (int r = 1; r < nvs[j]; ++r) {
        parents_fromRow[idx_v[j] + r] = idx_v[j] + r - 1;
        nvSubtree_fromRow[idx_v[j] + r] = nvSubtree[j] - r;
      }
    }
  }

  // crba.hxx:573-595 : compact (DFS-contiguous) ordering check
  bool isDescendant(int j, int root) const {
    if (j >= njoints) return false;
    if (j == 0) return root == 0;
    return j == root || isDescendant(parents[j], root);
  }
  bool isCompact() const {
    for (int i = 1; i < njoints - 1; ++i) {
      int k = i + 1;
      while (isDescendant(k, i)) ++k;
      for (; k < njoints; ++k)
        if (isDescendant(k, i)) return false;
    }
    return true;
  }
};

// ----------------------------------------------------------------------------------------
// Joint kernels
// ----------------------------------------------------------------------------------------
template<class S> struct JointCalc {
  SE3<S> M;          // joint transform
  Motion<S> v;       // joint velocity S * qdot
  // motion subspace columns (nv of them) in the joint frame
  int nv;
  Motion<S> Scol[6];
};

// joint-revolute.hpp:793-820 and joint-free-flyer.hpp:333-375
template<class S> inline JointCalc<S> jointCalc(const Model<S>& model, int i, const S* q, const S* v) {
  JointCalc<S> jc;
  const int t = model.jtype[i];
  const S* qi = q + model.idx_q[i];
  if (t == JT_FREEFLYER) {
    jc.nv = 6;
    const S x = qi[3], y = qi[4], z = qi[5], w = qi[6];
    // Eigen Quaternion::toRotationMatrix (not renormalised, joint-free-flyer.hpp:349-351)
    const S tx = S(2) * x, ty = S(2) * y, tz = S(2) * z;
    const S twx = tx * w, twy = ty * w, twz = tz * w;
    const S txx = tx * x, txy = ty * x, txz = tz * x;
    const S tyy = ty * y, tyz = tz * y, tzz = tz * z;
    M3<S>& R = jc.M.R;
    R(0, 0) = S(1) - (tyy + tzz); R(0, 1) = txy - twz;          R(0, 2) = txz + twy;
    R(1, 0) = txy + twz;          R(1, 1) = S(1) - (txx + tzz); R(1, 2) = tyz - twx;
    R(2, 0) = txz - twy;          R(2, 1) = tyz + twx;          R(2, 2) = S(1) - (txx + tyy);
    jc.M.p = v3(qi[0], qi[1], qi[2]);
    if (v) {
      const S* vi = v + model.idx_v[i];
      jc.v = Motion<S>{v3(vi[0], vi[1], vi[2]), v3(vi[3], vi[4], vi[5])};
    } else jc.v = zeroMotion<S>();
    for (int k = 0; k < 6; ++k) {
      S e[6] = {0, 0, 0, 0, 0, 0};
      e[k] = S(1);
      jc.Scol[k] = motionFrom(e);
    }
  } else {
    jc.nv = 1;
    const int ax = t - JT_RX;
    using std::cos; using std::sin;
    const S s = sin(qi[0]), c = cos(qi[0]);
    M3<S> R = eye3<S>();
    const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
    R(a1, a1) = c; R(a1, a2) = -s;
    R(a2, a1) = s; R(a2, a2) = c;
    jc.M = SE3<S>{R, zero3<S>()};
    Motion<S> Sc = zeroMotion<S>();
    Sc.ang[ax] = S(1);
    jc.Scol[0] = Sc;
    const S qd = v ? v[model.idx_v[i]] : S(0);
    jc.v = qd * Sc;
  }
  return jc;
}

// ----------------------------------------------------------------------------------------
// Data workspace (data.hpp:115-414), only the members the restated algorithms use
// ----------------------------------------------------------------------------------------
template<class S> struct Blk36;

template<class S> struct Data {
  int nj, nv;
  std::vector<SE3<S>> liMi, oMi;
  std::vector<Motion<S>> v, a, a_gf, ov, oa, oa_gf;
  std::vector<Force<S>> f, h, of, oh;
  std::vector<Inertia<S>> oinertias, oYcrb;
  std::vector<M6<S>> Yaba, oYaba, doYcrb;
  std::vector<JointCalc<S>> jc;
  // per-joint ABA quantities (U, Dinv, UDinv), stored as 6 x nv_j and nv_j x nv_j
  std::vector<Blk36<S>> U, Dinv, UDinv;
  std::vector<S> tau, u, ddq;
  Mat<S> J, dJ, dVdq, dAdq, dAdv, dFdq, dFdv, dFda, SDinv, Fcrb0;
  std::vector<Mat<S>> Fcrb;  // per joint 6 x nv
  Mat<S> M, Minv, dtau_dq, dtau_dv, Ufac;
  std::vector<S> D, Dinvd, tmp;

  explicit Data(const Model<S>& m) : nj(m.njoints), nv(m.nv) {
    liMi.assign(nj, SE3<S>::identity()); oMi = liMi;
    v.assign(nj, zeroMotion<S>()); a = a_gf = ov = oa = oa_gf = v;
    f.assign(nj, zeroForce<S>()); h = of = oh = f;
    oinertias.assign(nj, Inertia<S>::zero()); oYcrb = oinertias;
    Yaba.assign(nj, M6<S>::zero()); oYaba = doYcrb = Yaba;
    jc.resize(nj);
    U.resize(nj); Dinv.resize(nj); UDinv.resize(nj);
    tau.assign(nv, S(0)); u = ddq = tau;
    J = dJ = dVdq = dAdq = dAdv = dFdq = dFdv = dFda = SDinv = Fcrb0 = Mat<S>(6, nv);
    Fcrb.assign(nj, Mat<S>(6, nv));
    M = Minv = dtau_dq = dtau_dv = Ufac = Mat<S>(nv, nv);
    D.assign(nv, S(0)); Dinvd = tmp = D;
  }
};

// Fixed-size scratch for the per-joint ABA blocks (at most 6 x 6), allocation free.
template<class S> struct Blk36 {
  S a[36];
  S& operator[](size_t k) { return a[k]; }
  const S& operator[](size_t k) const { return a[k]; }
  void zero() {
    for (auto& e : a) e = S(0);
  }
};

// PerformStYSInversion (joint-common-operations.hpp:22-45): D^{-1} of a small SPD block.
// Restated as a Cholesky-based inverse (LLT solve against the identity).
template<class S> inline void spdInverse(const Blk36<S>& A, int n, Blk36<S>& Ainv) {
  Blk36<S> L;
  L.zero();
  for (int j = 0; j < n; ++j) {
    S s = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    if (!(s > S(0))) throw std::runtime_error("StYS block is not positive definite");
    using std::sqrt;
    const S ljj = sqrt(s);
    L[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      S t = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) t -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = t / ljj;
    }
  }
  Ainv.zero();
  for (int c = 0; c < n; ++c) {
    S y[6];
    for (int i = 0; i < n; ++i) {
      S t = (i == c) ? S(1) : S(0);
      for (int k = 0; k < i; ++k) t -= L[(size_t)i * n + k] * y[k];
      y[i] = t / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      S t = y[i];
      for (int k = i + 1; k < n; ++k) t -= L[(size_t)k * n + i] * Ainv[(size_t)k * n + c];
      Ainv[(size_t)i * n + c] = t / L[(size_t)i * n + i];
    }
  }
}

// ----------------------------------------------------------------------------------------
// forwardKinematics, kinematics.hxx:121-145 (first order) and 209-238 (second order):
// oMi = oMi[parent] liMi, v_i = vJ + liMi^-1 v_parent, a_i = S qdd + c + v_i x vJ
// + liMi^-1 a_parent (c = 0 for the joints restated here).  v and a may be null.
// ----------------------------------------------------------------------------------------
template<class S>
inline void forwardKinematics(const Model<S>& model, Data<S>& data, const S* q, const S* v, const S* a) {
  data.v[0] = zeroMotion<S>();
  data.a[0] = zeroMotion<S>();
  for (int i = 1; i < model.njoints; ++i) {
    const int p = model.parents[i];
    const JointCalc<S> jc = jointCalc(model, i, q, v);
    data.liMi[i] = model.placements[i] * jc.M;
    data.oMi[i] = p > 0 ? data.oMi[p] * data.liMi[i] : data.liMi[i];
    if (v) {
      data.v[i] = jc.v;
      if (p > 0) data.v[i] = data.v[i] + actInv(data.liMi[i], data.v[p]);
    }
    if (v && a) {
      Motion<S> ai = crossMM(data.v[i], jc.v);
      for (int k = 0; k < jc.nv; ++k) ai = ai + a[model.idx_v[i] + k] * jc.Scol[k];
      if (p > 0) ai = ai + actInv(data.liMi[i], data.a[p]);
      data.a[i] = ai;
    }
  }
}

// computeJointJacobians, jacobian.hxx:23-62: column of dof k of joint i = oMi.act(S_k)
// (WORLD frame); the 6 x nv result is data.J.
template<class S> inline void computeJointJacobians(const Model<S>& model, Data<S>& data, const S* q) {
  forwardKinematics(model, data, q, static_cast<const S*>(nullptr), static_cast<const S*>(nullptr));
  for (int i = 1; i < model.njoints; ++i) {
    const JointCalc<S> jc = jointCalc(model, i, q, static_cast<const S*>(nullptr));
    for (int k = 0; k < jc.nv; ++k) setCol(data.J, model.idx_v[i] + k, act(data.oMi[i], jc.Scol[k]));
  }
}

// computeJointJacobiansTimeVariation, jacobian.hxx:673-740 (JointJacobiansTimeVariation
// ForwardStep): J as above, ov_i = oMi.act(v_i) the joint's spatial velocity in WORLD, and
// each of joint i's dJ columns = ov_i x J_col (motion action).  dJ is 6 x nv like J.
template<class S>
inline void computeJointJacobiansTimeVariation(const Model<S>& model, Data<S>& data, const S* q, const S* v, Mat<S>& dJ) {
  forwardKinematics(model, data, q, v, static_cast<const S*>(nullptr));
  for (int i = 1; i < model.njoints; ++i) {
    const JointCalc<S> jc = jointCalc(model, i, q, static_cast<const S*>(nullptr));
    const Motion<S> ov = act(data.oMi[i], data.v[i]);
    for (int k = 0; k < jc.nv; ++k) {
      const Motion<S> Jc = act(data.oMi[i], jc.Scol[k]);
      setCol(data.J, model.idx_v[i] + k, Jc);
      setCol(dJ, model.idx_v[i] + k, crossMM(ov, Jc));
    }
  }
}

// ----------------------------------------------------------------------------------------
// RNEA, rnea.hxx:17-161 (LOCAL frames)
// ----------------------------------------------------------------------------------------
// fext (optional): an external force on each joint 1.., [linear; angular] in the joint's
// local frame, 6 per joint (rnea.hpp:60-79: f_i -= fext_i in the forward step)
template<class S>
inline void rnea(const Model<S>& model, Data<S>& data, const S* q, const S* v, const S* a, S* tau_out,
                 const S* fext = nullptr) {
  std::fill(data.tau.begin(), data.tau.end(), S(0));
  data.v[0] = zeroMotion<S>();
  data.a_gf[0] = S(-1) * model.gravity;
  for (int i = 1; i < model.njoints; ++i) {  // RneaForwardStep, rnea.hxx:44-79
    const int p = model.parents[i];
    const JointCalc<S> jc = jointCalc(model, i, q, v);
    data.liMi[i] = model.placements[i] * jc.M;
    data.v[i] = jc.v;
    if (p > 0) data.v[i] = data.v[i] + actInv(data.liMi[i], data.v[p]);
    Motion<S> ag = crossMM(data.v[i], jc.v);
    for (int k = 0; k < jc.nv; ++k) ag = ag + a[model.idx_v[i] + k] * jc.Scol[k];
    ag = ag + actInv(data.liMi[i], data.a_gf[p]);
    data.a_gf[i] = ag;
    data.h[i] = model.inertias[i] * data.v[i];
    data.f[i] = model.inertias[i] * data.a_gf[i] + crossMF(data.v[i], data.h[i]);
    if (fext) {
      const S* fe = fext + (size_t)(i - 1) * 6;
      data.f[i] = Force<S>{data.f[i].lin - v3(fe[0], fe[1], fe[2]), data.f[i].ang - v3(fe[3], fe[4], fe[5])};
    }
    data.jc[i] = jc;
  }
  for (int i = model.njoints - 1; i > 0; --i) {  // RneaBackwardStep, rnea.hxx:91-107
    const int p = model.parents[i];
    const JointCalc<S>& jc = data.jc[i];
    for (int k = 0; k < jc.nv; ++k) data.tau[model.idx_v[i] + k] += dot(jc.Scol[k], data.f[i]);
    if (p > 0) data.f[p] = data.f[p] + act(data.liMi[i], data.f[i]);
  }
  for (int k = 0; k < model.nv; ++k) data.tau[k] += model.armature[k] * a[k];  // rnea.hxx:158
  if (tau_out) std::copy(data.tau.begin(), data.tau.end(), tau_out);
}

// ----------------------------------------------------------------------------------------
// ABA, WORLD convention (aba.hxx:79-293), the convention abaInParallel uses
// ----------------------------------------------------------------------------------------
template<class S>
inline void abaWorld(const Model<S>& model, Data<S>& data, const S* q, const S* v, const S* tau, S* ddq_out) {
  const int nj = model.njoints;
  data.oa_gf[0] = S(-1) * model.gravity;
  data.of[0] = zeroForce<S>();
  std::copy(tau, tau + model.nv, data.u.begin());
  for (int i = 1; i < nj; ++i) {  // AbaWorldConventionForwardStep1, aba.hxx:100-138
    const int p = model.parents[i];
    const JointCalc<S> jc = jointCalc(model, i, q, v);
    data.jc[i] = jc;
    data.liMi[i] = model.placements[i] * jc.M;
    data.oMi[i] = p > 0 ? data.oMi[p] * data.liMi[i] : data.liMi[i];
    for (int k = 0; k < jc.nv; ++k) setCol(data.J, model.idx_v[i] + k, act(data.oMi[i], jc.Scol[k]));
    Motion<S> ov = act(data.oMi[i], jc.v);
    if (p > 0) ov = ov + data.ov[p];
    data.ov[i] = ov;
    data.oa_gf[i] = zeroMotion<S>();  // oMi.act(c), c = 0 for FreeFlyer and revolute
    if (p > 0) data.oa_gf[i] = data.oa_gf[i] + crossMM(data.ov[p], ov);
    data.oinertias[i] = data.oYcrb[i] = act(data.oMi[i], model.inertias[i]);
    data.oYaba[i] = data.oYcrb[i].matrix();
    data.oh[i] = data.oYcrb[i] * ov;
    data.of[i] = crossMF(ov, data.oh[i]);
  }
  for (int i = nj - 1; i > 0; --i) {  // AbaWorldConventionBackwardStep, aba.hxx:151-192
    const int p = model.parents[i];
    const int nvi = model.nvs[i], iv = model.idx_v[i];
    M6<S>& Ia = data.oYaba[i];
    S fi[6]; toVec(data.of[i], fi);
    Blk36<S>& U = data.U[i];
    U.zero();  // U is 6 x nvi, column-major
    Blk36<S> StU;
    for (int k = 0; k < nvi; ++k) {
      const S* Jk = &data.J.a[(size_t)(iv + k) * 6];
      S d = S(0);
      for (int r = 0; r < 6; ++r) d += Jk[r] * fi[r];
      data.u[iv + k] -= d;
      matvec6(Ia, Jk, &U[(size_t)k * 6]);
    }
    for (int r = 0; r < nvi; ++r)
      for (int c = 0; c < nvi; ++c) {
        S s = S(0);
        for (int e = 0; e < 6; ++e) s += data.J.a[(size_t)(iv + r) * 6 + e] * U[(size_t)c * 6 + e];
        StU[(size_t)r * nvi + c] = s;
      }
    for (int k = 0; k < nvi; ++k) StU[(size_t)k * nvi + k] += model.armature[iv + k];
    spdInverse(StU, nvi, data.Dinv[i]);
    Blk36<S>& UD = data.UDinv[i];
    UD.zero();
    for (int c = 0; c < nvi; ++c)
      for (int r = 0; r < 6; ++r) {
        S s = S(0);
        for (int k = 0; k < nvi; ++k) s += U[(size_t)k * 6 + r] * data.Dinv[i][(size_t)k * nvi + c];
        UD[(size_t)c * 6 + r] = s;
      }
    if (p > 0) {
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
          S s = S(0);
          for (int k = 0; k < nvi; ++k) s += UD[(size_t)k * 6 + r] * U[(size_t)k * 6 + c];
          Ia(r, c) -= s;
        }
      S ag[6]; toVec(data.oa_gf[i], ag);
      S t[6]; matvec6(Ia, ag, t);
      for (int r = 0; r < 6; ++r) {
        S s = t[r];
        for (int k = 0; k < nvi; ++k) s += UD[(size_t)k * 6 + r] * data.u[iv + k];
        fi[r] += s;
      }
      data.of[i] = forceFrom(fi);
      for (int e = 0; e < 36; ++e) data.oYaba[p].m[e] += Ia.m[e];
      data.of[p] = data.of[p] + data.of[i];
    }
  }
  for (int i = 1; i < nj; ++i) {  // AbaWorldConventionForwardStep2, aba.hxx:205-231
    const int p = model.parents[i];
    const int nvi = model.nvs[i], iv = model.idx_v[i];
    data.oa_gf[i] = data.oa_gf[i] + data.oa_gf[p];
    S ag[6]; toVec(data.oa_gf[i], ag);
    for (int r = 0; r < nvi; ++r) {
      S s = S(0);
      for (int k = 0; k < nvi; ++k) s += data.Dinv[i][(size_t)r * nvi + k] * data.u[iv + k];
      for (int e = 0; e < 6; ++e) s -= data.UDinv[i][(size_t)r * 6 + e] * ag[e];
      data.ddq[iv + r] = s;
    }
    for (int k = 0; k < nvi; ++k) data.oa_gf[i] = data.oa_gf[i] + data.ddq[iv + k] * colM(data.J, iv + k);
    data.oa[i] = data.oa_gf[i] + model.gravity;
    data.of[i] = data.oinertias[i] * data.oa_gf[i] + crossMF(data.ov[i], data.oh[i]);
  }
  for (int i = nj - 1; i > 0; --i) {  // aba.hxx:286-290
    const int p = model.parents[i];
    data.of[p] = data.of[p] + data.of[i];
  }
  if (ddq_out) std::copy(data.ddq.begin(), data.ddq.end(), ddq_out);
}

// ----------------------------------------------------------------------------------------
// ABA, LOCAL convention (aba.hxx:354-540), the single-configuration default
// ----------------------------------------------------------------------------------------
// fext as in rnea (aba.hpp:63-85: the bias force f_i = v_i x* I_i v_i - fext_i)
template<class S>
inline void abaLocal(const Model<S>& model, Data<S>& data, const S* q, const S* v, const S* tau, S* ddq_out,
                     const S* fext = nullptr) {
  const int nj = model.njoints;
  data.v[0] = zeroMotion<S>();
  data.a_gf[0] = S(-1) * model.gravity;
  data.f[0] = zeroForce<S>();
  std::copy(tau, tau + model.nv, data.u.begin());
  for (int i = 1; i < nj; ++i) {
    const int p = model.parents[i];
    const JointCalc<S> jc = jointCalc(model, i, q, v);
    data.jc[i] = jc;
    data.liMi[i] = model.placements[i] * jc.M;
    data.v[i] = jc.v;
    if (p > 0) data.v[i] = data.v[i] + actInv(data.liMi[i], data.v[p]);
    data.a_gf[i] = crossMM(data.v[i], jc.v);
    data.Yaba[i] = model.inertias[i].matrix();
    data.h[i] = model.inertias[i] * data.v[i];
    data.f[i] = crossMF(data.v[i], data.h[i]);
    if (fext) {
      const S* fe = fext + (size_t)(i - 1) * 6;
      data.f[i] = Force<S>{data.f[i].lin - v3(fe[0], fe[1], fe[2]), data.f[i].ang - v3(fe[3], fe[4], fe[5])};
    }
  }
  for (int i = nj - 1; i > 0; --i) {
    const int p = model.parents[i];
    const int nvi = model.nvs[i], iv = model.idx_v[i];
    const JointCalc<S>& jc = data.jc[i];
    M6<S>& Ia = data.Yaba[i];
    for (int k = 0; k < nvi; ++k) data.u[iv + k] -= dot(jc.Scol[k], data.f[i]);
    // calc_aba (joint-revolute.hpp:823-836, joint-free-flyer.hpp:377-391)
    Blk36<S>& U = data.U[i];
    U.zero();
    for (int k = 0; k < nvi; ++k) {
      S e[6]; toVec(jc.Scol[k], e);
      matvec6(Ia, e, &U[(size_t)k * 6]);
    }
    Blk36<S> StU;
    for (int r = 0; r < nvi; ++r)
      for (int c = 0; c < nvi; ++c) {
        S e[6]; toVec(jc.Scol[r], e);
        S s = S(0);
        for (int k = 0; k < 6; ++k) s += e[k] * U[(size_t)c * 6 + k];
        StU[(size_t)r * nvi + c] = s;
      }
    for (int k = 0; k < nvi; ++k) StU[(size_t)k * nvi + k] += model.armature[iv + k];
    spdInverse(StU, nvi, data.Dinv[i]);
    Blk36<S>& UD = data.UDinv[i];
    UD.zero();
    for (int c = 0; c < nvi; ++c)
      for (int r = 0; r < 6; ++r) {
        S s = S(0);
        for (int k = 0; k < nvi; ++k) s += U[(size_t)k * 6 + r] * data.Dinv[i][(size_t)k * nvi + c];
        UD[(size_t)c * 6 + r] = s;
      }
    if (p > 0) {
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
          S s = S(0);
          for (int k = 0; k < nvi; ++k) s += UD[(size_t)k * 6 + r] * U[(size_t)k * 6 + c];
          Ia(r, c) -= s;
        }
      S pa[6]; toVec(data.f[i], pa);
      S ag[6]; toVec(data.a_gf[i], ag);
      S t[6]; matvec6(Ia, ag, t);
      for (int r = 0; r < 6; ++r) {
        S s = t[r];
        for (int k = 0; k < nvi; ++k) s += UD[(size_t)k * 6 + r] * data.u[iv + k];
        pa[r] += s;
      }
      data.f[i] = forceFrom(pa);
      const M6<S> Ip = actOn6(data.liMi[i], Ia);
      for (int e = 0; e < 36; ++e) data.Yaba[p].m[e] += Ip.m[e];
      data.f[p] = data.f[p] + act(data.liMi[i], data.f[i]);
    }
  }
  for (int i = 1; i < nj; ++i) {
    const int p = model.parents[i];
    const int nvi = model.nvs[i], iv = model.idx_v[i];
    const JointCalc<S>& jc = data.jc[i];
    data.a_gf[i] = data.a_gf[i] + actInv(data.liMi[i], data.a_gf[p]);
    S ag[6]; toVec(data.a_gf[i], ag);
    for (int r = 0; r < nvi; ++r) {
      S s = S(0);
      for (int k = 0; k < nvi; ++k) s += data.Dinv[i][(size_t)r * nvi + k] * data.u[iv + k];
      for (int e = 0; e < 6; ++e) s -= data.UDinv[i][(size_t)r * 6 + e] * ag[e];
      data.ddq[iv + r] = s;
    }
    for (int k = 0; k < nvi; ++k) data.a_gf[i] = data.a_gf[i] + data.ddq[iv + k] * jc.Scol[k];
  }
  if (ddq_out) std::copy(data.ddq.begin(), data.ddq.end(), ddq_out);
}

// ----------------------------------------------------------------------------------------
// CRBA, WORLD convention (crba.hxx:20-99, 498-548).  Fills the upper triangle of M and
// adds the armature to the diagonal, exactly like the reference.
// ----------------------------------------------------------------------------------------
template<class S> inline void crbaWorld(const Model<S>& model, Data<S>& data, const S* q) {
  const int nj = model.njoints;
  data.M.setZero();
  data.oYcrb[0] = Inertia<S>::zero();
  for (int i = 1; i < nj; ++i) {
    const int p = model.parents[i];
    const JointCalc<S> jc = jointCalc(model, i, q, (const S*)nullptr);
    data.jc[i] = jc;
    data.liMi[i] = model.placements[i] * jc.M;
    data.oMi[i] = p > 0 ? data.oMi[p] * data.liMi[i] : data.liMi[i];
    for (int k = 0; k < jc.nv; ++k) setCol(data.J, model.idx_v[i] + k, act(data.oMi[i], jc.Scol[k]));
    data.oYcrb[i] = act(data.oMi[i], model.inertias[i]);
  }
  Mat<S>& Ag = data.dFda;  // centroidal momentum map columns
  for (int i = nj - 1; i > 0; --i) {
    const int p = model.parents[i];
    const int iv = model.idx_v[i], nvi = model.nvs[i];
    for (int k = 0; k < nvi; ++k) setCol(Ag, iv + k, data.oYcrb[i] * colM(data.J, iv + k));
    for (int r = 0; r < nvi; ++r)
      for (int c = iv; c < iv + model.nvSubtree[i]; ++c) data.M(iv + r, c) = colDot(data.J, iv + r, Ag, c);
    data.oYcrb[p] = data.oYcrb[p] + data.oYcrb[i];
  }
  for (int k = 0; k < model.nv; ++k) data.M(k, k) += model.armature[k];
}

// ----------------------------------------------------------------------------------------
// Tree-sparse Cholesky M = U D U^T (cholesky.hxx:17-62), solves (193-317, 466-520) and
// computeMinv (522-597).  The factor is stored in data.Ufac (unit upper triangular).
// ----------------------------------------------------------------------------------------
template<class S> inline void choleskyDecompose(const Model<S>& model, Data<S>& data) {
  const Mat<S>& M = data.M;
  Mat<S>& U = data.Ufac;
  U.setZero();
  for (int k = 0; k < model.nv; ++k) U(k, k) = S(1);
  for (int j = model.nv - 1; j >= 0; --j) {
    const int NVT = model.nvSubtree_fromRow[j] - 1;
    for (int k = 0; k < NVT; ++k) data.tmp[k] = U(j, j + 1 + k) * data.D[j + 1 + k];
    S s = M(j, j);
    for (int k = 0; k < NVT; ++k) s -= U(j, j + 1 + k) * data.tmp[k];
    data.D[j] = s;
    data.Dinvd[j] = S(1) / s;
    for (int i = model.parents_fromRow[j]; i >= 0; i = model.parents_fromRow[i]) {
      S t = M(i, j);
      for (int k = 0; k < NVT; ++k) t -= U(i, j + 1 + k) * data.tmp[k];
      U(i, j) = t * data.Dinvd[j];
    }
  }
}
template<class S> inline void choleskyUiv(const Model<S>& model, const Data<S>& data, S* x) {
  for (int k = model.nv - 2; k >= 0; --k) {
    const int n = model.nvSubtree_fromRow[k] - 1;
    S s = S(0);
    for (int e = 0; e < n; ++e) s += data.Ufac(k, k + 1 + e) * x[k + 1 + e];
    x[k] -= s;
  }
}
template<class S> inline void choleskyUtiv(const Model<S>& model, const Data<S>& data, S* x) {
  for (int k = 0; k < model.nv - 1; ++k) {
    const int n = model.nvSubtree_fromRow[k] - 1;
    for (int e = 0; e < n; ++e) x[k + 1 + e] -= data.Ufac(k, k + 1 + e) * x[k];
  }
}
template<class S> inline void choleskySolve(const Model<S>& model, const Data<S>& data, S* x) {
  choleskyUiv(model, data, x);
  for (int k = 0; k < model.nv; ++k) x[k] *= data.Dinvd[k];
  choleskyUtiv(model, data, x);
}
template<class S> inline void choleskyComputeMinv(const Model<S>& model, const Data<S>& data, Mat<S>& Minv) {
  const int nv = model.nv;
  std::vector<S> col(nv);
  for (int c = 0; c < nv; ++c) {
    std::fill(col.begin(), col.end(), S(0));
    col[c] = S(1);
    choleskySolve(model, data, col.data());
    for (int r = 0; r < nv; ++r) Minv(r, c) = col[r];
  }
}

// ----------------------------------------------------------------------------------------
// Frames: updateFramePlacements (frames.hxx:15-35), oMf = oMi[parent] * placement, and
// getFrameJacobian (frames.hxx:156-175 with translateJointJacobian, jacobian.hxx:175-300):
// the support columns of the parent joint's WORLD Jacobian, expressed in WORLD (as they
// are), LOCAL_WORLD_ALIGNED (linear part moved to the frame origin: lin -= p x ang) or LOCAL
// (oMf^-1 applied); the other columns are zero.  Needs computeJointJacobians first.
// ----------------------------------------------------------------------------------------
enum { RF_WORLD = 0, RF_LOCAL = 1, RF_LOCAL_WORLD_ALIGNED = 2 };
template<class S>
inline void frameJacobian(const Model<S>& model, const Data<S>& data, int parent, const SE3<S>& placement, int rf,
                          SE3<S>& oMf, S* J) {
  oMf = data.oMi[parent] * placement;
  std::fill(J, J + 6 * model.nv, S(0));
  for (int j = parent; j > 0; j = model.parents[j])
    for (int k = 0; k < model.nvs[j]; ++k) {
      const int c = model.idx_v[j] + k;
      Motion<S> m = colM(data.J, c);
      if (rf == RF_LOCAL_WORLD_ALIGNED) m.lin = m.lin - cross(oMf.p, m.ang);
      else if (rf == RF_LOCAL) m = actInv(oMf, m);
      toVec(m, J + 6 * c);
    }
}

// getFrameJacobianTimeVariation (frames.hxx:293-368) after
// computeJointJacobiansTimeVariation: the support columns of the joint dJ taken into the
// reference frame like frameJacobian takes J, then for LOCAL
//   dJ_c -= v_frame x (oMf^-1 . J_c),  v_frame = placement^-1 . v_parent,
// and for LOCAL_WORLD_ALIGNED
//   dJ_c.lin -= (ov_parent.lin + ov_parent.ang x oMf.p) x J_c.ang.
template<class S>
inline void frameJacobianTimeVariation(const Model<S>& model, const Data<S>& data, const Mat<S>& dJj, int parent,
                                       const SE3<S>& placement, int rf, S* dJ) {
  const SE3<S> oMf = data.oMi[parent] * placement;
  std::fill(dJ, dJ + 6 * model.nv, S(0));
  const Motion<S> vf = actInv(placement, data.v[parent]);
  const Motion<S> ov = act(data.oMi[parent], data.v[parent]);
  for (int j = parent; j > 0; j = model.parents[j])
    for (int k = 0; k < model.nvs[j]; ++k) {
      const int c = model.idx_v[j] + k;
      Motion<S> m = colM(dJj, c);
      const Motion<S> Jc = colM(data.J, c);
      if (rf == RF_LOCAL_WORLD_ALIGNED) {
        m.lin = m.lin - cross(oMf.p, m.ang);
        m.lin = m.lin - cross(ov.lin + cross(ov.ang, oMf.p), Jc.ang);
      } else if (rf == RF_LOCAL) {
        m = actInv(oMf, m);
        const Motion<S> x = crossMM(vf, actInv(oMf, Jc));
        m.lin = m.lin - x.lin;
        m.ang = m.ang - x.ang;
      }
      toVec(m, dJ + 6 * c);
    }
}

// ----------------------------------------------------------------------------------------
// Frame motions after forwardKinematics(q, v[, a]): getFrameVelocity (frames.hxx:71-99),
// getFrameAcceleration (frames.hxx:102-130), getFrameClassicalAcceleration
// (frames.hxx:133-149), for the frame (parent joint, placement):
//   LOCAL                  placement^-1 . m_parent
//   WORLD                  oMi[parent] . m_parent
//   LOCAL_WORLD_ALIGNED    (R_i (m.lin + m.ang x placement.p), R_i m.ang)
// with m = data.v / data.a of the parent joint; classical: acc.lin += vel.ang x vel.lin
// (both in the requested frame).
// ----------------------------------------------------------------------------------------
template<class S>
inline Motion<S> frameMotion(const Data<S>& data, int parent, const SE3<S>& placement, int rf, const Motion<S>& m) {
  const SE3<S>& oMi = data.oMi[parent];
  switch (rf) {
    case RF_LOCAL: return actInv(placement, m);
    case RF_WORLD: return act(oMi, m);
    case RF_LOCAL_WORLD_ALIGNED: return Motion<S>{oMi.R * (m.lin + cross(m.ang, placement.p)), oMi.R * m.ang};
  }
  throw std::invalid_argument("Bad reference frame.");
}
template<class S>
inline void frameMotions(const Data<S>& data, int parent, const SE3<S>& placement, int rf, bool withAcc, Motion<S>& vel,
                         Motion<S>& acc, Motion<S>& classical) {
  vel = frameMotion(data, parent, placement, rf, data.v[parent]);
  if (!withAcc) return;
  acc = frameMotion(data, parent, placement, rf, data.a[parent]);
  classical = acc;
  classical.lin = classical.lin + cross(vel.ang, vel.lin);
}

// ----------------------------------------------------------------------------------------
// forwardDynamics with acceleration constraints J ddq + gamma = 0
// (contact-dynamics.hxx:90-167): computeAllTerms gives M (CRBA) and nle (RNEA at a = 0),
// then cholesky::decompose and
//   torque_residual = M^-1 (tau - nle)
//   sDUiJt = D^-1/2 U^-1 J^T,  JMinvJt = sDUiJt^T sDUiJt + inv_damping I
//   lambda_c = JMinvJt^-1 (-J torque_residual - gamma)        (LLT, no pivoting)
//   ddq = M^-1 J^T lambda_c + torque_residual
// J is nc x nv, column major (element (i, k) at J[i + k nc]).
// ----------------------------------------------------------------------------------------
template<class S>
inline void constrainedForwardDynamics(const Model<S>& model, Data<S>& data, const S* q, const S* v, const S* tau,
                                       int nc, const S* J, const S* gamma, S inv_damping, S* ddq, S* lambda) {
  using std::sqrt;
  const int nv = model.nv;
  std::vector<S> r(nv), zero(nv, S(0));
  rnea(model, data, q, v, zero.data(), r.data());
  for (int k = 0; k < nv; ++k) r[k] = tau[k] - r[k];
  crbaWorld(model, data, q);
  choleskyDecompose(model, data);
  choleskySolve(model, data, r.data());
  std::vector<S> sd((size_t)nv * nc), G((size_t)nc * nc), L((size_t)nc * nc, S(0));
  for (int i = 0; i < nc; ++i) {
    S* col = &sd[(size_t)i * nv];
    for (int k = 0; k < nv; ++k) col[k] = J[i + (size_t)k * nc];
    choleskyUiv(model, data, col);
    for (int k = 0; k < nv; ++k) col[k] /= sqrt(data.D[k]);
  }
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j) {
      S g = S(0);
      for (int k = 0; k < nv; ++k) g += sd[(size_t)i * nv + k] * sd[(size_t)j * nv + k];
      G[i + (size_t)j * nc] = i == j ? g + inv_damping : g;
    }
  for (int i = 0; i < nc; ++i) {
    S s = S(0);
    for (int k = 0; k < nv; ++k) s += J[i + (size_t)k * nc] * r[k];
    lambda[i] = -s - gamma[i];
  }
  // G = L L^T, then the two triangular solves
  for (int j = 0; j < nc; ++j) {
    S d = G[j + (size_t)j * nc];
    for (int k = 0; k < j; ++k) d -= L[j + (size_t)k * nc] * L[j + (size_t)k * nc];
    L[j + (size_t)j * nc] = sqrt(d);
    for (int i = j + 1; i < nc; ++i) {
      S t = G[i + (size_t)j * nc];
      for (int k = 0; k < j; ++k) t -= L[i + (size_t)k * nc] * L[j + (size_t)k * nc];
      L[i + (size_t)j * nc] = t / L[j + (size_t)j * nc];
    }
  }
  for (int i = 0; i < nc; ++i) {
    S s = lambda[i];
    for (int k = 0; k < i; ++k) s -= L[i + (size_t)k * nc] * lambda[k];
    lambda[i] = s / L[i + (size_t)i * nc];
  }
  for (int i = nc - 1; i >= 0; --i) {
    S s = lambda[i];
    for (int k = i + 1; k < nc; ++k) s -= L[k + (size_t)i * nc] * lambda[k];
    lambda[i] = s / L[i + (size_t)i * nc];
  }
  for (int k = 0; k < nv; ++k) {
    S s = S(0);
    for (int i = 0; i < nc; ++i) s += J[i + (size_t)k * nc] * lambda[i];
    ddq[k] = s;
  }
  choleskySolve(model, data, ddq);
  for (int k = 0; k < nv; ++k) ddq[k] += r[k];
}

// ----------------------------------------------------------------------------------------
// impulseDynamics (contact-dynamics.hxx:264-331): crba, cholesky::decompose, then
//   sDUiJt = D^-1/2 U^-1 J^T,  JMinvJt = sDUiJt^T sDUiJt + inv_damping I
//   impulse_c = JMinvJt^-1 (-(1 + r_coeff) J v_before)         (LLT)
//   dq_after = M^-1 J^T impulse_c + v_before
// ----------------------------------------------------------------------------------------
template<class S>
inline void impulseDynamics(const Model<S>& model, Data<S>& data, const S* q, const S* v_before, int nc, const S* J,
                            S r_coeff, S inv_damping, S* dq_after, S* impulse) {
  using std::sqrt;
  const int nv = model.nv;
  crbaWorld(model, data, q);
  choleskyDecompose(model, data);
  std::vector<S> sd((size_t)nv * nc), G((size_t)nc * nc), L((size_t)nc * nc, S(0));
  for (int i = 0; i < nc; ++i) {
    S* col = &sd[(size_t)i * nv];
    for (int k = 0; k < nv; ++k) col[k] = J[i + (size_t)k * nc];
    choleskyUiv(model, data, col);
    for (int k = 0; k < nv; ++k) col[k] /= sqrt(data.D[k]);
  }
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j) {
      S g = S(0);
      for (int k = 0; k < nv; ++k) g += sd[(size_t)i * nv + k] * sd[(size_t)j * nv + k];
      G[i + (size_t)j * nc] = i == j ? g + inv_damping : g;
    }
  for (int i = 0; i < nc; ++i) {
    S s = S(0);
    for (int k = 0; k < nv; ++k) s += J[i + (size_t)k * nc] * v_before[k];
    impulse[i] = (-r_coeff - S(1)) * s;
  }
  for (int j = 0; j < nc; ++j) {
    S d = G[j + (size_t)j * nc];
    for (int k = 0; k < j; ++k) d -= L[j + (size_t)k * nc] * L[j + (size_t)k * nc];
    L[j + (size_t)j * nc] = sqrt(d);
    for (int i = j + 1; i < nc; ++i) {
      S t = G[i + (size_t)j * nc];
      for (int k = 0; k < j; ++k) t -= L[i + (size_t)k * nc] * L[j + (size_t)k * nc];
      L[i + (size_t)j * nc] = t / L[j + (size_t)j * nc];
    }
  }
  for (int i = 0; i < nc; ++i) {
    S s = impulse[i];
    for (int k = 0; k < i; ++k) s -= L[i + (size_t)k * nc] * impulse[k];
    impulse[i] = s / L[i + (size_t)i * nc];
  }
  for (int i = nc - 1; i >= 0; --i) {
    S s = impulse[i];
    for (int k = i + 1; k < nc; ++k) s -= L[k + (size_t)i * nc] * impulse[k];
    impulse[i] = s / L[i + (size_t)i * nc];
  }
  for (int k = 0; k < nv; ++k) {
    S s = S(0);
    for (int i = 0; i < nc; ++i) s += J[i + (size_t)k * nc] * impulse[i];
    dq_after[k] = s;
  }
  choleskySolve(model, data, dq_after);
  for (int k = 0; k < nv; ++k) dq_after[k] += v_before[k];
}

// ----------------------------------------------------------------------------------------
// constraintDynamics with rigid constraints (constrained-dynamics.hxx:193-500), restricted
// to CONTACT_3D / CONTACT_6D constraints expressed in LOCAL or LOCAL_WORLD_ALIGNED:
//   RigidConstraintModel::calc / jacobian (contact-info.hpp:426-700): oMc1 = oMi[j1] M1,
//   oMc2 = oMi[j2] M2, c1Mc2 = oMc1^-1 oMc2, and the constraint Jacobian from the WORLD joint
//   Jacobian columns, signed by which joint's support holds the column;
//   the contact drifts from ov, oa of the forward step at a = 0 (constrained-dynamics.hxx:
//   71-135: oa = oa_parent + ov_parent x ov, the joints here having c = 0), placement and
//   velocity errors, Baumgarte correction -Kp e - Kd de (log6 for 6D placements);
//   the KKT [[-mu I, J], [J^T, M]] [x; ddq] = [rhs + mu lambda_prox; tau - nle] solved
//   max_iter times (proximal iterations) with the reference's residuals:
//     absolute = |(J M^-1 J^T) x - rhs|_inf,  relative = |x + lambda_prox|_inf,
//   stopping when either is within its accuracy; lambda = -x, lambda_prox = -x.
// The reference factors the whole KKT matrix (contact Cholesky, contact-cholesky.hxx:
// 184-324); here the same system is solved through its Schur complement
// (J M^-1 J^T + mu I, LLT) and the tree factor of M -- the same solution up to rounding.
// ----------------------------------------------------------------------------------------
enum { CONTACT_3D = 0, CONTACT_6D = 1 };
template<class S> struct RigidConstraint {
  int type, joint1, joint2, rf;  // rf: RF_LOCAL or RF_LOCAL_WORLD_ALIGNED
  SE3<S> placement1, placement2;
  S Kp[6], Kd[6];                // corrector gains (first 3 for CONTACT_3D)
  S desired[6];                  // contact_acceleration_desired (linear, angular)
  int size() const { return type == CONTACT_6D ? 6 : 3; }
};
template<class S> struct ProximalSettings {
  S absolute_accuracy, relative_accuracy, mu;
  int max_iter;
};

// unSkew (spatial/skew.hpp): axial vector of the antisymmetric part of R
template<class S> inline V3<S> unSkew(const M3<S>& R) {
  return v3(S(0.5) * (R(2, 1) - R(1, 2)), S(0.5) * (R(0, 2) - R(2, 0)), S(0.5) * (R(1, 0) - R(0, 1)));
}
// log3 (spatial/log.hxx:58-183) away from theta = pi (the constraints' placement errors are
// small): theta from the trace (Taylor expansion near 0), w = theta / sin(theta) unSkew(R)
template<class S> inline V3<S> log3(const M3<S>& R, S& theta) {
  using std::acos; using std::sin; using std::sqrt; using std::cbrt;
  const S eps = std::numeric_limits<S>::epsilon();
  const S prec = cbrt(eps);  // TaylorSeriesExpansion<S>::precision<2>()
  const S tr = R(0, 0) + R(1, 1) + R(2, 2);
  const S cv = (tr - S(1)) / S(2);
  if (!(tr > S(-1) + prec)) throw std::invalid_argument("log3: rotation angle too close to pi");
  theta = tr <= S(3) - prec ? acos(cv) : sqrt(S(2) * (S(1) - cv) + eps * eps);
  const V3<S> a = unSkew(R);
  const S n2 = dot(a, a);
  const S t = theta >= prec ? theta / sin(theta) : S(1) + n2 / S(6) + n2 * n2 * S(3) / S(40);
  return t * a;
}
// log6 (spatial/log.hxx:189-230)
template<class S> inline Motion<S> log6(const SE3<S>& M) {
  using std::sin; using std::cos; using std::cbrt;
  const S prec = cbrt(std::numeric_limits<S>::epsilon());
  S theta;
  const V3<S> w = log3(M.R, theta);
  const V3<S> a = unSkew(M.R);
  const S t2 = dot(a, a);
  const S tr = M.R(0, 0) + M.R(1, 1) + M.R(2, 2);
  const S st = sin(theta), ct = cos(theta);
  const bool small = tr >= S(3) - prec;
  const S alpha = small ? S(1) - t2 / S(12) - t2 * t2 / S(720) : theta * st / (S(2) * (S(1) - ct));
  const S beta = small ? S(1) / S(12) + t2 / S(720) : S(1) / (theta * theta) - st / (S(2) * theta * (S(1) - ct));
  const V3<S>& p = M.p;
  return {alpha * p - S(0.5) * cross(w, p) + (beta * dot(w, p)) * w, w};
}
template<class S> inline SE3<S> inverse(const SE3<S>& M) { return SE3<S>{transpose(M.R), -mulT(M.R, M.p)}; }

// rhs rows (skipped when rhs is null) and Jacobian rows (nrows x nv, column major at
// Jc[i + k nrows]) of the constraints, after the forward step (data.oMi, data.ov, data.oa,
// data.J filled).
template<class S>
inline void constraintTerms(const Model<S>& model, const Data<S>& data, int ncon, const RigidConstraint<S>* cons,
                            int nrows, S* rhs, S* Jc) {
  const int nv = model.nv;
  std::fill(Jc, Jc + (size_t)nrows * nv, S(0));
  std::vector<char> sup1(nv), sup2(nv);
  int row = 0;
  for (int c = 0; c < ncon; ++c) {
    const RigidConstraint<S>& cm = cons[c];
    const int j1 = cm.joint1, j2 = cm.joint2, dim = cm.size();
    const SE3<S> oMc1 = j1 > 0 ? data.oMi[j1] * cm.placement1 : cm.placement1;
    const SE3<S> oMc2 = j2 > 0 ? data.oMi[j2] * cm.placement2 : cm.placement2;
    const SE3<S> c1Mc2 = inverse(oMc1) * oMc2;
    const Motion<S> zm = zeroMotion<S>();
    const Motion<S>& ov1 = j1 > 0 ? data.ov[j1] : zm;
    const Motion<S>& ov2 = j2 > 0 ? data.ov[j2] : zm;
    const Motion<S>& oa1 = j1 > 0 ? data.oa[j1] : zm;
    const Motion<S>& oa2 = j2 > 0 ? data.oa[j2] : zm;
    const Motion<S> vc1 = j1 > 0 ? actInv(oMc1, ov1) : zm;
    const Motion<S> vc2 = j2 > 0 ? actInv(oMc2, ov2) : zm;
    if (rhs) {  // drift, errors and correction (not needed for the Jacobian alone)
      const Motion<S> vc2in1 = act(c1Mc2, vc2);
      Motion<S> perr, verr, aerr = zm;
      if (cm.type == CONTACT_6D) {
        const Motion<S> l = log6(c1Mc2);
        perr = {-l.lin, -l.ang};
        verr = vc1 - vc2in1;
      } else {
        perr = {-c1Mc2.p, zero3<S>()};
        verr = {vc1.lin - c1Mc2.R * vc2.lin, zero3<S>()};
      }
      bool gains = false;
      for (int k = 0; k < dim; ++k) gains = gains || cm.Kp[k] != S(0) || cm.Kd[k] != S(0);
      if (gains) {
        for (int k = 0; k < 3; ++k) aerr.lin[k] = -cm.Kp[k] * perr.lin[k] - cm.Kd[k] * verr.lin[k];
        if (cm.type == CONTACT_6D)
          for (int k = 0; k < 3; ++k) aerr.ang[k] = -cm.Kp[3 + k] * perr.ang[k] - cm.Kd[3 + k] * verr.ang[k];
      }
      Motion<S> acc1, acc2;
      if (cm.rf == RF_LOCAL_WORLD_ALIGNED) {
        acc1.lin = oa1.lin + cross(oa1.ang, oMc1.p);
        if (cm.type == CONTACT_3D) acc1.lin += cross(ov1.ang, ov1.lin + cross(ov1.ang, oMc1.p));
        acc1.ang = oa1.ang;
        if (cm.type == CONTACT_3D) {
          acc2.lin = oa2.lin + cross(oa2.ang, oMc2.p) + cross(ov2.ang, ov2.lin + cross(ov2.ang, oMc2.p));
          acc2.ang = zero3<S>();
        } else {
          acc2.lin = oa2.lin + cross(oa2.ang, oMc1.p);  // (the reference takes c1's origin here)
          acc2.ang = oa2.ang;
        }
        aerr = {oMc1.R * aerr.lin, oMc1.R * aerr.ang};
      } else if (cm.rf == RF_LOCAL) {
        acc1 = actInv(oMc1, oa1);
        if (cm.type == CONTACT_3D) {
          acc1.lin += cross(vc1.ang, vc1.lin);
          acc1.ang = zero3<S>();
        } else {
          acc1 = acc1 + crossMM(verr, vc2in1);
        }
        if (cm.type == CONTACT_3D) {
          acc2.lin = mulT(oMc1.R, oa2.lin + cross(oa2.ang, oMc2.p) + cross(ov2.ang, ov2.lin + cross(ov2.ang, oMc2.p)));
          acc2.ang = zero3<S>();
        } else {
          acc2 = actInv(oMc1, oa2);
        }
      } else {
        throw std::invalid_argument("constraintDynamics operates from LOCAL or LOCAL_WORLD_ALIGNED frames");
      }
      S r6[6], a1[6], a2[6], e6[6];
      toVec(acc1, a1);
      toVec(acc2, a2);
      toVec(aerr, e6);
      for (int k = 0; k < 6; ++k) r6[k] = -a1[k] + a2[k] + e6[k] + cm.desired[k];
      for (int k = 0; k < dim; ++k) rhs[row + k] = r6[k];
    }
    // Jacobian (contact-info.hpp:551-700)
    std::fill(sup1.begin(), sup1.end(), 0);
    std::fill(sup2.begin(), sup2.end(), 0);
    for (int j = j1; j > 0; j = model.parents[j])
      for (int k = 0; k < model.nvs[j]; ++k) sup1[model.idx_v[j] + k] = 1;
    for (int j = j2; j > 0; j = model.parents[j])
      for (int k = 0; k < model.nvs[j]; ++k) sup2[model.idx_v[j] + k] = 1;
    for (int k = 0; k < nv; ++k) {
      if (!sup1[k] && !sup2[k]) continue;
      const int sign = sup1[k] != sup2[k] ? (sup1[k] ? 1 : -1) : 0;
      const Motion<S> Jcol = colM(data.J, k);
      S col[6] = {S(0), S(0), S(0), S(0), S(0), S(0)};
      if (cm.type == CONTACT_3D) {
        V3<S> l;
        if (cm.rf == RF_LOCAL) {
          if (sign == 0) l = actInv(oMc1, Jcol).lin - c1Mc2.R * actInv(oMc2, Jcol).lin;
          else if (sign == 1) l = actInv(oMc1, Jcol).lin;
          else l = -(c1Mc2.R * actInv(oMc2, Jcol).lin);
        } else {
          if (sign == 0) l = cross(oMc2.p - oMc1.p, Jcol.ang);
          else l = S(sign) * (Jcol.lin - cross(sign == 1 ? oMc1.p : oMc2.p, Jcol.ang));
        }
        for (int r = 0; r < 3; ++r) col[r] = l[r];
      } else {
        Motion<S> m = cm.rf == RF_LOCAL ? actInv(oMc1, Jcol) : Motion<S>{Jcol.lin - cross(oMc1.p, Jcol.ang), Jcol.ang};
        toVec(S(sign) * m, col);
      }
      for (int r = 0; r < dim; ++r) Jc[(row + r) + (size_t)k * nrows] = col[r];
    }
    row += dim;
  }
}

// forward step (constrained-dynamics.hxx:71-135) for ov, oa and the WORLD joint Jacobian,
// then the constraint rows (rhs, Jc) at q, v
template<class S>
inline void constraintForward(const Model<S>& model, Data<S>& data, const S* q, const S* v, int ncon,
                              const RigidConstraint<S>* cons, int nrows, S* rhs, S* Jc) {
  for (int i = 1; i < model.njoints; ++i) {
    const int p = model.parents[i];
    const JointCalc<S> jc = jointCalc(model, i, q, v);
    data.liMi[i] = model.placements[i] * jc.M;
    data.oMi[i] = p > 0 ? data.oMi[p] * data.liMi[i] : data.liMi[i];
    data.ov[i] = act(data.oMi[i], jc.v);
    if (p > 0) data.ov[i] = data.ov[i] + data.ov[p];
    data.oa[i] = p > 0 ? crossMM(data.ov[p], data.ov[i]) + data.oa[p] : zeroMotion<S>();
    for (int k = 0; k < jc.nv; ++k) setCol(data.J, model.idx_v[i] + k, act(data.oMi[i], jc.Scol[k]));
  }
  constraintTerms(model, data, ncon, cons, nrows, rhs, Jc);
}

// computeDelassusMatrix (delassus.hxx:118-300): the operational-space inverse inertia of the
// constraints, J (M + mu J^T J)^-1 J^T -- the reference adds a contact inertia of mass mu
// (rotational inertia mu I for 6-D constraints) at each constraint frame and runs its OSIM
// sweeps.  Restated as D0 = J M^-1 J^T (tree factor, one solve per row) and, for mu > 0,
// D0 (I + mu D0)^-1 (Woodbury).  G: nrows x nrows, column major.
template<class S>
inline void delassusMatrix(const Model<S>& model, Data<S>& data, const S* q, int ncon, const RigidConstraint<S>* cons,
                           S mu, S* G) {
  if (mu < S(0)) throw std::invalid_argument("mu has to be positive");
  const int nv = model.nv;
  int nrows = 0;
  for (int c = 0; c < ncon; ++c) nrows += cons[c].size();
  std::vector<S> Jc((size_t)nrows * nv), zero(nv, S(0)), Y((size_t)nv * nrows);
  constraintForward(model, data, q, zero.data(), ncon, cons, nrows, (S*)nullptr, Jc.data());
  crbaWorld(model, data, q);
  choleskyDecompose(model, data);
  for (int i = 0; i < nrows; ++i) {
    S* col = &Y[(size_t)i * nv];
    for (int k = 0; k < nv; ++k) col[k] = Jc[i + (size_t)k * nrows];
    choleskySolve(model, data, col);
  }
  std::vector<S> D0((size_t)nrows * nrows);
  for (int i = 0; i < nrows; ++i)
    for (int j = 0; j < nrows; ++j) {
      S g = S(0);
      for (int k = 0; k < nv; ++k) g += Jc[i + (size_t)k * nrows] * Y[(size_t)j * nv + k];
      D0[i + (size_t)j * nrows] = g;
    }
  if (mu == S(0)) {
    std::copy(D0.begin(), D0.end(), G);
    return;
  }
  // A = I + mu D0 (symmetric positive definite); G = A^-1 D0, column by column (LLT)
  std::vector<S> L((size_t)nrows * nrows, S(0));
  for (int j = 0; j < nrows; ++j) {
    S d = S(1) + mu * D0[j + (size_t)j * nrows];
    for (int k = 0; k < j; ++k) d -= L[j + (size_t)k * nrows] * L[j + (size_t)k * nrows];
    L[j + (size_t)j * nrows] = std::sqrt(d);
    for (int i = j + 1; i < nrows; ++i) {
      S t = mu * D0[i + (size_t)j * nrows];
      for (int k = 0; k < j; ++k) t -= L[i + (size_t)k * nrows] * L[j + (size_t)k * nrows];
      L[i + (size_t)j * nrows] = t / L[j + (size_t)j * nrows];
    }
  }
  for (int j = 0; j < nrows; ++j) {
    S* x = G + (size_t)j * nrows;
    for (int i = 0; i < nrows; ++i) x[i] = D0[i + (size_t)j * nrows];
    for (int i = 0; i < nrows; ++i) {
      S s = x[i];
      for (int k = 0; k < i; ++k) s -= L[i + (size_t)k * nrows] * x[k];
      x[i] = s / L[i + (size_t)i * nrows];
    }
    for (int i = nrows - 1; i >= 0; --i) {
      S s = x[i];
      for (int k = i + 1; k < nrows; ++k) s -= L[k + (size_t)i * nrows] * x[k];
      x[i] = s / L[i + (size_t)i * nrows];
    }
  }
}

// A solve with the KKT matrix of the contact Cholesky (contact-cholesky.hpp:292-400),
// K = [[-mu I, J], [J^T, M]] at q: x = K^-1 rhs, rhs and x of nrows + nv entries (the
// constraint rows first).  Restated densely -- K assembled from the constraint Jacobian and
// CRBA's M, then Gaussian elimination with partial pivoting -- where the reference factors K
// with its sparsity and the GPU path goes through the Schur complement.
template<class S>
inline void contactKktSolve(const Model<S>& model, Data<S>& data, const S* q, int ncon, const RigidConstraint<S>* cons,
                            S mu, const S* rhs, S* x) {
  using std::abs;
  const int nv = model.nv;
  int nrows = 0;
  for (int c = 0; c < ncon; ++c) nrows += cons[c].size();
  const int n = nrows + nv;
  std::vector<S> Jc((size_t)nrows * nv), zero(nv, S(0)), K((size_t)n * n, S(0)), b(rhs, rhs + n);
  constraintForward(model, data, q, zero.data(), ncon, cons, nrows, (S*)nullptr, Jc.data());
  crbaWorld(model, data, q);
  auto at = [&](int i, int j) -> S& { return K[(size_t)i * n + j]; };  // row major
  for (int i = 0; i < nrows; ++i) {
    at(i, i) = -mu;
    for (int k = 0; k < nv; ++k) at(i, nrows + k) = at(nrows + k, i) = Jc[i + (size_t)k * nrows];
  }
  for (int i = 0; i < nv; ++i)
    for (int j = 0; j < nv; ++j) at(nrows + i, nrows + j) = i <= j ? data.M(i, j) : data.M(j, i);
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (abs(at(r, c)) > abs(at(p, c))) p = r;
    if (p != c) {
      for (int j = 0; j < n; ++j) std::swap(at(c, j), at(p, j));
      std::swap(b[c], b[p]);
    }
    for (int r = c + 1; r < n; ++r) {
      const S f = at(r, c) / at(c, c);
      if (f == S(0)) continue;
      for (int j = c; j < n; ++j) at(r, j) -= f * at(c, j);
      b[r] -= f * b[c];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    S s = b[i];
    for (int j = i + 1; j < n; ++j) s -= at(i, j) * x[j];
    x[i] = s / at(i, i);
  }
}

template<class S>
inline void constraintDynamics(const Model<S>& model, Data<S>& data, const S* q, const S* v, const S* tau, int ncon,
                               const RigidConstraint<S>* cons, const ProximalSettings<S>& settings, S* ddq, S* lambda,
                               int* iterations) {
  using std::abs;
  using std::sqrt;
  const int nv = model.nv;
  if (settings.mu < S(0)) throw std::invalid_argument("mu has to be positive");
  int nrows = 0;
  for (int c = 0; c < ncon; ++c) nrows += cons[c].size();
  std::vector<S> rhs(nrows), Jc((size_t)nrows * nv), r(nv), zero(nv, S(0));
  constraintForward(model, data, q, v, ncon, cons, nrows, rhs.data(), Jc.data());
  // M, nle, tree factor; r = M^-1 (tau - nle); Y = M^-1 J^T; G = J Y + mu I
  rnea(model, data, q, v, zero.data(), r.data());
  for (int k = 0; k < nv; ++k) r[k] = tau[k] - r[k];
  crbaWorld(model, data, q);
  choleskyDecompose(model, data);
  std::vector<S> rv = r;
  choleskySolve(model, data, r.data());
  std::vector<S> Y((size_t)nv * nrows), G((size_t)nrows * nrows), L((size_t)nrows * nrows, S(0));
  for (int i = 0; i < nrows; ++i) {
    S* col = &Y[(size_t)i * nv];
    for (int k = 0; k < nv; ++k) col[k] = Jc[i + (size_t)k * nrows];
    choleskySolve(model, data, col);
  }
  for (int i = 0; i < nrows; ++i)
    for (int j = 0; j < nrows; ++j) {
      S g = S(0);
      for (int k = 0; k < nv; ++k) g += Jc[i + (size_t)k * nrows] * Y[(size_t)j * nv + k];
      G[i + (size_t)j * nrows] = g;
    }
  for (int j = 0; j < nrows; ++j) {
    S d = G[j + (size_t)j * nrows] + settings.mu;
    for (int k = 0; k < j; ++k) d -= L[j + (size_t)k * nrows] * L[j + (size_t)k * nrows];
    L[j + (size_t)j * nrows] = sqrt(d);
    for (int i = j + 1; i < nrows; ++i) {
      S t = G[i + (size_t)j * nrows];
      for (int k = 0; k < j; ++k) t -= L[i + (size_t)k * nrows] * L[j + (size_t)k * nrows];
      L[i + (size_t)j * nrows] = t / L[j + (size_t)j * nrows];
    }
  }
  // J r: the Schur complement's constant part
  std::vector<S> Jr(nrows), x(nrows), prox(nrows, S(0));
  for (int i = 0; i < nrows; ++i) {
    S s = S(0);
    for (int k = 0; k < nv; ++k) s += Jc[i + (size_t)k * nrows] * r[k];
    Jr[i] = s;
  }
  int it = 0;
  while (it < settings.max_iter) {
    ++it;
    // x = (G + mu I)^-1 (J r - rhs - mu lambda_prox)
    for (int i = 0; i < nrows; ++i) x[i] = Jr[i] - rhs[i] - settings.mu * prox[i];
    for (int i = 0; i < nrows; ++i) {
      S s = x[i];
      for (int k = 0; k < i; ++k) s -= L[i + (size_t)k * nrows] * x[k];
      x[i] = s / L[i + (size_t)i * nrows];
    }
    for (int i = nrows - 1; i >= 0; --i) {
      S s = x[i];
      for (int k = i + 1; k < nrows; ++k) s -= L[k + (size_t)i * nrows] * x[k];
      x[i] = s / L[i + (size_t)i * nrows];
    }
    S absr = S(0), relr = S(0);
    for (int i = 0; i < nrows; ++i) {
      S g = S(0);
      for (int j = 0; j < nrows; ++j) g += G[i + (size_t)j * nrows] * x[j];
      absr = std::max(absr, abs(g - rhs[i]));
      relr = std::max(relr, abs(x[i] + prox[i]));
    }
    for (int i = 0; i < nrows; ++i) prox[i] = -x[i];
    if (absr <= settings.absolute_accuracy || relr <= settings.relative_accuracy) break;
  }
  if (iterations) *iterations = it;
  // ddq = M^-1 (tau - nle - J^T x), lambda = -x
  for (int k = 0; k < nv; ++k) {
    S s = rv[k];
    for (int i = 0; i < nrows; ++i) s -= Jc[i + (size_t)k * nrows] * x[i];
    ddq[k] = s;
  }
  choleskySolve(model, data, ddq);
  for (int i = 0; i < nrows; ++i) lambda[i] = -x[i];
}

// ----------------------------------------------------------------------------------------
// computeRNEADerivatives (rnea-derivatives.hxx:242-541)
// ----------------------------------------------------------------------------------------
template<class S>
inline void rneaDerivatives(const Model<S>& model, Data<S>& data, const S* q, const S* v, const S* a) {
  const int nj = model.njoints;
  if (model.gravity.ang[0] != S(0) || model.gravity.ang[1] != S(0) || model.gravity.ang[2] != S(0))
    throw std::invalid_argument("The gravity must be a pure force vector, no angular part");
  data.dtau_dq.setZero(); data.dtau_dv.setZero(); data.M.setZero();
  data.oa_gf[0] = S(-1) * model.gravity;
  data.ov[0] = zeroMotion<S>();
  for (int i = 1; i < nj; ++i) {  // ComputeRNEADerivativesForwardStep
    const int p = model.parents[i];
    const JointCalc<S> jc = jointCalc(model, i, q, v);
    data.jc[i] = jc;
    data.liMi[i] = model.placements[i] * jc.M;
    data.v[i] = jc.v;
    if (p > 0) {
      data.oMi[i] = data.oMi[p] * data.liMi[i];
      data.v[i] = data.v[i] + actInv(data.liMi[i], data.v[p]);
    } else data.oMi[i] = data.liMi[i];
    Motion<S> ai = crossMM(data.v[i], jc.v);
    for (int k = 0; k < jc.nv; ++k) ai = ai + a[model.idx_v[i] + k] * jc.Scol[k];
    if (p > 0) ai = ai + actInv(data.liMi[i], data.a[p]);
    data.a[i] = ai;
    data.oYcrb[i] = data.oinertias[i] = act(data.oMi[i], model.inertias[i]);
    data.ov[i] = act(data.oMi[i], data.v[i]);
    data.oa[i] = act(data.oMi[i], data.a[i]);
    data.oa_gf[i] = data.oa[i] - model.gravity;
    data.oh[i] = data.oYcrb[i] * data.ov[i];
    data.of[i] = data.oYcrb[i] * data.oa_gf[i] + crossMF(data.ov[i], data.oh[i]);
    for (int k = 0; k < jc.nv; ++k) {
      const int c = model.idx_v[i] + k;
      const Motion<S> Jc = act(data.oMi[i], jc.Scol[k]);
      setCol(data.J, c, Jc);
      const Motion<S> dJ = crossMM(data.ov[i], Jc);
      setCol(data.dJ, c, dJ);
      Motion<S> dAdq = crossMM(data.oa_gf[p], Jc);
      Motion<S> dAdv = dJ;
      if (p > 0) {
        const Motion<S> dVdq = crossMM(data.ov[p], Jc);
        setCol(data.dVdq, c, dVdq);
        dAdq = dAdq + crossMM(data.ov[p], dVdq);
        dAdv = dAdv + dVdq;
      } else setCol(data.dVdq, c, zeroMotion<S>());
      setCol(data.dAdq, c, dAdq);
      setCol(data.dAdv, c, dAdv);
    }
    data.doYcrb[i] = data.oYcrb[i].variation(data.ov[i]);
    addForceCrossMatrix(data.oh[i], data.doYcrb[i]);
  }
  Mat<S>& dYtJ = data.Fcrb0;
  for (int i = nj - 1; i > 0; --i) {  // ComputeRNEADerivativesBackwardStep
    const int p = model.parents[i];
    const int iv = model.idx_v[i], nvi = model.nvs[i];
    const int nvsub = model.nvSubtree[i];
    for (int k = 0; k < nvi; ++k) data.tau[iv + k] = dot(colM(data.J, iv + k), data.of[i]);
    for (int k = 0; k < nvi; ++k) setCol(data.dFda, iv + k, data.oYcrb[i] * colM(data.J, iv + k));
    for (int r = 0; r < nvi; ++r)
      for (int c = iv; c < iv + nvsub; ++c) data.M(iv + r, c) = colDot(data.J, iv + r, data.dFda, c);
    for (int k = 0; k < nvi; ++k) {
      const int c = iv + k;
      S t1[6], t2[6];
      S dv[6]; toVec(colM(data.dVdq, c), dv);
      Force<S> F = data.oYcrb[i] * colM(data.dAdq, c);
      if (p > 0) {
        matvec6(data.doYcrb[i], dv, t1);
        F = F + forceFrom(t1);
      }
      setCol(data.dFdq, c, F);
      S jcv[6]; toVec(colM(data.J, c), jcv);
      matTvec6(data.doYcrb[i], jcv, t2);  // dYtJ_col = doYcrb^T J_col
      setCol(dYtJ, c, forceFrom(t2));
    }
    for (int r = iv + nvi; r < iv + nvsub; ++r)
      for (int k = 0; k < nvi; ++k)
        data.dtau_dq(r, iv + k) = colDot(data.dFda, r, data.dAdq, iv + k) + colDot(dYtJ, r, data.dVdq, iv + k);
    for (int r = 0; r < nvi; ++r)
      for (int c = iv; c < iv + nvsub; ++c) data.dtau_dq(iv + r, c) = colDot(data.J, iv + r, data.dFdq, c);
    for (int k = 0; k < nvi; ++k) {
      const int c = iv + k;
      setCol(data.dFdq, c, colF(data.dFdq, c) + crossMF(colM(data.J, c), data.of[i]));
    }
    for (int k = 0; k < nvi; ++k) {
      const int c = iv + k;
      S jcv[6], t[6]; toVec(colM(data.J, c), jcv);
      matvec6(data.doYcrb[i], jcv, t);
      setCol(data.dFdv, c, forceFrom(t) + data.oYcrb[i] * colM(data.dAdv, c));
    }
    for (int r = iv + nvi; r < iv + nvsub; ++r)
      for (int k = 0; k < nvi; ++k)
        data.dtau_dv(r, iv + k) = colDot(data.dFda, r, data.dAdv, iv + k) + colDot(dYtJ, r, data.J, iv + k);
    for (int r = 0; r < nvi; ++r)
      for (int c = iv; c < iv + nvsub; ++c) data.dtau_dv(iv + r, c) = colDot(data.J, iv + r, data.dFdv, c);
    if (p > 0) {
      data.oYcrb[p] = data.oYcrb[p] + data.oYcrb[i];
      for (int e = 0; e < 36; ++e) data.doYcrb[p].m[e] += data.doYcrb[i].m[e];
      data.of[p] = data.of[p] + data.of[i];
    }
  }
  for (int k = 0; k < model.nv; ++k) {  // rnea-derivatives.hxx:529-535
    Motion<S> dA = colM(data.dAdq, k);
    dA.lin = dA.lin + cross(model.gravity.lin, colM(data.J, k).ang);
    setCol(data.dAdq, k, dA);
  }
  for (int k = 0; k < model.nv; ++k) {
    data.tau[k] += model.armature[k] * a[k];
    data.M(k, k) += model.armature[k];
  }
}

// ----------------------------------------------------------------------------------------
// computeABADerivatives (aba-derivatives.hxx:16-453).  Outputs ddq_dq, ddq_dv and Minv
// (symmetrised), and data.ddq.
// ----------------------------------------------------------------------------------------
template<class S>
inline void abaDerivatives(const Model<S>& model, Data<S>& data, const S* q, const S* v, const S* tau,
                           Mat<S>& ddq_dq, Mat<S>& ddq_dv, Mat<S>& Minv) {
  const int nj = model.njoints, nv = model.nv;
  if (model.gravity.ang[0] != S(0) || model.gravity.ang[1] != S(0) || model.gravity.ang[2] != S(0))
    throw std::invalid_argument("The gravity must be a pure force vector, no angular part");
  data.oa_gf[0] = S(-1) * model.gravity;
  data.ov[0] = zeroMotion<S>();
  std::copy(tau, tau + nv, data.u.begin());
  Minv.setZero();
  data.dtau_dq.setZero(); data.dtau_dv.setZero();
  for (int i = 1; i < nj; ++i) {  // ComputeABADerivativesForwardStep1, aba-derivatives.hxx:22-79
    const int p = model.parents[i];
    const JointCalc<S> jc = jointCalc(model, i, q, v);
    data.jc[i] = jc;
    data.liMi[i] = model.placements[i] * jc.M;
    data.oMi[i] = p > 0 ? data.oMi[p] * data.liMi[i] : data.liMi[i];
    Motion<S> ov = act(data.oMi[i], jc.v);
    if (p > 0) ov = ov + data.ov[p];
    data.ov[i] = ov;
    data.oa_gf[i] = zeroMotion<S>();
    if (p > 0) data.oa_gf[i] = crossMM(data.ov[p], ov);
    data.oYcrb[i] = data.oinertias[i] = act(data.oMi[i], model.inertias[i]);
    data.oYaba[i] = data.oYcrb[i].matrix();
    data.oh[i] = data.oYcrb[i] * ov;
    data.of[i] = crossMF(ov, data.oh[i]);
    for (int k = 0; k < jc.nv; ++k) setCol(data.J, model.idx_v[i] + k, act(data.oMi[i], jc.Scol[k]));
  }
  Mat<S>& Fcrb = data.Fcrb0;
  Fcrb.setZero();
  for (int i = nj - 1; i > 0; --i) {  // ComputeABADerivativesBackwardStep1, aba-derivatives.hxx:97-170
    const int p = model.parents[i];
    const int iv = model.idx_v[i], nvi = model.nvs[i];
    const int nvsub = model.nvSubtree[i];
    M6<S>& Ia = data.oYaba[i];
    S fi[6]; toVec(data.of[i], fi);
    Blk36<S>& U = data.U[i];
    U.zero();
    for (int k = 0; k < nvi; ++k) {
      const S* Jk = &data.J.a[(size_t)(iv + k) * 6];
      S d = S(0);
      for (int r = 0; r < 6; ++r) d += Jk[r] * fi[r];
      data.u[iv + k] -= d;
      matvec6(Ia, Jk, &U[(size_t)k * 6]);
    }
    Blk36<S> StU;
    for (int r = 0; r < nvi; ++r)
      for (int c = 0; c < nvi; ++c) {
        S s = S(0);
        for (int e = 0; e < 6; ++e) s += data.J.a[(size_t)(iv + r) * 6 + e] * U[(size_t)c * 6 + e];
        StU[(size_t)r * nvi + c] = s;
      }
    for (int k = 0; k < nvi; ++k) StU[(size_t)k * nvi + k] += model.armature[iv + k];
    spdInverse(StU, nvi, data.Dinv[i]);
    Blk36<S>& UD = data.UDinv[i];
    UD.zero();
    for (int c = 0; c < nvi; ++c)
      for (int r = 0; r < 6; ++r) {
        S s = S(0);
        for (int k = 0; k < nvi; ++k) s += U[(size_t)k * 6 + r] * data.Dinv[i][(size_t)k * nvi + c];
        UD[(size_t)c * 6 + r] = s;
      }
    for (int r = 0; r < nvi; ++r)
      for (int c = 0; c < nvi; ++c) Minv(iv + r, iv + c) = data.Dinv[i][(size_t)r * nvi + c];
    const int nv_children = nvsub - nvi;
    if (nv_children > 0) {
      // SDinv = J Dinv (6 x nvi); Minv(iv rows, children cols) = -SDinv^T Fcrb(:, children)
      for (int c = 0; c < nvi; ++c)
        for (int r = 0; r < 6; ++r) {
          S s = S(0);
          for (int k = 0; k < nvi; ++k) s += data.J.a[(size_t)(iv + k) * 6 + r] * data.Dinv[i][(size_t)k * nvi + c];
          data.SDinv.a[(size_t)(iv + c) * 6 + r] = s;
        }
      for (int r = 0; r < nvi; ++r)
        for (int c = iv + nvi; c < iv + nvsub; ++c) Minv(iv + r, c) = -colDot(data.SDinv, iv + r, Fcrb, c);
      if (p > 0) {
        for (int c = iv; c < iv + nvsub; ++c)
          for (int r = 0; r < 6; ++r) {
            S s = S(0);
            for (int k = 0; k < nvi; ++k) s += U[(size_t)k * 6 + r] * Minv(iv + k, c);
            Fcrb.a[(size_t)c * 6 + r] += s;
          }
      }
    } else {
      for (int c = iv; c < iv + nvsub; ++c)
        for (int r = 0; r < 6; ++r) {
          S s = S(0);
          for (int k = 0; k < nvi; ++k) s += U[(size_t)k * 6 + r] * Minv(iv + k, c);
          Fcrb.a[(size_t)c * 6 + r] = s;
        }
    }
    if (p > 0) {
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
          S s = S(0);
          for (int k = 0; k < nvi; ++k) s += UD[(size_t)k * 6 + r] * U[(size_t)k * 6 + c];
          Ia(r, c) -= s;
        }
      S ag[6]; toVec(data.oa_gf[i], ag);
      S t[6]; matvec6(Ia, ag, t);
      for (int r = 0; r < 6; ++r) {
        S s = t[r];
        for (int k = 0; k < nvi; ++k) s += UD[(size_t)k * 6 + r] * data.u[iv + k];
        fi[r] += s;
      }
      data.of[i] = forceFrom(fi);
      for (int e = 0; e < 36; ++e) data.oYaba[p].m[e] += Ia.m[e];
      data.of[p] = data.of[p] + data.of[i];
    }
  }
  for (int i = 1; i < nj; ++i) {  // ComputeABADerivativesForwardStep2, aba-derivatives.hxx:187-256
    const int p = model.parents[i];
    const int iv = model.idx_v[i], nvi = model.nvs[i];
    data.oa_gf[i] = data.oa_gf[i] + data.oa_gf[p];
    S ag[6]; toVec(data.oa_gf[i], ag);
    for (int r = 0; r < nvi; ++r) {
      S s = S(0);
      for (int k = 0; k < nvi; ++k) s += data.Dinv[i][(size_t)r * nvi + k] * data.u[iv + k];
      for (int e = 0; e < 6; ++e) s -= data.UDinv[i][(size_t)r * 6 + e] * ag[e];
      data.ddq[iv + r] = s;
    }
    for (int k = 0; k < nvi; ++k) data.oa_gf[i] = data.oa_gf[i] + data.ddq[iv + k] * colM(data.J, iv + k);
    data.oa[i] = data.oa_gf[i] + model.gravity;
    data.of[i] = data.oYcrb[i] * data.oa_gf[i] + crossMF(data.ov[i], data.oh[i]);
    if (p > 0) {
      for (int r = 0; r < nvi; ++r)
        for (int c = iv; c < nv; ++c) {
          S s = S(0);
          for (int e = 0; e < 6; ++e) s += data.UDinv[i][(size_t)r * 6 + e] * data.Fcrb[p].a[(size_t)c * 6 + e];
          Minv(iv + r, c) -= s;
        }
    }
    for (int c = iv; c < nv; ++c)
      for (int r = 0; r < 6; ++r) {
        S s = S(0);
        for (int k = 0; k < nvi; ++k) s += data.J.a[(size_t)(iv + k) * 6 + r] * Minv(iv + k, c);
        if (p > 0) s += data.Fcrb[p].a[(size_t)c * 6 + r];
        data.Fcrb[i].a[(size_t)c * 6 + r] = s;
      }
    for (int k = 0; k < nvi; ++k) {
      const int c = iv + k;
      const Motion<S> Jc = colM(data.J, c);
      const Motion<S> dJ = crossMM(data.ov[i], Jc);
      setCol(data.dJ, c, dJ);
      Motion<S> dAdq = crossMM(data.oa_gf[p], Jc);
      Motion<S> dAdv = dJ;
      if (p > 0) {
        const Motion<S> dVdq = crossMM(data.ov[p], Jc);
        setCol(data.dVdq, c, dVdq);
        dAdq = dAdq + crossMM(data.ov[p], dVdq);
        dAdv = dAdv + dVdq;
      } else setCol(data.dVdq, c, zeroMotion<S>());
      setCol(data.dAdq, c, dAdq);
      setCol(data.dAdv, c, dAdv);
    }
    data.doYcrb[i] = data.oYcrb[i].variation(data.ov[i]);
    addForceCrossMatrix(data.oh[i], data.doYcrb[i]);
  }
  for (int i = nj - 1; i > 0; --i) {  // ComputeABADerivativesBackwardStep2, aba-derivatives.hxx:282-367
    const int p = model.parents[i];
    const int iv = model.idx_v[i], nvi = model.nvs[i];
    const int nvsub = model.nvSubtree[i];
    for (int k = 0; k < nvi; ++k) {
      const int c = iv + k;
      S jcv[6], t[6]; toVec(colM(data.J, c), jcv);
      matvec6(data.doYcrb[i], jcv, t);
      setCol(data.dFdv, c, data.oYcrb[i] * colM(data.dAdv, c) + forceFrom(t));
    }
    for (int r = 0; r < nvi; ++r)
      for (int c = iv; c < iv + nvsub; ++c) data.dtau_dv(iv + r, c) = colDot(data.J, iv + r, data.dFdv, c);
    for (int k = 0; k < nvi; ++k) {
      const int c = iv + k;
      Force<S> F = data.oYcrb[i] * colM(data.dAdq, c);
      if (p > 0) {
        S dv[6], t[6]; toVec(colM(data.dVdq, c), dv);
        matvec6(data.doYcrb[i], dv, t);
        F = F + forceFrom(t);
      }
      setCol(data.dFdq, c, F);
    }
    for (int r = 0; r < nvi; ++r)
      for (int c = iv; c < iv + nvsub; ++c) data.dtau_dq(iv + r, c) = colDot(data.J, iv + r, data.dFdq, c);
    for (int k = 0; k < nvi; ++k) {
      const int c = iv + k;
      setCol(data.dFdq, c, colF(data.dFdq, c) + crossMF(colM(data.J, c), data.of[i]));
      setCol(data.dFda, c, data.oYcrb[i] * colM(data.J, c));
    }
    if (p > 0) {
      for (int r = 0; r < nvi; ++r) {
        const int row = iv + r;
        S StdY[6];
        S jcv[6]; toVec(colM(data.J, row), jcv);
        matTvec6(data.doYcrb[i], jcv, StdY);  // row r of J^T doYcrb
        const Force<S> sdy = forceFrom(StdY);
        for (int j = model.parents_fromRow[iv]; j >= 0; j = model.parents_fromRow[j]) {
          data.dtau_dq(row, j) = colDot(data.dFda, row, data.dAdq, j) + dot(colM(data.dVdq, j), sdy);
          data.dtau_dv(row, j) = colDot(data.dFda, row, data.dAdv, j) + dot(colM(data.J, j), sdy);
        }
      }
      data.oYcrb[p] = data.oYcrb[p] + data.oYcrb[i];
      for (int e = 0; e < 36; ++e) data.doYcrb[p].m[e] += data.doYcrb[i].m[e];
      data.of[p] = data.of[p] + data.of[i];
    }
    for (int k = 0; k < nvi; ++k) {
      const int c = iv + k;
      Motion<S> dA = colM(data.dAdq, c);
      dA.lin = dA.lin + cross(model.gravity.lin, colM(data.J, c).ang);
      setCol(data.dAdq, c, dA);
    }
  }
  // aba-derivatives.hxx:448-452
  for (int c = 0; c < nv; ++c)
    for (int r = c + 1; r < nv; ++r) Minv(r, c) = Minv(c, r);
  for (int c = 0; c < nv; ++c)
    for (int r = 0; r < nv; ++r) {
      S s1 = S(0), s2 = S(0);
      for (int k = 0; k < nv; ++k) {
        s1 += Minv(r, k) * data.dtau_dq(k, c);
        s2 += Minv(r, k) * data.dtau_dv(k, c);
      }
      ddq_dq(r, c) = -s1;
      ddq_dv(r, c) = -s2;
    }
}

// ----------------------------------------------------------------------------------------
// Configuration space: integrate (special-euclidean.hpp:659-698 with exp6 from
// explog-quaternion.hpp:92-147; revolute joints are Euclidean)
// ----------------------------------------------------------------------------------------
template<class S> inline void quatMul(const S* a, const S* b, S* r) {  // [x, y, z, w]
  const S ax = a[0], ay = a[1], az = a[2], aw = a[3];
  const S bx = b[0], by = b[1], bz = b[2], bw = b[3];
  r[0] = aw * bx + ax * bw + ay * bz - az * by;
  r[1] = aw * by - ax * bz + ay * bw + az * bx;
  r[2] = aw * bz + ax * by - ay * bx + az * bw;
  r[3] = aw * bw - ax * bx - ay * by - az * bz;
}
template<class S> inline void integrate(const Model<S>& model, const S* q, const S* v, S* qout) {
  for (int i = 1; i < model.njoints; ++i) {
    const int iq = model.idx_q[i], iv = model.idx_v[i];
    if (model.jtype[i] == JT_FREEFLYER) {
      const V3<S> vl = v3(v[iv], v[iv + 1], v[iv + 2]);
      const V3<S> w = v3(v[iv + 3], v[iv + 4], v[iv + 5]);
      const S t2 = dot(w, w), t = std::sqrt(t2);
      // exp6: rotation quaternion and translation V(w) vl
      S qe[4];
      S a, b, c;  // V = a I + b [w]x + c [w]x^2
      if (t < S(1e-6)) {
        const S h = S(0.5);
        qe[0] = h * w[0]; qe[1] = h * w[1]; qe[2] = h * w[2]; qe[3] = S(1) - t2 / S(8);
        a = S(1); b = S(0.5) - t2 / S(24); c = S(1) / S(6) - t2 / S(120);
      } else {
        const S s = std::sin(t / S(2)) / t;
        qe[0] = s * w[0]; qe[1] = s * w[1]; qe[2] = s * w[2]; qe[3] = std::cos(t / S(2));
        a = S(1); b = (S(1) - std::cos(t)) / t2; c = (t - std::sin(t)) / (t2 * t);
      }
      const V3<S> wx = cross(w, vl);
      const V3<S> tr = a * vl + b * wx + c * cross(w, wx);
      // out translation = quat * tr + p
      SE3<S> M0;
      {
        S x = q[iq + 3], y = q[iq + 4], z = q[iq + 5], ww = q[iq + 6];
        const S tx = S(2) * x, ty = S(2) * y, tz = S(2) * z;
        M3<S>& R = M0.R;
        R(0, 0) = S(1) - (ty * y + tz * z); R(0, 1) = ty * x - tz * ww; R(0, 2) = tz * x + ty * ww;
        R(1, 0) = ty * x + tz * ww; R(1, 1) = S(1) - (tx * x + tz * z); R(1, 2) = tz * y - tx * ww;
        R(2, 0) = tz * x - ty * ww; R(2, 1) = tz * y + tx * ww; R(2, 2) = S(1) - (tx * x + ty * y);
      }
      const V3<S> pt = M0.R * tr;
      for (int k = 0; k < 3; ++k) qout[iq + k] = q[iq + k] + pt[k];
      S qr[4];
      quatMul(q + iq + 3, qe, qr);
      S dp = S(0);
      for (int k = 0; k < 4; ++k) dp += qr[k] * q[iq + 3 + k];
      if (dp < S(0)) for (int k = 0; k < 4; ++k) qr[k] = -qr[k];
      S n = std::sqrt(qr[0] * qr[0] + qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
      for (int k = 0; k < 4; ++k) qout[iq + 3 + k] = qr[k] / n;
    } else {
      qout[iq] = q[iq] + v[iv];
    }
  }
}

// ----------------------------------------------------------------------------------------
// Energies (energy.hpp): independent pins for M (kinetic) and gravity torques (potential)
// ----------------------------------------------------------------------------------------
template<class S> inline S kineticEnergy(const Model<S>& model, Data<S>& data, const S* q, const S* v) {
  S T = S(0);
  for (int i = 1; i < model.njoints; ++i) {
    const int p = model.parents[i];
    const JointCalc<S> jc = jointCalc(model, i, q, v);
    data.liMi[i] = model.placements[i] * jc.M;
    data.v[i] = jc.v;
    if (p > 0) data.v[i] = data.v[i] + actInv(data.liMi[i], data.v[p]);
    T += S(0.5) * dot(data.v[i], model.inertias[i] * data.v[i]);
  }
  return T;
}
template<class S> inline S potentialEnergy(const Model<S>& model, Data<S>& data, const S* q) {
  S V = S(0);
  for (int i = 1; i < model.njoints; ++i) {
    const int p = model.parents[i];
    const JointCalc<S> jc = jointCalc(model, i, q, (const S*)nullptr);
    data.liMi[i] = model.placements[i] * jc.M;
    data.oMi[i] = p > 0 ? data.oMi[p] * data.liMi[i] : data.liMi[i];
    const V3<S> com = data.oMi[i].R * model.inertias[i].c + data.oMi[i].p;
    V -= model.inertias[i].m * dot(model.gravity.lin, com);
  }
  return V;
}

}  // namespace oracle
