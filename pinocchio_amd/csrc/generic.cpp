// generic.cpp -- run-time topology programs and the launchers of the run-time-topology
// kernels (k_generic.hpp).
#include "generic.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

#include "k_generic.hpp"
#include "runtime.hpp"

namespace rbd {

// ---------------------------------------------------------------------------------------
// Program construction (what ModelTpl::addJoint and DataTpl's topology helpers derive,
// model.hxx:61-170, data.hxx:196-315, for an arbitrary tree)
// ---------------------------------------------------------------------------------------
GenericModel* buildGenericModel(int nj, const int* parents, const int* types) {
  auto g = std::make_unique<GenericModel>();
  g->nj = nj;
  std::vector<int> nvj(nj, 0), iq(nj, 0), iv(nj, 0), depth(nj, 0), nchild(nj, 0);
  for (int j = 1; j < nj; ++j) {
    const bool ff = types[j] == JT_FF;
    iq[j] = g->nq;
    iv[j] = g->nv;
    nvj[j] = ff ? 6 : 1;
    g->nq += ff ? 7 : 1;
    g->nv += nvj[j];
    depth[j] = parents[j] == 0 ? 1 : depth[parents[j]] + 1;
    g->nlev = std::max(g->nlev, depth[j]);
    if (parents[j] > 0) nchild[parents[j]] += 1;
  }
  const int nv = g->nv;
  // dof tree (parents_fromRow): a joint's first dof hangs below the last dof of its parent
  std::vector<int> dofParent(nv, -1);
  for (int j = 1; j < nj; ++j) {
    const int p = parents[j];
    dofParent[iv[j]] = p > 0 ? iv[p] + nvj[p] - 1 : -1;
    for (int r = 1; r < nvj[j]; ++r) dofParent[iv[j] + r] = iv[j] + r - 1;
  }
  // packed strictly-upper pattern, column by column, ancestors from the root down
  std::vector<int> colPtr(nv + 1, 0), patRow;
  for (int c = 0; c < nv; ++c) {
    std::vector<int> anc;
    for (int r = dofParent[c]; r >= 0; r = dofParent[r]) anc.push_back(r);
    colPtr[c] = (int)patRow.size();
    patRow.insert(patRow.end(), anc.rbegin(), anc.rend());
  }
  colPtr[nv] = (int)patRow.size();
  g->nnz = (int)patRow.size();
  g->nzW = std::max(1, (nv + 31) / 32);
  std::vector<uint32_t> nz((size_t)nv * g->nzW, 0u);
  auto setNz = [&](int r, int c) { nz[(size_t)c * g->nzW + (r >> 5)] |= 1u << (r & 31); };
  for (int c = 0; c < nv; ++c) {
    setNz(c, c);
    for (int k = colPtr[c]; k < colPtr[c + 1]; ++k) {
      setNz(patRow[k], c);
      setNz(c, patRow[k]);
    }
  }
  // DFS schedule, children in increasing index order: PRE(j), the children's subtrees,
  // CLOSE(j)
  std::vector<std::vector<int>> kids(nj);
  for (int j = 1; j < nj; ++j) kids[parents[j]].push_back(j);
  std::vector<int32_t> ev;
  std::vector<std::pair<int, size_t>> stack;  // (joint, next child)
  stack.push_back({0, 0});
  while (!stack.empty()) {
    auto& top = stack.back();
    if (top.second < kids[top.first].size()) {
      const int c = kids[top.first][top.second++];
      ev.push_back((c << 1) | EV_PRE);
      stack.push_back({c, 0});
    } else {
      if (top.first != 0) ev.push_back((top.first << 1) | EV_CLOSE);
      stack.pop_back();
    }
  }
  g->nev = (int)ev.size();
  std::vector<int32_t> jr((size_t)nj * GJ_REC, 0);
  for (int j = 1; j < nj; ++j) {
    int32_t* r = &jr[(size_t)j * GJ_REC];
    r[GJ_PARENT] = parents[j];
    r[GJ_TYPE] = types[j];
    r[GJ_IQ] = iq[j];
    r[GJ_IV] = iv[j];
    r[GJ_NV] = nvj[j];
    r[GJ_LEVEL] = depth[j] - 1;
    r[GJ_DOFLEVEL] = colPtr[iv[j] + 1] - colPtr[iv[j]];
    r[GJ_NCHILD] = nchild[j];
    g->nvlev = std::max(g->nvlev, r[GJ_DOFLEVEL] + nvj[j]);
  }
  // one device buffer: ev | jr | colPtr | patRow | nzMask
  std::vector<int32_t> tab;
  const size_t oEv = tab.size();
  tab.insert(tab.end(), ev.begin(), ev.end());
  const size_t oJr = tab.size();
  tab.insert(tab.end(), jr.begin(), jr.end());
  const size_t oCp = tab.size();
  tab.insert(tab.end(), colPtr.begin(), colPtr.end());
  const size_t oPr = tab.size();
  tab.insert(tab.end(), patRow.begin(), patRow.end());
  const size_t oNz = tab.size();
  for (uint32_t w : nz) tab.push_back((int32_t)w);
  tab.push_back(0);  // keeps every sub-table non-empty
  HIP_CHECK(hipMalloc(&g->dtab, tab.size() * sizeof(int32_t)));
  HIP_CHECK(hipMemcpy(g->dtab, tab.data(), tab.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  g->ev = g->dtab + oEv;
  g->jr = g->dtab + oJr;
  g->colPtr = g->dtab + oCp;
  g->patRow = g->dtab + oPr;
  g->nzMask = reinterpret_cast<const uint32_t*>(g->dtab + oNz);
  return g.release();
}

GenericModel::~GenericModel() {
  if (dtab) (void)hipFree(dtab);
  if (lanes) (void)hipFree(lanes);
  for (void* p : retired) (void)hipFree(p);
}

void* GenericModel::laneMem(size_t bytes) const {
  if (bytes > laneBytes) {
    const size_t want = std::max(bytes, 2 * laneBytes);
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, want));
    if (lanes) retired.push_back(lanes);
    lanes = p;
    laneBytes = want;
  }
  return lanes;
}

// ---------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------
namespace {

constexpr int64_t kGMaxGrid = 2048;        // persistent workgroups (eight per CU)
constexpr size_t kGLdsWave = 64 * 1024;    // lane memory kept in LDS up to this size per wave

GProg progOf(GM g) {
  return GProg{g->ev, g->jr, g->colPtr, g->patRow, g->nzMask, g->nev, g->nj, g->nq, g->nv, g->nlev, g->nvlev,
               g->nnz, g->nzW};
}
unsigned tilesGrid(int64_t B) { return (unsigned)std::min<int64_t>((B + kTile - 1) / kTile, kGMaxGrid); }

// dynamic LDS above the default 64 KB per workgroup needs the kernel attribute; false if
// the device refuses
bool allowLds(const void* fn, size_t bytes) {
  if (bytes <= 64 * 1024) return true;
  if (bytes > 160 * 1024) return false;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}

// per-wave lane memory bytes of each per-lane kernel
template<class S> size_t rneaLane(GM g) { return (size_t)g->nlev * kGRneaLvl * kTile * sizeof(S); }
template<class S> size_t abaLane(GM g) {
  return (size_t)(g->nlev * kGAbaLvl + g->nj * kGAbaJnt) * kTile * sizeof(S);
}
template<class S> size_t crbaLane(GM g) {
  return (size_t)(g->nlev * kGCrbaLvl + g->nvlev * kGCrbaDof) * kTile * sizeof(S);
}
template<class S> size_t rneadLane(GM g) {
  return (size_t)(g->nlev * kGRneadLvl + g->nvlev * kGRneadDof) * kTile * sizeof(S);
}
template<class S> size_t kinLane(GM g) { return (size_t)g->nlev * kGKinLvl * kTile * sizeof(S); }

// Launches K<S, true> with the lane memory in LDS when it fits, else K<S, false> over a
// device buffer of one region per workgroup.
#define RBD_G_LAUNCH(KERNEL, S, LANE, B, ...)                                                       \
  do {                                                                                              \
    if ((B) <= 0) break;                                                                            \
    const unsigned grid_ = tilesGrid(B);                                                            \
    const size_t lane_ = (LANE);                                                                    \
    if (lane_ <= kGLdsWave && allowLds((const void*)&KERNEL<S, true>, lane_))                       \
      hipLaunchKernelGGL((KERNEL<S, true>), dim3(grid_), dim3(kTile), lane_, st, __VA_ARGS__, (S*)nullptr); \
    else                                                                                            \
      hipLaunchKernelGGL((KERNEL<S, false>), dim3(grid_), dim3(kTile), 0, st, __VA_ARGS__,         \
                         (S*)g->laneMem(lane_ * grid_));                                            \
  } while (0)

template<class S>
void gRnea(hipStream_t st, GM g, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> a, View<S> out) {
  RBD_G_LAUNCH(g_rnea, S, rneaLane<S>(g), B, progOf(g), mc, B, q, v, a, CView<S>{}, out);
}
template<class S>
void gAba(hipStream_t st, GM g, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> tau, View<S> out) {
  RBD_G_LAUNCH(g_aba, S, abaLane<S>(g), B, progOf(g), mc, B, q, v, tau, CView<S>{}, out);
}
template<class S> void gCrba(hipStream_t st, GM g, const S* mc, int64_t B, CView<S> q, View<S> M, int upper) {
  RBD_G_LAUNCH(g_crba, S, crbaLane<S>(g), B, progOf(g), mc, B, q, M, upper);
}
template<class S>
void gRneaD(hipStream_t st, GM g, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> a, View<S> dq, View<S> dv,
            View<S> M, int mask, int /*waves*/) {
  RBD_G_LAUNCH(g_rnead, S, rneadLane<S>(g), B, progOf(g), mc, B, q, v, a, dq, dv, M, mask);
}
template<class S>
void gKin(hipStream_t st, GM g, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> a, View<S> oM, View<S> vo,
          View<S> ao, View<S> J) {
  RBD_G_LAUNCH(g_kin, S, kinLane<S>(g), B, progOf(g), mc, B, q, v, a, oM, vo, ao, J);
}

template<class SIN, class SC> void gCholT(hipStream_t st, GM g, int64_t B, CView<SIN> M, View<SC> U, View<SC> D) {
  if (B <= 0) return;
  hipLaunchKernelGGL((g_chol<SIN, SC>), dim3((unsigned)((B + kTile - 1) / kTile)), dim3(kTile), 0, st, progOf(g), B,
                     M, U, D);
}
template<class S> void gChol(hipStream_t st, GM g, int64_t B, CView<S> M, View<double> U, View<double> D) {
  gCholT<S, double>(st, g, B, M, U, D);
}
void gChol32(hipStream_t st, GM g, int64_t B, CView<float> M, View<float> U, View<float> D) {
  gCholT<float, float>(st, g, B, M, U, D);
}
template<class S> void gCholSolve(hipStream_t st, GM g, int64_t B, CView<double> U, CView<double> D, View<S> x) {
  if (B <= 0) return;
  const unsigned grid = tilesGrid(B);
  const size_t lane = (size_t)g->nv * kTile * sizeof(double);
  if (lane <= kGLdsWave && allowLds((const void*)&g_chol_solve<S, true>, lane))
    hipLaunchKernelGGL((g_chol_solve<S, true>), dim3(grid), dim3(kTile), lane, st, progOf(g), B, U, D, x,
                       (double*)nullptr);
  else
    hipLaunchKernelGGL((g_chol_solve<S, false>), dim3(grid), dim3(kTile), 0, st, progOf(g), B, U, D, x,
                       (double*)g->laneMem(lane * grid));
}
template<class S>
void gAbaDSolve(hipStream_t st, GM g, int64_t B, CView<S> U, CView<S> D, View<S> Xq, View<S> Xv, View<S> Minv,
                int jobBegin) {
  if (B <= 0) return;
  const unsigned grid = tilesGrid(B);
  const size_t xb = (size_t)kGSolveWaves * g->nv * kTile * sizeof(S);
  const size_t fb = (size_t)(g->nnz + g->nv) * kTile * sizeof(S);
  if (allowLds((const void*)&g_abad_solve<S, true>, xb + fb)) {
    hipLaunchKernelGGL((g_abad_solve<S, true>), dim3(grid), dim3(kTile * kGSolveWaves), xb + fb, st, progOf(g), B,
                       U, D, Xq, Xv, Minv, jobBegin);
  } else if (allowLds((const void*)&g_abad_solve<S, false>, xb)) {
    hipLaunchKernelGGL((g_abad_solve<S, false>), dim3(grid), dim3(kTile * kGSolveWaves), xb, st, progOf(g), B, U,
                       D, Xq, Xv, Minv, jobBegin);
  } else {
    throw Error(RBD_EUNSUPPORTED, "ABA-derivative solve: the model has too many dofs for the LDS");
  }
}

}  // namespace

void genericReserveDyn(const GenericModel* g, int64_t count) {
  if (count <= 0) return;
  const size_t grid = tilesGrid(count);
  size_t need = 0;
  if (rneaLane<float>(g) > kGLdsWave) need = std::max(need, rneaLane<float>(g) * grid);
  if (abaLane<float>(g) > kGLdsWave) need = std::max(need, abaLane<float>(g) * grid);
  if (need) (void)g->laneMem(need);
}

const TopoOps* genericOps() {
  static const TopoOps ops = [] {
    TopoOps o{};
    o.name = "generic";
    o.rnea_f32 = &gRnea<float>;
    o.rnea_f64 = &gRnea<double>;
    o.aba_f32 = &gAba<float>;
    o.aba_f64 = &gAba<double>;
    o.crba_f32 = &gCrba<float>;
    o.crba_f64 = &gCrba<double>;
    o.chol_f32 = &gChol<float>;
    o.chol_f64 = &gChol<double>;
    o.cholsolve_f32 = &gCholSolve<float>;
    o.cholsolve_f64 = &gCholSolve<double>;
    o.rnead_f32 = &gRneaD<float>;
    o.rnead_f64 = &gRneaD<double>;
    o.chol32_f32 = &gChol32;
    o.abad_solve_f32 = &gAbaDSolve<float>;
    o.abad_solve_f64 = &gAbaDSolve<double>;
    o.kin_f32 = &gKin<float>;
    o.kin_f64 = &gKin<double>;
    return o;
  }();
  return &ops;
}

// RNEA / ABA with external forces (run-time topology for every tree)
template<class S>
void launchRneaFext(hipStream_t st, const GenericModel* g, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> a,
                    CView<S> fext, View<S> out) {
  RBD_G_LAUNCH(g_rnea, S, rneaLane<S>(g), B, progOf(g), mc, B, q, v, a, fext, out);
}
template<class S>
void launchAbaFext(hipStream_t st, const GenericModel* g, const S* mc, int64_t B, CView<S> q, CView<S> v,
                   CView<S> tau, CView<S> fext, View<S> out) {
  RBD_G_LAUNCH(g_aba, S, abaLane<S>(g), B, progOf(g), mc, B, q, v, tau, fext, out);
}
template void launchRneaFext<float>(hipStream_t, const GenericModel*, const float*, int64_t, CView<float>, CView<float>,
                                    CView<float>, CView<float>, View<float>);
template void launchRneaFext<double>(hipStream_t, const GenericModel*, const double*, int64_t, CView<double>,
                                     CView<double>, CView<double>, CView<double>, View<double>);
template void launchAbaFext<float>(hipStream_t, const GenericModel*, const float*, int64_t, CView<float>, CView<float>,
                                   CView<float>, CView<float>, View<float>);
template void launchAbaFext<double>(hipStream_t, const GenericModel*, const double*, int64_t, CView<double>,
                                    CView<double>, CView<double>, CView<double>, View<double>);

}  // namespace rbd
