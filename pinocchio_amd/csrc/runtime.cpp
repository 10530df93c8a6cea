// runtime.cpp -- C ABI of the engine: contexts, model upload, argument checking, layout
// handling and kernel dispatch (include/rbd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <memory>
#include <type_traits>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rbd.h"
#include "k_dyn.hpp"
#include "k_frames.hpp"
#include "k_contact.hpp"
#include "k_kkt.hpp"
#include "k_ncp.hpp"
#include "ops.hpp"
#include "runtime.hpp"

namespace rbd {
#define RBD_REGISTER(NAME) const TopoOps* ops_##NAME();
#include "_gen/registry.inc"
#undef RBD_REGISTER
#define RBD_REGISTER_MIX(NAME) const MixOps* mix_##NAME();
#include "_gen/mixes.inc"
#undef RBD_REGISTER_MIX

static const TopoOps* const* registry(int* n) {
#define RBD_REGISTER(NAME) ops_##NAME(),
  static const TopoOps* const all[] = {
#include "_gen/registry.inc"
  };
#undef RBD_REGISTER
  *n = (int)(sizeof(all) / sizeof(all[0]));
  return all;
}
}  // namespace rbd

using namespace rbd;

namespace rbd_rt {
thread_local std::string g_err;
}  // namespace rbd_rt

// Scratch of the allocation-free entry point (rbd_aba_derivatives_noalloc): the tree
// factor of M and ddq of one call.  Every GPU the library runs on holds one copy until the
// process ends.  Host code never uses this name as an address: kernels get the per-device
// address taken at context creation (scratchPtr below).
constexpr size_t kScratchFloats = size_t(16) << 20;  // 64 MiB
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
float* g_scratchPtr[kMaxDevices] = {};
// Records the device address of `scratch` on `device` (the current device) once.
void initScratch(int device) {
  if (device >= kMaxDevices) return;  // the allocation-free entry point then reports an error
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (g_scratchPtr[device]) return;
  void* p = nullptr;
  HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratchPtr[device] = static_cast<float*>(p);
}
float* scratchPtr(int device) {
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  return device < kMaxDevices ? g_scratchPtr[device] : nullptr;
}
}  // namespace

// batched identity matrices: element (r + c nv, b) at p[(r + c nv) rs + b bs]
__global__ void k_identity(int nv, int64_t B, double* p, int64_t rs, int64_t bs) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t k = e / B, b = e - k * B;
  if (k >= (int64_t)nv * nv) return;
  p[k * rs + b * bs] = (k % nv == k / nv) ? 1.0 : 0.0;
}

namespace {

void setDevice(rbd_ctx* ctx) { HIP_CHECK(hipSetDevice(ctx->device)); }

std::atomic<uint64_t> g_modelSerial{0};

// releases the mixed-batch plans of a context (model == nullptr) or those using a model
void dropMultiPlans(rbd_ctx* ctx, const rbd_model* model);

// Host flattening of the model constants into per-joint records (topo.hpp MC_* offsets).
std::vector<double> buildConstants(const rbd_model_desc* d, int nv) {
  const int nj = d->njoints;
  std::vector<double> mc((size_t)nj * MC_REC + nv + 6, 0.0);
  for (int j = 0; j < nj; ++j) {
    double* r = &mc[(size_t)j * MC_REC];
    const double* P = d->placement + 12 * j;
    for (int k = 0; k < 12; ++k) r[MC_R + k] = P[k];
    const double* Y = d->inertia + 10 * j;
    const double m = Y[0], cx = Y[1], cy = Y[2], cz = Y[3];
    r[MC_M] = m;
    r[MC_C] = cx; r[MC_C + 1] = cy; r[MC_C + 2] = cz;
    for (int k = 0; k < 6; ++k) r[MC_I + k] = Y[4 + k];
    const double Ixx = Y[4], Ixy = Y[5], Iyy = Y[6], Ixz = Y[7], Iyz = Y[8], Izz = Y[9];
    // packed symmetric 6x6 of the body inertia, spatial.hpp symIdx order
    double S6[6][6] = {};
    for (int k = 0; k < 3; ++k) S6[k][k] = m;
    const double mc_[3] = {m * cx, m * cy, m * cz};
    // -m[c]x in block (0,3)
    S6[0][4] = mc_[2]; S6[0][5] = -mc_[1];
    S6[1][3] = -mc_[2]; S6[1][5] = mc_[0];
    S6[2][3] = mc_[1]; S6[2][4] = -mc_[0];
    const double Io[3][3] = {{Ixx + m * (cy * cy + cz * cz), Ixy - m * cx * cy, Ixz - m * cx * cz},
                             {Ixy - m * cx * cy, Iyy + m * (cx * cx + cz * cz), Iyz - m * cy * cz},
                             {Ixz - m * cx * cz, Iyz - m * cy * cz, Izz + m * (cx * cx + cy * cy)}};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) S6[3 + a][3 + b] = Io[a][b];
    int idx = 0;
    for (int a = 0; a < 6; ++a)
      for (int b = a; b < 6; ++b) r[MC_Y6 + idx++] = S6[a][b];
    r[MC_CI + 0] = m;
    r[MC_CI + 1] = mc_[0]; r[MC_CI + 2] = mc_[1]; r[MC_CI + 3] = mc_[2];
    r[MC_CI + 4] = Io[0][0]; r[MC_CI + 5] = Io[0][1]; r[MC_CI + 6] = Io[1][1];
    r[MC_CI + 7] = Io[0][2]; r[MC_CI + 8] = Io[1][2]; r[MC_CI + 9] = Io[2][2];
  }
  double* arm = &mc[(size_t)nj * MC_REC];
  for (int k = 0; k < nv; ++k) arm[k] = d->armature ? d->armature[k] : 0.0;
  double* g = arm + nv;
  const double g0[6] = {0, 0, -9.81, 0, 0, 0};
  for (int k = 0; k < 6; ++k) g[k] = d->gravity ? d->gravity[k] : g0[k];
  return mc;
}

// Validated strided view over a batched argument with `rows` rows.
template<class S> struct ArgView {
  int64_t rs, bs;
  size_t extent;  // elements spanned
};
template<class S> ArgView<S> argView(int rows, int64_t B, int64_t ld, uint32_t flags, const char* what) {
  ArgView<S> v{};
  if (flags & RBD_LAYOUT_BATCH) {
    if (ld < B) throw Error(RBD_EINVAL, std::string(what) + ": leading dimension smaller than the batch");
    v.rs = ld;
    v.bs = 1;
    v.extent = (size_t)(ld * (rows - 1) + B);
  } else {
    if (ld < rows) throw Error(RBD_EINVAL, std::string(what) + ": leading dimension smaller than the row count");
    v.rs = 1;
    v.bs = ld;
    v.extent = (size_t)(ld * (B - 1) + rows);
  }
  return v;
}

// Moves an argument to the device when the call uses host pointers.
struct Staged {
  rbd_ctx* ctx;
  bool host;
  int slot = 0;
  std::vector<std::pair<void*, std::pair<const void*, size_t>>> back;  // device, host, bytes
  const void* in(const void* p, size_t bytes) {
    if (!host || !p) return p;
    void* d = ctx->stageBuf(slot++, bytes);
    HIP_CHECK(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, ctx->stream));
    return d;
  }
  void* out(void* p, size_t bytes, bool preload) {
    if (!host || !p) return p;
    void* d = ctx->stageBuf(slot++, bytes);
    if (preload) HIP_CHECK(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, ctx->stream));
    back.push_back({d, {p, bytes}});
    return d;
  }
  void finish() {
    if (!host) return;
    for (auto& b : back)
      HIP_CHECK(hipMemcpyAsync(const_cast<void*>(b.second.first), b.first, b.second.second, hipMemcpyDeviceToHost,
                               ctx->stream));
  }
};

void finishCall(rbd_ctx* ctx, uint32_t flags, Staged& st) {
  HIP_CHECK(hipGetLastError());
  st.finish();
  if (ctx->capturing) return;  // recorded into the graph being captured, not run
  if (!(flags & RBD_ASYNC) || (flags & RBD_HOST_PTRS)) HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void checkCommon(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype) {
  RBD_CHECK_ARG(ctx != nullptr, "null context");
  RBD_CHECK_ARG(m != nullptr, "null model");
  RBD_CHECK_ARG(m->ctx == ctx, "model belongs to another context");
  RBD_CHECK_ARG(B >= 0, "negative batch size");
  RBD_CHECK_ARG(dtype == RBD_F32 || dtype == RBD_F64, "dtype must be RBD_F32 or RBD_F64");
}

// marks an input of vec3Op that is zero for every configuration
const void* const kZeroInput = reinterpret_cast<const void*>(&kZeroInput);

template<class S>
int vec3Op(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* x0, int64_t ld0, int r0, const void* x1,
           int64_t ld1, const void* x2, int64_t ld2, void* out, int64_t ldo, uint32_t flags, LaunchVec3<S> fn,
           const char* opname) {
  return guarded([&] {
    if (!fn) throw Error(RBD_EUNSUPPORTED, std::string(opname) + ": no kernel for this dtype/topology");
    // x1 / x2 == kZeroInput: the input is zero for every configuration (a stride-0 view of
    // the context's zero block; nothing is read from the caller)
    const bool z1 = x1 == kZeroInput, z2 = x2 == kZeroInput;
    RBD_CHECK_ARG(B == 0 || (x0 && x1 && x2 && out), "null array argument");
    const int nv = m->nv;
    auto v0 = argView<S>(r0, B, ld0, flags, "input 0");
    auto v1 = z1 ? ArgView<S>{0, 0, 1} : argView<S>(nv, B, ld1, flags, "input 1");
    auto v2 = z2 ? ArgView<S>{0, 0, 1} : argView<S>(nv, B, ld2, flags, "input 2");
    auto vo = argView<S>(nv, B, ldo, flags, "output");
    if (B == 0) return;
    setDevice(ctx);
    Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
    const S* d0 = (const S*)st.in(x0, v0.extent * sizeof(S));
    const S* d1 = z1 ? (const S*)ctx->zeros : (const S*)st.in(x1, v1.extent * sizeof(S));
    const S* d2 = z2 ? (const S*)ctx->zeros : (const S*)st.in(x2, v2.extent * sizeof(S));
    S* dout = (S*)st.out(out, vo.extent * sizeof(S), ldo != nv || (flags & RBD_LAYOUT_BATCH));
    const S* mc = std::is_same<S, float>::value ? (const S*)m->mc32 : (const S*)m->mc64;
    fn(ctx->stream, m->gen.get(), mc, B, CView<S>{d0, v0.rs, v0.bs}, CView<S>{d1, v1.rs, v1.bs}, CView<S>{d2, v2.rs, v2.bs},
       View<S>{dout, vo.rs, vo.bs});
    finishCall(ctx, flags, st);
  });
}

}  // namespace

// rnea / aba with external forces (run-time-topology kernels for every tree)
namespace {
template<class S>
void fextCall(rbd_ctx* ctx, const rbd_model* m, int alg, int64_t B, const void* q, int64_t ldq, const void* v,
              int64_t ldv, const void* x, int64_t ldx, const void* fext, int64_t ldf, void* out, int64_t ldo,
              uint32_t flags) {
  const int nv = m->nv;
  const bool zv = v == kZeroInput, zx = x == kZeroInput;  // zero for every configuration
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  auto vv = zv ? ArgView<S>{0, 0, 1} : argView<S>(nv, B, ldv, flags, "v");
  auto vx = zx ? ArgView<S>{0, 0, 1} : argView<S>(nv, B, ldx, flags, alg == 0 ? "a" : "tau");
  auto vf = argView<S>(6 * (m->njoints - 1), B, ldf, flags, "fext");
  auto vo = argView<S>(nv, B, ldo, flags, alg == 0 ? "tau" : "ddq");
  if (B == 0) return;
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const S* dq = (const S*)st.in(q, vq.extent * sizeof(S));
  const S* dv = zv ? (const S*)ctx->zeros : (const S*)st.in(v, vv.extent * sizeof(S));
  const S* dx = zx ? (const S*)ctx->zeros : (const S*)st.in(x, vx.extent * sizeof(S));
  const S* df = (const S*)st.in(fext, vf.extent * sizeof(S));
  S* dout = (S*)st.out(out, vo.extent * sizeof(S), ldo != nv || (flags & RBD_LAYOUT_BATCH));
  const S* mc = std::is_same<S, float>::value ? (const S*)m->mc32 : (const S*)m->mc64;
  const CView<S> cq{dq, vq.rs, vq.bs}, cv{dv, vv.rs, vv.bs}, cx{dx, vx.rs, vx.bs}, cf{df, vf.rs, vf.bs};
  if (alg == 0)
    launchRneaFext<S>(ctx->stream, m->gen.get(), mc, B, cq, cv, cx, cf, View<S>{dout, vo.rs, vo.bs});
  else
    launchAbaFext<S>(ctx->stream, m->gen.get(), mc, B, cq, cv, cx, cf, View<S>{dout, vo.rs, vo.bs});
  finishCall(ctx, flags, st);
}
}  // namespace

extern "C" {

const char* rbd_last_error(void) { return g_err.c_str(); }
const char* rbd_version(void) { return "pinocchio_amd 0.1 (gfx950)"; }

int rbd_ctx_create(int device, rbd_ctx** out) {
  return guarded([&] {
    RBD_CHECK_ARG(out != nullptr, "null output pointer");
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    RBD_CHECK_ARG(device >= 0 && device < n, "device index out of range");
    auto* c = new rbd_ctx();
    c->device = device;
    try {
      HIP_CHECK(hipSetDevice(device));
      // a blocking stream: work on it is ordered with the legacy default stream, which is
      // where PyTorch (and most callers) enqueue the producers of the inputs and the
      // consumers of the outputs
      HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamDefault));
      HIP_CHECK(hipMalloc(&c->zeros, 64));
      HIP_CHECK(hipMemset(c->zeros, 0, 64));
      initScratch(device);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int rbd_ctx_destroy(rbd_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    dropMultiPlans(ctx, nullptr);
    for (void* p : ctx->stage)
      if (p) (void)hipFree(p);
    if (ctx->work) (void)hipFree(ctx->work);
    if (ctx->zeros) (void)hipFree(ctx->zeros);
    for (void* p : ctx->retired) (void)hipFree(p);
    for (hipStream_t s : ctx->side) (void)hipStreamDestroy(s);
    for (hipEvent_t e : ctx->forkJoin) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

void* rbd_ctx_stream(rbd_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int rbd_ctx_synchronize(rbd_ctx* ctx) {
  return guarded([&] {
    RBD_CHECK_ARG(ctx, "null context");
    RBD_CHECK_ARG(!ctx->capturing, "capture in progress: nothing runs until rbd_graph_launch");
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

// ---- stream capture of engine calls --------------------------------------------------
struct rbd_graph {
  rbd_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int rbd_capture_begin(rbd_ctx* ctx) {
  return guarded([&] {
    RBD_CHECK_ARG(ctx, "null context");
    RBD_CHECK_ARG(!ctx->capturing, "a capture is already in progress on this context");
    HIP_CHECK(hipSetDevice(ctx->device));
    HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
  });
}

int rbd_capture_end(rbd_ctx* ctx, rbd_graph** out) {
  return guarded([&] {
    RBD_CHECK_ARG(ctx, "null context");
    RBD_CHECK_ARG(ctx->capturing, "no capture in progress on this context");
    ctx->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess || !g) {
      (void)hipGetLastError();
      if (g) (void)hipGraphDestroy(g);
      throw Error(RBD_EDEVICE, std::string("capture failed (a call inside it could not be recorded): ") +
                                   hipGetErrorString(e));
    }
    RBD_CHECK_ARG(out != nullptr, "null output pointer");
    auto* r = new rbd_graph();
    r->ctx = ctx;
    r->graph = g;
    const hipError_t ie = hipGraphInstantiate(&r->exec, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(g);
      delete r;
      throw Error(RBD_EDEVICE, std::string("graph instantiation failed: ") + hipGetErrorString(ie));
    }
    ++ctx->graphs;
    *out = r;
  });
}

int rbd_graph_launch(rbd_ctx* ctx, const rbd_graph* graph, uint32_t flags) {
  return guarded([&] {
    RBD_CHECK_ARG(ctx && graph, "null argument");
    RBD_CHECK_ARG(graph->ctx == ctx, "graph captured on another context");
    RBD_CHECK_ARG(!ctx->capturing, "capture in progress on this context");
    HIP_CHECK(hipSetDevice(ctx->device));
    HIP_CHECK(hipGraphLaunch(graph->exec, ctx->stream));
    if (!(flags & RBD_ASYNC)) HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int rbd_graph_destroy(rbd_graph* graph) {
  return guarded([&] {
    if (!graph) return;
    (void)hipSetDevice(graph->ctx->device);
    (void)hipStreamSynchronize(graph->ctx->stream);
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    --graph->ctx->graphs;
    delete graph;
  });
}


int rbd_rnea_fext(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq, const void* v,
                  int64_t ldv, const void* a, int64_t lda, const void* fext, int64_t ldf, void* tau, int64_t ldtau,
                  uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (q && v && a && fext && tau), "null array argument");
    if (dtype == RBD_F32) fextCall<float>(ctx, m, 0, B, q, ldq, v, ldv, a, lda, fext, ldf, tau, ldtau, flags);
    else fextCall<double>(ctx, m, 0, B, q, ldq, v, ldv, a, lda, fext, ldf, tau, ldtau, flags);
  });
}

int rbd_aba_fext(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq, const void* v,
                 int64_t ldv, const void* tau, int64_t ldtau, const void* fext, int64_t ldf, void* ddq, int64_t ldddq,
                 uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (q && v && tau && fext && ddq), "null array argument");
    if (dtype == RBD_F32) fextCall<float>(ctx, m, 1, B, q, ldq, v, ldv, tau, ldtau, fext, ldf, ddq, ldddq, flags);
    else fextCall<double>(ctx, m, 1, B, q, ldq, v, ldv, tau, ldtau, fext, ldf, ddq, ldddq, flags);
  });
}

int rbd_cholesky_apply(rbd_ctx* ctx, const rbd_model* m, int64_t B, int op, const double* U, int64_t ldU,
                       const double* D, int64_t ldD, double* x, int64_t ldx, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, RBD_F64);
    RBD_CHECK_ARG(op >= RBD_CHOL_UV && op <= RBD_CHOL_MV, "op must be RBD_CHOL_UV, _UTV, _UIV, _UTIV or _MV");
    RBD_CHECK_ARG(B == 0 || (U && x && (D || op != RBD_CHOL_MV)), "null array argument");
    if (!m->compact) throw Error(RBD_EUNSUPPORTED, "cholesky: the model is not DFS-compact");
    const int nv = m->nv, nnz = (int)m->patRows.size();
    auto vU = argView<double>(nnz > 0 ? nnz : 1, B, ldU, flags, "U");
    auto vD = argView<double>(nv, B, op == RBD_CHOL_MV ? ldD : (flags & RBD_LAYOUT_BATCH ? B : nv), flags, "D");
    auto vx = argView<double>(nv, B, ldx, flags, "x");
    if (B == 0) return;
    setDevice(ctx);
    Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
    const double* dU = (const double*)st.in(U, vU.extent * 8);
    const double* dD = op == RBD_CHOL_MV ? (const double*)st.in(D, vD.extent * 8) : nullptr;
    double* dx = (double*)st.out(x, vx.extent * 8, true);
    launchCholApply(ctx->stream, m->gen.get(), op, B, CView<double>{dU, vU.rs, vU.bs}, CView<double>{dD, vD.rs, vD.bs},
                    View<double>{dx, vx.rs, vx.bs});
    finishCall(ctx, flags, st);
  });
}

}  // extern "C"

namespace {
// entries of `t` that are not set take the run-time-topology launcher
void fillMissing(TopoOps& t, const TopoOps& g) {
#define RBD_FILL_OP(F) \
  if (!t.F) t.F = g.F;
  RBD_FILL_OP(rnea_f32) RBD_FILL_OP(rnea_f64) RBD_FILL_OP(aba_f32) RBD_FILL_OP(aba_f64)
  RBD_FILL_OP(crba_f32) RBD_FILL_OP(crba_f64) RBD_FILL_OP(chol_f32) RBD_FILL_OP(chol_f64)
  RBD_FILL_OP(cholsolve_f32) RBD_FILL_OP(cholsolve_f64) RBD_FILL_OP(rnead_f32) RBD_FILL_OP(rnead_f64)
  RBD_FILL_OP(chol32_f32) RBD_FILL_OP(abad_solve_f32) RBD_FILL_OP(abad_solve_f64) RBD_FILL_OP(kin_f32)
  RBD_FILL_OP(kin_f64) RBD_FILL_OP(kkt_f32) RBD_FILL_OP(kkt_f64)
#undef RBD_FILL_OP
}

// DFS pre-order with children in increasing index order is the index order
// (crba.hxx:573-595)
bool isCompact(const std::vector<int>& parents) {
  const int nj = (int)parents.size();
  std::vector<std::vector<int>> kids(nj);
  for (int j = 1; j < nj; ++j) kids[parents[j]].push_back(j);
  std::vector<int> stack, order;
  for (int k = (int)kids[0].size() - 1; k >= 0; --k) stack.push_back(kids[0][k]);
  while (!stack.empty()) {
    const int j = stack.back();
    stack.pop_back();
    order.push_back(j);
    for (int k = (int)kids[j].size() - 1; k >= 0; --k) stack.push_back(kids[j][k]);
  }
  for (int k = 0; k < (int)order.size(); ++k)
    if (order[k] != k + 1) return false;
  return true;
}

int createModel(rbd_ctx* ctx, const rbd_model_desc* d, uint32_t mflags, rbd_model** out) {
  return guarded([&] {
    RBD_CHECK_ARG(ctx && d && out, "null argument");
    RBD_CHECK_ARG(!ctx->capturing, "capture in progress: create models before rbd_capture_begin");
    RBD_CHECK_ARG(d->njoints >= 2, "model needs at least one moving joint");
    RBD_CHECK_ARG(d->parents && d->jtype && d->placement && d->inertia, "null model array");
    const int nj = d->njoints;
    auto m = std::make_unique<rbd_model>();
    m->ctx = ctx;
    m->serial = ++g_modelSerial;
    m->njoints = nj;
    m->parents.assign(d->parents, d->parents + nj);
    m->jtype.assign(d->jtype, d->jtype + nj);
    m->iq.assign(nj, 0);
    m->iv.assign(nj, 0);
    for (int j = 1; j < nj; ++j) {
      RBD_CHECK_ARG(m->parents[j] >= 0 && m->parents[j] < j, "parents[i] must precede i");
      const int t = m->jtype[j];
      if (t != RBD_JOINT_FREEFLYER && t != RBD_JOINT_RX && t != RBD_JOINT_RY && t != RBD_JOINT_RZ)
        throw Error(RBD_EUNSUPPORTED, "joint type not supported (FreeFlyer, RX, RY, RZ only)");
      if (t == RBD_JOINT_FREEFLYER && m->parents[j] != 0)
        throw Error(RBD_EUNSUPPORTED, "a free-flyer joint must be attached to the universe");
      m->iq[j] = m->nq;
      m->iv[j] = m->nv;
      m->nq += t == RBD_JOINT_FREEFLYER ? 7 : 1;
      m->nv += t == RBD_JOINT_FREEFLYER ? 6 : 1;
    }
    for (int j = 1; j < nj; ++j) {
      const double* Y = d->inertia + 10 * j;
      RBD_CHECK_ARG(std::isfinite(Y[0]) && Y[0] >= 0.0, "body mass must be finite and non-negative");
    }
    if (!(mflags & RBD_MODEL_RUNTIME_TOPOLOGY)) {
      int n = 0;
      const TopoOps* const* all = registry(&n);
      for (int k = 0; k < n && !m->prebuilt; ++k) {
        const TopoOps* o = all[k];
        if (o->njoints != nj) continue;
        if (std::memcmp(o->parent, m->parents.data(), sizeof(int) * nj) != 0) continue;
        bool same = true;
        for (int j = 1; j < nj; ++j) same = same && o->type[j] == m->jtype[j];
        if (same) m->prebuilt = o;
      }
    }
    m->compact = isCompact(m->parents) ? 1 : 0;
    m->gravity.assign(6, 0.0);
    const double g0[6] = {0, 0, -9.81, 0, 0, 0};
    for (int k = 0; k < 6; ++k) m->gravity[k] = d->gravity ? d->gravity[k] : g0[k];
    // structural pattern of the upper triangle: (r, c) with r an ancestor dof of c
    {
      std::vector<int> dofParent(m->nv, -1);
      for (int j = 1; j < nj; ++j) {
        const int p = m->parents[j];
        const int nvj = m->jtype[j] == RBD_JOINT_FREEFLYER ? 6 : 1;
        dofParent[m->iv[j]] = p > 0 ? m->iv[p] + (m->jtype[p] == RBD_JOINT_FREEFLYER ? 6 : 1) - 1 : -1;
        for (int r = 1; r < nvj; ++r) dofParent[m->iv[j] + r] = m->iv[j] + r - 1;
      }
      for (int c = 0; c < m->nv; ++c) {
        std::vector<int> anc;
        for (int r = dofParent[c]; r >= 0; r = dofParent[r]) anc.push_back(r);
        for (auto it = anc.rbegin(); it != anc.rend(); ++it) {
          m->patRows.push_back(*it);
          m->patCols.push_back(c);
        }
      }
    }
    const std::vector<double> mc = buildConstants(d, m->nv);
    std::vector<float> mcf(mc.begin(), mc.end());
    setDevice(ctx);
    m->gen.reset(buildGenericModel(nj, m->parents.data(), m->jtype.data()));
    m->tab = m->prebuilt ? *m->prebuilt : TopoOps{};
    fillMissing(m->tab, *genericOps());
    m->tab.name = m->prebuilt ? m->prebuilt->name : "generic";
    m->tab.njoints = nj;
    m->tab.parent = m->parents.data();
    m->tab.type = m->jtype.data();
    m->tab.nq = m->nq;
    m->tab.nv = m->nv;
    m->tab.compact = m->compact;
    m->tab.nnzU = (int)m->patRows.size();
    m->ops = &m->tab;
    HIP_CHECK(hipMalloc(&m->mc64, mc.size() * sizeof(double)));
    HIP_CHECK(hipMalloc(&m->mc32, mcf.size() * sizeof(float)));
    HIP_CHECK(hipMemcpy(m->mc64, mc.data(), mc.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->mc32, mcf.data(), mcf.size() * sizeof(float), hipMemcpyHostToDevice));
    *out = m.release();
  });
}

}  // namespace

extern "C" {

int rbd_model_create(rbd_ctx* ctx, const rbd_model_desc* d, rbd_model** out) { return createModel(ctx, d, 0, out); }

int rbd_model_create_ex(rbd_ctx* ctx, const rbd_model_desc* d, uint32_t model_flags, rbd_model** out) {
  return createModel(ctx, d, model_flags, out);
}

int rbd_model_kernels(const rbd_model* m, int32_t* prebuilt) {
  return guarded([&] {
    RBD_CHECK_ARG(m && prebuilt, "null argument");
    *prebuilt = m->prebuilt ? 1 : 0;
  });
}

int rbd_model_destroy(rbd_model* m) {
  return guarded([&] {
    if (!m) return;
    RBD_CHECK_ARG(!m->ctx->capturing, "capture in progress: destroy the model after rbd_capture_end");
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);  // kernels in flight may still read the constants
    dropMultiPlans(m->ctx, m);
    if (m->mc32) (void)hipFree(m->mc32);
    if (m->mc64) (void)hipFree(m->mc64);
    delete m;
  });
}

int rbd_model_dims(const rbd_model* m, int32_t* nq, int32_t* nv, int32_t* nj, int32_t* compact) {
  return guarded([&] {
    RBD_CHECK_ARG(m, "null model");
    if (nq) *nq = m->nq;
    if (nv) *nv = m->nv;
    if (nj) *nj = m->njoints;
    if (compact) *compact = m->compact;
  });
}

int rbd_rnea(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq, const void* v,
             int64_t ldv, const void* a, int64_t lda, void* tau, int64_t ldtau, uint32_t flags) {
  int rc = guarded([&] { checkCommon(ctx, m, B, dtype); });
  if (rc) return rc;
  if (dtype == RBD_F32)
    return vec3Op<float>(ctx, m, B, q, ldq, m->nq, v, ldv, a, lda, tau, ldtau, flags, m->ops->rnea_f32, "rnea");
  return vec3Op<double>(ctx, m, B, q, ldq, m->nq, v, ldv, a, lda, tau, ldtau, flags, m->ops->rnea_f64, "rnea");
}

int rbd_nonlinear_effects(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                          const void* v, int64_t ldv, void* nle, int64_t ldnle, uint32_t flags) {
  int rc = guarded([&] { checkCommon(ctx, m, B, dtype); });
  if (rc) return rc;
  if (dtype == RBD_F32)
    return vec3Op<float>(ctx, m, B, q, ldq, m->nq, v, ldv, kZeroInput, 0, nle, ldnle, flags, m->ops->rnea_f32,
                         "nonLinearEffects");
  return vec3Op<double>(ctx, m, B, q, ldq, m->nq, v, ldv, kZeroInput, 0, nle, ldnle, flags, m->ops->rnea_f64,
                        "nonLinearEffects");
}

int rbd_generalized_gravity(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                            void* g, int64_t ldg, uint32_t flags) {
  int rc = guarded([&] { checkCommon(ctx, m, B, dtype); });
  if (rc) return rc;
  if (dtype == RBD_F32)
    return vec3Op<float>(ctx, m, B, q, ldq, m->nq, kZeroInput, 0, kZeroInput, 0, g, ldg, flags, m->ops->rnea_f32,
                         "computeGeneralizedGravity");
  return vec3Op<double>(ctx, m, B, q, ldq, m->nq, kZeroInput, 0, kZeroInput, 0, g, ldg, flags, m->ops->rnea_f64,
                        "computeGeneralizedGravity");
}

int rbd_static_torque(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                      const void* fext, int64_t ldfext, void* tau, int64_t ldtau, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (q && fext && tau), "null array argument");
    if (dtype == RBD_F32)
      fextCall<float>(ctx, m, 0, B, q, ldq, kZeroInput, 0, kZeroInput, 0, fext, ldfext, tau, ldtau, flags);
    else
      fextCall<double>(ctx, m, 0, B, q, ldq, kZeroInput, 0, kZeroInput, 0, fext, ldfext, tau, ldtau, flags);
  });
}

int rbd_aba(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq, const void* v,
            int64_t ldv, const void* tau, int64_t ldtau, void* ddq, int64_t ldddq, uint32_t flags) {
  int rc = guarded([&] { checkCommon(ctx, m, B, dtype); });
  if (rc) return rc;
  if (dtype == RBD_F32)
    return vec3Op<float>(ctx, m, B, q, ldq, m->nq, v, ldv, tau, ldtau, ddq, ldddq, flags, m->ops->aba_f32, "aba");
  return vec3Op<double>(ctx, m, B, q, ldq, m->nq, v, ldv, tau, ldtau, ddq, ldddq, flags, m->ops->aba_f64, "aba");
}

int rbd_crba(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq, void* M,
             int64_t ldM, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (q && M), "null array argument");
    if (!m->compact) throw Error(RBD_EUNSUPPORTED, "crba: the model is not DFS-compact (crba.hxx:573-595)");
    const int nv = m->nv;
    const bool f32 = dtype == RBD_F32;
    const size_t es = f32 ? 4 : 8;
    if ((f32 && !m->ops->crba_f32) || (!f32 && !m->ops->crba_f64))
      throw Error(RBD_EUNSUPPORTED, "crba: no kernel for this dtype/topology");
    auto vq = argView<float>(m->nq, B, ldq, flags, "q");
    auto vM = argView<float>(nv * nv, B, ldM, flags, "M");
    if (B == 0) return;
    setDevice(ctx);
    Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
    const void* dq = st.in(q, vq.extent * es);
    void* dM = st.out(M, vM.extent * es, true);
    const int upper = (flags & RBD_UPPER_ONLY) ? 1 : 0;
    if (f32)
      m->ops->crba_f32(ctx->stream, m->gen.get(), m->mc32, B, CView<float>{(const float*)dq, vq.rs, vq.bs},
                       View<float>{(float*)dM, vM.rs, vM.bs}, upper);
    else
      m->ops->crba_f64(ctx->stream, m->gen.get(), m->mc64, B, CView<double>{(const double*)dq, vq.rs, vq.bs},
                       View<double>{(double*)dM, vM.rs, vM.bs}, upper);
    finishCall(ctx, flags, st);
  });
}

}  // extern "C"

namespace {
// forward kinematics / joint Jacobians through one launcher (k_kin)
template<class S>
void kinCall(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* q, int64_t ldq, const void* v, int64_t ldv,
             const void* a, int64_t lda, void* oM, int64_t ldo, void* vo, void* ao, int64_t ldva, void* J,
             int64_t ldJ, uint32_t flags) {
  constexpr bool f32 = std::is_same_v<S, float>;
  const auto launch = f32 ? (void*)m->ops->kin_f32 : (void*)m->ops->kin_f64;
  if (!launch) throw Error(RBD_EUNSUPPORTED, "kinematics: no kernel for this topology");
  const int nj = m->njoints, nv = m->nv;
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  ArgView<S> vv{}, va{}, vM{}, vva{}, vJ{};
  if (v) vv = argView<S>(nv, B, ldv, flags, "v");
  if (a) va = argView<S>(nv, B, lda, flags, "a");
  if (oM) vM = argView<S>(12 * (nj - 1), B, ldo, flags, "oMi");
  if (vo || ao) vva = argView<S>(6 * (nj - 1), B, ldva, flags, "v/a out");
  if (J) vJ = argView<S>(6 * nv, B, ldJ, flags, "J");
  if (B == 0) return;
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const S* dq = (const S*)st.in(q, vq.extent * sizeof(S));
  const S* dv = (const S*)st.in(v, vv.extent * sizeof(S));
  const S* da = (const S*)st.in(a, va.extent * sizeof(S));
  // host outputs with gaps between columns (ld > rows, or the batch-fastest layout) are
  // preloaded so the copy back leaves the gaps as the caller had them
  const bool bl = (flags & RBD_LAYOUT_BATCH) != 0;
  S* dM = (S*)st.out(oM, vM.extent * sizeof(S), bl || ldo != 12 * (nj - 1));
  S* dvo = (S*)st.out(vo, vva.extent * sizeof(S), bl || ldva != 6 * (nj - 1));
  S* dao = (S*)st.out(ao, vva.extent * sizeof(S), bl || ldva != 6 * (nj - 1));
  S* dJ = (S*)st.out(J, vJ.extent * sizeof(S), bl || ldJ != 6 * nv);
  const S* mc;
  if constexpr (f32) mc = m->mc32;
  else mc = m->mc64;
  reinterpret_cast<LaunchKin<S>>(launch)(ctx->stream, m->gen.get(), mc, B, CView<S>{dq, vq.rs, vq.bs}, CView<S>{dv, vv.rs, vv.bs},
                                        CView<S>{da, va.rs, va.bs}, View<S>{dM, vM.rs, vM.bs},
                                        View<S>{dvo, vva.rs, vva.bs}, View<S>{dao, vva.rs, vva.bs},
                                        View<S>{dJ, vJ.rs, vJ.bs});
  finishCall(ctx, flags, st);
}
}  // namespace

extern "C" {

int rbd_forward_kinematics(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                           const void* v, int64_t ldv, const void* a, int64_t lda, void* oMi, int64_t ldo, void* vout,
                           void* aout, int64_t ldva, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (q && oMi), "null array argument");
    RBD_CHECK_ARG(!aout || (v && a && vout), "accelerations need v, a and the velocity output");
    RBD_CHECK_ARG(!vout || v, "velocities need v");
    if (dtype == RBD_F32) kinCall<float>(ctx, m, B, q, ldq, v, ldv, a, lda, oMi, ldo, vout, aout, ldva, nullptr, 0, flags);
    else kinCall<double>(ctx, m, B, q, ldq, v, ldv, a, lda, oMi, ldo, vout, aout, ldva, nullptr, 0, flags);
  });
}

int rbd_joint_jacobians(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq, void* J,
                        int64_t ldJ, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (q && J), "null array argument");
    if (dtype == RBD_F32)
      kinCall<float>(ctx, m, B, q, ldq, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, nullptr, 0, J, ldJ, flags);
    else
      kinCall<double>(ctx, m, B, q, ldq, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, nullptr, 0, J, ldJ, flags);
  });
}

int rbd_cholesky_decompose(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* M, int64_t ldM,
                           double* U, int64_t ldU, double* D, int64_t ldD, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (M && U && D), "null array argument");
    if (!m->compact) throw Error(RBD_EUNSUPPORTED, "cholesky: the model is not DFS-compact");
    const int nv = m->nv, nnz = (int)m->patRows.size();
    const bool f32 = dtype == RBD_F32;
    const size_t es = f32 ? 4 : 8;
    auto vM = argView<float>(nv * nv, B, ldM, flags, "M");
    auto vU = argView<double>(nnz > 0 ? nnz : 1, B, ldU, flags, "U");
    auto vD = argView<double>(nv, B, ldD, flags, "D");
    if (B == 0) return;
    setDevice(ctx);
    Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
    const void* dM = st.in(M, vM.extent * es);
    double* dU = (double*)st.out(U, vU.extent * 8, true);
    double* dD = (double*)st.out(D, vD.extent * 8, true);
    if (f32)
      m->ops->chol_f32(ctx->stream, m->gen.get(), B, CView<float>{(const float*)dM, vM.rs, vM.bs}, View<double>{dU, vU.rs, vU.bs},
                       View<double>{dD, vD.rs, vD.bs});
    else
      m->ops->chol_f64(ctx->stream, m->gen.get(), B, CView<double>{(const double*)dM, vM.rs, vM.bs},
                       View<double>{dU, vU.rs, vU.bs}, View<double>{dD, vD.rs, vD.bs});
    finishCall(ctx, flags, st);
  });
}

int rbd_cholesky_solve(rbd_ctx* ctx, const rbd_model* m, int64_t B, const double* U, int64_t ldU, const double* D,
                       int64_t ldD, int dtype, void* x, int64_t ldx, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (U && D && x), "null array argument");
    if (!m->compact) throw Error(RBD_EUNSUPPORTED, "cholesky: the model is not DFS-compact");
    const int nv = m->nv, nnz = (int)m->patRows.size();
    const bool f32 = dtype == RBD_F32;
    const size_t es = f32 ? 4 : 8;
    auto vU = argView<double>(nnz > 0 ? nnz : 1, B, ldU, flags, "U");
    auto vD = argView<double>(nv, B, ldD, flags, "D");
    auto vx = argView<float>(nv, B, ldx, flags, "x");
    if (B == 0) return;
    setDevice(ctx);
    Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
    const double* dU = (const double*)st.in(U, vU.extent * 8);
    const double* dD = (const double*)st.in(D, vD.extent * 8);
    void* dx = st.out(x, vx.extent * es, true);
    if (f32)
      m->ops->cholsolve_f32(ctx->stream, m->gen.get(), B, CView<double>{dU, vU.rs, vU.bs}, CView<double>{dD, vD.rs, vD.bs},
                            View<float>{(float*)dx, vx.rs, vx.bs});
    else
      m->ops->cholsolve_f64(ctx->stream, m->gen.get(), B, CView<double>{dU, vU.rs, vU.bs}, CView<double>{dD, vD.rs, vD.bs},
                            View<double>{(double*)dx, vx.rs, vx.bs});
    finishCall(ctx, flags, st);
  });
}

}  // extern "C"

namespace {
void checkDerivativeModel(const rbd_model* m) {
  if (!m->compact) throw Error(RBD_EUNSUPPORTED, "derivatives: the model is not DFS-compact");
  if (m->gravity[3] != 0.0 || m->gravity[4] != 0.0 || m->gravity[5] != 0.0)
    throw Error(RBD_EINVAL, "The gravity must be a pure force vector, no angular part");
}

template<class S> const S* mcOf(const rbd_model* m) {
  if constexpr (std::is_same_v<S, float>) return m->mc32;
  else return m->mc64;
}
template<class S> LaunchRneaD<S> rneadOf(const rbd_model* m) {
  if constexpr (std::is_same_v<S, float>) return m->ops->rnead_f32;
  else return m->ops->rnead_f64;
}

template<class S>
void rneaDerivCall(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* q, int64_t ldq, const void* v, int64_t ldv,
                   const void* a, int64_t lda, void* dq, void* dv, void* da, int64_t ldout, uint32_t flags) {
  const int nv = m->nv;
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  auto vv = argView<S>(nv, B, ldv, flags, "v");
  auto va = argView<S>(nv, B, lda, flags, "a");
  auto vo = argView<S>(nv * nv, B, ldout, flags, "outputs");
  if (B == 0) return;
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const S* dq_ = (const S*)st.in(q, vq.extent * sizeof(S));
  const S* dv_ = (const S*)st.in(v, vv.extent * sizeof(S));
  const S* da_ = (const S*)st.in(a, va.extent * sizeof(S));
  S* o1 = (S*)st.out(dq, vo.extent * sizeof(S), true);
  S* o2 = (S*)st.out(dv, vo.extent * sizeof(S), true);
  S* o3 = (S*)st.out(da, vo.extent * sizeof(S), true);
  const int mask = (dq ? 1 : 0) | (dv ? 2 : 0) | (da ? 4 : 0);
  rneadOf<S>(m)(ctx->stream, m->gen.get(), mcOf<S>(m), B, CView<S>{dq_, vq.rs, vq.bs}, CView<S>{dv_, vv.rs, vv.bs},
                CView<S>{da_, va.rs, va.bs}, View<S>{o1, vo.rs, vo.bs}, View<S>{o2, vo.rs, vo.bs},
                View<S>{o3, vo.rs, vo.bs}, mask, 0);
  finishCall(ctx, flags, st);
}

// ABA (ddq), RNEA derivatives at a = ddq (dtau_dq, dtau_dv, M into the outputs), the tree
// factor of M, then Minv and -Minv dtau in place.  FP32: the prebuilt kernels where the
// topology has them; FP64: the run-time-topology kernels.
// fixedWs: a workspace of fixedWsElems scalars to use instead of the context's (which
// grows by allocation); the call then allocates nothing
template<class S>
void abaDerivCall(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* q, int64_t ldq, const void* v, int64_t ldv,
                  const void* tau, int64_t ldtau, void* ddq_dq, void* ddq_dv, void* ddq_dtau, int64_t ldout, void* ddq,
                  int64_t ldddq, uint32_t flags, S* fixedWs = nullptr, size_t fixedWsElems = 0) {
  constexpr bool f32 = std::is_same_v<S, float>;
  const int nv = m->nv, nnz = m->ops->nnzU;
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  auto vv = argView<S>(nv, B, ldv, flags, "v");
  auto vt = argView<S>(nv, B, ldtau, flags, "tau");
  auto vo = argView<S>(nv * nv, B, ldout, flags, "outputs");
  ArgView<S> va{};
  if (ddq) va = argView<S>(nv, B, ldddq, flags, "ddq");
  if (B == 0) return;
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const S* dq_ = (const S*)st.in(q, vq.extent * sizeof(S));
  const S* dv_ = (const S*)st.in(v, vv.extent * sizeof(S));
  const S* dt_ = (const S*)st.in(tau, vt.extent * sizeof(S));
  S* Xq = (S*)st.out(ddq_dq, vo.extent * sizeof(S), true);
  S* Xv = (S*)st.out(ddq_dv, vo.extent * sizeof(S), true);
  S* Mi = (S*)st.out(ddq_dtau, vo.extent * sizeof(S), true);
  // workspace: U and D of the factor, then ddq (batch-fastest nv x B) unless the caller wants it
  const size_t nU = (size_t)(nnz > 0 ? nnz : 1) * B, nD = (size_t)nv * B;
  const size_t need = nU + nD + (size_t)nv * B;
  if (fixedWs && need > fixedWsElems)
    throw Error(RBD_EINVAL, "aba_derivatives_noalloc: the batch needs " + std::to_string(need * sizeof(S)) +
                                " bytes of scratch, more than the " + std::to_string(fixedWsElems * sizeof(S)) +
                                " available (at most " + std::to_string(fixedWsElems / (nnz + 2 * nv)) +
                                " configurations of this model)");
  S* wU = fixedWs ? fixedWs : (S*)ctx->workspace(need * sizeof(S));
  S* wD = wU + nU;
  S* wA = wD + nD;
  View<S> vddq{wA, B, 1};
  if (ddq) {
    S* dd = (S*)st.out(ddq, va.extent * sizeof(S), true);
    vddq = View<S>{dd, va.rs, va.bs};
  }
  const TopoOps& o = *m->ops;
  const CView<S> cq{dq_, vq.rs, vq.bs}, cv{dv_, vv.rs, vv.bs};
  // 1. ddq = aba(q, v, tau)
  if constexpr (f32) o.aba_f32(ctx->stream, m->gen.get(), m->mc32, B, cq, cv, CView<S>{dt_, vt.rs, vt.bs}, vddq);
  else o.aba_f64(ctx->stream, m->gen.get(), m->mc64, B, cq, cv, CView<S>{dt_, vt.rs, vt.bs}, vddq);
  // 2. dtau_dq, dtau_dv, M at a = ddq, into the output buffers
  rneadOf<S>(m)(ctx->stream, m->gen.get(), mcOf<S>(m), B, cq, cv, CView<S>{vddq.p, vddq.rs, vddq.bs},
                View<S>{Xq, vo.rs, vo.bs}, View<S>{Xv, vo.rs, vo.bs}, View<S>{Mi, vo.rs, vo.bs}, 7, 0);
  // 3. tree factor of M, then Minv and -Minv dtau in place
  if constexpr (f32) {
    o.chol32_f32(ctx->stream, m->gen.get(), B, CView<float>{Mi, vo.rs, vo.bs}, View<float>{wU, B, 1},
                 View<float>{wD, B, 1});
    o.abad_solve_f32(ctx->stream, m->gen.get(), B, CView<float>{wU, B, 1}, CView<float>{wD, B, 1},
                     View<float>{Xq, vo.rs, vo.bs}, View<float>{Xv, vo.rs, vo.bs}, View<float>{Mi, vo.rs, vo.bs}, 0);
  } else {
    o.chol_f64(ctx->stream, m->gen.get(), B, CView<double>{Mi, vo.rs, vo.bs}, View<double>{wU, B, 1},
               View<double>{wD, B, 1});
    o.abad_solve_f64(ctx->stream, m->gen.get(), B, CView<double>{wU, B, 1}, CView<double>{wD, B, 1},
                     View<double>{Xq, vo.rs, vo.bs}, View<double>{Xv, vo.rs, vo.bs}, View<double>{Mi, vo.rs, vo.bs},
                     0);
  }
  finishCall(ctx, flags, st);
}
}  // namespace

extern "C" {

int rbd_rnea_derivatives(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                         const void* v, int64_t ldv, const void* a, int64_t lda, void* dq, void* dv, void* da,
                         int64_t ldout, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (q && v && a), "null input argument");
    checkDerivativeModel(m);
    if (dtype == RBD_F32) rneaDerivCall<float>(ctx, m, B, q, ldq, v, ldv, a, lda, dq, dv, da, ldout, flags);
    else rneaDerivCall<double>(ctx, m, B, q, ldq, v, ldv, a, lda, dq, dv, da, ldout, flags);
  });
}

int rbd_aba_derivatives(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                        const void* v, int64_t ldv, const void* tau, int64_t ldtau, void* ddq_dq, void* ddq_dv,
                        void* ddq_dtau, int64_t ldout, void* ddq, int64_t ldddq, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (q && v && tau && ddq_dq && ddq_dv && ddq_dtau), "null array argument");
    checkDerivativeModel(m);
    if (dtype == RBD_F32)
      abaDerivCall<float>(ctx, m, B, q, ldq, v, ldv, tau, ldtau, ddq_dq, ddq_dv, ddq_dtau, ldout, ddq, ldddq, flags);
    else
      abaDerivCall<double>(ctx, m, B, q, ldq, v, ldv, tau, ldtau, ddq_dq, ddq_dv, ddq_dtau, ldout, ddq, ldddq, flags);
  });
}

int rbd_aba_derivatives_noalloc(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                                const void* v, int64_t ldv, const void* tau, int64_t ldtau, void* ddq_dq,
                                void* ddq_dv, void* ddq_dtau, int64_t ldout, void* ddq, int64_t ldddq,
                                uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (q && v && tau && ddq_dq && ddq_dv && ddq_dtau), "null array argument");
    checkDerivativeModel(m);
    if (dtype != RBD_F32) throw Error(RBD_EUNSUPPORTED, "aba_derivatives_noalloc: FP32 only");
    if (flags & RBD_HOST_PTRS)
      throw Error(RBD_EINVAL, "aba_derivatives_noalloc: device pointers only (host staging allocates)");
    // the prebuilt kernels keep their state on chip; the run-time-topology ones would need
    // lane memory
    const TopoOps* pre = m->prebuilt;
    if (!pre || !pre->aba_f32 || !pre->rnead_f32 || !pre->chol32_f32 || !pre->abad_solve_f32)
      throw Error(RBD_EUNSUPPORTED, "aba_derivatives_noalloc: needs a model with prebuilt kernels");
    float* ws = scratchPtr(ctx->device);
    if (!ws) throw Error(RBD_EDEVICE, "aba_derivatives_noalloc: no scratch address for this device");
    abaDerivCall<float>(ctx, m, B, q, ldq, v, ldv, tau, ldtau, ddq_dq, ddq_dv, ddq_dtau, ldout, ddq, ldddq, flags, ws,
                        kScratchFloats);
  });
}

int rbd_compute_minverse(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                         void* Minv, int64_t ldM, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (q && Minv), "null array argument");
    if (!m->compact || !m->ops->crba_f32 || !m->ops->abad_solve_f32)
      throw Error(RBD_EUNSUPPORTED, "computeMinverse: no kernels for this topology");
    const int nv = m->nv, nnz = m->ops->nnzU;
    if (dtype == RBD_F32) {
      auto vq = argView<float>(m->nq, B, ldq, flags, "q");
      auto vo = argView<float>(nv * nv, B, ldM, flags, "Minv");
      if (B == 0) return;
      setDevice(ctx);
      Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
      const float* dq_ = (const float*)st.in(q, vq.extent * 4);
      // host outputs with gaps (padded leading dimension, batch-fastest layout) are
      // preloaded so the copy back leaves the gaps as the caller had them
      float* Mi = (float*)st.out(Minv, vo.extent * 4, (flags & RBD_LAYOUT_BATCH) || ldM != (int64_t)nv * nv);
      // M into the output buffer, its FP32 tree factor into the workspace, then the unit
      // right-hand sides of the solve overwrite M with its inverse
      const size_t nU = (size_t)(nnz > 0 ? nnz : 1) * B, nD = (size_t)nv * B;
      float* wU = (float*)ctx->workspace((nU + nD) * 4);
      float* wD = wU + nU;
      const View<float> vM{Mi, vo.rs, vo.bs};
      m->ops->crba_f32(ctx->stream, m->gen.get(), m->mc32, B, CView<float>{dq_, vq.rs, vq.bs}, vM, 0);
      m->ops->chol32_f32(ctx->stream, m->gen.get(), B, CView<float>{Mi, vo.rs, vo.bs}, View<float>{wU, B, 1}, View<float>{wD, B, 1});
      m->ops->abad_solve_f32(ctx->stream, m->gen.get(), B, CView<float>{wU, B, 1}, CView<float>{wD, B, 1}, vM, vM, vM, 2 * nv);
      finishCall(ctx, flags, st);
      return;
    }
    // FP64: FP64 CRBA and tree factor, then one solve per unit column
    auto vq = argView<double>(m->nq, B, ldq, flags, "q");
    auto vo = argView<double>(nv * nv, B, ldM, flags, "Minv");
    if (B == 0) return;
    setDevice(ctx);
    Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
    const double* dq_ = (const double*)st.in(q, vq.extent * 8);
    double* Mi = (double*)st.out(Minv, vo.extent * 8, (flags & RBD_LAYOUT_BATCH) || ldM != (int64_t)nv * nv);
    const size_t nU = (size_t)(nnz > 0 ? nnz : 1) * B, nD = (size_t)nv * B;
    double* wU = (double*)ctx->workspace((nU + nD) * 8);
    double* wD = wU + nU;
    m->ops->crba_f64(ctx->stream, m->gen.get(), m->mc64, B, CView<double>{dq_, vq.rs, vq.bs}, View<double>{Mi, vo.rs, vo.bs}, 0);
    m->ops->chol_f64(ctx->stream, m->gen.get(), B, CView<double>{Mi, vo.rs, vo.bs}, View<double>{wU, B, 1}, View<double>{wD, B, 1});
    hipLaunchKernelGGL(k_identity, dim3((unsigned)((nv * nv * B + 255) / 256)), dim3(256), 0, ctx->stream, nv, B, Mi,
                       vo.rs, vo.bs);
    for (int c = 0; c < nv; ++c)
      m->ops->cholsolve_f64(ctx->stream, m->gen.get(), B, CView<double>{wU, B, 1}, CView<double>{wD, B, 1},
                            View<double>{Mi + (int64_t)c * nv * vo.rs, vo.rs, vo.bs});
    finishCall(ctx, flags, st);
  });
}

}  // extern "C"

namespace {
// frame placements and Jacobians: k_kin's oMi and WORLD Jacobian into the workspace, then
// k_frames.  The frame tables are host arrays, copied during the call (which therefore
// always synchronises).
template<class S>
void framesCall(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* q, int64_t ldq, int nf,
                const int32_t* fjoint, const double* fplace, int rf, void* oMf, int64_t ldo, void* J, int64_t ldJ,
                uint32_t flags) {
  constexpr bool f32 = std::is_same_v<S, float>;
  const auto launch = f32 ? (void*)m->ops->kin_f32 : (void*)m->ops->kin_f64;
  if (!launch) throw Error(RBD_EUNSUPPORTED, "frames: no kinematics kernel for this topology");
  const int nj = m->njoints, nv = m->nv;
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  ArgView<S> vM{}, vJ{};
  if (oMf) vM = argView<S>(12 * nf, B, ldo, flags, "oMf");
  if (J) vJ = argView<S>(6 * nv * nf, B, ldJ, flags, "J");
  if (B == 0 || nf == 0 || (!oMf && !J)) return;
  // tables: parent joint, placement (in S), support flags of every joint
  std::vector<S> pl((size_t)12 * nf);
  for (size_t k = 0; k < pl.size(); ++k) pl[k] = S(fplace[k]);
  std::vector<uint8_t> sup((size_t)nj * nv, 0);
  for (int j = 1; j < nj; ++j)
    for (int a = j; a > 0; a = m->parents[a]) {
      const int nva = m->jtype[a] == RBD_JOINT_FREEFLYER ? 6 : 1;
      for (int k = 0; k < nva; ++k) sup[(size_t)j * nv + m->iv[a] + k] = 1;
    }
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const S* dq = (const S*)st.in(q, vq.extent * sizeof(S));
  const bool bl = (flags & RBD_LAYOUT_BATCH) != 0;
  S* dM = (S*)st.out(oMf, vM.extent * sizeof(S), bl || ldo != 12 * nf);
  S* dJ = (S*)st.out(J, vJ.extent * sizeof(S), bl || ldJ != 6 * nv * nf);
  auto table = [&](const void* h, size_t bytes) {
    void* d = ctx->stageBuf(st.slot++, bytes);
    HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    return d;
  };
  const int32_t* dpar = (const int32_t*)table(fjoint, sizeof(int32_t) * nf);
  const S* dpl = (const S*)table(pl.data(), sizeof(S) * pl.size());
  const uint8_t* dsup = (const uint8_t*)table(sup.data(), sup.size());
  // workspace, batch-fastest: oMi (12 (nj - 1) rows), then the WORLD Jacobian (6 nv rows)
  const size_t nO = (size_t)12 * (nj - 1) * B, nJ = J ? (size_t)6 * nv * B : 0;
  S* wO = (S*)ctx->workspace((nO + nJ) * sizeof(S));
  S* wJ = J ? wO + nO : nullptr;
  const S* mc;
  if constexpr (f32) mc = m->mc32;
  else mc = m->mc64;
  reinterpret_cast<LaunchKin<S>>(launch)(ctx->stream, m->gen.get(), mc, B, CView<S>{dq, vq.rs, vq.bs}, CView<S>{}, CView<S>{},
                                        View<S>{wO, B, 1}, View<S>{}, View<S>{}, View<S>{wJ, B, 1});
  hipLaunchKernelGGL(k_frames<S>, dim3((unsigned)((B + 255) / 256), (unsigned)nf), dim3(256), 0, ctx->stream, B, nv,
                     dpar, dpl, dsup, CView<S>{wO, B, 1}, CView<S>{wJ, B, 1}, rf, View<S>{dM, vM.rs, vM.bs},
                     View<S>{dJ, vJ.rs, vJ.bs});
  finishCall(ctx, flags & ~(uint32_t)RBD_ASYNC, st);
}

// Frame motions: k_kin (oMi, LOCAL joint velocities and accelerations into a workspace),
// then k_frame_motions.  Frame tables are host arrays, copied during the call (which
// therefore synchronises).
template<class S>
void frameMotionsCall(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* q, int64_t ldq, const void* v,
                      int64_t ldv, const void* a, int64_t lda, int nf, const int32_t* fjoint, const double* fplace,
                      int rf, void* vf, void* af, void* acf, int64_t ldm, uint32_t flags) {
  constexpr bool f32 = std::is_same_v<S, float>;
  const auto launch = f32 ? (void*)m->ops->kin_f32 : (void*)m->ops->kin_f64;
  if (!launch) throw Error(RBD_EUNSUPPORTED, "frame motions: no kinematics kernel for this topology");
  const int nj = m->njoints;
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  auto vv = argView<S>(m->nv, B, ldv, flags, "v");
  ArgView<S> va{};
  if (a) va = argView<S>(m->nv, B, lda, flags, "a");
  auto vo = argView<S>(6 * nf, B, ldm, flags, "vf");
  if (B == 0 || nf == 0) return;
  std::vector<S> pl((size_t)12 * nf);
  for (size_t k = 0; k < pl.size(); ++k) pl[k] = S(fplace[k]);
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const S* dq = (const S*)st.in(q, vq.extent * sizeof(S));
  const S* dv = (const S*)st.in(v, vv.extent * sizeof(S));
  const S* da = a ? (const S*)st.in(a, va.extent * sizeof(S)) : nullptr;
  const bool pre = (flags & RBD_LAYOUT_BATCH) != 0 || ldm != 6 * nf;
  S* dvf = (S*)st.out(vf, vo.extent * sizeof(S), pre);
  S* daf = af ? (S*)st.out(af, vo.extent * sizeof(S), pre) : nullptr;
  S* dacf = acf ? (S*)st.out(acf, vo.extent * sizeof(S), pre) : nullptr;
  auto table = [&](const void* h, size_t bytes) {
    void* d = ctx->stageBuf(st.slot++, bytes);
    HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    return d;
  };
  const int32_t* dpar = (const int32_t*)table(fjoint, sizeof(int32_t) * nf);
  const S* dpl = (const S*)table(pl.data(), sizeof(S) * pl.size());
  // workspace, batch-fastest: oMi (12 (nj - 1) rows), joint velocities, accelerations (6 (nj - 1) each)
  const size_t nO = (size_t)12 * (nj - 1) * B, nV = (size_t)6 * (nj - 1) * B;
  S* wO = (S*)ctx->workspace((nO + (a ? 2 : 1) * nV) * sizeof(S));
  S* wV = wO + nO;
  S* wA = a ? wV + nV : nullptr;
  const S* mc;
  if constexpr (f32) mc = m->mc32;
  else mc = m->mc64;
  reinterpret_cast<LaunchKin<S>>(launch)(ctx->stream, m->gen.get(), mc, B, CView<S>{dq, vq.rs, vq.bs},
                                        CView<S>{dv, vv.rs, vv.bs}, CView<S>{da, va.rs, va.bs}, View<S>{wO, B, 1},
                                        View<S>{wV, B, 1}, View<S>{wA, B, 1}, View<S>{});
  hipLaunchKernelGGL(k_frame_motions<S>, dim3((unsigned)((B + 255) / 256), (unsigned)nf), dim3(256), 0, ctx->stream,
                     B, dpar, dpl, CView<S>{wO, B, 1}, CView<S>{wV, B, 1}, CView<S>{wA, B, 1}, rf,
                     View<S>{dvf, vo.rs, vo.bs}, View<S>{daf, vo.rs, vo.bs}, View<S>{dacf, vo.rs, vo.bs});
  finishCall(ctx, flags & ~(uint32_t)RBD_ASYNC, st);
}

// Jacobian time variation: k_kin (oMi, LOCAL joint velocities, WORLD Jacobian into a
// workspace), then k_jacobian_dot.
template<class S>
void jacobianDotCall(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* q, int64_t ldq, const void* v,
                     int64_t ldv, void* J, int64_t ldJ, void* dJ, int64_t lddJ, uint32_t flags) {
  constexpr bool f32 = std::is_same_v<S, float>;
  const auto launch = f32 ? (void*)m->ops->kin_f32 : (void*)m->ops->kin_f64;
  if (!launch) throw Error(RBD_EUNSUPPORTED, "Jacobian time variation: no kinematics kernel for this topology");
  const int nj = m->njoints, nv = m->nv;
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  auto vv = argView<S>(nv, B, ldv, flags, "v");
  auto vd = argView<S>(6 * nv, B, lddJ, flags, "dJ");
  ArgView<S> vJ{};
  if (J) vJ = argView<S>(6 * nv, B, ldJ, flags, "J");
  if (B == 0) return;
  std::vector<int32_t> dofJoint(nv);
  for (int j = 1; j < nj; ++j) {
    const int nvj = m->jtype[j] == RBD_JOINT_FREEFLYER ? 6 : 1;
    for (int k = 0; k < nvj; ++k) dofJoint[m->iv[j] + k] = j;
  }
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const S* dq = (const S*)st.in(q, vq.extent * sizeof(S));
  const S* dv = (const S*)st.in(v, vv.extent * sizeof(S));
  const bool bl = (flags & RBD_LAYOUT_BATCH) != 0;
  S* ddJ = (S*)st.out(dJ, vd.extent * sizeof(S), bl || lddJ != 6 * nv);
  S* dJo = J ? (S*)st.out(J, vJ.extent * sizeof(S), bl || ldJ != 6 * nv) : nullptr;
  void* dtab = ctx->stageBuf(st.slot++, sizeof(int32_t) * nv);
  HIP_CHECK(hipMemcpyAsync(dtab, dofJoint.data(), sizeof(int32_t) * nv, hipMemcpyHostToDevice, ctx->stream));
  // workspace, batch-fastest: oMi (12 (nj - 1) rows), joint velocities (6 (nj - 1)), WORLD Jacobian (6 nv)
  const size_t nO = (size_t)12 * (nj - 1) * B, nV = (size_t)6 * (nj - 1) * B, nJ = (size_t)6 * nv * B;
  S* wO = (S*)ctx->workspace((nO + nV + nJ) * sizeof(S));
  S* wV = wO + nO;
  S* wJ = wV + nV;
  const S* mc;
  if constexpr (f32) mc = m->mc32;
  else mc = m->mc64;
  reinterpret_cast<LaunchKin<S>>(launch)(ctx->stream, m->gen.get(), mc, B, CView<S>{dq, vq.rs, vq.bs},
                                        CView<S>{dv, vv.rs, vv.bs}, CView<S>{}, View<S>{wO, B, 1}, View<S>{wV, B, 1},
                                        View<S>{}, View<S>{wJ, B, 1});
  hipLaunchKernelGGL(k_jacobian_dot<S>, dim3((unsigned)((B + 255) / 256), (unsigned)nv), dim3(256), 0, ctx->stream, B,
                     (const int32_t*)dtab, CView<S>{wO, B, 1}, CView<S>{wV, B, 1}, CView<S>{wJ, B, 1},
                     View<S>{ddJ, vd.rs, vd.bs}, View<S>{dJo, vJ.rs, vJ.bs});
  finishCall(ctx, flags & ~(uint32_t)RBD_ASYNC, st);
}

// Frame Jacobian time variation: k_kin (oMi, joint velocities, WORLD Jacobian),
// k_jacobian_dot (WORLD dJ), then k_frame_jacobian_dot.  Frame tables are host arrays,
// copied during the call (which therefore synchronises).
template<class S>
void frameJacobianDotCall(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* q, int64_t ldq, const void* v,
                          int64_t ldv, int nf, const int32_t* fjoint, const double* fplace, int rf, void* dJ,
                          int64_t lddJ, uint32_t flags) {
  constexpr bool f32 = std::is_same_v<S, float>;
  const auto launch = f32 ? (void*)m->ops->kin_f32 : (void*)m->ops->kin_f64;
  if (!launch) throw Error(RBD_EUNSUPPORTED, "frame Jacobian time variation: no kinematics kernel for this topology");
  const int nj = m->njoints, nv = m->nv;
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  auto vv = argView<S>(nv, B, ldv, flags, "v");
  auto vd = argView<S>(6 * nv * nf, B, lddJ, flags, "dJ");
  if (B == 0 || nf == 0) return;
  std::vector<S> pl((size_t)12 * nf);
  for (size_t k = 0; k < pl.size(); ++k) pl[k] = S(fplace[k]);
  std::vector<uint8_t> sup((size_t)nj * nv, 0);
  std::vector<int32_t> dofJoint(nv);
  for (int j = 1; j < nj; ++j) {
    const int nvj = m->jtype[j] == RBD_JOINT_FREEFLYER ? 6 : 1;
    for (int k = 0; k < nvj; ++k) dofJoint[m->iv[j] + k] = j;
    for (int a = j; a > 0; a = m->parents[a]) {
      const int nva = m->jtype[a] == RBD_JOINT_FREEFLYER ? 6 : 1;
      for (int k = 0; k < nva; ++k) sup[(size_t)j * nv + m->iv[a] + k] = 1;
    }
  }
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const S* dq = (const S*)st.in(q, vq.extent * sizeof(S));
  const S* dv = (const S*)st.in(v, vv.extent * sizeof(S));
  S* dd = (S*)st.out(dJ, vd.extent * sizeof(S), (flags & RBD_LAYOUT_BATCH) != 0 || lddJ != 6 * nv * nf);
  auto table = [&](const void* h, size_t bytes) {
    void* d = ctx->stageBuf(st.slot++, bytes);
    HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    return d;
  };
  const int32_t* dpar = (const int32_t*)table(fjoint, sizeof(int32_t) * nf);
  const S* dpl = (const S*)table(pl.data(), sizeof(S) * pl.size());
  const uint8_t* dsup = (const uint8_t*)table(sup.data(), sup.size());
  const int32_t* ddof = (const int32_t*)table(dofJoint.data(), sizeof(int32_t) * nv);
  // workspace, batch-fastest: oMi, joint velocities, WORLD J, WORLD dJ
  const size_t nO = (size_t)12 * (nj - 1) * B, nV = (size_t)6 * (nj - 1) * B, nJ = (size_t)6 * nv * B;
  S* wO = (S*)ctx->workspace((nO + nV + 2 * nJ) * sizeof(S));
  S* wV = wO + nO;
  S* wJ = wV + nV;
  S* wD = wJ + nJ;
  const S* mc;
  if constexpr (f32) mc = m->mc32;
  else mc = m->mc64;
  reinterpret_cast<LaunchKin<S>>(launch)(ctx->stream, m->gen.get(), mc, B, CView<S>{dq, vq.rs, vq.bs},
                                        CView<S>{dv, vv.rs, vv.bs}, CView<S>{}, View<S>{wO, B, 1}, View<S>{wV, B, 1},
                                        View<S>{}, View<S>{wJ, B, 1});
  const unsigned gb = (unsigned)((B + 255) / 256);
  hipLaunchKernelGGL(k_jacobian_dot<S>, dim3(gb, (unsigned)nv), dim3(256), 0, ctx->stream, B, ddof, CView<S>{wO, B, 1},
                     CView<S>{wV, B, 1}, CView<S>{wJ, B, 1}, View<S>{wD, B, 1}, View<S>{});
  hipLaunchKernelGGL(k_frame_jacobian_dot<S>, dim3(gb, (unsigned)nf), dim3(256), 0, ctx->stream, B, nv, dpar, dpl, dsup,
                     CView<S>{wO, B, 1}, CView<S>{wV, B, 1}, CView<S>{wJ, B, 1}, CView<S>{wD, B, 1}, rf,
                     View<S>{dd, vd.rs, vd.bs});
  finishCall(ctx, flags & ~(uint32_t)RBD_ASYNC, st);
}

template<class S> __global__ void k_fill_zero(S* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = S(0);
}

// forwardDynamics(model, data, q, v, tau, J, gamma, inv_damping): CRBA, FP64 tree factor,
// nle = RNEA at a = 0, then the KKT kernels (k_kkt.hpp)
template<class S>
void cfdCall(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* q, int64_t ldq, const void* v, int64_t ldv,
             const void* tau, int64_t ldtau, int nc, const void* J, int64_t ldJ, const void* gamma, int64_t ldg,
             double damping, void* ddq, int64_t ldddq, void* lambda, int64_t ldl, uint32_t flags) {
  constexpr bool f32 = std::is_same_v<S, float>;
  const int nv = m->nv, nnz = m->ops->nnzU;
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  auto vv = argView<S>(nv, B, ldv, flags, "v");
  auto vt = argView<S>(nv, B, ldtau, flags, "tau");
  auto vJ = argView<S>(nc * nv, B, ldJ, flags, "J");
  auto vg = argView<S>(nc, B, ldg, flags, "gamma");
  auto va = argView<S>(nv, B, ldddq, flags, "ddq");
  auto vl = argView<S>(nc, B, ldl, flags, "lambda");
  if (B == 0) return;
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const bool batchLayout = (flags & RBD_LAYOUT_BATCH) != 0;
  const S* dq = (const S*)st.in(q, vq.extent * sizeof(S));
  const S* dv = (const S*)st.in(v, vv.extent * sizeof(S));
  const S* dt = (const S*)st.in(tau, vt.extent * sizeof(S));
  const S* dJ = (const S*)st.in(J, vJ.extent * sizeof(S));
  const S* dg = (const S*)st.in(gamma, vg.extent * sizeof(S));
  S* da = (S*)st.out(ddq, va.extent * sizeof(S), ldddq != nv || batchLayout);
  S* dl = (S*)st.out(lambda, vl.extent * sizeof(S), ldl != nc || batchLayout);
  // workspace, batch-fastest: U, D (FP64), then M, nle and one zero (S)
  const size_t nU = (size_t)(nnz > 0 ? nnz : 1) * B, nD = (size_t)nv * B;
  const size_t nM = (size_t)nv * nv * B, nN = (size_t)nv * B;
  double* wU = (double*)ctx->workspace((nU + nD) * 8 + (nM + nN + 1) * sizeof(S));
  double* wD = wU + nU;
  S* wM = (S*)(wD + nD);
  S* wN = wM + nM;
  S* wZ = wN + nN;
  const S* mc = f32 ? (const S*)m->mc32 : (const S*)m->mc64;
  const TopoOps& o = *m->ops;
  const CView<S> cq{dq, vq.rs, vq.bs}, cv{dv, vv.rs, vv.bs};
  if constexpr (f32) {
    o.crba_f32(ctx->stream, m->gen.get(), mc, B, cq, View<S>{wM, B, 1}, 0);
    o.chol_f32(ctx->stream, m->gen.get(), B, CView<S>{wM, B, 1}, View<double>{wU, B, 1}, View<double>{wD, B, 1});
  } else {
    o.crba_f64(ctx->stream, m->gen.get(), mc, B, cq, View<S>{wM, B, 1}, 0);
    o.chol_f64(ctx->stream, m->gen.get(), B, CView<S>{wM, B, 1}, View<double>{wU, B, 1}, View<double>{wD, B, 1});
  }
  // nle = rnea(q, v, 0): the acceleration view repeats one zero (both strides 0)
  hipLaunchKernelGGL(k_fill_zero<S>, dim3(1), dim3(64), 0, ctx->stream, wZ, (int64_t)1);
  if constexpr (f32) o.rnea_f32(ctx->stream, m->gen.get(), mc, B, cq, cv, CView<S>{wZ, 0, 0}, View<S>{wN, B, 1});
  else o.rnea_f64(ctx->stream, m->gen.get(), mc, B, cq, cv, CView<S>{wZ, 0, 0}, View<S>{wN, B, 1});
  const auto kkt = f32 ? (LaunchKkt<S>)o.kkt_f32 : (LaunchKkt<S>)o.kkt_f64;
  kkt(ctx->stream, m->gen.get(), B, CView<double>{wU, B, 1}, CView<double>{wD, B, 1}, nc, CView<S>{dJ, vJ.rs, vJ.bs},
      CView<S>{dt, vt.rs, vt.bs}, CView<S>{wN, B, 1}, CView<S>{dg, vg.rs, vg.bs}, damping, 0, 0.0,
      KktProx{1, 0.0, 0.0, nullptr}, View<S>{da, va.rs, va.bs}, View<S>{dl, vl.rs, vl.bs});
  finishCall(ctx, flags, st);
}

// impulseDynamics(model, data, q, v_before, J, r_coeff, inv_damping): CRBA, FP64 tree
// factor, then k_kkt in impulse mode
template<class S>
void impulseCall(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* q, int64_t ldq, const void* v,
                 int64_t ldv, int nc, const void* J, int64_t ldJ, double rcoef, double damping, void* dqa,
                 int64_t lda, void* imp, int64_t ldi, uint32_t flags) {
  constexpr bool f32 = std::is_same_v<S, float>;
  const int nv = m->nv, nnz = m->ops->nnzU;
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  auto vv = argView<S>(nv, B, ldv, flags, "v_before");
  auto vJ = argView<S>(nc * nv, B, ldJ, flags, "J");
  auto va = argView<S>(nv, B, lda, flags, "dq_after");
  auto vi = argView<S>(nc, B, ldi, flags, "impulse");
  if (B == 0) return;
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const bool batchLayout = (flags & RBD_LAYOUT_BATCH) != 0;
  const S* dq = (const S*)st.in(q, vq.extent * sizeof(S));
  const S* dv = (const S*)st.in(v, vv.extent * sizeof(S));
  const S* dJ = (const S*)st.in(J, vJ.extent * sizeof(S));
  S* da = (S*)st.out(dqa, va.extent * sizeof(S), lda != nv || batchLayout);
  S* di = (S*)st.out(imp, vi.extent * sizeof(S), ldi != nc || batchLayout);
  const size_t nU = (size_t)(nnz > 0 ? nnz : 1) * B, nD = (size_t)nv * B, nM = (size_t)nv * nv * B;
  double* wU = (double*)ctx->workspace((nU + nD) * 8 + nM * sizeof(S));
  double* wD = wU + nU;
  S* wM = (S*)(wD + nD);
  const S* mc = f32 ? (const S*)m->mc32 : (const S*)m->mc64;
  const TopoOps& o = *m->ops;
  const CView<S> cq{dq, vq.rs, vq.bs};
  if constexpr (f32) {
    o.crba_f32(ctx->stream, m->gen.get(), mc, B, cq, View<S>{wM, B, 1}, 0);
    o.chol_f32(ctx->stream, m->gen.get(), B, CView<S>{wM, B, 1}, View<double>{wU, B, 1}, View<double>{wD, B, 1});
  } else {
    o.crba_f64(ctx->stream, m->gen.get(), mc, B, cq, View<S>{wM, B, 1}, 0);
    o.chol_f64(ctx->stream, m->gen.get(), B, CView<S>{wM, B, 1}, View<double>{wU, B, 1}, View<double>{wD, B, 1});
  }
  const auto kkt = f32 ? (LaunchKkt<S>)o.kkt_f32 : (LaunchKkt<S>)o.kkt_f64;
  kkt(ctx->stream, m->gen.get(), B, CView<double>{wU, B, 1}, CView<double>{wD, B, 1}, nc, CView<S>{dJ, vJ.rs, vJ.bs},
      CView<S>{dv, vv.rs, vv.bs}, CView<S>{}, CView<S>{}, damping, 1, rcoef, KktProx{1, 0.0, 0.0, nullptr},
      View<S>{da, va.rs, va.bs}, View<S>{di, vi.rs, vi.bs});
  finishCall(ctx, flags, st);
}

// constraintDynamics(model, data, q, v, tau, contact_models, contact_datas, settings): the
// kinematics pass at zero joint acceleration (oMi, local v and a, WORLD joint Jacobian),
// the constraint rows (k_contact.hpp), CRBA, the FP64 tree factor, nle = RNEA at a = 0,
// then the KKT kernel with mu as its damping and the proximal iterations.
// Device table of the constraints and their support flags (bit 0: the dof is in joint1's
// support, bit 1: in joint2's); returns the total row count.
int constraintTable(const rbd_model* m, int ncon, const rbd_rigid_constraint* cons, std::vector<RigidConstraintDev>& tab,
                    std::vector<uint8_t>& sup) {
  const int nv = m->nv;
  tab.assign(ncon, RigidConstraintDev{});
  sup.assign((size_t)ncon * nv, 0);
  int nrows = 0;
  for (int c = 0; c < ncon; ++c) {
    const rbd_rigid_constraint& k = cons[c];
    RigidConstraintDev& d = tab[c];
    d.type = k.type;
    d.j1 = k.joint1_id;
    d.j2 = k.joint2_id;
    d.rf = k.reference_frame;
    d.row = nrows;
    const int dim = k.type == RBD_CONTACT_6D ? 6 : 3;
    d.gains = 0;
    for (int e = 0; e < 6; ++e) {
      d.Kp[e] = k.Kp[e];
      d.Kd[e] = k.Kd[e];
      d.des[e] = k.desired_acceleration[e];
      if (e < dim && (k.Kp[e] != 0.0 || k.Kd[e] != 0.0)) d.gains = 1;
    }
    for (int e = 0; e < 12; ++e) {
      d.M1[e] = k.joint1_placement[e];
      d.M2[e] = k.joint2_placement[e];
    }
    for (int bit = 0; bit < 2; ++bit)
      for (int a = bit ? k.joint2_id : k.joint1_id; a > 0; a = m->parents[a]) {
        const int nva = m->jtype[a] == RBD_JOINT_FREEFLYER ? 6 : 1;
        for (int e = 0; e < nva; ++e) sup[(size_t)c * nv + m->iv[a] + e] |= (uint8_t)(1 << bit);
      }
    nrows += dim;
  }
  return nrows;
}

template<class S>
void cdCall(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* q, int64_t ldq, const void* v, int64_t ldv,
            const void* tau, int64_t ldtau, int ncon, const rbd_rigid_constraint* cons, const rbd_proximal_settings& ps,
            void* ddq, int64_t ldddq, void* lambda, int64_t ldl, int32_t* iters, uint32_t flags) {
  constexpr bool f32 = std::is_same_v<S, float>;
  const int nv = m->nv, nj = m->njoints, nnz = m->ops->nnzU;
  std::vector<RigidConstraintDev> tab;
  std::vector<uint8_t> sup;
  const int nrows = constraintTable(m, ncon, cons, tab, sup);
  if (nrows > kKktMaxConstraints) throw Error(RBD_EINVAL, "the constraints' total dimension must be at most 12");
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  auto vv = argView<S>(nv, B, ldv, flags, "v");
  auto vt = argView<S>(nv, B, ldtau, flags, "tau");
  auto va = argView<S>(nv, B, ldddq, flags, "ddq");
  auto vl = argView<S>(nrows, B, ldl, flags, "lambda");
  if (B == 0) return;
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const bool batchLayout = (flags & RBD_LAYOUT_BATCH) != 0;
  const S* dq = (const S*)st.in(q, vq.extent * sizeof(S));
  const S* dv = (const S*)st.in(v, vv.extent * sizeof(S));
  const S* dt = (const S*)st.in(tau, vt.extent * sizeof(S));
  S* da = (S*)st.out(ddq, va.extent * sizeof(S), ldddq != nv || batchLayout);
  S* dl = (S*)st.out(lambda, vl.extent * sizeof(S), ldl != nrows || batchLayout);
  int* di = (int*)st.out(iters, (size_t)B * sizeof(int32_t), false);
  auto table = [&](const void* h, size_t bytes) {
    void* d = ctx->stageBuf(st.slot++, bytes);
    HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    return d;
  };
  const RigidConstraintDev* dtab = (const RigidConstraintDev*)table(tab.data(), sizeof(RigidConstraintDev) * ncon);
  const uint8_t* dsup = (const uint8_t*)table(sup.data(), sup.size());
  // workspace, batch-fastest: U, D (FP64); oMi, v, a of the joints, the WORLD Jacobian, the
  // constraint rows and gamma, M, nle and one zero (S)
  const size_t nU = (size_t)(nnz > 0 ? nnz : 1) * B, nD = (size_t)nv * B;
  const size_t nO = (size_t)12 * (nj - 1) * B, nVA = (size_t)6 * (nj - 1) * B, nJw = (size_t)6 * nv * B;
  const size_t nJc = (size_t)nrows * nv * B, nG = (size_t)nrows * B, nM = (size_t)nv * nv * B, nN = (size_t)nv * B;
  double* wU = (double*)ctx->workspace((nU + nD) * 8 + (nO + 2 * nVA + nJw + nJc + nG + nM + nN + 1) * sizeof(S));
  double* wD = wU + nU;
  S* wO = (S*)(wD + nD);
  S* wV = wO + nO;
  S* wA = wV + nVA;
  S* wJw = wA + nVA;
  S* wJc = wJw + nJw;
  S* wG = wJc + nJc;
  S* wM = wG + nG;
  S* wN = wM + nM;
  S* wZ = wN + nN;
  const S* mc = f32 ? (const S*)m->mc32 : (const S*)m->mc64;
  const TopoOps& o = *m->ops;
  const CView<S> cq{dq, vq.rs, vq.bs}, cv{dv, vv.rs, vv.bs}, zero{wZ, 0, 0};
  hipLaunchKernelGGL(k_fill_zero<S>, dim3(1), dim3(64), 0, ctx->stream, wZ, (int64_t)1);
  const auto kin = f32 ? (LaunchKin<S>)o.kin_f32 : (LaunchKin<S>)o.kin_f64;
  kin(ctx->stream, m->gen.get(), mc, B, cq, cv, zero, View<S>{wO, B, 1}, View<S>{wV, B, 1}, View<S>{wA, B, 1},
      View<S>{wJw, B, 1});
  hipLaunchKernelGGL(k_contact_terms<S>, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, ctx->stream, B, nv, ncon,
                     nrows, dtab, dsup, CView<S>{wO, B, 1}, CView<S>{wV, B, 1}, CView<S>{wA, B, 1},
                     CView<S>{wJw, B, 1}, View<S>{wJc, B, 1}, View<S>{wG, B, 1});
  if constexpr (f32) {
    o.crba_f32(ctx->stream, m->gen.get(), mc, B, cq, View<S>{wM, B, 1}, 0);
    o.chol_f32(ctx->stream, m->gen.get(), B, CView<S>{wM, B, 1}, View<double>{wU, B, 1}, View<double>{wD, B, 1});
    o.rnea_f32(ctx->stream, m->gen.get(), mc, B, cq, cv, zero, View<S>{wN, B, 1});
  } else {
    o.crba_f64(ctx->stream, m->gen.get(), mc, B, cq, View<S>{wM, B, 1}, 0);
    o.chol_f64(ctx->stream, m->gen.get(), B, CView<S>{wM, B, 1}, View<double>{wU, B, 1}, View<double>{wD, B, 1});
    o.rnea_f64(ctx->stream, m->gen.get(), mc, B, cq, cv, zero, View<S>{wN, B, 1});
  }
  const auto kkt = f32 ? (LaunchKkt<S>)o.kkt_f32 : (LaunchKkt<S>)o.kkt_f64;
  kkt(ctx->stream, m->gen.get(), B, CView<double>{wU, B, 1}, CView<double>{wD, B, 1}, nrows, CView<S>{wJc, B, 1},
      CView<S>{dt, vt.rs, vt.bs}, CView<S>{wN, B, 1}, CView<S>{wG, B, 1}, ps.mu, 0, 0.0,
      KktProx{ps.max_iter, ps.absolute_accuracy, ps.relative_accuracy, di}, View<S>{da, va.rs, va.bs},
      View<S>{dl, vl.rs, vl.bs});
  finishCall(ctx, flags, st);
}

// The constraints' matrices at q: kinematics at zero velocity, the constraint rows, CRBA and
// the FP64 tree factor, then the KKT kernel in mode 2 (computeDelassusMatrix: out = G,
// nrows^2 rows), 4 (computeDampedDelassusMatrixInverse: out = scale (G + mu I)^-1) or 3 (a
// solve with the contact Cholesky's KKT matrix [[-mu I, J], [J^T, M]]: rhs and out hold
// nrows + nv rows, the constraint rows first).
template<class S>
void constraintMatrixCall(rbd_ctx* ctx, const rbd_model* m, int64_t B, const void* q, int64_t ldq, int ncon,
                          const rbd_rigid_constraint* cons, int mode, double mu, double scale, const void* rhs,
                          int64_t ldrhs, void* out, int64_t ldout, uint32_t flags) {
  constexpr bool f32 = std::is_same_v<S, float>;
  const int nv = m->nv, nj = m->njoints, nnz = m->ops->nnzU;
  std::vector<RigidConstraintDev> tab;
  std::vector<uint8_t> sup;
  const int nrows = constraintTable(m, ncon, cons, tab, sup);
  if (nrows > kKktMaxConstraints) throw Error(RBD_EINVAL, "the constraints' total dimension must be at most 12");
  const int outRows = mode == 3 ? nrows + nv : nrows * nrows;
  auto vq = argView<S>(m->nq, B, ldq, flags, "q");
  auto vo = argView<S>(outRows, B, ldout, flags, mode == 3 ? "x" : "G");
  ArgView<S> vr{};
  if (mode == 3) vr = argView<S>(nrows + nv, B, ldrhs, flags, "rhs");
  if (B == 0) return;
  setDevice(ctx);
  Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
  const bool batchLayout = (flags & RBD_LAYOUT_BATCH) != 0;
  const S* dq = (const S*)st.in(q, vq.extent * sizeof(S));
  const S* dr = mode == 3 ? (const S*)st.in(rhs, vr.extent * sizeof(S)) : nullptr;
  S* dout = (S*)st.out(out, vo.extent * sizeof(S), ldout != outRows || batchLayout);
  auto table = [&](const void* h, size_t bytes) {
    void* d = ctx->stageBuf(st.slot++, bytes);
    HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    return d;
  };
  const RigidConstraintDev* dtab = (const RigidConstraintDev*)table(tab.data(), sizeof(RigidConstraintDev) * ncon);
  const uint8_t* dsup = (const uint8_t*)table(sup.data(), sup.size());
  // workspace, batch-fastest: U, D (FP64); oMi, v, a of the joints, the WORLD Jacobian, the
  // constraint rows and their (unused) drifts, M and one zero (S)
  const size_t nU = (size_t)(nnz > 0 ? nnz : 1) * B, nD = (size_t)nv * B;
  const size_t nO = (size_t)12 * (nj - 1) * B, nVA = (size_t)6 * (nj - 1) * B, nJw = (size_t)6 * nv * B;
  const size_t nJc = (size_t)nrows * nv * B, nGm = (size_t)nrows * B, nM = (size_t)nv * nv * B;
  double* wU = (double*)ctx->workspace((nU + nD) * 8 + (nO + 2 * nVA + nJw + nJc + nGm + nM + 1) * sizeof(S));
  double* wD = wU + nU;
  S* wO = (S*)(wD + nD);
  S* wV = wO + nO;
  S* wA = wV + nVA;
  S* wJw = wA + nVA;
  S* wJc = wJw + nJw;
  S* wGm = wJc + nJc;
  S* wM = wGm + nGm;
  S* wZ = wM + nM;
  const S* mc = f32 ? (const S*)m->mc32 : (const S*)m->mc64;
  const TopoOps& o = *m->ops;
  const CView<S> cq{dq, vq.rs, vq.bs}, zero{wZ, 0, 0};
  hipLaunchKernelGGL(k_fill_zero<S>, dim3(1), dim3(64), 0, ctx->stream, wZ, (int64_t)1);
  const auto kin = f32 ? (LaunchKin<S>)o.kin_f32 : (LaunchKin<S>)o.kin_f64;
  kin(ctx->stream, m->gen.get(), mc, B, cq, zero, zero, View<S>{wO, B, 1}, View<S>{wV, B, 1}, View<S>{wA, B, 1},
      View<S>{wJw, B, 1});
  hipLaunchKernelGGL(k_contact_terms<S>, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, ctx->stream, B, nv, ncon,
                     nrows, dtab, dsup, CView<S>{wO, B, 1}, CView<S>{wV, B, 1}, CView<S>{wA, B, 1},
                     CView<S>{wJw, B, 1}, View<S>{wJc, B, 1}, View<S>{wGm, B, 1});
  if constexpr (f32) {
    o.crba_f32(ctx->stream, m->gen.get(), mc, B, cq, View<S>{wM, B, 1}, 0);
    o.chol_f32(ctx->stream, m->gen.get(), B, CView<S>{wM, B, 1}, View<double>{wU, B, 1}, View<double>{wD, B, 1});
  } else {
    o.crba_f64(ctx->stream, m->gen.get(), mc, B, cq, View<S>{wM, B, 1}, 0);
    o.chol_f64(ctx->stream, m->gen.get(), B, CView<S>{wM, B, 1}, View<double>{wU, B, 1}, View<double>{wD, B, 1});
  }
  const auto kkt = f32 ? (LaunchKkt<S>)o.kkt_f32 : (LaunchKkt<S>)o.kkt_f64;
  const CView<double> cU{wU, B, 1}, cD{wD, B, 1};
  if (mode == 3)
    kkt(ctx->stream, m->gen.get(), B, cU, cD, nrows, CView<S>{wJc, B, 1},
        CView<S>{dr + (int64_t)nrows * vr.rs, vr.rs, vr.bs}, zero, CView<S>{dr, vr.rs, vr.bs}, mu, 3, 0.0,
        KktProx{1, 0.0, 0.0, nullptr}, View<S>{dout + (int64_t)nrows * vo.rs, vo.rs, vo.bs},
        View<S>{dout, vo.rs, vo.bs});
  else
    kkt(ctx->stream, m->gen.get(), B, cU, cD, nrows, CView<S>{wJc, B, 1}, CView<S>{}, CView<S>{}, CView<S>{}, mu,
        mode, scale, KktProx{1, 0.0, 0.0, nullptr}, View<S>{}, View<S>{dout, vo.rs, vo.bs});
  finishCall(ctx, flags, st);
}

static_assert(kNcpStats == RBD_NCP_STATS, "per-problem statistics rows");

// checks shared by the constraint-matrix entry points
void checkConstraintSet(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, int ncontacts,
                        const rbd_rigid_constraint* contacts, double mu, const char* what) {
  checkCommon(ctx, m, B, dtype);
  RBD_CHECK_ARG(ncontacts >= 1 && ncontacts <= kKktMaxConstraints, "between 1 and 12 constraints");
  RBD_CHECK_ARG(contacts != nullptr, "null constraint table");
  for (int c = 0; c < ncontacts; ++c) {
    const rbd_rigid_constraint& k = contacts[c];
    RBD_CHECK_ARG(k.type == RBD_CONTACT_3D || k.type == RBD_CONTACT_6D, "constraint type must be RBD_CONTACT_3D or RBD_CONTACT_6D");
    RBD_CHECK_ARG(k.joint1_id >= 0 && k.joint1_id < m->njoints && k.joint2_id >= 0 && k.joint2_id < m->njoints,
                  "The index of the Joint is outside the bounds.");
    RBD_CHECK_ARG(k.reference_frame == RBD_LOCAL || k.reference_frame == RBD_LOCAL_WORLD_ALIGNED,
                  "the constraint rows are taken in LOCAL or LOCAL_WORLD_ALIGNED frames");
  }
  RBD_CHECK_ARG(mu >= 0.0, "mu has to be positive");
  if (!m->compact || !m->ops->kkt_f32 || !m->ops->crba_f32 || !m->ops->kin_f32)
    throw Error(RBD_EUNSUPPORTED, std::string(what) + ": no kernels for this topology");
}

// the cone count of a contact problem and its checks
void checkNcp(rbd_ctx* ctx, int64_t B, int ncones) {
  RBD_CHECK_ARG(ctx != nullptr, "null context");
  RBD_CHECK_ARG(B >= 0, "negative batch size");
  RBD_CHECK_ARG(ncones >= 1 && ncones <= kNcpMaxCones, "between 1 and 4 friction cones per problem");
}
}  // namespace

extern "C" {

int rbd_frames(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq, int nframes,
               const int32_t* frame_joint, const double* frame_placement, int reference_frame, void* oMf,
               int64_t ldoMf, void* J, int64_t ldJ, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(nframes >= 0 && nframes <= 65535, "the frame count must be between 0 and 65535");
    RBD_CHECK_ARG(nframes == 0 || (frame_joint && frame_placement), "null frame table");
    RBD_CHECK_ARG(reference_frame >= RBD_WORLD && reference_frame <= RBD_LOCAL_WORLD_ALIGNED,
                  "reference_frame must be RBD_WORLD, RBD_LOCAL or RBD_LOCAL_WORLD_ALIGNED");
    for (int f = 0; f < nframes; ++f)
      RBD_CHECK_ARG(frame_joint[f] >= 0 && frame_joint[f] < m->njoints, "The index of the Joint is outside the bounds.");
    RBD_CHECK_ARG(B == 0 || nframes == 0 || q, "null array argument");
    if (dtype == RBD_F32)
      framesCall<float>(ctx, m, B, q, ldq, nframes, frame_joint, frame_placement, reference_frame, oMf, ldoMf, J, ldJ,
                        flags);
    else
      framesCall<double>(ctx, m, B, q, ldq, nframes, frame_joint, frame_placement, reference_frame, oMf, ldoMf, J,
                         ldJ, flags);
  });
}

int rbd_joint_jacobians_time_variation(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q,
                                       int64_t ldq, const void* v, int64_t ldv, void* J, int64_t ldJ, void* dJ,
                                       int64_t lddJ, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(B == 0 || (q && v && dJ), "null array argument");
    if (dtype == RBD_F32) jacobianDotCall<float>(ctx, m, B, q, ldq, v, ldv, J, ldJ, dJ, lddJ, flags);
    else jacobianDotCall<double>(ctx, m, B, q, ldq, v, ldv, J, ldJ, dJ, lddJ, flags);
  });
}

int rbd_frame_jacobians_time_variation(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q,
                                       int64_t ldq, const void* v, int64_t ldv, int nframes, const int32_t* frame_joint,
                                       const double* frame_placement, int reference_frame, void* dJ, int64_t lddJ,
                                       uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(nframes >= 0 && nframes <= 65535, "the frame count must be between 0 and 65535");
    RBD_CHECK_ARG(nframes == 0 || (frame_joint && frame_placement), "null frame table");
    RBD_CHECK_ARG(reference_frame >= RBD_WORLD && reference_frame <= RBD_LOCAL_WORLD_ALIGNED, "Bad reference frame.");
    for (int f = 0; f < nframes; ++f)
      RBD_CHECK_ARG(frame_joint[f] >= 0 && frame_joint[f] < m->njoints, "The index of the Joint is outside the bounds.");
    RBD_CHECK_ARG(B == 0 || nframes == 0 || (q && v && dJ), "null array argument");
    if (dtype == RBD_F32)
      frameJacobianDotCall<float>(ctx, m, B, q, ldq, v, ldv, nframes, frame_joint, frame_placement, reference_frame, dJ,
                                  lddJ, flags);
    else
      frameJacobianDotCall<double>(ctx, m, B, q, ldq, v, ldv, nframes, frame_joint, frame_placement, reference_frame,
                                   dJ, lddJ, flags);
  });
}

int rbd_frame_motions(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                      const void* v, int64_t ldv, const void* a, int64_t lda, int nframes, const int32_t* frame_joint,
                      const double* frame_placement, int reference_frame, void* vf, void* af, void* acf, int64_t ldm,
                      uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(nframes >= 0 && nframes <= 65535, "the frame count must be between 0 and 65535");
    RBD_CHECK_ARG(nframes == 0 || (frame_joint && frame_placement), "null frame table");
    RBD_CHECK_ARG(reference_frame >= RBD_WORLD && reference_frame <= RBD_LOCAL_WORLD_ALIGNED, "Bad reference frame.");
    for (int f = 0; f < nframes; ++f)
      RBD_CHECK_ARG(frame_joint[f] >= 0 && frame_joint[f] < m->njoints, "The index of the Joint is outside the bounds.");
    RBD_CHECK_ARG(B == 0 || nframes == 0 || (q && v && vf), "null array argument");
    RBD_CHECK_ARG(a || (!af && !acf), "frame accelerations need the joint accelerations a");
    if (dtype == RBD_F32)
      frameMotionsCall<float>(ctx, m, B, q, ldq, v, ldv, a, lda, nframes, frame_joint, frame_placement, reference_frame,
                              vf, af, acf, ldm, flags);
    else
      frameMotionsCall<double>(ctx, m, B, q, ldq, v, ldv, a, lda, nframes, frame_joint, frame_placement,
                               reference_frame, vf, af, acf, ldm, flags);
  });
}

int rbd_impulse_dynamics(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                         const void* v_before, int64_t ldv, int nc, const void* J, int64_t ldJ, double r_coeff,
                         double inv_damping, void* dq_after, int64_t ldda, void* impulse, int64_t ldimpulse,
                         uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(nc >= 1 && nc <= kKktMaxConstraints, "the constraint count must be between 1 and 12");
    RBD_CHECK_ARG(B == 0 || (q && v_before && J && dq_after && impulse), "null array argument");
    if (!m->compact || !m->ops->kkt_f32 || !m->ops->crba_f32)
      throw Error(RBD_EUNSUPPORTED, "impulse dynamics: no kernels for this topology");
    if (dtype == RBD_F32)
      impulseCall<float>(ctx, m, B, q, ldq, v_before, ldv, nc, J, ldJ, r_coeff, inv_damping, dq_after, ldda, impulse,
                         ldimpulse, flags);
    else
      impulseCall<double>(ctx, m, B, q, ldq, v_before, ldv, nc, J, ldJ, r_coeff, inv_damping, dq_after, ldda,
                          impulse, ldimpulse, flags);
  });
}

int rbd_constraint_dynamics(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                            const void* v, int64_t ldv, const void* tau, int64_t ldtau, int ncontacts,
                            const rbd_rigid_constraint* contacts, const rbd_proximal_settings* settings, void* ddq,
                            int64_t ldddq, void* lambda, int64_t ldlambda, int32_t* iterations, uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(ncontacts >= 1 && ncontacts <= kKktMaxConstraints, "between 1 and 12 constraints");
    RBD_CHECK_ARG(contacts != nullptr, "null constraint table");
    for (int c = 0; c < ncontacts; ++c) {
      const rbd_rigid_constraint& k = contacts[c];
      RBD_CHECK_ARG(k.type == RBD_CONTACT_3D || k.type == RBD_CONTACT_6D, "constraint type must be RBD_CONTACT_3D or RBD_CONTACT_6D");
      RBD_CHECK_ARG(k.joint1_id >= 0 && k.joint1_id < m->njoints && k.joint2_id >= 0 && k.joint2_id < m->njoints,
                    "The index of the Joint is outside the bounds.");
      RBD_CHECK_ARG(k.reference_frame == RBD_LOCAL || k.reference_frame == RBD_LOCAL_WORLD_ALIGNED,
                    "constraintDynamics is only operating from LOCAL or LOCAL_WORLD_ALIGNED reference frames.");
    }
    rbd_proximal_settings ps{1e-12, 1e-12, 0.0, 1};
    if (settings) ps = *settings;
    RBD_CHECK_ARG(ps.mu >= 0.0, "mu has to be positive");
    RBD_CHECK_ARG(ps.max_iter >= 1, "max_iter must be at least 1");
    RBD_CHECK_ARG(B == 0 || (q && v && tau && ddq && lambda), "null array argument");
    if (!m->compact || !m->ops->kkt_f32 || !m->ops->crba_f32 || !m->ops->kin_f32)
      throw Error(RBD_EUNSUPPORTED, "constraint dynamics: no kernels for this topology");
    if (dtype == RBD_F32)
      cdCall<float>(ctx, m, B, q, ldq, v, ldv, tau, ldtau, ncontacts, contacts, ps, ddq, ldddq, lambda, ldlambda,
                    iterations, flags);
    else
      cdCall<double>(ctx, m, B, q, ldq, v, ldv, tau, ldtau, ncontacts, contacts, ps, ddq, ldddq, lambda, ldlambda,
                     iterations, flags);
  });
}

int rbd_delassus(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq, int ncontacts,
                 const rbd_rigid_constraint* contacts, double mu, void* G, int64_t ldG, uint32_t flags) {
  return guarded([&] {
    checkConstraintSet(ctx, m, B, dtype, ncontacts, contacts, mu, "Delassus matrix");
    RBD_CHECK_ARG(B == 0 || (q && G), "null array argument");
    if (dtype == RBD_F32)
      constraintMatrixCall<float>(ctx, m, B, q, ldq, ncontacts, contacts, 2, mu, 1.0, nullptr, 0, G, ldG, flags);
    else
      constraintMatrixCall<double>(ctx, m, B, q, ldq, ncontacts, contacts, 2, mu, 1.0, nullptr, 0, G, ldG, flags);
  });
}

int rbd_damped_delassus_inverse(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                                int ncontacts, const rbd_rigid_constraint* contacts, double mu, int scaled,
                                void* Ginv, int64_t ldG, uint32_t flags) {
  return guarded([&] {
    checkConstraintSet(ctx, m, B, dtype, ncontacts, contacts, mu, "damped Delassus inverse");
    RBD_CHECK_ARG(mu > 0.0, "mu must be strictly positive");
    RBD_CHECK_ARG(B == 0 || (q && Ginv), "null array argument");
    const double scale = scaled ? mu : 1.0;
    if (dtype == RBD_F32)
      constraintMatrixCall<float>(ctx, m, B, q, ldq, ncontacts, contacts, 4, mu, scale, nullptr, 0, Ginv, ldG, flags);
    else
      constraintMatrixCall<double>(ctx, m, B, q, ldq, ncontacts, contacts, 4, mu, scale, nullptr, 0, Ginv, ldG,
                                   flags);
  });
}

int rbd_contact_cholesky_solve(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q, int64_t ldq,
                               int ncontacts, const rbd_rigid_constraint* contacts, double mu, const void* rhs,
                               int64_t ldrhs, void* x, int64_t ldx, uint32_t flags) {
  return guarded([&] {
    checkConstraintSet(ctx, m, B, dtype, ncontacts, contacts, mu, "contact Cholesky solve");
    RBD_CHECK_ARG(B == 0 || (q && rhs && x), "null array argument");
    if (dtype == RBD_F32)
      constraintMatrixCall<float>(ctx, m, B, q, ldq, ncontacts, contacts, 3, mu, 1.0, rhs, ldrhs, x, ldx, flags);
    else
      constraintMatrixCall<double>(ctx, m, B, q, ldq, ncontacts, contacts, 3, mu, 1.0, rhs, ldrhs, x, ldx, flags);
  });
}

int rbd_pgs_solve(rbd_ctx* ctx, int64_t B, int ncones, const double* G, int64_t ldG, const double* g, int64_t ldg,
                  const double* mu, int64_t ldmu, double* x, int64_t ldx, const rbd_pgs_settings* settings,
                  double* stats, int64_t ldstats, uint32_t flags) {
  return guarded([&] {
    checkNcp(ctx, B, ncones);
    rbd_pgs_settings ps{1000, 1e-6, 1e-6, 1.0};
    if (settings) ps = *settings;
    RBD_CHECK_ARG(ps.over_relax < 2.0 && ps.over_relax > 0.0, "over_relax should lie in ]0,2[.");
    RBD_CHECK_ARG(ps.max_it > 0, "max_it should be greater than 0.");
    RBD_CHECK_ARG(ps.absolute_precision >= 0.0 && ps.relative_precision >= 0.0, "precisions should be positive.");
    RBD_CHECK_ARG(B == 0 || (G && g && mu && x && stats), "null array argument");
    const int n = 3 * ncones;
    auto vG = argView<double>(n * n, B, ldG, flags, "G");
    auto vg = argView<double>(n, B, ldg, flags, "g");
    auto vm = argView<double>(ncones, B, ldmu, flags, "mu");
    auto vx = argView<double>(n, B, ldx, flags, "x");
    auto vs = argView<double>(kNcpStats, B, ldstats, flags, "stats");
    if (B == 0) return;
    setDevice(ctx);
    Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
    const double* dG = (const double*)st.in(G, vG.extent * 8);
    const double* dg = (const double*)st.in(g, vg.extent * 8);
    const double* dm = (const double*)st.in(mu, vm.extent * 8);
    double* dx = (double*)st.out(x, vx.extent * 8, true);
    double* ds = (double*)st.out(stats, vs.extent * 8, true);
    launchPgs(ctx->stream, ncones, B, CView<double>{dG, vG.rs, vG.bs}, CView<double>{dg, vg.rs, vg.bs},
              CView<double>{dm, vm.rs, vm.bs}, View<double>{dx, vx.rs, vx.bs}, ps.over_relax, ps.max_it,
              ps.absolute_precision, ps.relative_precision, View<double>{ds, vs.rs, vs.bs});
    finishCall(ctx, flags, st);
  });
}

int rbd_admm_solve(rbd_ctx* ctx, int64_t B, int ncones, const double* G, int64_t ldG, const double* g, int64_t ldg,
                   const double* mu, int64_t ldmu, const double* R, int64_t ldR, const double* primal_guess,
                   const double* dual_guess, int64_t ldguess, const rbd_admm_settings* settings, double* y, double* z,
                   double* s, int64_t ldout, double* stats, int64_t ldstats, uint32_t flags) {
  return guarded([&] {
    checkNcp(ctx, B, ncones);
    rbd_admm_settings as{1e-6, 0.5, 0.2, 0.05, 10.0, 20, 1000, 1e-6, 1e-6};
    if (settings) as = *settings;
    RBD_CHECK_ARG(as.tau <= 1.0 && as.tau > 0.0, "tau should lie in ]0,1].");
    RBD_CHECK_ARG(as.mu_prox >= 0.0, "mu_prox should be positive.");
    RBD_CHECK_ARG(as.ratio_primal_dual > 0.0, "The ratio primal/dual should be positive strictly");
    RBD_CHECK_ARG(as.max_it > 0 && as.max_it_largest_eigen_value >= 1, "max_it should be greater than 0.");
    RBD_CHECK_ARG(B == 0 || (G && g && mu && R && y && z && stats), "null array argument");
    const int n = 3 * ncones;
    auto vG = argView<double>(n * n, B, ldG, flags, "G");
    auto vg = argView<double>(n, B, ldg, flags, "g");
    auto vm = argView<double>(ncones, B, ldmu, flags, "mu");
    auto vR = argView<double>(n, B, ldR, flags, "R");
    auto vu = argView<double>(n, B, ldguess, flags, "guess");
    auto vo = argView<double>(n, B, ldout, flags, "outputs");
    auto vs = argView<double>(kNcpStats, B, ldstats, flags, "stats");
    if (B == 0) return;
    setDevice(ctx);
    Staged st{ctx, (flags & RBD_HOST_PTRS) != 0};
    const double* dG = (const double*)st.in(G, vG.extent * 8);
    const double* dg = (const double*)st.in(g, vg.extent * 8);
    const double* dm = (const double*)st.in(mu, vm.extent * 8);
    const double* dR = (const double*)st.in(R, vR.extent * 8);
    const double* dpg = (const double*)st.in(primal_guess, vu.extent * 8);
    const double* ddg = (const double*)st.in(dual_guess, vu.extent * 8);
    double* dy = (double*)st.out(y, vo.extent * 8, true);
    double* dz = (double*)st.out(z, vo.extent * 8, true);
    double* dsh = (double*)st.out(s, vo.extent * 8, true);
    double* ds = (double*)st.out(stats, vs.extent * 8, true);
    const AdmmParams P{as.mu_prox, as.tau, as.rho_power, as.rho_power_factor, as.ratio_primal_dual,
                       as.absolute_precision, as.relative_precision, as.max_it, as.max_it_largest_eigen_value};
    launchAdmm(ctx->stream, ncones, B, CView<double>{dG, vG.rs, vG.bs}, CView<double>{dg, vg.rs, vg.bs},
               CView<double>{dm, vm.rs, vm.bs}, CView<double>{dR, vR.rs, vR.bs},
               CView<double>{dpg, vu.rs, vu.bs}, CView<double>{ddg, vu.rs, vu.bs}, P, View<double>{dy, vo.rs, vo.bs},
               View<double>{dz, vo.rs, vo.bs}, View<double>{dsh, vo.rs, vo.bs}, View<double>{ds, vs.rs, vs.bs});
    finishCall(ctx, flags, st);
  });
}

int rbd_constrained_forward_dynamics(rbd_ctx* ctx, const rbd_model* m, int64_t B, int dtype, const void* q,
                                     int64_t ldq, const void* v, int64_t ldv, const void* tau, int64_t ldtau, int nc,
                                     const void* J, int64_t ldJ, const void* gamma, int64_t ldgamma,
                                     double inv_damping, void* ddq, int64_t ldddq, void* lambda, int64_t ldlambda,
                                     uint32_t flags) {
  return guarded([&] {
    checkCommon(ctx, m, B, dtype);
    RBD_CHECK_ARG(nc >= 1 && nc <= kKktMaxConstraints, "the constraint count must be between 1 and 12");
    RBD_CHECK_ARG(B == 0 || (q && v && tau && J && gamma && ddq && lambda), "null array argument");
    if (!m->compact || !m->ops->kkt_f32 || !m->ops->crba_f32)
      throw Error(RBD_EUNSUPPORTED, "constrained forward dynamics: no kernels for this topology");
    if (dtype == RBD_F32)
      cfdCall<float>(ctx, m, B, q, ldq, v, ldv, tau, ldtau, nc, J, ldJ, gamma, ldgamma, inv_damping, ddq, ldddq,
                     lambda, ldlambda, flags);
    else
      cfdCall<double>(ctx, m, B, q, ldq, v, ldv, tau, ldtau, nc, J, ldJ, gamma, ldgamma, inv_damping, ddq, ldddq,
                      lambda, ldlambda, flags);
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Mixed-robot batches: plan (tile table + strategy), autotuner, graph replay
// ---------------------------------------------------------------------------------------
namespace {

struct MultiPlan {
  std::vector<unsigned char> key;  // the caller's group array byte for byte, then model serials
  std::vector<const rbd_model*> models;
  const MixOps* mix = nullptr;     // null: per-model launches only
  bool fused = true;
  bool concurrent = false;         // per-model launches: one stream per distinct model
  std::vector<int> width;          // chunks of 64 instances per workgroup, per group
  MultiGroup* dGroups = nullptr;
  MultiTile* dTiles = nullptr;
  int ntiles = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string report;
};

void releasePlan(MultiPlan& p) {
  if (p.exec) (void)hipGraphExecDestroy(p.exec);
  if (p.graph) (void)hipGraphDestroy(p.graph);
  if (p.dGroups) (void)hipFree(p.dGroups);
  if (p.dTiles) (void)hipFree(p.dTiles);
  p.exec = nullptr;
  p.graph = nullptr;
  p.dGroups = nullptr;
  p.dTiles = nullptr;
}

struct MultiCache {
  std::vector<std::unique_ptr<MultiPlan>> plans;
  std::string lastReport;
  ~MultiCache() {
    for (auto& p : plans) releasePlan(*p);
  }
};

std::mutex g_multiMutex;
std::vector<std::pair<rbd_ctx*, std::unique_ptr<MultiCache>>> g_multi;

MultiCache& cacheFor(rbd_ctx* ctx) {
  for (auto& e : g_multi)
    if (e.first == ctx) return *e.second;
  g_multi.emplace_back(ctx, std::make_unique<MultiCache>());
  return *g_multi.back().second;
}

void dropMultiPlans(rbd_ctx* ctx, const rbd_model* model) {
  std::lock_guard<std::mutex> lock(g_multiMutex);
  for (size_t e = 0; e < g_multi.size(); ++e) {
    if (g_multi[e].first != ctx) continue;
    if (!model) {
      g_multi.erase(g_multi.begin() + e);
      return;
    }
    auto& plans = g_multi[e].second->plans;
    for (size_t k = 0; k < plans.size();) {
      const auto& ms = plans[k]->models;
      if (std::find(ms.begin(), ms.end(), model) != ms.end()) {
        releasePlan(*plans[k]);
        plans.erase(plans.begin() + k);
      } else {
        ++k;
      }
    }
    return;
  }
}

const MixOps* findMix(const std::vector<const TopoOps*>& need) {
#define RBD_REGISTER_MIX(NAME) rbd::mix_##NAME(),
  static const MixOps* const mixes[] = {
#include "_gen/mixes.inc"
  };
#undef RBD_REGISTER_MIX
  for (const MixOps* mx : mixes) {
    bool all = true;
    for (const TopoOps* o : need) {
      bool found = false;
      for (int k = 0; k < mx->ntopo; ++k) found = found || mx->topo[k] == o;
      all = all && found;
    }
    if (all) return mx;
  }
  return nullptr;
}

// per-instance cost proxy of RNEA + ABA: linear in the number of joints
double costOf(const rbd_model* m) { return (double)m->njoints; }

void buildTiles(MultiPlan& p, const rbd_batch_group* g, int ng, std::vector<MultiTile>& tiles) {
  struct T {
    double cost;
    MultiTile t;
  };
  std::vector<T> all;
  for (int k = 0; k < ng; ++k) {
    const int64_t chunks = (g[k].count + kTile - 1) / kTile;
    const int w = std::max(1, p.width[k]);
    for (int64_t c = 0; c < chunks; c += w) {
      const int n = (int)std::min<int64_t>(w, chunks - c);
      all.push_back({costOf(g[k].model) * n, MultiTile{k, (int32_t)c, n, 0}});
    }
  }
  std::stable_sort(all.begin(), all.end(), [](const T& a, const T& b) { return a.cost > b.cost; });
  tiles.clear();
  for (auto& e : all) tiles.push_back(e.t);
}

void enqueue(rbd_ctx* ctx, MultiPlan& p, const rbd_batch_group* g, int ng) {
  if (p.fused) {
    if (p.ntiles > 0) p.mix->launch(ctx->stream, (unsigned)p.ntiles, p.dGroups, p.dTiles);
    return;
  }
  // per-model launches; concurrent: the groups of each distinct model on a side stream of
  // their own, forked from and joined back into the context's stream with events (a
  // small group's kernels fill few CUs, so the groups' kernels overlap; groups of one
  // model share its run-time-topology lane memory and stay in order on one stream)
  std::vector<const rbd_model*> distinct;
  auto streamOf = [&](const rbd_model* m) {
    if (!p.concurrent) return ctx->stream;
    int k = 0;
    while (k < (int)distinct.size() && distinct[k] != m) ++k;
    if (k == (int)distinct.size()) distinct.push_back(m);
    return ctx->sideStream(k);
  };
  if (p.concurrent) {
    HIP_CHECK(hipEventRecord(ctx->event(0), ctx->stream));
    for (int k = 0; k < ng; ++k) (void)streamOf(g[k].model);
    for (size_t s = 0; s < distinct.size(); ++s) HIP_CHECK(hipStreamWaitEvent(ctx->sideStream((int)s), ctx->event(0), 0));
  }
  for (int k = 0; k < ng; ++k) {
    const rbd_model* m = g[k].model;
    const int64_t n = g[k].count;
    if (n == 0) continue;
    hipStream_t st = streamOf(m);
    m->ops->rnea_f32(st, m->gen.get(), m->mc32, n, CView<float>{g[k].q, n, 1}, CView<float>{g[k].v, n, 1},
                     CView<float>{g[k].a, n, 1}, View<float>{g[k].tau_out, n, 1});
    m->ops->aba_f32(st, m->gen.get(), m->mc32, n, CView<float>{g[k].q, n, 1}, CView<float>{g[k].v, n, 1},
                    CView<float>{g[k].tau, n, 1}, View<float>{g[k].ddq_out, n, 1});
  }
  if (p.concurrent)
    for (size_t s = 0; s < distinct.size(); ++s) {
      HIP_CHECK(hipEventRecord(ctx->event(1 + (int)s), ctx->sideStream((int)s)));
      HIP_CHECK(hipStreamWaitEvent(ctx->stream, ctx->event(1 + (int)s), 0));
    }
}

void uploadTiles(MultiPlan& p, const rbd_batch_group* g, int ng) {
  std::vector<MultiTile> tiles;
  buildTiles(p, g, ng, tiles);
  if (p.dTiles) HIP_CHECK(hipFree(p.dTiles));
  p.dTiles = nullptr;
  p.ntiles = (int)tiles.size();
  if (!tiles.empty()) {
    HIP_CHECK(hipMalloc(&p.dTiles, tiles.size() * sizeof(MultiTile)));
    HIP_CHECK(hipMemcpy(p.dTiles, tiles.data(), tiles.size() * sizeof(MultiTile), hipMemcpyHostToDevice));
  }
}

float timeMs(rbd_ctx* ctx, MultiPlan& p, const rbd_batch_group* g, int ng, int reps) {
  hipEvent_t a, b;
  HIP_CHECK(hipEventCreate(&a));
  HIP_CHECK(hipEventCreate(&b));
  enqueue(ctx, p, g, ng);  // warm
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    HIP_CHECK(hipEventRecord(a, ctx->stream));
    enqueue(ctx, p, g, ng);
    HIP_CHECK(hipEventRecord(b, ctx->stream));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, ms);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

// Coordinate search over per-group tile widths {1, 2, 4, 8} for the fused launch (most
// expensive group first), then the per-model launch variant; the fastest wins.
void autotune(rbd_ctx* ctx, MultiPlan& p, const rbd_batch_group* g, int ng, bool tune) {
  std::string rep = "{\"candidates\": [";
  bool first = true;
  auto note = [&](const std::string& desc, float ms) {
    rep += std::string(first ? "" : ", ") + "{\"variant\": \"" + desc + "\", \"ms\": " + std::to_string(ms) + "}";
    first = false;
  };
  auto widthStr = [&](const std::vector<int>& w) {
    std::string s;
    for (int k = 0; k < ng; ++k) s += std::string(k ? "," : "") + g[k].model->ops->name + "=" + std::to_string(w[k]);
    return s;
  };
  p.width.assign(ng, 1);
  float best = 1e30f;
  bool bestFused = p.mix != nullptr;
  std::vector<int> bestW = p.width;
  if (p.mix) {
    p.fused = true;
    uploadTiles(p, g, ng);
    if (tune) {
      best = timeMs(ctx, p, g, ng, 3);
      note("fused " + widthStr(p.width), best);
      std::vector<int> order(ng);
      for (int k = 0; k < ng; ++k) order[k] = k;
      std::sort(order.begin(), order.end(), [&](int x, int y) {
        return costOf(g[x].model) * g[x].count > costOf(g[y].model) * g[y].count;
      });
      for (int k : order) {
        for (int w : {2, 4, 8}) {
          std::vector<int> trial = bestW;
          trial[k] = w;
          p.width = trial;
          uploadTiles(p, g, ng);
          const float ms = timeMs(ctx, p, g, ng, 3);
          note("fused " + widthStr(trial), ms);
          if (ms < best) {
            best = ms;
            bestW = trial;
          }
        }
      }
    }
  }
  bool bestConcurrent = false;
  if (tune || !p.mix) {
    p.fused = false;
    for (bool conc : {false, true}) {
      if (!tune && conc) break;
      p.concurrent = conc;
      const float ms = tune ? timeMs(ctx, p, g, ng, 3) : 0.f;
      if (tune) note(conc ? "concurrent per-model launches" : "per-model launches", ms);
      if ((!p.mix && !conc) || ms < best) {
        best = ms;
        bestFused = false;
        bestConcurrent = conc;
      }
    }
  }
  p.fused = bestFused;
  p.concurrent = !bestFused && bestConcurrent;
  p.width = bestW;
  if (p.fused) uploadTiles(p, g, ng);
  const std::string chosen = p.fused ? "fused " + widthStr(p.width)
                             : p.concurrent ? "concurrent per-model launches" : "per-model launches";
  rep += "], \"chosen\": \"" + chosen +
         "\", \"mix\": \"" + (p.mix ? p.mix->name : "none") + "\", \"tiles\": " + std::to_string(p.ntiles) + "}";
  p.report = rep;
}

}  // namespace

extern "C" {

int rbd_multi_rnea_aba(rbd_ctx* ctx, int32_t ngroups, const rbd_batch_group* groups, uint32_t flags) {
  return guarded([&] {
    RBD_CHECK_ARG(ctx != nullptr, "null context");
    RBD_CHECK_ARG(!ctx->capturing, "capture: mixed batches replay their own graph and cannot be recorded");
    RBD_CHECK_ARG(ngroups >= 1 && ngroups <= 8 && groups, "between 1 and 8 groups are supported");
    std::vector<const TopoOps*> need;
    for (int k = 0; k < ngroups; ++k) {
      const rbd_batch_group& g = groups[k];
      RBD_CHECK_ARG(g.model && g.model->ctx == ctx, "group model missing or from another context");
      RBD_CHECK_ARG(g.count >= 0, "negative group size");
      if (g.count > 0)
        RBD_CHECK_ARG(g.q && g.v && g.a && g.tau && g.tau_out && g.ddq_out, "null group array");
      if (!g.model->ops->rnea_f32 || !g.model->ops->aba_f32)
        throw Error(RBD_EUNSUPPORTED, "no RNEA/ABA kernels for a group's topology");
      // the autotuner runs the batch several times: outputs must not overwrite inputs
      for (const void* out : {(const void*)g.tau_out, (const void*)g.ddq_out})
        for (const void* in : {(const void*)g.q, (const void*)g.v, (const void*)g.a, (const void*)g.tau})
          RBD_CHECK_ARG(g.count == 0 || out != in, "group outputs must not alias its inputs");
      need.push_back(g.model->prebuilt);
    }
    setDevice(ctx);
    std::lock_guard<std::mutex> lock(g_multiMutex);
    MultiCache& cache = cacheFor(ctx);
    const unsigned char* kb = reinterpret_cast<const unsigned char*>(groups);
    std::vector<unsigned char> key(kb, kb + sizeof(rbd_batch_group) * ngroups);
    for (int k = 0; k < ngroups; ++k) {
      const unsigned char* sb = reinterpret_cast<const unsigned char*>(&groups[k].model->serial);
      key.insert(key.end(), sb, sb + sizeof(uint64_t));
    }
    MultiPlan* plan = nullptr;
    for (auto& p : cache.plans)
      if (p->key == key) plan = p.get();
    if (!plan) {
      if (cache.plans.size() >= 8) {
        releasePlan(*cache.plans.front());
        cache.plans.erase(cache.plans.begin());
      }
      auto p = std::make_unique<MultiPlan>();
      p->key = key;
      for (int k = 0; k < ngroups; ++k) p->models.push_back(groups[k].model);
      p->mix = findMix(need);
      if (p->mix) {
        std::vector<MultiGroup> hg(ngroups);
        for (int k = 0; k < ngroups; ++k) {
          const rbd_batch_group& g = groups[k];
          int topo = 0;
          for (int t = 0; t < p->mix->ntopo; ++t)
            if (p->mix->topo[t] == g.model->prebuilt) topo = t;
          hg[k] = MultiGroup{g.model->mc32, g.count, g.q, g.v, g.a, g.tau, g.tau_out, g.ddq_out, topo, 0};
        }
        HIP_CHECK(hipMalloc(&p->dGroups, sizeof(MultiGroup) * ngroups));
        HIP_CHECK(hipMemcpy(p->dGroups, hg.data(), sizeof(MultiGroup) * ngroups, hipMemcpyHostToDevice));
      }
      // run-time-topology kernels size their lane memory now: nothing may allocate
      // inside the graph capture below
      for (int k = 0; k < ngroups; ++k)
        if (!groups[k].model->prebuilt) genericReserveDyn(groups[k].model->gen.get(), groups[k].count);
      autotune(ctx, *p, groups, ngroups, !(flags & RBD_NO_AUTOTUNE));
      if (!(flags & RBD_NO_GRAPH)) {
        HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueue(ctx, *p, groups, ngroups);
        HIP_CHECK(hipStreamEndCapture(ctx->stream, &p->graph));
        HIP_CHECK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
      }
      cache.plans.push_back(std::move(p));
      plan = cache.plans.back().get();
    }
    cache.lastReport = plan->report;
    if (plan->exec && !(flags & RBD_NO_GRAPH)) HIP_CHECK(hipGraphLaunch(plan->exec, ctx->stream));
    else enqueue(ctx, *plan, groups, ngroups);
    HIP_CHECK(hipGetLastError());
    if (!(flags & RBD_ASYNC)) HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int rbd_multi_tuning(rbd_ctx* ctx, char* buf, int32_t len) {
  return guarded([&] {
    RBD_CHECK_ARG(ctx && buf && len > 0, "bad argument");
    std::lock_guard<std::mutex> lock(g_multiMutex);
    const std::string& s = cacheFor(ctx).lastReport;
    const size_t n = std::min<size_t>(s.size(), (size_t)len - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  });
}

int rbd_model_cholesky_pattern(const rbd_model* m, int32_t* nnz, int32_t* rows, int32_t* cols) {
  return guarded([&] {
    RBD_CHECK_ARG(m && nnz, "null argument");
    *nnz = (int32_t)m->patRows.size();
    if (rows) std::memcpy(rows, m->patRows.data(), m->patRows.size() * sizeof(int32_t));
    if (cols) std::memcpy(cols, m->patCols.data(), m->patCols.size() * sizeof(int32_t));
  });
}

}  // extern "C"
