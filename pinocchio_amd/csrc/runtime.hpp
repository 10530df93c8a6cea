// runtime.hpp -- internal to the library: error handling of the C ABI entry points and
// the context / model objects shared by its translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rbd.h"
#include "generic.hpp"
#include "ops.hpp"

namespace rbd_rt {
extern thread_local std::string g_err;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
#define RBD_CHECK_ARG(cond, msg) \
  do {                           \
    if (!(cond)) throw Error(RBD_EINVAL, msg); \
  } while (0)
#define HIP_CHECK(expr)                                                                            \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) throw Error(RBD_EDEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template<class F> inline int guarded(F&& f) {
  try {
    f();
    return RBD_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory";
    return RBD_ENOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return RBD_EINVAL;
  }
}
}  // namespace rbd_rt
using namespace rbd_rt;

struct rbd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // device staging for RBD_HOST_PTRS calls, grown on demand
  std::vector<void*> stage;
  std::vector<size_t> stageBytes;
  // device workspace for intermediate results (ABA-derivative factor, ddq), grown on demand
  void* work = nullptr;
  size_t workBytes = 0;
  // stream capture (rbd_capture_begin / _end): calls are recorded, not run; graphs that
  // exist may reference the workspace, so a grown workspace retires the old buffer
  // instead of freeing it (freed with the context)
  bool capturing = false;
  int graphs = 0;
  // 64 zero bytes on the device (allocated with the context): a stride-0 view of them is an
  // all-zero batch of any size, for the inputs that nonLinearEffects and
  // computeGeneralizedGravity fix at zero
  void* zeros = nullptr;
  std::vector<void*> retired;
  void* workspace(size_t bytes) {
    if (workBytes < bytes) {
      if (capturing)
        throw Error(RBD_EINVAL, "capture: the call needs more workspace than the calls before the capture "
                                "reserved (run it once, uncaptured, with the same batch first)");
      if (work) {
        if (graphs > 0) retired.push_back(work);
        else HIP_CHECK(hipFree(work));
      }
      work = nullptr;
      HIP_CHECK(hipMalloc(&work, bytes));
      workBytes = bytes;
    }
    return work;
  }
  // side streams and fork / join events for concurrent launches (default priority,
  // non-blocking), created on first use
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> forkJoin;
  hipStream_t sideStream(int k) {
    while ((int)side.size() <= k) {
      hipStream_t s = nullptr;
      HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      side.push_back(s);
    }
    return side[k];
  }
  hipEvent_t event(int k) {
    while ((int)forkJoin.size() <= k) {
      hipEvent_t e = nullptr;
      HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      forkJoin.push_back(e);
    }
    return forkJoin[k];
  }
  void* stageBuf(int slot, size_t bytes) {
    if (capturing)
      throw Error(RBD_EINVAL, "capture: host arrays and host-side tables (frames, constraints) cannot be "
                              "recorded; use device pointers");
    if ((int)stage.size() <= slot) {
      stage.resize(slot + 1, nullptr);
      stageBytes.resize(slot + 1, 0);
    }
    if (stageBytes[slot] < bytes) {
      if (stage[slot]) HIP_CHECK(hipFree(stage[slot]));
      stage[slot] = nullptr;
      HIP_CHECK(hipMalloc(&stage[slot], bytes));
      stageBytes[slot] = bytes;
    }
    return stage[slot];
  }
};

struct rbd_model {
  rbd_ctx* ctx = nullptr;
  uint64_t serial = 0;  // unique per created model: mixed-batch plans key on it
  // kernels: the prebuilt topology's launchers (null when the tree is not prebuilt or the
  // run-time-topology kernels were requested), completed by the run-time-topology
  // launchers in `tab`; `ops` points to `tab`
  const rbd::TopoOps* prebuilt = nullptr;
  rbd::TopoOps tab{};
  const rbd::TopoOps* ops = nullptr;
  std::unique_ptr<rbd::GenericModel> gen;
  int njoints = 0, nq = 0, nv = 0, compact = 0;
  std::vector<int> parents, jtype, iq, iv;
  std::vector<double> gravity;
  float* mc32 = nullptr;
  double* mc64 = nullptr;
  // packed strictly-upper structural non-zeros of M (tree pattern), column by column
  std::vector<int32_t> patRows, patCols;
};

