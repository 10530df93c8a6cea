// generic.hpp -- run-time topology program of one model and the launcher table of the
// run-time-topology kernels (k_generic.hpp).  Host side; internal to the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "ops.hpp"

namespace rbd {

// The tree of one model as device tables: DFS event schedule, joint records, packed
// tree pattern of M and its structural non-zero masks (k_generic.hpp GProg).  Built for
// every model at upload; its launchers serve every tree that has no prebuilt kernels and
// every entry a prebuilt topology lacks (FP64 ABA, FP64 derivatives).
struct GenericModel {
  int nj = 0, nq = 0, nv = 0, nev = 0, nlev = 0, nvlev = 0, nnz = 0, nzW = 0;
  int32_t* dtab = nullptr;  // one device allocation holding the tables below
  const int32_t* ev = nullptr;
  const int32_t* jr = nullptr;
  const int32_t* colPtr = nullptr;
  const int32_t* patRow = nullptr;
  const uint32_t* nzMask = nullptr;
  // Lane memory of the kernels whose per-lane state does not fit LDS, grown on demand.
  // A buffer that is outgrown stays allocated until the model is destroyed: graphs
  // captured by the mixed-batch launcher may still reference it.
  mutable void* lanes = nullptr;
  mutable size_t laneBytes = 0;
  mutable std::vector<void*> retired;
  void* laneMem(size_t bytes) const;

  GenericModel() = default;
  GenericModel(const GenericModel&) = delete;
  GenericModel& operator=(const GenericModel&) = delete;
  ~GenericModel();
};

// Builds the program of the tree (parents[j] < j, joint types) on the current device.
GenericModel* buildGenericModel(int nj, const int* parents, const int* types);

// Launchers of the run-time-topology kernels; every entry is set (name "generic").
const TopoOps* genericOps();

// RNEA / ABA with external forces (fext: 6 rows per joint 1.., [linear; angular] in the
// joint's frame), on the run-time-topology kernels for every tree
template<class S>
void launchRneaFext(hipStream_t st, const GenericModel* g, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> a,
                    CView<S> fext, View<S> out);
template<class S>
void launchAbaFext(hipStream_t st, const GenericModel* g, const S* mc, int64_t B, CView<S> q, CView<S> v,
                   CView<S> tau, CView<S> fext, View<S> out);

// In place on x: the products and triangular solves of the packed tree factor (cholops.cpp;
// op: RBD_CHOL_UV, _UTV, _UIV, _UTIV, _MV).
void launchCholApply(hipStream_t st, const GenericModel* g, int op, int64_t B, CView<double> U, CView<double> D,
                     View<double> x);

// Grows the lane memory of a model so that its RNEA and ABA launches over `count`
// configurations allocate nothing (called before a HIP graph capture).
void genericReserveDyn(const GenericModel* g, int64_t count);

}  // namespace rbd
