// ncp.cpp -- launchers of the batched frictional contact solvers (k_ncp.hpp): one kernel
// instantiation per cone count.
#include "k_ncp.hpp"
#include "ops.hpp"

namespace rbd {

template<int NC> static void pgsN(hipStream_t st, int64_t B, CView<double> G, CView<double> g, CView<double> mu,
                                  View<double> x, double over_relax, int max_it, double abs_prec, double rel_prec,
                                  View<double> stats) {
  constexpr int perBlock = 256 / kPgsLanes;
  hipLaunchKernelGGL(k_pgs<NC>, dim3((unsigned)((B + perBlock - 1) / perBlock)), dim3(256), 0, st, B, G, g, mu, x,
                     over_relax, max_it, abs_prec, rel_prec, stats);
}

void launchPgs(hipStream_t st, int nc, int64_t B, CView<double> G, CView<double> g, CView<double> mu, View<double> x,
               double over_relax, int max_it, double abs_prec, double rel_prec, View<double> stats) {
  switch (nc) {
    case 1: pgsN<1>(st, B, G, g, mu, x, over_relax, max_it, abs_prec, rel_prec, stats); break;
    case 2: pgsN<2>(st, B, G, g, mu, x, over_relax, max_it, abs_prec, rel_prec, stats); break;
    case 3: pgsN<3>(st, B, G, g, mu, x, over_relax, max_it, abs_prec, rel_prec, stats); break;
    case 4: pgsN<4>(st, B, G, g, mu, x, over_relax, max_it, abs_prec, rel_prec, stats); break;
  }
}

template<int NC> static void admmN(hipStream_t st, int64_t B, CView<double> G, CView<double> g, CView<double> mu,
                                   CView<double> R, CView<double> pg, CView<double> dg, const AdmmParams& P,
                                   View<double> y, View<double> z, View<double> s, View<double> stats) {
  constexpr int perBlock = 256 / kAdmmLanes;
  hipLaunchKernelGGL(k_admm<NC>, dim3((unsigned)((B + perBlock - 1) / perBlock)), dim3(256), 0, st, B, G, g, mu, R,
                     pg, dg, P, y, z, s, stats);
}

void launchAdmm(hipStream_t st, int nc, int64_t B, CView<double> G, CView<double> g, CView<double> mu, CView<double> R,
                CView<double> pg, CView<double> dg, const AdmmParams& P, View<double> y, View<double> z,
                View<double> s, View<double> stats) {
  switch (nc) {
    case 1: admmN<1>(st, B, G, g, mu, R, pg, dg, P, y, z, s, stats); break;
    case 2: admmN<2>(st, B, G, g, mu, R, pg, dg, P, y, z, s, stats); break;
    case 3: admmN<3>(st, B, G, g, mu, R, pg, dg, P, y, z, s, stats); break;
    case 4: admmN<4>(st, B, G, g, mu, R, pg, dg, P, y, z, s, stats); break;
  }
}

}  // namespace rbd
