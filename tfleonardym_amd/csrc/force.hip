// Fermion force kernel: out_mu(x) (+)= sum_p w_p [L(X_p, Y_p) + sym L(Y_p, X_p)]
// with L the Lie derivative of the Wilson operator (dirac.py:525-565), the
// sum n_flavor.force forms over rational-approximation terms
// (fermion_action.py:86-91).
//
// One thread per (site, direction).  A block holds kForceSites sites and the
// D directions as D groups of whole wavefronts (direction uniform per wave):
// the D waves read the same X(x), Y(x) back to back, so those hit in cache
// after the first, and each thread reads its link U_mu(x) once for all pairs.
// Site blocks are contiguous per XCD, so the neighbour x+mu of most sites was
// fetched into the same XCD's L2 by a nearby block.  Bound: HBM.
#include "force_site.h"
#include "lym_internal.h"

namespace lym {

constexpr int kForceSites = 64;

template <int D, int NC, bool ADJ>
__global__ __launch_bounds__(kForceSites * 4) void k_force(Lat lat, int gs, LinkView<double> lk,
                                                           LinkView<double> lkr, ForcePairs fp, uint32_t vbytes,
                                                           cpx<double>* out, uint32_t obytes) {
  const int mu = (int)threadIdx.x / kForceSites;
  const int tid = (int)threadIdx.x % kForceSites;
  int64_t s;
  int x[D];
  const SiteMap linear{0, 0, 0, 0, 0};
  if (!site_of<D>(lat, linear, (int)blockIdx.x, gs, tid, kForceSites, s, x)) return;
  force_site<D, NC, ADJ>(lat, mu, s, x, lk, lkr, fp, vbytes, out, obytes);
}

template <int D, int NC, bool ADJ>
static int launch_force(const OpDesc& d, const ForcePairs& fp, cpx<double>* out, hipStream_t st) {
  const int64_t V = d.lat.V;
  const int gs = grid_for(V, kForceSites);
  const int recon = d.link_su ? 1 : 0;
  LinkView<double> lk{(const cpx<double>*)d.links64, V, recon};
  LinkView<double> lkr{(const cpx<double>*)(d.linksr64 ? d.linksr64 : d.links64), V, recon};
  const uint32_t vbytes = (uint32_t)(field_elems(d) * 16);
  const uint32_t obytes = (uint32_t)((int64_t)D * NC * NC * V * 16);
  hipLaunchKernelGGL((k_force<D, NC, ADJ>), dim3(gs), dim3(kForceSites * D), 0, st, d.lat, gs, lk, lkr, fp, vbytes,
                     out, obytes);
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

int force_launch(const OpDesc& d, const ForcePairs& fp, void* out, hipStream_t st) {
  cpx<double>* o = (cpx<double>*)out;
  const int key = d.ndim * 100 + d.nc * 10 + d.adj;
  switch (key) {
    case 420: return launch_force<4, 2, false>(d, fp, o, st);
    case 430: return launch_force<4, 3, false>(d, fp, o, st);
    case 220: return launch_force<2, 2, false>(d, fp, o, st);
    case 230: return launch_force<2, 3, false>(d, fp, o, st);
    case 421: return launch_force<4, 2, true>(d, fp, o, st);
    case 431: return launch_force<4, 3, true>(d, fp, o, st);
    case 221: return launch_force<2, 2, true>(d, fp, o, st);
    case 231: return launch_force<2, 3, true>(d, fp, o, st);
  }
  set_error("force: unsupported (ndim, nc, representation)");
  return LYM_ERR_ARG;
}

}  // namespace lym
