// Fermion force kernel: out_mu(x) (+)= sum_p w_p [L(X_p, Y_p) + sym L(Y_p, X_p)]
// with L the Lie derivative of the Wilson operator (dirac.py:525-565), the
// sum n_flavor.force forms over rational-approximation terms
// (fermion_action.py:86-91).
//
// One thread per (site, direction).  A block holds its sites once per
// direction, as D groups of whole wavefronts (direction uniform per wave):
// the D waves read the same X(x), Y(x) back to back, so those hit in cache
// after the first.  Fundamental representation: the pair loop accumulates
// colour outer products only; the link U_mu(x) is read and applied once per
// thread at the end.  Bound: HBM.
#include "force_site.h"
#include "lym_internal.h"

namespace lym {

constexpr int kForceSites = 64;

// TILED: the Wilson kernels' XCD-tiled schedule (4D): a block covers k whole
// lines of the fastest direction, XCD x owns an x0 slab and runs its blocks
// x0 fastest, so the x+mu neighbours of a block were fetched into the same
// L2 by the block before.  Otherwise kForceSites consecutive sites per block,
// block ranges contiguous per XCD.
template <int D, int NC, bool ADJ, bool HERM, bool TILED>
__global__ __launch_bounds__(TILED ? 1024 : kForceSites * 4) void k_force(Lat lat, SiteMap map, int nthr, int gs,
                                                                          LinkView<double> lk, LinkView<double> lkr,
                                                                          ForcePairs fp, uint32_t vbytes,
                                                                          cpx<double>* out, uint32_t obytes) {
  const int mu = (int)threadIdx.x / nthr;
  const int tid = (int)threadIdx.x % nthr;
  int64_t s;
  int x[D];
  const SiteMap m = TILED ? map : SiteMap{0, 0, 0, 0, 0};
  if (!site_of<D>(lat, m, (int)blockIdx.x, gs, tid, nthr, s, x)) return;
  force_site<D, NC, ADJ, HERM>(lat, mu, s, x, lk, lkr, fp, vbytes, out, obytes);
}

template <int D, int NC, bool ADJ, bool HERM>
static int launch_force(const OpDesc& d, const ForcePairs& fp, cpx<double>* out, hipStream_t st) {
  const int64_t V = d.lat.V;
  const int recon = d.link_su ? 1 : 0;
  LinkView<double> lk{(const cpx<double>*)d.links64, V, recon};
  LinkView<double> lkr{(const cpx<double>*)(d.linksr64 ? d.linksr64 : d.links64), V, recon};
  const uint32_t vbytes = (uint32_t)(field_elems(d) * 16);
  const uint32_t obytes = (uint32_t)((int64_t)D * NC * NC * V * 16);
  // (the non-hermitian SU(3) form needs more registers than 1024-thread blocks allow)
  if constexpr (D == 4 && !ADJ && (HERM || NC == 2)) {
    if (d.map.tiled && d.wthreads * D <= 1024) {
      hipLaunchKernelGGL((k_force<D, NC, ADJ, HERM, true>), dim3(d.wblocks), dim3(d.wthreads * D), 0, st, d.lat,
                         d.map, d.wthreads, d.wblocks, lk, lkr, fp, vbytes, out, obytes);
      LYM_HIP(hipGetLastError());
      return LYM_OK;
    }
  }
  const int gs = grid_for(V, kForceSites);
  hipLaunchKernelGGL((k_force<D, NC, ADJ, HERM, false>), dim3(gs), dim3(kForceSites * D), 0, st, d.lat, d.map,
                     kForceSites, gs, lk, lkr, fp, vbytes, out, obytes);
  LYM_HIP(hipGetLastError());
  return LYM_OK;
}

template <int D, int NC, bool ADJ>
static int launch_force_h(const OpDesc& d, const ForcePairs& fp, cpx<double>* out, hipStream_t st) {
  if constexpr (ADJ) {
    return launch_force<D, NC, true, false>(d, fp, out, st);  // the adjoint path reads herm at run time
  } else {
    if (d.herm) return launch_force<D, NC, false, true>(d, fp, out, st);
    return launch_force<D, NC, false, false>(d, fp, out, st);
  }
}

int force_launch(const OpDesc& d, const ForcePairs& fp, void* out, hipStream_t st) {
  cpx<double>* o = (cpx<double>*)out;
  const int key = d.ndim * 100 + d.nc * 10 + d.adj;
  switch (key) {
    case 420: return launch_force_h<4, 2, false>(d, fp, o, st);
    case 430: return launch_force_h<4, 3, false>(d, fp, o, st);
    case 220: return launch_force_h<2, 2, false>(d, fp, o, st);
    case 230: return launch_force_h<2, 3, false>(d, fp, o, st);
    case 421: return launch_force_h<4, 2, true>(d, fp, o, st);
    case 431: return launch_force_h<4, 3, true>(d, fp, o, st);
    case 221: return launch_force_h<2, 2, true>(d, fp, o, st);
    case 231: return launch_force_h<2, 3, true>(d, fp, o, st);
  }
  set_error("force: unsupported (ndim, nc, representation)");
  return LYM_ERR_ARG;
}

}  // namespace lym
