// Host build of the per-site Wilson and force arithmetic used by the kernels
// (wilson_site.h, force_site.h), looped over all sites.  Test-only: lets the CPU test suite
// check site indexing, gamma tables, colour transport and the boundary factor
// against the oracle without a GPU.  Not part of liblym.so.
#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "force_site.h"
#include "lym_fin.h"
#include "wilson_site.h"

using namespace lym;

template <int D, int NC, bool ADJ, bool PROJ>
static void run(const Lat& lat, const double* links, const double* links_r, const double* in, double* out) {
  constexpr int NCOMP = Gam<D>::NS * (ADJ ? NC * NC : NC);
  SpinorView<P64> sin{(double2*)in, nullptr, lat.V};
  SpinorView<P64> sout{(double2*)out, nullptr, lat.V};
  LinkView<double> lk{(const cpx<double>*)links, lat.V};
  LinkView<double> lkr{(const cpx<double>*)(links_r ? links_r : links), lat.V};
  for (int64_t s = 0; s < lat.V; ++s) {
    cpx<double> hop[NCOMP], c[NCOMP], res[NCOMP];
    wilson_hop<D, NC, ADJ, PROJ, double>(lat, s, sin, lk, lkr, hop);
    sin.template load<NCOMP>(s, c);
    wilson_finish<NCOMP, double>(lat, c, hop, res);
    sout.template store<NCOMP>(s, res);
  }
}

template <int D, int NC, bool ADJ>
static int run_mode(int proj, const Lat& lat, const double* l, const double* lr, const double* in, double* out) {
  if (proj) run<D, NC, ADJ, true>(lat, l, lr, in, out);
  else run<D, NC, ADJ, false>(lat, l, lr, in, out);
  return 0;
}

static Lat make_lat(int ndim, int herm, const int* dims, double kappa, double bc_re, double bc_im) {
  Lat lat{};
  int64_t V = 1;
  for (int mu = 0; mu < 4; ++mu) lat.L[mu] = 1;
  for (int mu = 0; mu < ndim; ++mu) {
    lat.L[mu] = dims[mu];
    V *= dims[mu];
  }
  int64_t st = 1;
  for (int mu = ndim - 1; mu >= 0; --mu) {
    lat.stride[mu] = (int)st;
    st *= dims[mu];
  }
  lat.V = V;
  lat.kappa = kappa;
  lat.bc_re = bc_re;
  lat.bc_im = bc_im;
  lat.herm = herm;
  lat.has_bc = (bc_re != 1.0 || bc_im != 0.0);
  return lat;
}

template <int D, int NC, bool ADJ, bool HERM>
static int run_force_h(const Lat& lat, const double* links, const double* links_r, const ForcePairs& fp,
                       double* out) {
  LinkView<double> lk{(const cpx<double>*)links, lat.V};
  LinkView<double> lkr{(const cpx<double>*)(links_r ? links_r : links), lat.V};
  for (int64_t s = 0; s < lat.V; ++s) {
    int x[D];
    coords<D>(lat, s, x);
    for (int mu = 0; mu < D; ++mu)
      force_site<D, NC, ADJ, HERM>(lat, mu, s, x, lk, lkr, fp, 0, (cpx<double>*)out, 0);
  }
  return 0;
}

template <int D, int NC, bool ADJ>
static int run_force(const Lat& lat, const double* links, const double* links_r, const ForcePairs& fp,
                     double* out) {
  if (ADJ || !lat.herm) return run_force_h<D, NC, ADJ, false>(lat, links, links_r, fp, out);
  return run_force_h<D, NC, ADJ, true>(lat, links, links_r, fp, out);
}

// out (+)= sum_p w_p [L(x_p, y_p) + sym L(y_p, x_p)], the per-site arithmetic
// of the force kernel looped over all sites and directions
extern "C" int lymcheck_lie_derivative(int ndim, int nc, int adj, int herm, const int* dims, double kappa,
                                       double bc_re, double bc_im, const double* links, const double* links_r,
                                       int npairs, const double* const* xs, const double* const* ys,
                                       const double* weights, int sym, int accumulate, double* out) {
  if (npairs > kForceMaxPairs) return -2;
  const Lat lat = make_lat(ndim, herm, dims, kappa, bc_re, bc_im);
  ForcePairs fp{};
  fp.n = npairs;
  fp.sym = sym;
  fp.accumulate = accumulate;
  for (int p = 0; p < npairs; ++p) {
    fp.x[p] = (const cpx<double>*)xs[p];
    fp.y[p] = (const cpx<double>*)ys[p];
    fp.w[p] = weights[p];
  }
  const int key = ndim * 100 + nc * 10 + adj;
  switch (key) {
    case 420: return run_force<4, 2, false>(lat, links, links_r, fp, out);
    case 430: return run_force<4, 3, false>(lat, links, links_r, fp, out);
    case 220: return run_force<2, 2, false>(lat, links, links_r, fp, out);
    case 230: return run_force<2, 3, false>(lat, links, links_r, fp, out);
    case 421: return run_force<4, 2, true>(lat, links, links_r, fp, out);
    case 431: return run_force<4, 3, true>(lat, links, links_r, fp, out);
    case 221: return run_force<2, 2, true>(lat, links, links_r, fp, out);
    case 231: return run_force<2, 3, true>(lat, links, links_r, fp, out);
  }
  return -1;
}

extern "C" int lymcheck_wilson_apply(int ndim, int nc, int adj, int herm, const int* dims, double kappa, double bc_re,
                                     double bc_im, int proj, const double* links, const double* links_r,
                                     const double* in, double* out) {
  Lat lat{};
  int64_t V = 1;
  for (int mu = 0; mu < 4; ++mu) lat.L[mu] = 1;
  for (int mu = 0; mu < ndim; ++mu) {
    lat.L[mu] = dims[mu];
    V *= dims[mu];
  }
  int64_t st = 1;
  for (int mu = ndim - 1; mu >= 0; --mu) {
    lat.stride[mu] = (int)st;
    st *= dims[mu];
  }
  lat.V = V;
  lat.kappa = kappa;
  lat.bc_re = bc_re;
  lat.bc_im = bc_im;
  lat.herm = herm;
  lat.has_bc = (bc_re != 1.0 || bc_im != 0.0);
  const int key = ndim * 100 + nc * 10 + adj;
  switch (key) {
    case 420: return run_mode<4, 2, false>(proj, lat, links, links_r, in, out);
    case 430: return run_mode<4, 3, false>(proj, lat, links, links_r, in, out);
    case 220: return run_mode<2, 2, false>(proj, lat, links, links_r, in, out);
    case 230: return run_mode<2, 3, false>(proj, lat, links, links_r, in, out);
    case 421: return run_mode<4, 2, true>(proj, lat, links, links_r, in, out);
    case 431: return run_mode<4, 3, true>(proj, lat, links, links_r, in, out);
    case 221: return run_mode<2, 2, true>(proj, lat, links, links_r, in, out);
    case 231: return run_mode<2, 3, true>(proj, lat, links, links_r, in, out);
  }
  return -1;
}

// 16-bit storage path of the hop (fundamental): the input as int16 pairs with
// one float norm per site (value = int16 * norm), complex64 links and
// arithmetic, complex64 output.  Exercises the neighbour loads that take the
// integers and the norm apart and fold the norm into the accumulating FMAs.
template <int D, int NC, bool PROJ>
static int run_half(const Lat& lat, const float* links, const short* in, const float* norm, float* out) {
  constexpr int NCOMP = Gam<D>::NS * NC;
  SpinorView<P16> sin{(short2*)in, const_cast<float*>(norm), lat.V};
  SpinorView<P32> sout{(float2*)out, nullptr, lat.V};
  // the kernels read complex64 links packed in 16-byte units (lym_pack_links_f32)
  constexpr int QU = link_units32(NC);
  std::vector<float> packed((size_t)D * QU * lat.V * 4, 0.0f);
  for (int mu = 0; mu < D; ++mu)
    for (int e = 0; e < NC * NC; ++e)
      for (int64_t s = 0; s < lat.V; ++s) {
        const size_t dst = (((size_t)mu * QU + e / 2) * lat.V + s) * 4 + (e % 2) * 2;
        const size_t src = (((size_t)mu * NC * NC + e) * lat.V + s) * 2;
        packed[dst] = links[src];
        packed[dst + 1] = links[src + 1];
      }
  LinkView<float> lk{(const cpx<float>*)packed.data(), lat.V};
  for (int64_t s = 0; s < lat.V; ++s) {
    cpx<float> hop[NCOMP], c[NCOMP], res[NCOMP];
    wilson_hop<D, NC, false, PROJ, float>(lat, s, sin, lk, lk, hop);
    sin.template load<NCOMP>(s, c);
    wilson_finish<NCOMP, float>(lat, c, hop, res);
    sout.template store<NCOMP>(s, res);
  }
  return 0;
}

extern "C" int lymcheck_wilson_apply_half(int ndim, int nc, int herm, const int* dims, double kappa, double bc_re,
                                          double bc_im, int proj, const float* links, const short* in,
                                          const float* norm, float* out) {
  const Lat lat = make_lat(ndim, herm, dims, kappa, bc_re, bc_im);
  const int key = ndim * 10 + nc;
  auto go = [&](auto d, auto n) {
    constexpr int DD = decltype(d)::value, NN = decltype(n)::value;
    return proj ? run_half<DD, NN, true>(lat, links, in, norm, out) : run_half<DD, NN, false>(lat, links, in, norm, out);
  };
  switch (key) {
    case 42: return go(std::integral_constant<int, 4>{}, std::integral_constant<int, 2>{});
    case 43: return go(std::integral_constant<int, 4>{}, std::integral_constant<int, 3>{});
    case 22: return go(std::integral_constant<int, 2>{}, std::integral_constant<int, 2>{});
    case 23: return go(std::integral_constant<int, 2>{}, std::integral_constant<int, 3>{});
  }
  return -1;
}

// the buffer-range predicate of the host argument checks (lym_types.h)
// The deferred multishift update as the shift-update kernel applies it
// (fold_shift_block, then x += A p + sum_j d_j r_j, p = B p + sum_j e_j r_j),
// on n complex elements; r holds nb >= mk residual vectors, steps j >= mk
// contribute nothing (a shift that converged inside the block).
extern "C" int lymcheck_fold_shift_block(int mk, int nb, const double* a, const double* z, const double* b, int64_t n,
                                         const double* r, double* x, double* p) {
  if (mk < 0 || nb < mk || nb > kMaxBlock) return 1;
  double2 de[kMaxBlock];
  double A = 0.0, B = 1.0;
  fold_shift_block(a, z, b, mk, de, 1, A, B);
  for (int j = mk; j < nb; ++j) de[j] = make_double2(0.0, 0.0);
  for (int64_t i = 0; i < 2 * n; ++i) {
    double xv = fma(A, p[i], x[i]);
    double pv = p[i] * B;
    for (int j = 0; j < nb; ++j) {
      xv = fma(de[j].x, r[(int64_t)j * 2 * n + i], xv);
      pv = fma(de[j].y, r[(int64_t)j * 2 * n + i], pv);
    }
    x[i] = xv;
    p[i] = pv;
  }
  return 0;
}

extern "C" int cpucheck_range_inside(uint64_t base, uint64_t size, uint64_t p, uint64_t bytes) {
  return range_inside((uintptr_t)base, (size_t)size, (uintptr_t)p, (size_t)bytes) ? 1 : 0;
}
