// Host AddressSanitizer check of the kernels' per-site arithmetic (test-only,
// not part of liblym.so): builds with cpucheck.cpp under -fsanitize=address
// and runs every Wilson and force variant on buffers of exactly the sizes
// the kernels assume, so any read or write past a field (a wrong plane,
// site or neighbour index) aborts with a report.  `short` as the first
// argument makes the input spinor one site short: the expected report
// shows the sanitizer is active.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

extern "C" int lymcheck_wilson_apply(int ndim, int nc, int adj, int herm, const int* dims, double kappa, double bc_re,
                                     double bc_im, int proj, const double* links, const double* links_r,
                                     const double* in, double* out);
extern "C" int lymcheck_lie_derivative(int ndim, int nc, int adj, int herm, const int* dims, double kappa,
                                       double bc_re, double bc_im, const double* links, const double* links_r,
                                       int npairs, const double* const* xs, const double* const* ys,
                                       const double* weights, int sym, int accumulate, double* out);

static std::vector<double> field(size_t n_complex, unsigned seed) {
  std::vector<double> v(2 * n_complex);
  unsigned s = seed * 2654435761u + 1u;
  for (auto& x : v) {
    s = s * 1664525u + 1013904223u;
    x = (double)(s >> 8) / 16777216.0 - 0.5;
  }
  return v;
}

int main(int argc, char** argv) {
  const bool short_input = argc > 1 && std::strcmp(argv[1], "short") == 0;
  const int geoms[][4] = {{4, 6, 1, 1}, {3, 4, 1, 1}, {2, 4, 2, 4}, {4, 2, 3, 4}};
  int cases = 0;
  for (const auto& g : geoms) {
    const int ndim = g[2] == 1 ? 2 : 4;
    const long V = (long)g[0] * g[1] * g[2] * g[3];
    const int ns = ndim == 2 ? 2 : 4;
    for (int nc = 2; nc <= 3; ++nc)
      for (int adj = 0; adj <= 1; ++adj) {
        const long ncc = adj ? nc * nc : nc;
        auto links = field((size_t)ndim * nc * nc * V, 1);
        auto links_r = field((size_t)ndim * nc * nc * V, 2);
        auto x = field((size_t)ns * ncc * V - (short_input ? 1 : 0), 3);
        auto y = field((size_t)ns * ncc * V, 4);
        std::vector<double> out(2 * (size_t)ns * ncc * V);
        std::vector<double> force(2 * (size_t)ndim * nc * nc * V);
        for (int herm = 0; herm <= 1; ++herm)
          for (int bc = 0; bc <= 1; ++bc) {
            const double bre = bc ? -1.0 : 1.0;
            for (int proj = 0; proj <= 1; ++proj) {
              const int rc = lymcheck_wilson_apply(ndim, nc, adj, herm, g, 0.12, bre, 0.0, proj, links.data(),
                                                   adj ? links_r.data() : nullptr, x.data(), out.data());
              if (rc) return 2;
              ++cases;
            }
            const double* xs[2] = {x.data(), y.data()};
            const double* ys[2] = {y.data(), x.data()};
            const double w[2] = {0.5, -1.5};
            const int rc = lymcheck_lie_derivative(ndim, nc, adj, herm, g, 0.12, bre, 0.0, links.data(),
                                                   adj ? links_r.data() : nullptr, 2, xs, ys, w, 1, 0,
                                                   force.data());
            if (rc) return 3;
            ++cases;
          }
      }
  }
  std::printf("asan ok: %d cases\n", cases);
  return 0;
}
