// extern "C" entry points of liblym.so (declared in include/lym.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "force_site.h"
#include "lym_internal.h"

namespace lym {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return LYM_ERR_HIP;
}

int64_t field_elems(const OpDesc& d) { return (int64_t)d.ncomp * d.lat.V; }

static int check_one(const void* p, size_t bytes, const char* what) {
  if (!p || bytes == 0) return LYM_OK;
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  const hipError_t e = hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // not a device allocation the runtime tracks: cannot check
    return LYM_OK;
  }
  if (!range_inside((uintptr_t)base, size, (uintptr_t)p, bytes)) {
    set_error(std::string(what) + ": buffer too small (the call reads or writes " + std::to_string(bytes) +
              " bytes, the allocation holds " + std::to_string((uintptr_t)base + size - (uintptr_t)p) +
              " from this pointer)");
    return LYM_ERR_ARG;
  }
  return LYM_OK;
}

int check_buffers(const OpDesc* d, hipStream_t st, std::initializer_list<Buf> bufs) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return LYM_OK;
  (void)hipGetLastError();
  if (d) {
    const size_t lb = (size_t)d->ndim * d->nc * d->nc * d->lat.V;  // complex numbers per link field
    const size_t lb32 = (size_t)d->ndim * link_units32(d->nc) * d->lat.V * 16;  // packed complex64 links
    int rc = check_one(d->links64, lb * 16, "links");
    if (!rc) rc = check_one(d->linksr64, lb * 16, "links_right");
    if (!rc) rc = check_one(d->links32, lb32, "links_f32");
    if (!rc) rc = check_one(d->linksr32, lb32, "links_right_f32");
    if (rc) return rc;
  }
  for (const Buf& b : bufs) {
    const int rc = check_one(b.p, b.bytes, b.what);
    if (rc) return rc;
  }
  return LYM_OK;
}

// bytes of one field of the operator's layout at precision prec
static size_t fbytes(const OpDesc& d, int prec) {
  return (size_t)field_elems(d) * (prec == LYM_PREC_F64 ? 16 : prec == LYM_PREC_F32 ? 8 : 4);
}

// Every LDS tile a 4D lattice admits (L0 divisible by the XCD count):
// a0 | L0/8, a1 | L1, k | L2, a0*a1*k*L3 a multiple of 64 and at most
// max_threads sites, lds_per_site * sites within max_lds bytes.
static std::vector<TileMap> admissible_tiles(const lym_wilson* w, int adj, int64_t lds_per_site, int max_threads,
                                             int64_t max_lds) {
  (void)adj;
  std::vector<TileMap> out;
  if (w->ndim != 4 || w->dims[0] % kNumXcd) return out;
  const int L0 = w->dims[0], L1 = w->dims[1], L2 = w->dims[2], T = w->dims[3];
  const int64_t V = (int64_t)L0 * L1 * L2 * T;
  const int slab = L0 / kNumXcd;
  for (int a0 = 1; a0 <= slab; ++a0) {
    if (slab % a0) continue;
    for (int a1 = 1; a1 <= L1; ++a1) {
      if (L1 % a1) continue;
      for (int k = 1; k <= L2; ++k) {
        if (L2 % k) continue;
        const int64_t ns = (int64_t)a0 * a1 * k * T;
        if (ns % 64 || ns > max_threads || ns * lds_per_site > max_lds) continue;
        out.push_back(TileMap{1, a0, a1, k, T, slab, slab / a0, L1 / a1, L2 / k, (int)ns, (int)(V / ns), 1});
      }
    }
  }
  return out;
}

std::vector<TileMap> tile_candidates(const OpDesc& d, int64_t lds_per_site, int max_threads, int64_t max_lds) {
  lym_wilson w{};
  w.ndim = d.ndim;
  for (int mu = 0; mu < d.ndim; ++mu) w.dims[mu] = d.lat.L[mu];
  return admissible_tiles(&w, d.adj, lds_per_site, max_threads, max_lds);
}

// Tiles chosen by lym_tune_wilson, per (lattice, Nc, representation,
// precision index), for the whole process.
struct TuneKey {
  int ndim, nc, adj, pi, L[4];
  bool operator<(const TuneKey& o) const {
    return std::tie(ndim, nc, adj, pi, L[0], L[1], L[2], L[3]) <
           std::tie(o.ndim, o.nc, o.adj, o.pi, o.L[0], o.L[1], o.L[2], o.L[3]);
  }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TileMap> g_tuned;
static TuneKey tune_key(const lym_wilson* w, int adj, int pi) {
  TuneKey k{w->ndim, w->nc, adj, pi, {1, 1, 1, 1}};
  for (int mu = 0; mu < w->ndim && mu < 4; ++mu) k.L[mu] = w->dims[mu];
  return k;
}

int make_desc(const lym_wilson* w, OpDesc* d) {
  if (!w) {
    set_error("null operator description");
    return LYM_ERR_ARG;
  }
  if (w->ndim != 2 && w->ndim != 4) {
    set_error("Number of dimensions " + std::to_string(w->ndim) + " not supported for fermions!");
    return LYM_ERR_ARG;
  }
  if (w->nc != 2 && w->nc != 3) {
    set_error("number of colours must be 2 or 3");
    return LYM_ERR_UNSUPPORTED;
  }
  if (w->representation != LYM_REP_FUNDAMENTAL && w->representation != LYM_REP_ADJOINT) {
    set_error("unknown representation");
    return LYM_ERR_ARG;
  }
  OpDesc o;
  o.ndim = w->ndim;
  o.nc = w->nc;
  o.adj = w->representation == LYM_REP_ADJOINT;
  o.herm = w->hermitian ? 1 : 0;
  o.ns = w->ndim == 4 ? 4 : 2;
  o.ncc = o.adj ? w->nc * w->nc : w->nc;
  o.ncomp = o.ns * o.ncc;
  int64_t V = 1;
  for (int mu = 0; mu < 4; ++mu) o.lat.L[mu] = 1;
  for (int mu = 0; mu < w->ndim; ++mu) {
    if (w->dims[mu] < 1) {
      set_error("lattice extents must be positive");
      return LYM_ERR_ARG;
    }
    o.lat.L[mu] = w->dims[mu];
    V *= w->dims[mu];
  }
  if (V > (int64_t)1 << 31) {
    set_error("lattice volume above 2^31 sites");
    return LYM_ERR_ARG;
  }
  int64_t st = 1;
  for (int mu = w->ndim - 1; mu >= 0; --mu) {
    o.lat.stride[mu] = (int)st;
    st *= w->dims[mu];
  }
  o.lat.V = V;
  o.lat.kappa = w->kappa;
  o.lat.bc_re = w->bc_re;
  o.lat.bc_im = w->bc_im;
  o.lat.herm = o.herm;
  o.lat.has_bc = (w->bc_re != 1.0 || w->bc_im != 0.0) ? 1 : 0;
  o.links64 = w->links;
  o.linksr64 = w->links_right;
  o.links32 = w->links_f32;
  o.linksr32 = w->links_right_f32;
  if (w->link_format != LYM_LINKS_FULL && w->link_format != LYM_LINKS_SU) {
    set_error("unknown link format");
    return LYM_ERR_ARG;
  }
  o.link_su = w->link_format == LYM_LINKS_SU && !o.adj;
  // block-resident field bytes must stay below 4 GiB (32-bit buffer offsets)
  const int64_t link_bytes = (int64_t)w->ndim * w->nc * w->nc * V * 16;
  const int64_t field_bytes = (int64_t)o.ncomp * V * 16;
  if (link_bytes >= ((int64_t)1 << 32) || field_bytes >= ((int64_t)1 << 32)) {
    set_error("lattice too large for one device field (>= 4 GiB per field)");
    return LYM_ERR_ARG;
  }
  // Wilson block schedule: 4D tiled when the extents allow it, else linear
  o.map = SiteMap{0, 0, 0, 0, 0};
  o.wthreads = kBlock;
  o.wblocks = grid_for(V);
  if (w->ndim == 4) {
    const int T = w->dims[3], L2 = w->dims[2], L0 = w->dims[0];
    if (T <= kBlock && L0 % kNumXcd == 0) {
      for (int k = kBlock / T; k >= 1; --k) {
        if ((T * k) % 64 == 0 && L2 % k == 0) {
          o.map = SiteMap{1, T, k, L0 / kNumXcd, L2 / k};
          o.wthreads = T * k;
          o.wblocks = (int)(V / (T * k));
          break;
        }
      }
    }
  }
  // LYM_NO_TILE=1 in the environment forces the plain kernel (A/B checks);
  // LYM_TILE=a0,a1,k forces a tile shape when it is admissible (and
  // LYM_TILE=a0,a1,k,walk its walk order)
  const char* nt = std::getenv("LYM_NO_TILE");
  o.use_tile = (nt && nt[0] == '1') ? 0 : 1;
  int want[3] = {0, 0, 0}, want_walk = 0;
  if (const char* wt = std::getenv("LYM_TILE")) {
    if (std::sscanf(wt, "%d,%d,%d,%d", &want[0], &want[1], &want[2], &want_walk) < 3) want[0] = want[1] = want[2] = 0;
  }
  // LDS tile per precision (4D, L0 a multiple of the XCD count):
  // minimise global spinor reads per site, 1 + sum_{mu<3} 2/a_mu (0 for a
  // full extent), within 64 KiB of LDS and 640 threads (a multiple of 64).
  // Blocks of 256..512 sites are preferred when the lattice admits them
  // (smaller tiles re-read more, larger ones lose occupancy); ties go to the
  // deeper x2 extent (neighbours nearer in memory), then the smaller tile.
  // lds_per_site: LDS bytes a tile site needs; [lo, hi]: preferred tile sizes
  auto choose_tile = [&](int64_t lds_per_site, int max_threads, int lo, int hi, int64_t max_lds) {
    TileMap best{};
    const std::vector<TileMap> all = admissible_tiles(w, o.adj, lds_per_site, max_threads, max_lds);
    const int L0 = w->dims[0], L1 = w->dims[1], L2 = w->dims[2];
    auto part = [](int a, int L) { return a == L ? 0.0 : 2.0 / a; };
    for (int pass = 0; pass < 2 && !best.on; ++pass) {
      double best_cost = 1e30;
      for (const TileMap& t : all) {
        const bool forced = want[0] && t.a0 == want[0] && t.a1 == want[1] && t.k == want[2];
        if (pass == 0 && !forced && (t.nsite < lo || t.nsite > hi)) continue;
        const double cost = forced ? -1.0 : 1.0 + part(t.a0, L0) + part(t.a1, L1) + part(t.k, L2);
        const bool tie = std::fabs(cost - best_cost) < 1e-9;
        if (cost < best_cost - 1e-9 || (tie && (t.k > best.k || (t.k == best.k && t.nsite < best.nsite)))) {
          best_cost = cost;
          best = t;
        }
      }
    }
    return best;
  };
  const char* hf = std::getenv("LYM_HIF");
  const int want_hif = hf ? std::atoi(hf) : 0;
  for (int pi = 0; pi < 2; ++pi) {
    const int64_t es = pi == 0 ? 16 : 8;
    // adjoint tiles: at most kTileAdjThreads sites (register budget), the
    // largest admissible preferred
    TileMap best = o.adj ? choose_tile(o.ncomp * es, kTileAdjThreads, 128, kTileAdjThreads, kTileMaxLds)
                         : choose_tile(o.ncomp * es, pi == 0 ? kTileMaxThreads64 : kTileMaxThreads, 256, 512,
                                       kTileMaxLds);
    o.tile_tuned[pi] = 0;
    if (!want[0] && !want_hif && o.use_tile) {
      // a shape recorded by lym_tune_wilson for this lattice wins
      std::lock_guard<std::mutex> lock(g_tune_mu);
      auto it = g_tuned.find(tune_key(w, o.adj, pi));
      if (it != g_tuned.end()) {
        best = it->second;
        o.tile_tuned[pi] = 1;
      }
    }
    if (want_hif == 2 || want_hif == 4) {
      // the launcher's limits (wilson.hip): hops in flight need blocks of at
      // most tile_hif_threads sites, and 4 is built for complex64 only; the
      // recorded value is the one that runs
      int h = want_hif;
      if (pi == 0 && h == 4) h = 2;
      if (best.nsite > tile_hif_threads(pi == 0 ? P64 : P32)) h = 1;
      best.hif = h;
    }
    if (want[0] && best.on) best.walk = want_walk ? 1 : 0;
    o.tile[pi] = best;
  }
  // multi-RHS tile (complex64): kMrRhs spinors per site in LDS.  Opt-in
  // (LYM_MR=1 in the environment): kMrRhs spinor tiles leave LDS room for
  // only a few waves per CU, and measured on MI355X that costs more than the
  // saved link traffic (batched CG, 16^3x32 and 24^3x48: 8% and 17% slower).
  o.tile_mr = o.adj ? TileMap{} : choose_tile(o.ncomp * 8 * kMrRhs, kTileMaxThreadsMr, 128, 256, kTileMaxLdsMr);
  const char* mr = std::getenv("LYM_MR");
  o.use_mr = (mr && mr[0] == '1') ? 1 : 0;
  *d = o;
  return LYM_OK;
}

size_t solver_workspace_bytes(const OpDesc& d, int prec, int nshift, int nrhs);
int solve_cgm(const OpDesc& d, int prec, int nshift, const double* shifts, const void* b, void* const* xs, double tol,
              int max_iter, void* ws, size_t ws_bytes, lym_solve_stats* stats, hipStream_t st);
int solve_cgm_batched(const OpDesc& d, int prec, int nrhs, int nshift, const double* shifts, const void* const* bs,
                      void* const* xs, double tol, int max_iter, void* ws, size_t ws_bytes, lym_solve_stats* stats,
                      hipStream_t st);
size_t bicgstab_workspace_bytes(const OpDesc& d);
int solve_bicgstab(const OpDesc& d, int square, double shift, const void* b, const void* x0, void* x, int max_steps,
                   double precision, void* ws, size_t ws_bytes, lym_solve_stats* stats, hipStream_t st);
int solve_mr(const OpDesc& d, int nshift, const double* shifts, int first, const void* b, void* const* xs,
             int max_steps, double precision, double omega, void* ws, size_t ws_bytes, lym_solve_stats* stats,
             hipStream_t st);

}  // namespace lym

using namespace lym;

static size_t reduce_scratch(int64_t) {
  // partials (kMaxRed x 2048) + ticket + results, each 256-aligned
  return 256 * 4 + sizeof(double) * kMaxRed * 2048 + 64 + sizeof(double) * 16;
}

extern "C" {

int lym_abi_version(void) { return LYM_ABI_VERSION; }

int lym_init(int device) {
  int cur = -1;
  LYM_HIP(hipGetDevice(&cur));
  if (cur != device) LYM_HIP(hipSetDevice(device));
  void* sp = nullptr;  // device address of lym_dot_noalloc's static scratch on this device
  return noalloc_scratch(&sp);
}

const char* lym_last_error(void) { return g_err.c_str(); }

int lym_wilson_apply(const lym_wilson* op, int prec, int mode, const void* in, void* out, void* stream) {
  OpDesc d;
  int rc = make_desc(op, &d);
  if (rc) return rc;
  if ((rc = check_buffers(&d, (hipStream_t)stream, {{in, fbytes(d, prec), "lym_wilson_apply in"},
                                                     {out, fbytes(d, prec), "lym_wilson_apply out"}})))
    return rc;
  if (!in || !out || in == out) {
    set_error("lym_wilson_apply: in/out must be distinct non-null device pointers");
    return LYM_ERR_ARG;
  }
  if (prec != LYM_PREC_F64 && prec != LYM_PREC_F32) {
    set_error("lym_wilson_apply: precision must be F64 or F32");
    return LYM_ERR_ARG;
  }
  EpiArgs a;
  a.out.data = out;
  Field f;
  f.data = const_cast<void*>(in);
  return wilson_launch(EPI_STORE, d, prec, mode != LYM_MODE_SIMPLE, f, a, RedCtl{nullptr, nullptr},
                       (hipStream_t)stream);
}

int lym_wilson_apply_square(const lym_wilson* op, int prec, double shift, const void* in, void* out, void* tmp,
                            void* stream) {
  OpDesc d;
  int rc = make_desc(op, &d);
  if (rc) return rc;
  if (!in || !out || !tmp || in == out || in == tmp || out == tmp) {
    set_error("lym_wilson_apply_square: in/out/tmp must be distinct non-null device pointers");
    return LYM_ERR_ARG;
  }
  if (prec != LYM_PREC_F64 && prec != LYM_PREC_F32) {
    set_error("lym_wilson_apply_square: precision must be F64 or F32");
    return LYM_ERR_ARG;
  }
  if ((rc = check_buffers(&d, (hipStream_t)stream, {{in, fbytes(d, prec), "lym_wilson_apply_square in"},
                                                     {out, fbytes(d, prec), "lym_wilson_apply_square out"},
                                                     {tmp, fbytes(d, prec), "lym_wilson_apply_square tmp"}})))
    return rc;
  EpiArgs a;
  a.out.data = tmp;
  Field f;
  f.data = const_cast<void*>(in);
  rc = wilson_launch(EPI_STORE, d, prec, true, f, a, RedCtl{nullptr, nullptr}, (hipStream_t)stream);
  if (rc) return rc;
  EpiArgs a2;
  a2.out.data = out;
  a2.aux0.data = const_cast<void*>(in);
  a2.sigma = shift;
  Field t;
  t.data = tmp;
  return wilson_launch(EPI_AXPY, d, prec, true, t, a2, RedCtl{nullptr, nullptr}, (hipStream_t)stream);
}

size_t lym_reduce_scratch_bytes(int64_t n) { return reduce_scratch(n); }

int lym_dot(int prec, int64_t n, const void* x, const void* y, double* result, void* scratch, void* stream) {
  if (!x || !y || !result || !scratch || n < 0) {
    set_error("lym_dot: bad arguments");
    return LYM_ERR_ARG;
  }
  char* b = (char*)(((uintptr_t)scratch + 255) & ~(uintptr_t)255);
  double* partials = (double*)b;
  unsigned* ticket = (unsigned*)(b + (((sizeof(double) * kMaxRed * 2048) + 255) & ~(size_t)255));
  double* res = (double*)((char*)ticket + 256);
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) {
    result[0] = result[1] = 0.0;
    return LYM_OK;
  }
  int rc = blas_dot(prec, n, x, y, res, RedCtl{partials, ticket}, st);
  if (rc) return rc;
  LYM_HIP(hipMemcpyAsync(result, res, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  return LYM_OK;
}

int lym_dot_noalloc(int prec, int64_t n, const void* x, const void* y, double* result_dev, void* stream) {
  if (!x || !y || !result_dev || n < 0) {
    set_error("lym_dot_noalloc: bad arguments");
    return LYM_ERR_ARG;
  }
  return blas_dot_noalloc(prec, n, x, y, result_dev, (hipStream_t)stream);
}

int lym_axpby(int prec, int64_t n, double a_re, double a_im, const void* x, double b_re, double b_im, void* y,
              void* stream) {
  if (!x || !y || n < 0) {
    set_error("lym_axpby: bad arguments");
    return LYM_ERR_ARG;
  }
  if (n == 0) return LYM_OK;
  return blas_axpby(prec, n, a_re, a_im, x, b_re, b_im, y, (hipStream_t)stream);
}

static void fill_choice(const TileMap& t, int tuned, double us, lym_tile_choice* c) {
  c->a0 = t.on ? t.a0 : 0;
  c->a1 = t.on ? t.a1 : 0;
  c->k = t.on ? t.k : 0;
  c->hif = t.on ? (t.hif ? t.hif : 1) : 0;
  c->nsite = t.on ? t.nsite : 0;
  c->tuned = tuned;
  c->walk = t.on ? t.walk : 0;
  c->reserved = 0;
  c->us = us;
}

int lym_tile_info(const lym_wilson* op, int prec, lym_tile_choice* out) {
  OpDesc d;
  int rc = make_desc(op, &d);
  if (rc) return rc;
  if (!out || (prec != LYM_PREC_F64 && prec != LYM_PREC_F32 && prec != LYM_PREC_F16)) {
    set_error("lym_tile_info: bad arguments");
    return LYM_ERR_ARG;
  }
  const int pi = prec == LYM_PREC_F64 ? 0 : 1;
  fill_choice(wilson_tiled(d, prec) ? d.tile[pi] : TileMap{}, d.tile_tuned[pi], 0.0, out);
  return LYM_OK;
}

int lym_tune_wilson(const lym_wilson* op, int prec, int reps, const void* in, void* out, lym_tile_choice* best,
                    void* stream) {
  OpDesc d;
  int rc = make_desc(op, &d);
  if (rc) return rc;
  if (!in || !out || in == out || reps < 1 || (prec != LYM_PREC_F64 && prec != LYM_PREC_F32)) {
    set_error("lym_tune_wilson: bad arguments");
    return LYM_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const int pi = prec == LYM_PREC_F64 ? 0 : 1;
  // candidates: the untiled kernel, then every tile x hops in flight x walk
  // order (the order in which an XCD's blocks take the tiles changes which
  // halos and which memory pages are in flight together: at 32^3x64 x1
  // before x2 is 4-5% faster for complex128 and 7-9% slower for complex64)
  std::vector<TileMap> cands{TileMap{}};
  for (const TileMap& t : admissible_tiles(op, d.adj, (int64_t)d.ncomp * (pi == 0 ? 16 : 8),
                                           tile_max_threads(d.adj, pi == 0 ? P64 : P32), kTileMaxLds)) {
    for (int hif : {1, 2, 4}) {
      if (hif > 1 && d.adj) continue;  // adjoint tiles run one hop in flight
      if (hif == 4 && prec != LYM_PREC_F32) continue;
      // the launcher runs larger tiles with one hop in flight
      if (hif > 1 && t.nsite > tile_hif_threads(prec == LYM_PREC_F64 ? P64 : P32)) continue;
      for (int walk : {0, 1}) {
        if (walk && (t.n1 == 1 || t.n2 == 1)) continue;  // the same order
        TileMap c = t;
        c.hif = hif;
        c.walk = walk;
        cands.push_back(c);
      }
    }
  }
  hipEvent_t e0, e1;
  LYM_HIP(hipEventCreate(&e0));
  LYM_HIP(hipEventCreate(&e1));
  EpiArgs a;
  a.out.data = out;
  Field f;
  f.data = const_cast<void*>(in);
  double best_us = 1e30;
  TileMap best_t{};
  // each candidate: the best of 3 timed runs of `reps` launches (run-to-run
  // noise is a few percent, the spread between good shapes not much more);
  // the first candidate is warmed up twice as long
  for (size_t ci = 0; ci < cands.size() && !rc; ++ci) {
    const TileMap& c = cands[ci];
    OpDesc dc = d;
    dc.use_tile = c.on ? 1 : 0;
    dc.use_mr = 0;
    dc.tile[pi] = c;
    const int warm = ci == 0 ? 2 * reps : 2;
    for (int i = 0; i < warm && !rc; ++i) rc = wilson_launch(EPI_STORE, dc, prec, true, f, a, RedCtl{}, st);
    double us = 1e30;
    for (int trial = 0; trial < 3 && !rc; ++trial) {
      if (hipEventRecord(e0, st) != hipSuccess) rc = hip_fail(hipErrorUnknown, "hipEventRecord");
      for (int i = 0; i < reps && !rc; ++i) rc = wilson_launch(EPI_STORE, dc, prec, true, f, a, RedCtl{}, st);
      if (!rc && hipEventRecord(e1, st) != hipSuccess) rc = hip_fail(hipErrorUnknown, "hipEventRecord");
      float ms = 0.f;
      if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = hip_fail(hipErrorUnknown, "hipEventSynchronize");
      if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = hip_fail(hipErrorUnknown, "hipEventElapsedTime");
      if (!rc) us = std::min(us, 1e3 * ms / reps);
    }
    if (!rc && us < best_us) {
      best_us = us;
      best_t = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned[tune_key(op, d.adj, pi)] = best_t;
  }
  if (best) fill_choice(best_t, 1, best_us, best);
  return LYM_OK;
}

int lym_link_deviation(const lym_wilson* op, int prec, double* dev_out, void* stream) {
  OpDesc d;
  int rc = make_desc(op, &d);
  if (rc) return rc;
  if (!dev_out || (prec != LYM_PREC_F64 && prec != LYM_PREC_F32)) {
    set_error("lym_link_deviation: bad arguments");
    return LYM_ERR_ARG;
  }
  return link_deviation(d, prec, dev_out, (hipStream_t)stream);
}

int lym_lie_derivative(const lym_wilson* op, int npairs, const void* const* xs, const void* const* ys,
                       const double* weights, int symmetric, int accumulate, void* out, void* stream) {
  OpDesc d;
  int rc = make_desc(op, &d);
  if (rc) return rc;
  if (!out || npairs < 0 || (npairs > 0 && (!xs || !ys || !weights))) {
    set_error("lym_lie_derivative: bad arguments");
    return LYM_ERR_ARG;
  }
  for (int p = 0; p < npairs; ++p) {
    if (!xs[p] || !ys[p] || xs[p] == out || ys[p] == out || !std::isfinite(weights[p])) {
      set_error("lym_lie_derivative: null, aliased or non-finite pair " + std::to_string(p));
      return LYM_ERR_ARG;
    }
  }
  {
    const size_t ob = (size_t)d.ndim * d.nc * d.nc * d.lat.V * 16;  // (D, nc, nc, V) complex128
    if ((rc = check_buffers(&d, (hipStream_t)stream, {{out, ob, "lym_lie_derivative out"}}))) return rc;
    for (int p = 0; p < npairs; ++p)
      if ((rc = check_buffers(nullptr, (hipStream_t)stream, {{xs[p], fbytes(d, LYM_PREC_F64), "lym_lie_derivative xs"},
                                                              {ys[p], fbytes(d, LYM_PREC_F64), "lym_lie_derivative ys"}})))
        return rc;
  }
  if (npairs == 0 && accumulate) return LYM_OK;
  // launches of at most kForceMaxPairs pairs; the later ones accumulate
  int p0 = 0;
  do {
    ForcePairs fp{};
    fp.n = std::min(npairs - p0, kForceMaxPairs);
    fp.sym = symmetric ? 1 : 0;
    fp.accumulate = (p0 > 0 || accumulate) ? 1 : 0;
    for (int q = 0; q < fp.n; ++q) {
      fp.x[q] = (const cpx<double>*)xs[p0 + q];
      fp.y[q] = (const cpx<double>*)ys[p0 + q];
      fp.w[q] = weights[p0 + q];
    }
    rc = force_launch(d, fp, out, (hipStream_t)stream);
    if (rc) return rc;
    p0 += fp.n;
  } while (p0 < npairs);
  return LYM_OK;
}

int lym_slice_sums(const lym_wilson* op, int dir, int nvec, const void* const* vecs, void* scratch,
                   size_t scratch_bytes, double* out, void* stream) {
  OpDesc d;
  int rc = make_desc(op, &d);
  if (rc) return rc;
  if (dir < 0 || dir >= d.ndim || nvec < 0 || (nvec > 0 && !vecs) || !scratch || !out) {
    set_error("lym_slice_sums: bad arguments");
    return LYM_ERR_ARG;
  }
  const int64_t V = d.lat.V;
  const int L = d.lat.L[dir];
  if (scratch_bytes < (size_t)(V + L) * sizeof(double)) {
    set_error("lym_slice_sums: scratch smaller than (V + L_dir) doubles");
    return LYM_ERR_ARG;
  }
  for (int p = 0; p < nvec; ++p) {
    if (!vecs[p]) {
      set_error("lym_slice_sums: null vector " + std::to_string(p));
      return LYM_ERR_ARG;
    }
  }
  hipStream_t st = (hipStream_t)stream;
  double* dens = (double*)scratch;
  double* out_dev = dens + V;
  int p0 = 0;
  do {
    SliceVecs sv{};
    sv.n = std::min(nvec - p0, kSliceMaxVec);
    sv.accumulate = p0 > 0 ? 1 : 0;
    for (int q = 0; q < sv.n; ++q) sv.v[q] = (const cpx<double>*)vecs[p0 + q];
    rc = slice_density(d, sv, dens, st);
    if (rc) return rc;
    p0 += sv.n;
  } while (p0 < nvec);
  rc = slice_reduce(d, dir, dens, out_dev, st);
  if (rc) return rc;
  LYM_HIP(hipMemcpyAsync(out, out_dev, (size_t)L * sizeof(double), hipMemcpyDeviceToHost, st));
  LYM_HIP(hipStreamSynchronize(st));
  return LYM_OK;
}

int lym_unpack_ustar(const void* raw, int64_t volume, int ndim, void* links, void* stream) {
  if (!raw || !links || volume < 1 || ndim < 1 || ndim > 4 || volume > ((int64_t)1 << 31) ||
      volume * ndim * 9 * 16 >= ((int64_t)1 << 40)) {
    set_error("lym_unpack_ustar: bad arguments");
    return LYM_ERR_ARG;
  }
  return unpack_ustar(raw, volume, ndim, links, (hipStream_t)stream);
}

int lym_unitary_projection(int ndim, int nc, int64_t volume, const void* in, void* out, void* stream) {
  if (!in || !out || volume < 1 || ndim < 1 || ndim > 4 || volume > ((int64_t)1 << 31)) {
    set_error("lym_unitary_projection: bad arguments");
    return LYM_ERR_ARG;
  }
  return unitary_project(ndim, nc, volume, in, out, (hipStream_t)stream);
}

int lym_gamma5(int prec, int ns, int64_t n_per_spin, const void* in, void* out, void* stream) {
  if (!in || !out || ns < 2 || (ns & 1) || n_per_spin < 0) {
    set_error("lym_gamma5: bad arguments");
    return LYM_ERR_ARG;
  }
  if (n_per_spin == 0) return LYM_OK;
  return blas_gamma5(prec, ns, n_per_spin, in, out, (hipStream_t)stream);
}

int lym_spin_axpby(int ns, int64_t n_per_spin, const double* coef, const void* x, const void* y, void* out,
                   void* stream) {
  if (!coef || !x || !y || !out || ns < 2 || (ns & 1) || n_per_spin < 0) {
    set_error("lym_spin_axpby: bad arguments");
    return LYM_ERR_ARG;
  }
  if (n_per_spin == 0) return LYM_OK;
  SpinCoef c;
  for (int i = 0; i < 8; ++i) c.v[i] = coef[i];
  return blas_spin_axpby(ns, n_per_spin, c, x, y, out, (hipStream_t)stream);
}

int lym_convert(int prec_in, const void* in, int prec_out, void* out, int64_t n, void* stream) {
  if (!in || !out || n < 0) {
    set_error("lym_convert: bad arguments");
    return LYM_ERR_ARG;
  }
  if (n == 0) return LYM_OK;
  return blas_convert(prec_in, in, prec_out, out, n, (hipStream_t)stream);
}

size_t lym_links_f32_bytes(int32_t ndim, int32_t nc, int64_t volume) {
  if ((ndim != 2 && ndim != 4) || nc < 2 || nc > 3 || volume < 1) return 0;
  return (size_t)ndim * link_units32(nc) * (size_t)volume * 16;
}

int lym_pack_links_f32(int32_t ndim, int32_t nc, int64_t volume, const void* links, void* links_f32, void* stream) {
  const size_t nb = lym_links_f32_bytes(ndim, nc, volume);
  if (!links || !links_f32 || nb == 0) {
    set_error("lym_pack_links_f32: bad arguments");
    return LYM_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const int rc = check_buffers(nullptr, st, {{links, (size_t)ndim * nc * nc * (size_t)volume * 16, "links"},
                                             {links_f32, nb, "links_f32"}});
  if (rc) return rc;
  return blas_pack_links32(ndim, nc, volume, links, links_f32, st);
}

size_t lym_solver_workspace_bytes(const lym_wilson* op, int prec, int nshift) {
  OpDesc d;
  if (make_desc(op, &d)) return 0;
  if (nshift < 1 || nshift > kMaxShifts) return 0;
  return solver_workspace_bytes(d, prec, nshift, 1);
}

size_t lym_cgm_batched_workspace_bytes(const lym_wilson* op, int prec, int nshift, int nrhs) {
  OpDesc d;
  if (make_desc(op, &d)) return 0;
  if (nshift < 1 || nshift > kMaxShifts || nrhs < 1) return 0;
  return solver_workspace_bytes(d, prec, nshift, prec == LYM_PREC_F64 ? 1 : nrhs);
}

int lym_mr_multishift(const lym_wilson* op, int nshift, const double* shifts, int first, const void* b,
                      void* const* xs, int max_steps, double precision, double omega, void* workspace,
                      size_t workspace_bytes, lym_solve_stats* stats, void* stream) {
  OpDesc d;
  int rc = make_desc(op, &d);
  if (rc) return rc;
  if (!b || !xs || !shifts) {
    set_error("lym_mr_multishift: null argument");
    return LYM_ERR_ARG;
  }
  if ((rc = check_buffers(&d, (hipStream_t)stream, {{b, fbytes(d, LYM_PREC_F64), "lym_mr_multishift b"},
                                                     {workspace, workspace_bytes, "lym_mr_multishift workspace"}})))
    return rc;
  for (int k = first; k >= 0 && k < nshift && k < 64; ++k)
    if ((rc = check_buffers(nullptr, (hipStream_t)stream, {{xs[k], fbytes(d, LYM_PREC_F64), "lym_mr_multishift xs"}})))
      return rc;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  return solve_mr(d, nshift, shifts, first, b, xs, max_steps, precision, omega, workspace, workspace_bytes, stats,
                  (hipStream_t)stream);
}

int lym_cgm(const lym_wilson* op, int prec, int nshift, const double* shifts, const void* b, void* const* xs,
            double tol, int max_iter, void* workspace, size_t workspace_bytes, lym_solve_stats* stats, void* stream) {
  OpDesc d;
  int rc = make_desc(op, &d);
  if (rc) return rc;
  if (!b || !xs || !shifts || !(tol > 0.0) || max_iter < 1) {
    set_error("lym_cgm: bad arguments");
    return LYM_ERR_ARG;
  }
  if (nshift >= 1 && nshift <= 64) {
    if ((rc = check_buffers(&d, (hipStream_t)stream, {{b, fbytes(d, LYM_PREC_F64), "lym_cgm b"},
                                                       {workspace, workspace_bytes, "lym_cgm workspace"}})))
      return rc;
    for (int k = 0; k < nshift; ++k)
      if ((rc = check_buffers(nullptr, (hipStream_t)stream, {{xs[k], fbytes(d, LYM_PREC_F64), "lym_cgm xs[k]"}})))
        return rc;
  }
  if (stats) std::memset(stats, 0, sizeof(*stats));
  return solve_cgm(d, prec, nshift, shifts, b, xs, tol, max_iter, workspace, workspace_bytes, stats,
                   (hipStream_t)stream);
}

int lym_cgm_batched(const lym_wilson* op, int prec, int nrhs, int nshift, const double* shifts, const void* const* bs,
                    void* const* xs, double tol, int max_iter, void* workspace, size_t workspace_bytes,
                    lym_solve_stats* stats, void* stream) {
  OpDesc d;
  int rc = make_desc(op, &d);
  if (rc) return rc;
  if (!bs || !xs || !shifts || !(tol > 0.0) || max_iter < 1 || nrhs < 1) {
    set_error("lym_cgm_batched: bad arguments");
    return LYM_ERR_ARG;
  }
  if (nshift >= 1 && nshift <= 64) {
    if ((rc = check_buffers(&d, (hipStream_t)stream, {{workspace, workspace_bytes, "lym_cgm_batched workspace"}})))
      return rc;
    for (int r = 0; r < nrhs; ++r) {
      if ((rc = check_buffers(nullptr, (hipStream_t)stream, {{bs[r], fbytes(d, LYM_PREC_F64), "lym_cgm_batched bs[r]"}})))
        return rc;
      for (int k = 0; k < nshift; ++k)
        if ((rc = check_buffers(nullptr, (hipStream_t)stream,
                                {{xs[(size_t)r * nshift + k], fbytes(d, LYM_PREC_F64), "lym_cgm_batched xs"}})))
          return rc;
    }
  }
  if (stats) std::memset(stats, 0, sizeof(*stats) * nrhs);
  return solve_cgm_batched(d, prec, nrhs, nshift, shifts, bs, xs, tol, max_iter, workspace, workspace_bytes, stats,
                           (hipStream_t)stream);
}

size_t lym_bicgstab_workspace_bytes(const lym_wilson* op) {
  OpDesc d;
  if (make_desc(op, &d)) return 0;
  return bicgstab_workspace_bytes(d);
}

int lym_bicgstab(const lym_wilson* op, int square, double shift, const void* b, const void* x0, void* x,
                 int max_steps, double precision, void* workspace, size_t workspace_bytes, lym_solve_stats* stats,
                 void* stream) {
  OpDesc d;
  int rc = make_desc(op, &d);
  if (rc) return rc;
  if (!b || !x || max_steps < 1 || (square != 0 && square != 1) || !std::isfinite(shift) || x == b) {
    set_error("lym_bicgstab: bad arguments");
    return LYM_ERR_ARG;
  }
  if ((rc = check_buffers(&d, (hipStream_t)stream, {{b, fbytes(d, LYM_PREC_F64), "lym_bicgstab b"},
                                                     {x0, fbytes(d, LYM_PREC_F64), "lym_bicgstab x0"},
                                                     {x, fbytes(d, LYM_PREC_F64), "lym_bicgstab x"},
                                                     {workspace, workspace_bytes, "lym_bicgstab workspace"}})))
    return rc;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  return solve_bicgstab(d, square, shift, b, x0, x, max_steps, precision, workspace, workspace_bytes, stats,
                        (hipStream_t)stream);
}

int lym_kernel_timing(int enable) { return kernel_timing(enable); }

int lym_kernel_timing_read(lym_kernel_time* out, int max, int* count) {
  if ((!out && max > 0) || max < 0) {
    set_error("lym_kernel_timing_read: bad arguments");
    return LYM_ERR_ARG;
  }
  return kernel_timing_read(out, max, count);
}

}  // extern "C"
