// Internal interfaces shared by the translation units of liblym.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <initializer_list>
#include <string>
#include <vector>

#include "../../include/lym.h"
#include "lym_reduce.h"
#include "lym_state.h"
#include "lym_types.h"

namespace lym {

// A spinor field argument: data planes plus (F16 storage only) per-site norms.
struct Field {
  void* data = nullptr;
  float* norm = nullptr;
};

// Validated operator description.
struct OpDesc {
  int ndim = 0, nc = 0, adj = 0, herm = 0;
  int ns = 0, ncc = 0, ncomp = 0;  // spins, colour components per spin, total
  Lat lat{};
  SiteMap map{};     // block schedule of the plain Wilson kernel
  int wblocks = 0;   // plain Wilson grid size (blocks)
  int wthreads = 0;  // plain Wilson block size (threads, multiple of 64)
  TileMap tile[2]{};  // LDS-tiled kernel per precision: [0] complex128, [1] complex64
  int tile_tuned[2]{};  // the tile came from lym_tune_wilson
  int use_tile = 1;   // 0 forces the plain kernel
  TileMap tile_mr{};  // multi-RHS LDS tile (complex64, kMrRhs spinors per site)
  int use_mr = 0;     // 1: batched complex64 launches use the multi-RHS kernel
  int link_su = 0;    // links are SU(nc): fundamental kernels rebuild the last row
  const void* links64 = nullptr;
  const void* linksr64 = nullptr;
  const void* links32 = nullptr;
  const void* linksr32 = nullptr;
};

// Arguments of a fused Wilson epilogue (meaning depends on the epilogue).
struct EpiArgs {
  Field out;   // primary output
  Field aux0;  // secondary input (center values)
  Field aux1;  // secondary in/out
  double sigma = 0.0;
  SolverState* st = nullptr;
  // fast CG-M residual ring / mixed CG direction ring: nslot fields of
  // `slot_elems` elements each
  Field ring;
  int64_t slot_elems = 0;
  int nslot = 0;
  // batched launches: RHS i uses fields offset by i * rhs_stride elements,
  // rings by i * ring_rhs_stride, and SolverState st + i
  int64_t rhs_stride = 0;
  int64_t ring_rhs_stride = 0;
  const Field* bs = nullptr;  // device array of per-RHS sources (reliable update)
};

// error reporting (thread local message)
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
#define LYM_HIP(call)                                   \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return ::lym::hip_fail(_e, #call); \
  } while (0)

// Raise a kernel's dynamic-LDS limit to `bytes` on the current device, once
// per (kernel, device): `done` is that kernel's bitmask of devices.
inline int raise_lds_limit(const void* fn, int bytes, std::atomic<uint32_t>& done) {
  int dev = 0;
  LYM_HIP(hipGetDevice(&dev));
  const uint32_t bit = 1u << (dev & 31);
  if (done.load(std::memory_order_acquire) & bit) return LYM_OK;
  LYM_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  done.fetch_or(bit, std::memory_order_release);
  return LYM_OK;
}

int make_desc(const lym_wilson* w, OpDesc* d);

// A device buffer argument and the bytes the call will touch from it.
struct Buf {
  const void* p;
  size_t bytes;
  const char* what;
};
// Host-side check that every buffer (and the operator's links) lies inside
// the device allocation that contains its start, as the runtime reports it
// (hipMemGetAddressRange): a short buffer returns LYM_ERR_ARG instead of being
// read past its end.  Skipped while `st` is being captured into a graph, and
// for pointers the runtime does not know.
int check_buffers(const OpDesc* d, hipStream_t st, std::initializer_list<Buf> bufs);
int64_t field_elems(const OpDesc& d);  // complex numbers per spinor field

// ---- Wilson kernel launchers (wilson.hip) --------------------------------
enum EpiKind : int {
  EPI_STORE = 0,     // out = Op in
  EPI_AXPY = 1,      // out = Op in + sigma * aux0
  EPI_CGM_K1 = 2,    // out = Op in; reduce |out|^2, |in|^2; CG-M finalizer A
  EPI_CGM_K2 = 3,    // Ap = Op in + sigma0 aux0; aux1 -= alpha Ap; reduce |aux1|^2; finalizer B
  EPI_MR_K1 = 4,     // out = Op in (phase gated)
  EPI_MR_K2 = 5,     // out = Op in + sigma aux0; reduce <out,aux0>, <out,out>; MR finalizer
  EPI_CGM_K2F = 6,   // fast CG-M: ring[k+1] = ring[k] - alpha (Op in + sigma0 aux0); deferred finalizer
  EPI_CG_K2X = 7,    // mixed CG: aux1 -= alpha (Op in + sigma0 p_k from the ring); reduce |aux1|^2
  EPI_RU1 = 8,       // reliable update (complex128, gate RU): out = Op in
  EPI_RU2 = 9,       // reliable update: r = bs - (Op in + sigma0 aux0) -> out (complex64); reduce |r|^2, <r, p>
  EPI_RU2H = 10,     // the same with out in half storage (per-site norms)
  EPI_GSTORE = 11,   // out = Op in, only while st->phase == PH_K1 (BiCGStab)
  EPI_GAXPY = 12,    // out = Op in + sigma aux0, only while st->phase == PH_K1
  EPI_CG_K1R = 13,   // mixed CG: out = Op p_k, input = ring slot iter of each RHS; reduce |out|^2, |p|^2
};
int wilson_launch(EpiKind kind, const OpDesc& d, int prec, bool proj, Field in, const EpiArgs& a, RedCtl rc,
                  hipStream_t stream, int nrhs = 1);
// number of site blocks (= partials per RHS) of a Wilson launch at `prec`
inline bool wilson_tiled(const OpDesc& d, int prec) {
  // half storage decodes into the complex64 tile (same LDS footprint)
  return d.use_tile && (prec == P64 || prec == P32 || prec == P16) && d.tile[prec == P64 ? 0 : 1].on;
}
// batched complex64 launches on 4D fundamental lattices: kMrRhs RHS per block
// share every link load
constexpr int kMrRhs = 3;
inline bool wilson_mr(const OpDesc& d, int prec, int nrhs) {
  return d.use_tile && d.use_mr && prec == P32 && nrhs > 1 && d.tile_mr.on;
}
// every LDS tile the lattice of d admits (4D fundamental, L0 a multiple of
// the XCD count), hop grouping 1
std::vector<TileMap> tile_candidates(const OpDesc& d, int64_t lds_per_site, int max_threads, int64_t max_lds);
inline int wilson_blocks(const OpDesc& d, int prec, int nrhs = 1) {
  if (wilson_mr(d, prec, nrhs)) return d.tile_mr.blocks;
  return wilson_tiled(d, prec) ? d.tile[prec == P64 ? 0 : 1].blocks : d.wblocks;
}

// ---- finalizer kernel (solvers.hip): one block reduces the partials of the
// preceding kernel in a fixed order, then advances the solver state.
enum FinKind : int {
  FK_SUM = 0,         // out[0..nr) = grid sums (no state)
  FK_INIT = 1,        // bb = rr = sum; tolerances scaled by bb; DONE if bb == 0
  FK_CGM_A = 2,       // gate K1|KP: step length and shifted zetas
  FK_CGM_B = 3,       // gate K2: beta, roll, convergence (plain CG-M)
  FK_CGM_BF = 4,      // gate K2: as FK_CGM_B, deferred-update recording (fast)
  FK_CGM_K3ADV = 5,   // gate K3: phase -> K1 or DONE
  FK_FAST_K3ADV = 6,  // gate K3: clear block, phase -> KP or DONE
  FK_MR = 7,          // gate K1: MR alpha and shift coefficients
  FK_MR_RR = 8,       // gate K3: MR stop test
  FK_CG_A = 9,        // gate K1|KP: mixed CG step length (counts a breakdown as done)
  FK_CG_B = 10,       // gate K2: mixed CG beta / reliable-update decision
  FK_CG_RU = 11,      // gate RU: mixed CG after the FP64 true residual
  FK_BICG_INIT = 12,  // gate K1: BiCGStab rho_next = <r_hat, r>, beta
  FK_BICG_A = 13,     // gate K1: alpha = rho_next / <r_hat, nu>
  FK_BICG_W = 14,     // gate K1: omega = <t, s> / <t, t>
  FK_BICG_R = 15,     // gate K1: stop test on |<r, r>|, rho, beta
};
// nb finalizer blocks: block i handles state st + i, partials + i * nr * G, out + i * nr
int launch_finalize(FinKind kind, SolverState* st, const double* partials, int G, int nr, double* out,
                    hipStream_t stream, int nb = 1);

// ---- vector kernels (blas.hip / solvers.hip) ------------------------------
int blas_dot(int prec, int64_t n, const void* x, const void* y, double* partial_out_dev, RedCtl rc,
             hipStream_t st);
int blas_axpby(int prec, int64_t n, double ar, double ai, const void* x, double br, double bi, void* y,
               hipStream_t st);
// allocation-free dot (static device scratch, blas.hip): result to out_dev (2 doubles)
int blas_dot_noalloc(int prec, int64_t n, const void* x, const void* y, double* out_dev, hipStream_t st);
int noalloc_scratch(void** out);
int blas_convert(int pin, const void* in, int pout, void* out, int64_t n, hipStream_t st);
int blas_pack_links32(int ndim, int nc, int64_t V, const void* in, void* out, hipStream_t st);
int blas_gamma5(int prec, int ns, int64_t nps, const void* in, void* out, hipStream_t st);
// out = a x + b y with coefficients per spin half: v = {a_up, a_lo, b_up, b_lo} (re, im)
struct SpinCoef {
  double v[8];
};
int blas_spin_axpby(int ns, int64_t nps, const SpinCoef& c, const void* x, const void* y, void* out,
                    hipStream_t st);
// out_dev[0] = sum |a-b|^2, out_dev[1] = sum |a|^2
int blas_diff_norm2(int prec, int64_t n, const void* a, const void* b, double* out_dev, RedCtl rc, hipStream_t st);
// out_dev[0] = max |stored - rebuilt| over the last rows of every link
int link_deviation(const OpDesc& d, int prec, double* out_dev, hipStream_t st);

// ---- kernel timing of the fast multishift CG (solvers.hip) ------------------
int kernel_timing(int enable);
int kernel_timing_read(lym_kernel_time* out, int max, int* count);

// ---- measurements (measure.hip) ---------------------------------------------
constexpr int kSliceMaxVec = 32;
struct SliceVecs {
  const cpx<double>* v[kSliceMaxVec];
  int n;
  int accumulate;
};
// dens[s] (+)= sum over sv of the pion density of site s (V doubles)
int slice_density(const OpDesc& d, const SliceVecs& sv, double* dens, hipStream_t st);
// out_dev[z] = sum of dens over the sites with x_dir = z (L_dir doubles)
int slice_reduce(const OpDesc& d, int dir, const double* dens, double* out_dev, hipStream_t st);

// ---- gauge-field files (gauge_io.hip) ---------------------------------------
int unpack_ustar(const void* raw, int64_t V, int ndim, void* links, hipStream_t st);
int unitary_project(int ndim, int nc, int64_t V, const void* in, void* out, hipStream_t st);

// ---- fermion force (force.hip) ----------------------------------------------
struct ForcePairs;
int force_launch(const OpDesc& d, const ForcePairs& fp, void* out, hipStream_t st);

// Largest number of blocks any reducing kernel launches for volume V
// (Wilson blocks have at least 64 threads).
inline int64_t max_red_blocks(int64_t V) {
  const int64_t g = V / 64 + kNumXcd + 1;
  return g > 2048 ? g : 2048;
}
constexpr int kMaxRed = 4;  // most values any kernel reduces
constexpr int kTileMaxThreads = 1024;   // LDS-tiled Wilson block size limit (complex64 / 16-bit, one hop in flight)
constexpr int kTileMaxThreads64 = 640;  // the same for complex128 (166 VGPRs: at most 10 waves per block)
constexpr int kTileMaxThreadsMr = 640;  // multi-RHS tile
// blocks that keep several hops in flight need more registers per thread:
constexpr int kTileHifThreads32 = 512;  // complex64 / 16-bit, HIF 2 and 4
constexpr int kTileHifThreads64 = 320;  // complex128, HIF 2
// LDS bytes per tiled block (gfx950 has 160 KiB per CU; launches above the
// default 64 KiB raise the kernel's dynamic-LDS limit first)
constexpr int64_t kTileMaxLds = 150 * 1024;
constexpr int64_t kTileMaxLdsMr = 64 * 1024;  // multi-RHS tile
inline int tile_hif_threads(int prec) { return prec == P64 ? kTileHifThreads64 : kTileHifThreads32; }
// adjoint representation (matrix spinors, Ns x Nc x Nc components): its site
// arithmetic holds ~220 (complex64) / ~370 (complex128) VGPRs, so adjoint
// tiles are blocks of at most 256 sites with one hop in flight
constexpr int kTileAdjThreads = 256;
inline int tile_max_threads(int adj, int prec) {
  return adj ? kTileAdjThreads : (prec == P64 ? kTileMaxThreads64 : kTileMaxThreads);
}

}  // namespace lym
