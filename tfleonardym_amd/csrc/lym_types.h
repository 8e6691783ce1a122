// Complex arithmetic, field storage views and lattice indexing for the
// Wilson-fermion kernels.  Everything here is __host__ __device__ so the same
// site arithmetic can be compiled into the CPU check library.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace lym {

enum Prec : int { P64 = 0, P32 = 1, P16 = 2 };

template <typename R>
struct cpx {
  R re, im;
};

template <typename R>
__host__ __device__ __forceinline__ cpx<R> mk(R re, R im) {
  cpx<R> c;
  c.re = re;
  c.im = im;
  return c;
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> czero() {
  return mk<R>(R(0), R(0));
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) {
  return mk<R>(a.re + b.re, a.im + b.im);
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) {
  return mk<R>(a.re - b.re, a.im - b.im);
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R>& operator+=(cpx<R>& a, cpx<R> b) {
  a.re += b.re;
  a.im += b.im;
  return a;
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R>& operator-=(cpx<R>& a, cpx<R> b) {
  a.re -= b.re;
  a.im -= b.im;
  return a;
}
// a * b + c, one rounding (v_fma_f32 / v_fma_f64 on the device).  The
// accumulating complex products below chain two of these per real part
// (acc = fma(-z, w, fma(x, y, acc))): 4 instructions per complex
// multiply-add, where "acc += x*y - z*w" costs 6 (the sum of the products
// cannot be reassociated into the accumulator).
template <typename R>
__host__ __device__ __forceinline__ R fmad(R a, R b, R c) {
  if constexpr (sizeof(R) == 4) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}
// a * b
template <typename R>
__host__ __device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return mk<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// conj(a) * b
template <typename R>
__host__ __device__ __forceinline__ cpx<R> cmulc(cpx<R> a, cpx<R> b) {
  return mk<R>(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}
// a * conj(b)
template <typename R>
__host__ __device__ __forceinline__ cpx<R> cmul_c(cpx<R> a, cpx<R> b) {
  return mk<R>(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}
// acc += a * b
template <typename R>
__host__ __device__ __forceinline__ void cfma(cpx<R>& acc, cpx<R> a, cpx<R> b) {
  acc.re = fmad(-a.im, b.im, fmad(a.re, b.re, acc.re));
  acc.im = fmad(a.im, b.re, fmad(a.re, b.im, acc.im));
}
// acc += conj(a) * b
template <typename R>
__host__ __device__ __forceinline__ void cfmac(cpx<R>& acc, cpx<R> a, cpx<R> b) {
  acc.re = fmad(a.im, b.im, fmad(a.re, b.re, acc.re));
  acc.im = fmad(-a.im, b.re, fmad(a.re, b.im, acc.im));
}
// acc += a * conj(b)
template <typename R>
__host__ __device__ __forceinline__ void cfma_c(cpx<R>& acc, cpx<R> a, cpx<R> b) {
  acc.re = fmad(a.im, b.im, fmad(a.re, b.re, acc.re));
  acc.im = fmad(-a.re, b.im, fmad(a.im, b.re, acc.im));
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> cscale(R s, cpx<R> a) {
  return mk<R>(s * a.re, s * a.im);
}
// multiply by i^k (k = 0..3) -- exact: only swaps and sign flips
template <int K, typename R>
__host__ __device__ __forceinline__ cpx<R> ipow(cpx<R> a) {
  if constexpr ((K & 3) == 0) return a;
  else if constexpr ((K & 3) == 1) return mk<R>(-a.im, a.re);
  else if constexpr ((K & 3) == 2) return mk<R>(-a.re, -a.im);
  else return mk<R>(a.im, -a.re);
}

// Element s of component plane `plane`: the plane base is uniform across the
// wavefront and the per-lane part is a 32-bit byte offset, so the loads use a
// scalar base address plus one 32-bit VGPR per site (fields < 4 GiB/plane).
template <typename T>
__host__ __device__ __forceinline__ T& plane_at(T* base, int64_t plane, int64_t V, int64_t s) {
  char* p = (char*)(base + plane * V);
  return *(T*)(p + (uint32_t)((uint32_t)s * (uint32_t)sizeof(T)));
}

// Device loads/stores of plane elements through raw buffer instructions: the
// buffer resource (4 SGPRs) and the plane offset (soffset SGPR) are uniform,
// the site offset is one 32-bit VGPR; accesses past `bytes` read 0 / are
// dropped by the hardware range check.  Fields must be < 4 GiB.
template <typename T>
__device__ __forceinline__ T buf_ld(const T* base, uint32_t bytes, int64_t plane, int64_t V, int64_t s) {
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000);
  const int voff = (int)((uint32_t)s * (uint32_t)sizeof(T));
  const int soff = (int)((uint32_t)(plane * V) * (uint32_t)sizeof(T));
  T out;
  if constexpr (sizeof(T) == 16) {
    auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    __builtin_memcpy(&out, &v, 16);
  } else if constexpr (sizeof(T) == 8) {
    auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    __builtin_memcpy(&out, &v, 8);
  } else {
    static_assert(sizeof(T) == 4, "buffer element size");
    auto v = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
    __builtin_memcpy(&out, &v, 4);
  }
  return out;
}

// Cache policy of the field stores: nontemporal (the NT bit of the buffer
// instruction's cache-policy operand).  Outputs are written once per launch;
// measured on MI355X, nontemporal stores shorten the Wilson kernels by 3-6%
// (a D launch at 24^3x48 and 32^3x64, complex64), while nontemporal loads
// lengthen them (neighbours re-read the lines through L2).
constexpr int kStorePolicy = 2;

template <typename T>
__device__ __forceinline__ void buf_st(T* base, uint32_t bytes, int64_t plane, int64_t V, int64_t s, const T& val) {
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000);
  const int voff = (int)((uint32_t)s * (uint32_t)sizeof(T));
  const int soff = (int)((uint32_t)(plane * V) * (uint32_t)sizeof(T));
  if constexpr (sizeof(T) == 16) {
    // 16-byte stores take the whole offset in the VGPR and soffset = 0.
    // With an SGPR soffset the compiler inserts no wait state between the
    // store and a VALU write of its data registers, and on gfx950 such a
    // write right after buffer_store_dwordx4 was measured to corrupt the
    // stored value (16-bit packed stores from 512-thread blocks, two waves
    // per SIMD); with soffset = 0 the compiler's hazard guard applies.
    __attribute__((ext_vector_type(4))) unsigned int v;
    __builtin_memcpy(&v, &val, 16);
    __builtin_amdgcn_raw_buffer_store_b128(v, r, voff + soff, 0, kStorePolicy);
  } else if constexpr (sizeof(T) == 8) {
    __attribute__((ext_vector_type(2))) unsigned int v;
    __builtin_memcpy(&v, &val, 8);
    __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, kStorePolicy);
  } else {
    static_assert(sizeof(T) == 4, "buffer element size");
    unsigned int v;
    __builtin_memcpy(&v, &val, 4);
    __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, soff, kStorePolicy);
  }
}

// element load/store: buffer instructions in the device pass, plain C++ on the host
template <typename T>
__host__ __device__ __forceinline__ T ld(const T* base, uint32_t bytes, int64_t plane, int64_t V, int64_t s) {
#if defined(__HIP_DEVICE_COMPILE__)
  return buf_ld<T>(base, bytes, plane, V, s);
#else
  (void)bytes;
  return plane_at(base, plane, V, s);
#endif
}
template <typename T>
__host__ __device__ __forceinline__ void put(T* base, uint32_t bytes, int64_t plane, int64_t V, int64_t s, const T& v) {
#if defined(__HIP_DEVICE_COMPILE__)
  buf_st<T>(base, bytes, plane, V, s, v);
#else
  (void)bytes;
  plane_at(base, plane, V, s) = v;
#endif
}

// ---------------------------------------------------------------------------
// Storage of spinor fields: component planes, site index fastest.
//   P64 / P32: component c of site s lives at  base[c * V + s]
//   (complex128 / complex64, interleaved re/im).
// P16: 16-bit fixed point scaled per site (value = int16 * norm[s],
// norm[s] = max |component| / 32767): every component of a site keeps 15
// bits below the site maximum, 16x finer than IEEE half for the same 4 bytes
// per complex number.  The integers are packed G complex numbers per
// 4G-byte element (G = 4 when NCOMP is a multiple of 4, else 2): component c
// of site s is short2 number (c / G * V + s) * G + c % G, so one lane reads
// a site's NCOMP components in NCOMP / G loads of 16 (8) bytes and a wave
// reads 1 KiB (512 B) contiguous per load.  The layout is internal to the
// solvers (their Krylov vectors); the caller's fields are complex128.
// ---------------------------------------------------------------------------
template <int NCOMP>
struct Pack16 {
  static constexpr int G = (NCOMP % 4 == 0) ? 4 : 2;
  struct alignas(4 * G) unit {
    short v[2 * G];
  };
};
// short2 index of component c of site s (packed 16-bit layout)
__host__ __device__ __forceinline__ int64_t p16_index(int ncomp, int c, int64_t V, int64_t s) {
  const int g = (ncomp % 4 == 0) ? 4 : 2;
  return ((int64_t)(c / g) * V + s) * g + c % g;
}
template <int P>
struct Store;

// lds_elem: a decoded component as the LDS-tiled kernel keeps it
template <>
struct Store<P64> {
  using R = double;
  using elem = double2;
  using lds_elem = double2;
  static constexpr int bytes = 16;
};
template <>
struct Store<P32> {
  using R = float;
  using elem = float2;
  using lds_elem = float2;
  static constexpr int bytes = 8;
};
template <>
struct Store<P16> {
  using R = float;
  using elem = short2;
  using lds_elem = float2;
  static constexpr int bytes = 4;
};

template <int P>
struct SpinorView {
  using R = typename Store<P>::R;
  using E = typename Store<P>::elem;
  E* data;
  float* norm;  // P16 only
  int64_t V;

  template <int NCOMP>
  __host__ __device__ __forceinline__ void load(int64_t s, cpx<R> (&v)[NCOMP]) const {
    const uint32_t bytes = (uint32_t)(NCOMP * V * (int64_t)sizeof(E));
    if constexpr (P == P16) {
      const float n = ld<float>(norm, (uint32_t)(V * 4), 0, V, s);
      load_units<NCOMP>(s, v, n);
    } else {
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) {
        const E e = ld<E>(data, bytes, c, V, s);
        v[c] = mk<R>(e.x, e.y);
      }
    }
  }
  // neighbour load of a hop (mu, dir 0 = forward, 1 = backward): global here;
  // the LDS tile view overrides it
  template <int NCOMP>
  __host__ __device__ __forceinline__ void load_hop(int, int, int64_t s, cpx<R> (&v)[NCOMP]) const {
    load<NCOMP>(s, v);
  }
  // 16-bit storage: the hop kernels take a neighbour's integers as floats and
  // its norm apart (value = v * n) and apply n once to the transported
  // half-spinor, inside the accumulating FMAs (Wilson hops are linear in v)
  static constexpr bool kScaled = (P == P16);
  // the packed integers of site s as floats, each scaled by sc
  template <int NCOMP>
  __host__ __device__ __forceinline__ void load_units(int64_t s, cpx<R> (&v)[NCOMP], float sc) const {
    using U = typename Pack16<NCOMP>::unit;
    constexpr int G = Pack16<NCOMP>::G;
    const uint32_t bytes = (uint32_t)(NCOMP * V * (int64_t)sizeof(E));
#pragma unroll
    for (int q = 0; q < NCOMP / G; ++q) {
      const U u = ld<U>((const U*)data, bytes, q, V, s);
#pragma unroll
      for (int g = 0; g < G; ++g) v[q * G + g] = mk<R>((float)u.v[2 * g] * sc, (float)u.v[2 * g + 1] * sc);
    }
  }
  // 16-bit storage: site s's packed integer units and its norm, undecoded
  template <int NCOMP>
  __device__ __forceinline__ void load_packed(int64_t s, typename Pack16<NCOMP>::unit (&u)[NCOMP / Pack16<NCOMP>::G],
                                              float& n) const {
    using U = typename Pack16<NCOMP>::unit;
    const uint32_t bytes = (uint32_t)(NCOMP * V * (int64_t)sizeof(E));
    n = ld<float>(norm, (uint32_t)(V * 4), 0, V, s);
#pragma unroll
    for (int q = 0; q < NCOMP / Pack16<NCOMP>::G; ++q) u[q] = ld<U>((const U*)data, bytes, q, V, s);
  }
  template <int NCOMP>
  __host__ __device__ __forceinline__ void load_raw(int64_t s, cpx<R> (&v)[NCOMP], R& n) const {
    if constexpr (P == P16) {
      n = ld<float>(norm, (uint32_t)(V * 4), 0, V, s);
      load_units<NCOMP>(s, v, 1.0f);
    } else {
      n = R(1);
      load<NCOMP>(s, v);
    }
  }
  template <int NCOMP>
  __host__ __device__ __forceinline__ void load_hop_scaled(int, int, int64_t s, cpx<R> (&v)[NCOMP], R& n) const {
    load_raw<NCOMP>(s, v, n);
  }
  template <int NCOMP>
  __host__ __device__ __forceinline__ void store(int64_t s, const cpx<R> (&v)[NCOMP]) const {
    const uint32_t bytes = (uint32_t)(NCOMP * V * (int64_t)sizeof(E));
    if constexpr (P == P16) {
      float m = 0.f;
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) {
        m = fmaxf(m, fmaxf(fabsf(v[c].re), fabsf(v[c].im)));
      }
      // |v / scale| <= 32767 by construction; rintf rounds to nearest
      const float scale = m * (1.0f / 32767.0f);
      const float inv = m > 0.f ? 32767.0f / m : 0.f;
      put<float>(norm, (uint32_t)(V * 4), 0, V, s, scale);
      using U = typename Pack16<NCOMP>::unit;
      constexpr int G = Pack16<NCOMP>::G;
#pragma unroll
      for (int q = 0; q < NCOMP / G; ++q) {
        U u;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          u.v[2 * g] = (short)fminf(fmaxf(rintf(v[q * G + g].re * inv), -32767.f), 32767.f);
          u.v[2 * g + 1] = (short)fminf(fmaxf(rintf(v[q * G + g].im * inv), -32767.f), 32767.f);
        }
        put<U>((U*)data, bytes, q, V, s, u);
      }
    } else {
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) {
        E e;
        e.x = v[c].re;
        e.y = v[c].im;
        put<E>(data, bytes, c, V, s, e);
      }
    }
  }
};

// Links.  complex128: (D, Nc, Nc, V) complex, site fastest (the caller's
// layout).  complex64: packed in 16-byte units, (D, U, V) with U = ceil(Nc^2 / 2)
// units per link: unit q of (mu, s) holds elements 2q and 2q + 1 of the
// row-major matrix (the second half of the last unit is padding when Nc^2 is
// odd), so a lane reads a link in half the load instructions; rows 0..NC-2
// are the first NC (NC - 1) / 2 units (NC (NC - 1) is even).
// recon = 1: the links are special unitary; only rows 0..NC-2 are read and
// the last row is rebuilt (SU(3): u2 = conj(u0 x u1); SU(2): u1 = (-u01*, u00*)),
// one third (SU(3)) or half (SU(2)) fewer link bytes per hop.
__host__ __device__ constexpr int link_units32(int nc) { return (nc * nc + 1) / 2; }

template <typename R>
struct LinkView {
  const cpx<R>* data;
  int64_t V;
  int recon = 0;
  template <int NC>
  __host__ __device__ __forceinline__ void load(int mu, int64_t s, cpx<R> (&u)[NC][NC]) const {
    if constexpr (sizeof(R) == 4) {
      load_packed<NC>(mu, s, u);
    } else {
      const uint32_t bytes = (uint32_t)((int64_t)(mu + 1) * NC * NC * V * (int64_t)sizeof(cpx<R>));
#pragma unroll
      for (int i = 0; i < NC - 1; ++i)
#pragma unroll
        for (int j = 0; j < NC; ++j) u[i][j] = ld<cpx<R>>(data, bytes, (int64_t)(mu * NC + i) * NC + j, V, s);
      if (recon) {
        rebuild<NC>(u);
      } else {
#pragma unroll
        for (int j = 0; j < NC; ++j)
          u[NC - 1][j] = ld<cpx<R>>(data, bytes, (int64_t)(mu * NC + NC - 1) * NC + j, V, s);
      }
    }
  }
  // complex64 units (q0 <= q < q1) of link (mu, s) into u
  template <int NC>
  __host__ __device__ __forceinline__ void load_units(int mu, int64_t s, int q0, int q1, uint32_t bytes,
                                                      cpx<R> (&u)[NC][NC]) const {
    constexpr int QU = link_units32(NC);
#pragma unroll
    for (int q = q0; q < q1; ++q) {
      const float4 w = ld<float4>((const float4*)data, bytes, (int64_t)mu * QU + q, V, s);
      u[(2 * q) / NC][(2 * q) % NC] = mk<R>(w.x, w.y);
      if (2 * q + 1 < NC * NC) u[(2 * q + 1) / NC][(2 * q + 1) % NC] = mk<R>(w.z, w.w);
    }
  }
  template <int NC>
  __host__ __device__ __forceinline__ void load_packed(int mu, int64_t s, cpx<R> (&u)[NC][NC]) const {
    constexpr int QU = link_units32(NC), QR = NC * (NC - 1) / 2;
    const uint32_t bytes = (uint32_t)((int64_t)(mu + 1) * QU * V * 16);
    load_units<NC>(mu, s, 0, QR, bytes, u);
    if (recon)
      rebuild<NC>(u);
    else
      load_units<NC>(mu, s, QR, QU, bytes, u);
  }
  // the last row of a special unitary link from rows 0..NC-2
  template <int NC>
  __host__ __device__ __forceinline__ static void rebuild(cpx<R> (&u)[NC][NC]) {
    if constexpr (NC == 3) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        // u[2][j] = conj(u0a u1b - u0b u1a)
        const int a = (j + 1) % 3, b = (j + 2) % 3;
        const cpx<R> x = u[0][a], y = u[1][b], z = u[0][b], w = u[1][a];
        const R re = fmad(x.re, y.re, fmad(-x.im, y.im, fmad(-z.re, w.re, z.im * w.im)));
        const R im = fmad(z.re, w.im, fmad(z.im, w.re, fmad(-x.re, y.im, -(x.im * y.re))));
        u[2][j] = mk<R>(re, im);
      }
    } else if constexpr (NC == 2) {
      u[1][0] = mk<R>(-u[0][1].re, u[0][1].im);
      u[1][1] = mk<R>(u[0][0].re, -u[0][0].im);
    }
  }
};

// [p, p + bytes) lies inside the allocation [base, base + size) (the host
// check of buffer arguments against their allocations)
__host__ __device__ inline bool range_inside(uintptr_t base, size_t size, uintptr_t p, size_t bytes) {
  return p >= base && p - base <= size && bytes <= size - (p - base);
}

// Lattice description passed by value to every Wilson kernel.
struct Lat {
  int L[4];
  int stride[4];
  int64_t V;
  double kappa;
  double bc_re, bc_im;  // factor on U_{D-1}(x) at x_{D-1} = L_{D-1}-1
  int herm;             // 1: multiply the result by gamma5
  int has_bc;           // bc factor != 1
};

// Block schedule of the Wilson kernels.  tiled = 1 (4D lattices that allow it):
// a block covers k whole lines of the fastest direction (T = L_3 sites each,
// so loads stay coalesced), XCD x owns the slab x0 in [x*slab, (x+1)*slab),
// and inside an XCD the blocks run with x0 fastest, then groups of k x2
// values, then x1 -- every neighbour a block reads was touched by a block a
// few positions earlier on the same XCD, i.e. is still in that XCD's L2.
// tiled = 0: consecutive sites, block ranges contiguous per XCD.
struct SiteMap {
  int tiled;
  int T;     // L_3
  int k;     // lines per block
  int slab;  // x0 values per XCD
  int ng;    // L_2 / k
  int walk;  // order of an XCD's blocks after x0: 0 = x2 groups, then x1; 1 = x1, then x2 groups
};

// LDS-tiled Wilson kernel (4D): a block stages the input spinor of a tile of
// a0 x a1 x k x T sites (whole lines of the fastest direction) in LDS and
// serves every hop whose neighbour is inside the tile from there.  XCD x owns
// x0 in [x*slab, (x+1)*slab); within an XCD tiles run x0 fastest, then x2,
// then x1.  on = 0: not applicable (the plain kernel is used).
struct TileMap {
  int on;
  int a0, a1, k, T;
  int slab;
  int n0, n1, n2;  // tiles per XCD along x0, x1, x2
  int nsite;       // a0 * a1 * k * T (block size)
  int blocks;      // tiles in the lattice
  int hif;         // hops in flight per wave: 1, 2 or 4 (0 = 1)
  int walk;        // order of an XCD's tiles: 0 = x0 fastest, then x2, x1; 1 = x0, then x1, x2
};

template <int D>
__host__ __device__ __forceinline__ void coords(const Lat& lat, int64_t s, int (&x)[D]) {
  uint32_t r = (uint32_t)s;
#pragma unroll
  for (int mu = D - 1; mu >= 1; --mu) {
    const uint32_t L = (uint32_t)lat.L[mu];
    const uint32_t q = r / L;
    x[mu] = (int)(r - q * L);
    r = q;
  }
  x[0] = (int)r;
}

}  // namespace lym
