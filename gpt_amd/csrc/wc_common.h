// Shared device helpers for the even-odd Wilson-clover kernels (gfx950).
//
// Field layouts in HBM (R = double or float, C2 = R pair re/im).  Every
// per-site array is AoSoA with blocks of 64 sites: entry k of an NK-entry
// object at half-site h lives at ((h / 64) * NK + k) * 64 + h % 64, so one
// wave's 64 sites read each entry as 1 KB (FP64) contiguous and all of a
// wave's entries form one contiguous run.
//   half field  (one parity)  : NK = 12, k = 3*spin + colour
//   full field                : two half fields, even parity first
//   links, output parity p    : NK = 8*NL, k = dir*NL + e; dir 2*mu : U_mu(x),
//                               dir 2*mu+1 : U_mu(x-mu)^dag; NL = 9 (3x3) or
//                               6 (first two rows; row 2 = conj(row0 x row1))
//   clover / inverse, parity p: NK = 36, 18 slots per chiral block: 3 slots
//                               of packed real diagonals, then the 15
//                               strictly-lower entries (1,0),(2,0),(2,1),(3,0)..
// Half-site index h = (((t*Lz + z)*Ly + y)*Lx + x) / 2 (x collapsed); the
// half volume Vh must be a multiple of 64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

template <typename R> struct c2;
template <> struct c2<double> { using t = double2; };
template <> struct c2<float> { using t = float2; };

// 16-bit storage of a complex FP32 value (4 bytes), used for the search
// direction and the Schur intermediates of the mixed-precision inner solve;
// arithmetic is always done after widening to FP32.  c2h holds IEEE half
// precision (11-bit significand, range 6e-5 .. 65504), c2b bfloat16 (8-bit
// significand, the FP32 range).  Narrowing rounds to nearest even.
struct __align__(4) c2h {
  _Float16 x, y;
};
struct __align__(4) c2b {
  __bf16 x, y;
};
// storage codes of the 16-bit fields (kernel arguments `half`, HopLaunch
// half_in / half_out): 0 the working precision, 1 c2h, 2 c2b
enum { kS16None = 0, kS16Half = 1, kS16Bf16 = 2 };
template <typename S> constexpr bool is16() { return sizeof(S) == 4; }
// complex type a stored value is computed in
template <typename S> struct wide;
template <> struct wide<double2> { using t = double2; };
template <> struct wide<float2> { using t = float2; };
template <> struct wide<c2h> { using t = float2; };
template <> struct wide<c2b> { using t = float2; };
template <typename C> __host__ __device__ __forceinline__ C widen(C v) { return v; }
__device__ __forceinline__ float2 widen(c2h v) { return make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ float2 widen(c2b v) { return make_float2((float)v.x, (float)v.y); }
template <typename S, typename C> __device__ __forceinline__ S narrow(C v) { return v; }
template <> __device__ __forceinline__ c2h narrow<c2h, float2>(float2 v) {
  c2h o;
  o.x = (_Float16)v.x;
  o.y = (_Float16)v.y;
  return o;
}
template <> __device__ __forceinline__ c2b narrow<c2b, float2>(float2 v) {
  c2b o;
  o.x = (__bf16)v.x;
  o.y = (__bf16)v.y;
  return o;
}

template <typename C> __device__ __forceinline__ C cmk(decltype(C::x) a, decltype(C::x) b) {
  C r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
// Products are written with explicit fused multiply-adds and the kernels
// are compiled with contraction off, so every kernel instantiation (any
// launch shape, any epilogue) performs the same roundings and gives the
// same bits.
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return cmk<C>(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
// conj(a) * b
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
  return cmk<C>(fma(a.x, b.x, a.y * b.y), fma(a.x, b.y, -(a.y * b.x)));
}
// acc + a * b
template <typename C> __device__ __forceinline__ C cmadd(C acc, C a, C b) {
  return cmk<C>(fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y)));
}
// acc + conj(a) * b
template <typename C> __device__ __forceinline__ C cmaddc(C acc, C a, C b) {
  return cmk<C>(fma(a.x, b.x, fma(a.y, b.y, acc.x)), fma(a.x, b.y, fma(-a.y, b.x, acc.y)));
}
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) {
  return cmk<C>(a.x * s, a.y * s);
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmk<C>(a.x, -a.y); }
// i * a
template <typename C> __device__ __forceinline__ C cimul(C a) { return cmk<C>(-a.y, a.x); }

// A c2h value as a packed pair of halves: the 16-bit hop projects its FP16
// neighbour spinors in FP16 (packed adds; the per-RHS frames keep the values
// near 1, far from FP16's range limits) and widens the half spinors, so a
// direction converts 6 complex values instead of 12.
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ h2v cadd(h2v a, h2v b) { return a + b; }
__device__ __forceinline__ h2v csub(h2v a, h2v b) { return a - b; }
__device__ __forceinline__ h2v cimul(h2v a) { return __builtin_shufflevector(a, -a, 3, 0); }
__device__ __forceinline__ float2 widen(h2v v) { return make_float2((float)v.x, (float)v.y); }

// Lattice geometry shared by all kernels.
struct Geom {
  int L[4];   // Lx, Ly, Lz, Lt
  int Lxh;    // Lx / 2
  long Vh;    // half volume
  long V;     // full volume
};

constexpr int kBlk = 64;  // AoSoA block (sites)

// element index of entry k of an NK-entry object at half-site h
__host__ __device__ __forceinline__ long aidx(int k, long h, int NK) {
  return ((h >> 6) * (long)NK + k) * kBlk + (h & 63);
}
// same, split into the per-site part (for a VGPR offset) and the entry part
__host__ __device__ __forceinline__ long aidx_site(long h, int NK) { return (h >> 6) * (long)NK * kBlk + (h & 63); }

// coordinates of half-site h of parity p.  In 32-bit unsigned arithmetic:
// h < Vh < 2^32 (a half field must fit the 4 GB window), and a 64-bit
// division costs several times the instructions of a 32-bit one.
__device__ __forceinline__ void half_coords(const Geom &g, long h, int p, int c[4]) {
  const uint32_t hu = (uint32_t)h, lxh = (uint32_t)g.Lxh, ly = (uint32_t)g.L[1], lz = (uint32_t)g.L[2];
  const uint32_t r = hu / lxh;
  const int xh = (int)(hu - r * lxh);
  const uint32_t r2 = r / ly;
  c[1] = (int)(r - r2 * ly);
  const uint32_t r3 = r2 / lz;
  c[2] = (int)(r2 - r3 * lz);
  c[3] = (int)r3;
  c[0] = 2 * xh + ((c[1] + c[2] + c[3] + p) & 1);
}

__device__ __forceinline__ long half_index(const Geom &g, int x, int y, int z, int t) {
  return ((((long)t * g.L[2] + z) * g.L[1] + y) * g.L[0] + x) >> 1;
}

// Block index -> tile index so that each of the 8 XCDs (which receive
// workgroups round-robin) processes one contiguous chunk of the lattice and
// its L2 sees the neighbour reuse of that chunk.  Bijective for any grid size.
__device__ __forceinline__ long xcd_tile(long b, long nb) {
#if !WC_XCD_MAP
  return b;
#endif
  long q = nb >> 3, rem = nb & 7;
  long k = b & 7, i = b >> 3;
  return k * q + (k < rem ? k : rem) + i;
}

// Two-stage deterministic FP64 reduction: wave64 shuffle tree then LDS across
// the waves of the block.  Every lane of the block gets the block sum.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += lds[i];
    lds[NT / 64] = s;
  }
  __syncthreads();
  s = lds[NT / 64];
  __syncthreads();
  return s;
}

// ------------------------------------------------------------ buffer loads
// Raw buffer loads: a wave-uniform descriptor (SGPRs) plus a 32-bit per-lane
// byte offset and a uniform byte offset, so a 12-component gather costs one
// offset VGPR instead of twelve 64-bit addresses.  Out-of-range offsets read
// zero instead of faulting.  Dword 3 = 0x00020000 is the raw-buffer format
// word of the gfx9 family (gfx950 included).
typedef __amdgpu_buffer_rsrc_t wc_rsrc;

__device__ __forceinline__ wc_rsrc make_rsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

// AUX: cache-policy bits of the load (gfx950: bit 1 = nt, non-temporal)
template <typename C, int AUX = 0> struct bload_t;
template <int AUX> struct bload_t<double2, AUX> {
  static __device__ __forceinline__ double2 go(wc_rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, AUX));
  }
};
template <int AUX> struct bload_t<float2, AUX> {
  static __device__ __forceinline__ float2 go(wc_rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, AUX));
  }
};
template <int AUX> struct bload_t<c2h, AUX> {
  static __device__ __forceinline__ c2h go(wc_rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(c2h, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, AUX));
  }
};
template <int AUX> struct bload_t<c2b, AUX> {
  static __device__ __forceinline__ c2b go(wc_rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(c2b, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, AUX));
  }
};
template <typename C, int AUX = 0>
__device__ __forceinline__ C bload(wc_rsrc r, uint32_t voff, uint32_t soff) {
  return bload_t<C, AUX>::go(r, voff, soff);
}

typedef int wc_v4i __attribute__((ext_vector_type(4)));
typedef int wc_v2i __attribute__((ext_vector_type(2)));
template <typename C, int AUX = 0> struct bstore_t;
template <int AUX> struct bstore_t<double2, AUX> {
  static __device__ __forceinline__ void go(double2 v, wc_rsrc r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(wc_v4i, v), r, (int)voff, (int)soff, AUX);
  }
};
template <int AUX> struct bstore_t<float2, AUX> {
  static __device__ __forceinline__ void go(float2 v, wc_rsrc r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(wc_v2i, v), r, (int)voff, (int)soff, AUX);
  }
};
template <int AUX> struct bstore_t<c2h, AUX> {
  static __device__ __forceinline__ void go(c2h v, wc_rsrc r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)voff, (int)soff, AUX);
  }
};
template <int AUX> struct bstore_t<c2b, AUX> {
  static __device__ __forceinline__ void go(c2b v, wc_rsrc r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)voff, (int)soff, AUX);
  }
};

// compile-time tuning knobs of the hop kernel (defaults are the tuned values)
#ifndef WC_LINK_AUX
#define WC_LINK_AUX 2  // links are read once per hop: non-temporal, keeps L2 for the spinors
#endif
#ifndef WC_SPINOR_AUX
#define WC_SPINOR_AUX 0  // cache policy of the neighbour-spinor gathers
#endif
#ifndef WC_NT_STORE
#define WC_NT_STORE 1  // the hop output is not re-read by the hop: non-temporal stores
#endif
#ifndef WC_HOP_MINWAVES
#define WC_HOP_MINWAVES 1  // occupancy hint (waves per SIMD) of the hop kernel
#endif
#ifndef WC_XCD_MAP
#define WC_XCD_MAP 1  // XCD-contiguous block -> tile mapping
#endif
