// Shared device helpers for the even-odd Wilson-clover kernels (gfx950).
//
// Field layouts in HBM (R = double or float, C2 = R pair re/im):
//   half field  (one parity)  : C2 psi[12][Vh]        component k = 3*spin + colour
//   full field                : C2 psi[2][12][Vh]     parity-major
//   links, output parity p    : C2 W[p][8][NL][Vh]    dir 2*mu   : U_mu(x)
//                                                      dir 2*mu+1 : U_mu(x-mu)^dag
//                               NL = 9 (3x3) or 6 (first two rows; row 2 is
//                               rebuilt as conj(row0 x row1))
//   clover / inverse, parity p: C2 A[p][36][Vh]       18 slots per chiral block:
//                               3 slots of packed real diagonals, then the 15
//                               strictly-lower entries (1,0),(2,0),(2,1),(3,0)..
// Half-site index h = (((t*Lz + z)*Ly + y)*Lx + x) / 2 (x collapsed).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

template <typename R> struct c2;
template <> struct c2<double> { using t = double2; };
template <> struct c2<float> { using t = float2; };

template <typename C> __device__ __forceinline__ C cmk(decltype(C::x) a, decltype(C::x) b) {
  C r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
  return cmk<C>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) {
  return cmk<C>(a.x * s, a.y * s);
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmk<C>(a.x, -a.y); }
// i * a
template <typename C> __device__ __forceinline__ C cimul(C a) { return cmk<C>(-a.y, a.x); }

// Lattice geometry shared by all kernels.
struct Geom {
  int L[4];   // Lx, Ly, Lz, Lt
  int Lxh;    // Lx / 2
  long Vh;    // half volume
  long V;     // full volume
};

// coordinates of half-site h of parity p
__device__ __forceinline__ void half_coords(const Geom &g, long h, int p, int c[4]) {
  long r = h / g.Lxh;
  int xh = (int)(h - r * g.Lxh);
  long r2 = r / g.L[1];
  c[1] = (int)(r - r2 * g.L[1]);
  long r3 = r2 / g.L[2];
  c[2] = (int)(r2 - r3 * g.L[2]);
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

template <typename C> __device__ __forceinline__ C bload(wc_rsrc r, uint32_t voff, uint32_t soff);
template <> __device__ __forceinline__ double2 bload<double2>(wc_rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
}
template <> __device__ __forceinline__ float2 bload<float2>(wc_rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
