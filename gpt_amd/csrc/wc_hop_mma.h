// The multi-right-hand-side hop with FP16 storage on the matrix cores
// (gfx950 v_mfma_f32_16x16x32_f16).
//
// With N right-hand sides a site's hop is a small dense product: in every
// direction the site's 3x3 link multiplies the projected half spinors of its
// neighbour, 2 spins x N right-hand sides.  Written over real numbers, with
// K running over (direction, colour, re/im), the whole hop of a site is
//   D (6 x 4N) = [U_1 ... U_8] (6 x 48) . B (48 x 4N),
// where U_d is the real form [[Re U, -Im U], [Im U, Re U]] of the link and
// the columns of B are (spin, right-hand side): the upper spins hold the
// projected h_0, h_1 of (1 -+ gamma_mu) psi, and the lower spins the
// reconstruction's multiple (+-1 or +-i) of h_0 or h_1, which commutes with
// the colour product.  The sum over directions is the product's K, so the
// accumulator ends up holding all 4 spins of the hop.
//
// One MFMA (16 x 16 x 32) covers a PAIR of sites, one direction pair
// (mu, forward and backward) and 4 right-hand sides:
//   rows (16)   : site of the pair (2) x (colour out x re/im: 6, 2 zero rows)
//   K (32)      : 4 groups of 8, group g = (site g & 1, direction 2 mu + g / 2),
//                 each (colour in x re/im: 6, 2 zeros)
//   columns (16): spin (4) x right-hand side (4)
// A is block diagonal (a row's link entries are zero in the other site's
// K groups).  Every lane supplies A row l % 16 and B column l % 16 with the
// K group l / 16, so A and B agree on K whatever the order inside a group.
//
// The FP32 links (third row rebuilt, coefficient and boundary phase folded
// in, as the VALU hop applies them) are split into FP16 hi + lo parts: two
// MFMAs per product keep 22 significant bits of every link entry.  Both parts
// are scaled so that they stay normal FP16 numbers (the matrix cores flush
// FP16 subnormals, and the lo part of an entry below 1/4 is one): hi =
// fp16(2^8 u), lo = fp16(2^11 (2^8 u - hi)), accumulated separately and
// combined as 2^-8 acc_hi + 2^-19 acc_lo.  The neighbour spinors are the
// FP16 storage itself; their projection is done in packed FP16 with the same
// single roundings as the VALU hop's, and all products are exact in the FP32
// accumulators.  So the hop differs from the VALU hop by the order of its
// FP32 sums and the 2^-22 of the links' split.
//
// A workgroup (4 waves) owns 8 consecutive output half-sites (one pair per
// wave) and runs, per chunk of at most 16 right-hand sides:
//   1. once: the sites' prepared links and EPI_CLOV blocks into LDS;
//   2. the neighbour spinors gathered (16-byte loads: 4 right-hand sides of
//      one component), projected and written to LDS as B fragments;
//   3. the MFMAs of the wave's pair, accumulators in registers;
//   4. the accumulators to LDS per (site, right-hand side);
//   5. one item per (site, right-hand side): the epilogue (clover, axpy,
//      inner product) and the store, as in k_hop.
#pragma once
#include "wc_hop.h"

namespace wck {

typedef _Float16 mh8 __attribute__((ext_vector_type(8)));
typedef float mf4 __attribute__((ext_vector_type(4)));
typedef uint32_t mu4 __attribute__((ext_vector_type(4)));

constexpr int kMmaSites = 8;     // output half-sites per workgroup: 4 pairs, one per wave
constexpr int kMmaThreads = 256;
constexpr int kMmaChunk = 16;    // right-hand sides per pass: 4 column tiles of 4
// scales of the links' FP16 hi and lo parts (powers of two: exact)
constexpr float kMmaHiScale = 256.0f;
constexpr float kMmaLoScale = 2048.0f;

// neighbour half-site of output half-site hh (coordinates c0..c3) in
// direction dir = 2 mu + fb (fb 0: +mu, 1: -mu), periodic
__device__ __forceinline__ uint32_t mma_neighbour(const Geom &g, uint32_t hh, int c0, int c1, int c2, int c3,
                                                  int dir) {
  const int mu = dir >> 1, fb = dir & 1;
  if (mu == 0) {
    const int xn = fb == 0 ? (c0 == g.L[0] - 1 ? 0 : c0 + 1) : (c0 == 0 ? g.L[0] - 1 : c0 - 1);
    return hh - ((uint32_t)c0 >> 1) + ((uint32_t)xn >> 1);
  }
  const uint32_t st1 = (uint32_t)g.Lxh, st2 = st1 * (uint32_t)g.L[1], st3 = st2 * (uint32_t)g.L[2];
  const uint32_t stride = mu == 1 ? st1 : (mu == 2 ? st2 : st3);
  const int cm = mu == 1 ? c1 : (mu == 2 ? c2 : c3);
  const int Lm = mu == 1 ? g.L[1] : (mu == 2 ? g.L[2] : g.L[3]);
  const uint32_t wrap = (uint32_t)(Lm - 1) * stride;
  if (fb == 0) return (cm == Lm - 1) ? hh - wrap : hh + stride;
  return (cm == 0) ? hh + wrap : hh - stride;
}

// f * v for f in {1, -1, i, -i} on a complex FP16 pair held as one dword:
// swap (f = +-i) takes (re, im) to (-im, re), neg negates both halves
__device__ __forceinline__ uint32_t mma_times(uint32_t v, bool swap, bool neg) {
  const uint32_t iv = ((v >> 16) ^ 0x8000u) | (v << 16);  // i * (re, im) = (-im, re)
  const uint32_t t = swap ? iv : v;
  return neg ? t ^ 0x80008000u : t;
}
__device__ __forceinline__ uint32_t h2_add(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(h2v, a) + __builtin_bit_cast(h2v, b));
}

template <int NL, int DAG, int EPI, typename SO>
__global__ __launch_bounds__(kMmaThreads, 4) void k_hop_mma(const HopLaunch a) {
  using C = float2;
  if (a.halt && *a.halt) return;
  const Geom &g = a.g;
  const int N = a.nrhs;
  const int tid = threadIdx.x;
  const long site0 = xcd_tile(blockIdx.x, gridDim.x) * kMmaSites;
  __shared__ C lw[kMmaSites * 8 * 9];        // prepared links [site][dir][9]
  __shared__ C lc[36 * kMmaSites];           // EPI_CLOV blocks [slot][site]
  __shared__ mu4 lb[kMmaSites * 8 * 2 * kMmaChunk];  // B fragments [site][dir][spin 0/1][rhs], 16 B each
  C *lo = (C *)lb;                           // accumulators [site][rhs][12] (after the MFMAs)

  // 1. links of the workgroup's sites, prepared, and the clover blocks
  if (tid < 8 * kMmaSites) {
    const int dir = tid / kMmaSites, s = tid % kMmaSites;
    const long h = site0 + s;
    int c[4];
    half_coords(g, h, a.pout, c);
    C u[9];
#pragma unroll
    for (int e = 0; e < NL; ++e) u[e] = ((const C *)a.W)[aidx(dir * NL + e, h, 8 * NL)];
    prepare_link<float, NL>(a, dir >> 1, dir & 1, c, u);
#pragma unroll
    for (int e = 0; e < 9; ++e) lw[(s * 8 + dir) * 9 + e] = u[e];
  }
  const bool clov_staged = (EPI & EPI_CLOV) && a.clov && !a.clov_roff;
  if (clov_staged)
    for (int i = tid; i < 36 * kMmaSites; i += kMmaThreads)
      lc[i] = ((const C *)a.clov)[aidx(i / kMmaSites, site0 + i % kMmaSites, 36)];

  const wc_rsrc in = make_rsrc(a.in, 12u * (uint32_t)g.Vh * (uint32_t)N * (uint32_t)sizeof(c2h));
  const int w = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;

  for (int r0 = 0; r0 < N; r0 += kMmaChunk) {
    const int nc = min(kMmaChunk, N - r0);  // a multiple of 4
    const int nq = nc >> 2;
    // 2. gather and project: task (dir, upper spin su, site s, rhs quad q),
    // q fastest so that neighbouring lanes read neighbouring 16 bytes
    const int ntask = 8 * 2 * kMmaSites * nq;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int tau = tid + k * kMmaThreads;
      if (tau < ntask) {
        const int q = tau % nq, rest = tau / nq;
        const int s = rest % kMmaSites, su = (rest / kMmaSites) & 1, dir = rest / (2 * kMmaSites);
        const int mu = dir >> 1, fb = dir & 1;
        const long h = site0 + s;
        int c[4];
        half_coords(g, h, a.pout, c);
        const uint32_t nb = mma_neighbour(g, (uint32_t)h, c[0], c[1], c[2], c[3], dir);
        // h_su = p_su + f p_pt, the upper spin of (1 - S gamma_mu) psi
        const int S = ((fb == 0) != (DAG != 0)) ? 1 : -1;
        const int pt = mu < 2 ? 3 - su : 2 + su;
        const bool swap = (mu & 1) == 0;  // mu 0, 2: f = +-i
        bool neg;
        if (mu == 0) neg = S > 0;                     // f = -S i
        else if (mu == 1) neg = su == 0 ? S < 0 : S > 0;  // f = S, -S
        else if (mu == 2) neg = su == 0 ? S > 0 : S < 0;  // f = -S i, S i
        else neg = S > 0;                             // f = -S
        const uint32_t vo = ((uint32_t)aidx_site(nb, 12) * (uint32_t)N + (uint32_t)(r0 + 4 * q)) * 4u;
        const uint32_t ks = (uint32_t)kBlk * (uint32_t)N * 4u;  // bytes between components
        // the spins differ between lanes: their offsets go in the per-lane
        // (VGPR) offset, the uniform (SGPR) one stays 0
        const uint32_t vs = vo + (uint32_t)(su * 3) * ks, vt = vo + (uint32_t)(pt * 3) * ks;
        mu4 p[3], t[3];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          p[cc] = __builtin_bit_cast(
              mu4, __builtin_amdgcn_raw_buffer_load_b128(in, (int)(vs + (uint32_t)cc * ks), 0, WC_SPINOR_AUX));
          t[cc] = __builtin_bit_cast(
              mu4, __builtin_amdgcn_raw_buffer_load_b128(in, (int)(vt + (uint32_t)cc * ks), 0, WC_SPINOR_AUX));
        }
        mu4 *dst = lb + ((s * 8 + dir) * 2 + su) * kMmaChunk + 4 * q;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          mu4 o;
          o.x = h2_add(p[0][rr], mma_times(t[0][rr], swap, neg));
          o.y = h2_add(p[1][rr], mma_times(t[1][rr], swap, neg));
          o.z = h2_add(p[2][rr], mma_times(t[2][rr], swap, neg));
          o.w = 0u;
          dst[rr] = o;
        }
      }
    }
    __syncthreads();

    // 3. the MFMAs of wave w's site pair (sites 2w, 2w + 1), hi and lo link
    // parts into separate accumulators
    mf4 acc[4], acl[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = acl[t] = mf4{0.f, 0.f, 0.f, 0.f};
    {
      const int ks = lg & 1, dl = lg >> 1;          // K group: site of the pair, forward / backward
      const int sk = 2 * w + ks;                    // that site in the workgroup
      const int arow = (li & 7) >> 1, apart = li & 1;
      const bool aon = ((li >> 3) == ks) && arow < 3;  // A: the row's site owns this K group
      const int sp = li >> 2, rr = li & 3;          // B column: spin, right-hand side in the tile
      const int S = ((dl == 0) != (DAG != 0)) ? 1 : -1;
#pragma unroll
      for (int mu = 0; mu < 4; ++mu) {
        const int dir = 2 * mu + dl;
        float v[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (aon) {
          const C *u = lw + (sk * 8 + dir) * 9 + 3 * arow;
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            const C x = u[cc];
            v[2 * cc] = apart == 0 ? x.x : x.y;       // Re row: (Re U, -Im U); Im row: (Im U, Re U)
            v[2 * cc + 1] = apart == 0 ? -x.y : x.x;
          }
        }
        mh8 ahi, alo;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          const float x = v[j] * kMmaHiScale;
          const _Float16 hi = (_Float16)x;
          ahi[j] = hi;
          alo[j] = (_Float16)((x - (float)hi) * kMmaLoScale);
        }
        ahi[6] = ahi[7] = alo[6] = alo[7] = (_Float16)0.f;
        // B: which projected spin this column reads, and the factor f
        int su;
        bool swap = false, neg = false;
        if (sp < 2) {
          su = sp;
        } else if (mu == 0) {  // a2 += S i c1, a3 += S i c0
          su = 3 - sp;
          swap = true;
          neg = S < 0;
        } else if (mu == 1) {  // a2 -= S c1, a3 += S c0
          su = 3 - sp;
          neg = sp == 2 ? S > 0 : S < 0;
        } else if (mu == 2) {  // a2 += S i c0, a3 -= S i c1
          su = sp - 2;
          swap = true;
          neg = sp == 2 ? S < 0 : S > 0;
        } else {  // a2 -= S c0, a3 -= S c1
          su = sp - 2;
          neg = S > 0;
        }
        const mu4 *src = lb + ((sk * 8 + dir) * 2 + su) * kMmaChunk + rr;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (t < nq) {
            mu4 b = src[4 * t];
            b.x = mma_times(b.x, swap, neg);
            b.y = mma_times(b.y, swap, neg);
            b.z = mma_times(b.z, swap, neg);
            const mh8 bh = __builtin_bit_cast(mh8, b);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bh, acc[t], 0, 0, 0);
            acl[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, bh, acl[t], 0, 0, 0);
          }
        }
      }
    }
    __syncthreads();  // every wave is done with the B fragments

    // 4. accumulators to LDS: lane (column li, row group lg) holds rows
    // 4 lg .. 4 lg + 3 = site 2w + lg / 2, (c0, c1) or (c2, zero rows)
    {
      const int s = 2 * w + (lg >> 1), sp = li >> 2, rr = li & 3;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (t < nq) {
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[t][j] = acc[t][j] * (1.0f / kMmaHiScale) +
                                                  acl[t][j] * (1.0f / (kMmaHiScale * kMmaLoScale));
          C *o = lo + ((s * kMmaChunk + 4 * t + rr) * 12 + sp * 3);
          if ((lg & 1) == 0) {
            o[0] = make_float2(acc[t][0], acc[t][1]);
            o[1] = make_float2(acc[t][2], acc[t][3]);
          } else {
            o[2] = make_float2(acc[t][0], acc[t][1]);
          }
        }
      }
    }
    __syncthreads();

    // 5. epilogue per (site, right-hand side), as k_hop's
    if (tid < kMmaSites * nc) {
      const int s = tid / nc, r = tid % nc, rhs = r0 + r;
      const long h = site0 + s;
      C v[1][12];
#pragma unroll
      for (int k = 0; k < 12; ++k) v[0][k] = lo[(s * kMmaChunk + r) * 12 + k];
      bool zb = false;
      if (a.open_bc) {
        int c[4];
        half_coords(g, h, a.pout, c);
        zb = c[3] == 0 || c[3] == g.L[3] - 1;
      }
      if (zb) {
#pragma unroll
        for (int k = 0; k < 12; ++k) v[0][k] = cmk<C>(0, 0);
      } else if (EPI & EPI_CLOV) {
        if (clov_staged) {
          clov_apply_inplace_at<float, 1>(lc + s, (long)kMmaSites, v);
        } else if (a.clov) {
          // one set of blocks per right-hand side (a batch of several masses)
          C o[12];
          clov_apply<float>((const C *)a.clov + a.clov_roff[rhs], h, v[0], o);
#pragma unroll
          for (int k = 0; k < 12; ++k) v[0][k] = o[k];
        } else {
#pragma unroll
          for (int k = 0; k < 12; ++k) v[0][k] = cscale(v[0][k], (float)a.clov_scalar);
        }
      }
      if (EPI & EPI_AXPY) {
        C y[12];
        load12((const c2h *)a.y, h, rhs, N, y);
#pragma unroll
        for (int k = 0; k < 12; ++k) v[0][k] = csub(y[k], v[0][k]);
      }
      if (EPI & EPI_DOT) {
        C dv[12];
        load12((const c2h *)a.dotv, h, rhs, N, dv);
        double sum = 0.0;
#pragma unroll
        for (int k = 0; k < 12; ++k)
          sum = fma((double)dv[k].y, (double)v[0][k].y, fma((double)dv[k].x, (double)v[0][k].x, sum));
        a.dots[h * N + rhs] = sum;
      }
      store12_stream((SO *)a.out, g.Vh, h, rhs, N, v[0]);
    }
    __syncthreads();  // the next chunk overwrites lb
  }
}

template <int NL, int DAG, int EPI, typename SO> static bool mma_go(const HopLaunch &a, hipStream_t s) {
  const unsigned grid = (unsigned)(a.g.Vh / kMmaSites);
  k_hop_mma<NL, DAG, EPI, SO><<<grid, kMmaThreads, 0, s>>>(a);
  return true;
}

template <int NL, int DAG> static bool mma_epi(const HopLaunch &a, hipStream_t s) {
  if (a.half_out == kS16Half) {
    switch (a.epi) {
      case 0: return mma_go<NL, DAG, 0, c2h>(a, s);
      case EPI_CLOV: return mma_go<NL, DAG, EPI_CLOV, c2h>(a, s);
      case EPI_AXPY: return mma_go<NL, DAG, EPI_AXPY, c2h>(a, s);
      default: return false;
    }
  }
  if (a.half_out == 0 && a.epi == (EPI_CLOV | EPI_AXPY | EPI_DOT))
    return mma_go<NL, DAG, EPI_CLOV | EPI_AXPY | EPI_DOT, float2>(a, s);
  return false;
}

bool hop_mma_applies(const HopLaunch &a) {
  // in and out within one descriptor's 4 GB (the FP16 field; an FP32 output
  // of twice its size)
  const double e = 12.0 * (double)a.g.Vh * (double)a.nrhs;
  return a.mma && a.half_in == kS16Half && (a.half_out == kS16Half || a.half_out == 0) && a.nrhs >= 4 &&
         a.nrhs % 4 == 0 && !a.skip_t_faces && a.g.Vh % kMmaSites == 0 && (a.nl == 6 || a.nl == 9) &&
         e * (a.half_out ? 4.0 : 8.0) < 4294967296.0;
}

bool hop_mma(const HopLaunch &a, hipStream_t s) {
  if (!hop_mma_applies(a)) return false;
  if (a.nl == 9) return a.dag ? mma_epi<9, 1>(a, s) : mma_epi<9, 0>(a, s);
  return a.dag ? mma_epi<6, 1>(a, s) : mma_epi<6, 0>(a, s);
}

}  // namespace wck
