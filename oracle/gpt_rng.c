/*
 * ORACLE / TEST INFRASTRUCTURE ONLY.  Nothing in the product path links this
 * file; tests, the fixture generator and bench.py's CPU baseline load it
 * through ctypes to reproduce the random fields gpt's own tests draw.
 *
 * A C restatement of gpt's lattice random-number generator, written from the
 * reference's description of the algorithm so that fields drawn from a given
 * seed string match gpt's draws bit for bit:
 *
 *   - RANLUX24 subtract-with-borrow (base 2^24, lags r=24, s=10) with a
 *     luxury block of p=389 steps per 24 outputs
 *       lib/cgpt/lib/random/ranlux.h:28-110
 *   - seeding: SHA-256 over the little-endian uint64 seed words plus a
 *     counter, 24 words per scalar generator; 64 lanes seeded one after the
 *     other from the scalar generator, outputs interleaved lane by lane
 *       lib/cgpt/lib/random/vector.h:38-114
 *   - a 64-bit bit reservoir refilled 24 bits at a time, doubles from 53
 *     bits, Box-Muller normals with a stack holding the second value,
 *     complex normals drawing the imaginary part first
 *       lib/cgpt/lib/random/distribution.h:20-160
 *   - one generator per 2^nd block of sites, seeded with the seed string's
 *     characters, the grid's full and checkerboarded dimensions and the
 *     block's hash (block coordinates with dimension 0 slowest); inside a
 *     block the values go to sites in lexicographic order (x fastest)
 *       lib/cgpt/lib/random/engine.h:48-97, parallel.h:21-130
 *
 * SHA-256 is the FIPS 180-4 hash; it is implemented here from the standard.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- SHA-256 */

static const uint32_t K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4,
    0xab1c5ed5, 0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe,
    0x9bdc06a7, 0xc19bf174, 0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f,
    0x4a7484aa, 0x5cb0a9dc, 0x76f988da, 0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7,
    0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967, 0x27b70a85, 0x2e1b2138, 0x4d2c6dfc,
    0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85, 0xa2bfe8a1, 0xa81a664b,
    0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070, 0x19a4c116,
    0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7,
    0xc67178f2};

static inline uint32_t ror32(uint32_t v, int n) { return (v >> n) | (v << (32 - n)); }

static void sha256_compress(uint32_t st[8], const uint8_t blk[64]) {
  uint32_t w[64];
  for (int i = 0; i < 16; ++i)
    w[i] = ((uint32_t)blk[4 * i] << 24) | ((uint32_t)blk[4 * i + 1] << 16) |
           ((uint32_t)blk[4 * i + 2] << 8) | (uint32_t)blk[4 * i + 3];
  for (int i = 16; i < 64; ++i) {
    uint32_t s0 = ror32(w[i - 15], 7) ^ ror32(w[i - 15], 18) ^ (w[i - 15] >> 3);
    uint32_t s1 = ror32(w[i - 2], 17) ^ ror32(w[i - 2], 19) ^ (w[i - 2] >> 10);
    w[i] = w[i - 16] + s0 + w[i - 7] + s1;
  }
  uint32_t a = st[0], b = st[1], c = st[2], d = st[3], e = st[4], f = st[5], g = st[6],
           h = st[7];
  for (int i = 0; i < 64; ++i) {
    uint32_t S1 = ror32(e, 6) ^ ror32(e, 11) ^ ror32(e, 25);
    uint32_t ch = (e & f) ^ (~e & g);
    uint32_t t1 = h + S1 + ch + K256[i] + w[i];
    uint32_t S0 = ror32(a, 2) ^ ror32(a, 13) ^ ror32(a, 22);
    uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
    uint32_t t2 = S0 + mj;
    h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
  }
  st[0] += a; st[1] += b; st[2] += c; st[3] += d;
  st[4] += e; st[5] += f; st[6] += g; st[7] += h;
}

/* digest as eight 32-bit words H0..H7 */
void gr_sha256_words(const uint8_t *msg, size_t len, uint32_t out[8]) {
  uint32_t st[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                    0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
  size_t full = len / 64;
  for (size_t i = 0; i < full; ++i) sha256_compress(st, msg + 64 * i);
  uint8_t tail[128];
  size_t rem = len - 64 * full;
  memset(tail, 0, sizeof tail);
  memcpy(tail, msg + 64 * full, rem);
  tail[rem] = 0x80;
  size_t nblk = (rem + 1 + 8 <= 64) ? 1 : 2;
  uint64_t bits = (uint64_t)len * 8u;
  for (int i = 0; i < 8; ++i) tail[64 * nblk - 1 - i] = (uint8_t)(bits >> (8 * i));
  for (size_t i = 0; i < nblk; ++i) sha256_compress(st, tail + 64 * i);
  memcpy(out, st, sizeof st);
}

/* ------------------------------------------------ RANLUX24 (r=24, s=10) */

enum { SWB_R = 24, SWB_S = 10, SWB_MASK = (1u << 24) - 1u };

typedef struct {
  uint32_t x[SWB_R]; /* ring of the last 24 values                       */
  uint32_t borrow;   /* c_{n-1}                                           */
  int off;           /* ring offset: lag k lives in slot (k-1-off) mod 24  */
  int used;          /* outputs taken in the current luxury block          */
  int lux;           /* p: steps per luxury block                          */
} swb24;

static inline uint32_t *swb_lag(swb24 *g, int k) {
  int i = (k - 1 - g->off) % SWB_R;
  if (i < 0) i += SWB_R;
  return &g->x[i];
}

static void swb_seed(swb24 *g, const uint32_t seed[SWB_R], int lux) {
  for (int i = 0; i < SWB_R; ++i) g->x[i] = seed[i] & SWB_MASK;
  g->borrow = (seed[0] == 0u);
  g->off = 0;
  g->used = SWB_R - 1; /* the first output starts a fresh luxury block */
  g->lux = lux;
}

/* one subtract-with-borrow step: x_n = x_{n-s} - x_{n-r} - c_{n-1} mod 2^24 */
static inline uint32_t swb_step(swb24 *g) {
  g->off = (g->off + 1) % SWB_R;
  uint32_t d = *swb_lag(g, SWB_S) - *swb_lag(g, SWB_R) - g->borrow; /* mod 2^32 */
  g->borrow = (d > SWB_MASK); /* wrapped below zero */
  d = (d + (SWB_MASK + 1u)) & SWB_MASK;
  *swb_lag(g, SWB_R) = d; /* lag r's slot becomes lag 0 (the newest value) */
  return d;
}

static inline uint32_t swb_next(swb24 *g) {
  if (++g->used == SWB_R) {
    g->used = 0;
    for (int i = 0; i < g->lux - SWB_R; ++i) swb_step(g);
  }
  return swb_step(g);
}

/* ------------------------------------- 64-lane generator + bit reservoir */

enum { NLANE = 64 };

typedef struct {
  swb24 lane[NLANE];
  uint32_t buf[NLANE];
  int nbuf;
  uint64_t reservoir;
  int nbits;
  double stack[4]; /* Box-Muller second values */
  int nstack;
} gr_stream;

static void lanes_refill(gr_stream *s) {
  for (int l = 0; l < NLANE; ++l) s->buf[l] = swb_next(&s->lane[l]);
  s->nbuf = 0;
}

static uint32_t lanes_next(gr_stream *s) {
  if (s->nbuf == NLANE) lanes_refill(s);
  return s->buf[s->nbuf++];
}

static void reservoir_add(gr_stream *s) {
  s->reservoir = s->reservoir * (uint64_t)(1u << 24) + (uint64_t)lanes_next(s);
  s->nbits += 24;
  if (s->nbits > 64) s->nbits = 64;
}

static void stream_init(gr_stream *s, const uint64_t *seed, size_t nseed, int lux) {
  /* scalar generator seeded from SHA-256(seed words || counter) */
  uint32_t words[SWB_R];
  int nw = 0;
  uint64_t *tmp = (uint64_t *)malloc((nseed + 1) * sizeof(uint64_t));
  uint64_t counter = 0;
  while (nw < SWB_R) {
    memcpy(tmp, seed, nseed * sizeof(uint64_t));
    tmp[nseed] = counter++;
    uint8_t bytes[8 * 64];
    size_t nbytes = (nseed + 1) * 8;
    uint8_t *mb = nbytes <= sizeof bytes ? bytes : (uint8_t *)malloc(nbytes);
    for (size_t i = 0; i < nseed + 1; ++i)
      for (int b = 0; b < 8; ++b) mb[8 * i + b] = (uint8_t)(tmp[i] >> (8 * b));
    uint32_t h[8];
    gr_sha256_words(mb, nbytes, h);
    if (mb != bytes) free(mb);
    for (int w = 0; w < 8 && nw < SWB_R; ++w) words[nw++] = h[w];
  }
  free(tmp);
  swb24 scalar;
  swb_seed(&scalar, words, lux);
  /* every lane takes 24 consecutive scalar outputs, lane 0 first */
  for (int l = 0; l < NLANE; ++l) {
    uint32_t ls[SWB_R];
    for (int j = 0; j < SWB_R; ++j) ls[j] = swb_next(&scalar);
    swb_seed(&s->lane[l], ls, lux);
  }
  lanes_refill(s);
  s->reservoir = 0;
  s->nbits = 0;
  s->nstack = 0;
  reservoir_add(s);
}

static uint64_t take_bits(gr_stream *s, int bits) {
  while (bits > s->nbits) reservoir_add(s);
  uint64_t base = (uint64_t)1 << bits;
  uint64_t r = s->reservoir & (base - 1);
  s->reservoir /= base;
  s->nbits -= bits;
  return r;
}

static double take_u01(gr_stream *s) {
  return (double)take_bits(s, 53) / (double)((uint64_t)1 << 53);
}

static double take_normal(gr_stream *s) {
  if (s->nstack > 0) return s->stack[--s->nstack];
  const double tiny = 2.2250738585072014e-308; /* smallest normal double */
  const double two_pi = 2.0 * 3.14159265358979323846;
  double u1, u2;
  do {
    u1 = take_u01(s);
    u2 = take_u01(s);
  } while (u1 <= tiny);
  double rad = sqrt(-2.0 * log(u1));
  double z0 = rad * cos(two_pi * u2);
  double z1 = sqrt(-2.0 * log(u1)) * sin(two_pi * u2);
  s->stack[s->nstack++] = z1;
  return z0;
}

/* ----------------------------------------------- lattice-level engine */

typedef struct {
  int nd;
  long ldim[8];
  long nblocks, block_sites;
  gr_stream *streams; /* one per 2^nd block, block index x fastest */
} gr_engine;

/* dist: 0 = uniform [lo,hi) real, 1 = normal(mu=lo, sigma=hi) real,
 *       2 = complex normal (imaginary part drawn first)                    */
static void draw(gr_stream *s, int dist, double lo, double hi, double out[2]) {
  if (dist == 0) {
    out[0] = take_u01(s) * (hi - lo) + lo;
    out[1] = 0.0;
  } else if (dist == 1) {
    out[0] = take_normal(s) * hi + lo;
    out[1] = 0.0;
  } else {
    double im = take_normal(s) * hi + lo;
    double re = take_normal(s) * hi + lo;
    out[0] = re;
    out[1] = im;
  }
}

/* Create the per-block streams for one grid.  seed: the seed string's bytes;
 * fdims/gdims: full and (checkerboarded) grid dimensions, nd entries each.  */
gr_engine *gr_create(const char *seed, const long *fdims, const long *gdims, int nd, int lux) {
  gr_engine *e = (gr_engine *)calloc(1, sizeof(gr_engine));
  size_t ls = strlen(seed);
  e->nd = nd;
  long nb = 1;
  long bdim[8];
  for (int j = 0; j < nd; ++j) {
    e->ldim[j] = gdims[j];
    if (gdims[j] % 2 != 0) { free(e); return NULL; }
    bdim[j] = gdims[j] / 2;
    nb *= bdim[j];
  }
  e->nblocks = nb;
  e->block_sites = 1L << nd;
  e->streams = (gr_stream *)malloc(sizeof(gr_stream) * nb);
  size_t nseed = ls + 2 * nd + 1;
  uint64_t *sw = (uint64_t *)malloc(sizeof(uint64_t) * nseed);
  for (size_t i = 0; i < ls; ++i) sw[i] = (uint64_t)(unsigned char)seed[i];
  for (int j = 0; j < nd; ++j) sw[ls + j] = (uint64_t)fdims[j];
  for (int j = 0; j < nd; ++j) sw[ls + nd + j] = (uint64_t)gdims[j];
  for (long b = 0; b < nb; ++b) {
    long rest = b, bc[8];
    for (int j = 0; j < nd; ++j) { bc[j] = rest % bdim[j]; rest /= bdim[j]; }
    long hash = 0;
    for (int j = 0; j < nd; ++j) hash = hash * bdim[j] + bc[j];
    sw[nseed - 1] = (uint64_t)hash;
    stream_init(&e->streams[b], sw, nseed, lux);
  }
  free(sw);
  return e;
}

void gr_destroy(gr_engine *e) {
  if (!e) return;
  free(e->streams);
  free(e);
}

/* Fill n_per_site complex numbers per site; out is [sites][n_per_site][2]
 * doubles with sites in lexicographic order of the grid's dims (x fastest). */
void gr_fill(gr_engine *e, int dist, double lo, double hi, int n_per_site, double *out) {
  int nd = e->nd;
  long bdim[8];
  for (int j = 0; j < nd; ++j) bdim[j] = e->ldim[j] / 2;
  for (long b = 0; b < e->nblocks; ++b) {
    long rest = b, bc[8];
    for (int j = 0; j < nd; ++j) { bc[j] = rest % bdim[j]; rest /= bdim[j]; }
    gr_stream *s = &e->streams[b];
    for (long r = 0; r < e->block_sites; ++r) {
      long site = 0, stride = 1;
      for (int j = 0; j < nd; ++j) {
        long c = 2 * bc[j] + ((r >> j) & 1);
        site += c * stride;
        stride *= e->ldim[j];
      }
      for (int k = 0; k < n_per_site; ++k) draw(s, dist, lo, hi, out + 2 * (site * n_per_site + k));
    }
  }
}

/* raw lane outputs for known-answer checks of the generator itself */
void gr_raw(gr_engine *e, long block, long n, uint32_t *out) {
  for (long i = 0; i < n; ++i) out[i] = lanes_next(&e->streams[block]);
}
