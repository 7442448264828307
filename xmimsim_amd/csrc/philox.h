/*
 * philox.h -- Philox4x32-10 counter-based generator (Salmon, Moraes, Dror,
 * Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), host and device.
 *
 * Stream layout of the photon-history engine: every photon history owns an
 * independent set of streams, addressed purely by counters, so any kernel can
 * regenerate any history's numbers without stored state:
 *
 *   key     = { seed_lo, seed_hi }
 *   counter = { block, (sample << 12) | (interaction << 4) | substream,
 *               history_lo, history_hi }
 *
 * substream 0: source sampling (interaction 0)
 * substream 1: transport of interaction n (step, atom, type, kinematics)
 * substream 2: detection-weight estimator of interaction n
 * substream 3: solid-angle grid (history = cell index)
 * substream 4: transport of interaction n of the offspring photon that a
 *              cascade of a brute-force run starts (history = its parent's)
 *
 * Each block yields four 32-bit words = two doubles in [0,1) with 53 random
 * bits each.
 */
#ifndef XMIMSIM_AMD_PHILOX_H
#define XMIMSIM_AMD_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__)
#define PHX_FN __host__ __device__ static inline
#else
#define PHX_FN static inline
#endif

#define PHX_M0 0xD2511F53u
#define PHX_M1 0xCD9E8D57u
#define PHX_W0 0x9E3779B9u
#define PHX_W1 0xBB67AE85u

#define PHX_SUB_SOURCE 0u
#define PHX_SUB_TRANSPORT 1u
#define PHX_SUB_VARRED 2u
#define PHX_SUB_GRID 3u
#define PHX_SUB_OFFSPRING 4u

PHX_FN void phx_mulhilo(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

PHX_FN void phx_round(uint32_t c[4], const uint32_t k[2]) {
  uint32_t hi0, lo0, hi1, lo1;
  phx_mulhilo(PHX_M0, c[0], &hi0, &lo0);
  phx_mulhilo(PHX_M1, c[2], &hi1, &lo1);
  uint32_t r0 = hi1 ^ c[1] ^ k[0];
  uint32_t r1 = lo1;
  uint32_t r2 = hi0 ^ c[3] ^ k[1];
  uint32_t r3 = lo0;
  c[0] = r0;
  c[1] = r1;
  c[2] = r2;
  c[3] = r3;
}

/* in: counter ctr, key key; out: 4 words */
PHX_FN void phx_4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; r++) {
    phx_round(c, k);
    if (r < 9) {
      k[0] += PHX_W0;
      k[1] += PHX_W1;
    }
  }
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
  out[3] = c[3];
}

PHX_FN double phx_u01(uint32_t a, uint32_t b) {
  uint64_t u = (((uint64_t)a << 32) | (uint64_t)b) >> 11;
  return (double)u * (1.0 / 9007199254740992.0);
}

/* a uniform on (0, 1) from one 32-bit word, at the word's midpoint: for the
 * solid-angle cone trials, two per trial, so that one block serves two trials */
PHX_FN double phx_u32(uint32_t a) { return ((double)a + 0.5) * (1.0 / 4294967296.0); }

/* a sequential stream: regenerates blocks on demand */
typedef struct phx_stream {
  uint32_t key[2];
  uint32_t ctr[4];
  double buf[2];
  int avail;
} phx_stream;

PHX_FN void phx_stream_init(phx_stream *s, uint64_t seed, uint32_t sample, uint32_t interaction, uint32_t substream,
                            uint64_t history) {
  s->key[0] = (uint32_t)seed;
  s->key[1] = (uint32_t)(seed >> 32);
  s->ctr[0] = 0u;
  s->ctr[1] = (sample << 12) | ((interaction & 0xFFu) << 4) | (substream & 0xFu);
  s->ctr[2] = (uint32_t)history;
  s->ctr[3] = (uint32_t)(history >> 32);
  s->avail = 0;
}

PHX_FN double phx_uniform(phx_stream *s) {
  if (s->avail == 0) {
    uint32_t o[4];
    phx_4x32_10(s->ctr, s->key, o);
    s->ctr[0]++;
    s->buf[0] = phx_u01(o[0], o[1]);
    s->buf[1] = phx_u01(o[2], o[3]);
    s->avail = 2;
    return s->buf[0];
  }
  s->avail = 0;
  return s->buf[1];
}

/* a sequential stream of 32-bit words, four per block, for draws that need
 * at most 32 random bits (the estimator's Doppler trials: 31 bits for the
 * profile's random number, one for the sign of the momentum); it starts at
 * the counter it is given and does not share blocks with a phx_stream */
typedef struct phx_wstream {
  uint32_t key[2];
  uint32_t ctr[4];
  uint32_t buf[4];
  int avail;
} phx_wstream;

PHX_FN uint32_t phx_word(phx_wstream *s) {
  if (s->avail == 0) {
    phx_4x32_10(s->ctr, s->key, s->buf);
    s->ctr[0]++;
    s->avail = 4;
  }
  /* words in block order: buf[0] first, then shift */
  const uint32_t w = s->buf[0];
  s->buf[0] = s->buf[1];
  s->buf[1] = s->buf[2];
  s->buf[2] = s->buf[3];
  s->avail--;
  return w;
}

#endif
