/*
 * phx_oracle.h -- the oracle's own Philox4x32-10 (test infrastructure).
 *
 * Written independently of the product's generator so that the two can be
 * checked against each other and against the published known-answer vectors
 * of Salmon et al. (SC'11).  The counter layout is the engine's documented one:
 * key = seed, counter = {block, sample<<12 | interaction<<4 | substream,
 * history_lo, history_hi}; two 53-bit doubles per block.
 */
#ifndef XMI_ORACLE_PHX_H
#define XMI_ORACLE_PHX_H
#include <stdint.h>

/* substreams: 0 source, 1 transport of interaction n, 2 detection-weight
 * estimator of interaction n, 3 solid-angle grid, 4 transport of interaction
 * n of a cascade's offspring photon (brute-force runs) */
#define PHX_SUB_OFFSPRING 4u

typedef struct {
  uint32_t k0, k1;
  uint32_t c0, c1, c2, c3;
  double pending;
  int has_pending;
  uint32_t words[4]; /* word mode: the rest of the current block */
  int n_words;
} orc_rng;

static inline void orc_philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int round = 0; round < 10; round++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

static inline void orc_rng_init(orc_rng *r, uint64_t seed, uint32_t word1, uint64_t history) {
  r->k0 = (uint32_t)seed;
  r->k1 = (uint32_t)(seed >> 32);
  r->c0 = 0;
  r->c1 = word1;
  r->c2 = (uint32_t)history;
  r->c3 = (uint32_t)(history >> 32);
  r->has_pending = 0;
  r->n_words = 0;
}

static inline double orc_word_to_double(uint32_t hi, uint32_t lo) {
  return (double)((((uint64_t)hi << 32) | lo) >> 11) * 0x1.0p-53;
}

static inline double orc_uniform(orc_rng *r) {
  if (r->has_pending) {
    r->has_pending = 0;
    return r->pending;
  }
  uint32_t c[4] = {r->c0, r->c1, r->c2, r->c3};
  orc_philox(c, r->k0, r->k1);
  r->c0++;
  r->pending = orc_word_to_double(c[2], c[3]);
  r->has_pending = 1;
  return orc_word_to_double(c[0], c[1]);
}

/* a uniform on (0, 1) at the midpoint of one 32-bit word of orc_word (the
 * solid-angle cone trials: two words per trial, two trials per block) */
static inline uint32_t orc_word(orc_rng *r);
static inline double orc_u32(orc_rng *r) { return ((double)orc_word(r) + 0.5) * (1.0 / 4294967296.0); }

/* 32-bit words from the following blocks, four per block in order (the
 * estimator's Doppler trials draw one word each: 31 bits of random number and
 * a sign bit); a pending double of orc_uniform is dropped */
static inline uint32_t orc_word(orc_rng *r) {
  if (r->n_words == 0) {
    uint32_t c[4] = {r->c0, r->c1, r->c2, r->c3};
    orc_philox(c, r->k0, r->k1);
    r->c0++;
    for (int k = 0; k < 4; k++) r->words[k] = c[3 - k];
    r->n_words = 4;
    r->has_pending = 0;
  }
  return r->words[--r->n_words];
}

#endif
