/*
 * orc_rng.h -- CPU oracle: Philox4x32-10 block function and the stream
 * layout used for every random draw (test infrastructure only).
 */
#ifndef ORC_RNG_H
#define ORC_RNG_H
#include <stdint.h>

#define ORC_PHILOX_M0 0xD2511F53u
#define ORC_PHILOX_M1 0xCD9E8D57u
#define ORC_PHILOX_W0 0x9E3779B9u
#define ORC_PHILOX_W1 0xBB67AE85u

static inline void orc_philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)ORC_PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)ORC_PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += ORC_PHILOX_W0;
        k1 += ORC_PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* A per-path stream: key = seed, counter = (path id lo, path id hi, draw
 * index, 0).  Draws are consumed one u32 at a time, four per block. */
typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int avail;
} orc_rng;

static inline void orc_rng_init(orc_rng *r, uint64_t seed, uint64_t path_id, uint32_t stream)
{
    r->key[0] = (uint32_t)seed; r->key[1] = (uint32_t)(seed >> 32);
    r->ctr[0] = (uint32_t)path_id; r->ctr[1] = (uint32_t)(path_id >> 32);
    r->ctr[2] = 0; r->ctr[3] = stream;
    r->avail = 0;
}
static inline uint32_t orc_rng_u32(orc_rng *r)
{
    if (r->avail == 0) {
        orc_philox(r->ctr, r->key, r->buf);
        r->ctr[2] += 1;
        r->avail = 4;
    }
    return r->buf[4 - r->avail--];
}
/* Uniform f32 in [0, 1): the top 24 bits scaled by 2^-24 (exact). */
static inline float orc_rng_f32(orc_rng *r)
{
    return (float)(orc_rng_u32(r) >> 8) * (1.0f / 16777216.0f);
}
/* Uniform integer in [0, n) by multiply-shift (Range<usize>, replaced). */
static inline uint32_t orc_rng_below(orc_rng *r, uint32_t n)
{
    return (uint32_t)(((uint64_t)orc_rng_u32(r) * n) >> 32);
}
#endif
