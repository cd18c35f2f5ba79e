/*
 * orc_rng.c -- CPU oracle: Philox4x32-10 (Salmon et al., "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11), the counter-based generator that
 * replaces the reference's entropy-seeded XorShift (src/core/random.rs:13-20).
 * Test infrastructure only; pinned by the published Random123 known-answer
 * vectors in tests/test_oracle_rng.py.
 */
#include "gr_oracle.h"
#include "orc_rng.h"

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    orc_philox(ctr, key, out);
}
