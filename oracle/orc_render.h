/*
 * orc_render.h -- CPU oracle: samplers, lobes, materials, camera, film and
 * integrators (test infrastructure only; see gr_oracle.h).
 *
 * Random numbers: every draw is addressed, not streamed.  Draw (path, stream,
 * bounce, slot) is word (slot & 3) of Philox4x32-10 with key = seed and counter
 * = (path lo, path hi, bounce, stream << 16 | slot >> 2).  The GPU library uses
 * the same addressing, so for one seed both sides see the same numbers.
 */
#ifndef ORC_RENDER_H
#define ORC_RENDER_H

#include "orc_internal.h"
#include "orc_rng.h"

enum { ORC_STREAM_CAMERA_PATH = 0, ORC_STREAM_LIGHT_PATH = 1, ORC_STREAM_FILM = 2 };
enum {
    ORC_SLOT_LOBE = 0, ORC_SLOT_LOBE_U1 = 1, ORC_SLOT_LOBE_U2 = 2, ORC_SLOT_RR = 3,
    ORC_SLOT_LIGHT = 4, ORC_SLOT_LIGHT_TRI = 5, ORC_SLOT_LIGHT_U1 = 6, ORC_SLOT_LIGHT_U2 = 7,
    ORC_SLOT_LIGHT_DIR1 = 8, ORC_SLOT_LIGHT_DIR2 = 9
};

static inline uint32_t orc_draw_u32(uint64_t seed, uint64_t path, uint32_t stream, uint32_t bounce,
                                    uint32_t slot)
{
    uint32_t ctr[4] = {(uint32_t)path, (uint32_t)(path >> 32), bounce, (stream << 16) | (slot >> 2)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox(ctr, key, out);
    return out[slot & 3];
}
static inline float orc_u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }
static inline float orc_draw_f32(uint64_t seed, uint64_t path, uint32_t stream, uint32_t bounce,
                                 uint32_t slot)
{
    return orc_u01(orc_draw_u32(seed, path, stream, bounce, slot));
}

typedef struct {
    ovec result, outgoing;
    float pdf;
} orc_lobe_sample; /* LobeSample, src/material/lobes.rs:13-17 */

typedef struct {
    ovec emission, radiance, outgoing;
    float pdf;
    int kind;
} orc_mat_sample; /* MaterialSample, src/material/material.rs:12-18 */

int orc_lobe_kind(const orc_lobe *l);
ovec orc_lobe_f(const orc_lobe *l, ovec i, ovec o, int c2l);
float orc_lobe_pdf(const orc_lobe *l, ovec i, ovec o);
float orc_lobe_density(const orc_lobe *l, ovec i, ovec o);
orc_lobe_sample orc_lobe_sample_f(const orc_lobe *l, ovec i, int c2l, float u1, float u2);

ovec orc_f_world(const orc_material *m, ovec wi, ovec wo, const orc_frame *f, int c2l);
float orc_pdf_world(const orc_material *m, ovec wi, ovec wo, const orc_frame *f);
float orc_sample_density_world(const orc_material *m, ovec wi, ovec wo, const orc_frame *f);
ovec orc_light_world(const orc_material *m, ovec wi, const orc_frame *f);
orc_mat_sample orc_sample_world(const orc_material *m, ovec wi, const orc_frame *f, int c2l,
                                uint32_t u_lobe, float u1, float u2);
int orc_count_lobes(const orc_material *m, int mask);

void orc_area_sample_disk(float u1, float u2, float *x, float *y);
ovec orc_cosine_hemisphere(float u1, float u2, int flipped);
ovec orc_uniform_sphere(float u1, float u2);

#endif
