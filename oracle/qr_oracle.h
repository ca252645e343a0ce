/*
 * qr_oracle.h -- CPU restatement of QuadRay's ray tracer, used as the
 * checker for the MI355X renderer.  TEST INFRASTRUCTURE ONLY: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never does.
 *
 * The oracle renders from the same qr_scene_tables / qr_frame_desc the GPU
 * kernels read, one ray at a time in scalar fp32, following the operation
 * order of QuadRay's render0 (core/tracer/tracer.cpp:2086-7237).
 */
#ifndef QR_ORACLE_H
#define QR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "qr_soa.h"

#ifdef __cplusplus
extern "C" {
#endif

/* per-sample outputs of a render (all optional) */
typedef struct qr_oracle_aux
{
    int32_t  *ids;      /* first hit: (user id << 1 | side), -1 miss;
                           layout [y][x][sample] */
    uint8_t  *flags;    /* 1 where a quadric test on the primary ray had a
                           normalised discriminant below 1e-5 */
    uint64_t  ops;      /* fp32 arithmetic ops executed (add/mul/div/sqrt 1) */
    uint64_t  rays;     /* rays traced: camera + shadow + secondary */
    /* audit of the renderer's culling boxes: with boxes set (6 floats per
     * surface record, min then max, NaN for none), every accepted hit point
     * is checked against its surface's box */
    const float *boxes;     /* 7 floats per surface: centre, half extents,
                               frame (0 world, 1 the surface's tracing frame) */
    uint64_t  box_checked;
    uint64_t  box_outside;
    float     box_excess;   /* largest distance outside a box over the L1
                               distance from the ray origin to its centre */
    float     pad;
    /* op accounting under per-ray culling with the tables' culling program:
     * ops of list elements a ray's own box tests would skip, and ops of
     * the box tests it evaluates (results are unaffected) */
    uint64_t  ops_skip;
    uint64_t  ops_box;
    /* branch coverage: roots taken through the near-zero discriminant
     * path (D < d_eps), two-plane solves, apex snaps */
    uint64_t  n_dmask;
    uint64_t  n_twoplane;
    uint64_t  n_apex;
} qr_oracle_aux;

/* Render rows [y0, y1) of a Whitted frame into "frame" (stride x_row pixels,
 * row y at frame + y * x_row).  Returns 0 on success. */
int qr_oracle_render(const qr_scene_tables *tab, const qr_frame_desc *fd,
                     uint32_t *frame, int y0, int y1, qr_oracle_aux *aux);

/* Path tracer: one pass over rows [y0, y1): one sample per sample slot
 * folded into the running-mean colour planes (fd->pts_c counts this pass),
 * the slots' seeds stepped, and the frame resolved from the planes.  Planes
 * and seeds hold one value per slot, slot = (y * x_res + x) * ns + k.
 * Returns 0 on success. */
int qr_oracle_render_pt(const qr_scene_tables *tab, const qr_frame_desc *fd,
                        uint32_t *frame, float *plane_r, float *plane_g,
                        float *plane_b, uint32_t *seeds, int y0, int y1,
                        qr_oracle_aux *aux);

/* initial seed plane (n slots) */
void qr_oracle_pt_seeds(uint32_t *seeds, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* QR_ORACLE_H */
