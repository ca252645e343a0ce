/*
 * qr_oracle_pt.c -- CPU restatement of QuadRay's path tracer (checker only).
 */
#include "qr_oracle.h"

int qr_oracle_render_pt(const qr_scene_tables *tab, const qr_frame_desc *fd,
                        uint32_t *frame, float *plane_r, float *plane_g,
                        float *plane_b, uint32_t *seeds, int y0, int y1,
                        qr_oracle_aux *aux)
{
    (void)tab; (void)fd; (void)frame; (void)plane_r; (void)plane_g;
    (void)plane_b; (void)seeds; (void)y0; (void)y1; (void)aux;
    return -1;
}
