/*
 * refit.h -- the culling programs' boxes from per-element inputs, shared by
 * the host (program build, host-only scenes) and the device update stage
 * (update.hip), which must agree bit for bit: the same double-precision
 * operations in the same order on both sides.
 *
 * No counterpart in QuadRay, whose per-frame tiling (core/engine/engine.cpp:
 * 991-1139, 1985-2160) plays the culling programs' role on the CPU.
 */
#ifndef QR_REFIT_H
#define QR_REFIT_H

#include <math.h>
#include <stdint.h>

#include "qr_soa.h"

#if defined(__HIP__)
#define QR_HD __attribute__((host)) __attribute__((device))
#else
#define QR_HD
#endif

/* one element of the surface list as the programs' boxes see it */
typedef struct qr_refit_elem
{
    float   bmin[3], bmax[3];   /* the surface's extent within its clip box,
                                   in its own frame (recalc_minmax) */
    float   mtx[4][3];          /* its transform node's matrix, rows 0-3 */
    float   lc[3], lh[3];       /* its box in its tracing frame: centre and
                                   half extents (elements with a local box) */
    int32_t world;              /* 1: bounded, has a world box */
    int32_t trn;                /* 1: the world box goes through mtx */
    int32_t local;              /* 1: has a box in its tracing frame */
    int32_t pad;
} qr_refit_elem;

static inline QR_HD double refit_min(double a, double b) { return b < a ? b : a; }
static inline QR_HD double refit_max(double a, double b) { return a < b ? b : a; }
static inline QR_HD double refit_abs(double a) { return a < 0.0 ? -a : a; }

/* the inflated world box of an element: its extent mapped through its
 * transform node (the eight corners), grown by 2e-3 of the coordinates'
 * magnitude; false when it has none */
static inline QR_HD int refit_world_box(const qr_refit_elem *r, float *wmin, float *wmax)
{
    if (!r->world) return 0;
    double lo[3], hi[3];
    if (!r->trn)
    {
        for (int a = 0; a < 3; a++)
        {
            lo[a] = r->bmin[a];
            hi[a] = r->bmax[a];
        }
    }
    else
    {
        for (int a = 0; a < 3; a++)
        {
            lo[a] = +1e300;
            hi[a] = -1e300;
        }
        for (int c = 0; c < 8; c++)
        {
            const double p[3] = {(c & 1) ? r->bmax[0] : r->bmin[0],
                                 (c & 2) ? r->bmax[1] : r->bmin[1],
                                 (c & 4) ? r->bmax[2] : r->bmin[2]};
            for (int x = 0; x < 3; x++)
            {
                double w = r->mtx[3][x];
                for (int i = 0; i < 3; i++) w += p[i] * r->mtx[i][x];
                lo[x] = refit_min(lo[x], w);
                hi[x] = refit_max(hi[x], w);
            }
        }
    }
    double mag = 1.0;
    for (int a = 0; a < 3; a++)
    {
        mag = refit_max(mag, refit_max(refit_abs(lo[a]), refit_abs(hi[a])));
        mag = refit_max(mag, hi[a] - lo[a]);
    }
    const double m = 2e-3 * mag;
    for (int a = 0; a < 3; a++)
    {
        wmin[a] = (float)(lo[a] - m);
        wmax[a] = (float)(hi[a] + m);
    }
    return 1;
}

/* centre and half extents of a box, the half extents rounded up until the
 * float box covers both ends */
static inline QR_HD void refit_centre(const float *lo, const float *hi, float *c, float *h)
{
    for (int a = 0; a < 3; a++)
    {
        c[a] = 0.5f * (lo[a] + hi[a]);
        h[a] = 0.5f * (hi[a] - lo[a]);
        while (c[a] - h[a] > lo[a] || c[a] + h[a] < hi[a])
        {
            h[a] = nextafterf(h[a], 3.402823466e+38f);
        }
    }
}

/* a box as the kernels read it: bounds grown by QR_CULL_FUZZ times the L1
 * distance of its centre from the reference point ref (world boxes) or
 * from the surface's own origin (local boxes), rounded outwards; unbounded
 * axes open */
static inline QR_HD void refit_bounds(const float *c, const float *h, int world,
                                      const float *ref, float *lo, float *hi)
{
    double l1 = 0.0;
    for (int a = 0; a < 3; a++) l1 += refit_abs((double)c[a] - (world ? (double)ref[a] : 0.0));
    const double m = (double)QR_CULL_FUZZ * l1;
    for (int a = 0; a < 3; a++)
    {
        const double cc = c[a], hh = h[a];
        const double dl = cc - hh - m, dh = cc + hh + m;
        if (!__builtin_isfinite(cc) || !__builtin_isfinite(hh) || !__builtin_isfinite(m))
        {
            lo[a] = -__builtin_inff();
            hi[a] = __builtin_inff();
            continue;
        }
        float fl = (float)dl, fh = (float)dh;
        if ((double)fl > dl) fl = nextafterf(fl, -__builtin_inff());
        if ((double)fh < dh) fh = nextafterf(fh, __builtin_inff());
        lo[a] = fl;
        hi[a] = fh;
    }
}

/* a local box moved into its group's frame: the box of centre c + pos and
 * half extents h, about that frame's origin (the transform node's), grown
 * by QR_CULL_FUZZ times the L1 distance of its centre from that origin and
 * rounded outwards; unbounded axes open */
static inline QR_HD void refit_bounds_about(const float *c, const float *h, const float *pos,
                                            float *lo, float *hi)
{
    double cc[3], l1 = 0.0;
    for (int a = 0; a < 3; a++)
    {
        cc[a] = (double)c[a] + (double)pos[a];
        l1 += refit_abs(cc[a]);
    }
    const double m = (double)QR_CULL_FUZZ * l1;
    for (int a = 0; a < 3; a++)
    {
        const double hh = h[a];
        if (!__builtin_isfinite(cc[a]) || !__builtin_isfinite(hh) || !__builtin_isfinite(m))
        {
            lo[a] = -__builtin_inff();
            hi[a] = __builtin_inff();
            continue;
        }
        const double dl = cc[a] - hh - m, dh = cc[a] + hh + m;
        float fl = (float)dl, fh = (float)dh;
        if ((double)fl > dl) fl = nextafterf(fl, -__builtin_inff());
        if ((double)fh < dh) fh = nextafterf(fh, __builtin_inff());
        lo[a] = fl;
        hi[a] = fh;
    }
}

#endif /* QR_REFIT_H */
