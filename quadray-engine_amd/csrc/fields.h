/*
 * fields.h -- the per-record part of the update stage: from one object's
 * matrix, axis map and its static shape and clip data, its record of the
 * surface table, its clip box and its culling inputs.  The host engine
 * (engine.cpp) and the device update kernel (update.hip, one work-item per
 * record) run these same functions, so the device writes the bytes the host
 * would: plain fp32 operations in the same order on both sides (the build
 * keeps fused multiply-adds off and divisions and square roots correctly
 * rounded), doubles where the culling inputs take them.
 *
 * What stays on the host is what needs the host's libm or the object tree:
 * the hierarchical transforms (sines and cosines of the rotation angles),
 * the axis maps and transform nodes they decide, the clip lists and the
 * surface list.  Everything here depends on one record's inputs (and, for
 * the clip box of a surface with custom clippers, on its clippers' inputs),
 * never on the order in which records are computed.
 *
 * Field meanings follow QuadRay's surface update: matrices and axis maps
 * (core/engine/object.cpp:221-389), the per-type coefficients (2920-2977,
 * 3037-3088, 3145-3936), clip boxes (2533-2870) and the backend records
 * (core/tracer/tracer.h:869-1292, core/engine/engine.cpp:2839-3290).
 */
#ifndef QR_FIELDS_H
#define QR_FIELDS_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "qr_format.h"
#include "qr_soa.h"
#include "refit.h"

#define QRF static inline QR_HD

#define QRF_INF 3.402823466e+38f

/* one record's inputs: a surface, or an array that is a transform node */
typedef struct qr_geom
{
    float    mtx[4][4];     /* the object's matrix (relative to its transform
                               node when it has one and is not its own) */
    float    tmtx[4][4];    /* its transform node's matrix (culling inputs) */
    float    scl[4];        /* axis scalers, axis map and signs of a trivial */
    int32_t  map[4];        /* transform (identity under a transform node) */
    int32_t  sgn[4];
    float    smin[4];       /* the surface's axis clip values (rt_SURFACE) */
    float    smax[4];
    float    prm[2];        /* shape: rad (cylinder, sphere), rat (cone,
                               hyperboloid, hypercylinder) then hyp, par
                               (paraboloid, paracylinder), pr1 and pr2
                               (hyperparaboloid) */
    float    tex[2][4];     /* plane texture per side: the material's two
                               scalers, then its offsets on the mapped axes */
    int32_t  tex_map[2][2]; /* per side: texture x / y take local axis 0 / 1 */
    int32_t  tag;           /* RT_TAG_*, RT_TAG_ARRAY for a transform node */
    int32_t  tn_rec;        /* record of its transform node, -1 none */
    int32_t  self_tn;       /* 1: it is its own transform node */
    int32_t  obj_has_trm;   /* transform flags (scale 1, rotation 2) */
    int32_t  id;            /* user-visible surface id, -1 for an array */
    int32_t  elem;          /* its element in the surface list, -1 none */
    int32_t  mat[2];        /* material records, outer / inner */
    uint32_t props[2];
    int32_t  clip_beg;      /* its custom clip list in the clip table */
    int32_t  clip_cnt;
    int32_t  adjust;        /* 1: clip box tightened by its outer clippers
                               (RT_OPTS_ADJUST, clippers, a shared node) */
    int32_t  pad;
} qr_geom;

QRF float qf_max(float a, float b) { return a > b ? a : b; }
QRF float qf_min(float a, float b) { return a < b ? a : b; }
/* square root that is 0 for non-positive input (QuadRay RT_SQRT32) */
QRF float qf_sqrt(float a) { return a <= 0.0f ? 0.0f : sqrtf(a); }

/* inverse of the upper 3x3 of m (cofactors over the determinant) */
QRF void qf_inverse(const float (*m)[4], float (*inv)[4])
{
    for (int i = 0; i < 4; i++)
    {
        for (int j = 0; j < 4; j++) inv[i][j] = 0.0f;
    }
    float a = m[1][1] * m[2][2] - m[2][1] * m[1][2];
    float b = m[2][1] * m[0][2] - m[0][1] * m[2][2];
    float c = m[0][1] * m[1][2] - m[1][1] * m[0][2];
    float d = m[2][0] * m[1][2] - m[1][0] * m[2][2];
    float e = m[0][0] * m[2][2] - m[2][0] * m[0][2];
    float f = m[0][2] * m[1][0] - m[0][0] * m[1][2];
    float g = m[1][0] * m[2][1] - m[2][0] * m[1][1];
    float h = m[2][0] * m[0][1] - m[0][0] * m[2][1];
    float l = m[0][0] * m[1][1] - m[1][0] * m[0][1];
    float q = 1.0f / (m[0][0] * a + m[1][0] * b + m[2][0] * c);
    inv[0][0] = a * q; inv[0][1] = b * q; inv[0][2] = c * q;
    inv[1][0] = d * q; inv[1][1] = e * q; inv[1][2] = f * q;
    inv[2][0] = g * q; inv[2][1] = h * q; inv[2][2] = l * q;
}

/* quadric coefficients with the axis scalers applied: sci x, y, z, w and
 * scj x, y, z (not yet halved) */
QRF void qf_shape(const qr_geom &g, float *sci, float *scj)
{
    const int mi = g.map[RT_I], mj = g.map[RT_J], mk = g.map[RT_K];
    for (int a = 0; a < 4; a++) scj[a] = 0.0f;
    if (g.tag == RT_TAG_PLANE)
    {
        for (int a = 0; a < 4; a++) sci[a] = 0.0f;
        return;
    }
    sci[0] = sci[1] = sci[2] = 1.0f;
    sci[3] = 0.0f;
    const float p0 = g.prm[0], p1 = g.prm[1];
    switch (g.tag)
    {
        case RT_TAG_CYLINDER:
        sci[mk] = 0.0f;
        sci[RT_W] = p0 * p0;
        break;
        case RT_TAG_SPHERE:
        sci[RT_W] = p0 * p0;
        break;
        case RT_TAG_CONE:
        sci[mk] = -(p0 * p0);
        break;
        case RT_TAG_PARABOLOID:
        sci[mk] = 0.0f;
        scj[mk] = p0 * (float)g.sgn[RT_K];
        break;
        case RT_TAG_HYPERBOLOID:
        sci[mk] = -(p0 * p0);
        sci[RT_W] = p1;
        break;
        case RT_TAG_PARACYLINDER:
        sci[mj] = 0.0f;
        sci[mk] = 0.0f;
        scj[mk] = p0 * (float)g.sgn[RT_K];
        break;
        case RT_TAG_HYPERCYLINDER:
        sci[mj] = 0.0f;
        sci[mk] = -(p0 * p0);
        sci[RT_W] = p1;
        break;
        case RT_TAG_HYPERPARABOLOID:
        sci[mi] = 1.0f / +fabsf(p0);
        sci[mj] = 1.0f / -fabsf(p1);
        sci[mk] = 0.0f;
        scj[mk] = 1.0f * (float)g.sgn[RT_K];
        break;
        default:
        break;
    }
    /* the axis scalers (QuadRay rt_Quadric::commit_fields) */
    const float ix = 1.0f / g.scl[RT_X], iy = 1.0f / g.scl[RT_Y], iz = 1.0f / g.scl[RT_Z];
    sci[RT_X] *= ix * ix;
    sci[RT_Y] *= iy * iy;
    sci[RT_Z] *= iz * iz;
    scj[RT_X] *= ix;
    scj[RT_Y] *= iy;
    scj[RT_Z] *= iz;
}

/* a plane's texture transform of one side: xscal, yscal, xoffs, yoffs */
QRF void qf_plane_tex(const qr_geom &g, int side, float *out)
{
    const float asc[2] = {g.scl[g.map[RT_I]], g.scl[g.map[RT_J]]};
    const float isc[2] = {1.0f / asc[0], 1.0f / asc[1]};
    const int m0 = g.tex_map[side][0], m1 = g.tex_map[side][1];
    out[0] = g.tex[side][0] * isc[m0];
    out[1] = g.tex[side][1] * isc[m1];
    out[2] = g.tex[side][2] * asc[m0];
    out[3] = g.tex[side][3] * asc[m1];
}

/* the shape's extent within a box of its local frame (smin, smax), and the
 * axis clips that extent makes redundant (QuadRay's per-type
 * adjust_minmax); bmin/bmax and cmin/cmax may be null */
QRF void qf_adjust(const qr_geom &g, const float *smin, const float *smax, float *bmin,
                   float *bmax, float *cmin, float *cmax)
{
    const bool cb = cmin != nullptr && cmax != nullptr;
    const bool bb = bmin != nullptr && bmax != nullptr;
    if (cb)
    {
        for (int a = 0; a < 3; a++)
        {
            cmin[a] = smin[a] > g.smin[a] ? -QRF_INF : smin[a];
            cmax[a] = smax[a] < g.smax[a] ? +QRF_INF : smax[a];
        }
    }
    const int I = RT_I, J = RT_J, K = RT_K;
    switch (g.tag)
    {
        case RT_TAG_PLANE:
        if (cb)
        {
            cmin[K] = -QRF_INF;
            cmax[K] = +QRF_INF;
        }
        if (bb)
        {
            bmin[I] = smin[I]; bmin[J] = smin[J]; bmin[K] = 0.0f;
            bmax[I] = smax[I]; bmax[J] = smax[J]; bmax[K] = 0.0f;
        }
        break;

        case RT_TAG_CYLINDER:
        {
            const float rad = fabsf(g.prm[0]);
            if (cb)
            {
                cmin[I] = cmin[I] <= -rad ? -QRF_INF : cmin[I];
                cmin[J] = cmin[J] <= -rad ? -QRF_INF : cmin[J];
                cmax[I] = cmax[I] >= +rad ? +QRF_INF : cmax[I];
                cmax[J] = cmax[J] >= +rad ? +QRF_INF : cmax[J];
            }
            if (bb)
            {
                bmin[I] = qf_max(smin[I], -rad);
                bmin[J] = qf_max(smin[J], -rad);
                bmin[K] = smin[K];
                bmax[I] = qf_min(smax[I], +rad);
                bmax[J] = qf_min(smax[J], +rad);
                bmax[K] = smax[K];
            }
        }
        break;

        case RT_TAG_SPHERE:
        {
            const float r0 = g.prm[0];
            float r = fabsf(r0);
            float rad[3] = {r, r, r};
            for (int k = 0; k < 3; k++)
            {
                const float top = smin[k] > 0.0f ? +smin[k] : smax[k] < 0.0f ? -smax[k] : 0.0f;
                r = qf_sqrt(qf_max(r0 * r0 - top * top, 0.0f));
                const int i = (k + 1) % 3, j = (k + 2) % 3;
                if (rad[i] > r) rad[i] = r;
                if (rad[j] > r) rad[j] = r;
            }
            for (int a = 0; a < 3; a++)
            {
                if (cb)
                {
                    cmin[a] = cmin[a] <= -rad[a] ? -QRF_INF : cmin[a];
                    cmax[a] = cmax[a] >= +rad[a] ? +QRF_INF : cmax[a];
                }
                if (bb)
                {
                    bmin[a] = qf_max(smin[a], -rad[a]);
                    bmax[a] = qf_min(smax[a], +rad[a]);
                }
            }
        }
        break;

        case RT_TAG_CONE:
        {
            const float rat = fabsf(g.prm[0]);
            float top = qf_max(fabsf(smin[K]), fabsf(smax[K]));
            const float rad = top != QRF_INF ? top * rat : QRF_INF;
            const float mxi = qf_max(fabsf(smin[I]), fabsf(smax[I]));
            const float mxj = qf_max(fabsf(smin[J]), fabsf(smax[J]));
            top = qf_min(mxi != QRF_INF && mxj != QRF_INF ?
                         qf_sqrt(mxi * mxi + mxj * mxj) / rat : top, top);
            if (cb)
            {
                cmin[I] = cmin[I] <= -rad ? -QRF_INF : cmin[I];
                cmin[J] = cmin[J] <= -rad ? -QRF_INF : cmin[J];
                cmin[K] = cmin[K] <  -top ? -QRF_INF : cmin[K];
                cmax[I] = cmax[I] >= +rad ? +QRF_INF : cmax[I];
                cmax[J] = cmax[J] >= +rad ? +QRF_INF : cmax[J];
                cmax[K] = cmax[K] >  +top ? +QRF_INF : cmax[K];
            }
            if (bb)
            {
                bmin[I] = qf_max(smin[I], -rad);
                bmin[J] = qf_max(smin[J], -rad);
                bmin[K] = cb ? qf_max(smin[K], -top) : smin[K];
                bmax[I] = qf_min(smax[I], +rad);
                bmax[J] = qf_min(smax[J], +rad);
                bmax[K] = cb ? qf_min(smax[K], +top) : smax[K];
            }
        }
        break;

        case RT_TAG_PARABOLOID:
        case RT_TAG_PARACYLINDER:
        {
            const bool pc = g.tag == RT_TAG_PARACYLINDER;
            const float par = g.prm[0];
            float top = qf_max(par < 0.0f ? -smin[K] : +smax[K], 0.0f);
            const float rad = top != QRF_INF ? qf_sqrt(top * fabsf(par)) : QRF_INF;
            const float mxi = qf_max(fabsf(smin[I]), fabsf(smax[I]));
            if (pc)
            {
                top = qf_min(mxi != QRF_INF ? (mxi * mxi) / fabsf(par) : top, top);
            }
            else
            {
                const float mxj = qf_max(fabsf(smin[J]), fabsf(smax[J]));
                top = qf_min(mxi != QRF_INF && mxj != QRF_INF ?
                             (mxi * mxi + mxj * mxj) / fabsf(par) : top, top);
            }
            if (cb)
            {
                cmin[I] = cmin[I] <= -rad ? -QRF_INF : cmin[I];
                cmax[I] = cmax[I] >= +rad ? +QRF_INF : cmax[I];
                if (!pc)
                {
                    cmin[J] = cmin[J] <= -rad ? -QRF_INF : cmin[J];
                    cmax[J] = cmax[J] >= +rad ? +QRF_INF : cmax[J];
                }
                cmin[K] = (cmin[K] <= 0.0f && par > 0.0f) ||
                          (cmin[K] <  -top && par < 0.0f) ? -QRF_INF : cmin[K];
                cmax[K] = (cmax[K] >= 0.0f && par < 0.0f) ||
                          (cmax[K] >  +top && par > 0.0f) ? +QRF_INF : cmax[K];
            }
            if (bb)
            {
                bmin[I] = qf_max(smin[I], -rad);
                bmin[J] = pc ? smin[J] : qf_max(smin[J], -rad);
                bmin[K] = par > 0.0f ? qf_max(smin[K], 0.0f) :
                          cb ? qf_max(smin[K], -top) : smin[K];
                bmax[I] = qf_min(smax[I], +rad);
                bmax[J] = pc ? smax[J] : qf_min(smax[J], +rad);
                bmax[K] = par < 0.0f ? qf_min(smax[K], 0.0f) :
                          cb ? qf_min(smax[K], +top) : smax[K];
            }
        }
        break;

        case RT_TAG_HYPERBOLOID:
        case RT_TAG_HYPERCYLINDER:
        {
            const bool hc = g.tag == RT_TAG_HYPERCYLINDER;
            const float rat = fabsf(g.prm[0]);
            const float hyp = g.prm[1];
            float top = qf_max(fabsf(smin[K]), fabsf(smax[K]));
            const float rad = top != QRF_INF ? qf_sqrt(top * top * rat * rat + hyp) : QRF_INF;
            const float mxi = qf_max(fabsf(smin[I]), fabsf(smax[I]));
            if (hc)
            {
                top = qf_min(mxi != QRF_INF ? qf_sqrt(mxi * mxi - hyp) / rat : top, top);
            }
            else
            {
                const float mxj = qf_max(fabsf(smin[J]), fabsf(smax[J]));
                top = qf_min(mxi != QRF_INF && mxj != QRF_INF ?
                             qf_sqrt(mxi * mxi + mxj * mxj - hyp) / rat : top, top);
            }
            if (cb)
            {
                cmin[I] = cmin[I] <= -rad ? -QRF_INF : cmin[I];
                cmax[I] = cmax[I] >= +rad ? +QRF_INF : cmax[I];
                if (!hc)
                {
                    cmin[J] = cmin[J] <= -rad ? -QRF_INF : cmin[J];
                    cmax[J] = cmax[J] >= +rad ? +QRF_INF : cmax[J];
                }
                cmin[K] = cmin[K] <  -top ? -QRF_INF : cmin[K];
                cmax[K] = cmax[K] >  +top ? +QRF_INF : cmax[K];
            }
            if (bb)
            {
                bmin[I] = qf_max(smin[I], -rad);
                bmin[J] = hc ? smin[J] : qf_max(smin[J], -rad);
                bmin[K] = cb ? qf_max(smin[K], -top) : smin[K];
                bmax[I] = qf_min(smax[I], +rad);
                bmax[J] = hc ? smax[J] : qf_min(smax[J], +rad);
                bmax[K] = cb ? qf_min(smax[K], +top) : smax[K];
            }
        }
        break;

        case RT_TAG_HYPERPARABOLOID:
        {
            const float pr1 = g.prm[0], pr2 = g.prm[1];
            const float rd1 = qf_max(-smin[I], +smax[I]);
            const float rd2 = qf_max(-smin[J], +smax[J]);
            const float tp1 = rd1 * rd1 / fabsf(pr1);
            const float tp2 = rd2 * rd2 / fabsf(pr2);
            if (cb)
            {
                cmin[K] = cmin[K] <= -tp2 ? -QRF_INF : cmin[K];
                cmax[K] = cmax[K] >= +tp1 ? +QRF_INF : cmax[K];
            }
            if (bb)
            {
                bmin[I] = smin[I];
                bmin[J] = smin[J];
                bmin[K] = cb ? qf_max(smin[K], -tp2) : smin[K];
                bmax[I] = smax[I];
                bmax[J] = smax[J];
                bmax[K] = cb ? qf_min(smax[K], +tp1) : smax[K];
            }
        }
        break;

        default:
        break;
    }
}

/* the frame of the clip box: its origin (the object's position, or none
 * for its own transform node) */
QRF void qf_origin(const qr_geom &g, float *pps)
{
    for (int a = 0; a < 3; a++) pps[a] = g.self_tn ? 0.0f : g.mtx[3][a];
}

/* a box of the clip-box frame into the shape's local axes */
QRF void qf_invert(const qr_geom &g, const float *smin, const float *smax, float *dmin,
                   float *dmax)
{
    float pps[3], tmin[3], tmax[3];
    qf_origin(g, pps);
    for (int a = 0; a < 3; a++)
    {
        tmin[a] = smin[a] == -QRF_INF ? -QRF_INF : smin[a] - pps[a];
        tmax[a] = smax[a] == +QRF_INF ? +QRF_INF : smax[a] - pps[a];
    }
    for (int a = 0; a < 3; a++)
    {
        tmin[a] = tmin[a] == -QRF_INF ? -QRF_INF : tmin[a] / g.scl[a];
        tmax[a] = tmax[a] == +QRF_INF ? +QRF_INF : tmax[a] / g.scl[a];
    }
    for (int a = 0; a < 3; a++)
    {
        const int m = g.map[a];
        dmin[a] = g.sgn[a] > 0 ? +tmin[m] : -tmax[m];
        dmax[a] = g.sgn[a] > 0 ? +tmax[m] : -tmin[m];
    }
}

/* a box of the shape's local axes into the clip-box frame */
QRF void qf_direct(const qr_geom &g, const float *smin, const float *smax, float *dmin,
                   float *dmax)
{
    float pps[3], tmin[3], tmax[3];
    qf_origin(g, pps);
    for (int a = 0; a < 3; a++)
    {
        const int m = g.map[a];
        tmin[m] = g.sgn[a] > 0 ? +smin[a] : -smax[a];
        tmax[m] = g.sgn[a] > 0 ? +smax[a] : -smin[a];
    }
    for (int a = 0; a < 3; a++)
    {
        tmin[a] = tmin[a] == -QRF_INF ? -QRF_INF : tmin[a] * g.scl[a];
        tmax[a] = tmax[a] == +QRF_INF ? +QRF_INF : tmax[a] * g.scl[a];
    }
    for (int a = 0; a < 3; a++)
    {
        dmin[a] = tmin[a] == -QRF_INF ? -QRF_INF : tmin[a] + pps[a];
        dmax[a] = tmax[a] == +QRF_INF ? +QRF_INF : tmax[a] + pps[a];
    }
}

/* the three uses of QuadRay's recalc_minmax:
 *  - no box (smin null): bbox and clip toggles of the surface from its own
 *    axis clips;
 *  - a box and no bbox (bmin null): the surface is a clipper of that box's
 *    owner; its extent within the box tightens cmin/cmax (an accumulated
 *    box in the owner's frame);
 *  - a box and a clip box: the owner's bbox and toggles within the
 *    accumulated box */
QRF void qf_recalc(const qr_geom &g, const float *smin, const float *smax, float *bmin,
                   float *bmax, float *cmin, float *cmax)
{
    float tmin[4] = {0.0f, 0.0f, 0.0f, 0.0f}, tmax[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float lmin[4] = {0.0f, 0.0f, 0.0f, 0.0f}, lmax[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float *pmin = nullptr, *pmax = nullptr;
    if (smin != nullptr && smax != nullptr && bmin == nullptr && bmax == nullptr)
    {
        qf_invert(g, smin, smax, tmin, tmax);
        bmin = lmin;
        bmax = lmax;
        pmin = cmin;
        pmax = cmax;
        cmin = nullptr;
        cmax = nullptr;
    }
    else if (smin != nullptr && smax != nullptr && cmin != nullptr && cmax != nullptr)
    {
        qf_invert(g, smin, smax, tmin, tmax);
        for (int a = 0; a < 3; a++)
        {
            tmin[a] = qf_max(tmin[a], g.smin[a]);
            tmax[a] = qf_min(tmax[a], g.smax[a]);
        }
    }
    else if (smin == nullptr && smax == nullptr)
    {
        for (int a = 0; a < 3; a++)
        {
            tmin[a] = g.smin[a];
            tmax[a] = g.smax[a];
        }
    }
    qf_adjust(g, tmin, tmax, bmin, bmax, cmin, cmax);
    if (pmin != nullptr && pmax != nullptr)
    {
        for (int a = 0; a < 3; a++)
        {
            tmin[a] = tmin[a] == bmin[a] ? -QRF_INF : bmin[a];
            tmax[a] = tmax[a] == bmax[a] ? +QRF_INF : bmax[a];
        }
        qf_direct(g, tmin, tmax, tmin, tmax);
        for (int a = 0; a < 3; a++)
        {
            pmin[a] = qf_max(pmin[a], tmin[a]);
            pmax[a] = qf_min(pmax[a], tmax[a]);
        }
        bmin = nullptr;
        bmax = nullptr;
    }
    if (bmin != nullptr && bmax != nullptr) qf_direct(g, bmin, bmax, bmin, bmax);
    if (cmin != nullptr && cmax != nullptr) qf_direct(g, cmin, cmax, cmin, cmax);
}

/* a surface's bbox and clip toggles (bmin, bmax, cmin, cmax: 4 each), with
 * its custom clippers' tightening when g.adjust: the outer clippers (not
 * planes) that share its transform node, outside accumulator segments
 * (QuadRay rt_Surface::update_minmax, object.cpp:2533-2710) */
QRF void qf_bounds(const qr_geom *geoms, const qr_clip *clips, int rec, float *bmin,
                   float *bmax, float *cmin, float *cmax)
{
    const qr_geom &g = geoms[rec];
    for (int a = 0; a < 4; a++)
    {
        bmin[a] = bmax[a] = cmin[a] = cmax[a] = 0.0f;
    }
    if (!g.adjust)
    {
        qf_recalc(g, nullptr, nullptr, bmin, bmax, cmin, cmax);
        return;
    }
    qf_recalc(g, nullptr, nullptr, bmin, bmax, nullptr, nullptr);
    for (int a = 0; a < 3; a++)
    {
        cmin[a] = -QRF_INF;
        cmax[a] = +QRF_INF;
    }
    int skip = 0;
    for (int k = 0; k < g.clip_cnt; k++)
    {
        const qr_clip c = clips[g.clip_beg + k];
        if (c.kind == QR_CLIP_ENTER || c.kind == QR_CLIP_LEAVE)
        {
            skip = 1 - skip;
            continue;
        }
        if (skip == 1 || c.kind != QR_CLIP_SURFACE) continue;
        const qr_geom &cg = geoms[c.surf];
        if (cg.tag == RT_TAG_PLANE || cg.tn_rec != g.tn_rec || c.rel != RT_REL_MINUS_OUTER)
        {
            continue;
        }
        qf_recalc(cg, bmin, bmax, nullptr, nullptr, cmin, cmax);
    }
    float amin[4], amax[4];
    for (int a = 0; a < 4; a++)
    {
        amin[a] = cmin[a];
        amax[a] = cmax[a];
    }
    qf_recalc(g, amin, amax, bmin, bmax, cmin, cmax);
}

/* the extent of every hit the tracer can accept: the shape within the
 * axis clips that stay on (cmin/cmax toggles, bmin/bmax values), in the
 * clip-box frame.  With custom clippers the tightened bbox need not bound
 * the accepted hits (a clipper's region may be unbounded), so culling uses
 * this instead. */
QRF void qf_cull_box(const qr_geom &g, const float *bmin, const float *bmax,
                     const float *cmin, const float *cmax, float *omin, float *omax)
{
    float emin[4] = {0.0f, 0.0f, 0.0f, 0.0f}, emax[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float tmin[4] = {0.0f, 0.0f, 0.0f, 0.0f}, tmax[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int a = 0; a < 3; a++)
    {
        emin[a] = cmin[a] == -QRF_INF ? -QRF_INF : bmin[a];
        emax[a] = cmax[a] == +QRF_INF ? +QRF_INF : bmax[a];
    }
    qf_invert(g, emin, emax, tmin, tmax);
    qf_adjust(g, tmin, tmax, omin, omax, nullptr, nullptr);
    qf_direct(g, omin, omax, omin, omax);
}

/* the record of a transform-node array */
QRF void qf_array_record(const qr_geom &g, qr_surface &r)
{
    memset(&r, 0, sizeof(r));
    float inv[4][4];
    qf_inverse(g.mtx, inv);
    for (int a = 0; a < 3; a++) r.pos[a] = g.mtx[3][a];
    for (int i = 0; i < 3; i++)
    {
        for (int x = 0; x < 3; x++) r.tc[i][x] = inv[x][i];
    }
    for (int a = 0; a < 3; a++)
    {
        r.amap[a] = g.map[a];
        r.asgn[a] = g.sgn[a] >= 0 ? 0 : 1;
    }
    r.shift = g.tn_rec >= 0 ? 1 : 0;
    r.trm = g.obj_has_trm;
    r.d_eps = 0.00000000001f;
    r.t_eps = 0.0000001f;
    r.trnode = g.tn_rec;
    r.tag = RT_TAG_ARRAY;
    r.id = -1;
    r.solver = QR_SOLVE_NONE;
    r.mat[0] = r.mat[1] = -1;
}

/* the record of a surface from its coefficients and clip box */
QRF void qf_surface_record(const qr_geom &g, const float *sci, const float *scj,
                           const float *bmin, const float *bmax, const float *cmin,
                           const float *cmax, qr_surface &r)
{
    memset(&r, 0, sizeof(r));
    for (int a = 0; a < 3; a++) r.pos[a] = g.mtx[3][a];
    if (g.self_tn)
    {
        float inv[4][4];
        qf_inverse(g.mtx, inv);
        for (int i = 0; i < 3; i++)
        {
            for (int x = 0; x < 3; x++) r.tc[i][x] = inv[x][i];
        }
    }
    for (int a = 0; a < 3; a++)
    {
        r.amap[a] = g.map[a];
        r.asgn[a] = g.sgn[a] >= 0 ? 0 : 1;
    }
    r.shift = g.tn_rec >= 0 ? 1 : 0;
    r.trm = g.obj_has_trm;
    r.d_eps = 0.00000000001f;
    r.t_eps = 0.0000001f;
    r.trnode = g.tn_rec;
    r.tag = g.tag;
    r.id = g.id;
    for (int a = 0; a < 4; a++) r.sci[a] = sci[a];
    for (int a = 0; a < 3; a++) r.scj[a] = scj[a] * 0.5f;
    float pps[3];
    qf_origin(g, pps);
    for (int a = 0; a < 3; a++)
    {
        r.cmin_on |= (cmin[a] == -QRF_INF ? 0 : 1) << a;
        r.cmax_on |= (cmax[a] == +QRF_INF ? 0 : 1) << a;
        r.cmin[a] = bmin[a] - pps[a];
        r.cmax[a] = bmax[a] - pps[a];
    }
    /* solver, normal and clipper kinds (QuadRay update0, tracer.cpp:8115-8150) */
    const int tag = g.tag;
    const bool flat_hyp = tag == RT_TAG_HYPERCYLINDER && r.sci[3] == 0.0f;
    const bool lin = tag == RT_TAG_PARABOLOID || tag == RT_TAG_PARACYLINDER ||
                     tag == RT_TAG_HYPERPARABOLOID;
    r.solver = tag > RT_TAG_PLANE ? (flat_hyp ? QR_SOLVE_TWOPLANE : QR_SOLVE_QUADRIC)
                                  : QR_SOLVE_PLANE;
    r.nrm_kind = tag > RT_TAG_PLANE ? (lin ? 2 : 3) : 1;
    r.clp_kind = r.nrm_kind;
    r.apex = tag == RT_TAG_CONE || (tag == RT_TAG_HYPERBOLOID && r.sci[3] == 0.0f)
             ? 1 : flat_hyp ? 2 : 0;
    for (int side = 0; side < 2; side++)
    {
        r.mat[side] = g.mat[side];
        r.props[side] = g.props[side];
    }
    r.clip_beg = g.clip_beg;
    r.clip_cnt = g.clip_cnt;
}

/* a surface element's inputs of the culling programs' boxes: its extent
 * within its clip box (world box through its transform node) and, for a
 * transformed surface, that extent in its tracing frame (the node's axes,
 * origin at the surface) grown by 2e-3 of its magnitude */
QRF void qf_refit_elem(const qr_geom &g, const float *bmin, const float *bmax,
                       const float *cmin, const float *cmax, qr_refit_elem &r)
{
    memset(&r, 0, sizeof(r));
    float emin[4] = {0.0f, 0.0f, 0.0f, 0.0f}, emax[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    qf_cull_box(g, bmin, bmax, cmin, cmax, emin, emax);
    bool bounded = true;
    for (int a = 0; a < 3; a++)
    {
        if (!(emin[a] > -QRF_INF) || !(emax[a] < QRF_INF) || !(emin[a] <= emax[a]))
        {
            bounded = false;
        }
        r.bmin[a] = emin[a];
        r.bmax[a] = emax[a];
    }
    r.world = bounded ? 1 : 0;
    const bool trn = g.tn_rec >= 0;
    r.trn = trn ? 1 : 0;
    for (int i = 0; i < 4; i++)
    {
        for (int x = 0; x < 3; x++) r.mtx[i][x] = trn ? g.tmtx[i][x] : 0.0f;
    }
    if (bounded && trn && g.obj_has_trm != 0)
    {
        float pps[3];
        qf_origin(g, pps);
        double lo[3], hi[3], mag = 1.0;
        for (int a = 0; a < 3; a++)
        {
            lo[a] = (double)emin[a] - pps[a];
            hi[a] = (double)emax[a] - pps[a];
            mag = refit_max(mag, refit_max(refit_abs(lo[a]), refit_abs(hi[a])));
            mag = refit_max(mag, hi[a] - lo[a]);
        }
        const double m = 2e-3 * mag;
        float lmin[3], lmax[3];
        for (int a = 0; a < 3; a++)
        {
            lmin[a] = (float)(lo[a] - m);
            lmax[a] = (float)(hi[a] + m);
        }
        refit_centre(lmin, lmax, r.lc, r.lh);
        r.local = 1;
    }
}

#endif /* QR_FIELDS_H */
