/*
 * lct_oracle.c -- FP64 CPU restatement of the LISTCTRecons helical FDK path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product.  See lct_oracle.h for the arithmetic convention.
 *
 * Every function cites the reference file:line it restates.  The loop nests
 * are reorganised for speed (z-independent terms are evaluated once per
 * (x, y, view) and reused for every requested slice), but each voxel still
 * sums the same terms in the same ascending-view order as the reference.
 */
#include "lct_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define LCO_PI 3.141592653589793238462643383 /* Include/Common.h:3 */
#define LCO_BLOCK_SLICES 32                  /* Include/Common.h:4 */

int lco_num_threads(void) { return omp_get_max_threads(); }
void lco_set_num_threads(int n) { if (n > 0) omp_set_num_threads(n); }

/* ------------------------------------------------------------------ planner */

/* Src/DataStructs.cpp:12-34.  The reference tests key against array[idx]
 * before testing idx < n (it reads one element past the end when the key is
 * beyond the data); the bound is tested first here, which returns the same
 * index whenever the reference's read is in bounds. */
int lco_array_search(float key, const float *a, int n, int type)
{
    int idx = 0;
    switch (type) {
    case -1: while (idx < n && key < a[idx]) idx++; break;
    case 0: while (idx < n && key != a[idx]) idx++; break;
    case 1: while (idx < n && key > a[idx]) idx++; break;
    default: break;
    }
    return idx;
}

static int direction_of(float a, float b)
{
    /* int(fabs(b-a)/(b-a)), with the reference's fallback to 1 when the
     * quotient is not +-1 (Src/DataStructs.cpp:196-198, :82-84). */
    float d = b - a;
    if (d == 0.0f || d != d) return 1;
    int r = (int)(fabsf(d) / d);
    if (r != 1 && r != -1) r = 1;
    return r;
}

/* Src/DataStructs.cpp:121-250 (float32 evaluation order preserved). */
int lco_plan_helical(const lco_prm *p, const float *ang, const float *couch, lco_plan *o)
{
    memset(o, 0, sizeof(*o));
    if (!p || !ang || !couch || p->readings < 101 || p->ns <= 0 || p->nt <= 0 ||
        p->n_proj_turn <= 0 || p->dz <= 0.0f)
        return LCO_ERR_ARG;

    /* :123  max(cc, ns-cc)*ds*npt / (2.0f*dso*PI) + 10, truncated to int */
    float cc = p->central_channel;
    float other = (float)p->ns - cc;
    float m = (cc < other) ? other : cc;
    float num = m * p->ds;
    num = num * (float)p->n_proj_turn;
    float den_f = 2.0f * p->r_src_to_iso;
    int add = (int)((double)num / ((double)den_f * LCO_PI) + 10);

    /* :169-176 */
    float d100 = couch[100] - couch[0];
    if (d100 == 0.0f) return LCO_ERR_ARG;
    int zdir = (int)(fabsf(d100) / d100);
    float orbit;
    if ((ang[1] < 1 && ang[0] > 6) || (ang[1] > 6 && ang[0] < 1))
        orbit = ang[2] - ang[1];
    else
        orbit = ang[1] - ang[0];
    float source_z0 = -p->pitch_value / 2.0f;

    /* :179-195 */
    float data_begin = couch[0];
    float data_end = couch[p->readings - 1];
    float pad = p->pitch_value * (float)(p->n_proj_turn / 2 + add + 256);
    pad = pad / (float)p->n_proj_turn;
    float allowed_begin = data_begin + (float)zdir * pad;
    float allowed_end = data_end - (float)zdir * pad;
    float start_pos, end_pos;
    if (zdir > 0) {
        start_pos = ceilf(allowed_begin + (float)(1 * zdir));
        end_pos = floorf(allowed_end - (float)(1 * zdir));
    } else {
        start_pos = floorf(allowed_begin + (float)(1 * zdir));
        end_pos = ceilf(allowed_end - (float)(1 * zdir));
    }
    int rd = direction_of(start_pos, end_pos);

    /* :200-208 */
    float recon_start = start_pos - (float)rd * p->dz;
    float recon_end = end_pos + (float)rd * p->dz;
    int n_req = (int)(floorf(fabsf(recon_end - recon_start) / p->dz) + 1.0f);
    int nsb = LCO_BLOCK_SLICES;
    int n_recon = (n_req - 1) - (n_req - 1) % nsb;
    recon_end = recon_start + (float)(rd * (n_recon - 1)) * p->dz;
    int n_blocks = n_recon / nsb;

    /* :210-218 */
    if (((start_pos > allowed_begin) && (start_pos > allowed_end)) ||
        ((start_pos < allowed_begin) && (start_pos < allowed_end)))
        return LCO_ERR_RANGE;
    if (((end_pos > allowed_begin) && (end_pos > allowed_end)) ||
        ((end_pos < allowed_begin) && (end_pos < allowed_end)))
        return LCO_ERR_RANGE;

    /* :222-240 */
    float block_start = recon_start;
    float block_end = block_start + (float)(rd * (nsb - 1)) * p->dz;
    int bps = lco_array_search(block_start, couch, p->readings, zdir);
    int bpe = lco_array_search(block_end, couch, p->readings, zdir);
    if (bps > bpe) {
        bps = bpe - p->n_proj_turn / 2 - add;
        bpe = bps + p->n_proj_turn / 2 + add;
    } else {
        bps = bps - p->n_proj_turn / 2 - add;
        bpe = bpe + p->n_proj_turn / 2 + add;
    }
    int naa = bpe - bps + 100;
    int na = naa - 2 * add;

    /* :242-245 and Include/DataStructs.h:26-46, :112-123 */
    o->ctg.orbit = orbit;
    o->ctg.orbit_start = ang[0];
    o->ctg.ns = p->ns;
    o->ctg.nt = p->nt;
    o->ctg.ws = p->central_channel;
    o->ctg.wt = (float)(p->nt - 1) / 2.0f;
    o->ctg.n_angle = na;
    o->ctg.n_angle_add = naa;
    o->ctg.ds = p->ds;
    o->ctg.dt = p->dt;
    o->ctg.pitch = p->pitch_value;
    o->ctg.source_z0 = source_z0;
    o->ctg.dso = p->r_src_to_iso;
    o->ctg.n_proj_turn = p->n_proj_turn;
    o->ctg.z_dir = zdir;
    o->ctg.add_projections = add;

    o->ig.nx = p->nx;
    o->ig.ny = p->ny;
    o->ig.nz = nsb;
    o->ig.dx = p->dx;
    o->ig.dy = p->dx; /* :243 passes Dx for dy */
    o->ig.dz = p->dz;
    o->ig.wx = (float)(p->nx - 1) / 2.0f + p->offset_x;
    o->ig.wy = (float)(p->ny - 1) / 2.0f + p->offset_y;
    o->ig.wz = 0.0f;

    o->n_readings = p->readings;
    o->data_begin_pos = data_begin;
    o->data_end_pos = data_end;
    o->allow_begin = allowed_begin;
    o->allow_end = allowed_end;
    o->start_pos = start_pos;
    o->end_pos = end_pos;
    o->recon_start_pos = recon_start;
    o->recon_end_pos = recon_end;
    o->n_slices_recon = n_recon;
    o->n_slices_block = nsb;
    o->n_blocks = n_blocks;
    o->data_offset = p->data_offset;
    o->recon_direction = rd;
    return LCO_OK;
}

/* Src/DataStructs.cpp:77-108 (block b) and :222-238 (block 0). */
int lco_block_info(const lco_plan *pl, const float *couch, int b, int *proj_start,
                   int *proj_end, float *block_start, float *block_end)
{
    if (!pl || !couch || b < 0) return LCO_ERR_ARG;
    const lco_ct_geom *g = &pl->ctg;
    int rd = direction_of(pl->start_pos, pl->end_pos);
    float bs = pl->recon_start_pos +
               (float)(rd * b) * pl->ig.dz * (float)pl->n_slices_block;
    float be = bs + (float)(rd * (pl->n_slices_block - 1)) * pl->ig.dz;
    int s = lco_array_search(bs, couch, pl->n_readings, g->z_dir);
    int e = lco_array_search(be, couch, pl->n_readings, g->z_dir);
    int half = g->n_proj_turn / 2;
    if (b == 0) {
        if (s > e) { s = e - half - g->add_projections; e = s + half + g->add_projections; }
        else { s = s - half - g->add_projections; e = e + half + g->add_projections; }
    } else {
        if (s > e) s = e - half - g->add_projections;
        else s = s - half - g->add_projections;
        e = s + g->n_angle_add;
    }
    if (proj_start) *proj_start = s;
    if (proj_end) *proj_end = e;
    if (block_start) *block_start = bs;
    if (block_end) *block_end = be;
    return LCO_OK;
}

/* ------------------------------------------------------------------ filter */

/* Src/Filter.cpp:13-18 evaluated in double: tau * h_KakSlaney(n tau). */
static double ramlak_tap(long n, double ds)
{
    if (n == 0) return 0.25 / ds;
    double nn = (double)n;
    return (sin(nn * LCO_PI) / (nn * LCO_PI) +
            (cos(nn * LCO_PI) - 1.0) / (nn * nn * LCO_PI * LCO_PI)) / (2.0 * ds);
}

void lco_ramp_taps(int ns, double ds, int window, double *h)
{
    for (int i = 0; i < 2 * ns; i++) {
        long n = (long)i - ns;
        double v;
        if (window == 1) {
            /* Hann window 0.5(1+cos(pi f/fN)) = 0.5 + 0.5 cos(2 pi f tau): in
             * space the ramp taps convolved with (1/4, 1/2, 1/4). */
            v = 0.5 * ramlak_tap(n, ds) + 0.25 * (ramlak_tap(n - 1, ds) + ramlak_tap(n + 1, ds));
        } else if (window == 2) {
            /* Shepp-Logan: |f| sinc(f tau) -> -2 / (pi^2 tau (4n^2 - 1)). */
            double nn = (double)n;
            v = -2.0 / (LCO_PI * LCO_PI * ds * (4.0 * nn * nn - 1.0));
        } else {
            v = ramlak_tap(n, ds);
        }
        h[i] = v;
    }
}

/* CUDA/Filter.cu:10-50 (k ascending, taps h[is-k+ns]). */
void lco_filter_rows(int ns, const double *h, const double *in, double *out, long rows)
{
#pragma omp parallel for schedule(static)
    for (long r = 0; r < rows; r++) {
        const double *row = in + r * ns;
        double *o = out + r * ns;
        for (int is = 0; is < ns; is++) {
            const double *hh = h + is + ns; /* hh[-k] = h[is-k+ns] */
            double acc = 0.0;
            for (int k = 0; k < ns; k++) acc += hh[-k] * row[k];
            o[is] = acc;
        }
    }
}

/* ------------------------------------------------------------------ rebin */

static inline double fan_at(const lco_ct_geom *g, const float *fan, long view, int it, int is)
{
    if (view < 0 || view >= g->n_angle_add || it < 0 || it >= g->nt || is < 0 || is >= g->ns)
        return 0.0;
    return (double)fan[((size_t)view * g->nt + it) * g->ns + is];
}

/* CUDA/Rebiner.cu:82-103: linear in the view axis only, zero border. */
void lco_rebin_fan2par(const lco_ct_geom *g, const float *fan, double *par)
{
    const int ns = g->ns, nt = g->nt, na = g->n_angle;
#pragma omp parallel for schedule(static)
    for (int ia = 0; ia < na; ia++) {
        for (int is = 0; is < ns; is++) {
            double ss = ((double)is - (double)g->ws) * (double)g->ds;
            double alpha = ss / (double)g->dso;
            double v = (double)ia + alpha / (double)g->orbit + (double)g->add_projections;
            double fv = floor(v);
            long i0 = (long)fv;
            double f = v - fv;
            for (int it = 0; it < nt; it++) {
                double a = fan_at(g, fan, i0, it, is);
                double b = fan_at(g, fan, i0 + 1, it, is);
                par[((size_t)ia * nt + it) * ns + is] = (1.0 - f) * a + f * b;
            }
        }
    }
}

/* CUDA/Rebiner.cu:13-29. */
void lco_rebin_null(const lco_ct_geom *g, const float *fan, double *par)
{
    const int ns = g->ns, nt = g->nt, na = g->n_angle;
#pragma omp parallel for schedule(static)
    for (int ia = 0; ia < na; ia++)
        for (int it = 0; it < nt; it++)
            for (int is = 0; is < ns; is++)
                par[((size_t)ia * nt + it) * ns + is] =
                    fan_at(g, fan, ia + g->add_projections, it, is);
}

/* CUDA/Rebiner.cu:156-179: trilinear, zero border in all three axes. */
void lco_rebin_pi_ori(const lco_ct_geom *g, const float *fan, double *par)
{
    const int ns = g->ns, nt = g->nt, na = g->n_angle;
    const double P = g->pitch, dso = g->dso, ds = g->ds, dt = g->dt;
#pragma omp parallel for schedule(static)
    for (int ia = 0; ia < na; ia++) {
        for (int it = 0; it < nt; it++) {
            for (int is = 0; is < ns; is++) {
                double s = ((double)is - (double)g->ws) * ds;
                double gam = asin(s / dso);
                double pI = gam * dso / ds + (double)g->ws;
                double q = (((double)it - (double)g->wt) * P / (2.0 * (nt - 1)) -
                            gam * P / (2.0 * LCO_PI)) / cos(gam);
                double qI = q / dt + (double)g->wt;
                double bI = (double)ia + gam / (double)g->orbit + (double)g->add_projections;
                double fx = floor(pI), fy = floor(qI), fz = floor(bI);
                long x0 = (long)fx, y0 = (long)fy, z0 = (long)fz;
                double ax = pI - fx, ay = qI - fy, az = bI - fz;
                double acc = 0.0;
                for (int dz = 0; dz < 2; dz++) {
                    double wz = dz ? az : 1.0 - az;
                    for (int dy = 0; dy < 2; dy++) {
                        double wy = dy ? ay : 1.0 - ay;
                        for (int dx = 0; dx < 2; dx++) {
                            double wx = dx ? ax : 1.0 - ax;
                            long xx = x0 + dx, yy = y0 + dy;
                            double v = 0.0;
                            if (xx >= 0 && xx < ns && yy >= 0 && yy < nt)
                                v = fan_at(g, fan, z0 + dz, (int)yy, (int)xx);
                            acc += wz * wy * wx * v;
                        }
                    }
                }
                par[((size_t)ia * nt + it) * ns + is] = acc;
            }
        }
    }
}

/* CUDA/PreWeighting.cu:8-29 (t uses 2*nt, not 2*(nt-1)). */
void lco_weight_pi_ori(const lco_ct_geom *g, double *par)
{
    const int ns = g->ns, nt = g->nt, na = g->n_angle;
    const double P = g->pitch, dso = g->dso, ds = g->ds;
#pragma omp parallel for schedule(static)
    for (int ia = 0; ia < na; ia++) {
        for (int it = 0; it < nt; it++) {
            for (int is = 0; is < ns; is++) {
                double s = ((double)is - (double)g->ws) * ds;
                double t = ((double)it - (double)g->wt) * P / (2.0 * nt);
                double gam = asin(s / dso);
                double temp = dso * dso - s * s;
                double temp2 = t - gam * P / (2.0 * LCO_PI);
                double w = sqrt(temp) / sqrt(temp + temp2 * temp2);
                par[((size_t)ia * nt + it) * ns + is] *= w;
            }
        }
    }
}

/* ------------------------------------------------------------ backproject */

/* tex3D(proj, p+0.5, q+0.5, ia+0.5) with linear filtering and border
 * addressing: bilinear in (channel p, row q) of view ia, zero outside. */
static inline double tap2(const lco_ct_geom *g, const double *q, long ia, double p, double r)
{
    if (ia < 0 || ia >= g->n_angle) return 0.0;
    const int ns = g->ns, nt = g->nt;
    if (!(p > -1.0 && p < (double)ns && r > -1.0 && r < (double)nt)) return 0.0;
    double fp = floor(p), fr = floor(r);
    long p0 = (long)fp, r0 = (long)fr;
    double ap = p - fp, ar = r - fr;
    const double *v = q + (size_t)ia * nt * ns;
    double v00 = (p0 >= 0 && r0 >= 0) ? v[r0 * ns + p0] : 0.0;
    double v01 = (p0 + 1 < ns && r0 >= 0) ? v[r0 * ns + p0 + 1] : 0.0;
    double v10 = (p0 >= 0 && r0 + 1 < nt) ? v[(r0 + 1) * ns + p0] : 0.0;
    double v11 = (p0 + 1 < ns && r0 + 1 < nt) ? v[(r0 + 1) * ns + p0 + 1] : 0.0;
    return (1.0 - ar) * ((1.0 - ap) * v00 + ap * v01) + ar * ((1.0 - ap) * v10 + ap * v11);
}

/* a_middle / a_min / a_max of CUDA/Backprojector.cu:33-42 in float32. */
static void bp_window(const lco_ct_geom *g, const lco_img_geom *ig, int iz, int *am, int *amin,
                      int *amax)
{
    float zc_f = ((float)iz - ig->wz) * ig->dz;
    float dzv_f = g->pitch / (float)g->n_proj_turn; /* Backprojector.cu:149 */
    int h = g->n_proj_turn / 2;
    int a = (int)roundf((zc_f - g->source_z0) / dzv_f);
    a = a < 0 ? 0 : a;
    a = a > g->n_angle - 1 ? g->n_angle - 1 : a;
    int lo = (a - h >= 0) ? (a - h) : 0;
    int hi = (a + h < g->n_angle) ? (a + h) : g->n_angle - 1;
    *am = a;
    *amin = lo;
    *amax = hi;
}

/* CUDA/Backprojector.cu:12-99. */
void lco_bp_helical(const lco_ct_geom *g, const lco_img_geom *ig, const double *q,
                    const float *betas, const int *iz_list, int n_iz, int iy0, int iy1,
                    double *img)
{
    const int nx = ig->nx, ny = ig->ny, na = g->n_angle, h = g->n_proj_turn / 2;
    const double dso = g->dso, ds = g->ds, dt = g->dt, ws = g->ws, wt = g->wt;
    const double orbit = g->orbit, z0 = g->source_z0;
    const double dzv = (double)g->pitch / (double)g->n_proj_turn;
    const double zdir = (double)g->z_dir;
    if (n_iz <= 0 || iy1 <= iy0) return;

    int *amin = (int *)malloc(sizeof(int) * n_iz), *amax = (int *)malloc(sizeof(int) * n_iz);
    double *zc = (double *)malloc(sizeof(double) * n_iz);
    int vlo = na, vhi = 0;
    for (int k = 0; k < n_iz; k++) {
        int am;
        bp_window(g, ig, iz_list[k], &am, &amin[k], &amax[k]);
        zc[k] = ((double)iz_list[k] - (double)ig->wz) * (double)ig->dz;
        if (amin[k] < vlo) vlo = amin[k];
        if (amax[k] > vhi) vhi = amax[k];
    }
    /* cos/sin of every float32 angle this block may touch (ia2 can reach
     * ia + h beyond a_max on clipped windows). */
    int nb = vhi + h + 1; /* callers provide n_angle + npt/2 + 1 angles */
    double *cb = (double *)malloc(sizeof(double) * nb), *sb = (double *)malloc(sizeof(double) * nb);
    for (int i = 0; i < nb; i++) {
        cb[i] = cos((double)betas[i]);
        sb[i] = sin((double)betas[i]);
    }

#pragma omp parallel
    {
        double *acc = (double *)malloc(sizeof(double) * n_iz);
#pragma omp for collapse(2) schedule(dynamic, 4)
        for (int iy = iy0; iy < iy1; iy++) {
            for (int ix = 0; ix < nx; ix++) {
                double xc = ((double)ix - (double)ig->wx) * (double)ig->dx;
                double yc = ((double)iy - (double)ig->wy) * (double)ig->dy;
                for (int k = 0; k < n_iz; k++) acc[k] = 0.0;
                for (int ia = vlo; ia < vhi; ia++) {
                    double phat = -xc * cb[ia] - yc * sb[ia];
                    double theta = asin(phat / dso);
                    double delta = theta * dso / ds;
                    double p1 = delta + ws, p2 = -delta + ws;
                    double root = sqrt(dso * dso - phat * phat);
                    double zr = z0 + ((double)ia + theta / orbit) * dzv;
                    double lhat = root + xc * sb[ia] - yc * cb[ia];
                    for (int k = 0; k < n_iz; k++) {
                        if (ia < amin[k] || ia >= amax[k]) continue;
                        int ia2 = ia + h < amax[k] ? ia + h : ia - h;
                        ia2 = ia2 < 0 ? ia + h : ia2;
                        double zr2 = z0 + ((double)ia2 - theta / orbit) * dzv;
                        double lhat2 = root + xc * sb[ia2] - yc * cb[ia2];
                        double qh = (zc[k] - zr) * dso * zdir / (lhat * dt);
                        double qh2 = (zc[k] - zr2) * dso * zdir / (lhat2 * dt);
                        double w1, w2;
                        if (fabs(qh) > wt) { w1 = 0.0; w2 = 1.0; }
                        else if (fabs(qh2) > wt) { w1 = 1.0; w2 = 0.0; }
                        else { w1 = -qh2 / (qh - qh2); w2 = qh / (qh - qh2); }
                        double a1 = tap2(g, q, ia, p1, qh + wt) * w1;
                        double a2 = tap2(g, q, ia2, p2, qh2 + wt) * w2;
                        acc[k] += a1 + a2;
                    }
                }
                for (int k = 0; k < n_iz; k++)
                    img[((size_t)iz_list[k] * ny + iy) * nx + ix] =
                        acc[k] * 2.0 * LCO_PI / (double)(amax[k] - amin[k]);
            }
        }
        free(acc);
    }
    free(amin); free(amax); free(zc); free(cb); free(sb);
}

/* CUDA/Backprojector.cu:171-250.  Views at or beyond n_angle contribute 0
 * (the reference's border-addressed texture returns 0 there). */
void lco_bp_pi_ori(const lco_ct_geom *g, const lco_img_geom *ig, const double *q,
                   const float *betas, const int *iz_list, int n_iz, int iy0, int iy1,
                   double *img)
{
    const int nx = ig->nx, ny = ig->ny, h = g->n_proj_turn / 2;
    const double dso = g->dso, ds = g->ds, ws = g->ws, wt = g->wt, P = g->pitch;
    const double z0 = g->source_z0, zdir = (double)g->z_dir;
    const double dzv = (double)g->pitch / (double)g->n_proj_turn;
    const double temp2 = P / (2.0 * LCO_PI);
    const int nt = g->nt;
#pragma omp parallel for collapse(3) schedule(dynamic, 4)
    for (int k = 0; k < n_iz; k++) {
        for (int iy = iy0; iy < iy1; iy++) {
            for (int ix = 0; ix < nx; ix++) {
                int am, amin, amax;
                bp_window(g, ig, iz_list[k], &am, &amin, &amax);
                double xc = ((double)ix - (double)ig->wx) * (double)ig->dx;
                double yc = ((double)iy - (double)ig->wy) * (double)ig->dy;
                double zc = ((double)iz_list[k] - (double)ig->wz) * (double)ig->dz;
                double sum = 0.0;
                int ia;
                for (ia = am; ia >= 0; ia--) {
                    double cbv = cos((double)betas[ia]), sbv = sin((double)betas[ia]);
                    double s = -xc * cbv - yc * sbv;
                    double gam = asin(s / dso);
                    double temp1 = sqrt(dso * dso - s * s);
                    double zoff = zc - (z0 + dzv * ia);
                    double t = temp1 * (zoff - gam * temp2) / (temp1 + xc * sbv - yc * cbv) + gam * temp2;
                    t *= zdir;
                    if (t > P / 4.0 || t < -P / 4.0) break;
                    double is = s / ds + ws;
                    double it = t * 2.0 * (nt - 1) / P + wt;
                    sum += tap2(g, q, ia, is, it);
                }
                int a_min = ia;
                for (ia = am + 1; ia <= a_min + h; ia++) {
                    if (ia >= g->n_angle) break; /* border: every later term is 0 */
                    double cbv = cos((double)betas[ia]), sbv = sin((double)betas[ia]);
                    double s = -xc * cbv - yc * sbv;
                    double gam = asin(s / dso);
                    double temp1 = sqrt(dso * dso - s * s);
                    double zoff = zc - (z0 + dzv * ia);
                    double t = temp1 * (zoff - gam * temp2) / (temp1 + xc * sbv - yc * cbv) + gam * temp2;
                    t *= zdir;
                    double is = s / ds + ws;
                    double it = t * 2.0 * (nt - 1) / P + wt;
                    sum += tap2(g, q, ia, is, it);
                }
                img[((size_t)iz_list[k] * ny + iy) * nx + ix] = sum * LCO_PI / (double)h;
            }
        }
    }
}

/* ------------------------------------------------------------- pipeline */

/* Src/Reconstructor.cpp:33-48 with Src/DataReader.cpp:8-19 (64-bit offsets)
 * and the betas upload of CUDA/Backprojector.cu:145. */
int lco_reconstruct(const lco_plan *pl, const float *fan_file, long file_views,
                    const float *angles, const float *couch, int rebin_kind,
                    int weight_kind, int window, int bp_kind, int b0, int b1,
                    const int *iz_list, int n_iz, int iy0, int iy1, double *vol)
{
    const lco_ct_geom *g = &pl->ctg;
    const lco_img_geom *ig = &pl->ig;
    const size_t view_sz = (size_t)g->ns * g->nt;
    const size_t par_sz = view_sz * (size_t)g->n_angle;
    const size_t slab = (size_t)ig->nx * ig->ny;
    if (b0 < 0 || b1 > pl->n_blocks || b0 >= b1) return LCO_ERR_ARG;

    double *taps = (double *)malloc(sizeof(double) * 2 * g->ns);
    double *par = (double *)malloc(sizeof(double) * par_sz);
    double *qf = (double *)malloc(sizeof(double) * par_sz);
    double *img = (double *)calloc(slab * ig->nz, sizeof(double));
    float *betas = (float *)malloc(sizeof(float) * (g->n_angle + g->n_proj_turn / 2 + 1));
    if (!taps || !par || !qf || !img || !betas) {
        free(taps); free(par); free(qf); free(img); free(betas);
        return LCO_ERR_ARG;
    }
    lco_ramp_taps(g->ns, (double)g->ds, window, taps);
    int rc = LCO_OK;
    for (int b = b0; b < b1; b++) {
        int bps;
        lco_block_info(pl, couch, b, &bps, NULL, NULL, NULL);
        long first = (long)bps + pl->data_offset;
        if (first < 0 || first + g->n_angle_add > file_views || bps < 0 ||
            bps + g->add_projections + g->n_angle > pl->n_readings) {
            rc = LCO_ERR_RANGE;
            break;
        }
        const float *fan = fan_file + (size_t)first * view_sz;
        if (rebin_kind == 0) lco_rebin_null(g, fan, par);
        else if (rebin_kind == 2) lco_rebin_pi_ori(g, fan, par);
        else lco_rebin_fan2par(g, fan, par);
        if (weight_kind == 1) lco_weight_pi_ori(g, par);
        lco_filter_rows(g->ns, taps, par, qf, (long)g->n_angle * g->nt);
        /* betas[0..NA) = tubeAngles[bps+add ...] (Backprojector.cu:145); the
         * conjugate index of a clipped window can reach NA + npt/2, where the
         * reference reads an uninitialised part of its buffer: the true
         * angles are used when the file has them, 0 otherwise. */
        int nbeta = g->n_angle + g->n_proj_turn / 2 + 1;
        for (int i = 0; i < nbeta; i++) {
            long src = (long)bps + g->add_projections + i;
            betas[i] = src < pl->n_readings ? angles[src] : 0.0f;
        }
        if (bp_kind == 1) lco_bp_pi_ori(g, ig, qf, betas, iz_list, n_iz, iy0, iy1, img);
        else lco_bp_helical(g, ig, qf, betas, iz_list, n_iz, iy0, iy1, img);
        for (int k = 0; k < n_iz; k++) {
            size_t src = (size_t)iz_list[k] * slab;
            size_t dst = ((size_t)b * pl->n_slices_block + iz_list[k]) * slab;
            for (int iy = iy0; iy < iy1; iy++)
                memcpy(vol + dst + (size_t)iy * ig->nx, img + src + (size_t)iy * ig->nx,
                       sizeof(double) * ig->nx);
        }
    }
    free(taps); free(par); free(qf); free(img); free(betas);
    return rc;
}

/* ------------------------------------------------------ analytic phantom */

static inline double chord_sum(const double *ell, int n_ell, const double S[3], const double d[3])
{
    double total = 0.0;
    for (int e = 0; e < n_ell; e++) {
        const double *E = ell + (size_t)e * LCO_ELL_STRIDE;
        const double *R = E + 6;
        double r0 = S[0] - E[0], r1 = S[1] - E[1], r2 = S[2] - E[2];
        /* body coordinates: R^T (S - c) / axes and R^T d / axes */
        double p0 = (R[0] * r0 + R[3] * r1 + R[6] * r2) / E[3];
        double p1 = (R[1] * r0 + R[4] * r1 + R[7] * r2) / E[4];
        double p2 = (R[2] * r0 + R[5] * r1 + R[8] * r2) / E[5];
        double q0 = (R[0] * d[0] + R[3] * d[1] + R[6] * d[2]) / E[3];
        double q1 = (R[1] * d[0] + R[4] * d[1] + R[7] * d[2]) / E[4];
        double q2 = (R[2] * d[0] + R[5] * d[1] + R[8] * d[2]) / E[5];
        double A = q0 * q0 + q1 * q1 + q2 * q2;
        double B = p0 * q0 + p1 * q1 + p2 * q2;
        double C = p0 * p0 + p1 * p1 + p2 * p2 - 1.0;
        double disc = B * B - A * C;
        if (disc > 0.0) total += E[15] * 2.0 * sqrt(disc) / A;
    }
    return total;
}

void lco_project_ellipsoids(const lco_ct_geom *g, const float *angles, const float *couch,
                            long view0, long n_views, const double *ell, int n_ell,
                            double *out)
{
    const int ns = g->ns, nt = g->nt;
    const double dso = g->dso;
#pragma omp parallel for schedule(dynamic, 1)
    for (long v = 0; v < n_views; v++) {
        double b = (double)angles[view0 + v];
        double S[3] = {-dso * sin(b), dso * cos(b), (double)couch[view0 + v]};
        for (int is = 0; is < ns; is++) {
            double a = ((double)is - (double)g->ws) * (double)g->ds / dso;
            double ux = sin(b - a), uy = -cos(b - a);
            for (int it = 0; it < nt; it++) {
                double slope = ((double)it - (double)g->wt) * (double)g->dt / dso;
                double inv = 1.0 / sqrt(1.0 + slope * slope);
                double d[3] = {ux * inv, uy * inv, slope * inv};
                out[((size_t)v * nt + it) * ns + is] = chord_sum(ell, n_ell, S, d);
            }
        }
    }
}

void lco_voxelize(const lco_img_geom *ig, double z0, int nz, const double *ell, int n_ell,
                  int supersample, double *vol)
{
    const int nx = ig->nx, ny = ig->ny;
    const int ss = supersample < 1 ? 1 : supersample;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < nz; iz++) {
        for (int iy = 0; iy < ny; iy++) {
            for (int ix = 0; ix < nx; ix++) {
                double acc = 0.0;
                for (int sz = 0; sz < ss; sz++)
                    for (int sy = 0; sy < ss; sy++)
                        for (int sx = 0; sx < ss; sx++) {
                            double ox = ((sx + 0.5) / ss - 0.5), oy = ((sy + 0.5) / ss - 0.5),
                                   oz = ((sz + 0.5) / ss - 0.5);
                            double x = ((double)ix + ox - (double)ig->wx) * (double)ig->dx;
                            double y = ((double)iy + oy - (double)ig->wy) * (double)ig->dy;
                            double z = z0 + ((double)iz + oz) * (double)ig->dz;
                            for (int e = 0; e < n_ell; e++) {
                                const double *E = ell + (size_t)e * LCO_ELL_STRIDE;
                                const double *R = E + 6;
                                double r0 = x - E[0], r1 = y - E[1], r2 = z - E[2];
                                double p0 = (R[0] * r0 + R[3] * r1 + R[6] * r2) / E[3];
                                double p1 = (R[1] * r0 + R[4] * r1 + R[7] * r2) / E[4];
                                double p2 = (R[2] * r0 + R[5] * r1 + R[8] * r2) / E[5];
                                if (p0 * p0 + p1 * p1 + p2 * p2 <= 1.0) acc += E[15];
                            }
                        }
                vol[((size_t)iz * ny + iy) * nx + ix] = acc / (double)(ss * ss * ss);
            }
        }
    }
}

/* Trilinear sample at voxel-index coordinates (fx, fy, fz), zero outside. */
static double trilinear(const double *vol, int nx, int ny, int nz, double fx, double fy,
                        double fz)
{
    const double x0f = floor(fx), y0f = floor(fy), z0f = floor(fz);
    const int x0 = (int)x0f, y0 = (int)y0f, z0 = (int)z0f;
    const double ax = fx - x0f, ay = fy - y0f, az = fz - z0f;
    double acc = 0.0;
    for (int k = 0; k < 2; k++) {
        const int zz = z0 + k;
        if (zz < 0 || zz >= nz) continue;
        for (int j = 0; j < 2; j++) {
            const int yy = y0 + j;
            if (yy < 0 || yy >= ny) continue;
            for (int i = 0; i < 2; i++) {
                const int xx = x0 + i;
                if (xx < 0 || xx >= nx) continue;
                const double w = (k ? az : 1.0 - az) * (j ? ay : 1.0 - ay) * (i ? ax : 1.0 - ax);
                acc += w * vol[((size_t)zz * ny + yy) * nx + xx];
            }
        }
    }
    return acc;
}

/* Adjoint of trilinear(): adds val times each of the eight weights to the
 * voxels that trilinear() reads (the same zero-border rule). */
static void trilinear_scatter(double *vol, int nx, int ny, int nz, double fx, double fy,
                              double fz, double val)
{
    const double x0f = floor(fx), y0f = floor(fy), z0f = floor(fz);
    const int x0 = (int)x0f, y0 = (int)y0f, z0 = (int)z0f;
    const double ax = fx - x0f, ay = fy - y0f, az = fz - z0f;
    for (int k = 0; k < 2; k++) {
        const int zz = z0 + k;
        if (zz < 0 || zz >= nz) continue;
        for (int j = 0; j < 2; j++) {
            const int yy = y0 + j;
            if (yy < 0 || yy >= ny) continue;
            for (int i = 0; i < 2; i++) {
                const int xx = x0 + i;
                if (xx < 0 || xx >= nx) continue;
                const double w = (k ? az : 1.0 - az) * (j ? ay : 1.0 - ay) * (i ? ax : 1.0 - ax);
#pragma omp atomic
                vol[((size_t)zz * ny + yy) * nx + xx] += w * val;
            }
        }
    }
}

/* The ray of view (b, sz), row it, channel is (source S, unit direction d)
 * and its chord [t0, t1] through the box of voxel centres lo..hi; returns 0
 * when the ray misses the box. */
static int ray_chord(const lco_ct_geom *g, double b, double sz, int it, int is,
                     const double lo[3], const double hi[3], double S[3], double d[3],
                     double *t0_out, double *t1_out)
{
    const double dso = g->dso;
    const double slope = ((double)it - (double)g->wt) * (double)g->dt / dso;
    const double inv = 1.0 / sqrt(1.0 + slope * slope);
    const double a = ((double)is - (double)g->ws) * (double)g->ds / dso;
    S[0] = -dso * sin(b);
    S[1] = dso * cos(b);
    S[2] = sz;
    d[0] = sin(b - a) * inv;
    d[1] = -cos(b - a) * inv;
    d[2] = slope * inv;
    double t0 = -1e300, t1 = 1e300;
    for (int k = 0; k < 3; k++) {
        if (fabs(d[k]) < 1e-300) {
            if (S[k] < lo[k] || S[k] > hi[k]) t1 = -1e300;
            continue;
        }
        double ta = (lo[k] - S[k]) / d[k], tb = (hi[k] - S[k]) / d[k];
        if (ta > tb) { const double tmp = ta; ta = tb; tb = tmp; }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
    }
    *t0_out = t0;
    *t1_out = t1;
    return t1 > t0;
}

void lco_forward_project(const lco_ct_geom *g, const float *angles, const float *couch,
                         long view0, long n_views, const double *vol, const lco_img_geom *ig,
                         double z0, int nz, double step, double *out)
{
    const int ns = g->ns, nt = g->nt, nx = ig->nx, ny = ig->ny;
    const double dx = ig->dx, dy = ig->dy, dz = ig->dz;
    const double wx = ig->wx, wy = ig->wy;
    const double lo[3] = {-wx * dx, -wy * dy, z0};
    const double hi[3] = {(nx - 1 - wx) * dx, (ny - 1 - wy) * dy, z0 + (nz - 1) * dz};
#pragma omp parallel for schedule(dynamic, 1)
    for (long v = 0; v < n_views; v++) {
        const double b = (double)angles[view0 + v];
        for (int it = 0; it < nt; it++) {
            for (int is = 0; is < ns; is++) {
                double S[3], d[3], t0, t1;
                double sum = 0.0;
                if (ray_chord(g, b, (double)couch[view0 + v], it, is, lo, hi, S, d, &t0, &t1)) {
                    const long n = (long)ceil((t1 - t0) / step);
                    const double h = (t1 - t0) / (double)n;
                    for (long k = 0; k < n; k++) {
                        const double t = t0 + ((double)k + 0.5) * h;
                        sum += trilinear(vol, nx, ny, nz, (S[0] + t * d[0]) / dx + wx,
                                         (S[1] + t * d[1]) / dy + wy, (S[2] + t * d[2] - z0) / dz);
                    }
                    sum *= h;
                }
                out[((size_t)v * nt + it) * ns + is] = sum;
            }
        }
    }
}

void lco_forward_project_adjoint(const lco_ct_geom *g, const float *angles, const float *couch,
                                 long view0, long n_views, const double *proj,
                                 const lco_img_geom *ig, double z0, int nz, double step,
                                 double *vol)
{
    const int ns = g->ns, nt = g->nt, nx = ig->nx, ny = ig->ny;
    const double dx = ig->dx, dy = ig->dy, dz = ig->dz;
    const double wx = ig->wx, wy = ig->wy;
    const double lo[3] = {-wx * dx, -wy * dy, z0};
    const double hi[3] = {(nx - 1 - wx) * dx, (ny - 1 - wy) * dy, z0 + (nz - 1) * dz};
    memset(vol, 0, sizeof(double) * (size_t)nz * ny * nx);
#pragma omp parallel for schedule(dynamic, 1)
    for (long v = 0; v < n_views; v++) {
        const double b = (double)angles[view0 + v];
        for (int it = 0; it < nt; it++) {
            for (int is = 0; is < ns; is++) {
                double S[3], d[3], t0, t1;
                if (!ray_chord(g, b, (double)couch[view0 + v], it, is, lo, hi, S, d, &t0, &t1))
                    continue;
                const long n = (long)ceil((t1 - t0) / step);
                const double h = (t1 - t0) / (double)n;
                const double val = h * proj[((size_t)v * nt + it) * ns + is];
                for (long k = 0; k < n; k++) {
                    const double t = t0 + ((double)k + 0.5) * h;
                    trilinear_scatter(vol, nx, ny, nz, (S[0] + t * d[0]) / dx + wx,
                                      (S[1] + t * d[1]) / dy + wy, (S[2] + t * d[2] - z0) / dz,
                                      val);
                }
            }
        }
    }
}
