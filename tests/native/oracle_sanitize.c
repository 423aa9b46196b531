/* Drives every function of the FP64 oracle (oracle/lct_oracle.c) on a tiny
 * helical scan (64 x 8 detector, 96 views per turn, 1100 views, 32^2 x 32
 * volume) so that it can be built with AddressSanitizer and
 * UndefinedBehaviorSanitizer: the planner, block windows, taps and windows,
 * the three rebins and the pre-weight, the filter, both backprojectors, the
 * block loop, the ellipsoid projector, the voxeliser and the forward
 * projector with its adjoint.  Prints "ok" and exits 0, or names the failed
 * step and exits 1. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "lct_oracle.h"

#define FAIL(msg)                                \
    do {                                         \
        printf("FAIL %s\n", msg);                \
        return 1;                                \
    } while (0)

int main(void)
{
    const int ns = 64, nt = 8, npt = 96, readings = 1100;
    const double pi = 3.141592653589793;
    lco_prm p = {0};
    p.ns = ns;
    p.nt = nt;
    p.ds = p.dt = 2.0f;
    p.pitch_value = 1.0f * nt * 2.0f;
    p.nx = p.ny = 32;
    p.dx = 3.0f;
    p.dz = 2.0f;
    p.readings = readings;
    p.r_src_to_iso = 541.0f;
    p.central_channel = (ns - 1) / 2.0f;
    p.n_proj_turn = npt;
    float *ang = malloc(sizeof(float) * readings), *couch = malloc(sizeof(float) * readings);
    for (int g = 0; g < readings; g++) {
        ang[g] = (float)fmod(0.25 + g * (2.0 * pi / npt), 2.0 * pi);
        couch[g] = (float)(g * (double)p.pitch_value / npt);
    }
    lco_plan pl;
    if (lco_plan_helical(&p, ang, couch, &pl) != LCO_OK || pl.n_blocks < 1) FAIL("plan");
    for (int b = 0; b < pl.n_blocks; b++) {
        int s, e;
        float bs, be;
        if (lco_block_info(&pl, couch, b, &s, &e, &bs, &be) != LCO_OK) FAIL("block_info");
    }
    if (lco_array_search(couch[readings - 1] + 1.0f, couch, readings, 1) < -1) FAIL("search");

    /* projection file from one ellipsoid */
    double ell[LCO_ELL_STRIDE] = {5, -3, couch[readings / 2], 30, 25, 20, 1, 0, 0, 0, 1, 0, 0, 0, 1,
                                  0.02};
    double *proj = malloc(sizeof(double) * (size_t)readings * nt * ns);
    lco_project_ellipsoids(&pl.ctg, ang, couch, 0, readings, ell, 1, proj);
    float *fan = malloc(sizeof(float) * (size_t)readings * nt * ns);
    for (size_t i = 0; i < (size_t)readings * nt * ns; i++) fan[i] = (float)proj[i];

    /* the block loop with every stage kind */
    const size_t nvox = (size_t)pl.n_slices_recon * pl.ig.ny * pl.ig.nx;
    double *vol = malloc(sizeof(double) * nvox);
    int iz[32];
    for (int i = 0; i < 32; i++) iz[i] = i;
    const int kinds[4][4] = {{1, 0, 0, 0}, {0, 0, 1, 0}, {2, 1, 2, 1}, {1, 0, 0, 1}};
    for (int k = 0; k < 4; k++)
        if (lco_reconstruct(&pl, fan, readings, ang, couch, kinds[k][0], kinds[k][1], kinds[k][2],
                            kinds[k][3], 0, pl.n_blocks, iz, 32, 0, pl.ig.ny, vol) != LCO_OK)
            FAIL("reconstruct");
    double sum = 0;
    for (size_t i = 0; i < nvox; i++) sum += vol[i];
    if (!isfinite(sum)) FAIL("non-finite volume");

    /* stages on their own: taps, rebins, weight, filter */
    const lco_ct_geom *g = &pl.ctg;
    const size_t nview = (size_t)nt * ns;
    double *h = malloc(sizeof(double) * 2 * ns);
    double *par = malloc(sizeof(double) * g->n_angle * nview);
    double *flt = malloc(sizeof(double) * g->n_angle * nview);
    for (int w = 0; w < 3; w++) lco_ramp_taps(ns, g->ds, w, h);
    lco_rebin_null(g, fan, par);
    lco_rebin_pi_ori(g, fan, par);
    lco_weight_pi_ori(g, par);
    lco_rebin_fan2par(g, fan, par);
    lco_filter_rows(ns, h, par, flt, (long)g->n_angle * nt);
    double *img = malloc(sizeof(double) * 32 * pl.ig.ny * pl.ig.nx);
    lco_bp_helical(g, &pl.ig, flt, ang + g->add_projections, iz, 32, 0, pl.ig.ny, img);
    lco_bp_pi_ori(g, &pl.ig, flt, ang + g->add_projections, iz, 32, 0, pl.ig.ny, img);

    /* voxeliser, forward projector and its adjoint */
    lco_img_geom vg = {24, 20, 12, 11.5f, 9.5f, 0.0f, 4.0f, 4.0f, 3.0f};
    const double z0 = couch[500] - 18.0;
    double *v = malloc(sizeof(double) * 12 * 20 * 24);
    double *fp = malloc(sizeof(double) * 20 * nview);
    lco_voxelize(&vg, z0, 12, ell, 1, 2, v);
    lco_forward_project(g, ang, couch, 490, 20, v, &vg, z0, 12, 1.5, fp);
    lco_forward_project_adjoint(g, ang, couch, 490, 20, fp, &vg, z0, 12, 1.5, v);

    free(ang);
    free(couch);
    free(proj);
    free(fan);
    free(vol);
    free(h);
    free(par);
    free(flt);
    free(img);
    free(v);
    free(fp);
    printf("ok\n");
    return 0;
}
