// Host build of the fast scan's tile culling (phygas_amd/csrc/cull.h) for the CPU tests: every
// (star, cluster, tile) through level 1 and level 2, so that a test can compare the dropped tiles with the
// chi2 of every point computed directly.
#include "cull.h"

using namespace phx;

extern "C" int cc_scan(const double *strk, int ns, int snpts, const double *ctrk, const int *cn, const double *R2,
                       int nc, int cnpts, double chi2_max, unsigned char *p1, unsigned char *p2) {
    const int ntl = (snpts + TILE - 1) / TILE;
    TileRec1 *s1 = new TileRec1[(size_t)ns * ntl];
    TileRec2 *s2 = new TileRec2[(size_t)ns * ntl];
    for (int s = 0; s < ns; s++)
        for (int j = 0; j < ntl; j++)
            tile_records(strk + (size_t)s * snpts * 12, snpts, j, 0.0, s1[(size_t)s * ntl + j], s2[(size_t)s * ntl + j]);
    const double chi2c = chi2_max * (1.0 + CUT_REL);
    const float kchi = up(sqrt(chi2c));
    for (int c = 0; c < nc; c++) {
        const int n = cn[c], nt = (n + TILE - 1) / TILE;
        for (int j = 0; j < ntl; j++) {
            TileRec1 c1 = {};
            TileRec2 c2 = {};
            if (j < nt) tile_records(ctrk + (size_t)c * cnpts * 12, n, j, R2[c], c1, c2);
            const float t1 = tile_t1(n, j);
            for (int s = 0; s < ns; s++) {
                const size_t o = ((size_t)s * nc + c) * ntl + j;
                const bool a = j < nt && l1_pass(s1[(size_t)s * ntl + j], c1, t1, kchi);
                p1[o] = a;
                p2[o] = a && l2_pass(s2[(size_t)s * ntl + j], c2, (double)t1, chi2c);
            }
        }
    }
    delete[] s1;
    delete[] s2;
    return ntl;
}

// Largest difference of the level-2 rotation constants from cos / sin of the frame rotation per point.
extern "C" double cc_rot_check(void) {
    return fmax(fabs(ROT_C - cos(OMEGA_PT)), fabs(ROT_S - sin(OMEGA_PT)));
}

// Level-1 passes of the padding star record against every tile of every cluster (0 expected).
extern "C" long cc_pad_passes(const double *ctrk, const int *cn, const double *R2, int nc, int cnpts, double chi2_max) {
    const float kchi = up(sqrt(chi2_max * (1.0 + CUT_REL)));
    const TileRec1 pad = pad_rec1();
    long passes = 0;
    for (int c = 0; c < nc; c++) {
        const int n = cn[c], nt = (n + TILE - 1) / TILE;
        for (int j = 0; j < nt; j++) {
            TileRec1 c1;
            TileRec2 c2;
            tile_records(ctrk + (size_t)c * cnpts * 12, n, j, R2[c], c1, c2);
            passes += l1_pass(pad, c1, tile_t1(n, j), kchi) ? 1 : 0;
        }
    }
    return passes;
}
