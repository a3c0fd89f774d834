// Host build of the fast scan's tile culling (phygas_amd/csrc/cull.h) for the CPU tests: every
// (star, cluster, tile) through level 1 and level 2, so that a test can compare the dropped tiles with the
// chi2 of every point computed directly.
#include "cull.h"

using namespace phx;

// p2: level 2 of the fine tiles as the scan runs it (l2f_pass); *n_l2u: level-1 survivors that the single-U
// level 2 (l2_pass) would leave.
extern "C" int cc_scan(const double *strk, int ns, int snpts, const double *ctrk, const int *cn, const double *R2,
                       int nc, int cnpts, double chi2_max, unsigned char *p1, unsigned char *p2, long *n_l2u) {
    const int ntl = (snpts + TILE - 1) / TILE;
    TileRec1 *s1 = new TileRec1[(size_t)ns * ntl];
    TileRec2 *s2 = new TileRec2[(size_t)ns * ntl];
    TileRecF *sf = new TileRecF[(size_t)ns * ntl];
    for (int s = 0; s < ns; s++)
        for (int j = 0; j < ntl; j++) {
            TileRec1 u1;
            tile_records(strk + (size_t)s * snpts * 12, snpts, j, 0.0, u1, s2[(size_t)s * ntl + j]);   // one U
            tile_records(strk + (size_t)s * snpts * 12, snpts, j, 0.0, s1[(size_t)s * ntl + j], s2[(size_t)s * ntl + j],
                         &sf[(size_t)s * ntl + j]);
        }
    *n_l2u = 0;
    const double chi2c = chi2_max * (1.0 + CUT_REL);
    const float kchi = up(sqrt(chi2c));
    for (int c = 0; c < nc; c++) {
        const int n = cn[c], nt = (n + TILE - 1) / TILE;
        for (int j = 0; j < ntl; j++) {
            TileRec1 c1 = {}, u1;
            TileRec2 c2 = {};
            TileRecF cf = {};
            if (j < nt) {
                tile_records(ctrk + (size_t)c * cnpts * 12, n, j, R2[c], u1, c2);   // one U
                tile_records(ctrk + (size_t)c * cnpts * 12, n, j, R2[c], c1, c2, &cf);
            }
            const float t1 = tile_t1(n, j);
            for (int s = 0; s < ns; s++) {
                const size_t o = ((size_t)s * nc + c) * ntl + j;
                const bool a = j < nt && l1_pass(s1[(size_t)s * ntl + j], c1, t1, kchi);
                p1[o] = a;
                p2[o] = a && l2f_pass(sf[(size_t)s * ntl + j], cf, up(chi2c));
                *n_l2u += a && l2_pass(s2[(size_t)s * ntl + j], c2, (double)t1, chi2c);
            }
        }
    }
    delete[] s1;
    delete[] s2;
    delete[] sf;
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

// The scan's two-level order: a coarse tile (TILE2 points) through level 1 and the interpolated level 2
// first, then each of its two tiles through the fine level 2 when the coarse tile is left.  p2 as cc_scan's
// (tiles left); counts[0..3] = coarse level-1 tests, coarse level-2 tests, fine level-2 tests, fine tiles left.
extern "C" int cc_scan2(const double *strk, int ns, int snpts, const double *ctrk, const int *cn, const double *R2,
                        int nc, int cnpts, double chi2_max, unsigned char *p2, long *counts) {
    const int ntl = (snpts + TILE - 1) / TILE, ntl2 = (snpts + TILE2 - 1) / TILE2;
    TileRec1 *s1 = new TileRec1[(size_t)ns * ntl], *q1 = new TileRec1[(size_t)ns * ntl2];
    TileRecF *sf = new TileRecF[(size_t)ns * ntl], *qf = new TileRecF[(size_t)ns * ntl2];
    for (int s = 0; s < ns; s++) {
        const double *t = strk + (size_t)s * snpts * 12;
        TileRec2 r2;
        for (int j = 0; j < ntl; j++) tile_records(t, snpts, j, 0.0, s1[(size_t)s * ntl + j], r2, &sf[(size_t)s * ntl + j]);
        for (int J = 0; J < ntl2; J++)
            tile_records<TILE2>(t, snpts, J, 0.0, q1[(size_t)s * ntl2 + J], r2, &qf[(size_t)s * ntl2 + J]);
    }
    const double chi2c = chi2_max * (1.0 + CUT_REL);
    const float kchi = up(sqrt(chi2c)), chi2cf = up(chi2c);
    for (int i = 0; i < 4; i++) counts[i] = 0;
    for (int c = 0; c < nc; c++) {
        const int n = cn[c], nt = (n + TILE - 1) / TILE, nt2 = (n + TILE2 - 1) / TILE2;
        const double *t = ctrk + (size_t)c * cnpts * 12;
        for (int J = 0; J < ntl2; J++) {
            TileRec1 c1 = {}, d1;
            TileRec2 r2;
            TileRecF cq = {}, df[2] = {};
            for (int h = 0; h < 2; h++)
                if (2 * J + h < nt) tile_records(t, n, 2 * J + h, R2[c], d1, r2, &df[h]);
            if (J < nt2) tile_records<TILE2>(t, n, J, R2[c], c1, r2, &cq);
            for (int s = 0; s < ns; s++) {
                for (int h = 0; h < 2; h++)
                    if (2 * J + h < ntl) p2[((size_t)s * nc + c) * ntl + 2 * J + h] = 0;
                if (J >= nt2) continue;
                counts[0]++;
                if (!l1_pass<TILE2>(q1[(size_t)s * ntl2 + J], c1, tile_t1<TILE2>(n, J), kchi)) continue;
                counts[1]++;
                if (!l2f_pass<TILE2>(qf[(size_t)s * ntl2 + J], cq, chi2cf)) continue;
                for (int h = 0; h < 2; h++) {
                    const int j = 2 * J + h;
                    if (j >= nt) continue;
                    counts[2]++;
                    const bool f = l2f_pass(sf[(size_t)s * ntl + j], df[h], chi2cf);
                    p2[((size_t)s * nc + c) * ntl + j] = f;
                    counts[3] += f;
                }
            }
        }
    }
    delete[] s1; delete[] q1; delete[] sf; delete[] qf;
    return ntl;
}

// The records of tile j (T = 8) or coarse tile j (T = 16) of one object track, for the tests of their
// defining bounds: out = the 24 floats of TileRecF then lam (TileRec1).  Returns the tile's first point k0
// and writes its last point kp.
extern "C" int cc_recf(const double *trk, int n, int j, int coarse, double R2, float *out, int *kp_out) {
    TileRec1 r1;
    TileRec2 r2;
    TileRecF rf;
    if (coarse) tile_records<TILE2>(trk, n, j, R2, r1, r2, &rf);
    else tile_records<TILE>(trk, n, j, R2, r1, r2, &rf);
    const float *v = reinterpret_cast<const float *>(&rf);
    for (int u = 0; u < 24; u++) out[u] = v[u];
    out[24] = r1.lam;
    const int T = coarse ? TILE2 : TILE, k0 = j * T;
    *kp_out = (k0 + T < n ? k0 + T : n) - 1;
    return k0;
}
