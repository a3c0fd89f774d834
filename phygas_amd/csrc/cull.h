// Tile culling of the fast screening scan: per-object tile records and the pair tests.
//
// A tile j of an object holds the points k in [8j, 8j+8) of its track (the points whose chi2 the
// reference evaluates: reference src/phygas.c:395-403 / src/phygas_kernel_body.h:124-163), and the
// positions [8j, 8j+8] (the far end of the last point's segment).  Everything is expressed in a frame
// that co-rotates with the Galaxy: point k is rotated by R_k, the rotation by OMEGA_PT * k about the
// Galactic z axis, so that the circular part of the motion drops out and a tile's positions lie close to
// a straight line.
//
// Records of an object tile (built from its FP64 track; tiles of 8 points, and coarse tiles of 16 with the
// same construction):
//   a, b   least-squares line p'_k ~ a + b (k - tc), tc = k0 + T/2, through the rotated positions
//   e      largest distance of a rotated position from that line (the tube radius)
//   F, L   the rotated covariances of the tile's first and last points, each plus a slack E (the positive
//          part of the interpolation's largest error, along the ellipsoid's long axes), and eta, so that
//          R_k S_k R_k^T <= (1 - w_k) F + w_k L + eta I for every point (Loewner order), w_k = (k - k0) / np;
//          clusters add the pair's kernel R^2 I (TileRecF, stored in FP32 with its rounding in e and eta)
//   lam    upper bound of the largest eigenvalue over the tile: max(lam(F), lam(L)) + eta (TileRec1)
// (tile_records without TileRecF builds the older single bound U >= every R_k S_k R_k^T instead: the last
// point's covariance plus eta1 I, or last + (first - last)_+ plus eta2 I; the tests compare against it.)
//
// For a pair (star s, cluster c) and a point k of tile j, the reference's value is
// chi2_k = min(q(m0), q(ms)) with q(x) = x^T M_k^-1 x, M_k = S_s,k + S_c,k + R^2 I, m0 = p_s,k - p_c,k and
// ms a point of the segment m0 -> m1 = p_s,k+1 - p_c,k+1.  Rotated by R_k, m0 lies in the tube
// T = {da + db t + r : t in [-T/2, tl], |r| <= rho}, da = a_s - a_c, db = b_s - b_c, rho = e_s + e_c; m1
// rotated by R_k+1 lies in T too, so rotated by R_k it lies in R_1^-1 T (R_1 the rotation by OMEGA_PT),
// within ROT_W |m1| of T.  Every point of the segment is a convex combination of its ends.
//
//   level 1 (FP32): x^T M_k^-1 x >= |x|^2 / lam >= (dist(0, axis segment) - rho')^2 / (lam_s + lam_c),
//                   rho' = rho + ROT_W (|da| + 4 |db| + rho) covering the rotated segment ends (l1_pass)
//   level 2 (FP32, explicit rounding bounds): for every vector z, (z.x)^2 <= (x^T M^-1 x)(z^T M z); z.x is
//                   bounded below over the hull of T and R_1^-1 T by linear functions of the time
//                   parameter, and z^T M_k z above by the linear interpolation of the two objects' records,
//                   so the tile is dropped when a quadratic in the time parameter stays positive over the
//                   tile (l2f_pass); z = U^-1 x* with x* the U^-1-closest point of the axis segment.  One U
//                   for all points (l2_pass, FP64, kept as the tests' reference) inflates the narrow axes
//                   of an ellipsoid that turns or shears over the tile; the interpolation follows it.
//
// A tile that fails either test holds no point with chi2 < chi2_max (1 + CUT_REL), so it can hold neither
// the minimum of a kept pair nor change which pairs are kept.  CUT_REL covers the reference's own FP64
// rounding of chi2 for matrices with condition number below 1e9 (cond(M) <= trace(M) / R^2); level 2 does
// not prune pairs beyond that.
#pragma once

#include "force32.h"

namespace phx {

constexpr int TILE = 8;
constexpr int TILE2 = 16;   // coarse tile of the scan's first pass: two tiles
constexpr double OMEGA_PT = -0.5 * (220.0 / KMS_PER_KPCMYR) / 8.122;   // frame rotation per 0.5-Myr point
constexpr double ROT_W = 0.5 * (220.0 / KMS_PER_KPCMYR) / 8.122 * 1.000001;   // |R(w) x - x| <= ROT_W |x|
constexpr double CUT_REL = 1e-6;
// R_1, the rotation by OMEGA_PT about z: cos and sin (to double precision)
constexpr double ROT_C = 0.9999040755528862, ROT_S = -0.013850620662200556;

struct TileRec1 {   // FP32, level 1
    float ax, ay, az, e, bx, by, bz, lam;
};
struct TileRec2 {   // FP64, level 2
    double a[3], b[3], U[6], e, R2, t1, pad;
};
// Level 2 of the fine tiles, stored in FP32 (the test runs in FP64): the axis a, b and tube radius e, the
// rotated covariances of the first and last points with an anisotropic slack E added to both (F, L; the
// kernel term R^2 I included), and eta, the isotropic rest: every point k of the tile has its rotated
// covariance below (1 - w) F + w L + eta I, w = (k - k0) / np, np = kp - k0.  The FP32 rounding of a and b
// is part of e, and that of F and L part of eta (both rounded up), R2 is rounded down.  tp = kp - tc and
// tl = kl - tc (last point and last position; tc = k0 + T / 2) are exact; ispan = 1 / np (0 when np = 0).
struct TileRecF {
    float a[3], b[3], F[6], L[6], e, eta, tp, tl, ispan, R2;
};

// Level-1 record of a star slot past the end of a batch (the scan reads whole star blocks): far enough that
// level 1 rejects it against every cluster tile, with its squares finite in FP32.
PHX_HDF TileRec1 pad_rec1() { return TileRec1{1e12f, 1e12f, 1e12f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}; }

// The neighbouring float towards +inf (finite inputs).
PHX_HDF float next_up(float f) {
    if (f == 0.0f) return 1.4e-45f;
    union { float f; int i; } u = {f};
    u.i += f > 0.0f ? 1 : -1;
    return u.f;
}
PHX_HDF float up(double v) {   // float >= v
    const float f = (float)v;
    return (double)f < v ? next_up(f) : f;
}
PHX_HDF float down(double v) { return -up(-v); }   // float <= v

// Gershgorin bound of |lambda| of a symmetric 3x3 (xx xy xz yy yz zz).
PHX_HDF double sym3_abs_bound(const double *d) {
    const double g0 = fabs(d[0]) + fabs(d[1]) + fabs(d[2]), g1 = fabs(d[1]) + fabs(d[3]) + fabs(d[4]),
                 g2 = fabs(d[2]) + fabs(d[4]) + fabs(d[5]);
    return fmax(g0, fmax(g1, g2));
}


// Largest-eigenvalue upper bound of a symmetric PSD 3x3 (xx xy xz yy yz zz): min(trace, Gershgorin).
PHX_HDF double lam_bound(const double *S) {
    const double tr = S[0] + S[3] + S[5];
    const double g0 = S[0] + fabs(S[1]) + fabs(S[2]), g1 = S[3] + fabs(S[1]) + fabs(S[4]),
                 g2 = S[5] + fabs(S[2]) + fabs(S[4]);
    const double g = fmax(g0, fmax(g1, g2));
    return fmax(fmin(tr, g), 0.0);
}

// Upper bound of the largest eigenvalue of a symmetric 3x3 (xx xy xz yy yz zz), possibly indefinite: the
// closed form (trigonometric solution of the characteristic cubic) plus 1e-10 of the matrix's scale for its
// rounding, and never above the Gershgorin bound.
PHX_HDF double sym3_lmax_bound(const double *d) {
    const double g0 = d[0] + fabs(d[1]) + fabs(d[2]), g1 = d[3] + fabs(d[1]) + fabs(d[4]),
                 g2 = d[5] + fabs(d[2]) + fabs(d[4]);
    const double gersh = fmax(g0, fmax(g1, g2));
    const double p1 = d[1] * d[1] + d[2] * d[2] + d[4] * d[4];
    const double q = (d[0] + d[3] + d[5]) / 3.0;
    const double a = d[0] - q, b = d[3] - q, c = d[5] - q;
    const double p2 = a * a + b * b + c * c + 2.0 * p1;
    const double pp = sqrt(p2 / 6.0);
    double lmax = q;
    if (pp > 0.0) {
        // det((D - qI) / pp) / 2
        const double r = (a * (b * c - d[4] * d[4]) - d[1] * (d[1] * c - d[4] * d[2]) + d[2] * (d[1] * d[4] - b * d[2])) /
                         (2.0 * pp * pp * pp);
        const double phi = acos(fmin(fmax(r, -1.0), 1.0)) / 3.0;
        lmax = q + 2.0 * pp * cos(phi);
    }
    return fmin(gersh, lmax + 1e-10 * (fabs(q) + 2.0 * pp));
}

// R_k S R_k^T of a symmetric 3x3 (xx xy xz yy yz zz), R_k the rotation by OMEGA_PT k about z.
PHX_HDF void rot_cov(const double *S, double c, double s, double *o) {
    const double cc = c * c, ss = s * s, cs = c * s;
    o[0] = cc * S[0] - 2.0 * cs * S[1] + ss * S[3];
    o[1] = cs * (S[0] - S[3]) + (cc - ss) * S[1];
    o[2] = c * S[2] - s * S[4];
    o[3] = ss * S[0] + 2.0 * cs * S[1] + cc * S[3];
    o[4] = s * S[2] + c * S[4];
    o[5] = S[5];
}

// cos and sin of the frame rotation at point k (OMEGA_PT k): from the table cs ([k][2]) when given, else
// computed.  Only the consistency of one point's rotation of its position and its covariance matters to a
// record (chi2 is invariant under any invertible map applied to both).
PHX_HDF void rot_at(const double *cs, int k, double &c, double &s) {
    if (cs) {
        c = cs[2 * k];
        s = cs[2 * k + 1];
    } else {
        const double th = OMEGA_PT * k;
        c = cos(th);
        s = sin(th);
    }
}

// Positive part of a symmetric 3x3 (xx xy xz yy yz zz): V diag(max(lambda, 0)) V^T from cyclic Jacobi
// rotations.  Approximate (the callers bound what it misses).
PHX_HDF void sym3_pos_part(const double *D, double *P) {
    double a[3][3] = {{D[0], D[1], D[2]}, {D[1], D[3], D[4]}, {D[2], D[4], D[5]}};
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const double scale = fabs(D[0]) + fabs(D[3]) + fabs(D[5]) + fabs(D[1]) + fabs(D[2]) + fabs(D[4]);
    for (int sweep = 0; sweep < 6; sweep++) {
        for (int r = 0; r < 3; r++) {
            const int p = r == 2 ? 1 : 0, q = r == 0 ? 1 : 2;
            const double apq = a[p][q];
            if (!(fabs(apq) > 1e-18 * scale)) continue;
            const double tau = (a[q][q] - a[p][p]) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
            for (int k = 0; k < 3; k++) {   // A <- A J (columns p, q)
                const double akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - sn * akq;
                a[k][q] = sn * akp + c * akq;
            }
            for (int k = 0; k < 3; k++) {   // A <- J^T A (rows p, q)
                const double apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - sn * aqk;
                a[q][k] = sn * apk + c * aqk;
            }
            for (int k = 0; k < 3; k++) {   // V <- V J
                const double vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - sn * vkq;
                v[k][q] = sn * vkp + c * vkq;
            }
        }
    }
    const double l[3] = {fmax(a[0][0], 0.0), fmax(a[1][1], 0.0), fmax(a[2][2], 0.0)};
    P[0] = l[0] * v[0][0] * v[0][0] + l[1] * v[0][1] * v[0][1] + l[2] * v[0][2] * v[0][2];
    P[1] = l[0] * v[0][0] * v[1][0] + l[1] * v[0][1] * v[1][1] + l[2] * v[0][2] * v[1][2];
    P[2] = l[0] * v[0][0] * v[2][0] + l[1] * v[0][1] * v[2][1] + l[2] * v[0][2] * v[2][2];
    P[3] = l[0] * v[1][0] * v[1][0] + l[1] * v[1][1] * v[1][1] + l[2] * v[1][2] * v[1][2];
    P[4] = l[0] * v[1][0] * v[2][0] + l[1] * v[1][1] * v[2][1] + l[2] * v[1][2] * v[2][2];
    P[5] = l[0] * v[2][0] * v[2][0] + l[1] * v[2][1] * v[2][1] + l[2] * v[2][2] * v[2][2];
}

// tr(S^-1) of a symmetric 3x3 (xx xy xz yy yz zz); 0 when S is not positive definite.
PHX_HDF double inv_trace(const double *S) {
    const double c0 = S[3] * S[5] - S[4] * S[4], c3 = S[0] * S[5] - S[2] * S[2], c5 = S[0] * S[3] - S[1] * S[1];
    const double det = S[0] * c0 + S[1] * (S[2] * S[4] - S[1] * S[5]) + S[2] * (S[1] * S[4] - S[2] * S[3]);
    return det > 0.0 ? (c0 + c3 + c5) / det : 0.0;
}

// Axis of tile j (T points) of an object with n points: the line fit and tube radius fields of r1 / r2 and
// t1; returns the first and last point indices (k0, kp).  trk = its FP64 track [n][12] (pos, vel, S).
template <int T = TILE>
PHX_HDF void tile_line(const double *trk, int n, int j, TileRec1 &r1, TileRec2 &r2, int &k0_out, int &kp_out,
                       const double *cs = nullptr) {
    const int k0 = j * T;
    const int kl = k0 + T < n - 1 ? k0 + T : n - 1;   // last position
    const int kp = (k0 + T < n ? k0 + T : n) - 1;     // last point
    const int m = kl - k0 + 1;
    const double tc = k0 + 0.5 * T;
    // rotated position of point k, evaluated again where it is needed rather than kept in registers
    auto rpos = [&](int k, double &x, double &y, double &z) {
        const double *p = trk + (size_t)k * 12;
        double c, s;
        rot_at(cs, k, c, s);
        x = c * p[0] - s * p[1];
        y = s * p[0] + c * p[1];
        z = p[2];
    };
    double st = 0, stt = 0, sx = 0, sy = 0, sz = 0, stx = 0, sty = 0, stz = 0;
#pragma unroll
    for (int i = 0; i <= T; i++) {
        if (i < m) {
            double x, y, z;
            rpos(k0 + i, x, y, z);
            const double t = (k0 + i) - tc;
            st += t; stt += t * t;
            sx += x; sy += y; sz += z;
            stx += t * x; sty += t * y; stz += t * z;
        }
    }
    double b[3] = {0.0, 0.0, 0.0};
    const double den = m * stt - st * st;
    if (m > 1 && den > 0.0) {
        b[0] = (m * stx - st * sx) / den;
        b[1] = (m * sty - st * sy) / den;
        b[2] = (m * stz - st * sz) / den;
    }
    const double a[3] = {(sx - b[0] * st) / m, (sy - b[1] * st) / m, (sz - b[2] * st) / m};
    const float af[3] = {(float)a[0], (float)a[1], (float)a[2]}, bf[3] = {(float)b[0], (float)b[1], (float)b[2]};
    double e64 = 0.0, e32 = 0.0;
#pragma unroll
    for (int i = 0; i <= T; i++) {
        if (i < m) {
            double x, y, z;
            rpos(k0 + i, x, y, z);
            const double t = (k0 + i) - tc;
            const double rx = x - (a[0] + b[0] * t), ry = y - (a[1] + b[1] * t), rz = z - (a[2] + b[2] * t);
            e64 = fmax(e64, sqrt(rx * rx + ry * ry + rz * rz));
            const double fx = x - ((double)af[0] + (double)bf[0] * t), fy = y - ((double)af[1] + (double)bf[1] * t),
                         fz = z - ((double)af[2] + (double)bf[2] * t);
            e32 = fmax(e32, sqrt(fx * fx + fy * fy + fz * fz));
        }
    }
    r1.ax = af[0]; r1.ay = af[1]; r1.az = af[2];
    r1.bx = bf[0]; r1.by = bf[1]; r1.bz = bf[2];
    r1.e = up(e32 * (1.0 + 1e-6) + 1e-12);
#pragma unroll
    for (int u = 0; u < 3; u++) { r2.a[u] = a[u]; r2.b[u] = b[u]; }
    r2.e = e64 * (1.0 + 1e-12) + 1e-15;
    r2.t1 = (double)(kl - k0) - 0.5 * T;
    r2.pad = 0.0;
    k0_out = k0;
    kp_out = kp;
}

// The interpolated record of a tile from the rotated covariances q_k of its points k0 .. kp (q_at(k - k0, q)
// writes q_k) and the axis in r2.  D_k = q_k - ((1 - w_k) q_k0 + w_k q_kp) is the
// interpolation's error; over a tile the covariance follows a smooth curve, so the D_k are nearly multiples
// of one matrix whose positive part lies along the ellipsoid's long axes.  E = that positive part of the D_k
// with the largest eigenvalue bound goes into F and L, and eta bounds what is left (E is approximate; eta is
// an upper bound).  (first = q_k0, last = q_kp.)  q_k and D_k are evaluated again in each loop rather than
// kept, so a lane holds a few of them at a time.
template <int T = TILE, class QAt>
PHX_HDF void tile_record_f(QAt q_at, const double *first, const double *last, int k0, int kp, int n,
                           const TileRec2 &r2, double R2, TileRecF &rf, TileRec1 &r1) {
    const int kl = k0 + T < n - 1 ? k0 + T : n - 1;
    const double tc = k0 + 0.5 * T;
    const int np = kp - k0;
    double trmax = 0.0;   // >= every point's trace (taken in the first loop)
    auto d_at = [&](int i, double *D, bool tr) {   // D_{k0 + i}
        const double w = np > 0 ? (double)i / (double)np : 0.0;
        double q[6];
        q_at(i, q);
        if (tr) trmax = fmax(trmax, q[0] + q[3] + q[5]);
#pragma unroll
        for (int u = 0; u < 6; u++) D[u] = q[u] - ((1.0 - w) * first[u] + w * last[u]);
    };
    double Db[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, best = -1.0;
#pragma unroll
    for (int i = 0; i < T; i++) {
        if (i <= np) {
            double D[6];
            d_at(i, D, true);
            const double lm = sym3_lmax_bound(D);
            if (lm > best) {
                best = lm;
#pragma unroll
                for (int u = 0; u < 6; u++) Db[u] = D[u];
            }
        }
    }
    double E[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (best > 0.0) sym3_pos_part(Db, E);
    double eta = 0.0;
#pragma unroll
    for (int i = 0; i < T; i++) {
        if (i <= np) {
            double d[6];
            d_at(i, d, false);
#pragma unroll
            for (int u = 0; u < 6; u++) d[u] -= E[u];
            eta = fmax(eta, sym3_lmax_bound(d));
        }
    }
    // FP32 storage: F - fl(F) and L - fl(L) are exact in FP64 and below their Gershgorin bounds times I
    double F[6], L[6], dF[6], dL[6];
#pragma unroll
    for (int u = 0; u < 6; u++) {
        const double r = (u == 0 || u == 3 || u == 5) ? R2 : 0.0;
        F[u] = first[u] + E[u] + r;
        L[u] = last[u] + E[u] + r;
        rf.F[u] = (float)F[u];
        rf.L[u] = (float)L[u];
        dF[u] = F[u] - (double)rf.F[u];
        dL[u] = L[u] - (double)rf.L[u];
    }
    // rounding of the rotations, the interpolation and the differences
    eta = eta * (1.0 + 1e-12) + 1e-13 * trmax;
    // level 1: every M_k is below (1 - w) F + w L + eta I, whose largest eigenvalue is at most the larger of
    // F's and L's plus eta
    r1.lam = up((fmax(lam_bound(F), lam_bound(L)) + eta) * (1.0 + 1e-6));
    // and the storage
    rf.eta = up((eta + fmax(sym3_abs_bound(dF), sym3_abs_bound(dL))) * (1.0 + 1e-12));
    // the stored axis is within |a - fl(a)| + (T / 2) |b - fl(b)| of the FP64 one over t in [-T/2, T/2]
    double ea = 0.0, eb = 0.0;
#pragma unroll
    for (int u = 0; u < 3; u++) {
        rf.a[u] = (float)r2.a[u];
        rf.b[u] = (float)r2.b[u];
        const double xa = r2.a[u] - (double)rf.a[u], xb = r2.b[u] - (double)rf.b[u];
        ea += xa * xa;
        eb += xb * xb;
    }
    rf.e = up((r2.e + sqrt(ea) + 0.5 * T * sqrt(eb)) * (1.0 + 1e-12));
    rf.tp = (float)(kp - tc);
    rf.tl = (float)(kl - tc);
    rf.ispan = np > 0 ? (float)(1.0 / (double)np) : 0.0f;
    rf.R2 = down(R2);
}

// Records of tile j of an object with n points; trk = its FP64 track [n][12] (pos, vel, S).  With rf: the
// interpolated level-2 record and level 1 from it (r2.U not set); without: r2.U, the one Loewner bound of
// the tile, and level 1 from that.  cs: table of the frame rotation's cos and sin per point (or null).
template <int T = TILE>
PHX_HDF void tile_records(const double *trk, int n, int j, double R2, TileRec1 &r1, TileRec2 &r2,
                          TileRecF *rf = nullptr, const double *cs = nullptr) {
    int k0, kp;
    tile_line<T>(trk, n, j, r1, r2, k0, kp, cs);
    r2.R2 = R2;
    // rotated covariance of point k0 + i
    auto q_at = [&](int i, double *q) {
        double c, s;
        rot_at(cs, k0 + i, c, s);
        rot_cov(trk + (size_t)(k0 + i) * 12 + 6, c, s, q);
    };
    double first[6], ref[6], trmax = 0.0;
    if (rf) {
        q_at(0, first);
        q_at(kp - k0, ref);
        tile_record_f<T>(q_at, first, ref, k0, kp, n, r2, R2, *rf, r1);
        return;
    }
    double qs[T][6];   // rotated covariances of the tile's points
#pragma unroll
    for (int i = 0; i < T; i++) {
        if (k0 + i <= kp) {
            q_at(i, qs[i]);
            trmax = fmax(trmax, qs[i][0] + qs[i][3] + qs[i][5]);
        }
    }
    q_at(kp - k0, ref);
#pragma unroll
    for (int u = 0; u < 6; u++) first[u] = qs[0][u];
    // One Loewner bound of the rotated covariances of the tile's points: ref + eta1 I (ref = the last point's)
    // or, for tiles of more than one point, sum + eta2 I (sum = ref + the positive part of the first point's
    // minus ref), the one with the larger tr(U^-1); level 1 the smaller eigenvalue bound.
    double sum[6];
    {
        double d[6], pp[6];
#pragma unroll
        for (int u = 0; u < 6; u++) d[u] = first[u] - ref[u];
        sym3_pos_part(d, pp);
#pragma unroll
        for (int u = 0; u < 6; u++) sum[u] = ref[u] + pp[u];
    }
    double eta1 = 0.0, eta2 = 0.0;
    trmax = fmax(trmax, sum[0] + sum[3] + sum[5]);
#pragma unroll
    for (int i = 0; i < T; i++) {
        if (k0 + i <= kp) {
            double d1[6], d2[6];
#pragma unroll
            for (int u = 0; u < 6; u++) {
                d1[u] = qs[i][u] - ref[u];
                d2[u] = qs[i][u] - sum[u];
            }
            eta1 = fmax(eta1, sym3_lmax_bound(d1));
            eta2 = fmax(eta2, sym3_lmax_bound(d2));
        }
    }
    eta1 = eta1 * (1.0 + 1e-12) + 1e-13 * trmax;   // rounding of the rotations, sums and differences
    eta2 = eta2 * (1.0 + 1e-12) + 1e-13 * trmax;
    const double U1[6] = {ref[0] + eta1 + R2, ref[1], ref[2], ref[3] + eta1 + R2, ref[4], ref[5] + eta1 + R2};
    const double U2[6] = {sum[0] + eta2 + R2, sum[1], sum[2], sum[3] + eta2 + R2, sum[4], sum[5] + eta2 + R2};
    const double *U = kp > k0 && inv_trace(U2) > inv_trace(U1) ? U2 : U1;
    r1.lam = up(fmin(lam_bound(U1), lam_bound(U2)) * (1.0 + 1e-6));
#pragma unroll
    for (int u = 0; u < 6; u++) r2.U[u] = U[u];
}

// The axis parameter range of tile j of a cluster with n points: t in [-4, t1].
template <int T = TILE>
PHX_HDF float tile_t1(int n, int j) {
    const int k0 = j * T, kl = k0 + T < n - 1 ? k0 + T : n - 1;
    return (float)(kl - k0) - 0.5f * T;
}

// Square root to within 1 ulp (the hardware instruction on the device).
PHX_HDF float f_sqrt_fast(float x) { return f_sqrt_hw(x); }

// Level 1, FP32.  kchi = sqrt(chi2_max (1 + CUT_REL)) rounded up.  The test accepts when
// dist(0, axis segment) <= rho + kchi sqrt(lam_s + lam_c), compared as squares with relative margin 4e-5
// and an absolute one of 1e-5 kpc (plus 2e-6 |da|^2 for the rounding of the differences); FP32 rounding of
// every step is far below both.  The distance is |da + db t~|^2 at t~, the closest parameter -da.db / |db|^2
// through the hardware reciprocal, clamped to the segment: |t~ - t*| <= delta = 1e-6 (|t~| + 1), and the
// excess of the squared distance at t~ over the minimum is at most 3 |db|^2 delta^2 (the derivative at a
// clamped end is at most 2 |db|^2 delta when t* lies within delta of it), which is taken off.
template <int T = TILE>
PHX_HDF bool l1_pass(const TileRec1 &s, const TileRec1 &c, float t1, float kchi) {
    const float dax = s.ax - c.ax, day = s.ay - c.ay, daz = s.az - c.az;
    const float dbx = s.bx - c.bx, dby = s.by - c.by, dbz = s.bz - c.bz;
    const float t0 = -0.5f * T;
    const float dd = f_fma(dbx, dbx, f_fma(dby, dby, dbz * dbz));
    const float ad = f_fma(dax, dbx, f_fma(day, dby, daz * dbz));
    const float tt = fminf(fmaxf(-ad * f_rcp(dd), t0), t1);   // dd = 0: NaN or inf, clamped to an end
    const float xx = f_fma(dbx, tt, dax), xy = f_fma(dby, tt, day), xz = f_fma(dbz, tt, daz);
    const float dl = fabsf(tt) * 1e-6f + 1e-6f;
    const float d2 = f_fma(xx, xx, f_fma(xy, xy, xz * xz)) - 3.0f * dd * dl * dl;
    const float e = s.e + c.e;
    const float na = fabsf(dax) + fabsf(day) + fabsf(daz), nb = fabsf(dbx) + fabsf(dby) + fabsf(dbz);
    const float rho = e + (float)ROT_W * (na + 0.5f * T * nb + e);
    const float lim = (rho + kchi * f_sqrt_fast(s.lam + c.lam)) * 1.00001f + 1e-5f;
    const float lim2 = f_fma(lim, lim, 2e-6f * na * na);
    return d2 * 0.99996f <= lim2;
}

// Level 2, FP64.  chi2c = chi2_max (1 + CUT_REL).  For y = s z (any z, s > 0) the dual bound is
// 2 s L - s^2 z^T U z with L = min over the hull of z.x = min(lin(z), lin(R_1 z)) - rho |z|, largest at
// s = L / z^T U z: every point has q >= L^2 / z^T U z when L > 0.  z = adj(U) x* (x* the U^-1-closest point
// of the axis segment) up to a positive factor, which needs no division; any z is valid, so only the
// rounding of L and z^T U z is bounded (err_l, err_u).  U is positive definite: its diagonal holds the
// kernel term R^2 > 0 and the trace guard bounds its condition number.
template <int T = TILE>
PHX_HDF bool l2_pass(const TileRec2 &s, const TileRec2 &c, double t1, double chi2c) {
    const double da0 = s.a[0] - c.a[0], da1 = s.a[1] - c.a[1], da2 = s.a[2] - c.a[2];
    const double db0 = s.b[0] - c.b[0], db1 = s.b[1] - c.b[1], db2 = s.b[2] - c.b[2];
    double U[6];
#pragma unroll
    for (int u = 0; u < 6; u++) U[u] = s.U[u] + c.U[u];
    const double tr = U[0] + U[3] + U[5];
    if (!(tr < 1e9 * (s.R2 + c.R2))) return true;   // condition number not bounded: never pruned here
    const double dlt = 1e-13 * tr;                    // U rounded up over the rounding of the sum
    U[0] += dlt; U[3] += dlt; U[5] += dlt;
    const double rho = (s.e + c.e) * (1.0 + 1e-12);
    const double G0 = U[3] * U[5] - U[4] * U[4], G1 = U[2] * U[4] - U[1] * U[5], G2 = U[1] * U[4] - U[2] * U[3];
    const double G3 = U[0] * U[5] - U[2] * U[2], G4 = U[1] * U[2] - U[0] * U[4], G5 = U[0] * U[3] - U[1] * U[1];
    const double ga0 = G0 * da0 + G1 * da1 + G2 * da2, ga1 = G1 * da0 + G3 * da1 + G4 * da2,
                 ga2 = G2 * da0 + G4 * da1 + G5 * da2;
    const double gb0 = G0 * db0 + G1 * db1 + G2 * db2, gb1 = G1 * db0 + G3 * db1 + G4 * db2,
                 gb2 = G2 * db0 + G4 * db1 + G5 * db2;
    const double B = da0 * gb0 + da1 * gb1 + da2 * gb2, C = db0 * gb0 + db1 * gb1 + db2 * gb2;
    const double t0 = -0.5 * T;
    // z = C (ga + th gb) with th = -B / C inside [t0, t1], else ga + t gb at the nearer end
    double wa = C, wb = -B;
    if (!(C > 0.0) || -B <= t0 * C) { wa = 1.0; wb = t0; }
    else if (-B >= t1 * C) { wa = 1.0; wb = t1; }
    const double z0 = wa * ga0 + wb * gb0, z1 = wa * ga1 + wb * gb1, z2 = wa * ga2 + wb * gb2;
    const double za = z0 * da0 + z1 * da1 + z2 * da2, zb = z0 * db0 + z1 * db1 + z2 * db2;
    // the segment ends m1 lie in R_1^-1 T (R_1: the rotation by OMEGA_PT), where z.x takes the values of
    // (R_1 z).x' over T
    const double zr0 = ROT_C * z0 - ROT_S * z1, zr1 = ROT_S * z0 + ROT_C * z1;
    const double zra = zr0 * da0 + zr1 * da1 + z2 * da2, zrb = zr0 * db0 + zr1 * db1 + z2 * db2;
    const double lin = fmin(fmin(za + t0 * zb, za + t1 * zb), fmin(zra + t0 * zrb, zra + t1 * zrb));
    const double nz2 = z0 * z0 + z1 * z1 + z2 * z2;
    const double nz = sqrt(nz2) * (1.0 + 1e-12);
    const double u0 = U[0] * z0 + U[1] * z1 + U[2] * z2, u1 = U[1] * z0 + U[3] * z1 + U[4] * z2,
                 u2 = U[2] * z0 + U[4] * z1 + U[5] * z2;
    const double zuz = z0 * u0 + z1 * u1 + z2 * u2;
    const double na1 = fabs(da0) + fabs(da1) + fabs(da2), nb1 = fabs(db0) + fabs(db1) + fabs(db2);
    const double err_l = 1e-12 * nz * (na1 + 0.5 * T * nb1 + rho);
    const double err_u = 1e-12 * tr * nz2;
    const double L = lin - rho * nz - err_l;
    return !(L > 0.0 && L * L > chi2c * (zuz + err_u) * (1.0 + 1e-12));
}

// Level 2 is evaluated without early exits: every condition is computed and the results combined, so the
// lanes of a wavefront (different pairs) stay on one instruction stream.
//
// FP32 check that g(t) = l(t)^2 - c V(t) > 0 and l(t) > 0 over [ta, tb] for the true l(t) >= l0 + l1 t - el
// and V(t) <= v0 + v1 t + ev, el = 1e-6 (|l0| + |l1| tm) and ev = 1e-6 (|v0| + |v1| tm) (tm >= |t|) covering
// the rounding of l0, l1, v0, v1 and of their evaluation: l is linear, so its lower bound positive at both
// ends suffices; g is a quadratic, so its ends and, when it lies inside, its vertex (4 A C > B^2, division
// free), each against a margin well above the FP32 rounding of the comparison.
PHX_HDF bool quad_clear(float l0, float l1, float v0, float v1, float c, float ta, float tb, float tm) {
    const float el = 1e-6f * f_fma(fabsf(l1), tm, fabsf(l0)), ev = 1e-6f * f_fma(fabsf(v1), tm, fabsf(v0));
    const float la = f_fma(l1, ta, l0) - el, lb = f_fma(l1, tb, l0) - el;
    const float va = c * (f_fma(v1, ta, v0) + ev), vb = c * (f_fma(v1, tb, v0) + ev);
    bool ok = (la > 0.0f) & (lb > 0.0f);
    ok &= f_fma(la, la, -va) > 1e-5f * f_fma(la, la, va);
    ok &= f_fma(lb, lb, -vb) > 1e-5f * f_fma(lb, lb, vb);
    const float A = l1 * l1, B = f_fma(2.0f * l0, l1, -c * v1), C = f_fma(l0, l0, -c * v0);
    const bool inside = (A > 0.0f) & (-B > 2.0f * A * ta) & (-B < 2.0f * A * tb);
    const float lm = f_fma(fabsf(l1), tm, fabsf(l0)), vm = f_fma(fabsf(v1), tm, fabsf(v0));
    const float m = 1e-4f * f_fma(4.0f * A, f_fma(lm, lm, c * vm), B * B);
    const bool vok = f_fma(4.0f * A, C, -B * B) > m;
    return ok & (!inside | vok);
}

// Fine level 2 with the covariances interpolated in time (TileRecF), in FP32 with explicit error bounds.
// Segment k of the pair (k0 <= k <= kp, kp the cluster's, which ends no later than the star's) has its near
// end at time parameter tau_k = k - tc and its far end (the next position, seen from frame k through R_1) at
// tau_k + 1; every point x of it has z.x >= min(l_a(tau_k), l_b(tau_k + 1)), l_a(t) = z.(da + db t) - rho |z|
// and l_b the same with R_1 z, and its covariance M_k has z^T M_k z <= V(tau_k), V(t) = v0 + v1 t the sum of
// the two objects' interpolations (each linear in t between its own first and last points) plus eta |z|^2.
// With z.x > 0, q >= (z.x)^2 / z^T M_k z, so the tile is dropped when l_a(t)^2 > chi2c V(t) over [t0, tp]
// and l_b(t)^2 > chi2c V(t - 1) over [t0 + 1, tl].  z = U^-1 x* as in l2_pass with U the mean of the
// interpolation's ends, normalised (|z| = 1 to ~1e-7), so that every term is of the size of the geometry.
// Rounding: the differences da, db, the dot products and the quadratic forms of the PSD F, L are within
// ~1e-6 relative of their absolute-value bounds, far below the 1e-5 taken for them.
template <int T = TILE>
PHX_HDF bool l2f_pass(const TileRecF &s, const TileRecF &c, float chi2c) {
    const float da0 = s.a[0] - c.a[0], da1 = s.a[1] - c.a[1], da2 = s.a[2] - c.a[2];
    const float db0 = s.b[0] - c.b[0], db1 = s.b[1] - c.b[1], db2 = s.b[2] - c.b[2];
    const float eta = s.eta + c.eta;
    float U[6];
#pragma unroll
    for (int u = 0; u < 6; u++) U[u] = 0.5f * ((s.F[u] + s.L[u]) + (c.F[u] + c.L[u]));
    const float trF = (s.F[0] + s.F[3] + s.F[5]) + (c.F[0] + c.F[3] + c.F[5]);
    const float trL = (s.L[0] + s.L[3] + s.L[5]) + (c.L[0] + c.L[3] + c.L[5]);
    // every M_k has trace <= max(trF, trL) + 3 eta and smallest eigenvalue >= R^2: condition number bounded
    // (the bound taken 1% below 1e9 for the rounding)
    const bool guard = !(f_fma(3.0f, eta, fmaxf(trF, trL)) < 0.99e9f * (s.R2 + c.R2));
    U[0] += eta; U[3] += eta; U[5] += eta;
    const float t0 = -0.5f * T, tp = fminf(s.tp, c.tp), tl = fminf(s.tl, c.tl), tm = 0.5f * T;
    auto dot = [](float a0, float a1, float a2, float b0, float b1, float b2) {
        return f_fma(a0, b0, f_fma(a1, b1, a2 * b2));
    };
    // z = U^-1 x* through the Cholesky factor U = L L^T (backward stable in FP32, where an adjugate of an
    // elongated U is not): w = L^-1 (da + db t), t* minimising |w|^2 over [t0, tl], z = L^-T w(t*)
    const float i0 = f_rsqrt(U[0]), l10 = U[1] * i0, l20 = U[2] * i0;
    const float i1 = f_rsqrt(f_fma(-l10, l10, U[3])), l21 = f_fma(-l20, l10, U[4]) * i1;
    const float i2 = f_rsqrt(f_fma(-l21, l21, f_fma(-l20, l20, U[5])));
    const float wa0 = da0 * i0, wa1 = f_fma(-l10, wa0, da1) * i1, wa2 = f_fma(-l21, wa1, f_fma(-l20, wa0, da2)) * i2;
    const float wb0 = db0 * i0, wb1 = f_fma(-l10, wb0, db1) * i1, wb2 = f_fma(-l21, wb1, f_fma(-l20, wb0, db2)) * i2;
    const float B = dot(wa0, wa1, wa2, wb0, wb1, wb2), C = dot(wb0, wb1, wb2, wb0, wb1, wb2);
    float ts = C > 0.0f ? -B * f_rcp(C) : t0;
    ts = fminf(fmaxf(ts, t0), tl);
    const float x0 = f_fma(wb0, ts, wa0), x1 = f_fma(wb1, ts, wa1), x2 = f_fma(wb2, ts, wa2);
    float z2 = x2 * i2, z1 = f_fma(-l21, z2, x1) * i1, z0 = f_fma(-l20, z2, f_fma(-l10, z1, x0)) * i0;
    const float n2 = dot(z0, z1, z2, z0, z1, z2);
    const bool nbad = !((n2 > 1e-30f) & (n2 < 1e30f));
    const float r = f_rsqrt(n2);
    z0 *= r; z1 *= r; z2 *= r;   // |z| <= 1 + 1e-6
    const float zr0 = f_fma((float)ROT_C, z0, -(float)ROT_S * z1), zr1 = f_fma((float)ROT_S, z0, (float)ROT_C * z1);
    const float p00 = z0 * z0, p11 = z1 * z1, p22 = z2 * z2, p01 = 2.0f * z0 * z1, p02 = 2.0f * z0 * z2,
                p12 = 2.0f * z1 * z2;
    auto qf = [&](const float *M) {
        return f_fma(p00, M[0], f_fma(p11, M[3], f_fma(p22, M[5], f_fma(p01, M[1], f_fma(p02, M[2], p12 * M[4])))));
    };
    const float VFs = qf(s.F), VLs = qf(s.L), VFc = qf(c.F), VLc = qf(c.L);
    const float v1 = f_fma(VLs - VFs, s.ispan, (VLc - VFc) * c.ispan);
    // |z|^2 <= 1 + 3e-6; for a PSD M, |M_ij| <= sqrt(M_ii M_jj), so sum |p_ij M_ij| <= tr(M) |z|^2 and each
    // form is within 7 u tr(M) of its value; ispan is within u of 1 / np
    const float v0 = f_fma(-v1, t0, VFs + VFc) + f_fma(eta, 1.00001f, 2e-6f * (trF + trL + eta));
    const float rho = (s.e + c.e) * 1.00001f;
    const float na1 = fabsf(da0) + fabsf(da1) + fabsf(da2), nb1 = fabsf(db0) + fabsf(db1) + fabsf(db2);
    const float el = 1e-5f * (f_fma(tm, nb1, na1) + rho);   // the differences' and dot products' rounding
    const float la0 = dot(z0, z1, z2, da0, da1, da2) - (rho + el), la1 = dot(z0, z1, z2, db0, db1, db2);
    const float lb0 = dot(zr0, zr1, z2, da0, da1, da2) - (rho + el), lb1 = dot(zr0, zr1, z2, db0, db1, db2);
    const bool c1 = quad_clear(la0, la1, v0, v1, chi2c, t0, tp, tm);
    const bool c2 = (tl < t0 + 1.0f) | quad_clear(lb0, lb1, v0 - v1, v1, chi2c, t0 + 1.0f, tl, tm);   // no far end: kl = k0
    return guard | nbad | !(c1 & c2);
}

}  // namespace phx
