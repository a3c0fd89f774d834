// Host build of candidate arithmetic for the linearised star tracks (reference: src/phygas.c:286-323: the nominal
// orbit and six neighbours 1e-5 apart, kick-drift-kick, the state-transition matrix by finite differences), so
// that a CPU test can state each candidate's error against the oracle's FP64 tracks:
//
//   mode 0  FP64 force and accumulators: the reference's arithmetic (checks the harness itself)
//   mode 1  all seven orbits with the FP32 force, each position and velocity accumulated as a float-float
//           (Kahan-compensated) pair; the matrix by differencing the pairs in FP64
//   mode 2  the nominal as in mode 1; the six neighbours as offsets driven by the FP32 difference force, FP64
//           offset accumulators
//   mode 3  the nominal with the FP32 force and plain FP32 accumulators; offsets as in mode 2
//   mode 4  the nominal in FP64 (the reference's force); offsets as in mode 2 (the fast form's arithmetic,
//           with the host's FP64 force in place of the device's)
//
// Output: [n][npts][12] as the tracks: position, velocity, the six entries of S = A Sigma0 A^T.
#include <math.h>

#include "force32.h"

using namespace phx;

namespace {

struct FF3 { float hx, lx, hy, ly, hz, lz; };

FF3 ff_from(V3 a) {
    FF3 f;
    f.hx = (float)a.x; f.lx = (float)(a.x - (double)f.hx);
    f.hy = (float)a.y; f.ly = (float)(a.y - (double)f.hy);
    f.hz = (float)a.z; f.lz = (float)(a.z - (double)f.hz);
    return f;
}
V3 ff_val(const FF3 &f) { return v3((double)f.hx + f.lx, (double)f.hy + f.ly, (double)f.hz + f.lz); }
// f += v * s (FP32 increment, compensated sum)
void ff_axpy(FF3 &f, F3 v, float s) { ff_add(f.hx, f.lx, v.x * s); ff_add(f.hy, f.ly, v.y * s); ff_add(f.hz, f.lz, v.z * s); }
F3 ff_hi(const FF3 &f) { return F3{f.hx, f.hy, f.hz}; }

// One orbit in FP32 with float-float position and velocity; the force at the high part of the position.
struct OrbitFF {
    FF3 x, v;
    F3 a;
    void start(const PotF &Q, V3 p, V3 w) { x = ff_from(p); v = ff_from(w); a = force32(Q, x.hx, x.hy, x.hz); }
    void step(const PotF &Q, float dt) {
        const float h = 0.5f * dt;
        ff_axpy(v, a, h);
        ff_axpy(x, ff_hi(v), dt);
        ff_add(x.hx, x.lx, v.lx * dt); ff_add(x.hy, x.ly, v.ly * dt); ff_add(x.hz, x.lz, v.lz * dt);
        a = force32(Q, x.hx, x.hy, x.hz);
        ff_axpy(v, a, h);
    }
};

// The nominal in plain FP32.
struct Orbit32 {
    F3 x, v, a;
    void start(const PotF &Q, V3 p, V3 w) {
        x = F3{(float)p.x, (float)p.y, (float)p.z};
        v = F3{(float)w.x, (float)w.y, (float)w.z};
        a = force32(Q, x.x, x.y, x.z);
    }
    void step(const PotF &Q, float dt) {
        const float h = 0.5f * dt;
        v.x += a.x * h; v.y += a.y * h; v.z += a.z * h;
        x.x += v.x * dt; x.y += v.y * dt; x.z += v.z * dt;
        a = force32(Q, x.x, x.y, x.z);
        v.x += a.x * h; v.y += a.y * h; v.z += a.z * h;
    }
};

void assemble(const double A[3][6], const double *sg, double *S) {
    double AS[3][6];
    for (int r = 0; r < 3; r++)
        for (int b = 0; b < 6; b++) {
            double acc = 0.0;
            for (int a = 0; a < 6; a++) acc += A[r][a] * sg[a * 6 + b];
            AS[r][b] = acc;
        }
    int o = 0;
    for (int r = 0; r < 3; r++)
        for (int c = r; c < 3; c++) {
            double acc = 0.0;
            for (int b = 0; b < 6; b++) acc += AS[r][b] * A[c][b];
            S[o++] = acc;
        }
}

void write_point(double *out, V3 x, V3 v, const double A[3][6], const double *sg) {
    out[0] = x.x; out[1] = x.y; out[2] = x.z; out[3] = v.x; out[4] = v.y; out[5] = v.z;
    assemble(A, sg, out + 6);
}

void track_fd(int mode, const Pot &P, const PotF &Q, const double *p0, const double *sg, int npts, double *out) {
    const double eps = PH_LIN_EPS_POS, dt = -STEP_MYR / PH_LIN_SUB;
    V3 x0 = v3(p0[0], p0[1], p0[2]), w0 = v3(p0[3], p0[4], p0[5]);
    Body b[7];
    OrbitFF f[7];
    for (int j = 0; j < 7; j++) {
        V3 x = x0, w = w0;
        if (j > 0) {
            const int o = j - 1;
            if (o < 3) (&x.x)[o] += eps; else (&w.x)[o - 3] += eps;
        }
        if (mode == 0) b[j] = body_start(P, x, w); else f[j].start(Q, x, w);
    }
    for (int k = 0; k < npts; k++) {
        if (k > 0)
            for (int it = 0; it < PH_LIN_SUB; it++)
                for (int j = 0; j < 7; j++) {
                    if (mode == 0) kdk(P, b[j], dt); else f[j].step(Q, (float)dt);
                }
        V3 xs[7], vs[7];
        for (int j = 0; j < 7; j++) {
            xs[j] = mode == 0 ? b[j].pos : ff_val(f[j].x);
            vs[j] = mode == 0 ? b[j].vel : ff_val(f[j].v);
        }
        double A[3][6];
        for (int o = 0; o < 6; o++) {
            const V3 d = xs[o + 1] - xs[0];
            A[0][o] = d.x / eps; A[1][o] = d.y / eps; A[2][o] = d.z / eps;
        }
        write_point(out + (size_t)k * 12, xs[0], vs[0], A, sg);
    }
}

// Modes 2-4: the nominal (FP32 float-float, plain FP32 or FP64) and six offsets driven by the difference force.
void track_diff(int mode, const Pot &P, const PotF &Q, const double *p0, const double *sg, int npts, double *out) {
    const double eps = PH_LIN_EPS_POS, dt = -STEP_MYR / PH_LIN_SUB, h = 0.5 * dt;
    const V3 x0 = v3(p0[0], p0[1], p0[2]), w0 = v3(p0[3], p0[4], p0[5]);
    OrbitFF nf;
    Orbit32 n32;
    Body n64;
    if (mode == 2) nf.start(Q, x0, w0);
    if (mode == 3) n32.start(Q, x0, w0);
    if (mode == 4) n64 = body_start(P, x0, w0);
    auto nom_hi = [&]() -> F3 {
        if (mode == 2) return ff_hi(nf.x);
        if (mode == 3) return n32.x;
        return F3{(float)n64.pos.x, (float)n64.pos.y, (float)n64.pos.z};
    };
    double dx[6][3] = {}, dv[6][3] = {};
    for (int o = 0; o < 6; o++) { if (o < 3) dx[o][o] = eps; else dv[o][o - 3] = eps; }
    F3 da[6];
    F3 p = nom_hi();
    for (int o = 0; o < 6; o++) da[o] = dforce32(Q, p.x, p.y, p.z, (float)dx[o][0], (float)dx[o][1], (float)dx[o][2]);
    for (int k = 0; k < npts; k++) {
        if (k > 0)
            for (int it = 0; it < PH_LIN_SUB; it++) {
                for (int o = 0; o < 6; o++) {
                    dv[o][0] += (double)da[o].x * h; dv[o][1] += (double)da[o].y * h; dv[o][2] += (double)da[o].z * h;
                    for (int c = 0; c < 3; c++) dx[o][c] += dv[o][c] * dt;
                }
                if (mode == 2) nf.step(Q, (float)dt);
                if (mode == 3) n32.step(Q, (float)dt);
                if (mode == 4) kdk(P, n64, dt);
                p = nom_hi();
                for (int o = 0; o < 6; o++)
                    da[o] = dforce32(Q, p.x, p.y, p.z, (float)dx[o][0], (float)dx[o][1], (float)dx[o][2]);
                for (int o = 0; o < 6; o++) {
                    dv[o][0] += (double)da[o].x * h; dv[o][1] += (double)da[o].y * h; dv[o][2] += (double)da[o].z * h;
                }
            }
        double A[3][6];
        for (int o = 0; o < 6; o++)
            for (int r = 0; r < 3; r++) A[r][o] = dx[o][r] / eps;
        V3 x, v;
        if (mode == 2) { x = ff_val(nf.x); v = ff_val(nf.v); }
        else if (mode == 3) { x = v3(n32.x.x, n32.x.y, n32.x.z); v = v3(n32.v.x, n32.v.y, n32.v.z); }
        else { x = n64.pos; v = n64.vel; }
        write_point(out + (size_t)k * 12, x, v, A, sg);
    }
}

}  // namespace

extern "C" int pr_tracks(int mode, const double *ps0, const double *sig, int n, int npts, double *out) {
    if (mode < 0 || mode > 4 || n < 0 || npts < 1) return -1;
    const Pot P = default_pot();
    const PotF Q = make_potf(P);
    for (int s = 0; s < n; s++) {
        double *o = out + (size_t)s * npts * 12;
        if (mode <= 1) track_fd(mode, P, Q, ps0 + (size_t)s * 6, sig + (size_t)s * 36, npts, o);
        else track_diff(mode, P, Q, ps0 + (size_t)s * 6, sig + (size_t)s * 36, npts, o);
    }
    return 0;
}
