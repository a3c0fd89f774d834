// Plain FP64 path: the reference's GPU algorithm with its arithmetic kept operation by operation
// (reference: src/phygas_kernel_body.h:17-274 for what each kernel computes), laid out for CDNA4:
//   * orbit bundles (nominal + 6 neighbours) run 8 lanes per object, one orbit per lane, so a 20,000-star
//     batch fills 2,500 wavefronts instead of 313;
//   * pair scans and clones run one lane per pair / clone, reading the cluster tables as
//     structure-of-arrays so a wavefront's 64 neighbouring clusters are one coalesced read;
//   * every launch is asynchronous on the caller's stream (the pg_* ABI below adds the reference's
//     synchronous contract on top).
#include <stdlib.h>
#include <string.h>

#include <string>

#include "common.h"
#include "phygas_hip.h"

namespace phx {

static thread_local std::string g_err;

void set_error(const char *where, const char *what) {
    g_err = std::string(where) + ": " + what;
    fprintf(stderr, "phygas_hip: %s\n", g_err.c_str());
}
const char *last_error() { return g_err.c_str(); }

// ---------------------------------------------------------------- draws -> (ps0, Sigma0)
// One lane per object: n_samp seeded draws accumulated into the 6x6 covariance about the nominal, divided by
// the number of successful draws (reference: src/phygas.c:256-264, 274-284).  The nominal start state
// nom[n][6] comes from the host (nominal_states), drawn with the C library as the reference draws it: the
// finite-difference tracks amplify a one-ulp difference in a start state a hundred-thousand-fold.
// The 21 unique sums are accumulated in draw order, which is the reference's order for every entry.
template <class Rec, int IsCluster>
__global__ __launch_bounds__(128) void k_lin_init(const Rec *__restrict__ recs, const uint64_t *__restrict__ seeds,
                                                  const double *__restrict__ nom, int n, int n_samp, Pot P,
                                                  double *__restrict__ ps0, double *__restrict__ sig) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Rec r = recs[i];
    const double *q0 = nom + (size_t)i * 6;
    const V3 np = v3(q0[0], q0[1], q0[2]), nv = v3(q0[3], q0[4], q0[5]);
    double acc[21];
#pragma unroll
    for (int q = 0; q < 21; q++) acc[q] = 0.0;
    uint64_t st = seeds[i];
    int k = 0;
    for (int d = 0; d < n_samp; d++) {
        V3 p, v;
        int rc = IsCluster ? sample_cluster(P, (const PhCluster &)r, &st, p, v)
                           : sample_target(P, (const PhTarget &)r, &st, p, v);
        if (rc != 0) continue;
        double e[6] = {p.x - np.x, p.y - np.y, p.z - np.z, v.x - nv.x, v.y - nv.y, v.z - nv.z};
        int q = 0;
#pragma unroll
        for (int a = 0; a < 6; a++)
#pragma unroll
            for (int b = a; b < 6; b++) acc[q++] += e[a] * e[b];
        k++;
    }
    double div = k > 0 ? k : 1;
    double *o = sig + (size_t)i * 36;
    int q = 0;
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
        for (int b = a; b < 6; b++, q++) {
            double v = acc[q] / div;
            o[a * 6 + b] = v;
            o[b * 6 + a] = v;
        }
    double *p0 = ps0 + (size_t)i * 6;
    p0[0] = np.x; p0[1] = np.y; p0[2] = np.z; p0[3] = nv.x; p0[4] = nv.y; p0[5] = nv.z;
}

template <class Rec, int IsCluster>
static void nominal_states_of(const Pot &P, const Rec *r, int n, double *out) {
    for (int i = 0; i < n; i++) {
        V3 p = v3(0, 0, 0), v = v3(0, 0, 0);
        if (IsCluster) sample_cluster(P, (const PhCluster &)r[i], nullptr, p, v);
        else sample_target(P, (const PhTarget &)r[i], nullptr, p, v);
        double *o = out + (size_t)i * 6;
        o[0] = p.x; o[1] = p.y; o[2] = p.z; o[3] = v.x; o[4] = v.y; o[5] = v.z;
    }
}
void nominal_states(const Pot &P, const PhTarget *t, int n, double *out) { nominal_states_of<PhTarget, 0>(P, t, n, out); }
void nominal_states(const Pot &P, const PhCluster *c, int n, double *out) { nominal_states_of<PhCluster, 1>(P, c, n, out); }

int launch_lin_init_targets(const PhTarget *d_t, const uint64_t *d_seed, const double *d_nom, int n, int n_samp,
                            const Pot &P, double *d_ps0, double *d_sig, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL((k_lin_init<PhTarget, 0>), dim3(blocks_for(n, 128)), dim3(128), 0, st, d_t, d_seed, d_nom, n,
                       n_samp, P, d_ps0, d_sig);
    PHX_CK(hipGetLastError());
    return 0;
}

int launch_lin_init_clusters(const PhCluster *d_c, const uint64_t *d_seed, const double *d_nom, int n, int n_samp,
                             const Pot &P, double *d_ps0, double *d_sig, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL((k_lin_init<PhCluster, 1>), dim3(blocks_for(n, 128)), dim3(128), 0, st, d_c, d_seed, d_nom, n,
                       n_samp, P, d_ps0, d_sig);
    PHX_CK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- linear tracks
// Lane j of each 8-lane group integrates orbit j of one object (0 = nominal, 1..6 = +eps along x, y, z,
// vx, vy, vz; lane 7 idles).  At every 0.5-Myr point lane 0 gathers the neighbours' positions with
// cross-lane shuffles and assembles A (3x6 finite-difference STM rows), A*Sigma0 and S = A Sigma0 A^T in
// the reference's summation order (reference: src/phygas.c:286-323, device src/phygas_kernel_body.h:64-109).
__global__ __launch_bounds__(256) void k_lin_plain(const double *__restrict__ ps0, const double *__restrict__ sig,
                                                   double *__restrict__ state, double *__restrict__ trk, Pot P,
                                                   double eps_pos, double eps_vel, double dt, int sub, int k0, int nk,
                                                   int npts, int n, int soa_nc, const int *__restrict__ nobj) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = gid >> 3, j = gid & 7;
    const bool live = s < n && j < 7;
    Body b;
    b.pos = b.vel = b.acc = v3(0, 0, 0);
    if (live) {
        if (k0 == 0) {
            const double *p = ps0 + (size_t)s * 6;
            double c[6] = {p[0], p[1], p[2], p[3], p[4], p[5]};
            if (j > 0) c[j - 1] += (j - 1) < 3 ? eps_pos : eps_vel;
            b = body_start(P, v3(c[0], c[1], c[2]), v3(c[3], c[4], c[5]));
        } else {
            const double *q = state + (size_t)s * 63 + j * 9;
            b.pos = v3(q[0], q[1], q[2]);
            b.vel = v3(q[3], q[4], q[5]);
            b.acc = v3(q[6], q[7], q[8]);
        }
    }
    // every lane of a group has the same object, so the group's lanes run the same number of points
    const int kend = min(k0 + nk, (nobj && s < n) ? min(nobj[s], npts) : npts);
    for (int k = k0; k < kend; k++) {
        if (k > 0 && live)
            for (int i = 0; i < sub; i++) kdk(P, b, dt);
        // every lane takes part in the shuffles; lane 0 of the group receives orbit jj's position
        double ox[7], oy[7], oz[7];
#pragma unroll
        for (int jj = 0; jj < 7; jj++) {
            ox[jj] = __shfl(b.pos.x, jj, 8);
            oy[jj] = __shfl(b.pos.y, jj, 8);
            oz[jj] = __shfl(b.pos.z, jj, 8);
        }
        if (j != 0 || s >= n) continue;
        double A[3][6];
#pragma unroll
        for (int c = 0; c < 6; c++) {
            double e = c < 3 ? eps_pos : eps_vel;
            A[0][c] = (ox[c + 1] - ox[0]) / e;
            A[1][c] = (oy[c + 1] - oy[0]) / e;
            A[2][c] = (oz[c + 1] - oz[0]) / e;
        }
        const double *sg = sig + (size_t)s * 36;
        double AS[3][6];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 6; c++) {
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < 6; a++) acc += A[r][a] * sg[a * 6 + c];
                AS[r][c] = acc;
            }
        double S[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
                double acc = 0.0;
#pragma unroll
                for (int c = 0; c < 6; c++) acc += AS[r][c] * A[q][c];
                S[r][q] = acc;
            }
        const double v[12] = {b.pos.x, b.pos.y, b.pos.z, b.vel.x, b.vel.y, b.vel.z,
                              S[0][0], S[0][1], S[0][2], S[1][1], S[1][2], S[2][2]};
        if (soa_nc > 0) {
#pragma unroll
            for (int c = 0; c < 12; c++) trk[((size_t)k * 12 + c) * soa_nc + s] = v[c];
        } else {
            double *o = trk + ((size_t)s * npts + k) * 12;
#pragma unroll
            for (int c = 0; c < 12; c++) o[c] = v[c];
        }
    }
    if (state && live) {
        double *q = state + (size_t)s * 63 + j * 9;
        q[0] = b.pos.x; q[1] = b.pos.y; q[2] = b.pos.z;
        q[3] = b.vel.x; q[4] = b.vel.y; q[5] = b.vel.z;
        q[6] = b.acc.x; q[7] = b.acc.y; q[8] = b.acc.z;
    }
}

int launch_lin_plain(const double *d_ps0, const double *d_sig, double *d_state, double *d_trk, const Pot &P,
                     double eps_pos, double eps_vel, double dt, int sub, int k0, int nk, int npts, int n, int soa_nc,
                     const int *d_nobj, hipStream_t st) {
    if (n <= 0 || nk <= 0) return 0;
    PHX_REQUIRE(k0 == 0 || d_state, "a sliced track needs its state buffer");
    hipLaunchKernelGGL(k_lin_plain, dim3(blocks_for((long)n * 8, 256)), dim3(256), 0, st, d_ps0, d_sig, d_state,
                       d_trk, P, eps_pos, eps_vel, dt, sub, k0, nk, npts, n, soa_nc, d_nobj);
    PHX_CK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- full chi2 scan
#define CSOA(k, c, ci) csoa[((size_t)(k) * 12 + (c)) * nc + (ci)]

// One lane per (star, cluster) pair, every point scanned (reference: src/phygas.c:409-411 with the scan of
// :370-407; device src/phygas_kernel_body.h:125-163).
__global__ __launch_bounds__(256) void k_chi2_plain(const double *__restrict__ trk, const double *__restrict__ csoa,
                                                    const int *__restrict__ cn, const double *__restrict__ R2s, int nb,
                                                    int nc, int npts, long p0, long np, double *__restrict__ o_chi2,
                                                    int *__restrict__ o_kb, double *__restrict__ o_sb) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long id = p0 + t;
    if (t >= np || id >= (long)nb * nc) return;
    const int s = (int)(id / nc), ci = (int)(id % nc);
    const int n = min(npts, cn[ci]);
    const double R2 = R2s[ci];
    const double *tr = trk + (size_t)s * npts * 12;
    double best = __builtin_inf(), sb = 0.0;
    int kb = 0;
    for (int k = 0; k < n; k++) {
        double M[6];
#pragma unroll
        for (int i = 0; i < 6; i++) M[i] = tr[k * 12 + 6 + i] + CSOA(k, 6 + i, ci);
        M[0] += R2; M[3] += R2; M[5] += R2;
        V3 m0 = v3(tr[k * 12] - CSOA(k, 0, ci), tr[k * 12 + 1] - CSOA(k, 1, ci), tr[k * 12 + 2] - CSOA(k, 2, ci));
        double q = mahalanobis(M, m0.x, m0.y, m0.z), sv = 0.0;
        if (k + 1 < n) {
            V3 m1 = v3(tr[(k + 1) * 12] - CSOA(k + 1, 0, ci), tr[(k + 1) * 12 + 1] - CSOA(k + 1, 1, ci),
                       tr[(k + 1) * 12 + 2] - CSOA(k + 1, 2, ci));
            sv = seg_param(m0, m1);
            V3 ms = m0 + (m1 - m0) * sv;
            double qs = mahalanobis(M, ms.x, ms.y, ms.z);
            if (qs >= 0 && qs < q) q = qs; else sv = 0.0;
        }
        if (q >= 0 && q < best) { best = q; kb = k; sb = sv; }
    }
    o_chi2[id] = best;
    o_kb[id] = kb;
    o_sb[id] = sb;
}

int launch_chi2_plain(const double *d_trk, const double *d_csoa, const int *d_cn, const double *d_R2, int nb, int nc,
                      int npts, long p0, long np, double *d_chi2, int *d_kb, double *d_sb, hipStream_t st) {
    if (np <= 0) return 0;
    hipLaunchKernelGGL(k_chi2_plain, dim3(blocks_for(np, 256)), dim3(256), 0, st, d_trk, d_csoa, d_cn, d_R2, nb, nc,
                       npts, p0, np, d_chi2, d_kb, d_sb);
    PHX_CK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- finish
// One lane per selected pair: t, d_nom, vrel, sigma and the 512-draw probabilities
// (reference: src/phygas.c:413-444; device src/phygas_kernel_body.h:182-220).
__global__ __launch_bounds__(256) void k_finish_plain(const double *__restrict__ trk, const double *__restrict__ csoa,
                                                      const int *__restrict__ cn, const double *__restrict__ r50s,
                                                      const double *__restrict__ rts, int nc, int npts,
                                                      const int *__restrict__ hs, const int *__restrict__ hc,
                                                      const int *__restrict__ hkb, const double *__restrict__ hsb,
                                                      const uint64_t *__restrict__ hseed, int h0, int nh,
                                                      double *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nh) return;
    const int h = h0 + t;
    const int s = hs[h], ci = hc[h], kb = hkb[h];
    const double sb = hsb[h];
    const int n = min(npts, cn[ci]);
    const double *tr = trk + (size_t)s * npts * 12;
    const int k1 = kb + 1 < n ? kb + 1 : kb;
    V3 a = v3(tr[kb * 12] - CSOA(kb, 0, ci), tr[kb * 12 + 1] - CSOA(kb, 1, ci), tr[kb * 12 + 2] - CSOA(kb, 2, ci));
    V3 b = v3(tr[k1 * 12] - CSOA(k1, 0, ci), tr[k1 * 12 + 1] - CSOA(k1, 1, ci), tr[k1 * 12 + 2] - CSOA(k1, 2, ci));
    V3 m = a + (b - a) * sb;
    V3 dv = v3(tr[kb * 12 + 3] - CSOA(kb, 3, ci), tr[kb * 12 + 4] - CSOA(kb, 4, ci), tr[kb * 12 + 5] - CSOA(kb, 5, ci));
    double C[6];
#pragma unroll
    for (int i = 0; i < 6; i++) C[i] = tr[kb * 12 + 6 + i] + CSOA(kb, 6 + i, ci);
    int in50, intot;
    finish_counts(C, m, hseed[h], r50s[ci], rts[ci], in50, intot);
    double *o = out + (size_t)h * 6;
    o[0] = (kb + sb) * STEP_MYR;
    o[1] = norm(m) * 1000.0;
    o[2] = sqrt((C[0] + C[3] + C[5]) / 3.0) * 1000.0;
    o[3] = (double)in50 / 512;
    o[4] = (double)intot / 512;
    o[5] = norm(dv) * KMS_PER_KPCMYR;
}

int launch_finish_plain(const double *d_trk, const double *d_csoa, const int *d_cn, const double *d_r50,
                        const double *d_rt, int nc, int npts, const int *d_hs, const int *d_hc, const int *d_hkb,
                        const double *d_hsb, const uint64_t *d_hseed, int h0, int nh, double *d_out, hipStream_t st) {
    if (nh <= 0) return 0;
    hipLaunchKernelGGL(k_finish_plain, dim3(blocks_for(nh, 256)), dim3(256), 0, st, d_trk, d_csoa, d_cn, d_r50, d_rt,
                       nc, npts, d_hs, d_hc, d_hkb, d_hsb, d_hseed, h0, nh, d_out);
    PHX_CK(hipGetLastError());
    return 0;
}

// Finish of the pairs a device-side selection kept: entry h of idx[0 .. min(*count, cap)) is a pair id
// s*nc + ci.  The pair seed ph_seed(id, name) is completed on the device from the star's hash prefix
// "id|" and the cluster's name.  Grid-stride, so the launch needs no host-side count.
__global__ __launch_bounds__(256) void k_finish_sel_plain(
    const double *__restrict__ trk, const double *__restrict__ csoa, const int *__restrict__ cn,
    const double *__restrict__ r50s, const double *__restrict__ rts, const PhCluster *__restrict__ cl,
    const uint64_t *__restrict__ sprefix, int nc, int npts, const long *__restrict__ idx,
    const long *__restrict__ count, long cap, const int *__restrict__ kb_all, const double *__restrict__ sb_all,
    double *__restrict__ out) {
    const long nh = min(*count, cap);
    for (long h = (long)blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += (long)gridDim.x * blockDim.x) {
        const long id = idx[h];
        const int s = (int)(id / nc), ci = (int)(id % nc), kb = kb_all[id];
        const double sb = sb_all[id];
        const int n = min(npts, cn[ci]);
        const double *tr = trk + (size_t)s * npts * 12;
        const int k1 = kb + 1 < n ? kb + 1 : kb;
        V3 a = v3(tr[kb * 12] - CSOA(kb, 0, ci), tr[kb * 12 + 1] - CSOA(kb, 1, ci), tr[kb * 12 + 2] - CSOA(kb, 2, ci));
        V3 b = v3(tr[k1 * 12] - CSOA(k1, 0, ci), tr[k1 * 12 + 1] - CSOA(k1, 1, ci), tr[k1 * 12 + 2] - CSOA(k1, 2, ci));
        V3 m = a + (b - a) * sb;
        V3 dv = v3(tr[kb * 12 + 3] - CSOA(kb, 3, ci), tr[kb * 12 + 4] - CSOA(kb, 4, ci),
                   tr[kb * 12 + 5] - CSOA(kb, 5, ci));
        double C[6];
#pragma unroll
        for (int i = 0; i < 6; i++) C[i] = tr[kb * 12 + 6 + i] + CSOA(kb, 6 + i, ci);
        int in50, intot;
        finish_counts(C, m, fnv_feed(sprefix[s], cl[ci].name, 40), r50s[ci], rts[ci], in50, intot);
        double *o = out + (size_t)h * 6;
        o[0] = (kb + sb) * STEP_MYR;
        o[1] = norm(m) * 1000.0;
        o[2] = sqrt((C[0] + C[3] + C[5]) / 3.0) * 1000.0;
        o[3] = (double)in50 / 512;
        o[4] = (double)intot / 512;
        o[5] = norm(dv) * KMS_PER_KPCMYR;
    }
}

int launch_finish_sel_plain(const double *d_trk, const double *d_csoa, const int *d_cn, const double *d_r50,
                            const double *d_rt, const PhCluster *d_cl, const uint64_t *d_sprefix, int nc, int npts,
                            const long *d_idx, const long *d_count, long cap, const int *d_kb, const double *d_sb,
                            double *d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_finish_sel_plain, dim3(2048), dim3(256), 0, st, d_trk, d_csoa, d_cn, d_r50, d_rt, d_cl,
                       d_sprefix, nc, npts, d_idx, d_count, cap, d_kb, d_sb, d_out);
    PHX_CK(hipGetLastError());
    return 0;
}

// Output rows of the pairs a second selection kept (finished pairs with p_r50 >= p_min).
__global__ __launch_bounds__(256) void k_rows(const long *__restrict__ pair_idx, const long *__restrict__ ridx,
                                              const long *__restrict__ rcount, long rcap, int nc,
                                              const int *__restrict__ sgroup, const double *__restrict__ chi2_all,
                                              const int *__restrict__ kb_all, const double *__restrict__ sb_all,
                                              const double *__restrict__ fout, PhxRow *__restrict__ rows) {
    const long nr = min(*rcount, rcap);
    for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < nr; r += (long)gridDim.x * blockDim.x) {
        const long h = ridx[r], id = pair_idx[h];
        PhxRow o;
        o.star = (int)(id / nc);
        o.cluster = (int)(id % nc);
        o.kb = kb_all[id];
        o.flags = sgroup[o.star];
        o.chi2 = chi2_all[id];
        o.sb = sb_all[id];
        const double *f = fout + (size_t)h * 6;
        o.t = f[0]; o.d_nom_pc = f[1]; o.sig_pc = f[2]; o.p_r50 = f[3]; o.p_rtot = f[4]; o.vrel = f[5];
        rows[r] = o;
    }
}

int launch_rows(const long *d_pair_idx, const long *d_ridx, const long *d_rcount, long rcap, int nc,
                const int *d_sgroup, const double *d_chi2, const int *d_kb, const double *d_sb, const double *d_fout,
                PhxRow *d_rows, hipStream_t st) {
    hipLaunchKernelGGL(k_rows, dim3(256), dim3(256), 0, st, d_pair_idx, d_ridx, d_rcount, rcap, nc, d_sgroup, d_chi2,
                       d_kb, d_sb, d_fout, d_rows);
    PHX_CK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- pair-mode prefilter
// Nominal path (0.5-Myr points, 10 substeps each) and the RMS spread of n_spread draws about the nominal
// (reference: src/phygas.c:98-144), one lane per object; paths stored [obj][npts][3].
template <class Rec, int IsCluster>
__global__ __launch_bounds__(128) void k_track(const Rec *__restrict__ recs, const uint64_t *__restrict__ seeds,
                                               const int *__restrict__ nobj, int n, int n_spread, int npts, Pot P,
                                               double *__restrict__ path, double *__restrict__ spread) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Rec r = recs[i];
    V3 np = v3(0, 0, 0), nv = v3(0, 0, 0);
    if (IsCluster) sample_cluster(P, (const PhCluster &)r, nullptr, np, nv);
    else sample_target(P, (const PhTarget &)r, nullptr, np, nv);
    uint64_t st = seeds[i];
    double sp = 0, sv = 0;
    int k = 0;
    for (int d = 0; d < n_spread; d++) {
        V3 p, v;
        const int rc = IsCluster ? sample_cluster(P, (const PhCluster &)r, &st, p, v)
                                 : sample_target(P, (const PhTarget &)r, &st, p, v);
        if (rc != 0) continue;
        const V3 dp = p - np, dv = v - nv;
        sp += dot(dp, dp);
        sv += dot(dv, dv);
        k++;
    }
    spread[2 * i] = k ? sqrt(sp / k) : 0;
    spread[2 * i + 1] = k ? sqrt(sv / k) : 0;
    Body b = body_start(P, np, nv);
    const int n_i = nobj[i];
    double *o = path + (size_t)i * npts * 3;
    o[0] = b.pos.x; o[1] = b.pos.y; o[2] = b.pos.z;
    const double dt = -STEP_MYR / 10;
    for (int q = 1; q < n_i; q++) {
        for (int s = 0; s < 10; s++) kdk(P, b, dt);
        o[q * 3] = b.pos.x; o[q * 3 + 1] = b.pos.y; o[q * 3 + 2] = b.pos.z;
    }
}

// One lane per (target, cluster): can the target have come within reach (reference: src/phygas.c:158-168)?
__global__ __launch_bounds__(256) void k_prefilter(const double *__restrict__ tpath, const double *__restrict__ tspr,
                                                   const int *__restrict__ tn, const double *__restrict__ cpath,
                                                   const double *__restrict__ cspr, const int *__restrict__ cnn,
                                                   const PhCluster *__restrict__ cl, int nt, int nc, int npts,
                                                   unsigned char *__restrict__ mask) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nt * nc) return;
    const int j = (int)(id / nc), i = (int)(id % nc);
    const int n = min(tn[j], cnn[i]);
    const double sp = hypot(tspr[2 * j], cspr[2 * i]), sv = hypot(tspr[2 * j + 1], cspr[2 * i + 1]);
    const double r0 = (3.0 * cl[i].rtot_pc + 10.0) / 1000.0;
    const double *a = tpath + (size_t)j * npts * 3, *b = cpath + (size_t)i * npts * 3;
    unsigned char hit = 0;
    for (int k = 0; k + 1 < n && !hit; k++) {
        const V3 m0 = v3(a[k * 3] - b[k * 3], a[k * 3 + 1] - b[k * 3 + 1], a[k * 3 + 2] - b[k * 3 + 2]);
        const V3 m1 = v3(a[k * 3 + 3] - b[k * 3 + 3], a[k * 3 + 4] - b[k * 3 + 4], a[k * 3 + 5] - b[k * 3 + 5]);
        const double s = seg_param(m0, m1);
        const double d = norm(m0 + (m1 - m0) * s);
        const double t = (k + s) * STEP_MYR;
        if (d < r0 + 3.0 * (sp + sv * t)) hit = 1;
    }
    mask[id] = hit;
}

int launch_prefilter(const PhTarget *d_t, const uint64_t *d_tseed, const int *d_tn, int nt, const PhCluster *d_c,
                     const uint64_t *d_cseed, const int *d_cn, int nc, int npts, const Pot &P, double *d_tpath,
                     double *d_tspr, double *d_cpath, double *d_cspr, unsigned char *d_mask, hipStream_t st) {
    if (nt <= 0 || nc <= 0) return 0;
    hipLaunchKernelGGL((k_track<PhTarget, 0>), dim3(blocks_for(nt, 128)), dim3(128), 0, st, d_t, d_tseed, d_tn, nt, 32,
                       npts, P, d_tpath, d_tspr);
    hipLaunchKernelGGL((k_track<PhCluster, 1>), dim3(blocks_for(nc, 128)), dim3(128), 0, st, d_c, d_cseed, d_cn, nc,
                       32, npts, P, d_cpath, d_cspr);
    hipLaunchKernelGGL(k_prefilter, dim3(blocks_for((long)nt * nc, 256)), dim3(256), 0, st, d_tpath, d_tspr, d_tn,
                       d_cpath, d_cspr, d_cn, d_c, nt, nc, npts, d_mask);
    PHX_CK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- Monte Carlo clones
// One lane per clone: target and cluster orbits stepped together, closest approach on each step's
// segment, strict-< minimum (reference: src/phygas.c:200-216; device src/phygas_kernel_body.h:225-274).
// st keeps 21 doubles per clone (two bodies and the last separation) so launches can be sliced in time.
__global__ __launch_bounds__(256) void k_pair_plain(const double *__restrict__ ps, const int *__restrict__ steps,
                                                    const double *__restrict__ dts, double *__restrict__ stt,
                                                    double *__restrict__ res, Pot P, int s0, int nk, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    Body a, b;
    V3 r0;
    double best, tbest, vbest;
    double *S = stt + (size_t)k * 21;
    double *R = res + (size_t)k * 3;
    if (s0 == 0) {
        const double *p = ps + (size_t)k * 12;
        a = body_start(P, v3(p[0], p[1], p[2]), v3(p[3], p[4], p[5]));
        b = body_start(P, v3(p[6], p[7], p[8]), v3(p[9], p[10], p[11]));
        r0 = a.pos - b.pos;
        best = norm(r0);
        tbest = 0;
        vbest = norm(a.vel - b.vel);
    } else {
        a.pos = v3(S[0], S[1], S[2]); a.vel = v3(S[3], S[4], S[5]); a.acc = v3(S[6], S[7], S[8]);
        b.pos = v3(S[9], S[10], S[11]); b.vel = v3(S[12], S[13], S[14]); b.acc = v3(S[15], S[16], S[17]);
        r0 = v3(S[18], S[19], S[20]);
        best = R[0]; tbest = R[1]; vbest = R[2];
    }
    const int ns = steps[k];
    const double dt = dts[k];
    for (int i = s0 + 1; i <= s0 + nk && i <= ns; i++) {
        kdk(P, a, dt);
        kdk(P, b, dt);
        V3 r1 = a.pos - b.pos;
        double s = seg_param(r0, r1);
        double dm = norm(r0 + (r1 - r0) * s);
        if (dm < best) {
            best = dm;
            tbest = (i - 1 + s) * -dt;
            vbest = norm(a.vel - b.vel);
        }
        r0 = r1;
    }
    S[0] = a.pos.x; S[1] = a.pos.y; S[2] = a.pos.z; S[3] = a.vel.x; S[4] = a.vel.y; S[5] = a.vel.z;
    S[6] = a.acc.x; S[7] = a.acc.y; S[8] = a.acc.z;
    S[9] = b.pos.x; S[10] = b.pos.y; S[11] = b.pos.z; S[12] = b.vel.x; S[13] = b.vel.y; S[14] = b.vel.z;
    S[15] = b.acc.x; S[16] = b.acc.y; S[17] = b.acc.z;
    S[18] = r0.x; S[19] = r0.y; S[20] = r0.z;
    R[0] = best; R[1] = tbest; R[2] = vbest;
}

int launch_pair_plain(const double *d_ps, const int *d_steps, const double *d_dts, double *d_st, double *d_res,
                      const Pot &P, int s0, int nk, int n, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_pair_plain, dim3(blocks_for(n, 256)), dim3(256), 0, st, d_ps, d_steps, d_dts, d_st, d_res, P,
                       s0, nk, n);
    PHX_CK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- correctly rounded math, device side
// The device evaluation of crmath.h over given arguments (fn: 0 pow(x, y), 1 x^1.5, 2 hypot(x, y), 3 sin x,
// 4 cos x), for the parity check against the same source compiled for the host.
__global__ __launch_bounds__(256) void k_crmath(int fn, const double *__restrict__ x, const double *__restrict__ y,
                                                long n, double *__restrict__ out) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double s, c, v;
        switch (fn) {
        case 0: v = cr::pow_cr(x[i], y[i]); break;
        case 1: v = cr::pow15_cr(x[i]); break;
        case 2: v = cr::hypot_cr(x[i], y[i]); break;
        case 3: cr::sincos_cr(x[i], s, c); v = s; break;
        default: cr::sincos_cr(x[i], s, c); v = c; break;
        }
        out[i] = v;
    }
}

}  // namespace phx

// ================================================================ pg_* ABI (drop-in for the reference's
// GPU backend, src/phygas_gpu.h:16-31): synchronous calls, 0 = ok, messages on stderr.
using namespace phx;

static int pg_params(PgBuf params, Pot &P) {
    double p[9];
    PHX_CK(hipMemcpy(p, params, sizeof p, hipMemcpyDeviceToHost));
    P.bulge_gm = p[0]; P.bulge_b = p[1]; P.disk_gm = p[2]; P.disk_a = p[3]; P.disk_b = p[4];
    P.halo_gm = p[5]; P.halo_a = p[6]; P.halo_cut = p[7]; P.halo_gamma = p[8];
    return 0;
}

extern "C" int pg_init(int platform, int device) {
    (void)platform;
    int n = 0;
    PHX_CK(hipGetDeviceCount(&n));
    if (n <= 0) {
        set_error("pg_init", "no HIP device");
        return 1;
    }
    if (device < 0 || device >= n) device = 0;
    PHX_CK(hipSetDevice(device));
    hipDeviceProp_t prop;
    PHX_CK(hipGetDeviceProperties(&prop, device));
    fprintf(stderr, "phygas_hip: device %d: %s (%s, %d CUs, %.0f MB)\n", device, prop.name, prop.gcnArchName,
            prop.multiProcessorCount, prop.totalGlobalMem / 1048576.0);
    return 0;
}

extern "C" PgBuf pg_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

extern "C" int pg_put(PgBuf b, const void *src, size_t bytes) {
    PHX_CK(hipMemcpy(b, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int pg_get(PgBuf b, void *dst, size_t bytes) {
    PHX_CK(hipMemcpy(dst, b, bytes, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" void pg_free(PgBuf b) {
    if (b) (void)hipFree(b);
}

extern "C" int pg_lin(PgBuf ps0, PgBuf sig, PgBuf state, PgBuf trk, PgBuf params, double eps_pos, double eps_vel,
                      double dt, int sub, int k0, int nk, int npts, int ns) {
    Pot P;
    if (pg_params(params, P)) return 1;
    if (launch_lin_plain((const double *)ps0, (const double *)sig, (double *)state, (double *)trk, P, eps_pos, eps_vel,
                         dt, sub, k0, nk, npts, ns, 0, nullptr, nullptr))
        return 1;
    PHX_CK(hipDeviceSynchronize());
    return 0;
}

extern "C" int pg_chi2(PgBuf trk, PgBuf csoa, PgBuf cn, PgBuf R2s, int nb, int nc, int npts, int p0, int np,
                       PgBuf chi2, PgBuf kb, PgBuf sb) {
    if (launch_chi2_plain((const double *)trk, (const double *)csoa, (const int *)cn, (const double *)R2s, nb, nc, npts,
                          p0, np, (double *)chi2, (int *)kb, (double *)sb, nullptr))
        return 1;
    PHX_CK(hipDeviceSynchronize());
    return 0;
}

extern "C" int pg_finish(PgBuf trk, PgBuf csoa, PgBuf cn, PgBuf r50s, PgBuf rts, int nc, int npts, PgBuf hs, PgBuf hc,
                         PgBuf hkb, PgBuf hsb, PgBuf hseed, int h0, int nh, PgBuf out) {
    if (launch_finish_plain((const double *)trk, (const double *)csoa, (const int *)cn, (const double *)r50s,
                            (const double *)rts, nc, npts, (const int *)hs, (const int *)hc, (const int *)hkb,
                            (const double *)hsb, (const uint64_t *)hseed, h0, nh, (double *)out, nullptr))
        return 1;
    PHX_CK(hipDeviceSynchronize());
    return 0;
}

extern "C" int pg_pair(PgBuf ps, PgBuf steps, PgBuf dts, PgBuf st, PgBuf res, PgBuf params, int s0, int nk, int n) {
    Pot P;
    if (pg_params(params, P)) return 1;
    if (launch_pair_plain((const double *)ps, (const int *)steps, (const double *)dts, (double *)st, (double *)res, P,
                          s0, nk, n, nullptr))
        return 1;
    PHX_CK(hipDeviceSynchronize());
    return 0;
}

static double crmath_eval(int fn, double x, double y) {
    double s, c;
    switch (fn) {
    case 0: return cr::pow_cr(x, y);
    case 1: return cr::pow15_cr(x);
    case 2: return cr::hypot_cr(x, y);
    case 3: cr::sincos_cr(x, s, c); return s;
    default: cr::sincos_cr(x, s, c); return c;
    }
}

extern "C" int phx_check_crmath(int fn, int on_device, const double *x, const double *y, long n, double *out) {
    PHX_REQUIRE(fn >= 0 && fn <= 4 && n >= 0, "bad function or count");
    if (n == 0) return 0;
    if (!on_device) {   // the host compilation of the same source
        for (long i = 0; i < n; i++) out[i] = crmath_eval(fn, x[i], y[i]);
        return 0;
    }
    double *d = nullptr;
    PHX_CK(hipMalloc(&d, 24 * (size_t)n));
    int rc = 0;
    if (hipMemcpy(d, x, 8 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n, y, 8 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) {
        rc = 1;
    } else {
        hipLaunchKernelGGL(k_crmath, dim3(1024), dim3(256), 0, 0, fn, d, d + n, n, d + 2 * n);
        if (hipGetLastError() != hipSuccess || hipMemcpy(out, d + 2 * n, 8 * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
            rc = 1;
    }
    (void)hipFree(d);
    if (rc) set_error("phx_check_crmath", "device evaluation failed");
    return rc;
}
