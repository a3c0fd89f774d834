// Fast form of the screening pass and of the Monte Carlo clones.
//
// Screening, per pass over a resident star batch (ns stars, real and control fused) against the resident
// cluster tables:
//   1. draws         ps0 / Sigma0 per star (the plain FP64 kernel; 64 draws per star)
//   2. k_lin_fast    tracks: FP64 nominal orbit, the six neighbour offsets of the finite-difference STM
//                    propagated with the FP32 difference force (force32.h) and FP64 accumulators,
//                    S = A Sigma0 A^T assembled in FP64.  Two lanes per star, three offsets each.
//   3. k_star_tiles  per (star, 8-point tile): bounding box of the positions in a frame co-rotating with
//                    the Galaxy, and the radius term of the cut (below)
//   4. k_scan_tiles  persistent, length-bucketed: each wavefront takes 64 clusters (sorted by track length)
//                    x 32 stars, tests every (pair, tile) and appends the pairs with any tile that may hold
//                    a point under the cut, with a bit mask of those tiles
//   5. k_refine      one lane per candidate pair: FP32 point test, then the reference's FP64 evaluation
//                    (two Mahalanobis forms per point, strict-< first-index minimum) on surviving points;
//                    pairs with chi2 < chi2_max are appended to the kept list
//   6. k_finish_fast one lane per kept pair: the 512 Box-Muller draws of the pair's seed, screened draw
//                    by draw (below), exact counts; rows with p_r50 >= p_min appended
//
// Exactness of the cut.  For a point k of a pair, chi2_k = min(q(m0), q(ms)) >= d^2 / lambda_max(M_k) with
// d the distance of the step's segment m0 -> m1 from the origin (reference proof: src/phygas.c:360-369), and
// lambda_max(M_k) <= lambda_s + lambda_c + R^2 (Weyl) with lambda_o an upper bound of the largest eigenvalue
// of each covariance (min of the trace and the Gershgorin bound).  A tile's boxes are built from the points
// rotated by R_k (rotation angle w*k about the Galactic z axis); rotating both ends of a segment by R_k moves
// the far end by at most w*|m1| from the box, so d >= gap - w (gap + diag_s + diag_c), gap = distance of the
// origin from box_s - box_c.  When that exceeds sqrt(chi2_max) (sigma_s + sigma_c + R) every point of the
// tile has chi2 > chi2_max and cannot be the minimum of a kept pair.  Float rounding is covered by outward
// rounding of the boxes and a 1e-5 relative + 1e-5 kpc margin.  Kept pairs get the FP64 chi2, kb and sb of
// a full scan of the same tracks.
//
// Exactness of the finish.  splitmix64's state after i calls is seed + i*gamma, so draw i's six uniforms
// are known by index.  x0 = m0 + l00 z0 needs |x0| < r to count, i.e. z0 in an interval I; z0 = rho cos(.)
// with rho = sqrt(-2 ln u1), so a draw whose u1 exceeds exp(-dist(0, I)^2 / 2) cannot count and its other
// five uniforms are never generated.  Otherwise z0 is estimated in FP32 (hardware log2 / cos) and only
// draws whose estimate falls within 1e-3 of I are evaluated exactly (the reference's FP64 arithmetic).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fast.h"
#include "force32.h"

namespace phx {

// ------------------------------------------------------------------------------------------- layout
constexpr int TILE = 8;                 // points per culling tile
constexpr int SPW = 32;                 // stars per wavefront in the tile scan
constexpr double OMEGA_PT = -0.5 * (220.0 / KMS_PER_KPCMYR) / 8.122;   // frame rotation per 0.5-Myr point
constexpr float ROT_SLACK = (float)(0.5 * (220.0 / KMS_PER_KPCMYR) / 8.122 * 1.000001);
constexpr float REL_MARGIN = 1e-5f, ABS_MARGIN = 1e-5f;

struct Cand { int star, cluster; unsigned long long mask; };
struct Kept { int star, cluster, kb, pad; double sb, chi2; };

// Work buffers (FastWork::mem slots)
enum { W_PS0, W_SIG, W_TRK, W_TRKF, W_STILE, W_CAND, W_KEPT, W_ROWS, W_CNT, W_NSLOT };
// Table buffers (FastTables::mem slots)
enum { T_CTRK, T_CTRKF, T_CTLO, T_CTHI, T_CNT, T_CPERM, T_NSLOT };

static int ensure(void *&p, size_t &have, size_t need) {
    if (need <= have && p) return 0;
    if (p) PHX_CK(hipFree(p));
    p = nullptr;
    have = 0;
    PHX_CK(hipMalloc(&p, need ? need : 8));
    have = need ? need : 8;
    return 0;
}

// Largest-eigenvalue upper bound of a symmetric PSD 3x3 (xx xy xz yy yz zz): min(trace, Gershgorin).
__device__ __forceinline__ double lam_bound(const double *S) {
    double tr = S[0] + S[3] + S[5];
    double g0 = S[0] + fabs(S[1]) + fabs(S[2]), g1 = S[3] + fabs(S[1]) + fabs(S[4]), g2 = S[5] + fabs(S[2]) + fabs(S[4]);
    double g = fmax(g0, fmax(g1, g2));
    return fmax(fmin(tr, g), 0.0);
}
// Neighbouring floats (finite inputs): one step towards +inf / -inf.
__device__ __forceinline__ float next_up(float f) {
    if (f == 0.0f) return __int_as_float(1);
    const int b = __float_as_int(f);
    return __int_as_float(f > 0.0f ? b + 1 : b - 1);
}
__device__ __forceinline__ float next_down(float f) { return -next_up(-f); }
__device__ __forceinline__ float up(double v) {   // float >= v
    float f = (float)v;
    return (double)f < v ? next_up(f) : f;
}
__device__ __forceinline__ float down(double v) {   // float <= v
    float f = (float)v;
    return (double)f > v ? next_down(f) : f;
}

// ------------------------------------------------------------------------------------------- FP64 nominal
// Acceleration with x^1.5 as u*sqrt(u) and the halo powers from one log2 and two exp2; every operation in
// FP64 (~1 ulp from the reference's pow-based evaluation).
__device__ __forceinline__ V3 accel_fast64(const Pot &P, V3 p) {
    const double R2 = p.x * p.x + p.y * p.y, r2 = R2 + p.z * p.z;
    const double u = r2 + P.bulge_b * P.bulge_b;
    const double fb = -P.bulge_gm / (u * sqrt(u));
    const double zeta = sqrt(p.z * p.z + P.disk_b * P.disk_b);
    const double az = P.disk_a + zeta;
    const double D2 = R2 + az * az;
    const double fd = -P.disk_gm / (D2 * sqrt(D2));
    const double r = sqrt(r2);
    const double rs = r > 1e-12 ? r : 1e-12;
    const double xh = (rs < P.halo_cut ? rs : P.halo_cut) / P.halo_a;
    const double lx = log2(xh);
    const double gm = P.halo_gm * exp2(P.halo_gamma * lx) / (1.0 + exp2((P.halo_gamma - 1.0) * lx));
    const double fh = -gm / (rs * rs * rs);
    return v3((fb + fd + fh) * p.x, (fb + fd + fh) * p.y, (fb + fh) * p.z + fd * p.z * (az / zeta));
}

// ------------------------------------------------------------------------------------------- tracks
// Two lanes per star; lane q propagates offsets 3q .. 3q+2 (x, y, z for q = 0; vx, vy, vz for q = 1) and
// both lanes run the same FP64 nominal.  Track layout as the plain path: [star][npts][12] FP64, plus
// [star][npts] float4 (x, y, z, lambda bound of S) for the cut.
__global__ __launch_bounds__(256) void k_lin_fast(const double *__restrict__ ps0, const double *__restrict__ sig,
                                                  double *__restrict__ trk, float4 *__restrict__ trkf, Pot P, PotF Q,
                                                  int npts, int n) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = gid >> 1, q = gid & 1;
    const bool live = s < n;
    const double eps = PH_LIN_EPS_POS, dt = -STEP_MYR / PH_LIN_SUB, h = 0.5 * dt;
    V3 x = v3(0, 0, 0), v = v3(0, 0, 0);
    if (live) {
        const double *p = ps0 + (size_t)s * 6;
        x = v3(p[0], p[1], p[2]);
        v = v3(p[3], p[4], p[5]);
    }
    double dx[3][3], dv[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            dx[i][c] = (q == 0 && c == i) ? eps : 0.0;
            dv[i][c] = (q == 1 && c == i) ? eps : 0.0;
        }
    V3 a = accel_fast64(P, x);
    F3 da[3];
#pragma unroll
    for (int i = 0; i < 3; i++) da[i] = dforce32(Q, (float)x.x, (float)x.y, (float)x.z, (float)dx[i][0], (float)dx[i][1], (float)dx[i][2]);
    for (int k = 0; k < npts; k++) {
        if (k > 0) {
            for (int it = 0; it < PH_LIN_SUB; it++) {
                v = v + a * h;
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    dv[i][0] += (double)da[i].x * h; dv[i][1] += (double)da[i].y * h; dv[i][2] += (double)da[i].z * h;
                }
                x = x + v * dt;
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int c = 0; c < 3; c++) dx[i][c] += dv[i][c] * dt;
                a = accel_fast64(P, x);
                const float xf = (float)x.x, yf = (float)x.y, zf = (float)x.z;
#pragma unroll
                for (int i = 0; i < 3; i++) da[i] = dforce32(Q, xf, yf, zf, (float)dx[i][0], (float)dx[i][1], (float)dx[i][2]);
                v = v + a * h;
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    dv[i][0] += (double)da[i].x * h; dv[i][1] += (double)da[i].y * h; dv[i][2] += (double)da[i].z * h;
                }
            }
        }
        // A[r][3q+i] = dx[i][r] / eps; lane 0 receives lane 1's three columns
        double A[3][6];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const double mine = dx[i][r] / eps;
                const double other = __shfl_xor(mine, 1, 64);
                A[r][i] = q == 0 ? mine : other;
                A[r][3 + i] = q == 0 ? other : mine;
            }
        if (!live || q != 0) continue;
        const double *sg = sig + (size_t)s * 36;
        double AS[3][6];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 6; c++) {
                double acc = 0.0;
#pragma unroll
                for (int b = 0; b < 6; b++) acc += A[r][b] * sg[b * 6 + c];
                AS[r][c] = acc;
            }
        double S[6];
        int o = 0;
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = r; c < 3; c++) {
                double acc = 0.0;
#pragma unroll
                for (int b = 0; b < 6; b++) acc += AS[r][b] * A[c][b];
                S[o++] = acc;
            }
        double *out = trk + ((size_t)s * npts + k) * 12;
        out[0] = x.x; out[1] = x.y; out[2] = x.z; out[3] = v.x; out[4] = v.y; out[5] = v.z;
#pragma unroll
        for (int i = 0; i < 6; i++) out[6 + i] = S[i];
        trkf[(size_t)s * npts + k] = make_float4((float)x.x, (float)x.y, (float)x.z, up(lam_bound(S) * (1.0 + 1e-6)));
    }
}

// ------------------------------------------------------------------------------------------- tiles
// Tile j of an object covers points [jT, jT+T) of its scan and its box the points [jT, min(jT+T, n-1)]
// (each step's far end).  lo.w = e = (sqrt(chi2_max) sigma + w diag) / (1 - w), hi.w = diag.
__device__ void tile_box(const double *pos, size_t stride, int n, int j, double lam_sqrt_max, float kchi,
                         float4 &lo, float4 &hi) {
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    const int k0 = j * TILE, k1 = min(k0 + TILE, n - 1);
    for (int k = k0; k <= k1; k++) {
        const double *p = pos + (size_t)k * stride;
        const double ph = OMEGA_PT * k, c = cos(ph), sn = sin(ph);
        const double rx = c * p[0] - sn * p[1], ry = sn * p[0] + c * p[1], rz = p[2];
        mn[0] = fmin(mn[0], rx); mn[1] = fmin(mn[1], ry); mn[2] = fmin(mn[2], rz);
        mx[0] = fmax(mx[0], rx); mx[1] = fmax(mx[1], ry); mx[2] = fmax(mx[2], rz);
    }
    const double dg = sqrt((mx[0] - mn[0]) * (mx[0] - mn[0]) + (mx[1] - mn[1]) * (mx[1] - mn[1]) +
                           (mx[2] - mn[2]) * (mx[2] - mn[2]));
    const double w = ROT_SLACK;
    const double e = ((double)kchi * lam_sqrt_max + w * dg) / (1.0 - w);
    lo = make_float4(down(mn[0]), down(mn[1]), down(mn[2]), up(e * (1.0 + REL_MARGIN) + ABS_MARGIN));
    hi = make_float4(up(mx[0]), up(mx[1]), up(mx[2]), up(dg));
}

// One lane per (star, tile).
__global__ __launch_bounds__(256) void k_star_tiles(const double *__restrict__ trk, const float4 *__restrict__ trkf,
                                                    int npts, int ntiles, int n, float kchi, float4 *__restrict__ tlo,
                                                    float4 *__restrict__ thi) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)n * ntiles) return;
    const int s = (int)(id / ntiles), j = (int)(id % ntiles);
    const int k0 = j * TILE, k1 = min(k0 + TILE, npts) - 1;
    double lam = 0.0;
    for (int k = k0; k <= k1; k++) lam = fmax(lam, (double)trkf[(size_t)s * npts + k].w);
    float4 lo, hi;
    tile_box(trk + (size_t)s * npts * 12, 12, npts, j, sqrt(lam), kchi, lo, hi);
    tlo[id] = lo;
    thi[id] = hi;
}

// Cluster tables for the fast form: AoS FP64 tracks [c][npts][12], float4 tracks [c][npts], and tiles in
// structure-of-arrays [tile][nc_pad] over clusters sorted by track length (longest first).
__global__ __launch_bounds__(256) void k_cluster_aos(const double *__restrict__ csoa, const int *__restrict__ cn,
                                                     int nc, int npts, double *__restrict__ ctrk,
                                                     float4 *__restrict__ ctrkf) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nc * npts) return;
    const int c = (int)(id / npts), k = (int)(id % npts);
    double v[12];
#pragma unroll
    for (int q = 0; q < 12; q++) v[q] = k < cn[c] ? csoa[((size_t)k * 12 + q) * nc + c] : 0.0;
    double *o = ctrk + id * 12;
#pragma unroll
    for (int q = 0; q < 12; q++) o[q] = v[q];
    ctrkf[id] = make_float4((float)v[0], (float)v[1], (float)v[2], up(lam_bound(v + 6) * (1.0 + 1e-6)));
}

__global__ __launch_bounds__(256) void k_cluster_tiles(const double *__restrict__ ctrk,
                                                       const float4 *__restrict__ ctrkf, const int *__restrict__ cn,
                                                       const double *__restrict__ R2, const int *__restrict__ perm,
                                                       int nc, int nc_pad, int npts, int ntiles, float kchi,
                                                       float4 *__restrict__ tlo, float4 *__restrict__ thi) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nc_pad * ntiles) return;
    const int j = (int)(id / nc_pad), cs = (int)(id % nc_pad);
    float4 lo = make_float4(1e30f, 1e30f, 1e30f, 0.f), hi = make_float4(-1e30f, -1e30f, -1e30f, 0.f);
    if (cs < nc) {
        const int c = perm[cs], n = cn[c];
        if (j * TILE < n) {
            const int k0 = j * TILE, k1 = min(k0 + TILE, n) - 1;
            double lam = 0.0;
            for (int k = k0; k <= k1; k++) lam = fmax(lam, (double)ctrkf[(size_t)c * npts + k].w);
            tile_box(ctrk + (size_t)c * npts * 12, 12, n, j, sqrt(lam), kchi, lo, hi);
            // the kernel radius R of the pair enters through the cluster: e += sqrt(chi2_max) R / (1 - w)
            const double w = ROT_SLACK;
            lo.w = up(((double)lo.w + (double)kchi * sqrt(R2[c]) / (1.0 - w)) * (1.0 + REL_MARGIN));
        }
    }
    tlo[(size_t)j * nc_pad + cs] = lo;
    thi[(size_t)j * nc_pad + cs] = hi;
}

// ------------------------------------------------------------------------------------------- tile scan
// Persistent: each wavefront loops over tasks (cluster block of 64 sorted clusters, star block of SPW
// stars); tasks are ordered longest clusters first.
__global__ __launch_bounds__(256) void k_scan_tiles(const float4 *__restrict__ ctlo, const float4 *__restrict__ cthi,
                                                    const int *__restrict__ cnt_sorted, const int *__restrict__ perm,
                                                    int nc, int nc_pad, const float4 *__restrict__ stlo,
                                                    const float4 *__restrict__ sthi, int ns, int ntiles_s,
                                                    Cand *__restrict__ cand, unsigned long long *__restrict__ ncand,
                                                    long cap) {
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long nwaves = ((long)gridDim.x * blockDim.x) >> 6;
    const int ncb = (nc + 63) / 64, nsb = (ns + SPW - 1) / SPW;
    const long ntasks = (long)ncb * nsb;
    for (long task = wave; task < ntasks; task += nwaves) {
        const int cb = (int)(task / nsb), sb = (int)(task % nsb);
        const int cs = cb * 64 + lane;
        const int nt = cs < nc ? cnt_sorted[cs] : 0;
        const int ntmax = __shfl(nt, 0, 64);   // sorted descending: lane 0 has the block's longest
        const int s0 = sb * SPW;
        const int nsl = min(SPW, ns - s0);
        unsigned long long mask[SPW];
#pragma unroll
        for (int i = 0; i < SPW; i++) mask[i] = 0ull;
        for (int j = 0; j < ntmax; j++) {
            const float4 cl = ctlo[(size_t)j * nc_pad + cs], ch = cthi[(size_t)j * nc_pad + cs];
            const bool cv = j < nt;
#pragma unroll
            for (int i = 0; i < SPW; i++) {
                if (i < nsl) {
                    const size_t si = (size_t)(s0 + i) * ntiles_s + j;
                    const float4 sl = stlo[si], sh = sthi[si];
                    const float gx = fmaxf(fmaxf(sl.x - ch.x, cl.x - sh.x), 0.f);
                    const float gy = fmaxf(fmaxf(sl.y - ch.y, cl.y - sh.y), 0.f);
                    const float gz = fmaxf(fmaxf(sl.z - ch.z, cl.z - sh.z), 0.f);
                    const float g2 = f_fma(gx, gx, f_fma(gy, gy, gz * gz));
                    const float t = sl.w + cl.w;
                    if (cv && g2 <= t * t * (1.0f + REL_MARGIN)) mask[i] |= 1ull << j;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < SPW; i++) {
            if (i < nsl) {   // wave-uniform
                const bool hit = cs < nc && mask[i] != 0ull;
                const unsigned long long b = __ballot(hit);
                if (b != 0ull) {
                    unsigned long long base = 0;
                    if (lane == 0) base = atomicAdd(ncand, (unsigned long long)__popcll(b));
                    base = __shfl(base, 0, 64);
                    if (hit) {
                        const unsigned long long o = base + __popcll(b & ((1ull << lane) - 1ull));
                        if ((long)o < cap) cand[o] = Cand{s0 + i, perm[cs], mask[i]};
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------------------------------------- refine
__global__ __launch_bounds__(256) void k_refine(const Cand *__restrict__ cand,
                                                const unsigned long long *__restrict__ ncand, long cand_cap,
                                                const double *__restrict__ strk, const float4 *__restrict__ strkf,
                                                const double *__restrict__ ctrk, const float4 *__restrict__ ctrkf,
                                                const int *__restrict__ cn, const double *__restrict__ R2s, int npts,
                                                double chi2_max, Kept *__restrict__ kept,
                                                unsigned long long *__restrict__ nkept, long kept_cap) {
    const long nc_ = min((long)*ncand, cand_cap);
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < nc_; id += (long)gridDim.x * blockDim.x) {
        const Cand cd = cand[id];
        const int s = cd.star, c = cd.cluster;
        const int n = min(npts, cn[c]);
        const double R2 = R2s[c];
        const float R2f = (float)(R2 * (1.0 + 1e-6)) + 1e-30f;
        const float4 *sf = strkf + (size_t)s * npts;
        const float4 *cf = ctrkf + (size_t)c * npts;
        const double *sd = strk + (size_t)s * npts * 12;
        const double *cdp = ctrk + (size_t)c * npts * 12;
        double best = __builtin_inf(), sbest = 0.0;
        int kb = 0;
        unsigned long long mask = cd.mask;
        while (mask) {
            const int j = __builtin_ctzll(mask);
            mask &= mask - 1ull;
            const int k0 = j * TILE, k1 = min(k0 + TILE, n);
            for (int k = k0; k < k1; k++) {
                // FP32 point test: distance of the step's segment from the origin vs sqrt(chi2_max lambda)
                const float4 a0 = sf[k], b0 = cf[k];
                const float m0x = a0.x - b0.x, m0y = a0.y - b0.y, m0z = a0.z - b0.z;
                float d2 = f_fma(m0x, m0x, f_fma(m0y, m0y, m0z * m0z));
                if (k + 1 < n) {
                    const float4 a1 = sf[k + 1], b1 = cf[k + 1];
                    const float rx = (a1.x - b1.x) - m0x, ry = (a1.y - b1.y) - m0y, rz = (a1.z - b1.z) - m0z;
                    const float dd = f_fma(rx, rx, f_fma(ry, ry, rz * rz));
                    const float ad = f_fma(m0x, rx, f_fma(m0y, ry, m0z * rz));
                    if (ad < 0.f && dd > 0.f) {
                        const float sp = fminf(-ad / dd, 1.f);
                        const float px = f_fma(rx, sp, m0x), py = f_fma(ry, sp, m0y), pz = f_fma(rz, sp, m0z);
                        d2 = f_fma(px, px, f_fma(py, py, pz * pz));
                    }
                }
                const float lim = f_sqrt((float)chi2_max * (a0.w + b0.w + R2f)) * (1.0f + REL_MARGIN) + ABS_MARGIN;
                if (f_sqrt(d2) > lim) continue;
                // the reference's FP64 evaluation of the point (reference: src/phygas.c:395-403)
                const double *ps = sd + (size_t)k * 12, *pc = cdp + (size_t)k * 12;
                double M[6];
#pragma unroll
                for (int i = 0; i < 6; i++) M[i] = ps[6 + i] + pc[6 + i];
                M[0] += R2; M[3] += R2; M[5] += R2;
                const V3 m0 = v3(ps[0] - pc[0], ps[1] - pc[1], ps[2] - pc[2]);
                double qv = mahalanobis(M, m0.x, m0.y, m0.z), sv = 0.0;
                if (k + 1 < n) {
                    const V3 m1 = v3(ps[12] - pc[12], ps[13] - pc[13], ps[14] - pc[14]);
                    sv = seg_param(m0, m1);
                    const V3 ms = m0 + (m1 - m0) * sv;
                    const double qs = mahalanobis(M, ms.x, ms.y, ms.z);
                    if (qs >= 0 && qs < qv) qv = qs; else sv = 0.0;
                }
                if (qv >= 0 && qv < best) { best = qv; kb = k; sbest = sv; }
            }
        }
        if (best < chi2_max) {
            const unsigned long long o = atomicAdd(nkept, 1ull);
            if ((long)o < kept_cap) kept[o] = Kept{s, c, kb, 0, sbest, best};
        }
    }
}

// ------------------------------------------------------------------------------------------- finish
// FP32 estimate of z0 = sqrt(-2 ln u1) cos(2 pi u2) (hardware log2 and cos, cos argument in revolutions).
__device__ __forceinline__ float z0_estimate(double u1, double u2) {
    const float l = __builtin_amdgcn_logf((float)u1);                  // log2
    const float rho = __builtin_sqrtf(-2.0f * 0.6931471805599453f * l);
    const float c = __builtin_amdgcn_cosf((float)u2);                  // cos(2 pi u2)
    return rho * c;
}

// Count draws with |m + L z| < r among the pair's 512 (exact), screening with the x0 interval.
__device__ int count_inside(uint64_t seed, const double l[6], V3 m, double r, double r_alt, int *n_alt) {
    // l = l00 l10 l20 l11 l21 l22
    const double l00 = l[0];
    double lo, hi;   // z0 interval where |m.x + l00 z0| < r
    bool all = false;
    if (l00 > 0) {
        lo = (-r - m.x) / l00;
        hi = (r - m.x) / l00;
    } else {
        all = fabs(m.x) < r;
        lo = -1e300;
        hi = 1e300;
    }
    const double a = all ? 0.0 : (lo > 0 ? lo : (hi < 0 ? -hi : 0.0));
    const double ucut = exp(-0.5 * a * a * (1.0 - 1e-9)) * (1.0 + 1e-9);   // u1 above this: |z0| < a
    const float flo = (float)lo - 1e-3f, fhi = (float)hi + 1e-3f;
    int cnt = 0, alt = 0;
    // |z0| <= 9.42 for every representable u1; with l00 = 0, x0 = m.x for every draw
    if ((l00 > 0 && (lo >= 9.5 || hi <= -9.5)) || (l00 <= 0 && !all)) {
        if (n_alt) *n_alt = 0;
        return 0;
    }
    for (int i = 0; i < 512; i++) {
        const uint64_t b = seed + (uint64_t)(6 * i) * GOLDEN_GAMMA;
        const double u1 = to_unit(mix64(b + GOLDEN_GAMMA));
        if (u1 > ucut) continue;
        const double u2 = to_unit(mix64(b + 2 * GOLDEN_GAMMA));
        if (!all) {
            const float ze = z0_estimate(u1 <= 0.0 ? 1.0 / 18446744073709551616.0 : u1, u2);
            if (ze < flo || ze > fhi) continue;
        }
        const double z0 = box_muller(u1, u2);
        const double z1 = box_muller(to_unit(mix64(b + 3 * GOLDEN_GAMMA)), to_unit(mix64(b + 4 * GOLDEN_GAMMA)));
        const double z2 = box_muller(to_unit(mix64(b + 5 * GOLDEN_GAMMA)), to_unit(mix64(b + 6 * GOLDEN_GAMMA)));
        const double x = m.x + l00 * z0, y = m.y + l[1] * z0 + l[3] * z1, z = m.z + l[2] * z0 + l[4] * z1 + l[5] * z2;
        const double rr = sqrt(x * x + y * y + z * z);
        cnt += rr < r;
        alt += rr < r_alt;
    }
    if (n_alt) *n_alt = alt;
    return cnt;
}

__global__ __launch_bounds__(256) void k_finish_fast(const Kept *__restrict__ kept,
                                                     const unsigned long long *__restrict__ nkept, long kept_cap,
                                                     const double *__restrict__ strk, const double *__restrict__ ctrk,
                                                     const int *__restrict__ cn, const double *__restrict__ r50s,
                                                     const double *__restrict__ rts,
                                                     const PhCluster *__restrict__ cl,
                                                     const uint64_t *__restrict__ sprefix,
                                                     const int *__restrict__ sgroup, int npts, int n50_min,
                                                     PhxRow *__restrict__ rows, unsigned long long *__restrict__ nrows,
                                                     long row_cap) {
    const long nk = min((long)*nkept, kept_cap);
    for (long h = (long)blockIdx.x * blockDim.x + threadIdx.x; h < nk; h += (long)gridDim.x * blockDim.x) {
        const Kept kp = kept[h];
        const int s = kp.star, c = kp.cluster, kb = kp.kb;
        const int n = min(npts, cn[c]);
        const int k1 = kb + 1 < n ? kb + 1 : kb;
        const double *ps = strk + (size_t)s * npts * 12, *pc = ctrk + (size_t)c * npts * 12;
        const V3 a = v3(ps[kb * 12] - pc[kb * 12], ps[kb * 12 + 1] - pc[kb * 12 + 1], ps[kb * 12 + 2] - pc[kb * 12 + 2]);
        const V3 b = v3(ps[k1 * 12] - pc[k1 * 12], ps[k1 * 12 + 1] - pc[k1 * 12 + 1], ps[k1 * 12 + 2] - pc[k1 * 12 + 2]);
        const V3 m = a + (b - a) * kp.sb;
        double C[6];
#pragma unroll
        for (int i = 0; i < 6; i++) C[i] = ps[kb * 12 + 6 + i] + pc[kb * 12 + 6 + i];
        double L[6];
        L[0] = sqrt(fmax(C[0], 0.0));
        L[1] = L[0] > 0 ? C[1] / L[0] : 0;
        L[2] = L[0] > 0 ? C[2] / L[0] : 0;
        L[3] = sqrt(fmax(C[3] - L[1] * L[1], 0.0));
        L[4] = L[3] > 0 ? (C[4] - L[2] * L[1]) / L[3] : 0;
        L[5] = sqrt(fmax(C[5] - L[2] * L[2] - L[4] * L[4], 0.0));
        const uint64_t seed = fnv_feed(sprefix[s], cl[c].name, 40);
        const double r50 = r50s[c], rt = rts[c];
        int in50, intot;
        if (n50_min > 0) {
            in50 = count_inside(seed, L, m, r50, r50, nullptr);
            if (in50 < n50_min) continue;
            intot = count_inside(seed, L, m, rt, rt, nullptr);
        } else {   // every kept pair is a row: count both, screening with the larger radius
            const int n_big = count_inside(seed, L, m, fmax(rt, r50), rt < r50 ? rt : r50, &in50);
            if (rt < r50) { intot = in50; in50 = n_big; } else intot = n_big;
        }
        PhxRow o;
        o.star = s; o.cluster = c; o.kb = kb; o.flags = sgroup[s];
        o.chi2 = kp.chi2; o.sb = kp.sb;
        o.t = (kb + kp.sb) * STEP_MYR;
        o.d_nom_pc = norm(m) * 1000.0;
        o.sig_pc = sqrt((C[0] + C[3] + C[5]) / 3.0) * 1000.0;
        o.p_r50 = (double)in50 / 512;
        o.p_rtot = (double)intot / 512;
        const V3 dv = v3(ps[kb * 12 + 3] - pc[kb * 12 + 3], ps[kb * 12 + 4] - pc[kb * 12 + 4], ps[kb * 12 + 5] - pc[kb * 12 + 5]);
        o.vrel = norm(dv) * KMS_PER_KPCMYR;
        const unsigned long long ro = atomicAdd(nrows, 1ull);
        if ((long)ro < row_cap) rows[ro] = o;
    }
}

// float4 tracks (position, lambda bound) from FP64 tracks [n][npts][12] (the exact-track variant).
__global__ __launch_bounds__(256) void k_trkf(const double *__restrict__ trk, long npoints, float4 *__restrict__ trkf) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= npoints) return;
    const double *p = trk + id * 12;
    trkf[id] = make_float4((float)p[0], (float)p[1], (float)p[2], up(lam_bound(p + 6) * (1.0 + 1e-6)));
}

// Largest |z0_estimate - z0| over n (u1, u2) pairs drawn from splitmix64(seed): the margin check of the
// finish screen.  One lane per pair; the maximum is reduced with an atomic on its bit pattern (>= 0).
__global__ __launch_bounds__(256) void k_z0_check(uint64_t seed, long n, unsigned long long *__restrict__ worst) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t b = seed + (uint64_t)(2 * i) * GOLDEN_GAMMA;
    double u1 = to_unit(mix64(b + GOLDEN_GAMMA)), u2 = to_unit(mix64(b + 2 * GOLDEN_GAMMA));
    if (i % 3 == 1) u1 = ldexp(u1, -(int)(i % 53));   // also the tail: small u1, large |z0|
    const double z = box_muller(u1, u2);
    const float e = z0_estimate(u1 <= 0.0 ? 1.0 / 18446744073709551616.0 : u1, u2);
    const double d = fabs((double)e - z);
    atomicMax(worst, (unsigned long long)__double_as_longlong(d));
}

__global__ void k_zero_counters(unsigned long long *c, int n) {
    if (threadIdx.x < n) c[threadIdx.x] = 0ull;
}

// ------------------------------------------------------------------------------------------- host side
static int n_tiles(int npts) { return (npts + TILE - 1) / TILE; }

int fast_build_tables(FastTables &T, const double *d_csoa, const int *d_cn, const double *d_R2, int nc, int npts,
                      hipStream_t st) {
    T.nc = nc;
    T.npts = npts;
    T.ntiles = n_tiles(npts);
    T.chi2_built = -1.0;
    if (ensure(T.mem[T_CTRK], T.bytes[T_CTRK], sizeof(double) * 12 * (size_t)nc * npts) ||
        ensure(T.mem[T_CTRKF], T.bytes[T_CTRKF], sizeof(float4) * (size_t)nc * npts))
        return 1;
    hipLaunchKernelGGL(k_cluster_aos, dim3(blocks_for((long)nc * npts, 256)), dim3(256), 0, st, d_csoa, d_cn, nc, npts,
                       (double *)T.mem[T_CTRK], (float4 *)T.mem[T_CTRKF]);
    PHX_CK(hipGetLastError());
    // sort clusters by tile count, longest first (length buckets of the tile scan)
    std::vector<int> cn(nc), perm(nc), nt_sorted(nc);
    PHX_CK(hipMemcpyAsync(cn.data(), d_cn, 4 * (size_t)nc, hipMemcpyDeviceToHost, st));
    PHX_CK(hipStreamSynchronize(st));
    for (int i = 0; i < nc; i++) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return cn[a] > cn[b]; });
    for (int i = 0; i < nc; i++) nt_sorted[i] = n_tiles(cn[perm[i]]);
    T.nc_pad = (nc + 63) / 64 * 64;
    if (ensure(T.mem[T_CNT], T.bytes[T_CNT], 4 * (size_t)T.nc_pad) ||
        ensure(T.mem[T_CPERM], T.bytes[T_CPERM], 4 * (size_t)T.nc_pad))
        return 1;
    nt_sorted.resize(T.nc_pad, 0);
    perm.resize(T.nc_pad, 0);
    PHX_CK(hipMemcpyAsync(T.mem[T_CNT], nt_sorted.data(), 4 * (size_t)T.nc_pad, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(T.mem[T_CPERM], perm.data(), 4 * (size_t)T.nc_pad, hipMemcpyHostToDevice, st));
    PHX_CK(hipStreamSynchronize(st));
    T.d_cn = d_cn;
    T.d_R2 = d_R2;
    return 0;
}

static int build_cluster_tiles(FastTables &T, double chi2_max, hipStream_t st) {
    if (T.chi2_built == chi2_max) return 0;
    const size_t nt = (size_t)T.ntiles * T.nc_pad;
    if (ensure(T.mem[T_CTLO], T.bytes[T_CTLO], sizeof(float4) * nt) ||
        ensure(T.mem[T_CTHI], T.bytes[T_CTHI], sizeof(float4) * nt))
        return 1;
    hipLaunchKernelGGL(k_cluster_tiles, dim3(blocks_for((long)nt, 256)), dim3(256), 0, st,
                       (const double *)T.mem[T_CTRK], (const float4 *)T.mem[T_CTRKF], T.d_cn, T.d_R2,
                       (const int *)T.mem[T_CPERM], T.nc, T.nc_pad, T.npts, T.ntiles, (float)sqrt(chi2_max),
                       (float4 *)T.mem[T_CTLO], (float4 *)T.mem[T_CTHI]);
    PHX_CK(hipGetLastError());
    T.chi2_built = chi2_max;
    return 0;
}

static int fast_alloc(FastTables &T, FastWork &W, const FastArgs &a, hipStream_t st) {
    const int ns = a.ns, npts = a.npts, ntl = T.ntiles;
    const long npairs = (long)ns * a.nc;
    if (W.cand_cap == 0) W.cand_cap = std::max(1L << 20, npairs / 5);
    if (W.kept_cap == 0) W.kept_cap = std::max(1L << 18, npairs / 16);
    if (W.row_cap == 0) W.row_cap = std::max(1L << 16, W.kept_cap / 8);
    if (ensure(W.mem[W_PS0], W.bytes[W_PS0], 48 * (size_t)ns) || ensure(W.mem[W_SIG], W.bytes[W_SIG], 288 * (size_t)ns) ||
        ensure(W.mem[W_TRK], W.bytes[W_TRK], sizeof(double) * 12 * (size_t)ns * npts) ||
        ensure(W.mem[W_TRKF], W.bytes[W_TRKF], sizeof(float4) * (size_t)ns * npts) ||
        ensure(W.mem[W_STILE], W.bytes[W_STILE], 2 * sizeof(float4) * (size_t)ns * ntl) ||
        ensure(W.mem[W_CAND], W.bytes[W_CAND], sizeof(Cand) * (size_t)W.cand_cap) ||
        ensure(W.mem[W_KEPT], W.bytes[W_KEPT], sizeof(Kept) * (size_t)W.kept_cap) ||
        ensure(W.mem[W_ROWS], W.bytes[W_ROWS], sizeof(PhxRow) * (size_t)W.row_cap) ||
        ensure(W.mem[W_CNT], W.bytes[W_CNT], 4 * sizeof(unsigned long long)))
        return 1;
    return build_cluster_tiles(T, a.chi2_max, st);
}

// The pass's launches only (no allocation, no host synchronisation): capturable into a hipGraph.
static int fast_launch(FastTables &T, FastWork &W, const FastArgs &a, hipStream_t st, StageMark mark, void *mark_ctx) {
    const int ns = a.ns, npts = a.npts, ntl = T.ntiles;
    unsigned long long *cnt = (unsigned long long *)W.mem[W_CNT];
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, cnt, 4);
    PHX_CK(hipGetLastError());
    double *ps0 = (double *)W.mem[W_PS0], *sig = (double *)W.mem[W_SIG], *trk = (double *)W.mem[W_TRK];
    float4 *trkf = (float4 *)W.mem[W_TRKF];
    float4 *stlo = (float4 *)W.mem[W_STILE], *sthi = stlo + (size_t)ns * ntl;
    if (launch_lin_init_targets(a.stars, a.seeds, ns, a.n_samp, a.P, ps0, sig, st)) return 1;
    if (mark) mark(mark_ctx, "draws");
    if (a.exact_tracks) {   // the plain FP64 tracks: rows identical to the plain form's
        if (launch_lin_plain(ps0, sig, nullptr, trk, a.P, PH_LIN_EPS_POS, PH_LIN_EPS_VEL, -STEP_MYR / PH_LIN_SUB,
                             PH_LIN_SUB, 0, npts, npts, ns, 0, nullptr, st))
            return 1;
        hipLaunchKernelGGL(k_trkf, dim3(blocks_for((long)ns * npts, 256)), dim3(256), 0, st, trk, (long)ns * npts, trkf);
    } else {
        hipLaunchKernelGGL(k_lin_fast, dim3(blocks_for((long)ns * 2, 256)), dim3(256), 0, st, ps0, sig, trk, trkf, a.P,
                           make_potf(a.P), npts, ns);
    }
    PHX_CK(hipGetLastError());
    if (mark) mark(mark_ctx, "tracks");
    hipLaunchKernelGGL(k_star_tiles, dim3(blocks_for((long)ns * ntl, 256)), dim3(256), 0, st, trk, trkf, npts, ntl, ns,
                       (float)sqrt(a.chi2_max), stlo, sthi);
    PHX_CK(hipGetLastError());
    if (mark) mark(mark_ctx, "tiles");
    hipLaunchKernelGGL(k_scan_tiles, dim3(W.scan_blocks > 0 ? W.scan_blocks : 2048), dim3(256), 0, st,
                       (const float4 *)T.mem[T_CTLO], (const float4 *)T.mem[T_CTHI], (const int *)T.mem[T_CNT],
                       (const int *)T.mem[T_CPERM], T.nc, T.nc_pad, stlo, sthi, ns, ntl, (Cand *)W.mem[W_CAND], cnt,
                       W.cand_cap);
    PHX_CK(hipGetLastError());
    if (mark) mark(mark_ctx, "scan");
    hipLaunchKernelGGL(k_refine, dim3(4096), dim3(256), 0, st, (const Cand *)W.mem[W_CAND], cnt, W.cand_cap, trk, trkf,
                       (const double *)T.mem[T_CTRK], (const float4 *)T.mem[T_CTRKF], a.cn, a.R2, npts, a.chi2_max,
                       (Kept *)W.mem[W_KEPT], cnt + 1, W.kept_cap);
    PHX_CK(hipGetLastError());
    if (mark) mark(mark_ctx, "refine");
    const int n50_min = a.p_min > 0 ? (int)ceil(a.p_min * 512 - 1e-9) : 0;
    hipLaunchKernelGGL(k_finish_fast, dim3(4096), dim3(256), 0, st, (const Kept *)W.mem[W_KEPT], cnt + 1, W.kept_cap,
                       trk, (const double *)T.mem[T_CTRK], a.cn, a.r50, a.rt, a.cl, a.prefix, a.group, npts, n50_min,
                       (PhxRow *)W.mem[W_ROWS], cnt + 2, W.row_cap);
    PHX_CK(hipGetLastError());
    if (mark) mark(mark_ctx, "finish");
    return 0;
}

static void graph_drop(FastWork &W) {
    if (W.gexec) (void)hipGraphExecDestroy((hipGraphExec_t)W.gexec);
    if (W.graph) (void)hipGraphDestroy((hipGraph_t)W.graph);
    W.gexec = W.graph = nullptr;
}

int fast_screen(FastTables &T, FastWork &W, const FastArgs &a, hipStream_t st, StageMark mark, void *mark_ctx) {
    if (fast_alloc(T, W, a, st)) return 1;
    if (!W.use_graphs || mark) {   // eager launches (timed stages need their events between kernels)
        return fast_launch(T, W, a, st, mark, mark_ctx);
    }
    // the key of a captured pass: sizes, parameters and every buffer it touches
    uint64_t key = 1469598103934665603ULL;
    auto mix = [&](uint64_t v) { key = (key ^ v) * 1099511628211ULL; };
    mix((uint64_t)a.ns); mix((uint64_t)a.nc); mix((uint64_t)a.npts); mix((uint64_t)a.n_samp);
    mix((uint64_t)a.exact_tracks); mix((uint64_t)W.scan_blocks);
    double dv[2] = {a.chi2_max, a.p_min};
    uint64_t bits[2];
    memcpy(bits, dv, sizeof bits);
    mix(bits[0]); mix(bits[1]);
    const void *ptrs[] = {a.stars, a.seeds, a.prefix, a.group, a.cl, a.cn, a.R2, a.r50, a.rt};
    for (const void *p : ptrs) mix((uint64_t)(uintptr_t)p);
    for (int i = 0; i < 16; i++) { mix((uint64_t)(uintptr_t)W.mem[i]); mix((uint64_t)(uintptr_t)T.mem[i]); }
    mix((uint64_t)W.cand_cap); mix((uint64_t)W.kept_cap); mix((uint64_t)W.row_cap);
    if (!W.gexec || W.gkey != key) {
        graph_drop(W);
        hipGraph_t g = nullptr;
        PHX_CK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = fast_launch(T, W, a, st, nullptr, nullptr);
        const hipError_t ec = hipStreamEndCapture(st, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        PHX_CK(ec);
        hipGraphExec_t ex = nullptr;
        PHX_CK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        W.graph = g;
        W.gexec = ex;
        W.gkey = key;
        W.captures++;
    }
    PHX_CK(hipGraphLaunch((hipGraphExec_t)W.gexec, st));
    W.replays++;
    return 0;
}

int fast_counts(FastWork &W, hipStream_t st, long h[2], int *grow) {
    unsigned long long c[3];
    PHX_CK(hipMemcpyAsync(c, W.mem[W_CNT], sizeof c, hipMemcpyDeviceToHost, st));
    PHX_CK(hipStreamSynchronize(st));
    W.last_cand = (long)c[0];
    *grow = 0;
    if ((long)c[0] > W.cand_cap) { W.cand_cap = (long)c[0] + (long)c[0] / 4 + 4096; *grow = 1; }
    if ((long)c[1] > W.kept_cap) { W.kept_cap = (long)c[1] + (long)c[1] / 4 + 4096; *grow = 1; }
    if ((long)c[2] > W.row_cap) { W.row_cap = (long)c[2] + (long)c[2] / 4 + 4096; *grow = 1; }
    h[0] = (long)c[1];
    h[1] = (long)c[2];
    return 0;
}

const void *fast_rows_ptr(const FastWork &W) { return W.mem[W_ROWS]; }
const double *fast_tracks_ptr(const FastWork &W) { return (const double *)W.mem[W_TRK]; }

int fast_z0_check(uint64_t seed, long n, double *worst, hipStream_t st) {
    void *d = nullptr;
    PHX_CK(hipMalloc(&d, 8));
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, (unsigned long long *)d, 1);
    hipLaunchKernelGGL(k_z0_check, dim3(blocks_for(n, 256)), dim3(256), 0, st, seed, n, (unsigned long long *)d);
    PHX_CK(hipGetLastError());
    unsigned long long h = 0;
    PHX_CK(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, st));
    PHX_CK(hipStreamSynchronize(st));
    PHX_CK(hipFree(d));
    memcpy(worst, &h, 8);
    return 0;
}

void fast_release(FastTables &T, FastWork &W) {
    graph_drop(W);
    for (int i = 0; i < 16; i++) {
        if (T.mem[i]) (void)hipFree(T.mem[i]);
        if (W.mem[i]) (void)hipFree(W.mem[i]);
        T.mem[i] = W.mem[i] = nullptr;
        T.bytes[i] = W.bytes[i] = 0;
    }
}

// ------------------------------------------------------------------------------------------- clones
// One lane per clone (reference: src/phygas.c:200-216): both bodies kick-drift-kick with the FP32 force
// evaluated at the FP64 position, FP64 position / velocity accumulators, and the closest approach of each
// step's segment refined in FP64 with the reference's strict-< first-step rule.  Callers lay clones out by
// pair, longest trace-back first, so a wavefront's lanes run the same number of steps.
__global__ __launch_bounds__(256) void k_pair_fast(const double *__restrict__ ps, const int *__restrict__ steps,
                                                   const double *__restrict__ dts, double *__restrict__ res, PotF Q,
                                                   int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double *p = ps + (size_t)k * 12;
    V3 ap = v3(p[0], p[1], p[2]), av = v3(p[3], p[4], p[5]), bp = v3(p[6], p[7], p[8]), bv = v3(p[9], p[10], p[11]);
    F3 fa = force32(Q, (float)ap.x, (float)ap.y, (float)ap.z), fb = force32(Q, (float)bp.x, (float)bp.y, (float)bp.z);
    V3 r0 = ap - bp;
    double best = norm(r0), tbest = 0.0, vbest = norm(av - bv);
    const int ns = steps[k];
    const double dt = dts[k], h = 0.5 * dt;
    for (int i = 1; i <= ns; i++) {
        av = av + v3(fa.x, fa.y, fa.z) * h;
        bv = bv + v3(fb.x, fb.y, fb.z) * h;
        ap = ap + av * dt;
        bp = bp + bv * dt;
        fa = force32(Q, (float)ap.x, (float)ap.y, (float)ap.z);
        fb = force32(Q, (float)bp.x, (float)bp.y, (float)bp.z);
        av = av + v3(fa.x, fa.y, fa.z) * h;
        bv = bv + v3(fb.x, fb.y, fb.z) * h;
        const V3 r1 = ap - bp;
        const double s = seg_param(r0, r1);
        const double dm = norm(r0 + (r1 - r0) * s);
        if (dm < best) {
            best = dm;
            tbest = (i - 1 + s) * -dt;
            vbest = norm(av - bv);
        }
        r0 = r1;
    }
    double *R = res + (size_t)k * 3;
    R[0] = best; R[1] = tbest; R[2] = vbest;
}

int fast_pair_clones(const double *d_ps, const int *d_steps, const double *d_dts, double *d_res, const Pot &P, int n,
                     hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_pair_fast, dim3(blocks_for(n, 256)), dim3(256), 0, st, d_ps, d_steps, d_dts, d_res,
                       make_potf(P), n);
    PHX_CK(hipGetLastError());
    return 0;
}

}  // namespace phx
