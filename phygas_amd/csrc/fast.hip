// Fast form of the screening pass and of the Monte Carlo clones.
//
// Screening, per pass over a resident star batch (ns stars, real and control fused) against the resident
// cluster tables:
//   1. draws         ps0 / Sigma0 per star (the plain FP64 kernel; 64 draws per star)
//   2. k_lin_team    tracks: FP64 nominal orbit, the six neighbour offsets of the finite-difference STM
//                    propagated with the FP32 difference force (force32.h) and FP64 accumulators,
//                    S = A Sigma0 A^T assembled in FP64.  One workgroup per 64 stars: one wavefront integrates
//                    their nominal orbits (once per star) and the others their offsets, two per lane by
//                    default.  Writes the AoS tracks and their tiled copy.
//   3. k_star_recs   per (star, 8-point tile) and per (star, 16-point coarse tile): the tile records of
//                    cull.h (axis line, tube radius and Loewner covariance bounds in the co-rotating frame:
//                    one bound's largest eigenvalue for level 1, the interpolation between the tile's ends
//                    for level 2)
//   4. k_scan_tiles  persistent, length-bucketed: each wavefront takes 64 clusters (sorted by track length)
//                    x 16 stars, tests every (pair, coarse tile) with the FP32 tube test and the survivors
//                    with the Mahalanobis certificate on covariances interpolated over the tile (cull.h), then
//                    the tiles of the coarse tiles left with the same certificate, and appends the pairs with
//                    any tile left, with a bit mask of those tiles
//   5. k_expand      one item per (candidate, tile left), contiguous per candidate
//      k_items       eight lanes per item: the reference's FP64 evaluation of the tile's points (two
//                    Mahalanobis forms per point, strict-< first-index minimum)
//      k_reduce      eight lanes per candidate: minimum over its items; pairs with chi2 < chi2_max go to the
//                    kept list (64 segments, packed by k_kept_compact)
//   6. finish        k_fin_setup (frame and draw screen per kept pair), k_finish_fast (the 512 Box-Muller
//                    draws of each pair's seed, screened draw by draw, below), k_fin_draw (the draws left,
//                    one lane each), k_fin_rows / k_rows_rtot: exact counts; rows with p_r50 >= p_min
//
// Exactness of the cut: a tile dropped by the scan holds no point with chi2 < chi2_max (cull.h), so kept
// pairs get the FP64 chi2, kb and sb of a full scan of the same tracks.
//
// Exactness of the finish.  splitmix64's state after i calls is seed + i*gamma, so draw i's six uniforms
// are known by index.  A draw counts only when z = L^-1 (x - m) lies in the image of the counting ball, whose
// projection on normal coordinate k is an interval I_k; z_k = rho cos(.) with rho = sqrt(-2 ln u) from the
// draw's (2k+1)-th uniform u, so on the axis k whose interval lies farthest from 0 a draw whose u exceeds
// exp(-dist(0, I_k)^2 / 2) cannot count, and its other uniforms are never generated.  Otherwise z_k is
// estimated in FP32 (hardware log2 / cos); a draw whose estimate is more than 1e-3 outside I_k cannot count.
// For the rest all three z are estimated the same way, and the draw is counted (or not) from the estimated
// position when it lies more than 1e-3 sum|L| from the counting radius; only draws within that band are
// evaluated exactly (the reference's FP64 arithmetic).  The 1e-3 bound on |z estimate - z| holds with a
// factor ~40 to spare (phx_check_finish_screen measures it, tail draws included).
#include <math.h>
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "cull.h"
#include "fast.h"
#include "force32.h"

namespace phx {

// ------------------------------------------------------------------------------------------- layout
constexpr int SPW = 16;                 // stars per wavefront in the tile scan
constexpr int TREC = 75;                // doubles per object tile of the tiled tracks (k_tiled)

struct Cand { int star, cluster; unsigned long long mask; };
struct Kept { int star, cluster, kb, pad; double sb, chi2; };

// Work buffers (FastWork::mem slots)
enum { W_PS0, W_SIG, W_TRK, W_SREC1, W_CAND, W_KEPT, W_ROWS, W_CNT, W_ITEMS, W_IRES, W_ISTART,
       W_TILED, W_FSCR, W_FDRAW, W_FCNT, W_SURV, W_KSEG, W_SREC1C, W_SRECFC, W_SRECF, W_NSLOT };
static_assert(W_NSLOT <= 24, "FastWork holds 24 buffers");
// Counters (W_CNT): candidates, kept pairs, rows, items, points evaluated, level-2 tests of the scan, draws left
// for the exact evaluation by the finish screen, largest kept segment, level-1 tests of the scan
enum { C_CAND, C_KEPT, C_ROWS, C_ITEMS, C_POINTS, C_L2, C_SURV, C_KSEGMAX, C_L1, C_N };
// The reduce appends kept pairs to KSEG segments of the W_KSEG list, each with its own counter (counter i at
// C_KSEG0 + i KSEG_STRIDE, one per 256 bytes): one counter for every wavefront of the grid would take its
// atomics one after another.  k_kept_compact then packs the segments into W_KEPT.
constexpr int KSEG = 64, KSEG_STRIDE = 32, C_KSEG0 = 32;
// The scan's task queues, one per XCD (counter x at C_TASKQ0 + x KSEG_STRIDE).
constexpr int N_XCD = 8, C_TASKQ0 = C_KSEG0 + KSEG * KSEG_STRIDE, C_TOTAL = C_TASKQ0 + N_XCD * KSEG_STRIDE;
// Table buffers (FastTables::mem slots)
enum { T_CTRK, T_CTILED, T_CREC1, T_CNT, T_CPERM, T_CNS, T_CREC1C, T_CRECFC, T_CRECF, T_ROT, T_NSLOT };
static_assert(T_NSLOT <= 16, "FastTables holds 16 buffers");

static int ensure(void *&p, size_t &have, size_t need) {
    if (need <= have && p) return 0;
    if (p) PHX_CK(hipFree(p));
    p = nullptr;
    have = 0;
    PHX_CK(hipMalloc(&p, need ? need : 8));
    have = need ? need : 8;
    return 0;
}

// Wave-local ordering of LDS traffic between the lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------------------------------- FP64 nominal
// 1 / sqrt(u) and 1 / b in FP64 from the hardware estimates (about 2^-29 relative) and two Newton steps each,
// within ~1 ulp; u > 0, b != 0.
__device__ __forceinline__ double rsq64(double u) {
    double y = __builtin_amdgcn_rsq(u);
    const double h = 0.5 * u;
#pragma unroll
    for (int it = 0; it < 2; it++) {
        const double e = __builtin_fma(-h * y, y, 0.5);   // 0.5 (1 - u y^2)
        y = __builtin_fma(y, e, y);
    }
    return y;
}
__device__ __forceinline__ double rcp64(double b) {
    double y = __builtin_amdgcn_rcp(b);
#pragma unroll
    for (int it = 0; it < 2; it++) {
        const double e = __builtin_fma(-b, y, 1.0);
        y = __builtin_fma(y, e, y);
    }
    return y;
}

// Acceleration with every root and quotient through rsq64 / rcp64 (u^-1.5 as (u^-1/2)^3) and the halo powers
// from one log2 and one exp2; every operation in FP64 (a few ulp from the reference's pow-based evaluation).
// inv_ha = 1 / halo_a.  Every fused multiply-add is written out and nothing else is contracted: a sum of two
// products left to the compiler's contraction is fused on either product depending on the code around the
// call, and kernels that integrate the same orbit must get the same bits.
__device__ __forceinline__ V3 accel_fast64(const Pot &P, double inv_ha, V3 p) {
    const double R2 = __builtin_fma(p.y, p.y, p.x * p.x), r2 = __builtin_fma(p.z, p.z, R2);
    const double iu = rsq64(__builtin_fma(P.bulge_b, P.bulge_b, r2));
    const double fb = -P.bulge_gm * (iu * iu * iu);
    const double zz = __builtin_fma(p.z, p.z, P.disk_b * P.disk_b);
    const double izeta = rsq64(zz);
    const double az = __builtin_fma(zz, izeta, P.disk_a);
    const double iD = rsq64(__builtin_fma(az, az, R2));
    const double fd = -P.disk_gm * (iD * iD * iD);
    const double ir0 = rsq64(r2 > 0.0 ? r2 : 1.0), ir = r2 > 0.0 ? ir0 : 0.0;   // selects, no branch
    const double r = r2 * ir;
    const double irs = r > 1e-12 ? ir : 1e12;
    const double rl = r > 1e-12 ? r : 1e-12;
    const double xh = (rl < P.halo_cut ? rl : P.halo_cut) * inv_ha;
    const double lx = log2(xh);
    const double e1 = exp2((P.halo_gamma - 1.0) * lx);   // xh^(gamma - 1); xh^gamma = xh e1
    const double gm = P.halo_gm * (xh * e1) * rcp64(1.0 + e1);
    const double fh = -gm * (irs * irs * irs);
    const double kxy = fb + fd + fh;
    return v3(kxy * p.x, kxy * p.y, __builtin_fma(fb + fh, p.z, fd * p.z * (az * izeta)));
}

// ------------------------------------------------------------------------------------------- tracks
// One workgroup per 64 stars (lane = star) in 1 + 6 / NO wavefronts with different roles:
//   wavefront 0        the 64 FP64 nominal orbits, once per star; each sub-step's position, rounded to FP32 as
//                      the difference force takes it, goes to an LDS ring; it also assembles S = A Sigma0 A^T
//                      and writes the tracks
//   wavefront w >= 1   offsets NO (w - 1) .. NO w - 1 of the same 64 stars (FP32 difference force, FP64
//                      accumulators), reading the nominal's positions from the ring
// The workgroup advances one point (PH_LIN_SUB sub-steps) per barrier with the nominal one point ahead: in step
// j the nominal integrates point j + 1 and assembles S of point j - 1 from the columns of A that the offsets
// left in LDS in step j - 1, while the offsets integrate point j.  Every wavefront passes the same npts + 1
// barriers.  Every offset's arithmetic is the same whatever NO, so 1, 2 or 3 offsets per lane give the same
// bits; running a star's nominal in a wavefront of its own takes the (1 + 6 / NO)-fold repetition of the FP64
// orbit out of the offsets' lanes (a layout with the nominal on every lane of a star measured 4.5 against
// 3.3 ms on 40,000 stars).  Kick-drift-kick sub-steps with each second half kick merged into the next first
// one: K(h) D K(dt) D ... K(dt) D K(h); one check per sub-step sends the rare out-of-range terms of the
// difference force to its branching form.
template <int NO>
__global__ __launch_bounds__(64 * (1 + 6 / NO)) __attribute__((amdgpu_waves_per_eu(3)))
void k_lin_team(const double *__restrict__ ps0, const double *__restrict__ sig, double *__restrict__ trk, Pot P,
                PotF Q, int npts, int n, double *__restrict__ tiled, int ntl) {
#pragma clang fp contract(fast)
    static_assert(NO == 1 || NO == 2 || NO == 3, "1, 2 or 3 offsets per lane");
    __shared__ float ring[2][PH_LIN_SUB][3][64];   // nominal positions of the point being integrated
    __shared__ double colA[2][18][64];             // A[r][o] = dx_o[r] / eps at entry 3 o + r
    __shared__ double sg_s[36][64];                // Sigma0 of the workgroup's stars
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s = blockIdx.x * 64 + lane;
    const bool live = s < n;
    // lanes past the batch's end integrate the last star's orbit without writing it (n > 0 on every launch)
    const int se = live ? s : n - 1;
    const double eps = PH_LIN_EPS_POS, inv_eps = 1.0 / PH_LIN_EPS_POS, dt = -STEP_MYR / PH_LIN_SUB, h = 0.5 * dt;
    const double *p0 = ps0 + (size_t)se * 6;
    auto barrier = []() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    };
    if (w == 0) {
        const double *sgp = sig + (size_t)se * 36;
#pragma unroll
        for (int i = 0; i < 36; i++) sg_s[i][lane] = sgp[i];
        V3 x = v3(p0[0], p0[1], p0[2]), v = v3(p0[3], p0[4], p0[5]);
        const double inv_ha = 1.0 / P.halo_a;
        V3 a = accel_fast64(P, inv_ha, x);
        // the tracks' position and velocity of point k (and their tiled copy: point k in tile k / 8, and the
        // far end of tile k / 8 - 1; past the track the positions repeat the last one)
        auto put_xv = [&](int k) {
            if (!live) return;
            double *out = trk + ((size_t)s * npts + k) * 12;
            out[0] = x.x; out[1] = x.y; out[2] = x.z; out[3] = v.x; out[4] = v.y; out[5] = v.z;
            const int j = k >> 3, p = k & 7;
            double *tb = tiled + ((size_t)s * ntl + j) * TREC;
            tb[p] = x.x; tb[9 + p] = x.y; tb[18 + p] = x.z;
            if (p == 0 && j > 0) { tb[8 - TREC] = x.x; tb[17 - TREC] = x.y; tb[26 - TREC] = x.z; }
            if (k == npts - 1)
                for (int pp = p + 1; pp <= TILE; pp++) { tb[pp] = x.x; tb[9 + pp] = x.y; tb[18 + pp] = x.z; }
        };
        put_xv(0);
        for (int j = 0; j <= npts; j++) {
            if (j >= 1 && live) {   // S of point j - 1 (past the track the tiled covariances are 0)
                const int k = j - 1, sl = k & 1;
                double A[3][6];
#pragma unroll
                for (int o = 0; o < 6; o++)
#pragma unroll
                    for (int r = 0; r < 3; r++) A[r][o] = colA[sl][3 * o + r][lane];
                // A Sigma0 one column of Sigma0 at a time (the scheduler kept from loading all 36 entries ahead,
                // which would hold 72 more registers); each entry summed over b = 0 .. 5 in order
                double AS[3][6];
#pragma unroll
                for (int c = 0; c < 6; c++) {
                    double g[6];
#pragma unroll
                    for (int b = 0; b < 6; b++) g[b] = sg_s[b * 6 + c][lane];
#pragma unroll
                    for (int r = 0; r < 3; r++) {
                        double acc = 0.0;
#pragma unroll
                        for (int b = 0; b < 6; b++) acc += A[r][b] * g[b];
                        AS[r][c] = acc;
                    }
                    __builtin_amdgcn_sched_barrier(0);
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
#pragma unroll
                for (int i = 0; i < 6; i++) out[6 + i] = S[i];
                const int tj = k >> 3, p = k & 7;
                double *tb = tiled + ((size_t)s * ntl + tj) * TREC;
#pragma unroll
                for (int f = 0; f < 6; f++) tb[27 + f * TILE + p] = S[f];
                if (k == npts - 1)
                    for (int pp = p + 1; pp < TILE; pp++)
#pragma unroll
                        for (int f = 0; f < 6; f++) tb[27 + f * TILE + pp] = 0.0;
            }
            if (j + 1 < npts) {   // point j + 1
                const int sl = (j + 1) & 1;
                for (int it = 0; it < PH_LIN_SUB; it++) {
                    const double kick = it == 0 ? h : dt;
                    v = v + a * kick;
                    x = x + v * dt;
                    a = accel_fast64(P, inv_ha, x);
                    ring[sl][it][0][lane] = (float)x.x;
                    ring[sl][it][1][lane] = (float)x.y;
                    ring[sl][it][2][lane] = (float)x.z;
                }
                v = v + a * h;
                put_xv(j + 1);
            }
            barrier();
        }
    } else {
        const int o0 = NO * (w - 1);
        double dx[NO][3], dv[NO][3];
#pragma unroll
        for (int i = 0; i < NO; i++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const int o = o0 + i;
                dx[i][c] = (o < 3 && c == o) ? eps : 0.0;
                dv[i][c] = (o >= 3 && c == o - 3) ? eps : 0.0;
            }
        F3 da[NO];
        {
            const float xf = (float)p0[0], yf = (float)p0[1], zf = (float)p0[2];
#pragma unroll
            for (int i = 0; i < NO; i++)
                da[i] = dforce32(Q, xf, yf, zf, (float)dx[i][0], (float)dx[i][1], (float)dx[i][2]);
        }
        for (int j = 0; j <= npts; j++) {
            if (j >= 1 && j < npts) {
                const int sl = j & 1;
                for (int it = 0; it < PH_LIN_SUB; it++) {
                    const double kick = it == 0 ? h : dt;
#pragma unroll
                    for (int i = 0; i < NO; i++) {
                        dv[i][0] += (double)da[i].x * kick; dv[i][1] += (double)da[i].y * kick;
                        dv[i][2] += (double)da[i].z * kick;
                    }
#pragma unroll
                    for (int i = 0; i < NO; i++)
#pragma unroll
                        for (int c = 0; c < 3; c++) dx[i][c] += dv[i][c] * dt;
                    const float xf = ring[sl][it][0][lane], yf = ring[sl][it][1][lane], zf = ring[sl][it][2][lane];
                    float tmax = 0.0f, r2max = 0.0f;
#pragma unroll
                    for (int i = 0; i < NO; i++)
                        da[i] = dforce32_t<true>(Q, xf, yf, zf, (float)dx[i][0], (float)dx[i][1], (float)dx[i][2],
                                                 tmax, r2max);
                    if (!(tmax < 1e-2f && r2max < Q.halo_cut2)) {
#pragma unroll
                        for (int i = 0; i < NO; i++)
                            da[i] = dforce32(Q, xf, yf, zf, (float)dx[i][0], (float)dx[i][1], (float)dx[i][2]);
                    }
                }
#pragma unroll
                for (int i = 0; i < NO; i++) {
                    dv[i][0] += (double)da[i].x * h; dv[i][1] += (double)da[i].y * h;
                    dv[i][2] += (double)da[i].z * h;
                }
            }
            if (j < npts) {
                const int sl = j & 1;
#pragma unroll
                for (int i = 0; i < NO; i++)
#pragma unroll
                    for (int r = 0; r < 3; r++) colA[sl][3 * (o0 + i) + r][lane] = dx[i][r] * inv_eps;
            }
            barrier();
        }
    }
}

// ------------------------------------------------------------------------------------------- tile records
// One lane per (star, tile): the level-1 and interpolated level-2 records of the star tracks [ns][npts][12],
// srec1 / srecf [ns_pad][ntl] of the 8-point tiles (Tl = TILE) and the same of the 16-point coarse tiles
// (Tl = TILE2, [ns_pad][ntlc]); rows past the last star hold records no pair test accepts (the scan reads
// whole blocks of SPW stars).
template <int Tl>
__global__ __launch_bounds__(256) void k_star_recs(const double *__restrict__ trk, int npts, int ntl, int n, int n_pad,
                                                   const double *__restrict__ cs, TileRec1 *__restrict__ r1,
                                                   TileRecF *__restrict__ rf) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)n_pad * ntl) return;
    const int s = (int)(id / ntl), j = (int)(id % ntl);
    TileRec1 a = pad_rec1();
    TileRec2 b = {};
    TileRecF f = {};
    if (s < n) tile_records<Tl>(trk + (size_t)s * npts * 12, npts, j, 0.0, a, b, &f, cs);
    r1[id] = a;
    rf[id] = f;
}

// Cluster tables for the fast form: AoS FP64 tracks [c][npts][12].
__global__ __launch_bounds__(256) void k_cluster_aos(const double *__restrict__ csoa, const int *__restrict__ cn,
                                                     int nc, int npts, double *__restrict__ ctrk) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nc * npts) return;
    const int c = (int)(id / npts), k = (int)(id % npts);
    double v[12];
#pragma unroll
    for (int q = 0; q < 12; q++) v[q] = k < cn[c] ? csoa[((size_t)k * 12 + q) * nc + c] : 0.0;
    double *o = ctrk + id * 12;
#pragma unroll
    for (int q = 0; q < 12; q++) o[q] = v[q];
}

// Cluster tile records (Tl = TILE or TILE2 points) in structure-of-arrays [tile][nc_pad] over clusters sorted
// by track length (longest first); the pair's kernel radius enters through the cluster (R^2 I added).
template <int Tl>
__global__ __launch_bounds__(256) void k_cluster_recs(const double *__restrict__ ctrk, const int *__restrict__ cn,
                                                      const double *__restrict__ R2, const int *__restrict__ perm,
                                                      int nc, int nc_pad, int npts, int ntl, const double *__restrict__ rot,
                                                      TileRec1 *__restrict__ r1, TileRecF *__restrict__ rf) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nc_pad * ntl) return;
    const int j = (int)(id / nc_pad), cs = (int)(id % nc_pad);
    TileRec1 a = {};
    TileRec2 b = {};
    TileRecF f = {};
    if (cs < nc) {
        const int c = perm[cs], n = cn[c];
        if (j * Tl < n) tile_records<Tl>(ctrk + (size_t)c * npts * 12, n, j, R2[c], a, b, &f, rot);
    }
    r1[id] = a;
    rf[id] = f;
}

// ------------------------------------------------------------------------------------------- tile scan
// Persistent: each wavefront takes tasks (block of SPW stars x block of 64 clusters sorted by length) until
// none is left, lane = cluster.  The tasks are split over N_XCD queues by cluster block (queue x: the blocks
// cb = x mod N_XCD, star-block-major), and a workgroup starts on the queue of its XCD (workgroups go to the
// XCDs in turn): each XCD's L2 then holds the records of its eighth of the cluster blocks instead of all of
// them, and the star blocks pass through it in order.  A wavefront whose queue is empty takes tasks from the
// next queues, so every queue drains and every wavefront reaches the exit.  The order of the tasks changes
// nothing but the order of the candidate list.  Per coarse tile J (TILE2 = 2 TILE points: tiles 2J, 2J + 1):
//   level 1  every star of the block against the lane's cluster, on the coarse records (the star's record
//            is wave-uniform and read through scalar loads);
//   level 2  the level-1 survivors of the whole wavefront, queued in an LDS ring as (star, cluster, J)
//            entries across the task's coarse tiles and evaluated 64 at a time, one per lane;
//   fine     the coarse tiles left, queued again as (star, cluster, J) in a second LDS ring that collects
//            them across the task's coarse tiles and is evaluated 64 at a time (per coarse tile the fine
//            passes would be 40% full), each entry testing tiles 2J and 2J + 1 with the fine level 2
//            (l2f_pass, covariances interpolated over the tile) on their own records (level 1 would pass
//            nearly all of them).  The verdicts go to the task's tile masks in LDS ([star][cluster]).
// A coarse tile dropped holds no point under the cut, so neither of its tiles can (the fine tests only
// drop more).  Pairs with any tile left are appended with the bit mask of those tiles (MaskT: 32 or 64).
// Without the coarse level, the level-2 records of a wavefront's 64 clusters are staged in LDS field-major
// ([field][lane], the 24 FP32 fields of TileRecF): the lanes' writes are consecutive and reads of different
// clusters fall in different banks, where a [lane] array of 96-byte records would put the lanes' copies of a
// field in few banks.
constexpr int CFF = sizeof(TileRecF) / sizeof(float);
static_assert(CFF == 24, "TileRecF is 24 floats");
__device__ __forceinline__ void lds_put_recf(float (*d)[64], int i, const TileRecF &r) {
    const float *v = reinterpret_cast<const float *>(&r);
#pragma unroll
    for (int u = 0; u < CFF; u++) d[u][i] = v[u];
}
__device__ __forceinline__ TileRecF lds_get_recf(const float (*d)[64], int i) {
    TileRecF r;
    float *v = reinterpret_cast<float *>(&r);
#pragma unroll
    for (int u = 0; u < CFF; u++) v[u] = d[u][i];
    return r;
}

constexpr int L1G = 2;   // star records of the level-1 loop held at once (scalar registers)
// SPWT stars per task (8 or 16; the star records are padded to SPW = 16).  Coarse: the first level on 16-point
// coarse tiles, then the fine tiles of the coarse tiles left (as above); otherwise both levels on the 8-point
// tiles directly (more level-1 tests, no second queue).  The autotuner picks both; rows do not depend on them.
template <typename MaskT, int SPWT, bool Coarse>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_scan_tiles(const TileRec1 *__restrict__ crec1c,
                                                    const TileRecF *__restrict__ crecfc,
                                                    const TileRecF *__restrict__ crecf,
                                                    const TileRec1 *__restrict__ crec1,
                                                    const TileRec1 *__restrict__ srec1,
                                                    const int *__restrict__ ntc_sorted,
                                                    const int *__restrict__ cn_sorted, const int *__restrict__ perm,
                                                    int nc, int nc_pad, const TileRec1 *__restrict__ srec1c,
                                                    const TileRecF *__restrict__ srecfc,
                                                    const TileRecF *__restrict__ srecf, int ns, int ntl, int ntlc,
                                                    float kchi, double chi2c, Cand *__restrict__ cand,
                                                    unsigned long long *__restrict__ ncand, long cap,
                                                    unsigned long long *__restrict__ task_q,
                                                    unsigned long long *__restrict__ n_l1_total,
                                                    unsigned long long *__restrict__ n_l2_total) {
    static_assert(SPWT == 8 || SPWT == 16, "8 or 16 stars per task");
    // the level-1 survivors: per tile without the coarse level; with it, a ring across the task's coarse
    // tiles (at most 63 entries wait while a coarse tile adds at most SPWT x 64)
    constexpr int QN = Coarse ? 2048 : SPWT * 64;
    static_assert(!Coarse || QN >= 63 + SPWT * 64, "coarse ring");
    __shared__ unsigned short q_s[4][QN];
    __shared__ float c2_s[4][CFF][Coarse ? 1 : 64];
    __shared__ unsigned l2_s[4][64];
    __shared__ int nt_s[4][64];
    constexpr int FQ2 = 128;   // fine ring: at most 63 entries wait between the coarse passes, each adding <= 64
    __shared__ unsigned short fq_s[4][Coarse ? FQ2 : 1];
    // the task's tile masks, [star][cluster lane], 32 tiles from tile mbase (64-bit masks: tiles 0-31, then
    // 32-63 once the fine ring has been emptied and the first half moved to the lanes' masks)
    __shared__ unsigned mask_s[4][Coarse ? SPWT : 1][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const float chi2cf = up(chi2c);   // the fine level's FP32 cut, rounded up
    unsigned short *q = q_s[w];
    const int ncb = (nc + 63) / 64, nsb = (ns + SPWT - 1) / SPWT;
    const int xcd = blockIdx.x % N_XCD;
    int qi = 0;   // queues tried: own, then the following ones (wave-uniform)
    // exclusive prefix of a per-lane count over the wavefront; the total into *tot
    auto excl_prefix = [&](unsigned cnt, int *tot) {
        unsigned x = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        *tot = __builtin_amdgcn_readfirstlane(__shfl(x, 63, 64));
        return x - cnt;
    };
    auto enqueue = [&](unsigned bits, unsigned o, int tag) {   // entries tag << 10 | star << 6 | lane
        while (bits) {
            const int i = __builtin_ctz(bits);
            bits &= bits - 1u;
            q[o++ & (QN - 1)] = (unsigned short)(tag << 10 | i << 6 | lane);
        }
    };
    for (;;) {
        int sb = -1, cb = 0;
        while (qi < N_XCD) {   // wave-uniform
            const int qx = (xcd + qi) % N_XCD;
            const int nbq = qx < ncb ? (ncb - qx + N_XCD - 1) / N_XCD : 0;   // cluster blocks of queue qx
            int t = 0;
            if (lane == 0) {
                const unsigned long long v = atomicAdd(task_q + (size_t)qx * KSEG_STRIDE, 1ull);
                t = v < 0x7fffffffull ? (int)v : 0x7fffffff;
            }
            const int task = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
            if (task < nsb * nbq) {   // star-block-major within the queue
                sb = task / nbq;
                cb = qx + N_XCD * (task % nbq);
                break;
            }
            qi++;
        }
        if (sb < 0) break;
        const int cs = cb * 64 + lane;
        const int nt = cs < nc ? ntc_sorted[cs] : 0;
        const int ncl = cs < nc ? cn_sorted[cs] : 0;
        const int ntmax = __builtin_amdgcn_readfirstlane(nt);   // sorted descending: lane 0 has the longest
        const int s0 = sb * SPWT;
        const int nsv = min(SPWT, ns - s0);   // stars of the block (the rest is padding)
        nt_s[w][lane] = nt;
        unsigned n_l1 = 0, n_l2 = 0;   // tests of this task (wave-uniform)
        unsigned fine_l2 = 0;          // fine level-2 tests of this lane's queue entries
        MaskT mask[SPWT];
#pragma unroll
        for (int i = 0; i < SPWT; i++) mask[i] = 0;
        // the fine ring (entries J << 10 | star << 6 | cluster lane; head fh, tail ft, wave-uniform): one pass
        // of up to 64 entries, the verdicts of tiles 2J and 2J + 1 or'ed into mask_s
        unsigned short *fq = fq_s[w];
        int fh = 0, ft = 0, mbase = 0;
        int qh = 0, qt = 0;   // the coarse ring's head and tail (wave-uniform)
        auto fine_pass = [&]() {
            wave_sync();
            const int e = fh + lane;
            if (e < ft) {
                const int ent = fq[e & (FQ2 - 1)], J = ent >> 10, i = (ent >> 6) & 15, c = ent & 63;
                const int ntc = nt_s[w][c], csc = cb * 64 + c;
                unsigned bits = 0u;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int j = 2 * J + h;
                    if (j < ntc) {
                        fine_l2++;
                        if (l2f_pass(srecf[(size_t)(s0 + i) * ntl + j], crecf[(size_t)j * nc_pad + csc], chi2cf))
                            bits |= 1u << (j - mbase);
                    }
                }
                if (bits) atomicOr(&mask_s[w][i][c], bits);
            }
            fh = min(fh + 64, ft);
            wave_sync();
        };
        // one coarse level-2 pass over up to 64 entries of the coarse ring; the coarse tiles left go to the
        // fine ring, which takes a pass of its own whenever 64 wait
        const unsigned long long below = (1ull << lane) - 1ull;
        auto coarse_pass = [&]() {
            wave_sync();
            const int e = qh + lane;
            bool pass = false;
            int ent = 0;
            if (e < qt) {
                ent = q[e & (QN - 1)];
                const int J = ent >> 10, i = (ent >> 6) & 15, c = ent & 63;
                pass = l2f_pass<TILE2>(srecfc[(size_t)(s0 + i) * ntlc + J], crecfc[(size_t)J * nc_pad + cb * 64 + c],
                                       chi2cf);
            }
            const unsigned long long bal = __ballot(pass);
            if (pass) fq[(ft + __popcll(bal & below)) & (FQ2 - 1)] = (unsigned short)ent;
            ft += __popcll(bal);
            qh = min(qh + 64, qt);
            if (ft - fh >= 64) fine_pass();   // wave-uniform
        };
        if constexpr (Coarse) {
#pragma unroll
            for (int i = 0; i < SPWT; i++) mask_s[w][i][lane] = 0;
        }
        if constexpr (!Coarse) {
            float (*c2s)[64] = c2_s[w];
            // one level of 8-point tiles: level 1, then the certificate on the survivors, per tile j
            for (int j = 0; j < ntmax; j++) {
                const bool cv = j < nt;
                const TileRec1 c1 = crec1[(size_t)j * nc_pad + cs];
                const float t1 = cv ? tile_t1<TILE>(ncl, j) : 0.0f;
                n_l1 += (unsigned)nsv * (unsigned)__popcll(__ballot(cv));
                unsigned l1 = 0u;
#pragma unroll
                for (int i0 = 0; i0 < SPWT; i0 += L1G) {
                    TileRec1 sr[L1G];
#pragma unroll
                    for (int u = 0; u < L1G; u++) sr[u] = srec1[(size_t)(s0 + i0 + u) * ntl + j];
#pragma unroll
                    for (int u = 0; u < L1G; u++)
                        if (l1_pass<TILE>(sr[u], c1, t1, kchi)) l1 |= 1u << (i0 + u);
                }
                if (!cv) l1 = 0u;
                int nq = 0;
                const unsigned o1 = excl_prefix((unsigned)__popc(l1), &nq);
                if (nq == 0) continue;
                n_l2 += nq;
                enqueue(l1, o1, 0);
                if (l1) lds_put_recf(c2s, lane, crecf[(size_t)j * nc_pad + cs]);
                l2_s[w][lane] = 0u;
                wave_sync();
                for (int e0 = 0; e0 < nq; e0 += 64) {   // wave-uniform
                    const int e = e0 + lane;
                    if (e < nq) {
                        const int ent = q[e], i = ent >> 6, c = ent & 63;
                        if (l2f_pass(srecf[(size_t)(s0 + i) * ntl + j], lds_get_recf(c2s, c), chi2cf))
                            atomicOr(&l2_s[w][c], 1u << i);
                    }
                }
                wave_sync();
                const unsigned l2 = l2_s[w][lane];
#pragma unroll
                for (int i = 0; i < SPWT; i++) mask[i] |= (MaskT)((l2 >> i) & 1u) << j;
                wave_sync();   // q / c2s / l2_s are rewritten by the next tile
            }
        }
        for (int J = 0; Coarse && 2 * J < ntmax; J++) {
            if (sizeof(MaskT) > 4 && 2 * J == 32) {   // wave-uniform: tiles 32-63 next
                while (qt > qh) coarse_pass();
                while (ft > fh) fine_pass();
#pragma unroll
                for (int i = 0; i < SPWT; i++) {
                    mask[i] = mask_s[w][i][lane];
                    mask_s[w][i][lane] = 0u;
                }
                mbase = 32;
            }
            const bool cv = 2 * J < nt;
            const TileRec1 c1 = crec1c[(size_t)J * nc_pad + cs];
            const float t1 = cv ? tile_t1<TILE2>(ncl, J) : 0.0f;
            n_l1 += (unsigned)nsv * (unsigned)__popcll(__ballot(cv));
            unsigned l1 = 0u;
#pragma unroll
            for (int i0 = 0; i0 < SPWT; i0 += L1G) {   // star records of the padded block: wave-uniform loads
                TileRec1 sr[L1G];
#pragma unroll
                for (int u = 0; u < L1G; u++) sr[u] = srec1c[(size_t)(s0 + i0 + u) * ntlc + J];
#pragma unroll
                for (int u = 0; u < L1G; u++)
                    if (l1_pass<TILE2>(sr[u], c1, t1, kchi)) l1 |= 1u << (i0 + u);
            }
            if (!cv) l1 = 0u;
            // the wavefront's level-1 survivors to the coarse ring
            int nq = 0;
            const unsigned o1 = excl_prefix((unsigned)__popc(l1), &nq);
            if (nq == 0) continue;
            n_l2 += nq;
            enqueue(l1, (unsigned)qt + o1, J);
            qt += nq;
            while (qt - qh >= 64) coarse_pass();   // wave-uniform
        }
        if constexpr (Coarse) {
            while (qt > qh) coarse_pass();   // wave-uniform
            while (ft > fh) fine_pass();
#pragma unroll
            for (int i = 0; i < SPWT; i++) mask[i] |= (MaskT)mask_s[w][i][lane] << (mbase & (8 * sizeof(MaskT) - 1));
        }
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) fine_l2 += __shfl_xor(fine_l2, d, 64);
        n_l2 += fine_l2;
        if (lane == 0) {
            atomicAdd(n_l1_total, (unsigned long long)n_l1);
            if (n_l2) atomicAdd(n_l2_total, (unsigned long long)n_l2);
        }
        // the task's candidates take one range of the list (one atomic per task)
        unsigned total = 0;
#pragma unroll
        for (int i = 0; i < SPWT; i++) total += (unsigned)__popcll(__ballot(cs < nc && s0 + i < ns && mask[i] != 0));
        if (total == 0) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(ncand, (unsigned long long)total);
        base = __shfl(base, 0, 64);
        const int cl = cs < nc ? perm[cs] : 0;
#pragma unroll
        for (int i = 0; i < SPWT; i++) {
            const bool hit = cs < nc && s0 + i < ns && mask[i] != 0;
            const unsigned long long b = __ballot(hit);
            if (hit) {
                const unsigned long long o = base + __popcll(b & below);
                if ((long)o < cap) cand[o] = Cand{s0 + i, cl, (unsigned long long)mask[i]};
            }
            base += __popcll(b);
        }
    }
}

// ------------------------------------------------------------------------------------------- refine
// The reference's FP64 value of point k of a pair (reference: src/phygas.c:395-403): min of the Mahalanobis
// form at the step's start m0 and at the point of the step's segment m0 -> m1 closest to the origin (that
// one only when >= 0 and smaller), -1 when neither form is defined.  *sv: the segment parameter of the
// chosen form (0 for m0).  Ms: S_star + S_cluster (xx xy xz yy yz zz) before the kernel term R^2.
__device__ __forceinline__ double point_chi2(V3 m0, V3 m1, bool has_next, const double *Ms, double R2, double *sv_out) {
    double M[6];
#pragma unroll
    for (int i = 0; i < 6; i++) M[i] = Ms[i];
    M[0] += R2; M[3] += R2; M[5] += R2;
    double qv = mahalanobis(M, m0.x, m0.y, m0.z), sv = 0.0;
    if (has_next) {
        sv = seg_param(m0, m1);
        const V3 ms = m0 + (m1 - m0) * sv;
        const double qs = mahalanobis(M, ms.x, ms.y, ms.z);
        if (qs >= 0 && qs < qv) qv = qs; else sv = 0.0;
    }
    *sv_out = sv;
    return qv;
}

// Tiled tracks for the pair evaluation: per object and 8-point tile j, 75 doubles in structure-of-arrays
// order: x, y, z of the positions 8j .. 8j+8 (9 each: the last is the far end of point 8j+7's segment),
// then the six covariance entries of the points 8j .. 8j+7 (8 each).  Eight lanes reading the same field of
// eight consecutive points read 64 contiguous bytes.  Positions past the track repeat its last position and
// covariances past it are 0 (never read for a pair: its points stop at the cluster's length).
__global__ __launch_bounds__(256) void k_tiled(const double *__restrict__ trk, const int *__restrict__ cn, int n_const,
                                               int nobj, int npts, int ntl, double *__restrict__ out) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nobj * ntl * TREC) return;
    const int i = (int)(id % TREC);
    const long ot = id / TREC;
    const int j = (int)(ot % ntl), o = (int)(ot / ntl);
    const int n = cn ? cn[o] : n_const;
    const double *t = trk + (size_t)o * npts * 12;
    double v;
    if (i < 27) {
        const int f = i / 9, k = j * TILE + i % 9;
        v = t[(size_t)min(k, n - 1) * 12 + f];
    } else {
        const int f = (i - 27) / TILE, k = j * TILE + (i - 27) % TILE;
        v = k < n ? t[(size_t)k * 12 + 6 + f] : 0.0;
    }
    out[id] = v;
}

// Point 8j + p of a pair from the tiled tracks of the star (ts) and the cluster (tc) at tile j: m0, m1 and
// S_s + S_c.
__device__ __forceinline__ void tiled_point(const double *__restrict__ ts, const double *__restrict__ tc, int p, V3 &m0,
                                            V3 &m1, double *M) {
    m0 = v3(ts[p] - tc[p], ts[9 + p] - tc[9 + p], ts[18 + p] - tc[18 + p]);
    m1 = v3(ts[p + 1] - tc[p + 1], ts[10 + p] - tc[10 + p], ts[19 + p] - tc[19 + p]);
#pragma unroll
    for (int f = 0; f < 6; f++) M[f] = ts[27 + f * TILE + p] + tc[27 + f * TILE + p];
}

struct Item { int cand, star, cluster, j; };
struct IRes { double q, sv; int k, pad; };

// Items of the candidates, one per tile left by the scan, contiguous per candidate in increasing tile
// order: istart[c] = first item of candidate c.  Grid-stride over blocks of 256 x EXP_PER candidates (thread t
// takes EXP_PER consecutive ones, so the item list keeps the candidates' order); each block takes its range of the item list with one atomic (few
// enough that the counter is not contended).
constexpr int EXP_PER = 8;
__global__ __launch_bounds__(256) void k_expand(const Cand *__restrict__ cand, const unsigned long long *__restrict__ ncand,
                                                long cand_cap, Item *__restrict__ items,
                                                unsigned long long *__restrict__ nitems, long item_cap,
                                                long *__restrict__ istart, const int *__restrict__ cn, int npts,
                                                unsigned long long *__restrict__ npoints) {
    __shared__ unsigned wsum[4];
    __shared__ unsigned long long base_s;
    const long n = min((long)*ncand, cand_cap);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (long b0 = (long)blockIdx.x * blockDim.x * EXP_PER; b0 < n; b0 += (long)gridDim.x * blockDim.x * EXP_PER) {
        Cand cd[EXP_PER];
        unsigned cnt = 0;
#pragma unroll
        for (int u = 0; u < EXP_PER; u++) {
            const long id = b0 + threadIdx.x * EXP_PER + u;
            cd[u] = id < n ? cand[id] : Cand{0, 0, 0ull};
            cnt += (unsigned)__popcll(cd[u].mask);
        }
        unsigned x = cnt;   // inclusive scan within the wavefront
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        unsigned off = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            off += i < w ? wsum[i] : 0u;
            tot += wsum[i];
        }
        if (threadIdx.x == 0) base_s = atomicAdd(nitems, (unsigned long long)tot);
        __syncthreads();
        unsigned long long o = base_s + off + x - cnt;
        unsigned pts = 0;   // points of the items (the evaluation's work count)
#pragma unroll
        for (int u = 0; u < EXP_PER; u++) {
            const long id = b0 + threadIdx.x * EXP_PER + u;
            if (id < n) {
                istart[id] = (long)o;
                const int np = min(npts, cn[cd[u].cluster]);
                unsigned long long mm = cd[u].mask;
                while (mm) {
                    const int j = __builtin_ctzll(mm);
                    mm &= mm - 1ull;
                    if ((long)o < item_cap) items[o] = Item{(int)id, cd[u].star, cd[u].cluster, j};
                    o++;
                    pts += (unsigned)max(0, min(TILE, np - j * TILE));
                }
            }
        }
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) pts += __shfl_xor(pts, d, 64);
        if (lane == 0 && pts) atomicAdd(npoints, (unsigned long long)pts);
        __syncthreads();   // wsum / base_s reused by the next block of candidates
    }
}

// The lexicographic (q, k) minimum over the 8 lanes of each group (lane >> 3), with its sv, through DPP moves
// (VALU, no LDS round trip): lanes xor 1 and xor 2 within each quad, then each lane with its mirror in the
// 8-lane half row, which lies in the other quad.  Every lane of the wavefront must be active.
template <int CTRL> __device__ __forceinline__ unsigned dpp32(unsigned v) {
    return (unsigned)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}
template <int CTRL> __device__ __forceinline__ double dpp64(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    return __builtin_bit_cast(double, (uint64_t)dpp32<CTRL>((unsigned)u) | (uint64_t)dpp32<CTRL>((unsigned)(u >> 32)) << 32);
}
template <int CTRL> __device__ __forceinline__ void qk_min_dpp(double &q, double &sv, int &k) {
    const double qo = dpp64<CTRL>(q), so = dpp64<CTRL>(sv);
    const int ko = (int)dpp32<CTRL>((unsigned)k);
    if (qo < q || (qo == q && ko < k)) { q = qo; k = ko; sv = so; }
}
__device__ __forceinline__ void group8_min(double &q, double &sv, int &k) {
    qk_min_dpp<0xB1>(q, sv, k);    // quad_perm [1, 0, 3, 2]
    qk_min_dpp<0x4E>(q, sv, k);    // quad_perm [2, 3, 0, 1]
    qk_min_dpp<0x141>(q, sv, k);   // row_half_mirror
}

// Eight lanes per item, one per point of the tile: the reference's FP64 evaluation, then the strict-<
// first-index minimum over the tile (lexicographic (chi2, k) across the eight lanes).  The item list is cut
// into 8 contiguous chunks, chunk x taken by the workgroups the dispatcher places on XCD x (workgroup b on
// XCD b mod 8), so that each XCD works through the items of a range of scan tasks.
__global__ __launch_bounds__(256) void k_items(const Item *__restrict__ items, const unsigned long long *__restrict__ nitems,
                                               long item_cap, const double *__restrict__ stile,
                                               const double *__restrict__ ctile, const int *__restrict__ cn,
                                               const double *__restrict__ R2s, int npts, int ntl,
                                               IRes *__restrict__ res) {
    const long n = min((long)*nitems, item_cap);
    const int x = blockIdx.x & 7, nbx = gridDim.x >> 3;
    const long c0 = n * x / 8, c1 = n * (x + 1) / 8;
    const int p = threadIdx.x & 7;
    const int gpb = blockDim.x >> 3;   // items per workgroup pass
    for (long b0 = c0 + (long)(blockIdx.x >> 3) * gpb; b0 < c1; b0 += (long)nbx * gpb) {
        const long id = b0 + (threadIdx.x >> 3);
        double q = __builtin_inf(), sv = 0.0;
        int k = 0;
        if (id < c1) {
            const Item it = items[id];
            const int np = min(npts, cn[it.cluster]);
            k = it.j * TILE + p;
            if (k < np) {
                V3 m0, m1;
                double M[6];
                tiled_point(stile + ((size_t)it.star * ntl + it.j) * TREC, ctile + ((size_t)it.cluster * ntl + it.j) * TREC,
                            p, m0, m1, M);
                const double v = point_chi2(m0, m1, k + 1 < np, M, R2s[it.cluster], &sv);
                if (v >= 0) q = v;
            }
        }
        group8_min(q, sv, k);
        if (p == 0 && id < c1) res[id] = IRes{q, sv, k, 0};
    }
}

// Eight lanes per candidate: lane p takes its items p, p + 8, ...; the lexicographic (chi2, k) minimum over
// the lanes is the reference's strict-< first-index minimum (items are in tile order and carry the segment
// parameter of their minimum).  Kept pairs wait in a per-wavefront LDS ring and go to the wavefront's
// segment of the kept list 64 at a time (kept: KSEG segments of kept_cap / KSEG; nkept: their counters).
constexpr int KQ = 128;
__global__ __launch_bounds__(256) void k_reduce(const Cand *__restrict__ cand, const unsigned long long *__restrict__ ncand,
                                                long cand_cap, const long *__restrict__ istart,
                                                const IRes *__restrict__ res, const unsigned long long *__restrict__ nitems,
                                                long item_cap, double chi2_max, Kept *__restrict__ kept,
                                                unsigned long long *__restrict__ nkept, long kept_cap) {
    __shared__ Kept ring_s[4][KQ];
    const long n = min((long)*ncand, cand_cap);
    if ((long)*nitems > item_cap) return;   // the pass is run again with a larger item list
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, p = lane & 7, g = lane >> 3;
    Kept *ring = ring_s[w];
    const unsigned long long below = (1ull << lane) - 1ull;
    int head = 0, tail = 0;   // wave-uniform
    const int seg = (int)((((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6) % KSEG);
    const long seg_cap = kept_cap / KSEG;
    Kept *kseg = kept + seg * seg_cap;
    auto flush = [&]() {
        const int nf = min(tail - head, 64);
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(nkept + seg * KSEG_STRIDE, (unsigned long long)nf);
        base = __shfl(base, 0, 64);
        if (lane < nf && (long)(base + lane) < seg_cap) kseg[base + lane] = ring[(head + lane) & (KQ - 1)];
        head += nf;
        wave_sync();
    };
    const long stride = ((long)gridDim.x * blockDim.x) >> 3;
    for (long b0 = ((long)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) >> 3; b0 < n; b0 += stride) {   // wave-uniform
        const long id = b0 + g;
        double best = __builtin_inf(), sb = 0.0;
        int kb = 0x7fffffff;
        Cand cd = {0, 0, 0ull};
        if (id < n) {
            cd = cand[id];
            const long i0 = istart[id];
            const int ni = __popcll(cd.mask);
            for (int i = p; i < ni; i += 8) {
                const IRes q = res[i0 + i];
                if (q.q < best || (q.q == best && q.k < kb)) { best = q.q; kb = q.k; sb = q.sv; }
            }
        }
        group8_min(best, sb, kb);
        const bool hit = id < n && p == 0 && best < chi2_max;
        const unsigned long long bal = __ballot(hit);
        if (hit) ring[(tail + __popcll(bal & below)) & (KQ - 1)] = Kept{cd.star, cd.cluster, kb, 0, sb, best};
        tail += __popcll(bal);
        if (tail - head >= 64) {   // wave-uniform
            wave_sync();
            flush();
        }
    }
    if (tail > head) {   // wave-uniform
        wave_sync();
        flush();
    }
}

// Packs the KSEG segments of the reduce's list into the kept list; nkept[C_KEPT] = the total and
// nkept[C_KSEGMAX] = the largest segment count (a segment over kept_cap / KSEG lost pairs: the pass is run
// again with a larger list).
__global__ __launch_bounds__(256) void k_kept_compact(const Kept *__restrict__ kseg, long kept_cap,
                                                      unsigned long long *__restrict__ cnt, Kept *__restrict__ kept) {
    __shared__ long pre_s[KSEG + 1];
    __shared__ unsigned long long c_s[KSEG];
    const long seg_cap = kept_cap / KSEG;
    if (threadIdx.x < KSEG) c_s[threadIdx.x] = cnt[C_KSEG0 + threadIdx.x * KSEG_STRIDE];
    __syncthreads();
    if (threadIdx.x == 0) {
        long acc = 0;
        unsigned long long mx = 0;
        for (int i = 0; i < KSEG; i++) {
            const unsigned long long c = c_s[i];
            mx = c > mx ? c : mx;
            pre_s[i] = acc;
            acc += (long)min((long)c, seg_cap);
        }
        pre_s[KSEG] = acc;
        if (blockIdx.x == 0) {
            cnt[C_KEPT] = (unsigned long long)acc;
            cnt[C_KSEGMAX] = mx;
        }
    }
    __syncthreads();
    const long total = pre_s[KSEG];
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        int lo = 0, hi = KSEG;   // segment i with pre_s[i] <= e < pre_s[i + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pre_s[mid] <= e) lo = mid; else hi = mid;
        }
        kept[e] = kseg[lo * seg_cap + (e - pre_s[lo])];
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

// Count draws with |m + L z| < r among the pair's 512 (exact), screening with the x0 interval; the eight
// lanes of a group (p = lane mod 8) take draws p, p + 8, ... and the counts are summed over the group.
// u1 is screened as an integer: draw i cannot count when (mix >> 11) > u1max, u1max >= ucut 2^53.
__device__ int count_inside8(uint64_t seed, const double l[6], V3 m, double r, double r_alt, int *n_alt, int p) {
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
    // |z0| <= 9.42 for every representable u1; with l00 = 0, x0 = m.x for every draw
    if ((l00 > 0 && (lo >= 9.5 || hi <= -9.5)) || (l00 <= 0 && !all)) {
        if (n_alt) *n_alt = 0;
        return 0;
    }
    const double a = all ? 0.0 : (lo > 0 ? lo : (hi < 0 ? -hi : 0.0));
    const double ucut = exp(-0.5 * a * a * (1.0 - 1e-9)) * (1.0 + 1e-9);   // u1 above this: |z0| < a
    const uint64_t u1max = ucut >= 1.0 ? ~0ull : (uint64_t)(ucut * 9007199254740992.0) + 1;
    const float flo = (float)lo - 1e-3f, fhi = (float)hi + 1e-3f;
    // |position estimate - position| <= 1e-3 (sum of |L|) when every z estimate is within 1e-3 of z
    const double marg = 1e-3 * (fabs(l00) + fabs(l[1]) + fabs(l[2]) + fabs(l[3]) + fabs(l[4]) + fabs(l[5])) +
                        1e-12 * fmax(r, r_alt);
    int cnt = 0, alt = 0;
    for (int i = p; i < 512; i += 8) {
        const uint64_t b = seed + (uint64_t)(6 * i) * GOLDEN_GAMMA;
        const uint64_t w1 = mix64(b + GOLDEN_GAMMA) >> 11;
        if (w1 > u1max) continue;
        const double u1 = (double)w1 * (1.0 / 9007199254740992.0);
        const double u2 = to_unit(mix64(b + 2 * GOLDEN_GAMMA));
        const float ze0 = z0_estimate(u1 <= 0.0 ? 1.0 / 18446744073709551616.0 : u1, u2);
        if (!all && (ze0 < flo || ze0 > fhi)) continue;
        // FP32 estimate of the whole draw: inside / outside decided unless within the margin of a radius
        const double u3 = to_unit(mix64(b + 3 * GOLDEN_GAMMA)), u4 = to_unit(mix64(b + 4 * GOLDEN_GAMMA));
        const double u5 = to_unit(mix64(b + 5 * GOLDEN_GAMMA)), u6 = to_unit(mix64(b + 6 * GOLDEN_GAMMA));
        {
            const double e0 = ze0;
            const double e1 = z0_estimate(u3 <= 0.0 ? 1.0 / 18446744073709551616.0 : u3, u4);
            const double e2 = z0_estimate(u5 <= 0.0 ? 1.0 / 18446744073709551616.0 : u5, u6);
            const double xe = m.x + l00 * e0, ye = m.y + l[1] * e0 + l[3] * e1,
                         ze = m.z + l[2] * e0 + l[4] * e1 + l[5] * e2;
            const double re = sqrt(xe * xe + ye * ye + ze * ze);
            const bool sure_r = re + marg < r || re - marg >= r, sure_a = re + marg < r_alt || re - marg >= r_alt;
            if (sure_r && sure_a) {
                cnt += re + marg < r;
                alt += re + marg < r_alt;
                continue;
            }
        }
        const double z0 = box_muller(u1, u2);
        const double z1 = box_muller(u3, u4);
        const double z2 = box_muller(u5, u6);
        const double x = m.x + l00 * z0, y = m.y + l[1] * z0 + l[3] * z1, z = m.z + l[2] * z0 + l[4] * z1 + l[5] * z2;
        const double rr = sqrt(x * x + y * y + z * z);
        cnt += rr < r;
        alt += rr < r_alt;
    }
#pragma unroll
    for (int d = 1; d < 8; d <<= 1) {
        cnt += __shfl_xor(cnt, d, 64);
        alt += __shfl_xor(alt, d, 64);
    }
    if (n_alt) *n_alt = alt;
    return cnt;
}

// Per kept pair: the draw screen (FinScr) and the frame of the exact evaluation of a draw (FinDraw).
struct FinScr {
    uint64_t seed, u1max;
    float flo, fhi;
    int flags;   // bit 0: every draw passes the screen, bit 1: no draw can count, bits 2-3: the axis k
    int pad;
};
struct FinDraw {
    uint64_t seed;
    double mx, my, mz, l[6], r, ra, marg;
};

// Screens for one radius pair (r, ra <= r).  A draw counts only when z = L^-1 (x - m) lies in the ellipsoid
// E = c + L^-1 B(0, r), c = -L^-1 m, so its normal coordinate z_k lies in the projection of E on axis k,
// I_k = c_k +- r |row k of L^-1|.  z_k = rho_k cos(.) with rho_k = sqrt(-2 ln u) from the draw's (2k+1)-th
// uniform u: a draw whose u exceeds exp(-dist(0, I_k)^2 / 2) cannot count.  The axis with the farthest
// interval screens (one hash per draw for most draws), then the FP32 estimate of z_k against I_k.  With a
// singular L only the x axis is used (I_0 from |m.x + l00 z0| < r).  none = 1: no draw can count.
__device__ void fin_setup(FinScr &q, FinDraw &f, uint64_t seed, const double L[6], V3 m, double r, double ra) {
    f.seed = seed;
    q.seed = seed;
    f.mx = m.x; f.my = m.y; f.mz = m.z;
#pragma unroll
    for (int i = 0; i < 6; i++) f.l[i] = L[i];
    f.r = r; f.ra = ra;
    f.marg = 1e-3 * (fabs(L[0]) + fabs(L[1]) + fabs(L[2]) + fabs(L[3]) + fabs(L[4]) + fabs(L[5])) + 1e-12 * r;
    double lo, hi;
    bool all = false;
    int axis = 0;
    if (L[0] > 0 && L[3] > 0 && L[5] > 0) {
        // L^-1 (lower triangular) and the three projections
        const double i00 = 1.0 / L[0], i11 = 1.0 / L[3], i22 = 1.0 / L[5];
        const double i10 = -i11 * L[1] * i00, i21 = -i22 * L[4] * i11, i20 = -(i21 * L[1] + i22 * L[2]) * i00;
        const double c[3] = {-i00 * m.x, -(i10 * m.x + i11 * m.y), -(i20 * m.x + i21 * m.y + i22 * m.z)};
        const double h[3] = {r * i00, r * sqrt(i10 * i10 + i11 * i11), r * sqrt(i20 * i20 + i21 * i21 + i22 * i22)};
        // row magnitudes (including the terms i20 is summed from) times the size of m and of the computed
        // positions bound the rounding of c_k, h_k and of x = m + L z far below this padding
        const double mabs = fabs(m.x) + fabs(m.y) + fabs(m.z) + r +
                            10.0 * (fabs(L[0]) + fabs(L[1]) + fabs(L[2]) + fabs(L[3]) + fabs(L[4]) + fabs(L[5]));
        const double rowa[3] = {i00, fabs(i10) + i11,
                                (fabs(i21 * L[1]) + fabs(i22 * L[2])) * i00 + fabs(i21) + i22};
        double best = -1.0;
        for (int k = 0; k < 3; k++) {
            const double pad = 1e-9 * (fabs(c[k]) + h[k] + rowa[k] * mabs);
            const double lk = c[k] - h[k] - pad, hk = c[k] + h[k] + pad;
            const double dk = lk > 0 ? lk : (hk < 0 ? -hk : 0.0);
            if (dk > best) { best = dk; axis = k; lo = lk; hi = hk; }
        }
    } else if (L[0] > 0) {
        lo = (-r - m.x) / L[0];
        hi = (r - m.x) / L[0];
    } else {
        all = fabs(m.x) < r;
        lo = -1e300;
        hi = 1e300;
    }
    // |z_k| <= 9.42 for every representable u; with l00 = 0, x0 = m.x for every draw
    const bool none = (!all && (lo >= 9.5 || hi <= -9.5)) || (L[0] <= 0 && !all);
    q.flags = (all ? 1 : 0) | (none ? 2 : 0) | axis << 2;
    q.pad = 0;
    const double a = all ? 0.0 : (lo > 0 ? lo : (hi < 0 ? -hi : 0.0));
    const double ucut = exp(-0.5 * a * a * (1.0 - 1e-9)) * (1.0 + 1e-9);   // u above this: |z_k| < a
    q.u1max = ucut >= 1.0 ? ~0ull : (uint64_t)(ucut * 9007199254740992.0) + 1;
    q.flo = (float)lo - 1e-3f;
    q.fhi = (float)hi + 1e-3f;
}

// Draw i of a pair that passed the axis screen: bit 0 when it counts for r, bit 1 for ra.
__device__ int fin_draw(const FinDraw &f, int i) {
    const uint64_t b = f.seed + (uint64_t)(6 * i) * GOLDEN_GAMMA;
    const double u1 = to_unit(mix64(b + GOLDEN_GAMMA)), u2 = to_unit(mix64(b + 2 * GOLDEN_GAMMA));
    const double u3 = to_unit(mix64(b + 3 * GOLDEN_GAMMA)), u4 = to_unit(mix64(b + 4 * GOLDEN_GAMMA));
    const double u5 = to_unit(mix64(b + 5 * GOLDEN_GAMMA)), u6 = to_unit(mix64(b + 6 * GOLDEN_GAMMA));
    {   // FP32 estimate of the whole draw, decided unless within the margin of a radius
        const double e0 = z0_estimate(u1 <= 0.0 ? 1.0 / 18446744073709551616.0 : u1, u2);
        const double e1 = z0_estimate(u3 <= 0.0 ? 1.0 / 18446744073709551616.0 : u3, u4);
        const double e2 = z0_estimate(u5 <= 0.0 ? 1.0 / 18446744073709551616.0 : u5, u6);
        const double xe = f.mx + f.l[0] * e0, ye = f.my + f.l[1] * e0 + f.l[3] * e1,
                     ze = f.mz + f.l[2] * e0 + f.l[4] * e1 + f.l[5] * e2;
        const double re = sqrt(xe * xe + ye * ye + ze * ze);
        const bool sure_r = re + f.marg < f.r || re - f.marg >= f.r, sure_a = re + f.marg < f.ra || re - f.marg >= f.ra;
        if (sure_r && sure_a) return (re + f.marg < f.r ? 1 : 0) | (re + f.marg < f.ra ? 2 : 0);
    }
    const double z0 = box_muller(u1, u2), z1 = box_muller(u3, u4), z2 = box_muller(u5, u6);
    const double x = f.mx + f.l[0] * z0, y = f.my + f.l[1] * z0 + f.l[3] * z1,
                 z = f.mz + f.l[2] * z0 + f.l[4] * z1 + f.l[5] * z2;
    const double rr = sqrt(x * x + y * y + z * z);
    return (rr < f.r ? 1 : 0) | (rr < f.ra ? 2 : 0);
}

// fnv_feed(h, name, 40) of a cluster name (the first field of PhCluster, so 8-byte aligned): the 40 bytes
// through five 8-byte loads instead of a chain of byte loads, the bytes after the first NUL left out.
static_assert(offsetof(PhCluster, name) == 0 && sizeof(PhCluster) % 8 == 0, "cluster names are 8-byte aligned");
__device__ __forceinline__ uint64_t fnv_name40(uint64_t h, const char *__restrict__ name) {
    const uint64_t *w = reinterpret_cast<const uint64_t *>(name);
    uint64_t v[5];
#pragma unroll
    for (int i = 0; i < 5; i++) v[i] = w[i];
    bool live = true;
#pragma unroll
    for (int i = 0; i < 40; i++) {
        const uint64_t ch = (v[i >> 3] >> (8 * (i & 7))) & 0xffull;   // little-endian byte i
        live = live && ch != 0;
        h = live ? (h ^ ch) * FNV_PRIME : h;
    }
    return h;
}

// Frame of the draws of a kept pair (star s, cluster c, point kb, segment parameter sb): the mean m (the
// point of the step's segment), the covariance C = S_s + S_c at kb and its Cholesky factor L
// (l00 l10 l20 l11 l21 l22), the seed of the pair's generator.  One function for every kernel that needs
// it, so all see the same bits.
__device__ __forceinline__ void pair_frame(const double *__restrict__ strk, const double *__restrict__ ctrk,
                                           const int *__restrict__ cn, const PhCluster *__restrict__ cl,
                                           const uint64_t *__restrict__ sprefix, int npts, int s, int c, int kb,
                                           double sb, V3 &m, double C[6], double L[6], uint64_t &seed) {
    const int n = min(npts, cn[c]);
    const int k1 = kb + 1 < n ? kb + 1 : kb;
    const double *ps = strk + (size_t)s * npts * 12, *pc = ctrk + (size_t)c * npts * 12;
    const V3 a = v3(ps[kb * 12] - pc[kb * 12], ps[kb * 12 + 1] - pc[kb * 12 + 1], ps[kb * 12 + 2] - pc[kb * 12 + 2]);
    const V3 b = v3(ps[k1 * 12] - pc[k1 * 12], ps[k1 * 12 + 1] - pc[k1 * 12 + 1], ps[k1 * 12 + 2] - pc[k1 * 12 + 2]);
    m = a + (b - a) * sb;
#pragma unroll
    for (int i = 0; i < 6; i++) C[i] = ps[kb * 12 + 6 + i] + pc[kb * 12 + 6 + i];
    L[0] = sqrt(fmax(C[0], 0.0));
    L[1] = L[0] > 0 ? C[1] / L[0] : 0;
    L[2] = L[0] > 0 ? C[2] / L[0] : 0;
    L[3] = sqrt(fmax(C[3] - L[1] * L[1], 0.0));
    L[4] = L[3] > 0 ? (C[4] - L[2] * L[1]) / L[3] : 0;
    L[5] = sqrt(fmax(C[5] - L[2] * L[2] - L[4] * L[4], 0.0));
    seed = fnv_name40(sprefix[s], cl[c].name);
}

// The finish runs as four kernels so that the draw loop, which is most of its work, keeps few registers
// (and many wavefronts per SIMD):
//   k_fin_setup    one lane per kept pair: its frame, screen and counters (zeroed)
//   k_finish_fast  the draw screen (below); the draws it cannot decide go to one list
//   k_fin_draw     one lane per listed draw: fin_draw, counted with atomics into the pair's counters
//   k_fin_rows     one lane per kept pair: the row when the counts make one (frame rebuilt for rows only)
// and k_rows_rtot takes the r_tot counts of the rows of a p_min pass.
//
// Counting radii: r = max(r50, r_tot) and ra = min(r50, r_tot) (both r50 in the p_min mode).
__global__ __launch_bounds__(256) void k_fin_setup(const Kept *__restrict__ kept,
                                                   const unsigned long long *__restrict__ nkept, long kept_cap,
                                                   const double *__restrict__ strk, const double *__restrict__ ctrk,
                                                   const int *__restrict__ cn, const double *__restrict__ r50s,
                                                   const double *__restrict__ rts, const PhCluster *__restrict__ cl,
                                                   const uint64_t *__restrict__ sprefix, int npts, int n50_min,
                                                   FinScr *__restrict__ fscr, FinDraw *__restrict__ fdraw,
                                                   int2 *__restrict__ fcnt) {
    const long nk = min((long)*nkept, kept_cap);
    for (long h = (long)blockIdx.x * blockDim.x + threadIdx.x; h < nk; h += (long)gridDim.x * blockDim.x) {
        const Kept kp = kept[h];
        V3 m;
        double C[6], L[6];
        uint64_t seed;
        pair_frame(strk, ctrk, cn, cl, sprefix, npts, kp.star, kp.cluster, kp.kb, kp.sb, m, C, L, seed);
        const double r50 = r50s[kp.cluster], rt = rts[kp.cluster];
        const double rbig = n50_min > 0 ? r50 : fmax(rt, r50), rsmall = n50_min > 0 ? r50 : (rt < r50 ? rt : r50);
        FinScr q;
        FinDraw f;
        fin_setup(q, f, seed, L, m, rbig, rsmall);
        fscr[h] = q;
        fdraw[h] = f;
        fcnt[h] = make_int2(0, 0);
    }
}

// The draw screen: one wavefront per 8 kept pairs, lane p of group g takes draws p, p + 8, ... of pair g.
// Stage 1 hashes every draw's screening uniform (u of axis k) and queues the draws that pass the integer
// bound; stage 2 takes 64 queued draws at a time, hashes their second uniform and sends those whose FP32 z_k
// estimate lies in I_k (+- 1e-3) to the list of draws for k_fin_draw.  Both queues are per-wavefront LDS
// rings of FQ entries with wave-uniform heads and tails; the list takes one atomic per 64 draws.  About 13% of
// the draws pass stage 1, spread so that 95% of a wavefront's stage-1 steps queue at least one: a queue entry
// is one 16-byte record (the uniform's 53 bits, the draw's index) written at one address, and a lane's place
// in the queue comes from mbcnt.
constexpr int FQ = 128;
struct FinQ { uint64_t w1; unsigned idx, pad; };
__global__ __launch_bounds__(256) void k_finish_fast(const FinScr *__restrict__ fscr,
                                                     const unsigned long long *__restrict__ nkept, long kept_cap,
                                                     uint64_t *__restrict__ surv, unsigned long long *__restrict__ nsurv,
                                                     long surv_cap) {
    __shared__ FinScr par_s[4][8];
    __shared__ FinQ q1_s[4][FQ];
    __shared__ uint64_t qx_s[4][FQ];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 3, p = lane & 7;
    FinScr *par = par_s[w];
    FinQ *q1 = q1_s[w];
    uint64_t *qx = qx_s[w];
    const long nk = min((long)*nkept, kept_cap);
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwave = ((long)gridDim.x * blockDim.x) >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    // draws for k_fin_draw wait in a second ring (qx, head h2, tail t2) and go to the list 64 at a time
    int h2 = 0, t2 = 0;
    auto flush = [&]() {
        const int nf = min(t2 - h2, 64);
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(nsurv, (unsigned long long)nf);
        base = __shfl(base, 0, 64);
        if (lane < nf && (long)(base + lane) < surv_cap) surv[base + lane] = qx[(h2 + lane) & (FQ - 1)];
        h2 += nf;
        wave_sync();
    };
    for (long h0 = wave * 8; h0 < nk; h0 += nwave * 8) {   // wave-uniform
        const long h = h0 + g;
        const bool live = h < nk;
        FinScr q = fscr[live ? h : h0];
        if (!live) q.flags = 2;
        if (p == 0) par[g] = q;
        const bool none = (q.flags & 2) != 0;
        const uint64_t u1max = q.u1max;
        int h1 = 0, t1 = 0;
        // splitmix64 state before draw i's screening uniform: seed + (6 i + 2 k + 1) gamma, +48 gamma per step
        uint64_t st = q.seed + (uint64_t)(6 * p + 2 * (q.flags >> 2) + 1) * GOLDEN_GAMMA;
        const unsigned gp = (unsigned)(g << 9 | p);   // a queue entry's index: g << 9 | draw (p + 8 t)
        const unsigned long long live_mask = __builtin_amdgcn_ballot_w64(!none);
        wave_sync();
        for (int t = 0; t <= 64; t++, st += 48 * GOLDEN_GAMMA) {
            if (t < 64) {   // stage 1
                const uint64_t w1 = mix64(st) >> 11;
                const bool pass = !none && w1 <= u1max;   // a draw whose z_k lies outside I_k cannot count
                // the lanes of pass, as the comparison's own lane mask (ICmp predicate 37: unsigned <=)
                const unsigned long long bal = __builtin_amdgcn_uicmpl(w1, u1max, 37) & live_mask;
                if (pass) {
                    const unsigned below_n = __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                                       __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
                    FinQ &e = q1[(t1 + (int)below_n) & (FQ - 1)];
                    e.w1 = w1;
                    e.idx = gp + 8u * (unsigned)t;
                }
                t1 += __popcll(bal);
            }
            while (t1 - h1 >= 64 || (t == 64 && t1 > h1)) {   // stage 2, wave-uniform
                wave_sync();
                const int e = h1 + lane;
                bool pass = false;
                uint64_t id = 0;
                if (e < t1) {
                    const FinQ qe = q1[e & (FQ - 1)];
                    const int v = (int)qe.idx, gg = v >> 9, i = v & 511;
                    const FinScr &f = par[gg];
                    const uint64_t b = f.seed + (uint64_t)(6 * i + 2 * (f.flags >> 2) + 2) * GOLDEN_GAMMA;
                    const double u1 = (double)qe.w1 * (1.0 / 9007199254740992.0);
                    const float ze = z0_estimate(u1 <= 0.0 ? 1.0 / 18446744073709551616.0 : u1, to_unit(mix64(b)));
                    pass = (f.flags & 1) || (ze >= f.flo && ze <= f.fhi);
                    id = (uint64_t)(h0 + gg) << 9 | (uint64_t)i;
                }
                const unsigned long long bal = __ballot(pass);
                if (pass) qx[(t2 + __popcll(bal & below)) & (FQ - 1)] = id;
                t2 += __popcll(bal);
                h1 = min(h1 + 64, t1);
                wave_sync();
                if (t2 - h2 >= 64) flush();   // wave-uniform
            }
        }
        wave_sync();   // par is rewritten by the next batch
    }
    if (t2 > h2) flush();   // wave-uniform
}

// One lane per listed draw (pair h = id >> 9, draw id & 511): the whole-draw estimate, else the exact draw.
__global__ __launch_bounds__(256) void k_fin_draw(const uint64_t *__restrict__ surv,
                                                  const unsigned long long *__restrict__ nsurv, long surv_cap,
                                                  const FinDraw *__restrict__ fdraw, int2 *__restrict__ fcnt) {
    const long n = (long)*nsurv;
    if (n > surv_cap) return;   // the pass is run again with a longer list
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const uint64_t id = surv[e];
        const long h = (long)(id >> 9);
        const int r = fin_draw(fdraw[h], (int)(id & 511));
        if (r & 1) atomicAdd(&fcnt[h].x, 1);
        if (r & 2) atomicAdd(&fcnt[h].y, 1);
    }
}

// One lane per kept pair: its row when the counts make one (every pair without p_min; in50 >= n50_min with it,
// the r_tot count then left to k_rows_rtot).
__global__ __launch_bounds__(256) void k_fin_rows(const Kept *__restrict__ kept,
                                                  const unsigned long long *__restrict__ nkept, long kept_cap,
                                                  const int2 *__restrict__ fcnt, const double *__restrict__ strk,
                                                  const double *__restrict__ ctrk, const int *__restrict__ cn,
                                                  const double *__restrict__ r50s, const double *__restrict__ rts,
                                                  const PhCluster *__restrict__ cl,
                                                  const uint64_t *__restrict__ sprefix,
                                                  const int *__restrict__ sgroup, int npts, int n50_min,
                                                  PhxRow *__restrict__ rows, unsigned long long *__restrict__ nrows,
                                                  long row_cap) {
    const long nk = min((long)*nkept, kept_cap);
    for (long h = (long)blockIdx.x * blockDim.x + threadIdx.x; h < nk; h += (long)gridDim.x * blockDim.x) {
        const int2 k = fcnt[h];
        const int cbig = k.x, csmall = k.y;
        if (n50_min > 0 && cbig < n50_min) continue;
        const Kept kp = kept[h];
        const double r50 = r50s[kp.cluster], rt = rts[kp.cluster];
        int in50, intot;
        if (n50_min > 0) { in50 = cbig; intot = 0; }
        else if (rt < r50) { intot = csmall; in50 = cbig; }
        else { intot = cbig; in50 = csmall; }
        V3 m;
        double C[6], L[6];
        uint64_t seed;
        pair_frame(strk, ctrk, cn, cl, sprefix, npts, kp.star, kp.cluster, kp.kb, kp.sb, m, C, L, seed);
        const double *ps = strk + (size_t)kp.star * npts * 12, *pc = ctrk + (size_t)kp.cluster * npts * 12;
        PhxRow o;
        o.star = kp.star; o.cluster = kp.cluster; o.kb = kp.kb; o.flags = sgroup[kp.star];
        o.chi2 = kp.chi2; o.sb = kp.sb;
        o.t = (kp.kb + kp.sb) * STEP_MYR;
        o.d_nom_pc = norm(m) * 1000.0;
        o.sig_pc = sqrt((C[0] + C[3] + C[5]) / 3.0) * 1000.0;
        o.p_r50 = (double)in50 / 512;
        o.p_rtot = (double)intot / 512;
        const V3 dv = v3(ps[kp.kb * 12 + 3] - pc[kp.kb * 12 + 3], ps[kp.kb * 12 + 4] - pc[kp.kb * 12 + 4],
                         ps[kp.kb * 12 + 5] - pc[kp.kb * 12 + 5]);
        o.vrel = norm(dv) * KMS_PER_KPCMYR;
        const unsigned long long ro = atomicAdd(nrows, 1ull);
        if ((long)ro < row_cap) rows[ro] = o;
    }
}

// r_tot counts of the rows of a p_min pass: eight lanes per row (draws p, p + 8, ...), the pair's frame
// rebuilt from the row as the finish built it.
__global__ __launch_bounds__(256) void k_rows_rtot(const double *__restrict__ strk, const double *__restrict__ ctrk,
                                                   const int *__restrict__ cn, const double *__restrict__ rts,
                                                   const PhCluster *__restrict__ cl,
                                                   const uint64_t *__restrict__ sprefix, int npts,
                                                   PhxRow *__restrict__ rows,
                                                   const unsigned long long *__restrict__ nrows, long row_cap) {
    const long n = min((long)*nrows, row_cap);
    const int p = threadIdx.x & 7;
    for (long r0 = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 3; r0 < n;
         r0 += ((long)gridDim.x * blockDim.x) >> 3) {   // group-uniform
        const PhxRow o = rows[r0];
        V3 m;
        double C[6], L[6];
        uint64_t seed;
        pair_frame(strk, ctrk, cn, cl, sprefix, npts, o.star, o.cluster, o.kb, o.sb, m, C, L, seed);
        const double rt = rts[o.cluster];
        const int intot = count_inside8(seed, L, m, rt, rt, nullptr, p);
        if (p == 0) rows[r0].p_rtot = (double)intot / 512;
    }
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
    for (int i = threadIdx.x; i < n; i += blockDim.x) c[i] = 0ull;
}


// ------------------------------------------------------------------------------------------- host side
static int n_tiles(int npts) { return (npts + TILE - 1) / TILE; }
static int n_coarse(int ntl) { return (ntl + 1) / 2; }   // coarse tiles (TILE2 = 2 TILE points)
static int ns_pad(int ns) { return (ns + SPW - 1) / SPW * SPW; }

int fast_build_tables(FastTables &T, const double *d_csoa, const int *d_cn, const double *d_R2, int nc, int npts,
                      hipStream_t st) {
    PHX_REQUIRE(n_tiles(npts) <= 64, "fast forms: tracks longer than 512 points (256 Myr) need the plain form");
    T.nc = nc;
    T.npts = npts;
    T.ntiles = n_tiles(npts);
    {
        int dev = 0, ncu = 0;
        PHX_CK(hipGetDevice(&dev));
        PHX_CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        if (ncu > 0) T.n_cu = ncu;
    }
    const long ntiled = (long)nc * T.ntiles * TREC;
    if (ensure(T.mem[T_CTRK], T.bytes[T_CTRK], sizeof(double) * 12 * (size_t)nc * npts) ||
        ensure(T.mem[T_CTILED], T.bytes[T_CTILED], sizeof(double) * (size_t)ntiled) ||
        ensure(T.mem[T_ROT], T.bytes[T_ROT], sizeof(double) * 2 * (size_t)npts))
        return 1;
    // the tile records' frame rotation per point (the stream is synchronised before the vector goes)
    std::vector<double> rot(2 * (size_t)npts);
    for (int k = 0; k < npts; k++) {
        rot[2 * k] = cos(OMEGA_PT * k);
        rot[2 * k + 1] = sin(OMEGA_PT * k);
    }
    PHX_CK(hipMemcpyAsync(T.mem[T_ROT], rot.data(), sizeof(double) * 2 * (size_t)npts, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_cluster_aos, dim3(blocks_for((long)nc * npts, 256)), dim3(256), 0, st, d_csoa, d_cn, nc, npts,
                       (double *)T.mem[T_CTRK]);
    hipLaunchKernelGGL(k_tiled, dim3(blocks_for(ntiled, 256)), dim3(256), 0, st, (const double *)T.mem[T_CTRK], d_cn,
                       0, nc, npts, T.ntiles, (double *)T.mem[T_CTILED]);
    PHX_CK(hipGetLastError());
    // sort clusters by track length, longest first (length buckets of the tile scan)
    std::vector<int> cn(nc), perm(nc), nt_sorted(nc), cn_sorted(nc);
    PHX_CK(hipMemcpyAsync(cn.data(), d_cn, 4 * (size_t)nc, hipMemcpyDeviceToHost, st));
    PHX_CK(hipStreamSynchronize(st));
    for (int i = 0; i < nc; i++) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return cn[a] > cn[b]; });
    for (int i = 0; i < nc; i++) {
        nt_sorted[i] = n_tiles(cn[perm[i]]);
        cn_sorted[i] = cn[perm[i]];
    }
    T.nc_pad = (nc + 63) / 64 * 64;
    const size_t nrec = (size_t)T.ntiles * T.nc_pad, nrec_c = (size_t)n_coarse(T.ntiles) * T.nc_pad;
    if (ensure(T.mem[T_CNT], T.bytes[T_CNT], 4 * (size_t)T.nc_pad) ||
        ensure(T.mem[T_CPERM], T.bytes[T_CPERM], 4 * (size_t)T.nc_pad) ||
        ensure(T.mem[T_CNS], T.bytes[T_CNS], 4 * (size_t)T.nc_pad) ||
        ensure(T.mem[T_CREC1], T.bytes[T_CREC1], sizeof(TileRec1) * nrec) ||
        ensure(T.mem[T_CREC1C], T.bytes[T_CREC1C], sizeof(TileRec1) * nrec_c) ||
        ensure(T.mem[T_CRECFC], T.bytes[T_CRECFC], sizeof(TileRecF) * nrec_c) ||
        ensure(T.mem[T_CRECF], T.bytes[T_CRECF], sizeof(TileRecF) * nrec))
        return 1;
    nt_sorted.resize(T.nc_pad, 0);
    cn_sorted.resize(T.nc_pad, 0);
    perm.resize(T.nc_pad, 0);
    PHX_CK(hipMemcpyAsync(T.mem[T_CNT], nt_sorted.data(), 4 * (size_t)T.nc_pad, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(T.mem[T_CNS], cn_sorted.data(), 4 * (size_t)T.nc_pad, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(T.mem[T_CPERM], perm.data(), 4 * (size_t)T.nc_pad, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_cluster_recs<TILE>, dim3(blocks_for((long)nrec, 256)), dim3(256), 0, st,
                       (const double *)T.mem[T_CTRK], d_cn, d_R2, (const int *)T.mem[T_CPERM], nc, T.nc_pad, npts,
                       T.ntiles, (const double *)T.mem[T_ROT], (TileRec1 *)T.mem[T_CREC1], (TileRecF *)T.mem[T_CRECF]);
    hipLaunchKernelGGL(k_cluster_recs<TILE2>, dim3(blocks_for((long)nrec_c, 256)), dim3(256), 0, st,
                       (const double *)T.mem[T_CTRK], d_cn, d_R2, (const int *)T.mem[T_CPERM], nc, T.nc_pad, npts,
                       n_coarse(T.ntiles), (const double *)T.mem[T_ROT], (TileRec1 *)T.mem[T_CREC1C],
                       (TileRecF *)T.mem[T_CRECFC]);
    PHX_CK(hipGetLastError());
    PHX_CK(hipStreamSynchronize(st));
    T.d_cn = d_cn;
    T.d_R2 = d_R2;
    return 0;
}

static int fast_alloc(FastTables &T, FastWork &W, const FastArgs &a, hipStream_t st) {
    (void)st;
    const int ns = a.ns, npts = a.npts, ntl = T.ntiles;
    const long npairs = (long)ns * a.nc;
    if (W.cand_cap == 0) W.cand_cap = std::max(1L << 20, npairs / 6);
    if (W.kept_cap == 0) W.kept_cap = std::max(1L << 18, npairs / 16);
    if (W.row_cap == 0) W.row_cap = std::max(1L << 16, W.kept_cap / 8);
    if (W.item_cap == 0) W.item_cap = std::max(1L << 22, npairs);
    if (W.surv_cap == 0) W.surv_cap = std::max(1L << 20, 4 * W.kept_cap);
    if (ensure(W.mem[W_PS0], W.bytes[W_PS0], 48 * (size_t)ns) || ensure(W.mem[W_SIG], W.bytes[W_SIG], 288 * (size_t)ns) ||
        ensure(W.mem[W_TRK], W.bytes[W_TRK], sizeof(double) * 12 * (size_t)ns * npts) ||
        ensure(W.mem[W_TILED], W.bytes[W_TILED], sizeof(double) * TREC * (size_t)ns * ntl) ||
        ensure(W.mem[W_SREC1], W.bytes[W_SREC1], sizeof(TileRec1) * (size_t)ns_pad(ns) * ntl) ||
        ensure(W.mem[W_SREC1C], W.bytes[W_SREC1C], sizeof(TileRec1) * (size_t)ns_pad(ns) * n_coarse(ntl)) ||
        ensure(W.mem[W_SRECFC], W.bytes[W_SRECFC], sizeof(TileRecF) * (size_t)ns_pad(ns) * n_coarse(ntl)) ||
        ensure(W.mem[W_SRECF], W.bytes[W_SRECF], sizeof(TileRecF) * (size_t)ns_pad(ns) * ntl) ||
        ensure(W.mem[W_CAND], W.bytes[W_CAND], sizeof(Cand) * (size_t)W.cand_cap) ||
        ensure(W.mem[W_KEPT], W.bytes[W_KEPT], sizeof(Kept) * (size_t)W.kept_cap) ||
        ensure(W.mem[W_ROWS], W.bytes[W_ROWS], sizeof(PhxRow) * (size_t)W.row_cap) ||
        ensure(W.mem[W_ITEMS], W.bytes[W_ITEMS], sizeof(Item) * (size_t)W.item_cap) ||
        ensure(W.mem[W_IRES], W.bytes[W_IRES], sizeof(IRes) * (size_t)W.item_cap) ||
        ensure(W.mem[W_ISTART], W.bytes[W_ISTART], sizeof(long) * (size_t)W.cand_cap) ||
        ensure(W.mem[W_FSCR], W.bytes[W_FSCR], sizeof(FinScr) * (size_t)W.kept_cap) ||
        ensure(W.mem[W_FDRAW], W.bytes[W_FDRAW], sizeof(FinDraw) * (size_t)W.kept_cap) ||
        ensure(W.mem[W_FCNT], W.bytes[W_FCNT], sizeof(int2) * (size_t)W.kept_cap) ||
        ensure(W.mem[W_SURV], W.bytes[W_SURV], sizeof(uint64_t) * (size_t)W.surv_cap) ||

        ensure(W.mem[W_KSEG], W.bytes[W_KSEG], sizeof(Kept) * (size_t)W.kept_cap) ||
        ensure(W.mem[W_CNT], W.bytes[W_CNT], C_TOTAL * sizeof(unsigned long long)))
        return 1;
    return 0;
}

// The pass's launches only (no allocation, no host synchronisation): capturable into a hipGraph.  The head:
// counters, draws and tracks.
static int fast_launch_head(FastTables &T, FastWork &W, const FastArgs &a, hipStream_t st, StageMark mark,
                            void *mark_ctx) {
    const int ns = a.ns, npts = a.npts, ntl = T.ntiles;
    unsigned long long *cnt = (unsigned long long *)W.mem[W_CNT];
    double *ps0 = (double *)W.mem[W_PS0], *sig = (double *)W.mem[W_SIG], *trk = (double *)W.mem[W_TRK];
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(256), 0, st, cnt, C_TOTAL);
    PHX_CK(hipGetLastError());
    if (a.ps0_in) {
        ps0 = const_cast<double *>(a.ps0_in);
        sig = const_cast<double *>(a.sig_in);
    } else if (launch_lin_init_targets(a.stars, a.seeds, a.nom, ns, a.n_samp, a.P, ps0, sig, st)) {
        return 1;
    }
    if (mark) mark(mark_ctx, "draws");
    if (a.exact_tracks) {   // the plain FP64 tracks: rows identical to the plain form's
        if (launch_lin_plain(ps0, sig, nullptr, trk, a.P, PH_LIN_EPS_POS, PH_LIN_EPS_VEL, -STEP_MYR / PH_LIN_SUB,
                             PH_LIN_SUB, 0, npts, npts, ns, 0, nullptr, st))
            return 1;
        hipLaunchKernelGGL(k_tiled, dim3(blocks_for((long)ns * ntl * TREC, 256)), dim3(256), 0, st, trk, nullptr, npts,
                           ns, npts, ntl, (double *)W.mem[W_TILED]);
    } else {   // writes the tiled copy too
        const int no = a.track_offsets;
        PHX_REQUIRE(no == 1 || no == 2 || no == 3, "the tracks take 1, 2 or 3 offsets per lane");
        auto kern = no == 1 ? k_lin_team<1> : no == 2 ? k_lin_team<2> : k_lin_team<3>;
        if (ns > 0)
            hipLaunchKernelGGL(kern, dim3(blocks_for(ns, 64)), dim3(64 * (1 + 6 / no)), 0, st, ps0, sig, trk, a.P,
                               make_potf(a.P), npts, ns, (double *)W.mem[W_TILED], ntl);
    }
    PHX_CK(hipGetLastError());
    if (mark) mark(mark_ctx, "tracks");
    return 0;
}

// The rest of the pass, after its head on the same stream: records, scan, point evaluation, finish, rows.
static int fast_launch_tail(FastTables &T, FastWork &W, const FastArgs &a, hipStream_t st, StageMark mark,
                            void *mark_ctx) {
    const int ns = a.ns, npts = a.npts, ntl = T.ntiles;
    unsigned long long *cnt = (unsigned long long *)W.mem[W_CNT];
    double *trk = (double *)W.mem[W_TRK];
    TileRec1 *srec1 = (TileRec1 *)W.mem[W_SREC1];
    PHX_REQUIRE(npts <= T.npts, "star tracks longer than the cluster tables");
    const double *rot = (const double *)T.mem[T_ROT];
    hipLaunchKernelGGL(k_star_recs<TILE>, dim3(blocks_for((long)ns_pad(ns) * ntl, 256)), dim3(256), 0, st, trk, npts,
                       ntl, ns, ns_pad(ns), rot, srec1, (TileRecF *)W.mem[W_SRECF]);
    hipLaunchKernelGGL(k_star_recs<TILE2>, dim3(blocks_for((long)ns_pad(ns) * n_coarse(ntl), 256)), dim3(256), 0, st,
                       trk, npts, n_coarse(ntl), ns, ns_pad(ns), rot, (TileRec1 *)W.mem[W_SREC1C],
                       (TileRecF *)W.mem[W_SRECFC]);
    PHX_CK(hipGetLastError());
    if (mark) mark(mark_ctx, "tiles");
    const double chi2c = a.chi2_max * (1.0 + CUT_REL);
    const bool wide = ntl > 32, coarse = W.scan_coarse != 0;
    const int spw = W.scan_spw == 8 ? 8 : 16;
    auto scan = wide ? (coarse ? (spw == 8 ? k_scan_tiles<unsigned long long, 8, true> : k_scan_tiles<unsigned long long, 16, true>)
                               : (spw == 8 ? k_scan_tiles<unsigned long long, 8, false> : k_scan_tiles<unsigned long long, 16, false>))
                     : (coarse ? (spw == 8 ? k_scan_tiles<unsigned, 8, true> : k_scan_tiles<unsigned, 16, true>)
                               : (spw == 8 ? k_scan_tiles<unsigned, 8, false> : k_scan_tiles<unsigned, 16, false>));
    // default grid: 2 blocks (2 wavefronts per SIMD) per CU, below the 3 that fit (VGPRs): the scan's tasks
    // are taken dynamically, so a grid the device holds at once leaves no blocks waiting, and the room left
    // on each SIMD lets the other sub-pass's kernels run beside the scan instead of queueing behind it
    const int scan_grid = W.scan_blocks > 0 ? W.scan_blocks : T.n_cu * 2;
    hipLaunchKernelGGL(scan, dim3(scan_grid), dim3(256), 0, st,
                       (const TileRec1 *)T.mem[T_CREC1C], (const TileRecF *)T.mem[T_CRECFC],
                       (const TileRecF *)T.mem[T_CRECF], (const TileRec1 *)T.mem[T_CREC1], (const TileRec1 *)srec1,
                       (const int *)T.mem[T_CNT], (const int *)T.mem[T_CNS],
                       (const int *)T.mem[T_CPERM], T.nc, T.nc_pad, (const TileRec1 *)W.mem[W_SREC1C],
                       (const TileRecF *)W.mem[W_SRECFC], (const TileRecF *)W.mem[W_SRECF], ns, ntl, n_coarse(ntl),
                       up(sqrt(chi2c)), chi2c,
                       (Cand *)W.mem[W_CAND], cnt + C_CAND, W.cand_cap, cnt + C_TASKQ0, cnt + C_L1, cnt + C_L2);
    PHX_CK(hipGetLastError());
    if (mark) mark(mark_ctx, "scan");
    const Cand *cand = (const Cand *)W.mem[W_CAND];
    hipLaunchKernelGGL(k_expand, dim3(2048), dim3(256), 0, st, cand, cnt + C_CAND, W.cand_cap, (Item *)W.mem[W_ITEMS],
                       cnt + C_ITEMS, W.item_cap, (long *)W.mem[W_ISTART], a.cn, npts, cnt + C_POINTS);
    PHX_CK(hipGetLastError());
    if (mark) mark(mark_ctx, "expand");
    hipLaunchKernelGGL(k_items, dim3(W.items_blocks > 0 ? W.items_blocks : 65536), dim3(256), 0, st,
                       (const Item *)W.mem[W_ITEMS], cnt + C_ITEMS, W.item_cap,
                       (const double *)W.mem[W_TILED], (const double *)T.mem[T_CTILED], a.cn, a.R2, npts, ntl,
                       (IRes *)W.mem[W_IRES]);
    PHX_CK(hipGetLastError());
    if (mark) mark(mark_ctx, "items");
    hipLaunchKernelGGL(k_reduce, dim3(W.reduce_blocks > 0 ? W.reduce_blocks : 1024), dim3(256), 0, st, cand,
                       cnt + C_CAND, W.cand_cap, (const long *)W.mem[W_ISTART], (const IRes *)W.mem[W_IRES],
                       cnt + C_ITEMS, W.item_cap, a.chi2_max, (Kept *)W.mem[W_KSEG], cnt + C_KSEG0, W.kept_cap);
    PHX_CK(hipGetLastError());
    hipLaunchKernelGGL(k_kept_compact, dim3(1024), dim3(256), 0, st, (const Kept *)W.mem[W_KSEG], W.kept_cap, cnt,
                       (Kept *)W.mem[W_KEPT]);
    PHX_CK(hipGetLastError());
    if (mark) mark(mark_ctx, "reduce");
    const int n50_min = a.p_min > 0 ? (int)ceil(a.p_min * 512 - 1e-9) : 0;
    const Kept *kept = (const Kept *)W.mem[W_KEPT];
    FinScr *fscr = (FinScr *)W.mem[W_FSCR];
    FinDraw *fdraw = (FinDraw *)W.mem[W_FDRAW];
    int2 *fcnt = (int2 *)W.mem[W_FCNT];
    const double *ctrk = (const double *)T.mem[T_CTRK];
    const unsigned kgrid = (unsigned)std::min<long>(blocks_for(W.kept_cap, 256), 8192);
    hipLaunchKernelGGL(k_fin_setup, dim3(kgrid), dim3(256), 0, st, kept, cnt + C_KEPT, W.kept_cap, trk, ctrk, a.cn,
                       a.r50, a.rt, a.cl, a.prefix, npts, n50_min, fscr, fdraw, fcnt);
    PHX_CK(hipGetLastError());
    hipLaunchKernelGGL(k_finish_fast, dim3(W.finish_blocks > 0 ? W.finish_blocks : 8192), dim3(256), 0, st,
                       (const FinScr *)fscr, cnt + C_KEPT, W.kept_cap, (uint64_t *)W.mem[W_SURV], cnt + C_SURV,
                       W.surv_cap);
    PHX_CK(hipGetLastError());
    hipLaunchKernelGGL(k_fin_draw, dim3(4096), dim3(256), 0, st, (const uint64_t *)W.mem[W_SURV], cnt + C_SURV,
                       W.surv_cap, (const FinDraw *)fdraw, fcnt);
    PHX_CK(hipGetLastError());
    hipLaunchKernelGGL(k_fin_rows, dim3(kgrid), dim3(256), 0, st, kept, cnt + C_KEPT, W.kept_cap, (const int2 *)fcnt,
                       trk, ctrk, a.cn, a.r50, a.rt, a.cl, a.prefix, a.group, npts, n50_min, (PhxRow *)W.mem[W_ROWS],
                       cnt + C_ROWS, W.row_cap);
    PHX_CK(hipGetLastError());
    if (n50_min > 0) {
        hipLaunchKernelGGL(k_rows_rtot, dim3(256), dim3(256), 0, st, trk, (const double *)T.mem[T_CTRK], a.cn, a.rt,
                           a.cl, a.prefix, npts, (PhxRow *)W.mem[W_ROWS], cnt + C_ROWS, W.row_cap);
        PHX_CK(hipGetLastError());
    }
    if (mark) mark(mark_ctx, "finish");
    return 0;
}

static int fast_launch(FastTables &T, FastWork &W, const FastArgs &a, hipStream_t st, StageMark mark, void *mark_ctx,
                       int part = FAST_ALL) {
    if (part != FAST_TAIL && fast_launch_head(T, W, a, st, mark, mark_ctx)) return 1;
    if (part != FAST_HEAD && fast_launch_tail(T, W, a, st, mark, mark_ctx)) return 1;
    return 0;
}

static void graph_drop(FastWork &W) {
    if (W.gexec) (void)hipGraphExecDestroy((hipGraphExec_t)W.gexec);
    if (W.graph) (void)hipGraphDestroy((hipGraph_t)W.graph);
    W.gexec = W.graph = nullptr;
}

int fast_screen(FastTables &T, FastWork &W, const FastArgs &a, hipStream_t st, StageMark mark, void *mark_ctx,
                int part) {
    if (!W.use_graphs || mark) {   // eager launches (timed stages need their events between kernels)
        if (part != FAST_TAIL && fast_alloc(T, W, a, st)) return 1;
        return fast_launch(T, W, a, st, mark, mark_ctx, part);
    }
    if (part == FAST_TAIL) return 0;   // a captured pass is launched whole with its head
    if (fast_alloc(T, W, a, st)) return 1;
    // the key of a captured pass: sizes, parameters and every buffer it touches
    uint64_t key = 1469598103934665603ULL;
    auto mix = [&](uint64_t v) { key = (key ^ v) * 1099511628211ULL; };
    mix((uint64_t)a.ns); mix((uint64_t)a.nc); mix((uint64_t)a.npts); mix((uint64_t)a.n_samp);
    mix((uint64_t)a.exact_tracks); mix((uint64_t)a.track_offsets); mix((uint64_t)W.scan_blocks); mix((uint64_t)W.items_blocks);
    mix((uint64_t)W.reduce_blocks); mix((uint64_t)W.finish_blocks); mix((uint64_t)W.scan_spw);
    mix((uint64_t)W.scan_coarse);
    double dv[2] = {a.chi2_max, a.p_min};
    uint64_t bits[2];
    memcpy(bits, dv, sizeof bits);
    mix(bits[0]); mix(bits[1]);
    const void *ptrs[] = {a.stars, a.seeds, a.prefix, a.nom, a.group, a.cl, a.cn, a.R2, a.r50, a.rt, a.ps0_in, a.sig_in};
    for (const void *p : ptrs) mix((uint64_t)(uintptr_t)p);
    for (int i = 0; i < 24; i++) mix((uint64_t)(uintptr_t)W.mem[i]);
    for (int i = 0; i < 16; i++) mix((uint64_t)(uintptr_t)T.mem[i]);
    mix((uint64_t)W.cand_cap); mix((uint64_t)W.kept_cap); mix((uint64_t)W.row_cap); mix((uint64_t)W.item_cap);
    mix((uint64_t)W.surv_cap);
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
    unsigned long long c[C_N];
    PHX_CK(hipMemcpyAsync(c, W.mem[W_CNT], sizeof c, hipMemcpyDeviceToHost, st));
    PHX_CK(hipStreamSynchronize(st));
    W.last_cand = (long)c[C_CAND];
    W.last_items = (long)c[C_ITEMS];
    W.last_points = (long)c[C_POINTS];
    W.last_l2 = (long)c[C_L2];
    W.last_l1 = (long)c[C_L1];
    *grow = 0;
    if ((long)c[C_CAND] > W.cand_cap) { W.cand_cap = (long)c[C_CAND] + (long)c[C_CAND] / 4 + 4096; *grow = 1; }
    if ((long)c[C_ITEMS] > W.item_cap) { W.item_cap = (long)c[C_ITEMS] + (long)c[C_ITEMS] / 4 + 4096; *grow = 1; }
    if ((long)c[C_KSEGMAX] > W.kept_cap / KSEG) {   // a segment of the reduce's list overflowed
        W.kept_cap = std::max(W.kept_cap, ((long)c[C_KSEGMAX] + (long)c[C_KSEGMAX] / 4 + 64) * KSEG);
        *grow = 1;
    }
    if ((long)c[C_ROWS] > W.row_cap) { W.row_cap = (long)c[C_ROWS] + (long)c[C_ROWS] / 4 + 4096; *grow = 1; }
    if ((long)c[C_SURV] > W.surv_cap) { W.surv_cap = (long)c[C_SURV] + (long)c[C_SURV] / 4 + 4096; *grow = 1; }
    h[0] = (long)c[C_KEPT];
    h[1] = (long)c[C_ROWS];
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

void fast_release_work(FastWork &W) {
    graph_drop(W);
    for (int i = 0; i < 24; i++) {
        if (W.mem[i]) (void)hipFree(W.mem[i]);
        W.mem[i] = nullptr;
        W.bytes[i] = 0;
    }
}

void fast_release(FastTables &T, FastWork &W) {
    for (int i = 0; i < 16; i++) {
        if (T.mem[i]) (void)hipFree(T.mem[i]);
        T.mem[i] = nullptr;
        T.bytes[i] = 0;
    }
    fast_release_work(W);
}

}  // namespace phx
