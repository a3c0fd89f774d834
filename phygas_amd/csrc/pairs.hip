// Monte Carlo pair search on the device (reference: src/phygas.c:170-249 ph_pair and its parts; the reference's
// GPU app batches clones across pairs the same way, src/main_phygas_boinc.cpp:187-301).
//
// A batch of pairs runs as four launches on one stream, capturable into a hipGraph:
//   1. k_pair_draw    clone start states drawn on the device from each pair's splitmix64 stream (or uploaded
//                     host draws), pair p owning the clone slots [off[p], off[p] + m[p])
//   2. k_pair_steps   every slot's step count and step (0 steps for unused slots)
//   3. k_clones       the traces: one lane per clone (two, target and cluster, for small batches), a persistent
//                     grid whose wavefronts take chunks of 64 (32) clones from one queue in slot order; pairs are
//                     laid out longest trace-back first, so the longest chunks start first and the short ones
//                     fill the tail, and a chunk's lanes (neighbouring clones of one pair or of two pairs of
//                     similar length) run nearly the same number of steps
//   4. reduce         per pair: counts within r50 / rtot and the 5/50/95 % quantiles of distance, time and
//                     speed, exactly as the reference's sort-and-interpolate (LDS bitonic sort up to 4,096
//                     clones, an 8-pass radix selection of the six order statistics above that)
//
// Buffers of a batch are carved from one block (pair_layout): the engine allocates it, the allocation-free
// entry point passes the library's static scratch array instead.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "common.h"
#include "force32.h"
#include "pairs.h"

namespace phx {

// ---------------------------------------------------------------- layout
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

PairLayout pair_layout(char *base, long slots, int npairs, int ntg, int ncl) {
    PairLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + o : nullptr;
        o += align256(bytes ? bytes : 8);
        return (void *)p;
    };
    const size_t n = (size_t)std::max(slots, 1L), np = (size_t)std::max(npairs, 1);
    L.ps = (double *)take(96 * n);
    L.steps = (int *)take(4 * n);
    L.dts = (double *)take(8 * n);
    L.res = (double *)take(24 * n);
    L.off = (long *)take(8 * np);
    L.m = (int *)take(4 * np);
    L.psteps = (int *)take(4 * np);
    L.pdt = (double *)take(8 * np);
    L.r50 = (double *)take(8 * np);
    L.rt = (double *)take(8 * np);
    L.ti = (int *)take(4 * np);
    L.ci = (int *)take(4 * np);
    L.seed = (uint64_t *)take(8 * np);
    L.out = (PhResult *)take(sizeof(PhResult) * np);
    L.tg = ntg > 0 ? (PhTarget *)take(sizeof(PhTarget) * (size_t)ntg) : nullptr;
    L.cl = ncl > 0 ? (PhCluster *)take(sizeof(PhCluster) * (size_t)ncl) : nullptr;
    L.next = (unsigned *)take(8);
    L.total = o;
    return L;
}

// ---------------------------------------------------------------- draws
// The clones of a pair are drawn from one splitmix64 stream in clone order (reference: src/phygas.c:182-198):
// target draw then cluster draw, each retried while the parallax / distance is not positive, a clone that
// fails either draw skipped.  splitmix64's state after i draws is seed + i*gamma, so one wavefront per pair
// draws 64 consecutive attempts at once, each assuming that every attempt before it used the regular
// 16 uniforms (8 per body: parallax or distance, pmra, pmdec, RV).  Attempts up to and including the first
// one that used a different count are exact; the next round starts after that attempt's true end state.
// Output order and values are those of the sequential loop.
constexpr uint64_t GAMMA_INV = 0xF1DE83E19937733DULL;   // gamma^-1 mod 2^64

__global__ __launch_bounds__(64) void k_pair_draw(const PhTarget *__restrict__ tg, const PhCluster *__restrict__ cl,
                                                 const int *__restrict__ ti, const int *__restrict__ ci,
                                                 const uint64_t *__restrict__ seeds, const long *__restrict__ off,
                                                 int n_clones, Pot P, double *__restrict__ ps, int *__restrict__ m_out) {
    const int p = blockIdx.x, lane = threadIdx.x;
    const PhTarget T = tg[ti[p]];
    const PhCluster Cc = cl[ci[p]];
    double *out = ps + (size_t)off[p] * 12;
    uint64_t base = seeds[p];
    int attempt = 0, made = 0;
    while (attempt < n_clones) {
        const int k = attempt + lane;
        uint64_t st = base + (uint64_t)(16 * lane) * GOLDEN_GAMMA;
        const uint64_t st0 = st;
        V3 ap, av, bp, bv;
        int ok = 0;
        if (k < n_clones) ok = (sample_target(P, T, &st, ap, av) == 0 && sample_cluster(P, Cc, &st, bp, bv) == 0);
        const uint64_t used = (st - st0) * GAMMA_INV;
        const bool irregular = k < n_clones && used != 16;
        const unsigned long long irr = __ballot(irregular);
        const int last = irr ? __builtin_ctzll(irr) : 63;          // lanes 0..last are exact
        const unsigned long long oks = __ballot(ok && lane <= last);
        if (ok && lane <= last) {
            double *o = out + (size_t)(made + __popcll(oks & ((1ull << lane) - 1ull))) * 12;
            o[0] = ap.x; o[1] = ap.y; o[2] = ap.z; o[3] = av.x; o[4] = av.y; o[5] = av.z;
            o[6] = bp.x; o[7] = bp.y; o[8] = bp.z; o[9] = bv.x; o[10] = bv.y; o[11] = bv.z;
        }
        made += __popcll(oks);
        // next round starts after lane `last`'s true end state
        const uint64_t end_last = __shfl(st, last, 64);
        base = irr ? end_last : base + (uint64_t)(16 * 64) * GOLDEN_GAMMA;
        attempt += last + 1;
    }
    if (lane == 0) m_out[p] = made;
}

// Steps and dt of every clone slot from its pair: pair p owns slots [off[p], off[p] + n_clones), the first
// m[p] hold clones, the rest get 0 steps.
__global__ __launch_bounds__(256) void k_pair_steps(const long *__restrict__ off, const int *__restrict__ m,
                                                    const int *__restrict__ psteps, const double *__restrict__ pdt,
                                                    int n_clones, int *__restrict__ steps, double *__restrict__ dts) {
    const int p = blockIdx.x;
    for (int k = threadIdx.x; k < n_clones; k += blockDim.x) {
        steps[off[p] + k] = k < m[p] ? psteps[p] : 0;
        dts[off[p] + k] = pdt[p];
    }
}

__global__ void k_zero_u32(unsigned *p) { *p = 0u; }

// ---------------------------------------------------------------- traces
// One clone per lane (reference: src/phygas.c:200-216): target and cluster orbits stepped together, the
// closest approach on each step's segment, strict-< first minimum.  Fast = 0: the reference's FP64
// kick-drift-kick (physics.h).  Fast = 1: the FP32 force at the FP64 position with FP64 accumulators (each
// kick and drift one fused multiply-add) and the FP64 segment refinement.

// a * s + c per component, one rounding each
__device__ __forceinline__ V3 fma3(V3 a, double s, V3 c) {
    return v3(__builtin_fma(a.x, s, c.x), __builtin_fma(a.y, s, c.y), __builtin_fma(a.z, s, c.z));
}
__device__ __forceinline__ V3 dbl3(F3 f) { return v3(f.x, f.y, f.z); }

template <int Fast>
__device__ __forceinline__ void trace_clone(const double *__restrict__ p, int ns, double dt, const Pot &P,
                                            const PotF &Q, double *__restrict__ R) {
    double best, tbest, vbest;
    if (Fast) {
        V3 ap = v3(p[0], p[1], p[2]), av = v3(p[3], p[4], p[5]), bp = v3(p[6], p[7], p[8]), bv = v3(p[9], p[10], p[11]);
        F3 fa = force32(Q, (float)ap.x, (float)ap.y, (float)ap.z), fb = force32(Q, (float)bp.x, (float)bp.y, (float)bp.z);
        V3 r0 = ap - bp;
        best = norm(r0);
        tbest = 0.0;
        vbest = norm(av - bv);
        const double h = 0.5 * dt;
        for (int i = 1; i <= ns; i++) {
            av = fma3(dbl3(fa), h, av);
            bv = fma3(dbl3(fb), h, bv);
            ap = fma3(av, dt, ap);
            bp = fma3(bv, dt, bp);
            fa = force32(Q, (float)ap.x, (float)ap.y, (float)ap.z);
            fb = force32(Q, (float)bp.x, (float)bp.y, (float)bp.z);
            av = fma3(dbl3(fa), h, av);
            bv = fma3(dbl3(fb), h, bv);
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
    } else {
        Body a = body_start(P, v3(p[0], p[1], p[2]), v3(p[3], p[4], p[5]));
        Body b = body_start(P, v3(p[6], p[7], p[8]), v3(p[9], p[10], p[11]));
        V3 r0 = a.pos - b.pos;
        best = norm(r0);
        tbest = 0.0;
        vbest = norm(a.vel - b.vel);
        for (int i = 1; i <= ns; i++) {
            kdk(P, a, dt);
            kdk(P, b, dt);
            const V3 r1 = a.pos - b.pos;
            const double s = seg_param(r0, r1);
            const double dm = norm(r0 + (r1 - r0) * s);
            if (dm < best) {
                best = dm;
                tbest = (i - 1 + s) * -dt;
                vbest = norm(a.vel - b.vel);
            }
            r0 = r1;
        }
    }
    R[0] = best;
    R[1] = tbest;
    R[2] = vbest;
}

// Two lanes per clone: the even lane steps the target, the odd lane the cluster, with the operations of
// trace_clone for that orbit.  After each step the lanes swap positions (and, at a new minimum, velocities) and
// both evaluate the closest approach from the same values, so both take every branch together and the even
// lane holds trace_clone's bits.  Half the instructions per lane and twice the wavefronts: for FP64 batches too
// small to give every SIMD several wavefronts at one lane per clone.
// The value of the neighbouring lane (lane ^ 1): a DPP quad permutation [1, 0, 3, 2] of each 32-bit half, a VALU
// move with no LDS round trip.  Both lanes of a clone are active wherever it is used.
__device__ __forceinline__ double swap_lane_pair(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, 0xB1, 0xF, 0xF, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), 0xB1, 0xF, 0xF, true);
    return __builtin_bit_cast(double, (uint64_t)lo | (uint64_t)hi << 32);
}

template <int Fast>
__device__ __forceinline__ void trace_clone2(const double *__restrict__ p, int ns, double dt, const Pot &P,
                                             const PotF &Q, int side, double *__restrict__ R) {
    auto other = [](V3 a) { return v3(swap_lane_pair(a.x), swap_lane_pair(a.y), swap_lane_pair(a.z)); };
    auto tmc = [side](V3 own, V3 oth) { return side == 0 ? own - oth : oth - own; };   // target minus cluster
    const double *q = p + 6 * side;
    V3 x = v3(q[0], q[1], q[2]), v = v3(q[3], q[4], q[5]);
    Body b;
    F3 f = {0.0f, 0.0f, 0.0f};
    if (Fast) f = force32(Q, (float)x.x, (float)x.y, (float)x.z);
    else b = body_start(P, x, v);
    V3 r0 = tmc(x, other(x));
    double best = norm(r0), tbest = 0.0, vbest = norm(tmc(v, other(v)));
    const double h = 0.5 * dt;
    for (int i = 1; i <= ns; i++) {
        if (Fast) {
            v = fma3(dbl3(f), h, v);
            x = fma3(v, dt, x);
            f = force32(Q, (float)x.x, (float)x.y, (float)x.z);
            v = fma3(dbl3(f), h, v);
        } else {
            kdk(P, b, dt);
            x = b.pos;
            v = b.vel;
        }
        const V3 r1 = tmc(x, other(x));
        const double s = seg_param(r0, r1);
        const double dm = norm(r0 + (r1 - r0) * s);
        if (dm < best) {   // the same decision in both lanes of the clone
            best = dm;
            tbest = (i - 1 + s) * -dt;
            vbest = norm(tmc(v, other(v)));
        }
        r0 = r1;
    }
    if (side == 0) {
        R[0] = best;
        R[1] = tbest;
        R[2] = vbest;
    }
}

// Persistent grid: each wavefront takes the next bucket of cw chunks (64 / LPC clones each, LPC lanes per
// clone) from *next (zeroed before the launch) and traces them one after another, until the buckets run out;
// every wavefront reaches that exit.  Wider buckets take fewer queue atomics, narrower ones balance the tail
// better (the autotuner picks cw and LPC).
template <int Fast, int LPC>
__global__ __launch_bounds__(256) void k_clones(const double *__restrict__ ps, const int *__restrict__ steps,
                                                const double *__restrict__ dts, double *__restrict__ res, Pot P,
                                                PotF Q, long n, int cw, unsigned *__restrict__ next) {
    static_assert(LPC == 1 || LPC == 2, "one or two lanes per clone");
    constexpr int CPC = 64 / LPC;   // clones per chunk
    const int lane = threadIdx.x & 63;
    const long nchunks = (n + CPC - 1) / CPC, nbuckets = (nchunks + cw - 1) / cw;
    for (;;) {
        unsigned b = 0;
        if (lane == 0) b = atomicAdd(next, 1u);
        b = __shfl(b, 0, 64);
        if ((long)b >= nbuckets) break;
        for (int u = 0; u < cw; u++) {
            const long k = ((long)b * cw + u) * CPC + lane / LPC;   // the same k in both lanes of a clone
            if (k < n) {
                if (LPC == 1) trace_clone<Fast>(ps + (size_t)k * 12, steps[k], dts[k], P, Q, res + (size_t)k * 3);
                else trace_clone2<Fast>(ps + (size_t)k * 12, steps[k], dts[k], P, Q, lane & 1, res + (size_t)k * 3);
            }
        }
    }
}

// ---------------------------------------------------------------- reduction
// One workgroup per pair: the m clone results are converted to pc / Myr / km/s, counted against r50 and
// rtot, and each column is sorted in LDS (bitonic, padded with +inf to a power of two) so the 5/50/95 %
// quantiles interpolate exactly as the reference's qsort + quantile (reference: src/phygas.c:170-180,
// 218-238).
__device__ void bitonic_sort_lds(double *v, int n2) {
    for (int k = 2; k <= n2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                int l = i ^ j;
                if (l > i) {
                    bool up = (i & k) == 0;
                    double a = v[i], b = v[l];
                    if ((a > b) == up) { v[i] = b; v[l] = a; }
                }
            }
            __syncthreads();
        }
}

__device__ double quantile_sorted(const double *v, int n, double q) {
    double pos = q * (n - 1);
    int i = (int)pos;
    if (i >= n - 1) return v[n - 1];
    return v[i] + (pos - i) * (v[i + 1] - v[i]);
}

__device__ __forceinline__ double column_value(const double *R, int k, int col) {
    double x = R[k * 3 + col];
    if (col == 0) x = x * 1000.0;
    else if (col == 2) x = x * KMS_PER_KPCMYR;
    return x;
}

__device__ void count_radii(const double *R, int m, double r50, double rt, int *cnt, PhResult &o) {
    int a = 0, b = 0;
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        double d = R[k * 3] * 1000.0;
        a += d < r50;
        b += d < rt;
    }
    atomicAdd(&cnt[0], a);
    atomicAdd(&cnt[1], b);
    __syncthreads();
    if (m > 0) {
        o.f_r50 = (double)cnt[0] / m;
        o.f_rtot = (double)cnt[1] / m;
    }
}

__device__ void set_quantiles(PhResult &o, int col, double q05, double q50, double q95) {
    if (col == 0) { o.d05 = q05; o.d50 = q50; o.d95 = q95; }
    else if (col == 1) { o.t05 = q05; o.t50 = q50; o.t95 = q95; }
    else { o.v05 = q05; o.v50 = q50; o.v95 = q95; }
}

__global__ __launch_bounds__(256) void k_pair_reduce(const double *__restrict__ res, const long *__restrict__ off,
                                                     const int *__restrict__ m_arr, const double *__restrict__ r50_pc,
                                                     const double *__restrict__ rtot_pc, PhResult *__restrict__ out) {
    __shared__ double v[PAIR_RED_LDS];
    __shared__ int cnt[2];
    const int p = blockIdx.x, m = m_arr[p];
    const double *R = res + (size_t)off[p] * 3;
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
    __syncthreads();
    PhResult o;
    memset(&o, 0, sizeof o);
    o.n = m;
    count_radii(R, m, r50_pc[p], rtot_pc[p], cnt, o);
    int n2 = 1;
    while (n2 < m) n2 <<= 1;
    for (int col = 0; col < 3; col++) {
        __syncthreads();
        for (int k = threadIdx.x; k < n2; k += blockDim.x) v[k] = k < m ? column_value(R, k, col) : __builtin_inf();
        __syncthreads();
        bitonic_sort_lds(v, n2);
        if (threadIdx.x == 0 && m > 0)
            set_quantiles(o, col, quantile_sorted(v, m, 0.05), quantile_sorted(v, m, 0.5), quantile_sorted(v, m, 0.95));
    }
    __syncthreads();
    if (threadIdx.x == 0) out[p] = o;
}

// Order-preserving map of doubles to unsigned integers (negative values reversed below the positive ones).
__device__ __forceinline__ uint64_t order_key(double x) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}
__device__ __forceinline__ double from_order_key(uint64_t k) {
    return __builtin_bit_cast(double, (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFULL) : ~k);
}

// Any number of clones: the six order statistics the three quantiles interpolate between (ranks i and i + 1
// of q (m - 1), q = 0.05, 0.5, 0.95) selected digit by digit, 8 bits per pass over the column, with one
// 256-bin histogram per rank in LDS.  The values are the ones a sort would put at those ranks, so the
// quantiles are the reference's.
__global__ __launch_bounds__(256) void k_pair_reduce_select(const double *__restrict__ res, const long *__restrict__ off,
                                                            const int *__restrict__ m_arr,
                                                            const double *__restrict__ r50_pc,
                                                            const double *__restrict__ rtot_pc,
                                                            PhResult *__restrict__ out) {
    __shared__ unsigned hist[6][256];
    __shared__ uint64_t prefix[6];
    __shared__ long rank[6];
    __shared__ int cnt[2];
    const int p = blockIdx.x, m = m_arr[p];
    const double *R = res + (size_t)off[p] * 3;
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
    __syncthreads();
    PhResult o;
    memset(&o, 0, sizeof o);
    o.n = m;
    count_radii(R, m, r50_pc[p], rtot_pc[p], cnt, o);
    const double qs[3] = {0.05, 0.5, 0.95};
    for (int col = 0; col < 3 && m > 0; col++) {
        __syncthreads();
        if (threadIdx.x < 6) {
            const double pos = qs[threadIdx.x >> 1] * (m - 1);
            const long i = (long)pos;
            rank[threadIdx.x] = std::min(i + (threadIdx.x & 1), (long)m - 1);
            prefix[threadIdx.x] = 0;
        }
        for (int shift = 56; shift >= 0; shift -= 8) {
            const uint64_t hi_mask = shift == 56 ? 0ULL : ~0ULL << (shift + 8);
            for (int i = threadIdx.x; i < 6 * 256; i += blockDim.x) hist[i >> 8][i & 255] = 0;
            __syncthreads();
            uint64_t pf[6];
#pragma unroll
            for (int r = 0; r < 6; r++) pf[r] = prefix[r];
            for (int k = threadIdx.x; k < m; k += blockDim.x) {
                const uint64_t key = order_key(column_value(R, k, col));
                const unsigned dig = (unsigned)(key >> shift) & 255u;
#pragma unroll
                for (int r = 0; r < 6; r++)
                    if ((key & hi_mask) == pf[r]) atomicAdd(&hist[r][dig], 1u);
            }
            __syncthreads();
            if (threadIdx.x < 6) {
                const int r = threadIdx.x;
                long want = rank[r], below = 0;
                int d = 0;
                for (; d < 255; d++) {
                    if (below + (long)hist[r][d] > want) break;
                    below += hist[r][d];
                }
                rank[r] = want - below;
                prefix[r] |= (uint64_t)d << shift;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double q[3];
            for (int j = 0; j < 3; j++) {
                const double pos = qs[j] * (m - 1);
                const int i = (int)pos;
                const double a = from_order_key(prefix[2 * j]), b = from_order_key(prefix[2 * j + 1]);
                q[j] = i >= m - 1 ? b : a + (pos - i) * (b - a);
            }
            set_quantiles(o, col, q[0], q[1], q[2]);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) out[p] = o;
}

// ---------------------------------------------------------------- launches
// Lanes per clone when the caller leaves it open (lanes = 0).  A small batch's kernel time is the longest
// clone's chain of steps.  The FP32 form's step is latency-bound: one lane already overlaps the two orbits'
// forces, and two lanes only add the lane exchange to the chain, so it keeps one.  The FP64 step is long
// enough to be issue-bound for a wavefront alone on its SIMD: two lanes halve it, while one lane per clone
// would give the device fewer than two wavefronts per SIMD.
int clone_lanes(int form, long n, int lanes) {
    if (lanes == 1 || lanes == 2) return lanes;
    if (form == PHX_FORM_FAST) return 1;
    static int simds = 0;
    if (!simds) {
        int dev = 0;
        hipDeviceProp_t prop;
        simds = hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess
                    ? 4 * prop.multiProcessorCount
                    : 1024;
    }
    return n < 2L * 64 * simds ? 2 : 1;
}

template <int Fast, int LPC>
static int clone_blocks_per_cu() {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_clones<Fast, LPC>, 256, 0) != hipSuccess) return 0;
    return std::max(b, 1);
}

static int clone_grid(int fast, int lpc, long n, int cw) {
    static int per_cu[2][2] = {{0, 0}, {0, 0}}, cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
        per_cu[0][0] = clone_blocks_per_cu<0, 1>();
        per_cu[0][1] = clone_blocks_per_cu<0, 2>();
        per_cu[1][0] = clone_blocks_per_cu<1, 1>();
        per_cu[1][1] = clone_blocks_per_cu<1, 2>();
        if (!per_cu[0][0] || !per_cu[0][1] || !per_cu[1][0] || !per_cu[1][1]) return 0;
        cus = prop.multiProcessorCount;
    }
    const int cpc = 64 / lpc;
    const long buckets = ((n + cpc - 1) / cpc + cw - 1) / cw, blocks_needed = (buckets + 3) / 4;
    return (int)std::min<long>((long)per_cu[fast][lpc - 1] * cus, std::max(blocks_needed, 1L));
}

int launch_clones(int form, const double *d_ps, const int *d_steps, const double *d_dts, double *d_res, const Pot &P,
                  long n, int cw, int lanes, unsigned *d_next, hipStream_t st) {
    if (n <= 0) return 0;
    PHX_REQUIRE(cw >= 1 && cw <= 64, "clone bucket width must be 1..64 chunks");
    PHX_REQUIRE(lanes >= 0 && lanes <= 2, "lanes per clone must be 0 (chosen), 1 or 2");
    const int fast = form == PHX_FORM_FAST ? 1 : 0, lpc = clone_lanes(form, n, lanes);
    const int grid = clone_grid(fast, lpc, n, cw);
    PHX_REQUIRE(grid > 0, "cannot size the clone grid");
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, st, d_next);
    const PotF Q = make_potf(P);
    auto kern = fast ? (lpc == 2 ? k_clones<1, 2> : k_clones<1, 1>) : (lpc == 2 ? k_clones<0, 2> : k_clones<0, 1>);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, st, d_ps, d_steps, d_dts, d_res, P, Q, n, cw, d_next);
    PHX_CK(hipGetLastError());
    return 0;
}

int launch_pair_reduce(const double *d_res, const long *d_off, const int *d_m, const double *d_r50_pc,
                       const double *d_rtot_pc, int npairs, int max_m, PhResult *d_out, hipStream_t st) {
    if (npairs <= 0) return 0;
    if (max_m <= PAIR_RED_LDS)
        hipLaunchKernelGGL(k_pair_reduce, dim3(npairs), dim3(256), 0, st, d_res, d_off, d_m, d_r50_pc, d_rtot_pc, d_out);
    else
        hipLaunchKernelGGL(k_pair_reduce_select, dim3(npairs), dim3(256), 0, st, d_res, d_off, d_m, d_r50_pc,
                           d_rtot_pc, d_out);
    PHX_CK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- a batch of pairs
void pair_plan(const PhTarget *t, const PhCluster *c, int npairs, const int *ti, const int *ci, int n_clones,
               double dt_myr, PairPlan &plan) {
    plan.npairs = npairs;
    plan.n_clones = n_clones;
    plan.seed.resize(npairs);
    plan.psteps.resize(npairs);
    plan.pdt.resize(npairs);
    plan.r50.resize(npairs);
    plan.rt.resize(npairs);
    plan.off.resize(npairs);
    std::vector<int> order(npairs);
    for (int p = 0; p < npairs; p++) {
        const PhTarget &T = t[ti[p]];
        const PhCluster &C = c[ci[p]];
        plan.seed[p] = phx_seed(T.id, C.name);
        plan.psteps[p] = (int)ceil(C.tmax_myr / dt_myr);   // reference: src/phygas.c:185-187
        plan.pdt[p] = -C.tmax_myr / plan.psteps[p];
        plan.r50[p] = C.r50_pc;
        plan.rt[p] = C.rtot_pc;
        order[p] = p;
    }
    // longest trace-back first: the clone queue hands out the long chunks before the short ones
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return plan.psteps[a] > plan.psteps[b]; });
    for (int r = 0; r < npairs; r++) plan.off[order[r]] = (long)r * n_clones;
    plan.slots = (long)npairs * n_clones;
    plan.max_steps = npairs ? *std::max_element(plan.psteps.begin(), plan.psteps.end()) : 0;
}

int pair_upload_plan(const PairLayout &L, const PairPlan &plan, const int *ti, const int *ci, hipStream_t st) {
    const size_t np = (size_t)plan.npairs;
    PHX_CK(hipMemcpyAsync(L.ti, ti, 4 * np, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(L.ci, ci, 4 * np, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(L.seed, plan.seed.data(), 8 * np, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(L.psteps, plan.psteps.data(), 4 * np, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(L.pdt, plan.pdt.data(), 8 * np, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(L.off, plan.off.data(), 8 * np, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(L.r50, plan.r50.data(), 8 * np, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(L.rt, plan.rt.data(), 8 * np, hipMemcpyHostToDevice, st));
    return 0;
}

int pair_launch_draws(const PairLayout &L, const PairPlan &plan, const PhTarget *d_t, const PhCluster *d_c,
                      const Pot &P, hipStream_t st) {
    const int np = plan.npairs;
    if (np <= 0) return 0;
    hipLaunchKernelGGL(k_pair_draw, dim3(np), dim3(64), 0, st, d_t, d_c, L.ti, L.ci, L.seed, L.off, plan.n_clones, P,
                       L.ps, L.m);
    hipLaunchKernelGGL(k_pair_steps, dim3(np), dim3(256), 0, st, L.off, L.m, L.psteps, L.pdt, plan.n_clones, L.steps,
                       L.dts);
    PHX_CK(hipGetLastError());
    return 0;
}

int pair_launch_device(const PairLayout &L, const PairPlan &plan, int form, const PhTarget *d_t, const PhCluster *d_c,
                       const Pot &P, hipStream_t st, PairMark mark, void *mark_ctx) {
    const int np = plan.npairs;
    if (np <= 0) return 0;
    if (pair_launch_draws(L, plan, d_t, d_c, P, st)) return 1;
    if (mark) mark(mark_ctx, "draws");
    if (launch_clones(form, L.ps, L.steps, L.dts, L.res, P, plan.slots, plan.bucket, plan.lanes, L.next, st))
        return 1;
    if (mark) mark(mark_ctx, "clones");
    if (launch_pair_reduce(L.res, L.off, L.m, L.r50, L.rt, np, plan.n_clones, L.out, st)) return 1;
    if (mark) mark(mark_ctx, "reduce");
    return 0;
}

}  // namespace phx

// ================================================================ allocation-free Monte Carlo batch
// For callers that cannot allocate device memory (code under stream capture, tight serving loops, harnesses
// that forbid hipMalloc): the batch's buffers come from `scratch`, a static device array of PAIR_SCRATCH_BYTES
// per GPU (held until the process ends).  In host code the array's name is a stand-in, not a device address:
// its address on each device is taken once, with that device current, by hipGetSymbolAddress and kept in
// g_scratch[device]; only that pointer is handed to kernels and copies.
namespace phx {

constexpr size_t PAIR_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[PAIR_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static char *g_scratch[SCRATCH_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_ptr(int device, char **out) {
    PHX_REQUIRE(device >= 0 && device < SCRATCH_MAX_DEVICES, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        PHX_CK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PHX_REQUIRE(p != nullptr, "no device address for the scratch array");
        g_scratch[device] = (char *)p;
    }
    *out = g_scratch[device];
    return 0;
}

// Per-pair plan on the device, in the caller's pair order: seed ph_seed(id, name) (FNV over "id|name"),
// steps ceil(tmax / dt_myr), step -tmax / steps (reference: src/phygas.c:185-187, 452-458), radii, and
// n_clones slots per pair.
__global__ __launch_bounds__(256) void k_pair_plan(const PhTarget *__restrict__ tg, const PhCluster *__restrict__ cl,
                                                   const int *__restrict__ ti, const int *__restrict__ ci, int npairs,
                                                   int n_clones, double dt_myr, uint64_t *__restrict__ seed,
                                                   int *__restrict__ psteps, double *__restrict__ pdt,
                                                   double *__restrict__ r50, double *__restrict__ rt,
                                                   long *__restrict__ off) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const PhTarget &T = tg[ti[p]];
    const PhCluster &C = cl[ci[p]];
    seed[p] = fnv_feed(seed_prefix(T.id), C.name, 40);
    const int s = (int)ceil(C.tmax_myr / dt_myr);
    psteps[p] = s;
    pdt[p] = -C.tmax_myr / s;
    r50[p] = C.r50_pc;
    rt[p] = C.rtot_pc;
    off[p] = (long)p * n_clones;
}

}  // namespace phx

using namespace phx;

extern "C" size_t phx_pair_noalloc_limit(void) { return PAIR_SCRATCH_BYTES; }

extern "C" int phx_pair_noalloc(int form, const PhTarget *t, const PhCluster *c, int npairs, const int *ti,
                                const int *ci, int n_clones, double dt_myr, PhResult *res, void *stream) {
    PHX_REQUIRE(npairs >= 0 && n_clones >= 0, "negative size");
    PHX_REQUIRE(form == PHX_FORM_PLAIN || form == PHX_FORM_FAST || form == PHX_FORM_FAST64, "unknown form");
    PHX_REQUIRE(dt_myr > 0, "dt must be positive");
    if (form == PHX_FORM_FAST64) form = PHX_FORM_PLAIN;
    if (npairs == 0) return 0;
    int nt = 0, ncl = 0;
    for (int p = 0; p < npairs; p++) {
        PHX_REQUIRE(ti[p] >= 0 && ci[p] >= 0, "negative record index");
        nt = std::max(nt, ti[p] + 1);
        ncl = std::max(ncl, ci[p] + 1);
    }
    const long slots = (long)npairs * n_clones;
    PHX_REQUIRE((double)npairs * n_clones < 2.0e9, "too many clones in one batch");
    const PairLayout sz = pair_layout(nullptr, slots, npairs, nt, ncl);
    if (sz.total > PAIR_SCRATCH_BYTES) {
        char msg[192];
        snprintf(msg, sizeof msg, "batch needs %zu bytes of scratch, above the %zu-byte limit (phx_pair_noalloc_limit)",
                 sz.total, PAIR_SCRATCH_BYTES);
        set_error("phx_pair_noalloc", msg);
        return 2;
    }
    int dev = 0;
    PHX_CK(hipGetDevice(&dev));
    char *base = nullptr;
    if (scratch_ptr(dev, &base)) return 1;
    const PairLayout L = pair_layout(base, slots, npairs, nt, ncl);
    hipStream_t st = (hipStream_t)stream;
    const Pot P = default_pot();
    PHX_CK(hipMemcpyAsync(L.tg, t, sizeof(PhTarget) * (size_t)nt, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(L.cl, c, sizeof(PhCluster) * (size_t)ncl, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(L.ti, ti, 4 * (size_t)npairs, hipMemcpyHostToDevice, st));
    PHX_CK(hipMemcpyAsync(L.ci, ci, 4 * (size_t)npairs, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_pair_plan, dim3(blocks_for(npairs, 256)), dim3(256), 0, st, L.tg, L.cl, L.ti, L.ci, npairs,
                       n_clones, dt_myr, L.seed, L.psteps, L.pdt, L.r50, L.rt, L.off);
    PHX_CK(hipGetLastError());
    PairPlan plan;   // sizes only: the per-pair arrays were built on the device
    plan.npairs = npairs;
    plan.n_clones = n_clones;
    plan.slots = slots;
    if (pair_launch_device(L, plan, form, L.tg, L.cl, P, st, nullptr, nullptr)) return 1;
    PHX_CK(hipMemcpyAsync(res, L.out, sizeof(PhResult) * (size_t)npairs, hipMemcpyDeviceToHost, st));
    PHX_CK(hipStreamSynchronize(st));
    return 0;
}
