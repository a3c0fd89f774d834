// Device-wide order-preserving selection: the indices i in [0, n) whose value passes a threshold test,
// written in increasing order.  Three launches: per-tile counts, one-block scan of the tile counts,
// per-tile scan and scatter.  Tiles are 256 threads x 8 consecutive items, so the count pass reads HBM
// in coalesced 2 KB runs per wavefront.
#include "common.h"

namespace phx {

constexpr int SEL_THREADS = 256, SEL_ITEMS = 8, SEL_TILE = SEL_THREADS * SEL_ITEMS;

struct LessThan {
    const double *v;
    long stride;
    double thr;
    __device__ bool operator()(long i) const { return v[i * stride] < thr; }
};
struct AtLeastCounted {
    const double *v;
    long stride;
    const long *n;
    double thr;
    __device__ bool operator()(long i) const { return i < *n && v[i * stride] >= thr; }
};

__device__ __forceinline__ int wave_incl_scan(int x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive prefix of x over the block (256 threads = 4 wavefronts); *total gets the block's sum.
__device__ __forceinline__ int block_excl_scan(int x, int *total) {
    __shared__ int wsum[SEL_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = wave_incl_scan(x);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < SEL_THREADS / 64; q++) {
        if (q < w) base += wsum[q];
        tot += wsum[q];
    }
    __syncthreads();
    *total = tot;
    return base + inc - x;
}

template <class Pred>
__global__ __launch_bounds__(SEL_THREADS) void k_sel_count(Pred pred, long n, long *__restrict__ tile_cnt) {
    const long base = (long)blockIdx.x * SEL_TILE + (long)threadIdx.x * SEL_ITEMS;
    int c = 0;
#pragma unroll
    for (int q = 0; q < SEL_ITEMS; q++) c += (base + q < n && pred(base + q)) ? 1 : 0;
    int tot;
    block_excl_scan(c, &tot);
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = tot;
}

// One block: exclusive scan of the tile counts in place; the grand total into *count.
__global__ __launch_bounds__(1024) void k_sel_scan(long *__restrict__ tile_cnt, long ntiles, long *__restrict__ count) {
    __shared__ long part[1024];
    long carry = 0;
    for (long b0 = 0; b0 < ntiles; b0 += 1024) {
        long i = b0 + threadIdx.x;
        long v = i < ntiles ? tile_cnt[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            long y = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
            __syncthreads();
            part[threadIdx.x] += y;
            __syncthreads();
        }
        if (i < ntiles) tile_cnt[i] = carry + part[threadIdx.x] - v;
        long chunk = part[1023];
        __syncthreads();
        carry += chunk;
    }
    if (threadIdx.x == 0) *count = carry;
}

template <class Pred>
__global__ __launch_bounds__(SEL_THREADS) void k_sel_scatter(Pred pred, long n, const long *__restrict__ tile_off,
                                                             long *__restrict__ idx, long cap) {
    const long base = (long)blockIdx.x * SEL_TILE + (long)threadIdx.x * SEL_ITEMS;
    bool f[SEL_ITEMS];
    int c = 0;
#pragma unroll
    for (int q = 0; q < SEL_ITEMS; q++) {
        f[q] = base + q < n && pred(base + q);
        c += f[q] ? 1 : 0;
    }
    int tot;
    long o = tile_off[blockIdx.x] + block_excl_scan(c, &tot);
#pragma unroll
    for (int q = 0; q < SEL_ITEMS; q++)
        if (f[q]) {
            if (o < cap) idx[o] = base + q;
            o++;
        }
}

size_t select_tmp_bytes(long n) { return sizeof(long) * (size_t)((n + SEL_TILE - 1) / SEL_TILE + 1); }

__global__ void k_zero_long(long *p) { *p = 0; }

template <class Pred>
static int run_select(Pred pred, long n, long *d_idx, long cap, long *d_count, void *d_tmp, hipStream_t st) {
    long ntiles = (n + SEL_TILE - 1) / SEL_TILE;
    long *tiles = (long *)d_tmp;
    if (ntiles == 0) {
        hipLaunchKernelGGL(k_zero_long, dim3(1), dim3(1), 0, st, d_count);
        PHX_CK(hipGetLastError());
        return 0;
    }
    hipLaunchKernelGGL(k_sel_count<Pred>, dim3((unsigned)ntiles), dim3(SEL_THREADS), 0, st, pred, n, tiles);
    PHX_CK(hipGetLastError());
    hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(1024), 0, st, tiles, ntiles, d_count);
    PHX_CK(hipGetLastError());
    hipLaunchKernelGGL(k_sel_scatter<Pred>, dim3((unsigned)ntiles), dim3(SEL_THREADS), 0, st, pred, n, tiles, d_idx, cap);
    PHX_CK(hipGetLastError());
    return 0;
}

int launch_select_less(const double *d_val, long n, double thr, long *d_idx, long cap, long *d_count, void *d_tmp,
                       hipStream_t st) {
    return run_select(LessThan{d_val, 1, thr}, n, d_idx, cap, d_count, d_tmp, st);
}

int launch_select_geq_counted(const double *d_val, long stride, const long *d_n, long n_cap, double thr, long *d_idx,
                              long cap, long *d_count, void *d_tmp, hipStream_t st) {
    return run_select(AtLeastCounted{d_val, stride, d_n, thr}, n_cap, d_idx, cap, d_count, d_tmp, st);
}

}  // namespace phx
