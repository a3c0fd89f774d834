// Fast form of the screening pass and the Monte Carlo clones (fast.hip): interface used by ctx.hip.
#pragma once

#include "common.h"

namespace phx {

// Per-device tables derived from the resident cluster tracks (built once per cluster load).
struct FastTables {
    void *mem[16] = {};
    size_t bytes[16] = {};
    int nc = 0, npts = 0, ntiles = 0, nc_pad = 0;
    int n_cu = 256;                    // compute units of the device (the scan's default grid follows it)
    const int *d_cn = nullptr;         // resident cluster point counts and kernel radii (owned by the context)
    const double *d_R2 = nullptr;
};

// Work buffers of one pass (grown on demand, reused across passes).
struct FastWork {
    void *mem[24] = {};
    size_t bytes[24] = {};
    long cand_cap = 0, kept_cap = 0, row_cap = 0, item_cap = 0, surv_cap = 0;
    long last_cand = 0, last_items = 0, last_l1 = 0, last_l2 = 0, last_points = 0;
    // launch grids (0: the defaults, chosen on the 40,000-star x 4,087-cluster unit run as sub-passes);
    // phx_autotune picks them by timing whole passes on the loaded batch
    int scan_blocks = 0, items_blocks = 0, reduce_blocks = 0, finish_blocks = 0;
    // the scan's stars per task (8 or 16) and its first level (1: 16-point coarse tiles, then their 8-point
    // tiles; 0: the 8-point tiles directly)
    int scan_spw = 16, scan_coarse = 1;
    // hipGraph of the pass's launches, replayed while sizes, parameters and buffers stay the same
    bool use_graphs = false;
    void *graph = nullptr, *gexec = nullptr;
    uint64_t gkey = 0;
    long captures = 0, replays = 0;
};

struct FastArgs {
    Pot P;
    int ns, nc, npts, n_samp;
    double chi2_max, p_min;
    int exact_tracks;                  // 1: the plain FP64 tracks instead of the FP32-offset tracks
    const PhTarget *stars;
    const uint64_t *seeds, *prefix;
    const double *nom;                 // nominal start states [ns][6] (host-drawn)
    const int *group;
    const PhCluster *cl;
    const double *csoa;
    const int *cn;
    const double *R2, *r50, *rt;
    // ps0 / Sigma0 of the batch's stars when already drawn (the sub-passes of a split pass share one draw
    // kernel); null: the pass draws them into its own buffers
    const double *ps0_in = nullptr, *sig_in = nullptr;
    // offsets per lane of the tracks kernel's offset wavefronts (1, 2 or 3; the same bits whatever the value)
    int track_offsets = 2;
};

typedef void (*StageMark)(void *ctx, const char *name);

int fast_build_tables(FastTables &T, const double *d_csoa, const int *d_cn, const double *d_R2, int nc, int npts,
                      hipStream_t st);
// part: FAST_ALL (the whole pass), FAST_HEAD (counters, draws and tracks) or FAST_TAIL (the rest, after the head
// on the same stream).  A caller running several passes on several streams launches every head first, so that
// each pass's first kernels do not wait behind the host's launches of the others' later ones.  A pass captured
// into a hipGraph is launched whole with its head (its tail is then nothing).
enum { FAST_ALL = 0, FAST_HEAD = 1, FAST_TAIL = 2 };
int fast_screen(FastTables &T, FastWork &W, const FastArgs &a, hipStream_t st, StageMark mark, void *mark_ctx,
                int part = FAST_ALL);
// Kept-pair and row counts of the last pass into h[0], h[1]; *grow = 1 when a buffer overflowed (it has
// been enlarged and the pass must be run again).
int fast_counts(FastWork &W, hipStream_t st, long h[2], int *grow);
const void *fast_rows_ptr(const FastWork &W);
const double *fast_tracks_ptr(const FastWork &W);
// Largest error of the FP32 Box-Muller estimate used to screen draws, over n uniforms pairs.
int fast_z0_check(uint64_t seed, long n, double *worst, hipStream_t st);
void fast_release(FastTables &T, FastWork &W);
void fast_release_work(FastWork &W);   // the work buffers and graph alone


}  // namespace phx
