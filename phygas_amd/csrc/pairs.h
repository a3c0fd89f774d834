// Monte Carlo pair batches on the device (pairs.hip): buffer layout, the host-side plan of a batch, launches.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "physics.h"

namespace phx {

// Largest clone count per pair the LDS sort of the reduction takes; above it the reduction selects the
// quantiles' order statistics by radix passes (any count).
constexpr int PAIR_RED_LDS = 4096;

// Device buffers of a batch of npairs pairs with `slots` clone slots, carved from one block of `total` bytes
// (256-byte aligned pieces).  tg / cl: record copies for host-given records (null when the records are
// already resident).
struct PairLayout {
    double *ps = nullptr;      // [slots][12] clone start states: target pos, vel, cluster pos, vel
    int *steps = nullptr;      // [slots]
    double *dts = nullptr;     // [slots]
    double *res = nullptr;     // [slots][3] closest approach (kpc), its time (Myr ago), relative speed (kpc/Myr)
    long *off = nullptr;       // [npairs] first slot of each pair
    int *m = nullptr;          // [npairs] clones drawn
    int *psteps = nullptr;     // [npairs]
    double *pdt = nullptr;     // [npairs]
    double *r50 = nullptr, *rt = nullptr;   // [npairs] pc
    int *ti = nullptr, *ci = nullptr;       // [npairs] record indices
    uint64_t *seed = nullptr;  // [npairs]
    PhResult *out = nullptr;   // [npairs]
    PhTarget *tg = nullptr;
    PhCluster *cl = nullptr;
    unsigned *next = nullptr;  // the clone queue's counter
    size_t total = 0;
};
// base == nullptr: sizes only (total).
PairLayout pair_layout(char *base, long slots, int npairs, int ntg, int ncl);

// Host-side plan: per-pair seeds ph_seed(id, name), step counts and steps (reference: src/phygas.c:185-187),
// radii, and slot offsets with the pairs ranked longest trace-back first (n_clones slots each).
struct PairPlan {
    int npairs = 0, n_clones = 0, max_steps = 0;
    int bucket = 1;   // chunks (one wavefront's clones) per queue grab of the clone kernel
    int lanes = 0;    // lanes per clone: 1, 2, or 0 (chosen from the form and the batch size)
    long slots = 0;
    std::vector<uint64_t> seed;
    std::vector<int> psteps;
    std::vector<double> pdt, r50, rt;
    std::vector<long> off;
};
void pair_plan(const PhTarget *t, const PhCluster *c, int npairs, const int *ti, const int *ci, int n_clones,
               double dt_myr, PairPlan &plan);
int pair_upload_plan(const PairLayout &L, const PairPlan &plan, const int *ti, const int *ci, hipStream_t st);
// Draws (device), steps, traces and reduction of a planned batch from device records d_t / d_c (indexed by
// L.ti / L.ci); results in L.out in the caller's pair order.  Launches only: capturable into a hipGraph.  mark
// (nullable) is called after the draws, the traces and the reduction (stage timing).
typedef void (*PairMark)(void *ctx, const char *name);
int pair_launch_device(const PairLayout &L, const PairPlan &plan, int form, const PhTarget *d_t, const PhCluster *d_c,
                       const Pot &P, hipStream_t st, PairMark mark, void *mark_ctx);
int pair_launch_draws(const PairLayout &L, const PairPlan &plan, const PhTarget *d_t, const PhCluster *d_c,
                      const Pot &P, hipStream_t st);
// Traces of n clones (persistent, queue counter d_next zeroed by a kernel first; each queue grab takes cw
// chunks of one wavefront's clones).  form: PHX_FORM_FAST = FP32 force, otherwise FP64.  lanes per clone: 1,
// 2 or 0 (chosen from n); the results do not depend on it.
int launch_clones(int form, const double *d_ps, const int *d_steps, const double *d_dts, double *d_res, const Pot &P,
                  long n, int cw, int lanes, unsigned *d_next, hipStream_t st);
// The lanes per clone a launch of n clones of the given form takes for the setting lanes (0: one for the FP32
// form; for FP64 two below 2 x 64 x SIMDs clones, else one).
int clone_lanes(int form, long n, int lanes);
// Quantiles and fractions of each pair's clones; max_m bounds every m[p] (chooses the LDS sort or selection).
int launch_pair_reduce(const double *d_res, const long *d_off, const int *d_m, const double *d_r50_pc,
                       const double *d_rtot_pc, int npairs, int max_m, PhResult *d_out, hipStream_t st);

}  // namespace phx
