/* C ABI of the MI355X (gfx950) implementation of the Phygas trace-back path.
 *
 * Plain C: pointers, sizes and the record types of phygas_types.h.  Two layers:
 *
 *  1. pg_*  -- drop-in for the reference's GPU backend interface (reference: src/phygas_gpu.h:16-31,
 *     implemented there by src/phygas_cuda.cu:25-77 and src/phygas_opencl.c).  Same signatures, buffer
 *     layouts, slicing state and synchronous 0 = ok contract, so the reference's BOINC driver
 *     (src/main_phygas_boinc.cpp:59-301) links against this library unchanged.  Plain FP64.
 *
 *  2. phx_* -- the batched engine: a context per device with the cluster tables resident in HBM, star
 *     batches screened in one device pass (draws, tracks, scan, 512-draw finish and row compaction all on
 *     the GPU), Monte Carlo pairs with device-side quantiles, in a plain FP64 form and the fast form.
 *     Replaces the host loops of src/phygas_screen.c:48-67 and src/main_phygas_boinc.cpp:59-301.
 *
 * Every function returns 0 on success and a nonzero code on failure (message: phx_last_error()). */
#ifndef PHYGAS_HIP_H
#define PHYGAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "phygas_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- 1. reference GPU backend ABI */
typedef void *PgBuf;

/* reference: src/phygas_gpu.h:21 (src/phygas_cuda.cu:25-36) */
int pg_init(int platform, int device);
/* reference: src/phygas_gpu.h:22-25 (src/phygas_cuda.cu:38-45); 0 bytes allocates 8 */
PgBuf pg_alloc(size_t bytes);
int pg_put(PgBuf b, const void *src, size_t bytes);
int pg_get(PgBuf b, void *dst, size_t bytes);
void pg_free(PgBuf b);
/* reference: src/phygas_gpu.h:28-29, device src/phygas_kernel_body.h:64-109 */
int pg_lin(PgBuf ps0, PgBuf sig, PgBuf state, PgBuf trk, PgBuf params, double eps_pos, double eps_vel, double dt,
           int sub, int k0, int nk, int npts, int ns);
/* reference: src/phygas_gpu.h:30-31, device src/phygas_kernel_body.h:124-163 */
int pg_chi2(PgBuf trk, PgBuf csoa, PgBuf cn, PgBuf R2s, int nb, int nc, int npts, int p0, int np, PgBuf chi2,
            PgBuf kb, PgBuf sb);
/* reference: src/phygas_gpu.h:32-33, device src/phygas_kernel_body.h:181-220 */
int pg_finish(PgBuf trk, PgBuf csoa, PgBuf cn, PgBuf r50s, PgBuf rts, int nc, int npts, PgBuf hs, PgBuf hc, PgBuf hkb,
              PgBuf hsb, PgBuf hseed, int h0, int nh, PgBuf out);
/* reference: src/phygas_gpu.h:34, device src/phygas_kernel_body.h:222-274 */
int pg_pair(PgBuf ps, PgBuf steps, PgBuf dts, PgBuf st, PgBuf res, PgBuf params, int s0, int nk, int n);

/* ---------------------------------------------------------------- 2. batched engine */
typedef struct PhxCtx PhxCtx;

/* PLAIN: the reference's GPU algorithm in FP64 (full scan, every kept pair finished with FP64 draws).
 * FAST: FP32-offset tracks, exact culled scan, screened finish (the chosen form).
 * FAST64: the plain FP64 tracks with FAST's culled scan and screened finish: rows identical to PLAIN's. */
enum { PHX_FORM_PLAIN = 0, PHX_FORM_FAST = 1, PHX_FORM_FAST64 = 2 };

/* One screened (star, cluster) pair.  flags bit 0: the star belongs to the control group. */
typedef struct {
    int star, cluster, kb, flags;
    double chi2, sb, t, d_nom_pc, sig_pc, p_r50, p_rtot, vrel;
} PhxRow;

/* Context on a device with its own stream (default priority).  Not thread-safe: one host thread per
 * context.  No counterpart in the reference (its backend keeps one global device). */
int phx_create(int device, PhxCtx **out);
void phx_destroy(PhxCtx *ctx);
const char *phx_last_error(void);
int phx_sync(PhxCtx *ctx);
/* The context's stream (a hipStream_t) for callers that time or order work around it. */
void *phx_stream(PhxCtx *ctx);

/* Cluster tables: n_samp draws (seed "lin"|name) and the linear tracks of every cluster, computed on the
 * device and kept resident; r_kernel as the reference's ph_screen_r (0 = total radius).
 * Replaces the precompute loop of src/phygas_screen.c:32-33 / src/main_phygas_boinc.cpp:402-407 and the
 * table packing of src/main_phygas_boinc.cpp:61-90. */
int phx_load_clusters(PhxCtx *ctx, const PhCluster *cl, int nc, int n_samp, double r_kernel);
/* Star batch resident on the device: append ns records to group (0 = real, 1 = control). */
int phx_load_stars(PhxCtx *ctx, const PhTarget *st, int ns, int group);
int phx_clear_stars(PhxCtx *ctx);

/* One screening pass over every loaded star against the resident clusters (reference semantics of
 * src/phygas_screen.c:50-62 with the work-unit cut p_r50 >= p_min of src/main_phygas_boinc.cpp:441):
 * asynchronous on the context's stream.  form: PHX_FORM_PLAIN (reference GPU algorithm, FP64) or
 * PHX_FORM_FAST. */
int phx_screen(PhxCtx *ctx, int form, double chi2_max, int n_samp, double p_min);
/* Waits for the pass and copies its rows out, ordered by (group, star, cluster). *n = rows available
 * (may exceed cap, then only cap are copied). */
int phx_rows(PhxCtx *ctx, PhxRow *rows, long cap, long *n);
/* Pairs that passed the chi2 cut in the last pass (finished pairs). */
int phx_kept(PhxCtx *ctx, long *n);
/* Pairs the last pass evaluated point by point: every pair for the plain form, the pairs left by the tile
 * scan for the fast forms. */
int phx_candidates(PhxCtx *ctx, long *n);
/* Work counts of the last pass into st[0 .. min(cap, 7)): the scan's first-level tests (pair x 16-point
 * coarse tile), its certificate tests (coarse tiles and their 8-point tiles), candidate pairs (any tile
 * left), point-evaluation items (pair x tile left), kept pairs, rows, points evaluated.  For the plain form
 * the first and third entries are the pair count and the second, fourth and seventh 0. */
int phx_pass_stats(PhxCtx *ctx, long *st, int cap);
/* Durations (ms, HIP events on the context's stream) of the last pass's stages; returns the count, names
 * into names[] (static strings). */
int phx_stage_times(PhxCtx *ctx, double *ms, const char **names, int cap);
void phx_set_timing(PhxCtx *ctx, int on);
/* Fast-form passes captured into a hipGraph and replayed while batch sizes, parameters and buffers stay the
 * same (a pass with timing on runs eagerly, its stage events between the kernels). */
void phx_set_graphs(PhxCtx *ctx, int on);
int phx_graph_stats(PhxCtx *ctx, long *captures, long *replays);
/* A fast pass over a batch of at least 4,096 stars runs as `parts` sub-passes (1..4; default 3) over
 * contiguous star ranges, each on its own stream, so that one's latency-bound tracks overlap another's scan;
 * the last takes last_share_pct of the stars (phx_set_tuning, default 35%) and the others share the rest
 * equally.  Rows are the same.  A pass with timing on runs whole.
 * No counterpart in the reference (its screen is one host loop over the targets, src/phygas_screen.c:48-67). */
int phx_set_split(PhxCtx *ctx, int parts);
/* Offsets per lane of the fast tracks kernel (1, 2 or 3; 0 restores the default, 2).  The kernel runs one
 * workgroup per 64 stars: one wavefront integrates their FP64 nominal orbits, the others their six
 * finite-difference offsets, this many per lane.  Tracks and rows do not depend on it.  No counterpart in the
 * reference. */
int phx_set_track_offsets(PhxCtx *ctx, int offsets);

/* Launch grids of the fast pass's tile scan, item evaluation, kept-pair reduction and finish, chosen one
 * after another by timing whole passes over the loaded batch as the library runs them (results do not
 * depend on them); grids[4] and ms[4] receive the choices and their pass times.  No counterpart in the
 * reference. */
int phx_autotune(PhxCtx *ctx, int form, double chi2_max, int n_samp, double p_min, int *grids, double *ms);

/* Set those four launch grids directly (grids[i] <= 0 keeps the library's default for that stage), e.g. to
 * apply an earlier phx_autotune result without timing again.  No counterpart in the reference. */
int phx_set_grids(PhxCtx *ctx, const int *grids);

/* Every tunable setting of the fast pass.  stars_per_task: stars of one scan task (8 or 16; each task is one
 * wavefront over 64 clusters of similar track length, the clusters sorted by length into buckets of 64);
 * coarse_tiles: the scan's first level on 16-point time tiles, then their 8-point tiles (1, default), or on
 * the 8-point tiles directly (0); sub_passes: as phx_set_split (1..4, default 3); last_share_pct: the last
 * sub-pass's share of the stars (20..60 %, default 35); track_offsets: as phx_set_track_offsets (1..3, default
 * 2).  In phx_set_tuning sub_passes, last_share_pct and track_offsets <= 0 keep the current setting.  Rows do
 * not depend on any setting.  No counterpart in the reference. */
typedef struct {
    int scan_blocks, items_blocks, reduce_blocks, finish_blocks;   /* launch grids; <= 0: the library's default */
    int stars_per_task;
    int coarse_tiles;
    int sub_passes;
    int last_share_pct;
    int track_offsets;
    double pass_ms;   /* phx_autotune_ex: the pass time with the settings chosen */
} PhxTuning;
/* phx_autotune over every setting (sub-passes and the last one's share, the tracks' offsets per lane, the time
 * tile of the scan's first level, stars per task, the four grids); the choices stay set on the context and are returned in *out. */
int phx_autotune_ex(PhxCtx *ctx, int form, double chi2_max, int n_samp, double p_min, PhxTuning *out);
int phx_set_tuning(PhxCtx *ctx, const PhxTuning *t);
int phx_get_tuning(PhxCtx *ctx, PhxTuning *t);

/* Parity access to intermediate device results of the last plain pass / the cluster tables. */
int phx_cluster_tracks(PhxCtx *ctx, int *n, double *pos, double *vel, double *S, int npts_cap);
int phx_star_tracks(PhxCtx *ctx, int first, int count, double *trk, int npts_cap);
int phx_chi2_matrix(PhxCtx *ctx, double *chi2, int *kb, double *sb);
int phx_npts(PhxCtx *ctx);
/* Largest error of the FP32 Box-Muller estimate the fast finish screens draws with, over n uniform pairs of
 * splitmix64(seed) (the screen keeps draws within 1e-3 of a counting interval). */
int phx_check_finish_screen(PhxCtx *ctx, uint64_t seed, long n, double *worst);
/* The correctly rounded pow / x^1.5 / hypot / sin / cos that the device uses in the plain path's force and
 * coordinates, over n arguments: fn 0 pow(x, y), 1 x^1.5, 2 hypot(x, y), 3 sin x, 4 cos x (y unused for 1, 3,
 * 4); on_device 0 evaluates the host compilation of the same source (no device calls). */
int phx_check_crmath(int fn, int on_device, const double *x, const double *y, long n, double *out);

/* Monte Carlo clones already drawn (ps[n][12] = target pos/vel, cluster pos/vel): closest approach (kpc),
 * its time (Myr ago) and relative speed (kpc/Myr), per clone (reference: src/phygas.c:200-216). */
int phx_pair_clones(PhxCtx *ctx, int form, const double *ps, const int *steps, const double *dts, int n, double *best,
                    double *tbest, double *vbest);
/* Full Monte Carlo pairs (reference: src/phygas.c:240-249): clones drawn on the host with the reference
 * RNG (seed ph_seed(id, name)), traced on the device, quantiles and fractions reduced on the device. */
int phx_pair(PhxCtx *ctx, int form, const PhTarget *t, const PhCluster *c, int npairs, const int *ti, const int *ci,
             int n_clones, double dt_myr, PhResult *res);
/* phx_pair with options.  PHX_DRAW_ON_DEVICE: the clones are drawn on the GPU from the same splitmix64
 * stream and in the same order as the reference's host loop (values differ from host draws only by the
 * device libm's last bits); for batches whose host draws would dominate (thousands of clones per pair). */
enum { PHX_DRAW_ON_DEVICE = 1 };
int phx_pair_ex(PhxCtx *ctx, int form, int flags, const PhTarget *t, const PhCluster *c, int npairs, const int *ti,
                const int *ci, int n_clones, double dt_myr, PhResult *res);
/* The Monte Carlo pair search for pairs of the loaded stars and the resident clusters (star = index in the
 * loaded batch, in load order; cluster = index in the cluster table): e.g. stage 2 over the rows of a screening
 * pass above a p_r50 cut, which the reference's GPU app runs over a pair list with its clones batched across
 * pairs (reference: src/main_phygas_boinc.cpp:187-301).  Clones are drawn on the device from the resident
 * records (the reference RNG stream of ph_seed(id, name)), traced in one persistent launch ordered longest
 * trace-back first and reduced on the device; res in the caller's pair order.  Any clone count (pairs above
 * 4,096 clones are reduced by radix selection).  With graphs on (phx_set_graphs) the batch's launches are
 * captured once and replayed while its shape stays the same.  Pair forms: PHX_FORM_PLAIN (FP64) or
 * PHX_FORM_FAST (FP32 force); PHX_FORM_FAST64 runs the FP64 clones. */
int phx_pair_loaded(PhxCtx *ctx, int form, int npairs, const int *star, const int *cluster, int n_clones,
                    double dt_myr, PhResult *res);
/* Captures and replays of the device-draw Monte Carlo graphs. */
int phx_pair_graph_stats(PhxCtx *ctx, long *captures, long *replays);
/* Bucket width of the persistent clone kernel: chunks (one wavefront's clones) per queue grab (1..64; default 1).
 * Wider buckets take fewer queue atomics, narrower ones balance the tail better; results do not depend on it.
 * No counterpart in the reference. */
int phx_set_pair_bucket(PhxCtx *ctx, int chunks);
/* Lanes per clone of the clone kernel: 1 (target and cluster in one lane), 2 (one lane each), 0 (default: 1 for
 * the FP32 form; for FP64, 2 below 2 x 64 x SIMDs clones, else 1).  Results do not depend on it.  No
 * counterpart in the reference. */
int phx_set_pair_lanes(PhxCtx *ctx, int lanes);
/* The lanes per clone (1 or 2) a batch of n_clones clone slots of the given form runs with under the context's
 * setting; -1 for an unknown form. */
int phx_pair_lanes_for(PhxCtx *ctx, int form, long n_clones);
/* Lanes per clone and bucket width chosen by timing phx_pair_loaded on the caller's batch over 1 and 2 lanes x 1,
 * 2, 4 and 8 chunks; both stay set on the context.  *bucket, *ms: the bucket chosen and the time.  No counterpart
 * in the reference. */
int phx_autotune_pairs(PhxCtx *ctx, int form, int npairs, const int *star, const int *cluster, int n_clones,
                       double dt_myr, int *bucket, double *ms);

/* ---------------------------------------------------------------- 3. serving loop
 * Requests (screening chunks, Monte Carlo pair jobs) submitted as they arrive and batched continuously onto
 * the GPU by slot threads (each with its own stream and resident cluster tables), admitted by HBM footprint.
 * A request's results are the same whether it runs alone or batched with others.  No counterpart in the
 * reference (its app runs one work unit per process, src/main_phygas_boinc.cpp:336-471). */
typedef struct PhxServer PhxServer;
typedef struct {
    int form;                 /* PHX_FORM_FAST or PHX_FORM_PLAIN (pairs: FAST = FP32 clones, PLAIN = FP64) */
    int screen_slots;         /* concurrent screening batches */
    int pair_slots;           /* concurrent Monte Carlo batches */
    size_t hbm_budget;        /* bytes the batches in flight may hold by their footprint estimate */
    long max_batch_stars;     /* stars merged into one screening pass */
    long max_batch_clones;    /* clones merged into one Monte Carlo pass */
    double chi2_max, p_min, r_kernel;
    int n_samp;
} PhxServeConfig;

void phx_serve_defaults(PhxServeConfig *cfg);
int phx_server_create(int device, const PhCluster *cl, int nc, const PhxServeConfig *cfg, PhxServer **out);
/* Stops the slots after the batches in flight; requests still queued are dropped. */
void phx_server_destroy(PhxServer *srv);
/* Copies the request; returns its id (>= 0) or a negative code. */
long phx_server_submit_screen(PhxServer *srv, const PhTarget *stars, int ns);
long phx_server_submit_pair(PhxServer *srv, const PhTarget *target, const int *cluster_idx, int n, int n_clones,
                            double dt_myr);
/* 0 = done, 1 = still pending after timeout_s (< 0: wait forever), < 0 = the request failed. */
int phx_server_wait(PhxServer *srv, long id, double timeout_s);
/* Rows of a finished screening request (star = index within the request), ordered by (star, cluster). */
int phx_server_screen_rows(PhxServer *srv, long id, PhxRow *rows, long cap, long *n);
/* PhResult per requested cluster, in request order. */
int phx_server_pair_results(PhxServer *srv, long id, PhResult *res, int cap);
/* Arrival, start and completion times (s, monotonic clock) and the number of requests in its batch;
 * returns the request's status (0 queued, 1 running, 2 done, < 0 failed). */
int phx_server_info(PhxServer *srv, long id, double t[3], int *batch_reqs);
int phx_server_release(PhxServer *srv, long id);

/* ---------------------------------------------------------------- 4. work units
 * The reference app's task (reference: src/main_phygas_boinc.cpp:336-471) over the batched engine: the `in` file
 * (header "PHYGAS 1 <screen|pair> p1 p2 p3 p4", targets with optional clusters= lists) and the cluster file in,
 * the app's rows (10 columns for screen, 15 for pair) and "# done <nt>" out, in target order.  Screen mode: p1
 * chi2_max, p2 n_samp, p3 kernel radius (<= 0: 2), p4 p_min.  Pair mode: p1 clones, p2 dt (Myr); targets without
 * a clusters= list take the device prefilter.  After each batch of targets a checkpoint line
 * "phygas_checkpoint 2 <targets_done> <lines_done>" is written atomically; a resumed run keeps exactly that many
 * complete lines of `out`, so no row is lost or duplicated. */
typedef struct {
    int form;             /* screening: PHX_FORM_FAST64 (default: rows byte-identical to PHX_FORM_PLAIN's, the
                             reference's results), PHX_FORM_PLAIN or PHX_FORM_FAST (FP32 tracks: opt-in) */
    int batch_targets;    /* targets per GPU batch (checkpoint interval); default 20,000 */
    int max_batches;      /* > 0: stop (return PHX_WU_STOPPED) after this many batches, checkpoint written */
    int device;
    int draw_on_device;   /* pair mode: clones drawn on the GPU (PHX_DRAW_ON_DEVICE) instead of the host */
    int pair_form;        /* pair mode: PHX_FORM_PLAIN (default: FP64 clones) or PHX_FORM_FAST (FP32 force) */
} PhxWuOptions;
void phx_wu_defaults(PhxWuOptions *opt);
/* Return codes of phx_wu_run besides 0 (done: out ends with "# done"): PHX_WU_STOPPED after max_batches (the
 * checkpoint is written); PHX_WU_CKPT_UNREADABLE: the checkpoint exists but cannot be read (nothing touched:
 * retry later); any other nonzero code is an error (phx_last_error()), a device failure included.
 * *targets_done: targets in out. */
enum { PHX_WU_CKPT_UNREADABLE = 4, PHX_WU_STOPPED = 6 };
int phx_wu_run(const char *in_path, const char *clusters_path, const char *out_path, const char *ckpt_path,
               const PhxWuOptions *opt, long *targets_done);
/* The resume step alone (host only): reads the checkpoint, trims `out` to the lines it counts (or truncates it
 * for a fresh start) and returns where to continue.  No counterpart in the reference (its app resumes from a bare
 * target count and keeps every row written after it). */
int phx_wu_resume(const char *out_path, const char *ckpt_path, long nt, long *targets_done, long *lines_kept);

/* Pair-mode prefilter on the device (reference: src/phygas.c:98-168 ph_track_* and ph_prefilter, as
 * src/phygas_cli.c:103-117 runs them): mask[j * nc + i] = 1 when target j may have reached cluster i. */
int phx_prefilter(PhxCtx *ctx, const PhTarget *t, int nt, const PhCluster *c, int nc, unsigned char *mask);

/* ---------------------------------------------------------------- 5. allocation-free Monte Carlo batch
 * No counterpart in the reference.  For callers that cannot allocate device memory (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc): phx_pair_ex with PHX_DRAW_ON_DEVICE on the
 * current device, without a context, calling no allocation routine.  The batch's buffers come from a static
 * device array of phx_pair_noalloc_limit() bytes per GPU (64 MiB, held until the process ends); a batch that
 * needs more (about 132 bytes per clone slot plus the records) is rejected with code 2 before any device work.
 * Every call on one GPU shares that array, which keeps the last call's contents: calls that overlap on one GPU
 * (from several threads or streams) race.  stream: a hipStream_t (null: the default stream); the call returns
 * once the results are in res.  Results are those of phx_pair_ex. */
size_t phx_pair_noalloc_limit(void);
int phx_pair_noalloc(int form, const PhTarget *t, const PhCluster *c, int npairs, const int *ti, const int *ci,
                     int n_clones, double dt_myr, PhResult *res, void *stream);

/* Host helpers with the reference's semantics. */
uint64_t phx_seed(const char *a, const char *b); /* reference: src/phygas.c:452-458 */
int phx_pair_draw(const PhTarget *t, const PhCluster *c, int n_clones, double dt_myr, uint64_t seed, double *ps,
                  int *steps, double *dt); /* reference: src/phygas.c:182-198 */

#ifdef __cplusplus
}
#endif
#endif
