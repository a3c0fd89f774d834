// Batched engine (phx_* C ABI): one context per device, cluster tables resident in HBM, star batches
// screened in a single device pass, Monte Carlo pairs traced and reduced on the device.
//
// HBM layout of the resident tables (nc clusters, P = ceil(max tmax / 0.5) + 1 points):
//   csoa [P][12][nc] f64   pos xyz, vel xyz, S (xx xy xz yy yz zz) per point, clusters innermost so a
//                          wavefront reading 64 neighbouring clusters is one coalesced 512-byte load
//   cn [nc] i32, R2 / r50 / rtot [nc] f64 (kpc), records [nc] PhCluster (names seed the finish)
// Star batch: records [ns] PhTarget, "lin"|id seeds [ns] u64, "id|" hash prefixes [ns] u64, group [ns].
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <string>
#include <vector>

#include "common.h"
#include "fast.h"
#include "pairs.h"

using namespace phx;

namespace {

struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap && p) return 0;
        if (p) PHX_CK(hipFree(p));
        p = nullptr;
        cap = 0;
        size_t b = bytes ? bytes : 8;
        PHX_CK(hipMalloc(&p, b));
        cap = b;
        return 0;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return (T *)p; }
};

constexpr int MAX_STAGES = 12;
// A fast pass over a batch of at least 2 SPLIT_MIN stars runs as up to MAX_SPLIT sub-passes (contiguous star
// ranges) on their own streams, so that one sub-pass's latency-bound tracks overlap another's scan and items.
constexpr int MAX_SPLIT = 4, SPLIT_MIN = 2048;

}  // namespace

struct PhxCtx {
    int device = 0;
    hipStream_t st = nullptr;
    Pot P;
    // clusters
    int nc = 0, npts = 0, n_samp_c = 64;
    double r_kernel = 2.0;
    std::vector<int> cn_h;
    std::vector<PhCluster> cl_h;
    DBuf cl, cseed, cnom, cps0, csig, csoa, cn, R2, r50, rt;
    FastTables fast;   // fast-form tables (fast.hip), built on first use
    bool fast_ready = false;
    // stars
    std::vector<PhTarget> st_h;
    std::vector<int> group_h;
    bool stars_dirty = true;
    DBuf sr, sseed, sprefix, sgroup, snom;
    // plain pass
    DBuf ps0, sig, trk, chi2, kb, sb, seltmp, idx, counts, fout, ridx, rows;
    long kept_cap = 0, row_cap = 0;
    int last_form = -1;
    double last_chi2_max = 25.0, last_p_min = 0.0;
    int last_n_samp = 64;
    FastWork fw;       // fast-form work buffers (part 0 of a split pass)
    int split = 3;     // sub-passes of a large fast pass (phx_set_split)
    int last_pct = 35; // the last sub-pass's share of the stars, percent (phx_set_tuning)
    int track_offsets = 2; // offsets per lane of the fast tracks' offset wavefronts (phx_set_track_offsets)
    int nparts = 1;    // sub-passes of the last fast pass, over stars [part_base[i], part_base[i + 1])
    int part_base[MAX_SPLIT + 1] = {};
    long part_rows[MAX_SPLIT] = {};
    FastWork pfw[MAX_SPLIT];                 // parts 1.. (pfw[0] unused)
    hipStream_t part_stream[MAX_SPLIT] = {}; // their streams (part_stream[0] unused)
    hipEvent_t ev_fork = nullptr, ev_join[MAX_SPLIT] = {};
    // timing
    bool timing = false;
    int nstage = 0;
    hipEvent_t ev[MAX_STAGES + 1] = {};
    const char *stage_name[MAX_STAGES] = {};
    // Monte Carlo batches: one buffer block (pair_layout) and the graph of the last device-draw batch
    DBuf pmem;
    hipGraphExec_t pgexec = nullptr;
    uint64_t pgkey = 0;
    long pcaptures = 0, preplays = 0;
    int pair_bucket = 1;   // chunks per queue grab of the clone kernel (phx_set_pair_bucket)
    int pair_lanes = 0;    // lanes per clone, 0 = chosen from the batch (phx_set_pair_lanes)
};

static FastWork &part_fw(PhxCtx *c, int i) { return i == 0 ? c->fw : c->pfw[i]; }
static hipStream_t part_st(PhxCtx *c, int i) { return i == 0 ? c->st : c->part_stream[i]; }

static uint64_t fnv_str(uint64_t h, const char *s) {
    for (const unsigned char *p = (const unsigned char *)s; *p; p++) h = (h ^ *p) * FNV_PRIME;
    return h;
}

extern "C" uint64_t phx_seed(const char *a, const char *b) {
    uint64_t h = fnv_str(FNV_BASIS, a);
    h = (h ^ (unsigned char)'|') * FNV_PRIME;
    return fnv_str(h, b);
}

extern "C" const char *phx_last_error(void) { return last_error(); }

extern "C" int phx_create(int device, PhxCtx **out) {
    *out = nullptr;
    int n = 0;
    PHX_CK(hipGetDeviceCount(&n));
    PHX_REQUIRE(n > 0, "no HIP device");
    PHX_REQUIRE(device >= 0 && device < n, "device index out of range");
    PHX_CK(hipSetDevice(device));
    PhxCtx *c = new PhxCtx();
    c->device = device;
    c->P = default_pot();
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_error("phx_create", "stream creation failed");
        return 1;
    }
    for (int i = 0; i <= MAX_STAGES; i++)
        if (hipEventCreate(&c->ev[i]) != hipSuccess) {
            set_error("phx_create", "event creation failed");
            return 1;
        }
    *out = c;
    return 0;
}

extern "C" void phx_destroy(PhxCtx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->st);
    DBuf *bufs[] = {&c->cl, &c->cseed, &c->cnom, &c->snom, &c->cps0, &c->csig, &c->csoa, &c->cn, &c->R2, &c->r50, &c->rt,
                    &c->sr, &c->sseed, &c->sprefix, &c->sgroup, &c->ps0, &c->sig, &c->trk, &c->chi2,
                    &c->kb, &c->sb, &c->seltmp, &c->idx, &c->counts, &c->fout, &c->ridx, &c->rows, &c->pmem};
    if (c->pgexec) (void)hipGraphExecDestroy(c->pgexec);
    for (DBuf *b : bufs) b->release();
    fast_release(c->fast, c->fw);
    for (int i = 1; i < MAX_SPLIT; i++) {
        if (c->part_stream[i]) (void)hipStreamSynchronize(c->part_stream[i]);
        fast_release_work(c->pfw[i]);
        if (c->part_stream[i]) (void)hipStreamDestroy(c->part_stream[i]);
        if (c->ev_join[i]) (void)hipEventDestroy(c->ev_join[i]);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    for (int i = 0; i <= MAX_STAGES; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    (void)hipStreamDestroy(c->st);
    delete c;
}

extern "C" int phx_sync(PhxCtx *c) {
    PHX_CK(hipStreamSynchronize(c->st));
    return 0;
}

extern "C" void *phx_stream(PhxCtx *c) { return (void *)c->st; }

extern "C" void phx_set_timing(PhxCtx *c, int on) { c->timing = on != 0; }

extern "C" void phx_set_graphs(PhxCtx *c, int on) { c->fw.use_graphs = on != 0; }

extern "C" int phx_set_track_offsets(PhxCtx *c, int offsets) {
    PHX_REQUIRE(offsets >= 0 && offsets <= 3, "offsets per lane must be 1, 2 or 3 (0: the default, 2)");
    c->track_offsets = offsets ? offsets : 2;
    return 0;
}

extern "C" int phx_set_split(PhxCtx *c, int parts) {
    PHX_REQUIRE(parts >= 1 && parts <= MAX_SPLIT, "split must be 1..4");
    c->split = parts;
    return 0;
}

extern "C" int phx_graph_stats(PhxCtx *c, long *captures, long *replays) {
    *captures = *replays = 0;
    for (int i = 0; i < MAX_SPLIT; i++) {
        *captures += part_fw(c, i).captures;
        *replays += part_fw(c, i).replays;
    }
    return 0;
}

static void stage_begin(PhxCtx *c) {
    c->nstage = 0;
    if (c->timing) (void)hipEventRecord(c->ev[0], c->st);
}
static void stage_mark(PhxCtx *c, const char *name) {
    if (!c->timing || c->nstage >= MAX_STAGES) return;
    c->stage_name[c->nstage] = name;
    c->nstage++;
    (void)hipEventRecord(c->ev[c->nstage], c->st);
}

extern "C" int phx_stage_times(PhxCtx *c, double *ms, const char **names, int cap) {
    PHX_CK(hipStreamSynchronize(c->st));
    int n = std::min(c->nstage, cap);
    for (int i = 0; i < n; i++) {
        float f = 0;
        PHX_CK(hipEventElapsedTime(&f, c->ev[i], c->ev[i + 1]));
        ms[i] = f;
        if (names) names[i] = c->stage_name[i];
    }
    return n;
}

// ------------------------------------------------------------------------------------- clusters
extern "C" int phx_load_clusters(PhxCtx *c, const PhCluster *cl, int nc, int n_samp, double r_kernel) {
    PHX_REQUIRE(nc > 0, "no clusters");
    PHX_CK(hipSetDevice(c->device));
    c->nc = nc;
    c->n_samp_c = n_samp;
    c->r_kernel = r_kernel;
    c->fast_ready = false;
    c->cl_h.assign(cl, cl + nc);
    double tmax = 0;
    for (int i = 0; i < nc; i++) tmax = std::max(tmax, cl[i].tmax_myr);
    c->npts = (int)ceil(tmax / STEP_MYR) + 1;   // reference: src/main_phygas_boinc.cpp:61
    c->cn_h.resize(nc);
    std::vector<uint64_t> seeds(nc);
    std::vector<double> R2(nc), r50(nc), rt(nc);
    for (int i = 0; i < nc; i++) {
        c->cn_h[i] = (int)ceil(cl[i].tmax_myr / STEP_MYR) + 1;   // reference: src/phygas.c:297
        seeds[i] = phx_seed("lin", cl[i].name);
        double R = (r_kernel > 0 ? r_kernel * cl[i].r50_pc : cl[i].rtot_pc) / 1000.0;   // src/phygas.c:374
        R2[i] = R * R;
        r50[i] = cl[i].r50_pc / 1000.0;
        rt[i] = cl[i].rtot_pc / 1000.0;
    }
    std::vector<double> nom((size_t)nc * 6);
    nominal_states(c->P, cl, nc, nom.data());
    if (c->cl.ensure(sizeof(PhCluster) * nc) || c->cseed.ensure(8 * nc) || c->cnom.ensure(48 * nc) ||
        c->cps0.ensure(48 * nc) ||
        c->csig.ensure(288 * nc) || c->csoa.ensure(sizeof(double) * 12 * (size_t)c->npts * nc) ||
        c->cn.ensure(4 * nc) || c->R2.ensure(8 * nc) || c->r50.ensure(8 * nc) || c->rt.ensure(8 * nc))
        return 1;
    PHX_CK(hipMemcpyAsync(c->cl.p, cl, sizeof(PhCluster) * nc, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(c->cseed.p, seeds.data(), 8 * nc, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(c->cnom.p, nom.data(), 48 * (size_t)nc, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(c->cn.p, c->cn_h.data(), 4 * nc, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(c->R2.p, R2.data(), 8 * nc, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(c->r50.p, r50.data(), 8 * nc, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(c->rt.p, rt.data(), 8 * nc, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemsetAsync(c->csoa.p, 0, c->csoa.cap, c->st));
    if (launch_lin_init_clusters(c->cl.as<PhCluster>(), c->cseed.as<uint64_t>(), c->cnom.as<double>(), nc, n_samp, c->P,
                                 c->cps0.as<double>(), c->csig.as<double>(), c->st))
        return 1;
    if (launch_lin_plain(c->cps0.as<double>(), c->csig.as<double>(), nullptr, c->csoa.as<double>(), c->P,
                         PH_LIN_EPS_POS, PH_LIN_EPS_VEL, -STEP_MYR / PH_LIN_SUB, PH_LIN_SUB, 0, c->npts, c->npts, nc,
                         nc, c->cn.as<int>(), c->st))
        return 1;
    PHX_CK(hipStreamSynchronize(c->st));
    return 0;
}

extern "C" int phx_npts(PhxCtx *c) { return c->npts; }

extern "C" int phx_cluster_tracks(PhxCtx *c, int *n, double *pos, double *vel, double *S, int npts_cap) {
    PHX_REQUIRE(c->nc > 0, "no clusters loaded");
    PHX_REQUIRE(npts_cap >= c->npts, "npts_cap too small");
    std::vector<double> h(c->csoa.cap / 8);
    PHX_CK(hipMemcpyAsync(h.data(), c->csoa.p, c->csoa.cap, hipMemcpyDeviceToHost, c->st));
    PHX_CK(hipStreamSynchronize(c->st));
    const int nc = c->nc;
    for (int i = 0; i < nc; i++) {
        n[i] = c->cn_h[i];
        for (int k = 0; k < c->cn_h[i]; k++) {
            size_t o = (size_t)i * npts_cap + k;
            for (int q = 0; q < 3; q++) {
                pos[o * 3 + q] = h[((size_t)k * 12 + q) * nc + i];
                vel[o * 3 + q] = h[((size_t)k * 12 + 3 + q) * nc + i];
            }
            for (int q = 0; q < 6; q++) S[o * 6 + q] = h[((size_t)k * 12 + 6 + q) * nc + i];
        }
    }
    return 0;
}

// ------------------------------------------------------------------------------------- stars
extern "C" int phx_clear_stars(PhxCtx *c) {
    c->st_h.clear();
    c->group_h.clear();
    c->stars_dirty = true;
    return 0;
}

extern "C" int phx_load_stars(PhxCtx *c, const PhTarget *s, int ns, int group) {
    PHX_REQUIRE(ns >= 0, "negative star count");
    c->st_h.insert(c->st_h.end(), s, s + ns);
    c->group_h.insert(c->group_h.end(), (size_t)ns, group);
    c->stars_dirty = true;
    return 0;
}

static int upload_stars(PhxCtx *c) {
    if (!c->stars_dirty) return 0;
    const int ns = (int)c->st_h.size();
    std::vector<uint64_t> seed(ns), prefix(ns);
    std::vector<double> nom((size_t)ns * 6);
    for (int i = 0; i < ns; i++) {
        seed[i] = phx_seed("lin", c->st_h[i].id);
        prefix[i] = seed_prefix(c->st_h[i].id);
    }
    nominal_states(c->P, c->st_h.data(), ns, nom.data());
    if (c->sr.ensure(sizeof(PhTarget) * ns) || c->sseed.ensure(8 * ns) || c->sprefix.ensure(8 * ns) ||
        c->sgroup.ensure(4 * ns) || c->snom.ensure(48 * (size_t)ns))
        return 1;
    PHX_CK(hipMemcpyAsync(c->sr.p, c->st_h.data(), sizeof(PhTarget) * ns, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(c->sseed.p, seed.data(), 8 * ns, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(c->sprefix.p, prefix.data(), 8 * ns, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(c->sgroup.p, c->group_h.data(), 4 * ns, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(c->snom.p, nom.data(), 48 * (size_t)ns, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipStreamSynchronize(c->st));
    c->stars_dirty = false;
    return 0;
}

// ------------------------------------------------------------------------------------- screening
static int fast_prepare(PhxCtx *c) {
    if (c->fast_ready) return 0;
    if (fast_build_tables(c->fast, c->csoa.as<double>(), c->cn.as<int>(), c->R2.as<double>(), c->nc, c->npts, c->st))
        return 1;
    c->fast_ready = true;
    return 0;
}

static int screen_plain(PhxCtx *c, double chi2_max, int n_samp, double p_min) {
    const int ns = (int)c->st_h.size(), nc = c->nc, npts = c->npts;
    const long npairs = (long)ns * nc;
    if (c->kept_cap == 0) c->kept_cap = std::max(1L << 20, npairs / 8);
    if (c->row_cap == 0) c->row_cap = std::max(1L << 16, c->kept_cap / 16);
    c->kept_cap = std::min(c->kept_cap, std::max(npairs, 1L));
    if (c->ps0.ensure(48 * (size_t)ns) || c->sig.ensure(288 * (size_t)ns) ||
        c->trk.ensure(sizeof(double) * 12 * (size_t)npts * ns) || c->chi2.ensure(8 * (size_t)npairs) ||
        c->kb.ensure(4 * (size_t)npairs) || c->sb.ensure(8 * (size_t)npairs) ||
        c->seltmp.ensure(select_tmp_bytes(std::max(npairs, c->kept_cap))) || c->idx.ensure(8 * (size_t)c->kept_cap) ||
        c->counts.ensure(16) || c->fout.ensure(48 * (size_t)c->kept_cap) || c->ridx.ensure(8 * (size_t)c->row_cap) ||
        c->rows.ensure(sizeof(PhxRow) * (size_t)c->row_cap))
        return 1;
    long *cnt = c->counts.as<long>();
    stage_begin(c);
    if (launch_lin_init_targets(c->sr.as<PhTarget>(), c->sseed.as<uint64_t>(), c->snom.as<double>(), ns, n_samp, c->P,
                                c->ps0.as<double>(), c->sig.as<double>(), c->st))
        return 1;
    stage_mark(c, "draws");
    if (launch_lin_plain(c->ps0.as<double>(), c->sig.as<double>(), nullptr, c->trk.as<double>(), c->P, PH_LIN_EPS_POS,
                         PH_LIN_EPS_VEL, -STEP_MYR / PH_LIN_SUB, PH_LIN_SUB, 0, npts, npts, ns, 0, nullptr, c->st))
        return 1;
    stage_mark(c, "tracks");
    if (launch_chi2_plain(c->trk.as<double>(), c->csoa.as<double>(), c->cn.as<int>(), c->R2.as<double>(), ns, nc, npts,
                          0, npairs, c->chi2.as<double>(), c->kb.as<int>(), c->sb.as<double>(), c->st))
        return 1;
    stage_mark(c, "scan");
    if (launch_select_less(c->chi2.as<double>(), npairs, chi2_max, c->idx.as<long>(), c->kept_cap, cnt,
                           c->seltmp.p, c->st))
        return 1;
    stage_mark(c, "select");
    if (launch_finish_sel_plain(c->trk.as<double>(), c->csoa.as<double>(), c->cn.as<int>(), c->r50.as<double>(),
                                c->rt.as<double>(), c->cl.as<PhCluster>(), c->sprefix.as<uint64_t>(), nc, npts,
                                c->idx.as<long>(), cnt, c->kept_cap, c->kb.as<int>(), c->sb.as<double>(),
                                c->fout.as<double>(), c->st))
        return 1;
    stage_mark(c, "finish");
    if (launch_select_geq_counted(c->fout.as<double>() + 3, 6, cnt, c->kept_cap, p_min, c->ridx.as<long>(), c->row_cap,
                                  cnt + 1, c->seltmp.p, c->st) ||
        launch_rows(c->idx.as<long>(), c->ridx.as<long>(), cnt + 1, c->row_cap, nc, c->sgroup.as<int>(),
                    c->chi2.as<double>(), c->kb.as<int>(), c->sb.as<double>(), c->fout.as<double>(),
                    c->rows.as<PhxRow>(), c->st))
        return 1;
    stage_mark(c, "rows");
    return 0;
}

static int screen_pass(PhxCtx *c, int form, double chi2_max, int n_samp, double p_min) {
    if (form == PHX_FORM_PLAIN) return screen_plain(c, chi2_max, n_samp, p_min);
    if (fast_prepare(c)) return 1;
    stage_begin(c);
    FastArgs a;
    a.P = c->P; a.ns = (int)c->st_h.size(); a.nc = c->nc; a.npts = c->npts; a.n_samp = n_samp;
    a.chi2_max = chi2_max; a.p_min = p_min; a.exact_tracks = form == PHX_FORM_FAST64;
    a.stars = c->sr.as<PhTarget>(); a.seeds = c->sseed.as<uint64_t>(); a.prefix = c->sprefix.as<uint64_t>();
    a.nom = c->snom.as<double>();
    a.group = c->sgroup.as<int>(); a.cl = c->cl.as<PhCluster>(); a.csoa = c->csoa.as<double>();
    a.cn = c->cn.as<int>(); a.R2 = c->R2.as<double>(); a.r50 = c->r50.as<double>(); a.rt = c->rt.as<double>();
    auto mark = [](void *ctx, const char *name) { stage_mark((PhxCtx *)ctx, name); };
    const int ns = a.ns;
    int parts = c->timing || ns < 2 * SPLIT_MIN ? 1 : std::min(c->split, MAX_SPLIT);
    c->nparts = parts;
    // the last sub-pass takes last_pct of the stars and the others share the rest equally; by default three
    // sub-passes of 32.5/32.5/35%: measured on the 40,000-star unit, 14.5-14.7 ms per pass against 15.0-15.2
    // for two halves (the pass time moves with how the sub-passes' kernels meet, not smoothly with the sizes;
    // the autotuner tries other counts and shares on the caller's batch)
    const int last = parts > 1 ? (int)((long)ns * c->last_pct / 100) : ns;
    for (int i = 0; i < parts; i++) c->part_base[i] = parts > 1 ? (int)((long)(ns - last) * i / (parts - 1)) : 0;
    c->part_base[parts] = ns;
    if (parts == 1) {
        a.track_offsets = c->track_offsets;
        return fast_screen(c->fast, c->fw, a, c->st, c->timing ? (StageMark)mark : nullptr, c);
    }
    if (!c->ev_fork) PHX_CK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    for (int i = 1; i < parts; i++) {
        if (!c->part_stream[i]) PHX_CK(hipStreamCreateWithFlags(&c->part_stream[i], hipStreamNonBlocking));
        if (!c->ev_join[i]) PHX_CK(hipEventCreateWithFlags(&c->ev_join[i], hipEventDisableTiming));
    }
    // the draws of every sub-pass in one kernel before the fork: a sub-pass's tracks then start without
    // waiting for its own draw kernel to find room beside the other sub-passes' tracks
    if (c->ps0.ensure(48 * (size_t)ns) || c->sig.ensure(288 * (size_t)ns)) return 1;
    if (launch_lin_init_targets(a.stars, a.seeds, a.nom, ns, n_samp, c->P, c->ps0.as<double>(), c->sig.as<double>(),
                                c->st))
        return 1;
    PHX_CK(hipEventRecord(c->ev_fork, c->st));   // uploads, cluster tables and draws first
    for (int i = 1; i < parts; i++) PHX_CK(hipStreamWaitEvent(c->part_stream[i], c->ev_fork, 0));
    // every sub-pass's head (counters, tracks) is launched before any sub-pass's tail, so that the last
    // sub-pass's tracks, the start of the pass's longest chain, do not wait for the host to launch the other
    // sub-passes' fifteen kernels each
    FastArgs parg[MAX_SPLIT];
    for (int i = 0; i < parts; i++) {
        const int b = c->part_base[i];
        FastArgs ai = a;
        ai.ns = c->part_base[i + 1] - b;
        ai.stars += b; ai.seeds += b; ai.prefix += b; ai.group += b; ai.nom += (size_t)b * 6;
        ai.ps0_in = c->ps0.as<double>() + (size_t)b * 6;
        ai.sig_in = c->sig.as<double>() + (size_t)b * 36;
        ai.track_offsets = c->track_offsets;
        FastWork &W = part_fw(c, i);
        if (i > 0) {   // the launch settings of part 0
            W.scan_blocks = c->fw.scan_blocks; W.items_blocks = c->fw.items_blocks;
            W.reduce_blocks = c->fw.reduce_blocks; W.finish_blocks = c->fw.finish_blocks;
            W.scan_spw = c->fw.scan_spw; W.scan_coarse = c->fw.scan_coarse;
            W.use_graphs = c->fw.use_graphs;
        }
        parg[i] = ai;
        if (fast_screen(c->fast, W, ai, part_st(c, i), nullptr, nullptr, FAST_HEAD)) return 1;
    }
    for (int i = 0; i < parts; i++)
        if (fast_screen(c->fast, part_fw(c, i), parg[i], part_st(c, i), nullptr, nullptr, FAST_TAIL)) return 1;
    for (int i = 1; i < parts; i++) {   // the context's stream waits for every part
        PHX_CK(hipEventRecord(c->ev_join[i], c->part_stream[i]));
        PHX_CK(hipStreamWaitEvent(c->st, c->ev_join[i], 0));
    }
    return 0;
}

extern "C" int phx_screen(PhxCtx *c, int form, double chi2_max, int n_samp, double p_min) {
    PHX_REQUIRE(c->nc > 0, "load clusters first");
    PHX_REQUIRE(form == PHX_FORM_PLAIN || form == PHX_FORM_FAST || form == PHX_FORM_FAST64, "unknown form");
    PHX_CK(hipSetDevice(c->device));
    if (upload_stars(c)) return 1;
    c->last_form = form;
    c->last_chi2_max = chi2_max;
    c->last_n_samp = n_samp;
    c->last_p_min = p_min;
    if (c->st_h.empty()) return 0;
    return screen_pass(c, form, chi2_max, n_samp, p_min);
}

// Kept / row counts of the last pass; reruns the pass with larger buffers when a selection overflowed.
static int last_counts(PhxCtx *c, long *kept, long *nrows) {
    if (c->st_h.empty() || c->last_form < 0) {
        *kept = *nrows = 0;
        return 0;
    }
    // each rerun can find one more overflowing stage (candidates, items, kept segments, listed draws, rows)
    for (int attempt = 0; attempt < 8; attempt++) {
        long h[2];
        if (c->last_form == PHX_FORM_PLAIN) {
            PHX_CK(hipMemcpyAsync(h, c->counts.p, sizeof h, hipMemcpyDeviceToHost, c->st));
            PHX_CK(hipStreamSynchronize(c->st));
            if (h[0] > c->kept_cap || h[1] > c->row_cap) {
                c->kept_cap = std::max(c->kept_cap, h[0] + h[0] / 4 + 1024);
                c->row_cap = std::max(c->row_cap, h[1] + h[1] / 4 + 1024);
                if (screen_plain(c, c->last_chi2_max, c->last_n_samp, c->last_p_min)) return 1;
                continue;
            }
        } else {
            int grow = 0;
            h[0] = h[1] = 0;
            for (int i = 0; i < c->nparts; i++) {   // the stream has waited for every part
                long hp[2];
                int g = 0;
                if (fast_counts(part_fw(c, i), c->st, hp, &g)) return 1;
                grow |= g;
                h[0] += hp[0];
                h[1] += hp[1];
                c->part_rows[i] = hp[1];
            }
            if (grow) {
                if (screen_pass(c, c->last_form, c->last_chi2_max, c->last_n_samp, c->last_p_min)) return 1;
                continue;
            }
        }
        *kept = h[0];
        *nrows = h[1];
        return 0;
    }
    set_error("phx_rows", "selection buffers kept overflowing");
    return 3;
}

extern "C" int phx_kept(PhxCtx *c, long *n) {
    long r;
    return last_counts(c, n, &r);
}

// Launch settings of the fast pass chosen by timing whole passes on the loaded batch as the library runs them
// (sub-passes overlapping), one knob after another with the others held, the mean of five passes per option
// after one not counted, each pass checked complete (overlapping sub-passes vary from pass to pass: the best or
// the median of a few would pick luck; throughput is the mean): the number of sub-passes and the last one's
// share of the stars, the tracks' offsets per lane, the scan's first level (16-point coarse tiles or the 8-point
// tiles), its stars per task,
// and the launch grids of the scan, the item evaluation, the kept-pair reduction and the finish.  Rows do not depend on any of them.  knob_ms[k]: the
// best pass time while knob k was tuned.
constexpr int N_KNOBS = 9;
static int autotune_fast(PhxCtx *c, int form, double chi2_max, int n_samp, double p_min, double knob_ms[N_KNOBS]) {
    PHX_REQUIRE(form == PHX_FORM_FAST || form == PHX_FORM_FAST64, "the autotuner tunes the fast forms");
    PHX_REQUIRE(!c->st_h.empty(), "load stars first");
    if (phx_screen(c, form, chi2_max, n_samp, p_min)) return 1;
    long kept, nr;
    if (last_counts(c, &kept, &nr)) return 1;   // sizes the buffers
    struct Knob { int *v; int opts[5]; };
    Knob knobs[N_KNOBS] = {{&c->split, {3, 2, 4, 1, -1}},
                     {&c->last_pct, {35, 40, 45, 50, 30}},
                     {&c->track_offsets, {2, 1, 3, -1, -1}},
                     {&c->fw.scan_coarse, {1, 0, -1, -1, -1}},
                     {&c->fw.scan_spw, {16, 8, -1, -1, -1}},
                     {&c->fw.scan_blocks, {512, 576, 768, 1024, 2048}},
                     {&c->fw.items_blocks, {4096, 8192, 16384, 32768, 65536}},
                     {&c->fw.reduce_blocks, {1024, 2048, 4096, 8192, -1}},
                     {&c->fw.finish_blocks, {1024, 2048, 4096, 8192, -1}}};
    const bool timing = c->timing;
    c->timing = false;
    int rc = 0;
    for (int q = 0; q < N_KNOBS && rc == 0; q++) {
        int best = *knobs[q].v;
        double best_ms = 1e300;
        for (int o = 0; o < 5 && knobs[q].opts[o] >= 0; o++) {
            *knobs[q].v = knobs[q].opts[o];
            constexpr int NRUN = 5;
            double runs[NRUN];
            for (int rep = 0; rep <= NRUN && rc == 0; rep++) {   // the first pass with new settings is not counted
                const auto t0 = std::chrono::steady_clock::now();
                if ((rc = phx_screen(c, form, chi2_max, n_samp, p_min)) != 0) break;
                // complete passes only: a setting whose buffers overflow is run again with larger ones (the
                // first, uncounted pass of an option grows them), as a caller's pass would be
                if ((rc = last_counts(c, &kept, &nr)) != 0) break;
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (rep > 0) runs[rep - 1] = ms;
            }
            if (rc) break;
            double t = 0.0;
            for (double r : runs) t += r / NRUN;
            if (rc == 0 && t < best_ms) { best_ms = t; best = knobs[q].opts[o]; }
        }
        *knobs[q].v = best;
        knob_ms[q] = best_ms;
    }
    c->timing = timing;
    return rc;
}

extern "C" int phx_autotune(PhxCtx *c, int form, double chi2_max, int n_samp, double p_min, int *grids, double *ms) {
    double kms[N_KNOBS];
    if (int rc = autotune_fast(c, form, chi2_max, n_samp, p_min, kms)) return rc;
    const int g[4] = {c->fw.scan_blocks, c->fw.items_blocks, c->fw.reduce_blocks, c->fw.finish_blocks};
    for (int i = 0; i < 4; i++) {
        if (grids) grids[i] = g[i];
        if (ms) ms[i] = kms[N_KNOBS - 4 + i];
    }
    return 0;
}

extern "C" int phx_set_grids(PhxCtx *c, const int *grids) {
    PHX_REQUIRE(grids != nullptr, "grids is null");
    for (int i = 0; i < 4; i++) PHX_REQUIRE(grids[i] <= (1 << 20), "grid too large");
    c->fw.scan_blocks = grids[0];
    c->fw.items_blocks = grids[1];
    c->fw.reduce_blocks = grids[2];
    c->fw.finish_blocks = grids[3];
    return 0;
}

extern "C" int phx_get_tuning(PhxCtx *c, PhxTuning *t) {
    t->scan_blocks = c->fw.scan_blocks;
    t->items_blocks = c->fw.items_blocks;
    t->reduce_blocks = c->fw.reduce_blocks;
    t->finish_blocks = c->fw.finish_blocks;
    t->stars_per_task = c->fw.scan_spw;
    t->coarse_tiles = c->fw.scan_coarse;
    t->sub_passes = c->split;
    t->last_share_pct = c->last_pct;
    t->track_offsets = c->track_offsets;
    t->pass_ms = 0;
    return 0;
}

extern "C" int phx_autotune_ex(PhxCtx *c, int form, double chi2_max, int n_samp, double p_min, PhxTuning *out) {
    double kms[N_KNOBS];
    if (int rc = autotune_fast(c, form, chi2_max, n_samp, p_min, kms)) return rc;
    if (out) {
        phx_get_tuning(c, out);
        out->pass_ms = kms[N_KNOBS - 1];
    }
    return 0;
}

extern "C" int phx_set_tuning(PhxCtx *c, const PhxTuning *t) {
    PHX_REQUIRE(t != nullptr, "tuning is null");
    PHX_REQUIRE(t->stars_per_task == 8 || t->stars_per_task == 16, "stars per task must be 8 or 16");
    PHX_REQUIRE(t->coarse_tiles == 0 || t->coarse_tiles == 1, "coarse_tiles must be 0 or 1");
    PHX_REQUIRE(t->sub_passes <= MAX_SPLIT, "sub_passes must be 1..4 (0 keeps the current setting)");
    PHX_REQUIRE(t->last_share_pct <= 0 || (t->last_share_pct >= 20 && t->last_share_pct <= 60),
                "last_share_pct must be 20..60 (0 keeps the current setting)");
    PHX_REQUIRE(t->track_offsets <= 3, "track_offsets must be 1..3 (0 keeps the current setting)");
    const int g[4] = {t->scan_blocks, t->items_blocks, t->reduce_blocks, t->finish_blocks};
    if (int rc = phx_set_grids(c, g)) return rc;
    c->fw.scan_spw = t->stars_per_task;
    c->fw.scan_coarse = t->coarse_tiles;
    if (t->sub_passes > 0) c->split = t->sub_passes;
    if (t->last_share_pct > 0) c->last_pct = t->last_share_pct;
    if (t->track_offsets > 0) c->track_offsets = t->track_offsets;
    return 0;
}

extern "C" int phx_pass_stats(PhxCtx *c, long *st, int cap) {
    long kept, nr;
    if (last_counts(c, &kept, &nr)) return 1;
    long v[7] = {0, 0, 0, 0, kept, nr, 0};
    const long ns = (long)c->st_h.size();
    if (c->last_form == PHX_FORM_PLAIN) {
        v[0] = ns * c->nc;   // every pair is scanned point by point
        v[2] = ns * c->nc;
    } else {
        for (int i = 0; i < c->nparts; i++) {
            v[0] += part_fw(c, i).last_l1;
            v[1] += part_fw(c, i).last_l2;
            v[2] += part_fw(c, i).last_cand;
            v[3] += part_fw(c, i).last_items;
            v[6] += part_fw(c, i).last_points;
        }
    }
    for (int i = 0; i < cap && i < 7; i++) st[i] = v[i];
    return 0;
}

extern "C" int phx_candidates(PhxCtx *c, long *n) {
    long kept, r;
    if (last_counts(c, &kept, &r)) return 1;
    if (c->last_form == PHX_FORM_PLAIN) {
        *n = (long)c->st_h.size() * c->nc;
    } else {
        *n = 0;
        for (int i = 0; i < c->nparts; i++) *n += part_fw(c, i).last_cand;
    }
    return 0;
}

extern "C" int phx_rows(PhxCtx *c, PhxRow *rows, long cap, long *n) {
    long kept, nr;
    if (last_counts(c, &kept, &nr)) return 1;
    *n = nr;
    long m = std::min(cap, nr);
    if (m > 0 && c->last_form == PHX_FORM_PLAIN) {
        PHX_CK(hipMemcpyAsync(rows, c->rows.p, sizeof(PhxRow) * m, hipMemcpyDeviceToHost, c->st));
        PHX_CK(hipStreamSynchronize(c->st));
    } else if (m > 0) {
        // every part's rows (its star indices are local to the part, its rows in append order), put into
        // (star, cluster) order before the first m are returned: counted by star, then each star's few rows
        // ordered by cluster ((star, cluster) is unique per row)
        std::unique_ptr<PhxRow[]> all(new PhxRow[nr]);
        long o = 0;
        for (int i = 0; i < c->nparts; i++) {
            const long mi = c->part_rows[i];
            if (mi <= 0) continue;
            PHX_CK(hipMemcpyAsync(all.get() + o, fast_rows_ptr(part_fw(c, i)), sizeof(PhxRow) * mi,
                                  hipMemcpyDeviceToHost, c->st));
            PHX_CK(hipStreamSynchronize(c->st));
            for (long r = o; r < o + mi; r++) all[r].star += c->part_base[i];
            o += mi;
        }
        const int ns = c->part_base[c->nparts];   // the pass's batch (the loaded stars may have changed since)
        std::vector<long> at(ns + 1, 0);
        for (long r = 0; r < o; r++) {
            PHX_REQUIRE(all[r].star >= 0 && all[r].star < ns, "row star index out of the batch");
            at[all[r].star + 1]++;
        }
        for (int s = 0; s < ns; s++) at[s + 1] += at[s];
        std::vector<long> idx(o);
        {
            std::vector<long> put(at.begin(), at.end() - 1);
            for (long r = 0; r < o; r++) idx[put[all[r].star]++] = r;
        }
        for (int s = 0; s < ns; s++)
            if (at[s + 1] - at[s] > 1)
                std::sort(idx.begin() + at[s], idx.begin() + at[s + 1],
                          [&](long a, long b) { return all[a].cluster < all[b].cluster; });
        for (long k = 0; k < m; k++) rows[k] = all[idx[k]];
    }
    return 0;
}

extern "C" int phx_star_tracks(PhxCtx *c, int first, int count, double *trk, int npts_cap) {
    PHX_REQUIRE(c->last_form >= 0, "no pass has run");
    PHX_REQUIRE(npts_cap == c->npts, "npts_cap must equal phx_npts()");
    PHX_REQUIRE(first >= 0 && first + count <= (int)c->st_h.size(), "star range out of the batch");
    if (c->last_form == PHX_FORM_PLAIN) {
        PHX_CK(hipMemcpyAsync(trk, c->trk.as<double>() + (size_t)first * c->npts * 12,
                              sizeof(double) * 12 * (size_t)c->npts * count, hipMemcpyDeviceToHost, c->st));
    } else {
        for (int i = 0; i < c->nparts; i++) {   // the parts' pieces of [first, first + count)
            const int lo = std::max(first, c->part_base[i]), hi = std::min(first + count, c->part_base[i + 1]);
            if (lo >= hi) continue;
            PHX_CK(hipMemcpyAsync(trk + (size_t)(lo - first) * c->npts * 12,
                                  fast_tracks_ptr(part_fw(c, i)) + (size_t)(lo - c->part_base[i]) * c->npts * 12,
                                  sizeof(double) * 12 * (size_t)c->npts * (hi - lo), hipMemcpyDeviceToHost, c->st));
        }
    }
    PHX_CK(hipStreamSynchronize(c->st));
    return 0;
}

extern "C" int phx_chi2_matrix(PhxCtx *c, double *chi2, int *kb, double *sb) {
    PHX_REQUIRE(c->last_form == PHX_FORM_PLAIN, "the chi2 matrix is kept by the plain pass");
    size_t n = (size_t)c->st_h.size() * c->nc;
    PHX_CK(hipMemcpyAsync(chi2, c->chi2.p, 8 * n, hipMemcpyDeviceToHost, c->st));
    PHX_CK(hipMemcpyAsync(kb, c->kb.p, 4 * n, hipMemcpyDeviceToHost, c->st));
    PHX_CK(hipMemcpyAsync(sb, c->sb.p, 8 * n, hipMemcpyDeviceToHost, c->st));
    PHX_CK(hipStreamSynchronize(c->st));
    return 0;
}

extern "C" int phx_check_finish_screen(PhxCtx *c, uint64_t seed, long n, double *worst) {
    PHX_CK(hipSetDevice(c->device));
    return fast_z0_check(seed, n, worst, c->st);
}

// ------------------------------------------------------------------------------------- Monte Carlo pairs
// Clone draws on the host with the reference RNG (reference: src/phygas.c:182-198); the physics is the
// same FP64 code the device runs, compiled for the host against the C library's libm.
extern "C" int phx_pair_draw(const PhTarget *t, const PhCluster *cl, int n_clones, double dt_myr, uint64_t seed,
                             double *ps, int *steps_out, double *dt_out) {
    const Pot P = default_pot();
    uint64_t st = seed;
    int steps = (int)ceil(cl->tmax_myr / dt_myr);
    *steps_out = steps;
    *dt_out = -cl->tmax_myr / steps;
    int m = 0;
    for (int k = 0; k < n_clones; k++) {
        V3 ap, av, bp, bv;
        if (sample_target(P, *t, &st, ap, av) || sample_cluster(P, *cl, &st, bp, bv)) continue;
        double *o = ps + (size_t)m * 12;
        o[0] = ap.x; o[1] = ap.y; o[2] = ap.z; o[3] = av.x; o[4] = av.y; o[5] = av.z;
        o[6] = bp.x; o[7] = bp.y; o[8] = bp.z; o[9] = bv.x; o[10] = bv.y; o[11] = bv.z;
        m++;
    }
    return m;
}

// Pair-mode prefilter on the device (reference: src/phygas_cli.c:103-117, src/phygas.c:98-168): target paths
// over the longest cluster trace-back, cluster paths over their own; mask[j * nc + i].
extern "C" int phx_prefilter(PhxCtx *c, const PhTarget *t, int nt, const PhCluster *cl, int nc, unsigned char *mask) {
    PHX_REQUIRE(nt >= 0 && nc >= 0, "negative size");
    PHX_CK(hipSetDevice(c->device));
    if (nt == 0 || nc == 0) return 0;
    double tmax = 0;
    for (int i = 0; i < nc; i++) tmax = std::max(tmax, cl[i].tmax_myr);
    const int npts = (int)ceil(tmax / STEP_MYR) + 1;
    std::vector<uint64_t> ts(nt), cs(nc);
    std::vector<int> tn(nt, npts), cn(nc);
    for (int j = 0; j < nt; j++) ts[j] = phx_seed("track", t[j].id);
    for (int i = 0; i < nc; i++) {
        cs[i] = phx_seed("track", cl[i].name);
        cn[i] = (int)ceil(cl[i].tmax_myr / STEP_MYR) + 1;
    }
    DBuf dt, dc, dts, dcs, dtn, dcn, tp, tsp, cp, csp, dm;
    int rc = dt.ensure(sizeof(PhTarget) * nt) || dc.ensure(sizeof(PhCluster) * nc) || dts.ensure(8 * (size_t)nt) ||
             dcs.ensure(8 * (size_t)nc) || dtn.ensure(4 * (size_t)nt) || dcn.ensure(4 * (size_t)nc) ||
             tp.ensure(24 * (size_t)nt * npts) || tsp.ensure(16 * (size_t)nt) || cp.ensure(24 * (size_t)nc * npts) ||
             csp.ensure(16 * (size_t)nc) || dm.ensure((size_t)nt * nc);
    if (!rc) {
        PHX_CK(hipMemcpyAsync(dt.p, t, sizeof(PhTarget) * nt, hipMemcpyHostToDevice, c->st));
        PHX_CK(hipMemcpyAsync(dc.p, cl, sizeof(PhCluster) * nc, hipMemcpyHostToDevice, c->st));
        PHX_CK(hipMemcpyAsync(dts.p, ts.data(), 8 * (size_t)nt, hipMemcpyHostToDevice, c->st));
        PHX_CK(hipMemcpyAsync(dcs.p, cs.data(), 8 * (size_t)nc, hipMemcpyHostToDevice, c->st));
        PHX_CK(hipMemcpyAsync(dtn.p, tn.data(), 4 * (size_t)nt, hipMemcpyHostToDevice, c->st));
        PHX_CK(hipMemcpyAsync(dcn.p, cn.data(), 4 * (size_t)nc, hipMemcpyHostToDevice, c->st));
        rc = launch_prefilter(dt.as<PhTarget>(), dts.as<uint64_t>(), dtn.as<int>(), nt, dc.as<PhCluster>(),
                              dcs.as<uint64_t>(), dcn.as<int>(), nc, npts, c->P, tp.as<double>(), tsp.as<double>(),
                              cp.as<double>(), csp.as<double>(), dm.as<unsigned char>(), c->st);
        if (!rc) {
            PHX_CK(hipMemcpyAsync(mask, dm.p, (size_t)nt * nc, hipMemcpyDeviceToHost, c->st));
            PHX_CK(hipStreamSynchronize(c->st));
        }
    }
    DBuf *all[] = {&dt, &dc, &dts, &dcs, &dtn, &dcn, &tp, &tsp, &cp, &csp, &dm};
    for (DBuf *b : all) b->release();
    return rc;
}

// Device buffers of the context's Monte Carlo batches: one block, carved by pair_layout.
static int pair_buffers(PhxCtx *c, long slots, int npairs, int ntg, int ncl, PairLayout &L) {
    const size_t need = pair_layout(nullptr, slots, npairs, ntg, ncl).total;
    if (c->pmem.ensure(need)) return 1;
    L = pair_layout((char *)c->pmem.p, slots, npairs, ntg, ncl);
    return 0;
}

static int check_pair_args(int &form, int npairs, int n_clones) {
    PHX_REQUIRE(npairs >= 0 && n_clones >= 0, "negative size");
    PHX_REQUIRE(form == PHX_FORM_PLAIN || form == PHX_FORM_FAST || form == PHX_FORM_FAST64, "unknown form");
    PHX_REQUIRE((double)npairs * n_clones < 4.0e9, "too many clones in one batch (at most 4e9 slots)");
    if (form == PHX_FORM_FAST64) form = PHX_FORM_PLAIN;   // FAST64's clones are the plain FP64 ones
    return 0;
}

extern "C" int phx_pair_clones(PhxCtx *c, int form, const double *ps, const int *steps, const double *dts, int n,
                               double *best, double *tbest, double *vbest) {
    if (check_pair_args(form, 0, 0)) return 2;
    PHX_REQUIRE(n >= 0, "negative clone count");
    PHX_CK(hipSetDevice(c->device));
    if (n == 0) return 0;
    PairLayout L;
    if (pair_buffers(c, n, 0, 0, 0, L)) return 1;
    PHX_CK(hipMemcpyAsync(L.ps, ps, 96 * (size_t)n, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(L.steps, steps, 4 * (size_t)n, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(L.dts, dts, 8 * (size_t)n, hipMemcpyHostToDevice, c->st));
    stage_begin(c);
    if (launch_clones(form, L.ps, L.steps, L.dts, L.res, c->P, n, c->pair_bucket, c->pair_lanes, L.next, c->st))
        return 1;
    stage_mark(c, "clones");
    std::vector<double> r(3 * (size_t)n);
    PHX_CK(hipMemcpyAsync(r.data(), L.res, 24 * (size_t)n, hipMemcpyDeviceToHost, c->st));
    PHX_CK(hipStreamSynchronize(c->st));
    for (int i = 0; i < n; i++) {
        best[i] = r[3 * i];
        tbest[i] = r[3 * i + 1];
        vbest[i] = r[3 * i + 2];
    }
    return 0;
}

// The launches of a planned device-draw batch, captured into a hipGraph on first use and replayed while the
// batch's sizes, form and buffers stay the same (when graphs are on).
static int pair_run_launches(PhxCtx *c, const PairLayout &L, const PairPlan &plan, int form, const PhTarget *d_t,
                             const PhCluster *d_c) {
    stage_begin(c);
    if (!c->fw.use_graphs || c->timing) {
        auto mark = [](void *ctx, const char *name) { stage_mark((PhxCtx *)ctx, name); };
        return pair_launch_device(L, plan, form, d_t, d_c, c->P, c->st, c->timing ? (PairMark)mark : nullptr, c);
    }
    uint64_t key = FNV_BASIS;
    auto mix = [&](uint64_t v) { key = (key ^ v) * FNV_PRIME; };
    mix((uint64_t)plan.npairs); mix((uint64_t)plan.n_clones); mix((uint64_t)plan.slots); mix((uint64_t)form);
    mix((uint64_t)plan.bucket); mix((uint64_t)plan.lanes);
    mix((uint64_t)(uintptr_t)L.ps); mix((uint64_t)(uintptr_t)d_t); mix((uint64_t)(uintptr_t)d_c);
    if (!c->pgexec || c->pgkey != key) {
        if (c->pgexec) (void)hipGraphExecDestroy(c->pgexec);
        c->pgexec = nullptr;
        hipGraph_t g = nullptr;
        PHX_CK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        const int rc = pair_launch_device(L, plan, form, d_t, d_c, c->P, c->st, nullptr, nullptr);
        const hipError_t ec = hipStreamEndCapture(c->st, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        PHX_CK(ec);
        const hipError_t ei = hipGraphInstantiate(&c->pgexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PHX_CK(ei);
        c->pgkey = key;
        c->pcaptures++;
    } else {
        c->preplays++;
    }
    PHX_CK(hipGraphLaunch(c->pgexec, c->st));
    stage_mark(c, "pairs");
    return 0;
}

// Clone draws on the device from host-given records (copied into the batch's buffers).
static int pair_device_draws(PhxCtx *c, int form, const PhTarget *t, const PhCluster *cl, int npairs, const int *ti,
                             const int *ci, int n_clones, double dt_myr, PhResult *res) {
    int nt = 0, ncl = 0;
    for (int p = 0; p < npairs; p++) {
        nt = std::max(nt, ti[p] + 1);
        ncl = std::max(ncl, ci[p] + 1);
    }
    PairPlan plan;
    pair_plan(t, cl, npairs, ti, ci, n_clones, dt_myr, plan);
    plan.bucket = c->pair_bucket;
    plan.lanes = c->pair_lanes;
    PairLayout L;
    if (pair_buffers(c, plan.slots, npairs, nt, ncl, L)) return 1;
    PHX_CK(hipMemcpyAsync(L.tg, t, sizeof(PhTarget) * nt, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(L.cl, cl, sizeof(PhCluster) * ncl, hipMemcpyHostToDevice, c->st));
    if (pair_upload_plan(L, plan, ti, ci, c->st)) return 1;
    if (pair_run_launches(c, L, plan, form, L.tg, L.cl)) return 1;
    PHX_CK(hipMemcpyAsync(res, L.out, sizeof(PhResult) * (size_t)npairs, hipMemcpyDeviceToHost, c->st));
    PHX_CK(hipStreamSynchronize(c->st));
    return 0;
}

extern "C" int phx_pair_ex(PhxCtx *c, int form, int flags, const PhTarget *t, const PhCluster *cl, int npairs,
                           const int *ti, const int *ci, int n_clones, double dt_myr, PhResult *res) {
    if (check_pair_args(form, npairs, n_clones)) return 2;
    PHX_CK(hipSetDevice(c->device));
    if (npairs == 0) return 0;
    if (flags & PHX_DRAW_ON_DEVICE) return pair_device_draws(c, form, t, cl, npairs, ti, ci, n_clones, dt_myr, res);
    return phx_pair(c, form, t, cl, npairs, ti, ci, n_clones, dt_myr, res);
}

extern "C" int phx_pair_loaded(PhxCtx *c, int form, int npairs, const int *star, const int *cluster, int n_clones,
                               double dt_myr, PhResult *res) {
    if (check_pair_args(form, npairs, n_clones)) return 2;
    PHX_REQUIRE(c->nc > 0, "load clusters first");
    PHX_CK(hipSetDevice(c->device));
    if (npairs == 0) return 0;
    const int ns = (int)c->st_h.size();
    for (int p = 0; p < npairs; p++)
        PHX_REQUIRE(star[p] >= 0 && star[p] < ns && cluster[p] >= 0 && cluster[p] < c->nc, "pair index out of range");
    if (upload_stars(c)) return 1;
    PairPlan plan;
    pair_plan(c->st_h.data(), c->cl_h.data(), npairs, star, cluster, n_clones, dt_myr, plan);
    plan.bucket = c->pair_bucket;
    plan.lanes = c->pair_lanes;
    PairLayout L;
    if (pair_buffers(c, plan.slots, npairs, 0, 0, L)) return 1;
    if (pair_upload_plan(L, plan, star, cluster, c->st)) return 1;
    if (pair_run_launches(c, L, plan, form, c->sr.as<PhTarget>(), c->cl.as<PhCluster>())) return 1;
    PHX_CK(hipMemcpyAsync(res, L.out, sizeof(PhResult) * (size_t)npairs, hipMemcpyDeviceToHost, c->st));
    PHX_CK(hipStreamSynchronize(c->st));
    return 0;
}

extern "C" int phx_set_pair_bucket(PhxCtx *c, int chunks) {
    PHX_REQUIRE(chunks >= 1 && chunks <= 64, "bucket width must be 1..64 chunks");
    c->pair_bucket = chunks;
    return 0;
}

extern "C" int phx_set_pair_lanes(PhxCtx *c, int lanes) {
    PHX_REQUIRE(lanes >= 0 && lanes <= 2, "lanes per clone must be 0 (chosen from the batch), 1 or 2");
    c->pair_lanes = lanes;
    return 0;
}

// The clone kernel's lanes per clone and bucket width chosen by timing phx_pair_loaded on the caller's batch
// (results do not depend on them); the choice stays set.  No counterpart in the reference.
extern "C" int phx_autotune_pairs(PhxCtx *c, int form, int npairs, const int *star, const int *cluster, int n_clones,
                                  double dt_myr, int *bucket, double *ms) {
    PHX_REQUIRE(npairs > 0, "no pairs to tune on");
    std::vector<PhResult> res(npairs);
    const int opts[4] = {1, 2, 4, 8};
    int best = c->pair_bucket, best_lanes = c->pair_lanes;
    double best_ms = 1e300;
    for (int lanes = 1; lanes <= 2; lanes++)
        for (int o : opts) {
            c->pair_bucket = o;
            c->pair_lanes = lanes;
            double t = 1e300;
            for (int rep = 0; rep < 3; rep++) {   // the first call with a new setting is not counted
                const auto t0 = std::chrono::steady_clock::now();
                if (int rc = phx_pair_loaded(c, form, npairs, star, cluster, n_clones, dt_myr, res.data())) return rc;
                const double dt =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (rep > 0) t = std::min(t, dt);
            }
            if (t < best_ms) { best_ms = t; best = o; best_lanes = lanes; }
        }
    c->pair_bucket = best;
    c->pair_lanes = best_lanes;
    if (bucket) *bucket = best;
    if (ms) *ms = best_ms;
    return 0;
}

extern "C" int phx_pair_lanes_for(PhxCtx *c, int form, long n_clones) {
    if (check_pair_args(form, 0, 0)) return -1;
    PHX_CK(hipSetDevice(c->device));
    return clone_lanes(form, n_clones, c->pair_lanes);
}

extern "C" int phx_pair_graph_stats(PhxCtx *c, long *captures, long *replays) {
    *captures = c->pcaptures;
    *replays = c->preplays;
    return 0;
}

extern "C" int phx_pair(PhxCtx *c, int form, const PhTarget *t, const PhCluster *cl, int npairs, const int *ti,
                        const int *ci, int n_clones, double dt_myr, PhResult *res) {
    if (check_pair_args(form, npairs, n_clones)) return 2;
    PHX_CK(hipSetDevice(c->device));
    if (npairs == 0) return 0;
    // draws per pair on the host (reference RNG), into the slots of the device layout: pairs longest trace-back
    // first, so the clone queue hands out the long chunks first; results stay in the caller's pair order
    PairPlan plan;
    pair_plan(t, cl, npairs, ti, ci, n_clones, dt_myr, plan);
    const size_t n = (size_t)plan.slots;
    std::vector<double> ps(n * 12, 0.0), dts(n, 0.0);
    std::vector<int> steps(n, 0), m(npairs);
    for (int p = 0; p < npairs; p++) {
        int st_p = 0;
        double dt_p = 0;
        m[p] = phx_pair_draw(&t[ti[p]], &cl[ci[p]], n_clones, dt_myr, plan.seed[p], ps.data() + (size_t)plan.off[p] * 12,
                             &st_p, &dt_p);
        for (int k = 0; k < n_clones; k++) {
            steps[plan.off[p] + k] = k < m[p] ? st_p : 0;
            dts[plan.off[p] + k] = dt_p;
        }
    }
    PairLayout L;
    if (pair_buffers(c, plan.slots, npairs, 0, 0, L)) return 1;
    if (pair_upload_plan(L, plan, ti, ci, c->st)) return 1;
    PHX_CK(hipMemcpyAsync(L.ps, ps.data(), 96 * n, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(L.steps, steps.data(), 4 * n, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(L.dts, dts.data(), 8 * n, hipMemcpyHostToDevice, c->st));
    PHX_CK(hipMemcpyAsync(L.m, m.data(), 4 * (size_t)npairs, hipMemcpyHostToDevice, c->st));
    stage_begin(c);
    if (launch_clones(form, L.ps, L.steps, L.dts, L.res, c->P, plan.slots, c->pair_bucket, c->pair_lanes, L.next,
                      c->st))
        return 1;
    stage_mark(c, "clones");
    if (launch_pair_reduce(L.res, L.off, L.m, L.r50, L.rt, npairs, n_clones, L.out, c->st)) return 1;
    stage_mark(c, "reduce");
    PHX_CK(hipMemcpyAsync(res, L.out, sizeof(PhResult) * (size_t)npairs, hipMemcpyDeviceToHost, c->st));
    PHX_CK(hipStreamSynchronize(c->st));
    return 0;
}
