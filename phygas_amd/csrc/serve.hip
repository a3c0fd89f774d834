// Serving loop: screening chunks and Monte Carlo pair jobs submitted as they arrive, batched continuously
// onto the GPU.
//
// Slots.  Each slot is a host thread driving its own context (its own stream, default priority): screening
// slots hold the cluster tables resident in HBM; pair slots trace clones drawn on the device.  Several slots
// keep several batches in flight on the device at once, so a request that arrives while others run starts
// as soon as a slot and memory free up, and its kernels overlap theirs.
//
// Batching.  A free slot takes every queued request of its kind, oldest first, that fits its batch limit
// (stars, or clones) and the HBM budget (the footprint estimate below), merges them into one pass (one
// star batch; one clone batch with per-pair seeds) and splits the results back per request.  Every result
// depends only on its own inputs and per-pair seeds, so a request's rows are the same alone or in any mix.
//
// Admission.  A request's footprint is estimated from its size (screening: tracks, tiles and candidate
// buffers per star; pairs: 300 bytes per clone slot, about twice the Monte Carlo layout's 132); a batch is
// started only while the footprints of the batches in flight stay under the budget.  A single request larger
// than the budget runs alone.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.h"

using namespace phx;

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
    long id = 0;
    int kind = 0;   // 0 screening, 1 pairs
    std::vector<PhTarget> stars;
    PhTarget target{};
    std::vector<int> cidx;
    int n_clones = 0;
    double dt = 0.05;
    size_t bytes = 0;
    double t_arrive = 0, t_start = 0, t_done = 0;
    int status = 0;   // 0 queued, 1 running, 2 done, < 0 failed
    int batch_reqs = 0;
    std::vector<PhxRow> rows;
    std::vector<PhResult> results;
};

}  // namespace

struct PhxServer {
    int device = 0;
    PhxServeConfig cfg{};
    std::vector<PhCluster> clusters;
    int npts = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> q[2];
    std::unordered_map<long, std::unique_ptr<Req>> reqs;
    size_t in_flight = 0;
    long next_id = 0;
    bool stop = false;
    int ready = 0, failed = 0;
    std::string error;
    std::vector<std::thread> workers;
};

static size_t screen_bytes(const PhxServer *s, int ns) {
    const size_t nc = s->clusters.size();
    return (size_t)ns * ((size_t)s->npts * 112 + 2048) + (size_t)ns * nc * 8;
}
static size_t pair_bytes(int npairs, int n_clones) { return (size_t)npairs * n_clones * 300; }

// Take the batch a free slot of this kind runs next (caller holds the lock); empty when nothing fits.
static std::vector<Req *> take_batch(PhxServer *s, int kind) {
    std::vector<Req *> b;
    auto &q = s->q[kind];
    long units = 0;
    size_t bytes = 0;
    const long limit = kind == 0 ? s->cfg.max_batch_stars : s->cfg.max_batch_clones;
    while (!q.empty()) {
        Req *r = q.front();
        if (kind == 1 && !b.empty() && (r->n_clones != b[0]->n_clones || r->dt != b[0]->dt)) break;
        const long u = kind == 0 ? (long)r->stars.size() : (long)r->cidx.size() * r->n_clones;
        const bool alone = b.empty() && s->in_flight == 0;   // oversized requests still run, alone
        if (!alone && (units + u > limit || s->in_flight + bytes + r->bytes > s->cfg.hbm_budget)) break;
        q.pop_front();
        b.push_back(r);
        units += u;
        bytes += r->bytes;
    }
    return b;
}

static void finish_batch(PhxServer *s, std::vector<Req *> &b, int rc) {
    std::lock_guard<std::mutex> lk(s->mu);
    const double t = now_s();
    for (Req *r : b) {
        r->status = rc == 0 ? 2 : -1;
        r->t_done = t;
        r->batch_reqs = (int)b.size();
        s->in_flight -= r->bytes;
    }
    if (rc) s->error = last_error();
    s->cv_done.notify_all();
    s->cv_work.notify_all();
}

static void screen_slot(PhxServer *s) {
    PhxCtx *ctx = nullptr;
    int rc = phx_create(s->device, &ctx);
    if (!rc) rc = phx_load_clusters(ctx, s->clusters.data(), (int)s->clusters.size(), s->cfg.n_samp, s->cfg.r_kernel);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (rc) { s->failed++; s->error = last_error(); } else s->ready++;
        s->cv_done.notify_all();
    }
    if (rc) { phx_destroy(ctx); return; }
    std::vector<PhxRow> rows;
    for (;;) {
        std::vector<Req *> b;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                if (s->stop) return true;
                b = take_batch(s, 0);
                return !b.empty();
            });
            if (s->stop && b.empty()) break;
            const double t = now_s();
            for (Req *r : b) { r->status = 1; r->t_start = t; s->in_flight += r->bytes; }
        }
        std::vector<int> first(b.size() + 1, 0);
        rc = phx_clear_stars(ctx);
        for (size_t i = 0; i < b.size() && !rc; i++) {
            rc = phx_load_stars(ctx, b[i]->stars.data(), (int)b[i]->stars.size(), 0);
            first[i + 1] = first[i] + (int)b[i]->stars.size();
        }
        long n = 0;
        if (!rc) rc = phx_screen(ctx, s->cfg.form, s->cfg.chi2_max, s->cfg.n_samp, s->cfg.p_min);
        if (!rc) rc = phx_rows(ctx, nullptr, 0, &n);
        if (!rc) {
            rows.resize(n);
            rc = phx_rows(ctx, rows.data(), n, &n);
        }
        if (!rc) {   // rows are ordered by star: split at the requests' star offsets
            size_t i = 0;
            for (Req *r : b) r->rows.clear();
            for (const PhxRow &row : rows) {
                while (i + 1 < b.size() && row.star >= first[i + 1]) i++;
                PhxRow o = row;
                o.star -= first[i];
                b[i]->rows.push_back(o);
            }
        }
        finish_batch(s, b, rc);
    }
    phx_destroy(ctx);
}

static void pair_slot(PhxServer *s) {
    PhxCtx *ctx = nullptr;
    int rc = phx_create(s->device, &ctx);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (rc) { s->failed++; s->error = last_error(); } else s->ready++;
        s->cv_done.notify_all();
    }
    if (rc) return;
    for (;;) {
        std::vector<Req *> b;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                if (s->stop) return true;
                b = take_batch(s, 1);
                return !b.empty();
            });
            if (s->stop && b.empty()) break;
            const double t = now_s();
            for (Req *r : b) { r->status = 1; r->t_start = t; s->in_flight += r->bytes; }
        }
        std::vector<PhTarget> tg(b.size());
        std::vector<int> ti, ci;
        for (size_t j = 0; j < b.size(); j++) {
            tg[j] = b[j]->target;
            for (int c : b[j]->cidx) { ti.push_back((int)j); ci.push_back(c); }
        }
        std::vector<PhResult> res(ti.size());
        rc = phx_pair_ex(ctx, s->cfg.form == PHX_FORM_PLAIN ? PHX_FORM_PLAIN : PHX_FORM_FAST, PHX_DRAW_ON_DEVICE,
                         tg.data(), s->clusters.data(), (int)ti.size(), ti.data(), ci.data(), b[0]->n_clones, b[0]->dt,
                         res.data());
        if (!rc) {
            size_t o = 0;
            for (Req *r : b) {
                r->results.assign(res.begin() + o, res.begin() + o + r->cidx.size());
                o += r->cidx.size();
            }
        }
        finish_batch(s, b, rc);
    }
    phx_destroy(ctx);
}

extern "C" void phx_serve_defaults(PhxServeConfig *c) {
    c->form = PHX_FORM_FAST;
    c->screen_slots = 2;
    c->pair_slots = 1;
    c->hbm_budget = (size_t)64 << 30;
    c->max_batch_stars = 40000;
    c->max_batch_clones = 1L << 20;
    c->chi2_max = 25.0;
    c->p_min = 0.001;
    c->n_samp = 64;
    c->r_kernel = 2.0;
}

extern "C" int phx_server_create(int device, const PhCluster *cl, int nc, const PhxServeConfig *cfg, PhxServer **out) {
    *out = nullptr;
    PHX_REQUIRE(nc > 0, "no clusters");
    PHX_REQUIRE(cfg->screen_slots >= 1 && cfg->pair_slots >= 0, "need at least one screening slot");
    PhxServer *s = new PhxServer();
    s->device = device;
    s->cfg = *cfg;
    s->clusters.assign(cl, cl + nc);
    double tmax = 0;
    for (int i = 0; i < nc; i++) tmax = std::max(tmax, cl[i].tmax_myr);
    s->npts = (int)ceil(tmax / STEP_MYR) + 1;
    for (int i = 0; i < cfg->screen_slots; i++) s->workers.emplace_back(screen_slot, s);
    for (int i = 0; i < cfg->pair_slots; i++) s->workers.emplace_back(pair_slot, s);
    {
        std::unique_lock<std::mutex> lk(s->mu);
        const int want = cfg->screen_slots + cfg->pair_slots;
        s->cv_done.wait(lk, [&] { return s->ready + s->failed == want; });
        if (s->failed) {
            std::string e = s->error;
            lk.unlock();
            phx_server_destroy(s);
            set_error("phx_server_create", e.c_str());
            return 1;
        }
    }
    *out = s;
    return 0;
}

extern "C" void phx_server_destroy(PhxServer *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    for (auto &t : s->workers) t.join();
    delete s;
}

static long submit(PhxServer *s, std::unique_ptr<Req> r) {
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->t_arrive = now_s();
    Req *p = r.get();
    s->reqs[p->id] = std::move(r);
    s->q[p->kind].push_back(p);
    s->cv_work.notify_all();
    return p->id;
}

extern "C" long phx_server_submit_screen(PhxServer *s, const PhTarget *stars, int ns) {
    if (ns <= 0 || !stars) { set_error("phx_server_submit_screen", "empty request"); return -2; }
    std::unique_ptr<Req> r(new Req());
    r->kind = 0;
    r->stars.assign(stars, stars + ns);
    r->bytes = screen_bytes(s, ns);
    return submit(s, std::move(r));
}

extern "C" long phx_server_submit_pair(PhxServer *s, const PhTarget *target, const int *cluster_idx, int n,
                                       int n_clones, double dt_myr) {
    if (n <= 0 || n_clones <= 0) { set_error("phx_server_submit_pair", "bad request size"); return -2; }
    if (s->cfg.pair_slots <= 0) { set_error("phx_server_submit_pair", "the server has no pair slot"); return -2; }
    for (int i = 0; i < n; i++)
        if (cluster_idx[i] < 0 || cluster_idx[i] >= (int)s->clusters.size()) {
            set_error("phx_server_submit_pair", "cluster index out of range");
            return -2;
        }
    std::unique_ptr<Req> r(new Req());
    r->kind = 1;
    r->target = *target;
    r->cidx.assign(cluster_idx, cluster_idx + n);
    r->n_clones = n_clones;
    r->dt = dt_myr;
    r->bytes = pair_bytes(n, n_clones);
    return submit(s, std::move(r));
}

extern "C" int phx_server_wait(PhxServer *s, long id, double timeout_s) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    PHX_REQUIRE(it != s->reqs.end(), "unknown request");
    Req *r = it->second.get();
    auto done = [&] { return r->status == 2 || r->status < 0; };
    if (timeout_s < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) return 1;
    if (r->status < 0) {
        set_error("phx_server_wait", s->error.c_str());
        return -1;
    }
    return 0;
}

extern "C" int phx_server_screen_rows(PhxServer *s, long id, PhxRow *rows, long cap, long *n) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    PHX_REQUIRE(it != s->reqs.end() && it->second->kind == 0, "unknown screening request");
    Req *r = it->second.get();
    PHX_REQUIRE(r->status == 2, "request not done");
    *n = (long)r->rows.size();
    const long m = std::min(cap, *n);
    if (m > 0) memcpy(rows, r->rows.data(), sizeof(PhxRow) * m);
    return 0;
}

extern "C" int phx_server_pair_results(PhxServer *s, long id, PhResult *res, int cap) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    PHX_REQUIRE(it != s->reqs.end() && it->second->kind == 1, "unknown pair request");
    Req *r = it->second.get();
    PHX_REQUIRE(r->status == 2, "request not done");
    const int m = std::min(cap, (int)r->results.size());
    if (m > 0) memcpy(res, r->results.data(), sizeof(PhResult) * m);
    return 0;
}

extern "C" int phx_server_info(PhxServer *s, long id, double t[3], int *batch_reqs) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    PHX_REQUIRE(it != s->reqs.end(), "unknown request");
    Req *r = it->second.get();
    t[0] = r->t_arrive;
    t[1] = r->t_start;
    t[2] = r->t_done;
    if (batch_reqs) *batch_reqs = r->batch_reqs;
    return r->status;
}

extern "C" int phx_server_release(PhxServer *s, long id) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    PHX_REQUIRE(it != s->reqs.end(), "unknown request");
    PHX_REQUIRE(it->second->status == 2 || it->second->status < 0, "request still queued or running");
    s->reqs.erase(it);
    return 0;
}
