// Work-unit runner: the reference app's task (reference: src/main_phygas_boinc.cpp:336-471) over the batched
// engine.  Reads the `in` file (header "PHYGAS 1 <screen|pair> p1 p2 p3 p4", then one target per line with an
// optional clusters=a,b,... list) and the cluster file, runs the targets in batches on the GPU, and writes the
// app's rows to `out` in target order (clusters in file order), then "# done <nt>".
//
// Checkpoint/resume keeps `out` consistent with the checkpoint (the fixes the reference's own issue notes
// propose for its app: rows after the last checkpoint duplicated on resume, a checkpoint written in place and
// not range-checked, an unopenable checkpoint restarting from zero):
//   * after every batch, "phygas_checkpoint 2 <targets_done> <lines_done>" is written to a temporary file,
//     flushed and checked, then renamed over the checkpoint;
//   * on resume, `out` keeps exactly its first <lines_done> complete lines (torn or extra rows dropped);
//   * a checkpoint that does not parse, is of another version, is out of range, or counts more lines than
//     `out` holds means a fresh start with `out` truncated;
//   * a checkpoint that exists but cannot be read leaves everything alone and reports a temporary failure.
// Every row depends only on its seeds, so recomputed rows are byte-identical to the dropped ones.
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <string>
#include <vector>

#include "common.h"

namespace phx {
namespace {

// Whitespace tokens with sscanf-like field widths: a %Ns field takes at most N characters and the rest of the
// token starts the next field (reference: the %39s / %3s / %31s conversions of src/phygas.c:462-506).
std::vector<std::string> tokens(const char *line, const int *widths, int nw) {
    std::vector<std::string> out;
    const char *p = line;
    while (*p) {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\v' || *p == '\f') p++;
        if (!*p) break;
        const char *q = p;
        while (*q && !(*q == ' ' || *q == '\t' || *q == '\r' || *q == '\n' || *q == '\v' || *q == '\f')) q++;
        std::string tok(p, q - p);
        for (;;) {
            const int f = (int)out.size();
            const int w = f < nw ? widths[f] : 0;
            if (w > 0 && (int)tok.size() > w) {
                out.push_back(tok.substr(0, w));
                tok = tok.substr(w);
                continue;
            }
            out.push_back(tok);
            break;
        }
        p = q;
    }
    return out;
}

bool num(const std::string &s, double *v, bool nan_ok) {
    if (nan_ok && s == "nan") {
        *v = NAN;
        return true;
    }
    char *end = nullptr;
    *v = strtod(s.c_str(), &end);
    return end && *end == 0 && !s.empty();
}

void copy_str(char *dst, size_t cap, const std::string &s) {
    snprintf(dst, cap, "%s", s.c_str());
}

}  // namespace

// Cluster line: index name kind ra dec dist e_dist pmra e_pmra pmdec e_pmdec rv e_rv r50 rtot age16 age50 age84
// tmax (reference: src/phygas.c:462-485).  Lines that do not parse are skipped.
bool parse_cluster(const char *line, PhCluster *c) {
    static const int w[19] = {0, 39, 3, 0, 0, 0, 0, 0, 0, 0, 0, 31, 31, 0, 0, 0, 0, 0, 0};
    const std::vector<std::string> t = tokens(line, w, 19);
    if (t.size() < 19) return false;
    double v[19];
    char *end = nullptr;
    strtol(t[0].c_str(), &end, 10);
    if (!end || *end) return false;
    for (int i = 3; i < 19; i++)
        if (!num(t[i], &v[i], i == 11 || i == 12)) return false;
    memset(c, 0, sizeof *c);
    copy_str(c->name, sizeof c->name, t[1]);
    c->kind = t[2][0];
    c->ra = v[3]; c->dec = v[4]; c->dist_pc = v[5]; c->e_dist_pc = v[6];
    c->pmra = v[7]; c->e_pmra = v[8]; c->pmdec = v[9]; c->e_pmdec = v[10];
    c->rv = v[11]; c->e_rv = v[12];
    c->has_rv = !isnan(c->rv) && !isnan(c->e_rv);
    c->r50_pc = v[13]; c->rtot_pc = v[14]; c->age16 = v[15]; c->age50 = v[16]; c->age84 = v[17]; c->tmax_myr = v[18];
    return true;
}

// Target line of the `in` file: id kind ra dec plx e_plx dist e_lndist pmra e_pmra pmdec e_pmdec rv e_rv
// [clusters=a,b,...] (reference: src/main_phygas_boinc.cpp:309-327).
bool parse_target(const char *line, PhTarget *t, std::string *only) {
    static const int w[14] = {39, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 31, 31};
    const std::vector<std::string> k = tokens(line, w, 14);
    if (k.size() < 14) return false;
    double v[14];
    for (int i = 2; i < 14; i++)
        if (!num(k[i], &v[i], i >= 12)) return false;
    memset(t, 0, sizeof *t);
    copy_str(t->id, sizeof t->id, k[0]);
    t->kind = k[1][0];
    t->ra = v[2]; t->dec = v[3]; t->plx = v[4]; t->e_plx = v[5]; t->dist_pc = v[6]; t->e_lndist = v[7];
    t->pmra = v[8]; t->e_pmra = v[9]; t->pmdec = v[10]; t->e_pmdec = v[11];
    t->rv = v[12]; t->e_rv = v[13];
    t->has_rv = !isnan(t->rv) && !isnan(t->e_rv);
    only->clear();
    for (size_t i = 14; i < k.size(); i++)
        if (k[i].compare(0, 9, "clusters=") == 0) {
            *only = k[i].substr(9);
            break;
        }
    return true;
}

namespace {

// name is one element of the comma-separated list
bool in_list(const char *name, const std::string &list) {
    size_t a = 0;
    while (a <= list.size()) {
        size_t b = list.find(',', a);
        if (b == std::string::npos) b = list.size();
        if (list.compare(a, b - a, name) == 0 && strlen(name) == b - a) return true;
        a = b + 1;
    }
    return false;
}

}  // namespace

bool read_lines(const char *path, std::vector<std::string> &lines) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    std::string cur;
    int ch;
    while ((ch = fgetc(f)) != EOF) {
        cur.push_back((char)ch);
        if (ch == '\n') {
            lines.push_back(cur);
            cur.clear();
        }
    }
    if (!cur.empty()) lines.push_back(cur);   // a last line without its newline
    fclose(f);
    return true;
}

namespace {

bool file_exists(const char *path) {
    struct stat st;
    return stat(path, &st) == 0;
}

// Whole-file replacement through a temporary file; every write checked.
bool write_atomic(const char *path, const std::string &data) {
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return false;
    bool ok = fwrite(data.data(), 1, data.size(), f) == data.size();
    ok = fflush(f) == 0 && ok;
    ok = fclose(f) == 0 && ok;
    if (!ok || rename(tmp.c_str(), path) != 0) {
        remove(tmp.c_str());
        return false;
    }
    return true;
}

}  // namespace
}  // namespace phx

using namespace phx;

extern "C" int phx_wu_resume(const char *out_path, const char *ckpt_path, long nt, long *targets_done,
                             long *lines_kept) {
    *targets_done = 0;
    *lines_kept = 0;
    long t = 0, l = 0;
    bool resume = false;
    if (file_exists(ckpt_path)) {
        FILE *f = fopen(ckpt_path, "rb");
        if (!f) {
            set_error("phx_wu_resume", "the checkpoint exists but cannot be read: retry later (out left alone)");
            return PHX_WU_CKPT_UNREADABLE;
        }
        char buf[256] = {0};
        const size_t n = fread(buf, 1, sizeof buf - 1, f);
        const bool err = ferror(f) != 0;
        fclose(f);
        if (err) {
            set_error("phx_wu_resume", "the checkpoint exists but cannot be read: retry later (out left alone)");
            return PHX_WU_CKPT_UNREADABLE;
        }
        buf[n] = 0;
        char tag[32] = {0};
        int ver = 0;
        if (sscanf(buf, "%31s %d %ld %ld", tag, &ver, &t, &l) == 4 && strcmp(tag, "phygas_checkpoint") == 0 &&
            ver == 2 && t >= 0 && t <= nt && l >= 0)
            resume = true;
    }
    std::vector<std::string> lines;
    if (resume) {
        if (!read_lines(out_path, lines)) lines.clear();
        long complete = 0;
        for (const std::string &s : lines) {
            if (!s.empty() && s.back() == '\n') complete++;
            else break;
        }
        if (complete < l) resume = false;   // out is missing rows the checkpoint counts: start over
    }
    std::string keep;
    if (resume)
        for (long i = 0; i < l; i++) keep += lines[i];
    if (!write_atomic(out_path, keep)) {
        set_error("phx_wu_resume", "cannot rewrite the output file");
        return 5;
    }
    if (resume) {
        *targets_done = t;
        *lines_kept = l;
    }
    return 0;
}

// Defaults give the reference's results: screening rows byte-identical to the plain FP64 form (FAST64: the
// plain tracks, the exact culled scan and the screened finish), pair rows from FP64 clones.  The FP32 forms
// change results in the last printed digits, so they are opt-in, as the reference keeps result-changing
// variants behind a switch (its replicas are compared by a validator).
extern "C" void phx_wu_defaults(PhxWuOptions *o) {
    memset(o, 0, sizeof *o);
    o->form = PHX_FORM_FAST64;
    o->batch_targets = 20000;
    o->max_batches = 0;
    o->device = 0;
    o->draw_on_device = 0;
    o->pair_form = PHX_FORM_PLAIN;
}

namespace {

struct Out {
    FILE *f = nullptr;
    long lines = 0;
    bool ok = true;
    void row(const char *s) {
        ok = ok && fputs(s, f) >= 0;
        lines++;
    }
};

int checkpoint(Out &o, const char *ckpt, long done) {
    if (fflush(o.f) != 0 || !o.ok) {
        set_error("phx_wu_run", "writing the output file failed");
        return 5;
    }
    char buf[96];
    snprintf(buf, sizeof buf, "phygas_checkpoint 2 %ld %ld\n", done, o.lines);
    if (!write_atomic(ckpt, buf)) {
        set_error("phx_wu_run", "writing the checkpoint failed");
        return 5;
    }
    return 0;
}

}  // namespace

extern "C" int phx_wu_run(const char *in_path, const char *clusters_path, const char *out_path, const char *ckpt_path,
                          const PhxWuOptions *opt, long *targets_done) {
    PhxWuOptions o;
    if (opt) o = *opt; else phx_wu_defaults(&o);
    *targets_done = 0;
    // clusters
    std::vector<PhCluster> cl;
    {
        std::vector<std::string> lines;
        PHX_REQUIRE(read_lines(clusters_path, lines), "cannot read the cluster file");
        for (const std::string &s : lines) {
            PhCluster c;
            if (parse_cluster(s.c_str(), &c)) cl.push_back(c);
        }
        PHX_REQUIRE(!cl.empty(), "no cluster in the cluster file");
    }
    // header and targets
    std::vector<PhTarget> tg;
    std::vector<std::string> only;
    char mode[16] = {0};
    double p[4] = {0, 0, 0, 0};
    {
        std::vector<std::string> lines;
        PHX_REQUIRE(read_lines(in_path, lines) && !lines.empty(), "cannot read the input file");
        PHX_REQUIRE(sscanf(lines[0].c_str(), "PHYGAS 1 %15s %lf %lf %lf %lf", mode, &p[0], &p[1], &p[2], &p[3]) >= 3,
                    "bad header");
        PHX_REQUIRE(!strcmp(mode, "screen") || !strcmp(mode, "pair"), "unknown mode");
        for (size_t i = 1; i < lines.size(); i++) {
            PhTarget t;
            std::string lst;
            if (!parse_target(lines[i].c_str(), &t, &lst)) continue;
            tg.push_back(t);
            only.push_back(lst);
        }
    }
    const bool screen = strcmp(mode, "screen") == 0;
    const long nt = (long)tg.size();
    long start = 0, lines0 = 0;
    if (int rc = phx_wu_resume(out_path, ckpt_path, nt, &start, &lines0)) return rc;
    Out out;
    out.f = fopen(out_path, "ab");
    PHX_REQUIRE(out.f, "cannot open the output file");
    out.lines = lines0;
    PhxCtx *ctx = nullptr;
    if (int rc = phx_create(o.device, &ctx)) { fclose(out.f); return rc; }
    int rc = 0;
    const int nc = (int)cl.size();
    const long B = o.batch_targets > 0 ? o.batch_targets : 20000;
    int batches = 0;
    char line[512];
    if (screen) {
        const double chi2_max = p[0], p_min = p[3];
        const int n_samp = (int)p[1];
        rc = phx_load_clusters(ctx, cl.data(), nc, n_samp, p[2] > 0 ? p[2] : 2.0);
        std::vector<PhxRow> rows;
        for (long b0 = start; rc == 0 && b0 < nt; b0 += B) {
            if (o.max_batches > 0 && batches == o.max_batches) { rc = PHX_WU_STOPPED; break; }
            const int nb = (int)std::min(B, nt - b0);
            long n = 0;
            if ((rc = phx_clear_stars(ctx)) || (rc = phx_load_stars(ctx, tg.data() + b0, nb, 0)) ||
                (rc = phx_screen(ctx, o.form, chi2_max, n_samp, p_min)) || (rc = phx_rows(ctx, nullptr, 0, &n)))
                break;
            rows.resize(n);
            if (n > 0 && (rc = phx_rows(ctx, rows.data(), n, &n))) break;
            for (const PhxRow &r : rows) {   // ordered by (star, cluster): targets in order, clusters in file order
                const PhTarget &t = tg[b0 + r.star];
                const PhCluster &c = cl[r.cluster];
                snprintf(line, sizeof line, "%s %s %c %.3f %.3f %.2f %.2f %.4f %.4f %.2f\n", t.id, c.name, c.kind, r.chi2,
                         r.t, r.d_nom_pc, r.sig_pc, r.p_r50, r.p_rtot, r.vrel);
                out.row(line);
            }
            if ((rc = checkpoint(out, ckpt_path, b0 + nb))) break;
            *targets_done = b0 + nb;
            batches++;
        }
    } else {
        const int n_clones = (int)p[0];
        const double dt_myr = p[1];
        std::vector<unsigned char> mask;
        std::vector<int> ti, ci;
        std::vector<PhResult> res;
        for (long b0 = start; rc == 0 && b0 < nt; b0 += B) {
            if (o.max_batches > 0 && batches == o.max_batches) { rc = PHX_WU_STOPPED; break; }
            const int nb = (int)std::min(B, nt - b0);
            mask.assign((size_t)nb * nc, 0);
            if ((rc = phx_prefilter(ctx, tg.data() + b0, nb, cl.data(), nc, mask.data()))) break;
            ti.clear();
            ci.clear();
            for (int j = 0; j < nb; j++)
                for (int i = 0; i < nc; i++) {
                    const std::string &lst = only[b0 + j];
                    if (lst.empty() ? mask[(size_t)j * nc + i] != 0 : in_list(cl[i].name, lst)) {
                        ti.push_back((int)(b0 + j));
                        ci.push_back(i);
                    }
                }
            res.resize(ti.size());
            if (!ti.empty() && (rc = phx_pair_ex(ctx, o.pair_form, o.draw_on_device ? PHX_DRAW_ON_DEVICE : 0, tg.data(),
                                                 cl.data(), (int)ti.size(), ti.data(), ci.data(), n_clones, dt_myr,
                                                 res.data())))
                break;
            for (size_t k = 0; k < ti.size(); k++) {
                const PhTarget &t = tg[ti[k]];
                const PhCluster &c = cl[ci[k]];
                const PhResult &r = res[k];
                snprintf(line, sizeof line, "%s %s %c %d %.3f %.3f %.3f %.4f %.4f %.4f %.2f %.2f %.2f %.4f %.4f\n", t.id,
                         c.name, c.kind, r.n, r.d05, r.d50, r.d95, r.t05, r.t50, r.t95, r.v05, r.v50, r.v95, r.f_r50,
                         r.f_rtot);
                out.row(line);
            }
            if ((rc = checkpoint(out, ckpt_path, b0 + nb))) break;
            *targets_done = b0 + nb;
            batches++;
        }
    }
    if (rc == 0) {
        snprintf(line, sizeof line, "# done %ld\n", nt);
        out.row(line);
        if (fflush(out.f) != 0 || !out.ok) rc = 5;
        *targets_done = nt;
    }
    if (fclose(out.f) != 0 && rc == 0) rc = 5;
    phx_destroy(ctx);
    return rc;
}
