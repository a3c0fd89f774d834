// Shared host-side helpers of the HIP library: error checks and the launcher interface between
// translation units (each .hip file holds its own kernels; host code calls them through these).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "physics.h"
#include "phygas_hip.h"

namespace phx {

// Record a failed HIP call (file:line and the runtime's message) in a thread-local buffer, print it to
// stderr, and make the enclosing function return its error code.
void set_error(const char *where, const char *what);
const char *last_error();

#define PHX_CK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            char w_[256];                                                                          \
            snprintf(w_, sizeof w_, "%s:%d %s", __FILE__, __LINE__, #x);                           \
            ::phx::set_error(w_, hipGetErrorString(e_));                                           \
            return 1;                                                                              \
        }                                                                                          \
    } while (0)

#define PHX_REQUIRE(cond, msg)                                                                     \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            ::phx::set_error(__func__, msg);                                                       \
            return 2;                                                                              \
        }                                                                                          \
    } while (0)

inline unsigned blocks_for(long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// ---- plain FP64 launchers (plain.hip) -----------------------------------------------------------
// Draws of every object: the start state ps0[n][6] (a copy of the host's nominal nom[n][6]) and the 6x6
// covariance sig[n][36] of n_samp draws about it.
int launch_lin_init_targets(const PhTarget *d_t, const uint64_t *d_seed, const double *d_nom, int n, int n_samp,
                            const Pot &P, double *d_ps0, double *d_sig, hipStream_t st);
int launch_lin_init_clusters(const PhCluster *d_c, const uint64_t *d_seed, const double *d_nom, int n, int n_samp,
                             const Pot &P, double *d_ps0, double *d_sig, hipStream_t st);
// Nominal start states of n records on the host with the C library's transcendentals, as the reference
// draws them (reference: src/phygas.c:50-82 with rng == NULL): out[n][6]; zeros where every retry fails.
void nominal_states(const Pot &P, const PhTarget *t, int n, double *out);
void nominal_states(const Pot &P, const PhCluster *c, int n, double *out);
// Nominal + 6 neighbour orbits, points [k0, k0+nk).  soa_nc == 0: track layout [obj][npts][12];
// soa_nc > 0: cluster layout [point][12][soa_nc].  d_state (63 doubles/object) may be null when k0 == 0
// and nk covers every point.
// d_nobj (nullable): points of each object (fewer than npts stop early; the rest of the slot is unused).
int launch_lin_plain(const double *d_ps0, const double *d_sig, double *d_state, double *d_trk, const Pot &P,
                     double eps_pos, double eps_vel, double dt, int sub, int k0, int nk, int npts, int n, int soa_nc,
                     const int *d_nobj, hipStream_t st);
// Full chi2 scan of pairs [p0, p0+np) of nb stars x nc clusters.
int launch_chi2_plain(const double *d_trk, const double *d_csoa, const int *d_cn, const double *d_R2, int nb, int nc,
                      int npts, long p0, long np, double *d_chi2, int *d_kb, double *d_sb, hipStream_t st);
// 512-draw finish of selected pairs [h0, h0+nh): out[h][6] = t, d_nom_pc, sig_pc, p_r50, p_rtot, vrel.
int launch_finish_plain(const double *d_trk, const double *d_csoa, const int *d_cn, const double *d_r50,
                        const double *d_rt, int nc, int npts, const int *d_hs, const int *d_hc, const int *d_hkb,
                        const double *d_hsb, const uint64_t *d_hseed, int h0, int nh, double *d_out, hipStream_t st);
// Monte Carlo clones: steps (s0, s0+nk] of clones [0, n).
int launch_pair_plain(const double *d_ps, const int *d_steps, const double *d_dts, double *d_st, double *d_res,
                      const Pot &P, int s0, int nk, int n, hipStream_t st);
// Finish of the pairs selected into idx[0 .. min(*count, cap)) (pair id = star * nc + cluster).
int launch_finish_sel_plain(const double *d_trk, const double *d_csoa, const int *d_cn, const double *d_r50,
                            const double *d_rt, const PhCluster *d_cl, const uint64_t *d_sprefix, int nc, int npts,
                            const long *d_idx, const long *d_count, long cap, const int *d_kb, const double *d_sb,
                            double *d_out, hipStream_t st);
// Rows of finished pairs selected into ridx[0 .. min(*rcount, rcap)).
int launch_rows(const long *d_pair_idx, const long *d_ridx, const long *d_rcount, long rcap, int nc,
                const int *d_sgroup, const double *d_chi2, const int *d_kb, const double *d_sb, const double *d_fout,
                PhxRow *d_rows, hipStream_t st);
// Pair-mode prefilter of nt targets x nc clusters: nominal paths with 32-draw spreads (seeds "track"|name),
// mask[j * nc + i] = 1 when target j may reach cluster i.  Paths [obj][npts][3].
int launch_prefilter(const PhTarget *d_t, const uint64_t *d_tseed, const int *d_tn, int nt, const PhCluster *d_c,
                     const uint64_t *d_cseed, const int *d_cn, int nc, int npts, const Pot &P, double *d_tpath,
                     double *d_tspr, double *d_cpath, double *d_cspr, unsigned char *d_mask, hipStream_t st);
// ---- record files (wu.hip)
// Cluster line (19 fields) and target line (14 fields, then an optional clusters=a,b,... list into *only), with
// the reference's field widths (reference: src/phygas.c:460-506); false for a line that does not parse.
bool parse_cluster(const char *line, PhCluster *c);
bool parse_target(const char *line, PhTarget *t, std::string *only);
// Every line of a file (newlines kept; a last line without one too); false when it cannot be opened.
bool read_lines(const char *path, std::vector<std::string> &lines);

// ---- device-wide selection (select.hip) ---------------------------------------------------------
// Indices i in [0, n) with val[i] < thr (strictly; NaN never), in increasing order, into d_idx[0 .. cap);
// the full count into *d_count (device; may exceed cap, then only cap indices are written).
// d_tmp must hold select_tmp_bytes(n) bytes.
size_t select_tmp_bytes(long n);
int launch_select_less(const double *d_val, long n, double thr, long *d_idx, long cap, long *d_count, void *d_tmp,
                       hipStream_t st);
// The same over i < min(*d_n, n_cap) with val[i * stride] >= thr.
int launch_select_geq_counted(const double *d_val, long stride, const long *d_n, long n_cap, double thr, long *d_idx,
                              long cap, long *d_count, void *d_tmp, hipStream_t st);

}  // namespace phx
