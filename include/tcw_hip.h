/*
 * tcw_hip.h -- C ABI of the MI355X (gfx950) transient-CW F-statistic map library
 * (libtcw_hip.so).  Plain C types and pointers only; device buffers are raw HIP
 * device pointers owned by the caller, streams are hipStream_t passed as void*.
 *
 * The library replaces, one for one, the device-side work of PyFstat's
 * tCWFstatMapVersion="pycuda" backend:
 *   pyfstat/tcw_fstat_map_funcs.py:656-834  pycuda_compute_transient_fstat_map
 *   pyfstat/tcw_fstat_map_funcs.py:837-916  ..._rect (H2D, kernel, D2H)
 *   pyfstat/tcw_fstat_map_funcs.py:919-998  ..._exp
 *   pyfstat/pyCUDAkernels/cudaTransientFstatRectWindow.cu:1-130
 *   pyfstat/pyCUDAkernels/cudaTransientFstatExpWindow.cu:1-138
 * plus the host-side steps the reference runs around them on the CPU:
 *   lalpulsar.mergeMultiFstatAtomsBinned (called at tcw_fstat_map_funcs.py:706),
 *   max/argmax (tcw_fstat_map_funcs.py:809-815),
 *   get_lnBtSG / get_t0_max_posterior / get_tau_max_posterior
 *   (tcw_fstat_map_funcs.py:196-287, called at 824-828).
 *
 * Error convention: every entry point returns 0 (TCW_OK) or a TCW_ERR_* code;
 * tcw_last_error() then gives a message (thread-local, valid until the next
 * call on the same thread).
 *
 * Threading: one handle per host thread / stream.  Calls on one handle are not
 * re-entrant; distinct handles are independent.
 */
#ifndef TCW_HIP_H
#define TCW_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCW_NCOL 7          /* a2, b2, ab, Fa_re, Fa_im, Fb_re, Fb_im */
#define TCW_MAX_DETECTORS 10 /* lalpulsar PULSAR_MAX_DETECTORS */

/* lalpulsar transientWindowType_t (pyfstat/core.py:829-833) */
#define TCW_TRANSIENT_NONE 0u
#define TCW_TRANSIENT_RECTANGULAR 1u
#define TCW_TRANSIENT_EXPONENTIAL 2u
#define TCW_TRANSIENT_LAST 3u

/* return codes */
#define TCW_OK 0
#define TCW_ERR_INVALID 1     /* bad argument (maps to ValueError)            */
#define TCW_ERR_DEVICE 2      /* HIP runtime failure (maps to RuntimeError)   */
#define TCW_ERR_SINGLE_ATOM 3 /* lal-strict mode: a window spans one atom     */
#define TCW_ERR_NOMEM 4       /* device allocation failed                     */

/* flags */
#define TCW_FLAG_BTSG 1u       /* also lnBtSG and the (t0, tau) posterior modes */
#define TCW_FLAG_LAL_STRICT 2u /* fail like lalpulsar (XLAL_EDOM) on i_t1==i_t0 */
#define TCW_FLAG_GENERIC 4u    /* force the generic per-cell kernels (testing)  */

/* transientWindowRange_t: UINT4 fields (pyfstat/core.py:1447 "TYPE UINT4") */
typedef struct {
    uint32_t type, t0, t0Band, dt0, tau, tauBand, dtau;
} tcw_window;

/* One template's atoms, detector-concatenated, already on the device.
 * ts[det_off[X] .. det_off[X+1]) are detector X's timestamps (ascending);
 * cols[c*col_stride + k] is column c of atom k.  det_off, t_min, t_max live on
 * the host (t_min/t_max: smallest first / largest last timestamp over all
 * detectors, which fix the binned length). */
typedef struct {
    uint32_t numDet;
    uint32_t TAtom;
    const uint32_t *ts;   /* device */
    const float *cols;    /* device, [TCW_NCOL][col_stride] */
    int64_t col_stride;
    int64_t det_off[TCW_MAX_DETECTORS + 1];
    uint32_t t_min, t_max;
} tcw_atoms;

/* Per-map results (FstatMap fields, tcw_fstat_map_funcs.py:134-144). */
typedef struct {
    float maxF;              /* F, not 2F */
    int32_t single_atom;     /* 1 if any cell has i_t1 == i_t0 */
    int64_t m_ML, n_ML;      /* first row-major argmax */
    int64_t N_t0, N_tau;
    int64_t t0_ML, tau_ML;   /* effective window t0 + m*dt0, tau + n*dtau */
    double lnBtSG;           /* NaN unless TCW_FLAG_BTSG */
    int64_t m_MP, n_MP;      /* posterior-mode bins, -1 unless TCW_FLAG_BTSG */
    double t0_MP, tau_MP;    /* caller's window: t0 + (m_MP+0.5)*t0Band/N_t0 ... */
    uint32_t numAtoms;       /* binned atoms */
    uint32_t t0_data;
} tcw_stats;

typedef struct tcw_handle tcw_handle;

/* Library version string. */
const char *tcw_version(void);

/* Thread-local message of the last failing call on this thread. */
const char *tcw_last_error(void);

/* Number of visible HIP devices (0 when none); never fails. */
int tcw_device_count(void);

/* Marketing name of device `device` (hipDeviceProp_t::name) into buf. */
int tcw_device_name(int device, char *buf, int buflen);

/* Map shape after the TRANSIENT_NONE conversion (tcw_fstat_map_funcs.py:742-780):
 * N_t0 = floor(t0Band/dt0)+1, N_tau = floor(tauBand/dtau)+1. */
int tcw_map_shape(uint32_t numAtoms, uint32_t TAtom, uint32_t t0_data,
                  const tcw_window *win, int64_t *N_t0, int64_t *N_tau);

/* Handle owning device scratch (merged atoms, prefix sums, reduction partials)
 * on `device`; replaces pycuda's context (tcw_fstat_map_funcs.py:361-492). */
int tcw_create(int device, tcw_handle **out);
int tcw_destroy(tcw_handle *h);

/* One transient F-statistic map: merge + window sums + F + fused reductions.
 * Replaces pycuda_compute_transient_fstat_map (tcw_fstat_map_funcs.py:656-834).
 * F_mn_dev: device float32 [N_t0][N_tau] (F, not 2F), caller-allocated.
 * stats: host struct, filled before return (the call synchronises `stream`). */
int tcw_map(tcw_handle *h, const tcw_atoms *atoms, const tcw_window *win,
            float *F_mn_dev, tcw_stats *stats, uint32_t flags, void *stream);

/* Bytes of the library's static device scratch used by tcw_map_noalloc (one
 * copy per device, held from library load until the process ends). */
#define TCW_NOALLOC_SCRATCH_BYTES (32u << 20)

/* tcw_map without any device or pinned-memory allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels as tcw_map, with every intermediate buffer carved from the static
 * scratch of h's device (looked up once by tcw_create).
 * stats_dev: DEVICE tcw_stats, written by the last kernel; the call does not
 * wait for the device (read stats_dev after synchronising `stream`; under
 * TCW_FLAG_LAL_STRICT check stats_dev->single_atom yourself).
 * Fails with TCW_ERR_INVALID, before queueing anything, when the map needs
 * more than TCW_NOALLOC_SCRATCH_BYTES (e.g. an exp map of 2000 x 2000 cells on
 * 4000 atoms needs ~9 MB; a rect map of 2000 x 2000 cells ~0.6 MB).
 * All calls on one device share the scratch: calls that overlap in time on
 * that device (different streams or host threads) race and must be ordered
 * by the caller; the descriptors are uploaded with hipMemcpyAsync on `stream`. */
int tcw_map_noalloc(tcw_handle *h, const tcw_atoms *atoms, const tcw_window *win,
                    float *F_mn_dev, tcw_stats *stats_dev, uint32_t flags, void *stream);

/* ---------------------------------------------------------------------------
 * Batched maps: many templates (ragged atom counts and windows) in one pass.
 * Replaces the per-template Python loop of TransientGridSearch.run
 * (pyfstat/grid_based_searches.py:1186-1189) over get_transient_detstats
 * (pyfstat/core.py:1412-1471).
 *
 * Template t owns detector segments seg_begin[t] .. seg_begin[t+1]-1; segment
 * s owns atoms seg_off[s] .. seg_off[s+1]-1 of ts/cols.  All host arrays are
 * read during tcw_batch_plan only.
 * ------------------------------------------------------------------------- */
typedef struct {
    int64_t T;                 /* templates */
    uint32_t TAtom;            /* common atom length */
    const int64_t *seg_begin;  /* host [T+1] */
    const int64_t *seg_off;    /* host [S+1] */
    const uint32_t *seg_tfirst;/* host [S]: first timestamp of each segment */
    const uint32_t *seg_tlast; /* host [S]: last timestamp of each segment */
    const tcw_window *windows; /* host [T] */
} tcw_batch_desc;

typedef struct tcw_plan tcw_plan;

/* Plan a batch: shapes, output offsets, work items (cost-sorted), scratch.
 * F_off (host, [T+1], may be NULL) receives each map's float offset into the
 * concatenated output; the total float count is F_off[T]. */
int tcw_batch_plan(tcw_handle *h, const tcw_batch_desc *desc, uint32_t flags,
                   tcw_plan **out, int64_t *F_off);

/* tcw_batch_plan with the rows (t0) per work item chosen per kernel mode:
 * rows_by_mode (host [4], may be NULL = all 0) for modes 0 (rect, regular
 * grid), 1 (rect, generic), 2 (exp, direct) and 3 (exp on the atom grid);
 * each 0 (the library's choice for the batch) or a multiple of 16 in
 * [16, 512] ([16, 128] for mode 3).  The setting belongs to this plan only.
 * What the autotuner varies; no counterpart in the reference. */
int tcw_batch_plan_rows(tcw_handle *h, const tcw_batch_desc *desc, uint32_t flags,
                        const int32_t *rows_by_mode, tcw_plan **out, int64_t *F_off);

/* Device bytes of scratch the plan holds (prefix sums, partials, tables). */
int64_t tcw_plan_scratch_bytes(const tcw_plan *p);

/* Work items per kernel mode (items[0..3]: rect regular grid, rect generic,
 * exp direct, exp on the atom grid); what the autotuner tunes.  No
 * counterpart in the reference. */
int tcw_plan_items(const tcw_plan *p, int64_t *items);

/* Run a planned batch on device-resident atoms.  ts/cols follow the plan's
 * segment layout (cols is [TCW_NCOL][col_stride]); F_out is the concatenated
 * float32 output (F_off layout); stats_dev is a device array of T tcw_stats.
 * Asynchronous on `stream`.  With use_graph != 0 the launch sequence is
 * captured into a hipGraph on the first call and replayed afterwards while
 * the pointers stay the same. */
int tcw_batch_run(tcw_handle *h, tcw_plan *p, const uint32_t *ts,
                  const float *cols, int64_t col_stride, float *F_out,
                  tcw_stats *stats_dev, int use_graph, void *stream);

/* tcw_batch_run without a graph, with HIP events recorded on `stream` around
 * the stage kernel, the map kernels and the finalize kernel; waits for them
 * and returns the three elapsed times in ms_out[0..2] (milliseconds).  For
 * benchmarks and profiling; no counterpart in the reference. */
int tcw_batch_run_timed(tcw_handle *h, tcw_plan *p, const uint32_t *ts,
                        const float *cols, int64_t col_stride, float *F_out,
                        tcw_stats *stats_dev, double *ms_out, void *stream);

int tcw_plan_destroy(tcw_plan *p);

/* ---------------------------------------------------------------------------
 * Binned detector merge for many templates (lalpulsar
 * mergeMultiFstatAtomsBinned, called at pyfstat/tcw_fstat_map_funcs.py:706):
 * template t's bins k = 0 .. N_t-1 cover [t_data[t] + k*TAtom, +TAtom), with
 * t_data[t] the smallest first timestamp of its segments and
 * N_t = (largest last timestamp - t_data[t]) / TAtom + 1; each bin holds the
 * float32 sum of the atoms falling in it, in detector order.
 *
 * m_off (host [T+1]) and t_data (host [T]) are always filled; merged bins of
 * template t go to merged[c*merged_stride + m_off[t] + k].  With merged ==
 * NULL only the shapes are computed.  atom_bin (device [total atoms], may be
 * NULL) receives each input atom's absolute merged index m_off[t] + k, which
 * tcw_gather_bins uses to pull merged-bin gradients back to the atoms.
 * desc->windows is not read.  The call synchronises `stream`.
 * ------------------------------------------------------------------------- */
int tcw_merge_batch(tcw_handle *h, const tcw_batch_desc *desc, const uint32_t *ts,
                    const float *cols, int64_t col_stride, float *merged,
                    int64_t merged_stride, int64_t *m_off, uint32_t *t_data,
                    int64_t *atom_bin, void *stream);

/* out[c*out_stride + j] = in[c*in_stride + atom_bin[j]] for the 7 columns and
 * j < n_atoms: the adjoint of the merge (each atom's gradient is that of its
 * bin).  Asynchronous on `stream`. */
int tcw_gather_bins(const float *in, int64_t in_stride, const int64_t *atom_bin,
                    int64_t n_atoms, float *out, int64_t out_stride, void *stream);

/* ---------------------------------------------------------------------------
 * Exponential-window candidates on a fixed atom support, with derivatives.
 *
 * A candidate is a continuous (t0, tau) of one template together with a fixed
 * support: merged bins k_lo .. k_hi of that template.  Its window sums are
 *   A, B, C        = sum_k w_k^2 (a2, b2, ab)_k,
 *   Fa, Fb         = sum_k w_k   (Fa, Fb)_k,     w_k = exp(-(t_k - t0)/tau),
 * t_k = t_data + k*TAtom, and F follows from them as in the map kernels
 * (cond < 1e4 guard, F = 2 otherwise).  At a grid cell with the support the
 * map uses (bins with t0 <= t_k <= t0 + 3 tau, the window of
 * pyfstat/pyCUDAkernels/cudaTransientFstatExpWindow.cu:82-102) F equals the
 * map value.  lnBtSG of the candidate is that of a 1x1 map,
 * log(70) + F (tcw_fstat_map_funcs.py:196-217 with N_t0 = N_tau = 1).
 *
 * On a fixed support every weight carries the common factor exp(t0/tau), and
 * F is homogeneous of degree 0 in (A, B, C; Fa, Fb) with weights (2; 1), so
 * dF/dt0 vanishes identically; it is still returned (it is the sum the
 * backward uses and measures rounding).
 *
 * Reference analogue: MCMCTransientSearch evaluates a 1x1 map per likelihood
 * call (pyfstat/mcmc_based_searches.py:3511-3516); the reference has no
 * derivatives.
 * ------------------------------------------------------------------------- */
#define TCW_REFINE_SAVED 10 /* doubles per candidate in `saved` */

typedef struct {
    int64_t T;               /* templates */
    int64_t n_cand;          /* candidates, grouped by template */
    uint32_t TAtom;
    uint32_t reserved;
    const int64_t *m_off;    /* host [T+1]: template t owns merged bins m_off[t] .. m_off[t+1]-1 */
    const uint32_t *t_data;  /* host [T]: timestamp of bin 0 of each template */
    const int64_t *cand_off; /* host [T+1]: template t owns candidates cand_off[t] .. cand_off[t+1]-1 */
    const int32_t *k_lo;     /* host [n_cand]: first support bin, relative to m_off[t] */
    const int32_t *k_hi;     /* host [n_cand]: last support bin (inclusive) */
} tcw_refine_desc;

typedef struct tcw_refine_plan tcw_refine_plan;

/* Plan: uploads the supports and a work schedule (host arrays read here
 * only). */
int tcw_refine_plan_create(tcw_handle *h, const tcw_refine_desc *desc,
                           tcw_refine_plan **out);
int tcw_refine_plan_destroy(tcw_refine_plan *p);

/* Forward: twoF[i] = 2F, lnBtSG[i] = log(70) + F (may be NULL), and
 * saved[i*TCW_REFINE_SAVED + j] = dF/dt0, dF/dtau, dF/dA, dF/dB, dF/dC,
 * dF/dFa_re, dF/dFa_im, dF/dFb_re, dF/dFb_im, F (may be NULL).
 * cols: device merged atoms [TCW_NCOL][col_stride] (tcw_merge_batch layout);
 * t0, tau: device double [n_cand], tau > 0.  Asynchronous on `stream`. */
int tcw_refine_forward(tcw_refine_plan *p, const float *cols, int64_t col_stride,
                       const double *t0, const double *tau, double *twoF, double *lnBtSG,
                       double *saved, void *stream);

/* Backward of tcw_refine_forward for upstream gradients g_twoF, g_lnBtSG
 * (device double [n_cand], either may be NULL): with g = 2 g_twoF + g_lnBtSG,
 *   grad_t0[i]  = g_i dF_i/dt0,   grad_tau[i] = g_i dF_i/dtau   (may be NULL),
 *   grad_cols[c*col_stride + m_off[t] + k] = sum over the template's
 *       candidates i whose support holds k of g_i dF_i/dS_c w_ik^{p_c}
 *   (p = 2 for a2, b2, ab; 1 for Fa, Fb), float32; every bin of every
 *   template is written (zero outside the supports); may be NULL.
 * saved is the forward's.  Asynchronous on `stream`; uses the plan's
 * per-candidate scratch, so backward calls on one plan must not overlap. */
int tcw_refine_backward(tcw_refine_plan *p, const double *t0, const double *tau,
                        const double *saved, const double *g_twoF, const double *g_lnBtSG,
                        float *grad_cols, int64_t col_stride, double *grad_t0,
                        double *grad_tau, void *stream);

/* n_steps of gradient ascent of F in log(tau) per candidate, on the fixed
 * support (all steps of a candidate in one kernel): a step proposes
 * s' = s + eta dF/ds (s = log tau, clamped to [log tau_lo, log tau_hi];
 * tau_lo/tau_hi may be NULL for no bound), is accepted if F rises
 * (eta *= 1.5) and rejected otherwise
 * (eta *= 0.5); the first step moves s by step0.  t0 is not moved (dF/dt0 = 0
 * on a fixed support).  Starts from tau_in and writes the final tau to
 * tau_out (the two may be the same array); twoF receives the final 2F
 * and n_accept (may be NULL) the accepted steps.  Asynchronous on `stream`:
 * short supports run on a stream of the plan that starts after the work
 * queued on `stream` and that `stream` waits for before its later work, so
 * ascent calls on one plan must not overlap (from several host threads). */
int tcw_refine_ascent(tcw_refine_plan *p, const float *cols, int64_t col_stride,
                      const double *t0, const double *tau_in, double *tau_out,
                      const double *tau_lo, const double *tau_hi, int n_steps, double step0,
                      double *twoF, int32_t *n_accept, void *stream);

/* ---------------------------------------------------------------------------
 * F-statistic atoms on the device (the step before the map path).
 *
 * Replaces, for one sky position and many (f, fdot) templates, the per-point
 * host call lalpulsar.ComputeFstat(..., FSTATQ_ATOMS_PER_DET) that
 * pyfstat/core.py:671-673 requests and core.py:1359-1365 makes once per
 * Doppler point (lalpulsar is not part of the reference source; the sum is
 * the resampling-free "demodulation" F-statistic restated from its
 * definition, not lalpulsar's code).
 *
 * Per SFT alpha (start t_alpha, length Tsft, midpoint t_alpha + Tsft/2) and
 * template (f, fdot) at reference time ref_time in the solar-system barycentre:
 *   dtau  = t_alpha + Tsft/2 + delay_alpha - ref_time
 *   phase = dtau * (f + fdot dtau / 2)                (cycles, FP64)
 *   kappa = (f + fdot dtau) (1 + delay_rate_alpha) Tsft   (observed bin)
 *   X     = exp(-2 pi i phase) sum_k y_k K_k,  K_k = (-1)^kf sin(pi delta) / (pi (kappa - k))
 * over the 2*dterms bins k = kf-dterms+1 .. kf+dterms around kappa (kf =
 * floor(kappa), delta = kappa - kf; delta below 1e-30 takes bin kf alone), y_k the
 * SFT bins normalised to unit noise variance (x_k / sqrt(S Tsft / 2)); bins
 * outside the stored band count as zero.  The atoms are a2 = a^2, b2 = b^2,
 * ab = a b, Fa = a X, Fb = b X with the SFT's antenna patterns a, b, and
 * timestamp t_alpha, so that window sums give the transient F of the map
 * kernels (E[2F] = 4 sum K^2 ~ 4 in Gaussian noise).
 * ------------------------------------------------------------------------- */
typedef struct {
    int64_t n_sft;             /* SFTs of all detectors, concatenated */
    const uint32_t *t_start;   /* device [n_sft]: SFT start times (GPS s) */
    const float *bins;         /* device [n_sft][n_bins][2]: normalised bins (re, im) */
    int32_t n_bins;            /* stored bins per SFT */
    int32_t k_min;             /* stored bin j is frequency (k_min + j) / Tsft */
    uint32_t Tsft;             /* SFT length (s), the atoms' TAtom */
    uint32_t reserved;
    const double *delay;       /* device [n_sft]: SSB minus detector time at the midpoint (s) */
    const double *delay_rate;  /* device [n_sft]: d(delay)/dt at the midpoint */
    const float *a, *b;        /* device [n_sft]: antenna patterns at the midpoint */
} tcw_sft_set;

/* Atoms of n_tmpl templates (freq, fdot: device double [n_tmpl]) from one SFT
 * set: template t's atom of SFT alpha is element t*n_sft + alpha of ts_out
 * (device uint32) and of the 7 columns cols_out[c*col_stride + ...]
 * (a2, b2, ab, Fa_re, Fa_im, Fb_re, Fb_im; col_stride >= n_tmpl*n_sft) --
 * the layout tcw_batch_run reads, one segment per template and detector.
 * dterms in 1..64.  Asynchronous on `stream`. */
int tcw_demod_atoms(tcw_handle *h, const tcw_sft_set *sfts, int64_t n_tmpl, const double *freq,
                    const double *fdot, double ref_time, int32_t dterms, uint32_t *ts_out,
                    float *cols_out, int64_t col_stride, void *stream);

/* tcw_demod_atoms for n_sky sky positions in one launch (the reference calls
 * lalpulsar.ComputeFstat once per Doppler point, pyfstat/core.py:1359-1365;
 * here the (f, fdot) grid of every sky point at once): the SFT bins and start
 * times come from `sfts`, the per-SFT delay, delay_rate, a and b of sky point
 * s from the device arrays delay[s*n_sft + alpha] etc. (sfts' own delay/a/b
 * pointers are not read).  Sky s, template t, SFT alpha is element
 * (s*n_tmpl + t)*n_sft + alpha of ts_out and of each column
 * (col_stride >= n_sky*n_tmpl*n_sft); n_sky in 1..65535. */
int tcw_demod_atoms_sky(tcw_handle *h, const tcw_sft_set *sfts, int64_t n_sky,
                        const double *delay, const double *delay_rate, const float *a,
                        const float *b, int64_t n_tmpl, const double *freq, const double *fdot,
                        double ref_time, int32_t dterms, uint32_t *ts_out, float *cols_out,
                        int64_t col_stride, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TCW_HIP_H */
