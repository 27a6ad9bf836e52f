/*
 * TEST INFRASTRUCTURE ONLY -- never linked into, loaded by or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may use this file, and only to check results or to time them.
 *
 * FP64 CPU restatement of the transient-CW F-statistic map of lalpulsar's
 * XLALComputeTransientFstatMap() (LALSuite, TransientCW_utils.c; not vendored in
 * the reference, lalsuite[lalpulsar]>=7.13 per setup.py:45), as the reference
 * mirrors it in
 *   pyfstat/tcw_fstat_map_funcs.py:656-834   (host orchestration, N_t0/N_tau)
 *   pyfstat/pyCUDAkernels/cudaTransientFstatRectWindow.cu:14-124
 *   pyfstat/pyCUDAkernels/cudaTransientFstatExpWindow.cu:15-134
 * plus the binned detector merge (lalpulsar mergeMultiFstatAtomsBinned, called at
 * pyfstat/tcw_fstat_map_funcs.py:706) and the Bayes factor / posterior modes
 * (pyfstat/tcw_fstat_map_funcs.py:196-287).
 *
 * Differences from lalpulsar that are deliberate (this is the FP64 oracle):
 *   - window sums, determinant, condition number and F are all FP64
 *     (lalpulsar accumulates REAL4; pycuda accumulates float);
 *   - the exponential window uses libm exp() (lalpulsar uses a lookup table,
 *     which the reference test allows ~10% deviation for,
 *     tests/test_tcw_fstat_map_funcs.py:99-100).
 * Everything else -- uint32 wrap-around index maths, clamping, the incremental
 * rectangular sum, loop order, the cond<1e4 guard, the F=2 fallback, the
 * single-atom (i_t1 == i_t0) check -- follows the reference line by line.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_TRANSIENT_NONE 0u
#define ORC_TRANSIENT_RECTANGULAR 1u
#define ORC_TRANSIENT_EXPONENTIAL 2u
#define ORC_TRANSIENT_LAST 3u
#define ORC_EXP_EFOLDING 3u /* exp.cu:22 */
#define ORC_NCOL 7

typedef struct {
    uint32_t type, t0, t0Band, dt0, tau, tauBand, dtau;
} orc_window;

typedef struct {
    double maxF;
    int64_t m_max, n_max;   /* first row-major argmax */
    int64_t N_t0, N_tau;
    uint32_t t0, dt0, tau, dtau; /* effective window after the NONE conversion */
    int64_t bad_m, bad_n;   /* first single-atom cell (i_t1 == i_t0), -1 if none */
} orc_result;

/* Number of binned atoms: lalpulsar mergeMultiFstatAtomsBinned, NBinnedAtoms =
 * floor((tMax - tMin)/deltaT) + 1 over all detectors (called at tcw.py:706). */
int64_t orc_merge_size(int numDet, const int64_t *det_off, const uint32_t *ts,
                       uint32_t TAtom, uint32_t *tmin_out)
{
    uint32_t tmin = 0xFFFFFFFFu, tmax = 0;
    int any = 0;
    for (int X = 0; X < numDet; X++) {
        int64_t a = det_off[X], b = det_off[X + 1];
        if (b <= a) return -1;
        if (ts[a] < tmin) tmin = ts[a];
        if (ts[b - 1] > tmax) tmax = ts[b - 1];
        any = 1;
    }
    if (!any || TAtom == 0) return -1;
    *tmin_out = tmin;
    return (int64_t)((tmax - tmin) / TAtom) + 1;
}

/* Binned merge: every atom of every detector is added, in detector order and
 * then atom order, onto bin floor((t - tMin)/TAtom) of a zeroed float32 vector.
 * float32 accumulation on purpose: the merged atoms are the map's input and
 * are reproduced bit-exactly by the device merge. cols is [7][col_stride]. */
int orc_merge(int numDet, const int64_t *det_off, const uint32_t *ts,
              const float *cols, int64_t col_stride, uint32_t TAtom,
              uint32_t tmin, int64_t nbins, float *out /* [7][nbins] */)
{
    memset(out, 0, sizeof(float) * ORC_NCOL * (size_t)nbins);
    for (int X = 0; X < numDet; X++) {
        for (int64_t k = det_off[X]; k < det_off[X + 1]; k++) {
            int64_t i = (int64_t)((ts[k] - tmin) / TAtom);
            if (i < 0 || i >= nbins) return -1;
            for (int c = 0; c < ORC_NCOL; c++)
                out[c * nbins + i] += cols[c * col_stride + k];
        }
    }
    return 0;
}

/* N_t0 = floor(t0Band/dt0)+1, N_tau = floor(tauBand/dtau)+1 (tcw.py:775-780),
 * after the TRANSIENT_NONE -> full-span RECT conversion (tcw.py:742-749), done
 * on a copy. */
int orc_map_shape(uint32_t numAtoms, uint32_t TAtom, uint32_t t0_data,
                  const orc_window *win_in, orc_window *eff,
                  int64_t *N_t0, int64_t *N_tau)
{
    orc_window w = *win_in;
    if (w.type >= ORC_TRANSIENT_LAST) return -1;
    if (w.type == ORC_TRANSIENT_NONE) {
        w.type = ORC_TRANSIENT_RECTANGULAR;
        w.t0 = t0_data;
        w.t0Band = 0;
        w.dt0 = TAtom;
        w.tau = numAtoms * TAtom;
        w.tauBand = 0;
        w.dtau = TAtom;
    }
    if (w.dt0 == 0 || w.dtau == 0) return -2;
    *eff = w;
    *N_t0 = (int64_t)(w.t0Band / w.dt0) + 1;
    *N_tau = (int64_t)(w.tauBand / w.dtau) + 1;
    return 0;
}

/* i = (t - t0_data + TAtomHalf)/TAtom in uint32 with wrap-around, stored to a
 * signed int, clamped to [0, numAtoms-1]  (rect.cu:21-31). */
static inline uint32_t orc_idx_t0(uint32_t t, uint32_t t0_data, uint32_t TAtom,
                                  uint32_t numAtoms)
{
    uint32_t half = TAtom / 2;
    int32_t i_tmp = (int32_t)((uint32_t)(t - t0_data + half) / TAtom);
    if (i_tmp < 0) i_tmp = 0;
    uint32_t i = (uint32_t)i_tmp;
    if (i >= numAtoms) i = numAtoms - 1;
    return i;
}

/* i = (t1 - t0_data + TAtomHalf)/TAtom - 1, same maths (rect.cu:62-69). */
static inline uint32_t orc_idx_t1(uint32_t t1, uint32_t t0_data, uint32_t TAtom,
                                  uint32_t numAtoms)
{
    uint32_t half = TAtom / 2;
    int32_t i_tmp = (int32_t)((uint32_t)(t1 - t0_data + half) / TAtom - 1u);
    if (i_tmp < 0) i_tmp = 0;
    uint32_t i = (uint32_t)i_tmp;
    if (i >= numAtoms) i = numAtoms - 1;
    return i;
}

/* cond-guarded F from window sums (rect.cu:98-118, exp.cu:109-129), FP64.
 * The condition number is also returned through *cond_out. */
static inline double orc_F(double A, double B, double C, double Fa_re,
                           double Fa_im, double Fb_re, double Fb_im,
                           double *cond_out)
{
    double sumAB = A + B;
    double diffAB = A - B;
    double disc = sqrt(diffAB * diffAB + 4.0 * C * C);
    double denom = sumAB - disc;
    double cond = (denom > 0) ? ((sumAB + disc) / denom) : INFINITY;
    *cond_out = cond;
    double DdInv = 0.0;
    if (cond < 1e4) DdInv = 1.0 / (A * B - C * C);
    double F = 2.0;
    if (DdInv > 0)
        F = DdInv * (B * (Fa_re * Fa_re + Fa_im * Fa_im)
                     + A * (Fb_re * Fb_re + Fb_im * Fb_im)
                     - 2.0 * C * (Fa_re * Fb_re + Fa_im * Fb_im));
    return F;
}

/* The map itself.  atoms is [7][numAtoms] (a2,b2,ab,Fa_re,Fa_im,Fb_re,Fb_im) of
 * the merged, binned vector whose bin i has timestamp t0_data + i*TAtom.
 * F_out is [N_t0][N_tau] row-major (F, not 2F). Returns 0, or -1 for a bad
 * window. Single-atom cells are recorded in res->bad_m/bad_n (the caller
 * decides between lalpulsar's XLAL_EDOM and pycuda's silent evaluation). */
int orc_map(const float *atoms, uint32_t numAtoms, uint32_t TAtom,
            uint32_t t0_data, const orc_window *win_in, double *F_out,
            double *cond_out, orc_result *res)
{
    orc_window w;
    int64_t N_t0, N_tau;
    if (numAtoms == 0) return -1;
    if (orc_map_shape(numAtoms, TAtom, t0_data, win_in, &w, &N_t0, &N_tau))
        return -1;
    const float *a2 = atoms, *b2 = atoms + numAtoms, *ab = atoms + 2 * (size_t)numAtoms;
    const float *Fare = atoms + 3 * (size_t)numAtoms, *Faim = atoms + 4 * (size_t)numAtoms;
    const float *Fbre = atoms + 5 * (size_t)numAtoms, *Fbim = atoms + 6 * (size_t)numAtoms;

    res->maxF = -1.0; /* tcw.py:139 */
    res->m_max = 0;
    res->n_max = 0;
    res->N_t0 = N_t0;
    res->N_tau = N_tau;
    res->t0 = w.t0;
    res->dt0 = w.dt0;
    res->tau = w.tau;
    res->dtau = w.dtau;
    res->bad_m = -1;
    res->bad_n = -1;
    int have_max = 0;

    for (int64_t m = 0; m < N_t0; m++) {
        uint32_t t0 = w.t0 + (uint32_t)m * w.dt0;
        uint32_t i_t0 = orc_idx_t0(t0, t0_data, TAtom, numAtoms);
        double A = 0, B = 0, C = 0, Far = 0, Fai = 0, Fbr = 0, Fbi = 0;
        uint32_t i_t1_last = i_t0;
        for (int64_t n = 0; n < N_tau; n++) {
            uint32_t tau = w.tau + (uint32_t)n * w.dtau;
            uint32_t t1 = (w.type == ORC_TRANSIENT_EXPONENTIAL)
                              ? t0 + ORC_EXP_EFOLDING * tau
                              : t0 + tau;
            uint32_t i_t1 = orc_idx_t1(t1, t0_data, TAtom, numAtoms);
            if (i_t1 == i_t0 && res->bad_m < 0) {
                res->bad_m = m;
                res->bad_n = n;
            }
            if (w.type == ORC_TRANSIENT_RECTANGULAR) {
                /* incremental sum from i_t1_last (rect.cu:75-91) */
                for (uint32_t i = i_t1_last; i <= i_t1; i++) {
                    A += a2[i];
                    B += b2[i];
                    C += ab[i];
                    Far += Fare[i];
                    Fai += Faim[i];
                    Fbr += Fbre[i];
                    Fbi += Fbim[i];
                    i_t1_last = i_t1 + 1;
                }
            } else {
                /* direct exp-weighted sum over [i_t0, i_t1] (exp.cu:82-102) */
                A = B = C = Far = Fai = Fbr = Fbi = 0;
                for (uint32_t i = i_t0; i <= i_t1; i++) {
                    uint32_t t_i = t0_data + i * TAtom;
                    double win = 0.0;
                    if (t_i >= t0 && t_i <= t1) {
                        double x = (double)(t_i - t0) / (double)tau;
                        win = exp(-x);
                    }
                    double win2 = win * win;
                    A += a2[i] * win2;
                    B += b2[i] * win2;
                    C += ab[i] * win2;
                    Far += Fare[i] * win;
                    Fai += Faim[i] * win;
                    Fbr += Fbre[i] * win;
                    Fbi += Fbim[i] * win;
                }
            }
            double cond;
            double F = orc_F(A, B, C, Far, Fai, Fbr, Fbi, &cond);
            F_out[m * N_tau + n] = F;
            if (cond_out) cond_out[m * N_tau + n] = cond;
            if (!have_max || F > res->maxF) {
                /* first occurrence in row-major order (tcw.py:810-813) */
                res->maxF = F;
                res->m_max = m;
                res->n_max = n;
                have_max = 1;
            }
        }
    }
    return 0;
}

/* lnBtSG = log(70/(N_t0 N_tau)) + maxF + log sum exp(F - maxF)
 * (tcw.py:196-217, port of ComputeTransientBstat). */
double orc_lnBtSG(const double *F, int64_t N_t0, int64_t N_tau, double maxF)
{
    double s = 0.0;
    for (int64_t k = 0; k < N_t0 * N_tau; k++) s += exp(-(maxF - F[k]));
    return log(70.0 / ((double)N_t0 * (double)N_tau)) + maxF + log(s);
}

/* Posterior modes (tcw.py:219-287): argmax of row (t0) / column (tau) sums of
 * exp(F - maxF), returned at the bin centre x0 + (k+0.5)*band/N. */
void orc_posterior_modes(const double *F, int64_t N_t0, int64_t N_tau,
                         double maxF, const orc_window *win, double *t0_MP,
                         double *tau_MP)
{
    double best = -1.0;
    int64_t kbest = 0;
    for (int64_t m = 0; m < N_t0; m++) {
        double s = 0.0;
        for (int64_t n = 0; n < N_tau; n++) s += exp(-(maxF - F[m * N_tau + n]));
        if (s > best) {
            best = s;
            kbest = m;
        }
    }
    *t0_MP = (double)win->t0 + ((double)kbest + 0.5) * ((double)win->t0Band / (double)N_t0);
    double *cs = (double *)calloc((size_t)N_tau, sizeof(double));
    for (int64_t m = 0; m < N_t0; m++)
        for (int64_t n = 0; n < N_tau; n++) cs[n] += exp(-(maxF - F[m * N_tau + n]));
    best = -1.0;
    kbest = 0;
    for (int64_t n = 0; n < N_tau; n++)
        if (cs[n] > best) {
            best = cs[n];
            kbest = n;
        }
    free(cs);
    *tau_MP = (double)win->tau + ((double)kbest + 0.5) * ((double)win->tauBand / (double)N_tau);
}

/* Many independent maps, one per template, OpenMP over templates: the CPU
 * baseline for batched workloads.  atoms_off[t] indexes atoms (each template a
 * [7][numAtoms[t]] block), F_off[t] indexes F_out.  Per-template maxF, argmax
 * and lnBtSG go to the out arrays. */
int orc_map_batch(int64_t T, const float *atoms, const int64_t *atoms_off,
                  const uint32_t *numAtoms, const uint32_t *TAtom,
                  const uint32_t *t0_data, const orc_window *wins,
                  double *F_out, const int64_t *F_off, double *maxF,
                  int64_t *m_max, int64_t *n_max, double *lnBtSG)
{
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int64_t t = 0; t < T; t++) {
        orc_result r;
        int rc = orc_map(atoms + atoms_off[t], numAtoms[t], TAtom[t], t0_data[t],
                         &wins[t], F_out + F_off[t], NULL, &r);
        if (rc) {
            err |= 1;
            continue;
        }
        maxF[t] = r.maxF;
        m_max[t] = r.m_max;
        n_max[t] = r.n_max;
        if (lnBtSG)
            lnBtSG[t] = orc_lnBtSG(F_out + F_off[t], r.N_t0, r.N_tau, r.maxF);
    }
    return err ? -1 : 0;
}

/* Threads for the OpenMP loops above (CPU baseline: single core vs all
 * cores); returns the previous maximum. */
#ifdef _OPENMP
#include <omp.h>
#endif
int orc_set_threads(int n)
{
#ifdef _OPENMP
    int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
#else
    (void)n;
    return 1;
#endif
}

/* Exponential-window candidates on a fixed support (the refinement's
 * likelihood call; the same window sums as the map, exp.cu:82-129, at a
 * continuous (t0, tau)): per candidate i of template cand_tmpl[i], bins
 * k_lo[i]..k_hi[i] of that template's merged atoms (cols is [7][col_stride],
 * template t at columns m_off[t]..), bin k at t_data[t] + k*TAtom:
 *   S_c = sum_k w^p_c x_ck,  U_c = sum_k w^p_c (t_k - t0) x_ck,
 *   w = exp(-(t_k - t0)/tau), p = (2,2,2,1,1,1,1),
 * F with the cond < 1e4 guard (F = 2 and zero derivatives otherwise), and
 *   dF/dS_c (7), dF/dt0 = sum p S g / tau, dF/dtau = sum p U g / tau^2.
 * A C restatement of oracle/refine_ref.py's evaluate() (FP64, direct exp per
 * bin), OpenMP over candidates. */
int orc_refine_eval_batch(int64_t n, const float *cols, int64_t col_stride,
                          const int64_t *m_off, const uint32_t *t_data, uint32_t TAtom,
                          const int64_t *cand_tmpl, const int64_t *k_lo,
                          const int64_t *k_hi, const double *t0, const double *tau,
                          double *F_out, double *dt0_out, double *dtau_out,
                          double *g_out /* [n][7] */)
{
    static const int P[ORC_NCOL] = {2, 2, 2, 1, 1, 1, 1};
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        const int64_t t = cand_tmpl[i];
        const float *x = cols + m_off[t];
        double S[ORC_NCOL] = {0}, U[ORC_NCOL] = {0};
        for (int64_t k = k_lo[i]; k <= k_hi[i]; k++) {
            const double u = ((double)t_data[t] + (double)k * (double)TAtom) - t0[i];
            const double w = exp(-u / tau[i]);
            for (int c = 0; c < ORC_NCOL; c++) {
                const double wp = P[c] == 2 ? w * w : w;
                const double v = wp * (double)x[(int64_t)c * col_stride + k];
                S[c] += v;
                U[c] += v * u;
            }
        }
        const double A = S[0], B = S[1], C = S[2];
        const double far = S[3], fai = S[4], fbr = S[5], fbi = S[6];
        double cond;
        const double F = orc_F(A, B, C, far, fai, fbr, fbi, &cond);
        double g[ORC_NCOL] = {0};
        double d0 = 0.0, dt = 0.0;
        if (cond < 1e4) {
            const double D = A * B - C * C;
            const double p = far * far + fai * fai;
            const double q = fbr * fbr + fbi * fbi;
            const double r = far * fbr + fai * fbi;
            g[0] = (q - F * B) / D;
            g[1] = (p - F * A) / D;
            g[2] = 2.0 * (F * C - r) / D;
            g[3] = 2.0 * (B * far - C * fbr) / D;
            g[4] = 2.0 * (B * fai - C * fbi) / D;
            g[5] = 2.0 * (A * fbr - C * far) / D;
            g[6] = 2.0 * (A * fbi - C * fai) / D;
            for (int c = 0; c < ORC_NCOL; c++) {
                d0 += P[c] * S[c] * g[c];
                dt += P[c] * U[c] * g[c];
            }
            d0 /= tau[i];
            dt /= tau[i] * tau[i];
        }
        F_out[i] = F;
        dt0_out[i] = d0;
        dtau_out[i] = dt;
        for (int c = 0; c < ORC_NCOL; c++) g_out[i * ORC_NCOL + c] = g[c];
    }
    return 0;
}

/* F-statistic atoms of the resampling-free demodulation sum (the formula of
 * oracle/demod_ref.py's demod_atoms, restated in C for timing): for template
 * t and SFT a,
 *   dtau  = t_start + Tsft/2 + delay - ref_time,
 *   phase = dtau (f + fdot dtau / 2)   (cycles),
 *   kappa = (f + fdot dtau) (1 + delay_rate) Tsft,
 *   X     = exp(-2 pi i frac(phase)) sum_k y_k (-1)^kf sin(pi delta) / (pi (kappa - k))
 * over k = kf-D+1 .. kf+D (kf = floor(kappa), delta = kappa - kf; delta = 0:
 * bin kf alone), out-of-band bins zero.  FP64; bins are float32 (re, im)
 * pairs [n_sft][n_bins][2].  Writes Fa = a X and Fb = b X as (re, im) pairs
 * [n_tmpl][n_sft][2].  OpenMP over templates and SFTs. */
int orc_demod_atoms(int64_t n_sft, const uint32_t *t_start, const float *bins, int32_t n_bins,
                    int32_t k_min, double Tsft, const double *delay, const double *delay_rate,
                    const float *a, const float *b, int64_t n_tmpl, const double *freq,
                    const double *fdot, double ref_time, int32_t dterms, double *Fa,
                    double *Fb)
{
    if (dterms < 1 || n_bins < 1) return 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t t = 0; t < n_tmpl; t++) {
        for (int64_t s = 0; s < n_sft; s++) {
            const double dtau = (double)t_start[s] + 0.5 * Tsft + delay[s] - ref_time;
            const double phase = dtau * (freq[t] + 0.5 * fdot[t] * dtau);
            const double dop = (1.0 + delay_rate[s]) * Tsft;
            const double kappa = (freq[t] + fdot[t] * dtau) * dop;
            const double kf = floor(kappa);
            const double delta = kappa - kf;
            const int64_t kfi = (int64_t)kf;
            const double sign = (kfi % 2 != 0) ? -1.0 : 1.0;
            const float *row = bins + s * (int64_t)n_bins * 2;
            double sr = 0.0, si = 0.0;
            if (delta == 0.0) {
                const int64_t j = kfi - k_min;
                if (j >= 0 && j < n_bins) {
                    sr = sign * (double)row[2 * j];
                    si = sign * (double)row[2 * j + 1];
                }
            } else {
                const double amp = sign * sin(M_PI * delta);
                for (int i = 0; i < 2 * dterms; i++) {
                    const int64_t k = kfi - dterms + 1 + i;
                    const int64_t j = k - k_min;
                    if (j < 0 || j >= n_bins) continue;
                    const double K = amp / (M_PI * (kappa - (double)k));
                    sr += K * (double)row[2 * j];
                    si += K * (double)row[2 * j + 1];
                }
            }
            const double frac = phase - floor(phase);
            const double c = cos(2.0 * M_PI * frac), sn = sin(2.0 * M_PI * frac);
            const double xr = sr * c + si * sn, xi = si * c - sr * sn;
            const int64_t o = (t * n_sft + s) * 2;
            Fa[o] = (double)a[s] * xr;
            Fa[o + 1] = (double)a[s] * xi;
            Fb[o] = (double)b[s] * xr;
            Fb[o + 1] = (double)b[s] * xi;
        }
    }
    return 0;
}
