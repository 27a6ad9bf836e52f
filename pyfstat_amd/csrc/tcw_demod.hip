// tcw_demod.hip -- F-statistic atoms on the device (gfx950): for one sky
// position, the per-SFT atoms (a2, b2, ab, Fa, Fb) of many (f, fdot)
// templates, written in the layout the batched map reads, so a grid search
// goes from SFTs to transient maps without a host round trip.
//
// In the reference these atoms come from lalpulsar.ComputeFstat(...,
// FSTATQ_ATOMS_PER_DET), requested at pyfstat/core.py:671-673 and called once
// per Doppler point at core.py:1359-1365; lalpulsar is outside the reference
// source.  The sum here is the resampling-free demodulation F-statistic from
// its definition (the matched filter of a CW signal that is a pure tone
// across each SFT): the Dirichlet-kernel weights of the 2*dterms bins around
// the Doppler-shifted frequency, and the signal phase at the SFT midpoint.
//
// Tiles of TT templates x 64 SFTs per workgroup, two phases:
//   compute  lanes are templates of one SFT (TT = 64; the SFTs' dtau and
//            Doppler factors are formed once per workgroup into LDS), so the
//            lanes read the same few SFT bins, since neighbouring templates
//            fall in the same bins; the atom's complex X goes to LDS.
//   write    lanes are SFTs of one template: each of the 8 output columns is
//            written as 64 consecutive 4-byte values per wave (full lines,
//            nontemporal stores).
// The phase is FP64 (f * dtau reaches 1e10 cycles); the kernel weights and
// the complex sum are float32 (the atoms are float32), summed as one rational
// function per run of up to 8 bins (one reciprocal per run instead of one
// per bin); sin(pi delta) and the rotation are Taylor polynomials on reduced
// arguments.  Bound: the 32 bytes written per atom (HBM); the bins read are
// a few MB shared by every template and stay in the caches.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "tcw_common.h"
#include "tcw_hip.h"

using tcwc::fail;
using tcwc::set_device;

namespace {

constexpr int DEMOD_WG = 256;
constexpr int DEMOD_WAVES = DEMOD_WG / 64;
constexpr int SFT_TILE = 64;  // SFTs per workgroup
constexpr int DTERMS_MAX = 64;
constexpr int RUN = 8;  // bins per rational run: |prod (delta + c)| <= 65^8 fits float
constexpr unsigned DEMOD_GRID_Y = 65535;

// Taylor coefficients, evaluated at compile time from pi:
//   sin(w x) = sum_k sin_coef(w, k) x^(2k+1),  cos(w x) = sum_k cos_coef(w, k) x^(2k)
constexpr float sin_coef(double w, int k) {
    double c = w;
    for (int i = 1; i <= k; i++) c *= -w * w / ((2.0 * i) * (2.0 * i + 1.0));
    return (float)c;
}
constexpr float cos_coef(double w, int k) {
    double c = 1.0;
    for (int i = 1; i <= k; i++) c *= -w * w / ((2.0 * i - 1.0) * (2.0 * i));
    return (float)c;
}
constexpr double TWO_PI = 2.0 * M_PI;

// sin(2 pi v) and cos(2 pi v) for |v| <= 1/8 (|2 pi v| <= pi/4): Taylor
// polynomials through v^9 and v^10, truncation below 2e-9.
__device__ __forceinline__ float sin_2pi_small(float v, float v2) {
    float p = sin_coef(TWO_PI, 4);
    p = fmaf(p, v2, sin_coef(TWO_PI, 3));
    p = fmaf(p, v2, sin_coef(TWO_PI, 2));
    p = fmaf(p, v2, sin_coef(TWO_PI, 1));
    p = fmaf(p, v2, sin_coef(TWO_PI, 0));
    return p * v;
}
__device__ __forceinline__ float cos_2pi_small(float v2) {
    float p = cos_coef(TWO_PI, 5);
    p = fmaf(p, v2, cos_coef(TWO_PI, 4));
    p = fmaf(p, v2, cos_coef(TWO_PI, 3));
    p = fmaf(p, v2, cos_coef(TWO_PI, 2));
    p = fmaf(p, v2, cos_coef(TWO_PI, 1));
    return fmaf(p, v2, 1.0f);
}

// sin(pi x) / pi for 0 <= x <= 1/2: x times the Taylor series of
// sin(pi x) / (pi x) through x^12, relative truncation below 1e-9.
__device__ __forceinline__ float sinpi_over_pi(float x) {
    const float x2 = x * x;
    float p = cos_coef(M_PI, 6) / 13.0f;  // (-1)^k pi^(2k) / (2k+1)!, k = 6
    p = fmaf(p, x2, (float)((double)cos_coef(M_PI, 5) / 11.0));
    p = fmaf(p, x2, (float)((double)cos_coef(M_PI, 4) / 9.0));
    p = fmaf(p, x2, (float)((double)cos_coef(M_PI, 3) / 7.0));
    p = fmaf(p, x2, (float)((double)cos_coef(M_PI, 2) / 5.0));
    p = fmaf(p, x2, (float)((double)cos_coef(M_PI, 1) / 3.0));
    p = fmaf(p, x2, 1.0f);
    return p * x;
}

// One run of up to RUN bins at stored indices j, j+Step, ... (Guard: only the
// first cnt of them; Checked: out-of-band bins are zero): adds the sum of
// y / s with s = bm - Step * m for the bin's position m.  The run is
// accumulated as one fraction n/d (n <- n s + y d, d <- d s) and divided once.
template <bool Checked, int Step, bool Guard>
__device__ __forceinline__ void dirichlet_run(const float2 *__restrict__ row, int j, int cnt,
                                              float bm, int n_bins, float &re, float &im) {
    float nr, ni, d;  // cnt >= 1: the first bin starts the fraction
#pragma unroll
    for (int m = 0; m < RUN; m++) {
        if (!Guard || m < cnt) {
            const int jj = j + Step * m;
            float2 y;
            if (Checked)
                y = (jj >= 0 && jj < n_bins) ? row[jj] : make_float2(0.0f, 0.0f);
            else
                y = row[jj];
            const float s = bm - (float)(Step * m);
            if (m == 0) {
                nr = y.x;
                ni = y.y;
                d = s;
            } else {
                nr = fmaf(nr, s, y.x * d);
                ni = fmaf(ni, s, y.y * d);
                d *= s;
            }
        }
    }
    // |d| lies between |delta| or |1 - delta| times a product of small integers
    // and 65^8: a normal float, where the hardware reciprocal is within 1 ulp
    const float r = __builtin_amdgcn_rcpf(d);
    re = fmaf(nr, r, re);
    im = fmaf(ni, r, im);
}

// Half of the window: dterms bins from stored index j stepping by Step, with
// s = base - Step * pos, starting at the bin nearest kappa so that every s is
// formed from base with at most two roundings relative to |s| (small |s|
// keeps its relative accuracy).  DT: dterms at compile time, or 0.
template <bool Checked, int Step, int DT>
__device__ __forceinline__ void dirichlet_half(const float2 *__restrict__ row, int j, int dterms,
                                               float base, int n_bins, float &re, float &im) {
    if constexpr (DT != 0) {
#pragma unroll
        for (int m0 = 0; m0 < DT; m0 += RUN)
            dirichlet_run<Checked, Step, (DT % RUN != 0)>(row, j + Step * m0, DT - m0,
                                                          base - (float)(Step * m0), n_bins, re,
                                                          im);
    } else {
        for (int m0 = 0; m0 < dterms; m0 += RUN)
            dirichlet_run<Checked, Step, true>(row, j + Step * m0, dterms - m0,
                                               base - (float)(Step * m0), n_bins, re, im);
    }
}

// X of template (f, fd) in one SFT, before the antenna factors.
template <int DT>
__device__ __forceinline__ float2 demod_one(const float2 *__restrict__ row, int n_bins,
                                            int k_min, double dtau, double dop, double f,
                                            double fd, int dterms_rt) {
    const int dterms = DT != 0 ? DT : dterms_rt;
    const double phase = dtau * fma(0.5 * fd, dtau, f);  // cycles
    const double kappa = fma(fd, dtau, f) * dop;           // observed frequency in bins
    const double kf = floor(kappa);
    const double delta = kappa - kf;
    // saturating conversion: a kappa beyond the int range lands out of band
    const int kfi = (int)fmin(fmax(kf, -2.0e9), 2.0e9);
    float re = 0.0f, im = 0.0f;
    if (delta < 1e-30) {
        // kappa on a bin (to within 1e-30 of one, where the other weights are
        // below 1e-30 and float could not hold delta): that bin alone
        const int j = kfi - k_min;
        if (j >= 0 && j < n_bins) {
            re = row[j].x;
            im = row[j].y;
        }
    } else {
        // K_k = (-1)^kf sin(pi delta) / (pi (kappa - k)) over k = kf-D+1 .. kf+D:
        // kappa - k runs delta, delta + 1, .. over the lower half (from the bin
        // at kappa down) and delta - 1, delta - 2, .. over the upper half;
        // delta - 1 is formed in FP64, so that delta near 1 keeps its distance
        // from the bin instead of rounding to zero
        const int j0 = kfi - dterms + 1 - k_min;
        const float lo = (float)delta;
        const float hi = (float)(delta - 1.0);
        if (j0 >= 0 && j0 <= n_bins - 2 * dterms) {
            dirichlet_half<false, -1, DT>(row, j0 + dterms - 1, dterms, lo, n_bins, re, im);
            dirichlet_half<false, 1, DT>(row, j0 + dterms, dterms, hi, n_bins, re, im);
        } else {
            dirichlet_half<true, -1, DT>(row, j0 + dterms - 1, dterms, lo, n_bins, re, im);
            dirichlet_half<true, 1, DT>(row, j0 + dterms, dterms, hi, n_bins, re, im);
        }
        // sin(pi delta) = sin(pi (1 - delta)): the argument nearer zero, in float
        const float amp = sinpi_over_pi(delta <= 0.5 ? lo : -hi);
        re *= amp;
        im *= amp;
    }
    // rotate by (-1)^kf exp(-2 pi i phase) = exp(-i (2 pi u + kf pi)): the
    // cycle fraction u in [-1/2, 1/2] in FP64, then 2 pi u = 2 pi v + q pi/2
    // with |v| <= 1/8 in float, and the angle 2 pi v + qi pi/2 with
    // qi = q + 2 kf (mod 4)
    const float u = (float)(phase - rint(phase));
    const float q = rintf(4.0f * u);
    const float v = fmaf(-0.25f, q, u);
    const float v2 = v * v;
    const float s0 = sin_2pi_small(v, v2), c0 = cos_2pi_small(v2);
    const int qi = ((int)q + 2 * kfi) & 3;
    float sn = (qi & 1) ? c0 : s0;
    float cs = (qi & 1) ? s0 : c0;
    if (qi & 2) sn = -sn;
    if ((qi + 1) & 2) cs = -cs;
    return make_float2(fmaf(re, cs, im * sn), fmaf(im, cs, -(re * sn)));
}

// TT templates x SFT_TILE SFTs per workgroup; grid x over SFT tiles, grid y
// strides over template tiles, grid z is the sky point (its per-SFT delay,
// rate and antenna patterns, and its block of n_tmpl templates in the
// output).  DT: dterms at compile time, or 0 (runtime).
template <int TT, int DT>
__global__ __launch_bounds__(DEMOD_WG) void demod_atoms_kernel(
    int64_t n_sft, int64_t n_tmpl, const uint32_t *__restrict__ t_start,
    const float2 *__restrict__ bins, int n_bins, int k_min, double Tsft,
    const double *__restrict__ delay, const double *__restrict__ delay_rate,
    const float *__restrict__ pa, const float *__restrict__ pb, const double *__restrict__ freq,
    const double *__restrict__ fdot, double ref_time, int dterms, uint32_t *__restrict__ ts_out,
    float *__restrict__ cols, int64_t cstride) {
    constexpr int LD = TT + 1;  // odd row stride: both LDS phases conflict-free
    constexpr int SFT_PER_PASS = 64 / TT * DEMOD_WAVES;
    __shared__ float xr_s[SFT_TILE * LD];
    __shared__ float xi_s[SFT_TILE * LD];
    __shared__ double dtau_s[SFT_TILE], dop_s[SFT_TILE];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t a0 = (int64_t)blockIdx.x * SFT_TILE;
    {  // this workgroup's sky point
        const int64_t sky = blockIdx.z;
        delay += sky * n_sft;
        delay_rate += sky * n_sft;
        pa += sky * n_sft;
        pb += sky * n_sft;
        ts_out += sky * n_tmpl * n_sft;
        cols += sky * n_tmpl * n_sft;
    }
    // write phase: this lane's SFT (the wave's template rows are uniform, so
    // every store is a uniform row pointer plus the lane's 32-bit offset)
    const int64_t wa = a0 + lane;
    const bool w_ok = wa < n_sft;
    uint32_t w_ts = 0;
    float w_a = 0.0f, w_b = 0.0f;
    if (w_ok) {
        w_ts = t_start[wa];
        w_a = pa[wa];
        w_b = pb[wa];
    }
    const float w_a2 = w_a * w_a, w_b2 = w_b * w_b, w_ab = w_a * w_b;
    // an (empty) use of the loaded values here makes the compiler wait for
    // those loads once, before the tiles, instead of at the top of every row
    // of stores below (where the wait would also drain the previous row's
    // stores)
    asm volatile("" ::"v"(w_ts), "v"(w_a2), "v"(w_b2), "v"(w_ab), "v"(w_a), "v"(w_b));
    // the tile's SFT midpoints relative to ref_time and Doppler factors, once
    if (threadIdx.x < SFT_TILE && w_ok) {
        dtau_s[lane] = (double)w_ts + 0.5 * Tsft + delay[wa] - ref_time;
        dop_s[lane] = (1.0 + delay_rate[wa]) * Tsft;
    }
    __syncthreads();
    // compute phase: this lane's template slot and SFT within each pass
    const int r = lane % TT;
    const int isub = wave * (64 / TT) + lane / TT;
    for (int64_t t0 = (int64_t)blockIdx.y * TT; t0 < n_tmpl; t0 += (int64_t)gridDim.y * TT) {
        const int64_t t = t0 + r;
        const bool t_ok = t < n_tmpl;
        const double f = t_ok ? freq[t] : 0.0;
        const double fd = t_ok ? fdot[t] : 0.0;
        for (int i = isub; i < SFT_TILE; i += SFT_PER_PASS) {
            const int64_t a = a0 + i;
            float2 x = make_float2(0.0f, 0.0f);
            if (t_ok && a < n_sft) {
                x = demod_one<DT>(bins + a * n_bins, n_bins, k_min, dtau_s[i], dop_s[i], f, fd,
                                  dterms);
            }
            xr_s[i * LD + r] = x.x;
            xi_s[i * LD + r] = x.y;
        }
        __syncthreads();
        // rows of this tile that exist (uniform: a scalar loop bound)
        const int rows = (int)min<int64_t>(TT, n_tmpl - t0);
        if (w_ok) {
            const unsigned off = (unsigned)lane;
            for (int rr = wave; rr < rows; rr += DEMOD_WAVES) {
                const int64_t orow = (t0 + rr) * n_sft + a0;  // uniform
                uint32_t *trow = ts_out + orow;
                float *crow = cols + orow;
                const float xr = xr_s[lane * LD + rr], xi = xi_s[lane * LD + rr];
                __builtin_nontemporal_store(w_ts, trow + off);
                __builtin_nontemporal_store(w_a2, crow + off);
                __builtin_nontemporal_store(w_b2, (crow + cstride) + off);
                __builtin_nontemporal_store(w_ab, (crow + 2 * cstride) + off);
                __builtin_nontemporal_store(w_a * xr, (crow + 3 * cstride) + off);
                __builtin_nontemporal_store(w_a * xi, (crow + 4 * cstride) + off);
                __builtin_nontemporal_store(w_b * xr, (crow + 5 * cstride) + off);
                __builtin_nontemporal_store(w_b * xi, (crow + 6 * cstride) + off);
            }
        }
        __syncthreads();
    }
}

// Both entry points: argument checks, then the launch (n_sky sky points on
// grid z; delay .. b are [n_sky][n_sft]).
int demod_launch(tcw_handle *h, const tcw_sft_set *sfts, int64_t n_sky, const double *delay,
                 const double *delay_rate, const float *a, const float *b, int64_t n_tmpl,
                 const double *freq, const double *fdot, double ref_time, int32_t dterms,
                 uint32_t *ts_out, float *cols_out, int64_t col_stride, void *stream) {
    if (!h || !sfts) return fail(TCW_ERR_INVALID, "tcw_demod_atoms: null argument");
    if (sfts->n_sft < 1) return fail(TCW_ERR_INVALID, "tcw_demod_atoms: no SFTs");
    if (sfts->n_bins < 1) return fail(TCW_ERR_INVALID, "tcw_demod_atoms: no SFT bins");
    // bin indices are int32 in the kernel: the band must lie in 0 .. 2^31 - 1
    if (sfts->k_min < 0 || (int64_t)sfts->k_min + sfts->n_bins > INT32_MAX)
        return fail(TCW_ERR_INVALID, "SFT band must lie in bins 0 .. 2^31-1");
    if (sfts->Tsft < 2) return fail(TCW_ERR_INVALID, "Tsft must be at least 2 seconds.");
    if (dterms < 1 || dterms > DTERMS_MAX)
        return fail(TCW_ERR_INVALID, "dterms must be in 1.." + std::to_string(DTERMS_MAX));
    if (n_sky < 1 || n_sky > (int64_t)DEMOD_GRID_Y)
        return fail(TCW_ERR_INVALID, "sky points must number 1.." + std::to_string(DEMOD_GRID_Y));
    if (n_tmpl < 0) return fail(TCW_ERR_INVALID, "tcw_demod_atoms: negative template count");
    if (n_tmpl == 0) return TCW_OK;
    if (!sfts->t_start || !sfts->bins || !delay || !delay_rate || !a || !b || !freq || !fdot ||
        !ts_out || !cols_out)
        return fail(TCW_ERR_INVALID, "tcw_demod_atoms: null array");
    if (n_tmpl > INT64_MAX / sfts->n_sft / n_sky || col_stride < n_sky * n_tmpl * sfts->n_sft)
        return fail(TCW_ERR_INVALID, "col_stride smaller than n_sky * n_tmpl * n_sft");
    if (!std::isfinite(ref_time)) return fail(TCW_ERR_INVALID, "ref_time must be finite");
    const int rc = set_device(tcwc::handle_device(h));
    if (rc) return rc;
    const int64_t bx = (sfts->n_sft + SFT_TILE - 1) / SFT_TILE;
    if (bx > INT32_MAX) return fail(TCW_ERR_INVALID, "tcw_demod_atoms: too many SFTs");
    // template slots per tile: 64 for a bank of templates, fewer lanes of one
    // SFT when there are only a few templates
    const int tt = n_tmpl >= 32 ? 64 : n_tmpl >= 4 ? 8 : 1;
    const unsigned gy = (unsigned)std::min<int64_t>((n_tmpl + tt - 1) / tt, DEMOD_GRID_Y);
    const dim3 grid((unsigned)bx, gy, (unsigned)n_sky);
    const auto *bins = (const float2 *)sfts->bins;
    const double Tsft = (double)sfts->Tsft;
#define TCW_DEMOD_LAUNCH(TT, DT)                                                               \
    hipLaunchKernelGGL((demod_atoms_kernel<TT, DT>), grid, dim3(DEMOD_WG), 0,                  \
                       (hipStream_t)stream, sfts->n_sft, n_tmpl, sfts->t_start, bins,          \
                       sfts->n_bins, sfts->k_min, Tsft, delay, delay_rate, a, b, freq, fdot,   \
                       ref_time, dterms, ts_out, cols_out, col_stride)
    // dterms 8 (lalpulsar's default) unrolled at compile time
    if (tt == 64) {
        if (dterms == 8) TCW_DEMOD_LAUNCH(64, 8); else TCW_DEMOD_LAUNCH(64, 0);
    } else if (tt == 8) {
        if (dterms == 8) TCW_DEMOD_LAUNCH(8, 8); else TCW_DEMOD_LAUNCH(8, 0);
    } else {
        if (dterms == 8) TCW_DEMOD_LAUNCH(1, 8); else TCW_DEMOD_LAUNCH(1, 0);
    }
#undef TCW_DEMOD_LAUNCH
    TCW_HIP_TRY(hipGetLastError());
    return TCW_OK;
}

}  // namespace

extern "C" {

int tcw_demod_atoms(tcw_handle *h, const tcw_sft_set *sfts, int64_t n_tmpl, const double *freq,
                    const double *fdot, double ref_time, int32_t dterms, uint32_t *ts_out,
                    float *cols_out, int64_t col_stride, void *stream) {
    if (!h || !sfts) return fail(TCW_ERR_INVALID, "tcw_demod_atoms: null argument");
    return demod_launch(h, sfts, 1, sfts->delay, sfts->delay_rate, sfts->a, sfts->b, n_tmpl,
                        freq, fdot, ref_time, dterms, ts_out, cols_out, col_stride, stream);
}

int tcw_demod_atoms_sky(tcw_handle *h, const tcw_sft_set *sfts, int64_t n_sky,
                        const double *delay, const double *delay_rate, const float *a,
                        const float *b, int64_t n_tmpl, const double *freq, const double *fdot,
                        double ref_time, int32_t dterms, uint32_t *ts_out, float *cols_out,
                        int64_t col_stride, void *stream) {
    return demod_launch(h, sfts, n_sky, delay, delay_rate, a, b, n_tmpl, freq, fdot, ref_time,
                        dterms, ts_out, cols_out, col_stride, stream);
}

}  // extern "C"
