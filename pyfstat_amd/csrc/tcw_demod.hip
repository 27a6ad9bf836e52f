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
// One thread per (template, SFT): the SFT's timing, Doppler factor and
// antenna patterns are read once and kept for the templates of its column of
// blocks (grid y strides over the templates); the phase is FP64 (f * dtau
// reaches 1e10 cycles), the kernel weights and the complex sum float32 (the
// atoms are float32).  Bound: the 32 bytes of output per atom (HBM writes);
// the bins read are shared by neighbouring templates and stay in L2.

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
constexpr int DTERMS_MAX = 64;
constexpr unsigned DEMOD_GRID_Y = 65535;

__global__ __launch_bounds__(DEMOD_WG) void demod_atoms_kernel(
    int64_t n_sft, int64_t n_tmpl, const uint32_t *__restrict__ t_start,
    const float2 *__restrict__ bins, int n_bins, int k_min, double Tsft,
    const double *__restrict__ delay, const double *__restrict__ delay_rate,
    const float *__restrict__ pa, const float *__restrict__ pb, const double *__restrict__ freq,
    const double *__restrict__ fdot, double ref_time, int dterms, uint32_t *__restrict__ ts_out,
    float *__restrict__ cols, int64_t cstride) {
    const int64_t a = (int64_t)blockIdx.x * DEMOD_WG + threadIdx.x;
    if (a >= n_sft) return;
    const uint32_t ts = t_start[a];
    const double dtau = (double)ts + 0.5 * Tsft + delay[a] - ref_time;
    const double dop = (1.0 + delay_rate[a]) * Tsft;
    const float av = pa[a], bv = pb[a];
    const float2 *__restrict__ row = bins + a * n_bins;
    for (int64_t t = blockIdx.y; t < n_tmpl; t += gridDim.y) {
        const double f = freq[t], fd = fdot[t];
        const double phase = dtau * fma(0.5 * fd, dtau, f);  // cycles
        const double kappa = fma(fd, dtau, f) * dop;           // observed frequency in bins
        const double kf = floor(kappa);
        const double delta = kappa - kf;
        const int64_t kfi = (int64_t)kf;
        // first stored index of the window k = kf-dterms+1 .. kf+dterms
        const int64_t j0 = kfi - dterms + 1 - k_min;
        float re = 0.0f, im = 0.0f;
        if (delta == 0.0) {  // kappa on a bin: the kernel is that bin alone
            const int64_t j = kfi - k_min;
            if (j >= 0 && j < n_bins) {
                const float s = (kfi & 1) ? -1.0f : 1.0f;
                re = s * row[j].x;
                im = s * row[j].y;
            }
        } else {
            // K = (-1)^kf sin(pi delta) / (pi (kappa - k)), kappa - k = delta + dterms - 1 - i
            const float amp = (float)(sinpi(delta) * (1.0 / M_PI)) * ((kfi & 1) ? -1.0f : 1.0f);
            const float dl = (float)delta;
            for (int i = 0; i < 2 * dterms; i++) {
                const int64_t j = j0 + i;
                if (j >= 0 && j < n_bins) {
                    const float w = amp / (dl + (float)(dterms - 1 - i));
                    const float2 y = row[j];
                    re = fmaf(w, y.x, re);
                    im = fmaf(w, y.y, im);
                }
            }
        }
        // rotate by exp(-2 pi i phase): the fractional cycle in FP64, its
        // sine and cosine in float
        const float fr = (float)(phase - floor(phase));
        float sn, cs;
        sincospif(2.0f * fr, &sn, &cs);
        const float xr = fmaf(re, cs, im * sn);
        const float xi = fmaf(im, cs, -(re * sn));
        const int64_t o = t * n_sft + a;
        ts_out[o] = ts;
        cols[o] = av * av;
        cols[cstride + o] = bv * bv;
        cols[2 * cstride + o] = av * bv;
        cols[3 * cstride + o] = av * xr;
        cols[4 * cstride + o] = av * xi;
        cols[5 * cstride + o] = bv * xr;
        cols[6 * cstride + o] = bv * xi;
    }
}

}  // namespace

extern "C" {

int tcw_demod_atoms(tcw_handle *h, const tcw_sft_set *sfts, int64_t n_tmpl, const double *freq,
                    const double *fdot, double ref_time, int32_t dterms, uint32_t *ts_out,
                    float *cols_out, int64_t col_stride, void *stream) {
    if (!h || !sfts) return fail(TCW_ERR_INVALID, "tcw_demod_atoms: null argument");
    if (sfts->n_sft < 1) return fail(TCW_ERR_INVALID, "tcw_demod_atoms: no SFTs");
    if (sfts->n_bins < 1) return fail(TCW_ERR_INVALID, "tcw_demod_atoms: no SFT bins");
    if (sfts->Tsft < 2) return fail(TCW_ERR_INVALID, "Tsft must be at least 2 seconds.");
    if (dterms < 1 || dterms > DTERMS_MAX)
        return fail(TCW_ERR_INVALID, "dterms must be in 1.." + std::to_string(DTERMS_MAX));
    if (n_tmpl < 0) return fail(TCW_ERR_INVALID, "tcw_demod_atoms: negative template count");
    if (n_tmpl == 0) return TCW_OK;
    if (!sfts->t_start || !sfts->bins || !sfts->delay || !sfts->delay_rate || !sfts->a ||
        !sfts->b || !freq || !fdot || !ts_out || !cols_out)
        return fail(TCW_ERR_INVALID, "tcw_demod_atoms: null array");
    if (n_tmpl > INT64_MAX / sfts->n_sft || col_stride < n_tmpl * sfts->n_sft)
        return fail(TCW_ERR_INVALID, "col_stride smaller than n_tmpl * n_sft");
    if (!std::isfinite(ref_time)) return fail(TCW_ERR_INVALID, "ref_time must be finite");
    const int rc = set_device(tcwc::handle_device(h));
    if (rc) return rc;
    const int64_t bx = (sfts->n_sft + DEMOD_WG - 1) / DEMOD_WG;
    if (bx > INT32_MAX) return fail(TCW_ERR_INVALID, "tcw_demod_atoms: too many SFTs");
    const unsigned gy = (unsigned)std::min<int64_t>(n_tmpl, DEMOD_GRID_Y);
    hipLaunchKernelGGL(demod_atoms_kernel, dim3((unsigned)bx, gy), dim3(DEMOD_WG), 0,
                       (hipStream_t)stream, sfts->n_sft, n_tmpl, sfts->t_start,
                       (const float2 *)sfts->bins, sfts->n_bins, sfts->k_min,
                       (double)sfts->Tsft, sfts->delay, sfts->delay_rate, sfts->a, sfts->b, freq,
                       fdot, ref_time, dterms, ts_out, cols_out, col_stride);
    TCW_HIP_TRY(hipGetLastError());
    return TCW_OK;
}

}  // extern "C"
