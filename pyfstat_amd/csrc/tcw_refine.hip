// tcw_refine.hip -- exponential-window candidates on a fixed atom support
// (gfx950): 2F and lnBtSG with analytic first derivatives, the backward pass
// to the merged atoms, and a multi-step gradient ascent in tau (all steps of a
// candidate in one kernel).
//
// Window sums of a candidate (t0, tau) with support bins k_lo..k_hi:
//   S_c = sum_k w_k^{p_c} x_{c,k},  w_k = exp(-(t_k - t0)/tau),  p = (2,2,2,1,1,1,1)
// for x = (a2, b2, ab, Fa_re, Fa_im, Fb_re, Fb_im), the weighting of
// pyfstat/pyCUDAkernels/cudaTransientFstatExpWindow.cu:82-102 (win for Fa/Fb,
// win^2 for A/B/C), and F from them as at :109-129 (cond < 1e4 guard, F = 2
// otherwise).  Derivatives, with D = AB - C^2 and F = (B|Fa|^2 + A|Fb|^2 -
// 2C Re(Fa Fb*))/D:
//   dF/dA = (|Fb|^2 - F B)/D, dF/dB = (|Fa|^2 - F A)/D, dF/dC = 2(F C - Re(Fa Fb*))/D,
//   dF/dFa = 2(B Fa - C Fb)/D, dF/dFb = 2(A Fb - C Fa)/D   (re and im parts),
//   dS_c/dt0 = p_c S_c / tau,  dS_c/dtau = p_c U_c / tau^2,  U_c = sum_k w_k^{p_c} (t_k - t0) x_{c,k}.
// All sums and the 2x2 algebra run in FP64 (a wave's FP64 FMA issues at the
// rate of an unpacked FP32 one here; FP32 sums would carry an eps32*cond
// error, up to ~6e-4 at cond = 1e4, into F and its derivatives).  t_k is a
// regular grid, so a lane's weights are w0 r^j along its bins and its sums are
// two Horner recurrences (two FMAs a bin and column, group_sums); the lanes
// of a candidate then add their shares with DPP exchanges.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "tcw_common.h"
#include "tcw_hip.h"

using tcwc::DevBuf;
using tcwc::fail;
using tcwc::set_device;

namespace {

constexpr int NC = TCW_NCOL;
constexpr int RWG = 256;             // threads per block
// Supports longer than SHORT4_MAX bins are walked by LONG_G lanes per
// candidate, those longer than WIDE_MIN - 1 bins by WIDE_G lanes, so that no
// lane walks much more than 25 bins per evaluation (the longest serial chains
// of the launch).
constexpr int LONG_G = 16;
constexpr int WIDE_G = 32;
constexpr int WIDE_MIN = 385;
constexpr int CAND_PER_ITEM = RWG / LONG_G;  // most candidates of an LDS item (one block)
constexpr int LDS_ATOMS_CAP = 1696;  // staged bins: 32 B each, 54272 B (three blocks a CU)
constexpr int BWD_WG = 256;          // backward block: BWD_PER bins a thread,
constexpr int BWD_PER = 4;           // BWD_WG apart (coalesced stores)
constexpr int BWD_BINS = BWD_WG * BWD_PER;  // bins per backward block
// Short supports are held in registers, RB bins per lane, loaded once per
// launch: up to RB bins on one lane per candidate, up to 4*RB bins on four
// lanes (two DPP steps instead of four), longer ones on the 16-lane groups
// reading the item's support union from LDS.  Grid maxima mostly sit at short
// windows (median support ~6 bins), where a 16-lane group idles in the bin
// loop and repeats the per-evaluation exp / division / 2x2 algebra 16 times.
constexpr int RB = 8;
constexpr int SHORT1_MAX = RB;
constexpr int SHORT4_MAX = 4 * RB;
constexpr double COND_MAX = 1e4;
constexpr double QCOND = (COND_MAX + 1.0) * (COND_MAX + 1.0) / COND_MAX;  // cond < K <=> T^2 < Q D
constexpr double LN70 = 4.248495242049359;  // log(70)
constexpr int NSAVED = TCW_REFINE_SAVED;
static_assert(NSAVED == NC + 3, "saved row: dF/dt0, dF/dtau, dF/dS_c, F");

struct RTmpl {
    int64_t m_off;             // first merged bin
    int64_t cand_beg, cand_end;
    uint32_t t_data, N;
};

// A work item of the LDS kernels: up to CAND_PER_ITEM long-support
// candidates of one or more templates, each template's support union (a
// segment) staged in LDS, or a single candidate whose support does not fit,
// read from global memory (lds_bins = 0).
struct RItem {
    int64_t cbeg;      // first slot in the plan's candidate order
    int32_t ncand;     // candidates in the item
    int32_t seg_beg;   // first segment
    int32_t nseg;      // segments
    int32_t lds_bins;  // bins staged in LDS; 0: global item
    int32_t g;         // lanes per candidate: LONG_G or WIDE_G
    int32_t pad;
};

// A template's bins u_lo .. u_lo+u_len-1, staged from LDS bin lds_off on.
struct RSeg {
    int32_t t, u_lo, u_len, lds_off;
};

struct BItem {
    int64_t t;
    int32_t b0, nb;  // bins b0 .. b0+nb-1 of template t
};

static_assert(sizeof(RItem) == 32, "RItem layout");

struct Eval {
    double F, dFdt0, dFdtau;
    double dFdlog;  // tau dF/dtau
    double g[NC];
};

// DPP exchanges within a 16-lane row (gfx9 encodings): each pairs every lane
// with exactly one partner, so a + b on one side and b + a on the other give
// bit-identical sums; mirror, half-mirror, xor 2, xor 1 together cover the
// 16 lanes of the group.
constexpr int DPP_ROW_MIRROR = 0x140;
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_QUAD_XOR2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int DPP_QUAD_XOR1 = 0xB1;  // quad_perm [1,0,3,2]

template <int CTRL>
__device__ __forceinline__ double dpp_pair(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One exchange step for all sums at once (the moves of one step read values
// written a step earlier, so they need no wait states between them).  The
// adds must not be contracted with a product that feeds them: a lane would
// add its own unrounded product to its partner's rounded one, the two sides
// of a pair would differ in the last bit, and the lanes of a group would walk
// the ascent on different taus.
template <int CTRL, int N>
__device__ __forceinline__ void group_step(double (&s)[N]) {
#pragma clang fp contract(off)
    double t[N];
#pragma unroll
    for (int i = 0; i < N; i++) t[i] = dpp_pair<CTRL>(s[i]);
#pragma unroll
    for (int i = 0; i < N; i++) s[i] += t[i];
}

// Lane l with lane l ^ 16 of its 32-lane half: ds_swizzle in bit-mask mode
// (source lane = ((l & and) | or) ^ xor with and = 0x1F, or = 0, xor = 0x10).
constexpr int SWZ_XOR16 = 0x1F | (0x10 << 10);

__device__ __forceinline__ double swz_xor16(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_ds_swizzle((int)b, SWZ_XOR16);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), SWZ_XOR16);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int N>
__device__ __forceinline__ void group_step_xor16(double (&s)[N]) {
#pragma clang fp contract(off)
    double t[N];
#pragma unroll
    for (int i = 0; i < N; i++) t[i] = swz_xor16(s[i]);
#pragma unroll
    for (int i = 0; i < N; i++) s[i] += t[i];
}

// Sum each value over the G lanes of an aligned group (G = 1, 4, 8, 16, 32);
// every lane gets the same totals.
template <int G, int N>
__device__ __forceinline__ void group_sum(double (&s)[N]) {
    static_assert(G == 1 || G == 4 || G == 8 || G == 16 || G == 32, "group size");
    if constexpr (G == 32) group_step_xor16(s);
    if constexpr (G >= 16) group_step<DPP_ROW_MIRROR>(s);
    if constexpr (G >= 8) group_step<DPP_ROW_HALF_MIRROR>(s);
    if constexpr (G >= 4) {
        group_step<DPP_QUAD_XOR2>(s);
        group_step<DPP_QUAD_XOR1>(s);
    }
}

// Bin k's seven atoms from an item's LDS copy: two planes of float4 (columns
// 0-3, then columns 4-6 and a pad), `plane` float4s apart, so the lanes of a
// group read consecutive 16-byte words; template-relative bin k at LDS bin
// k - base.
struct LdsBins {
    const float4 *b;
    int base, plane;
    __device__ __forceinline__ int first() const { return base; }
    __device__ __forceinline__ void get(int k, float (&x)[NC]) const {
        const float4 p = b[k - base], q = b[plane + k - base];
        x[0] = p.x;
        x[1] = p.y;
        x[2] = p.z;
        x[3] = p.w;
        x[4] = q.x;
        x[5] = q.y;
        x[6] = q.z;
    }
};

// Bin k's seven atoms from the template's columns in global memory.
struct ColBins {
    const float *g;
    int64_t stride;
    __device__ __forceinline__ int first() const { return 0; }
    __device__ __forceinline__ void get(int k, float (&x)[NC]) const {
#pragma unroll
        for (int c = 0; c < NC; c++) x[c] = g[c * stride + k];
    }
};

// S_c and U_c from a lane's Horner sums: S_c = w0^{p_c} P_c, U_c = w0^{p_c}
// ((u0 - G TA) P_c + G TA Z_c).
__device__ __forceinline__ void horner_finish(const double (&P)[NC], const double (&Z)[NC],
                                              double w0, double a, double b,
                                              double (&s)[2 * NC]) {
    const double w2 = w0 * w0;
    const double bw1 = b * w0, bw2 = b * w2;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        s[c] = (c < 3 ? w2 : w0) * P[c];
        s[NC + c] = fma(a, s[c], (c < 3 ? bw2 : bw1) * Z[c]);
    }
}

// A lane's share of the window sums of its group's candidate: s[0..6] = S_c,
// s[7..13] = U_c over the lane's bins k0 + G j (k0 = kb + gl, j = 0 .. n-1);
// an empty share (k0 > ke) gives zeros.  With w_j = w0 r^j (r = exp(-G TA /
// tau)) and rho_c = r^{p_c}, the lane's bins enter through two Horner sums
// run from its last bin down to its first,
//   P_c = sum_j rho_c^j x_{c,j},   Z_c = sum_j (j+1) rho_c^j x_{c,j},
// (P <- rho P + x_j, Z <- rho Z + P): two FMAs a bin and column, no weight
// recurrence.  Then S_c = w0^{p_c} P_c and, with u_j = u0 + G TA j,
// U_c = w0^{p_c} (u0 P_c + G TA (Z_c - P_c)).
template <int G, typename Bins>
__device__ __forceinline__ void group_sums(const Bins &src, int kb, int ke, double t_data,
                                           double TA, double t0, double inv_tau, int gl,
                                           double (&s)[2 * NC]) {
    const int k0 = kb + gl;
    const double u0 = fma((double)k0, TA, t_data - t0);  // t_k0 - t0
    const double GTA = (double)G * TA;
    // an empty share keeps zeros even where its first weight would overflow
    const bool any = k0 <= ke;
    const double e0 = exp(-u0 * inv_tau);
    const double w0 = any ? e0 : 0.0;
    const double r1 = exp(-GTA * inv_tau);
    const double r2 = r1 * r1;
    // the lane's last bin opens the sums (P = Z = x); an empty share reads the
    // source's first bin and keeps zeros
    const int top = any ? k0 + G * ((ke - k0) / G) : src.first();
    double P[NC], Z[NC];
    {
        float a[NC];
        src.get(top, a);
#pragma unroll
        for (int c = 0; c < NC; c++) Z[c] = P[c] = (double)(any ? a[c] : 0.0f);
    }
    // then down to its first bin (lanes with fewer bins leave the loop early)
#pragma unroll 2
    for (int k = top - G; k >= k0; k -= G) {
        float a[NC];
        src.get(k, a);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            P[c] = fma(c < 3 ? r2 : r1, P[c], (double)a[c]);
            Z[c] = fma(c < 3 ? r2 : r1, Z[c], P[c]);
        }
    }
    horner_finish(P, Z, w0, u0 - GTA, GTA, s);
}

// 1/x for a positive normal x (tau, and D where the cond test passes):
// v_rcp_f64 and two Newton steps, ~1 ulp, half the instructions of a full
// FP64 division
__device__ __forceinline__ double recip_pos(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// F and its derivatives from the lanes' shares of the window sums: the S_c
// are summed over the group; the U_c enter only through dF/dtau =
// sum_c p_c U_c dF/dS_c / tau^2, so each lane folds its shares into one
// number first and 8 values cross lanes instead of 14.  Group-uniform result.
template <int G>
__device__ __forceinline__ Eval eval_from_sums(const double (&s)[2 * NC], double inv_tau) {
    Eval e;
    double S[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) S[c] = s[c];
    group_sum<G>(S);
    const double A = S[0], B = S[1], C = S[2];
    const double far = S[3], fai = S[4], fbr = S[5], fbi = S[6];
    const double D = fma(A, B, -(C * C));
    const double T = A + B;
    const bool ok = (T > 0.0) && (T * T < QCOND * D);
    const double iD = ok ? recip_pos(D) : 0.0;
    const double p = fma(far, far, fai * fai);
    const double q = fma(fbr, fbr, fbi * fbi);
    const double rr = fma(far, fbr, fai * fbi);
    const double F = fma(B, p, fma(A, q, (-2.0 * C) * rr)) * iD;
    e.F = ok ? F : 2.0;
    e.g[0] = (q - F * B) * iD;
    e.g[1] = (p - F * A) * iD;
    e.g[2] = 2.0 * (F * C - rr) * iD;
    e.g[3] = 2.0 * (B * far - C * fbr) * iD;
    e.g[4] = 2.0 * (B * fai - C * fbi) * iD;
    e.g[5] = 2.0 * (A * fbr - C * far) * iD;
    e.g[6] = 2.0 * (A * fbi - C * fai) * iD;
    // sum_c p_c X_c g_c as 2 (terms c < 3) + (terms c >= 3): two short chains
    double d0a = S[0] * e.g[0], d0b = S[3] * e.g[3];
    double d1a = s[NC] * e.g[0], d1b = s[NC + 3] * e.g[3];
#pragma unroll
    for (int c = 1; c < 3; c++) {
        d0a = fma(S[c], e.g[c], d0a);
        d1a = fma(s[NC + c], e.g[c], d1a);
    }
#pragma unroll
    for (int c = 4; c < NC; c++) {
        d0b = fma(S[c], e.g[c], d0b);
        d1b = fma(s[NC + c], e.g[c], d1b);
    }
    double d1[1] = {fma(2.0, d1a, d1b)};
    group_sum<G>(d1);
    e.dFdt0 = fma(2.0, d0a, d0b) * inv_tau;
    e.dFdlog = d1[0] * inv_tau;
    e.dFdtau = e.dFdlog * inv_tau;
    return e;
}

// A candidate's saved row: dF/dt0, dF/dtau, dF/dS_c (7), F (fixed indices,
// so the values stay in registers).
__device__ __forceinline__ void write_saved(double *__restrict__ row, const Eval &e) {
    row[0] = e.dFdt0;
    row[1] = e.dFdtau;
#pragma unroll
    for (int c = 0; c < NC; c++) row[2 + c] = e.g[c];
    row[2 + NC] = e.F;
}

template <int G, typename Bins>
__device__ __forceinline__ Eval eval_bins(const Bins &src, int kb, int ke, double t_data,
                                          double TA, double t0, double inv_tau, int gl) {
    double s[2 * NC];
    group_sums<G>(src, kb, ke, t_data, TA, t0, inv_tau, gl, s);
    return eval_from_sums<G>(s, inv_tau);
}

struct Walk {
    double tau, F;
    int acc;
};

// The ascent of one candidate in s = log(tau): proposal s' = clamp(s + eta g,
// lo, hi) with g = dF/ds, accepted when F rises (then eta *= 1.5, else eta *=
// 0.5); a proposal that cannot move (zero gradient or clamped) is rejected like
// a worse one.  ev(1/tau) evaluates the candidate; proposals are evaluated at
// 1/tau = exp(-s') and the result's tau is exp(s) of the last accepted
// proposal (tau0 itself when none was).  Every group evaluates every step
// (no divergent skips).
template <typename EvalFn>
__device__ __forceinline__ Walk ascend(const EvalFn &ev, double tau0, double lo, double hi,
                                       int n_steps, double step0) {
    const Eval e0 = ev(recip_pos(tau0));
    double s = log(tau0);
    double F = e0.F, g = e0.dFdlog;
    double eta = g != 0.0 ? step0 / fabs(g) : 0.0;
    int acc = 0;
    for (int step = 0; step < n_steps; step++) {
        const double sn = fmin(fmax(fma(eta, g, s), lo), hi);
        const bool moved = sn != s;
        const Eval en = ev(exp(-sn));
        const bool accept = moved && en.F > F;
        if (accept) {
            s = sn;
            F = en.F;
            g = en.dFdlog;
            acc++;
        }
        eta *= accept ? 1.5 : 0.5;
    }
    return Walk{acc > 0 ? exp(s) : tau0, F, acc};
}

// Stage an item's segments into LDS (the two planes of LdsBins, `plane`
// float4s apart; NT threads per block).  Global items stage nothing and need
// no barrier.
template <int NT>
__device__ __forceinline__ void stage_segments(const RItem &it, const RSeg *__restrict__ segs,
                                               const RTmpl *__restrict__ tm,
                                               const float *__restrict__ cols, int64_t cstride,
                                               float *lds, int plane) {
    if (it.lds_bins == 0) return;  // block-uniform
    for (int q = 0; q < it.nseg; q++) {
        const RSeg sg = segs[it.seg_beg + q];
        const float *g = cols + tm[sg.t].m_off + sg.u_lo;
        float4 *b = (float4 *)lds + sg.lds_off;
        for (int k = threadIdx.x; k < sg.u_len; k += NT) {
            float x[NC];
#pragma unroll
            for (int c = 0; c < NC; c++) x[c] = g[c * cstride + k];
            b[k] = make_float4(x[0], x[1], x[2], x[3]);
            b[plane + k] = make_float4(x[4], x[5], x[6], 0.0f);
        }
    }
    __syncthreads();
}

// Candidate slot of a lane: G-lane group j = thread / G handles the item's
// candidate order[cbeg + j], whose template-relative bin k sits at LDS bin
// k + shift[cbeg + j]; lanes without a candidate get an empty support.
struct Slot {
    int64_t ci;
    bool has;
    int kb, ke, shift;
    RTmpl tp;
};

template <int G>
__device__ __forceinline__ Slot lane_slot(const RItem &it, const int32_t *__restrict__ order,
                                          const int32_t *__restrict__ shift,
                                          const int32_t *__restrict__ cand_t,
                                          const RTmpl *__restrict__ tm,
                                          const int32_t *__restrict__ k_lo,
                                          const int32_t *__restrict__ k_hi) {
    const int j = (int)threadIdx.x / G;
    Slot sl;
    sl.has = j < it.ncand;
    const int64_t slot = it.cbeg + (sl.has ? j : 0);
    sl.ci = order[slot];
    sl.shift = shift[slot];
    sl.kb = sl.has ? k_lo[sl.ci] : 1;
    sl.ke = sl.has ? k_hi[sl.ci] : 0;
    sl.tp = tm[cand_t[sl.ci]];
    return sl;
}

template <int G>
__device__ __forceinline__ void forward_block(
    const RItem &it, const float *lds, int plane, const RTmpl *__restrict__ tm,
    const int32_t *__restrict__ order, const int32_t *__restrict__ shift,
    const int32_t *__restrict__ cand_t, const int32_t *__restrict__ k_lo,
    const int32_t *__restrict__ k_hi, const float *__restrict__ cols, int64_t cstride, double TA,
    const double *__restrict__ t0v, const double *__restrict__ tauv, double *__restrict__ twoF,
    double *__restrict__ lnB, double *__restrict__ saved) {
    // waves without a candidate leave (no barrier follows the staging one)
    if (((int)threadIdx.x & ~63) / G >= it.ncand) return;
    const int gl = (int)threadIdx.x % G;
    const Slot sl = lane_slot<G>(it, order, shift, cand_t, tm, k_lo, k_hi);
    const double t0 = sl.has ? t0v[sl.ci] : 0.0;
    const double itau = recip_pos(sl.has ? tauv[sl.ci] : 1.0);
    const double t_data = (double)sl.tp.t_data;
    const Eval e = it.lds_bins > 0  // block-uniform
        ? eval_bins<G>(LdsBins{(const float4 *)lds, -sl.shift, plane}, sl.kb, sl.ke, t_data, TA, t0, itau, gl)
        : eval_bins<G>(ColBins{cols + sl.tp.m_off, cstride}, sl.kb, sl.ke, t_data, TA, t0, itau, gl);
    if (!sl.has) return;
    if (gl == 0) {
        twoF[sl.ci] = 2.0 * e.F;
        if (lnB) lnB[sl.ci] = LN70 + e.F;
        if (saved) write_saved(saved + sl.ci * NSAVED, e);
    }
}

__global__ __launch_bounds__(RWG) void refine_forward_kernel(
    const RTmpl *__restrict__ tm, const RItem *__restrict__ items, const RSeg *__restrict__ segs,
    int plane,
    const int32_t *__restrict__ order, const int32_t *__restrict__ shift,
    const int32_t *__restrict__ cand_t, const int32_t *__restrict__ k_lo,
    const int32_t *__restrict__ k_hi, const float *__restrict__ cols, int64_t cstride, double TA,
    const double *__restrict__ t0v, const double *__restrict__ tauv, double *__restrict__ twoF,
    double *__restrict__ lnB, double *__restrict__ saved) {
    extern __shared__ float lds[];
    const RItem it = items[blockIdx.x];
    stage_segments<RWG>(it, segs, tm, cols, cstride, lds, plane);
    if (it.g == WIDE_G)  // block-uniform
        forward_block<WIDE_G>(it, lds, plane, tm, order, shift, cand_t, k_lo, k_hi, cols, cstride, TA,
                              t0v, tauv, twoF, lnB, saved);
    else
        forward_block<LONG_G>(it, lds, plane, tm, order, shift, cand_t, k_lo, k_hi, cols, cstride, TA,
                              t0v, tauv, twoF, lnB, saved);
}

template <int G>
__device__ __forceinline__ void ascent_block(
    const RItem &it, const float *lds, int plane, const RTmpl *__restrict__ tm,
    const int32_t *__restrict__ order, const int32_t *__restrict__ shift,
    const int32_t *__restrict__ cand_t, const int32_t *__restrict__ k_lo,
    const int32_t *__restrict__ k_hi, const float *__restrict__ cols, int64_t cstride, double TA,
    const double *__restrict__ t0v, const double *tau_iv, double *tau_ov,
    const double *__restrict__ tau_lo, const double *__restrict__ tau_hi, int n_steps,
    double step0, double *__restrict__ twoF, int32_t *__restrict__ n_accept) {
    // waves without a candidate leave (no barrier follows the staging one)
    if (((int)threadIdx.x & ~63) / G >= it.ncand) return;
    const int gl = (int)threadIdx.x % G;
    const Slot sl = lane_slot<G>(it, order, shift, cand_t, tm, k_lo, k_hi);
    // issue ahead of the register kernel's waves, which run beside this
    // kernel to fill its idle SIMD cycles (this kernel's walks are the longer
    // serial chains)
    __builtin_amdgcn_s_setprio(2);
    const double t0 = sl.has ? t0v[sl.ci] : 0.0;
    const double tau_in = sl.has ? tau_iv[sl.ci] : 1.0;
    const bool live = sl.has && tau_in > 0.0;
    const double lo = (live && tau_lo) ? log(tau_lo[sl.ci]) : -INFINITY;
    const double hi = (live && tau_hi) ? log(tau_hi[sl.ci]) : INFINITY;
    const int kb = live ? sl.kb : 1, ke = live ? sl.ke : 0;
    const double t_data = (double)sl.tp.t_data;
    const double tau0 = live ? tau_in : 1.0;
    Walk wk;
    if (it.lds_bins > 0) {  // block-uniform: one step loop per atom source
        const LdsBins src{(const float4 *)lds, -sl.shift, plane};
        wk = ascend([&](double iv) { return eval_bins<G>(src, kb, ke, t_data, TA, t0, iv, gl); },
                    tau0, lo, hi, n_steps, step0);
    } else {
        const ColBins src{cols + sl.tp.m_off, cstride};
        wk = ascend([&](double iv) { return eval_bins<G>(src, kb, ke, t_data, TA, t0, iv, gl); },
                    tau0, lo, hi, n_steps, step0);
    }
    if (sl.has && gl == 0) {
        tau_ov[sl.ci] = live ? wk.tau : tau_in;
        twoF[sl.ci] = live ? 2.0 * wk.F : NAN;
        if (n_accept) n_accept[sl.ci] = wk.acc;
    }
}

__global__ __launch_bounds__(RWG) void refine_ascent_kernel(
    const RTmpl *__restrict__ tm, const RItem *__restrict__ items, const RSeg *__restrict__ segs,
    int plane,
    const int32_t *__restrict__ order, const int32_t *__restrict__ shift,
    const int32_t *__restrict__ cand_t, const int32_t *__restrict__ k_lo,
    const int32_t *__restrict__ k_hi, const float *__restrict__ cols, int64_t cstride, double TA,
    const double *__restrict__ t0v, const double *tau_iv, double *tau_ov,
    const double *__restrict__ tau_lo, const double *__restrict__ tau_hi, int n_steps,
    double step0, double *__restrict__ twoF, int32_t *__restrict__ n_accept) {
    extern __shared__ float lds[];
    const RItem it = items[blockIdx.x];
    stage_segments<RWG>(it, segs, tm, cols, cstride, lds, plane);
    if (it.g == WIDE_G)  // block-uniform
        ascent_block<WIDE_G>(it, lds, plane, tm, order, shift, cand_t, k_lo, k_hi, cols, cstride, TA,
                             t0v, tau_iv, tau_ov, tau_lo, tau_hi, n_steps, step0, twoF, n_accept);
    else
        ascent_block<LONG_G>(it, lds, plane, tm, order, shift, cand_t, k_lo, k_hi, cols, cstride, TA,
                             t0v, tau_iv, tau_ov, tau_lo, tau_hi, n_steps, step0, twoF, n_accept);
}

// ---- short supports in registers: G lanes per candidate (G = 1 or 4) ----
// Lane gl of a group holds bins kb + gl + G*j, j < RB; positions past ke hold
// zeros and add nothing to the sums.
template <int G>
struct RegSlot {
    int64_t ci;
    bool has;
    int kb, ke, gl;
    int nw;  // bins per lane to walk: the maximum over the wave (wave-uniform)
    float x[RB][NC];
};

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m));
    return __builtin_amdgcn_readfirstlane(v);
}

// Called by every lane of the block (the wave maximum needs them all).
template <int G>
__device__ __forceinline__ void reg_slot(RegSlot<G> &r, int64_t blk, const RTmpl *__restrict__ tm,
                                         const int32_t *__restrict__ cand_t,
                                         const int32_t *__restrict__ list, int64_t n_list,
                                         const int32_t *__restrict__ k_lo,
                                         const int32_t *__restrict__ k_hi,
                                         const float *__restrict__ cols, int64_t cstride,
                                         RTmpl &tp) {
    const int64_t idx = blk * RWG + threadIdx.x;
    const int64_t slot = idx / G;
    r.gl = (int)(idx % G);
    r.has = slot < n_list;
    r.ci = r.has ? list[slot] : 0;
    r.kb = r.has ? k_lo[r.ci] : 1;
    r.ke = r.has ? k_hi[r.ci] : 0;
    tp = tm[r.has ? cand_t[r.ci] : 0];
    const float *src = cols + tp.m_off;
    const int first = r.kb + r.gl;
    const int mine = r.ke >= first ? (r.ke - first) / G + 1 : 0;
#pragma unroll
    for (int j = 0; j < RB; j++) {
        const int k = first + G * j;
        if (j < mine) {
#pragma unroll
            for (int c = 0; c < NC; c++) r.x[j][c] = src[c * cstride + k];
        } else {
#pragma unroll
            for (int c = 0; c < NC; c++) r.x[j][c] = 0.0f;
        }
    }
    r.nw = wave_max_int(mine);
}

// The lane's sums in the Horner form of group_sums over its register bins
// (positions past its share hold zeros, which add nothing at the top of a
// Horner sum).
template <int G>
__device__ __forceinline__ Eval eval_regs(const RegSlot<G> &r, double t_data, double TA, double t0,
                                          double inv_tau) {
    const int k0 = r.kb + r.gl;
    const double u0 = fma((double)k0, TA, t_data - t0);  // t_k0 - t0
    const double GTA = (double)G * TA;
    const double w0 = k0 <= r.ke ? exp(-u0 * inv_tau) : 0.0;
    const double r1 = exp(-GTA * inv_tau);
    const double r2 = r1 * r1;
    double P[NC], Z[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) P[c] = Z[c] = 0.0;
#pragma unroll
    for (int j = RB - 1; j >= 0; j--) {
        if (j < r.nw) {
#pragma unroll
            for (int c = 0; c < NC; c++) {
                P[c] = fma(c < 3 ? r2 : r1, P[c], (double)r.x[j][c]);
                Z[c] = fma(c < 3 ? r2 : r1, Z[c], P[c]);
            }
        }
    }
    double s[2 * NC];
    horner_finish(P, Z, w0, u0 - GTA, GTA, s);
    return eval_from_sums<G>(s, inv_tau);
}

// Block blk of a register class (G lanes per candidate of `list`).
template <int G>
__device__ __forceinline__ void forward_regs(
    int64_t blk, const RTmpl *__restrict__ tm, const int32_t *__restrict__ cand_t,
    const int32_t *__restrict__ list, int64_t n_list, const int32_t *__restrict__ k_lo,
    const int32_t *__restrict__ k_hi, const float *__restrict__ cols, int64_t cstride, double TA,
    const double *__restrict__ t0v, const double *__restrict__ tauv, double *__restrict__ twoF,
    double *__restrict__ lnB, double *__restrict__ saved) {
    RegSlot<G> r;
    RTmpl tp;
    reg_slot<G>(r, blk, tm, cand_t, list, n_list, k_lo, k_hi, cols, cstride, tp);
    const double t0 = r.has ? t0v[r.ci] : 0.0;
    const double tau = r.has ? tauv[r.ci] : 1.0;
    const Eval e = eval_regs<G>(r, (double)tp.t_data, TA, t0, recip_pos(tau));
    if (!r.has) return;
    if (r.gl == 0) {
        twoF[r.ci] = 2.0 * e.F;
        if (lnB) lnB[r.ci] = LN70 + e.F;
        if (saved) write_saved(saved + r.ci * NSAVED, e);
    }
}

// The ascent of refine_ascent_kernel, step for step, on register supports.
template <int G>
__device__ __forceinline__ void ascent_regs(
    int64_t blk, const RTmpl *__restrict__ tm, const int32_t *__restrict__ cand_t,
    const int32_t *__restrict__ list, int64_t n_list, const int32_t *__restrict__ k_lo,
    const int32_t *__restrict__ k_hi, const float *__restrict__ cols, int64_t cstride, double TA,
    const double *__restrict__ t0v, const double *tau_iv, double *tau_ov,
    const double *__restrict__ tau_lo, const double *__restrict__ tau_hi, int n_steps,
    double step0, double *__restrict__ twoF, int32_t *__restrict__ n_accept) {
    RegSlot<G> r;
    RTmpl tp;
    reg_slot<G>(r, blk, tm, cand_t, list, n_list, k_lo, k_hi, cols, cstride, tp);
    const double t_data = (double)tp.t_data;
    const double t0 = r.has ? t0v[r.ci] : 0.0;
    const double tau_in = r.has ? tau_iv[r.ci] : 1.0;
    const bool live = r.has && tau_in > 0.0;
    const double lo = (live && tau_lo) ? log(tau_lo[r.ci]) : -INFINITY;
    const double hi = (live && tau_hi) ? log(tau_hi[r.ci]) : INFINITY;
    if (!live) {  // an empty support: F = 2 and no move, as in the group kernel
        r.kb = 1;
        r.ke = 0;
#pragma unroll
        for (int j = 0; j < RB; j++)
#pragma unroll
            for (int c = 0; c < NC; c++) r.x[j][c] = 0.0f;
    }
    const Walk wk = ascend([&](double it) { return eval_regs<G>(r, t_data, TA, t0, it); },
                           live ? tau_in : 1.0, lo, hi, n_steps, step0);
    if (r.has && r.gl == 0) {
        tau_ov[r.ci] = live ? wk.tau : tau_in;
        twoF[r.ci] = live ? 2.0 * wk.F : NAN;
        if (n_accept) n_accept[r.ci] = wk.acc;
    }
}

// One launch for both register classes: blocks below mid_blocks take the
// four-lane list (supports of SHORT1_MAX+1 .. SHORT4_MAX bins), the rest the
// one-lane list.
__global__ __launch_bounds__(RWG) void refine_forward_regs_kernel(
    const RTmpl *__restrict__ tm, const int32_t *__restrict__ cand_t,
    const int32_t *__restrict__ mids, int64_t n_mid, unsigned mid_blocks,
    const int32_t *__restrict__ shorts, int64_t n_short, const int32_t *__restrict__ k_lo,
    const int32_t *__restrict__ k_hi, const float *__restrict__ cols, int64_t cstride, double TA,
    const double *__restrict__ t0v, const double *__restrict__ tauv, double *__restrict__ twoF,
    double *__restrict__ lnB, double *__restrict__ saved) {
    if (blockIdx.x < mid_blocks)  // block-uniform
        forward_regs<4>(blockIdx.x, tm, cand_t, mids, n_mid, k_lo, k_hi, cols, cstride, TA, t0v,
                        tauv, twoF, lnB, saved);
    else
        forward_regs<1>(blockIdx.x - mid_blocks, tm, cand_t, shorts, n_short, k_lo, k_hi, cols,
                        cstride, TA, t0v, tauv, twoF, lnB, saved);
}

__global__ __launch_bounds__(RWG) void refine_ascent_regs_kernel(
    const RTmpl *__restrict__ tm, const int32_t *__restrict__ cand_t,
    const int32_t *__restrict__ mids, int64_t n_mid, unsigned mid_blocks,
    const int32_t *__restrict__ shorts, int64_t n_short, const int32_t *__restrict__ k_lo,
    const int32_t *__restrict__ k_hi, const float *__restrict__ cols, int64_t cstride, double TA,
    const double *__restrict__ t0v, const double *tau_iv, double *tau_ov,
    const double *__restrict__ tau_lo, const double *__restrict__ tau_hi, int n_steps,
    double step0, double *__restrict__ twoF, int32_t *__restrict__ n_accept) {
    if (blockIdx.x < mid_blocks)  // block-uniform
        ascent_regs<4>(blockIdx.x, tm, cand_t, mids, n_mid, k_lo, k_hi, cols, cstride, TA, t0v,
                       tau_iv, tau_ov, tau_lo, tau_hi, n_steps, step0, twoF, n_accept);
    else
        ascent_regs<1>(blockIdx.x - mid_blocks, tm, cand_t, shorts, n_short, k_lo, k_hi, cols,
                       cstride, TA, t0v, tau_iv, tau_ov, tau_lo, tau_hi, n_steps, step0, twoF,
                       n_accept);
}

// Per candidate, once per backward: g = 2 g_twoF + g_lnBtSG, the t0 / tau
// gradients, and what the atom-gradient kernel needs: w_k = exp(-(x0 + k dx))
// with x0 = (t_data - t0)/tau, dx = TAtom/tau, and coef_c = g dF/dS_c.
struct CandCoef {
    double x0, dx;
    float coef[NC];
    float pad;
};
static_assert(sizeof(CandCoef) == 48, "CandCoef layout");

__global__ __launch_bounds__(256) void refine_coef_kernel(
    int64_t n, const RTmpl *__restrict__ tm, const int32_t *__restrict__ cand_t, double TA, const double *__restrict__ t0v,
    const double *__restrict__ tauv, const double *__restrict__ saved,
    const double *__restrict__ g2F, const double *__restrict__ glnB, double *__restrict__ gt0,
    double *__restrict__ gtau, CandCoef *__restrict__ cc) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double gF = (g2F ? 2.0 * g2F[i] : 0.0) + (glnB ? glnB[i] : 0.0);
    if (gt0) gt0[i] = gF * saved[i * NSAVED + 0];
    if (gtau) gtau[i] = gF * saved[i * NSAVED + 1];
    if (cc) {
        const double inv = 1.0 / tauv[i];
        CandCoef c;
        c.x0 = ((double)tm[cand_t[i]].t_data - t0v[i]) * inv;
        c.dx = TA * inv;
#pragma unroll
        for (int k = 0; k < NC; k++) c.coef[k] = (float)(gF * saved[i * NSAVED + 2 + k]);
        c.pad = 0.0f;
        cc[i] = c;
    }
}

// grad_cols for bins b0 .. b0+nb-1 of one template: sum over the template's
// candidates, in order, of coef_c w_k^{p_c} (FP32; w from an FP64 argument
// rounded once).  The candidates' records are staged in LDS BWD_WG at a time
// with one coalesced load; a candidate whose support misses the block is
// skipped (block-uniform test on LDS broadcasts).
__global__ __launch_bounds__(BWD_WG) void refine_backward_kernel(
    const RTmpl *__restrict__ tm, const BItem *__restrict__ bitems, const int32_t *__restrict__ k_lo,
    const int32_t *__restrict__ k_hi, const CandCoef *__restrict__ cc, float *__restrict__ grad,
    int64_t cstride) {
    __shared__ int s_kb[BWD_WG], s_ke[BWD_WG];
    __shared__ double s_x0[BWD_WG], s_dx[BWD_WG];
    __shared__ float s_coef[NC][BWD_WG];
    const BItem bi = bitems[blockIdx.x];
    const RTmpl tp = tm[bi.t];
    const int tid = (int)threadIdx.x;
    float acc[BWD_PER][NC];
#pragma unroll
    for (int j = 0; j < BWD_PER; j++)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[j][c] = 0.0f;
    const int blo = bi.b0, bhi = bi.b0 + bi.nb - 1;
    for (int64_t cb = tp.cand_beg; cb < tp.cand_end; cb += BWD_WG) {
        const int n = (int)min<int64_t>(BWD_WG, tp.cand_end - cb);
        if (tid < n) {
            const int64_t i = cb + tid;
            const CandCoef c = cc[i];
            s_kb[tid] = k_lo[i];
            s_ke[tid] = k_hi[i];
            s_x0[tid] = c.x0;
            s_dx[tid] = c.dx;
#pragma unroll
            for (int q = 0; q < NC; q++) s_coef[q][tid] = c.coef[q];
        }
        __syncthreads();
        for (int i = 0; i < n; i++) {
            const int kb = s_kb[i], ke = s_ke[i];
            if (ke < blo || kb > bhi) continue;  // block-uniform
            const double x0 = s_x0[i], dx = s_dx[i];
            float cf[NC];
#pragma unroll
            for (int q = 0; q < NC; q++) cf[q] = s_coef[q][i];
#pragma unroll
            for (int j = 0; j < BWD_PER; j++) {
                const int k = bi.b0 + tid + j * BWD_WG;
                if (k >= kb && k <= ke) {
                    const float w = expf((float)-fma((double)k, dx, x0));
                    const float w2 = w * w;
#pragma unroll
                    for (int q = 0; q < NC; q++) acc[j][q] = fmaf(cf[q], q < 3 ? w2 : w, acc[j][q]);
                }
            }
        }
        __syncthreads();  // before the next chunk overwrites the records
    }
#pragma unroll
    for (int j = 0; j < BWD_PER; j++) {
        const int r = tid + j * BWD_WG;
        if (r < bi.nb) {
#pragma unroll
            for (int c = 0; c < NC; c++) grad[c * cstride + tp.m_off + bi.b0 + r] = acc[j][c];
        }
    }
}

}  // namespace

struct tcw_refine_plan {
    int device = 0;
    int64_t T = 0, n_cand = 0, total_bins = 0;
    uint32_t TAtom = 0;
    int lds_bytes = 0, lds_plane = 0;  // LDS per block; float4s between the planes
    int64_t n_items = 0, n_bitems = 0, n_long = 0, n_mid = 0;
    // d_order: candidate indices; long supports item by item (within an item
    // by decreasing length), then supports of SHORT1_MAX+1 .. SHORT4_MAX
    // bins (four lanes each), then those of up to SHORT1_MAX bins (one lane
    // each), the last two classes by decreasing length.  d_shift: for each
    // long slot, LDS bin minus template bin of its item's copy.
    DevBuf d_tmpl, d_items, d_segs, d_shift, d_bitems, d_klo, d_khi, d_ctmpl, d_coef, d_order;
    // the register kernels run on `side`, beside the LDS kernel on the
    // caller's stream, so their blocks fill the LDS kernel's tail
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    ~tcw_refine_plan() {
        if (side) (void)hipStreamDestroy(side);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        d_order.release();
        d_ctmpl.release();
        d_coef.release();
        d_tmpl.release();
        d_items.release();
        d_segs.release();
        d_shift.release();
        d_bitems.release();
        d_klo.release();
        d_khi.release();
    }
};

namespace {

// Consecutive long-support candidates of one template, longs[first ..
// first+n), and the union of their supports, template bins lo .. lo+len-1.
struct Chunk {
    int32_t t, first, n, lo, len;
    int32_t g;  // lanes per candidate
};

// Pack chunks into items of one lane count each.  A chunk whose union fits in
// LDS goes (chunks with more candidates first) into the open item of its lane
// count with the fewest free candidate slots that still takes its candidates
// (RWG / g a block) and bins (LDS_ATOMS_CAP), looking at the latest item of
// each free-slot count only (linear time); a chunk that does not fit in LDS
// is an item of its own, read from global memory.  Items run by decreasing total
// support bins (the heaviest blocks start first); within an item the
// candidates run by decreasing length, so the groups of a wave walk similar
// bin counts.  Appends the items, their segments, and the long candidates'
// order and LDS shifts; returns the most bins an item stages.
int pack_items(const std::vector<Chunk> &chunks, const std::vector<int32_t> &longs,
               const int32_t *k_lo, const int32_t *k_hi, std::vector<RItem> &items,
               std::vector<RSeg> &segs, std::vector<int32_t> &order,
               std::vector<int32_t> &shifts) {
    struct Pack {
        std::vector<int32_t> chunk;
        int ncand = 0, bins = 0;  // bins = 0: a global item
        int g = LONG_G;
        int64_t work = 0;
    };
    const auto len = [&](int32_t c) { return k_hi[c] - k_lo[c] + 1; };
    std::vector<int32_t> ids(chunks.size());
    for (size_t i = 0; i < ids.size(); i++) ids[i] = (int32_t)i;
    std::stable_sort(ids.begin(), ids.end(),
                     [&](int32_t a, int32_t b) { return chunks[a].n > chunks[b].n; });
    std::vector<Pack> packs;
    // open items by lane count (0: LONG_G, 1: WIDE_G) and free slots
    std::vector<std::vector<int32_t>> open[2];
    open[0].resize((size_t)CAND_PER_ITEM + 1);
    open[1].resize((size_t)CAND_PER_ITEM + 1);
    for (const int32_t id : ids) {
        const Chunk &c = chunks[(size_t)id];
        const int cap = RWG / c.g;
        std::vector<std::vector<int32_t>> &op = open[c.g == WIDE_G ? 1 : 0];
        int64_t w = 0;
        for (int k = 0; k < c.n; k++) w += len(longs[(size_t)(c.first + k)]);
        const bool in_lds = c.len <= LDS_ATOMS_CAP;
        int32_t di = -1;
        for (int f = c.n; in_lds && di < 0 && f <= cap; f++) {
            std::vector<int32_t> &st = op[(size_t)f];
            if (!st.empty() && packs[(size_t)st.back()].bins + c.len <= LDS_ATOMS_CAP) {
                di = st.back();
                st.pop_back();
            }
        }
        if (di < 0) {
            di = (int32_t)packs.size();
            packs.emplace_back();
            packs.back().g = c.g;
        }
        Pack &q = packs[(size_t)di];
        q.chunk.push_back(id);
        q.ncand += c.n;
        q.bins += in_lds ? c.len : 0;
        q.work += w;
        if (in_lds && q.ncand < cap) op[(size_t)(cap - q.ncand)].push_back(di);
    }
    std::stable_sort(packs.begin(), packs.end(),
                     [](const Pack &a, const Pack &b) { return a.work > b.work; });
    int max_bins = 0;
    std::vector<std::pair<int32_t, int32_t>> cs;  // (candidate, shift)
    for (const Pack &q : packs) {
        RItem it{};
        it.cbeg = (int64_t)order.size();
        it.ncand = q.ncand;
        it.seg_beg = (int32_t)segs.size();
        it.nseg = (int32_t)q.chunk.size();
        it.lds_bins = q.bins;
        it.g = q.g;
        cs.clear();
        int off = 0;
        for (const int32_t id : q.chunk) {
            const Chunk &c = chunks[(size_t)id];
            segs.push_back(RSeg{c.t, c.lo, c.len, off});
            for (int k = 0; k < c.n; k++) cs.emplace_back(longs[(size_t)(c.first + k)], off - c.lo);
            off += c.len;
        }
        std::stable_sort(cs.begin(), cs.end(),
                         [&](const std::pair<int32_t, int32_t> &a,
                             const std::pair<int32_t, int32_t> &b) { return len(a.first) > len(b.first); });
        for (const auto &x : cs) {
            order.push_back(x.first);
            shifts.push_back(q.bins > 0 ? x.second : 0);
        }
        max_bins = std::max(max_bins, q.bins);
        items.push_back(it);
    }
    return max_bins;
}

// blocks for n candidates at G lanes each
unsigned reg_blocks(int64_t n, int G) { return (unsigned)((n * G + RWG - 1) / RWG); }

// The stream for the register kernels: the plan's side stream, ordered after
// the work already queued on st, when both kinds of kernel run; else st.
int fork_side(tcw_refine_plan *p, hipStream_t st, hipStream_t *out) {
    *out = st;
    if (p->n_items == 0 || p->n_long == p->n_cand) return TCW_OK;
    TCW_HIP_TRY(hipEventRecord(p->ev_fork, st));
    TCW_HIP_TRY(hipStreamWaitEvent(p->side, p->ev_fork, 0));
    *out = p->side;
    return TCW_OK;
}

// Work queued on st after this waits for the side stream's kernels.
int join_side(tcw_refine_plan *p, hipStream_t st, hipStream_t side) {
    if (side == st) return TCW_OK;
    TCW_HIP_TRY(hipEventRecord(p->ev_join, side));
    TCW_HIP_TRY(hipStreamWaitEvent(st, p->ev_join, 0));
    return TCW_OK;
}

int check_plan(const tcw_refine_plan *p) {
    if (!p) return fail(TCW_ERR_INVALID, "null refine plan");
    return set_device(p->device);
}

}  // namespace

extern "C" {

int tcw_refine_plan_create(tcw_handle *h, const tcw_refine_desc *d, tcw_refine_plan **out) {
    if (!h || !d || !out) return fail(TCW_ERR_INVALID, "tcw_refine_plan_create: null argument");
    if (d->T <= 0 || d->T > INT32_MAX) return fail(TCW_ERR_INVALID, "template count out of range");
    if (d->n_cand < 0 || d->n_cand > INT32_MAX)
        return fail(TCW_ERR_INVALID, "candidate count out of range");
    if (d->TAtom < 2) return fail(TCW_ERR_INVALID, "TAtom must be at least 2 seconds.");
    if (!d->m_off || !d->t_data || !d->cand_off || (d->n_cand > 0 && (!d->k_lo || !d->k_hi)))
        return fail(TCW_ERR_INVALID, "tcw_refine_plan_create: null array");
    if (d->m_off[0] < 0 || d->cand_off[0] != 0 || d->cand_off[d->T] != d->n_cand)
        return fail(TCW_ERR_INVALID, "cand_off must run from 0 to n_cand");
    std::vector<RTmpl> tm((size_t)d->T);
    std::vector<int32_t> longs, mids, shorts;
    std::vector<Chunk> chunks;
    std::vector<BItem> bitems;
    for (int64_t t = 0; t < d->T; t++) {
        const int64_t N = d->m_off[t + 1] - d->m_off[t];
        if (N <= 0 || N > INT32_MAX / 2)
            return fail(TCW_ERR_INVALID, "template " + std::to_string(t) + ": bad bin range");
        const int64_t c0 = d->cand_off[t], c1 = d->cand_off[t + 1];
        if (c1 < c0) return fail(TCW_ERR_INVALID, "cand_off not ascending");
        tm[(size_t)t] = RTmpl{d->m_off[t], c0, c1, d->t_data[t], (uint32_t)N};
        const double t_last = (double)d->t_data[t] + (double)(N - 1) * d->TAtom;
        if (t_last >= 4294967296.0)
            return fail(TCW_ERR_INVALID, "template " + std::to_string(t) + ": bins pass 2^32 s");
        for (int64_t i = c0; i < c1; i++)
            if (d->k_lo[i] < 0 || d->k_hi[i] < d->k_lo[i] || d->k_hi[i] >= N)
                return fail(TCW_ERR_INVALID, "candidate " + std::to_string(i) +
                                                 ": support outside its template's bins");
        // short supports go to the register lists; the long ones of a
        // template, by lane count, form chunks of consecutive candidates
        // whose union fits in LDS (a support too long for LDS starts a chunk
        // read from global memory)
        for (const int g : {LONG_G, WIDE_G}) {
            const size_t l0 = longs.size();
            for (int64_t i = c0; i < c1; i++) {
                const int len = d->k_hi[i] - d->k_lo[i] + 1;
                if (g == LONG_G && len <= SHORT1_MAX)
                    shorts.push_back((int32_t)i);
                else if (g == LONG_G && len <= SHORT4_MAX)
                    mids.push_back((int32_t)i);
                else if (len > SHORT4_MAX && (len >= WIDE_MIN) == (g == WIDE_G))
                    longs.push_back((int32_t)i);
            }
            size_t i = l0;
            while (i < longs.size()) {
                int lo = d->k_lo[longs[i]], hi = d->k_hi[longs[i]];
                size_t j = i + 1;
                while (j < longs.size() && (int)(j - i) < RWG / g) {
                    const int32_t cj = longs[j];
                    const int nlo = std::min(lo, d->k_lo[cj]), nhi = std::max(hi, d->k_hi[cj]);
                    if (nhi - nlo + 1 > LDS_ATOMS_CAP && hi - lo + 1 <= LDS_ATOMS_CAP) break;
                    lo = nlo;
                    hi = nhi;
                    j++;
                }
                chunks.push_back(Chunk{(int32_t)t, (int32_t)i, (int32_t)(j - i), lo, hi - lo + 1, g});
                i = j;
            }
        }
        // every bin gets a backward slot (templates without candidates get zeros)
        for (int64_t b = 0; b < N; b += BWD_BINS)
            bitems.push_back(BItem{t, (int32_t)b, (int32_t)std::min<int64_t>(BWD_BINS, N - b)});
    }
    std::vector<RItem> items;
    std::vector<RSeg> segs;
    std::vector<int32_t> order, shifts;
    const int max_bins = pack_items(chunks, longs, d->k_lo, d->k_hi, items, segs, order, shifts);
    const int device = tcwc::handle_device(h);
    int rc = set_device(device);
    if (rc) return rc;
    auto *p = new tcw_refine_plan;
    p->device = device;
    p->T = d->T;
    p->n_cand = d->n_cand;
    p->TAtom = d->TAtom;
    p->total_bins = d->m_off[d->T];
    p->lds_plane = max_bins;
    p->lds_bytes = (int)(max_bins * 2 * sizeof(float4));
    p->n_items = (int64_t)items.size();
    p->n_bitems = (int64_t)bitems.size();
    p->n_long = (int64_t)order.size();
    p->n_mid = (int64_t)mids.size();
    // by decreasing length, so the lanes of a wave walk similar bin counts
    const auto longer = [d](int32_t a, int32_t b) {
        return d->k_hi[a] - d->k_lo[a] > d->k_hi[b] - d->k_lo[b];
    };
    std::stable_sort(mids.begin(), mids.end(), longer);
    std::stable_sort(shorts.begin(), shorts.end(), longer);
    order.insert(order.end(), mids.begin(), mids.end());
    order.insert(order.end(), shorts.begin(), shorts.end());
    const size_t nc = (size_t)std::max<int64_t>(d->n_cand, 1);
    if ((rc = p->d_tmpl.ensure(tm.size() * sizeof(RTmpl))) ||
        (rc = p->d_items.ensure(std::max<size_t>(items.size(), 1) * sizeof(RItem))) ||
        (rc = p->d_segs.ensure(std::max<size_t>(segs.size(), 1) * sizeof(RSeg))) ||
        (rc = p->d_shift.ensure(std::max<size_t>(shifts.size(), 1) * sizeof(int32_t))) ||
        (rc = p->d_bitems.ensure(bitems.size() * sizeof(BItem))) ||
        (rc = p->d_klo.ensure(nc * sizeof(int32_t))) || (rc = p->d_khi.ensure(nc * sizeof(int32_t))) ||
        (rc = p->d_ctmpl.ensure(nc * sizeof(int32_t))) || (rc = p->d_coef.ensure(nc * sizeof(CandCoef))) ||
        (rc = p->d_order.ensure(nc * sizeof(int32_t)))) {
        delete p;
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_tmpl.p, tm.data(), tm.size() * sizeof(RTmpl), hipMemcpyHostToDevice);
    if (e == hipSuccess && !items.empty())
        e = hipMemcpy(p->d_items.p, items.data(), items.size() * sizeof(RItem), hipMemcpyHostToDevice);
    if (e == hipSuccess && !segs.empty())
        e = hipMemcpy(p->d_segs.p, segs.data(), segs.size() * sizeof(RSeg), hipMemcpyHostToDevice);
    if (e == hipSuccess && !shifts.empty())
        e = hipMemcpy(p->d_shift.p, shifts.data(), shifts.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_bitems.p, bitems.data(), bitems.size() * sizeof(BItem),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess && d->n_cand > 0)
        e = hipMemcpy(p->d_klo.p, d->k_lo, (size_t)d->n_cand * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && d->n_cand > 0)
        e = hipMemcpy(p->d_khi.p, d->k_hi, (size_t)d->n_cand * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && d->n_cand > 0)
        e = hipMemcpy(p->d_order.p, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && d->n_cand > 0) {
        std::vector<int32_t> ct((size_t)d->n_cand);
        for (int64_t t = 0; t < d->T; t++)
            for (int64_t i = d->cand_off[t]; i < d->cand_off[t + 1]; i++) ct[(size_t)i] = (int32_t)t;
        e = hipMemcpy(p->d_ctmpl.p, ct.data(), ct.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        delete p;
        return fail(TCW_ERR_DEVICE, std::string("refine plan upload: ") + hipGetErrorString(e));
    }
    *out = p;
    return TCW_OK;
}

int tcw_refine_plan_destroy(tcw_refine_plan *p) {
    if (p) (void)set_device(p->device);
    delete p;
    return TCW_OK;
}

int tcw_refine_forward(tcw_refine_plan *p, const float *cols, int64_t col_stride, const double *t0,
                       const double *tau, double *twoF, double *lnBtSG, double *saved,
                       void *stream) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (p->n_cand == 0) return TCW_OK;
    if (!cols || !t0 || !tau || !twoF) return fail(TCW_ERR_INVALID, "tcw_refine_forward: null argument");
    if (col_stride < p->total_bins) return fail(TCW_ERR_INVALID, "col_stride smaller than the merged bins");
    const int32_t *order = (const int32_t *)p->d_order.p;
    const hipStream_t st = (hipStream_t)stream;
    // one evaluation: both launches on st (a fork and join to the side stream
    // costs more than their overlap saves)
    if (p->n_items > 0)
        hipLaunchKernelGGL(refine_forward_kernel, dim3((unsigned)p->n_items),
                           dim3(RWG), p->lds_bytes, st, (const RTmpl *)p->d_tmpl.p,
                           (const RItem *)p->d_items.p, (const RSeg *)p->d_segs.p, p->lds_plane, order,
                           (const int32_t *)p->d_shift.p, (const int32_t *)p->d_ctmpl.p,
                           (const int32_t *)p->d_klo.p,
                           (const int32_t *)p->d_khi.p, cols, col_stride, (double)p->TAtom, t0, tau,
                           twoF, lnBtSG, saved);
    const int64_t n_short = p->n_cand - p->n_long - p->n_mid;
    const unsigned mid_blocks = reg_blocks(p->n_mid, 4), short_blocks = reg_blocks(n_short, 1);
    if (mid_blocks + short_blocks > 0)
        hipLaunchKernelGGL(refine_forward_regs_kernel, dim3(mid_blocks + short_blocks), dim3(RWG), 0,
                           st, (const RTmpl *)p->d_tmpl.p, (const int32_t *)p->d_ctmpl.p,
                           order + p->n_long, p->n_mid, mid_blocks, order + p->n_long + p->n_mid,
                           n_short, (const int32_t *)p->d_klo.p, (const int32_t *)p->d_khi.p, cols,
                           col_stride, (double)p->TAtom, t0, tau, twoF, lnBtSG, saved);
    TCW_HIP_TRY(hipGetLastError());
    return TCW_OK;
}

int tcw_refine_backward(tcw_refine_plan *p, const double *t0, const double *tau, const double *saved,
                        const double *g_twoF, const double *g_lnBtSG, float *grad_cols,
                        int64_t col_stride, double *grad_t0, double *grad_tau, void *stream) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!t0 || !tau || !saved) return fail(TCW_ERR_INVALID, "tcw_refine_backward: null argument");
    hipStream_t st = (hipStream_t)stream;
    if (grad_cols && col_stride < p->total_bins)
        return fail(TCW_ERR_INVALID, "col_stride smaller than the merged bins");
    if (p->n_cand > 0 && (grad_cols || grad_t0 || grad_tau))
        hipLaunchKernelGGL(refine_coef_kernel, dim3((unsigned)((p->n_cand + 255) / 256)), dim3(256),
                           0, st, p->n_cand, (const RTmpl *)p->d_tmpl.p,
                           (const int32_t *)p->d_ctmpl.p, (double)p->TAtom, t0, tau, saved, g_twoF,
                           g_lnBtSG, grad_t0, grad_tau,
                           grad_cols ? (CandCoef *)p->d_coef.p : nullptr);
    if (grad_cols)
        hipLaunchKernelGGL(refine_backward_kernel, dim3((unsigned)p->n_bitems), dim3(BWD_WG), 0, st,
                           (const RTmpl *)p->d_tmpl.p, (const BItem *)p->d_bitems.p,
                           (const int32_t *)p->d_klo.p, (const int32_t *)p->d_khi.p,
                           (const CandCoef *)p->d_coef.p, grad_cols, col_stride);
    TCW_HIP_TRY(hipGetLastError());
    return TCW_OK;
}

int tcw_refine_ascent(tcw_refine_plan *p, const float *cols, int64_t col_stride, const double *t0,
                      const double *tau_in, double *tau_out, const double *tau_lo,
                      const double *tau_hi, int n_steps, double step0, double *twoF,
                      int32_t *n_accept, void *stream) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (p->n_cand == 0) return TCW_OK;
    if (!cols || !t0 || !tau_in || !tau_out || !twoF)
        return fail(TCW_ERR_INVALID, "tcw_refine_ascent: null argument");
    if (n_steps < 0) return fail(TCW_ERR_INVALID, "n_steps must be >= 0");
    if (!(step0 > 0.0) || !std::isfinite(step0)) return fail(TCW_ERR_INVALID, "step0 must be > 0");
    if (col_stride < p->total_bins) return fail(TCW_ERR_INVALID, "col_stride smaller than the merged bins");
    const int32_t *order = (const int32_t *)p->d_order.p;
    const hipStream_t st = (hipStream_t)stream;
    hipStream_t side;
    if ((rc = fork_side(p, st, &side))) return rc;
    if (p->n_items > 0)
        hipLaunchKernelGGL(refine_ascent_kernel, dim3((unsigned)p->n_items),
                           dim3(RWG), p->lds_bytes, st, (const RTmpl *)p->d_tmpl.p,
                           (const RItem *)p->d_items.p, (const RSeg *)p->d_segs.p, p->lds_plane, order,
                           (const int32_t *)p->d_shift.p, (const int32_t *)p->d_ctmpl.p,
                           (const int32_t *)p->d_klo.p,
                           (const int32_t *)p->d_khi.p, cols, col_stride, (double)p->TAtom, t0,
                           tau_in, tau_out, tau_lo, tau_hi, n_steps, step0, twoF, n_accept);
    const int64_t n_short = p->n_cand - p->n_long - p->n_mid;
    const unsigned mid_blocks = reg_blocks(p->n_mid, 4), short_blocks = reg_blocks(n_short, 1);
    if (mid_blocks + short_blocks > 0)
        hipLaunchKernelGGL(refine_ascent_regs_kernel, dim3(mid_blocks + short_blocks), dim3(RWG), 0,
                           side, (const RTmpl *)p->d_tmpl.p, (const int32_t *)p->d_ctmpl.p,
                           order + p->n_long, p->n_mid, mid_blocks, order + p->n_long + p->n_mid,
                           n_short, (const int32_t *)p->d_klo.p, (const int32_t *)p->d_khi.p, cols,
                           col_stride, (double)p->TAtom, t0, tau_in, tau_out, tau_lo, tau_hi,
                           n_steps, step0, twoF, n_accept);
    TCW_HIP_TRY(hipGetLastError());
    return join_side(p, st, side);
}

}  // extern "C"
