// tcw_map.hip -- transient-CW F-statistic maps on MI355X (gfx950).
//
// Device path for one or many templates:
//   1. stage_kernel    binned detector merge (lalpulsar mergeMultiFstatAtomsBinned,
//                      called at pyfstat/tcw_fstat_map_funcs.py:706) fused with an
//                      FP64 prefix scan of the float32 merged bins (rect windows);
//                      stage_exp_kernel: the same merge into a float32 atom table
//                      (exp windows);
//   2. map kernels     window sums -> cond-guarded F per (t0, tau) cell, with the
//                      max/argmax, logsumexp and row/column posterior sums fused
//                      into the producer (pyfstat/tcw_fstat_map_funcs.py:809-828);
//   3. finalize_kernel per-template combination of the partials into FstatMap
//                      fields (maxF, t0_ML, tau_ML, lnBtSG, t0_MP, tau_MP).
// Host side: the C ABI declared in include/tcw_hip.h.
//
// Index maths and window semantics follow the reference kernels
// pyfstat/pyCUDAkernels/cudaTransientFstatRectWindow.cu:21-124 and
// pyfstat/pyCUDAkernels/cudaTransientFstatExpWindow.cu:27-134 (uint32
// wrap-around, clamping, incremental rect sums, cond < 1e4 guard, F = 2
// fallback); the arithmetic is new (prefix differences, FP64 determinant).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "tcw_common.h"
#include "tcw_hip.h"

// Device memory of tcw_map_noalloc: the single-map plan's buffers (descriptors,
// prefix or atom tables, reduction partials) are carved out of this array, so
// that call allocates nothing.  Each device holds its own copy from module
// load until the process ends; host code never uses this name as an address
// (it is a host-side stand-in there), only the pointer hipGetSymbolAddress
// returns for the current device (noalloc_scratch_ptr below).
static __device__ __align__(256) unsigned char scratch[TCW_NOALLOC_SCRATCH_BYTES];

namespace {

// ----------------------------------------------------------------------------
// constants and device-side descriptors
// ----------------------------------------------------------------------------

constexpr int NC = TCW_NCOL;
constexpr int WG = 256;           // threads per workgroup (4 wave64)
constexpr int TILE_COLS = 256;    // tau columns per work item
constexpr int MERGED_STRIDE = 8;  // floats per atom in the atom table (7 columns + pad)
constexpr double COND_MAX = 1e4;  // rect.cu:104
// cond < K  <=>  T^2 < Q*D  with T = A+B, D = AB-C^2 > 0, Q = (K+1)^2/K, because
// cond = (T+s)/(T-s), s^2 = T^2-4D, and (1+cond)^2/cond = T^2/D grows with cond.
constexpr double QCOND = (COND_MAX + 1.0) * (COND_MAX + 1.0) / COND_MAX;

enum Mode : int32_t {
    MODE_RECT_FAST = 0,     // regular grid: dt0 == dtau == a*TAtom, no wrap-around
    MODE_RECT_GENERIC = 1,  // any rectangular window range, exact uint32 semantics
    MODE_EXP_DIRECT = 2,    // exponential window, direct weighted sum (any grid)
    MODE_EXP_FAST = 3,      // exponential window on the atom grid: recurrence in t0
    MODE_COUNT = 4
};

constexpr int RB_DEFAULT[MODE_COUNT] = {128, 16, 16, 128};
constexpr int64_t SMALL_BATCH_ITEMS = 512;  // rect items aimed at when rows are not tuned
constexpr int EXP_SMALL_BATCH_ROWS = 64;    // exp rows when untuned and under 256 column blocks

struct TmplDev {
    int64_t seg_begin, seg_end;  // detector segments of this template
    int64_t merged_off;          // floats, atom table [bins][8]      (exp modes)
    int64_t prefix_off;          // doubles, prefix entries [N+1][7]  (rect modes)
    int64_t F_off;               // floats, map [N_t0][N_tau]
    int64_t item_begin;          // first work item of this template
    int64_t rowpart_off;         // float, [n_colblk][N_t0], relative to the item max
    int64_t colpart_off;         // float, [n_rowblk][N_tau], relative to the item max
    uint32_t N, TAtom, t_min;    // binned atoms, atom length, t0_data
    uint32_t type, t0, dt0, tau, dtau;  // effective window (after NONE -> RECT)
    uint32_t N_t0, N_tau;
    uint32_t n_rowblk, n_colblk;
    int32_t mode;
    int32_t rb_rows;             // rows per work item
    int32_t q0, q00, a;          // MODE_RECT_FAST index maths; a also MODE_EXP_FAST
    int32_t k00, L0, Ls, H;      // MODE_EXP_FAST: first start atom, window lengths
                                 // L(n) = L0 + Ls*n atoms, atoms [0, H) reachable
    int32_t n_seg;               // MODE_EXP_FAST: row chunks (>= n_rowblk)
    int32_t init_empty;          // MODE_EXP_FAST: the top chunk's direct sum is empty
    int64_t etab_off;            // MODE_EXP_FAST: doubles, [n_seg + INIT_PARTS][7][N_tau]
    uint32_t o_t0, o_t0Band, o_tau, o_tauBand;  // caller's window (posterior modes)
};

struct SegDev {
    int64_t begin, end;
};

// the Python binding (pyfstat_amd/_hiplib.py) mirrors these layouts
static_assert(sizeof(tcw_stats) == 104, "tcw_stats layout");
static_assert(sizeof(tcw_atoms) == 128, "tcw_atoms layout");
static_assert(sizeof(tcw_window) == 28, "tcw_window layout");

struct ItemDev {
    int32_t t, rb, cb, pad;
};

struct ItemPart {
    float maxF;
    int32_t idx;
    float M, S;  // M = the item's max F times LOG2E (float); S = sum of exp(F - M/LOG2E)
};

// ----------------------------------------------------------------------------
// stage: binned merge + FP64 prefix (rect) / float32 atom table (exp)
// ----------------------------------------------------------------------------

constexpr int SR_WG = 256;                        // 4 wave64
constexpr int SR_BINS = 4;                        // consecutive bins per thread per round
constexpr int SR_ROUND = SR_WG * SR_BINS;         // 1024 bins per round
constexpr int SR_CHUNK = 1024;                    // atoms staged in LDS per pass
constexpr int SR_PASS_LANES = 32;                 // lanes per output pass
constexpr int SR_PASS_BINS = SR_PASS_LANES * SR_BINS;  // 128 bins per wave per pass

// Rectangular-window templates: the exclusive prefix P[b] = sum of merged
// bins < b per column, [N+1][7] doubles.  One workgroup per template, rounds
// of 1024 bins, 4 consecutive bins per thread.
//  - merge: per detector, the workgroup copies the round's atoms into LDS in
//    chunks of 1024 with coalesced loads (a cursor per detector carries the
//    first atom of the next round); each thread finds its first bin's atoms
//    by binary search in LDS and sums its bins.  A bin's merged value is the
//    float32 sum over detectors X = 0,1,... (in that order) of the detector's
//    atoms in the bin (in atom order), from 0 -- the additions of
//    lalpulsar's binned merge, in its order, so merged atoms are
//    bit-identical;
//  - scan (FP64): thread totals, a Kogge-Stone scan within the wave, the 4
//    wave totals and the carry of earlier rounds in LDS.  A prefix value
//    carries ~20 roundings of 2^-53 relative to the largest partial sum;
//  - output: the LDS buffer is reused to transpose the prefix values, 32 lanes
//    of each wave per pass, so every global store is a contiguous run of the
//    [N+1][7] table (a thread's own 4 bins span 224 bytes).
__global__ __launch_bounds__(SR_WG) void stage_kernel(
    const TmplDev *__restrict__ tds, const SegDev *__restrict__ segs,
    const uint32_t *__restrict__ ts, const float *__restrict__ cols, int64_t cstride,
    double *__restrict__ prefix, int *__restrict__ flags) {
    // one block per template, before any map kernel: clear its single-atom flag
    if (threadIdx.x == 0) flags[blockIdx.x] = 0;
    const TmplDev td = tds[blockIdx.x];
    if (td.mode != MODE_RECT_FAST && td.mode != MODE_RECT_GENERIC) return;  // block-uniform
    // staging: ts [SR_CHUNK] uint32, then NC float planes [SR_CHUNK];
    // output: per wave SR_PASS_BINS x NC doubles
    __shared__ double s_buf[SR_CHUNK * (NC + 1) / 2];
    static_assert(4 * SR_PASS_BINS * NC <= SR_CHUNK * (NC + 1) / 2, "output buffer");
    uint32_t *s_ts = reinterpret_cast<uint32_t *>(s_buf);
    float *s_x = reinterpret_cast<float *>(s_buf) + SR_CHUNK;
    __shared__ double s_off[NC][SR_WG / 64];  // wave totals, then wave offsets
    __shared__ double s_carry[NC];            // sum of all earlier rounds
    __shared__ int64_t s_ka;

    const uint32_t N = td.N;
    const uint64_t TA = td.TAtom;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int nseg = (int)(td.seg_end - td.seg_begin);
    double *__restrict__ P = prefix + td.prefix_off;
    // thread X < nseg keeps detector X's cursor: its first atom not yet merged
    int64_t cur = tid < nseg ? segs[td.seg_begin + tid].begin : 0;
    if (tid < NC) s_carry[tid] = 0.0;

    for (uint32_t base = 0; base < N; base += SR_ROUND) {
        const uint32_t bend = min(N, base + (uint32_t)SR_ROUND);
        const uint32_t b0 = base + (uint32_t)tid * SR_BINS;
        const uint64_t t_lo = (uint64_t)td.t_min + (uint64_t)b0 * TA;
        const uint64_t tb_hi = (uint64_t)td.t_min + (uint64_t)bend * TA;  // later rounds' atoms
        float v[SR_BINS][NC];
#pragma unroll
        for (int j = 0; j < SR_BINS; j++)
#pragma unroll
            for (int c = 0; c < NC; c++) v[j][c] = 0.0f;
        for (int X = 0; X < nseg; X++) {
            const int64_t seg_end = segs[td.seg_begin + X].end;
            if (tid == X) s_ka = cur;
            __syncthreads();
            int64_t c0 = s_ka;
            __syncthreads();  // s_ka is rewritten for the next detector
            int64_t next = seg_end;
            while (c0 < seg_end) {  // block-uniform
                const int na = (int)min((int64_t)SR_CHUNK, seg_end - c0);
                for (int i = tid; i < na; i += SR_WG) {
                    s_ts[i] = ts[c0 + i];
#pragma unroll
                    for (int c = 0; c < NC; c++) s_x[c * SR_CHUNK + i] = cols[c * cstride + c0 + i];
                }
                __syncthreads();
                if (b0 < bend) {
                    int lo = 0, hi = na;  // first staged atom with ts >= t_lo
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if ((uint64_t)s_ts[mid] < t_lo)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    int k = lo;
#pragma unroll
                    for (int j = 0; j < SR_BINS; j++) {
                        const uint64_t t_hi = t_lo + (uint64_t)(j + 1) * TA;
                        for (; k < na && (uint64_t)s_ts[k] < t_hi; k++) {
#pragma unroll
                            for (int c = 0; c < NC; c++) v[j][c] += s_x[c * SR_CHUNK + k];
                        }
                    }
                }
                int lo = 0, hi = na;  // first staged atom of a later round (same on all threads)
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if ((uint64_t)s_ts[mid] < tb_hi)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                __syncthreads();  // staged atoms consumed before the next chunk
                if (lo < na) {
                    next = c0 + lo;
                    break;
                }
                c0 += na;
            }
            if (tid == X) cur = next;
        }
        // inclusive scan of the thread totals within the wave (bins >= N hold 0)
        double x[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < SR_BINS; j++) t += (double)v[j][c];
            x[c] = t;
        }
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const double y = __shfl_up(x[c], d);
                if (lane >= d) x[c] += y;
            }
        }
        if (lane == 63) {
#pragma unroll
            for (int c = 0; c < NC; c++) s_off[c][wave] = x[c];
        }
        __syncthreads();
        if (tid < NC) {  // exclusive scan of the wave totals, seeded by the carry
            double run = s_carry[tid];
            for (int w = 0; w < SR_WG / 64; w++) {
                const double t = s_off[tid][w];
                s_off[tid][w] = run;
                run += t;
            }
            s_carry[tid] = run;
        }
        __syncthreads();
        double ex[NC];  // exclusive prefix at the thread's first bin
#pragma unroll
        for (int c = 0; c < NC; c++) {
            double e = __shfl_up(x[c], 1);
            if (lane == 0) e = 0.0;
            ex[c] = s_off[c][wave] + e;
        }
        double *out = s_buf + wave * (SR_PASS_BINS * NC);
#pragma unroll 1
        for (int pass = 0; pass < 64 / SR_PASS_LANES; pass++) {
            if (lane / SR_PASS_LANES == pass && b0 < bend) {
                const int q = lane % SR_PASS_LANES;
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    double r = ex[c];
#pragma unroll
                    for (int j = 0; j < SR_BINS; j++) {
                        out[(q * SR_BINS + j) * NC + c] = r;
                        r += (double)v[j][c];
                    }
                }
            }
            __syncthreads();
            const int64_t wb = (int64_t)base + wave * (64 * SR_BINS) + pass * SR_PASS_BINS;
            const int cnt = (int)max((int64_t)0, min((int64_t)SR_PASS_BINS, (int64_t)bend - wb));
            for (int i = lane; i < cnt * NC; i += 64) P[wb * NC + i] = out[i];
            __syncthreads();  // copied out before the next pass (or round) rewrites
        }
    }
    if (tid < NC) P[(int64_t)N * NC + tid] = s_carry[tid];
}

// Bins of a template's atom table: the recurrence path reads none at or
// past H (the first atom no window reaches), the direct path any of the N.
__host__ __device__ __forceinline__ uint32_t exp_table_bins(const TmplDev &td) {
    return td.mode == MODE_EXP_FAST ? (uint32_t)td.H : td.N;
}

// Exponential-window templates: the merged atoms as the float32 table
// [exp_table_bins][8] (7 columns + padding), the same float32 detector-order
// sums as stage_kernel (the kernels that read it widen them to FP64, exactly).
// A workgroup owns 256 * SXB consecutive bins of one template: per detector,
// a block-wide search bounds the block's atoms, which are copied
// into LDS in chunks with coalesced loads; each thread then sums its SXB bins
// from LDS (atoms in order, detectors in order).  Grid: (bin blocks, exp
// templates), y-strided.  SXB = 4 (1024-bin blocks) for batches, 1 (256-bin
// blocks) when 1024-bin blocks would give fewer than SX_NARROW_BELOW
// workgroups (one large map would otherwise occupy a handful of CUs).
constexpr int SX_NARROW_BELOW = 256;
constexpr int SX_CHUNK = 1024;  // atoms staged per pass
// the finished rows are restaged in LDS (row stride SX_ROW floats, odd so the
// lanes' writes spread over the banks) and written to the table with
// consecutive lanes on consecutive floats
constexpr int SX_ROW = MERGED_STRIDE + 1;

// For sorted ts[lo0, hi0): the first index with ts >= key0 into s_r[0][0]
// and with ts >= key1 into s_r[1][0] (hi0 if none).  The two halves of a
// 256-thread block search one key each, 128-way: each round samples 128
// evenly spaced timestamps of the remaining range (one load per thread, all
// in flight together) and keeps the gap where the key falls, so a range of
// n atoms takes about log_128(n) rounds of loads instead of log_2(n)
// dependent ones.  Every thread of the block calls it; the results are valid
// on return.
__device__ __forceinline__ void block_lower_bound2(const uint32_t *__restrict__ ts, int64_t lo0,
                                                   int64_t hi0, uint64_t key0, uint64_t key1,
                                                   int64_t (&s_r)[2][2], int (&s_c)[4]) {
    const int tid = threadIdx.x, half = tid >> 7, t = tid & 127, wave = tid >> 6;
    const uint64_t key = half ? key1 : key0;
    if (tid < 2) {
        s_r[tid][0] = lo0;
        s_r[tid][1] = hi0;
    }
    __syncthreads();
    // the answer lies in [lo, hi]; n = hi - lo timestamps are left to test
    while (s_r[0][1] > s_r[0][0] || s_r[1][1] > s_r[1][0]) {  // block-uniform
        const int64_t lo = s_r[half][0], hi = s_r[half][1];
        const int64_t n = hi - lo;
        const int64_t sd = n <= 128 ? 1 : (n + 127) / 128;
        const bool pred = n > 0 && t * sd < n && (uint64_t)ts[lo + t * sd] < key;
        const uint64_t b = __ballot(pred);
        if ((tid & 63) == 0) s_c[wave] = __popcll(b);
        __syncthreads();
        if (t == 0 && n > 0) {
            // c samples lie below the key: the answer is in (sample c-1, sample c]
            const int c = s_c[2 * half] + s_c[2 * half + 1];
            if (c == 0) {
                s_r[half][1] = lo;
            } else {
                s_r[half][0] = lo + (int64_t)(c - 1) * sd + 1;
                s_r[half][1] = min(lo + (int64_t)c * sd, hi);
            }
        }
        __syncthreads();
    }
}

template <int SXB>
__global__ __launch_bounds__(256) void stage_exp_kernel(
    const TmplDev *__restrict__ tds, const int32_t *__restrict__ exp_tmpls, int64_t n_exp,
    const SegDev *__restrict__ segs, const uint32_t *__restrict__ ts,
    const float *__restrict__ cols, int64_t cstride, float *__restrict__ merged,
    int *__restrict__ flags) {
    // atom staging during the merge, the finished rows afterwards
    constexpr int SX_RAW = (NC + 1) * SX_CHUNK > 256 * SXB * SX_ROW ? (NC + 1) * SX_CHUNK
                                                                     : 256 * SXB * SX_ROW;
    __shared__ float s_raw[SX_RAW];
    uint32_t *s_ts = reinterpret_cast<uint32_t *>(s_raw);
    float(*s_x)[SX_CHUNK] = reinterpret_cast<float(*)[SX_CHUNK]>(s_raw + SX_CHUNK);
    __shared__ int64_t s_rng[2][2];
    __shared__ int s_cnt[4];
    const int tid = threadIdx.x;
    for (int64_t e = blockIdx.y; e < n_exp; e += gridDim.y) {
        // before any map kernel: clear the template's single-atom flag
        if (blockIdx.x == 0 && tid == 0) flags[exp_tmpls[e]] = 0;
        const TmplDev td = tds[exp_tmpls[e]];
        constexpr int SX_BINS = SXB;
        constexpr int SX_BINS_BLOCK = 256 * SXB;
        const uint32_t B0 = blockIdx.x * (uint32_t)SX_BINS_BLOCK;
        const uint32_t NB = exp_table_bins(td);
        if (B0 >= NB) continue;  // block-uniform
        const uint32_t B1 = min(NB, B0 + (uint32_t)SX_BINS_BLOCK);
        const uint64_t TA = td.TAtom;
        const uint64_t tb_lo = (uint64_t)td.t_min + (uint64_t)B0 * TA;
        const uint64_t tb_hi = (uint64_t)td.t_min + (uint64_t)B1 * TA;
        const uint32_t b0 = B0 + (uint32_t)tid * SX_BINS;  // this thread's first bin
        float v[SX_BINS][NC];
#pragma unroll
        for (int j = 0; j < SX_BINS; j++)
#pragma unroll
            for (int c = 0; c < NC; c++) v[j][c] = 0.0f;
        for (int64_t X = td.seg_begin; X < td.seg_end; X++) {
            const SegDev sg = segs[X];
            // first atom with ts >= tb_lo, resp. >= tb_hi
            block_lower_bound2(ts, sg.begin, sg.end, tb_lo, tb_hi, s_rng, s_cnt);
            const int64_t ka = s_rng[0][0], kb = s_rng[1][0];
            __syncthreads();  // s_rng is rewritten by the next detector
            for (int64_t c0 = ka; c0 < kb; c0 += SX_CHUNK) {
                const int na = (int)min((int64_t)SX_CHUNK, kb - c0);
                for (int i = tid; i < na; i += 256) {
                    s_ts[i] = ts[c0 + i];
#pragma unroll
                    for (int c = 0; c < NC; c++) s_x[c][i] = cols[c * cstride + c0 + i];
                }
                __syncthreads();
                if (b0 < B1) {
                    const uint64_t t_lo = (uint64_t)td.t_min + (uint64_t)b0 * TA;
                    int lo = 0, hi = na;  // first staged atom with ts >= t_lo
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if ((uint64_t)s_ts[mid] < t_lo)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    int k = lo;
#pragma unroll
                    for (int j = 0; j < SX_BINS; j++) {
                        const uint64_t t_hi = t_lo + (uint64_t)(j + 1) * TA;
                        for (; k < na && (uint64_t)s_ts[k] < t_hi; k++) {
#pragma unroll
                            for (int c = 0; c < NC; c++) v[j][c] += s_x[c][k];
                        }
                    }
                }
                __syncthreads();
            }
        }
        // the last staging pass ended with a barrier: s_raw is free
#pragma unroll
        for (int j = 0; j < SX_BINS; j++) {
            float *r = s_raw + (tid * SX_BINS + j) * SX_ROW;
#pragma unroll
            for (int c = 0; c < NC; c++) r[c] = v[j][c];
            r[NC] = 0.0f;
        }
        __syncthreads();
        float *base = merged + td.merged_off + (int64_t)B0 * MERGED_STRIDE;
        const int nq = (int)(B1 - B0) * MERGED_STRIDE;
        for (int q = tid; q < nq; q += 256)
            base[q] = s_raw[(q / MERGED_STRIDE) * SX_ROW + q % MERGED_STRIDE];
        __syncthreads();  // s_raw is staged again for the next template
    }
}

// ----------------------------------------------------------------------------
// per-cell F
// ----------------------------------------------------------------------------

// F from window sums: determinant, numerator and condition guard in FP64
// (rect.cu:98-118 / exp.cu:109-129; F = 2 when the guard fails); the final
// quotient in float32 (the stored map is float32).
__device__ __forceinline__ float cell_F(double A, double B, double C, double far,
                                        double fai, double fbr, double fbi) {
    const double D = fma(A, B, -(C * C));
    const double T = A + B;
    const double p = fma(far, far, fai * fai);
    const double q = fma(fbr, fbr, fbi * fbi);
    const double r = fma(far, fbr, fai * fbi);
    const double num = fma(B, p, fma(A, q, (-2.0 * C) * r));
    // the map is float32: the quotient of the FP64 numerator and determinant
    // (each rounded to float) through v_rcp_f32 is within ~3e-7 relative, and
    // v_rcp_f64 issues at ~1/3.5 the rate of an FP64 FMA.  The condition test
    // is FP64 (two multiplies and a compare: one conversion fewer than on
    // float T and D).  A, B >= 0 (sums of squares), so T <= 0 only for an
    // empty window, where D = 0 and the test fails as well.
    const bool ok = T * T < QCOND * D;
    const float F = (float)num * __builtin_amdgcn_rcpf((float)D);
    return ok ? F : 2.0f;
}

// uint32 index maths of rect.cu:21-31 (start) and rect.cu:59-69 (end).
__device__ __forceinline__ uint32_t idx_start(uint32_t t, uint32_t t_min, uint32_t TA,
                                              uint32_t N) {
    int32_t i = (int32_t)((uint32_t)(t - t_min + TA / 2) / TA);
    if (i < 0) i = 0;
    uint32_t u = (uint32_t)i;
    return u >= N ? N - 1 : u;
}

__device__ __forceinline__ uint32_t idx_end_q(uint32_t q, uint32_t N) {
    int32_t i = (int32_t)(q - 1u);
    if (i < 0) i = 0;
    uint32_t u = (uint32_t)i;
    return u >= N ? N - 1 : u;
}

// ----------------------------------------------------------------------------
// fused reductions
// ----------------------------------------------------------------------------

constexpr float LOG2E = 1.4426950408889634f;

// 2^(aL - bL) = exp(a - b) for references kept in log2 units (aL = a*LOG2E);
// exp2(-inf) = 0
__device__ __forceinline__ float exp2_diff(float aL, float bL) {
    return __builtin_amdgcn_exp2f(aL - bL);
}


// Cross-lane moves without LDS: DPP within 16-lane rows (quad permutes and
// row rotations; every source lane lies inside the row, so all lanes are
// written) and gfx950's row swaps across rows.
constexpr int DPP_QUAD_XOR1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_XOR2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int DPP_ROW_ROR4 = 0x124;  // row_ror:4
constexpr int DPP_ROW_ROR8 = 0x128;  // row_ror:8

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Combine over the four 16-lane rows at each row position: v_permlane16_swap
// pairs rows (0,1) and (2,3), v_permlane32_swap the two halves; every lane of
// a position ends with the same bits.
template <typename Op>
__device__ __forceinline__ float rows16_reduce(float v, Op op) {
    const unsigned u = __float_as_uint(v);
    const auto p = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    const float s = op(__uint_as_float(p[0]), __uint_as_float(p[1]));
    const unsigned us = __float_as_uint(s);
    const auto q = __builtin_amdgcn_permlane32_swap(us, us, false, false);
    return op(__uint_as_float(q[0]), __uint_as_float(q[1]));
}

__device__ __forceinline__ float rows16_sum(float v) {
    return rows16_reduce(v, [](float a, float b) { return a + b; });
}

// Sum over the 4 quads of a 16-lane row at each quad position.
__device__ __forceinline__ float quads_sum(float v) {
    v += dpp_f<DPP_ROW_ROR4>(v);
    return v + dpp_f<DPP_ROW_ROR8>(v);
}

__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_f<DPP_QUAD_XOR1>(v);
    v += dpp_f<DPP_QUAD_XOR2>(v);
    return rows16_sum(quads_sum(v));
}

// wave maximum, the same value on every lane
__device__ __forceinline__ float wave_max(float v) {
    v = fmaxf(v, dpp_f<DPP_QUAD_XOR1>(v));
    v = fmaxf(v, dpp_f<DPP_QUAD_XOR2>(v));
    v = fmaxf(v, dpp_f<DPP_ROW_ROR4>(v));
    v = fmaxf(v, dpp_f<DPP_ROW_ROR8>(v));
    return rows16_reduce(v, [](float a, float b) { return fmaxf(a, b); });
}

__device__ __forceinline__ void argmax_merge(float &bv, int &bi, float v, int i) {
    if (v > bv || (v == bv && i < bi)) {
        bv = v;
        bi = i;
    }
}

__device__ __forceinline__ void wave_argmax(float &bv, int &bi) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        float ov = __shfl_xor(bv, o);
        int oi = __shfl_xor(bi, o);
        argmax_merge(bv, bi, ov, oi);
    }
}

constexpr int MAX_RB = 512;
constexpr float REF_SLACK = 60.0f;  // e^60 * cells per item stays far below FLT_MAX

// Wave sums of R per-lane values.  Returns the sum of one of them, and in
// `row` which one (-1: this lane holds none).  R = 1: lane 0 holds the sum.
// R = 4: two halving exchange steps within each quad first (lanes pair up and
// keep half of the rows each), then one value per lane is summed over the
// quads (row rotations) and the rows (row swaps); row r ends up on lane
// 2*(r&1) + (r>>1).
template <int R>
__device__ __forceinline__ float rows_sum(const float (&v)[R], int lane, int &row);

template <>
__device__ __forceinline__ float rows_sum<1>(const float (&v)[1], int lane, int &row) {
    row = lane == 0 ? 0 : -1;
    return wave_sum(v[0]);
}

template <>
__device__ __forceinline__ float rows_sum<4>(const float (&v)[4], int lane, int &row) {
    const bool b = lane & 1;
    // step 1: b = 0 keeps rows {0,1}, b = 1 keeps rows {2,3}
    const float s0 = b ? v[0] : v[2];
    const float s1 = b ? v[1] : v[3];
    const float u0 = (b ? v[2] : v[0]) + dpp_f<DPP_QUAD_XOR1>(s0);
    const float u1 = (b ? v[3] : v[1]) + dpp_f<DPP_QUAD_XOR1>(s1);
    // step 2: c = 0 keeps the first of the pair, c = 1 the second
    const bool c = lane & 2;
    const float z = (c ? u1 : u0) + dpp_f<DPP_QUAD_XOR2>(c ? u0 : u1);
    row = lane < 4 ? 2 * (b ? 1 : 0) + (c ? 1 : 0) : -1;
    return rows16_sum(quads_sum(z));
}

// Per-work-item reduction state.  A lane owns 4 columns; every sub-tile brings
// R wave-uniform rows x those 4 columns, with invalid cells set to -inf.
//  - argmax: lane maximum of the sub-tile; only when it beats the lane's best
//    is the first (row-major) position searched;
//  - BtSG: sum of 2^(F*LOG2E - ML), ML = Mw*LOG2E rounded to float, Mw the
//    wave's running maximum (accumulators rescaled when it grows) -- one FMA
//    and one exp2 per cell; each sum is relative to its reference in log2
//    units (ML), i.e. to exp(F - ML/LOG2E).  Column sums stay in registers
//    across sub-tiles (their sum is the item total), row sums are
//    wave-reduced and parked in LDS with ML.
template <bool BTSG, int NCOLS = 4>
struct Reducer {
    float best;
    int bidx;
    // a lower bound of the wave's largest value so far (the wave maximum of
    // best at the last search): a lane whose block maximum is below it holds
    // no candidate for the item's first maximum and skips the search
    float wbest;
    float Mw;
    float ML;
    float Scol[NCOLS];

    __device__ void init() {
        best = -INFINITY;
        bidx = 0x7fffffff;
        wbest = -INFINITY;
        Mw = -INFINITY;
        ML = -INFINITY;
#pragma unroll
        for (int c = 0; c < NCOLS; c++) Scol[c] = 0.0f;
    }

    // the lane's sum over all its cells (relative to ML)
    __device__ float total() const {
        float s = Scol[0];
#pragma unroll
        for (int c = 1; c < NCOLS; c++) s += Scol[c];
        return s;
    }

    // F[r][c]: cell values (-inf when invalid); row index m0 + r; colidx[c]
    // column index of the lane's c-th column; rl0: item-local row of r = 0.
    template <int R>
    __device__ void add(const float (&F)[R][NCOLS], int m0, const int (&colidx)[NCOLS], int N_tau,
                        int lane, int rl0, float *sh_rowM, float *sh_rowS) {
        float lm = -INFINITY;
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int c = 0; c < NCOLS; c++) lm = fmaxf(lm, F[r][c]);
        // a block holding the item's largest value V always searches (V >=
        // wbest, V >= best); a block below wbest has a larger value in
        // another lane of the wave.  Only for 16-cell blocks: for 4 cells the
        // wave maximum that updates wbest costs more than the search
        constexpr bool FILTER = R * NCOLS >= 16;
        const bool search = (!FILTER || lm >= wbest) &&
                            (lm > best || (lm == best && lm != -INFINITY));
        if (search) {
            // the first cell (row-major) equal to lm: selects from the last
            // cell to the first, no branches (a search with an early exit
            // compiles to a chain of masked branches, one per cell)
            int pos = 0;
#pragma unroll
            for (int q = R * NCOLS - 1; q >= 0; q--) pos = F[q / NCOLS][q % NCOLS] == lm ? q : pos;
            const int rr = pos / NCOLS, cc = pos % NCOLS;
            // a lane's columns are evenly spaced (1 or 64 apart): arithmetic,
            // not a selection the compiler would turn into a private array
            int col = colidx[0];
            if constexpr (NCOLS > 1) col += cc * (colidx[1] - colidx[0]);
            argmax_merge(best, bidx, lm, (m0 + rr) * N_tau + col);
        }
        if (FILTER && __any(search)) wbest = wave_max(best);
        if (!BTSG) return;
        // exp(F - Mw) only has to stay finite, so the wave reference Mw is
        // raised (wave maximum, accumulators rescaled) only when some lane
        // sees a value more than REF_SLACK above it
        if (__any(lm > Mw + REF_SLACK)) {
            const float wm = wave_max(lm);
            const float wmL = wm * LOG2E;
            const float sc = exp2_diff(ML, wmL);
#pragma unroll
            for (int c = 0; c < NCOLS; c++) Scol[c] *= sc;
            Mw = wm;
            ML = wmL;
        }
        if (Mw == -INFINITY) return;  // wave-uniform: no valid cell yet
        float sr[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            sr[r] = 0.0f;
#pragma unroll
            for (int c = 0; c < NCOLS; c++) {
                const float e = __builtin_amdgcn_exp2f(fmaf(F[r][c], LOG2E, -ML));
                sr[r] = c == 0 ? e : sr[r] + e;
                Scol[c] += e;
            }
        }
        int row;
        const float srw = rows_sum<R>(sr, lane, row);
        if (row >= 0) {
            sh_rowM[rl0 + row] = ML;
            sh_rowS[rl0 + row] = srw;
        }
    }
};

// End of a work item: combine the 4 waves, then write the item partial
// {max, argmax, M, S} and, for BtSG, the row and column partial sums rescaled
// to the item reference M (so each is a single float).  M is in log2 units:
// the partial sums are relative to exp(M / LOG2E).
template <bool BTSG, int NCOLS, typename ColMap>
__device__ void finish_item(Reducer<BTSG, NCOLS> &red, int wave, int lane, const TmplDev &td,
                            const ItemDev &it, int64_t item_id, int nrows_item,
                            ItemPart *ipart, float *rowpart, float *colpart,
                            const float *sh_rowM, const float *sh_rowS, ColMap colmap) {
    __shared__ float sh_best[4];
    __shared__ int sh_bidx[4];
    __shared__ float sh_M[4];
    __shared__ float sh_S[4];
    __shared__ float sh_col[4][TILE_COLS];

    wave_argmax(red.best, red.bidx);
    float stot = 0.0f;
    if (BTSG) stot = wave_sum(red.total());
    if (lane == 0) {
        sh_best[wave] = red.best;
        sh_bidx[wave] = red.bidx;
        sh_M[wave] = red.ML;
        sh_S[wave] = stot;
    }
    if (BTSG) {
#pragma unroll
        for (int c = 0; c < NCOLS; c++) sh_col[wave][colmap(lane, c)] = red.Scol[c];
    }
    __syncthreads();
    const int tid = threadIdx.x;
    const float M = fmaxf(fmaxf(sh_M[0], sh_M[1]), fmaxf(sh_M[2], sh_M[3]));
    if (BTSG) {
        const int n = it.cb * TILE_COLS + tid;
        if (n < (int)td.N_tau) {
            float S = 0.0f;
#pragma unroll
            for (int w = 0; w < 4; w++)
                if (sh_M[w] != -INFINITY) S += sh_col[w][tid] * exp2_diff(sh_M[w], M);
            colpart[td.colpart_off + (int64_t)it.rb * td.N_tau + n] = S;
        }
        for (int rl = tid; rl < nrows_item; rl += WG) {
            const int m = it.rb * td.rb_rows + rl;
            rowpart[td.rowpart_off + (int64_t)it.cb * td.N_t0 + m] =
                sh_rowS[rl] * exp2_diff(sh_rowM[rl], M);
        }
    }
    if (tid == 0) {
        float bv = sh_best[0];
        int bi = sh_bidx[0];
#pragma unroll
        for (int w = 1; w < 4; w++) argmax_merge(bv, bi, sh_best[w], sh_bidx[w]);
        float S = 0.0f;
        if (BTSG) {
#pragma unroll
            for (int w = 0; w < 4; w++)
                if (sh_M[w] != -INFINITY) S += sh_S[w] * exp2_diff(sh_M[w], M);
        }
        ipart[item_id] = ItemPart{bv, bi, M, S};
    }
}

// ----------------------------------------------------------------------------
// rectangular window, regular grid: prefix differences with diagonal reuse
// ----------------------------------------------------------------------------
//
// With dt0 = dtau = a*TAtom and no uint32 wrap-around, the window of cell
// (m, n) is atoms [i(m), j(m+n)) of the prefix table, where
//   i(m)   = min(q0 + a*m, N-1)                      (rect.cu:21-31)
//   j(D)   = min(max(q00 + a*D - 1, 0), N-1) + 1     (rect.cu:54-69), D = m+n,
// and j >= i always, so the incremental sum of rect.cu:75-91 equals
// P[j] - P[i].  The item's j(D) range is staged in LDS in groups of 4
// diagonals, each group's 7 components x 4 diagonals contiguous plus one pad
// double (LDS_GS doubles: a lane's 4-group stride of 58 words spreads its
// 8-byte reads over all banks); each lane computes a 4x4 block whose 16
// cells touch 7 diagonals -- its own group and 3 entries of the next, all at
// constant offsets from one address -- loading 4 entries for the first row
// and one more per following row.
constexpr int LDS_GS = 4 * NC + 1;
// entry (diagonal d, component k) of the staged range
__host__ __device__ __forceinline__ int lds_at(int d, int k) {
    return (d >> 2) * LDS_GS + k * 4 + (d & 3);
}
// doubles of LDS for nd diagonals
__host__ __device__ __forceinline__ int lds_doubles(int nd) { return (nd + 3) / 4 * LDS_GS; }

// Word 3 of a raw buffer resource: 32-bit data format, as gfx9 raw buffer
// accesses require.
constexpr int BUF_RSRC_WORD3 = 0x00020000;

// Work-item order.  Blocks go to the 8 XCDs round-robin (block b on XCD
// b % 8) and every XCD has its own L2, so XCD x is given a contiguous run of
// work items: the items of one template, which read the same prefix table,
// then share one L2.
#ifndef TCW_XCD_REMAP
#define TCW_XCD_REMAP 1
#endif
__device__ __forceinline__ int64_t block_slot() {
#if TCW_XCD_REMAP
    const int64_t G = gridDim.x, b = blockIdx.x;
    const int64_t q = G >> 3, r = G & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
#else
    return blockIdx.x;
#endif
}

#ifndef TCW_FAST_MIN_WAVES
#define TCW_FAST_MIN_WAVES 4
#endif

template <bool BTSG, bool STRICT>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(TCW_FAST_MIN_WAVES, 8)))
void rect_fast_kernel(
    const TmplDev *__restrict__ tds, const ItemDev *__restrict__ items, int64_t item0,
    const double *__restrict__ prefix, float *__restrict__ Fout,
    ItemPart *__restrict__ ipart, float *__restrict__ rowpart,
    float *__restrict__ colpart, int *__restrict__ flags) {
    extern __shared__ double lds[];
    __shared__ float sh_rowM[MAX_RB];
    __shared__ float sh_rowS[MAX_RB];
    const int64_t item_id = item0 + block_slot();
    const ItemDev it = items[item_id];
    const TmplDev td = tds[it.t];
    const int RB = td.rb_rows;
    const int row0 = it.rb * RB;
    const int col0 = it.cb * TILE_COLS;
    const int N_t0 = (int)td.N_t0, N_tau = (int)td.N_tau;
    const int nd = RB + TILE_COLS - 1;
    const int D0 = row0 + col0;
    const double *__restrict__ P = prefix + td.prefix_off;
    const int Nm1 = (int)td.N - 1;

    for (int d = threadIdx.x; d < nd; d += WG) {
        int64_t q = (int64_t)td.q00 + (int64_t)td.a * (D0 + d) - 1;
        q = min(max(q, (int64_t)0), (int64_t)Nm1);
        const double *__restrict__ src = P + (q + 1) * NC;
#pragma unroll
        for (int c = 0; c < NC; c++) lds[lds_at(d, c)] = src[c];
    }
    __syncthreads();

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    Reducer<BTSG> red;
    red.init();
    bool single = false;
    float *__restrict__ Ft = Fout + td.F_off;
    const int c0 = col0 + lane * 4;  // the lane's first column
    const int nvc = min(max(N_tau - c0, 0), 4);
    int colidx[4];
#pragma unroll
    for (int c = 0; c < 4; c++) colidx[c] = c0 + c;
    const bool cols_full = col0 + TILE_COLS <= N_tau;
    // partial sub-tiles store through a buffer resource over their rows
    // inside the map: the lane's column c at byte voff[c] of a row, or far
    // outside the resource for a column past N_tau (the hardware drops the
    // store; rows past the map lie beyond the resource as well), so no
    // branch per cell
    const uint32_t rowb = (uint32_t)N_tau * (uint32_t)sizeof(float);
    uint32_t voff[4];
#pragma unroll
    for (int c = 0; c < 4; c++) voff[c] = c < nvc ? (uint32_t)(c0 + c) * 4u : 0x80000000u;

    for (int s = 0; s < RB / 16; s++) {
        const int rl0 = s * 16 + wave * 4;
        const int m0 = row0 + rl0;
        if (m0 >= N_t0) break;
        const int nvr = min(N_t0 - m0, 4);
        // lane's diagonals d = rl0 + 4*lane + x, x = 0..6; rl0 + 4*lane is a
        // multiple of 4, so lds_at(d, k) = lb + lds_at(x, k)
        const double *__restrict__ lb = lds + lds_at(rl0 + lane * 4, 0);
        double w[4][NC];
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int k = 0; k < NC; k++) w[c][k] = lb[lds_at(c, k)];
        float F[4][4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if (r > 0) {
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int k = 0; k < NC; k++) w[c][k] = w[c + 1][k];
                const int x = r + 3;
#pragma unroll
                for (int k = 0; k < NC; k++) w[3][k] = lb[lds_at(x, k)];
            }
            // window start: wave-uniform, read through the scalar cache
            const int i = min(td.q0 + td.a * (m0 + r), Nm1);
            double Pi[NC];
#pragma unroll
            for (int k = 0; k < NC; k++) Pi[k] = P[(int64_t)i * NC + k];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                F[r][c] = cell_F(w[c][0] - Pi[0], w[c][1] - Pi[1], w[c][2] - Pi[2],
                                 w[c][3] - Pi[3], w[c][4] - Pi[4], w[c][5] - Pi[5],
                                 w[c][6] - Pi[6]);
                if (STRICT) {
                    const int m = m0 + r;
                    int64_t q = (int64_t)td.q00 + (int64_t)td.a * (m + colidx[c]) - 1;
                    q = min(max(q, (int64_t)0), (int64_t)Nm1);
                    single |= (q == i) && (r < nvr) && (c < nvc);
                }
            }
        }
        if (cols_full && nvr == 4) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float *Frow = Ft + (int64_t)(m0 + r) * N_tau + c0;
#pragma unroll
                for (int c = 0; c < 4; c++) Frow[c] = F[r][c];
            }
        } else {
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                Ft + (int64_t)m0 * N_tau, (short)0, (int)((uint32_t)nvr * rowb), BUF_RSRC_WORD3);
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(F[r][c]), rs,
                                                          (int)voff[c], (int)(r * rowb), 0);
                    F[r][c] = (r < nvr) && (c < nvc) ? F[r][c] : -INFINITY;
                }
        }
        red.template add<4>(F, m0, colidx, N_tau, lane, rl0, sh_rowM, sh_rowS);
    }
    if (STRICT && single) atomicOr(flags + it.t, 1);
    __syncthreads();
    finish_item(red, wave, lane, td, it, item_id, min(RB, N_t0 - row0), ipart, rowpart,
                      colpart, sh_rowM, sh_rowS, [](int ln, int c) { return ln * 4 + c; });
}

// ----------------------------------------------------------------------------
// generic per-cell kernels (rect with exact wrap-around semantics; exp)
// ----------------------------------------------------------------------------
//
// Lanes own columns col0 + c*64 + lane (c = 0..3); wave w owns item rows
// w, w+4, w+8, ...  Row quantities are wave-uniform.

struct RowRect {
    uint32_t i0;       // window start index
    uint32_t x0;       // (t0_m + tau - t_min + TAtom/2) mod 2^32 at n = 0
    uint32_t nwrap;    // first n whose x wrapped past 2^32 (N_tau if none)
    uint32_t i1_pre;   // i_t1 just before the wrap (running maximum after it)
};

__device__ __forceinline__ RowRect rect_row(const TmplDev &td, uint32_t m) {
    RowRect rr;
    const uint32_t t0m = td.t0 + m * td.dt0;
    rr.i0 = idx_start(t0m, td.t_min, td.TAtom, td.N);
    rr.x0 = t0m + td.tau - td.t_min + td.TAtom / 2;
    const uint64_t room = (1ull << 32) - (uint64_t)rr.x0;  // n*dtau >= room wraps
    const uint64_t nw = (room + td.dtau - 1) / td.dtau;
    rr.nwrap = nw < (uint64_t)td.N_tau ? (uint32_t)nw : td.N_tau;
    rr.i1_pre = 0;
    if (rr.nwrap < td.N_tau) {
        const uint32_t xp = rr.x0 + (rr.nwrap - 1) * td.dtau;
        rr.i1_pre = idx_end_q(xp / td.TAtom, td.N);
    }
    return rr;
}

template <bool BTSG, bool STRICT>
__global__ __launch_bounds__(WG) void rect_generic_kernel(
    const TmplDev *__restrict__ tds, const ItemDev *__restrict__ items, int64_t item0,
    const double *__restrict__ prefix, float *__restrict__ Fout,
    ItemPart *__restrict__ ipart, float *__restrict__ rowpart,
    float *__restrict__ colpart, int *__restrict__ flags) {
    __shared__ float sh_rowM[MAX_RB];
    __shared__ float sh_rowS[MAX_RB];
    const int64_t item_id = item0 + block_slot();
    const ItemDev it = items[item_id];
    const TmplDev td = tds[it.t];
    const int RB = td.rb_rows;
    const int row0 = it.rb * RB;
    const int col0 = it.cb * TILE_COLS;
    const int N_t0 = (int)td.N_t0, N_tau = (int)td.N_tau;
    const double *__restrict__ P = prefix + td.prefix_off;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    Reducer<BTSG> red;
    red.init();
    bool single = false;
    float *__restrict__ Ft = Fout + td.F_off;
    int colidx[4];
#pragma unroll
    for (int c = 0; c < 4; c++) colidx[c] = col0 + c * 64 + lane;

    for (int rl = wave; rl < RB; rl += 4) {
        const int m = row0 + rl;
        if (m >= N_t0) break;
        const RowRect rr = rect_row(td, (uint32_t)m);
        double Pi[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) Pi[k] = P[(int64_t)rr.i0 * NC + k];
        float F[1][4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const bool valid = colidx[c] < N_tau;
            const uint32_t n = valid ? (uint32_t)colidx[c] : 0u;
            const uint32_t x = rr.x0 + n * td.dtau;
            const uint32_t i1 = idx_end_q(x / td.TAtom, td.N);
            // running maximum of i_t1 over n' <= n (the i_t1_last recursion)
            const uint32_t run = (n >= rr.nwrap) ? max(i1, rr.i1_pre) : i1;
            const uint32_t j = max(rr.i0, run + 1);
            const double *Pj = P + (int64_t)j * NC;
            const float f = cell_F(Pj[0] - Pi[0], Pj[1] - Pi[1], Pj[2] - Pi[2], Pj[3] - Pi[3],
                                   Pj[4] - Pi[4], Pj[5] - Pi[5], Pj[6] - Pi[6]);
            if (STRICT) single |= valid && (i1 == rr.i0);
            if (valid) Ft[(int64_t)m * N_tau + colidx[c]] = f;
            F[0][c] = valid ? f : -INFINITY;
        }
        red.template add<1>(F, m, colidx, N_tau, lane, rl, sh_rowM, sh_rowS);
    }
    if (STRICT && single) atomicOr(flags + it.t, 1);
    __syncthreads();
    finish_item(red, wave, lane, td, it, item_id, min(RB, N_t0 - row0), ipart, rowpart,
                      colpart, sh_rowM, sh_rowS, [](int ln, int c) { return c * 64 + ln; });
}

// Exponential window, direct weighted sum over the truncated support
// (exp.cu:39-102): atoms i in [i_t0, i_t1] with t0 <= t_i <= t1 = t0 + 3 tau,
// weight w_i = exp(-(t_i - t0)/tau), A/B/C weighted by w^2, Fa/Fb by w.  All
// lanes of a wave share the row (t0) and therefore the atom stream, which is
// read once per wave through uniform loads; each lane carries its own tau.
// The weights are advanced by rho = exp(-TAtom/tau) per atom (FP64, relative
// error ~ k * 1e-16 after k atoms), the sums are FP64.
template <bool BTSG, bool STRICT>
__global__ __launch_bounds__(WG) void exp_direct_kernel(
    const TmplDev *__restrict__ tds, const ItemDev *__restrict__ items, int64_t item0,
    const float *__restrict__ merged, float *__restrict__ Fout,
    ItemPart *__restrict__ ipart, float *__restrict__ rowpart,
    float *__restrict__ colpart, int *__restrict__ flags) {
    __shared__ float sh_rowM[MAX_RB];
    __shared__ float sh_rowS[MAX_RB];
    const int64_t item_id = item0 + block_slot();
    const ItemDev it = items[item_id];
    const TmplDev td = tds[it.t];
    const int RB = td.rb_rows;
    const int row0 = it.rb * RB;
    const int col0 = it.cb * TILE_COLS;
    const int N_t0 = (int)td.N_t0, N_tau = (int)td.N_tau;
    const float *__restrict__ X = merged + td.merged_off;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    Reducer<BTSG> red;
    red.init();
    bool single = false;
    float *__restrict__ Ft = Fout + td.F_off;
    const uint32_t TA = td.TAtom;
    int colidx[4];
#pragma unroll
    for (int c = 0; c < 4; c++) colidx[c] = col0 + c * 64 + lane;

    for (int rl = wave; rl < RB; rl += 4) {
        const int m = row0 + rl;
        if (m >= N_t0) break;
        const uint32_t t0m = td.t0 + (uint32_t)m * td.dt0;
        const uint32_t i0 = idx_start(t0m, td.t_min, TA, td.N);
        // first atom with t_i >= t0m (timestamps t_min + i*TA never wrap)
        uint32_t ilo = 0;
        if (t0m > td.t_min) ilo = (uint32_t)(((uint64_t)(t0m - td.t_min) + TA - 1) / TA);
        const uint32_t lo = max(i0, ilo);
        float F[1][4];
        for (int c = 0; c < 4; c++) {
            const bool valid = colidx[c] < N_tau;
            const uint32_t n = valid ? (uint32_t)colidx[c] : 0u;
            const uint32_t tau = td.tau + n * td.dtau;
            const uint32_t t1 = t0m + 3u * tau;
            const uint32_t i1 = idx_end_q((uint32_t)(t1 - td.t_min + TA / 2) / TA, td.N);
            if (STRICT) single |= valid && (i1 == i0);
            // last atom with t_i <= t1
            int64_t ihi = -1;
            if (t1 >= td.t_min) ihi = (int64_t)((t1 - td.t_min) / TA);
            const int64_t hi = min((int64_t)i1, ihi);
            double acc[NC];
#pragma unroll
            for (int k = 0; k < NC; k++) acc[k] = 0.0;
            const double rtau = 1.0 / (double)tau;
            const double rho = exp(-(double)TA * rtau);
            double w = exp(-(double)(td.t_min + lo * TA - t0m) * rtau);
            // wave-uniform loop bounds: the longest lane decides
            int64_t hi_w = hi;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1)
                hi_w = max(hi_w, (int64_t)__shfl_xor((long long)hi_w, o));
            for (int64_t i = lo; i <= hi_w; i++) {
                const double wi = (i <= hi) ? w : 0.0;
                const double w2 = wi * wi;
                const float *xi = X + i * MERGED_STRIDE;
                acc[0] = fma(xi[0], w2, acc[0]);
                acc[1] = fma(xi[1], w2, acc[1]);
                acc[2] = fma(xi[2], w2, acc[2]);
                acc[3] = fma(xi[3], wi, acc[3]);
                acc[4] = fma(xi[4], wi, acc[4]);
                acc[5] = fma(xi[5], wi, acc[5]);
                acc[6] = fma(xi[6], wi, acc[6]);
                w *= rho;
            }
            const float f = cell_F(acc[0], acc[1], acc[2], acc[3], acc[4], acc[5], acc[6]);
            if (valid) Ft[(int64_t)m * N_tau + colidx[c]] = f;
            F[0][c] = valid ? f : -INFINITY;
        }
        red.template add<1>(F, m, colidx, N_tau, lane, rl, sh_rowM, sh_rowS);
    }
    if (STRICT && single) atomicOr(flags + it.t, 1);
    __syncthreads();
    finish_item(red, wave, lane, td, it, item_id, min(RB, N_t0 - row0), ipart, rowpart,
                      colpart, sh_rowM, sh_rowS, [](int ln, int c) { return c * 64 + ln; });
}

// ----------------------------------------------------------------------------
// exponential window on the atom grid: sliding recurrence along t0
// ----------------------------------------------------------------------------
//
// When t0 - t0_data, dt0, tau and dtau are whole multiples of TAtom (and
// nothing wraps around 2^32), the window of cell (m, n) covers the atoms
// k(m) = k00 + a*m .. k(m) + L(n) - 1, L(n) = 3*tau_n/TAtom, clipped at the
// data end, and atom k(m) + l has weight rho^l, rho = exp(-TAtom/tau_n)
// (exp.cu:53-90: t1 = t0 + 3 tau, so i_t1 = k + L - 1, and t0 <= t_i <= t1).
// The weighted window sum E(k) = sum_{l < L} rho^l x_{k+l} slides by one atom
// at a time:
//     E(k - 1) = rho E(k) + x_{k-1} - rho^L x_{k-1+L}
// (rho^2 and rho^2L = e^-6 for the a^2, b^2, ab columns; rho^L = e^-3), so a
// cell costs O(1) instead of O(L).  This first-order linear recurrence runs
// down the rows in chunks of RB rows (the map's row blocks), in parallel:
//   exp_init_kernel  E at the top row of the top chunk, by the direct O(L) sum;
//   exp_seg_kernel   Y(s): the recurrence across chunk s started from E = 0;
//   exp_scan_kernel  per tau column, E at every chunk top:
//                    E_top(s-1) = rho^(a RB) E_top(s) + Y(s);
//   exp_fast_kernel  each chunk walks its rows from E_top, evaluating F.
// Chunk s spans rows s*RB .. (s+1)*RB - 1 (rows past N_t0 are walked but not
// output).  Batches with many (template, column block) pairs use one chunk per
// column (RB up to 1024 rows): no exp_seg or exp_scan, and exp_fast adds the
// direct sum's parts itself.  Atoms at or past H (no output window reaches
// them) count as zero everywhere, which keeps the recurrence exact for every
// output row.  Atoms entering and leaving the windows of a chunk are staged
// in LDS, pass by pass; one thread per tau column.

struct ExpCol {
    double rho[2];   // rho^2 (a2, b2, ab), rho (Fa, Fb)
    double rhoL[2];  // rho^(2L) = exp(-6), rho^L = exp(-3)
    int L;
};

__device__ __forceinline__ ExpCol exp_col(const TmplDev &td, int n) {
    ExpCol ec;
    ec.L = td.L0 + td.Ls * n;
    const double tau = (double)td.tau + (double)n * (double)td.dtau;
    const double r1 = exp(-(double)td.TAtom / tau);
    ec.rho[0] = r1 * r1;
    ec.rho[1] = r1;
    ec.rhoL[1] = exp(-3.0);
    ec.rhoL[0] = exp(-6.0);
    return ec;
}

// Rows per work item of the recurrence at most (the walk stages its atoms in
// passes, so this bounds the item's pass bookkeeping, not its LDS).
constexpr int MAX_RB_EXPF = 1024;
// batches with fewer (template, column block) pairs than this walk the rows
// above the map instead of summing the top chunk directly
constexpr int64_t DIRECT_INIT_MIN_BLOCKS = 256;
// Untuned batches with at least DIRECT_INIT_MIN_BLOCKS pairs: row chunks per
// column block so that the map kernel has about this many work items.  When
// the pairs alone reach it, every column is one chunk: the map kernel walks
// it from the direct sum at its top and no chunk walk from zero (exp_seg) or
// chunk scan is needed.
constexpr int64_t EXP_TARGET_ITEMS = 512;
constexpr int INIT_PARTS = 8;  // blocks sharing one column block's direct sum
// exp_init beside exp_seg on a second stream only when the two launches hold
// at least this many blocks together (a fork and join costs ~10 µs)
constexpr int64_t EXP_FORK_MIN_BLOCKS = 1024;

// Atoms a walk stages in LDS, per pass of up to EXPW_PASS rows (dt0 == TAtom).
// Entering (head) atoms are the same on every lane of a row step: atom-major
// [j][8] doubles, read as broadcasts.  Leaving (tail) atoms differ per lane (L
// grows with the column): component planes [c][EXPW_TAIL_CAP] doubles, lanes
// Ls doubles apart (Ls = 3 for dtau = TAtom: 16 lanes hit 32 distinct banks).
// FP64 in LDS spares the walk a conversion per component and step.  Both
// ranges start one atom below the pass's last step, so every row reads the
// next row's atoms ahead (in range) while its own F is evaluated.
// 896 x 7 x 8 + 136 x 64 B = 58.9 KB: two workgroups per CU.
constexpr int EXPW_PASS = 128;
constexpr int EXPW_TAIL_CAP = 896;
constexpr int EXPW_HEAD_CAP = EXPW_PASS + 8;
constexpr int EXPF_LDS_BYTES = (NC * EXPW_TAIL_CAP + EXPW_HEAD_CAP * MERGED_STRIDE) *
                               (int)sizeof(double);
constexpr int EXPW_MAX_PASSES = MAX_RB_EXPF / 16;

// Rows per LDS pass of a walk over `cols` columns, or 0 when its atoms do not
// fit (dt0 > TAtom, or windows spread too far): the walk then reads the
// table directly.  A multiple of 4, at least 16.
__host__ __device__ __forceinline__ int expw_pass_rows(int a, int Ls, int cols) {
    if (a != 1) return 0;
    const int64_t room = (int64_t)EXPW_TAIL_CAP - 1 - (int64_t)Ls * (cols - 1);
    const int64_t pr = (room < EXPW_PASS ? room : (int64_t)EXPW_PASS) & ~(int64_t)3;
    return pr >= 16 ? (int)pr : 0;
}

struct ExpWalk {
    const float *X;   // the template's atom table (float32 merged sums)
    double *tl;       // LDS tail planes [c][EXPW_TAIL_CAP]
    double *hd;       // LDS head atoms [j][MERGED_STRIDE]
    int64_t head_lo;  // atom in head slot 0
    int64_t tail_lo;  // atom in tail slot 0
    int H;
};

__device__ __forceinline__ ExpWalk expw_init(const TmplDev &td, const float *X, double *lds) {
    ExpWalk wk;
    wk.X = X;
    wk.tl = lds;
    wk.hd = lds + NC * EXPW_TAIL_CAP;
    wk.head_lo = 0;
    wk.tail_lo = 0;
    wk.H = td.H;
    return wk;
}

// Staging of one pass: the atoms of the steps at rows sb..st (dt0 == TAtom)
// of columns col0..col_last -- head atoms k(m) - 1 and tail atoms k(m) - 1 +
// L(n), each range with one more atom below (read ahead at row sb); atoms
// outside [0, H) count as zero.  Thread tid moves component tid & 7 of atoms
// (tid >> 3) + 32 k (consecutive threads read consecutive doubles of the
// table), through registers: expw_fetch issues the loads, expw_put writes
// LDS, so the next pass's loads stay in flight during the current walk.
constexpr int EXPW_HEAD_PER = (EXPW_HEAD_CAP + 31) / 32;
constexpr int EXPW_TAIL_PER = EXPW_TAIL_CAP / 32;
static_assert(EXPW_TAIL_CAP % 32 == 0 && WG == 256, "staging decomposition");

struct ExpPass {
    int64_t head_lo, tail_lo;
    int nhead, ntail;
    float h[EXPW_HEAD_PER];  // exact: the table holds float32 sums
    float t[EXPW_TAIL_PER];
};

__device__ __forceinline__ void expw_fetch(const ExpWalk &wk, const TmplDev &td, int st, int sb,
                                           int col0, int col_last, ExpPass &ps) {
    ps.head_lo = (int64_t)td.k00 + sb - 2;
    ps.tail_lo = ps.head_lo + td.L0 + (int64_t)td.Ls * col0;
    ps.nhead = st - sb + 2;
    ps.ntail = ps.nhead + td.Ls * (col_last - col0);
    const int c = threadIdx.x & 7, j0 = threadIdx.x >> 3;
#pragma unroll
    for (int k = 0; k < EXPW_HEAD_PER; k++) {
        const int j = j0 + 32 * k;
        const int64_t atom = ps.head_lo + j;
        ps.h[k] = j < ps.nhead && atom >= 0 && atom < wk.H ? wk.X[atom * MERGED_STRIDE + c] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < EXPW_TAIL_PER; k++) {
        const int j = j0 + 32 * k;
        const int64_t atom = ps.tail_lo + j;
        ps.t[k] = c < NC && j < ps.ntail && atom >= 0 && atom < wk.H
                      ? wk.X[atom * MERGED_STRIDE + c]
                      : 0.0f;
    }
}

// The caller synchronises before the walk reads LDS.
__device__ __forceinline__ void expw_put(ExpWalk &wk, const ExpPass &ps) {
    wk.head_lo = ps.head_lo;
    wk.tail_lo = ps.tail_lo;
    const int c = threadIdx.x & 7, j0 = threadIdx.x >> 3;
#pragma unroll
    for (int k = 0; k < EXPW_HEAD_PER; k++) {
        const int j = j0 + 32 * k;
        if (j < ps.nhead) wk.hd[j * MERGED_STRIDE + c] = (double)ps.h[k];
    }
#pragma unroll
    for (int k = 0; k < EXPW_TAIL_PER; k++) {
        const int j = j0 + 32 * k;
        if (c < NC && j < ps.ntail) wk.tl[c * EXPW_TAIL_CAP + j] = (double)ps.t[k];
    }
}

// Operands of one row step from LDS: head slot jh (wave-uniform), tail slot jt.
__device__ __forceinline__ void expw_load(const ExpWalk &wk, int jh, int jt, double (&xh)[NC],
                                          double (&xt)[NC]) {
    const double *h = wk.hd + jh * MERGED_STRIDE;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        xh[c] = h[c];
        xt[c] = wk.tl[c * EXPW_TAIL_CAP + jt];
    }
}

// E(k - 1) = rho E(k) + (x_{k-1} - rho^L x_{k-1+L})
__device__ __forceinline__ void expw_step(const ExpCol &ec, const double (&xh)[NC],
                                          const double (&xt)[NC], double (&E)[NC]) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int g = c < 3 ? 0 : 1;
        const double d = fma(-ec.rhoL[g], xt[c], xh[c]);
        E[c] = fma(ec.rho[g], E[c], d);
    }
}

// One row step read from the table: k -> k - a (a atom steps).  The entering
// atoms k - l >= 0 for every step a walk takes.
__device__ __forceinline__ void expw_step_table(const ExpWalk &wk, const ExpCol &ec, int a,
                                                int64_t k, double (&E)[NC]) {
    for (int l = 1; l <= a; l++) {
        const int64_t kk = k - l;     // entering atom (wave-uniform)
        const int64_t kt = kk + ec.L;  // leaving atom
        const float *ph = wk.X + min(kk, (int64_t)wk.H - 1) * MERGED_STRIDE;
        const float *pt = wk.X + min(kt, (int64_t)wk.H - 1) * MERGED_STRIDE;
        double xh[NC], xt[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            xh[c] = kk < wk.H ? (double)ph[c] : 0.0;
            xt[c] = kt < wk.H ? (double)pt[c] : 0.0;
        }
        expw_step(ec, xh, xt, E);
    }
}

// Etab layout per template: [segment][component][tau column]
__device__ __forceinline__ double *expf_etab(double *etab, const TmplDev &td, int s, int c) {
    return etab + td.etab_off + ((int64_t)s * NC + c) * td.N_tau;
}

// Y(s), s >= 1: the recurrence across chunk s (from its top row down to the
// top row of chunk s-1) started from zero.  Work items (t, s, cb).
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(2, 2))) void exp_seg_kernel(
    const TmplDev *__restrict__ tds, const ItemDev *__restrict__ items,
    const float *__restrict__ merged, double *__restrict__ etab, int lds_ok) {
    extern __shared__ double lds_raw[];  // one dynamic LDS symbol per file
    const ItemDev it = items[block_slot()];
    const TmplDev td = tds[it.t];
    const int RB = td.rb_rows;
    const int col0 = it.cb * TILE_COLS;
    const int n = col0 + threadIdx.x;
    const int N_tau = (int)td.N_tau;
    const bool valid = n < N_tau;
    const int col_last = min(col0 + TILE_COLS, N_tau) - 1;
    // steps at rows r_top .. r_bot + 1
    const int r_top = (it.rb + 1) * RB - 1, r_bot = it.rb * RB - 1;
    const int ncol = valid ? n : col0;  // lanes past the map follow the first column
    const ExpCol ec = exp_col(td, ncol);
    ExpWalk wk = expw_init(td, merged + td.merged_off, lds_raw);
    double E[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) E[c] = 0.0;
    const int pr = lds_ok ? expw_pass_rows(td.a, td.Ls, col_last - col0 + 1) : 0;  // uniform
    if (pr > 0) {
        ExpPass ps;
        expw_fetch(wk, td, r_top, max(r_bot + 1, r_top - pr + 1), col0, col_last, ps);
        for (int st = r_top; st > r_bot; st -= pr) {
            const int sb = max(r_bot + 1, st - pr + 1);
            expw_put(wk, ps);
            __syncthreads();
            if (st - pr > r_bot)  // the next pass's atoms, in flight during this walk
                expw_fetch(wk, td, st - pr, max(r_bot + 1, st - 2 * pr + 1), col0, col_last, ps);
            int jh = st - sb + 1;
            int jt = jh + td.Ls * (ncol - col0);
            double xh[NC], xt[NC];
            expw_load(wk, jh, jt, xh, xt);
#pragma unroll 2
            for (int m = st; m >= sb; m--) {
                double nh[NC], nt[NC];
                expw_load(wk, jh - 1, jt - 1, nh, nt);  // the next row's atoms
                expw_step(ec, xh, xt, E);
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    xh[c] = nh[c];
                    xt[c] = nt[c];
                }
                jh--;
                jt--;
            }
            __syncthreads();  // the next pass restages
        }
    } else {
        int64_t k = (int64_t)td.k00 + (int64_t)td.a * r_top;
        for (int m = r_top; m > r_bot; m--) {
            expw_step_table(wk, ec, td.a, k, E);
            k -= td.a;
        }
    }
    if (valid) {
#pragma unroll
        for (int c = 0; c < NC; c++) expf_etab(etab, td, it.rb, c)[n] = E[c];
    }
}

// E at the top row of the top chunk, r = n_seg*RB - 1, by the direct sum
// E(k) = sum_{l < L} rho^l x_{k+l} (atoms at or past H count as zero, as in
// the walks), split over INIT_PARTS waves per (template, column block): part
// p sums its share of the block's l range into slot n_seg + p; the scan adds
// the parts.  Each lane sums INIT_GROUPS columns (its lane, + 64, + 128, +
// 192 of the block), so every atom, read from LDS as a broadcast, feeds up to
// 4 x 9 FMA/MUL instructions.  The wave stages its atoms in LDS in chunks of
// INIT_CHUNK with coalesced loads, the next chunk in flight in registers
// while the current one is summed.  The l range runs in segments that end
// where a column group's longest window ends (groups whose windows have all
// ended drop out); within 63 Ls atoms of that end the lanes of the group
// mask their weights.
constexpr int INIT_WG = 64;
constexpr int INIT_GROUPS = TILE_COLS / INIT_WG;
constexpr int INIT_CHUNK = 64;                                // atoms per LDS chunk (4 KB)
constexpr int INIT_PER = INIT_CHUNK * MERGED_STRIDE / INIT_WG;  // doubles per lane per chunk

struct InitState {
    double E[INIT_GROUPS][NC];
    double w[INIT_GROUPS][2];
    double rho[INIT_GROUPS][2];
    int64_t lr[INIT_GROUPS];  // this lane's window end, per group
    int64_t lo[INIT_GROUPS];  // the group's shortest window (uniform)
};

template <bool MASK>
__device__ __forceinline__ void init_add(InitState &S, int i, int64_t l, const double *x) {
    double wa = S.w[i][0], wb = S.w[i][1];
    if (MASK) {
        const bool in = l < S.lr[i];
        wa = in ? wa : 0.0;
        wb = in ? wb : 0.0;
    }
#pragma unroll
    for (int c = 0; c < NC; c++) S.E[i][c] = fma(c < 3 ? wa : wb, x[c], S.E[i][c]);
    S.w[i][0] *= S.rho[i][0];
    S.w[i][1] *= S.rho[i][1];
}

// atoms [a, b) of the chunk starting at atom c0, for the column groups I0 ..
// INIT_GROUPS - 1; group I0 masks its lanes' weights when MASK0, the later
// groups never need to (their shortest window reaches past b)
template <int I0, bool MASK0>
__device__ __forceinline__ void init_segment(InitState &S, const double *buf, int64_t c0,
                                             int64_t a, int64_t b) {
#pragma unroll 2
    for (int64_t l = a; l < b; l++) {
        double x[NC];
        const double *p = buf + (l - c0) * MERGED_STRIDE;  // broadcast
#pragma unroll
        for (int c = 0; c < NC; c++) x[c] = p[c];
        init_add<MASK0>(S, I0, l, x);
#pragma unroll
        for (int i = I0 + 1; i < INIT_GROUPS; i++) init_add<false>(S, i, l, x);
    }
}

__global__ __launch_bounds__(INIT_WG) void exp_init_kernel(const TmplDev *__restrict__ tds,
                                                           const ItemDev *__restrict__ items,
                                                           const float *__restrict__ merged,
                                                           double *__restrict__ etab) {
    __shared__ double buf[INIT_CHUNK * MERGED_STRIDE];
    const ItemDev it = items[blockIdx.x];  // (t, part, column block)
    const int part = it.rb;
    const TmplDev td = tds[it.t];
    const int col0 = it.cb * TILE_COLS;
    const int col_last = min(col0 + TILE_COLS, (int)td.N_tau) - 1;
    const int lane = (int)threadIdx.x;
    const int64_t k0 = (int64_t)td.k00 + (int64_t)td.a * ((int64_t)td.n_seg * td.rb_rows - 1);
    const int64_t room = (int64_t)td.H - k0;
    auto clip = [&](int64_t L) { return max(min(L, room), (int64_t)0); };
    // longest window of the block (L grows with n), clipped at H: uniform
    const int64_t lr_blk = clip((int64_t)td.L0 + (int64_t)td.Ls * col_last);
    const int64_t q = (lr_blk + INIT_PARTS - 1) / INIT_PARTS;
    const int64_t l0 = part * q, l1 = min(lr_blk, l0 + q);
    InitState S;
    int64_t hi[INIT_GROUPS];  // each group's longest window (uniform)
#pragma unroll
    for (int i = 0; i < INIT_GROUPS; i++) {
#pragma unroll
        for (int c = 0; c < NC; c++) S.E[i][c] = 0.0;
        const int g0 = col0 + INIT_WG * i;
        const int glast = min(g0 + INIT_WG - 1, col_last);
        const int n = min(g0 + lane, col_last);  // lanes past the map follow the last column
        const bool any = g0 <= col_last;
        S.lr[i] = any ? clip((int64_t)td.L0 + (int64_t)td.Ls * n) : 0;
        S.lo[i] = any ? clip((int64_t)td.L0 + (int64_t)td.Ls * g0) : 0;
        hi[i] = any ? clip((int64_t)td.L0 + (int64_t)td.Ls * glast) : 0;
        const double tau = (double)td.tau + (double)n * (double)td.dtau;
        const double r1 = exp(-(double)td.TAtom / tau);
        S.rho[i][0] = r1 * r1;
        S.rho[i][1] = r1;
        const double w1 = exp(-(double)l0 * (double)td.TAtom / tau);  // rho^l0
        S.w[i][0] = w1 * w1;
        S.w[i][1] = w1;
    }
    // Windows grow with the column: group i's shortest window (lo) is at
    // least group i-1's longest (hi).  Atoms below lo[0] lie inside every
    // window; segment I0 = [hi[I0-1], hi[I0]) needs group I0 masked and no
    // group before it.
    const int64_t seg_end[5] = {S.lo[0], hi[0], hi[1], hi[2], hi[3]};
    // rows k0 + l, l < l1 <= room, lie below H
    const float *__restrict__ X = merged + td.merged_off + k0 * MERGED_STRIDE;
    float r[INIT_PER];  // exact: the table holds float32 sums
    auto fetch = [&](int64_t c0) {
        const int64_t nq = min((int64_t)INIT_CHUNK, l1 - c0) * MERGED_STRIDE;
#pragma unroll
        for (int k = 0; k < INIT_PER; k++) {
            const int64_t qd = (int64_t)k * INIT_WG + lane;
            r[k] = qd < nq ? X[c0 * MERGED_STRIDE + qd] : 0.0f;
        }
    };
    if (l0 < l1) fetch(l0);
    for (int64_t c0 = l0; c0 < l1; c0 += INIT_CHUNK) {
        const int64_t c1 = min(l1, c0 + INIT_CHUNK);
        __syncthreads();  // the previous chunk is consumed
#pragma unroll
        for (int k = 0; k < INIT_PER; k++) buf[k * INIT_WG + lane] = (double)r[k];
        __syncthreads();
        if (c1 < l1) fetch(c1);  // in flight while this chunk is summed
        int64_t sa = c0;
#pragma unroll
        for (int sgm = 0; sgm < 5; sgm++) {
            const int64_t sb = min(c1, max(sa, seg_end[sgm]));
            if (sa < sb) {  // uniform
                if (sgm == 0) init_segment<0, false>(S, buf, c0, sa, sb);
                if (sgm == 1) init_segment<0, true>(S, buf, c0, sa, sb);
                if (sgm == 2) init_segment<1, true>(S, buf, c0, sa, sb);
                if (sgm == 3) init_segment<2, true>(S, buf, c0, sa, sb);
                if (sgm == 4) init_segment<3, true>(S, buf, c0, sa, sb);
            }
            sa = sb;
        }
    }
#pragma unroll
    for (int i = 0; i < INIT_GROUPS; i++) {
        const int n = col0 + INIT_WG * i + lane;
        if (n <= col_last) {
#pragma unroll
            for (int c = 0; c < NC; c++) expf_etab(etab, td, td.n_seg + part, c)[n] = S.E[i][c];
        }
    }
}

// Per tau column: E at every chunk top.  Work items (t, -, cb); one wave per
// column, 4 columns per workgroup.  With f_s(x) = mult x + Y(s) (mult = the
// weight factor of one chunk of rows), the chunk tops are V(n_seg-1) = the
// direct sum and V(s-1) = f_s(V(s)), a first-order linear recurrence with a
// constant coefficient: lane l of a 64-chunk tile holds Y(s_hi - l) and a
// log-depth (Kogge-Stone) scan, x_l += mult^d x_{l-d} for d = 1 .. 32, gives
// V(s_hi - l) = mult^l V(s_hi) + x_{l-1}; tiles run from the top chunk down,
// each carrying V(s_hi - 64) = mult^64 V(s_hi) + x_63.  Slot s holds Y(s) on
// entry and V(s) on exit (slot 0 has no Y).
constexpr int SCAN_COLS = WG / 64;  // columns per workgroup

__global__ __launch_bounds__(WG) void exp_scan_kernel(const TmplDev *__restrict__ tds,
                                                      const ItemDev *__restrict__ items,
                                                      double *__restrict__ etab) {
    const ItemDev it = items[blockIdx.x / (TILE_COLS / SCAN_COLS)];
    const TmplDev td = tds[it.t];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int n = it.cb * TILE_COLS + (int)(blockIdx.x % (TILE_COLS / SCAN_COLS)) * SCAN_COLS + wave;
    if (n >= (int)td.N_tau) return;  // wave-uniform
    const double tau = (double)td.tau + (double)n * (double)td.dtau;
    const double xstep = (double)td.TAtom * (double)td.a * (double)td.rb_rows / tau;
    // mult[k]^d for the scan steps d = 1, 2, ..., 32 (k = 0: A, B, C; k = 1: Fa, Fb)
    double pw[2][6];
    pw[1][0] = exp(-xstep);
    pw[0][0] = pw[1][0] * pw[1][0];
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int d = 1; d < 6; d++) pw[k][d] = pw[k][d - 1] * pw[k][d - 1];
    // mult^l for this lane, and mult^64
    double ml[2], m64[2];
    ml[1] = exp(-xstep * (double)lane);
    ml[0] = ml[1] * ml[1];
    m64[0] = pw[0][5] * pw[0][5];
    m64[1] = pw[1][5] * pw[1][5];
    double carry[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {  // top chunk: the direct sum's parts
        carry[c] = 0.0;
        if (!td.init_empty)
            for (int pp = 0; pp < INIT_PARTS; pp++)
                carry[c] += expf_etab(etab, td, td.n_seg + pp, c)[n];
    }
    for (int s_hi = td.n_seg - 1; s_hi >= 0; s_hi -= 64) {
        const int sl = s_hi - lane;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int k = c < 3 ? 0 : 1;
            double *slot = expf_etab(etab, td, max(sl, 0), c) + n;
            double x = sl >= 1 ? *slot : 0.0;
#pragma unroll
            for (int d = 0; d < 6; d++) {
                const double y = __shfl_up(x, 1 << d);
                if (lane >= (1 << d)) x = fma(pw[k][d], y, x);
            }
            double sx = __shfl_up(x, 1);
            if (lane == 0) sx = 0.0;
            if (sl >= 0) *slot = fma(ml[k], carry[c], sx);
            const double x63 = __shfl(x, 63);
            carry[c] = fma(m64[k], carry[c], x63);
        }
    }
}

// The same scan, sequential along the chunks: one thread per tau column (its
// loads and stores coalesce across the columns).  For batches of many
// templates with few chunks each, where one wave per column would leave most
// lanes idle.
constexpr int SCAN_WG = 64;
constexpr int SCAN_SPLIT = TILE_COLS / SCAN_WG;

__global__ __launch_bounds__(SCAN_WG) void exp_scan_seq_kernel(const TmplDev *__restrict__ tds,
                                                               const ItemDev *__restrict__ items,
                                                               double *__restrict__ etab) {
    const ItemDev it = items[blockIdx.x / SCAN_SPLIT];
    const TmplDev td = tds[it.t];
    const int n = it.cb * TILE_COLS + (int)(blockIdx.x % SCAN_SPLIT) * SCAN_WG + threadIdx.x;
    if (n >= (int)td.N_tau) return;
    const double tau = (double)td.tau + (double)n * (double)td.dtau;
    const double steps = (double)td.a * td.rb_rows;
    const double m1 = exp(-(double)td.TAtom * steps / tau);
    const double mult[2] = {m1 * m1, m1};
    double E[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {  // top chunk: the direct sum's parts
        E[c] = 0.0;
        if (!td.init_empty)
            for (int pp = 0; pp < INIT_PARTS; pp++) E[c] += expf_etab(etab, td, td.n_seg + pp, c)[n];
    }
    for (int s = td.n_seg - 1; s >= 1; s--) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            double *slot = expf_etab(etab, td, s, c) + n;
            const double Y = *slot;
            *slot = E[c];
            E[c] = fma(mult[c < 3 ? 0 : 1], E[c], Y);
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) expf_etab(etab, td, 0, c)[n] = E[c];
}


// Rows st .. sb of a chunk (one pass) in groups of 4 (the pass's rows are a
// multiple of 4), reduced together.  Rows past the map (m > m_hi, only in a
// template's last chunk) are evaluated but neither stored nor reduced, which
// keeps the loop free of branches.  LDS: the pass's atoms are staged (dt0 ==
// TAtom), jt_off = Ls * (the
// lane's column - the first column); otherwise each step reads the table.
template <bool BTSG, bool STRICT, bool LDS>
__device__ __forceinline__ void expw_rows(const TmplDev &td, const ExpWalk &wk, const ExpCol &ec,
                                          double (&E)[NC], Reducer<BTSG, 1> &red, bool &single,
                                          float *__restrict__ Ft, int st, int sb, int row0,
                                          int m_hi, int n, bool valid, int lane, int jt_off,
                                          float *sh_rowM, float *sh_rowS) {
    const int N_tau = (int)td.N_tau;
    const int colidx[1] = {n};
    int64_t k = (int64_t)td.k00 + (int64_t)td.a * st;  // first window atom of row m
    // stores without a branch: one buffer resource over the pass's rows inside
    // the map, a byte offset per lane that falls outside it for lanes past the
    // last column (2^31 and below it by less than 2^31 minus the resource's
    // size, as the plan bounds N_tau); rows past the map lie beyond it too
    const int nv = min(st, m_hi) - sb + 1;
    const uint32_t rowb = (uint32_t)N_tau * (uint32_t)sizeof(float);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        Ft + (int64_t)sb * N_tau, (short)0, nv > 0 ? (int)((uint32_t)nv * rowb) : 0,
        BUF_RSRC_WORD3);
    uint32_t vo = valid ? (uint32_t)(st - sb) * rowb + (uint32_t)n * (uint32_t)sizeof(float)
                        : 0x80000000u;  // row st
    // LDS path: head and tail slots of the step at the group's first row;
    // the group's rows address slots below them by constants
    int jh = st - sb + 1;
    int jt = jh + jt_off;
    double xh[NC], xt[NC];
    if (LDS) expw_load(wk, jh, jt, xh, xt);
    for (int g = st; g >= sb; g -= 4) {
        const bool last = g - 3 == row0;
        float F4[4][1];  // F4[r] is row g - 3 + r
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int m = g - j;
            double nh[NC], nt[NC];
            if (LDS) expw_load(wk, jh - 1 - j, jt - 1 - j, nh, nt);  // the next row's atoms
            // every lane evaluates F (no branch around it, so the compiler
            // can overlap the rows' evaluations); stores and masks follow
            F4[3 - j][0] = cell_F(E[0], E[1], E[2], E[3], E[4], E[5], E[6]);
            if (STRICT) single |= valid && m <= m_hi && (k >= (int64_t)td.N - 1);
            // the chunk's final row takes no step: from the table it could
            // read below atom 0; from LDS the step reads staged slots and its
            // result is never used, so it runs unconditionally (no branch)
            if (LDS)
                expw_step(ec, xh, xt, E);
            else if (j < 3 || !last)
                expw_step_table(wk, ec, td.a, k, E);
            if (LDS) {
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    xh[c] = nh[c];
                    xt[c] = nt[c];
                }
            }
            k -= td.a;
        }
        jh -= 4;
        jt -= 4;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = g - 3 + r;
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(F4[r][0]), rs,
                                                  (int)(vo - (uint32_t)(3 - r) * rowb), 0, 0);
            F4[r][0] = valid && m <= m_hi ? F4[r][0] : -INFINITY;
        }
        vo -= 4 * rowb;
        if (g - 3 <= m_hi)
            red.template add<4>(F4, g - 3, colidx, N_tau, lane, g - 3 - sb, sh_rowM, sh_rowS);
    }
}

// The map: chunk rb walks rows (rb+1)*RB - 1 .. rb*RB from E_top(rb), in
// passes.  BTSG row sums: after each pass the 4 waves' sums of its rows are
// combined relative to the pass reference (the waves' largest reference then)
// and written; at the end, the rows of passes whose reference lies below the
// item's are rescaled by the threads that wrote them.
template <bool BTSG, bool STRICT>
// Two workgroups share a CU (LDS): registers for two waves per SIMD.
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(2, 2))) void exp_fast_kernel(
    const TmplDev *__restrict__ tds, const ItemDev *__restrict__ items, int64_t item0,
    const float *__restrict__ merged, const double *__restrict__ etab,
    float *__restrict__ Fout, ItemPart *__restrict__ ipart, float *__restrict__ rowpart,
    float *__restrict__ colpart, int *__restrict__ flags, int lds_ok) {
    extern __shared__ double lds_raw[];  // one dynamic LDS symbol per file
    __shared__ float sh_rowM[4][EXPW_PASS];
    __shared__ float sh_rowS[4][EXPW_PASS];
    __shared__ float sh_passM[EXPW_MAX_PASSES];
    __shared__ float sh_best[4];
    __shared__ int sh_bidx[4];
    __shared__ float sh_M[4];
    __shared__ float sh_S[4];
    const int64_t item_id = item0 + block_slot();
    const ItemDev it = items[item_id];
    const TmplDev td = tds[it.t];
    const int RB = td.rb_rows;
    const int row0 = it.rb * RB;
    const int N_t0 = (int)td.N_t0, N_tau = (int)td.N_tau;
    const int r_top = row0 + RB - 1;
    const int m_hi = min(r_top, N_t0 - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int col0 = it.cb * TILE_COLS;
    const int n = col0 + threadIdx.x;
    const bool valid = n < N_tau;
    const int col_last = min(col0 + TILE_COLS, N_tau) - 1;
    const int ncol = valid ? n : col0;
    const ExpCol ec = exp_col(td, ncol);
    ExpWalk wk = expw_init(td, merged + td.merged_off, lds_raw);
    double E[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) E[c] = 0.0;
    if (valid && td.n_seg > 1) {
#pragma unroll
        for (int c = 0; c < NC; c++) E[c] = expf_etab((double *)etab, td, it.rb, c)[n];
    } else if (valid && !td.init_empty) {  // one chunk: E_top is the direct sum's parts
#pragma unroll
        for (int pp = 0; pp < INIT_PARTS; pp++)
#pragma unroll
            for (int c = 0; c < NC; c++) E[c] += expf_etab((double *)etab, td, 1 + pp, c)[n];
    }

    Reducer<BTSG, 1> red;
    red.init();
    bool single = false;
    float *__restrict__ Ft = Fout + td.F_off;
    float *__restrict__ rp = rowpart + td.rowpart_off + (int64_t)it.cb * td.N_t0;
    const int prl = lds_ok ? expw_pass_rows(td.a, td.Ls, col_last - col0 + 1) : 0;  // uniform
    const int pr = prl > 0 ? prl : EXPW_PASS;
    int npass = 0;
    ExpPass ps;
    if (prl > 0) expw_fetch(wk, td, r_top, max(row0, r_top - pr + 1), col0, col_last, ps);
    for (int st = r_top; st >= row0; st -= pr, npass++) {
        const int sb = max(row0, st - pr + 1);
        if (prl > 0) {
            expw_put(wk, ps);
            __syncthreads();
            if (st - pr >= row0)  // the next pass's atoms, in flight during this walk
                expw_fetch(wk, td, st - pr, max(row0, st - 2 * pr + 1), col0, col_last, ps);
            expw_rows<BTSG, STRICT, true>(td, wk, ec, E, red, single, Ft, st, sb, row0, m_hi, n,
                                          valid, lane, td.Ls * (ncol - col0), sh_rowM[wave],
                                          sh_rowS[wave]);
        } else {
            expw_rows<BTSG, STRICT, false>(td, wk, ec, E, red, single, Ft, st, sb, row0, m_hi, n,
                                           valid, lane, 0, sh_rowM[wave], sh_rowS[wave]);
        }
        const int nv = min(st, m_hi) - sb + 1;  // rows of the pass inside the map
        if (BTSG && nv > 0) {  // block-uniform
            if (lane == 0) sh_M[wave] = red.ML;
            __syncthreads();
            const float Ms = fmaxf(fmaxf(sh_M[0], sh_M[1]), fmaxf(sh_M[2], sh_M[3]));
            for (int rl = threadIdx.x; rl < nv; rl += WG) {
                float S = 0.0f;
#pragma unroll
                for (int w = 0; w < 4; w++)
                    if (sh_M[w] != -INFINITY) S += sh_rowS[w][rl] * exp2_diff(sh_rowM[w][rl], Ms);
                rp[sb + rl] = S;
            }
            if (threadIdx.x == 0) sh_passM[npass] = Ms;
        }
        __syncthreads();  // staging, row sums and sh_M are rewritten by the next pass
    }
    if (STRICT && single) atomicOr(flags + it.t, 1);
    // combine: each column belongs to one lane; each row to all 4 waves
    wave_argmax(red.best, red.bidx);
    float stot = 0.0f;
    if (BTSG) stot = wave_sum(red.total());
    if (lane == 0) {
        sh_best[wave] = red.best;
        sh_bidx[wave] = red.bidx;
        sh_M[wave] = red.ML;
        sh_S[wave] = stot;
    }
    __syncthreads();
    const float M = fmaxf(fmaxf(sh_M[0], sh_M[1]), fmaxf(sh_M[2], sh_M[3]));
    if (BTSG) {
        if (valid)
            colpart[td.colpart_off + (int64_t)it.rb * td.N_tau + n] =
                red.ML == -INFINITY ? 0.0f : red.Scol[0] * exp2_diff(red.ML, M);
        // rows written relative to an earlier, lower reference: the same
        // threads (same rows) rescale them
        for (int j = 0; j < npass; j++) {
            const int st = r_top - j * pr;
            const int sb = max(row0, st - pr + 1);
            const int nv = min(st, m_hi) - sb + 1;
            if (nv <= 0) continue;
            const float Ms = sh_passM[j];
            if (Ms == M || Ms == -INFINITY) continue;  // block-uniform
            const float sc = exp2_diff(Ms, M);
            for (int rl = threadIdx.x; rl < nv; rl += WG) rp[sb + rl] *= sc;
        }
    }
    if (threadIdx.x == 0) {
        float bv = sh_best[0];
        int bi = sh_bidx[0];
#pragma unroll
        for (int w = 1; w < 4; w++) argmax_merge(bv, bi, sh_best[w], sh_bidx[w]);
        float S = 0.0f;
        if (BTSG) {
#pragma unroll
            for (int w = 0; w < 4; w++)
                if (sh_M[w] != -INFINITY) S += sh_S[w] * exp2_diff(sh_M[w], M);
        }
        ipart[item_id] = ItemPart{bv, bi, M, S};
    }
}

// ----------------------------------------------------------------------------
// finalize: per-template stats from the partials
// ----------------------------------------------------------------------------

template <int NT>
__device__ double block_sum_d(double v) {
    __shared__ double sd[NT / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sd[wave] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < NT / 64; w++) s += sd[w];
    __syncthreads();
    return s;
}

constexpr int FIN_SCALE_CAP = 4096;        // per-item scale factors kept in LDS
constexpr int64_t FIN_WIDE_MAX_T = 256;  // batches smaller than this finalize with 1024 threads

// Block-wide: argmax (first index on ties) of (bv, bi) and maximum of mm.
template <int NT>
__device__ void block_argmax_max(float &bv, int &bi, float &mm) {
    __shared__ float sv[NT / 64], sm[NT / 64];
    __shared__ int si[NT / 64];
    wave_argmax(bv, bi);
    mm = wave_max(mm);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        sv[wave] = bv;
        si[wave] = bi;
        sm[wave] = mm;
    }
    __syncthreads();
    bv = sv[0];
    bi = si[0];
    mm = sm[0];
    for (int w = 1; w < NT / 64; w++) {
        argmax_merge(bv, bi, sv[w], si[w]);
        mm = fmaxf(mm, sm[w]);
    }
    __syncthreads();
}

// Block-wide argmax over doubles (first index on ties) of two pairs at once.
template <int NT>
__device__ void block_argmax2_d(double &av, int &ai, double &bv, int &bi) {
    __shared__ double sav[NT / 64], sbv[NT / 64];
    __shared__ int sai[NT / 64], sbi[NT / 64];
    auto merge = [](double &v, int &i, double ov, int oi) {
        if (ov > v || (ov == v && oi < i)) {
            v = ov;
            i = oi;
        }
    };
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double oa = __shfl_xor(av, o), ob = __shfl_xor(bv, o);
        const int oai = __shfl_xor(ai, o), obi = __shfl_xor(bi, o);
        merge(av, ai, oa, oai);
        merge(bv, bi, ob, obi);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        sav[wave] = av;
        sai[wave] = ai;
        sbv[wave] = bv;
        sbi[wave] = bi;
    }
    __syncthreads();
    av = sav[0];
    ai = sai[0];
    bv = sbv[0];
    bi = sbi[0];
    for (int w = 1; w < NT / 64; w++) {
        merge(av, ai, sav[w], sai[w]);
        merge(bv, bi, sbv[w], sbi[w]);
    }
    __syncthreads();
}

// One workgroup of NT threads per template (NT = 1024 for small batches, whose
// maps are large, 256 otherwise).  Three block reductions: the argmax of F
// with the largest item reference Mref (log2 units), the lnBtSG sum, and the
// two posterior argmaxes together.  Item k's partials are relative to
// 2^(M_k) (M_k = -inf: no valid cell, partials 0), rescaled here by
// exp((M_k - Mref) / LOG2E) <= 1, so no item reference needs maxF; the row
// and column partials of a thread are read in one pass.
template <bool BTSG, int NT>
__global__ __launch_bounds__(NT) void finalize_kernel(
    const TmplDev *__restrict__ tds, const ItemPart *__restrict__ ipart,
    const float *__restrict__ rowpart, const float *__restrict__ colpart,
    const int *__restrict__ flags, tcw_stats *__restrict__ stats) {
    __shared__ float sh_scale[FIN_SCALE_CAP];
    // the wide variant (few, large maps: many partials per row and column)
    // unrolls the partial sums by 8 so their independent loads issue
    // together; the narrow one (many templates) is faster as the compiler
    // schedules it by itself
    constexpr bool FIN_UNROLL = NT >= 1024;
    constexpr double INV_LOG2E = 1.0 / (double)LOG2E;
    const int t = blockIdx.x;
    const TmplDev td = tds[t];
    const int64_t nitems = (int64_t)td.n_rowblk * td.n_colblk;
    const ItemPart *ip = ipart + td.item_begin;
    float bv = -INFINITY, Mref = -INFINITY;
    int bi = 0x7fffffff;
    for (int64_t k = threadIdx.x; k < nitems; k += NT) {
        const ItemPart q = ip[k];
        argmax_merge(bv, bi, q.maxF, q.idx);
        Mref = fmaxf(Mref, q.M);
    }
    block_argmax_max<NT>(bv, bi, Mref);
    const float maxF = bv;
    const int64_t m_ML = bi / (int)td.N_tau;
    const int64_t n_ML = bi - m_ML * (int64_t)td.N_tau;

    double lnB = NAN, t0_MP = NAN, tau_MP = NAN;
    int64_t m_MP = -1, n_MP = -1;
    if (BTSG) {
        auto item_scale = [&](int64_t k) -> float {
            const float Mk = ip[k].M;
            return Mk == -INFINITY ? 0.0f : __expf((float)((double)(Mk - Mref) * INV_LOG2E));
        };
        double s = 0.0;
        for (int64_t k = threadIdx.x; k < nitems; k += NT) {
            const float sc = item_scale(k);
            if (k < FIN_SCALE_CAP) sh_scale[k] = sc;
            s += (double)ip[k].S * (double)sc;
        }
        s = block_sum_d<NT>(s);  // contains a barrier: sh_scale is complete
        lnB = log(70.0 / ((double)td.N_t0 * (double)td.N_tau)) + (double)Mref * INV_LOG2E + log(s);
        auto scale = [&](int64_t k) -> float {
            return k < FIN_SCALE_CAP ? sh_scale[k] : item_scale(k);
        };
        double rv = -1.0, cv = -1.0;
        int ri = 0x7fffffff, ci = 0x7fffffff;
        const int nmax = max((int)td.N_t0, (int)td.N_tau);
        for (int i = threadIdx.x; i < nmax; i += NT) {
            double R = -1.0, Cs = -1.0;
            if (i < (int)td.N_t0) {  // row i: [colblk][N_t0]
                const int rb = i / td.rb_rows;
                const float *rp = rowpart + td.rowpart_off + i;
                auto term = [&](uint32_t cb) {
                    return (double)(rp[(int64_t)cb * td.N_t0] *
                                    scale((int64_t)rb * td.n_colblk + cb));
                };
                R = 0.0;
                if constexpr (FIN_UNROLL) {
#pragma unroll 8
                    for (uint32_t cb = 0; cb < td.n_colblk; cb++) R += term(cb);
                } else {
                    for (uint32_t cb = 0; cb < td.n_colblk; cb++) R += term(cb);
                }
            }
            if (i < (int)td.N_tau) {  // column i: [rowblk][N_tau]
                const int cb = i / TILE_COLS;
                const float *cp = colpart + td.colpart_off + i;
                auto term = [&](uint32_t rb) {
                    return (double)(cp[(int64_t)rb * td.N_tau] *
                                    scale((int64_t)rb * td.n_colblk + cb));
                };
                Cs = 0.0;
                if constexpr (FIN_UNROLL) {
#pragma unroll 8
                    for (uint32_t rb = 0; rb < td.n_rowblk; rb++) Cs += term(rb);
                } else {
                    for (uint32_t rb = 0; rb < td.n_rowblk; rb++) Cs += term(rb);
                }
            }
            if (R > rv) {  // i grows within a thread: the first maximum stays
                rv = R;
                ri = i;
            }
            if (Cs > cv) {
                cv = Cs;
                ci = i;
            }
        }
        block_argmax2_d<NT>(rv, ri, cv, ci);
        m_MP = ri;
        n_MP = ci;
        t0_MP = (double)td.o_t0 + ((double)m_MP + 0.5) * ((double)td.o_t0Band / (double)td.N_t0);
        tau_MP = (double)td.o_tau + ((double)n_MP + 0.5) * ((double)td.o_tauBand / (double)td.N_tau);
    }
    if (threadIdx.x == 0) {
        tcw_stats st;
        st.maxF = maxF;
        st.single_atom = flags[t];
        st.m_ML = m_ML;
        st.n_ML = n_ML;
        st.N_t0 = td.N_t0;
        st.N_tau = td.N_tau;
        st.t0_ML = (int64_t)td.t0 + m_ML * (int64_t)td.dt0;
        st.tau_ML = (int64_t)td.tau + n_ML * (int64_t)td.dtau;
        st.lnBtSG = lnB;
        st.m_MP = m_MP;
        st.n_MP = n_MP;
        st.t0_MP = t0_MP;
        st.tau_MP = tau_MP;
        st.numAtoms = td.N;
        st.t0_data = td.t_min;
        stats[t] = st;
    }
}

// ----------------------------------------------------------------------------
// binned merge for many templates (tcw_merge_batch) and its adjoint
// ----------------------------------------------------------------------------

struct MergeTmpl {
    int64_t seg_begin, seg_end;  // into the SegDev table
    int64_t out_off;             // first merged bin of the template
    uint32_t N, t_min;
};

// One thread per merged bin: per detector segment a binary search for the
// bin's first atom, then float32 sums over the bin's atoms, detectors in
// order -- the accumulation order of mergeMultiFstatAtomsBinned, so the
// result is bit-identical to the stage kernel's merge.
__global__ __launch_bounds__(256) void merge_kernel(
    const MergeTmpl *__restrict__ mt, const SegDev *__restrict__ segs, int64_t T,
    const uint32_t *__restrict__ ts, const float *__restrict__ cols, int64_t cstride,
    uint32_t TA, float *__restrict__ out, int64_t ostride) {
    for (int64_t t = blockIdx.y; t < T; t += gridDim.y) {
        const MergeTmpl m = mt[t];
        const uint32_t b = blockIdx.x * 256u + threadIdx.x;
        if (b >= m.N) continue;
        float v[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) v[c] = 0.0f;
        const uint64_t t_lo = (uint64_t)m.t_min + (uint64_t)b * TA, t_hi = t_lo + TA;
        for (int64_t s = m.seg_begin; s < m.seg_end; s++) {
            const SegDev sg = segs[s];
            int64_t lo = sg.begin, hi = sg.end;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if ((uint64_t)ts[mid] < t_lo)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            for (int64_t k = lo; k < sg.end && (uint64_t)ts[k] < t_hi; k++) {
#pragma unroll
                for (int c = 0; c < NC; c++) v[c] += cols[c * cstride + k];
            }
        }
#pragma unroll
        for (int c = 0; c < NC; c++) out[c * ostride + m.out_off + b] = v[c];
    }
}

// atom_bin[k] = out_off + (ts[k] - t_min) / TA for every atom of every segment;
// segment s of template t is blockIdx.y-strided, atoms over x.
struct MergeSeg {
    int64_t begin, end, out_off;
    uint32_t t_min, pad;
};

__global__ __launch_bounds__(256) void atom_bin_kernel(const MergeSeg *__restrict__ ms,
                                                       int64_t S, const uint32_t *__restrict__ ts,
                                                       uint32_t TA, int64_t *__restrict__ atom_bin) {
    for (int64_t s = blockIdx.y; s < S; s += gridDim.y) {
        const MergeSeg g = ms[s];
        for (int64_t k = g.begin + (int64_t)blockIdx.x * 256 + threadIdx.x; k < g.end;
             k += (int64_t)gridDim.x * 256)
            atom_bin[k] = g.out_off + (int64_t)((ts[k] - g.t_min) / TA);
    }
}

__global__ __launch_bounds__(256) void gather_bins_kernel(const float *__restrict__ in,
                                                          int64_t istride,
                                                          const int64_t *__restrict__ atom_bin,
                                                          int64_t n, float *__restrict__ out,
                                                          int64_t ostride) {
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * 256) {
        const int64_t b = atom_bin[j];
#pragma unroll
        for (int c = 0; c < NC; c++) out[c * ostride + j] = in[c * istride + b];
    }
}

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------

using tcwc::DevBuf;
using tcwc::fail;
using tcwc::set_device;


}  // namespace

struct tcw_plan {
    int device = 0;
    uint32_t flags = 0;
    int64_t T = 0;
    std::vector<TmplDev> tmpl;
    std::vector<SegDev> segs;
    std::vector<ItemDev> items;      // grouped by mode
    std::vector<ItemDev> seg_items;  // MODE_EXP_FAST (t, chunk, column block), chunk >= 1
    std::vector<ItemDev> col_items;  // MODE_EXP_FAST (t, -, column block)
    std::vector<ItemDev> init_items;  // exp_init waves (t, part, column block), heaviest first
    std::vector<int32_t> exp_tmpls;  // templates needing the atom table
    int64_t exp_max_N = 0;
    bool exp_init_needed = false;  // some exp template has a non-empty direct sum
    bool exp_scan_wave = false;    // the wave-per-column chunk scan (few, long columns)
    int64_t mode_begin[MODE_COUNT + 1] = {0, 0, 0, 0, 0};
    int64_t prefix_doubles = 0, merged_floats = 0, rowpart_n = 0, colpart_n = 0;
    int64_t etab_doubles = 0;
    int64_t F_total = 0;
    int max_lds_bytes = 0;  // rect_fast_kernel dynamic LDS
    int max_lds_expf = 0;   // exp_fast_kernel dynamic LDS
    DevBuf d_tmpl, d_segs, d_items, d_prefix, d_merged, d_ipart, d_rowpart, d_colpart,
        d_flags, d_seg_items, d_col_items, d_etab, d_exp_tmpls, d_init_items;
    hipGraphExec_t graph = nullptr;
    hipStream_t cap_stream = nullptr;  // private stream the graph is captured on
    // exp_init_kernel runs here beside exp_seg_kernel (they are independent)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    const void *g_ts = nullptr, *g_cols = nullptr, *g_F = nullptr, *g_stats = nullptr;
    int64_t g_cstride = -1;
    // rows per work item requested for each kernel mode (0: built-in choice)
    int32_t rows_req[MODE_COUNT] = {0, 0, 0, 0};
    // when set, device buffers are views into this memory (tcw_map_noalloc)
    char *arena = nullptr;
    size_t arena_bytes = 0;
    void reset() {
        tmpl.clear();
        segs.clear();
        items.clear();
        seg_items.clear();
        col_items.clear();
        init_items.clear();
        exp_tmpls.clear();
        exp_max_N = 0;
        exp_init_needed = false;
        exp_scan_wave = false;
        prefix_doubles = merged_floats = rowpart_n = colpart_n = F_total = 0;
        etab_doubles = 0;
        max_lds_bytes = 0;
        max_lds_expf = 0;
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
    ~tcw_plan() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        if (side) (void)hipStreamDestroy(side);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        d_tmpl.release();
        d_segs.release();
        d_items.release();
        d_prefix.release();
        d_merged.release();
        d_ipart.release();
        d_rowpart.release();
        d_colpart.release();
        d_flags.release();
        d_seg_items.release();
        d_col_items.release();
        d_init_items.release();
        d_etab.release();
        d_exp_tmpls.release();
    }
};

// The handle keeps one plan and one stats slot for single-map calls, reused
// (grow-only) across calls so tcw_map allocates nothing in steady state.
struct tcw_handle {
    int device = 0;
    tcw_plan single;
    tcw_plan noalloc;  // tcw_map_noalloc: buffers are views into the static scratch
    tcw_stats *d_stats = nullptr;
    tcw_stats *h_stats = nullptr;  // pinned
    DevBuf d_mtab, d_mseg, d_msegs;  // tcw_merge_batch tables
    ~tcw_handle() {
        if (d_stats) (void)hipFree(d_stats);
        if (h_stats) (void)hipHostFree(h_stats);
        d_mtab.release();
        d_mseg.release();
        d_msegs.release();
    }
};

int tcwc::handle_device(const tcw_handle *h) { return h->device; }

namespace {

// Per-device address of the static scratch, looked up once per device (with
// that device current) by tcw_create; hipGetSymbolAddress's status is kept,
// so tcw_map_noalloc reports a failed lookup instead of using a bad address.
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
hipError_t g_scratch_err[NOALLOC_MAX_DEVICES] = {};
bool g_scratch_done[NOALLOC_MAX_DEVICES] = {};

void noalloc_scratch_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
    g_scratch_done[device] = true;
}

int noalloc_scratch_ptr(int device, char **out) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return fail(TCW_ERR_INVALID, "tcw_map_noalloc: device " + std::to_string(device) +
                                         " beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_done[device])
        return fail(TCW_ERR_INVALID, "tcw_map_noalloc: no handle was created on device " +
                                         std::to_string(device));
    if (g_scratch_err[device] != hipSuccess || !g_scratch_ptr[device])
        return fail(TCW_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                        hipGetErrorString(g_scratch_err[device]));
    *out = (char *)g_scratch_ptr[device];
    return TCW_OK;
}

}  // namespace

namespace {

// Effective window and shape (tcw_fstat_map_funcs.py:742-780, on a copy).
int effective_window(uint32_t numAtoms, uint32_t TAtom, uint32_t t0_data,
                     const tcw_window &in, tcw_window &w, int64_t &N_t0, int64_t &N_tau) {
    w = in;
    if (w.type >= TCW_TRANSIENT_LAST)
        return fail(TCW_ERR_INVALID, "Unknown window-type (" + std::to_string(w.type) +
                                         ") passed as input. Allowed are [0," +
                                         std::to_string(TCW_TRANSIENT_LAST - 1) + "].");
    if (w.type == TCW_TRANSIENT_NONE) {
        w.type = TCW_TRANSIENT_RECTANGULAR;
        w.t0 = t0_data;
        w.t0Band = 0;
        w.dt0 = TAtom;
        w.tau = numAtoms * TAtom;
        w.tauBand = 0;
        w.dtau = TAtom;
    }
    if (w.dt0 == 0 || w.dtau == 0)
        return fail(TCW_ERR_INVALID, "windowRange.dt0 and windowRange.dtau must be positive.");
    N_t0 = (int64_t)(w.t0Band / w.dt0) + 1;
    N_tau = (int64_t)(w.tauBand / w.dtau) + 1;
    if (N_t0 * N_tau >= (int64_t)0x7fffffff)
        return fail(TCW_ERR_INVALID, "map of " + std::to_string(N_t0) + " x " +
                                         std::to_string(N_tau) + " cells exceeds 2^31-1 cells.");
    return TCW_OK;
}

// Atom-column products an exp_init wave does for (template, column block,
// part): the kernel's l range of the part, each column group counted up to its
// longest window (the order the waves are launched in: heaviest first).
int32_t init_work(const TmplDev &td, uint32_t cb, int part) {
    const int64_t k0 = (int64_t)td.k00 + (int64_t)td.a * ((int64_t)td.n_seg * td.rb_rows - 1);
    const int64_t room = (int64_t)td.H - k0;
    auto clip = [&](int64_t L) { return std::max<int64_t>(std::min<int64_t>(L, room), 0); };
    const int64_t col0 = (int64_t)cb * TILE_COLS;
    const int64_t col_last = std::min<int64_t>(col0 + TILE_COLS, td.N_tau) - 1;
    const int64_t lr_blk = clip((int64_t)td.L0 + (int64_t)td.Ls * col_last);
    const int64_t q = (lr_blk + INIT_PARTS - 1) / INIT_PARTS;
    const int64_t l0 = part * q, l1 = std::min<int64_t>(lr_blk, l0 + q);
    int64_t w = 0;
    for (int i = 0; i < INIT_GROUPS; i++) {
        const int64_t g0 = col0 + (int64_t)INIT_WG * i;
        if (g0 > col_last) break;
        const int64_t glast = std::min<int64_t>(g0 + INIT_WG - 1, col_last);
        const int64_t hi = clip((int64_t)td.L0 + (int64_t)td.Ls * glast);
        w += std::max<int64_t>(0, std::min(l1, hi) - l0);
    }
    return (int32_t)std::min<int64_t>(w, 0x7fffffff);
}

// Build one template's descriptor and items.  seg range and N/t_min precomputed.
int plan_template(tcw_plan &p, int64_t t, const tcw_window &win_in, uint32_t N,
                  uint32_t TAtom, uint32_t t_min, int64_t seg_begin, int64_t seg_end,
                  std::vector<ItemDev> (&items_by_mode)[MODE_COUNT]) {
    tcw_window w;
    int64_t N_t0, N_tau;
    int rc = effective_window(N, TAtom, t_min, win_in, w, N_t0, N_tau);
    if (rc) return rc;
    if (seg_end - seg_begin > SR_WG)  // the merge keeps one cursor per detector and thread
        return tcwc::fail(TCW_ERR_INVALID, "template " + std::to_string(t) + " has " +
                                           std::to_string(seg_end - seg_begin) +
                                           " detector segments; at most " +
                                           std::to_string(SR_WG) + " are supported.");
    TmplDev td;
    std::memset(&td, 0, sizeof(td));
    td.seg_begin = seg_begin;
    td.seg_end = seg_end;
    td.N = N;
    td.TAtom = TAtom;
    td.t_min = t_min;
    td.type = w.type;
    td.t0 = w.t0;
    td.dt0 = w.dt0;
    td.tau = w.tau;
    td.dtau = w.dtau;
    td.N_t0 = (uint32_t)N_t0;
    td.N_tau = (uint32_t)N_tau;
    td.o_t0 = win_in.t0;
    td.o_t0Band = win_in.t0Band;
    td.o_tau = win_in.tau;
    td.o_tauBand = win_in.tauBand;

    int mode;
    if (w.type == TCW_TRANSIENT_EXPONENTIAL) {
        mode = MODE_EXP_DIRECT;
        // on the atom grid and free of uint32 wrap-around: the recurrence path
        const uint64_t end = (uint64_t)w.t0 + w.t0Band + 3ull * ((uint64_t)w.tau + w.tauBand) +
                             TAtom;
        const bool on_grid = w.t0 >= t_min && (w.t0 - t_min) % TAtom == 0 &&
                             w.dt0 % TAtom == 0 && w.tau % TAtom == 0 &&
                             w.dtau % TAtom == 0 && w.tau >= TAtom;
        // (rows of at most 2^20 cells: the map kernel's row stores address a
        // pass of rows by 32-bit byte offsets)
        if (!(p.flags & TCW_FLAG_GENERIC) && on_grid && end < (1ull << 32) &&
            N_tau <= (int64_t)1 << 20) {
            mode = MODE_EXP_FAST;
            td.k00 = (int32_t)((w.t0 - t_min) / TAtom);
            td.a = (int32_t)(w.dt0 / TAtom);
            td.L0 = (int32_t)(3ull * w.tau / TAtom);
            td.Ls = (int32_t)(3ull * w.dtau / TAtom);
            const int64_t kmax = (int64_t)td.k00 + (int64_t)td.a * (N_t0 - 1);
            const int64_t Lmax = (int64_t)td.L0 + (int64_t)td.Ls * (N_tau - 1);
            td.H = (int32_t)std::min<int64_t>((int64_t)N, kmax + Lmax);
        }
    } else {
        mode = MODE_RECT_GENERIC;
        const uint64_t half = TAtom / 2;
        const bool regular = (w.dt0 % TAtom == 0) && (w.dtau % TAtom == 0) && (w.dt0 == w.dtau);
        const uint64_t x_lo = (uint64_t)w.t0 + half;  // must not be below t_min
        const uint64_t x_hi = (uint64_t)w.t0 + w.t0Band + w.tau + w.tauBand + half;
        // (rows of at most 2^20 cells: partial sub-tiles store by 32-bit byte
        // offsets within their four rows)
        if (!(p.flags & TCW_FLAG_GENERIC) && regular && x_lo >= t_min && x_hi < (1ull << 32) &&
            N_tau <= (int64_t)1 << 20) {
            mode = MODE_RECT_FAST;
            td.a = (int32_t)(w.dt0 / TAtom);
            td.q0 = (int32_t)(((uint64_t)w.t0 - t_min + half) / TAtom);
            td.q00 = (int32_t)(((uint64_t)w.t0 + w.tau - t_min + half) / TAtom);
        }
    }
    td.mode = mode;
    const int req = p.rows_req[mode];  // the plan's tuned rows, 0 = built-in choice
    int rb = req > 0 ? req : RB_DEFAULT[mode];
    if (mode == MODE_RECT_FAST && req <= 0) {
        // few or small maps (the single-map call): fewer rows per item, so
        // that the batch's items cover the CUs (about SMALL_BATCH_ITEMS of them
        // if the templates are alike)
        const double ncb = (double)((N_tau + TILE_COLS - 1) / TILE_COLS);
        const double want = (double)N_t0 * ncb * (double)p.T / (double)SMALL_BATCH_ITEMS;
        rb = want >= RB_DEFAULT[mode] ? RB_DEFAULT[mode]
                                      : std::max(16, ((int)std::ceil(want) + 15) / 16 * 16);
    }
    if (mode == MODE_EXP_FAST && req <= 0) {
        const int64_t pairs = p.T * ((N_tau + TILE_COLS - 1) / TILE_COLS);
        if (pairs < DIRECT_INIT_MIN_BLOCKS) {
            rb = EXP_SMALL_BATCH_ROWS;  // few column blocks: shorter chunks, more of them
        } else {
            // about EXP_TARGET_ITEMS work items: one chunk per column when the
            // batch has that many column blocks
            const int64_t chunks = std::max<int64_t>(1, (EXP_TARGET_ITEMS + pairs - 1) / pairs);
            const int64_t per = (N_t0 + chunks - 1) / chunks;
            rb = (int)std::min<int64_t>(MAX_RB_EXPF, (per + 15) / 16 * 16);
        }
    }
    if (mode == MODE_EXP_FAST)  // rows past the map are walked: no longer chunks than the map
        rb = (int)std::min<int64_t>(rb, (N_t0 + 15) / 16 * 16);
    td.rb_rows = rb;
    td.n_rowblk = (uint32_t)((N_t0 + rb - 1) / rb);
    td.n_colblk = (uint32_t)((N_tau + TILE_COLS - 1) / TILE_COLS);
    if (mode == MODE_EXP_DIRECT || mode == MODE_EXP_FAST) {
        td.mode = mode;  // exp_table_bins reads it
        if (mode == MODE_EXP_FAST) {
            // E at the top chunk's top row comes from a direct O(L) sum
            // (exp_init_kernel, slot n_seg of the table).  With enough
            // independent column blocks in the batch the chunks are just the
            // map's row blocks (least total work); otherwise extra chunks above
            // the map reach H, where E = 0 and the direct sum is empty, so that
            // no lane runs a long serial sum.
            int64_t nseg = std::max<int64_t>((int64_t)td.n_rowblk, 1);
            if (p.T * (int64_t)td.n_colblk < DIRECT_INIT_MIN_BLOCKS) {
                const int64_t steps = ((int64_t)td.H - td.k00 + td.a - 1) / td.a;  // rows to H
                nseg = std::max<int64_t>(nseg, (std::max<int64_t>(steps, 0) + rb) / rb);
            }
            td.n_seg = (int32_t)nseg;
            // the top chunk's top row lies at or beyond H: its direct sum is
            // empty, exp_scan starts from 0 and exp_init need not run
            td.init_empty = (int64_t)td.H - ((int64_t)td.k00 + (int64_t)td.a * (nseg * rb - 1)) <= 0;
            if (!td.init_empty) {
                p.exp_init_needed = true;
                for (uint32_t cbk = 0; cbk < td.n_colblk; cbk++)
                    for (int part = 0; part < INIT_PARTS; part++)
                        p.init_items.push_back(ItemDev{(int32_t)t, part, (int32_t)cbk,
                                                       init_work(td, cbk, part)});
            }
            // few (template, column block) pairs, each with many chunks: a
            // wave per column scans the chunks in log depth; otherwise a
            // thread per column walks them (coalesced, no idle lanes)
            if (p.T * (int64_t)td.n_colblk < DIRECT_INIT_MIN_BLOCKS && nseg >= 16)
                p.exp_scan_wave = true;
            td.etab_off = p.etab_doubles;
            p.etab_doubles += (nseg + INIT_PARTS) * NC * N_tau;
            for (int64_t sg = 1; sg < nseg; sg++)
                for (uint32_t cbk = 0; cbk < td.n_colblk; cbk++)
                    p.seg_items.push_back(ItemDev{(int32_t)t, (int32_t)sg, (int32_t)cbk, 0});
            if (nseg > 1)  // one chunk: the map kernel adds the direct sum's parts
                for (uint32_t cbk = 0; cbk < td.n_colblk; cbk++)
                    p.col_items.push_back(ItemDev{(int32_t)t, 0, (int32_t)cbk, 0});
            if (expw_pass_rows(td.a, td.Ls, (int)std::min<int64_t>(TILE_COLS, N_tau)) > 0)
                p.max_lds_expf = EXPF_LDS_BYTES;  // fixed plane strides
        }
        const int64_t nbins = exp_table_bins(td);
        td.merged_off = p.merged_floats;
        p.merged_floats += nbins * MERGED_STRIDE;
        p.exp_tmpls.push_back((int32_t)t);
        p.exp_max_N = std::max<int64_t>(p.exp_max_N, nbins);
    } else {
        td.prefix_off = p.prefix_doubles;
        p.prefix_doubles += ((int64_t)N + 1) * NC;
    }
    td.F_off = p.F_total;
    p.F_total += N_t0 * N_tau;
    td.rowpart_off = p.rowpart_n;
    p.rowpart_n += N_t0 * (int64_t)td.n_colblk;
    td.colpart_off = p.colpart_n;
    p.colpart_n += N_tau * (int64_t)td.n_rowblk;
    // items of a template are contiguous in item-id space; item_begin fixed below
    for (uint32_t rbk = 0; rbk < td.n_rowblk; rbk++)
        for (uint32_t cbk = 0; cbk < td.n_colblk; cbk++)
            items_by_mode[mode].push_back(ItemDev{(int32_t)t, (int32_t)rbk, (int32_t)cbk, 0});
    if (mode == MODE_RECT_FAST) {
        const int nd = rb + TILE_COLS - 1;
        const int bytes = (int)sizeof(double) * lds_doubles(nd);
        p.max_lds_bytes = std::max(p.max_lds_bytes, bytes);
    }
    p.tmpl.push_back(td);
    return TCW_OK;
}

// Every device buffer of a plan with the bytes the plan needs in it.
template <typename F>
void plan_buffers(tcw_plan &p, F f) {
    f(p.d_tmpl, sizeof(TmplDev) * std::max<size_t>(p.tmpl.size(), 1));
    f(p.d_segs, sizeof(SegDev) * std::max<size_t>(p.segs.size(), 1));
    f(p.d_items, sizeof(ItemDev) * std::max<size_t>(p.items.size(), 1));
    f(p.d_prefix, sizeof(double) * std::max<int64_t>(p.prefix_doubles, 1));
    f(p.d_merged, sizeof(float) * std::max<int64_t>(p.merged_floats, 1));
    f(p.d_ipart, sizeof(ItemPart) * std::max<size_t>(p.items.size(), 1));
    f(p.d_rowpart, sizeof(float) * std::max<int64_t>(p.rowpart_n, 1));
    f(p.d_colpart, sizeof(float) * std::max<int64_t>(p.colpart_n, 1));
    f(p.d_flags, sizeof(int) * std::max<int64_t>(p.T, 1));
    f(p.d_seg_items, sizeof(ItemDev) * std::max<size_t>(p.seg_items.size(), 1));
    f(p.d_col_items, sizeof(ItemDev) * std::max<size_t>(p.col_items.size(), 1));
    f(p.d_init_items, sizeof(ItemDev) * std::max<size_t>(p.init_items.size(), 1));
    f(p.d_etab, sizeof(double) * std::max<int64_t>(p.etab_doubles, 1));
    f(p.d_exp_tmpls, sizeof(int32_t) * std::max<size_t>(p.exp_tmpls.size(), 1));
}

// After all templates: lay out items mode by mode, fix item_begin (the items
// of one template stay contiguous, rb-major), size scratch (own buffers, or
// views into p.arena when set), upload descriptors (asynchronously on `st`
// when given, else synchronously).
int plan_finish(tcw_plan &p, std::vector<ItemDev> (&items_by_mode)[MODE_COUNT],
                hipStream_t st, bool async) {
    // exp_init waves heaviest first (work estimate in .pad): the last waves
    // to start are the short ones
    std::stable_sort(p.init_items.begin(), p.init_items.end(),
                     [](const ItemDev &x, const ItemDev &y) { return x.pad > y.pad; });
    p.items.clear();
    for (int md = 0; md < MODE_COUNT; md++) {
        p.mode_begin[md] = (int64_t)p.items.size();
        for (size_t k = 0; k < items_by_mode[md].size(); k++) {
            const ItemDev &it = items_by_mode[md][k];
            TmplDev &td = p.tmpl[it.t];
            if (it.rb == 0 && it.cb == 0) td.item_begin = (int64_t)p.items.size();
            p.items.push_back(it);
        }
    }
    p.mode_begin[MODE_COUNT] = (int64_t)p.items.size();
    int rc = TCW_OK;
    if (p.arena) {
        // carve every buffer out of the caller-provided arena (no allocation)
        size_t off = 0;
        plan_buffers(p, [&](DevBuf &b, size_t bytes) {
            const size_t a = (bytes + 255) / 256 * 256;
            if (off + a <= p.arena_bytes) b.set_view(p.arena + off, a);
            off += a;
        });
        if (off > p.arena_bytes)
            return fail(TCW_ERR_INVALID, "this map needs " + std::to_string(off) +
                                             " bytes of scratch; the static scratch holds " +
                                             std::to_string(p.arena_bytes) + " bytes");
    } else {
        plan_buffers(p, [&](DevBuf &b, size_t bytes) {
            if (!rc) rc = b.ensure(bytes);
        });
        if (rc) return rc;
    }
    auto up = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
        if (bytes == 0) return hipSuccess;
        return async ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st)
                     : hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    };
    TCW_HIP_TRY(up(p.d_tmpl.p, p.tmpl.data(), sizeof(TmplDev) * p.tmpl.size()));
    TCW_HIP_TRY(up(p.d_segs.p, p.segs.data(), sizeof(SegDev) * p.segs.size()));
    TCW_HIP_TRY(up(p.d_items.p, p.items.data(), sizeof(ItemDev) * p.items.size()));
    TCW_HIP_TRY(up(p.d_seg_items.p, p.seg_items.data(), sizeof(ItemDev) * p.seg_items.size()));
    TCW_HIP_TRY(up(p.d_col_items.p, p.col_items.data(), sizeof(ItemDev) * p.col_items.size()));
    TCW_HIP_TRY(up(p.d_init_items.p, p.init_items.data(), sizeof(ItemDev) * p.init_items.size()));
    TCW_HIP_TRY(up(p.d_exp_tmpls.p, p.exp_tmpls.data(), sizeof(int32_t) * p.exp_tmpls.size()));
    return TCW_OK;
}

// Rows per work item a caller may request for kernel mode `mode`.
int check_rows(int mode, int rows) {
    const int maxr = mode == MODE_EXP_FAST ? MAX_RB_EXPF : MAX_RB;
    if (rows != 0 && (rows < 16 || rows > maxr || rows % 16))
        return fail(TCW_ERR_INVALID, "rows per item for mode " + std::to_string(mode) +
                                         " must be 0 or a multiple of 16 in [16, " +
                                         std::to_string(maxr) + "], not " + std::to_string(rows));
    return TCW_OK;
}

// Plan `d` into `p` (buffers reused, grow-only).  rows_by_mode: host [4] or
// null (built-in choices).
int plan_build(tcw_plan &p, int device, const tcw_batch_desc *d, uint32_t flags,
               int64_t *F_off, hipStream_t st, bool async,
               const int32_t *rows_by_mode = nullptr) {
    if (d->T <= 0) return fail(TCW_ERR_INVALID, "no templates");
    if (d->TAtom < 2) return fail(TCW_ERR_INVALID, "TAtom must be at least 2 seconds.");
    for (int md = 0; md < MODE_COUNT; md++) {
        const int r = rows_by_mode ? rows_by_mode[md] : 0;
        int rc = check_rows(md, r);
        if (rc) return rc;
        p.rows_req[md] = r;
    }
    p.reset();
    p.device = device;
    p.flags = flags;
    p.T = d->T;
    std::vector<ItemDev> items_by_mode[MODE_COUNT];
    for (int64_t t = 0; t < d->T; t++) {
        const int64_t s0 = d->seg_begin[t], s1 = d->seg_begin[t + 1];
        if (s1 <= s0 || s1 - s0 > TCW_MAX_DETECTORS)
            return fail(TCW_ERR_INVALID, "template " + std::to_string(t) + ": needs 1.." +
                                             std::to_string(TCW_MAX_DETECTORS) +
                                             " detector segments");
        uint32_t tmin = 0xffffffffu, tmax = 0;
        const int64_t first_seg = (int64_t)p.segs.size();
        for (int64_t s = s0; s < s1; s++) {
            if (d->seg_off[s + 1] <= d->seg_off[s])
                return fail(TCW_ERR_INVALID, "template " + std::to_string(t) + ": empty detector");
            if (d->seg_tlast[s] < d->seg_tfirst[s])
                return fail(TCW_ERR_INVALID, "template " + std::to_string(t) +
                                                 ": timestamps not in ascending order");
            tmin = std::min(tmin, d->seg_tfirst[s]);
            tmax = std::max(tmax, d->seg_tlast[s]);
            p.segs.push_back(SegDev{d->seg_off[s], d->seg_off[s + 1]});
        }
        const uint64_t N = (uint64_t)(tmax - tmin) / d->TAtom + 1;
        if (F_off) F_off[t] = p.F_total;
        int rc = plan_template(p, t, d->windows[t], (uint32_t)N, d->TAtom, tmin, first_seg,
                               (int64_t)p.segs.size(), items_by_mode);
        if (rc) return rc;
    }
    if (F_off) F_off[d->T] = p.F_total;
    return plan_finish(p, items_by_mode, st, async);
}

template <bool BTSG, bool STRICT>
void launch_maps(tcw_plan &p, float *F, hipStream_t st) {
    auto *tds = (const TmplDev *)p.d_tmpl.p;
    auto *items = (const ItemDev *)p.d_items.p;
    auto *prefix = (const double *)p.d_prefix.p;
    auto *merged = (const float *)p.d_merged.p;
    auto *ipart = (ItemPart *)p.d_ipart.p;
    auto *rowpart = (float *)p.d_rowpart.p;
    auto *colpart = (float *)p.d_colpart.p;
    auto *flags = (int *)p.d_flags.p;
    int64_t n;
    if ((n = p.mode_begin[MODE_RECT_FAST + 1] - p.mode_begin[MODE_RECT_FAST]) > 0)
        hipLaunchKernelGGL((rect_fast_kernel<BTSG, STRICT>), dim3((unsigned)n), dim3(WG),
                           p.max_lds_bytes, st, tds, items, p.mode_begin[MODE_RECT_FAST],
                           prefix, F, ipart, rowpart, colpart, flags);
    if ((n = p.mode_begin[MODE_RECT_GENERIC + 1] - p.mode_begin[MODE_RECT_GENERIC]) > 0)
        hipLaunchKernelGGL((rect_generic_kernel<BTSG, STRICT>), dim3((unsigned)n), dim3(WG), 0,
                           st, tds, items, p.mode_begin[MODE_RECT_GENERIC], prefix, F, ipart,
                           rowpart, colpart, flags);
    if ((n = p.mode_begin[MODE_EXP_FAST + 1] - p.mode_begin[MODE_EXP_FAST]) > 0)
        hipLaunchKernelGGL((exp_fast_kernel<BTSG, STRICT>), dim3((unsigned)n), dim3(WG),
                           p.max_lds_expf, st, tds, items, p.mode_begin[MODE_EXP_FAST], merged,
                           (const double *)p.d_etab.p, F, ipart, rowpart, colpart, flags,
                           p.max_lds_expf > 0 ? 1 : 0);
    if ((n = p.mode_begin[MODE_EXP_DIRECT + 1] - p.mode_begin[MODE_EXP_DIRECT]) > 0)
        hipLaunchKernelGGL((exp_direct_kernel<BTSG, STRICT>), dim3((unsigned)n), dim3(WG), 0,
                           st, tds, items, p.mode_begin[MODE_EXP_DIRECT], merged, F, ipart,
                           rowpart, colpart, flags);
}

// Optional per-stage event marks: ev[0] before the stage kernel, ev[1] after
// it, ev[2] after the map kernels, ev[3] after finalize.
// check_errors = false while a stream is being captured (hipGetLastError is
// not permitted then; launch errors surface at instantiate/launch instead).
int launch_all(tcw_plan &p, const uint32_t *ts, const float *cols, int64_t cstride,
               float *F, tcw_stats *stats_dev, hipStream_t st, hipEvent_t *ev = nullptr,
               bool check_errors = true) {
    const bool btsg = p.flags & TCW_FLAG_BTSG;
    const bool strict = p.flags & TCW_FLAG_LAL_STRICT;
    auto *flags = (int *)p.d_flags.p;
    if (ev) TCW_HIP_TRY(hipEventRecord(ev[0], st));
    // (stage_exp_kernel clears the single-atom flags of the exp templates, so
    // a batch of exp templates only launches no stage_kernel)
    if ((int64_t)p.exp_tmpls.size() < p.T)
        hipLaunchKernelGGL(stage_kernel, dim3((unsigned)p.T), dim3(SR_WG), 0, st,
                           (const TmplDev *)p.d_tmpl.p, (const SegDev *)p.d_segs.p, ts, cols,
                           cstride, (double *)p.d_prefix.p, flags);
    if (!p.exp_tmpls.empty()) {
        const int64_t n_exp = (int64_t)p.exp_tmpls.size();
        const int64_t wide_blocks = (p.exp_max_N + 1023) / 1024;
        const unsigned gy = (unsigned)std::min<int64_t>(n_exp, 65535);
        if (n_exp > 0 && wide_blocks * n_exp < SX_NARROW_BELOW)
            hipLaunchKernelGGL(stage_exp_kernel<1>, dim3((unsigned)((p.exp_max_N + 255) / 256), gy),
                               dim3(256), 0, st, (const TmplDev *)p.d_tmpl.p,
                               (const int32_t *)p.d_exp_tmpls.p, n_exp, (const SegDev *)p.d_segs.p,
                               ts, cols, cstride, (float *)p.d_merged.p, flags);
        else if (n_exp > 0)
            hipLaunchKernelGGL(stage_exp_kernel<4>, dim3((unsigned)wide_blocks, gy), dim3(256), 0,
                               st, (const TmplDev *)p.d_tmpl.p, (const int32_t *)p.d_exp_tmpls.p,
                               n_exp, (const SegDev *)p.d_segs.p, ts, cols, cstride,
                               (float *)p.d_merged.p, flags);
    }
    // exp_init and exp_seg read the atom table and write different slots of the
    // E table: with enough blocks for both, exp_init runs on the side stream
    // exp_init runs one wave per item; blocks of 4 waves as the unit here
    const int64_t init_blocks = p.exp_init_needed ? (int64_t)p.init_items.size() / INIT_GROUPS : 0;
    const bool fork = p.side && init_blocks > 0 && !p.seg_items.empty() &&
                      init_blocks + (int64_t)p.seg_items.size() >= EXP_FORK_MIN_BLOCKS;
    if (fork) {
        TCW_HIP_TRY(hipEventRecord(p.ev_fork, st));
        TCW_HIP_TRY(hipStreamWaitEvent(p.side, p.ev_fork, 0));
    }
    if (p.exp_init_needed)
        hipLaunchKernelGGL(exp_init_kernel, dim3((unsigned)p.init_items.size()), dim3(INIT_WG), 0,
                           fork ? p.side : st, (const TmplDev *)p.d_tmpl.p,
                           (const ItemDev *)p.d_init_items.p, (const float *)p.d_merged.p,
                           (double *)p.d_etab.p);
    if (!p.seg_items.empty())
        hipLaunchKernelGGL(exp_seg_kernel, dim3((unsigned)p.seg_items.size()), dim3(WG),
                           p.max_lds_expf, st, (const TmplDev *)p.d_tmpl.p,
                           (const ItemDev *)p.d_seg_items.p, (const float *)p.d_merged.p,
                           (double *)p.d_etab.p, p.max_lds_expf > 0 ? 1 : 0);
    if (fork) {
        TCW_HIP_TRY(hipEventRecord(p.ev_join, p.side));
        TCW_HIP_TRY(hipStreamWaitEvent(st, p.ev_join, 0));
    }
    if (!p.col_items.empty() && !p.exp_scan_wave)
        hipLaunchKernelGGL(exp_scan_seq_kernel, dim3((unsigned)(p.col_items.size() * SCAN_SPLIT)),
                           dim3(SCAN_WG), 0, st, (const TmplDev *)p.d_tmpl.p,
                           (const ItemDev *)p.d_col_items.p, (double *)p.d_etab.p);
    if (!p.col_items.empty() && p.exp_scan_wave)
        hipLaunchKernelGGL(exp_scan_kernel,
                           dim3((unsigned)(p.col_items.size() * (TILE_COLS / SCAN_COLS))),
                           dim3(WG), 0, st,
                           (const TmplDev *)p.d_tmpl.p, (const ItemDev *)p.d_col_items.p,
                           (double *)p.d_etab.p);
    if (ev) TCW_HIP_TRY(hipEventRecord(ev[1], st));
    if (btsg) {
        if (strict)
            launch_maps<true, true>(p, F, st);
        else
            launch_maps<true, false>(p, F, st);
        if (ev) TCW_HIP_TRY(hipEventRecord(ev[2], st));
        if (p.T < FIN_WIDE_MAX_T)
            hipLaunchKernelGGL((finalize_kernel<true, 1024>), dim3((unsigned)p.T), dim3(1024), 0, st,
                           (const TmplDev *)p.d_tmpl.p, (const ItemPart *)p.d_ipart.p,
                           (const float *)p.d_rowpart.p, (const float *)p.d_colpart.p,
                           (const int *)p.d_flags.p, stats_dev);
        else
            hipLaunchKernelGGL((finalize_kernel<true, WG>), dim3((unsigned)p.T), dim3(WG), 0, st,
                           (const TmplDev *)p.d_tmpl.p, (const ItemPart *)p.d_ipart.p,
                           (const float *)p.d_rowpart.p, (const float *)p.d_colpart.p,
                           (const int *)p.d_flags.p, stats_dev);
    } else {
        if (strict)
            launch_maps<false, true>(p, F, st);
        else
            launch_maps<false, false>(p, F, st);
        if (ev) TCW_HIP_TRY(hipEventRecord(ev[2], st));
        if (p.T < FIN_WIDE_MAX_T)
            hipLaunchKernelGGL((finalize_kernel<false, 1024>), dim3((unsigned)p.T), dim3(1024), 0, st,
                           (const TmplDev *)p.d_tmpl.p, (const ItemPart *)p.d_ipart.p,
                           (const float *)p.d_rowpart.p, (const float *)p.d_colpart.p,
                           (const int *)p.d_flags.p, stats_dev);
        else
            hipLaunchKernelGGL((finalize_kernel<false, WG>), dim3((unsigned)p.T), dim3(WG), 0, st,
                           (const TmplDev *)p.d_tmpl.p, (const ItemPart *)p.d_ipart.p,
                           (const float *)p.d_rowpart.p, (const float *)p.d_colpart.p,
                           (const int *)p.d_flags.p, stats_dev);
    }
    if (ev) TCW_HIP_TRY(hipEventRecord(ev[3], st));
    if (check_errors) TCW_HIP_TRY(hipGetLastError());
    return TCW_OK;
}

// The exp side stream and its events, made on first use (never while a
// stream is being captured).
int ensure_side(tcw_plan *p) {
    if (p->side || p->col_items.empty() || p->seg_items.empty()) return TCW_OK;
    hipStream_t s = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&a, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (s) (void)hipStreamDestroy(s);
        if (a) (void)hipEventDestroy(a);
        return fail(TCW_ERR_DEVICE, std::string("exp side stream: ") + hipGetErrorString(e));
    }
    p->side = s;
    p->ev_fork = a;
    p->ev_join = b;
    return TCW_OK;
}

}  // namespace

extern "C" {

const char *tcw_version(void) { return "tcw_hip 0.1 (gfx950)"; }

const char *tcw_last_error(void) { return tcwc::g_err.c_str(); }

int tcw_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int tcw_device_name(int device, char *buf, int buflen) {
    if (!buf || buflen <= 0) return fail(TCW_ERR_INVALID, "tcw_device_name: no buffer");
    hipDeviceProp_t prop;
    TCW_HIP_TRY(hipGetDeviceProperties(&prop, device));
    std::snprintf(buf, (size_t)buflen, "%s", prop.name);
    return TCW_OK;
}

int tcw_map_shape(uint32_t numAtoms, uint32_t TAtom, uint32_t t0_data, const tcw_window *win,
                  int64_t *N_t0, int64_t *N_tau) {
    if (!win || !N_t0 || !N_tau) return fail(TCW_ERR_INVALID, "tcw_map_shape: null argument");
    tcw_window w;
    return effective_window(numAtoms, TAtom, t0_data, *win, w, *N_t0, *N_tau);
}

int tcw_create(int device, tcw_handle **out) {
    if (!out) return fail(TCW_ERR_INVALID, "tcw_create: null output");
    int n = tcw_device_count();
    if (device < 0 || device >= n)
        return fail(TCW_ERR_INVALID, "Requested HIP device number " + std::to_string(device) +
                                         " exceeds number of available devices (" +
                                         std::to_string(n) + ").");
    int rc = set_device(device);
    if (rc) return rc;
    auto *h = new tcw_handle;
    h->device = device;
    noalloc_scratch_setup(device);  // with `device` current
    if (hipMalloc(&h->d_stats, sizeof(tcw_stats)) != hipSuccess ||
        hipHostMalloc(&h->h_stats, sizeof(tcw_stats), hipHostMallocDefault) != hipSuccess) {
        delete h;
        return fail(TCW_ERR_NOMEM, "tcw_create: stats slot allocation failed");
    }
    *out = h;
    return TCW_OK;
}

int tcw_destroy(tcw_handle *h) {
    if (h) (void)set_device(h->device);
    delete h;
    return TCW_OK;
}

int tcw_batch_plan(tcw_handle *h, const tcw_batch_desc *d, uint32_t flags, tcw_plan **out,
                   int64_t *F_off) {
    return tcw_batch_plan_rows(h, d, flags, nullptr, out, F_off);
}

int tcw_batch_plan_rows(tcw_handle *h, const tcw_batch_desc *d, uint32_t flags,
                        const int32_t *rows_by_mode, tcw_plan **out, int64_t *F_off) {
    if (!h || !d || !out) return fail(TCW_ERR_INVALID, "tcw_batch_plan: null argument");
    int rc = set_device(h->device);
    if (rc) return rc;
    auto *p = new tcw_plan;
    rc = plan_build(*p, h->device, d, flags, F_off, nullptr, false, rows_by_mode);
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return TCW_OK;
}

int64_t tcw_plan_scratch_bytes(const tcw_plan *p) {
    if (!p) return 0;
    return (int64_t)(p->d_tmpl.cap + p->d_segs.cap + p->d_items.cap + p->d_prefix.cap +
                     p->d_merged.cap + p->d_ipart.cap + p->d_rowpart.cap + p->d_colpart.cap +
                     p->d_flags.cap + p->d_seg_items.cap + p->d_col_items.cap +
                     p->d_etab.cap + p->d_init_items.cap);
}

int tcw_plan_items(const tcw_plan *p, int64_t *items) {
    if (!p || !items) return fail(TCW_ERR_INVALID, "tcw_plan_items: null argument");
    for (int md = 0; md < MODE_COUNT; md++) items[md] = p->mode_begin[md + 1] - p->mode_begin[md];
    return TCW_OK;
}

int tcw_batch_run(tcw_handle *h, tcw_plan *p, const uint32_t *ts, const float *cols,
                  int64_t col_stride, float *F_out, tcw_stats *stats_dev, int use_graph,
                  void *stream) {
    if (!h || !p || !ts || !cols || !F_out || !stats_dev)
        return fail(TCW_ERR_INVALID, "tcw_batch_run: null argument");
    int rc = set_device(p->device);
    if (rc) return rc;
    if ((rc = ensure_side(p))) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (!use_graph) return launch_all(*p, ts, cols, col_stride, F_out, stats_dev, st);
    const bool same = p->graph && p->g_ts == ts && p->g_cols == cols && p->g_F == F_out &&
                      p->g_stats == stats_dev && p->g_cstride == col_stride;
    if (!same) {
        if (p->graph) {
            (void)hipGraphExecDestroy(p->graph);
            p->graph = nullptr;
        }
        // Capture on a private stream: the caller's stream may be the legacy
        // default stream, which cannot be captured.  The graph is then
        // launched on the caller's stream.
        if (!p->cap_stream)
            TCW_HIP_TRY(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        TCW_HIP_TRY(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        rc = launch_all(*p, ts, cols, col_stride, F_out, stats_dev, p->cap_stream, nullptr,
                        false);
        hipError_t e = hipStreamEndCapture(p->cap_stream, &g);  // always end the capture
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess)
            return fail(TCW_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            p->graph = nullptr;
            return fail(TCW_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        p->g_ts = ts;
        p->g_cols = cols;
        p->g_F = F_out;
        p->g_stats = stats_dev;
        p->g_cstride = col_stride;
    }
    TCW_HIP_TRY(hipGraphLaunch(p->graph, st));
    return TCW_OK;
}

int tcw_batch_run_timed(tcw_handle *h, tcw_plan *p, const uint32_t *ts, const float *cols,
                        int64_t col_stride, float *F_out, tcw_stats *stats_dev,
                        double *ms_out, void *stream) {
    if (!h || !p || !ts || !cols || !F_out || !stats_dev || !ms_out)
        return fail(TCW_ERR_INVALID, "tcw_batch_run_timed: null argument");
    int rc = set_device(p->device);
    if (rc) return rc;
    if ((rc = ensure_side(p))) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t ev[4];
    for (int k = 0; k < 4; k++) TCW_HIP_TRY(hipEventCreate(&ev[k]));
    rc = launch_all(*p, ts, cols, col_stride, F_out, stats_dev, st, ev);
    hipError_t e = hipSuccess;
    if (!rc) e = hipEventSynchronize(ev[3]);
    if (!rc && e == hipSuccess) {
        float a = 0, b = 0, c = 0;
        (void)hipEventElapsedTime(&a, ev[0], ev[1]);
        (void)hipEventElapsedTime(&b, ev[1], ev[2]);
        (void)hipEventElapsedTime(&c, ev[2], ev[3]);
        ms_out[0] = a;
        ms_out[1] = b;
        ms_out[2] = c;
    }
    for (int k = 0; k < 4; k++) (void)hipEventDestroy(ev[k]);
    if (rc) return rc;
    if (e != hipSuccess) return fail(TCW_ERR_DEVICE, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
    return TCW_OK;
}

int tcw_plan_destroy(tcw_plan *p) {
    delete p;
    return TCW_OK;
}

}  // extern "C"

namespace {

// One template as a batch of one (tcw_map and tcw_map_noalloc): checks the
// atoms and plans `p` on `st` (descriptor uploads queued on `st`).
int plan_single(tcw_handle *h, tcw_plan &p, const tcw_atoms *atoms, const tcw_window *win,
                uint32_t flags, hipStream_t st, const char *who) {
    if (atoms->numDet < 1 || atoms->numDet > TCW_MAX_DETECTORS)
        return fail(TCW_ERR_INVALID,
                    std::string(who) + ": numDet must be 1.." + std::to_string(TCW_MAX_DETECTORS));
    if (!atoms->ts || !atoms->cols) return fail(TCW_ERR_INVALID, std::string(who) + ": null atoms");
    if (atoms->t_max < atoms->t_min)
        return fail(TCW_ERR_INVALID, std::string(who) + ": t_max < t_min");
    int rc = set_device(h->device);
    if (rc) return rc;
    int64_t seg_begin[2] = {0, (int64_t)atoms->numDet};
    uint32_t tfirst[TCW_MAX_DETECTORS], tlast[TCW_MAX_DETECTORS];
    // only the overall min/max matter for the binned length
    for (uint32_t X = 0; X < atoms->numDet; X++) {
        tfirst[X] = atoms->t_min;
        tlast[X] = atoms->t_max;
    }
    tcw_batch_desc d;
    d.T = 1;
    d.TAtom = atoms->TAtom;
    d.seg_begin = seg_begin;
    d.seg_off = atoms->det_off;
    d.seg_tfirst = tfirst;
    d.seg_tlast = tlast;
    d.windows = win;
    return plan_build(p, h->device, &d, flags, nullptr, st, true);
}

}  // namespace

extern "C" {

int tcw_map(tcw_handle *h, const tcw_atoms *atoms, const tcw_window *win, float *F_mn_dev,
            tcw_stats *stats, uint32_t flags, void *stream) {
    if (!h || !atoms || !win || !F_mn_dev || !stats)
        return fail(TCW_ERR_INVALID, "tcw_map: null argument");
    hipStream_t st = (hipStream_t)stream;
    int rc = plan_single(h, h->single, atoms, win, flags, st, "tcw_map");
    if (rc) return rc;
    rc = launch_all(h->single, atoms->ts, atoms->cols, atoms->col_stride, F_mn_dev,
                    h->d_stats, st);
    if (rc) return rc;
    TCW_HIP_TRY(hipMemcpyAsync(h->h_stats, h->d_stats, sizeof(tcw_stats),
                               hipMemcpyDeviceToHost, st));
    TCW_HIP_TRY(hipStreamSynchronize(st));
    *stats = *h->h_stats;
    if ((flags & TCW_FLAG_LAL_STRICT) && stats->single_atom)
        return fail(TCW_ERR_SINGLE_ATOM,
                    "encountered a single-atom Fstat-calculation. This is degenerate and cannot "
                    "be computed! t0 must stay away *at least* 2*TAtom from the end of the data.");
    return TCW_OK;
}

int tcw_map_noalloc(tcw_handle *h, const tcw_atoms *atoms, const tcw_window *win,
                    float *F_mn_dev, tcw_stats *stats_dev, uint32_t flags, void *stream) {
    if (!h || !atoms || !win || !F_mn_dev || !stats_dev)
        return fail(TCW_ERR_INVALID, "tcw_map_noalloc: null argument");
    char *base = nullptr;
    int rc = noalloc_scratch_ptr(h->device, &base);
    if (rc) return rc;
    h->noalloc.arena = base;
    h->noalloc.arena_bytes = TCW_NOALLOC_SCRATCH_BYTES;
    hipStream_t st = (hipStream_t)stream;
    // the plan's buffers become views into the scratch (plan_finish), or the
    // call fails before anything is queued when the map needs more
    rc = plan_single(h, h->noalloc, atoms, win, flags, st, "tcw_map_noalloc");
    if (rc) return rc;
    // every scratch byte a kernel reads is written earlier in the same call
    // (the stage kernel also clears the single-atom flag), so nothing is cleared
    return launch_all(h->noalloc, atoms->ts, atoms->cols, atoms->col_stride, F_mn_dev,
                      stats_dev, st);
}

int tcw_merge_batch(tcw_handle *h, const tcw_batch_desc *d, const uint32_t *ts,
                    const float *cols, int64_t col_stride, float *merged, int64_t merged_stride,
                    int64_t *m_off, uint32_t *t_data, int64_t *atom_bin, void *stream) {
    if (!h || !d || !m_off || !t_data) return fail(TCW_ERR_INVALID, "tcw_merge_batch: null argument");
    if (d->T <= 0) return fail(TCW_ERR_INVALID, "no templates");
    if (d->TAtom < 2) return fail(TCW_ERR_INVALID, "TAtom must be at least 2 seconds.");
    std::vector<MergeTmpl> tab((size_t)d->T);
    std::vector<SegDev> segs;
    std::vector<MergeSeg> mseg;
    int64_t off = 0, maxN = 0;
    for (int64_t t = 0; t < d->T; t++) {
        const int64_t s0 = d->seg_begin[t], s1 = d->seg_begin[t + 1];
        if (s1 <= s0 || s1 - s0 > TCW_MAX_DETECTORS)
            return fail(TCW_ERR_INVALID, "template " + std::to_string(t) + ": needs 1.." +
                                             std::to_string(TCW_MAX_DETECTORS) +
                                             " detector segments");
        uint32_t tmin = 0xffffffffu, tmax = 0;
        for (int64_t s = s0; s < s1; s++) {
            if (d->seg_off[s + 1] <= d->seg_off[s])
                return fail(TCW_ERR_INVALID, "template " + std::to_string(t) + ": empty detector");
            if (d->seg_tlast[s] < d->seg_tfirst[s])
                return fail(TCW_ERR_INVALID, "template " + std::to_string(t) +
                                                 ": timestamps not in ascending order");
            tmin = std::min(tmin, d->seg_tfirst[s]);
            tmax = std::max(tmax, d->seg_tlast[s]);
        }
        const int64_t N = (int64_t)((tmax - tmin) / d->TAtom) + 1;
        tab[(size_t)t] = MergeTmpl{(int64_t)segs.size(), (int64_t)segs.size() + (s1 - s0), off,
                                   (uint32_t)N, tmin};
        for (int64_t s = s0; s < s1; s++) {
            segs.push_back(SegDev{d->seg_off[s], d->seg_off[s + 1]});
            mseg.push_back(MergeSeg{d->seg_off[s], d->seg_off[s + 1], off, tmin, 0});
        }
        m_off[t] = off;
        t_data[t] = tmin;
        off += N;
        maxN = std::max(maxN, N);
    }
    m_off[d->T] = off;
    if (!merged) return TCW_OK;
    if (!ts || !cols) return fail(TCW_ERR_INVALID, "tcw_merge_batch: null atoms");
    if (merged_stride < off)
        return fail(TCW_ERR_INVALID, "merged_stride " + std::to_string(merged_stride) +
                                         " < merged bins " + std::to_string(off));
    int64_t max_end = 0;
    for (const auto &g : segs) max_end = std::max(max_end, g.end);
    if (col_stride < max_end) return fail(TCW_ERR_INVALID, "col_stride smaller than the atom count");
    int rc = set_device(h->device);
    if (rc) return rc;
    if ((rc = h->d_mtab.ensure(tab.size() * sizeof(MergeTmpl))) ||
        (rc = h->d_msegs.ensure(segs.size() * sizeof(SegDev))) ||
        (rc = h->d_mseg.ensure(mseg.size() * sizeof(MergeSeg))))
        return rc;
    hipStream_t st = (hipStream_t)stream;
    // the previous call's kernels may still read the tables
    TCW_HIP_TRY(hipStreamSynchronize(st));
    TCW_HIP_TRY(hipMemcpy(h->d_mtab.p, tab.data(), tab.size() * sizeof(MergeTmpl),
                          hipMemcpyHostToDevice));
    TCW_HIP_TRY(hipMemcpy(h->d_msegs.p, segs.data(), segs.size() * sizeof(SegDev),
                          hipMemcpyHostToDevice));
    TCW_HIP_TRY(hipMemcpy(h->d_mseg.p, mseg.data(), mseg.size() * sizeof(MergeSeg),
                          hipMemcpyHostToDevice));
    const unsigned gy = (unsigned)std::min<int64_t>(d->T, 65535);
    hipLaunchKernelGGL(merge_kernel, dim3((unsigned)((maxN + 255) / 256), gy), dim3(256), 0, st,
                       (const MergeTmpl *)h->d_mtab.p, (const SegDev *)h->d_msegs.p, d->T, ts,
                       cols, col_stride, d->TAtom, merged, merged_stride);
    if (atom_bin) {
        int64_t maxS = 0;
        for (const auto &g : mseg) maxS = std::max(maxS, g.end - g.begin);
        const unsigned sy = (unsigned)std::min<size_t>(mseg.size(), 65535);
        const unsigned sx = (unsigned)std::min<int64_t>((maxS + 255) / 256, 1024);
        hipLaunchKernelGGL(atom_bin_kernel, dim3(sx, sy), dim3(256), 0, st,
                           (const MergeSeg *)h->d_mseg.p, (int64_t)mseg.size(), ts, d->TAtom,
                           atom_bin);
    }
    TCW_HIP_TRY(hipGetLastError());
    TCW_HIP_TRY(hipStreamSynchronize(st));
    return TCW_OK;
}

int tcw_gather_bins(const float *in, int64_t in_stride, const int64_t *atom_bin, int64_t n_atoms,
                    float *out, int64_t out_stride, void *stream) {
    if (n_atoms < 0 || out_stride < n_atoms)
        return fail(TCW_ERR_INVALID, "tcw_gather_bins: bad sizes");
    if (n_atoms == 0) return TCW_OK;
    if (!in || !atom_bin || !out) return fail(TCW_ERR_INVALID, "tcw_gather_bins: null argument");
    const unsigned g = (unsigned)std::min<int64_t>((n_atoms + 255) / 256, 65536);
    hipLaunchKernelGGL(gather_bins_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, in,
                       in_stride, atom_bin, n_atoms, out, out_stride);
    TCW_HIP_TRY(hipGetLastError());
    return TCW_OK;
}

}  // extern "C"
