// Banded Cholesky factorisation and substitutions of the reduced camera system
// for the reassociating precisions (FP64-fast, mixed).  They replace
// matrix::decompose_cholesky / matrix::solve_cholesky
// (source/matrix_decomposition_cholesky.hpp:26-95, called at
// factor_graph.hpp:782-793) when the reduced system is a narrow band: row r of S
// is zero left of column first[r], and band = max_r (r - first[r] + 1) is at
// most kBandMaxW.  Cholesky fill stays inside that row profile (L_rj = 0 for
// j < first[r]), so the factor is a band of the same width.
//
// One workgroup of four waves.  The band lives in LDS as a window of rows: row r
// in slot r mod kRing holds its entries (r, c) for c = r - band + 1 .. r at
// offsets 0 .. band - 1, its right-hand side at offset band, and a scratch entry
// at band + 1 that masked lanes write instead of branching.  Per panel of
// kBandPanel columns c0 ..:
//   phase 1  wave 0 holds the panel's rows in registers (lane l: rows c0 + l and
//            c0 + 64 + l) and factors it right-looking: pivot d, y = 1/sqrt(d)
//            (the hardware estimate and two Newton steps), L_rj = a_rj y, the
//            remaining panel columns updated with L_cj broadcast by readlane;
//            the forward substitution runs along (y_j = b_j y, b_r -= L_rj y_j).
//            A short last panel is padded with identity columns.  Meanwhile
//            waves 1-3 store the previous panel's rows (the factor as a compact
//            band Lb: row r's entries at Lb[r (band + 1) + 0 .. band - 1] and
//            1 / L_rr at Lb[r (band + 1) + band]; the forward solution to x) and
//            load the rows the next panel reaches, every load of a thread issued
//            before its first store, so that memory latency hides behind the
//            factor;
//   phase 2  all waves update the trailing triangle of the window with the
//            panel (4 x 4 blocks, operands from the transposed panel LpT).
// Then the backward substitution L^T x = y runs over 64-row blocks from the
// bottom, each block's rows of Lb staged in LDS one block ahead by waves 1-3:
// wave 0 runs the block's recurrence, then all waves apply the block's solved
// values to the rows above it.
//
// Failure rules are the reference's: the decomposition fails when a pivot is not
// greater than 1e-14 max|S_ii| (NaN fails), the solve when some |L_ii| is not
// greater than 1e-7 max|L_ii|.  With floor_rel > 0 every diagonal entry is first
// raised by floor_rel max|S_ii| (in the window; S itself is only read).  Entries
// of S's lower triangle outside the row profile must be zero (the reduced-system
// kernels clear S before forming it).
//
// Selected by factor path 3 only: on the 100-keyframe workload (n = 600, band 72)
// the single-wave recurrences cost 85-175 cycles per step, and the factor plus
// both substitutions take 0.41 ms against 0.32 ms for the tiled one-workgroup
// factorisation and its substitution.
#include "common.hpp"

#include <cmath>
#include <cstdint>

namespace zb {
namespace {

constexpr int kP = kBandPanel;
constexpr int kWmax = kBandMaxW;
constexpr int kStMax = kWmax + 2;      // window slot stride: entries, right-hand side, scratch
constexpr int kRing = 3 * kP + kWmax;  // rows resident: previous panel, this panel's reach, next panel's
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kXRing = 256;            // solved / pending values of the substitutions (>= 64 + kWmax)
constexpr int kLpN = 136;              // panel rows kept for the trailing update
constexpr int kMaxT = (kWmax + 2) / 4; // 4-row blocks of the largest trailing triangle
constexpr int kMaxTiles = kMaxT * (kMaxT + 1) / 2;
constexpr int kBlk = 64;               // substitution block rows
constexpr int kGroup = 8;              // recurrence steps whose entries load together
constexpr int kBlkBuf = kBlk * (kWmax + 1) + 1;  // one substitution block of Lb rows
static_assert(kP + kWmax - 1 <= 128, "the panel's rows fit two rows per lane of one wave");
static_assert(kXRing >= kBlk + kWmax, "the substitution ring holds a block and the band above it");
static_assert(kP + 4 * kMaxT <= kLpN, "trailing blocks read panel rows inside LpT");
static_assert(2 * kBlkBuf <= kRing * kStMax, "two substitution blocks fit in the window's memory");

__device__ __forceinline__ double lane_value(double v, int j) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), j);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), j);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// 1 / sqrt(d) within about an ulp for positive normal d: the hardware estimate
// (relative error near 2^-23), then two Newton steps y += y (1 - d y^2) / 2.
__device__ __forceinline__ double rsqrt_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = fma(-(d * y), y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}

__device__ __forceinline__ double dmax_nan(double a, double b) { return (a < b) ? b : a; }

// The lane index, hidden from loop-invariant code motion: the per-step lane masks
// of the unrolled recurrences are formed where they are used instead of being
// hoisted out and kept live in scalar registers (which spills them).
__device__ __forceinline__ int opaque_lane() {
    int lane = static_cast<int>(threadIdx.x & 63);
    asm volatile("" : "+v"(lane));
    return lane;
}

// Copies cnt doubles (element q: ld(q), then stv(q, value)) with kBatch loads of
// a thread issued before its first store; threads tid0 .. tid0 + nthr - 1 take
// part.
template <int kBatch, class Ld, class St>
__device__ __forceinline__ void copy_batched(int cnt, int tid0, int nthr, Ld ld, St stv) {
    for (int q0 = static_cast<int>(threadIdx.x) - tid0; q0 < cnt; q0 += kBatch * nthr) {
        double v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int q = q0 + u * nthr;
            v[u] = (q < cnt) ? ld(q) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int q = q0 + u * nthr;
            if (q < cnt) stv(q, v[u]);
        }
    }
}

// Rows [lo, hi) of the compact band Lb into a substitution block buffer (same
// layout, stride wb + 1) and their right-hand sides v[lo ..] into vb.
__device__ __forceinline__ void load_block(double* buf, double* vb, int lo, int hi, int wb, const double* Lb,
                                           const double* v, int tid0, int nthr) {
    const double* src = Lb + static_cast<int64_t>(lo) * (wb + 1);
    copy_batched<8>((hi - lo) * (wb + 1), tid0, nthr, [&](int q) { return src[q]; },
                    [&](int q, double val) { buf[q] = val; });
    for (int i = static_cast<int>(threadIdx.x) - tid0; i < hi - lo; i += nthr) vb[i] = v[lo + i];
}

// Backward substitution L^T x = y over blocks of kBlk rows from the bottom.  On
// entry blk[(nb - 1) & 1] and vbk[(nb - 1) & 1] hold the last block (load_block
// with y); the other blocks come from Lb and x (= y).  x receives the solution.
// blk0, blk1: kBlkBuf each; xr, ya: rings of kXRing.
__device__ __forceinline__ void band_backward(int n, int wb, const double* Lb, double* x, double* blk0, double* blk1,
                                              double (*vbk)[kBlk], double* xr, double* ya, double (*part)[64]) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int st = wb + 1;
    const int nb = (n + kBlk - 1) / kBlk;
    double* flat = &part[0][0];
    for (int b = nb - 1; b >= 0; --b) {
        const int i0 = b * kBlk;
        const int mb = (n - i0) < kBlk ? (n - i0) : kBlk;
        double* cur = (b & 1) ? blk1 : blk0;
        double* nxt = (b & 1) ? blk0 : blk1;
        if (wave == 0) {
            const int i = i0 + lane;
            const bool real = lane < mb;
            // pending updates from the blocks below reach rows i >= i0 + kBlk + 1 - wb
            const bool touched = (b < nb - 1) && (i >= i0 + kBlk + 1 - wb);
            double yv = real ? vbk[b & 1][lane] - (touched ? ya[i % kXRing] : 0.0) : 0.0;
            const double dl = real ? cur[lane * st + wb] : 0.0;
            // step k: x_k from lane k, then y_i -= L_{i0+k, i} x_k for i < k (row i0 + k
            // of the block); the next kGroup steps' entries load during these
            auto entry = [&](int k) {
                const int ln = opaque_lane();
                const int off = ln - k + wb - 1;
                const bool in = k < mb && ln < k && off >= 0;
                const double t = cur[k * st + (in ? off : 0)];
                return in ? t : 0.0;
            };
            double tc[kGroup], tn[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) tc[u] = entry(kBlk - 1 - u);
#pragma unroll
            for (int g = 0; g < kBlk / kGroup; ++g) {
#pragma unroll
                for (int u = 0; u < kGroup; ++u)
                    tn[u] = (g + 1 < kBlk / kGroup) ? entry(kBlk - 1 - (g + 1) * kGroup - u) : 0.0;
#pragma unroll
                for (int u = 0; u < kGroup; ++u) {
                    const int k = kBlk - 1 - g * kGroup - u;
                    const int ln = opaque_lane();
                    const double xk = lane_value(yv * dl, k);
                    yv = (ln == k) ? xk : fma(-tc[u], xk, yv);
                }
#pragma unroll
                for (int u = 0; u < kGroup; ++u) tc[u] = tn[u];
            }
            if (real) {
                x[i] = yv;
                xr[i % kXRing] = yv;
            }
        } else if (b > 0) {
            load_block(nxt, vbk[(b - 1) & 1], i0 - kBlk, i0, wb, Lb, x, 64, kThreads - 64);
        }
        __syncthreads();
        if (b == 0) break;
        // rows i in [max(0, i0 - wb + 1), i0): pending += sum over block rows r of L_ri x_r
        {
            const int t = tid & 127, h = tid >> 7;
            const int i = i0 - 1 - t;
            double acc = 0.0;
            if (i >= 0 && i >= i0 - wb + 1) {
                const int rlo = i0 + 32 * h, rhi = (i0 + 32 * h + 32) < i0 + mb ? (i0 + 32 * h + 32) : i0 + mb;
                for (int r = rlo; r < rhi; ++r) {
                    const int off = i - r + wb - 1;
                    if (off >= 0) acc = fma(cur[(r - i0) * st + off], xr[r % kXRing], acc);
                }
            }
            flat[tid] = acc;
        }
        __syncthreads();
        if (tid < 128) {
            const int i = i0 - 1 - tid;
            if (i >= 0 && i >= i0 - wb + 1) {
                const double s = flat[tid] + flat[128 + tid];
                const bool fresh = (b == nb - 1) || (i < i0 + kBlk + 1 - wb);
                ya[i % kXRing] = fresh ? s : ya[i % kXRing] + s;
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kThreads) k_chol_band(int n, int wb, const double* A, const double* b, double* x,
                                                        double* Lb, double floor_rel, double* eps_out,
                                                        int32_t* status, lm_state* lm, int gated) {
    if (gated && lm->done) return;
    __shared__ double ring[kRing * kStMax];
    __shared__ __attribute__((aligned(16))) double LpT[kP][kLpN];
    __shared__ double xr[kXRing], ya[kXRing];
    __shared__ double part[kWaves][64];
    __shared__ double vbk[2][kBlk];
    __shared__ int tiles[kMaxTiles];
    __shared__ double dstage[2][kP];
    __shared__ int flags[2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int st = wb + 2, sb = wb + 1;
    auto slot = [&](int r) -> double* { return ring + (r % kRing) * st; };
    double* flat = &part[0][0];

    // max |S_ii| (NaN never wins), the floor and the pivot threshold; the
    // trailing blocks' (I, J) table; the panel copy cleared
    double m = 0.0;
    for (int i = tid; i < n; i += kThreads) m = dmax_nan(m, fabs(A[static_cast<int64_t>(i) * n + i]));
    flat[tid] = m;
    for (int q = tid; q < kMaxTiles; q += kThreads) {
        int I = 0;
        while ((I + 1) * (I + 2) / 2 <= q) ++I;
        tiles[q] = (I << 8) | (q - I * (I + 1) / 2);
    }
    for (int q = tid; q < kP * kLpN; q += kThreads) (&LpT[0][0])[q] = 0.0;
    if (tid == 0) flags[0] = 0;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (tid < w) flat[tid] = dmax_nan(flat[tid], flat[tid + w]);
        __syncthreads();
    }
    const double mx = flat[0];
    const double lift = floor_rel > 0.0 ? floor_rel * mx : 0.0;
    const double eps = (mx + lift) * 1e-14;
    if (tid == 0) {
        eps_out[0] = eps;
        status[0] = 0;
    }
    // rows [lo, hi) from S into the window (threads tid0 ..)
    auto load_rows = [&](int lo, int hi, int tid0, int nthr) {
        copy_batched<4>(
            (hi - lo) * st, tid0, nthr,
            [&](int q) {
                const int i = q / st, o = q - i * st, r = lo + i;
                const int c = r - wb + 1 + o;
                double v = 0.0;
                if (o < wb) {
                    v = (c >= 0) ? A[static_cast<int64_t>(r) * n + c] : 0.0;
                    if (c == r) v += lift;
                } else if (o == wb) {
                    v = b[r];
                }
                return v;
            },
            [&](int q, double v) {
                const int i = q / st;
                slot(lo + i)[q - i * st] = v;
            });
    };
    // rows [lo, hi) of the factor to Lb with their 1 / L_ii (from dstage[s]), their
    // forward solution to x
    auto store_rows = [&](int lo, int hi, int s, int tid0, int nthr) {
        double* dst = Lb + static_cast<int64_t>(lo) * sb;
        const int cnt = (hi - lo) * sb;
        for (int q = tid - tid0; q < cnt; q += nthr) {
            const int i = q / sb, o = q - i * sb;
            dst[q] = (o < wb) ? slot(lo + i)[o] : dstage[s][i];
        }
        for (int i = tid - tid0; i < hi - lo; i += nthr) x[lo + i] = slot(lo + i)[wb];
    };
    int loaded = (kP + wb - 1) < n ? (kP + wb - 1) : n;
    load_rows(0, loaded, 0, kThreads);
    __syncthreads();

    bool ok = true;  // wave 0
    double lmax = 0.0, lmin = INFINITY;
    int k = 0;
    for (int c0 = 0; c0 < n; c0 += kP, ++k) {
        const int nk = (n - c0) < kP ? (n - c0) : kP;
        const int re = (c0 + nk + wb - 1) < n ? (c0 + nk + wb - 1) : n;  // rows reaching the panel
        const int want = (c0 + 2 * kP + wb - 1) < n ? (c0 + 2 * kP + wb - 1) : n;
        if (wave == 0) {
            const int r0 = c0 + lane, r1 = r0 + 64;
            const bool v0 = r0 < re, v1 = r1 < re;
            double* R0 = slot(v0 ? r0 : c0);
            double* R1 = slot(v1 ? r1 : c0);
            double a0[kP], a1[kP];
            int o0[kP], o1[kP];
#pragma unroll
            for (int j = 0; j < kP; ++j) {
                const int off0 = j - lane + wb - 1, off1 = off0 - 64;
                const bool in0 = v0 && off0 >= 0 && j <= lane, in1 = v1 && off1 >= 0;
                o0[j] = in0 ? off0 : wb + 1;
                o1[j] = in1 ? off1 : wb + 1;
                const double t0 = R0[o0[j]], t1 = R1[o1[j]];
                // a short last panel: identity rows and columns past n
                a0[j] = in0 ? t0 : ((r0 >= n && lane == j) ? 1.0 : 0.0);
                a1[j] = in1 ? t1 : 0.0;
            }
            double b0 = v0 ? R0[wb] : 0.0, b1 = v1 ? R1[wb] : 0.0;
            double yv = 0.0;
#pragma unroll
            for (int j = 0; j < kP; ++j) {
                const double d = lane_value(a0[j], j);
                ok &= (j >= nk) | (d > eps);
                const double y = rsqrt_nr(d);
                const double l = d * y;
                if (j < nk) {
                    lmax = dmax_nan(lmax, fabs(l));
                    lmin = fmin(lmin, fabs(l));
                }
                a0[j] = (lane == j) ? l : ((lane > j) ? a0[j] * y : 0.0);
                a1[j] = a1[j] * y;
                const double yj = lane_value(b0, j) * y;
                b0 = (lane == j) ? yj : ((lane > j) ? fma(-a0[j], yj, b0) : b0);
                b1 = fma(-a1[j], yj, b1);
                yv = (lane == j) ? y : yv;
#pragma unroll
                for (int jj = j + 1; jj < kP; ++jj) {
                    const double t = lane_value(a0[j], jj);
                    a0[jj] = (lane >= jj) ? fma(-a0[j], t, a0[jj]) : a0[jj];
                    a1[jj] = fma(-a1[j], t, a1[jj]);
                }
            }
#pragma unroll
            for (int j = 0; j < kP; ++j) {
                R0[o0[j]] = a0[j];
                R1[o1[j]] = a1[j];
                LpT[j][lane] = a0[j];
                LpT[j][64 + lane] = a1[j];
            }
            R0[v0 ? wb : wb + 1] = b0;
            R1[v1 ? wb : wb + 1] = b1;
            if (lane < kP) dstage[k & 1][lane] = yv;
            if (lane == 0 && !ok) flags[0] = 1;
        } else {
            // waves 1-3: the previous panel's rows out, the next panel's rows in
            if (c0 > 0) store_rows(c0 - kP, c0, (k - 1) & 1, 64, kThreads - 64);
            if (want > loaded) load_rows(loaded, want, 64, kThreads - 64);
        }
        loaded = want > loaded ? want : loaded;
        __syncthreads();
        if (flags[0]) break;
        // trailing triangle: rows and columns t0 .. re - 1
        const int t0 = c0 + nk, mt = re - t0;
        if (mt > 0) {
            const int T = (mt + 3) / 4, ntile = T * (T + 1) / 2;
            for (int q = tid; q < ntile; q += kThreads) {
                const int packed = tiles[q];
                const int i0 = 4 * (packed >> 8), j0 = 4 * (packed & 255);
                double* Rw[4];
                int off[4][4];
                double acc[4][4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int i = i0 + a;
                    Rw[a] = slot(t0 + (i < mt ? i : mt - 1));
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const int j = j0 + c;
                        off[a][c] = (i < mt && j <= i) ? j - i + wb - 1 : wb + 1;
                        acc[a][c] = Rw[a][off[a][c]];
                    }
                }
#pragma unroll
                for (int kk = 0; kk < kP; ++kk) {
                    const double2 li0 = *reinterpret_cast<const double2*>(&LpT[kk][nk + i0]);
                    const double2 li1 = *reinterpret_cast<const double2*>(&LpT[kk][nk + i0 + 2]);
                    const double2 lj0 = *reinterpret_cast<const double2*>(&LpT[kk][nk + j0]);
                    const double2 lj1 = *reinterpret_cast<const double2*>(&LpT[kk][nk + j0 + 2]);
                    const double li[4] = {li0.x, li0.y, li1.x, li1.y};
                    const double lj[4] = {lj0.x, lj0.y, lj1.x, lj1.y};
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int c = 0; c < 4; ++c) acc[a][c] = fma(-li[a], lj[c], acc[a][c]);
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int c = 0; c < 4; ++c) Rw[a][off[a][c]] = acc[a][c];
            }
        }
        __syncthreads();
    }
    if (wave == 0 && lane == 0) flags[1] = (ok && !(lmin > 1e-7 * lmax)) ? 1 : 0;
    // the last panel's rows out
    if (!flags[0] && k > 0) store_rows((k - 1) * kP, n, (k - 1) & 1, 0, kThreads);
    __syncthreads();
    if (flags[0] || flags[1]) {
        if (tid == 0) {
            status[0] = flags[0] ? 1 : 2;
            lm->solve_ok = 0;
        }
        return;
    }
    // backward substitution; the block buffers take the window's memory
    double* blk0 = ring;
    double* blk1 = ring + kBlkBuf;
    const int nb = (n + kBlk - 1) / kBlk;
    load_block(((nb - 1) & 1) ? blk1 : blk0, vbk[(nb - 1) & 1], (nb - 1) * kBlk, n, wb, Lb, x, 0, kThreads);
    __syncthreads();
    band_backward(n, wb, Lb, x, blk0, blk1, vbk, xr, ya, part);
    if (tid == 0) lm->solve_ok = 1;
}

// x = S^-1 b with the factor of k_chol_band in Lb: the forward substitution over
// blocks of kBlk rows from the top (rows staged one block ahead), then
// band_backward.
__global__ void __launch_bounds__(kThreads) k_band_solve(int n, int wb, const double* Lb, const double* b, double* x,
                                                         const int32_t* status, lm_state* lm, int gated) {
    if (gated && lm->done) return;
    if (*status != 0) {
        if (threadIdx.x == 0) lm->solve_ok = 0;
        return;
    }
    __shared__ double blkmem[2 * kBlkBuf];
    __shared__ double vbk[2][kBlk];
    __shared__ double xr[kXRing], ya[kXRing];
    __shared__ double part[kWaves][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int st = wb + 1;
    double* blk0 = blkmem;
    double* blk1 = blkmem + kBlkBuf;
    const int nb = (n + kBlk - 1) / kBlk;
    load_block(blk0, vbk[0], 0, (n < kBlk ? n : kBlk), wb, Lb, b, 0, kThreads);
    __syncthreads();
    for (int bi = 0; bi < nb; ++bi) {
        const int i0 = bi * kBlk;
        const int mb = (n - i0) < kBlk ? (n - i0) : kBlk;
        double* cur = (bi & 1) ? blk1 : blk0;
        double* nxt = (bi & 1) ? blk0 : blk1;
        // rows r of the block: partial = sum over j < i0 of L_rj y_j (a quarter of j per wave)
        {
            const int r = i0 + lane;
            double acc = 0.0;
            if (lane < mb && i0 > 0) {
                const int lo = (r - wb + 1) > 0 ? (r - wb + 1) : 0;
                for (int j = lo + wave; j < i0; j += kWaves)
                    acc = fma(cur[lane * st + (j - r + wb - 1)], xr[j % kXRing], acc);
            }
            part[wave][lane] = acc;
        }
        __syncthreads();
        if (wave == 0) {
            const int r = i0 + lane;
            const bool real = lane < mb;
            const double* row = cur + (real ? lane : 0) * st;
            double bv = real ? vbk[bi & 1][lane] - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]))
                             : 0.0;
            const double dl = real ? row[wb] : 0.0;
            // step k: y_k from lane k, then b_r -= L_{r, i0+k} y_k for r > k; the next
            // kGroup steps' entries load during these
            auto entry = [&](int k) {
                const int ln = opaque_lane();
                const int off = k - ln + wb - 1;
                const bool in = real && k < ln && off >= 0;
                const double t = row[in ? off : 0];
                return in ? t : 0.0;
            };
            double tc[kGroup], tn[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u) tc[u] = entry(u);
#pragma unroll
            for (int g = 0; g < kBlk / kGroup; ++g) {
#pragma unroll
                for (int u = 0; u < kGroup; ++u) tn[u] = (g + 1 < kBlk / kGroup) ? entry((g + 1) * kGroup + u) : 0.0;
#pragma unroll
                for (int u = 0; u < kGroup; ++u) {
                    const int k = g * kGroup + u;
                    const int ln = opaque_lane();
                    const double yk = lane_value(bv * dl, k);
                    bv = (ln == k) ? yk : fma(-tc[u], yk, bv);
                }
#pragma unroll
                for (int u = 0; u < kGroup; ++u) tc[u] = tn[u];
            }
            if (real) {
                x[r] = bv;
                xr[r % kXRing] = bv;
                vbk[bi & 1][lane] = bv;  // the backward pass starts from this block
            }
        } else if (bi + 1 < nb) {
            const int hi = (i0 + 2 * kBlk) < n ? (i0 + 2 * kBlk) : n;
            load_block(nxt, vbk[(bi + 1) & 1], i0 + kBlk, hi, wb, Lb, b, 64, kThreads - 64);
        }
        __syncthreads();
    }
    band_backward(n, wb, Lb, x, blk0, blk1, vbk, xr, ya, part);
    if (tid == 0) lm->solve_ok = 1;
}

}  // namespace

bool band_path(int32_t n, const dev_work64& w, bool exact) {
    if (exact || n <= 0 || (n & 1) || w.band <= 0 || w.band > kBandMaxW || !w.Lt) return false;
    return w.factor_path == 3;
}

void launch_chol_band(int32_t n, const dev_work64& w, hipStream_t s, bool gated, double floor_rel) {
    k_chol_band<<<1, kThreads, 0, s>>>(n, w.band, w.S, w.bs, w.dxp, w.Lt, floor_rel, w.chol_scale, w.chol_status, w.lm,
                                       gated ? 1 : 0);
}

void launch_band_solve(int32_t n, const dev_work64& w, const double* b, double* x, hipStream_t s, bool gated) {
    k_band_solve<<<1, kThreads, 0, s>>>(n, w.band, w.Lt, b, x, w.chol_status, w.lm, gated ? 1 : 0);
}

}  // namespace zb
