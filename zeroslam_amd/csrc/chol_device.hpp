// Workgroup-level dense Cholesky routines shared by the per-graph kernels
// (kernels_chol.hip) and the batched per-window solver (kernels_batch.hip).
// Every routine is called by all threads of one workgroup (blockDim.x a multiple
// of 64, at most 1024) and synchronises with workgroup barriers only.
//
// The factor follows the reference's row-Crout sequence
// (matrix_decomposition_cholesky.hpp:26-59): every element receives its
// a -= L_ik L_jk subtractions in ascending k and its division / square root last.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace zb {
namespace wg {

constexpr int kTile = 32;
constexpr int NB = kTile;

// Lane-uniform read of lane j's value.
__device__ __forceinline__ double lane_value(double v, int j) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), j);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), j);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// The lane index, hidden from loop-invariant code motion: values derived from it
// inside a loop (lane masks, per-lane constants) are then formed where they are
// used instead of being hoisted out and kept live, which spills registers.
__device__ __forceinline__ int opaque_lane() {
    int lane = static_cast<int>(threadIdx.x & 63);
    asm volatile("" : "+v"(lane));
    return lane;
}

// Block-wide maximum of one value per thread (NaN never wins).
__device__ inline double block_max(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) {
            const double a = sh[t + w];
            sh[t] = (sh[t] < a) ? a : sh[t];
        }
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Correctly rounded a / b from y = RN(1 / b): q = RN(a y) is within an ulp of
// a / b, e = a - q b is exact (one FMA) and RN(q + e y) is RN(a / b)
// (Markstein's correction; also checked against the division on 5e8 operand
// pairs).  A zero numerator gives q = a y, the signed zero a / b has for b > 0.
// Outside the range where the correction is proven (non-finite or
// near-subnormal quotients) it divides.
__device__ __forceinline__ double div_cr(double a, double b, double y) {
    const double q = a * y;
    const double e = fma(-q, b, a);
    double r = fma(e, y, q);
    if (a == 0.0) r = q;
    else if (!(fabs(r) < 1e300) || fabs(r) < 1e-290) r = a / b;
    return r;
}

// div_cr without its fallback, for straight-line code: sets bad where the
// fallback would have divided, so that the caller can redo the work with div_cr.
__device__ __forceinline__ double div_cr_nofb(double a, double b, double y, bool& bad) {
    const double q = a * y;
    const double e = fma(-q, b, a);
    const double r = fma(e, y, q);
    const double ar = fabs(r);
    bad |= (a != 0.0) & ((!(ar < 1e300)) | (ar < 1e-290));
    return (a == 0.0) ? q : r;
}

// Row-Crout factorisation of a 32 x 32 tile held one row per lane (lane r holds
// row r; rows at or past nr are identity rows, which leave the others untouched:
// their entries in real columns are zero and real rows are zero in theirs; lanes
// 32..63 repeat rows 0..31 and are never read).  Column j: every lane reads the
// pivot A_jj, forms L_jj = sqrt, its reciprocal and its own L_rj, then subtracts
// L_rj L_cj from its entries c > j.  L_{j+1,j}, which the next pivot needs, comes
// from lane j + 1 by readlane; the other L_cj are broadcast through col[j] in
// LDS (col[j][c] = L_cj for c >= j: the factor transposed), so no column's
// broadcasts pile up in scalar registers.  No branches: bad
// is set where a division needs div_cr's fallback.  recips[j] receives
// RN(1 / L_jj).  Returns false (uniformly) when a pivot of a real row is not
// greater than eps.
template <int kChunk>
__device__ __forceinline__ bool factor_rows(double (&row)[NB], int nr, double eps, double* recips, bool& bad,
                                            double (*col)[64]) {
    const int lane = opaque_lane();
    bool ok = true;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const double d = lane_value(row[j], j);
        ok &= (j >= nr) | (d > eps);
        const double l = sqrt(d);
        const double y = 1.0 / l;
        const double q = div_cr_nofb(row[j], l, y, bad);
        row[j] = (lane == j) ? l : q;
        recips[j] = y;
        col[j][lane] = row[j];
        if (j + 1 < NB) {
            row[j + 1] = row[j + 1] - row[j] * lane_value(row[j], j + 1);
#pragma unroll
            for (int c0 = j + 2; c0 < NB; c0 += kChunk) {
                double t[kChunk];  // the column's broadcasts, kChunk loaded together before any use
#pragma unroll
                for (int u = 0; u < kChunk; ++u)
                    if (c0 + u < NB) t[u] = col[j][c0 + u];
#pragma unroll
                for (int u = 0; u < kChunk; ++u)
                    if (c0 + u < NB) row[c0 + u] = row[c0 + u] - row[j] * t[u];
                if (kChunk < NB) __builtin_amdgcn_sched_barrier(0);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    return ok;
}

// factor_rows for the reassociating precisions: 1 / sqrt(A_jj) from the hardware
// estimate and two Newton steps, L_jj = A_jj / sqrt(A_jj), L_rj = A_rj / L_jj as a
// product with that reciprocal, updates as fused multiply-adds.  recips[j]
// receives the reciprocal.
template <int kChunk>
__device__ __forceinline__ bool factor_rows_fast(double (&row)[NB], int nr, double eps, double* recips,
                                                 double (*col)[64]) {
    const int lane = opaque_lane();
    bool ok = true;
    if (kChunk == NB) {
        // lagged updates: column j - 1's broadcasts, loaded during step j - 1, are
        // applied during step j, so their LDS round trip overlaps step j's pivot
        // chain; entry j gets column j - 1's update first, by readlane.  Each
        // element still receives its updates in ascending column order.
        double tp[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j >= 1) row[j] = fma(-row[j - 1], lane_value(row[j - 1], j), row[j]);
            const double d = lane_value(row[j], j);
            ok &= (j >= nr) | (d > eps);
            double y = __builtin_amdgcn_rsq(d);
            y = y * fma(-0.5 * d * y, y, 1.5);
            y = y * fma(-0.5 * d * y, y, 1.5);
            row[j] = (lane == j) ? d * y : row[j] * y;
            recips[j] = y;
            col[j][lane] = row[j];
            if (j >= 1) {
#pragma unroll
                for (int c = j + 1; c < NB; ++c) row[c] = fma(-row[j - 1], tp[c], row[c]);
            }
#pragma unroll
            for (int c = j + 2; c < NB; ++c) tp[c] = col[j][c];
            __builtin_amdgcn_sched_barrier(0);
        }
        // the last column's update has no entries left to reach
        return ok;
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const double d = lane_value(row[j], j);
        ok &= (j >= nr) | (d > eps);
        double y = __builtin_amdgcn_rsq(d);
        y = y * fma(-0.5 * d * y, y, 1.5);
        y = y * fma(-0.5 * d * y, y, 1.5);
        row[j] = (lane == j) ? d * y : row[j] * y;
        recips[j] = y;
        col[j][lane] = row[j];
        if (j + 1 < NB) {
            row[j + 1] = fma(-row[j], lane_value(row[j], j + 1), row[j + 1]);
#pragma unroll
            for (int c0 = j + 2; c0 < NB; c0 += kChunk) {
                double t[kChunk];
#pragma unroll
                for (int u = 0; u < kChunk; ++u)
                    if (c0 + u < NB) t[u] = col[j][c0 + u];
#pragma unroll
                for (int u = 0; u < kChunk; ++u)
                    if (c0 + u < NB) row[c0 + u] = fma(-row[j], t[u], row[c0 + u]);
                if (kChunk < NB) __builtin_amdgcn_sched_barrier(0);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    return ok;
}

// The same factorisation in place in LDS with div_cr (fallback included), one
// row per lane of wave 0: the rare path when factor_rows reports bad.
__device__ __noinline__ bool factor_tile_lds(double (*D)[NB + 1], int nr, double eps, double* recips) {
    const int lane = threadIdx.x & 63;
    bool ok = true;
    for (int j = 0; j < nr; ++j) {
        const double d = D[j][j];
        ok = ok && (d > eps);
        const double l = sqrt(d);
        const double y = 1.0 / l;
        if (lane == j) {
            D[j][j] = l;
            recips[j] = y;
        } else if (lane > j && lane < nr) {
            D[lane][j] = div_cr(D[lane][j], l, y);
        }
        if (lane > j && lane < nr) {
            for (int c = j + 1; c <= lane; ++c) D[lane][c] = D[lane][c] - D[lane][j] * D[c][j];
        }
    }
    return ok;
}

// Factorisation of a diagonal tile (in LDS, lower triangle, zero above) by wave
// 0; rows at or past nr are overwritten with identity rows.  recips[j] receives
// RN(1 / L_jj); col is NB x 64 doubles of LDS scratch.  Returns false
// (uniformly) when a pivot is not greater than eps.
__device__ __forceinline__ bool factor_diag_tile(double (*D)[NB + 1], int nr, double eps, int* ok_sh, double* recips,
                                                 double (*col)[64]) {
    if (threadIdx.x < 64) {
        const int lane = opaque_lane();
        const int lr = lane & (NB - 1);
        if (nr < NB && lane >= nr && lane < NB) {
#pragma unroll
            for (int c = 0; c < NB; ++c) D[lane][c] = 0.0;
            D[lane][lane] = 1.0;
        }
        double row[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) row[c] = D[lr][c];
        bool bad = false;
        bool ok = factor_rows<NB>(row, nr, eps, recips, bad, col);
        if (!__any(bad)) {
            if (lane < nr) {
#pragma unroll
                for (int c = 0; c < NB; ++c) D[lane][c] = row[c];
            }
        } else {
            ok = factor_tile_lds(D, nr, eps, recips);
        }
        if (lane == 0) *ok_sh = ok ? 1 : 0;
    }
    __syncthreads();
    return *ok_sh != 0;
}

// factor_diag_tile for the reassociating precisions (factor_rows_fast); recips[j]
// receives 1 / L_jj.
__device__ __forceinline__ bool factor_diag_tile_fast(double (*D)[NB + 1], int nr, double eps, int* ok_sh,
                                                      double* recips, double (*col)[64]) {
    if (threadIdx.x < 64) {
        const int lane = opaque_lane();
        const int lr = lane & (NB - 1);
        if (nr < NB && lane >= nr && lane < NB) {
#pragma unroll
            for (int c = 0; c < NB; ++c) D[lane][c] = 0.0;
            D[lane][lane] = 1.0;
        }
        double row[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) row[c] = D[lr][c];
        const bool ok = factor_rows_fast<NB>(row, nr, eps, recips, col);
        if (lane < nr) {
#pragma unroll
            for (int c = 0; c < NB; ++c) D[lane][c] = row[c];
        }
        if (lane == 0) *ok_sh = ok ? 1 : 0;
    }
    __syncthreads();
    return *ok_sh != 0;
}

// Row i of a panel below the diagonal tile: L_ic = (A_ic - sum_{m<c} L_im L_cm) / L_cc
// in ascending c, subtractions in ascending m (the row-Crout order), divisions by
// the reciprocals of the diagonal (div_cr, correctly rounded).  The row is solved
// in place in LDS; only full 32-column panels have rows below their diagonal
// tile.
__device__ __noinline__ void solve_panel_row_slow(double* row, const double (*D)[NB + 1]) {
    for (int c = 0; c < NB; ++c) {
        double v = row[c];
        for (int m = 0; m < c; ++m) v = v - row[m] * D[c][m];
        row[c] = v / D[c][c];
    }
}

__device__ __forceinline__ void solve_panel_row(double* row, const double (*D)[NB + 1], const double* recips) {
    double x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = row[c];
    bool bad = false;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        double v = x[c];
#pragma unroll
        for (int m = 0; m < c; ++m) v = v - x[m] * D[c][m];
        x[c] = div_cr_nofb(v, D[c][c], recips[c], bad);
        __builtin_amdgcn_sched_barrier(0);  // keep the LDS reads of later columns from being hoisted
    }
    if (!bad) {
#pragma unroll
        for (int c = 0; c < NB; ++c) row[c] = x[c];
    } else {
        solve_panel_row_slow(row, D);
    }
}

// solve_panel_row for the reassociating precisions: fused multiply-adds and
// products with the reciprocals of the diagonal.
__device__ __forceinline__ void solve_panel_row_fast(double* row, const double (*D)[NB + 1], const double* recips) {
    double x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = row[c];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        double v = x[c];
#pragma unroll
        for (int m = 0; m < c; ++m) v = fma(-x[m], D[c][m], v);
        x[c] = v * recips[c];
        asm volatile("" ::: "memory");  // keep later columns' LDS reads from being hoisted
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) row[c] = x[c];
}

// ---- whole factorisation in one workgroup -------------------------------------
// Per 32-column panel k (rows k0 .. the last row tile of the envelope in column
// k, already updated by the panels before it):
//   1. wave 0 loads the diagonal tile into registers and factors it
//      (factor_rows: L_kk transposed in col, reciprocals of the diagonal in
//      recips), writing L_kk back; meanwhile every other thread loads the row
//      below the tile it will solve;
//   2. each row below is solved right-looking against col: x_m = a_m / L_mm
//      (div_cr), then a_c -= x_m L_cm for c > m, so each element receives its
//      subtractions in ascending m and its division last; the row goes back to
//      memory and, transposed, into LDS (PT, column q of the panel tile contiguous
//      in the row);
//   3. the trailing tiles (I, J), k < J <= I, are updated in memory from PT,
//      each thread holding a 2 x 4 block: a -= L_Iq L_Jq for q ascending.
// Workgroup barriers order the phases; nothing waits on another workgroup.
// Every panel's rows below its diagonal tile need a thread of waves 1..: the
// caller keeps (tallest panel - 1) x 32 <= blockDim.x - 64 (returns false
// otherwise).
// kFast (the reassociating precisions): reciprocal square roots, products with
// reciprocals and fused multiply-adds instead of the reference's operation
// sequence (factor_rows_fast).
// Returns false (uniformly) when a pivot fails.  PT holds kTiles - 1 transposed
// tiles; every column panel of the envelope must be at most kTiles tiles tall.
constexpr int kPT = NB + 2;  // PT row stride: 16-byte aligned rows

// A row below the diagonal tile whose fast solve needed div_cr's fallback:
// solved again in memory with divisions.
__device__ __noinline__ void solve_row_slow(double* arow, const double (*col)[64]) {
    for (int c = 0; c < NB; ++c) {
        double v = arow[c];
        for (int m = 0; m < c; ++m) v = v - arow[m] * col[m][c];
        arow[c] = v / col[c][c];
    }
}

template <int kTiles, bool kFast, int kChunk = NB>
__device__ bool fused_factor(int n, double* A, const int32_t* tile_last, double eps, double (*PT)[NB][kPT],
                             int* ok_sh, double* recips, double (*col)[64]) {
    const int T = (n + NB - 1) / NB;
    const int tid = threadIdx.x;
    const int rt = tid - 64;  // row-solve thread index (waves 1..)
    auto rows_below = [&](int k) {  // rows below diagonal tile k in its panel (full tiles only)
        const int k0 = k * NB, nk = (n - k0) < NB ? (n - k0) : NB;
        const int rows_end = (tile_last[k] + 1) * NB < n ? (tile_last[k] + 1) * NB : n;
        return rows_end - k0 - nk;
    };
    // Wave 0: diagonal tile k from memory into registers, factored (col: L_kk
    // transposed, recips: reciprocals of its diagonal), written back; ok_sh[0] =
    // the pivots passed.  Exact path: when a division needs div_cr's fallback it
    // returns true without writing back, for diag_redo.
    auto diag = [&](int k, double (&v)[NB]) -> bool {
        const int k0 = k * NB, nk = (n - k0) < NB ? (n - k0) : NB;
        const int lane = opaque_lane();
        const int lr = lane & (NB - 1);
        const bool real = lr < nk;
        const double* src = A + static_cast<int64_t>(k0 + (real ? lr : 0)) * n + k0;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            double a = 0.0;
            if (real && c <= lr) a = src[c];
            v[c] = real ? a : ((c == lr) ? 1.0 : 0.0);
        }
        bool bad = false;
        const bool ok = kFast ? factor_rows_fast<kChunk>(v, nk, eps, recips, col)
                              : factor_rows<kChunk>(v, nk, eps, recips, bad, col);
        if (!kFast && __any(bad)) return true;
        if (lane < nk) {
            double* dst = A + static_cast<int64_t>(k0 + lane) * n + k0;
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c <= lane) dst[c] = v[c];
        }
        if (lane == 0) ok_sh[0] = ok ? 1 : 0;
        return false;
    };
    // Wave 0: the rare redo of diagonal tile k in LDS (PT[0] as scratch) with
    // true divisions, from the unfactored rows still in memory.
    auto diag_redo = [&](int k, double (&v)[NB]) {
        const int k0 = k * NB, nk = (n - k0) < NB ? (n - k0) : NB;
        const int lane = opaque_lane();
        const int lr = lane & (NB - 1);
        const bool real = lr < nk;
        const double* src = A + static_cast<int64_t>(k0 + (real ? lr : 0)) * n + k0;
        double (*D)[NB + 1] = reinterpret_cast<double (*)[NB + 1]>(&PT[0][0][0]);
        if (lane < NB) {
#pragma unroll
            for (int c = 0; c < NB; ++c) D[lane][c] = (real && c <= lr) ? src[c] : ((c == lr) ? 1.0 : 0.0);
        }
        const bool ok = factor_tile_lds(D, nk, eps, recips);
        if (lane < NB) {
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                v[c] = D[lane][c];
                col[c][lane] = D[lane][c];
            }
        }
        if (lane < nk) {
            double* dst = A + static_cast<int64_t>(k0 + lane) * n + k0;
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c <= lane) dst[c] = v[c];
        }
        if (lane == 0) ok_sh[0] = ok ? 1 : 0;
    };
    // Block b (2 x 4 entries) of trailing tile (k + 1 + Ip, k + 1 + Jp): a -= L_Iq L_Jq
    // over panel k's columns q in ascending order, from PT.
    auto tile_block = [&](int k, int rows_end, int Ip, int Jp, int b) {
        const int k0 = k * NB;
        const int r0 = (b >> 3) * 2, c0 = (b & 7) * 4;
        if (Ip == Jp && c0 > r0 + 1) return;  // above the diagonal
        const int gr = k0 + NB * (1 + Ip) + r0, gc = k0 + NB * (1 + Jp) + c0;
        double acc[2][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const bool in = (Ip != Jp || c0 + c <= r0 + a) && gr + a < rows_end && gc + c < rows_end;
                acc[a][c] = in ? A[static_cast<int64_t>(gr + a) * n + gc + c] : 0.0;
            }
        const double (*LI)[kPT] = PT[Ip];
        const double (*LJ)[kPT] = PT[Jp];
#pragma unroll 1
        for (int q0 = 0; q0 < NB; q0 += 4) {
            double li[4][2], lj[4][4];  // four steps' operands, loaded before any use
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                li[u][0] = LI[q0 + u][r0];
                li[u][1] = LI[q0 + u][r0 + 1];
#pragma unroll
                for (int c = 0; c < 4; ++c) lj[u][c] = LJ[q0 + u][c0 + c];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        acc[a][c] = kFast ? fma(-li[u][a], lj[u][c], acc[a][c]) : acc[a][c] - li[u][a] * lj[u][c];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if ((Ip != Jp || c0 + c <= r0 + a) && gr + a < rows_end && gc + c < rows_end)
                    A[static_cast<int64_t>(gr + a) * n + gc + c] = acc[a][c];
    };

    double v[NB];  // wave 0: the diagonal tile's rows; waves 1..: a row below it
    {
        const int nb0 = rows_below(0);
        if (nb0 > static_cast<int>(blockDim.x) - 64) return false;  // callers keep panels within the row threads
        if (tid < 64) {
            if (diag(0, v)) diag_redo(0, v);
        } else if (rt < nb0) {
            const double* src = A + static_cast<int64_t>(NB + rt) * n;
#pragma unroll
            for (int c = 0; c < NB; ++c) v[c] = src[c];
        }
        __syncthreads();
        if (ok_sh[0] == 0) return false;
    }
    // Per panel k (diagonal tile factored, rows below loaded): the rows are solved
    // (waves 1..); panel k's update goes first to column k + 1 (all threads); then
    // wave 0 factors diagonal tile k + 1 while waves 1.. load the rows below it
    // and update the rest of the trailing tiles.  Every element receives the same
    // operations in the same order as without the overlap.
    for (int k = 0; k < T; ++k) {
        const int k0 = k * NB;
        const int nk = (n - k0) < NB ? (n - k0) : NB;
        const int rows_end = (tile_last[k] + 1) * NB < n ? (tile_last[k] + 1) * NB : n;
        const int nbelow = rows_end - k0 - nk;
        if (tid >= 64 && rt < nbelow) {
            bool bad = false;
            if (kChunk == NB) {
                // column m + 1 of L_kk (and its reciprocal) loaded while column m is applied
                double t[NB], tn[NB];
                double y = recips[0], yn = 0.0;
#pragma unroll
                for (int c = 0; c < NB; ++c) t[c] = col[0][c];
#pragma unroll
                for (int m = 0; m < NB; ++m) {
                    if (m + 1 < NB) {
#pragma unroll
                        for (int c = m + 1; c < NB; ++c) tn[c] = col[m + 1][c];
                        yn = recips[m + 1];
                    }
                    v[m] = kFast ? v[m] * y : div_cr_nofb(v[m], t[m], y, bad);
#pragma unroll
                    for (int c = m + 1; c < NB; ++c) v[c] = kFast ? fma(-v[m], t[c], v[c]) : v[c] - v[m] * t[c];
#pragma unroll
                    for (int c = m + 1; c < NB; ++c) t[c] = tn[c];
                    y = yn;
                    __builtin_amdgcn_sched_barrier(0);
                }
            } else {
#pragma unroll
                for (int m = 0; m < NB; ++m) {
                    v[m] = kFast ? v[m] * recips[m] : div_cr_nofb(v[m], col[m][m], recips[m], bad);
#pragma unroll
                    for (int c0 = m + 1; c0 < NB; c0 += kChunk) {
                        double t[kChunk];  // column m of L_kk, kChunk loaded together before any use
#pragma unroll
                        for (int u = 0; u < kChunk; ++u)
                            if (c0 + u < NB) t[u] = col[m][c0 + u];
#pragma unroll
                        for (int u = 0; u < kChunk; ++u)
                            if (c0 + u < NB) v[c0 + u] = kFast ? fma(-v[m], t[u], v[c0 + u]) : v[c0 + u] - v[m] * t[u];
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            double* arow = A + static_cast<int64_t>(k0 + nk + rt) * n + k0;
            if (bad) {
                solve_row_slow(arow, col);
#pragma unroll
                for (int c = 0; c < NB; ++c) v[c] = arow[c];
            } else {
#pragma unroll
                for (int c = 0; c < NB; ++c) arow[c] = v[c];
            }
            const int t = rt >> 5, rr = rt & (NB - 1);
#pragma unroll
            for (int c = 0; c < NB; ++c) PT[t][c][rr] = v[c];
        }
        __syncthreads();
        if (k + 1 >= T) break;
        const int m = (nbelow + NB - 1) / NB;
        for (int bidx = tid; bidx < m * 128; bidx += blockDim.x) tile_block(k, rows_end, bidx >> 7, 0, bidx & 127);
        __syncthreads();
        const int nb1 = rows_below(k + 1);
        if (nb1 > static_cast<int>(blockDim.x) - 64) return false;
        if (tid < 64) {
            const bool redo = diag(k + 1, v);
            if (!kFast && opaque_lane() == 0) ok_sh[1] = redo ? 1 : 0;
        } else {
            if (rt < nb1) {
                const double* src = A + static_cast<int64_t>((k + 2) * NB + rt) * n + (k + 1) * NB;
#pragma unroll
                for (int c = 0; c < NB; ++c) v[c] = src[c];
            }
            // trailing tiles with 1 <= Jp <= Ip < m
            const int pairs = m * (m - 1) / 2;
            for (int bidx = rt; bidx < pairs * 128; bidx += static_cast<int>(blockDim.x) - 64) {
                const int pr = bidx >> 7;
                int Ip = 1;
                while (Ip * (Ip + 1) / 2 <= pr) ++Ip;
                tile_block(k, rows_end, Ip, 1 + pr - (Ip - 1) * Ip / 2, bidx & 127);
            }
        }
        __syncthreads();
        if (!kFast && ok_sh[1]) {
            if (tid < 64) diag_redo(k + 1, v);
            __syncthreads();
        }
        if (ok_sh[0] == 0) return false;
    }
    return true;
}

// Forward (and unless forward_only, backward) substitution with the factor L.
// Per diagonal tile, wave 0 runs the recurrence with x_j broadcast by readlane
// and the divisions by L_jj taken through reciprocals formed in parallel
// (div_cr, correctly rounded); the rows the tile touches are then updated by all
// threads.  Forward substitution subtracts L_ij x_j in ascending j, the
// reference's sequence (matrix_decomposition_cholesky.hpp:76-84).  Returns 0, or
// 2 when some |L_ii| is not greater than 1e-7 max |L_ii| (the reference's solve
// failure, matrix_decomposition_cholesky.hpp:70-74).  x may be LDS or global.
__device__ inline int triangular_solve(int n, const double* L, const double* b, double* x, const int32_t* tile_first,
                                       const int32_t* tile_last, double (*T)[NB + 1], double* sh, int* flag,
                                       bool forward_only) {
    const int tid = threadIdx.x;
    double m = 0;
    for (int i = tid; i < n; i += blockDim.x) {
        const double a = fabs(L[static_cast<int64_t>(i) * n + i]);
        m = (m < a) ? a : m;
    }
    const double mx = block_max(m, sh);
    if (tid == 0) *flag = 0;
    __syncthreads();
    const double eps = mx * 1e-7;
    for (int i = tid; i < n; i += blockDim.x) {
        if (!(fabs(L[static_cast<int64_t>(i) * n + i]) > eps)) *flag = 1;
        x[i] = b[i];
    }
    __syncthreads();
    if (*flag) return 2;
    const int tiles = (n + NB - 1) / NB;
    const int lane = tid & 63;
    for (int t = 0; t < tiles; ++t) {
        const int i0 = t * NB, nb = (n - i0) < NB ? (n - i0) : NB;
        for (int idx = tid; idx < nb * nb; idx += blockDim.x) {
            const int r = idx / nb, c = idx % nb;
            T[r][c] = (c <= r) ? L[static_cast<int64_t>(i0 + r) * n + i0 + c] : 0.0;
        }
        __syncthreads();
        if (tid < 64) {
            double xi = (lane < nb) ? x[i0 + lane] : 0.0;
            const double dl = (lane < nb) ? T[lane][lane] : 1.0;
            const double yl = 1.0 / dl;
            for (int j = 0; j < nb; ++j) {
                if (lane == j) xi = div_cr(xi, dl, yl);
                const double xj = lane_value(xi, j);
                if (lane > j) xi = xi - T[lane < NB ? lane : 0][j] * xj;
            }
            if (lane < nb) x[i0 + lane] = xi;
        }
        __syncthreads();
        const int iend = (tile_last[t] + 1) * NB < n ? (tile_last[t] + 1) * NB : n;
        for (int i = i0 + nb + tid; i < iend; i += blockDim.x) {
            // only full tiles have rows below them (nb == NB); operands loaded before any use
            const double* Li = L + static_cast<int64_t>(i) * n + i0;
            double v = x[i];
#pragma unroll
            for (int h = 0; h < NB; h += NB / 2) {
                double lv[NB / 2], xv[NB / 2];
#pragma unroll
                for (int j = 0; j < NB / 2; ++j) {
                    lv[j] = Li[h + j];
                    xv[j] = x[i0 + h + j];
                }
#pragma unroll
                for (int j = 0; j < NB / 2; ++j) v = v - lv[j] * xv[j];
                __builtin_amdgcn_sched_barrier(0);
            }
            x[i] = v;
        }
        __syncthreads();
    }
    if (forward_only) return 0;
    for (int t = tiles - 1; t >= 0; --t) {
        const int i0 = t * NB, nb = (n - i0) < NB ? (n - i0) : NB;
        for (int idx = tid; idx < nb * nb; idx += blockDim.x) {
            const int r = idx / nb, c = idx % nb;
            T[r][c] = (c <= r) ? L[static_cast<int64_t>(i0 + r) * n + i0 + c] : 0.0;
        }
        __syncthreads();
        if (tid < 64) {
            double xi = (lane < nb) ? x[i0 + lane] : 0.0;
            const double dl = (lane < nb) ? T[lane][lane] : 1.0;
            const double yl = 1.0 / dl;
            for (int j = nb - 1; j >= 0; --j) {
                if (lane == j) xi = div_cr(xi, dl, yl);
                const double xj = lane_value(xi, j);
                if (lane < j) xi = xi - T[j][lane] * xj;
            }
            if (lane < nb) x[i0 + lane] = xi;
        }
        __syncthreads();
        // columns left of the tile, from the first column row tile t touches
        for (int i = tile_first[t] * NB + tid; i < i0; i += blockDim.x) {
            double v = x[i];
            for (int j = 0; j < nb; ++j) v = v - L[static_cast<int64_t>(i0 + j) * n + i] * x[i0 + j];
            x[i] = v;
        }
        __syncthreads();
    }
    return 0;
}

// ---- pipelined substitution for the reassociating precisions -----------------
// Left-looking by 32-row tiles with one workgroup barrier per tile, each tile's
// rows scaled by their reciprocal pivots (unit-diagonal recurrences).  While wave
// 0 runs tile t's recurrence (and the product of tile t's solution with the one
// factor tile that couples it to the next tile), the other waves form the
// contributions of every earlier, already final unknown to the next tile and
// stage the next tile's two factor tiles in LDS.  The sums' association is fixed
// (kSlots interleaved slots of the earlier unknowns, each summed in ascending
// order, then the slots summed in order by wave 0), so the result does not
// depend on the workgroup size.  Scratch: 6 LDS tiles (two diagonal, two
// coupling tiles, two tiles of partial sums); blockDim.x == kThreads >= 128.
constexpr int kSlots = 30;

// Tile (r0.., c0..) of L, zero outside nr x nc (and above the diagonal when
// lower), staged into dst by threads first, first + stride, ...: stage_load
// issues the loads into v, stage_store writes them (kPer >= NB * NB / stride).
template <int kPer>
__device__ __forceinline__ void stage_load(double (&v)[kPer], const double* L, int n, int r0, int c0, int nr, int nc,
                                           bool lower, int first, int stride) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int idx = first + u * stride;
        const int r = idx >> 5, c = idx & (NB - 1);
        v[u] = (idx < NB * NB && r < nr && c < nc && (!lower || c <= r)) ? L[static_cast<int64_t>(r0 + r) * n + c0 + c]
                                                                           : 0.0;
    }
}
template <int kPer>
__device__ __forceinline__ void stage_store(double (*dst)[NB + 1], const double (&v)[kPer], int first, int stride) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int idx = first + u * stride;
        if (idx < NB * NB) dst[idx >> 5][idx & (NB - 1)] = v[u];
    }
}
__device__ __forceinline__ void stage_tile(double (*dst)[NB + 1], const double* L, int n, int r0, int c0, int nr,
                                           int nc, bool lower, int first, int stride) {
    constexpr int kMax = NB * NB / 64;  // stride is at least one wave
    double v[kMax];
    stage_load(v, L, n, r0, c0, nr, nc, lower, first, stride);
    stage_store(dst, v, first, stride);
}

// Slot s of the partial sums for the next tile: sum over j = j0 + s, j0 + s +
// kSlots, ... < j1 of a(j) * x[j], in ascending j, four loads in flight.
template <class F>
__device__ __forceinline__ double slot_sum(F a, const double* x, int j0, int j1) {
    double acc = 0.0;
    int j = j0;
    for (; j + 3 * kSlots < j1; j += 4 * kSlots) {
        const double l0 = a(j), l1 = a(j + kSlots), l2 = a(j + 2 * kSlots), l3 = a(j + 3 * kSlots);
        const double x0 = x[j], x1 = x[j + kSlots], x2 = x[j + 2 * kSlots], x3 = x[j + 3 * kSlots];
        acc = fma(l0, x0, acc);
        acc = fma(l1, x1, acc);
        acc = fma(l2, x2, acc);
        acc = fma(l3, x3, acc);
    }
    for (; j < j1; j += kSlots) acc = fma(a(j), x[j], acc);
    return acc;
}

// Sum of the kSlots partial sums of one unknown (stride NB), in a fixed order.
__device__ __forceinline__ double slot_total(const double* ps) {
    double pv[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) pv[s] = ps[s * NB];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < kSlots; ++s) a[s & 3] += pv[s];
    return (a[0] + a[1]) + (a[2] + a[3]);
}

template <int kThreads>
__device__ __forceinline__ int triangular_solve_fast_ld(int n, int ld, const double* L, const double* b, double* x,
                                                        const int32_t* tile_first, const int32_t* tile_last,
                                                        double (*buf)[NB][NB + 1], double* sh, int* flag, int passes,
                                                        bool check) {
    static_assert(kThreads >= 128 && kThreads % 64 == 0, "one wave for the recurrences, at least one worker wave");
    constexpr int nwk = kThreads / 64 - 1;                           // worker waves
    constexpr int kPer = (NB * NB + nwk * 64 - 1) / (nwk * 64);      // staged elements per worker thread
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (check) {
        double m = 0;
        for (int i = tid; i < n; i += blockDim.x) {
            const double a = fabs(L[static_cast<int64_t>(i) * ld + i]);
            m = (m < a) ? a : m;
        }
        const double mx = block_max(m, sh);
        if (tid == 0) *flag = 0;
        __syncthreads();
        const double eps = mx * 1e-7;
        for (int i = tid; i < n; i += blockDim.x)
            if (!(fabs(L[static_cast<int64_t>(i) * ld + i]) > eps)) *flag = 1;
    }
    if (b != x)
        for (int i = tid; i < n; i += blockDim.x) x[i] = b[i];
    const int T = (n + NB - 1) / NB;
    auto rows_of = [n](int t) { return (n - t * NB) < NB ? (n - t * NB) : NB; };
    double* part = &buf[4][0][0];  // forward [2][NB]; backward [2][kSlots][NB]
    double own = 0.0;              // wave 0: the coupling tile's product for the next tile
    if (passes & 1) {
    stage_tile(buf[0], L, ld, 0, 0, rows_of(0), rows_of(0), true, tid, blockDim.x);
    if (T > 1 && tile_first[1] <= 0) stage_tile(buf[2], L, ld, NB, 0, rows_of(1), NB, false, tid, blockDim.x);
    __syncthreads();
    if (check && *flag) return 2;

    // forward: L y = b
    for (int t = 0; t < T; ++t) {
        const int p = t & 1, i0 = t * NB, nb = rows_of(t);
        if (wave == 0) {
            const int lane = opaque_lane(), lr = lane & (NB - 1);
            const double (*D)[NB + 1] = buf[p];
            double xi = 0.0;
            if (lane < nb) {
                double red = 0.0;
                if (t > 0) red = slot_total(part + p * kSlots * NB + lane);
                xi = x[i0 + lane] - (red + own);
            }
            __builtin_amdgcn_sched_barrier(0);
            // unit-diagonal form: row i scaled by 1 / L_ii, so that step j is one
            // broadcast and one fma
            const double yl = 1.0 / ((lane < nb) ? D[lr][lr] : 1.0);
            double rowv[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) rowv[c] = D[lr][c] * ((c < lane) ? yl : 0.0);  // loads not predicated
            xi *= yl;
#pragma unroll
            for (int j = 0; j < NB; ++j) xi = fma(-rowv[j], lane_value(xi, j), xi);
            if (lane < nb) x[i0 + lane] = xi;
            own = 0.0;
            if (t + 1 < T && tile_first[t + 1] <= t) {
                const double (*O)[NB + 1] = buf[2 + p];
                double ov[NB];
#pragma unroll
                for (int c = 0; c < NB; ++c) ov[c] = O[lr][c];
                double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
                for (int c = 0; c < NB; c += 4) {
                    a0 = fma(ov[c], lane_value(xi, c), a0);
                    a1 = fma(ov[c + 1], lane_value(xi, c + 1), a1);
                    a2 = fma(ov[c + 2], lane_value(xi, c + 2), a2);
                    a3 = fma(ov[c + 3], lane_value(xi, c + 3), a3);
                    if ((c & 7) == 4) __builtin_amdgcn_sched_barrier(0);  // bound the broadcasts in flight
                }
                own = (a0 + a1) + (a2 + a3);
            }
        } else if (t + 1 < T) {
            const int u = t + 1, iu = u * NB, nbu = rows_of(u), c0 = tile_first[u] * NB;
            const int first = tid - 64, stride = nwk * 64;
            // the next tile's factor tiles, loaded first (they do not depend on x)
            double sd[kPer], so[kPer];
            const bool has_o = u + 1 < T && tile_first[u + 1] <= u;
            stage_load(sd, L, ld, iu, iu, nbu, nbu, true, first, stride);
            if (has_o) stage_load(so, L, ld, iu + NB, iu, rows_of(u + 1), NB, false, first, stride);
            // slot s, row r: columns c0 + s, c0 + s + kSlots, ... left of this tile
            const int r = lane & (NB - 1);
            const double* Lr = L + static_cast<int64_t>(iu + (r < nbu ? r : 0)) * ld;
            for (int s = 2 * (wave - 1) + (lane >> 5); s < kSlots; s += 2 * nwk) {
                const double acc = (r < nbu) ? slot_sum([Lr](int j) { return Lr[j]; }, x, c0 + s, i0) : 0.0;
                part[(u & 1) * kSlots * NB + s * NB + r] = acc;
            }
            stage_store(buf[u & 1], sd, first, stride);
            if (has_o) stage_store(buf[2 + (u & 1)], so, first, stride);
        }
        __syncthreads();
    }

    }
    if (!(passes & 2)) return 0;
    if (!(passes & 1)) {
        __syncthreads();
        if (check && *flag) return 2;
    }

    // backward: L^T x = y
    {
        const int t = T - 1;
        stage_tile(buf[t & 1], L, ld, t * NB, t * NB, rows_of(t), rows_of(t), true, tid, blockDim.x);
        if (t >= 1 && tile_last[t - 1] >= t)
            stage_tile(buf[2 + (t & 1)], L, ld, t * NB, (t - 1) * NB, rows_of(t), NB, false, tid, blockDim.x);
    }
    __syncthreads();
    own = 0.0;
    for (int t = T - 1; t >= 0; --t) {
        const int p = t & 1, i0 = t * NB, nb = rows_of(t);
        if (wave == 0) {
            const int lane = opaque_lane(), lr = lane & (NB - 1);
            const double (*D)[NB + 1] = buf[p];
            double xc = 0.0;
            if (lane < nb) {
                double red = 0.0;
                if (t < T - 1) red = slot_total(part + p * kSlots * NB + lane);
                xc = x[i0 + lane] - (red + own);
            }
            __builtin_amdgcn_sched_barrier(0);
            const double yl = 1.0 / ((lane < nb) ? D[lr][lr] : 1.0);
            double colv[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) colv[j] = D[j][lr] * ((j > lane) ? yl : 0.0);  // loads not predicated
            xc *= yl;
#pragma unroll
            for (int j = NB - 1; j >= 0; --j) xc = fma(-colv[j], lane_value(xc, j), xc);
            if (lane < nb) x[i0 + lane] = xc;
            own = 0.0;
            if (t >= 1 && tile_last[t - 1] >= t) {
                const double (*O)[NB + 1] = buf[2 + p];
                double ov[NB];
#pragma unroll
                for (int j = 0; j < NB; ++j) ov[j] = O[j][lr];
                double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j += 4) {
                    a0 = fma(ov[j], lane_value(xc, j), a0);
                    a1 = fma(ov[j + 1], lane_value(xc, j + 1), a1);
                    a2 = fma(ov[j + 2], lane_value(xc, j + 2), a2);
                    a3 = fma(ov[j + 3], lane_value(xc, j + 3), a3);
                    if ((j & 7) == 4) __builtin_amdgcn_sched_barrier(0);  // bound the broadcasts in flight
                }
                own = (a0 + a1) + (a2 + a3);
            }
        } else if (t >= 1) {
            const int u = t - 1, iu = u * NB;
            const int jbeg = i0 + NB;
            const int jend = (tile_last[u] + 1) * NB < n ? (tile_last[u] + 1) * NB : n;
            const int first = tid - 64, stride = nwk * 64;
            double sd[kPer], so[kPer];
            const bool has_o = u >= 1 && tile_last[u - 1] >= u;
            stage_load(sd, L, ld, iu, iu, NB, NB, true, first, stride);
            if (has_o) stage_load(so, L, ld, iu, iu - NB, NB, NB, false, first, stride);
            const int c = lane & (NB - 1);
            const double* Lc = L + iu + c;
            for (int s = 2 * (wave - 1) + (lane >> 5); s < kSlots; s += 2 * nwk) {
                const double acc = slot_sum([Lc, ld](int j) { return Lc[static_cast<int64_t>(j) * ld]; }, x, jbeg + s, jend);
                part[(u & 1) * kSlots * NB + s * NB + c] = acc;
            }
            stage_store(buf[u & 1], sd, first, stride);
            if (has_o) stage_store(buf[2 + (u & 1)], so, first, stride);
        }
        __syncthreads();
    }
    return 0;
}


// The whole system (leading dimension n), both passes, with the failure check.
template <int kThreads>
__device__ __forceinline__ int triangular_solve_fast(int n, const double* L, const double* b, double* x,
                                                     const int32_t* tile_first, const int32_t* tile_last,
                                                     double (*buf)[NB][NB + 1], double* sh, int* flag) {
    return triangular_solve_fast_ld<kThreads>(n, n, L, b, x, tile_first, tile_last, buf, sh, flag, 3, true);
}

}  // namespace wg
}  // namespace zb
