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
// LDS, so no column's broadcasts pile up in scalar registers.  No branches: bad
// is set where a division needs div_cr's fallback.  recips[j] receives
// RN(1 / L_jj).  Returns false (uniformly) when a pivot of a real row is not
// greater than eps.
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
        if (j + 1 < NB) {
            col[j][lane] = row[j];
            row[j + 1] = row[j + 1] - row[j] * lane_value(row[j], j + 1);
#pragma unroll
            for (int c = j + 2; c < NB; ++c) row[c] = row[c] - row[j] * col[j][c];
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
        bool ok = factor_rows(row, nr, eps, recips, bad, col);
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

// ---- whole factorisation in one workgroup -------------------------------------
// Panel k (row tiles k .. tile_last[k] of tile column k, already updated by the
// panels before it) is staged in LDS, its diagonal tile factored by wave 0, the
// tiles below solved one row per thread, the panel written back, and the trailing
// tiles it touches updated element by element from LDS.  Workgroup barriers
// order the panels; nothing waits on another workgroup.
// Returns false (uniformly) when a pivot fails.  Pn holds kTiles staged tiles;
// every column panel of the envelope must be at most kTiles tiles tall.
template <int kTiles>
__device__ bool fused_factor(int n, double* A, const int32_t* tile_last, double eps, double (*Pn)[NB][NB + 1],
                             int* ok_sh, double* recips, double (*col)[64]) {
    const int T = (n + NB - 1) / NB;
    const int tid = threadIdx.x;
    for (int k = 0; k < T; ++k) {
        const int k0 = k * NB;
        const int nk = (n - k0) < NB ? (n - k0) : NB;
        const int h = tile_last[k] - k + 1;
        const int rows_end = (tile_last[k] + 1) * NB < n ? (tile_last[k] + 1) * NB : n;
        const int rows = rows_end - k0;  // panel rows, the diagonal tile's first
#pragma unroll 4
        for (int idx = tid; idx < rows * NB; idx += blockDim.x) {
            const int r = idx / NB, c = idx % NB;
            const int t = r / NB, rr = r % NB;
            Pn[t][rr][c] = (c < nk && (t > 0 || c <= rr)) ? A[static_cast<int64_t>(k0 + r) * n + k0 + c] : 0.0;
        }
        __syncthreads();
        if (!factor_diag_tile(Pn[0], nk, eps, ok_sh, recips, col)) return false;
        // rows below: L_r = A_r L_kk^-T, element c gets -= x_m L_cm (m < c) then / L_cc
        if (tid < rows - nk) {
            const int r = nk + tid;
            const int t = r / NB, rr = r % NB;
            solve_panel_row(Pn[t][rr], Pn[0], recips);
        }
        __syncthreads();
#pragma unroll 4
        for (int idx = tid; idx < rows * NB; idx += blockDim.x) {
            const int r = idx / NB, c = idx % NB;
            const int t = r / NB, rr = r % NB;
            if (c < nk && (t > 0 || c <= rr)) A[static_cast<int64_t>(k0 + r) * n + k0 + c] = Pn[t][rr][c];
        }
        // trailing tiles (I, J), k < J <= I <= tile_last[k]: a -= L_I,k L_J,k^T
        // over the panel's 32 columns in order (only full panels have tiles below).
        // Each thread loads kChunk elements before updating them.
        constexpr int kChunk = 4;
        const int m = h - 1;
        const int64_t elems = static_cast<int64_t>(m) * (m + 1) / 2 * (NB * NB);
        const double* P = &Pn[0][0][0];
        for (int64_t e0 = tid; e0 < elems; e0 += static_cast<int64_t>(kChunk) * blockDim.x) {
            int64_t off[kChunk];
            int li[kChunk], lj[kChunk];
            double v[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const int64_t e = e0 + static_cast<int64_t>(u) * blockDim.x;
                off[u] = -1;
                li[u] = lj[u] = 0;
                if (e < elems) {
                    const int pair = static_cast<int>(e / (NB * NB));
                    const int within = static_cast<int>(e % (NB * NB));
                    int Ip = 0;
                    while ((Ip + 1) * (Ip + 2) / 2 <= pair) ++Ip;
                    const int Jp = pair - Ip * (Ip + 1) / 2;
                    const int rr = within / NB, cc = within % NB;
                    const int pr = (Ip + 1) * NB + rr, pc = (Jp + 1) * NB + cc;  // panel row, column
                    if (pr < rows && pc < rows && (Ip != Jp || cc <= rr)) {
                        off[u] = static_cast<int64_t>(k0 + pr) * n + (k0 + pc);
                        li[u] = ((Ip + 1) * NB + rr) * (NB + 1);
                        lj[u] = ((Jp + 1) * NB + cc) * (NB + 1);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kChunk; ++u) v[u] = (off[u] >= 0) ? A[off[u]] : 0.0;
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                if (off[u] >= 0) {
                    double a = v[u];
#pragma unroll 8
                    for (int q = 0; q < NB; ++q) a = a - P[li[u] + q] * P[lj[u] + q];
                    A[off[u]] = a;
                }
            }
        }
        __syncthreads();
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
            double v = x[i];
            const double* Li = L + static_cast<int64_t>(i) * n + i0;
            for (int j = 0; j < nb; ++j) v = v - Li[j] * x[i0 + j];
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
// Left-looking by 32-row tiles with one workgroup barrier per tile.  While wave
// 0 runs tile t's recurrence (and the product of tile t's solution with the one
// factor tile that couples it to the next tile), the other waves form the
// contributions of every earlier, already final unknown to the next tile and
// stage the next tile's two factor tiles in LDS.  The sums' association is fixed
// (lane-strided column sums reduced by a butterfly for the forward pass, kSlots
// row slots summed in order for the backward pass), so the result does not depend
// on the workgroup size.  Scratch: 6 LDS tiles (two diagonal, two coupling tiles,
// two tiles of partial sums); blockDim.x is at least 128.
constexpr int kSlots = 30;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

// dst = tile (r0.., c0..) of L, zero outside nr x nc (and above the diagonal
// when lower), by threads first, first + stride, ...
__device__ __forceinline__ void stage_tile(double (*dst)[NB + 1], const double* L, int n, int r0, int c0, int nr,
                                           int nc, bool lower, int first, int stride) {
    for (int idx = first; idx < NB * NB; idx += stride) {
        const int r = idx >> 5, c = idx & (NB - 1);
        dst[r][c] = (r < nr && c < nc && (!lower || c <= r)) ? L[static_cast<int64_t>(r0 + r) * n + c0 + c] : 0.0;
    }
}

__device__ __forceinline__ int triangular_solve_fast(int n, const double* L, const double* b, double* x,
                                            const int32_t* tile_first, const int32_t* tile_last,
                                            double (*buf)[NB][NB + 1], double* sh, int* flag) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nwk = static_cast<int>(blockDim.x >> 6) - 1;  // worker waves
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
    const int T = (n + NB - 1) / NB;
    auto rows_of = [n](int t) { return (n - t * NB) < NB ? (n - t * NB) : NB; };
    double* part = &buf[4][0][0];  // forward [2][NB]; backward [2][kSlots][NB]
    stage_tile(buf[0], L, n, 0, 0, rows_of(0), rows_of(0), true, tid, blockDim.x);
    if (T > 1 && tile_first[1] <= 0) stage_tile(buf[2], L, n, NB, 0, rows_of(1), NB, false, tid, blockDim.x);
    __syncthreads();
    if (*flag) return 2;

    // forward: L y = b
    double own = 0.0;  // wave 0: the coupling tile's product for the next tile
    for (int t = 0; t < T; ++t) {
        const int p = t & 1, i0 = t * NB, nb = rows_of(t);
        if (wave == 0) {
            const int lane = opaque_lane(), lr = lane & (NB - 1);
            const double (*D)[NB + 1] = buf[p];
            double rowv[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) rowv[c] = (lane < NB && c < lane) ? D[lr][c] : 0.0;
            const double yl = 1.0 / ((lane < nb) ? D[lr][lr] : 1.0);
            double xi = 0.0;
            if (lane < nb) xi = x[i0 + lane] - (((t > 0) ? part[p * NB + lane] : 0.0) + own);
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                xi = (lane == j) ? xi * yl : xi;
                xi = fma(-rowv[j], lane_value(xi, j), xi);
            }
            if (lane < nb) x[i0 + lane] = xi;
            own = 0.0;
            if (t + 1 < T && tile_first[t + 1] <= t) {
                const double (*O)[NB + 1] = buf[2 + p];
                double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
                for (int c = 0; c < NB; c += 4) {
                    a0 = fma(O[lr][c], lane_value(xi, c), a0);
                    a1 = fma(O[lr][c + 1], lane_value(xi, c + 1), a1);
                    a2 = fma(O[lr][c + 2], lane_value(xi, c + 2), a2);
                    a3 = fma(O[lr][c + 3], lane_value(xi, c + 3), a3);
                }
                own = (a0 + a1) + (a2 + a3);
            }
        } else if (t + 1 < T) {
            const int u = t + 1, iu = u * NB, nbu = rows_of(u), c0 = tile_first[u] * NB;
            for (int r = wave - 1; r < nbu; r += nwk) {
                const double* Lr = L + static_cast<int64_t>(iu + r) * n;
                double acc = 0.0;
                for (int j = c0 + lane; j < i0; j += 64) acc = fma(Lr[j], x[j], acc);
                acc = wave_sum(acc);
                if (lane == 0) part[(u & 1) * NB + r] = acc;
            }
            const int first = tid - 64, stride = nwk * 64;
            stage_tile(buf[u & 1], L, n, iu, iu, nbu, nbu, true, first, stride);
            if (u + 1 < T && tile_first[u + 1] <= u)
                stage_tile(buf[2 + (u & 1)], L, n, iu + NB, iu, rows_of(u + 1), NB, false, first, stride);
        }
        __syncthreads();
    }

    // backward: L^T x = y
    {
        const int t = T - 1;
        stage_tile(buf[t & 1], L, n, t * NB, t * NB, rows_of(t), rows_of(t), true, tid, blockDim.x);
        if (t >= 1 && tile_last[t - 1] >= t)
            stage_tile(buf[2 + (t & 1)], L, n, t * NB, (t - 1) * NB, rows_of(t), NB, false, tid, blockDim.x);
    }
    __syncthreads();
    own = 0.0;
    for (int t = T - 1; t >= 0; --t) {
        const int p = t & 1, i0 = t * NB, nb = rows_of(t);
        if (wave == 0) {
            const int lane = opaque_lane(), lr = lane & (NB - 1);
            const double (*D)[NB + 1] = buf[p];
            double colv[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) colv[j] = (lane < NB && j > lane) ? D[j][lr] : 0.0;
            const double yl = 1.0 / ((lane < nb) ? D[lr][lr] : 1.0);
            double xc = 0.0;
            if (lane < nb) {
                double red = 0.0;
                if (t < T - 1) {
                    const double* ps = part + p * kSlots * NB + lane;
                    for (int s = 0; s < kSlots; ++s) red += ps[s * NB];
                }
                xc = x[i0 + lane] - (red + own);
            }
#pragma unroll
            for (int j = NB - 1; j >= 0; --j) {
                xc = (lane == j) ? xc * yl : xc;
                xc = fma(-colv[j], lane_value(xc, j), xc);
            }
            if (lane < nb) x[i0 + lane] = xc;
            own = 0.0;
            if (t >= 1 && tile_last[t - 1] >= t) {
                const double (*O)[NB + 1] = buf[2 + p];
                double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j += 4) {
                    a0 = fma(O[j][lr], lane_value(xc, j), a0);
                    a1 = fma(O[j + 1][lr], lane_value(xc, j + 1), a1);
                    a2 = fma(O[j + 2][lr], lane_value(xc, j + 2), a2);
                    a3 = fma(O[j + 3][lr], lane_value(xc, j + 3), a3);
                }
                own = (a0 + a1) + (a2 + a3);
            }
        } else if (t >= 1) {
            const int u = t - 1, iu = u * NB;
            const int jbeg = i0 + NB;
            const int jend = (tile_last[u] + 1) * NB < n ? (tile_last[u] + 1) * NB : n;
            const int c = lane & (NB - 1);
            for (int s = 2 * (wave - 1) + (lane >> 5); s < kSlots; s += 2 * nwk) {
                double acc = 0.0;
                for (int j = jbeg + s; j < jend; j += kSlots) acc = fma(L[static_cast<int64_t>(j) * n + iu + c], x[j], acc);
                part[(u & 1) * kSlots * NB + s * NB + c] = acc;
            }
            const int first = tid - 64, stride = nwk * 64;
            stage_tile(buf[u & 1], L, n, iu, iu, NB, NB, true, first, stride);
            if (u >= 1 && tile_last[u - 1] >= u)
                stage_tile(buf[2 + (u & 1)], L, n, iu, iu - NB, NB, NB, false, first, stride);
        }
        __syncthreads();
    }
    return 0;
}

}  // namespace wg
}  // namespace zb
