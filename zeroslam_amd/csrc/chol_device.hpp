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

// Factorisation of a diagonal tile by wave 0 with row r in the registers of lane
// r; column j is broadcast lane by lane.  Every lane forms L_jj and its
// reciprocal; the updates run on all lanes without masks, because entries above
// the diagonal are never read.  Returns false (uniformly) when a pivot is not
// greater than eps.
__device__ inline bool factor_diag_tile(double (*D)[NB + 1], int nr, double eps, int* ok_sh) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        double row[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) row[c] = (lane < nr) ? D[lane][c] : 0.0;
        int ok = 1;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < nr) {
                const double d = lane_value(row[j], j);
                ok = ok && (d > eps);
                const double l = sqrt(d);
                const double y = 1.0 / l;
                row[j] = (lane == j) ? l : div_cr(row[j], l, y);
#pragma unroll
                for (int c = j + 1; c < NB; ++c) {
                    const double lcj = lane_value(row[j], c);
                    row[c] = row[c] - row[j] * lcj;
                }
            }
        }
        if (lane < nr) {
#pragma unroll
            for (int c = 0; c < NB; ++c) D[lane][c] = row[c];
        }
        if (lane == 0) *ok_sh = ok;
    }
    __syncthreads();
    return *ok_sh != 0;
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
                             int* ok_sh) {
    const int T = (n + NB - 1) / NB;
    for (int k = 0; k < T; ++k) {
        const int k0 = k * NB;
        const int nk = (n - k0) < NB ? (n - k0) : NB;
        const int h = tile_last[k] - k + 1;
        const int rows_end = (tile_last[k] + 1) * NB < n ? (tile_last[k] + 1) * NB : n;
        const int rows = rows_end - k0;  // panel rows, the diagonal tile's first
        for (int idx = threadIdx.x; idx < rows * NB; idx += blockDim.x) {
            const int r = idx / NB, c = idx % NB;
            const int t = r / NB, rr = r % NB;
            Pn[t][rr][c] = (c < nk && (t > 0 || c <= rr)) ? A[static_cast<int64_t>(k0 + r) * n + k0 + c] : 0.0;
        }
        __syncthreads();
        if (!factor_diag_tile(Pn[0], nk, eps, ok_sh)) return false;
        // rows below: L_r = A_r L_kk^-T, element c gets -= x_m L_cm (m < c) then / L_cc
        if (static_cast<int>(threadIdx.x) < rows - nk) {
            const int r = nk + threadIdx.x;
            const int t = r / NB, rr = r % NB;
            double x[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) x[c] = Pn[t][rr][c];
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                if (c < nk) {
                    double v = x[c];
#pragma unroll
                    for (int m = 0; m < c; ++m) v = v - x[m] * Pn[0][c][m];
                    x[c] = v / Pn[0][c][c];
                }
            }
#pragma unroll
            for (int c = 0; c < NB; ++c) Pn[t][rr][c] = x[c];
        }
        __syncthreads();
        for (int idx = threadIdx.x; idx < rows * NB; idx += blockDim.x) {
            const int r = idx / NB, c = idx % NB;
            const int t = r / NB, rr = r % NB;
            if (c < nk && (t > 0 || c <= rr)) A[static_cast<int64_t>(k0 + r) * n + k0 + c] = Pn[t][rr][c];
        }
        // trailing tiles (I, J), k < J <= I <= tile_last[k]: a -= L_I,k L_J,k^T
        // over the panel's 32 columns in order (only full panels have tiles below)
        const int m = h - 1;
        const int64_t elems = static_cast<int64_t>(m) * (m + 1) / 2 * (NB * NB);
        for (int64_t e = threadIdx.x; e < elems; e += blockDim.x) {
            const int pair = static_cast<int>(e / (NB * NB));
            const int within = static_cast<int>(e % (NB * NB));
            int Ip = 0;
            while ((Ip + 1) * (Ip + 2) / 2 <= pair) ++Ip;
            const int Jp = pair - Ip * (Ip + 1) / 2;
            const int rr = within / NB, cc = within % NB;
            const int pr = (Ip + 1) * NB + rr, pc = (Jp + 1) * NB + cc;  // panel rows of the element's row, column
            if (pr >= rows || pc >= rows || (Ip == Jp && cc > rr)) continue;
            double* out = A + static_cast<int64_t>(k0 + pr) * n + (k0 + pc);
            double v = *out;
            const double* Li = Pn[Ip + 1][rr];
            const double* Lj = Pn[Jp + 1][cc];
#pragma unroll 8
            for (int q = 0; q < NB; ++q) v = v - Li[q] * Lj[q];
            *out = v;
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

}  // namespace wg
}  // namespace zb
