// Dense binary64 Cholesky factorisation and solve of the reduced camera system
// (replaces matrix::decompose_cholesky / matrix::solve_cholesky,
// source/matrix_decomposition_cholesky.hpp:26-95, as called at
// factor_graph.hpp:782-793).
//
// Blocked factorisation over 32-wide panels, in place on the lower triangle of S
// (row-major n x n): factor the diagonal tile, solve the panel below it, update
// the trailing lower triangle tile by tile.  Every element receives its
// a -= L_ik L_jk subtractions in ascending k and its division / square root
// last, the operation sequence of the reference's row-Crout loop, so the
// factor equals the reference's bit for bit.
//
// Envelope.  Row r of S is zero left of column 6 * (first pose coupled to r's
// pose), and Cholesky fill stays inside that envelope, so every product with a
// factor entry left of it is an exact zero (a - 0 = a).  The kernels skip those
// 32 x 32 tiles: tile_first[I] is the first non-zero tile column of row tile I,
// tile_last[k] the last row tile with a non-zero tile in column k.  On the
// forward-walk problems the envelope holds a few percent of the dense flops.
//
// Failure rules are the reference's: decomposition fails when a pivot is not
// greater than 1e-14 * max|S_ii| (NaN fails), the solve fails when any |L_ii| is
// not greater than 1e-7 * max|L_ii|.
#include "chol_device.hpp"
#include "common.hpp"

#include <cmath>

namespace zb {
namespace {

constexpr int NB = wg::kTile;
using wg::factor_diag_tile;
using wg::lane_value;

__device__ __forceinline__ bool skip(const lm_state* lm, int gated, const int32_t* status) {
    return (gated && lm->done) || *status != 0;
}

// eps = 1e-14 * max|S_ii| with math::max semantics (NaN entries never win).
// floor_rel > 0 first raises every diagonal entry by floor_rel * max|S_ii|.
__global__ void __launch_bounds__(1024) k_chol_begin(int n, double* S, double* eps, int32_t* status,
                                                     const lm_state* lm, int gated, double floor_rel) {
    if (gated && lm->done) return;
    __shared__ double sh[1024];
    double m = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double a = fabs(S[static_cast<int64_t>(i) * n + i]);
        m = (m < a) ? a : m;
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double a = sh[threadIdx.x + w];
            sh[threadIdx.x] = (sh[threadIdx.x] < a) ? a : sh[threadIdx.x];
        }
        __syncthreads();
    }
    const double mx = sh[0];
    const double lift = floor_rel * mx;
    if (floor_rel > 0.0) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) S[static_cast<int64_t>(i) * n + i] += lift;
    }
    if (threadIdx.x == 0) {
        eps[0] = (mx + lift) * 1e-14;
        status[0] = 0;
    }
}

// ---- panel-lookahead factorisation -------------------------------------------
// Lower-triangle tiles (I, J), I >= J, of 32 x 32.  Launch k (k = 0 .. T-1) does
// two independent jobs that only read panel k-1, which launch k-1 finished:
//   * column k: for every row tile I >= k, apply the panel k-1 update to (I, k),
//     factor the diagonal tile (k, k) (each workgroup factors its own copy, so no
//     workgroup waits for another) and solve (I, k) against it;
//   * every tile (I, J) with J > k: apply the panel k-1 update.
// Each element therefore receives a -= L_ij * L_kj for panel columns in
// ascending order and its division / square root last: the operation sequence
// of the reference's row-Crout loop.  Ordering comes from kernel boundaries only.

// own -= L_I L_J^T over one 32-column panel, column by column (LDS operands).
__device__ __forceinline__ void panel_update(double (*own)[NB + 1], const double (*LI)[NB + 1], const double (*LJ)[NB + 1],
                                             int nr, int nc, bool lower_only) {
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        const int r = idx / NB, c = idx % NB;
        if (r >= nr || c >= nc || (lower_only && c > r)) continue;
        double v = own[r][c];
#pragma unroll 8
        for (int m = 0; m < NB; ++m) v = v - LI[r][m] * LJ[c][m];
        own[r][c] = v;
    }
}

typedef double mfma_d4 __attribute__((ext_vector_type(4)));

// panel_update on FP64 matrix cores (the reassociating precisions; 256 threads):
// wave w forms quadrant (w / 2, w % 2) of LI LJ^T with v_mfma_f64_16x16x4f64
// (lane l: operand row l % 16 at depth l / 16; result element v at row
// l / 16 + 4 v, column l % 16) and subtracts it; lower_only skips the upper
// quadrant and the entries above the diagonal.
__device__ __forceinline__ void panel_update_mfma(double (*own)[NB + 1], const double (*LI)[NB + 1],
                                                  const double (*LJ)[NB + 1], int nr, int nc, bool lower_only) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int qi = wave >> 1, qj = wave & 1;
    if (lower_only && qj > qi) return;
    const int ar = qi * 16 + (lane & 15), br = qj * 16 + (lane & 15), kq = lane >> 4;
    mfma_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < NB / 4; ++kk)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(LI[ar][kk * 4 + kq], LJ[br][kk * 4 + kq], acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = qi * 16 + kq + 4 * v;
        if (r < nr && br < nc && (!lower_only || br <= r)) own[r][br] -= acc[v];
    }
}

__device__ __forceinline__ void load_tile(double (*dst)[NB + 1], const double* A, int n, int r0, int c0, int nr, int nc,
                                          bool lower_only) {
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        const int r = idx / NB, c = idx % NB;
        dst[r][c] = (r < nr && c < nc && (!lower_only || c <= r)) ? A[static_cast<int64_t>(r0 + r) * n + c0 + c] : 0.0;
    }
}
__device__ __forceinline__ void store_tile(const double (*src)[NB + 1], double* A, int n, int r0, int c0, int nr, int nc,
                                           bool lower_only) {
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        const int r = idx / NB, c = idx % NB;
        if (r < nr && c < nc && (!lower_only || c <= r)) A[static_cast<int64_t>(r0 + r) * n + c0 + c] = src[r][c];
    }
}

// Blocked form (kb0 > 0 or kb1 < T): launch k only updates the tiles of the
// columns kb0 .. kb1 - 1; the panels of that block reach the tiles right of it
// through k_chol_syrk, which runs after the block's last launch, so launch kb0
// applies no panel (the block before it arrived through k_chol_syrk).
__host__ __device__ inline int64_t step_trailing_tiles(int64_t m, int64_t w) {
    return w * (w + 1) / 2 + (m - w) * w;  // tiles k < J < kb1, J <= I < T (m = T - k - 1, w = kb1 - k - 1)
}

// kFast (the reassociating precisions): every tile's loads issued together,
// panel updates on matrix cores, factor_diag_tile_fast and solve_panel_row_fast.
template <bool kFast>
__global__ void __launch_bounds__(256) k_chol_step(int n, double* A, const double* eps_p, int32_t* status, int k,
                                                   int kb0, int kb1, const int32_t* tile_first, const lm_state* lm,
                                                   int gated) {
    if ((gated && lm->done) || status[0] != 0) return;
    __shared__ double own[NB][NB + 1];
    __shared__ double D[NB][NB + 1];
    __shared__ double LI[NB][NB + 1];
    __shared__ double LJ[NB][NB + 1];
    __shared__ double recips[NB];
    __shared__ double col[NB][64];
    __shared__ int ok_sh;
    const int T = (n + NB - 1) / NB;
    const int np_wg = T - k;  // column-k workgroups
    const int k0 = k * NB, kp0 = (k - 1) * NB;
    const int nk = (n - k0) < NB ? (n - k0) : NB;
    if (static_cast<int>(blockIdx.x) < np_wg) {
        const int I = k + blockIdx.x;
        if (tile_first[I] > k) return;  // tile (I, k) lies outside the envelope
        const int r0 = I * NB;
        const int nr = (n - r0) < NB ? (n - r0) : NB;
        if (kFast) {
            const bool upd_k = k > kb0 && tile_first[k] <= k - 1;
            const bool upd_i = I != k && upd_k && tile_first[I] <= k - 1;
            constexpr int kPer = NB * NB / 256;
            double vd[kPer], vj[kPer], vo[kPer], vi[kPer];
#pragma unroll
            for (int e = 0; e < kPer; ++e) {
                const int idx = threadIdx.x + e * 256, r = idx / NB, c = idx % NB;
                vd[e] = (r < nk && c <= r) ? A[static_cast<int64_t>(k0 + r) * n + k0 + c] : 0.0;
                vj[e] = (upd_k && r < nk) ? A[static_cast<int64_t>(k0 + r) * n + kp0 + c] : 0.0;
                vo[e] = (I != k && r < nr) ? A[static_cast<int64_t>(r0 + r) * n + k0 + c] : 0.0;
                vi[e] = (upd_i && r < nr) ? A[static_cast<int64_t>(r0 + r) * n + kp0 + c] : 0.0;
            }
#pragma unroll
            for (int e = 0; e < kPer; ++e) {
                const int idx = threadIdx.x + e * 256, r = idx / NB, c = idx % NB;
                D[r][c] = vd[e];
                LJ[r][c] = vj[e];
                own[r][c] = vo[e];
                LI[r][c] = vi[e];
            }
            __syncthreads();
            if (upd_k) panel_update_mfma(D, LJ, LJ, nk, nk, true);
            if (upd_i) panel_update_mfma(own, LI, LJ, nr, nk, false);
            __syncthreads();
            if (!wg::factor_diag_tile_fast(D, nk, eps_p[0], &ok_sh, recips, col)) {
                if (threadIdx.x == 0) atomicCAS(status, 0, 1);
                return;
            }
            if (I == k) {
                store_tile(D, A, n, k0, k0, nk, nk, true);
                return;
            }
            if (threadIdx.x < nr) wg::solve_panel_row_fast(own[threadIdx.x], D, recips);
            __syncthreads();
            store_tile(own, A, n, r0, k0, nr, nk, false);
            return;
        }
        // diagonal tile (k, k) with its panel k-1 update
        load_tile(D, A, n, k0, k0, nk, nk, true);
        if (k > kb0 && tile_first[k] <= k - 1) {
            load_tile(LJ, A, n, k0, kp0, nk, NB, false);
            __syncthreads();
            panel_update(D, LJ, LJ, nk, nk, true);
        }
        __syncthreads();
        if (!factor_diag_tile(D, nk, eps_p[0], &ok_sh, recips, col)) {
            if (threadIdx.x == 0) atomicCAS(status, 0, 1);
            return;
        }
        if (I == k) {
            store_tile(D, A, n, k0, k0, nk, nk, true);
            return;
        }
        // row tile (I, k): panel k-1 update, then solve against L_kk^T
        load_tile(own, A, n, r0, k0, nr, nk, false);
        if (k > kb0 && tile_first[I] <= k - 1 && tile_first[k] <= k - 1) {
            load_tile(LI, A, n, r0, kp0, nr, NB, false);
            __syncthreads();
            panel_update(own, LI, LJ, nr, nk, false);
        }
        __syncthreads();
        if (threadIdx.x < nr) {  // I > k, so column k is a full panel
            wg::solve_panel_row(own[threadIdx.x], D, recips);
        }
        __syncthreads();
        store_tile(own, A, n, r0, k0, nr, nk, false);
        return;
    }
    if (k == kb0) return;
    // trailing tile (I, J), k < J < kb1: panel k-1 update.  The first w rows form
    // a triangle, the rows below it full rows of w tiles.
    const int64_t u = static_cast<int64_t>(blockIdx.x) - np_wg;
    const int64_t w = kb1 - 1 - k, tri = w * (w + 1) / 2;
    int64_t Ip, Jp;
    if (u < tri) {
        Ip = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(u) + 1.0) - 1.0) / 2.0);
        while ((Ip + 1) * (Ip + 2) / 2 <= u) ++Ip;
        while (Ip * (Ip + 1) / 2 > u) --Ip;
        Jp = u - Ip * (Ip + 1) / 2;
    } else {
        Ip = w + (u - tri) / w;
        Jp = (u - tri) % w;
    }
    const int I = k + 1 + static_cast<int>(Ip), J = k + 1 + static_cast<int>(Jp);
    if (tile_first[I] > k - 1 || tile_first[J] > k - 1) return;  // panel k-1 is zero in row tile I or J
    const int r0 = I * NB, c0 = J * NB;
    const int nr = (n - r0) < NB ? (n - r0) : NB, nc = (n - c0) < NB ? (n - c0) : NB;
    load_tile(own, A, n, r0, c0, nr, nc, I == J);
    load_tile(LI, A, n, r0, kp0, nr, NB, false);
    load_tile(LJ, A, n, c0, kp0, nc, NB, false);
    __syncthreads();
    if (kFast)
        panel_update_mfma(own, LI, LJ, nr, nc, I == J);
    else
        panel_update(own, LI, LJ, nr, nc, I == J);
    __syncthreads();
    store_tile(own, A, n, r0, c0, nr, nc, I == J);
}

// ---- block update on FP64 matrix cores ----------------------------------------
// A(I, J) -= L(I, P) L(J, P)^T for every lower tile (I, J) with J >= kb1, where P
// are the panel columns kb0 .. kb1 - 1 (only those inside both rows' envelopes).
// Four waves, each one 16 x 16 quadrant of the 32 x 32 tile, accumulate with
// v_mfma_f64_16x16x4f64 (lane l: A operand row l % 16, depth l / 16, B operand
// depth l / 16, column l % 16; result element v: row l / 16 + 4 v, column l % 16).  The depth is walked 32 columns
// at a time through LDS, the next chunk's loads in flight during the products.
// The sums are reassociated (depth order 4 at a time), so the blocked path
// serves the reassociating precisions only.
constexpr int kBlockTiles = 8;  // panels per block of the blocked factorisation

__global__ void __launch_bounds__(256) k_chol_syrk(int n, double* A, int kb0, int kb1, const int32_t* tile_first,
                                                   const int32_t* status, const lm_state* lm, int gated) {
    if ((gated && lm->done) || status[0] != 0) return;
    __shared__ double LI[NB][NB + 1];
    __shared__ double LJ[NB][NB + 1];
    const int64_t u = blockIdx.x;
    int64_t Ip = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(u) + 1.0) - 1.0) / 2.0);
    while ((Ip + 1) * (Ip + 2) / 2 <= u) ++Ip;
    while (Ip * (Ip + 1) / 2 > u) --Ip;
    const int64_t Jp = u - Ip * (Ip + 1) / 2;
    const int I = kb1 + static_cast<int>(Ip), J = kb1 + static_cast<int>(Jp);
    int c_lo = tile_first[I] > tile_first[J] ? tile_first[I] : tile_first[J];
    c_lo = c_lo > kb0 ? c_lo : kb0;
    if (c_lo >= kb1) return;  // the block's columns are zero in row tile I or J
    const int r0 = I * NB, c0 = J * NB;
    const int nr = (n - r0) < NB ? (n - r0) : NB, nc = (n - c0) < NB ? (n - c0) : NB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int qi = wave >> 1, qj = wave & 1;
    const bool active = !(I == J && qj > qi);  // the diagonal tile's upper quadrant
    constexpr int kPer = NB * NB / 256;
    double pa[kPer], pb[kPer];
    auto fetch = [&](int c) {
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = tid + e * 256, r = idx / NB, cc = idx % NB;
            pa[e] = r < nr ? A[static_cast<int64_t>(r0 + r) * n + c * NB + cc] : 0.0;
            pb[e] = r < nc ? A[static_cast<int64_t>(c0 + r) * n + c * NB + cc] : 0.0;
        }
    };
    fetch(c_lo);
    mfma_d4 acc = {0.0, 0.0, 0.0, 0.0};
    const int ar = qi * 16 + (lane & 15), br = qj * 16 + (lane & 15), kq = lane >> 4;
    for (int c = c_lo; c < kb1; ++c) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = tid + e * 256, r = idx / NB, cc = idx % NB;
            LI[r][cc] = pa[e];
            LJ[r][cc] = pb[e];
        }
        __syncthreads();
        if (c + 1 < kb1) fetch(c + 1);
        if (active) {
#pragma unroll
            for (int kk = 0; kk < NB / 4; ++kk)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(LI[ar][kk * 4 + kq], LJ[br][kk * 4 + kq], acc, 0, 0, 0);
        }
    }
    if (!active) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = qi * 16 + kq + 4 * v, cc = qj * 16 + (lane & 15);
        if (r < nr && cc < nc && (I != J || cc <= r)) {
            double* a = A + static_cast<int64_t>(r0 + r) * n + c0 + cc;
            *a = *a - acc[v];
        }
    }
}

// ---- whole factorisation in one workgroup -------------------------------------
// Used when every column panel of the envelope is at most kFusedTiles tiles tall
// (wg::fused_factor: one thread of waves 1.. per row below the diagonal tile).
constexpr int kFusedTiles = 15;
static_assert((kFusedTiles - 1) * NB <= 512 - 64, "a row thread for every row below the diagonal tile");

template <bool kFast>
__global__ void __launch_bounds__(512) k_chol_fused(int n, double* A, const int32_t* tile_last, const double* eps_p,
                                                    int32_t* status, const lm_state* lm, int gated) {
    if ((gated && lm->done) || status[0] != 0) return;
    __shared__ double col[NB][64];
    __shared__ double recips[NB];
    __shared__ double PT[kFusedTiles - 1][NB][wg::kPT];
    __shared__ int ok_sh[2];
    if (!wg::fused_factor<kFusedTiles, kFast>(n, A, tile_last, eps_p[0], PT, ok_sh, recips, col) && threadIdx.x == 0)
        status[0] = 1;
}

__device__ __forceinline__ double readlane_d(double v, int j) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), j);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), j);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Lt[i][j] = L[j][i] for j > i (the strict upper triangle of L^T).
__global__ void __launch_bounds__(256) k_transpose_lower(int n, const double* L, double* Lt, const int32_t* tile_last,
                                                         const int32_t* status, const lm_state* lm, int gated) {
    if (skip(lm, gated, status)) return;
    __shared__ double T[32][33];
    const int bi = blockIdx.y * 32, bj = blockIdx.x * 32;  // source tile rows bi.., cols bj..
    if (bj > bi + 31) return;
    if (static_cast<int>(blockIdx.y) > tile_last[blockIdx.x]) return;  // never read by the backward pass
    for (int t = threadIdx.x; t < 32 * 32; t += blockDim.x) {
        const int r = t / 32, c = t % 32;
        T[r][c] = (bi + r < n && bj + c < n) ? L[static_cast<int64_t>(bi + r) * n + bj + c] : 0.0;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 32 * 32; t += blockDim.x) {
        const int r = t / 32, c = t % 32;  // destination row bj + r, column bi + c
        if (bj + r < n && bi + c < n && bi + c > bj + r) Lt[static_cast<int64_t>(bj + r) * n + bi + c] = T[c][r];
    }
}

// Backward substitution L^T x = y in the reference's order
// (matrix_decomposition_cholesky.hpp:85-93): x_i -= L_ji x_j for j = i+1 .. n-1
// ascending, then x_i /= L_ii.  One wave with x in LDS: while lane 0 runs the
// ordered subtractions for row i, the wave has already staged row i-1 of L^T.
constexpr int kBackLds = 6144;  // largest n handled from LDS
__global__ void __launch_bounds__(64) k_backward_exact_lds(int n, const double* L, const double* Lt, double* x,
                                                           const int32_t* tile_last, const int32_t* status, lm_state* lm,
                                                           int gated) {
    if (skip(lm, gated, status)) return;
    __shared__ double xs[kBackLds];
    __shared__ double prod[kBackLds];
    const int lane = threadIdx.x;
    for (int j = lane; j < n; j += 64) xs[j] = x[j];
    __syncthreads();
    for (int i = n - 1; i >= 0; --i) {
        // products L_ji x_j (x_j final for j > i), formed in parallel; column i
        // of L is zero below row tile tile_last[i / 32]
        const int jend = (tile_last[i / NB] + 1) * NB < n ? (tile_last[i / NB] + 1) * NB : n;
        const double* u = Lt + static_cast<int64_t>(i) * n;
        for (int j = i + 1 + lane; j < jend; j += 64) prod[j] = u[j] * xs[j];
        __syncthreads();
        if (lane == 0) {
            double v = xs[i];
            int j = i + 1;
            for (; j + 32 <= jend; j += 32) {
#pragma unroll
                for (int k = 0; k < 32; ++k) v = v - prod[j + k];
            }
            for (; j < jend; ++j) v = v - prod[j];
            xs[i] = v / L[static_cast<int64_t>(i) * n + i];
        }
        __syncthreads();
    }
    for (int j = lane; j < n; j += 64) x[j] = xs[j];
    if (lane == 0) lm->solve_ok = 1;
}

__global__ void __launch_bounds__(64) k_backward_exact(int n, const double* L, const double* Lt, double* x,
                                                       const int32_t* tile_last, const int32_t* status, lm_state* lm,
                                                       int gated) {
    if (skip(lm, gated, status)) return;
    const int lane = threadIdx.x;
    for (int i = n - 1; i >= 0; --i) {
        double v = x[i];
        const int jend = (tile_last[i / NB] + 1) * NB < n ? (tile_last[i / NB] + 1) * NB : n;
        const double* u = Lt + static_cast<int64_t>(i) * n;
        for (int j0 = i + 1; j0 < jend; j0 += 64) {
            const int j = j0 + lane;
            const double pj = (j < jend) ? u[j] * x[j] : 0.0;
            const int cnt = (jend - j0) < 64 ? (jend - j0) : 64;
            if (cnt == 64) {
#pragma unroll
                for (int k = 0; k < 64; ++k) v = v - readlane_d(pj, k);
            } else {
                for (int k = 0; k < cnt; ++k) v = v - readlane_d(pj, k);
            }
        }
        v = v / L[static_cast<int64_t>(i) * n + i];
        if (lane == 0) x[i] = v;
        __threadfence_block();
        __syncthreads();
    }
    if (lane == 0) lm->solve_ok = 1;
}

// ---- blocked substitution for large systems (reassociating precisions) ------
// Blocks of kSolveBlockTiles row tiles.  Forward: per block, one workgroup
// multiplies the block's unknowns (which already hold every earlier block's
// terms) by the inverse of the diagonal block, formed once per factorisation;
// then a grid-wide launch subtracts the block's terms from every later row of
// the envelope, one wave per row.  Backward: per block from the last, the
// product with the transposed inverse, then a grid-wide launch subtracts its
// terms from every earlier unknown the block's rows reach, one column per lane.
// The L tiles a solve reads never leave the envelope; every sum has a fixed
// order.  Failure rule as in the one-workgroup solve (|L_ii| > 1e-7 max |L_ii|).
constexpr int kSolveBlockTiles = 8;
constexpr int kBlockedSolveTiles = 64;  // systems of more tiles take the blocked substitution

__global__ void __launch_bounds__(1024) k_trsv_begin(int n, const double* L, const double* b, double* x,
                                                     int32_t* status, lm_state* lm, int gated) {
    if (gated && lm->done) return;
    if (*status != 0) {
        if (threadIdx.x == 0) lm->solve_ok = 0;
        return;
    }
    __shared__ double sh[1024];
    __shared__ int bad;
    double m = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double a = fabs(L[static_cast<int64_t>(i) * n + i]);
        m = (m < a) ? a : m;
    }
    const double eps = wg::block_max(m, sh) * 1e-7;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        if (!(fabs(L[static_cast<int64_t>(i) * n + i]) > eps)) bad = 1;
        x[i] = b[i];
    }
    __syncthreads();
    if (bad && threadIdx.x == 0) {
        *status = 2;
        lm->solve_ok = 0;
    }
}

// Rows below block p: x_i -= L_i,block x_block over the row's envelope, one wave
// per row (lane-strided products, then a butterfly sum).
__global__ void __launch_bounds__(256) k_trsv_rows_update(int n, const double* L, double* x, int p,
                                                          const int32_t* tile_first, const int32_t* status,
                                                          const lm_state* lm, int gated) {
    if ((gated && lm->done) || *status != 0) return;
    const int T = (n + NB - 1) / NB, t0 = p * kSolveBlockTiles;
    const int t1 = (t0 + kSolveBlockTiles) < T ? (t0 + kSolveBlockTiles) : T;
    const int c1 = (t1 * NB) < n ? (t1 * NB) : n;
    const int i = c1 + static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const int tf = tile_first[i / NB];
    if (tf >= t1) return;  // the row is zero over the block
    const int cs = (tf * NB) > (t0 * NB) ? (tf * NB) : (t0 * NB);
    const double* Li = L + static_cast<int64_t>(i) * n;
    double acc = 0.0;
    for (int j = cs + lane; j < c1; j += 64) acc = fma(Li[j], x[j], acc);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) x[i] -= acc;
}

// Unknowns left of block p: x_j -= sum_{i in block} L_ij x_i over the rows whose
// envelope reaches column j.  Lane = column; wave w sums the rows of tile t0 + w
// in four interleaved chains; the eight tile sums are added in a fixed order.
__global__ void __launch_bounds__(512) k_trsv_cols_update(int n, const double* L, double* x, int p,
                                                          const int32_t* tile_first, const int32_t* status,
                                                          const lm_state* lm, int gated) {
    static_assert(kSolveBlockTiles == 8, "eight waves of one tile");
    if ((gated && lm->done) || *status != 0) return;
    __shared__ double red[8][64];
    const int T = (n + NB - 1) / NB, t0 = p * kSolveBlockTiles;
    const int t1 = (t0 + kSolveBlockTiles) < T ? (t0 + kSolveBlockTiles) : T;
    const int c0 = t0 * NB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = static_cast<int>(blockIdx.x) * 64 + lane;
    const int t = t0 + wave;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    if (j < c0 && t < t1 && tile_first[t] * NB <= j) {
        const int i0 = t * NB, nb = (n - i0) < NB ? (n - i0) : NB;
        const double* Lj = L + static_cast<int64_t>(i0) * n + j;
        int r = 0;
        for (; r + 3 < nb; r += 4) {
            a0 = fma(Lj[static_cast<int64_t>(r) * n], x[i0 + r], a0);
            a1 = fma(Lj[static_cast<int64_t>(r + 1) * n], x[i0 + r + 1], a1);
            a2 = fma(Lj[static_cast<int64_t>(r + 2) * n], x[i0 + r + 2], a2);
            a3 = fma(Lj[static_cast<int64_t>(r + 3) * n], x[i0 + r + 3], a3);
        }
        for (; r < nb; ++r) a0 = fma(Lj[static_cast<int64_t>(r) * n], x[i0 + r], a0);
    }
    red[wave][lane] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    if (wave == 0 && j < c0)
        x[j] -= ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane])) +
                ((red[4][lane] + red[5][lane]) + (red[6][lane] + red[7][lane]));
}

// Inverses of the diagonal blocks, formed once per factorisation so that the
// diagonal steps of every solve with that factor (the step and the mixed
// path's refinement corrections) are matrix-vector products instead of
// dependent substitutions.  Block p (rows r0 .. r0 + m - 1, m <= 256) gets a
// 256 x 256 slot holding X = inv(L_bb) column-major (slot p) and row-major
// (slot P + p), zeros above the diagonal, so that both products read
// coalesced columns.  X is formed by 32-row tiles: the tiles' own inverses
// X_ii = inv(L_ii), then down each tile column j
//   X_ij = -X_ii sum_{k = j}^{i-1} L_ik X_kj   (i > j),
// a chain of at most 8 tile steps.  Entries whose row or column is m or more
// are never read by the products.
constexpr int kInvRows = kSolveBlockTiles * NB;
constexpr int64_t kInvSlot = static_cast<int64_t>(kInvRows) * kInvRows;

// One wave per diagonal tile: column c of inv(L_ii) by forward substitution
// of e_c, two lanes per column summing alternate terms, the tile's columns in
// LDS (each lane reads back only its own stores: both lanes of a column hold
// the same value).
__global__ void __launch_bounds__(64) k_inv_diag_tiles(int n, const double* L, double* Linv, int P, const int32_t* status,
                                                       const lm_state* lm, int gated) {
    if ((gated && lm->done) || *status != 0) return;
    __shared__ double ys[NB * NB];
    const int p = blockIdx.x / kSolveBlockTiles, t = blockIdx.x % kSolveBlockTiles;
    const int R = p * kInvRows + t * NB;  // first row of the tile in the system
    if (R >= n) return;
    const int mt = (n - R) < NB ? (n - R) : NB;
    const int col = threadIdx.x & (NB - 1), half = threadIdx.x >> 5;
    double* Bc = Linv + p * kInvSlot;
    double* Br = Linv + (P + p) * kInvSlot;
    const double* Lt = L + static_cast<int64_t>(R) * n + R;
    for (int i = 0; i < NB; ++i) {
        double acc = 0.0;
        if (i < mt) {
            const double* Li = Lt + static_cast<int64_t>(i) * n;
            for (int k = half; k < i; k += 2) acc = fma(Li[k], ys[k * NB + col], acc);
        }
        const double other = __shfl_xor(acc, 32);
        const double sum = half == 0 ? acc + other : other + acc;
        double y = 0.0;
        if (i < mt && col < mt && col <= i) {
            const double d = Lt[static_cast<int64_t>(i) * n + i];
            y = col == i ? 1.0 / d : -sum / d;
        }
        ys[i * NB + col] = y;
        if (half == 0) {
            const int gi = t * NB + i, gj = t * NB + col;
            Bc[gj * kInvRows + gi] = y;
            Br[gi * kInvRows + gj] = y;
        }
    }
}

// One workgroup (256 threads) per tile column j of a block: the tiles above the
// diagonal set to zero, then X_ij for i = j + 1 .. in order.  Thread (r, c4)
// owns row r, columns c4 .. c4 + 3 of the tile being formed.  X_jj is read from
// the row-major slot (formed by the previous launch), the tiles formed here
// from LDS.
constexpr int kInvKeep = kSolveBlockTiles - 2;  // tiles X_ij kept in LDS for later steps

__global__ void __launch_bounds__(256) k_inv_offdiag_tiles(int n, const double* L, double* Linv, int P,
                                                          const int32_t* status, const lm_state* lm, int gated) {
    if ((gated && lm->done) || *status != 0) return;
    __shared__ double xs[kInvKeep][NB][NB];
    __shared__ double ts[NB][NB + 1];
    const int p = blockIdx.x / kSolveBlockTiles, j = blockIdx.x % kSolveBlockTiles;
    const int r0 = p * kInvRows;
    const int m = (n - r0) < kInvRows ? (n - r0) : kInvRows;
    const int nt = (m + NB - 1) / NB;
    if (j >= nt) return;
    const int r = threadIdx.x >> 3, c4 = (threadIdx.x & 7) * 4;
    double* Bc = Linv + p * kInvSlot;
    double* Br = Linv + (P + p) * kInvSlot;
    auto put = [&](int ti, const double (&v)[4]) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int gi = ti * NB + r, gj = j * NB + c4 + u;
            Bc[gj * kInvRows + gi] = v[u];
            Br[gi * kInvRows + gj] = v[u];
        }
    };
    const double zero[4] = {0.0, 0.0, 0.0, 0.0};
    for (int ti = 0; ti < j; ++ti) put(ti, zero);
    for (int i = j + 1; i < nt; ++i) {
        // T = sum_k L_ik X_kj, row r of it in acc
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const int gr = r0 + i * NB + r;  // row of L
        for (int k = j; k < i; ++k) {
            const int gc = r0 + k * NB;  // first column of tile k
#pragma unroll 8
            for (int kk = 0; kk < NB; ++kk) {
                const double a = (gr < n && gc + kk < n) ? L[static_cast<int64_t>(gr) * n + gc + kk] : 0.0;
                double x[4];
                if (k == j) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) x[u] = Br[(j * NB + kk) * kInvRows + j * NB + c4 + u];
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u) x[u] = xs[k - j - 1][kk][c4 + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) acc[u] = fma(a, x[u], acc[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) ts[r][c4 + u] = acc[u];
        __syncthreads();
        // X_ij = -X_ii T
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        for (int q = 0; q < NB; ++q) {
            const double e = Br[(i * NB + r) * kInvRows + i * NB + q];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = fma(e, ts[q][c4 + u], v[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = -v[u];
        put(i, v);
        if (i - j - 1 < kInvKeep) {
#pragma unroll
            for (int u = 0; u < 4; ++u) xs[i - j - 1][r][c4 + u] = v[u];
        }
        __syncthreads();
    }
}

// out_r = sum_q M[q][r] v_q for r < m: thread pair per r
// (alternate q, eight loads in flight), M column-major in a 256 x 256 slot so
// that the threads of a wave read consecutive addresses; the two partial sums
// added in a fixed order.  Forward (lower): M = inv(L_bb) column-major, q <= r;
// backward (upper): M = inv(L_bb) row-major, i.e. inv(L_bb)^T column-major, all q.
template <bool kForward>
__device__ __forceinline__ void block_inverse_product(int m, const double* M, double* x, double* xs, double* part) {
    for (int i = threadIdx.x; i < m; i += 512) xs[i] = x[i];
    __syncthreads();
    const int r = threadIdx.x & (kInvRows - 1), half = threadIdx.x / kInvRows;
    double acc = 0.0;
    if (r < m) {
        // backward: from q = 0 too (inv[q][r] = 0 for q < r), so that the
        // threads of a wave read the same row of the slot at each step
        const int qb = half, qe = kForward ? r + 1 : m;
        int q = qb;
        for (; q + 14 < qe; q += 16) {
            double a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = M[(q + 2 * u) * kInvRows + r];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = fma(a[u], xs[q + 2 * u], acc);
        }
        for (; q < qe; q += 2) acc = fma(M[q * kInvRows + r], xs[q], acc);
    }
    if (half == 1) part[r] = acc;
    __syncthreads();
    if (half == 0 && r < m) x[r] = acc + part[r];
}

__global__ void __launch_bounds__(512) k_trsv_block_fwd_inv(int n, const double* Linv, double* x, int p,
                                                            const int32_t* status, const lm_state* lm, int gated) {
    if ((gated && lm->done) || *status != 0) return;
    __shared__ double xs[kInvRows], part[kInvRows];
    const int r0 = p * kInvRows;
    const int m = (n - r0) < kInvRows ? (n - r0) : kInvRows;
    block_inverse_product<true>(m, Linv + p * kInvSlot, x + r0, xs, part);
}

__global__ void __launch_bounds__(512) k_trsv_block_bwd_inv(int n, const double* Linv, double* x, int p, int P,
                                                            const int32_t* status, lm_state* lm, int gated) {
    if ((gated && lm->done) || *status != 0) return;
    __shared__ double xs[kInvRows], part[kInvRows];
    const int r0 = p * kInvRows;
    const int m = (n - r0) < kInvRows ? (n - r0) : kInvRows;
    block_inverse_product<false>(m, Linv + (P + p) * kInvSlot, x + r0, xs, part);
    if (p == 0 && threadIdx.x == 0) lm->solve_ok = 1;  // the solve's last launch
}

void launch_trsv_blocked(int n, const double* L, const double* b, double* x, const dev_work64& w, hipStream_t s,
                         int gated) {
    const int T = (n + NB - 1) / NB;
    const int P = (T + kSolveBlockTiles - 1) / kSolveBlockTiles;
    k_trsv_begin<<<1, 1024, 0, s>>>(n, L, b, x, w.chol_status, w.lm, gated);
    for (int p = 0; p < P; ++p) {
        k_trsv_block_fwd_inv<<<1, 512, 0, s>>>(n, w.Lt, x, p, w.chol_status, w.lm, gated);
        const int c1 = ((p + 1) * kSolveBlockTiles * NB) < n ? ((p + 1) * kSolveBlockTiles * NB) : n;
        if (c1 < n)
            k_trsv_rows_update<<<(n - c1 + 3) / 4, 256, 0, s>>>(n, L, x, p, w.tile_first, w.chol_status, w.lm, gated);
    }
    for (int p = P - 1; p >= 0; --p) {
        k_trsv_block_bwd_inv<<<1, 512, 0, s>>>(n, w.Lt, x, p, P, w.chol_status, w.lm, gated);
        const int c0 = p * kSolveBlockTiles * NB;
        if (c0 > 0)
            k_trsv_cols_update<<<(c0 + 63) / 64, 512, 0, s>>>(n, L, x, p, w.tile_first, w.chol_status, w.lm, gated);
    }
}

// Substitution in one workgroup; sets lm->solve_ok after a full solve.
// forward_only: the forward pass in the reference's order (wg::triangular_solve,
// the binary64 parity path); otherwise both passes pipelined
// (wg::triangular_solve_fast).  x lives in LDS for n <= kTrsvLds, otherwise in
// the output.
constexpr int kTrsvLds = 8192;

template <bool kLds>
__global__ void __launch_bounds__(512) k_trsv(int n, const double* L, const double* b, double* xout,
                                               const int32_t* tile_first, const int32_t* tile_last, int32_t* status,
                                               lm_state* lm, int gated, int forward_only) {
    if (gated && lm->done) return;
    if (*status != 0) {
        if (threadIdx.x == 0) lm->solve_ok = 0;
        return;
    }
    __shared__ double sh[1024];
    __shared__ double buf[6][NB][NB + 1];
    __shared__ double xl[kLds ? kTrsvLds : 1];
    __shared__ int flag;
    double* x = kLds ? xl : xout;
    const int rc = forward_only ? wg::triangular_solve(n, L, b, x, tile_first, tile_last, buf[0], sh, &flag, true)
                                : wg::triangular_solve_fast<512>(n, L, b, x, tile_first, tile_last, buf, sh, &flag);
    if (rc != 0) {
        if (threadIdx.x == 0) {
            *status = rc;
            lm->solve_ok = 0;
        }
        return;
    }
    if (kLds) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) xout[i] = x[i];
    }
    if (threadIdx.x == 0 && !forward_only) lm->solve_ok = 1;
}

void launch_trsv(int n, const double* L, const double* b, double* x, const dev_work64& w, hipStream_t s, int gated,
                 int forward_only) {
    if (!forward_only && (n + NB - 1) / NB > kBlockedSolveTiles) {
        launch_trsv_blocked(n, L, b, x, w, s, gated);
        return;
    }
    if (n <= kTrsvLds)
        k_trsv<true><<<1, 512, 0, s>>>(n, L, b, x, w.tile_first, w.tile_last, w.chol_status, w.lm, gated, forward_only);
    else
        k_trsv<false><<<1, 512, 0, s>>>(n, L, b, x, w.tile_first, w.tile_last, w.chol_status, w.lm, gated, forward_only);
}

// A system without poses: the reference's 0 x 0 decompose and solve succeed.
__global__ void k_empty_ok(lm_state* lm, int gated) {
    if (gated && lm->done) return;
    lm->solve_ok = 1;
}

}  // namespace

void launch_cholesky_solve64(int32_t n, const dev_work64& w, hipStream_t s, bool gated, bool exact, double floor_rel) {
    const int g = gated ? 1 : 0;
    if (n <= 0) {
        k_empty_ok<<<1, 1, 0, s>>>(w.lm, g);
        return;
    }
    if (band_path(n, w, exact)) {
        ZB_TIME_BEGIN(w, K_CHOLESKY, s);
        launch_chol_band(n, w, s, gated, floor_rel);
        ZB_TIME_END(w, s);
        return;
    }
    ZB_TIME_BEGIN(w, K_CHOLESKY, s);
    k_chol_begin<<<1, 1024, 0, s>>>(n, w.S, w.chol_scale, w.chol_status, w.lm, g, floor_rel);
    const int64_t tiles = (n + NB - 1) / NB;
    const bool fused = w.factor_path == 1 || (w.factor_path == 0 && w.panel_tiles <= kFusedTiles);
    if (fused && w.panel_tiles <= kFusedTiles) {
        if (exact)
            k_chol_fused<false><<<1, 512, 0, s>>>(n, w.S, w.tile_last, w.chol_scale, w.chol_status, w.lm, g);
        else
            k_chol_fused<true><<<1, 512, 0, s>>>(n, w.S, w.tile_last, w.chol_scale, w.chol_status, w.lm, g);
    } else {
        // exact: one block holding every panel (the reference's operation order);
        // otherwise blocks of kBlockTiles panels, each followed by its MFMA update
        const int64_t bt = exact ? tiles : kBlockTiles;
        for (int64_t kb0 = 0; kb0 < tiles; kb0 += bt) {
            const int64_t kb1 = (kb0 + bt) < tiles ? (kb0 + bt) : tiles;
            for (int64_t k = kb0; k < kb1; ++k) {
                const int64_t m = tiles - k - 1, wd = kb1 - 1 - k;
                const int64_t grid = (tiles - k) + (k > kb0 ? step_trailing_tiles(m, wd) : 0);
                const int ik = static_cast<int>(k), ib0 = static_cast<int>(kb0), ib1 = static_cast<int>(kb1);
                if (exact)
                    k_chol_step<false><<<static_cast<unsigned>(grid), 256, 0, s>>>(
                        n, w.S, w.chol_scale, w.chol_status, ik, ib0, ib1, w.tile_first, w.lm, g);
                else
                    k_chol_step<true><<<static_cast<unsigned>(grid), 256, 0, s>>>(
                        n, w.S, w.chol_scale, w.chol_status, ik, ib0, ib1, w.tile_first, w.lm, g);
            }
            const int64_t m = tiles - kb1;
            if (m > 0)
                k_chol_syrk<<<static_cast<unsigned>(m * (m + 1) / 2), 256, 0, s>>>(
                    n, w.S, static_cast<int>(kb0), static_cast<int>(kb1), w.tile_first, w.chol_status, w.lm, g);
        }
    }
    if (!exact && tiles > kBlockedSolveTiles) {
        const int64_t P = (tiles + kSolveBlockTiles - 1) / kSolveBlockTiles;
        const unsigned tiles_p = static_cast<unsigned>(P * kSolveBlockTiles);
        k_inv_diag_tiles<<<tiles_p, 64, 0, s>>>(n, w.S, w.Lt, static_cast<int>(P), w.chol_status, w.lm, g);
        k_inv_offdiag_tiles<<<tiles_p, 256, 0, s>>>(n, w.S, w.Lt, static_cast<int>(P), w.chol_status, w.lm, g);
    }
    ZB_TIME_END(w, s);
    ZB_TIME_BEGIN(w, K_TRSV, s);
    launch_trsv(n, w.S, w.bs, w.dxp, w, s, g, exact ? 1 : 0);
    if (exact) {
        const int tiles32 = (n + 31) / 32;
        k_transpose_lower<<<dim3(tiles32, tiles32), 256, 0, s>>>(n, w.S, w.Lt, w.tile_last, w.chol_status, w.lm, g);
        if (n <= kBackLds)
            k_backward_exact_lds<<<1, 64, 0, s>>>(n, w.S, w.Lt, w.dxp, w.tile_last, w.chol_status, w.lm, g);
        else
            k_backward_exact<<<1, 64, 0, s>>>(n, w.S, w.Lt, w.dxp, w.tile_last, w.chol_status, w.lm, g);
    }
    ZB_TIME_END(w, s);
}

void launch_cholesky_resolve64(int32_t n, const dev_work64& w, const double* b, double* x, hipStream_t s, bool gated) {
    if (n <= 0) return;
    if (band_path(n, w, false)) {
        launch_band_solve(n, w, b, x, s, gated);
        return;
    }
    launch_trsv(n, w.S, b, x, w, s, gated ? 1 : 0, 0);
}

}  // namespace zb
