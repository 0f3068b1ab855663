// Binary64 bundle-adjustment kernels for gfx950 (compiled with -ffp-contract=off).
//
// Per-observation arithmetic is the reference's IEEE sequence (ba_math.hpp).
// The per-vertex sums that make_hessian forms edge by edge (H_pp, H_ll, b, W;
// factor_graph.hpp:664-727) are formed here by one lane per sum walking the
// vertex's observations in insertion order, so they equal the reference's sums
// bit for bit.  Sums over all observations (chi2, the LM scale) and over a
// reduced-system block (the Schur products) use fixed-shape trees: deterministic
// run to run, different from the reference's sequential order only by rounding.
//
// Every kernel of an LM attempt first reads the device-side controller state and
// returns at once when the attempt does not apply (solve finished, step not
// applied, ...), so the host can enqueue attempts without reading anything back.
#include "ba_math.hpp"
#include "common.hpp"
#include "lm_control.hpp"
#include "mixed_items.hpp"

#include <cmath>

namespace zb {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ const double* cam_of(const dev_problem& p, int32_t param) {
    return p.cams + 4 * (param & 0xFFFF);
}
__device__ __forceinline__ void loss_of(const dev_problem& p, int32_t param, int32_t& kind, double& delta) {
    const int li = param >> 16;
    kind = p.loss_kind[li];
    delta = p.loss_delta[li];
}
__device__ __forceinline__ void info_of(const dev_problem& p, int s, double W[4]) {
    if (p.identity_info) {
        W[0] = 1.0; W[1] = 0.0; W[2] = 0.0; W[3] = 1.0;
    } else {
        const double* w = p.obs_info + 4L * s;
        W[0] = w[0]; W[1] = w[1]; W[2] = w[2]; W[3] = w[3];
    }
}
__device__ __forceinline__ void load3(const double* a, double o[3]) { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; }
__device__ __forceinline__ void load7(const double* a, double o[7]) {
#pragma unroll
    for (int i = 0; i < 7; ++i) o[i] = a[i];
}

// Block-wide sum of one value per thread into out (fixed tree, deterministic).
__device__ double block_sum(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] = sh[t] + sh[t + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// ---- cost: chi2 = sum of robust costs (factor_graph.hpp:618-629) -------------
__global__ void __launch_bounds__(kBlock) k_cost(dev_problem p, double* pose, double* point, double* resid,
                                                 double* partial, const lm_state* lm, int gate, int recompute,
                                                 double* terms) {
    if (gate == 1 && lm->done) return;
    if (gate == 2 && (lm->done || !lm->step_applied)) return;
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < p.no; s += gridDim.x * blockDim.x) {
        const int32_t param = p.obs_param[s];
        double r[2];
        if (recompute) {
            double P[7], X[3], Pc[3], uv[2] = {p.obs_uv[2L * s], p.obs_uv[2L * s + 1]};
            load7(pose + 7L * p.obs_pose[s], P);
            load3(point + 3L * p.obs_point[s], X);
            pose_apply(P, X, Pc);
            residual(Pc, uv, cam_of(p, param), r);
            resid[2L * s] = r[0];
            resid[2L * s + 1] = r[1];
        } else {
            r[0] = resid[2L * s];
            r[1] = resid[2L * s + 1];
        }
        double W[4];
        info_of(p, s, W);
        int32_t kind;
        double delta;
        loss_of(p, param, kind, delta);
        const double c = robust_chi2(kind, delta, r, W);
        if (terms) terms[p.obs_orig[s]] = c;
        else acc = acc + c;
    }
    if (terms) return;
    const double b = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// Lane-uniform read of lane j's value.
__device__ __forceinline__ double readlane_d(double v, int j) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), j);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), j);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Sequential sum a[0] + a[1] + ... from 0.0, in index order, as the reference's
// accumulation loops (factor_graph.hpp:556-560, 624-628).  One wave: all lanes
// stream the next 1024 values into LDS while lane 0 adds the current 1024, so
// only the dependent additions remain on the critical path.
constexpr int kSeqChunk = 1024;
__global__ void __launch_bounds__(64) k_seqsum(const double* a, int64_t n, lm_state* lm, int target, int gate,
                                               double* out) {
    if (gate == 1 && lm->done) return;
    if (gate == 2 && (lm->done || !lm->step_applied)) return;
    __shared__ double buf[2][kSeqChunk];
    const int lane = threadIdx.x;
    double reg[kSeqChunk / 64];
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kSeqChunk / 64; ++k) {
        const int64_t i = k * 64 + lane;
        buf[0][k * 64 + lane] = (i < n) ? a[i] : 0.0;
    }
    for (int64_t base = 0, c = 0; base < n; base += kSeqChunk, ++c) {
        const int64_t nbase = base + kSeqChunk;
        if (nbase < n) {
#pragma unroll
            for (int k = 0; k < kSeqChunk / 64; ++k) {
                const int64_t i = nbase + k * 64 + lane;
                reg[k] = (i < n) ? a[i] : 0.0;
            }
        }
        __syncthreads();
        if (lane == 0) {
            const double* b = buf[c & 1];
            const int cnt = (n - base) < kSeqChunk ? static_cast<int>(n - base) : kSeqChunk;
            if (cnt == kSeqChunk) {
#pragma unroll 32
                for (int k = 0; k < kSeqChunk; ++k) acc = acc + b[k];
            } else {
                for (int k = 0; k < cnt; ++k) acc = acc + b[k];
            }
        }
        __syncthreads();
        if (nbase < n) {
#pragma unroll
            for (int k = 0; k < kSeqChunk / 64; ++k) buf[(c + 1) & 1][k * 64 + lane] = reg[k];
        }
    }
    if (lane == 0) {
        if (target == 0) lm->chi2 = acc;
        else if (target == 1) lm->chi2_new = acc;
        else if (target == 2) lm->scale = acc + 1e-3;
        else if (out) *out = acc;
    }
}

__global__ void __launch_bounds__(1024) k_sum_partials(const double* partial, int n, double* out, lm_state* lm,
                                                       int target, int gate, int decide = 0) {
    if (gate == 1 && lm->done) return;
    if (gate == 2 && (lm->done || !lm->step_applied)) return;
    __shared__ double sh[1024];
    const double v = threadIdx.x < n ? partial[threadIdx.x] : 0.0;
    const double total = block_sum(v, sh);
    if (threadIdx.x == 0) {
        if (target == 0) lm->chi2 = total;
        else if (target == 1) lm->chi2_new = total;
        else if (target == 2) lm->scale = total + 1e-3;  // factor_graph.hpp:556-560
        else if (out) *out = total;
        if (decide) lm_decide(lm);  // the attempt's accept / reject on the new cost
    }
}

// ---- linearisation: make_hessian (factor_graph.hpp:648-732) -------------------
// One thread per point walks its observations in insertion order: H_ll, b_l and
// the W blocks accumulate in registers/HBM exactly as the reference adds them;
// the pose-side terms of each observation go to pose_terms for k_reduce_pose.
__global__ void __launch_bounds__(kBlock) k_linearise_points(dev_problem p, const double* pose, const double* point,
                                                             double* pterms, double* Hll, double* bl, double* Wbuf,
                                                             const lm_state* lm, int gated, uint8_t* frozen) {
    if (gated && (lm->done || !lm->need_linearize)) return;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= p.nl) return;
    const bool lfixed = p.point_fixed[l] != 0;
    double X[3];
    load3(point + 3L * l, X);
    double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double b[3] = {0, 0, 0};
    bool behind = false;
    const int s0 = p.lm_start[l], s1 = p.lm_start[l + 1];
    for (int s = s0; s < s1; ++s) {
        const int ip = p.obs_pose[s];
        const bool pfixed = p.pose_fixed[ip] != 0;
        if (pfixed && lfixed) continue;
        const int32_t param = p.obs_param[s];
        const double* K = cam_of(p, param);
        double P[7], Pc[3], q[4], r[2], Jp[12], Jl[6], W[4], Wr[4], drho;
        load7(pose + 7L * ip, P);
        pose_apply(P, X, Pc);
        behind = behind || behind_camera(Pc);
        const double uv[2] = {p.obs_uv[2L * s], p.obs_uv[2L * s + 1]};
        residual(Pc, uv, K, r);
        pose_quat(P, q);
        jacobians(Pc, q, K, Jp, Jl);
        info_of(p, s, W);
        int32_t kind;
        double delta;
        loss_of(p, param, kind, delta);
        robust_info(kind, delta, W, r, drho, Wr);
        if (!pfixed) {
            double H21[21], Wpl[18], v[6];
            pose_terms(Jp, Jl, Wr, W, r, drho, H21, Wpl, v);
            double* out = pterms + 27L * p.obs_pose_pos[s];
#pragma unroll
            for (int i = 0; i < 21; ++i) out[i] = H21[i];
#pragma unroll
            for (int i = 0; i < 6; ++i) out[21 + i] = v[i];
            if (!lfixed) {
                const int pr = p.obs_pair[s];
                double* Wp = Wbuf + 18L * pr;
                if (p.pair_first[pr] == s) {
#pragma unroll
                    for (int i = 0; i < 18; ++i) Wp[i] = Wpl[i];
                } else {
#pragma unroll
                    for (int i = 0; i < 18; ++i) Wp[i] = Wp[i] + Wpl[i];
                }
            }
        }
        if (!lfixed) {
            double Hl[9], v[3];
            point_terms(Jl, Wr, W, r, drho, Hl, v);
#pragma unroll
            for (int i = 0; i < 9; ++i) H[i] = H[i] + Hl[i];
#pragma unroll
            for (int i = 0; i < 3; ++i) b[i] = b[i] - v[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) Hll[9L * l + i] = H[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) bl[3L * l + i] = b[i];
    if (frozen) frozen[l] = behind ? 1 : 0;
}

// One wave per pose; lane j < 27 sums term j over the pose's observations in
// insertion order (H_pp += H; b_p -= v from zero, factor_graph.hpp:696, 725).
__global__ void __launch_bounds__(256) k_reduce_pose(dev_problem p, const double* pose_terms, double* Hpp, double* bp,
                                                     const lm_state* lm, int gated) {
    if (gated && (lm->done || !lm->need_linearize)) return;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= p.np || lane >= 27) return;
    double acc = 0.0;
    if (!p.pose_fixed[wave]) {
        // terms of the pose's observations are contiguous, in insertion order
        const int t0 = p.pose_start[wave], t1 = p.pose_start[wave + 1];
        const double* src = pose_terms + lane;
        const double sign = lane < 21 ? 1.0 : -1.0;  // H_pp += H;  b_p -= v
        constexpr int U = 16;
        int t = t0;
        for (; t + U <= t1; t += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = src[27L * (t + u)];
            if (sign > 0) {
#pragma unroll
                for (int u = 0; u < U; ++u) acc = acc + v[u];
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) acc = acc - v[u];
            }
        }
        for (; t < t1; ++t) acc = sign > 0 ? acc + src[27L * t] : acc - src[27L * t];
    }
    if (lane < 21) Hpp[21L * wave + lane] = acc;
    else bp[6L * wave + (lane - 21)] = acc;
}

// The same sums in kPoseChunks fixed chunks (reassociating precision): one wave
// per (pose, chunk), lane j < 27 summing term j, then one thread per (pose, term)
// adding the chunks in order.
__global__ void __launch_bounds__(256) k_reduce_pose_part(dev_problem p, const double* pose_terms, double* part,
                                                          const lm_state* lm, int gated) {
    if (gated && (lm->done || !lm->need_linearize)) return;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= p.np * kPoseChunks || lane >= 27) return;
    const int i = wave / kPoseChunks, c = wave % kPoseChunks;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (!p.pose_fixed[i]) {
        int t0, t1;
        mixed::chunk_range(p.pose_start[i], p.pose_start[i + 1], c, t0, t1);
        const double* src = pose_terms + lane;
        constexpr int U = 8;
        int t = t0;
        for (; t + U <= t1; t += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = src[27L * (t + u)];
#pragma unroll
            for (int u = 0; u < U; ++u) acc[u & 3] += v[u];
        }
        for (; t < t1; ++t) acc[0] += src[27L * t];
    }
    part[static_cast<int64_t>(wave) * 27 + lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

__global__ void __launch_bounds__(256) k_reduce_pose_chunks(dev_problem p, const double* part, double* Hpp,
                                                            double* bp, const lm_state* lm, int gated) {
    if (gated && (lm->done || !lm->need_linearize)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.np * 27) return;
    const int i = t / 27, j = t % 27;
    const double* src = part + static_cast<int64_t>(i) * kPoseChunks * 27 + j;
    double v[kPoseChunks];
#pragma unroll
    for (int c = 0; c < kPoseChunks; ++c) v[c] = src[27 * c];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < kPoseChunks; ++c) a[c & 3] += v[c];
    const double sum = (a[0] + a[1]) + (a[2] + a[3]);
    if (j < 21) Hpp[21L * i + j] = sum;  // H_pp += H
    else bp[6L * i + (j - 21)] = -sum;   // b_p -= v
}

__global__ void k_linearise_done(lm_state* lm) { lm_linearise_done(lm); }
__global__ void k_linearise_done_end(lm_state* lm) {
    lm_linearise_done(lm);
    lm_end_attempt(lm);
}

// ---- LM initialisation, factor_graph.hpp:511-522 -----------------------------
template <class TS>
__global__ void __launch_bounds__(1024) k_init_lm(dev_problem p, const TS* Hpp, const TS* Hll, lm_state* lm) {
    __shared__ max_scan sh[1024];
    const int64_t total = 6L * p.np + 3L * p.nl;
    const int64_t per = (total + blockDim.x - 1) / blockDim.x;
    const int64_t i0 = threadIdx.x * per, i1 = (i0 + per < total) ? i0 + per : total;
    max_scan m = {0.0, 0, 0, 1};
    for (int64_t i = i0; i < i1; ++i) {
        double h;
        if (i < 6L * p.np) {
            const int64_t pose = i / 6, k = i % 6;
            h = static_cast<double>(Hpp[21 * pose + k * (k + 1) / 2 + k]);
        } else {
            const int64_t j = i - 6L * p.np;
            h = static_cast<double>(Hll[9 * (j / 3) + 4 * (j % 3)]);
        }
        m = scan_push(m, h);
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        max_scan acc = {0.0, 0, 0, 1};
        for (unsigned t = 0; t < blockDim.x; ++t) acc = scan_combine(acc, sh[t]);
        lm_init(lm, scan_max_diagonal(acc));
    }
}

// ---- damping + point-block inverses, factor_graph.hpp:742-765 ------------------
__global__ void __launch_bounds__(kBlock) k_damp_invert(dev_problem p, const double* Hll, double* Hinv, const lm_state* lm,
                                                        const double* lambda_override, const uint8_t* frozen) {
    if (!lambda_override && lm->done) return;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= p.nl) return;
    const double lam = lambda_override ? *lambda_override : lm->lambda;
    double H[9], inv[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = Hll[9L * l + i];
    H[0] += lam; H[4] += lam; H[8] += lam;
    if (p.point_fixed[l] || (frozen && frozen[l])) {
#pragma unroll
        for (int i = 0; i < 9; ++i) inv[i] = 0.0;
    } else {
        invert3(H, inv);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) Hinv[9L * l + i] = inv[i];
}

// Y = W * H_ll^-1 per W block (matrix_sparse_block_diagonal.hpp:107-122).
__global__ void __launch_bounds__(kBlock) k_form_Y(dev_problem p, const double* Wbuf, const double* Hinv, double* Y,
                                                   const lm_state* lm, const double* lambda_override) {
    if (!lambda_override && lm->done) return;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= p.npair) return;
    double Wa[18], Hi[9], o[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) Wa[i] = Wbuf[18L * a + i];
    const int l = p.pair_point[a];
#pragma unroll
    for (int i = 0; i < 9; ++i) Hi[i] = Hinv[9L * l + i];
    mul<6, 3, 3>(Wa, Hi, o);
#pragma unroll
    for (int i = 0; i < 18; ++i) Y[18L * a + i] = o[i];
}

__global__ void __launch_bounds__(kBlock) k_fill_zero(double* a, int64_t n, const lm_state* lm, int gated) {
    if (gated && lm->done) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = 0.0;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

// Reduced camera system, one workgroup per 6x6 block (row >= col):
//   S_ik = [i == k] (H_pp,i + lambda I) - sum_l Y_il W_kl^T
//   b_s,i = b_p,i - sum_l Y_il b_l
// (factor_graph.hpp:768-781 with matrix_sparse_block.hpp:158-196), by the
// binary64 instance of the mixed path's block code (mixed::schur_block: four
// waves over the block's contributions, a reduce-scatter per wave, the waves'
// totals added in order), which the batched binary64 windows use too.
// kInFlight changes the registers in use, not the result.  Blocks are numbered
// row by row, so the XCD chunks of 64 consecutive blocks share their rows' Y.
template <int kInFlight>
__global__ void __launch_bounds__(mixed::kSchurThreads) k_schur(dev_problem p, const double* Y, const double* Wbuf,
                                                                 const double* Hpp, const double* bp, const double* bl,
                                                                 double* S, double* bs, const lm_state* lm,
                                                                 const double* lambda_override) {
    if (!lambda_override && lm->done) return;
    __shared__ mixed::schur_lds<double> red;
    const int blk = xcd_block<64>(blockIdx.x, gridDim.x);
    if (blk < p.nblk)
        mixed::schur_block<double, kInFlight>(p, Y, Wbuf, Hpp, bp, bl, S, bs,
                                              lambda_override ? *lambda_override : lm->lambda, blk,
                                              static_cast<int>(threadIdx.x), red);
}

__device__ __forceinline__ int order_set(const lm_state* lm) { return lm->lin_count <= 1 ? 0 : 1; }

// Reduced camera system in the reference's order, one wave per block, lane e < 36
// owning entry (e / 6, e % 6):  P = Y_a W_c^T for the first contribution, then
// P = P + Y_a W_c^T in the order the reference's Y map iterates
// (matrix_sparse_block.hpp:164-177); S = (H_pp + lambda I) - P on the diagonal,
// S = -P off it (matrix_sparse_block_diagonal.hpp:129-140).
__global__ void __launch_bounds__(256) k_schur_exact(dev_problem p, const double* Y, const double* Wbuf,
                                                     const double* Hpp, double* S, const lm_state* lm) {
    if (lm->done) return;
    const int blk = (xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (blk >= p.nblk || lane >= 36) return;
    const int row = p.blk_row[blk], col = p.blk_col[blk];
    const int r = lane / 6, c = lane % 6;
    const int2* contrib = p.contrib_x[order_set(lm)];
    const int64_t c0 = p.blk_start[blk], c1 = p.blk_start[blk + 1];
    double P = 0.0;
    constexpr int U = 8;  // operands of U contributions are loaded before the ordered additions
    int64_t t = c0;
    for (; t + U <= c1; t += U) {
        double y[U][3], wv[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int2 ab = contrib[t + u];
            const double* Ya = Y + 18L * ab.x + 3 * r;
            const double* Wc = Wbuf + 18L * ab.y + 3 * c;
            y[u][0] = Ya[0]; y[u][1] = Ya[1]; y[u][2] = Ya[2];
            wv[u][0] = Wc[0]; wv[u][1] = Wc[1]; wv[u][2] = Wc[2];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            double prod = 0;
            prod += y[u][0] * wv[u][0];
            prod += y[u][1] * wv[u][1];
            prod += y[u][2] * wv[u][2];
            P = (t + u == c0) ? prod : P + prod;
        }
    }
    for (; t < c1; ++t) {
        const int2 ab = contrib[t];
        const double* Ya = Y + 18L * ab.x + 3 * r;
        const double* Wc = Wbuf + 18L * ab.y + 3 * c;
        double prod = 0;
        prod += Ya[0] * Wc[0];
        prod += Ya[1] * Wc[1];
        prod += Ya[2] * Wc[2];
        P = (t == c0) ? prod : P + prod;
    }
    const int64_t n = p.n;
    double* out = S + (6L * row + r) * n + 6L * col + c;
    if (row == col) {
        if (c <= r) {
            double base = Hpp[21L * row + r * (r + 1) / 2 + c];
            if (r == c) base += lm->lambda;
            *out = (c1 > c0) ? base - P : base;
        }
    } else {
        *out = -P;
    }
}

// b_s = b_p - Y b_l with Y b_l accumulated from zero in Y-map order
// (matrix_sparse_block.hpp:184-196), one wave per pose, lanes 0..5.
__global__ void __launch_bounds__(256) k_bs_exact(dev_problem p, const double* Y, const double* bp, const double* bl,
                                                  double* bs, const lm_state* lm) {
    if (lm->done) return;
    const int pose = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (pose >= p.np || lane >= 6) return;
    const int32_t* pairs = p.pose_pairs_x[order_set(lm)];
    double yb = 0.0;
    constexpr int U = 8;
    const int t0 = p.pose_pair_start[pose], t1 = p.pose_pair_start[pose + 1];
    int t = t0;
    for (; t + U <= t1; t += U) {
        double y[U][3], bv[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int a = pairs[t + u];
            const int l = p.pair_point[a];
            const double* Ya = Y + 18L * a + 3 * lane;
            y[u][0] = Ya[0]; y[u][1] = Ya[1]; y[u][2] = Ya[2];
            bv[u][0] = bl[3L * l]; bv[u][1] = bl[3L * l + 1]; bv[u][2] = bl[3L * l + 2];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            double prod = 0;
            prod += y[u][0] * bv[u][0];
            prod += y[u][1] * bv[u][1];
            prod += y[u][2] * bv[u][2];
            yb = yb + prod;
        }
    }
    for (; t < t1; ++t) {
        const int a = pairs[t];
        const int l = p.pair_point[a];
        const double* Ya = Y + 18L * a + 3 * lane;
        double prod = 0;
        prod += Ya[0] * bl[3L * l];
        prod += Ya[1] * bl[3L * l + 1];
        prod += Ya[2] * bl[3L * l + 2];
        yb = yb + prod;
    }
    bs[6L * pose + lane] = bp[6L * pose + lane] - yb;
}

__global__ void k_after_solve(lm_state* lm) { lm_after_solve(lm); }

// ---- back-substitution + state update + LM scale -------------------------------
// Threads [0, np) update poses (T <- exp(dx) T), threads [np, np + nl) solve
// dx_l = H_ll^-1 (b_l - sum W^T dx_p) (factor_graph.hpp:799-800) and update
// points (factor_graph.hpp:809-830).  Each thread also adds its slots to the LM
// scale sum_i dx_i (lambda dx_i + b_i) (factor_graph.hpp:556-559).
// TS is the storage type of the W and H_ll^-1 blocks (binary32 on the mixed
// path); the gradient and the arithmetic are binary64 either way.
template <class TS>
__global__ void __launch_bounds__(kBlock) k_backsub_update(dev_problem p, double* pose, double* pose_backup, double* point,
                                                           double* point_backup, const TS* Wbuf, const TS* Hinv,
                                                           const double* dxp, double* dxl, const double* bp,
                                                           const double* bl, double* partial, const lm_state* lm,
                                                           int gated, int apply, double* sterms, int given_dxl = 0) {
    if (gated && (lm->done || !lm->step_applied)) return;
    __shared__ double sh[kBlock];
    const double lam = lm->lambda;
    double acc = 0.0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < p.np + p.nl; t += gridDim.x * blockDim.x) {
    if (t < p.np) {
        double d[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) d[i] = dxp[6L * t + i];
        if (sterms) {
#pragma unroll
            for (int i = 0; i < 6; ++i) sterms[6L * t + i] = d[i] * (lam * d[i] + bp[6L * t + i]);
        } else {
#pragma unroll
            for (int i = 0; i < 6; ++i) acc = acc + d[i] * (lam * d[i] + bp[6L * t + i]);
        }
        if (apply && !p.pose_fixed[t]) {
            double P[7];
            load7(pose + 7L * t, P);
#pragma unroll
            for (int i = 0; i < 7; ++i) pose_backup[7L * t + i] = P[i];
            pose_plus(P, d);
#pragma unroll
            for (int i = 0; i < 7; ++i) pose[7L * t + i] = P[i];
        }
    } else if (t < p.np + p.nl) {
        const int l = t - p.np;
        double dl[3];
        if (given_dxl) {
            load3(dxl + 3L * l, dl);
        } else {
        double tt[3] = {0.0, 0.0, 0.0};
        const int32_t* order = sterms ? p.lm_pairs_x[order_set(lm)] : nullptr;
        for (int t2 = p.lm_pair_start[l]; t2 < p.lm_pair_start[l + 1]; ++t2) {
            const int a = order ? order[t2] : t2;
            double Wa[18], lp[18], d[6], prod[3];
#pragma unroll
            for (int i = 0; i < 18; ++i) Wa[i] = static_cast<double>(Wbuf[18L * a + i]);
            transpose<6, 3>(Wa, lp);
            const int k = p.pair_pose[a];
#pragma unroll
            for (int i = 0; i < 6; ++i) d[i] = dxp[6L * k + i];
            mul<3, 6, 1>(lp, d, prod);
#pragma unroll
            for (int i = 0; i < 3; ++i) tt[i] = tt[i] + prod[i];
        }
        double v[3], Hi[9], prod[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) v[i] = bl[3L * l + i] - tt[i];
#pragma unroll
        for (int i = 0; i < 9; ++i) Hi[i] = static_cast<double>(Hinv[9L * l + i]);
        mul<3, 3, 1>(Hi, v, prod);
#pragma unroll
        for (int i = 0; i < 3; ++i) dl[i] = 0.0 + prod[i];
#pragma unroll
        for (int i = 0; i < 3; ++i) dxl[3L * l + i] = dl[i];
        }
        if (sterms) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
                sterms[6L * p.np + 3L * l + i] = dl[i] * (lam * dl[i] + bl[3L * l + i]);
        } else {
#pragma unroll
            for (int i = 0; i < 3; ++i) acc = acc + dl[i] * (lam * dl[i] + bl[3L * l + i]);
        }
        if (apply && !p.point_fixed[l]) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                point_backup[3L * l + i] = point[3L * l + i];
                point[3L * l + i] = point[3L * l + i] + dl[i];
            }
        }
    }
    }
    if (partial && !sterms) {
        const double b = block_sum(acc, sh);
        if (threadIdx.x == 0) partial[blockIdx.x] = b;
    }
}

// ---- iterative refinement of the mixed solve in binary64 (mixed_items.hpp) ----
__device__ __forceinline__ bool refine_skip(const lm_state* lm, int gated, const int32_t* status) {
    return (gated && lm->done) || *status != 0;
}

template <class TS>
__global__ void __launch_bounds__(kBlock) k_point_step(dev_problem p, const TS* Wbuf, const TS* Hinv, const double* bl,
                                                       const double* dxp, double* dxl, const lm_state* lm, int gated,
                                                       const int32_t* status) {
    if (refine_skip(lm, gated, status)) return;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < p.nl) mixed::point_step_item(p, Wbuf, Hinv, bl, dxp, dxl, l);
}

template <class TS>
__global__ void __launch_bounds__(kBlock) k_full_residual(dev_problem p, const double* pose, const double* point,
                                                          const double* dxp, const double* dxl, const double* bl,
                                                          const TS* Hinv, double* pslots, double* rl, double* gl,
                                                          const lm_state* lm, const double* lambda_override,
                                                          const int32_t* status) {
    if (refine_skip(lm, lambda_override ? 0 : 1, status)) return;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < p.nl)
        mixed::full_residual_item(p, pose, point, dxp, dxl, bl, Hinv, pslots, rl, gl,
                                  lambda_override ? *lambda_override : lm->lambda, l);
}

// The refinement residual by linearisation blocks (mixed::refine_block), then one
// wave per pose (mixed::refine_pose_finish).
template <class TS>
__global__ void __launch_bounds__(kLinBlock) k_refine_blocks(dev_problem p, const double* pose, const double* point,
                                                             const TS* Wbuf, const TS* Hinv, const double* bl,
                                                             const double* dxp, double* dxl, double* rl, double* gl,
                                                             double* part, const lm_state* lm,
                                                             const double* lambda_override, const int32_t* status,
                                                             int step_first) {
    if (refine_skip(lm, lambda_override ? 0 : 1, status)) return;
    __shared__ mixed::refine_stage sh;
    mixed::refine_block(p, pose, point, Wbuf, Hinv, bl, dxp, dxl, rl, gl, part,
                        lambda_override ? *lambda_override : lm->lambda, step_first != 0, static_cast<int>(blockIdx.x),
                        sh);
}

__global__ void __launch_bounds__(64) k_refine_finish(dev_problem p, const double* part, const double* bp,
                                                      const double* dxp, double* rs, const lm_state* lm,
                                                      const double* lambda_override, const int32_t* status) {
    if (refine_skip(lm, lambda_override ? 0 : 1, status)) return;
    mixed::refine_pose_finish(p, part, bp, dxp, rs, lambda_override ? *lambda_override : lm->lambda,
                              static_cast<int>(blockIdx.x), static_cast<int>(threadIdx.x));
}

// One wave per (pose, chunk), then one thread per (pose, row).
template <class TS>
__global__ void __launch_bounds__(256) k_pose_residual_part(dev_problem p, const double* pslots, const TS* Wbuf,
                                                            const double* gl, double* part, const lm_state* lm,
                                                            const double* lambda_override, const int32_t* status) {
    if (refine_skip(lm, lambda_override ? 0 : 1, status)) return;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave < p.np * kPoseChunks)
        mixed::pose_residual_chunk(p, pslots, Wbuf, gl, part, wave / kPoseChunks, wave % kPoseChunks, threadIdx.x & 63);
}

__global__ void __launch_bounds__(256) k_pose_residual(dev_problem p, const double* part, const double* bp,
                                                       const double* dxp, double* rs, const lm_state* lm,
                                                       const double* lambda_override, const int32_t* status) {
    if (refine_skip(lm, lambda_override ? 0 : 1, status)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < p.np * 6)
        mixed::pose_residual_finish(p, part, bp, dxp, rs, lambda_override ? *lambda_override : lm->lambda, t / 6, t % 6);
}

// Threads [0, nl): dx_l += H_ll^-1 (r_l - W^T delta_p); threads [nl, nl + n): dx_p += delta_p.
template <class TS>
__global__ void __launch_bounds__(kBlock) k_refine_apply(dev_problem p, const TS* Wbuf, const TS* Hinv, const double* rl,
                                                         const double* corr, double* dxp, double* dxl, const lm_state* lm,
                                                         int gated, const int32_t* status) {
    if (refine_skip(lm, gated, status)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < p.nl) mixed::refine_point_item(p, Wbuf, Hinv, rl, corr, dxl, t);
    else if (t < p.nl + p.n) dxp[t - p.nl] += corr[t - p.nl];
}

__global__ void k_decide(lm_state* lm) { lm_decide(lm); }

__global__ void __launch_bounds__(kBlock) k_rollback(dev_problem p, double* pose, const double* pose_backup, double* point,
                                                     const double* point_backup, const lm_state* lm) {
    if (!lm->need_rollback) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < p.np) {
        if (!p.pose_fixed[t])
            for (int i = 0; i < 7; ++i) pose[7L * t + i] = pose_backup[7L * t + i];
    } else if (t < p.np + p.nl) {
        const int l = t - p.np;
        if (!p.point_fixed[l])
            for (int i = 0; i < 3; ++i) point[3L * l + i] = point_backup[3L * l + i];
    }
}

__global__ void k_end_attempt(lm_state* lm) { lm_end_attempt(lm); }

// ---- diagnostics ---------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_eval_edges(dev_problem p, const double* pose, const double* point, double* out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= p.no) return;
    const int32_t param = p.obs_param[s];
    const double* K = cam_of(p, param);
    double P[7], X[3], Pc[3], q[4], r[2], Jp[12], Jl[6], W[4], Wr[4], drho;
    load7(pose + 7L * p.obs_pose[s], P);
    load3(point + 3L * p.obs_point[s], X);
    pose_apply(P, X, Pc);
    const double uv[2] = {p.obs_uv[2L * s], p.obs_uv[2L * s + 1]};
    residual(Pc, uv, K, r);
    pose_quat(P, q);
    jacobians(Pc, q, K, Jp, Jl);
    info_of(p, s, W);
    int32_t kind;
    double delta;
    loss_of(p, param, kind, delta);
    robust_info(kind, delta, W, r, drho, Wr);
    double* o = out + 27L * p.obs_orig[s];
    o[0] = r[0]; o[1] = r[1];
    for (int i = 0; i < 12; ++i) o[2 + i] = Jp[i];
    for (int i = 0; i < 6; ++i) o[14 + i] = Jl[i];
    o[20] = drho;
    for (int i = 0; i < 4; ++i) o[21 + i] = Wr[i];
    o[25] = chi2(r, W);
    o[26] = robust_chi2(kind, delta, r, W);
}

__global__ void __launch_bounds__(kBlock) k_inliers(dev_problem p, const double* pose, const double* point, uint8_t* flags) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= p.no) return;
    const int32_t param = p.obs_param[s];
    double P[7], X[3], Pc[3], r[2], W[4];
    load7(pose + 7L * p.obs_pose[s], P);
    load3(point + 3L * p.obs_point[s], X);
    pose_apply(P, X, Pc);
    const double uv[2] = {p.obs_uv[2L * s], p.obs_uv[2L * s + 1]};
    residual(Pc, uv, cam_of(p, param), r);
    info_of(p, s, W);
    int32_t kind;
    double delta;
    loss_of(p, param, kind, delta);
    flags[s] = inlier(kind, delta, r, W) ? 1 : 0;
}

__global__ void k_math_probe(int fn, int n, const double* in, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = in[i];
    out[i] = fn == 0 ? cr_sin(x) : fn == 1 ? cr_cos(x) : fn == 2 ? cr_cube(x) : fn == 3 ? sqrt(x) : x;
}

inline int grid_for(int64_t n, int block) { return static_cast<int>((n + block - 1) / block); }

// Below this many reduced-system blocks k_schur keeps two contributions in
// flight per lane (more latency hidden); above it, one (higher occupancy).
constexpr int kSchurSmallGrid = 2048;
inline int cost_grid(const dev_problem& p) {
    const int g = grid_for(p.no, kBlock);
    return g < kPartials ? (g > 0 ? g : 1) : kPartials;
}

}  // namespace

// target: 0 -> lm->chi2, 1 -> lm->chi2_new (gated on an applied step), 3 -> partial[kPartials]
void launch_cost64(const dev_problem& p, const dev_work64& w, hipStream_t s, int target, bool only_if_step, bool exact,
                   bool decide) {
    const int g = cost_grid(p);
    const int gate = only_if_step ? 2 : (target == 3 ? 0 : 1);
    ZB_TIME_BEGIN(w, K_COST, s);
    if (exact) {
        k_cost<<<grid_for(p.no, kBlock), kBlock, 0, s>>>(p, w.pose, w.point, w.residual, w.partial, w.lm, gate, 1, w.cbuf);
        k_seqsum<<<1, 64, 0, s>>>(w.cbuf, p.no, w.lm, target, gate, w.partial + kPartials);
    } else {
        k_cost<<<g, kBlock, 0, s>>>(p, w.pose, w.point, w.residual, w.partial, w.lm, gate, 1, nullptr);
        k_sum_partials<<<1, 1024, 0, s>>>(w.partial, g, w.partial + kPartials, w.lm, target, gate, decide ? 1 : 0);
    }
    ZB_TIME_END(w, s);
    if (exact && decide) launch_decide64(w, s);
}

// chi2 of the stored residuals (get_current_chi(false)), summed in insertion order.
void launch_cost_stored64(const dev_problem& p, const dev_work64& w, hipStream_t s) {
    k_cost<<<grid_for(p.no, kBlock), kBlock, 0, s>>>(p, w.pose, w.point, w.residual, w.partial, w.lm, 0, 0, w.cbuf);
    k_seqsum<<<1, 64, 0, s>>>(w.cbuf, p.no, w.lm, 3, 0, w.partial + kPartials);
}

void launch_math_probe(int fn, int n, const double* in, double* out, hipStream_t s) {
    if (n > 0) k_math_probe<<<grid_for(n, kBlock), kBlock, 0, s>>>(fn, n, in, out);
}

void launch_linearise64(const dev_problem& p, const dev_work64& w, hipStream_t s, bool gated, bool exact,
                        bool end_attempt) {
    if (p.nl > 0) {
        ZB_TIME_BEGIN(w, K_LINEARISE, s);
        k_linearise_points<<<grid_for(p.nl, kBlock), kBlock, 0, s>>>(p, w.pose, w.point, w.pose_terms, w.Hll, w.bl, w.W,
                                                                      w.lm, gated ? 1 : 0, w.frozen);
        ZB_TIME_END(w, s);
    }
    if (p.np > 0) {
        ZB_TIME_BEGIN(w, K_POSE_REDUCE, s);
        if (exact) {
            k_reduce_pose<<<grid_for(64L * p.np, 256), 256, 0, s>>>(p, w.pose_terms, w.Hpp, w.bp, w.lm, gated ? 1 : 0);
        } else {
            k_reduce_pose_part<<<grid_for(64L * p.np * kPoseChunks, 256), 256, 0, s>>>(p, w.pose_terms, w.pose_part,
                                                                                     w.lm, gated ? 1 : 0);
            k_reduce_pose_chunks<<<grid_for(27L * p.np, 256), 256, 0, s>>>(p, w.pose_part, w.Hpp, w.bp, w.lm,
                                                                          gated ? 1 : 0);
        }
        ZB_TIME_END(w, s);
    }
    if (gated && end_attempt) k_linearise_done_end<<<1, 1, 0, s>>>(w.lm);
    else if (gated) k_linearise_done<<<1, 1, 0, s>>>(w.lm);
}

void launch_init_lm64(const dev_problem& p, const dev_work64& w, hipStream_t s) {
    k_init_lm<double><<<1, 1024, 0, s>>>(p, w.Hpp, w.Hll, w.lm);
}

void launch_init_lm_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s) {
    k_init_lm<float><<<1, 1024, 0, s>>>(p, w.Hpp_f, w.Hll_f, w.lm);
}

void launch_refine_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s, int steps, const double* lambda_override) {
    if (p.n <= 0 && p.nl <= 0) return;
    const int gated = lambda_override ? 0 : 1;
    // graphs with linearisation blocks form the point steps and the residual by
    // blocks of observations
    const bool by_blocks = p.n_lblk > 0 && steps > 0;
    ZB_TIME_BEGIN(w, K_REFINE, s);
    if (p.nl > 0 && !by_blocks)
        k_point_step<float><<<grid_for(p.nl, kBlock), kBlock, 0, s>>>(p, w.W_f, w.Hinv_f, w.bl_m, w.dxp, w.dxl, w.lm, gated,
                                                                      w.chol_status);
    ZB_TIME_END(w, s);
    for (int k = 0; k < steps; ++k) {
        ZB_TIME_BEGIN(w, K_REFINE, s);
        if (by_blocks) {
            k_refine_blocks<float><<<p.n_lblk, kLinBlock, 0, s>>>(p, w.pose, w.point, w.W_f, w.Hinv_f, w.bl_m, w.dxp,
                                                                  w.dxl, w.zl, w.gl, w.lin_part_g, w.lm, lambda_override,
                                                                  w.chol_status, k == 0 ? 1 : 0);
            if (p.np > 0)
                k_refine_finish<<<p.np, 64, 0, s>>>(p, w.lin_part_g, w.bp_m, w.dxp, w.rvec, w.lm, lambda_override,
                                                    w.chol_status);
        } else if (p.nl > 0)
            k_full_residual<float><<<grid_for(p.nl, kBlock), kBlock, 0, s>>>(p, w.pose, w.point, w.dxp, w.dxl, w.bl_m,
                                                                             w.Hinv_f, w.pose_grad, w.zl, w.gl, w.lm,
                                                                             lambda_override, w.chol_status);
        if (p.np > 0 && !by_blocks) {
            k_pose_residual_part<float><<<grid_for(64L * p.np * kPoseChunks, 256), 256, 0, s>>>(
                p, w.pose_grad, w.W_f, w.gl, w.pose_part, w.lm, lambda_override, w.chol_status);
            k_pose_residual<<<grid_for(6L * p.np, 256), 256, 0, s>>>(p, w.pose_part, w.bp_m, w.dxp, w.rvec, w.lm,
                                                                     lambda_override, w.chol_status);
        }
        ZB_TIME_END(w, s);
        ZB_TIME_BEGIN(w, K_TRSV, s);
        launch_cholesky_resolve64(p.n, w, w.rvec, w.corr, s, gated != 0);
        ZB_TIME_END(w, s);
        ZB_TIME_BEGIN(w, K_REFINE, s);
        k_refine_apply<float><<<grid_for(static_cast<int64_t>(p.nl) + p.n, kBlock), kBlock, 0, s>>>(
            p, w.W_f, w.Hinv_f, w.zl, w.corr, w.dxp, w.dxl, w.lm, gated, w.chol_status);
        ZB_TIME_END(w, s);
    }
}

void launch_reduced_system64(const dev_problem& p, const dev_work64& w, hipStream_t s, const double* lambda_override,
                             bool exact) {
    if (p.nl > 0) {
        ZB_TIME_BEGIN(w, K_INVERT, s);
        k_damp_invert<<<grid_for(p.nl, kBlock), kBlock, 0, s>>>(p, w.Hll, w.Hinv, w.lm, lambda_override, w.frozen);
        ZB_TIME_END(w, s);
    }
    if (p.npair > 0) {
        ZB_TIME_BEGIN(w, K_FORM_Y, s);
        k_form_Y<<<grid_for(p.npair, kBlock), kBlock, 0, s>>>(p, w.W, w.Hinv, w.Y, w.lm, lambda_override);
        ZB_TIME_END(w, s);
    }
    const int64_t nn = static_cast<int64_t>(p.n) * p.n;
    int zg = grid_for(nn, kBlock);
    if (zg > 4096) zg = 4096;
    ZB_TIME_BEGIN(w, K_FILL, s);
    k_fill_zero<<<zg, kBlock, 0, s>>>(w.S, nn, w.lm, lambda_override ? 0 : 1);
    ZB_TIME_END(w, s);
    if (p.nblk > 0) {
        ZB_TIME_BEGIN(w, K_SCHUR, s);
        if (exact && !lambda_override) {
            k_schur_exact<<<grid_for(64L * p.nblk, 256), 256, 0, s>>>(p, w.Y, w.W, w.Hpp, w.S, w.lm);
            k_bs_exact<<<grid_for(64L * p.np, 256), 256, 0, s>>>(p, w.Y, w.bp, w.bl, w.bs, w.lm);
        } else {
            if (p.nblk < kSchurSmallGrid)
                k_schur<2><<<p.nblk, mixed::kSchurThreads, 0, s>>>(p, w.Y, w.W, w.Hpp, w.bp, w.bl, w.S, w.bs, w.lm,
                                                                   lambda_override);
            else
                k_schur<1><<<p.nblk, mixed::kSchurThreads, 0, s>>>(p, w.Y, w.W, w.Hpp, w.bp, w.bl, w.S, w.bs, w.lm,
                                                                   lambda_override);
        }
        ZB_TIME_END(w, s);
    }
}

void launch_after_solve64(const dev_work64& w, hipStream_t s) {
    ZB_TIME_BEGIN(w, K_CONTROL, s);
    k_after_solve<<<1, 1, 0, s>>>(w.lm);
    ZB_TIME_END(w, s);
}

void launch_backsub_update64(const dev_problem& p, const dev_work64& w, hipStream_t s, bool exact) {
    int g = grid_for(p.np + p.nl, kBlock);
    if (g > kPartials) g = kPartials;
    if (g < 1) g = 1;
    double* partial = w.partial + 2 * kPartials;
    ZB_TIME_BEGIN(w, K_BACKSUB, s);
    k_backsub_update<double><<<g, kBlock, 0, s>>>(p, w.pose, w.pose_backup, w.point, w.point_backup, w.W, w.Hinv, w.dxp,
                                                  w.dxl, w.bp, w.bl, partial, w.lm, 1, 1, exact ? w.sbuf : nullptr);
    if (exact) k_seqsum<<<1, 64, 0, s>>>(w.sbuf, 6L * p.np + 3L * p.nl, w.lm, 2, 2, nullptr);
    else k_sum_partials<<<1, 1024, 0, s>>>(partial, g, nullptr, w.lm, 2, 2);
    ZB_TIME_END(w, s);
}

void launch_backsub_only64(const dev_problem& p, const dev_work64& w, hipStream_t s) {
    int g = grid_for(p.np + p.nl, kBlock);
    if (g > kPartials) g = kPartials;
    if (g < 1) g = 1;
    k_backsub_update<double><<<g, kBlock, 0, s>>>(p, w.pose, w.pose_backup, w.point, w.point_backup, w.W, w.Hinv, w.dxp,
                                                  w.dxl, w.bp, w.bl, nullptr, w.lm, 0, 0, nullptr);
}

void launch_gn_apply64(const dev_problem& p, const dev_work64& w, hipStream_t s) {
    int g = grid_for(p.np + p.nl, kBlock);
    if (g > kPartials) g = kPartials;
    if (g < 1) g = 1;
    k_backsub_update<double><<<g, kBlock, 0, s>>>(p, w.pose, w.pose_backup, w.point, w.point_backup, w.W, w.Hinv, w.dxp,
                                                  w.dxl, w.bp, w.bl, nullptr, w.lm, 0, 1, nullptr);
}

void launch_backsub_update_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s) {
    int g = grid_for(p.np + p.nl, kBlock);
    if (g > kPartials) g = kPartials;
    if (g < 1) g = 1;
    double* partial = w.partial + 2 * kPartials;
    ZB_TIME_BEGIN(w, K_BACKSUB, s);
    k_backsub_update<float><<<g, kBlock, 0, s>>>(p, w.pose, w.pose_backup, w.point, w.point_backup, w.W_f, w.Hinv_f, w.dxp,
                                                 w.dxl, w.bp_m, w.bl_m, partial, w.lm, 1, 1, nullptr, 1);
    k_sum_partials<<<1, 1024, 0, s>>>(partial, g, nullptr, w.lm, 2, 2);
    ZB_TIME_END(w, s);
}

void launch_backsub_only_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s) {
    int g = grid_for(p.np + p.nl, kBlock);
    if (g > kPartials) g = kPartials;
    if (g < 1) g = 1;
    k_backsub_update<float><<<g, kBlock, 0, s>>>(p, w.pose, w.pose_backup, w.point, w.point_backup, w.W_f, w.Hinv_f, w.dxp,
                                                 w.dxl, w.bp_m, w.bl_m, nullptr, w.lm, 0, 0, nullptr, 1);
}

void launch_decide64(const dev_work64& w, hipStream_t s) {
    ZB_TIME_BEGIN(w, K_CONTROL, s);
    k_decide<<<1, 1, 0, s>>>(w.lm);
    ZB_TIME_END(w, s);
}

void launch_rollback64(const dev_problem& p, const dev_work64& w, hipStream_t s) {
    ZB_TIME_BEGIN(w, K_CONTROL, s);
    k_rollback<<<grid_for(p.np + p.nl, kBlock), kBlock, 0, s>>>(p, w.pose, w.pose_backup, w.point, w.point_backup, w.lm);
    ZB_TIME_END(w, s);
}

void launch_end_attempt64(const dev_work64& w, hipStream_t s) {
    ZB_TIME_BEGIN(w, K_CONTROL, s);
    k_end_attempt<<<1, 1, 0, s>>>(w.lm);
    ZB_TIME_END(w, s);
}

void launch_eval_edges64(const dev_problem& p, const dev_work64& w, double* out, hipStream_t s) {
    if (p.no > 0) k_eval_edges<<<grid_for(p.no, kBlock), kBlock, 0, s>>>(p, w.pose, w.point, out);
}

// map::cull's reprojection test (map.hpp:287-312) for point l: over its
// observations in insertion order, the camera-frame point R X + t with R the
// matrix of the pose quaternion (lie::so3::get_matrix, lie.hpp:188-204, the
// frame rotation map.hpp:243 sets), |uv - project(.)| rounded to binary32 and
// added in binary32, `threshold` for a point the camera cannot project
// (camera.hpp:97-99, map.hpp:297-299); the mean, and erase when there is no
// observation or the mean reaches the threshold (map.hpp:305-312).
__global__ void __launch_bounds__(kBlock) k_reprojection_cull(dev_problem p, const double* pose, const double* point,
                                                              float threshold, float* mean, uint8_t* erase) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= p.nl) return;
    double X[3];
    load3(point + 3L * l, X);
    float err = 0.0f;
    int count = 0;
    for (int s = p.lm_start[l]; s < p.lm_start[l + 1]; ++s) {
        double P[7], q[4], R[9], m[3];
#pragma unroll
        for (int i = 0; i < 7; ++i) P[i] = pose[7L * p.obs_pose[s] + i];
        pose_quat(P, q);
        qmatrix(q, R);
        mul<3, 3, 1>(R, X, m);
#pragma unroll
        for (int i = 0; i < 3; ++i) m[i] = m[i] + P[i];
        if (m[2] < 1.0e-12) {
            err = err + threshold;
        } else {
            const double uv[2] = {p.obs_uv[2L * s], p.obs_uv[2L * s + 1]};
            double r[2];
            residual(m, uv, p.cams + 4 * (p.obs_param[s] & 0xFFFF), r);
            double l2 = 0.0;  // matrix::get_length_squared, matrix.hpp:177-183
            l2 = l2 + r[0] * r[0];
            l2 = l2 + r[1] * r[1];
            err = err + static_cast<float>(sqrt(l2));
        }
        ++count;
    }
    if (count == 0) {
        mean[l] = INFINITY;
        erase[l] = 1;
    } else {
        const float mm = err / static_cast<float>(count);
        mean[l] = mm;
        erase[l] = (mm >= threshold) ? 1 : 0;
    }
}

void launch_reprojection_cull(const dev_problem& p, const dev_work64& w, float threshold, float* mean, uint8_t* erase,
                              hipStream_t s) {
    if (p.nl > 0)
        k_reprojection_cull<<<grid_for(p.nl, kBlock), kBlock, 0, s>>>(p, w.pose, w.point, threshold, mean, erase);
}

// ---- the whole solve of a graph with every point fixed, in one launch ---------
// map::optimise(n, true, ...) (map.hpp:79-192) builds such graphs for tracking:
// landmarks fixed, edges only to free frames.  The reduced system is then the
// damped H_pp alone, block-diagonal in 6x6 blocks, and the reference's dense
// row-Crout factorisation and substitutions of it (matrix_decomposition_
// cholesky.hpp:26-95) act block by block: every product with an entry outside
// a block is an exact zero, and the tolerances are the reference's global ones.
// One workgroup runs the LM loop of factor_graph::solve (factor_graph.hpp:
// 489-615) with no host round trip: residuals, costs and per-pose sums in the
// reference's order (sequential from zero, insertion order), so the result is the
// reference's bit for bit, as the per-graph launch sequence gives it.
constexpr int kMotionThreads = 256;
constexpr int kSeqStage = 4096;  // doubles staged in LDS for the sequential sums

// Sum over observations in insertion order of their robust cost at the current
// state (residuals stored), broadcast to every thread (factor_graph.hpp:618-629).
__device__ double motion_cost(const dev_problem& p, const dev_work64& w, double* stage, double* out) {
    for (int s = threadIdx.x; s < p.no; s += blockDim.x) {
        const int32_t param = p.obs_param[s];
        double P[7], X[3], Pc[3], r[2], W[4];
        const double uv[2] = {p.obs_uv[2L * s], p.obs_uv[2L * s + 1]};
        load7(w.pose + 7L * p.obs_pose[s], P);
        load3(w.point + 3L * p.obs_point[s], X);
        pose_apply(P, X, Pc);
        residual(Pc, uv, cam_of(p, param), r);
        w.residual[2L * s] = r[0];
        w.residual[2L * s + 1] = r[1];
        info_of(p, s, W);
        int32_t kind;
        double delta;
        loss_of(p, param, kind, delta);
        w.cbuf[p.obs_orig[s]] = robust_chi2(kind, delta, r, W);
    }
    __syncthreads();
    double total = 0.0;
    for (int c0 = 0; c0 < p.no; c0 += kSeqStage) {
        const int m = min(kSeqStage, p.no - c0);
        for (int i = threadIdx.x; i < m; i += blockDim.x) stage[i] = w.cbuf[c0 + i];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < m; ++i) total = total + stage[i];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = total;
    __syncthreads();
    const double r = *out;
    __syncthreads();
    return r;
}

// make_hessian's pose blocks (factor_graph.hpp:648-732) with every point fixed:
// per observation of a free pose its 21 H_pp and 6 gradient terms, then per
// (pose, term) the sum in insertion order from zero (H_pp += H, b_p -= v), the
// pose's terms (contiguous, pose-major) staged through LDS.
__device__ void motion_linearise(const dev_problem& p, const dev_work64& w, double* stage) {
    for (int s = threadIdx.x; s < p.no; s += blockDim.x) {
        const int ip = p.obs_pose[s];
        if (p.pose_fixed[ip]) continue;
        const int32_t param = p.obs_param[s];
        const double* K = cam_of(p, param);
        double P[7], X[3], Pc[3], q[4], r[2], Jp[12], Jl[6], W[4], Wr[4], drho;
        load7(w.pose + 7L * ip, P);
        load3(w.point + 3L * p.obs_point[s], X);
        pose_apply(P, X, Pc);
        const double uv[2] = {p.obs_uv[2L * s], p.obs_uv[2L * s + 1]};
        residual(Pc, uv, K, r);
        pose_quat(P, q);
        jacobians(Pc, q, K, Jp, Jl);
        info_of(p, s, W);
        int32_t kind;
        double delta;
        loss_of(p, param, kind, delta);
        robust_info(kind, delta, W, r, drho, Wr);
        double H21[21], Wpl[18], v[6];
        pose_terms(Jp, Jl, Wr, W, r, drho, H21, Wpl, v);
        double* out = w.pose_terms + 27L * p.obs_pose_pos[s];
#pragma unroll
        for (int i = 0; i < 21; ++i) out[i] = H21[i];
#pragma unroll
        for (int i = 0; i < 6; ++i) out[21 + i] = v[i];
    }
    __syncthreads();
    constexpr int kChunk = kSeqStage / 27;  // observations per staged chunk
    for (int i = 0; i < p.np; ++i) {
        const int j = threadIdx.x;  // term (threads 0..26)
        double acc = 0.0;
        if (!p.pose_fixed[i]) {
            const int t0 = p.pose_start[i], t1 = p.pose_start[i + 1];
            for (int c0 = t0; c0 < t1; c0 += kChunk) {
                const int m = min(kChunk, t1 - c0);
                const double* src = w.pose_terms + 27L * c0;
                for (int k = threadIdx.x; k < 27 * m; k += blockDim.x) stage[k] = src[k];
                __syncthreads();
                if (j < 27)
                    for (int t = 0; t < m; ++t) acc = j < 21 ? acc + stage[27 * t + j] : acc - stage[27 * t + j];
                __syncthreads();
            }
        }
        if (j < 21) w.Hpp[21L * i + j] = acc;
        else if (j < 27) w.bp[6L * i + (j - 21)] = acc;
    }
    __syncthreads();
}

// std::max(scale, |a|) of the reference's decompose / solve (math.hpp:235-237).
__device__ __forceinline__ double ref_max_abs(double scale, double a) {
    const double x = fabs(a);
    return (scale < x) ? x : scale;
}

// solve_linear_system (factor_graph.hpp:736-806) for the block-diagonal system:
// damping, decompose_cholesky and solve_cholesky per 6x6 block with the global
// tolerances.  Returns whether both succeeded; dx in w.dxp.
__device__ bool motion_linear_solve(const dev_problem& p, const dev_work64& w, double lam, double (*L)[36],
                                    double* sh, int* flag) {
    const int np = p.np, n = 6 * np, tid = threadIdx.x;
    // damped blocks (lower triangle; dense entry = 0 + (H + lambda on the diagonal))
    for (int i = tid; i < np; i += blockDim.x) {
        const double* H = w.Hpp + 21L * i;
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b <= a; ++b) {
                double v = H[a * (a + 1) / 2 + b];
                if (a == b) v += lam;
                L[i][6 * a + b] = 0 + v;
            }
    }
    if (tid == 0) *flag = 0;
    __syncthreads();
    if (tid == 0) {
        double scale = 0;
        for (int r = 0; r < n; ++r) scale = ref_max_abs(scale, L[r / 6][7 * (r % 6)]);
        sh[0] = scale * 1e-14;
    }
    __syncthreads();
    const double eps = sh[0];
    for (int i = tid; i < np; i += blockDim.x) {
        double* A = L[i];  // factorised in place, row by row (A's lower entries are read before written)
        bool ok = true;
        for (int r = 0; r < 6 && ok; ++r) {
            for (int k = 0; k < r; ++k) {
                double v = A[6 * r + k];
                for (int j = 0; j < k; ++j) v -= A[6 * r + j] * A[6 * k + j];
                A[6 * r + k] = v / A[7 * k];
            }
            double v = A[7 * r];
            for (int j = 0; j < r; ++j) v -= A[6 * r + j] * A[6 * r + j];
            if (!(v > eps)) ok = false;
            else A[7 * r] = sqrt(v);
        }
        if (!ok) atomicOr(flag, 1);
    }
    __syncthreads();
    if (*flag) return false;
    if (tid == 0) {
        double scale = 0;
        for (int r = 0; r < n; ++r) scale = ref_max_abs(scale, L[r / 6][7 * (r % 6)]);
        sh[0] = scale * 1e-7;
    }
    __syncthreads();
    const double eps2 = sh[0];
    for (int i = tid; i < np; i += blockDim.x) {
        const double* A = L[i];
        double x[6];
        bool ok = true;
        for (int r = 0; r < 6; ++r) {
            x[r] = w.bp[6L * i + r];  // b_s = b_p (no Schur terms)
            for (int j = 0; j < r; ++j) x[r] -= A[6 * r + j] * x[j];
            if (!(fabs(A[7 * r]) > eps2)) ok = false;
            x[r] /= A[7 * r];
        }
        for (int r = 6; r-- > 0;) {
            for (int j = r + 1; j < 6; ++j) x[r] -= A[6 * j + r] * x[j];
            if (!(fabs(A[7 * r]) > eps2)) ok = false;
            x[r] /= A[7 * r];
        }
        for (int r = 0; r < 6; ++r) w.dxp[6L * i + r] = x[r];
        if (!ok) atomicOr(flag, 1);
    }
    __syncthreads();
    return *flag == 0;
}

__global__ void __launch_bounds__(kMotionThreads) k_motion_solve(dev_problem p, dev_work64 w) {
    __shared__ double L[kMotionMaxPoses][36];
    __shared__ double stage[kSeqStage];
    __shared__ double sh[2];
    __shared__ int flag;
    lm_state* lm = w.lm;
    const int tid = threadIdx.x, np = p.np, n = 6 * np;
    // initial cost, linearisation and damping (factor_graph.hpp:500-522)
    const double c0 = motion_cost(p, w, stage, sh + 1);
    motion_linearise(p, w, stage);
    if (tid == 0) {
        lm->chi2 = c0;
        max_scan m = {0.0, 0, 0, 1};
        for (int r = 0; r < n; ++r) m = scan_push(m, w.Hpp[21L * (r / 6) + (r % 6) * (r % 6 + 1) / 2 + r % 6]);
        if (p.nl > 0) m = scan_push(m, 0.0);  // the fixed points' empty blocks
        lm_init(lm, scan_max_diagonal(m));
    }
    __syncthreads();
    const int max_attempts = 11 * lm->iterations + 1;
    for (int attempt = 0; attempt < max_attempts; ++attempt) {
        if (lm->done) break;
        const double lam = lm->lambda;
        const bool ok = motion_linear_solve(p, w, lam, L, sh, &flag);
        if (tid == 0) {
            lm->solve_ok = ok ? 1 : 0;
            lm_after_solve(lm);
        }
        __syncthreads();
        if (lm->step_applied) {
            // update_states and the LM scale over every slot (points: zero steps)
            for (int i = tid; i < np; i += blockDim.x) {
                if (p.pose_fixed[i]) continue;
                double P[7], d[6];
                load7(w.pose + 7L * i, P);
                for (int k = 0; k < 7; ++k) w.pose_backup[7L * i + k] = P[k];
                for (int k = 0; k < 6; ++k) d[k] = w.dxp[6L * i + k];
                pose_plus(P, d);
                for (int k = 0; k < 7; ++k) w.pose[7L * i + k] = P[k];
            }
            if (tid == 0) {
                double scale = 0;
                for (int r = 0; r < n; ++r) scale += w.dxp[r] * (lam * w.dxp[r] + w.bp[r]);
                lm->scale = scale + 1e-3;
            }
            __syncthreads();
            const double c = motion_cost(p, w, stage, sh + 1);
            if (tid == 0) {
                lm->chi2_new = c;
                lm_decide(lm);
            }
            __syncthreads();
            const bool rollback = lm->need_rollback != 0;
            const bool relinearise = !lm->done && lm->need_linearize;
            if (rollback)
                for (int i = tid; i < np; i += blockDim.x)
                    if (!p.pose_fixed[i])
                        for (int k = 0; k < 7; ++k) w.pose[7L * i + k] = w.pose_backup[7L * i + k];
            __syncthreads();
            if (relinearise) {
                motion_linearise(p, w, stage);
                if (tid == 0) lm_linearise_done(lm);
            }
        }
        __syncthreads();
        if (tid == 0) lm_end_attempt(lm);
        __syncthreads();
    }
}

void launch_motion_solve(const dev_problem& p, const dev_work64& w, hipStream_t s) {
    if (p.np > 0 && p.np <= kMotionMaxPoses) k_motion_solve<<<1, kMotionThreads, 0, s>>>(p, w);
}

void launch_inliers64(const dev_problem& p, const dev_work64& w, uint8_t* flags_sorted, hipStream_t s) {
    if (p.no > 0) k_inliers<<<grid_for(p.no, kBlock), kBlock, 0, s>>>(p, w.pose, w.point, flags_sorted);
}

}  // namespace zb
