// Binary32 linearisation and Schur elimination of the mixed-precision path.
//
// The per-vertex Hessian blocks H_pp, W, H_ll (make_hessian,
// factor_graph.hpp:648-732), the damped 3x3 point inverses and the Schur
// products Y W^T (factor_graph.hpp:742-781) run in binary32 from binary32
// Jacobians, with binary32 accumulation.  These carry almost all of the bytes
// and flops.  Four things stay binary64:
//  * the residual and the robust weight of each observation, in the reference's
//    operation order (compiled without contraction).  The robust information
//    gains its 2 rho'' (W r)(W r)^T term only when rho' + 2 rho'' e^2 > 0
//    (factor_graph.hpp:265), an expression that is exactly zero in real
//    arithmetic for every Huber outlier, so its sign is a rounding artefact that
//    changes the outlier's 2x2 information from rho' W to rho' (W - r r^T / e^2).
//    Evaluated in binary32 the branch flips for about half of the outliers and
//    the reduced system moves by tens of percent; evaluated as the binary64 path
//    does, the two paths take the same branch on the same state.  A binary32
//    residual would also carry the 3e-5 px spacing of binary32 pixel
//    coordinates into the gradient, which sets the optimum;
//  * the camera-frame point T X (a translation of tens of metres against the
//    rotated point), from which both the residual and the Jacobians are formed;
//  * the gradient b = -sum J^T rho' Omega r and its sums.  The Hessian only
//    shapes the step; the gradient fixes where the iteration stops.  With a
//    binary32 gradient (per-pose sums over ~1600 observations) the solver
//    stalls about 1e-5 above the binary64 cost on the 100-keyframe problem;
//  * the assembled reduced system S = (H_pp + lambda I) - sum Y W^T, which the
//    binary64 Cholesky factorises and the binary64 refinement
//    (kernels_fp64.hip) corrects.
// Data layouts are those of the binary64 kernels.
#include "ba_math.hpp"
#include "common.hpp"

namespace zb {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

// One thread per point walks its observations: H_ll and the W blocks accumulate
// in binary32, b_l in binary64; the pose-side terms of each observation go to
// pose-major slots (21 binary32 Hessian terms, 6 binary64 gradient terms) for
// k_reduce_pose_f.
__global__ void __launch_bounds__(kBlock) k_lin_points_f(dev_problem p, const double* pose, const double* point,
                                                         float* pterms, double* pgrad, float* Hll, double* bl,
                                                         float* Wbuf, const lm_state* lm, int gated) {
    if (gated && (lm->done || !lm->need_linearize)) return;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= p.nl) return;
    const bool lfixed = p.point_fixed[l] != 0;
    const double X[3] = {point[3L * l], point[3L * l + 1], point[3L * l + 2]};
    float H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double b[3] = {0, 0, 0};
    const int s0 = p.lm_start[l], s1 = p.lm_start[l + 1];
    for (int s = s0; s < s1; ++s) {
        const int ip = p.obs_pose[s];
        const bool pfixed = p.pose_fixed[ip] != 0;
        if (pfixed && lfixed) continue;
        const int32_t param = p.obs_param[s];
        const double* K = p.cams + 4 * (param & 0xFFFF);
        double P[7], Pc[3], q[4];
#pragma unroll
        for (int i = 0; i < 7; ++i) P[i] = pose[7L * ip + i];
        pose_apply(P, X, Pc);
        pose_quat(P, q);
        // residual, robust weight, Jacobians: binary64, reference order
        double r[2], W[4], Wr[4], drho, Jp[12], Jl[6];
        const double uv[2] = {p.obs_uv[2L * s], p.obs_uv[2L * s + 1]};
        residual(Pc, uv, K, r);
        if (p.identity_info) {
            W[0] = 1.0; W[1] = 0.0; W[2] = 0.0; W[3] = 1.0;
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) W[i] = p.obs_info[4L * s + i];
        }
        const int li = param >> 16;
        robust_info(p.loss_kind[li], p.loss_delta[li], W, r, drho, Wr);
        jacobians(Pc, q, K, Jp, Jl);
        // Hessian blocks: binary32
        float Jpf[12], Jlf[6], Wrf[4];
#pragma unroll
        for (int i = 0; i < 12; ++i) Jpf[i] = static_cast<float>(Jp[i]);
#pragma unroll
        for (int i = 0; i < 6; ++i) Jlf[i] = static_cast<float>(Jl[i]);
#pragma unroll
        for (int i = 0; i < 4; ++i) Wrf[i] = static_cast<float>(Wr[i]);
        if (!pfixed) {
            float H21[21], Wpl[18];
            double v[6];
            pose_hessian(Jpf, Jlf, Wrf, H21, Wpl);
            pose_grad(Jp, W, r, drho, v);
            const int64_t slot = p.obs_pose_pos[s];
#pragma unroll
            for (int i = 0; i < 21; ++i) pterms[21 * slot + i] = H21[i];
#pragma unroll
            for (int i = 0; i < 6; ++i) pgrad[6 * slot + i] = v[i];
            if (!lfixed) {
                const int pr = p.obs_pair[s];
                float* Wp = Wbuf + 18L * pr;
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
            float Hl[9];
            double v[3];
            point_hessian(Jlf, Wrf, Hl);
            point_grad(Jl, W, r, drho, v);
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
}

// One wave per pose: lanes 0..20 sum the binary32 H_pp terms, lanes 21..26 the
// binary64 gradient terms, over the pose's observations (H_pp += H; b_p -= v).
__global__ void __launch_bounds__(256) k_reduce_pose_f(dev_problem p, const float* pterms, const double* pgrad,
                                                       float* Hpp, double* bp, const lm_state* lm, int gated) {
    if (gated && (lm->done || !lm->need_linearize)) return;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= p.np || lane >= 27) return;
    const bool fixed = p.pose_fixed[wave] != 0;
    const int t0 = p.pose_start[wave], t1 = fixed ? t0 : p.pose_start[wave + 1];
    constexpr int U = 16;
    if (lane < 21) {
        const float* src = pterms + lane;
        float part[4] = {0.f, 0.f, 0.f, 0.f};
        int t = t0;
        for (; t + U <= t1; t += U) {
            float v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = src[21L * (t + u)];
#pragma unroll
            for (int u = 0; u < U; ++u) part[u & 3] += v[u];
        }
        for (; t < t1; ++t) part[0] += src[21L * t];
        Hpp[21L * wave + lane] = (part[0] + part[1]) + (part[2] + part[3]);
    } else {
        const double* src = pgrad + (lane - 21);
        double part[4] = {0.0, 0.0, 0.0, 0.0};
        int t = t0;
        for (; t + U <= t1; t += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = src[6L * (t + u)];
#pragma unroll
            for (int u = 0; u < U; ++u) part[u & 3] += v[u];
        }
        for (; t < t1; ++t) part[0] += src[6L * t];
        bp[6L * wave + (lane - 21)] = -((part[0] + part[1]) + (part[2] + part[3]));
    }
}

__global__ void k_linearise_done_f(lm_state* lm) {
    lm->need_linearize = 0;
    lm->lin_count += 1;
}

// Damped point blocks and their inverses (factor_graph.hpp:742-765), binary32.
__global__ void __launch_bounds__(kBlock) k_damp_invert_f(dev_problem p, const float* Hll, float* Hinv, const lm_state* lm,
                                                          const double* lambda_override) {
    if (!lambda_override && lm->done) return;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= p.nl) return;
    const float lam = static_cast<float>(lambda_override ? *lambda_override : lm->lambda);
    float H[9], inv[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = Hll[9L * l + i];
    H[0] += lam; H[4] += lam; H[8] += lam;
    if (p.point_fixed[l]) {
#pragma unroll
        for (int i = 0; i < 9; ++i) inv[i] = 0.f;
    } else {
        invert3(H, inv);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) Hinv[9L * l + i] = inv[i];
}

// Y = W H_ll^-1 per W block (matrix_sparse_block_diagonal.hpp:107-122), binary32.
__global__ void __launch_bounds__(kBlock) k_form_Y_f(dev_problem p, const float* Wbuf, const float* Hinv, float* Y,
                                                     const lm_state* lm, const double* lambda_override) {
    if (!lambda_override && lm->done) return;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= p.npair) return;
    float Wa[18], Hi[9], o[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) Wa[i] = Wbuf[18L * a + i];
    const int l = p.pair_point[a];
#pragma unroll
    for (int i = 0; i < 9; ++i) Hi[i] = Hinv[9L * l + i];
    mul<6, 3, 3>(Wa, Hi, o);
#pragma unroll
    for (int i = 0; i < 18; ++i) Y[18L * a + i] = o[i];
}

// Reduced camera system, one wave per 6x6 lower-triangle block: the products
// Y_a W_c^T accumulate in binary32 per lane and across the wave, Y b_l in
// binary64; the block is assembled in binary64:
//   S_ik = [i == k] (H_pp,i + lambda I) - sum Y W^T,  b_s,i = b_p,i - sum Y b_l.
__global__ void __launch_bounds__(256) k_schur_f(dev_problem p, const float* Y, const float* Wbuf, const float* Hpp,
                                                 const double* bp, const double* bl, double* S, double* bs,
                                                 const lm_state* lm, const double* lambda_override) {
    if (!lambda_override && lm->done) return;
    const int blk = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (blk >= p.nblk) return;
    const int row = p.blk_row[blk], col = p.blk_col[blk];
    const bool diag = row == col;
    float acc[36];
#pragma unroll
    for (int i = 0; i < 36; ++i) acc[i] = 0.f;
    double accb[6] = {0, 0, 0, 0, 0, 0};
    const int64_t c0 = p.blk_start[blk], c1 = p.blk_start[blk + 1];
    for (int64_t c = c0 + lane; c < c1; c += 64) {
        const int2 ab = p.contrib[c];
        float Ya[18], Wc[18];
        const float2* ya = reinterpret_cast<const float2*>(Y + 18L * ab.x);
        const float2* wc = reinterpret_cast<const float2*>(Wbuf + 18L * ab.y);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const float2 t = ya[i];
            Ya[2 * i] = t.x; Ya[2 * i + 1] = t.y;
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const float2 t = wc[i];
            Wc[2 * i] = t.x; Wc[2 * i + 1] = t.y;
        }
        // (Y W^T)_rc = sum_k Y[r][k] W[c][k]
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int cc = 0; cc < 6; ++cc)
                acc[r * 6 + cc] += Ya[r * 3] * Wc[cc * 3] + Ya[r * 3 + 1] * Wc[cc * 3 + 1] + Ya[r * 3 + 2] * Wc[cc * 3 + 2];
        if (diag) {
            const int l = p.pair_point[ab.x];
            const double b0 = bl[3L * l], b1 = bl[3L * l + 1], b2 = bl[3L * l + 2];
#pragma unroll
            for (int r = 0; r < 6; ++r)
                accb[r] += static_cast<double>(Ya[r * 3]) * b0 + static_cast<double>(Ya[r * 3 + 1]) * b1 +
                           static_cast<double>(Ya[r * 3 + 2]) * b2;
        }
    }
    const double lam = lambda_override ? *lambda_override : lm->lambda;
    const int64_t n = p.n;
    double* Sblk = S + (6L * row) * n + 6L * col;
#pragma unroll
    for (int i = 0; i < 36; ++i) {
        const float tot = wave_sum_f(acc[i]);
        const int r = i / 6, c = i % 6;
        if (lane == 0) {
            if (diag) {
                if (c <= r) {
                    double base = static_cast<double>(Hpp[21L * row + r * (r + 1) / 2 + c]);
                    if (r == c) base += lam;
                    Sblk[r * n + c] = base - static_cast<double>(tot);
                }
            } else {
                Sblk[r * n + c] = -static_cast<double>(tot);
            }
        }
    }
    if (diag) {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const double tot = wave_sum_d(accb[i]);
            if (lane == 0) bs[6L * row + i] = bp[6L * row + i] - tot;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_fill_zero_d(double* a, int64_t n, const lm_state* lm, int gated) {
    if (gated && lm->done) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = 0.0;
}

inline int grid_of(int64_t n, int block) { return static_cast<int>((n + block - 1) / block); }

}  // namespace

void launch_linearise_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s, bool gated) {
    if (p.nl > 0) {
        ZB_TIME_BEGIN(w, K_LINEARISE, s);
        k_lin_points_f<<<grid_of(p.nl, kBlock), kBlock, 0, s>>>(p, w.pose, w.point, w.pose_terms_f, w.pose_grad, w.Hll_f,
                                                                 w.bl_m, w.W_f, w.lm, gated ? 1 : 0);
        ZB_TIME_END(w, s);
    }
    if (p.np > 0) {
        ZB_TIME_BEGIN(w, K_POSE_REDUCE, s);
        k_reduce_pose_f<<<grid_of(64L * p.np, 256), 256, 0, s>>>(p, w.pose_terms_f, w.pose_grad, w.Hpp_f, w.bp_m, w.lm,
                                                                  gated ? 1 : 0);
        ZB_TIME_END(w, s);
    }
    if (gated) k_linearise_done_f<<<1, 1, 0, s>>>(w.lm);
}

void launch_reduced_system_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s, const double* lambda_override) {
    if (p.nl > 0) {
        ZB_TIME_BEGIN(w, K_INVERT, s);
        k_damp_invert_f<<<grid_of(p.nl, kBlock), kBlock, 0, s>>>(p, w.Hll_f, w.Hinv_f, w.lm, lambda_override);
        ZB_TIME_END(w, s);
    }
    if (p.npair > 0) {
        ZB_TIME_BEGIN(w, K_FORM_Y, s);
        k_form_Y_f<<<grid_of(p.npair, kBlock), kBlock, 0, s>>>(p, w.W_f, w.Hinv_f, w.Y_f, w.lm, lambda_override);
        ZB_TIME_END(w, s);
    }
    const int64_t nn = static_cast<int64_t>(p.n) * p.n;
    int zg = grid_of(nn, kBlock);
    if (zg > 4096) zg = 4096;
    ZB_TIME_BEGIN(w, K_FILL, s);
    k_fill_zero_d<<<zg, kBlock, 0, s>>>(w.S, nn, w.lm, lambda_override ? 0 : 1);
    ZB_TIME_END(w, s);
    if (p.nblk > 0) {
        ZB_TIME_BEGIN(w, K_SCHUR, s);
        k_schur_f<<<grid_of(64L * p.nblk, 256), 256, 0, s>>>(p, w.Y_f, w.W_f, w.Hpp_f, w.bp_m, w.bl_m, w.S, w.bs, w.lm,
                                                              lambda_override);
        ZB_TIME_END(w, s);
    }
}

}  // namespace zb
