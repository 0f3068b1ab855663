// Adjoint of the converged bundle adjustment (implicit differentiation).
//
// At the converged state x* the gradient g(x, theta) = sum_e v_e of the robust
// cost vanishes, where v_e = ((rho' J_e^T) Omega) r_e are the terms make_hessian
// subtracts into b (factor_graph.hpp:680-689, 725) and theta are the observed
// pixel coordinates and the pinhole intrinsics.  For a loss L(x*) the implicit
// function theorem gives
//   dL/dtheta = -lambda^T dg/dtheta,   H lambda = dL/dx,   H = dg/dx at x*,
// with H the exact Hessian of the robust cost in the retraction's local
// coordinates (pose: T <- exp(delta) T, point: X <- X + delta).  H v and dg/dtheta
// are evaluated exactly with forward-mode dual numbers through the reference's
// own retraction, residual, Jacobian, robust-weight and gradient arithmetic
// (ba_math.hpp); H lambda = dL/dx is solved by conjugate gradients
// preconditioned with the Gauss-Newton reduced system's Cholesky factor (the
// Schur factor the solver already forms).  Vectors are ordered [poses 6 np |
// points 3 np], rows of fixed vertices are zero.
#include "ba_math.hpp"
#include "common.hpp"
#include "dual.hpp"

namespace zb {
namespace {

constexpr int kBlock = 256;

struct k_tangent {
    double k[4];
};

// A point is held constant by the adjoint when it is fixed or when its undamped
// 3x3 block is singular (zero inverse by matrix::invert<3>, matrix.hpp:749-774):
// undamped steps never move it, so its position does not depend on theta.
__device__ __forceinline__ bool point_frozen(const dev_problem& p, const uint8_t* frozen, int l) {
    return p.point_fixed[l] != 0 || (frozen && frozen[l] != 0);
}

__global__ void __launch_bounds__(kBlock) k_mark_frozen(dev_problem p, const double* Hinv, uint8_t* frozen) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= p.nl) return;
    bool zero = true;
#pragma unroll
    for (int i = 0; i < 9; ++i) zero = zero && Hinv[9L * l + i] == 0.0;
    frozen[l] = (p.point_fixed[l] || zero) ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_zero_frozen_rows(dev_problem p, const uint8_t* frozen, double* v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 6 * p.np) {
        if (p.pose_fixed[t / 6]) v[t] = 0.0;
    } else if (t < 6 * p.np + 3 * p.nl) {
        if (frozen[(t - 6 * p.np) / 3]) v[t] = 0.0;
    }
}

__device__ __forceinline__ void obs_W(const dev_problem& p, int s, dual W[4]) {
    if (p.identity_info) {
        W[0] = 1.0; W[1] = 0.0; W[2] = 0.0; W[3] = 1.0;
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) W[i] = p.obs_info[4L * s + i];
    }
}

// Gradient terms of observation s at the state moved by eps along the tangents
// (pose dp[6], point dl[3], intrinsics dK[4], pixel duv[2]); returns the dual
// pose and point terms.
__device__ inline void grad_terms(const dev_problem& p, const double* pose, const double* point, int s, const double* dp,
                                  const double* dl, const double* dK, const double* duv, dual vp[6], dual vl[3],
                                  bool pfixed, bool lfixed) {
    const int ip = p.obs_pose[s], l = p.obs_point[s];
    const int32_t param = p.obs_param[s];
    const double* K0 = p.cams + 4 * (param & 0xFFFF);
    dual K[4], X[3], P[7], Pc[3], q[4], r[2], Jp[12], Jl[6], W[4], Wr[4], drho;
#pragma unroll
    for (int i = 0; i < 4; ++i) K[i] = dual(K0[i], dK ? dK[i] : 0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i) X[i] = dual(point[3L * l + i], (dl && !lfixed) ? dl[i] : 0.0);
#pragma unroll
    for (int i = 0; i < 7; ++i) P[i] = dual(pose[7L * ip + i]);
    if (dp && !pfixed) {
        dual d[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) d[i] = dual(0.0, dp[i]);
        pose_plus(P, d);
    }
    pose_apply(P, X, Pc);
    const dual uv[2] = {dual(p.obs_uv[2L * s], duv ? duv[0] : 0.0), dual(p.obs_uv[2L * s + 1], duv ? duv[1] : 0.0)};
    residual(Pc, uv, K, r);
    pose_quat(P, q);
    jacobians(Pc, q, K, Jp, Jl);
    obs_W(p, s, W);
    const int li = param >> 16;
    robust_info(p.loss_kind[li], dual(p.loss_delta[li]), W, r, drho, Wr);
    pose_grad(Jp, W, r, drho, vp);
    point_grad(Jl, W, r, drho, vl);
}

// Tangent of g along (vp, vl, dK): point rows directly, pose rows through the
// observations' pose-major slots (summed by k_pose_slots).
__global__ void __launch_bounds__(kBlock) k_grad_tangent(dev_problem p, const double* pose, const double* point,
                                                         const double* vpose, const double* vpoint, k_tangent kt,
                                                         int use_k, const uint8_t* frozen, double* pslots,
                                                         double* out_point) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= p.nl) return;
    const bool lfixed = point_frozen(p, frozen, l);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int s = p.lm_start[l]; s < p.lm_start[l + 1]; ++s) {
        const int ip = p.obs_pose[s];
        const bool pfixed = p.pose_fixed[ip] != 0;
        if (pfixed && lfixed) continue;
        dual vp[6], vl[3];
        grad_terms(p, pose, point, s, vpose ? vpose + 6L * ip : nullptr, vpoint ? vpoint + 3L * l : nullptr,
                   use_k ? kt.k : nullptr, nullptr, vp, vl, pfixed, lfixed);
        if (!pfixed) {
            double* o = pslots + 6L * p.obs_pose_pos[s];
#pragma unroll
            for (int k = 0; k < 6; ++k) o[k] = vp[k].d;
        }
        if (!lfixed) {
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[k] += vl[k].d;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) out_point[3L * l + k] = lfixed ? 0.0 : acc[k];
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

// One wave per pose: out[6 i + k] = sum of the pose's slots (0 for fixed poses).
__global__ void __launch_bounds__(256) k_pose_slots(dev_problem p, const double* pslots, double* out_pose) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (i >= p.np) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (!p.pose_fixed[i]) {
        for (int t = p.pose_start[i] + lane; t < p.pose_start[i + 1]; t += 64) {
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[k] += pslots[6L * t + k];
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[k] = wave_sum(acc[k]);
    if (lane < 6) {
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) a = (q == lane) ? acc[q] : a;
        out_pose[6L * i + lane] = a;
    }
}

// Gauss-Newton preconditioner z = M^-1 r by block elimination of the points:
// g_l = H_ll^-1 r_l; rs = r_p - W g; z_p = S^-1 rs; z_l = H_ll^-1 (r_l - W^T z_p).
__global__ void __launch_bounds__(kBlock) k_prec_points(dev_problem p, const double* Hinv, const double* rl, double* gl) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= p.nl) return;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double o = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) o += Hinv[9L * l + 3 * r + c] * rl[3L * l + c];
        gl[3L * l + r] = o;
    }
}

__global__ void __launch_bounds__(256) k_prec_poses(dev_problem p, const double* Wbuf, const double* rp, const double* gl,
                                                    double* rs) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (i >= p.np) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int t = p.pose_pair_start[i] + lane; t < p.pose_pair_start[i + 1]; t += 64) {
        const int a = p.pose_pairs[t];
        const int l = p.pair_point[a];
#pragma unroll
        for (int r = 0; r < 6; ++r)
            acc[r] += Wbuf[18L * a + 3 * r] * gl[3L * l] + Wbuf[18L * a + 3 * r + 1] * gl[3L * l + 1] +
                      Wbuf[18L * a + 3 * r + 2] * gl[3L * l + 2];
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) acc[r] = wave_sum(acc[r]);
    if (lane < 6) {
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) a = (q == lane) ? acc[q] : a;
        rs[6L * i + lane] = p.pose_fixed[i] ? 0.0 : rp[6L * i + lane] - a;
    }
}

__global__ void __launch_bounds__(kBlock) k_prec_back(dev_problem p, const double* Wbuf, const double* Hinv, const double* rl,
                                                      double* zp, double* zl) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < p.nl) {
        const int l = t;
        double tt[3] = {0.0, 0.0, 0.0};
        for (int a = p.lm_pair_start[l]; a < p.lm_pair_start[l + 1]; ++a) {
            const int k = p.pair_pose[a];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double v = 0.0;
#pragma unroll
                for (int r = 0; r < 6; ++r) v += Wbuf[18L * a + 3 * r + c] * zp[6L * k + r];
                tt[c] += v;
            }
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double o = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) o += Hinv[9L * l + 3 * r + c] * (rl[3L * l + c] - tt[c]);
            zl[3L * l + r] = p.point_fixed[l] ? 0.0 : o;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_zero_fixed_poses(dev_problem p, double* v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 6 * p.np && p.pose_fixed[t / 6]) v[t] = 0.0;
}

// Block partial sums of a . b (fixed tree per block), then one block adds them.
__global__ void __launch_bounds__(kBlock) k_dot(const double* a, const double* b, int64_t n, double* partial) {
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) acc += a[i] * b[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}
__global__ void __launch_bounds__(kBlock) k_sum(const double* partial, int n, double* out) {
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) acc += partial[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// y = a x + b y
__global__ void __launch_bounds__(kBlock) k_axpby(double a, const double* x, double b, double* y, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

// dL/duv_s = -(lambda_p . dv_p/duv + lambda_l . dv_l/duv) for both pixel axes.
__global__ void __launch_bounds__(kBlock) k_obs_grad(dev_problem p, const double* pose, const double* point,
                                                     const double* lam, const uint8_t* frozen, double* dobs) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= p.no) return;
    const int ip = p.obs_pose[s], l = p.obs_point[s];
    const bool pfixed = p.pose_fixed[ip] != 0, lfixed = point_frozen(p, frozen, l);
    double out[2] = {0.0, 0.0};
    if (!(pfixed && lfixed)) {
        for (int axis = 0; axis < 2; ++axis) {
            const double duv[2] = {axis == 0 ? 1.0 : 0.0, axis == 1 ? 1.0 : 0.0};
            dual vp[6], vl[3];
            grad_terms(p, pose, point, s, nullptr, nullptr, nullptr, duv, vp, vl, pfixed, lfixed);
            double t = 0.0;
            if (!pfixed) {
#pragma unroll
                for (int k = 0; k < 6; ++k) t += lam[6L * ip + k] * vp[k].d;
            }
            if (!lfixed) {
#pragma unroll
                for (int k = 0; k < 3; ++k) t += lam[6L * p.np + 3L * l + k] * vl[k].d;
            }
            out[axis] = -t;
        }
    }
    dobs[2L * p.obs_orig[s]] = out[0];
    dobs[2L * p.obs_orig[s] + 1] = out[1];
}

inline int grid_of(int64_t n, int block) { return static_cast<int>((n + block - 1) / block); }

}  // namespace

void launch_grad_tangent(const dev_problem& p, const dev_work64& w, const double* v, const double* dK, double* out,
                         hipStream_t s) {
    k_tangent kt{};
    if (dK)
        for (int i = 0; i < 4; ++i) kt.k[i] = dK[i];
    if (p.nl > 0)
        k_grad_tangent<<<grid_of(p.nl, kBlock), kBlock, 0, s>>>(p, w.pose, w.point, v, v ? v + 6L * p.np : nullptr, kt,
                                                                dK ? 1 : 0, w.adj_frozen, w.pose_grad, out + 6L * p.np);
    if (p.np > 0) k_pose_slots<<<grid_of(64L * p.np, 256), 256, 0, s>>>(p, w.pose_grad, out);
}

void launch_adjoint_precond(const dev_problem& p, const dev_work64& w, const double* r, double* z, hipStream_t s) {
    const double* rl = r + 6L * p.np;
    if (p.nl > 0) k_prec_points<<<grid_of(p.nl, kBlock), kBlock, 0, s>>>(p, w.Hinv, rl, w.gl);
    if (p.np > 0) {
        k_prec_poses<<<grid_of(64L * p.np, 256), 256, 0, s>>>(p, w.W, r, w.gl, w.rvec);
        launch_cholesky_resolve64(p.n, w, w.rvec, z, s, false);
        k_zero_fixed_poses<<<grid_of(6L * p.np, kBlock), kBlock, 0, s>>>(p, z);
    }
    if (p.nl > 0) k_prec_back<<<grid_of(p.nl, kBlock), kBlock, 0, s>>>(p, w.W, w.Hinv, rl, z, z + 6L * p.np);
}

void launch_adjoint_frozen(const dev_problem& p, const dev_work64& w, hipStream_t s) {
    if (p.nl > 0) k_mark_frozen<<<grid_of(p.nl, kBlock), kBlock, 0, s>>>(p, w.Hinv, w.adj_frozen);
}

void launch_zero_frozen_rows(const dev_problem& p, const dev_work64& w, double* v, hipStream_t s) {
    const int64_t N = 6L * p.np + 3L * p.nl;
    if (N > 0) k_zero_frozen_rows<<<grid_of(N, kBlock), kBlock, 0, s>>>(p, w.adj_frozen, v);
}

void launch_dot(const double* a, const double* b, int64_t n, double* partial, double* out, hipStream_t s) {
    int g = grid_of(n, kBlock);
    if (g > 512) g = 512;
    if (g < 1) g = 1;
    k_dot<<<g, kBlock, 0, s>>>(a, b, n, partial);
    k_sum<<<1, kBlock, 0, s>>>(partial, g, out);
}

void launch_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s) {
    if (n > 0) k_axpby<<<grid_of(n, kBlock), kBlock, 0, s>>>(a, x, b, y, n);
}

void launch_obs_grad(const dev_problem& p, const dev_work64& w, const double* lam, double* dobs_by_orig, hipStream_t s) {
    if (p.no > 0) k_obs_grad<<<grid_of(p.no, kBlock), kBlock, 0, s>>>(p, w.pose, w.point, lam, w.adj_frozen, dobs_by_orig);
}

}  // namespace zb
