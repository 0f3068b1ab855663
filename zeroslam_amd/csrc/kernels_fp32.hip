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
#include "lm_control.hpp"
#include "mixed_items.hpp"

namespace zb {
namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) k_lin_points_f(dev_problem p, const double* pose, const double* point,
                                                         float* pterms, double* pgrad, float* Hll, double* bl,
                                                         float* Wbuf, const lm_state* lm, int gated, uint8_t* frozen) {
    if (gated && (lm->done || !lm->need_linearize)) return;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < p.nl) mixed::lin_point(p, pose, point, pterms, pgrad, Hll, bl, Wbuf, frozen, l);
}

// Linearisation by blocks of observations, then one thread per (pose, term).
__global__ void __launch_bounds__(kLinBlock) k_lin_blocks_f(dev_problem p, const double* pose, const double* point,
                                                            float* Hll, double* bl, float* Wbuf, float* part_h,
                                                            double* part_g, float* wscratch, const lm_state* lm,
                                                            int gated, uint8_t* frozen) {
    if (gated && (lm->done || !lm->need_linearize)) return;
    __shared__ mixed::lin_stage<float> sh;
    mixed::lin_block(p, pose, point, Hll, bl, Wbuf, frozen, part_h, part_g, wscratch, static_cast<int>(blockIdx.x), sh);
}

__global__ void __launch_bounds__(mixed::kFinishThreads) k_lin_finish_f(dev_problem p, const float* part_h,
                                                                       const double* part_g, float* Hpp, double* bp,
                                                                       const lm_state* lm, int gated) {
    if (gated && (lm->done || !lm->need_linearize)) return;
    __shared__ double red[4][27];
    mixed::lin_pose_finish(p, part_h, part_g, Hpp, bp, static_cast<int>(blockIdx.x), red);
}

// One wave per (pose, chunk), then one thread per (pose, term).
__global__ void __launch_bounds__(256) k_reduce_pose_part_f(dev_problem p, const float* pterms, const double* pgrad,
                                                            double* part, const lm_state* lm, int gated) {
    if (gated && (lm->done || !lm->need_linearize)) return;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave < p.np * kPoseChunks)
        mixed::reduce_pose_chunk(p, pterms, pgrad, part, wave / kPoseChunks, wave % kPoseChunks, threadIdx.x & 63);
}

__global__ void __launch_bounds__(256) k_reduce_pose_f(dev_problem p, const double* part, float* Hpp, double* bp,
                                                       const lm_state* lm, int gated) {
    if (gated && (lm->done || !lm->need_linearize)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < p.np * 27) mixed::reduce_pose_finish(part, Hpp, bp, t / 27, t % 27);
}

__global__ void k_linearise_done_f(lm_state* lm) { lm_linearise_done(lm); }
__global__ void k_linearise_done_end_f(lm_state* lm) {
    lm_linearise_done(lm);
    lm_end_attempt(lm);
}

__global__ void __launch_bounds__(kBlock) k_damp_invert_f(dev_problem p, const float* Hll, float* Hinv, const lm_state* lm,
                                                          const double* lambda_override, const uint8_t* frozen) {
    if (!lambda_override && lm->done) return;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < p.nl) mixed::damp_invert_item(p, Hll, Hinv, frozen, lambda_override ? *lambda_override : lm->lambda, l);
}

__global__ void __launch_bounds__(kBlock) k_form_Y_f(dev_problem p, const float* Wbuf, const float* Hinv, float* Y,
                                                     const lm_state* lm, const double* lambda_override) {
    if (!lambda_override && lm->done) return;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < p.npair) mixed::form_Y_item(p, Wbuf, Hinv, Y, a);
}

// One workgroup per reduced-system block; kInFlight as in mixed::schur_block
// (the same result either way).  Blocks are numbered row by row, so a chunk of
// consecutive blocks shares its block rows' Y blocks: 64 per XCD chunk.
constexpr int kSchurChunk = 64;
template <int kInFlight>
__global__ void __launch_bounds__(mixed::kSchurThreads) k_schur_f(dev_problem p, const float* Y, const float* Wbuf,
                                                                   const float* Hpp, const double* bp, const double* bl,
                                                                   double* S, double* bs, const lm_state* lm,
                                                                   const double* lambda_override) {
    if (!lambda_override && lm->done) return;
    __shared__ mixed::schur_lds<float> red;
    const int blk = xcd_block<kSchurChunk>(blockIdx.x, gridDim.x);
    if (blk < p.nblk)
        mixed::schur_block<float, kInFlight>(p, Y, Wbuf, Hpp, bp, bl, S, bs,
                                             lambda_override ? *lambda_override : lm->lambda, blk, threadIdx.x, red);
}

__global__ void __launch_bounds__(kBlock) k_fill_zero_d(double* a, int64_t n, const lm_state* lm, int gated) {
    if (gated && lm->done) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = 0.0;
}

inline int grid_of(int64_t n, int block) { return static_cast<int>((n + block - 1) / block); }

// Below this many reduced-system blocks the Schur kernel keeps two contributions
// in flight per lane.
constexpr int kSchurSmallGrid = 2048;

}  // namespace

void launch_linearise_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s, bool gated, bool end_attempt) {
    if (p.n_lblk > 0) {
        float* part_h = static_cast<float*>(w.lin_part_h);
        ZB_TIME_BEGIN(w, K_LINEARISE, s);
        k_lin_blocks_f<<<p.n_lblk, kLinBlock, 0, s>>>(p, w.pose, w.point, w.Hll_f, w.bl_m, w.W_f, part_h, w.lin_part_g,
                                                      w.pose_terms_f, w.lm, gated ? 1 : 0, w.frozen);
        ZB_TIME_END(w, s);
        if (p.np > 0) {
            ZB_TIME_BEGIN(w, K_POSE_REDUCE, s);
            k_lin_finish_f<<<p.np, mixed::kFinishThreads, 0, s>>>(p, part_h, w.lin_part_g, w.Hpp_f, w.bp_m, w.lm,
                                                                 gated ? 1 : 0);
            ZB_TIME_END(w, s);
        }
        if (gated && end_attempt) k_linearise_done_end_f<<<1, 1, 0, s>>>(w.lm);
        else if (gated) k_linearise_done_f<<<1, 1, 0, s>>>(w.lm);
        return;
    }
    if (p.nl > 0) {
        ZB_TIME_BEGIN(w, K_LINEARISE, s);
        k_lin_points_f<<<grid_of(p.nl, kBlock), kBlock, 0, s>>>(p, w.pose, w.point, w.pose_terms_f, w.pose_grad, w.Hll_f,
                                                                 w.bl_m, w.W_f, w.lm, gated ? 1 : 0, w.frozen);
        ZB_TIME_END(w, s);
    }
    if (p.np > 0) {
        ZB_TIME_BEGIN(w, K_POSE_REDUCE, s);
        k_reduce_pose_part_f<<<grid_of(64L * p.np * kPoseChunks, 256), 256, 0, s>>>(p, w.pose_terms_f, w.pose_grad,
                                                                                     w.pose_part, w.lm, gated ? 1 : 0);
        k_reduce_pose_f<<<grid_of(27L * p.np, 256), 256, 0, s>>>(p, w.pose_part, w.Hpp_f, w.bp_m, w.lm, gated ? 1 : 0);
        ZB_TIME_END(w, s);
    }
    if (gated && end_attempt) k_linearise_done_end_f<<<1, 1, 0, s>>>(w.lm);
    else if (gated) k_linearise_done_f<<<1, 1, 0, s>>>(w.lm);
}

void launch_reduced_system_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s, const double* lambda_override) {
    if (p.nl > 0) {
        ZB_TIME_BEGIN(w, K_INVERT, s);
        k_damp_invert_f<<<grid_of(p.nl, kBlock), kBlock, 0, s>>>(p, w.Hll_f, w.Hinv_f, w.lm, lambda_override,
                                                                 w.frozen);
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
        // Two contributions in flight per lane hide more latency while the grid
        // is small; with more workgroups than the GPU holds at once, one (and
        // the higher occupancy it allows) is faster.
        if (p.nblk < kSchurSmallGrid)
            k_schur_f<2><<<p.nblk, mixed::kSchurThreads, 0, s>>>(p, w.Y_f, w.W_f, w.Hpp_f, w.bp_m, w.bl_m, w.S, w.bs,
                                                                 w.lm, lambda_override);
        else
            k_schur_f<1><<<p.nblk, mixed::kSchurThreads, 0, s>>>(p, w.Y_f, w.W_f, w.Hpp_f, w.bp_m, w.bl_m, w.S, w.bs,
                                                                 w.lm, lambda_override);
        ZB_TIME_END(w, s);
    }
}

}  // namespace zb
