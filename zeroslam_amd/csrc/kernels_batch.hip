// Batched local bundle adjustment (factor_graph::solve, factor_graph.hpp:489-615,
// on many independent windows), in lockstep.  Each step, every unfinished window
// makes one LM attempt (after its linearisation, when its last attempt was
// accepted), and each phase of the attempt is one grid-wide launch over every
// window's items: points, (pose, point) pairs, reduced-system blocks, pose rows,
// observations.  The GPU is shared out by item count, so a window of 40k
// observations and one of 5k finish a phase together, and every phase runs at
// full occupancy.  The factorisation and the substitutions stay one workgroup per
// window.
//
// A workgroup finds its window and its index inside the window in a table built
// on the host (batch_grid).  The workgroups of one window are dealt to one XCD
// (workgroup i runs on XCD i mod 8; windows go to XCDs balanced by observation
// count), so a window's Y and W blocks, read again by every Schur contribution,
// stay in that XCD's L2.
//
// The per-item arithmetic is the shared mixed_items.hpp code, the controller the
// shared lm_control.hpp code, and every sum keeps the association of the
// per-graph kernels (block trees of 256 threads over the per-graph grid, then a
// 1024-wide tree), so every window's result equals its per-graph solve.
#include "ba_math.hpp"
#include "chol_device.hpp"
#include "common.hpp"
#include "lm_control.hpp"
#include "mixed_items.hpp"

namespace zb {
namespace {

constexpr int kBatchTiles = kBatchMaxN / wg::kTile;

// Hessian-block buffers of a window for storage type TS.
template <class TS>
struct blocks;
template <>
struct blocks<float> {
    float *pterms, *Hpp, *Hll, *W, *Hinv, *Y;
    __device__ explicit blocks(const dev_work64& w)
        : pterms(w.pose_terms_f), Hpp(w.Hpp_f), Hll(w.Hll_f), W(w.W_f), Hinv(w.Hinv_f), Y(w.Y_f) {}
};
template <>
struct blocks<double> {
    double *pterms, *Hpp, *Hll, *W, *Hinv, *Y;
    __device__ explicit blocks(const dev_work64& w)
        : pterms(w.pose_terms), Hpp(w.Hpp), Hll(w.Hll), W(w.W), Hinv(w.Hinv), Y(w.Y) {}
};

__device__ __forceinline__ bool needs_linearisation(const lm_state* lm) { return !lm->done && lm->need_linearize; }

// Phase clock: workgroup 0 of every launch stamps the device clock (100 MHz), and
// the time since the previous launch's stamp goes to that launch's phase.
// pc[0] last stamp, pc[1] its phase, pc[2 + k] ticks of phase k.
__device__ __forceinline__ void stamp(uint64_t* pc, int phase) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint64_t t = wall_clock64();
        if (pc[0] != 0) pc[2 + pc[1]] += t - pc[0];
        pc[0] = t;
        pc[1] = static_cast<uint64_t>(phase);
    }
}

// Block-wide halving-tree sum of one value per thread (the per-graph kernels'
// tree).
__device__ double tree_sum(double v, double* sh) {
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

// Workgroups of the per-graph k_cost / k_backsub_update grids over n items.
__device__ __forceinline__ int64_t sum_grid(int64_t n) {
    const int64_t G = (n + 255) / 256;
    return G < 1 ? 1 : (G > kPartials ? kPartials : G);
}

// The same sum formed by one workgroup (a multiple of 256 threads), blockDim / 256
// of the per-graph blocks at a time: the initial cost of k_batch_begin.  sh holds
// blockDim values, part 1024.
template <class F>
__device__ double grid_sum(int64_t n, F item, double* sh, double* part) {
    const int tid = threadIdx.x, t = tid & 255, per = blockDim.x >> 8;
    const int64_t G = sum_grid(n);
    for (int64_t b0 = 0; b0 < G; b0 += per) {
        const int64_t b = b0 + (tid >> 8);
        double acc = 0.0;
        if (b < G)
            for (int64_t i = b * 256 + t; i < n; i += G * 256) acc = acc + item(i);
        sh[tid] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w) sh[tid] = sh[tid] + sh[tid + w];
            __syncthreads();
        }
        if (t == 0 && b < G) part[b] = sh[tid];
        __syncthreads();
    }
    for (int i = tid; i < 1024; i += blockDim.x)
        if (i >= G) part[i] = 0.0;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        for (int i = tid; i < w; i += blockDim.x) part[i] = part[i] + part[i + w];
        __syncthreads();
    }
    const double r = part[0];
    __syncthreads();
    return r;
}

// The workgroup factorisation and solve, inlined (as calls they save and restore
// the kernel's registers through scratch memory).  The factorisation loads its
// broadcasts 8 at a time: the same arithmetic as the per-graph kernel's 32 at a
// time, in fewer registers.
__device__ __forceinline__ bool batch_factor(int n, double* S, const int32_t* tile_last, double eps,
                                             double (*PT)[wg::kTile][wg::kPT], int* flag, double* recips,
                                             double (*col)[64]) {
    return wg::fused_factor<kBatchTiles, true, 8>(n, S, tile_last, eps, PT, flag, recips, col);
}

template <int kThreads>
__device__ __forceinline__ int batch_solve(int n, const double* S, const double* b, double* x, const int32_t* tile_first,
                                           const int32_t* tile_last, double (*buf)[wg::kTile][wg::kTile + 1], double* sh,
                                           int* flag) {
    return wg::triangular_solve_fast<kThreads>(n, S, b, x, tile_first, tile_last, buf, sh, flag);
}

// ---- start of a solve --------------------------------------------------------
// One workgroup per window: controller reset, the initial cost, the first
// linearisation requested.  Workgroup 0 clears the phase clock first.
__global__ void __launch_bounds__(256) k_batch_begin(const batch_window* wins, uint64_t* pc) {
    __shared__ double sh[256];
    __shared__ double part[1024];
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && tid == 0)
        for (int i = 0; i < 2 + kBatchPhases; ++i) pc[i] = 0;
    stamp(pc, PH_COST);
    const batch_window& bw = wins[blockIdx.x];
    const dev_problem& p = bw.p;
    const dev_work64& w = bw.w;
    lm_state* lm = w.lm;
    if (tid == 0) {
        *lm = lm_state{};
        lm->iterations = bw.iterations;
        lm->use_relative = bw.use_relative;
        lm->require_descent = bw.require_descent;
        lm->need_linearize = 1;
    }
    const double c = grid_sum(
        p.no, [&](int64_t s) { return mixed::cost_item(p, w.pose, w.point, w.residual, static_cast<int>(s)); }, sh,
        part);
    if (tid == 0) lm->chi2 = c;
}

// ---- linearisation of the windows whose last attempt was accepted --------------
// Windows with linearisation blocks (mixed): one workgroup per block, then one
// per pose.  The others: one thread per point, one wave per pose chunk (the pose
// terms formed again from the observations and summed in a fixed order), one
// thread per pose-block entry.
template <class TS>
__global__ void __launch_bounds__(kLinBlock) k_batch_lin_blocks(const batch_window* wins, const int2* map,
                                                                uint64_t* pc) {
    stamp(pc, PH_LINEARISE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (!needs_linearisation(bw.w.lm)) return;
    __shared__ mixed::lin_stage<TS> sh;
    const blocks<TS> B(bw.w);
    mixed::lin_block(bw.p, bw.w.pose, bw.w.point, B.Hll, bw.w.bl_m, B.W, bw.w.frozen, static_cast<TS*>(bw.w.lin_part_h),
                     bw.w.lin_part_g, B.pterms, m.y, sh);
}

template <class TS>
__global__ void __launch_bounds__(mixed::kFinishThreads) k_batch_lin_finish(const batch_window* wins, const int2* map,
                                                                           uint64_t* pc) {
    stamp(pc, PH_LINEARISE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (!needs_linearisation(bw.w.lm)) return;
    __shared__ double red[4][27];
    const blocks<TS> B(bw.w);
    mixed::lin_pose_finish(bw.p, static_cast<const TS*>(bw.w.lin_part_h), bw.w.lin_part_g, B.Hpp, bw.w.bp_m, m.y, red);
}

template <class TS>
__global__ void __launch_bounds__(256) k_batch_lin_points(const batch_window* wins, const int2* map, uint64_t* pc) {
    stamp(pc, PH_LINEARISE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (!needs_linearisation(bw.w.lm)) return;
    const int l = m.y * 256 + static_cast<int>(threadIdx.x);
    if (l >= bw.p.nl) return;
    const blocks<TS> B(bw.w);
    mixed::lin_point<TS, false>(bw.p, bw.w.pose, bw.w.point, B.pterms, bw.w.pose_grad, B.Hll, bw.w.bl_m, B.W,
                                bw.w.frozen, l);
}

template <class TS>
__global__ void __launch_bounds__(256) k_batch_lin_chunks(const batch_window* wins, const int2* map, uint64_t* pc) {
    stamp(pc, PH_LINEARISE);
    __shared__ TS ht[4][64][21];
    __shared__ double gt[4][64][6];
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (!needs_linearisation(bw.w.lm)) return;
    const int wave = threadIdx.x >> 6;
    const int ic = m.y * 4 + wave;  // one wave per chunk
    if (ic >= bw.p.np * kPoseChunks) return;
    mixed::reduce_pose_chunk_recompute(bw.p, bw.w.pose, bw.w.point, bw.w.pose_part, ic / kPoseChunks, ic % kPoseChunks,
                                       threadIdx.x & 63, ht[wave], gt[wave]);
}

template <class TS>
__global__ void __launch_bounds__(256) k_batch_lin_entries(const batch_window* wins, const int2* map, uint64_t* pc) {
    stamp(pc, PH_LINEARISE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (!needs_linearisation(bw.w.lm)) return;
    const int ij = m.y * 256 + static_cast<int>(threadIdx.x);
    if (ij >= 27 * bw.p.np) return;
    const blocks<TS> B(bw.w);
    mixed::reduce_pose_finish(bw.w.pose_part, B.Hpp, bw.w.bp_m, ij / 27, ij % 27);
}

// ---- one LM attempt of every unfinished window --------------------------------
// Start of the attempt, one workgroup per window: after the first linearisation
// the controller starts from the linearised system (max diagonal, lm_init);
// after a later one, the accepted attempt that asked for it ends here.  Every
// other attempt of the window ended at the start of this one.
template <class TS>
__global__ void __launch_bounds__(256) k_batch_step_begin(const batch_window* wins, const int2* map, uint64_t* pc) {
    stamp(pc, PH_UPDATE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    const dev_problem& p = bw.p;
    lm_state* lm = bw.w.lm;
    __shared__ max_scan ms[256];
    const int tid = threadIdx.x;
    const bool done = lm->done != 0, first = lm->lin_count == 0;
    if (first && !done) {
        // max_diagonal over the pose then the point diagonal entries, in order
        const blocks<TS> B(bw.w);
        const int64_t total = 6L * p.np + 3L * p.nl;
        const int64_t per = (total + blockDim.x - 1) / blockDim.x;
        const int64_t i0 = tid * per, i1 = (i0 + per < total) ? i0 + per : total;
        max_scan acc = {0.0, 0, 0, 1};
        for (int64_t i = i0; i < i1; ++i) {
            const double h = (i < 6L * p.np)
                                 ? static_cast<double>(B.Hpp[21 * (i / 6) + (i % 6) * (i % 6 + 1) / 2 + i % 6])
                                 : static_cast<double>(B.Hll[9 * ((i - 6L * p.np) / 3) + 4 * ((i - 6L * p.np) % 3)]);
            acc = scan_push(acc, h);
        }
        ms[tid] = acc;
        __syncthreads();
        if (tid == 0) {
            max_scan all = {0.0, 0, 0, 1};
            for (unsigned t = 0; t < blockDim.x; ++t) all = scan_combine(all, ms[t]);
            lm_init(lm, scan_max_diagonal(all));
        }
    } else if (tid == 0) {
        if (!done && lm->need_linearize) lm_linearise_done(lm);
        lm_end_attempt(lm);
    }
}

// Damping and point-block inverses (items [0, nl)), and the reduced system
// cleared (items [nl, nl + n^2)).
template <class TS>
__global__ void __launch_bounds__(256) k_batch_invert(const batch_window* wins, const int2* map, uint64_t* pc) {
    stamp(pc, PH_INVERT);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    const lm_state* lm = bw.w.lm;
    if (lm->done) return;
    const dev_problem& p = bw.p;
    const int64_t i = static_cast<int64_t>(m.y) * 256 + threadIdx.x;
    if (i < p.nl) {
        const blocks<TS> B(bw.w);
        mixed::damp_invert_item(p, B.Hll, B.Hinv, bw.w.frozen, lm->lambda, static_cast<int>(i));
    } else if (i - p.nl < static_cast<int64_t>(p.n) * p.n) {
        bw.w.S[i - p.nl] = 0.0;
    }
}

template <class TS>
__global__ void __launch_bounds__(256) k_batch_form_Y(const batch_window* wins, const int2* map, uint64_t* pc) {
    stamp(pc, PH_INVERT);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (bw.w.lm->done) return;
    const int a = m.y * 256 + static_cast<int>(threadIdx.x);
    if (a >= bw.p.npair) return;
    const blocks<TS> B(bw.w);
    mixed::form_Y_item(bw.p, B.W, B.Hinv, B.Y, a);
}

// One workgroup per reduced-system block.
template <class TS, int kInFlight>
__global__ void __launch_bounds__(mixed::kSchurThreads) k_batch_schur(const batch_window* wins, const int2* map,
                                                                       uint64_t* pc) {
    stamp(pc, PH_SCHUR);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    const lm_state* lm = bw.w.lm;
    if (lm->done) return;
    const int blk = m.y;
    if (blk >= bw.p.nblk) return;
    __shared__ mixed::schur_lds<TS> red;
    const blocks<TS> B(bw.w);
    const dev_work64& w = bw.w;
    mixed::schur_block<TS, kInFlight>(bw.p, B.Y, B.W, B.Hpp, w.bp_m, w.bl_m, w.S, w.bs, lm->lambda, blk,
                                      static_cast<int>(threadIdx.x), red);
}

// One workgroup per window: the diagonal floor, the envelope factorisation and
// the solve for dx_p.  Without refinement the controller's failure bookkeeping
// follows here.
template <int kThreads>
__global__ void __launch_bounds__(kThreads) k_batch_factor(const batch_window* wins, const int2* map, uint64_t* pc,
                                                           int mixed_path) {
    stamp(pc, PH_FACTOR);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    const dev_problem& p = bw.p;
    const dev_work64& w = bw.w;
    lm_state* lm = w.lm;
    if (lm->done) return;
    __shared__ double sh[kThreads];
    __shared__ double PT[kBatchTiles - 1][wg::kTile][wg::kPT];  // factor panels, then the solve's tiles
    static_assert((kBatchTiles - 1) * wg::kPT >= 6 * (wg::kTile + 1), "solve scratch");
    auto solve_buf = reinterpret_cast<double (*)[wg::kTile][wg::kTile + 1]>(&PT[0][0][0]);
    __shared__ double recips[wg::kTile];
    __shared__ double col[wg::kTile][64];
    __shared__ double xs[kBatchMaxN];
    __shared__ int flag;
    const int tid = threadIdx.x, n = p.n;
    double mx = 0.0;
    for (int i = tid; i < n; i += blockDim.x) {
        const double a = fabs(w.S[static_cast<int64_t>(i) * n + i]);
        mx = (mx < a) ? a : mx;
    }
    mx = wg::block_max(mx, sh);
    const double lift = bw.floor_rel * mx;
    if (bw.floor_rel > 0.0)
        for (int i = tid; i < n; i += blockDim.x) w.S[static_cast<int64_t>(i) * n + i] += lift;
    __syncthreads();
    bool ok = batch_factor(n, w.S, w.tile_last, (mx + lift) * 1e-14, PT, &flag, recips, col);
    if (ok) ok = batch_solve<kThreads>(n, w.S, w.bs, xs, w.tile_first, w.tile_last, solve_buf, sh, &flag) == 0;
    if (ok)
        for (int i = tid; i < n; i += blockDim.x) w.dxp[i] = xs[i];
    if (tid == 0) {
        lm->solve_ok = ok ? 1 : 0;
        if (!mixed_path || bw.refine_steps <= 0) lm_after_solve(lm);
    }
}

// dx_l = H_ll^-1 (b_l - sum W^T dx_p), one thread per point (mixed windows with
// linearisation blocks form it in their first refinement round).
template <class TS>
__global__ void __launch_bounds__(256) k_batch_point_step(const batch_window* wins, const int2* map, uint64_t* pc) {
    stamp(pc, PH_UPDATE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    const lm_state* lm = bw.w.lm;
    if (lm->done || !lm->solve_ok) return;
    if (sizeof(TS) == 4 && bw.p.n_lblk > 0 && bw.refine_steps > 0) return;  // k_batch_refine_blocks forms them
    const int l = m.y * 256 + static_cast<int>(threadIdx.x);
    if (l >= bw.p.nl) return;
    const blocks<TS> B(bw.w);
    mixed::point_step_item(bw.p, B.W, B.Hinv, bw.w.bl_m, bw.w.dxp, bw.w.dxl, l);
}

// ---- binary64 refinement round k of the mixed solve (mixed_items.hpp) ----------
__device__ __forceinline__ bool refining(const batch_window& bw, int k) {
    const lm_state* lm = bw.w.lm;
    return !lm->done && lm->solve_ok && k < bw.refine_steps;
}

// Windows with linearisation blocks: one workgroup per block, then one wave per
// pose (the per-graph k_refine_blocks, k_refine_finish).
template <class TS>
__global__ void __launch_bounds__(kLinBlock) k_batch_refine_blocks(const batch_window* wins, const int2* map,
                                                                   uint64_t* pc, int k) {
    stamp(pc, PH_REFINE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (!refining(bw, k)) return;
    __shared__ mixed::refine_stage sh;
    const dev_work64& w = bw.w;
    const blocks<TS> B(w);
    mixed::refine_block(bw.p, w.pose, w.point, B.W, B.Hinv, w.bl_m, w.dxp, w.dxl, w.zl, w.gl, w.lin_part_g,
                        w.lm->lambda, k == 0, m.y, sh);
}

__global__ void __launch_bounds__(64) k_batch_refine_finish(const batch_window* wins, const int2* map, uint64_t* pc,
                                                            int k) {
    stamp(pc, PH_REFINE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (!refining(bw, k)) return;
    const dev_work64& w = bw.w;
    mixed::refine_pose_finish(bw.p, w.lin_part_g, w.bp_m, w.dxp, w.rvec, w.lm->lambda, m.y,
                              static_cast<int>(threadIdx.x));
}

// The other windows: one thread per point, one wave per pose chunk, one thread
// per pose row.
template <class TS>
__global__ void __launch_bounds__(256) k_batch_full_residual(const batch_window* wins, const int2* map, uint64_t* pc,
                                                             int k) {
    stamp(pc, PH_REFINE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (!refining(bw, k)) return;
    const int l = m.y * 256 + static_cast<int>(threadIdx.x);
    if (l >= bw.p.nl) return;
    const dev_work64& w = bw.w;
    const blocks<TS> B(w);
    mixed::full_residual_item(bw.p, w.pose, w.point, w.dxp, w.dxl, w.bl_m, B.Hinv, w.pose_grad, w.zl, w.gl,
                              w.lm->lambda, l);
}

template <class TS>
__global__ void __launch_bounds__(256) k_batch_pose_chunk(const batch_window* wins, const int2* map, uint64_t* pc,
                                                          int k) {
    stamp(pc, PH_REFINE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (!refining(bw, k)) return;
    const int ic = m.y * 4 + static_cast<int>(threadIdx.x >> 6);  // one wave per (pose, chunk)
    if (ic >= bw.p.np * kPoseChunks) return;
    const dev_work64& w = bw.w;
    const blocks<TS> B(w);
    mixed::pose_residual_chunk(bw.p, w.pose_grad, B.W, w.gl, w.pose_part, ic / kPoseChunks, ic % kPoseChunks,
                               threadIdx.x & 63);
}

__global__ void __launch_bounds__(256) k_batch_pose_rows(const batch_window* wins, const int2* map, uint64_t* pc,
                                                         int k) {
    stamp(pc, PH_REFINE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (!refining(bw, k)) return;
    const int ir = m.y * 256 + static_cast<int>(threadIdx.x);
    if (ir >= 6 * bw.p.np) return;
    const dev_work64& w = bw.w;
    mixed::pose_residual_finish(bw.p, w.pose_part, w.bp_m, w.dxp, w.rvec, w.lm->lambda, ir / 6, ir % 6);
}

// One workgroup per window: the correction solve with the factor, and after the
// window's last round the controller's failure bookkeeping.
template <int kThreads>
__global__ void __launch_bounds__(kThreads) k_batch_resolve(const batch_window* wins, const int2* map, uint64_t* pc,
                                                            int k) {
    stamp(pc, PH_SOLVE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    const dev_work64& w = bw.w;
    lm_state* lm = w.lm;
    if (lm->done) return;
    const bool active = lm->solve_ok && k < bw.refine_steps, last = k == bw.refine_steps - 1;
    if (!active && !last) return;
    __shared__ double sh[kThreads];
    __shared__ double buf[6][wg::kTile][wg::kTile + 1];
    __shared__ double xs[kBatchMaxN];
    __shared__ int flag;
    const int n = bw.p.n;
    bool ok = true;
    if (active) {
        ok = batch_solve<kThreads>(n, w.S, w.rvec, xs, w.tile_first, w.tile_last, buf, sh, &flag) == 0;
        if (ok)
            for (int i = threadIdx.x; i < n; i += blockDim.x) w.corr[i] = xs[i];
    }
    if (threadIdx.x == 0) {
        if (active && !ok) lm->solve_ok = 0;
        if (last) lm_after_solve(lm);
    }
}

// dx_l += H_ll^-1 (r_l - W^T delta_p) (items [0, nl)); dx_p += delta_p (items
// [nl, nl + n)).
template <class TS>
__global__ void __launch_bounds__(256) k_batch_refine_apply(const batch_window* wins, const int2* map, uint64_t* pc,
                                                            int k) {
    stamp(pc, PH_REFINE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    if (!refining(bw, k)) return;
    const dev_work64& w = bw.w;
    const int t = m.y * 256 + static_cast<int>(threadIdx.x);
    if (t < bw.p.nl) {
        const blocks<TS> B(w);
        mixed::refine_point_item(bw.p, B.W, B.Hinv, w.zl, w.corr, w.dxl, t);
    } else if (t - bw.p.nl < bw.p.n) {
        w.dxp[t - bw.p.nl] += w.corr[t - bw.p.nl];
    }
}

// The step applied (poses then points) with the LM scale terms, workgroup b of
// the window's per-graph grid of sum_grid(np + nl) blocks; block sums to
// partial[2 kPartials + b].
__global__ void __launch_bounds__(256) k_batch_apply(const batch_window* wins, const int2* map, uint64_t* pc) {
    stamp(pc, PH_UPDATE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    const dev_problem& p = bw.p;
    const dev_work64& w = bw.w;
    const lm_state* lm = w.lm;
    if (lm->done || !lm->step_applied) return;
    __shared__ double sh[256];
    const int64_t slots = static_cast<int64_t>(p.np) + p.nl, G = sum_grid(slots);
    const double lam = lm->lambda;
    double acc = 0.0;
    for (int64_t i = static_cast<int64_t>(m.y) * 256 + threadIdx.x; i < slots; i += G * 256)
        acc = acc + mixed::apply_step_item(p, w.pose, w.pose_backup, w.point, w.point_backup, w.dxp, w.dxl, w.bp_m, w.bl_m,
                                           lam, static_cast<int>(i));
    const double b = tree_sum(acc, sh);
    if (threadIdx.x == 0) w.partial[2 * kPartials + m.y] = b;
}

// The cost at the updated state, workgroup b of the per-graph grid of
// sum_grid(no) blocks; block sums to partial[b].
__global__ void __launch_bounds__(256) k_batch_cost(const batch_window* wins, const int2* map, uint64_t* pc) {
    stamp(pc, PH_COST);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    const dev_problem& p = bw.p;
    const dev_work64& w = bw.w;
    const lm_state* lm = w.lm;
    if (lm->done || !lm->step_applied) return;
    __shared__ double sh[256];
    const int64_t G = sum_grid(p.no);
    double acc = 0.0;
    for (int64_t s = static_cast<int64_t>(m.y) * 256 + threadIdx.x; s < p.no; s += G * 256)
        acc = acc + mixed::cost_item(p, w.pose, w.point, w.residual, static_cast<int>(s));
    const double b = tree_sum(acc, sh);
    if (threadIdx.x == 0) w.partial[m.y] = b;
}

// One workgroup of 1024 threads per window: the two sums' 1024-wide trees (as
// k_sum_partials), then accept or reject.
__global__ void __launch_bounds__(1024) k_batch_decide(const batch_window* wins, const int2* map, uint64_t* pc) {
    stamp(pc, PH_UPDATE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    const dev_problem& p = bw.p;
    const dev_work64& w = bw.w;
    lm_state* lm = w.lm;
    if (lm->done || !lm->step_applied) return;
    __shared__ double sh[1024];
    const int tid = threadIdx.x;
    const int64_t Ga = sum_grid(static_cast<int64_t>(p.np) + p.nl), Gc = sum_grid(p.no);
    const double scale = tree_sum(tid < Ga ? w.partial[2 * kPartials + tid] : 0.0, sh);
    const double c = tree_sum(tid < Gc ? w.partial[tid] : 0.0, sh);
    if (tid == 0) {
        lm->scale = scale + 1e-3;  // factor_graph.hpp:556-560
        lm->chi2_new = c;
        lm_decide(lm);
    }
}

__global__ void __launch_bounds__(256) k_batch_rollback(const batch_window* wins, const int2* map, uint64_t* pc) {
    stamp(pc, PH_UPDATE);
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const batch_window& bw = wins[m.x];
    const dev_work64& w = bw.w;
    if (!w.lm->need_rollback) return;
    const int t = m.y * 256 + static_cast<int>(threadIdx.x);
    mixed::rollback_item(bw.p, w.pose, w.pose_backup, w.point, w.point_backup, t);
}

// End of a run of steps: finished windows end their last attempt; every
// controller is copied to out; *all_done = 1 when every window has finished.
__global__ void __launch_bounds__(256) k_batch_finish(const batch_window* wins, int count, lm_state* out,
                                                      int32_t* all_done, uint64_t* pc) {
    stamp(pc, PH_UPDATE);
    __shared__ int busy;
    if (threadIdx.x == 0) busy = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        lm_state* lm = wins[i].w.lm;
        if (lm->done) lm_end_attempt(lm);
        else atomicOr(&busy, 1);
        out[i] = *lm;
    }
    __syncthreads();
    if (threadIdx.x == 0) *all_done = busy ? 0 : 1;
}

inline unsigned wgs(const batch_grid& g, int k) { return static_cast<unsigned>(g.n_wg[k]); }

template <int kThreads>
void launch_factor(const batch_window* wins, const batch_grid& g, uint64_t* pc, bool mixed_path, hipStream_t s) {
    k_batch_factor<kThreads><<<wgs(g, BM_WIN), kThreads, 0, s>>>(wins, g.map[BM_WIN], pc, mixed_path ? 1 : 0);
}

template <int kThreads>
void launch_resolve(const batch_window* wins, const batch_grid& g, uint64_t* pc, int k, hipStream_t s) {
    k_batch_resolve<kThreads><<<wgs(g, BM_WIN), kThreads, 0, s>>>(wins, g.map[BM_WIN], pc, k);
}

template <class TS>
void enqueue_step(const batch_window* wins, const batch_grid& g, int threads, int inflight, int refine_rounds,
                  uint64_t* pc, hipStream_t s) {
    constexpr bool mixed_path = sizeof(TS) == 4;
    // linearisation of the windows that need one
    if (wgs(g, BM_LBLK)) k_batch_lin_blocks<TS><<<wgs(g, BM_LBLK), kLinBlock, 0, s>>>(wins, g.map[BM_LBLK], pc);
    if (wgs(g, BM_LPTS)) k_batch_lin_points<TS><<<wgs(g, BM_LPTS), 256, 0, s>>>(wins, g.map[BM_LPTS], pc);
    if (wgs(g, BM_LCHUNK)) k_batch_lin_chunks<TS><<<wgs(g, BM_LCHUNK), 256, 0, s>>>(wins, g.map[BM_LCHUNK], pc);
    if (wgs(g, BM_LENTRY)) k_batch_lin_entries<TS><<<wgs(g, BM_LENTRY), 256, 0, s>>>(wins, g.map[BM_LENTRY], pc);
    if (wgs(g, BM_LPOSE))
        k_batch_lin_finish<TS><<<wgs(g, BM_LPOSE), mixed::kFinishThreads, 0, s>>>(wins, g.map[BM_LPOSE], pc);
    // the attempt
    k_batch_step_begin<TS><<<wgs(g, BM_WIN), 256, 0, s>>>(wins, g.map[BM_WIN], pc);
    if (wgs(g, BM_INVZ)) k_batch_invert<TS><<<wgs(g, BM_INVZ), 256, 0, s>>>(wins, g.map[BM_INVZ], pc);
    if (wgs(g, BM_PAIR)) k_batch_form_Y<TS><<<wgs(g, BM_PAIR), 256, 0, s>>>(wins, g.map[BM_PAIR], pc);
    if (wgs(g, BM_SBLK)) {
        if (inflight == 2)
            k_batch_schur<TS, 2><<<wgs(g, BM_SBLK), mixed::kSchurThreads, 0, s>>>(wins, g.map[BM_SBLK], pc);
        else
            k_batch_schur<TS, 1><<<wgs(g, BM_SBLK), mixed::kSchurThreads, 0, s>>>(wins, g.map[BM_SBLK], pc);
    }
    if (threads == 256) launch_factor<256>(wins, g, pc, mixed_path, s);
    else if (threads == 512) launch_factor<512>(wins, g, pc, mixed_path, s);
    else launch_factor<1024>(wins, g, pc, mixed_path, s);
    if (wgs(g, BM_PTS)) k_batch_point_step<TS><<<wgs(g, BM_PTS), 256, 0, s>>>(wins, g.map[BM_PTS], pc);
    for (int k = 0; mixed_path && k < refine_rounds; ++k) {
        if (wgs(g, BM_LBLK))
            k_batch_refine_blocks<TS><<<wgs(g, BM_LBLK), kLinBlock, 0, s>>>(wins, g.map[BM_LBLK], pc, k);
        if (wgs(g, BM_LPOSE)) k_batch_refine_finish<<<wgs(g, BM_LPOSE), 64, 0, s>>>(wins, g.map[BM_LPOSE], pc, k);
        // windows linearised per point (the pose-row launch reuses the entry table:
        // 27 np >= 6 np items)
        if (wgs(g, BM_LPTS)) k_batch_full_residual<TS><<<wgs(g, BM_LPTS), 256, 0, s>>>(wins, g.map[BM_LPTS], pc, k);
        if (wgs(g, BM_LCHUNK)) k_batch_pose_chunk<TS><<<wgs(g, BM_LCHUNK), 256, 0, s>>>(wins, g.map[BM_LCHUNK], pc, k);
        if (wgs(g, BM_LENTRY)) k_batch_pose_rows<<<wgs(g, BM_LENTRY), 256, 0, s>>>(wins, g.map[BM_LENTRY], pc, k);
        if (threads == 256) launch_resolve<256>(wins, g, pc, k, s);
        else if (threads == 512) launch_resolve<512>(wins, g, pc, k, s);
        else launch_resolve<1024>(wins, g, pc, k, s);
        if (wgs(g, BM_PTSN)) k_batch_refine_apply<TS><<<wgs(g, BM_PTSN), 256, 0, s>>>(wins, g.map[BM_PTSN], pc, k);
    }
    if (wgs(g, BM_APPLY)) k_batch_apply<<<wgs(g, BM_APPLY), 256, 0, s>>>(wins, g.map[BM_APPLY], pc);
    if (wgs(g, BM_COST)) k_batch_cost<<<wgs(g, BM_COST), 256, 0, s>>>(wins, g.map[BM_COST], pc);
    k_batch_decide<<<wgs(g, BM_WIN), 1024, 0, s>>>(wins, g.map[BM_WIN], pc);
    if (wgs(g, BM_SLOT)) k_batch_rollback<<<wgs(g, BM_SLOT), 256, 0, s>>>(wins, g.map[BM_SLOT], pc);
}

}  // namespace

void launch_batch_begin(const batch_window* wins, int count, uint64_t* pc, hipStream_t s) {
    if (count > 0) k_batch_begin<<<count, 256, 0, s>>>(wins, pc);
}

void launch_batch_steps(const batch_window* wins, const batch_grid& g, int threads, int inflight, bool mixed_path,
                        int refine_rounds, int steps, uint64_t* pc, hipStream_t s) {
    if (g.n_wg[BM_WIN] <= 0) return;
    for (int i = 0; i < steps; ++i) {
        if (mixed_path) enqueue_step<float>(wins, g, threads, inflight, refine_rounds, pc, s);
        else enqueue_step<double>(wins, g, threads, inflight, 0, pc, s);
    }
}

void launch_batch_finish(const batch_window* wins, int count, lm_state* out, int32_t* all_done, uint64_t* pc,
                         hipStream_t s) {
    if (count > 0) k_batch_finish<<<1, 256, 0, s>>>(wins, count, out, all_done, pc);
}

}  // namespace zb
