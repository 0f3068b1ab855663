// Batched local bundle adjustment (factor_graph::solve, factor_graph.hpp:489-615,
// on many independent windows).  The LM loop of each window runs in one
// workgroup: damping and point inverses, Schur complement, envelope Cholesky,
// triangular solves, binary64 refinement (mixed), update, cost, accept / reject
// and rollback, with workgroup barriers between the phases.  The reference
// linearises once per outer iteration (after an accepted step), so a window's
// workgroup pauses there, and the linearisations of all paused windows run as
// grid-wide launches over every window's points and poses at once, the
// HBM-bound phase spread over the whole GPU.  A solve is a fixed sequence of
// rounds (linearisation, then attempts), captured once in a hipGraph; windows
// never wait on each other inside a launch.  Each window uses the device buffers
// of its own graph handle; the controller is the shared lm_control.hpp code and
// the per-item arithmetic the shared mixed_items.hpp code, so every window's
// result equals its per-graph solve.
#include "ba_math.hpp"
#include "chol_device.hpp"
#include "common.hpp"
#include "lm_control.hpp"
#include "mixed_items.hpp"

namespace zb {
namespace {

constexpr int kBatchTiles = kBatchMaxN / wg::kTile;

// Hessian-block buffers of a window for storage type TS.
template <class TS> struct blocks;
template <> struct blocks<float> {
    float *pterms, *Hpp, *Hll, *W, *Hinv, *Y;
    __device__ explicit blocks(const dev_work64& w)
        : pterms(w.pose_terms_f), Hpp(w.Hpp_f), Hll(w.Hll_f), W(w.W_f), Hinv(w.Hinv_f), Y(w.Y_f) {}
};
template <> struct blocks<double> {
    double *pterms, *Hpp, *Hll, *W, *Hinv, *Y;
    __device__ explicit blocks(const dev_work64& w)
        : pterms(w.pose_terms), Hpp(w.Hpp), Hll(w.Hll), W(w.W), Hinv(w.Hinv), Y(w.Y) {}
};

// Window holding global item g: the k with prefix[k] <= g < prefix[k + 1].
__device__ __forceinline__ int find_window(const int32_t* prefix, int count, int g) {
    int lo = 0, hi = count;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ bool needs_linearisation(const lm_state* lm) { return !lm->done && lm->need_linearize; }

// Mixed windows linearise by blocks of observations (as the per-graph mixed path
// does); binary64 windows per point and pose chunk (as the per-graph binary64
// path sums).
template <class TS> constexpr bool kBlocks = false;
template <> constexpr bool kBlocks<float> = true;

// Grid-wide linearisation, three launches: per point (H_ll, b_l and W), per pose
// chunk (the pose terms formed again from the observations and summed in a fixed
// order, so they never go through memory), per pose-block entry.
template <class TS>
__global__ void __launch_bounds__(256) k_batch_lin_points(const batch_window* wins, int count, const int32_t* prefix,
                                                          int total, uint64_t* lin_clock, int round) {
    if (threadIdx.x == 0) atomicMin(reinterpret_cast<unsigned long long*>(&lin_clock[2 * round]),
                                    static_cast<unsigned long long>(wall_clock64()));
    const int g = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= total) return;
    const int k = find_window(prefix, count, g);
    const batch_window& bw = wins[k];
    if ((kBlocks<TS> && bw.p.n_lblk > 0) || !needs_linearisation(bw.w.lm)) return;
    const blocks<TS> B(bw.w);
    mixed::lin_point<TS, false>(bw.p, bw.w.pose, bw.w.point, B.pterms, bw.w.pose_grad, B.Hll, bw.w.bl_m, B.W,
                                bw.w.frozen, g - prefix[k]);
}

template <class TS>
__global__ void __launch_bounds__(256) k_batch_lin_chunks(const batch_window* wins, int count, const int32_t* prefix,
                                                          int total) {
    __shared__ TS ht[4][64][21];
    __shared__ double gt[4][64][6];
    const int wave = threadIdx.x >> 6;
    const int g = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);  // one wave per chunk
    if (g >= total) return;
    const int k = find_window(prefix, count, g);
    const batch_window& bw = wins[k];
    if ((kBlocks<TS> && bw.p.n_lblk > 0) || !needs_linearisation(bw.w.lm)) return;
    const int ic = g - prefix[k];
    mixed::reduce_pose_chunk_recompute(bw.p, bw.w.pose, bw.w.point, bw.w.pose_part, ic / kPoseChunks,
                                       ic % kPoseChunks, threadIdx.x & 63, ht[wave], gt[wave]);
}

template <class TS>
__global__ void __launch_bounds__(256) k_batch_lin_entries(const batch_window* wins, int count, const int32_t* prefix,
                                                           int total) {
    const int g = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= total) return;
    const int k = find_window(prefix, count, g);
    const batch_window& bw = wins[k];
    if (!needs_linearisation(bw.w.lm)) return;
    const blocks<TS> B(bw.w);
    const int ij = g - prefix[k];
    if (kBlocks<TS> && bw.p.n_lblk > 0) return;  // k_batch_lin_finish
    mixed::reduce_pose_finish(bw.w.pose_part, B.Hpp, bw.w.bp_m, ij / 27, ij % 27);
}

// The windows linearised by blocks: one workgroup per pose (global pose g, window
// k with entries[k] = 27 x poses before it).
template <class TS>
__global__ void __launch_bounds__(mixed::kFinishThreads) k_batch_lin_finish(const batch_window* wins, int count,
                                                                           const int32_t* entries, int total) {
    const int g = static_cast<int>(blockIdx.x);
    if (g >= total) return;
    int lo = 0, hi = count;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (entries[mid] <= 27 * g) lo = mid;
        else hi = mid;
    }
    const batch_window& bw = wins[lo];
    if (!(kBlocks<TS> && bw.p.n_lblk > 0) || !needs_linearisation(bw.w.lm)) return;
    __shared__ double red[4][27];
    const blocks<TS> B(bw.w);
    mixed::lin_pose_finish(bw.p, static_cast<const TS*>(bw.w.lin_part_h), bw.w.lin_part_g, B.Hpp, bw.w.bp_m,
                           g - entries[lo] / 27, red);
}

// The windows linearised by blocks of observations: one workgroup per block.
template <class TS>
__global__ void __launch_bounds__(kLinBlock) k_batch_lin_blocks(const batch_window* wins, int count,
                                                                const int32_t* prefix, int total, uint64_t* lin_clock,
                                                                int round) {
    // the first workgroup dispatched stamps the round's start (one atomic, not one
    // per workgroup of a grid of tens of thousands)
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicMin(reinterpret_cast<unsigned long long*>(&lin_clock[2 * round]),
                  static_cast<unsigned long long>(wall_clock64()));
    const int g = static_cast<int>(blockIdx.x);
    if (g >= total) return;
    const int k = find_window(prefix, count, g);
    const batch_window& bw = wins[k];
    if (!needs_linearisation(bw.w.lm)) return;
    __shared__ mixed::lin_stage<TS> sh;
    const blocks<TS> B(bw.w);
    mixed::lin_block(bw.p, bw.w.pose, bw.w.point, B.Hll, bw.w.bl_m, B.W, bw.w.frozen, static_cast<TS*>(bw.w.lin_part_h),
                     bw.w.lin_part_g, B.pterms, g - prefix[k], sh);
}

// The sum the per-graph launches form over n items (k_cost and k_backsub_update,
// each followed by k_sum_partials): a grid of G = min(ceil(n / 256), 1024)
// blocks of 256 threads, thread (b, t) adding items b 256 + t, + 256 G, ... in
// turn from 0, each block's 256 values added by a halving tree, and the G block
// sums, padded with zeros to 1024, by a halving tree.  One workgroup (a
// multiple of 256 threads) forms it here, blockDim / 256 blocks at a time, so
// that the batched solver's costs and LM scales, and with them its accept
// decisions, equal the per-graph solver's bit for bit.  item(i) is called once
// per item.  sh holds blockDim values, part 1024.
template <class F>
__device__ double grid_sum(int64_t n, F item, double* sh, double* part) {
    const int tid = threadIdx.x, t = tid & 255, per = blockDim.x >> 8;
    int64_t G = (n + 255) / 256;
    G = G < 1 ? 1 : (G > 1024 ? 1024 : G);
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

__device__ double cost(const dev_problem& p, const dev_work64& w, double* sh, double* part) {
    return grid_sum(p.no, [&](int64_t s) { return mixed::cost_item(p, w.pose, w.point, w.residual, static_cast<int>(s)); },
                    sh, part);
}

// Phase clocks (thread 0, after the phase's closing barrier), added per window
// to phase_ticks[window * kBatchPhases + phase] in 100 MHz ticks.
struct phase_clock {
    int64_t* out;
    int64_t last;
    int64_t acc[kBatchPhases];
    __device__ explicit phase_clock(int64_t* o) : out(o), last(wall_clock64()) {
        for (int i = 0; i < kBatchPhases; ++i) acc[i] = 0;
    }
    __device__ void mark(int phase) {
        const int64_t t = wall_clock64();
        acc[phase] += t - last;
        last = t;
    }
    __device__ void flush() {
        for (int i = 0; i < kBatchPhases; ++i) out[i] += acc[i];
    }
};

// The workgroup factorisation and solve, inlined (as calls they saved and
// restored the megakernel's registers through scratch memory, several times
// slower).  The factorisation loads its broadcasts 8 at a time: the same
// arithmetic as the per-graph kernel's 32 at a time, in fewer registers.
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

// Start of a solve: controller reset, phase clocks cleared, the initial cost, and
// the first linearisation requested.
__global__ void __launch_bounds__(256) k_batch_begin(const batch_window* wins, int64_t* phase_ticks,
                                                     uint64_t* lin_clock, int lin_slots) {
    __shared__ double sh[256];
    __shared__ double part[1024];
    const int tid = threadIdx.x;
    if (blockIdx.x == 0)
        for (int i = tid; i < lin_slots; i += blockDim.x) lin_clock[i] = ~0ULL;
    const batch_window& bw = wins[blockIdx.x];
    const dev_problem& p = bw.p;
    const dev_work64& w = bw.w;
    lm_state* lm = w.lm;
    int64_t* ticks = phase_ticks + static_cast<int64_t>(blockIdx.x) * kBatchPhases;
    if (tid == 0) {
        for (int i = 0; i < kBatchPhases; ++i) ticks[i] = 0;
        *lm = lm_state{};
        lm->iterations = bw.iterations;
        lm->use_relative = bw.use_relative;
        lm->need_linearize = 1;
    }
    phase_clock clk(ticks);
    __syncthreads();
    const double c = cost(p, w, sh, part);
    clk.mark(PH_COST);
    if (tid == 0) {
        lm->chi2 = c;
        clk.flush();
    }
}

// A window's attempts up to its next accepted step that asks for a
// linearisation (the workgroup pauses there; the next round resumes it) or to the
// end of its solve.  The first round starts the controller from the linearised
// system (max diagonal, lm_init).
template <class TS, int kThreads>
__global__ void __launch_bounds__(kThreads) k_batch_attempts(const batch_window* wins, lm_state* out,
                                                             int64_t* phase_ticks, uint64_t* lin_clock, int round) {
    if (threadIdx.x == 0) atomicMin(reinterpret_cast<unsigned long long*>(&lin_clock[2 * round + 1]),
                                    static_cast<unsigned long long>(wall_clock64()));
    const batch_window& bw = wins[blockIdx.x];
    const dev_problem& p = bw.p;
    const dev_work64& w = bw.w;
    const blocks<TS> B(w);
    lm_state* lm = w.lm;
    if (lm->done) return;
    __shared__ double sh[kThreads];
    __shared__ max_scan ms[kThreads];
    __shared__ double PT[kBatchTiles - 1][wg::kTile][wg::kPT];  // factor panels, then the solve's tiles
    static_assert((kBatchTiles - 1) * wg::kPT >= 6 * (wg::kTile + 1), "solve scratch");
    auto solve_buf = reinterpret_cast<double (*)[wg::kTile][wg::kTile + 1]>(&PT[0][0][0]);
    __shared__ double recips[wg::kTile];
    __shared__ double col[wg::kTile][64];
    __shared__ double xs[kBatchMaxN];
    __shared__ int flag;
    const int tid = threadIdx.x, lane = tid & 63, nwaves = blockDim.x >> 6;
    const int n = p.n;
    const bool mixed_path = sizeof(TS) == 4;
    phase_clock clk(phase_ticks + static_cast<int64_t>(blockIdx.x) * kBatchPhases);

    if (lm->lin_count == 0) {
        // max_diagonal over pose then point diagonal entries, in order
        const int64_t total = 6L * p.np + 3L * p.nl;
        const int64_t per = (total + blockDim.x - 1) / blockDim.x;
        const int64_t i0 = tid * per, i1 = (i0 + per < total) ? i0 + per : total;
        max_scan m = {0.0, 0, 0, 1};
        for (int64_t i = i0; i < i1; ++i) {
            const double h = (i < 6L * p.np)
                                 ? static_cast<double>(B.Hpp[21 * (i / 6) + (i % 6) * (i % 6 + 1) / 2 + i % 6])
                                 : static_cast<double>(B.Hll[9 * ((i - 6L * p.np) / 3) + 4 * ((i - 6L * p.np) % 3)]);
            m = scan_push(m, h);
        }
        ms[tid] = m;
        __syncthreads();
        if (tid == 0) {
            max_scan acc = {0.0, 0, 0, 1};
            for (unsigned t = 0; t < blockDim.x; ++t) acc = scan_combine(acc, ms[t]);
            lm_init(lm, scan_max_diagonal(acc));
        }
    } else if (tid == 0) {
        // the accepted attempt that paused for the linearisation ends here
        lm_linearise_done(lm);
        lm_end_attempt(lm);
    }
    __syncthreads();
    clk.mark(PH_UPDATE);
    const int max_attempts = 11 * bw.iterations + 1;
    for (int attempt = 0; attempt < max_attempts; ++attempt) {
        if (lm->done) break;
        const double lam = lm->lambda;
        // reduced camera system
        for (int l = tid; l < p.nl; l += blockDim.x) mixed::damp_invert_item(p, B.Hll, B.Hinv, w.frozen, lam, l);
        for (int64_t i = tid; i < static_cast<int64_t>(n) * n; i += blockDim.x) w.S[i] = 0.0;
        __syncthreads();
        for (int a = tid; a < p.npair; a += blockDim.x) mixed::form_Y_item(p, B.W, B.Hinv, B.Y, a);
        __syncthreads();
        clk.mark(PH_INVERT);
        for (int blk = tid >> 6; blk < p.nblk; blk += nwaves)
            mixed::schur_block<TS, 1>(p, B.Y, B.W, B.Hpp, w.bp_m, w.bl_m, w.S, w.bs, lam, blk, lane);
        __syncthreads();
        clk.mark(PH_SCHUR);
        // factorisation and solve
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
        clk.mark(PH_FACTOR);
        if (ok) ok = batch_solve<kThreads>(n, w.S, w.bs, xs, w.tile_first, w.tile_last, solve_buf, sh, &flag) == 0;
        clk.mark(PH_SOLVE);
        if (ok) {
            for (int i = tid; i < n; i += blockDim.x) w.dxp[i] = xs[i];
            __syncthreads();
            for (int l = tid; l < p.nl; l += blockDim.x) mixed::point_step_item(p, B.W, B.Hinv, w.bl_m, w.dxp, w.dxl, l);
            __syncthreads();
            clk.mark(PH_UPDATE);
            const int steps = mixed_path ? bw.refine_steps : 0;
            for (int k = 0; k < steps; ++k) {
                for (int l = tid; l < p.nl; l += blockDim.x)
                    mixed::full_residual_item(p, w.pose, w.point, w.dxp, w.dxl, w.bl_m, B.Hinv, w.pose_grad, w.zl, w.gl,
                                              lam, l);
                __syncthreads();
                for (int ic = tid >> 6; ic < p.np * kPoseChunks; ic += nwaves)
                    mixed::pose_residual_chunk(p, w.pose_grad, B.W, w.gl, w.pose_part, ic / kPoseChunks,
                                               ic % kPoseChunks, lane);
                __syncthreads();
                for (int ir = tid; ir < p.np * 6; ir += blockDim.x)
                    mixed::pose_residual_finish(p, w.pose_part, w.bp_m, w.dxp, w.rvec, lam, ir / 6, ir % 6);
                __syncthreads();
                if (batch_solve<kThreads>(n, w.S, w.rvec, xs, w.tile_first, w.tile_last, solve_buf, sh, &flag) != 0) {
                    ok = false;
                    break;
                }
                for (int l = tid; l < p.nl; l += blockDim.x) mixed::refine_point_item(p, B.W, B.Hinv, w.zl, xs, w.dxl, l);
                for (int i = tid; i < n; i += blockDim.x) w.dxp[i] += xs[i];
                __syncthreads();
            }
            clk.mark(PH_REFINE);
        }
        if (tid == 0) {
            lm->solve_ok = ok ? 1 : 0;
            lm_after_solve(lm);
        }
        __syncthreads();
        if (lm->step_applied) {
            // the factorisation's column buffer is free here: the 1024 block sums
            double* part = &col[0][0];
            const double scale = grid_sum(
                static_cast<int64_t>(p.np) + p.nl,
                [&](int64_t t) {
                    return mixed::apply_step_item(p, w.pose, w.pose_backup, w.point, w.point_backup, w.dxp, w.dxl, w.bp_m,
                                                  w.bl_m, lam, static_cast<int>(t));
                },
                sh, part);
            clk.mark(PH_UPDATE);
            const double c = cost(p, w, sh, part);
            clk.mark(PH_COST);
            if (tid == 0) {
                lm->scale = scale + 1e-3;  // factor_graph.hpp:556-560
                lm->chi2_new = c;
                lm_decide(lm);
            }
            __syncthreads();
            // Every thread takes the decision into registers; lm_end_attempt below
            // clears need_rollback only after the barrier that follows these reads.
            const bool rollback = lm->need_rollback != 0;
            const bool pause = !lm->done && lm->need_linearize;
            if (rollback) {
                for (int t = tid; t < p.np + p.nl; t += blockDim.x)
                    mixed::rollback_item(p, w.pose, w.pose_backup, w.point, w.point_backup, t);
            }
            if (pause) break;  // the next round linearises
        }
        __syncthreads();
        if (tid == 0) lm_end_attempt(lm);
        __syncthreads();
        clk.mark(PH_UPDATE);
    }
    __syncthreads();
    clk.mark(PH_UPDATE);
    if (tid == 0) {
        out[blockIdx.x] = *lm;
        clk.flush();
    }
}

}  // namespace


void launch_batch_solve(const batch_window* wins, int count, int threads, bool mixed_path, lm_state* out,
                        int64_t* phase_ticks, const batch_grid& grid, uint64_t* lin_clock, int rounds,
                        hipStream_t s) {
    if (count <= 0) return;
    auto blocks_for = [](int64_t items) { return static_cast<unsigned>(items > 0 ? (items + 255) / 256 : 1); };
    k_batch_begin<<<count, 256, 0, s>>>(wins, phase_ticks, lin_clock, 2 * rounds);
    for (int r = 0; r < rounds; ++r) {
        if (mixed_path) {
            if (grid.n_lblocks > 0)
                k_batch_lin_blocks<float><<<grid.n_lblocks, kLinBlock, 0, s>>>(wins, count, grid.lblocks,
                                                                             grid.n_lblocks, lin_clock, r);
            if (grid.n_per_point > 0) {
                k_batch_lin_points<float><<<blocks_for(grid.n_points), 256, 0, s>>>(wins, count, grid.points,
                                                                                  grid.n_points, lin_clock, r);
                k_batch_lin_chunks<float><<<blocks_for(64L * grid.n_chunks), 256, 0, s>>>(wins, count, grid.chunks,
                                                                                        grid.n_chunks);
                k_batch_lin_entries<float><<<blocks_for(grid.n_entries), 256, 0, s>>>(wins, count, grid.entries,
                                                                                    grid.n_entries);
            }
            if (grid.n_lblocks > 0 && grid.n_entries > 0)
                k_batch_lin_finish<float><<<grid.n_entries / 27, mixed::kFinishThreads, 0, s>>>(
                    wins, count, grid.entries, grid.n_entries / 27);
            if (threads == 256) k_batch_attempts<float, 256><<<count, 256, 0, s>>>(wins, out, phase_ticks, lin_clock, r);
            else if (threads == 512) k_batch_attempts<float, 512><<<count, 512, 0, s>>>(wins, out, phase_ticks, lin_clock, r);
            else k_batch_attempts<float, 1024><<<count, 1024, 0, s>>>(wins, out, phase_ticks, lin_clock, r);
        } else {
            k_batch_lin_points<double><<<blocks_for(grid.n_points), 256, 0, s>>>(wins, count, grid.points, grid.n_points,
                                                                               lin_clock, r);
            k_batch_lin_chunks<double><<<blocks_for(64L * grid.n_chunks), 256, 0, s>>>(wins, count, grid.chunks,
                                                                                     grid.n_chunks);
            k_batch_lin_entries<double><<<blocks_for(grid.n_entries), 256, 0, s>>>(wins, count, grid.entries,
                                                                                 grid.n_entries);
            if (threads == 256) k_batch_attempts<double, 256><<<count, 256, 0, s>>>(wins, out, phase_ticks, lin_clock, r);
            else if (threads == 512) k_batch_attempts<double, 512><<<count, 512, 0, s>>>(wins, out, phase_ticks, lin_clock, r);
            else k_batch_attempts<double, 1024><<<count, 1024, 0, s>>>(wins, out, phase_ticks, lin_clock, r);
        }
    }
}

}  // namespace zb
