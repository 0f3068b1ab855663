// Device data layout of one bundle-adjustment problem and the launchers the
// host engine calls.  Everything lives in HBM as structure-of-arrays.
//
// Observation order on the device ("sorted order", index s): grouped by point
// slot, and within a point by insertion order, so each point's track is a
// contiguous segment [lm_start[l], lm_start[l+1]).  obs_orig[s] is the
// observation's index in insertion order (the reference's edge order).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace zb {

// The workgroups of a launch are dealt to the 8 XCDs in turn (workgroup i to
// XCD i mod 8), and each XCD has its own L2.  This renumbering gives each XCD
// chunks of kChunk consecutive logical workgroups, the chunks dealt in turn, so
// neighbouring work items, which share operands, share an L2 while the chunks
// keep the XCDs evenly loaded.  It is a bijection of [0, n) (the tail that
// does not fill a round of chunks keeps its numbers): the same work is done,
// by other workgroups.
template <int kChunk = 16>
__device__ __forceinline__ int xcd_block(int bid, int n) {
    constexpr int kXcds = 8;
    const int full = n / (kXcds * kChunk) * (kXcds * kChunk);
    if (bid >= full) return bid;
    const int x = bid % kXcds, i = bid / kXcds;
    return ((i / kChunk) * kXcds + x) * kChunk + i % kChunk;
}

// Per-kernel-class device time from HIP events recorded around launches on the
// launching stream (enabled by zsba_options.timing).
struct kernel_timer {
    static constexpr int kClasses = 16;
    struct record {
        hipEvent_t start, stop;
        int id;
    };
    bool on = false;
    uint32_t mask = ~0u;  // classes timed (bit per class id)
    bool open = false;    // a begin() recorded and awaiting its end()
    std::vector<record> pending;
    std::vector<hipEvent_t> pool;
    double ms[kClasses] = {};
    int64_t launches[kClasses] = {};

    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    void begin(int id, hipStream_t s) {
        if (!on || !((mask >> id) & 1u)) return;
        record r{take(), take(), id};
        (void)hipEventRecord(r.start, s);
        pending.push_back(r);
        open = true;
    }
    void end(hipStream_t s) {
        if (!on || !open) return;
        (void)hipEventRecord(pending.back().stop, s);
        open = false;
    }
    // Call after the stream has been synchronised.
    void collect() {
        for (const record& r : pending) {
            float t = 0.0f;
            if (hipEventElapsedTime(&t, r.start, r.stop) == hipSuccess) {
                ms[r.id] += t;
                launches[r.id] += 1;
            }
            pool.push_back(r.start);
            pool.push_back(r.stop);
        }
        pending.clear();
    }
    void reset() {
        collect();
        for (int i = 0; i < kClasses; ++i) { ms[i] = 0.0; launches[i] = 0; }
    }
    ~kernel_timer() {
        for (const record& r : pending) { (void)hipEventDestroy(r.start); (void)hipEventDestroy(r.stop); }
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

enum kernel_class : int {
    K_LINEARISE = 0, K_POSE_REDUCE = 1, K_COST = 2, K_INVERT = 3, K_FORM_Y = 4, K_FILL = 5, K_SCHUR = 6,
    K_CHOLESKY = 7, K_TRSV = 8, K_BACKSUB = 9, K_CONTROL = 10, K_REFINE = 11
};

// The fields of one observation that the linearisation reads together,
// packed so that a thread walking a track (or gathering a pose's observations)
// touches one 32-byte record per observation instead of one line per array.
struct obs_record {
    int32_t pose;   // pose slot
    int32_t point;  // point slot
    int32_t param;  // camera index | (loss index << 16)
    int32_t pairf;  // 2 * W block + 1 if this is the block's first observation; -1 without a block
    double u, v;    // observed pixel
};

// One W block (6x3 pose-point coupling) exists per distinct (pose, point) pair
// whose pose and point are both free (factor_graph.hpp:669-709).
struct dev_problem {
    int32_t np = 0, nl = 0, no = 0;  // live poses, points, observations
    int32_t npair = 0;               // W blocks
    int32_t nblk = 0;                // lower-triangle 6x6 blocks of the reduced system
    int64_t ncontrib = 0;            // (pair a, pair b) products feeding those blocks
    int32_t n = 0;                   // 6 * np
    int32_t identity_info = 1;       // all information matrices are I
    // observations, sorted order
    const int32_t* obs_pose = nullptr;   // [no] pose slot
    const int32_t* obs_point = nullptr;  // [no] point slot
    const int32_t* obs_pair = nullptr;   // [no] W block of the observation, -1 if none
    const int32_t* obs_param = nullptr;  // [no] camera index | (loss index << 16)
    const int32_t* obs_orig = nullptr;   // [no] insertion-order index
    const double* obs_uv = nullptr;      // [no*2]
    const double* obs_info = nullptr;    // [no*4] (unused when identity_info)
    const obs_record* obs_rec = nullptr; // [no] the fields above, packed (sorted order)
    // segments
    const int32_t* lm_start = nullptr;   // [nl+1] observations of a point
    const int32_t* pose_start = nullptr; // [np+1] into pose_list
    const int32_t* pose_list = nullptr;  // [no] sorted indices of a pose's observations, insertion order
    const int32_t* obs_pose_pos = nullptr; // [no] position of observation s in pose_list (pose-major slot)
    const int32_t* lm_pair_start = nullptr; // [nl+1] W blocks of a point (first-occurrence order)
    const int32_t* pair_pose = nullptr;  // [npair]
    const int32_t* pair_point = nullptr; // [npair]
    const int32_t* pair_first = nullptr; // [npair] sorted index of the pair's first observation
    // reduced-system blocks: block b = (blk_row[b], blk_col[b]), row >= col,
    // fed by contributions [blk_start[b], blk_start[b+1]) of (pair a of row, pair c of col)
    const int32_t* blk_row = nullptr;
    const int32_t* blk_col = nullptr;
    const int64_t* blk_start = nullptr;
    const int2* contrib = nullptr;
    const int32_t* diag_blk = nullptr;   // [np] block index of (i, i)
    // Reference summation orders (exact binary64 mode), two sets: [0] for the
    // first linearisation of a solve, [1] for every later one.  The order is the
    // iteration order of the reference's hash-map containers.
    const int2* contrib_x[2] = {nullptr, nullptr};      // per block, ordered as the Y map
    const int32_t* pose_pair_start = nullptr;           // [np+1]
    const int32_t* pose_pairs = nullptr;                // [npair] W blocks grouped by pose, ascending
    const int32_t* pose_pairs_x[2] = {nullptr, nullptr};// W blocks of a pose, ordered as the Y map
    const int32_t* lm_pairs_x[2] = {nullptr, nullptr};  // W blocks of a point, ordered as the h_lp map
    // vertices
    const uint8_t* pose_fixed = nullptr;
    const uint8_t* point_fixed = nullptr;
    // Linearisation blocks (mixed path): runs of whole points, at most kLinBlock
    // observations each (n_lblk = 0 when some point has more: the per-point
    // kernels then linearise).  Entry e is one free pose a block touches: its
    // observations' offsets in the block lo_list[lo_start[e] .. lo_start[e+1]) in
    // order; block b's entries are [lp_start[b], lp_start[b+1]).  Entry e's partial
    // sums go to slot e_slot[e]; pose i's slots are [pe_start[i], pe_start[i+1]),
    // in block order.
    int32_t n_lblk = 0, n_lent = 0;
    const int32_t* lb_start = nullptr;   // [n_lblk+1] observation (sorted index) ranges
    const int32_t* lb_point = nullptr;   // [n_lblk+1] point slot ranges
    const int32_t* lp_start = nullptr;   // [n_lblk+1]
    const int32_t* lo_start = nullptr;   // [n_lent+1]
    const uint8_t* lo_list = nullptr;    // [observations of free poses]
    const int32_t* e_slot = nullptr;     // [n_lent]
    const int32_t* pe_start = nullptr;   // [np+1]
    const uint8_t* pair_multi = nullptr; // [npair] the W block has more than one observation
    // parameter tables
    const double* cams = nullptr;        // [ncam*4]
    const int32_t* loss_kind = nullptr;  // [nloss]
    const double* loss_delta = nullptr;  // [nloss]
};

// LM controller state, kept on the device so that attempts need no host round
// trip (factor_graph.hpp:511-611).
struct lm_state {
    double lambda, nu, chi2, last_chi2, chi2_new, scale, rho, initial_chi2, legacy_threshold;
    int32_t iterations, use_relative;
    int32_t iter, failures, success, attempts, rejected, solve_failures;
    int32_t done, need_linearize, need_rollback, solve_ok, step_applied, executed_iterations;
    int32_t lin_count;       // linearisation sequences ended (the per-graph path ends one per attempt)
    int32_t linearisations;  // linearisations performed in this solve (make_hessian calls)
    // 1 on the reassociating paths: a step is accepted only if it lowers the
    // cost, besides rho > 0 (the two agree whenever the step solves the damped
    // system well enough that the predicted decrease is positive)
    int32_t require_descent;
    int32_t pad;
};

// Working buffers of the binary64 path.
struct dev_work64 {
    double* pose = nullptr;         // [np*7] current parameters
    double* pose_backup = nullptr;
    double* point = nullptr;        // [nl*3]
    double* point_backup = nullptr;
    double* residual = nullptr;     // [no*2] residuals of the last evaluation, sorted order
    double* pose_terms = nullptr;   // [no*27] per-observation H_pp lower (21) + v_p (6), pose-major order
    double* Hpp = nullptr;          // [np*21]
    double* bp = nullptr;           // [np*6]
    double* Hll = nullptr;          // [nl*9]
    double* bl = nullptr;           // [nl*3]
    double* W = nullptr;            // [npair*18]
    double* Hinv = nullptr;         // [nl*9]
    double* Y = nullptr;            // [npair*18]
    double* S = nullptr;            // [n*n] reduced system, factorised in place
    double* bs = nullptr;           // [n]
    double* dxp = nullptr;          // [n]
    double* dxl = nullptr;          // [nl*3]
    double* partial = nullptr;      // reduction partials
    double* cbuf = nullptr;         // [no] per-observation robust chi2, insertion order (exact sums)
    double* sbuf = nullptr;         // [6np + 3nl] per-slot LM scale terms (exact sums)
    double* Lt = nullptr;           // [n*n + n] transposed factor (exact backward substitution)
    double* Lband = nullptr;        // [band_buffer_doubles] band factor (kernels_band.hip), zeroed at allocation
    double* chol_scale = nullptr;   // [4] Cholesky epsilon scales
    int32_t* chol_status = nullptr; // [0] 0 ok, 1 decompose failed, 2 solve failed
    lm_state* lm = nullptr;
    kernel_timer* timer = nullptr;  // host object; null = no timing
    // mixed path: binary32 Hessian blocks, binary64 gradient
    float* pose_terms_f = nullptr;  // [no*21] per-observation H_pp terms, pose-major
    double* pose_grad = nullptr;    // [no*6] per-observation b_p terms, pose-major
    float* Hpp_f = nullptr;         // [np*21]
    double* bp_m = nullptr;         // [np*6]
    float* Hll_f = nullptr;         // [nl*9]
    double* bl_m = nullptr;         // [nl*3]
    float* W_f = nullptr;           // [npair*18]
    float* Hinv_f = nullptr;        // [nl*9]
    float* Y_f = nullptr;           // [npair*18]
    // adjoint (conjugate gradients): 6 vectors of 6 np + 3 nl, reduction scratch
    double* adj = nullptr;
    double* adj_partial = nullptr;  // [513]
    double* adj_obs = nullptr;      // [2 no] dL/duv by compacted observation
    uint8_t* adj_frozen = nullptr;  // [nl] point held constant by the adjoint
    // Behind-camera freeze (not the reference's rule; off on the reference-order
    // path): null, or [nl] flags set by each linearisation for the points with an
    // observation the camera cannot project (camera.hpp:98-100).  Until the next
    // linearisation such a point is held as a fixed one: zero H_ll^-1, so no Schur
    // terms and no step.
    uint8_t* frozen = nullptr;
    // envelope of the reduced system in 32 x 32 tiles: row tile I is zero left of
    // tile column tile_first[I]; column tile k is zero below row tile tile_last[k]
    const int32_t* tile_first = nullptr;  // [T]
    const int32_t* tile_last = nullptr;   // [T]
    int32_t panel_tiles = 0;              // max over k of tile_last[k] - k + 1
    int32_t factor_path = 0;              // 0 automatic, 1 one workgroup, 2 panel launches, 3 band
    // band of the reduced system: row r is zero left of column r - band + 1
    int32_t band = 0;
    // mixed-path refinement (pose_grad doubles as the per-observation slots)
    double* zl = nullptr;           // [nl*3] point rows r_l of the binary64 residual
    double* gl = nullptr;           // [nl*3] H_ll^-1 r_l
    double* rvec = nullptr;         // [n] reduced residual r_p - W H_ll^-1 r_l
    double* corr = nullptr;         // [n] pose correction
    double* pose_part = nullptr;    // [np*kPoseChunks*27] per-chunk partial sums of the per-pose reductions
    // per-entry partial sums of the linearisation blocks: 21 H_pp terms (storage
    // type of the blocks) and 6 gradient terms
    void* lin_part_h = nullptr;     // [n_lent*21]
    double* lin_part_g = nullptr;   // [n_lent*6]
};

#define ZB_TIME_BEGIN(w, id, s) do { if ((w).timer) (w).timer->begin((id), (s)); } while (0)
#define ZB_TIME_END(w, s) do { if ((w).timer) (w).timer->end((s)); } while (0)

constexpr int kPartials = 1024;  // reduction partial slots
// The mixed path's per-pose sums over a pose's observations (and pairs) are split
// into kPoseChunks contiguous chunks, one wave each, whose partial sums are then
// added in chunk order: a fixed association, independent of the launch shape.
constexpr int kPoseChunks = 16;
// Observations per linearisation block (one thread each).
constexpr int kLinBlock = 128;

// ---- binary64 kernels (kernels_fp64.hip) ----------------------------------
// `exact`: reproduce the reference's summation orders (sequential sums over all
// observations / slots, hash-map orders in the Schur products); otherwise use
// fixed-shape parallel reductions.
// decide: the accept / reject decision follows in the same launch as the sum
// (the LM attempt's sequence; a launch of its own on the exact path).
void launch_cost64(const dev_problem& p, const dev_work64& w, hipStream_t s, int target /*0 chi2, 1 chi2_new*/,
                   bool only_if_step, bool exact = false, bool decide = false);
// exact: per-pose sums in insertion order (the reference's); otherwise in the
// fixed chunks of the mixed path.  end_attempt (gated runs): the attempt's
// controller reset follows the linearisation's bookkeeping in one launch.
void launch_linearise64(const dev_problem& p, const dev_work64& w, hipStream_t s, bool gated, bool exact = true,
                        bool end_attempt = false);
void launch_init_lm64(const dev_problem& p, const dev_work64& w, hipStream_t s);
void launch_reduced_system64(const dev_problem& p, const dev_work64& w, hipStream_t s, const double* lambda_override,
                             bool exact = false);
void launch_after_solve64(const dev_work64& w, hipStream_t s);
void launch_backsub_update64(const dev_problem& p, const dev_work64& w, hipStream_t s, bool exact = false);
void launch_decide64(const dev_work64& w, hipStream_t s);
void launch_rollback64(const dev_problem& p, const dev_work64& w, hipStream_t s);
void launch_eval_edges64(const dev_problem& p, const dev_work64& w, double* out, hipStream_t s);
void launch_inliers64(const dev_problem& p, const dev_work64& w, uint8_t* flags_sorted, hipStream_t s);
// map::cull's reprojection test per point slot (mean error, erase flag).
void launch_reprojection_cull(const dev_problem& p, const dev_work64& w, float threshold, float* mean, uint8_t* erase,
                              hipStream_t s);
void launch_backsub_only64(const dev_problem& p, const dev_work64& w, hipStream_t s);
// Point step dx_l for the solved dx_p and the update of every free vertex (no LM gating).
void launch_gn_apply64(const dev_problem& p, const dev_work64& w, hipStream_t s);

// ---- mixed path: binary32 linearisation and Schur (kernels_fp32.hip) -----------
void launch_linearise_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s, bool gated,
                            bool end_attempt = false);
void launch_reduced_system_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s, const double* lambda_override);
// ---- mixed path: binary64 arithmetic over the binary32 blocks (kernels_fp64.hip)
void launch_init_lm_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s);
// dx_l for the solved dx_p, then `steps` rounds of refinement of (dx_p, dx_l)
// against the binary64 normal equations; the factor of S must be in w.S.
void launch_refine_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s, int steps, const double* lambda_override);
void launch_backsub_update_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s);
void launch_backsub_only_mixed(const dev_problem& p, const dev_work64& w, hipStream_t s);

// ---- adjoint of the converged solve (kernels_adjoint.hip) ----------------------
// Vectors are [poses 6 np | points 3 nl].  out = d/d eps g(x + eps v, K + eps dK)
// (v or dK may be null): H v, or dg/dK along dK.
void launch_grad_tangent(const dev_problem& p, const dev_work64& w, const double* v, const double* dK, double* out,
                         hipStream_t s);
// z = M^-1 r with the Gauss-Newton reduced-system factor in w.S and H_ll^-1 in w.Hinv.
void launch_adjoint_precond(const dev_problem& p, const dev_work64& w, const double* r, double* z, hipStream_t s);
// Marks the points the adjoint holds constant (fixed, or singular undamped block:
// w.Hinv of the undamped linearisation is zero) and zeroes their rows of v.
void launch_adjoint_frozen(const dev_problem& p, const dev_work64& w, hipStream_t s);
void launch_zero_frozen_rows(const dev_problem& p, const dev_work64& w, double* v, hipStream_t s);
void launch_dot(const double* a, const double* b, int64_t n, double* partial /*512*/, double* out, hipStream_t s);
void launch_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s);
// dL/duv per observation, indexed by compacted observation e (2 per observation).
void launch_obs_grad(const dev_problem& p, const dev_work64& w, const double* lam, double* dobs_by_e, hipStream_t s);

// ---- batched local BA (kernels_batch.hip) --------------------------------------
// Windows solve in lockstep: each step every unfinished window linearises (when
// its last attempt was accepted) and makes one LM attempt, each phase one
// grid-wide launch over every window's items; the factorisation and the
// substitutions run one workgroup per window.  Windows are limited to reduced
// systems of kBatchMaxN rows (42 poses).
constexpr int kBatchMaxN = 256;
struct batch_window {
    dev_problem p;
    dev_work64 w;
    int32_t iterations = 0, use_relative = 0, refine_steps = 0, require_descent = 1;
    double floor_rel = 0.0;
};
// Phases of the batched solver's device clock.
enum batch_phase : int {
    PH_COST = 0, PH_LINEARISE = 1, PH_INVERT = 2, PH_SCHUR = 3, PH_FACTOR = 4, PH_SOLVE = 5, PH_REFINE = 6, PH_UPDATE = 7
};
constexpr int kBatchPhases = 8;
// Workgroup tables of the batched launches.  Table k holds, per workgroup of its
// launch, the window (-1: none) and the workgroup's index inside the window's
// share: BM_WIN one per window; BM_LBLK one per linearisation block and BM_LPOSE
// one per pose (mixed windows with linearisation blocks: their linearisation and
// refinement); BM_LPTS 256 points, BM_LCHUNK 4 pose chunks, BM_LENTRY 256
// pose-block entries (the other windows); BM_PTS 256 points; BM_INVZ 256 of the
// nl + n^2 items of the damping and the cleared reduced system; BM_PAIR 256
// (pose, point) pairs; BM_SBLK one reduced-system block; BM_PTSN 256 of nl + n;
// BM_APPLY and BM_COST one per block of the per-graph grids of the update (np + nl
// items) and of the cost (no items); BM_SLOT 256 of np + nl.
enum batch_map : int {
    BM_WIN = 0, BM_LBLK, BM_LPOSE, BM_LPTS, BM_LCHUNK, BM_LENTRY, BM_PTS, BM_INVZ, BM_PAIR, BM_SBLK, BM_PTSN, BM_APPLY,
    BM_COST, BM_SLOT, kBatchMaps
};
struct batch_grid {
    const int2* map[kBatchMaps] = {};
    int32_t n_wg[kBatchMaps] = {};
};
// k_batch_begin: every window's controller reset, initial cost, first
// linearisation requested; the phase clock pc (2 + kBatchPhases values) cleared.
void launch_batch_begin(const batch_window* wins, int count, uint64_t* pc, hipStream_t s);
// `steps` lockstep steps.  threads: the factorisation and solve workgroup size
// (256, 512 or 1024); inflight: Schur contributions in flight per lane (1 or 2);
// refine_rounds: the largest refine_steps of the windows (mixed).
void launch_batch_steps(const batch_window* wins, const batch_grid& g, int threads, int inflight, bool mixed_path,
                        int refine_rounds, int steps, uint64_t* pc, hipStream_t s);
// Every controller to out[i]; *all_done = 1 when every window has finished.
void launch_batch_finish(const batch_window* wins, int count, lm_state* out, int32_t* all_done, uint64_t* pc,
                         hipStream_t s);

// ---- dense Cholesky of the reduced system (kernels_chol.hip) ------------------
// Factorises w.S (n x n, lower triangle) in place and solves for w.dxp from
// w.bs.  Sets *w.chol_status and lm->solve_ok; skips everything when the
// controller is done (gated) or a previous stage failed.
void launch_cholesky_solve64(int32_t n, const dev_work64& w, hipStream_t s, bool gated, bool exact = false,
                             double floor_rel = 0.0);
// Forward + backward substitution with the factor in w.S: x = S^-1 b.
void launch_cholesky_resolve64(int32_t n, const dev_work64& w, const double* b, double* x, hipStream_t s, bool gated);

// ---- banded factorisation and substitutions (kernels_band.hip) -----------------
// For the reassociating precisions when the reduced system's band (w.band) is at
// most kBandMaxW columns: one workgroup factorises w.S into the band factor in
// w.Lt (by columns, by rows, and the reciprocal pivots) and solves for w.dxp
// from w.bs (same status and solve_ok rules).  Factor path 3 forces it; the
// automatic choice takes it when the system has at least kBandMinRowsPerWindow
// windows of rows (a narrow band, not a dense matrix).
constexpr int kBandPanel = 8;
// The substitutions hold 128 unknowns and refill the 16 slots a block of steps
// frees at the block's end, so a factor row may reach at most 128 - 15 columns.
constexpr int kBandMaxW = 113;
constexpr int kBandMinRowsPerWindow = 3;
// Window (a multiple of 16, 64 .. 128) for a band, 0 if the band is too wide.
int band_window(int32_t band);
// Doubles of w.Lt the band factor needs (0 if the band is too wide).
size_t band_buffer_doubles(int32_t n, int32_t band);
bool band_path(int32_t n, const dev_work64& w, bool exact);
void launch_chol_band(int32_t n, const dev_work64& w, hipStream_t s, bool gated, double floor_rel);
void launch_band_solve(int32_t n, const dev_work64& w, const double* b, double* x, hipStream_t s, bool gated);

// ---- graphs with every point fixed (the tracking step's motion-only BA) ---------
// The whole FP64 (reference-order) solve in one launch of one workgroup, for at
// most kMotionMaxPoses poses (kernels_fp64.hip).  w.lm must hold iterations and
// use_relative with every other field zero.
constexpr int kMotionMaxPoses = 64;
void launch_motion_solve(const dev_problem& p, const dev_work64& w, hipStream_t s);

// ---- diagnostics ---------------------------------------------------------------
void launch_math_probe(int fn, int n, const double* in, double* out, hipStream_t s);

}  // namespace zb
