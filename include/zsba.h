/*
 * zsba.h — C ABI of the MI355X bundle-adjustment engine (libzsba.so).
 *
 * The engine is a drop-in for the Levenberg–Marquardt / Schur bundle-adjustment
 * optimiser of ZeroSLAM's factor graph:
 *   factor_graph::factor_graph::{add_vertex, add_edge, remove_vertex, remove_edge,
 *   solve, get_current_chi}            (reference source/factor_graph.hpp:385-629)
 * specialised, as the reference's solver itself is, to SE(3) pose vertices
 * (vertex_pose, factor_graph.hpp:938-961), 3-D point vertices (vertex_point_xyz,
 * factor_graph.hpp:963-969, always marginalised) and 2-D pinhole reprojection
 * edges (edge_reprojection<camera::pinhole>, factor_graph.hpp:973-1049).
 *
 * Conventions
 *   - Pose parameters are the reference layout [tx, ty, tz, qx, qy, qz, qw] of a
 *     world-to-camera transform (factor_graph.hpp:941, 953-959).
 *   - Ids are insertion indices (0, 1, 2, ...) per kind and are never reused.
 *     Slot order in the solve is insertion order of live items, as
 *     factor_graph::set_ordering (factor_graph.hpp:633-645) assigns it.
 *   - Observation order is insertion order (factor_graph.hpp:444-454, 664).
 *   - No function throws or aborts.  Functions returning int32_t return a value
 *     >= 0 on success and a negative ZSBA_E* code on failure; zsba_last_error()
 *     then holds a message.  Handles are not thread-safe: use one per host thread.
 *   - All input arrays are host memory and are copied; the caller keeps ownership.
 */
#ifndef ZSBA_H
#define ZSBA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZSBA_ABI_VERSION 2

/* Error codes (negative returns). */
#define ZSBA_EINVAL   (-1) /* bad argument (null pointer, id out of range, NaN size ...) */
#define ZSBA_ESTATE   (-2) /* call not valid in the handle's current state            */
#define ZSBA_EDEVICE  (-3) /* HIP runtime error (no device, launch failure, ...)      */
#define ZSBA_ENOMEM   (-4) /* device or host allocation failed                         */
#define ZSBA_EUNSUP   (-5) /* a graph shape the engine does not support                */

/* Robust loss kinds: factor_graph.hpp:851-934.  ZSBA_LOSS_NONE is an edge with no
 * loss object (loss_function == nullptr, factor_graph.hpp:222-224, 255-258). */
#define ZSBA_LOSS_NONE    0
#define ZSBA_LOSS_TRIVIAL 1
#define ZSBA_LOSS_HUBER   2
#define ZSBA_LOSS_CAUCHY  3
#define ZSBA_LOSS_TUKEY   4

/* Arithmetic of a solve.
 *   FP64:      binary64 in the reference's operation order throughout: per
 *              observation, per vertex sum, in the iteration order of the
 *              reference's hash maps for the Schur sums, sequential cost and
 *              scale sums, and a Cholesky with the row-Crout update order.  The
 *              result is the reference optimiser's, bit for bit (the C library's
 *              sin/cos/pow are matched by correctly rounded evaluations).
 *   MIXED:     residual/Jacobian/Hessian-block/Schur work in binary32 with binary32
 *              accumulation; the LM cost, its comparison, the reduced-camera
 *              Cholesky and iterative refinement of the reduced solve stay binary64.
 *   FP64_FAST: binary64 with parallel (fixed-shape, deterministic) reductions;
 *              equal to FP64 up to rounding of the sums. */
#define ZSBA_PRECISION_FP64      0
#define ZSBA_PRECISION_MIXED     1
#define ZSBA_PRECISION_FP64_FAST 2

typedef struct zsba_options {
    int32_t device;       /* HIP device ordinal (default 0)                          */
    int32_t verbose;      /* 1: print one summary line per solve to stderr (iterations,
                             attempts, accepted, rejected, failed solves, chi2 before
                             and after, final damping); the reference's per-attempt
                             [INIT]/[GOOD]/[ BAD] trace is not reproduced            */
    int32_t refine_steps; /* MIXED: refinement steps of the reduced solve: 0 = default
                             (1), negative = none                                      */
    int32_t timing;       /* 1: time kernel launches with HIP events (stats)          */
    int32_t floor_exp;    /* MIXED: the binary32-block reduced system is factorised
                             with its diagonal raised by 2^-floor_exp max|S_ii| (its
                             rounding floor); the binary64 refinement uses the true
                             damping.  0 = default, negative = no floor              */
    int32_t factor_path;  /* reduced-system factorisation: 0 = automatic, 1 = one
                             workgroup (envelope panels up to 16 tiles of 32 rows),
                             2 = one launch per 32-column panel, 3 = banded (FP64
                             fast and MIXED when every row of the reduced system
                             spans at most 120 columns; otherwise automatic)      */
    int32_t timing_classes; /* with timing: bit mask of the kernel classes timed
                             (bit ZSBA_K_*); 0 = every class                        */
    int32_t freeze_behind;  /* behind-camera freeze, not a rule of the reference: at
                             each linearisation, a point with an observation whose
                             camera cannot project it (camera.hpp:98-100, the
                             edge's behind-camera residual -0.3 z) is held fixed
                             until the next linearisation (zero inverse block, no
                             step).  Such points otherwise sit at the z = 0 edge,
                             where the residual jumps, and keep the shared damping
                             from falling (the solve stalls).  0 = default: on for
                             MIXED, off for FP64_FAST; 1 = on for MIXED and
                             FP64_FAST; -1 = off.  Never applied to FP64.        */
    int32_t attempt_graph;  /* 0 = default: each LM attempt's launch sequence is
                             captured once per structure and precision as a HIP
                             graph and replayed (one graph launch per attempt);
                             -1 = launch the kernels one by one.  Timing (above)
                             always launches one by one.                        */
} zsba_options;

/* Kernel classes timed when zsba_options.timing is set (zsba_stats.kernel_*). */
#define ZSBA_K_LINEARISE   0  /* per-point linearisation: residual, Jacobians, H_ll, b_l, W */
#define ZSBA_K_POSE_REDUCE 1  /* per-pose H_pp / b_p sums                                   */
#define ZSBA_K_COST        2  /* residuals + robust chi2 (+ its reduction)                  */
#define ZSBA_K_INVERT      3  /* damping + 3x3 point-block inverses                         */
#define ZSBA_K_FORM_Y      4  /* Y = W H_ll^-1                                              */
#define ZSBA_K_FILL        5  /* clearing the reduced system                                */
#define ZSBA_K_SCHUR       6  /* reduced camera system blocks and b_s                       */
#define ZSBA_K_CHOLESKY    7  /* dense factorisation of the reduced system                  */
#define ZSBA_K_TRSV        8  /* forward / backward substitution                            */
#define ZSBA_K_BACKSUB     9  /* point back-substitution, state update, LM scale            */
#define ZSBA_K_CONTROL    10  /* LM controller, rollback and bookkeeping                    */
#define ZSBA_K_REFINE     11  /* mixed path: binary64 residual of the reduced system        */
#define ZSBA_K_COUNT      16

/* Per-solve statistics (filled by zsba_get_stats after zsba_solve). */
typedef struct zsba_stats {
    int32_t iterations;      /* outer LM iterations executed (factor_graph.hpp:538)   */
    int32_t attempts;        /* linear solves attempted                               */
    int32_t accepted;        /* accepted steps (the value zsba_solve returns)         */
    int32_t rejected;        /* steps evaluated and rejected                          */
    int32_t solve_failures;  /* attempts whose Cholesky decompose/solve failed        */
    int32_t precision;       /* precision of the last solve                           */
    int64_t n_obs, n_poses, n_points; /* live sizes of the last solve                 */
    double initial_chi2;     /* chi2 at the start of the last solve                   */
    double final_chi2;       /* chi2 after the last accepted step                     */
    double lambda;           /* damping at exit                                       */
    double solve_ms;         /* wall time of the numeric solve, device-synchronised   */
    double setup_ms;         /* upload + structure analysis (0 when cached)           */
    double kernel_ms[ZSBA_K_COUNT];        /* summed device time per kernel class     */
    int64_t kernel_launches[ZSBA_K_COUNT]; /* launches per kernel class               */
    int32_t linearisations;  /* linearisations (make_hessian) in the last solve          */
    int32_t frozen_points;   /* points held by the behind-camera freeze at the last
                                linearisation (0 when the freeze is off)               */
    double reserved[7];
} zsba_stats;

typedef struct zsba_graph zsba_graph;

/* ---- lifetime ---------------------------------------------------------------- */
int32_t     zsba_abi_version(void);
zsba_graph* zsba_create(const zsba_options* options /* NULL = defaults */);
void        zsba_destroy(zsba_graph* graph);
const char* zsba_last_error(const zsba_graph* graph);

/* ---- graph construction --------------------------------------------------------
 * zsba_add_pose / zsba_add_point replace factor_graph::add_vertex
 * (factor_graph.hpp:385-402) for vertex_pose (not marginalised) and
 * vertex_point_xyz (marginalised, map.hpp:158-162).  They return the new id. */
int32_t zsba_add_pose(zsba_graph* graph, const double params[7], int32_t fixed);
int32_t zsba_add_point(zsba_graph* graph, const double xyz[3], int32_t fixed);
int32_t zsba_add_poses(zsba_graph* graph, int32_t count, const double* params /* count*7 */,
                       const uint8_t* fixed /* count, NULL = all free */);
int32_t zsba_add_points(zsba_graph* graph, int32_t count, const double* xyz /* count*3 */,
                        const uint8_t* fixed /* count, NULL = all free */);

/* zsba_add_observation replaces constructing edge_reprojection<pinhole>(camera)
 * with set_observation / set_information / set_loss_function / add_vertex(pose),
 * add_vertex(point) and factor_graph::add_edge (factor_graph.hpp:198-276, 444-454,
 * 984-987).  info = row-major 2x2 information (NULL = identity,
 * factor_graph.hpp:181); K = (fx, fy, cx, cy) (camera.hpp:71-80).  Returns the id. */
int32_t zsba_add_observation(zsba_graph* graph, int32_t pose_id, int32_t point_id,
                             const double uv[2], const double info[4] /* NULL = I */,
                             const double K[4], int32_t loss_kind, double loss_delta);
/* Bulk form with one camera and one loss for all observations.  Returns first id. */
int32_t zsba_add_observations(zsba_graph* graph, int32_t count, const int32_t* pose_ids,
                              const int32_t* point_ids, const double* uv /* count*2 */,
                              const double* info /* count*4, NULL = identity */,
                              const double K[4], int32_t loss_kind, double loss_delta);

/* Removal: factor_graph::remove_edge / remove_vertex (factor_graph.hpp:405-470).
 * Removing a vertex removes its observations.  Return 0, or ZSBA_EINVAL if the id
 * is unknown or already removed (the reference returns false). */
int32_t zsba_remove_observation(zsba_graph* graph, int32_t obs_id);
int32_t zsba_remove_pose(zsba_graph* graph, int32_t pose_id);
int32_t zsba_remove_point(zsba_graph* graph, int32_t point_id);

/* Parameter access (vertex_base::set_parameters / get_parameters / set_fixed,
 * factor_graph.hpp:107-129). */
int32_t zsba_set_pose(zsba_graph* graph, int32_t pose_id, const double params[7]);
int32_t zsba_set_point(zsba_graph* graph, int32_t point_id, const double xyz[3]);
int32_t zsba_set_pose_fixed(zsba_graph* graph, int32_t pose_id, int32_t fixed);
int32_t zsba_set_point_fixed(zsba_graph* graph, int32_t point_id, int32_t fixed);
/* counts[0..2] = poses, points, observations ever added (array sizes for getters). */
int32_t zsba_counts(const zsba_graph* graph, int64_t counts[3]);
/* Copy current parameters, indexed by id (removed ids are left untouched). */
int32_t zsba_get_poses(const zsba_graph* graph, double* params /* n_added*7 */);
int32_t zsba_get_points(const zsba_graph* graph, double* xyz /* n_added*3 */);

/* ---- optimisation -------------------------------------------------------------
 * zsba_solve replaces factor_graph::solve(iterations, use_relative_convergence)
 * (factor_graph.hpp:489-615) and returns the accepted-step count; 0 for a graph
 * without observations or vertices (factor_graph.hpp:494-499). */
int32_t zsba_solve(zsba_graph* graph, int32_t iterations, int32_t use_relative_convergence,
                   int32_t precision);
/* factor_graph::get_current_chi(recompute) (factor_graph.hpp:618-629).  With
 * recompute = 0 it sums the residuals stored by the last evaluation.  Returns NaN
 * on error. */
double  zsba_current_chi(zsba_graph* graph, int32_t recompute);
/* Inlier flag per observation id: 1 when the robust loss is in its quadratic
 * region (e2 <= delta^2 for Huber, factor_graph.hpp:873) at the current state,
 * 1 for every observation without a loss.  Removed ids are left untouched. */
int32_t zsba_get_inlier_flags(zsba_graph* graph, uint8_t* flags /* n_obs_added */);
/* The reprojection test of map::cull (map.hpp:287-312), the part that needs the
 * optimised state, for every live point at the current state: the mean over its
 * observations, in insertion order, of the reprojection error in pixels, each
 * rounded to binary32 and added in binary32, with the camera-frame point taken as
 * R X + t for R the matrix of the pose quaternion (the frame rotation map.hpp:243
 * sets) and `threshold` counted for a point the camera cannot project
 * (map.hpp:297-299; the reference uses 5.991 for both).  mean_error[id] receives
 * the mean (+inf for a point without observations), erase[id] 1 when the
 * reference erases the point on this test: no observation, or a mean of at least
 * `threshold` (map.hpp:305-312).  Removed ids are left untouched.  The id and
 * recency rules of map.hpp:271-285 stay with the caller.  Returns the number of
 * points flagged. */
int32_t zsba_reprojection_cull(zsba_graph* graph, float threshold, float* mean_error /* n_points_added */,
                               uint8_t* erase /* n_points_added */);
/* Per-observation residuals (obs - projection, factor_graph.hpp:990-1009) at the
 * current state, indexed by id.  */
int32_t zsba_get_residuals(zsba_graph* graph, double* residuals /* n_obs_added*2 */);
int32_t zsba_get_stats(const zsba_graph* graph, zsba_stats* stats);

/* ---- workload control (benchmarks) -------------------------------------------
 * zsba_snapshot copies the current device parameters aside; zsba_restore puts
 * them back without any host transfer, so repeated solves start from the same
 * state with their inputs already resident in device memory. */
int32_t zsba_snapshot(zsba_graph* graph);
int32_t zsba_restore(zsba_graph* graph);

/* ---- batched local BA (no reference counterpart: many independent graphs) -----
 * A batch solves many independent graphs (local BA windows) with the same
 * iteration count, convergence rule and precision.  Every window whose reduced
 * system has at most 256 rows (42 poses) runs its LM attempts in one workgroup;
 * the windows' linearisations (once per accepted step, factor_graph.hpp:598-611)
 * run between them as grid-wide launches over every window.  The sequence of
 * launches is captured in a hipGraph and replayed by later zsba_batch_solve
 * calls.  The attempts' workgroup size is chosen by timing the candidates on the
 * first solve of each batch shape.
 * Larger windows, and ZSBA_PRECISION_FP64 (the reference's operation order), are
 * solved one by one with zsba_solve.  accepted[i] receives the accepted-step
 * count of graphs[i]; per-graph zsba_get_stats reflect the batch solve.  The
 * graphs must all use one device and stay alive for the life of the batch. */
typedef struct zsba_batch zsba_batch;
typedef struct zsba_batch_info {
    int32_t windows;          /* graphs solved by the batched launches               */
    int32_t individual;       /* graphs solved one by one                            */
    int32_t threads;          /* workgroup size of the factorisation and solves      */
    int32_t graph_launches;   /* hipGraph launches by this solve                     */
    float tune_ms[3];         /* autotuner: fastest trial with 256, 512, 1024 threads
                                 (0 = not tried)                                     */
    float pad;
    double autotune_wall_ms;  /* time spent autotuning in this solve (0 when cached) */
    double kernel_wall_ms;    /* batched launches, host-measured, synchronised       */
    double solve_ms;          /* whole zsba_batch_solve call                         */
    double phase_ms[8];       /* device time of the batched launches by phase (all
                                 windows together): cost, linearisation, damping +
                                 inverses + Y, Schur, factorisation, refinement
                                 solves, refinement, update + controller             */
    double lin_ms;            /* phase_ms[1]: the linearisation launches             */
    int32_t steps;            /* lockstep steps launched (one attempt per unfinished
                                 window each)                                        */
    int32_t schur_inflight;   /* Schur contributions in flight per lane (autotuned)  */
    double reserved[2];
} zsba_batch_info;
zsba_batch* zsba_batch_create(zsba_graph* const* graphs, int32_t count);
void        zsba_batch_destroy(zsba_batch* batch);
const char* zsba_batch_last_error(const zsba_batch* batch);
int32_t     zsba_batch_solve(zsba_batch* batch, int32_t iterations, int32_t use_relative_convergence,
                             int32_t precision, int32_t* accepted /* count */);
int32_t     zsba_batch_get_info(const zsba_batch* batch, zsba_batch_info* info);
/* One-shot form: create, solve, destroy. */
int32_t zsba_solve_batch(zsba_graph* const* graphs, int32_t count, int32_t iterations,
                         int32_t use_relative_convergence, int32_t precision, int32_t* accepted);

/* ---- data association by projection (slam.hpp:341-384) ------------------------
 * A new frame's matched keypoints against the map's landmarks: the loop of
 * slam.hpp:341-384 with feature::distance<256> (feature.hpp:1134-1188).  Inputs,
 * in the reference's terms:
 *   frame:              frame_current.rotation (row-major) and .translation, the
 *                       pinhole intrinsics, image_grey's cols and rows;
 *   landmark_xyz[o]:    the o-th landmark in the map's iteration order
 *                       (reconstruction.landmarks), 3 doubles;
 *   landmark_descriptor[o]: the descriptor of its first observation, 32 bytes
 *                       (slam.hpp:374);
 *   landmark_observed[o]: 1 when frame_current already observes it (slam.hpp:352-362);
 *   keypoint_xy[i]:     match_point_current[i];
 *   keypoint_descriptor[i]: descriptor of match_index_current[i], 32 bytes;
 *   keypoint_key[i]:    match_index_previous[i], in [0, n_keys);
 *   key_taken[k]:       1 when frame_previous_points holds k before the loop.
 * The matches are written in the order the reference calls add_observation:
 * (match_landmark[j], match_keypoint[j]) = (o, i); the caller adds the
 * observation of landmark o with keypoint match_index_current[i] and marks
 * match_index_previous[i] taken.  At most max_matches are written; the return
 * value is the number of matches (it may exceed max_matches) or a negative
 * ZSBA_E* code (ZSBA_EUNSUP for more than 2^31 - 1 landmark-keypoint pairs
 * when a projection is NaN or there are more than 4096 candidates).  A matcher keeps its device buffers between calls; calls on one
 * matcher must not overlap. */
typedef struct zsba_matcher zsba_matcher;
typedef struct zsba_match_frame {
    double rotation[9];     /* row-major R: camera = R * world + t */
    double translation[3];
    double K[4];            /* fx, fy, cx, cy */
    int64_t cols, rows;     /* image size; the border is 10 px (slam.hpp:348) */
} zsba_match_frame;
zsba_matcher* zsba_matcher_create(int32_t device);
void          zsba_matcher_destroy(zsba_matcher* matcher);
const char*   zsba_matcher_last_error(const zsba_matcher* matcher);
int32_t zsba_match_by_projection(zsba_matcher* matcher, const zsba_match_frame* frame,
                                 int32_t n_landmarks, const double* landmark_xyz /* n*3 */,
                                 const uint8_t* landmark_descriptor /* n*32 */,
                                 const uint8_t* landmark_observed /* n */,
                                 int32_t n_keypoints, const double* keypoint_xy /* n*2 */,
                                 const uint8_t* keypoint_descriptor /* n*32 */,
                                 const int32_t* keypoint_key /* n */,
                                 int32_t n_keys, const uint8_t* key_taken /* n_keys */,
                                 int32_t max_matches, int32_t* match_landmark, int32_t* match_keypoint);
/* Device time of the last call (from its first kernel to its last, including the
 * read-back of the candidate count) and its number of candidate pairs. */
int32_t zsba_matcher_stats(const zsba_matcher* matcher, double* device_ms, int32_t* candidates);
/* zsba_match_by_projection_noalloc: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc): the search of
 * zsba_match_by_projection with every input and output array in device memory
 * of the current device, enqueued on `stream` (a hipStream_t of that device;
 * NULL = the null stream) with no allocation, no memset, no host
 * synchronisation and no read-back.  Its intermediate buffers are one static
 * device array per GPU of at most 10 MiB, held until the process ends: calls
 * that overlap on one GPU race (the array keeps the last call's contents; each
 * call clears what must start at zero).  Limits, else ZSBA_EUNSUP at once:
 * n_landmarks <= 262144, n_keypoints <= 16384, n_keys <= 65536, frame rows <=
 * 2048.  *match_count (device) receives the number of matches (the first
 * max_matches are written, in the order of zsba_match_by_projection) or, when
 * the banded search cannot finish the call, a negative code: ZSBA_EUNSUP for a
 * landmark whose projection is not finite or more than 4096 candidate pairs
 * (zsba_match_by_projection handles both), ZSBA_EINVAL for a keypoint_key
 * outside [0, n_keys).  Returns 0 once the work is enqueued, else a negative
 * ZSBA_E* code (nothing is enqueued then). */
int32_t zsba_match_by_projection_noalloc(const zsba_match_frame* frame,
                                         int32_t n_landmarks, const double* landmark_xyz /* device n*3 */,
                                         const uint8_t* landmark_descriptor /* device n*32 */,
                                         const uint8_t* landmark_observed /* device n */,
                                         int32_t n_keypoints, const double* keypoint_xy /* device n*2 */,
                                         const uint8_t* keypoint_descriptor /* device n*32 */,
                                         const int32_t* keypoint_key /* device n */,
                                         int32_t n_keys, const uint8_t* key_taken /* device n_keys */,
                                         int32_t max_matches, int32_t* match_landmark /* device */,
                                         int32_t* match_keypoint /* device */, int32_t* match_count /* device, 1 */,
                                         void* stream /* hipStream_t */);

/* ---- implicit differentiation (no reference counterpart) ----------------------
 * zsba_gauss_newton applies `steps` undamped Gauss-Newton steps unconditionally
 * (lambda = 0; no cost test), polishing a converged solve to the stationary
 * point of the binary64 cost.  Returns 0 or ZSBA_ESTATE when the normal
 * equations are not positive definite.
 *
 * zsba_adjoint: at a stationary state x* (after zsba_solve to convergence and
 * zsba_gauss_newton), propagate dL/d(points) (and optionally dL/d(pose local
 * increments, rotation then translation as in vertex_pose::plus,
 * factor_graph.hpp:947-960)) to every observation's pixel coordinates and to the
 * pinhole intrinsics (fx, fy, cx, cy), by the implicit function theorem on the
 * optimality condition g(x*, theta) = 0 of the robust cost:
 *   dL/dtheta = -lambda^T dg/dtheta,   H lambda = dL/dx,
 * with H the exact Hessian (forward-mode derivatives of the reference's
 * gradient arithmetic) solved by conjugate gradients preconditioned with the
 * Gauss-Newton reduced system's Cholesky factor.  dL_dK is the derivative for
 * intrinsics applied to every observation's camera.  Outputs are indexed by id
 * (NULL outputs are skipped).  Returns the conjugate-gradient iterations. */
int32_t zsba_gauss_newton(zsba_graph* graph, int32_t steps);
/* zsba_hold_indefinite_points: at the current state, set fixed every free point
 * that has no local optimum in its own coordinates: its undamped Gauss-Newton
 * block is singular under the reference's inverse rule (matrix.hpp:749-774), or
 * its block of the exact Hessian (forward-mode derivatives of the gradient) is
 * not positive definite, as for landmarks seen without parallax: a pivot of the
 * block's symmetric part is at most rel_tol times its trace (rel_tol in [0, 1);
 * 0 = 1e-12).  Held points
 * keep their parameters through later solves, Gauss-Newton steps and the
 * adjoint (which then treats them as constants).  held[id] is set to 1 for each
 * (may be NULL; other entries are left untouched).  Returns the number held.
 * No reference counterpart. */
int32_t zsba_hold_indefinite_points(zsba_graph* graph, double rel_tol, uint8_t* held /* n_points_added */);
int32_t zsba_adjoint(zsba_graph* graph, const double* dL_dpoints /* n_points_added*3 */,
                     const double* dL_dposes /* n_poses_added*6 or NULL */,
                     double* dL_dobs /* n_obs_added*2 or NULL */, double dL_dK[4] /* or NULL */);

/* Adjoint operators on a vector ordered [live poses 6 each | live points 3 each]
 * in slot order: Hv = exact Hessian product, Minv_v = Gauss-Newton Schur
 * preconditioner applied (either output may be NULL). */
int32_t zsba_debug_adjoint_ops(zsba_graph* graph, const double* v, double* Hv, double* Minv_v);

/* ---- diagnostics used by the parity tests -------------------------------------
 * Per-observation FP64 evaluation on the device at the current state, in
 * observation-id order, 27 doubles each:
 *   [0..1] residual, [2..13] pose Jacobian 2x6, [14..19] point Jacobian 2x3,
 *   [20] rho' (drho), [21..24] robust information 2x2, [25] chi2, [26] robust chi2
 * (factor_graph.hpp:217-269, 990-1048). */
#define ZSBA_EDGE_EVAL_STRIDE 27
int32_t zsba_debug_eval_edges(zsba_graph* graph, double* out /* n_obs_added*27 */);
/* Builds the damped reduced camera system at the current state with damping
 * lambda, in the given precision: S (n x n row-major, lower triangle valid),
 * b_s (n), the pose step (n) and the point step (3 * n_points), where
 * n = 6 * live poses.  Returns 1 if the Cholesky decompose/solve succeeded, 0 if
 * it failed (factor_graph.hpp:736-806). */
int32_t zsba_debug_reduced_system(zsba_graph* graph, double lambda, int32_t precision,
                                  double* S, double* b_s, double* dx_pose, double* dx_point);
/* The engine's correctly rounded elementary functions: fn 0 sin, 1 cos,
 * 2 cube (pow(x, 3.0)), 3 sqrt; evaluated by the host code (on_device = 0) or
 * by a device kernel (on_device = 1). */
int32_t zsba_debug_math(int32_t fn, int32_t n, const double* in, double* out, int32_t on_device);

/* ---- output formats (reference source/main.cpp:132-289, tools/trajectory_alignment)
 * zsba_pose_to_frame: the frame pose map::optimise writes back from an
 * optimised pose vertex (map.hpp:239-244): rotation (row-major 3x3, camera =
 * R world + t) = lie::so3(qw, qx, qy, qz).get_matrix(), translation = t.
 * zsba_write_trajectory_tum: save_trajectory_as_txt (main.cpp:132-156), one
 * line "id x y z qx qy qz qw" per frame in ascending id, with the camera centre
 * -R^T t and the quaternion of R as lie::so3(R) forms it; byte-identical to the
 * reference's file for the same frames.
 * zsba_write_map_ply: save_trajectory_and_map_as_ply (main.cpp:158-289): the
 * landmarks (landmark_colour in [0,1], NULL = black) and five vertices and
 * eight edges of each camera's frustum, in the order given (pass the map's
 * iteration order of its landmarks and frames to reproduce its file byte for
 * byte); K = (fx, fy, cx, cy).
 * zsba_trajectory_ate: the alignment and error statistics of
 * tools/trajectory_alignment/main.cpp:195-276: a similarity transform of the
 * estimated camera centres onto the ground truth (absolute_orientation.hpp;
 * first_pose_overlap = its --first option), then the error and path-length
 * statistics of the aligned trajectory over n_poses >= 3 pose pairs. */
typedef struct zsba_ate {
    double scale, rotation[9], translation[3];   /* ground truth = scale R estimate + t */
    double max, min, mean, median, rmse, sse, std; /* position errors (m)         */
    double distance_ground_truth, distance_estimated; /* path lengths (m)          */
    int32_t poses, pad;
} zsba_ate;
int32_t zsba_pose_to_frame(const double pose[7], double rotation[9], double translation[3]);
int32_t zsba_write_trajectory_tum(const char* path, int32_t n_frames, const int32_t* frame_ids,
                                  const double* rotation /* n*9 */, const double* translation /* n*3 */);
int32_t zsba_write_map_ply(const char* path, int32_t image_width, int32_t image_height, int32_t n_landmarks,
                           const double* landmark_xyz /* n*3 */, const double* landmark_colour /* n*3 or NULL */,
                           int32_t n_frames, const int32_t* frame_ids, const double* rotation /* n*9 */,
                           const double* translation /* n*3 */, const double K[4]);
int32_t zsba_trajectory_ate(int32_t n_poses, const double* ground_truth_xyz /* n*3 */,
                            const double* estimated_xyz /* n*3 */, int32_t first_pose_overlap, zsba_ate* out);

/* ---- synthetic problems (benchmark and test workloads) ------------------------ */
typedef struct zsba_synth_params {
    int32_t layout;       /* 0 circle look-at, 1 forward walk, 2 double loop facing out */
    int32_t n_frames;
    int32_t n_points;
    int32_t track_min;    /* track length range of a landmark (frames)             */
    int32_t track_max;
    int32_t track_geometric; /* 1: geometric track lengths on [min,max], mean track_mean */
    double  track_mean;
    double  revisit_rate; /* layout 2: share of landmarks also seen on the other lap  */
    double  noise_px;     /* sigma of the pixel noise                                   */
    double  outlier_rate; /* share of observations replaced by uniform image points    */
    int32_t loss_kind;
    double  loss_delta;
    double  rot_perturb;  /* rad, applied to non-gauge poses                            */
    double  trans_perturb;/* m                                                          */
    double  point_perturb;/* sigma (m) per coordinate                                   */
    int32_t n_fixed;      /* gauge: the first n_fixed poses are fixed                   */
    double  K[4];
    int32_t width, height;
    double  min_parallax; /* layouts 1-2: a landmark is accepted only when the first
                             and last camera centres of its track subtend at least
                             this angle at it (rad; 0 = no test)                    */
    int32_t reserved[2];
} zsba_synth_params;

typedef struct zsba_synth zsba_synth;
/* Defaults for the five benchmark configurations (0..4) of the engine. */
int32_t     zsba_synth_default_params(int32_t config, zsba_synth_params* params);
zsba_synth* zsba_synth_create(const zsba_synth_params* params, uint64_t seed);
void        zsba_synth_destroy(zsba_synth* synth);
/* sizes[0..2] = poses, points, observations */
void        zsba_synth_sizes(const zsba_synth* synth, int64_t sizes[3]);
/* Any output pointer may be NULL.  true_* are the unperturbed values; outlier
 * flags mark observations replaced by uniform image points. */
void        zsba_synth_copy(const zsba_synth* synth, double* poses, uint8_t* pose_fixed,
                            double* points, uint8_t* point_fixed, int32_t* obs_pose,
                            int32_t* obs_point, double* obs_uv, double* true_poses,
                            double* true_points, uint8_t* outlier);

#ifdef __cplusplus
}
#endif

#endif /* ZSBA_H */
