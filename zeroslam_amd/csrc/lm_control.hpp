// Levenberg-Marquardt controller of factor_graph::solve (factor_graph.hpp:489-615)
// as device functions on the lm_state record, shared by the per-graph kernel
// sequence and the batched per-window solver.  Each function is run by one
// thread; the kernels around it read lm_state to decide whether to work.
#pragma once

#include "ba_math.hpp"
#include "common.hpp"

namespace zb {

// ---- LM initialisation, factor_graph.hpp:511-522 -----------------------------
// max_diagonal = max over all pose then point diagonal entries with
// std::max(|h|, m) semantics (a NaN is kept only until the next entry), reduced
// over chunks with this associative combine.
struct max_scan {
    double tail;   // max of the entries after the last NaN
    int has_nan;
    int last_nan;  // the final entry was NaN
    int empty;
};
__device__ inline max_scan scan_combine(max_scan a, max_scan b) {
    if (b.empty) return a;
    if (a.empty) return b;
    if (b.has_nan) return b;
    max_scan r;
    r.has_nan = a.has_nan;
    r.last_nan = 0;
    r.empty = 0;
    r.tail = (a.tail < b.tail) ? b.tail : a.tail;
    if (a.last_nan) r.tail = b.tail;
    return r;
}
__device__ inline max_scan scan_push(max_scan m, double h) {
    const double a = fabs(h);
    if (isnan(a)) {
        m.has_nan = 1; m.last_nan = 1; m.empty = 0; m.tail = 0.0;
    } else {
        if (m.empty || m.last_nan) { m.tail = a; } else { m.tail = (a < m.tail) ? m.tail : a; }
        m.last_nan = 0; m.empty = 0;
    }
    return m;
}
// std::max(|h|, m) starting from m = 0: a final NaN leaves NaN, otherwise the max
// of the entries after the last NaN (and of 0 when none).
__device__ inline double scan_max_diagonal(max_scan acc) {
    if (acc.empty) return 0.0;
    if (acc.last_nan) return NAN;
    return acc.has_nan ? acc.tail : ((acc.tail < 0.0) ? 0.0 : acc.tail);
}

// lm->chi2 holds the initial cost; lm->iterations / use_relative are set.
__device__ inline void lm_init(lm_state* lm, double max_diagonal) {
    max_diagonal = (max_diagonal < 5e10) ? max_diagonal : 5e10;  // std::min(5e10, m)
    const double tau = 1e-5;
    lm->lambda = tau * max_diagonal;
    lm->nu = 2.0;
    lm->legacy_threshold = 1e-10 * lm->chi2;
    lm->initial_chi2 = lm->chi2;
    lm->last_chi2 = lm->chi2;
    lm->iter = 0;
    lm->failures = 0;
    lm->success = 0;
    lm->attempts = 0;
    lm->rejected = 0;
    lm->solve_failures = 0;
    lm->executed_iterations = 0;
    lm->need_linearize = 0;
    lm->need_rollback = 0;
    lm->step_applied = 0;
    lm->solve_ok = 0;
    lm->lin_count = 1;
    lm->linearisations = 1;
    lm->done = (lm->iterations <= 0) ? 1 : 0;
}

// ---- end of an outer iteration, factor_graph.hpp:598-611 ------------------------
__device__ inline void lm_outer_end(lm_state* lm) {
    const double floor_chi2 = 1e-12;
    const double m = (lm->last_chi2 < floor_chi2) ? floor_chi2 : lm->last_chi2;  // std::max(last, 1e-12)
    const double threshold = lm->use_relative ? (1e-6 * m) : lm->legacy_threshold;
    const bool stop = (lm->failures >= 10) || ((lm->last_chi2 - lm->chi2) < threshold) || !isfinite(lm->lambda);
    lm->executed_iterations += 1;
    lm->iter += 1;
    if (stop || lm->iter >= lm->iterations) {
        lm->done = 1;
    } else {
        lm->last_chi2 = lm->chi2;
        lm->failures = 0;
    }
}

// ---- after the reduced solve: failure bookkeeping, factor_graph.hpp:542-551 ----
__device__ inline void lm_after_solve(lm_state* lm) {
    if (lm->done) return;
    lm->attempts += 1;
    if (lm->solve_ok) {
        lm->step_applied = 1;
        return;
    }
    lm->step_applied = 0;
    lm->solve_failures += 1;
    lm->lambda *= lm->nu;
    lm->nu *= 2.0;
    lm->failures += 1;
    if (!isfinite(lm->lambda) || lm->failures >= 10) lm_outer_end(lm);
}

// ---- accept / reject, factor_graph.hpp:566-605 ---------------------------------
// The reference accepts a step when rho = (chi2 - chi2_new) / scale > 0.  With an
// exact solve of the damped system the predicted decrease (scale) is positive, so
// that means the cost fell.  The reassociating paths (require_descent) also ask
// for chi2_new < chi2: with binary32 blocks and almost no damping their step can
// miss the system by enough to make scale negative, and the reference's test
// alone would then accept a step that raises the cost.
__device__ inline void lm_decide(lm_state* lm) {
    if (lm->done || !lm->step_applied) return;
    const double temp_chi = lm->chi2_new;
    const double rho = (lm->chi2 - temp_chi) / lm->scale;
    lm->rho = rho;
    bool good = false;
    if ((rho > 0) && isfinite(temp_chi) && (!lm->require_descent || temp_chi < lm->chi2)) {
        double alpha = 1.0 - cr_cube(2.0 * rho - 1.0);  // math::pow(x, 3.0), correctly rounded
        alpha = (2.0 / 3.0 < alpha) ? 2.0 / 3.0 : alpha;
        const double scale_factor = (1.0 / 3.0 < alpha) ? alpha : 1.0 / 3.0;
        lm->lambda *= scale_factor;
        lm->nu = 2.0;
        lm->chi2 = temp_chi;
        good = true;
        lm->success += 1;
    } else {
        lm->lambda *= lm->nu;
        lm->nu *= 2.0;
        lm->rejected += 1;
    }
    if (!isfinite(lm->lambda)) {  // the rejected step stays applied (factor_graph.hpp:592-594)
        lm_outer_end(lm);
        return;
    }
    if (good) {
        lm_outer_end(lm);
        if (!lm->done) lm->need_linearize = 1;
        return;
    }
    lm->need_rollback = 1;
    lm->failures += 1;
    if (lm->failures >= 10) lm_outer_end(lm);
}

__device__ inline void lm_end_attempt(lm_state* lm) {
    lm->need_rollback = 0;
    lm->step_applied = 0;
    lm->solve_ok = 0;
}

__device__ inline void lm_linearise_done(lm_state* lm) {
    if (lm->need_linearize) lm->linearisations += 1;
    lm->need_linearize = 0;
    lm->lin_count += 1;
}

}  // namespace zb
