// =============================================================================
// TEST INFRASTRUCTURE — the CPU oracle.  Nothing in the product links or calls
// this file; only tests/, the smoke check and the bench's CPU-baseline leg load
// it, and only to check or time against it.
//
// A binary64 restatement, in new code, of the bundle-adjustment path of ZeroSLAM:
//   factor_graph::factor_graph::solve and everything it calls for vertex_pose /
//   vertex_point_xyz / edge_reprojection<camera::pinhole>
//   (source/factor_graph.hpp:40-1049, source/lie.hpp:133-531 (so3/se3),
//   source/camera.hpp:91-145, source/matrix.hpp:394-625, 749-774,
//   source/matrix_sparse_block.hpp:38-211, source/matrix_sparse_block_diagonal.hpp:40-143,
//   source/matrix_decomposition_cholesky.hpp:26-95).
//
// Every floating-point expression keeps the reference's operation order
// (products start from sum = 0 and accumulate k ascending, matrix.hpp:499-523;
// scalar factors are applied element-wise), and the block-sparse Schur products
// iterate std::unordered_map containers keyed and hashed as the reference's
// sparse_block is (matrix_sparse_block.hpp:47-68) and filled/reserved in the
// same sequence, so the summation order of every Schur sum is the reference's.
// Built with -O3 -ffp-contract=off and no -march the oracle reproduces the
// compiled reference bit for bit on the same machine; the tests pin that.
//
// The Schur product itself visits only the point blocks that exist instead of
// probing every pose column (matrix_sparse_block.hpp:166-168), which yields the
// identical sequence of additions at O(pairs) instead of O(blocks * poses) cost.
// =============================================================================
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

// ---- loss functions: factor_graph.hpp:851-934 -------------------------------
enum { LOSS_NONE = 0, LOSS_TRIVIAL = 1, LOSS_HUBER = 2, LOSS_CAUCHY = 3, LOSS_TUKEY = 4 };

void loss_compute(int kind, double delta, double e2, double rho[3]) {
    switch (kind) {
    case LOSS_TRIVIAL:  // factor_graph.hpp:854-858
        rho[0] = e2; rho[1] = 1; rho[2] = 0;
        return;
    case LOSS_HUBER: {  // factor_graph.hpp:871-884
        const double d2 = delta * delta;
        if (e2 <= d2) {
            rho[0] = e2; rho[1] = 1.0; rho[2] = 0.0;
        } else {
            const double e = std::sqrt(e2);
            rho[0] = 2.0 * e * delta - d2;
            rho[1] = delta / e;
            rho[2] = -0.5 * rho[1] / e2;
        }
        return;
    }
    case LOSS_CAUCHY: {  // factor_graph.hpp:897-904
        const double d2 = delta * delta;
        const double d2r = 1.0 / d2;
        const double aux = d2r * e2 + 1.0;
        rho[0] = d2 * std::log(aux);
        rho[1] = 1.0 / aux;
        rho[2] = -d2r * (rho[1] * rho[1]);
        return;
    }
    case LOSS_TUKEY: {  // factor_graph.hpp:917-933
        const double d2 = delta * delta;
        if (e2 <= d2) {
            const double aux = e2 / d2;
            const double oma = 1.0 - aux;
            const double oma2 = oma * oma;
            const double oma3 = oma2 * oma;
            rho[0] = d2 * (1.0 - oma3) / 3.0;
            rho[1] = oma2;
            rho[2] = -2.0 * oma / d2;
        } else {
            rho[0] = d2 / 3.0; rho[1] = 0; rho[2] = 0;
        }
        return;
    }
    default:
        rho[0] = e2; rho[1] = 1; rho[2] = 0;
        return;
    }
}

// ---- dense products: matrix.hpp:492-527 (sum = 0, k ascending) ---------------
template <int R, int K, int C>
inline void mul(const double* a, const double* b, double* out) {
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) {
            double s = 0;
            for (int k = 0; k < K; ++k) s += a[r * K + k] * b[k * C + c];
            out[r * C + c] = s;
        }
}
template <int R, int C>
inline void transpose(const double* a, double* out) {
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) out[c * R + r] = a[r * C + c];
}

// ---- so3 / se3: lie.hpp:133-531 ----------------------------------------------
// Quaternion (w, x, y, z) Hamilton product, lie.hpp:337-344.
inline void qmul(const double a[4], const double b[4], double o[4]) {
    o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    o[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    o[2] = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    o[3] = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
}
// Point action (q * (0,p)) * conj(q), lie.hpp:347-362.
inline void qrotate(const double q[4], const double p[3], double o[3]) {
    const double pq[4] = {0, p[0], p[1], p[2]};
    const double qc[4] = {+q[0], -q[1], -q[2], -q[3]};
    double t[4], r[4];
    qmul(q, pq, t);
    qmul(t, qc, r);
    o[0] = r[1]; o[1] = r[2]; o[2] = r[3];
}
// Rotation matrix of q, lie.hpp:188-204.
inline void qmatrix(const double q[4], double m[9]) {
    const double tx = 2.0 * q[1], ty = 2.0 * q[2], tz = 2.0 * q[3];
    const double txx = tx * q[1], txy = tx * q[2], txz = tx * q[3], txw = tx * q[0];
    const double tyy = ty * q[2], tyz = ty * q[3], tyw = ty * q[0];
    const double tzz = tz * q[3], tzw = tz * q[0];
    m[0] = 1.0 - (tyy + tzz); m[1] = txy - tzw;         m[2] = txz + tyw;
    m[3] = txy + tzw;         m[4] = 1.0 - (txx + tzz); m[5] = tyz - txw;
    m[6] = txz - tyw;         m[7] = tyz + txw;         m[8] = 1.0 - (txx + tyy);
}
inline double length_squared3(const double v[3]) {  // matrix.hpp:177-183
    double s = 0;
    for (int i = 0; i < 3; ++i) s += v[i] * v[i];
    return s;
}
// so3::exp, lie.hpp:247-264.
inline void so3_exp(const double w[3], double q[4]) {
    const double t2 = length_squared3(w);
    double real, imag;
    if (t2 < 1e-6 * 1e-6) {
        const double t4 = t2 * t2;
        real = 1.0 - (1.0 / 8.0) * t2 + (1.0 / 384.0) * t4;
        imag = 0.5 - (1.0 / 48.0) * t2 + (1.0 / 3840.0) * t4;
    } else {
        const double theta = std::sqrt(t2);
        const double half = 0.5 * theta;
        real = std::cos(half);
        imag = std::sin(half) / theta;
    }
    q[0] = real; q[1] = w[0] * imag; q[2] = w[1] * imag; q[3] = w[2] * imag;
}
// so3::left_jacobian, lie.hpp:284-294 (element-wise (I + a*W) + b*(W*W)).
inline void so3_left_jacobian(const double w[3], double J[9]) {
    const double t2 = length_squared3(w);
    const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double WW[9];
    mul<3, 3, 3>(W, W, WW);
    double a, b;
    if (t2 < 1e-6 * 1e-6) {
        a = 0.5;
        b = 1.0 / 6.0;
    } else {
        const double theta = std::sqrt(t2);
        a = (1 - std::cos(theta)) / (t2);
        b = (theta - std::sin(theta)) / (t2 * theta);
    }
    for (int i = 0; i < 9; ++i) {
        const double I = (i % 4 == 0) ? 1.0 : 0.0;
        J[i] = (I + W[i] * a) + WW[i] * b;
    }
}
// vertex_pose::plus: T <- exp(delta) * T, factor_graph.hpp:947-960 with
// se3::exp (lie.hpp:446-450) and se3::operator* (lie.hpp:521-526).
inline void pose_plus(double p[7], const double d[6]) {
    const double q[4] = {p[6], p[3], p[4], p[5]};
    const double t[3] = {p[0], p[1], p[2]};
    double qe[4], J[9], te[3];
    so3_exp(d, qe);
    so3_left_jacobian(d, J);
    mul<3, 3, 1>(J, d + 3, te);
    double qn[4], tr[3];
    qmul(qe, q, qn);
    qrotate(qe, t, tr);
    p[0] = tr[0] + te[0]; p[1] = tr[1] + te[1]; p[2] = tr[2] + te[2];
    p[3] = qn[1]; p[4] = qn[2]; p[5] = qn[3]; p[6] = qn[0];
}
// Camera-frame point T * X, lie.hpp:529-531.
inline void pose_apply(const double p[7], const double X[3], double Pc[3]) {
    const double q[4] = {p[6], p[3], p[4], p[5]};
    double r[3];
    qrotate(q, X, r);
    Pc[0] = r[0] + p[0]; Pc[1] = r[1] + p[1]; Pc[2] = r[2] + p[2];
}

// ---- 3x3 inverse: matrix.hpp:749-774 ----------------------------------------
bool invert3(const double* v, double* out) {
    const double det = +v[0] * (v[4] * v[8] - v[7] * v[5]) - v[1] * (v[3] * v[8] - v[5] * v[6]) +
                       v[2] * (v[3] * v[7] - v[4] * v[6]);
    double scale = 0;
    for (int i = 0; i < 9; ++i) {
        const double a = std::fabs(v[i]);
        scale = scale < a ? a : scale;  // math::max, math.hpp:235-237
    }
    if (!(std::fabs(det) > scale * scale * scale * 1e-12)) {
        for (int i = 0; i < 9; ++i) out[i] = 0;
        return false;
    }
    const double di = 1.0 / det;
    out[0] = (v[4] * v[8] - v[7] * v[5]) * di;
    out[1] = (v[2] * v[7] - v[1] * v[8]) * di;
    out[2] = (v[1] * v[5] - v[2] * v[4]) * di;
    out[3] = (v[5] * v[6] - v[3] * v[8]) * di;
    out[4] = (v[0] * v[8] - v[2] * v[6]) * di;
    out[5] = (v[3] * v[2] - v[0] * v[5]) * di;
    out[6] = (v[3] * v[7] - v[6] * v[4]) * di;
    out[7] = (v[6] * v[1] - v[0] * v[7]) * di;
    out[8] = (v[0] * v[4] - v[3] * v[1]) * di;
    return true;
}

// ---- dense Cholesky: matrix_decomposition_cholesky.hpp:26-95 ----------------
bool decompose_cholesky(const double* A, int n, double* L) {
    for (long i = 0; i < static_cast<long>(n) * n; ++i) L[i] = 0;
    double scale = 0;
    for (int i = 0; i < n; ++i) {
        const double a = std::fabs(A[static_cast<long>(i) * n + i]);
        scale = scale < a ? a : scale;
    }
    const double eps = scale * 1e-14;
    for (int i = 0; i < n; ++i) {
        double* Li = L + static_cast<long>(i) * n;
        for (int k = 0; k < i; ++k) {
            const double* Lk = L + static_cast<long>(k) * n;
            double value = A[static_cast<long>(i) * n + k];
            for (int j = 0; j < k; ++j) value -= Li[j] * Lk[j];
            Li[k] = value / Lk[k];
        }
        double value = A[static_cast<long>(i) * n + i];
        for (int j = 0; j < i; ++j) value -= Li[j] * Li[j];
        if (!(value > eps)) return false;
        Li[i] = std::sqrt(value);
    }
    return true;
}
bool solve_cholesky(const double* L, const double* b, int n, double* x) {
    double scale = 0;
    for (int i = 0; i < n; ++i) {
        const double a = std::fabs(L[static_cast<long>(i) * n + i]);
        scale = scale < a ? a : scale;
    }
    const double eps = scale * 1e-7;
    for (int r = 0; r < n; ++r) {
        x[r] = b[r];
        for (int j = 0; j < r; ++j) x[r] -= L[static_cast<long>(r) * n + j] * x[j];
        if (!(std::fabs(L[static_cast<long>(r) * n + r]) > eps)) return false;
        x[r] /= L[static_cast<long>(r) * n + r];
    }
    for (int i = n; i-- > 0;) {
        for (int j = i + 1; j < n; ++j) x[i] -= L[static_cast<long>(j) * n + i] * x[j];
        if (!(std::fabs(L[static_cast<long>(i) * n + i]) > eps)) return false;
        x[i] /= L[static_cast<long>(i) * n + i];
    }
    return true;
}

// ---- block-sparse containers: matrix_sparse_block.hpp:47-68 -----------------
struct block_key {
    size_t first, second;
    bool operator==(const block_key& o) const { return first == o.first && second == o.second; }
};
// Same hash as the reference's block_key (matrix_sparse_block.hpp:56-58); not
// noexcept, so the container caches hash codes as the reference's does.
struct block_hash {
    std::size_t operator()(const block_key& k) const {
        return ((std::hash<size_t>()(k.first) ^ (std::hash<size_t>()(k.second) << 1)) >> 1);
    }
};
template <int N> struct block { double v[N] = {}; };
using map63 = std::unordered_map<block_key, block<18>, block_hash>;
using map36 = std::unordered_map<block_key, block<18>, block_hash>;
using map66 = std::unordered_map<block_key, block<36>, block_hash>;

struct edge_state {
    double r[2] = {0, 0};
    double Jp[12] = {};
    double Jl[6] = {};
};

struct graph {
    int np = 0, nl = 0, no = 0;
    std::vector<double> pose, pose_backup;      // np * 7
    std::vector<uint8_t> pose_fixed;
    std::vector<double> point, point_backup;    // nl * 3
    std::vector<uint8_t> point_fixed;
    std::vector<int32_t> obs_pose, obs_point;
    std::vector<double> obs_uv, obs_info, obs_K;
    std::vector<int32_t> obs_loss;
    std::vector<double> obs_delta;
    std::vector<edge_state> edges;
    int verbose = 0;
    std::string trace;

    // Solver state, factor_graph.hpp:346-375.
    double lambda = 0, nu = 0, chi2 = 0;
    std::vector<double> h_pp;            // np * 36, sparse_block_diagonal<6>
    std::vector<double> h_ll;            // nl * 9,  sparse_block_diagonal<3>
    std::vector<double> h_ll_inverse;    // nl * 9
    map63 h_pl;                          // persists across make_hessian calls (clear keeps buckets)
    map36 h_lp;
    std::vector<double> b, dx;           // 6np + 3nl
    int P = 0, M = 0;                    // count_general_params, count_marginalised_params
    // statistics of the last solve
    int32_t stat_iterations = 0, stat_attempts = 0, stat_accepted = 0, stat_rejected = 0, stat_failures = 0;
    double stat_initial_chi2 = 0;

    void log(const char* fmt, ...) {
        if (!verbose) return;
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        trace += buf;
        if (verbose > 1) std::fputs(buf, stderr);
    }

    const double* K(int e) const { return obs_K.data() + 4L * e; }
    const double* info(int e) const { return obs_info.data() + 4L * e; }

    // edge_reprojection::compute_residual, factor_graph.hpp:990-1009, with
    // camera::pinhole::project, camera.hpp:91-111.
    void compute_residual(int e) {
        double Pc[3];
        pose_apply(&pose[7L * obs_pose[e]], &point[3L * obs_point[e]], Pc);
        edge_state& s = edges[e];
        if (Pc[2] < 1.0e-12) {
            const double penalty = -0.3 * Pc[2];
            s.r[0] = penalty; s.r[1] = penalty;
            return;
        }
        const double* k = K(e);
        const double iz = 1.0 / Pc[2];
        const double x = Pc[0] * iz, y = Pc[1] * iz;
        const double u = k[0] * x + k[2], v = k[1] * y + k[3];
        s.r[0] = obs_uv[2L * e] - u;
        s.r[1] = obs_uv[2L * e + 1] - v;
    }
    // edge_reprojection::compute_jacobians, factor_graph.hpp:1012-1048 with
    // camera.hpp:113-126.
    void compute_jacobians(int e) {
        const double* p = &pose[7L * obs_pose[e]];
        double Pc[3];
        pose_apply(p, &point[3L * obs_point[e]], Pc);
        double Jc[6];
        if (Pc[2] < 1.0e-12) {
            Jc[0] = 0.0; Jc[1] = 0.0; Jc[2] = 0.3;
            Jc[3] = 0.0; Jc[4] = 0.0; Jc[5] = 0.3;
        } else {
            const double* k = K(e);
            const double iz = 1.0 / Pc[2];
            const double iz2 = iz * iz;
            Jc[0] = k[0] * iz; Jc[1] = 0.0; Jc[2] = -k[0] * Pc[0] * iz2;
            Jc[3] = 0.0; Jc[4] = k[1] * iz; Jc[5] = -k[1] * Pc[1] * iz2;
        }
        const double X = Pc[0], Y = Pc[1], Z = Pc[2];
        const double Jpp[18] = {0, Z, -Y, 1, 0, 0, -Z, 0, X, 0, 1, 0, Y, -X, 0, 0, 0, 1};
        edge_state& s = edges[e];
        double t[12];
        mul<2, 3, 6>(Jc, Jpp, t);
        for (int i = 0; i < 12; ++i) s.Jp[i] = -t[i];
        const double q[4] = {p[6], p[3], p[4], p[5]};
        double R[9], u[6];
        qmatrix(q, R);
        mul<2, 3, 3>(Jc, R, u);
        for (int i = 0; i < 6; ++i) s.Jl[i] = -u[i];
    }
    // edge_base::chi2, factor_graph.hpp:217-219: (r^T * Omega) * r.
    double edge_chi2(int e) const {
        const double* r = edges[e].r;
        const double* W = info(e);
        double t[2];
        mul<1, 2, 2>(r, W, t);
        double s = 0;
        for (int k = 0; k < 2; ++k) s += t[k] * r[k];
        return s;
    }
    // edge_base::robust_chi2, factor_graph.hpp:221-228.
    double edge_robust_chi2(int e) const {
        if (obs_loss[e] == LOSS_NONE) return edge_chi2(e);
        double rho[3] = {0, 0, 0};
        loss_compute(obs_loss[e], obs_delta[e], edge_chi2(e), rho);
        return rho[0];
    }
    // edge_base::robust_info, factor_graph.hpp:254-269.
    void robust_info(int e, double& drho, double Wr[4]) const {
        const double* W = info(e);
        if (obs_loss[e] == LOSS_NONE) {
            drho = 1.0;
            for (int i = 0; i < 4; ++i) Wr[i] = W[i];
            return;
        }
        const double e2 = edge_chi2(e);
        double rho[3];
        loss_compute(obs_loss[e], obs_delta[e], e2, rho);
        drho = rho[1];
        for (int i = 0; i < 4; ++i) Wr[i] = W[i] * rho[1];
        if ((rho[1] + 2.0 * rho[2] * e2) > 0.0) {
            double we[2];
            mul<2, 2, 1>(W, edges[e].r, we);
            const double s = 2.0 * rho[2];
            const double swe[2] = {we[0] * s, we[1] * s};
            double outer[4];
            mul<2, 1, 2>(swe, we, outer);
            for (int i = 0; i < 4; ++i) Wr[i] = Wr[i] + outer[i];
        }
    }

    // factor_graph::get_current_chi, factor_graph.hpp:618-629.
    double current_chi(bool recompute) {
        if (recompute)
            for (int e = 0; e < no; ++e) compute_residual(e);
        double c = 0.0;
        for (int e = 0; e < no; ++e) c += edge_robust_chi2(e);
        return c;
    }

    // factor_graph::set_ordering, factor_graph.hpp:633-645.
    void set_ordering() {
        P = 6 * np;
        M = 3 * nl;
    }

    // factor_graph::make_hessian, factor_graph.hpp:648-732.
    void make_hessian() {
        h_pp.assign(36L * np, 0.0);
        h_ll.assign(9L * nl, 0.0);
        h_pl.clear();
        h_ll_inverse.assign(9L * nl, 0.0);
        b.assign(static_cast<size_t>(P + M), 0.0);
        dx.assign(static_cast<size_t>(P + M), 0.0);
        for (int e = 0; e < no; ++e) {
            compute_jacobians(e);
            const edge_state& s = edges[e];
            const int ip = obs_pose[e], il = obs_point[e];
            const bool fp = pose_fixed[ip] != 0, fl = point_fixed[il] != 0;
            const double* W = info(e);
            const double* r = s.r;
            // i = pose vertex
            if (!fp) {
                double drho, Wr[4];
                robust_info(e, drho, Wr);
                double JpT[12], jtw[12];
                transpose<2, 6>(s.Jp, JpT);
                mul<6, 2, 2>(JpT, Wr, jtw);
                double H[36];
                mul<6, 2, 6>(jtw, s.Jp, H);
                double* hp = &h_pp[36L * ip];
                for (int i = 0; i < 36; ++i) hp[i] = hp[i] + H[i];
                if (!fl) {
                    double Hpl[18];
                    mul<6, 2, 3>(jtw, s.Jl, Hpl);
                    const block_key key{static_cast<size_t>(ip), static_cast<size_t>(il)};
                    auto it = h_pl.find(key);
                    if (it == h_pl.end()) {
                        block<18>& blk = h_pl[key];
                        for (int i = 0; i < 18; ++i) blk.v[i] = Hpl[i];
                    } else {
                        for (int i = 0; i < 18; ++i) it->second.v[i] = it->second.v[i] + Hpl[i];
                    }
                }
                // b_i -= ((drho * J^T) * Omega) * r, factor_graph.hpp:725.
                double dJ[12], t[12], v[6];
                for (int i = 0; i < 12; ++i) dJ[i] = JpT[i] * drho;
                mul<6, 2, 2>(dJ, W, t);
                mul<6, 2, 1>(t, r, v);
                double* bb = &b[6L * ip];
                for (int i = 0; i < 6; ++i) bb[i] = bb[i] - v[i];
            }
            // i = point vertex
            if (!fl) {
                double drho, Wr[4];
                robust_info(e, drho, Wr);
                double JlT[6], jtw[6];
                transpose<2, 3>(s.Jl, JlT);
                mul<3, 2, 2>(JlT, Wr, jtw);
                double H[9];
                mul<3, 2, 3>(jtw, s.Jl, H);
                double* hl = &h_ll[9L * il];
                for (int i = 0; i < 9; ++i) hl[i] = hl[i] + H[i];
                double dJ[6], t[6], v[3];
                for (int i = 0; i < 6; ++i) dJ[i] = JlT[i] * drho;
                mul<3, 2, 2>(dJ, W, t);
                mul<3, 2, 1>(t, r, v);
                double* bb = &b[static_cast<size_t>(P) + 3L * il];
                for (int i = 0; i < 3; ++i) bb[i] = bb[i] - v[i];
            }
        }
        // h_lp = h_pl.get_transpose(), matrix_sparse_block.hpp:132-139.
        map36 t;
        t.reserve(h_pl.size());
        for (auto it = h_pl.begin(); it != h_pl.end(); ++it) {
            block<18>& blk = t[block_key{it->first.second, it->first.first}];
            transpose<6, 3>(it->second.v, blk.v);
        }
        h_lp = std::move(t);
    }

    // Reduced camera system of solve_linear_system, factor_graph.hpp:736-781,
    // at damping `lam`: fills the dense S (P x P) and b_s.  h_pp/h_ll are damped in
    // place (the caller restores them) and h_ll_inverse is filled.
    void reduced_system(double lam, std::vector<double>& dense, std::vector<double>& b_s) {
        for (int d = 0; d < np; ++d)
            for (int i = 0; i < 6; ++i) h_pp[36L * d + 7 * i] += lam;
        for (int d = 0; d < nl; ++d)
            for (int i = 0; i < 3; ++i) h_ll[9L * d + 4 * i] += lam;
        for (int l = 0; l < nl; ++l) {
            if (point_fixed[l]) continue;
            double inv[9];
            invert3(&h_ll[9L * l], inv);
            std::memcpy(&h_ll_inverse[9L * l], inv, sizeof inv);
        }
        // Y = h_pl * h_ll_inverse, matrix_sparse_block_diagonal.hpp:107-122.
        map63 Y;
        Y.reserve(h_pl.size());
        for (auto it = h_pl.begin(); it != h_pl.end(); ++it) {
            auto f = Y.find(it->first);
            double prod[18];
            mul<6, 3, 3>(it->second.v, &h_ll_inverse[9L * it->first.second], prod);
            if (f == Y.end()) {
                block<18>& blk = Y[it->first];
                std::memcpy(blk.v, prod, sizeof prod);
            } else {
                for (int i = 0; i < 18; ++i) f->second.v[i] = f->second.v[i] + prod[i];
            }
        }
        // b_s = b_pp - Y * b_ll, matrix_sparse_block.hpp:184-196.
        std::vector<double> yb(static_cast<size_t>(P), 0.0);
        for (auto it = Y.begin(); it != Y.end(); ++it) {
            double prod[6];
            mul<6, 3, 1>(it->second.v, &b[static_cast<size_t>(P) + 3 * it->first.second], prod);
            double* o = &yb[6 * it->first.first];
            for (int i = 0; i < 6; ++i) o[i] = o[i] + prod[i];
        }
        b_s.resize(static_cast<size_t>(P));
        for (int i = 0; i < P; ++i) b_s[i] = b[i] - yb[i];
        // Poses observing each point, ascending: the rhs columns that
        // sparse_block::multiply finds (matrix_sparse_block.hpp:166-169).
        std::vector<std::vector<std::pair<int, const double*>>> cols(static_cast<size_t>(nl));
        for (auto it = h_lp.begin(); it != h_lp.end(); ++it)
            cols[it->first.first].push_back({static_cast<int>(it->first.second), it->second.v});
        for (auto& c : cols) std::sort(c.begin(), c.end(), [](const auto& a, const auto& b2) { return a.first < b2.first; });
        // Y * h_lp, matrix_sparse_block.hpp:158-182.
        map66 prodmap;
        prodmap.reserve(Y.size());
        for (auto it = Y.begin(); it != Y.end(); ++it) {
            const size_t row = it->first.first;
            for (const auto& c : cols[it->first.second]) {
                double prod[36];
                mul<6, 3, 6>(it->second.v, c.second, prod);
                const block_key key{row, static_cast<size_t>(c.first)};
                auto f = prodmap.find(key);
                if (f == prodmap.end()) {
                    block<36>& blk = prodmap[key];
                    std::memcpy(blk.v, prod, sizeof prod);
                } else {
                    for (int i = 0; i < 36; ++i) f->second.v[i] = f->second.v[i] + prod[i];
                }
            }
        }
        // S = subtract(h_pp, Y * h_lp), matrix_sparse_block_diagonal.hpp:124-143.
        map66 S;
        S.reserve(prodmap.size());
        for (int d = 0; d < np; ++d) {
            block<36>& blk = S[block_key{static_cast<size_t>(d), static_cast<size_t>(d)}];
            std::memcpy(blk.v, &h_pp[36L * d], sizeof blk.v);
        }
        for (auto it = prodmap.begin(); it != prodmap.end(); ++it) {
            auto f = S.find(it->first);
            if (f == S.end()) {
                block<36>& blk = S[it->first];
                for (int i = 0; i < 36; ++i) blk.v[i] = -it->second.v[i];
            } else {
                for (int i = 0; i < 36; ++i) f->second.v[i] = f->second.v[i] - it->second.v[i];
            }
        }
        // Densify, factor_graph.hpp:774-781.
        dense.assign(static_cast<size_t>(P) * P, 0.0);
        for (auto it = S.begin(); it != S.end(); ++it)
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) {
                    double& d = dense[(it->first.first * 6 + i) * static_cast<size_t>(P) + it->first.second * 6 + j];
                    d = d + it->second.v[i * 6 + j];
                }
    }

    // factor_graph::solve_linear_system, factor_graph.hpp:736-806.
    bool solve_linear_system() {
        const std::vector<double> hpp_backup = h_pp, hll_backup = h_ll;
        std::vector<double> dense, b_s;
        reduced_system(lambda, dense, b_s);
        std::vector<double> L(static_cast<size_t>(P) * P), xp(static_cast<size_t>(P), 0.0);
        if (!decompose_cholesky(dense.data(), P, L.data()) || !solve_cholesky(L.data(), b_s.data(), P, xp.data())) {
            log("Cholesky solver failed!\n");
            h_pp = hpp_backup;
            h_ll = hll_backup;
            dx.assign(static_cast<size_t>(P + M), 0.0);
            return false;
        }
        for (int i = 0; i < P; ++i) dx[i] = xp[i];
        back_substitute(xp);
        h_pp = hpp_backup;
        h_ll = hll_backup;
        return true;
    }
    // delta_x_l = h_ll_inverse * (b_ll - h_lp * delta_x_p), factor_graph.hpp:799-800.
    void back_substitute(const std::vector<double>& xp) {
        std::vector<double> t(static_cast<size_t>(M), 0.0);
        for (auto it = h_lp.begin(); it != h_lp.end(); ++it) {
            double prod[3];
            mul<3, 6, 1>(it->second.v, &xp[6 * it->first.second], prod);
            double* o = &t[3 * it->first.first];
            for (int i = 0; i < 3; ++i) o[i] = o[i] + prod[i];
        }
        for (int l = 0; l < nl; ++l) {
            double v[3], prod[3];
            for (int i = 0; i < 3; ++i) v[i] = b[static_cast<size_t>(P) + 3 * l + i] - t[3 * l + i];
            mul<3, 3, 1>(&h_ll_inverse[9L * l], v, prod);
            for (int i = 0; i < 3; ++i) dx[static_cast<size_t>(P) + 3 * l + i] = 0 + prod[i];
        }
    }

    // factor_graph::update_states / rollback_states, factor_graph.hpp:809-846.
    void update_states() {
        for (int i = 0; i < np; ++i) {
            if (pose_fixed[i]) continue;
            std::memcpy(&pose_backup[7L * i], &pose[7L * i], 7 * sizeof(double));
            pose_plus(&pose[7L * i], &dx[6L * i]);
        }
        for (int l = 0; l < nl; ++l) {
            if (point_fixed[l]) continue;
            std::memcpy(&point_backup[3L * l], &point[3L * l], 3 * sizeof(double));
            for (int k = 0; k < 3; ++k) point[3L * l + k] += dx[static_cast<size_t>(P) + 3L * l + k];
        }
    }
    void rollback_states() {
        for (int i = 0; i < np; ++i)
            if (!pose_fixed[i]) std::memcpy(&pose[7L * i], &pose_backup[7L * i], 7 * sizeof(double));
        for (int l = 0; l < nl; ++l)
            if (!point_fixed[l]) std::memcpy(&point[3L * l], &point_backup[3L * l], 3 * sizeof(double));
    }

    // factor_graph::solve, factor_graph.hpp:489-615.
    int solve(int iterations, bool use_relative_convergence) {
        stat_iterations = stat_attempts = stat_accepted = stat_rejected = stat_failures = 0;
        if (no == 0 || (np == 0 && nl == 0)) {
            log("Cannot solve problem without edges or vertices\n");
            return 0;
        }
        for (int e = 0; e < no; ++e) compute_residual(e);
        set_ordering();
        make_hessian();
        nu = 2.0;
        chi2 = current_chi(true);
        stat_initial_chi2 = chi2;
        double max_diagonal = 0;
        for (int i = 0; i < P; ++i) {
            const double a = std::fabs(h_pp[36L * (i / 6) + 7 * (i % 6)]);
            max_diagonal = std::max(a, max_diagonal);
        }
        for (int i = 0; i < M; ++i) {
            const double a = std::fabs(h_ll[9L * (i / 3) + 4 * (i % 3)]);
            max_diagonal = std::max(a, max_diagonal);
        }
        max_diagonal = std::min(5e10, max_diagonal);
        const double tau = 1e-5;
        lambda = tau * max_diagonal;
        log("[INIT] iter: XXX, attempt XXX, chi = % 7.7f, rho = XXXX.XXXXXXX, base lambda = % 8.7f\n", chi2, lambda);
        const int max_failures = 10;
        const double legacy_stop_threshold = 1e-10 * chi2;
        const double relative_tolerance = 1e-6;
        const double chi_squared_floor = 1e-12;
        int success = 0;
        for (int iter = 0; iter < iterations; ++iter) {
            ++stat_iterations;
            const double last_chi2 = chi2;
            int failures = 0;
            while (failures < max_failures) {
                ++stat_attempts;
                if (!solve_linear_system()) {
                    ++stat_failures;
                    lambda *= nu;
                    nu *= 2.0;
                    ++failures;
                    if (!std::isfinite(lambda)) break;
                    continue;
                }
                update_states();
                double scale = 0;
                for (int i = 0; i < P + M; ++i) scale += dx[i] * (lambda * dx[i] + b[i]);
                scale += 1e-3;
                for (int e = 0; e < no; ++e) compute_residual(e);
                const double temp_chi = current_chi(true);
                const double rho = (chi2 - temp_chi) / scale;
                bool good = false;
                if ((rho > 0) && std::isfinite(temp_chi)) {
                    double alpha = 1.0 - std::pow((2.0 * rho - 1.0), 3.0);
                    alpha = (2.0 / 3.0 < alpha) ? 2.0 / 3.0 : alpha;                 // math::min
                    const double scale_factor = (1.0 / 3.0 < alpha) ? alpha : 1.0 / 3.0;  // math::max
                    lambda *= scale_factor;
                    nu = 2.0;
                    chi2 = temp_chi;
                    good = true;
                    ++success;
                    ++stat_accepted;
                    log("[GOOD] iter: % 3d, attempt % 3d, chi2 = % 7.7f, rho = % 7.7f, next lambda = % 8.7f\n", iter,
                        failures, temp_chi, rho, lambda);
                } else {
                    lambda *= nu;
                    nu *= 2.0;
                    ++stat_rejected;
                    log("[ BAD] iter: % 3d, attempt % 3d, chi2 = % 7.7f, rho = % 7.7f, next lambda = % 8.7f\n", iter,
                        failures, temp_chi, rho, lambda);
                }
                if (!std::isfinite(lambda)) break;
                if (good) {
                    make_hessian();
                    break;
                }
                rollback_states();
                ++failures;
            }
            const double stop_threshold =
                use_relative_convergence ? (relative_tolerance * std::max(last_chi2, chi_squared_floor)) : legacy_stop_threshold;
            if ((failures >= max_failures) || ((last_chi2 - chi2) < stop_threshold) || (!std::isfinite(lambda))) break;
        }
        return success;
    }
};

}  // namespace

// ---- C ABI for ctypes (same signatures as oracle/ref_driver.cpp) -------------
extern "C" {

void* zo_create(int32_t np, const double* poses, const uint8_t* pose_fixed, int32_t nl, const double* points,
                const uint8_t* point_fixed, int32_t no, const int32_t* obs_pose, const int32_t* obs_point,
                const double* obs_uv, const double* obs_info, const double* obs_K, const int32_t* obs_loss,
                const double* obs_delta, int32_t verbose) {
    graph* g = new graph();
    g->np = np; g->nl = nl; g->no = no;
    g->pose.assign(poses, poses + 7L * np);
    g->pose_backup = g->pose;
    g->pose_fixed.assign(pose_fixed, pose_fixed + np);
    g->point.assign(points, points + 3L * nl);
    g->point_backup = g->point;
    g->point_fixed.assign(point_fixed, point_fixed + nl);
    g->obs_pose.assign(obs_pose, obs_pose + no);
    g->obs_point.assign(obs_point, obs_point + no);
    g->obs_uv.assign(obs_uv, obs_uv + 2L * no);
    if (obs_info) {
        g->obs_info.assign(obs_info, obs_info + 4L * no);
    } else {
        g->obs_info.resize(4L * no);
        for (int e = 0; e < no; ++e) {
            g->obs_info[4L * e] = 1; g->obs_info[4L * e + 1] = 0; g->obs_info[4L * e + 2] = 0; g->obs_info[4L * e + 3] = 1;
        }
    }
    g->obs_K.assign(obs_K, obs_K + 4L * no);
    g->obs_loss.assign(obs_loss, obs_loss + no);
    g->obs_delta.assign(obs_delta, obs_delta + no);
    g->edges.resize(static_cast<size_t>(no));
    g->verbose = verbose;
    return g;
}

void zo_destroy(void* h) { delete static_cast<graph*>(h); }

int32_t zo_solve(void* h, int32_t iterations, int32_t use_rel) {
    return static_cast<graph*>(h)->solve(iterations, use_rel != 0);
}

double zo_current_chi(void* h, int32_t recompute) { return static_cast<graph*>(h)->current_chi(recompute != 0); }

void zo_get_state(void* h, double* poses, double* points) {
    graph* g = static_cast<graph*>(h);
    if (poses) std::memcpy(poses, g->pose.data(), g->pose.size() * sizeof(double));
    if (points) std::memcpy(points, g->point.data(), g->point.size() * sizeof(double));
}

void zo_set_state(void* h, const double* poses, const double* points) {
    graph* g = static_cast<graph*>(h);
    if (poses) std::memcpy(g->pose.data(), poses, g->pose.size() * sizeof(double));
    if (points) std::memcpy(g->point.data(), points, g->point.size() * sizeof(double));
}

// Residual, Jacobians and robust information of one edge at the current state,
// 27 doubles: r(2) Jp(12) Jl(6) drho(1) Omega'(4) chi2(1) robust_chi2(1).
void zo_edge_eval(void* h, int32_t e, double* out) {
    graph* g = static_cast<graph*>(h);
    g->compute_residual(e);
    g->compute_jacobians(e);
    const edge_state& s = g->edges[e];
    std::memcpy(out, s.r, 2 * sizeof(double));
    std::memcpy(out + 2, s.Jp, 12 * sizeof(double));
    std::memcpy(out + 14, s.Jl, 6 * sizeof(double));
    double drho, Wr[4];
    g->robust_info(e, drho, Wr);
    out[20] = drho;
    std::memcpy(out + 21, Wr, 4 * sizeof(double));
    out[25] = g->edge_chi2(e);
    out[26] = g->edge_robust_chi2(e);
}

// Reduced camera system at the current state with damping lambda: S (P*P),
// b_s (P), pose step (P) and point step (M).  Returns 1 if the Cholesky
// decompose and solve succeeded.
int32_t zo_reduced_system(void* h, double lambda, double* S, double* b_s, double* dx_pose, double* dx_point) {
    graph* g = static_cast<graph*>(h);
    for (int e = 0; e < g->no; ++e) g->compute_residual(e);
    g->set_ordering();
    g->make_hessian();
    const std::vector<double> hpp = g->h_pp, hll = g->h_ll;
    std::vector<double> dense, bs;
    g->reduced_system(lambda, dense, bs);
    if (S) std::memcpy(S, dense.data(), dense.size() * sizeof(double));
    if (b_s) std::memcpy(b_s, bs.data(), bs.size() * sizeof(double));
    std::vector<double> L(static_cast<size_t>(g->P) * g->P), xp(static_cast<size_t>(g->P), 0.0);
    const bool ok = decompose_cholesky(dense.data(), g->P, L.data()) && solve_cholesky(L.data(), bs.data(), g->P, xp.data());
    if (ok) {
        g->back_substitute(xp);
        if (dx_pose) std::memcpy(dx_pose, xp.data(), xp.size() * sizeof(double));
        if (dx_point) std::memcpy(dx_point, g->dx.data() + g->P, static_cast<size_t>(g->M) * sizeof(double));
    }
    g->h_pp = hpp;
    g->h_ll = hll;
    return ok ? 1 : 0;
}

// Normal-equation blocks of make_hessian at the current state: H_pp (np*36),
// H_ll (nl*9), b (6np + 3nl).
void zo_hessian(void* h, double* hpp, double* hll, double* b) {
    graph* g = static_cast<graph*>(h);
    for (int e = 0; e < g->no; ++e) g->compute_residual(e);
    g->set_ordering();
    g->make_hessian();
    if (hpp) std::memcpy(hpp, g->h_pp.data(), g->h_pp.size() * sizeof(double));
    if (hll) std::memcpy(hll, g->h_ll.data(), g->h_ll.size() * sizeof(double));
    if (b) std::memcpy(b, g->b.data(), g->b.size() * sizeof(double));
}

// Statistics of the last solve: iterations, attempts, accepted, rejected, failures.
void zo_stats(void* h, int32_t out[5], double* initial_chi2, double* lambda) {
    graph* g = static_cast<graph*>(h);
    out[0] = g->stat_iterations; out[1] = g->stat_attempts; out[2] = g->stat_accepted;
    out[3] = g->stat_rejected; out[4] = g->stat_failures;
    if (initial_chi2) *initial_chi2 = g->stat_initial_chi2;
    if (lambda) *lambda = g->lambda;
}

// The verbose lines the reference prints for the same solve (factor_graph.hpp:526,
// 581, 589, 786), accumulated since creation.  Returns the length.
int64_t zo_trace(void* h, char* buf, int64_t size) {
    graph* g = static_cast<graph*>(h);
    if (buf && size > 0) {
        const size_t n = std::min(static_cast<size_t>(size - 1), g->trace.size());
        std::memcpy(buf, g->trace.data(), n);
        buf[n] = 0;
    }
    return static_cast<int64_t>(g->trace.size());
}

// Standalone kernels for known-answer tests.
// map::cull's reprojection test (map.hpp:287-312) for every point: over its
// observations in insertion order, mapped = R X + t with R = so3::get_matrix of
// the pose quaternion (lie.hpp:188-204; the frame rotation set at map.hpp:243),
// the product in matrix.hpp's order (sum from zero, ascending k) and the
// translation added after; `threshold` for a point pinhole::project rejects
// (z < 1e-12, camera.hpp:97-99), otherwise float(sqrt(|uv - projection|^2)) with
// the squared length summed from zero (matrix.hpp:177-183); the error summed in
// binary32, the mean in binary32; erase without observations or when the mean
// reaches the threshold.  Returns the number of points flagged.
int32_t zo_reprojection_cull(void* h, float threshold, float* mean, uint8_t* erase) {
    const graph* g = static_cast<const graph*>(h);
    std::vector<float> err(static_cast<size_t>(g->nl), 0.0f);
    std::vector<int64_t> count(static_cast<size_t>(g->nl), 0);
    for (int32_t e = 0; e < g->no; ++e) {
        const int32_t l = g->obs_point[e];
        const double* P = &g->pose[7L * g->obs_pose[e]];
        const double* X = &g->point[3L * l];
        const double w = P[6], x = P[3], y = P[4], z = P[5];
        const double tx = 2.0 * x, ty = 2.0 * y, tz = 2.0 * z;
        const double txx = tx * x, txy = tx * y, txz = tx * z, txw = tx * w;
        const double tyy = ty * y, tyz = ty * z, tyw = ty * w;
        const double tzz = tz * z, tzw = tz * w;
        const double R[9] = {1.0 - (tyy + tzz), txy - tzw, txz + tyw,
                             txy + tzw, 1.0 - (txx + tzz), tyz - txw,
                             txz - tyw, tyz + txw, 1.0 - (txx + tyy)};
        double m[3];
        for (int i = 0; i < 3; ++i) {
            double sum = 0;
            for (int k = 0; k < 3; ++k) sum += R[3 * i + k] * X[k];
            m[i] = sum + P[i];
        }
        if (m[2] < 1.0e-12) {
            err[l] += threshold;
        } else {
            const double* K = &g->obs_K[4L * e];
            const double iz = 1.0 / m[2];
            const double u = K[0] * (m[0] * iz) + K[2];
            const double v = K[1] * (m[1] * iz) + K[3];
            const double d[2] = {g->obs_uv[2L * e] - u, g->obs_uv[2L * e + 1] - v};
            double l2 = 0;
            for (int i = 0; i < 2; ++i) l2 += d[i] * d[i];
            err[l] += static_cast<float>(std::sqrt(l2));
        }
        ++count[l];
    }
    int32_t flagged = 0;
    for (int32_t l = 0; l < g->nl; ++l) {
        if (count[l] == 0) {
            mean[l] = INFINITY;
            erase[l] = 1;
        } else {
            const float mm = err[l] / static_cast<float>(count[l]);
            mean[l] = mm;
            erase[l] = (mm >= threshold) ? 1 : 0;
        }
        flagged += erase[l];
    }
    return flagged;
}

// feature::distance<256> (feature.hpp:1134-1188): the number of differing bits
// of two 32-byte descriptors, as four 64-bit popcounts of the XOR.
int32_t zo_hamming256(const uint8_t* a, const uint8_t* b) {
    int32_t d = 0;
    for (int w = 0; w < 4; ++w) {
        uint64_t x, y;
        std::memcpy(&x, a + 8 * w, 8);
        std::memcpy(&y, b + 8 * w, 8);
        d += __builtin_popcountll(x ^ y);
    }
    return d;
}

// The projection search of slam.hpp:341-384, literally: landmarks in the given
// order; mapped = R X + t (products summed from zero in ascending k, the
// translation added after); pinhole::project (camera.hpp:97-112: no projection
// below z = 1e-12); the 10-pixel border against cols - 11 and rows - 11 in the
// reference's size_t arithmetic (slam.hpp:348-350); landmarks the frame already
// observes skipped (:352-362); then every keypoint in index order: skipped when
// its squared distance to the projection, summed from zero, is above 4
// (:365-368), when its previous-frame index is taken (:369-373), or when its
// descriptor is 64 or more bits from the landmark's (:374-376); otherwise a match,
// and its previous-frame index is taken (:377-379).  Writes at most max_matches
// (landmark, keypoint) pairs in that order; returns the number of matches.
int32_t zo_match_by_projection(const double* R, const double* t, const double* K, int64_t cols, int64_t rows,
                               int32_t n_landmarks, const double* xyz, const uint8_t* lm_desc, const uint8_t* observed,
                               int32_t n_keypoints, const double* kp_xy, const uint8_t* kp_desc, const int32_t* key,
                               int32_t n_keys, const uint8_t* key_taken, int32_t max_matches, int32_t* match_lm,
                               int32_t* match_kp) {
    std::vector<uint8_t> taken(key_taken, key_taken + n_keys);
    const double u_max = static_cast<double>(static_cast<uint64_t>(cols) - 11u);
    const double v_max = static_cast<double>(static_cast<uint64_t>(rows) - 11u);
    int32_t n = 0;
    for (int32_t o = 0; o < n_landmarks; ++o) {
        const double* X = &xyz[3L * o];
        double m[3];
        for (int i = 0; i < 3; ++i) {
            double sum = 0;
            for (int k = 0; k < 3; ++k) sum += R[3 * i + k] * X[k];
            m[i] = sum + t[i];
        }
        if (m[2] < 1.0e-12) continue;
        const double iz = 1.0 / m[2];
        const double ux = m[0] * iz, uy = m[1] * iz;
        const double u = K[0] * ux + K[2], v = K[1] * uy + K[3];
        if ((u < 10) || (u > u_max) || (v < 10) || (v > v_max)) continue;
        if (observed[o]) continue;
        for (int32_t i = 0; i < n_keypoints; ++i) {
            const double d[2] = {kp_xy[2L * i] - u, kp_xy[2L * i + 1] - v};
            double l2 = 0;
            for (int k = 0; k < 2; ++k) l2 += d[k] * d[k];
            if (l2 > 4) continue;
            if (taken[key[i]]) continue;
            if (zo_hamming256(&lm_desc[32L * o], &kp_desc[32L * i]) < 64) {
                if (n < max_matches) {
                    match_lm[n] = o;
                    match_kp[n] = i;
                }
                ++n;
                taken[key[i]] = 1;
            }
        }
    }
    return n;
}

int32_t zo_cholesky(const double* A, int32_t n, double* L) { return decompose_cholesky(A, n, L) ? 1 : 0; }
int32_t zo_cholesky_solve(const double* L, const double* b, int32_t n, double* x) { return solve_cholesky(L, b, n, x) ? 1 : 0; }
int32_t zo_invert3(const double* A, double* out) { return invert3(A, out) ? 1 : 0; }
void zo_pose_plus(double* p, const double* d) { pose_plus(p, d); }
void zo_pose_apply(const double* p, const double* X, double* Pc) { pose_apply(p, X, Pc); }
void zo_so3_exp(const double* w, double* q) { so3_exp(w, q); }
void zo_so3_left_jacobian(const double* w, double* J) { so3_left_jacobian(w, J); }
void zo_loss(int32_t kind, double delta, double e2, double* rho) { loss_compute(kind, delta, e2, rho); }

}  // extern "C"
