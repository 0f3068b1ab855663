// Synthetic bundle-adjustment workloads (host code).
//
// Deterministic problems shaped like the engine's benchmark configurations:
// keyframe trajectories, landmarks observed by runs of consecutive keyframes,
// pinhole observations with Gaussian pixel noise and uniform outliers, and the
// perturbation of the initial state that the optimiser has to undo.  The random
// source is the PCG32 (XSH-RR) generator of the reference tests
// (tests/factor_graph.test.cpp:62-83), seeded per problem.  Every arithmetic step
// is IEEE binary64 and the library is built without contraction, so one seed gives
// the same bytes wherever the same libm is used.
#include "zsba.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

namespace {

struct pcg32 {
    uint64_t state;
    uint64_t increment;
    explicit pcg32(uint64_t seed)
        : state(0x853C49E6748FEA9Bull ^ (seed * 0x9E3779B97F4A7C15ull)),
          increment(0xDA3E39CB94B95BDBull) {
        for (int i = 0; i < 4; ++i) next();
    }
    uint32_t next() {
        const uint64_t previous = state;
        state = previous * 0x5851F42D4C957F2Dull + increment;
        const uint32_t xorshifted = static_cast<uint32_t>(((previous >> 18u) ^ previous) >> 27u);
        const uint32_t rotation = static_cast<uint32_t>(previous >> 59u);
        return (xorshifted >> rotation) | (xorshifted << ((0u - rotation) & 31u));
    }
    // Uniform in [0, 1) with 53 random bits.
    double uniform() {
        const uint64_t hi = next() >> 5;   // 27 bits
        const uint64_t lo = next() >> 6;   // 26 bits
        return static_cast<double>((hi << 26) | lo) * (1.0 / 9007199254740992.0);
    }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    int32_t integer(int32_t lo, int32_t hi) {  // inclusive
        const uint32_t span = static_cast<uint32_t>(hi - lo + 1);
        return lo + static_cast<int32_t>(next() % span);
    }
    // Box-Muller, one value per call.
    double normal() {
        const double u1 = 1.0 - uniform();  // (0, 1]
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct vec3 { double x, y, z; };
static vec3 add(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static vec3 sub(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static vec3 scale(vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static double dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static vec3 cross(vec3 a, vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static vec3 normalise(vec3 a) { return scale(a, 1.0 / std::sqrt(dot(a, a))); }

// Unit quaternion (w, x, y, z) and its rotation matrix.
struct quat { double w, x, y, z; };
static quat qmul(quat a, quat b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
            a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
static quat qnormalise(quat q) {
    const double s = 1.0 / std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    return {q.w * s, q.x * s, q.y * s, q.z * s};
}
static void qmatrix(quat q, double r[9]) {
    r[0] = 1 - 2 * (q.y * q.y + q.z * q.z); r[1] = 2 * (q.x * q.y - q.z * q.w); r[2] = 2 * (q.x * q.z + q.y * q.w);
    r[3] = 2 * (q.x * q.y + q.z * q.w); r[4] = 1 - 2 * (q.x * q.x + q.z * q.z); r[5] = 2 * (q.y * q.z - q.x * q.w);
    r[6] = 2 * (q.x * q.z - q.y * q.w); r[7] = 2 * (q.y * q.z + q.x * q.w); r[8] = 1 - 2 * (q.x * q.x + q.y * q.y);
}
// Quaternion of a rotation matrix (Shepperd's method: largest-component branch).
static quat qfrom_matrix(const double m[9]) {
    const double tr = m[0] + m[4] + m[8];
    quat q;
    if (tr > 0) {
        const double s = 2.0 * std::sqrt(tr + 1.0);
        q = {0.25 * s, (m[7] - m[5]) / s, (m[2] - m[6]) / s, (m[3] - m[1]) / s};
    } else if (m[0] > m[4] && m[0] > m[8]) {
        const double s = 2.0 * std::sqrt(1.0 + m[0] - m[4] - m[8]);
        q = {(m[7] - m[5]) / s, 0.25 * s, (m[1] + m[3]) / s, (m[2] + m[6]) / s};
    } else if (m[4] > m[8]) {
        const double s = 2.0 * std::sqrt(1.0 + m[4] - m[0] - m[8]);
        q = {(m[2] - m[6]) / s, (m[1] + m[3]) / s, 0.25 * s, (m[5] + m[7]) / s};
    } else {
        const double s = 2.0 * std::sqrt(1.0 + m[8] - m[0] - m[4]);
        q = {(m[3] - m[1]) / s, (m[2] + m[6]) / s, (m[5] + m[7]) / s, 0.25 * s};
    }
    return qnormalise(q);
}
static quat qaxis_angle(vec3 axis, double angle) {
    const double s = std::sin(0.5 * angle);
    return {std::cos(0.5 * angle), axis.x * s, axis.y * s, axis.z * s};
}

// World-to-camera transform.
struct pose { quat q; vec3 t; };
static vec3 transform(const pose& T, vec3 p) {
    double r[9];
    qmatrix(T.q, r);
    return {r[0] * p.x + r[1] * p.y + r[2] * p.z + T.t.x,
            r[3] * p.x + r[4] * p.y + r[5] * p.z + T.t.y,
            r[6] * p.x + r[7] * p.y + r[8] * p.z + T.t.z};
}
// Camera at centre c looking along `forward`, image y axis towards world +y.
static pose look_along(vec3 c, vec3 forward) {
    const vec3 z = normalise(forward);
    const vec3 x = normalise(cross(vec3{0.0, 1.0, 0.0}, z));
    const vec3 y = cross(z, x);
    const double r_cw[9] = {x.x, x.y, x.z, y.x, y.y, y.z, z.x, z.y, z.z};
    pose T;
    T.q = qfrom_matrix(r_cw);
    double r[9];
    qmatrix(T.q, r);
    T.t = {-(r[0] * c.x + r[1] * c.y + r[2] * c.z), -(r[3] * c.x + r[4] * c.y + r[5] * c.z),
           -(r[6] * c.x + r[7] * c.y + r[8] * c.z)};
    return T;
}
static vec3 random_direction(pcg32& rng) {
    for (;;) {
        const vec3 v = {rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)};
        const double n2 = dot(v, v);
        if (n2 > 1e-6 && n2 <= 1.0) return scale(v, 1.0 / std::sqrt(n2));
    }
}
// Left perturbation T <- exp(delta) * T with |omega| = rot, |upsilon| = trans.
// The translation increment is applied directly (V(omega) ~ I at 0.01 rad).
static pose perturb(const pose& T, double rot, double trans, pcg32& rng) {
    const quat dq = qaxis_angle(random_direction(rng), rot);
    const vec3 dt = scale(random_direction(rng), trans);
    pose out;
    out.q = qnormalise(qmul(dq, T.q));
    double r[9];
    qmatrix(dq, r);
    out.t = add(vec3{r[0] * T.t.x + r[1] * T.t.y + r[2] * T.t.z, r[3] * T.t.x + r[4] * T.t.y + r[5] * T.t.z,
                     r[6] * T.t.x + r[7] * T.t.y + r[8] * T.t.z},
                dt);
    return out;
}

}  // namespace

struct zsba_synth {
    zsba_synth_params params;
    std::vector<pose> truth_poses, poses;
    std::vector<vec3> truth_points, points;
    std::vector<int32_t> obs_pose, obs_point;
    std::vector<double> obs_uv;
    std::vector<uint8_t> outlier;
};

namespace {

bool project(const zsba_synth_params& p, const pose& T, vec3 X, double& u, double& v) {
    const vec3 c = transform(T, X);
    if (c.z < 0.1) return false;
    u = p.K[0] * (c.x / c.z) + p.K[2];
    v = p.K[1] * (c.y / c.z) + p.K[3];
    return u >= 0.0 && u < p.width && v >= 0.0 && v < p.height;
}

// Camera centre of a world-to-camera transform: -R^T t.
vec3 centre(const pose& T) {
    double r[9];
    qmatrix(T.q, r);
    return {-(r[0] * T.t.x + r[3] * T.t.y + r[6] * T.t.z), -(r[1] * T.t.x + r[4] * T.t.y + r[7] * T.t.z),
            -(r[2] * T.t.x + r[5] * T.t.y + r[8] * T.t.z)};
}

// The angle at X between the rays from the first and the last camera centre of
// the track is at least min_parallax (always true when the test is off).
bool parallax_ok(const zsba_synth& s, vec3 X, const std::vector<int32_t>& frames) {
    if (!(s.params.min_parallax > 0.0) || frames.size() < 2) return true;
    const vec3 a = normalise(sub(X, centre(s.truth_poses[static_cast<size_t>(frames.front())])));
    const vec3 b = normalise(sub(X, centre(s.truth_poses[static_cast<size_t>(frames.back())])));
    const double c = dot(a, b);
    return std::acos(c < 1.0 ? c : 1.0) >= s.params.min_parallax;
}

bool visible_in_run(const zsba_synth& s, vec3 X, const std::vector<int32_t>& frames) {
    double u, v;
    for (int32_t f : frames)
        if (!project(s.params, s.truth_poses[static_cast<size_t>(f)], X, u, v)) return false;
    return true;
}

void build_trajectory(zsba_synth& s, pcg32& rng) {
    const zsba_synth_params& p = s.params;
    const int32_t n = p.n_frames;
    s.truth_poses.resize(static_cast<size_t>(n));
    if (p.layout == 0) {
        // Keyframes on a circle of radius 2 m in the y = 0 plane, looking at the origin.
        for (int32_t i = 0; i < n; ++i) {
            const double phi = 6.283185307179586 * i / n;
            const vec3 c = {2.0 * std::cos(phi), 0.0, 2.0 * std::sin(phi)};
            s.truth_poses[static_cast<size_t>(i)] = look_along(c, scale(c, -1.0));
        }
    } else if (p.layout == 1) {
        // Forward random walk: 0.2 m steps along a slowly turning heading.
        vec3 c = {0.0, 0.0, 0.0};
        double heading = 0.0;
        for (int32_t i = 0; i < n; ++i) {
            if (i > 0) {
                heading += 0.02 * rng.normal();
                c = add(c, vec3{0.2 * std::sin(heading), 0.01 * rng.normal(), 0.2 * std::cos(heading)});
            }
            s.truth_poses[static_cast<size_t>(i)] = look_along(c, vec3{std::sin(heading), 0.0, std::cos(heading)});
        }
    } else {
        // Two laps around a loop of 100 m circumference (200 m in total), cameras
        // facing outwards; the second lap runs 0.3 m further out and 0.2 m lower.
        const int32_t lap = (n + 1) / 2;
        const double radius = 100.0 / 6.283185307179586;
        for (int32_t i = 0; i < n; ++i) {
            const int32_t k = i % lap;
            const bool second = i >= lap;
            const double phi = 6.283185307179586 * k / lap;
            const double r = radius + (second ? 0.3 : 0.0);
            const vec3 c = {r * std::cos(phi), second ? 0.2 : 0.0, r * std::sin(phi)};
            s.truth_poses[static_cast<size_t>(i)] = look_along(c, vec3{std::cos(phi), 0.0, std::sin(phi)});
        }
    }
}

int32_t draw_track_length(const zsba_synth_params& p, pcg32& rng) {
    const int32_t hi = p.track_max < p.n_frames ? p.track_max : p.n_frames;
    const int32_t lo = p.track_min < hi ? p.track_min : hi;
    if (!p.track_geometric) return rng.integer(lo, hi);
    // Geometric on [lo, hi] with mean ~track_mean (rejection of the tail).
    const double q = 1.0 - 1.0 / (p.track_mean - lo + 1.0);
    for (;;) {
        int32_t len = lo;
        while (rng.uniform() < q) ++len;
        if (len <= hi) return len;
    }
}

// Uniform point in the volume the configuration draws landmarks from.
vec3 draw_point(const zsba_synth& s, const std::vector<int32_t>& frames, pcg32& rng) {
    const zsba_synth_params& p = s.params;
    if (p.layout == 0) return {rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)};
    if (p.layout == 1) {
        // 30 x 10 x 30 m box ahead of the first keyframe of the run.
        const pose& T = s.truth_poses[static_cast<size_t>(frames.front())];
        double r[9];
        qmatrix(T.q, r);
        // Camera centre and heading in the world frame.
        const vec3 c = {-(r[0] * T.t.x + r[3] * T.t.y + r[6] * T.t.z), -(r[1] * T.t.x + r[4] * T.t.y + r[7] * T.t.z),
                        -(r[2] * T.t.x + r[5] * T.t.y + r[8] * T.t.z)};
        const vec3 fwd = {r[6], r[7], r[8]};
        const vec3 right = {r[0], r[1], r[2]};
        const double a = rng.uniform(-15, 15), b = rng.uniform(-5, 5), d = rng.uniform(2, 32);
        return add(add(add(c, scale(right, a)), vec3{0.0, b, 0.0}), scale(fwd, d));
    }
    // Annulus 5..25 m outside the loop, around the run's angular position.
    const int32_t lap = (p.n_frames + 1) / 2;
    const int32_t k0 = frames.front() % lap, k1 = frames.back() % lap;
    double phi0 = 6.283185307179586 * k0 / lap, phi1 = 6.283185307179586 * k1 / lap;
    if (phi1 < phi0) phi1 += 6.283185307179586;
    const double phi = rng.uniform(phi0 - 0.3, phi1 + 0.3);
    const double radius = 100.0 / 6.283185307179586 + rng.uniform(5, 25);
    return {radius * std::cos(phi), rng.uniform(-5, 5), radius * std::sin(phi)};
}

void generate(zsba_synth& s, uint64_t seed) {
    pcg32 rng(seed);
    const zsba_synth_params& p = s.params;
    build_trajectory(s, rng);
    const int32_t n = p.n_frames;
    std::vector<int32_t> run, other;
    s.truth_points.reserve(static_cast<size_t>(p.n_points));
    for (int32_t j = 0; j < p.n_points; ++j) {
        vec3 X{};
        run.clear();
        other.clear();
        if (p.layout == 0) {
            // Every keyframe that sees the point observes it.
            for (int attempt = 0;; ++attempt) {
                X = draw_point(s, run, rng);
                run.clear();
                double u, v;
                for (int32_t f = 0; f < n; ++f)
                    if (project(p, s.truth_poses[static_cast<size_t>(f)], X, u, v)) run.push_back(f);
                if (!run.empty() || attempt > 100) break;
            }
        } else {
            for (int attempt = 0;; ++attempt) {
                const int32_t len = draw_track_length(p, rng);
                const int32_t start = rng.integer(0, n - len);
                run.clear();
                for (int32_t f = start; f < start + len; ++f) run.push_back(f);
                bool found = false;
                for (int tries = 0; tries < 200; ++tries) {
                    X = draw_point(s, run, rng);
                    if (visible_in_run(s, X, run) && parallax_ok(s, X, run)) { found = true; break; }
                }
                if (found || attempt > 1000) break;
            }
            if (p.layout == 2 && p.revisit_rate > 0.0 && rng.uniform() < p.revisit_rate) {
                const int32_t lap = (n + 1) / 2;
                for (int32_t f : run) {
                    const int32_t g = f < lap ? f + lap : f - lap;
                    if (g >= 0 && g < n) other.push_back(g);
                }
                if (!visible_in_run(s, X, other)) other.clear();
            }
        }
        s.truth_points.push_back(X);
        std::vector<int32_t> frames = run;
        frames.insert(frames.end(), other.begin(), other.end());
        for (int32_t f : frames) {
            double u, v;
            const vec3 c = transform(s.truth_poses[static_cast<size_t>(f)], X);
            u = p.K[0] * (c.x / c.z) + p.K[2];
            v = p.K[1] * (c.y / c.z) + p.K[3];
            uint8_t is_outlier = 0;
            if (p.outlier_rate > 0.0 && rng.uniform() < p.outlier_rate) {
                u = rng.uniform(0, p.width);
                v = rng.uniform(0, p.height);
                is_outlier = 1;
            } else {
                u += p.noise_px * rng.normal();
                v += p.noise_px * rng.normal();
            }
            s.obs_pose.push_back(f);
            s.obs_point.push_back(j);
            s.obs_uv.push_back(u);
            s.obs_uv.push_back(v);
            s.outlier.push_back(is_outlier);
        }
    }
    // Perturb the initial state: non-gauge poses by a left SE(3) increment, all
    // points by isotropic Gaussian noise.
    s.poses = s.truth_poses;
    for (int32_t i = p.n_fixed; i < n; ++i)
        s.poses[static_cast<size_t>(i)] = perturb(s.truth_poses[static_cast<size_t>(i)], p.rot_perturb, p.trans_perturb, rng);
    s.points = s.truth_points;
    for (vec3& X : s.points) {
        X.x += p.point_perturb * rng.normal();
        X.y += p.point_perturb * rng.normal();
        X.z += p.point_perturb * rng.normal();
    }
}

void write_pose(const pose& T, double* out) {
    out[0] = T.t.x; out[1] = T.t.y; out[2] = T.t.z;
    out[3] = T.q.x; out[4] = T.q.y; out[5] = T.q.z; out[6] = T.q.w;
}

}  // namespace

extern "C" {

int32_t zsba_synth_default_params(int32_t config, zsba_synth_params* p) {
    if (!p || config < 0 || config > 4) return ZSBA_EINVAL;
    std::memset(p, 0, sizeof(*p));
    p->K[0] = 525.0; p->K[1] = 525.0; p->K[2] = 320.0; p->K[3] = 240.0;
    p->width = 640; p->height = 480;
    p->rot_perturb = 0.01; p->trans_perturb = 0.02; p->point_perturb = 0.05;
    p->n_fixed = 2;
    p->track_mean = 8.0;
    switch (config) {
    case 0:  // 10 keyframes on a circle, 500 points in [-1,1]^3, 0.5 px noise, no loss
        p->layout = 0; p->n_frames = 10; p->n_points = 500; p->track_min = 1; p->track_max = 10;
        p->noise_px = 0.5; p->loss_kind = ZSBA_LOSS_NONE; p->loss_delta = 0.0;
        break;
    case 1:  // 100 keyframes forward walk, 20k points, tracks 4..12, 1 px, 5% outliers, Huber 2
        p->layout = 1; p->n_frames = 100; p->n_points = 20000; p->track_min = 4; p->track_max = 12;
        p->noise_px = 1.0; p->outlier_rate = 0.05; p->loss_kind = ZSBA_LOSS_HUBER; p->loss_delta = 2.0;
        break;
    case 2:  // one local window of the batched workload (sizes drawn per window by the caller)
        p->layout = 1; p->n_frames = 15; p->n_points = 1500; p->track_min = 2; p->track_max = 15;
        p->noise_px = 1.0; p->outlier_rate = 0.05; p->loss_kind = ZSBA_LOSS_HUBER; p->loss_delta = 2.0;
        break;
    case 3:  // 50 keyframes, 5k points, ~40k observations, no outliers, no loss; landmarks
             // seen with at least 2 degrees of parallax over their track, so that each
             // has a finite optimum to differentiate (the focus of expansion has none)
        p->layout = 1; p->n_frames = 50; p->n_points = 5000; p->track_min = 4; p->track_max = 12;
        p->noise_px = 1.0; p->loss_kind = ZSBA_LOSS_NONE; p->loss_delta = 0.0;
        p->min_parallax = 2.0 * 3.14159265358979323846 / 180.0;
        break;
    case 4:  // 1000 keyframes on a 2-lap loop, 200k points, geometric tracks 3..40
        p->layout = 2; p->n_frames = 1000; p->n_points = 200000; p->track_min = 3; p->track_max = 40;
        p->track_geometric = 1; p->track_mean = 8.5; p->revisit_rate = 0.2;
        p->noise_px = 1.0; p->outlier_rate = 0.05; p->loss_kind = ZSBA_LOSS_HUBER; p->loss_delta = 2.0;
        break;
    }
    return 0;
}

zsba_synth* zsba_synth_create(const zsba_synth_params* params, uint64_t seed) {
    if (!params || params->n_frames <= 0 || params->n_points < 0 || params->layout < 0 || params->layout > 2)
        return nullptr;
    zsba_synth* s = new (std::nothrow) zsba_synth();
    if (!s) return nullptr;
    s->params = *params;
    if (s->params.width <= 0) s->params.width = 640;
    if (s->params.height <= 0) s->params.height = 480;
    generate(*s, seed);
    return s;
}

void zsba_synth_destroy(zsba_synth* s) { delete s; }

void zsba_synth_sizes(const zsba_synth* s, int64_t sizes[3]) {
    sizes[0] = static_cast<int64_t>(s->poses.size());
    sizes[1] = static_cast<int64_t>(s->points.size());
    sizes[2] = static_cast<int64_t>(s->obs_pose.size());
}

void zsba_synth_copy(const zsba_synth* s, double* poses, uint8_t* pose_fixed, double* points,
                     uint8_t* point_fixed, int32_t* obs_pose, int32_t* obs_point, double* obs_uv,
                     double* true_poses, double* true_points, uint8_t* outlier) {
    const size_t np = s->poses.size(), nl = s->points.size(), no = s->obs_pose.size();
    for (size_t i = 0; i < np; ++i) {
        if (poses) write_pose(s->poses[i], poses + 7 * i);
        if (true_poses) write_pose(s->truth_poses[i], true_poses + 7 * i);
        if (pose_fixed) pose_fixed[i] = static_cast<uint8_t>(static_cast<int32_t>(i) < s->params.n_fixed);
    }
    for (size_t j = 0; j < nl; ++j) {
        if (points) { points[3 * j] = s->points[j].x; points[3 * j + 1] = s->points[j].y; points[3 * j + 2] = s->points[j].z; }
        if (true_points) {
            true_points[3 * j] = s->truth_points[j].x; true_points[3 * j + 1] = s->truth_points[j].y;
            true_points[3 * j + 2] = s->truth_points[j].z;
        }
        if (point_fixed) point_fixed[j] = 0;
    }
    if (obs_pose) std::memcpy(obs_pose, s->obs_pose.data(), no * sizeof(int32_t));
    if (obs_point) std::memcpy(obs_point, s->obs_point.data(), no * sizeof(int32_t));
    if (obs_uv) std::memcpy(obs_uv, s->obs_uv.data(), no * 2 * sizeof(double));
    if (outlier) std::memcpy(outlier, s->outlier.data(), no);
}

}  // extern "C"
