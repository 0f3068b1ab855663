// =============================================================================
// TEST INFRASTRUCTURE — runs ZeroSLAM's own map (source/map.hpp, read from the
// reference checkout at build time; nothing of it is copied here) on a scenario
// file, so that the same map::optimise / map::cull calls can be made once with
// the reference's factor graph (oracle/_ref/map_ref) and once with the engine's
// drop-in header include/dropin/factor_graph.hpp (oracle/_ref/map_zsba).  Only
// the map's public members are used: frames / landmarks / add_frame /
// add_landmark / add_observation / optimise / cull (map.hpp:38-323), with frames
// and landmarks filled as frame.hpp:36-47 and landmark.hpp:24-55 declare them.
// Built by oracle/Makefile into oracle/_ref/ only.
//
// Usage: map_driver <scenario> <out>
// Scenario (little-endian):
//   int32 n_frames, n_landmarks, n_observations, n_calls
//   double K[4]                                  pinhole fx, fy, cx, cy
//   n_frames x  { int32 id; double R[9] (row-major, camera = R world + t);
//                 double t[3]; int32 n_keypoints; float xy[2 n_keypoints] }
//   n_landmarks x { int32 id; double xyz[3] }
//   n_observations x { int32 frame_id, landmark_id, keypoint_index }   (in order)
//   n_calls x { int32 op (0 optimise, 1 cull), local_window, fix_landmarks,
//               rounds, use_relative_convergence }
// Output, after every call: int32 n_frames, then per frame in ascending id
// { int32 id; double R[9]; double t[3] }; int32 n_landmarks, then per landmark
// in ascending id { int32 id; double xyz[3] }.  map.hpp's own printf lines go
// to stdout; the wall time of each call goes to stderr ("call k op ms").  Exit
// status 0, or 2 for an unreadable scenario.
// =============================================================================
#include "map.hpp"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <vector>

namespace {

template <class T>
bool get(std::FILE* f, T* v, size_t n = 1) {
    return std::fread(v, sizeof(T), n, f) == n;
}
template <class T>
void put(std::FILE* f, const T* v, size_t n = 1) {
    std::fwrite(v, sizeof(T), n, f);
}

void dump(std::FILE* out, const map::map& m) {
    std::vector<int> fids, lids;
    for (const auto& kv : m.frames) fids.push_back(kv.first);
    for (const auto& kv : m.landmarks) lids.push_back(kv.first);
    std::sort(fids.begin(), fids.end());
    std::sort(lids.begin(), lids.end());
    const int32_t nf = static_cast<int32_t>(fids.size()), nl = static_cast<int32_t>(lids.size());
    put(out, &nf);
    for (int id : fids) {
        const frame::frame& f = m.frames.at(id);
        const int32_t i = id;
        double R[9], t[3];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) R[3 * r + c] = f.rotation[r][c];
            t[r] = f.translation[r];
        }
        put(out, &i);
        put(out, R, 9);
        put(out, t, 3);
    }
    put(out, &nl);
    for (int id : lids) {
        const landmark::point& p = m.landmarks.at(id);
        const int32_t i = id;
        const double x[3] = {p.location[0], p.location[1], p.location[2]};
        put(out, &i);
        put(out, x, 3);
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    std::FILE* in = std::fopen(argv[1], "rb");
    std::FILE* out = std::fopen(argv[2], "wb");
    if (!in || !out) return 2;
    int32_t head[4];
    double K[4];
    if (!get(in, head, 4) || !get(in, K, 4)) return 2;
    const camera::pinhole cam(K, 4);
    map::map m;
    std::vector<int32_t> frame_ids;
    int max_frame = -1, max_landmark = -1;
    for (int32_t i = 0; i < head[0]; ++i) {
        int32_t id, nk;
        double R[9], t[3];
        if (!get(in, &id) || !get(in, R, 9) || !get(in, t, 3) || !get(in, &nk) || nk < 0) return 2;
        std::vector<float> xy(2 * static_cast<size_t>(nk));
        if (nk && !get(in, xy.data(), xy.size())) return 2;
        frame::frame f;
        f.id = id;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) f.rotation[r][c] = R[3 * r + c];
            f.translation[r] = t[r];
        }
        f.camera = cam;
        std::vector<feature::point> kps(static_cast<size_t>(nk));
        for (int32_t k = 0; k < nk; ++k) kps[k] = feature::point{xy[2 * k], xy[2 * k + 1], 0.0f, 0.0f};
        f.keypoint_pyramid = {kps};
        m.add_frame(f);
        max_frame = std::max(max_frame, static_cast<int>(id));
    }
    for (int32_t i = 0; i < head[1]; ++i) {
        int32_t id;
        double x[3];
        if (!get(in, &id) || !get(in, x, 3)) return 2;
        landmark::point p;
        p.id = id;
        for (int k = 0; k < 3; ++k) p.location[k] = x[k];
        m.add_landmark(p);
        max_landmark = std::max(max_landmark, static_cast<int>(id));
    }
    // the map's id generators stand where the frames and landmarks left them
    frame::frame::id_generator = max_frame + 1;
    landmark::point::id_generator = max_landmark + 1;
    for (int32_t i = 0; i < head[2]; ++i) {
        int32_t o[3];
        if (!get(in, o, 3)) return 2;
        frame::frame f;
        f.id = o[0];
        landmark::point p;
        p.id = o[1];
        m.add_observation(f, p, static_cast<size_t>(o[2]));
    }
    for (int32_t c = 0; c < head[3]; ++c) {
        int32_t call[5];
        if (!get(in, call, 5)) return 2;
        const auto t0 = std::chrono::steady_clock::now();
        if (call[0] == 0) m.optimise(call[1], call[2] != 0, call[3], call[4] != 0);
        else m.cull();
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::fflush(stdout);
        std::fprintf(stderr, "call %d %d %.4f\n", static_cast<int>(c), static_cast<int>(call[0]), ms);
        dump(out, m);
    }
    std::fclose(in);
    std::fclose(out);
    return 0;
}
