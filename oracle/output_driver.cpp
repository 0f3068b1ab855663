// =============================================================================
// TEST INFRASTRUCTURE — the reference front end's output writers
// save_trajectory_as_txt and save_trajectory_and_map_as_ply (source/main.cpp:
// 132-289, read from the reference checkout at build time with its main()
// renamed; nothing of it is copied here) on a map filled from a scenario file of
// oracle/map_driver.cpp's format.  Landmark colours are set from the id,
// colour = (1.2 (id mod 7) / 6 - 0.1, 1.2 (id mod 11) / 10 - 0.1,
// 1.2 (id mod 13) / 12 - 0.1), so that clamping is exercised.  Built by
// oracle/Makefile into oracle/_ref/ only.
//
// Usage: output_driver <scenario> <out.tum> <out.ply> <order> <width> <height>
// <order> receives int32 n_landmarks, the landmark ids in the map's iteration
// order, int32 n_frames, the frame ids in the map's iteration order.
// =============================================================================
#define main zeroslam_reference_main
#include "main.cpp"
#undef main

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

template <class T>
bool get(std::FILE* f, T* v, size_t n = 1) {
    return std::fread(v, sizeof(T), n, f) == n;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 7) return 2;
    std::FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int32_t head[4];
    double K[4];
    if (!get(in, head, 4) || !get(in, K, 4)) return 2;
    const camera::pinhole cam(K, 4);
    map::map m;
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
        m.add_frame(f);
    }
    for (int32_t i = 0; i < head[1]; ++i) {
        int32_t id;
        double x[3];
        if (!get(in, &id) || !get(in, x, 3)) return 2;
        landmark::point p;
        p.id = id;
        for (int k = 0; k < 3; ++k) p.location[k] = x[k];
        p.colour[0] = 1.2 * (id % 7) / 6.0 - 0.1;
        p.colour[1] = 1.2 * (id % 11) / 10.0 - 0.1;
        p.colour[2] = 1.2 * (id % 13) / 12.0 - 0.1;
        m.add_landmark(p);
    }
    std::fclose(in);
    if (!save_trajectory_as_txt(argv[2], m)) return 3;
    if (!save_trajectory_and_map_as_ply(argv[3], std::atoi(argv[5]), std::atoi(argv[6]), m)) return 3;
    std::FILE* order = std::fopen(argv[4], "wb");
    if (!order) return 3;
    const int32_t nl = static_cast<int32_t>(m.landmarks.size()), nf = static_cast<int32_t>(m.frames.size());
    std::fwrite(&nl, 4, 1, order);
    for (const auto& kv : m.landmarks) std::fwrite(&kv.first, 4, 1, order);
    std::fwrite(&nf, 4, 1, order);
    for (const auto& kv : m.frames) std::fwrite(&kv.first, 4, 1, order);
    std::fclose(order);
    return 0;
}
