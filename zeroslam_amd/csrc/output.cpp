// Output formats over the engine's results, host code: the map's frame pose
// from an optimised pose vertex (map.hpp:239-244), the TUM trajectory
// (main.cpp:132-156) and the PLY map with camera frusta (main.cpp:158-289) in
// the reference's exact text and byte formats, and the absolute trajectory
// error of tools/trajectory_alignment (main.cpp:195-276 with
// absolute_orientation.hpp).  The arithmetic follows the reference's operation
// order (products summed from zero in ascending index, left-to-right sums), so
// the files are byte-identical to the reference's for the same frames and
// landmarks.  The alignment's 3x3 singular value decomposition is this file's
// own one-sided Jacobi iteration; its rotation and scale are unique for a
// covariance of rank 2 or more, so the statistics agree with the reference tool
// to rounding.
#include "zsba.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <utility>
#include <vector>

namespace {

// so3(w, x, y, z).get_matrix(), lie.hpp:188-204.
void quat_to_matrix(double w, double x, double y, double z, double R[9]) {
    const double two_x = 2.0 * x, two_y = 2.0 * y, two_z = 2.0 * z;
    const double two_x_x = two_x * x, two_x_y = two_x * y, two_x_z = two_x * z, two_x_w = two_x * w;
    const double two_y_y = two_y * y, two_y_z = two_y * z, two_y_w = two_y * w;
    const double two_z_z = two_z * z, two_z_w = two_z * w;
    R[0] = 1.0 - (two_y_y + two_z_z); R[1] = two_x_y - two_z_w;         R[2] = two_x_z + two_y_w;
    R[3] = two_x_y + two_z_w;         R[4] = 1.0 - (two_x_x + two_z_z); R[5] = two_y_z - two_x_w;
    R[6] = two_x_z - two_y_w;         R[7] = two_y_z + two_x_w;         R[8] = 1.0 - (two_x_x + two_y_y);
}

// so3(rotation matrix), lie.hpp:160-182: q = (w, x, y, z).
void matrix_to_quat(const double R[9], double q[4]) {
    const double trace = R[0] + R[4] + R[8];
    if (trace > 0) {
        const double s = std::sqrt(trace + 1);
        const double inv = 0.5 / s;
        q[0] = 0.5 * s;
        q[1] = (R[7] - R[5]) * inv;
        q[2] = (R[2] - R[6]) * inv;
        q[3] = (R[3] - R[1]) * inv;
        return;
    }
    const int i = (R[0] < R[4]) ? (1 + (R[4] < R[8])) : (2 * (R[0] < R[8]));
    const int j = (i + 1) % 3;
    const int k = (j + 1) % 3;
    const double s = std::sqrt(R[3 * i + i] - R[3 * j + j] - R[3 * k + k] + 1.0);
    const double inv = 0.5 / s;
    q[0] = (R[3 * j + k] - R[3 * k + j]) * inv;
    q[1 + i] = 0.5 * s;
    q[1 + j] = (R[3 * i + j] + R[3 * j + i]) * inv;
    q[1 + k] = (R[3 * i + k] + R[3 * k + i]) * inv;
}

// (-R^T) v + add (matrix product summed from zero in ascending k, then the sum).
void neg_rt_times(const double R[9], const double v[3], double out[3]) {
    for (int r = 0; r < 3; ++r) {
        double sum = 0;
        for (int k = 0; k < 3; ++k) sum += (-R[3 * k + r]) * v[k];
        out[r] = sum;
    }
}

void write_vertex(std::FILE* f, float x, float y, float z, unsigned char r, unsigned char g, unsigned char b) {
    std::fwrite(&x, sizeof(float), 1, f);
    std::fwrite(&y, sizeof(float), 1, f);
    std::fwrite(&z, sizeof(float), 1, f);
    std::fwrite(&r, 1, 1, f);
    std::fwrite(&g, 1, 1, f);
    std::fwrite(&b, 1, 1, f);
}

// math::max(0.0, math::min(c * 255.0, 255.0)) as an unsigned char
// (math.hpp:230-237: min = rhs < lhs ? rhs : lhs, max = lhs < rhs ? rhs : lhs).
unsigned char colour_byte(double c) {
    const double a = c * 255.0;
    const double lo = (255.0 < a) ? 255.0 : a;
    const double v = (0.0 < lo) ? lo : 0.0;
    return static_cast<unsigned char>(v);
}

// One-sided Jacobi SVD of a 3x3 matrix A = U diag(s) V^T, s descending.
void svd3(const double A[3][3], double U[3][3], double s[3], double V[3][3]) {
    double B[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            B[r][c] = A[r][c];
            V[r][c] = r == c ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int r = 0; r < 3; ++r) {
                    alpha += B[r][p] * B[r][p];
                    beta += B[r][q] * B[r][q];
                    gamma += B[r][p] * B[r][q];
                }
                if (gamma == 0.0) continue;
                off = std::max(off, std::fabs(gamma) / std::sqrt(alpha * beta));
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                for (int r = 0; r < 3; ++r) {
                    const double bp = B[r][p], bq = B[r][q];
                    B[r][p] = c * bp - sn * bq;
                    B[r][q] = sn * bp + c * bq;
                    const double vp = V[r][p], vq = V[r][q];
                    V[r][p] = c * vp - sn * vq;
                    V[r][q] = sn * vp + c * vq;
                }
            }
        if (off < 1e-15) break;
    }
    int order[3] = {0, 1, 2};
    double norm[3];
    for (int c = 0; c < 3; ++c) norm[c] = std::sqrt(B[0][c] * B[0][c] + B[1][c] * B[1][c] + B[2][c] * B[2][c]);
    std::sort(order, order + 3, [&](int a, int b) { return norm[a] > norm[b]; });
    double Vs[3][3], Bs[3][3];
    for (int c = 0; c < 3; ++c) {
        s[c] = norm[order[c]];
        for (int r = 0; r < 3; ++r) {
            Vs[r][c] = V[r][order[c]];
            Bs[r][c] = B[r][order[c]];
        }
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) V[r][c] = Vs[r][c];
    for (int c = 0; c < 2; ++c)
        for (int r = 0; r < 3; ++r) U[r][c] = s[c] > 0.0 ? Bs[r][c] / s[c] : (r == c ? 1.0 : 0.0);
    // third left vector: from the column when it carries information, else the
    // cross product of the first two (a covariance of rank 2)
    if (s[2] > 1e-12 * s[0]) {
        for (int r = 0; r < 3; ++r) U[r][2] = Bs[r][2] / s[2];
    } else {
        U[0][2] = U[1][0] * U[2][1] - U[2][0] * U[1][1];
        U[1][2] = U[2][0] * U[0][1] - U[0][0] * U[2][1];
        U[2][2] = U[0][0] * U[1][1] - U[1][0] * U[0][1];
    }
}

}  // namespace

extern "C" {

int32_t zsba_pose_to_frame(const double pose[7], double rotation[9], double translation[3]) {
    if (!pose || !rotation || !translation) return ZSBA_EINVAL;
    // lie::se3(lie::so3(qw, qx, qy, qz), t): rotation().get_matrix(), translation()
    quat_to_matrix(pose[6], pose[3], pose[4], pose[5], rotation);
    for (int i = 0; i < 3; ++i) translation[i] = pose[i];
    return 0;
}

int32_t zsba_write_trajectory_tum(const char* path, int32_t n_frames, const int32_t* frame_ids, const double* rotation,
                                  const double* translation) {
    if (!path || n_frames < 0 || (n_frames > 0 && (!frame_ids || !rotation || !translation))) return ZSBA_EINVAL;
    std::vector<std::pair<int32_t, int32_t>> order;  // (id, index): the frames in ascending id (std::map)
    for (int32_t i = 0; i < n_frames; ++i) order.push_back({frame_ids[i], i});
    std::sort(order.begin(), order.end());
    for (size_t i = 1; i < order.size(); ++i)
        if (order[i].first == order[i - 1].first) return ZSBA_EINVAL;  // map keys are unique
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return ZSBA_EINVAL;
    for (const auto& [id, i] : order) {
        const double* R = rotation + 9L * i;
        double centre[3], q[4];
        neg_rt_times(R, translation + 3L * i, centre);
        matrix_to_quat(R, q);
        std::fprintf(f, "%d %f %f %f %f %f %f %f\n", id, centre[0], centre[1], centre[2], q[1], q[2], q[3], q[0]);
    }
    std::fclose(f);
    return 0;
}

int32_t zsba_write_map_ply(const char* path, int32_t image_width, int32_t image_height, int32_t n_landmarks,
                           const double* landmark_xyz, const double* landmark_colour, int32_t n_frames,
                           const int32_t* frame_ids, const double* rotation, const double* translation,
                           const double K[4]) {
    if (!path || n_landmarks < 0 || n_frames < 0 || (n_landmarks > 0 && !landmark_xyz) ||
        (n_frames > 0 && (!frame_ids || !rotation || !translation || !K)))
        return ZSBA_EINVAL;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return ZSBA_EINVAL;
    const size_t nl = static_cast<size_t>(n_landmarks), nc = static_cast<size_t>(n_frames);
    std::fprintf(f, "ply\n");
    std::fprintf(f, "format binary_little_endian 1.0\n");
    std::fprintf(f, "comment Created using ZeroSLAM by Geoffrey Daniels\n");
    std::fprintf(f, "element vertex %zu\n", nl + nc * 5);
    std::fprintf(f, "property float x\n");
    std::fprintf(f, "property float y\n");
    std::fprintf(f, "property float z\n");
    std::fprintf(f, "property uchar red\n");
    std::fprintf(f, "property uchar green\n");
    std::fprintf(f, "property uchar blue\n");
    std::fprintf(f, "element edge %zu\n", nc * 8);
    std::fprintf(f, "property int vertex1\n");
    std::fprintf(f, "property int vertex2\n");
    std::fprintf(f, "end_header\n");
    for (size_t l = 0; l < nl; ++l) {
        const double* p = landmark_xyz + 3 * l;
        const double zero[3] = {0.0, 0.0, 0.0};
        const double* c = landmark_colour ? landmark_colour + 3 * l : zero;
        write_vertex(f, static_cast<float>(p[0]), static_cast<float>(p[1]), static_cast<float>(p[2]), colour_byte(c[0]),
                     colour_byte(c[1]), colour_byte(c[2]));
    }
    const double corners_px[4][2] = {{0, 0},
                                     {static_cast<double>(image_width), 0},
                                     {static_cast<double>(image_width), static_cast<double>(image_height)},
                                     {0, static_cast<double>(image_height)}};
    for (size_t i = 0; i < nc; ++i) {
        const double* R = rotation + 9 * i;
        // pinhole::unproject (camera.hpp:147-162), times the frustum scale 1
        double corner[4][3];
        for (int c = 0; c < 4; ++c) {
            corner[c][0] = (corners_px[c][0] - K[2]) / K[0] * 1.0;
            corner[c][1] = (corners_px[c][1] - K[3]) / K[1] * 1.0;
            corner[c][2] = 1.0 * 1.0;
        }
        double centre[3];
        neg_rt_times(R, translation + 3 * i, centre);
        // the camera colour: 255 * (id / frame count) in binary32, truncated
        const float shade = 255.0f * (static_cast<float>(frame_ids[i]) / static_cast<float>(nc));
        const unsigned char g = static_cast<unsigned char>(static_cast<int>(shade));
        write_vertex(f, static_cast<float>(centre[0]), static_cast<float>(centre[1]), static_cast<float>(centre[2]), 0, g,
                     255);
        for (int c = 0; c < 4; ++c) {
            double w[3];
            for (int r = 0; r < 3; ++r) {
                double sum = 0;
                for (int k = 0; k < 3; ++k) sum += R[3 * k + r] * corner[c][k];
                w[r] = sum + centre[r];
            }
            write_vertex(f, static_cast<float>(w[0]), static_cast<float>(w[1]), static_cast<float>(w[2]), 0, g, 255);
        }
    }
    for (size_t i = 0; i < nc; ++i) {
        const int base = static_cast<int>(nl + i * 5);
        for (int j = 0; j < 4; ++j) {
            const int e[2] = {base, base + 1 + j};
            std::fwrite(e, sizeof(int), 2, f);
        }
        for (int j = 0; j < 4; ++j) {
            const int e[2] = {base + 1 + j, base + 1 + ((j + 1) % 4)};
            std::fwrite(e, sizeof(int), 2, f);
        }
    }
    std::fclose(f);
    return 0;
}

int32_t zsba_trajectory_ate(int32_t n_poses, const double* ground_truth_xyz, const double* estimated_xyz,
                            int32_t first_pose_overlap, zsba_ate* out) {
    if (n_poses < 3 || !ground_truth_xyz || !estimated_xyz || !out) return ZSBA_EINVAL;
    const size_t n = static_cast<size_t>(n_poses);
    const double* g = ground_truth_xyz;
    const double* e = estimated_xyz;
    // centroids (absolute_orientation.hpp:26-73)
    double gc[3], ec[3];
    if (first_pose_overlap) {
        for (int a = 0; a < 3; ++a) {
            gc[a] = g[a];
            ec[a] = e[a];
        }
    } else {
        double gs[3] = {0, 0, 0}, es[3] = {0, 0, 0};
        for (size_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                gs[a] += g[3 * i + a];
                es[a] += e[3 * i + a];
            }
        for (int a = 0; a < 3; ++a) {
            gc[a] = gs[a] / static_cast<double>(n);
            ec[a] = es[a] / static_cast<double>(n);
        }
    }
    // covariance of the estimated against the ground-truth offsets (:75-113)
    double C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (size_t i = 0; i < n; ++i) {
        double de[3], dg[3];
        for (int a = 0; a < 3; ++a) {
            de[a] = e[3 * i + a] - ec[a];
            dg[a] = g[3 * i + a] - gc[a];
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) C[r][c] += de[r] * dg[c];
    }
    double U[3][3], s[3], V[3][3];
    svd3(C, U, s, V);
    if (!(std::isfinite(s[0]) && std::isfinite(s[1]) && std::isfinite(s[2]))) return ZSBA_EUNSUP;
    // rotation V D U^T with D fixing the determinant (:115-170)
    double VUt[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            VUt[r][c] = 0;
            for (int k = 0; k < 3; ++k) VUt[r][c] += V[r][k] * U[c][k];
        }
    const double det = VUt[0][0] * (VUt[1][1] * VUt[2][2] - VUt[2][1] * VUt[1][2]) -
                       VUt[0][1] * (VUt[1][0] * VUt[2][2] - VUt[1][2] * VUt[2][0]) +
                       VUt[0][2] * (VUt[1][0] * VUt[2][1] - VUt[1][1] * VUt[2][0]);
    const double d = det > 0.0 ? 1.0 : -1.0;
    double R[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            R[r][c] = 0;
            for (int k = 0; k < 3; ++k) R[r][c] += V[r][k] * (k == 2 ? d : 1.0) * U[c][k];
        }
    // scale: the sum of the singular values over the estimated spread (:172-237)
    double var = 0;
    for (size_t i = 0; i < n; ++i) {
        const double dx = e[3 * i] - ec[0], dy = e[3 * i + 1] - ec[1], dz = e[3 * i + 2] - ec[2];
        var += dx * dx + dy * dy + dz * dz;
    }
    const double scale = (var > 0.0) ? ((s[0] + s[1] + s[2]) / var) : 1.0;
    double t[3];
    for (int r = 0; r < 3; ++r) t[r] = gc[r] - scale * (R[r][0] * ec[0] + R[r][1] * ec[1] + R[r][2] * ec[2]);
    // aligned trajectory and the error statistics (main.cpp:234-276)
    std::vector<double> aligned(3 * n), err(n);
    double sum = 0.0, sse = 0.0;
    for (size_t i = 0; i < n; ++i) {
        for (int r = 0; r < 3; ++r)
            aligned[3 * i + r] = (R[r][0] * e[3 * i] + R[r][1] * e[3 * i + 1] + R[r][2] * e[3 * i + 2]) * scale + t[r];
        const double dx = aligned[3 * i] - g[3 * i], dy = aligned[3 * i + 1] - g[3 * i + 1],
                     dz = aligned[3 * i + 2] - g[3 * i + 2];
        err[i] = std::sqrt(dx * dx + dy * dy + dz * dz);
        sum += err[i];
        sse += err[i] * err[i];
    }
    std::vector<double> sorted = err;
    std::sort(sorted.begin(), sorted.end());
    out->scale = scale;
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) out->rotation[3 * r + c] = R[r][c];
        out->translation[r] = t[r];
    }
    out->min = sorted.front();
    out->max = sorted.back();
    out->median = sorted[n / 2];
    out->mean = sum / static_cast<double>(n);
    out->rmse = std::sqrt(sse / static_cast<double>(n));
    out->sse = sse;
    out->std = std::sqrt(std::max(0.0, sse / static_cast<double>(n) - out->mean * out->mean));
    double de = 0.0, dg = 0.0;
    for (size_t i = 0; i + 1 < n; ++i) {
        const double ax = aligned[3 * (i + 1)] - aligned[3 * i], ay = aligned[3 * (i + 1) + 1] - aligned[3 * i + 1],
                     az = aligned[3 * (i + 1) + 2] - aligned[3 * i + 2];
        de += std::sqrt(ax * ax + ay * ay + az * az);
        const double gx = g[3 * (i + 1)] - g[3 * i], gy = g[3 * (i + 1) + 1] - g[3 * i + 1],
                     gz = g[3 * (i + 1) + 2] - g[3 * i + 2];
        dg += std::sqrt(gx * gx + gy * gy + gz * gz);
    }
    out->distance_estimated = de;
    out->distance_ground_truth = dg;
    out->poses = n_poses;
    return 0;
}

}  // extern "C"
