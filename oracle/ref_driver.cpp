// =============================================================================
// TEST INFRASTRUCTURE — driver for the compiled reference (oracle/_ref).
//
// Builds ZeroSLAM's own factor graph (source/factor_graph.hpp, source/camera.hpp,
// read from the reference checkout at build time; nothing of it is copied here)
// from plain arrays, through the reference's public API only:
//   vertex_pose / vertex_point_xyz + set_parameters / set_fixed / set_marginalised
//   (factor_graph.hpp:95-137, 938-969), edge_reprojection<camera::pinhole> +
//   set_observation / set_information / set_loss_function / add_vertex
//   (factor_graph.hpp:198-276, 973-987), factor_graph::add_vertex / add_edge /
//   solve / get_current_chi (factor_graph.hpp:385-629).
// The C ABI mirrors oracle/zsba_oracle.cpp (zr_* for zo_*) so the tests drive both
// the same way.  Built by oracle/Makefile into oracle/_ref/ only.
// =============================================================================
#include "camera.hpp"
#include "factor_graph.hpp"

#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct ref_graph {
    factor_graph::factor_graph* ba = nullptr;
    std::vector<factor_graph::vertex_base*> poses, points;
    std::vector<factor_graph::edge_base*> edges;
    std::vector<factor_graph::loss_function_base*> losses;
    ~ref_graph() {
        delete ba;
        for (auto* v : poses) delete v;
        for (auto* v : points) delete v;
        for (auto* e : edges) delete e;
        for (auto* l : losses) delete l;
    }
};

factor_graph::loss_function_base* make_loss(int32_t kind, double delta) {
    switch (kind) {
    case 1: return new factor_graph::loss_trivial();
    case 2: return new factor_graph::loss_huber(delta);
    case 3: return new factor_graph::loss_cauchy(delta);
    case 4: return new factor_graph::loss_tukey(delta);
    default: return nullptr;
    }
}

}  // namespace

extern "C" {

void* zr_create(int32_t np, const double* poses, const uint8_t* pose_fixed, int32_t nl, const double* points,
                const uint8_t* point_fixed, int32_t no, const int32_t* obs_pose, const int32_t* obs_point,
                const double* obs_uv, const double* obs_info, const double* obs_K, const int32_t* obs_loss,
                const double* obs_delta, int32_t verbose) {
    ref_graph* g = new ref_graph();
    g->ba = new factor_graph::factor_graph(verbose != 0);
    for (int32_t i = 0; i < np; ++i) {
        factor_graph::vertex_base* v = new factor_graph::vertex_pose();
        v->set_parameters(matrix::matrix<double, 0, 0>(7, 1, poses + 7L * i));
        v->set_fixed(pose_fixed[i] != 0);
        g->ba->add_vertex(v);
        g->poses.push_back(v);
    }
    for (int32_t j = 0; j < nl; ++j) {
        factor_graph::vertex_base* v = new factor_graph::vertex_point_xyz();
        v->set_parameters(matrix::matrix<double, 0, 0>(3, 1, points + 3L * j));
        v->set_fixed(point_fixed[j] != 0);
        v->set_marginalised(true);
        g->ba->add_vertex(v);
        g->points.push_back(v);
    }
    for (int32_t e = 0; e < no; ++e) {
        camera::pinhole cam(obs_K + 4L * e, 4);
        factor_graph::edge_base* m = new factor_graph::edge_reprojection<camera::pinhole>(cam);
        m->set_observation(matrix::matrix<double, 0, 0>(2, 1, obs_uv + 2L * e));
        if (obs_info) m->set_information(matrix::matrix<double, 0, 0>(2, 2, obs_info + 4L * e));
        factor_graph::loss_function_base* loss = make_loss(obs_loss[e], obs_delta[e]);
        if (loss) {
            g->losses.push_back(loss);
            m->set_loss_function(loss);
        }
        m->add_vertex(g->poses[static_cast<size_t>(obs_pose[e])]);
        m->add_vertex(g->points[static_cast<size_t>(obs_point[e])]);
        g->ba->add_edge(m);
        g->edges.push_back(m);
    }
    return g;
}

void zr_destroy(void* h) { delete static_cast<ref_graph*>(h); }

int32_t zr_solve(void* h, int32_t iterations, int32_t use_rel) {
    return static_cast<ref_graph*>(h)->ba->solve(iterations, use_rel != 0);
}

double zr_current_chi(void* h, int32_t recompute) {
    return static_cast<ref_graph*>(h)->ba->get_current_chi(recompute != 0);
}

void zr_get_state(void* h, double* poses, double* points) {
    ref_graph* g = static_cast<ref_graph*>(h);
    if (poses)
        for (size_t i = 0; i < g->poses.size(); ++i)
            std::memcpy(poses + 7 * i, g->poses[i]->get_parameters().data(), 7 * sizeof(double));
    if (points)
        for (size_t j = 0; j < g->points.size(); ++j)
            std::memcpy(points + 3 * j, g->points[j]->get_parameters().data(), 3 * sizeof(double));
}

void zr_set_state(void* h, const double* poses, const double* points) {
    ref_graph* g = static_cast<ref_graph*>(h);
    if (poses)
        for (size_t i = 0; i < g->poses.size(); ++i)
            g->poses[i]->set_parameters(matrix::matrix<double, 0, 0>(7, 1, poses + 7 * i));
    if (points)
        for (size_t j = 0; j < g->points.size(); ++j)
            g->points[j]->set_parameters(matrix::matrix<double, 0, 0>(3, 1, points + 3 * j));
}

// r(2) Jp(12) Jl(6) drho(1) Omega'(4) chi2(1) robust_chi2(1) of one edge.
void zr_edge_eval(void* h, int32_t e, double* out) {
    ref_graph* g = static_cast<ref_graph*>(h);
    factor_graph::edge_base* m = g->edges[static_cast<size_t>(e)];
    m->compute_residual();
    m->compute_jacobians();
    std::memcpy(out, m->get_residual().data(), 2 * sizeof(double));
    std::memcpy(out + 2, m->get_jacobians()[0].data(), 12 * sizeof(double));
    std::memcpy(out + 14, m->get_jacobians()[1].data(), 6 * sizeof(double));
    double drho = 0;
    matrix::matrix<double, 0, 0> info(2, 2);
    m->robust_info(drho, info);
    out[20] = drho;
    std::memcpy(out + 21, info.data(), 4 * sizeof(double));
    out[25] = m->chi2();
    out[26] = m->robust_chi2();
}

}  // extern "C"
