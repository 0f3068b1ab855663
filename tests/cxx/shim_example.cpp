// A map::optimise-shaped caller (map.hpp:65-267) of the C++ drop-in
// zsba_factor_graph.hpp: vertices, edges and the loss are allocated by the caller
// and passed as non-owning pointers, parameters go in and out through the
// caller's matrix type.  The same problem is also solved directly through the C
// ABI; the two must agree bit for bit.  Exit status 0 = agree, 1 = differ,
// 2 = the engine could not solve (for example no GPU).
//
// Usage: shim_example [config] [iterations]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "zsba_factor_graph.hpp"

namespace fg = zsba::factor_graph;
using fg_matrix = fg::parameter_matrix;

int main(int argc, char** argv) {
    const int config = argc > 1 ? std::atoi(argv[1]) : 0;
    const int iterations = argc > 2 ? std::atoi(argv[2]) : 10;
    zsba_synth_params sp;
    if (zsba_synth_default_params(config, &sp) != 0) return 2;
    if (config == 1) {
        sp.n_frames = 20;
        sp.n_points = 1500;
    }
    zsba_synth* syn = zsba_synth_create(&sp, static_cast<uint64_t>(config));
    int64_t sizes[3];
    zsba_synth_sizes(syn, sizes);
    const int64_t np = sizes[0], nl = sizes[1], no = sizes[2];
    std::vector<double> poses(7 * np), points(3 * nl), uv(2 * no), tp(7 * np), tl(3 * nl);
    std::vector<uint8_t> pfix(np), lfix(nl), outlier(no);
    std::vector<int32_t> opose(no), opoint(no);
    zsba_synth_copy(syn, poses.data(), pfix.data(), points.data(), lfix.data(), opose.data(), opoint.data(), uv.data(),
                    tp.data(), tl.data(), outlier.data());
    zsba_synth_destroy(syn);

    // ---- through the drop-in, as map::optimise builds its graph ----
    fg::loss_function_base* loss = nullptr;
    if (sp.loss_kind == ZSBA_LOSS_HUBER) loss = new fg::loss_huber(sp.loss_delta);
    fg::factor_graph ba(false);
    std::vector<fg::vertex_base*> cams, lms;
    std::vector<fg::edge_base*> edges;
    for (int64_t i = 0; i < np; ++i) {
        fg::vertex_base* c = new fg::vertex_pose();
        c->set_parameters(fg_matrix(7, 1, &poses[7 * i]));
        c->set_fixed(pfix[i] != 0);
        if (!ba.add_vertex(c)) return 2;
        cams.push_back(c);
    }
    if (ba.add_vertex(cams[0])) return 1;  // duplicates are refused, factor_graph.hpp:386-393
    for (int64_t j = 0; j < nl; ++j) {
        fg::vertex_base* l = new fg::vertex_point_xyz();
        l->set_parameters(fg_matrix(3, 1, &points[3 * j]));
        l->set_fixed(lfix[j] != 0);
        l->set_marginalised(true);
        ba.add_vertex(l);
        lms.push_back(l);
    }
    const zsba::pinhole_camera camera_model(sp.K, 4);
    for (int64_t e = 0; e < no; ++e) {
        fg::edge_base* m = new fg::edge_reprojection<zsba::pinhole_camera>(camera_model);
        m->set_observation(fg_matrix(2, 1, &uv[2 * e]));
        m->add_vertex(cams[opose[e]]);
        m->add_vertex(lms[opoint[e]]);
        if (loss) m->set_loss_function(loss);
        ba.add_edge(m);
        edges.push_back(m);
    }
    const int accepted = ba.solve(iterations, false);
    if (!ba.last_error().empty()) {
        std::printf("shim solve failed: %s\n", ba.last_error().c_str());
        return 2;
    }
    const double chi = ba.get_current_chi();

    // ---- the same problem directly through the C ABI ----
    zsba_graph* g = zsba_create(nullptr);
    zsba_add_poses(g, static_cast<int32_t>(np), poses.data(), pfix.data());
    zsba_add_points(g, static_cast<int32_t>(nl), points.data(), lfix.data());
    zsba_add_observations(g, static_cast<int32_t>(no), opose.data(), opoint.data(), uv.data(), nullptr, sp.K,
                          sp.loss_kind, sp.loss_delta);
    const int accepted_c = zsba_solve(g, iterations, 0, ZSBA_PRECISION_FP64);
    std::vector<double> P(7 * np), X(3 * nl);
    zsba_get_poses(g, P.data());
    zsba_get_points(g, X.data());
    const double chi_c = zsba_current_chi(g, 1);
    zsba_destroy(g);

    bool same = accepted == accepted_c && std::memcmp(&chi, &chi_c, sizeof chi) == 0;
    for (int64_t i = 0; i < np && same; ++i) {
        const fg_matrix p = cams[i]->get_parameters();
        for (int k = 0; k < 7; ++k) same = same && std::memcmp(&p[k][0], &P[7 * i + k], sizeof(double)) == 0;
    }
    for (int64_t j = 0; j < nl && same; ++j) {
        const fg_matrix p = lms[j]->get_parameters();
        for (int k = 0; k < 3; ++k) same = same && std::memcmp(&p[k][0], &X[3 * j + k], sizeof(double)) == 0;
    }
    std::printf("shim: accepted %d chi2 %.17g | C ABI: accepted %d chi2 %.17g | identical %d\n", accepted, chi,
                accepted_c, chi_c, same ? 1 : 0);
    for (fg::edge_base* e : edges) delete e;
    for (fg::vertex_base* v : cams) delete v;
    for (fg::vertex_base* v : lms) delete v;
    delete loss;
    return same ? 0 : 1;
}
