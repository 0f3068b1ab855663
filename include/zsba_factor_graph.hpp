/* Header-only C++ drop-in for the reference's factor graph (source/factor_graph.hpp)
 * on top of the engine's C ABI (zsba.h).
 *
 * The classes carry the reference's names and member functions:
 * factor_graph::factor_graph (add_vertex / remove_vertex / add_edge / remove_edge /
 * solve / get_current_chi, factor_graph.hpp:341-629), vertex_pose /
 * vertex_point_xyz (938-969), edge_reprojection<camera_type> (973-1049) and the
 * losses loss_trivial / loss_huber / loss_cauchy / loss_tukey (851-934).  Ownership
 * is the reference's: the graph stores non-owning pointers, the caller allocates
 * and deletes vertices, edges and losses.  solve() sends the vertices and edges in
 * insertion order to one engine handle kept for the life of the graph (so repeated
 * solves behave as repeated calls on the reference object), solves on the GPU and
 * writes the results back into the vertices.
 *
 * Parameter, observation and information arguments may be any matrix type indexed
 * as m[row][col] (the reference's matrix::matrix<double, 0, 0> included);
 * get_parameters() returns ZSBA_SHIM_MATRIX, constructed as (rows, 1, const
 * double* data).  Define ZSBA_SHIM_MATRIX before including this header to return
 * the caller's own type (for example matrix::matrix<double, 0, 0>); the default is
 * zsba::column.  Cameras provide get_parameters(double* out, size_t 4) giving
 * (fx, fy, cx, cy), as camera::pinhole does (camera.hpp:83-90).
 *
 * No exceptions: failures are reported by return values (add_* / remove_* return
 * false, solve returns 0) and last_error().  Not thread-safe, as the reference.
 */
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "zsba.h"

namespace zsba {

// Dense column vector with the (rows, cols, data) constructor and m[i][0] access.
class column {
public:
    class row {
    public:
        explicit row(double* p) : p_(p) {}
        double& operator[](size_t) const { return *p_; }

    private:
        double* p_;
    };
    class const_row {
    public:
        explicit const_row(const double* p) : p_(p) {}
        const double& operator[](size_t) const { return *p_; }

    private:
        const double* p_;
    };
    column() = default;
    column(size_t rows, size_t /*cols*/, const double* data) : v_(data, data + rows) {}
    column(std::initializer_list<double> values) : v_(values) {}
    row operator[](size_t i) { return row(&v_[i]); }
    const_row operator[](size_t i) const { return const_row(&v_[i]); }
    size_t rows() const { return v_.size(); }
    const double* data() const { return v_.data(); }

private:
    std::vector<double> v_;
};

// Pinhole intrinsics with the reference camera's parameter interface.
class pinhole_camera {
public:
    pinhole_camera(double fx, double fy, double cx, double cy) : k_{fx, fy, cx, cy} {}
    pinhole_camera(const double* parameters, size_t /*length*/)
        : k_{parameters[0], parameters[1], parameters[2], parameters[3]} {}
    void get_parameters(double* out, size_t /*length*/) const {
        for (int i = 0; i < 4; ++i) out[i] = k_[i];
    }

private:
    double k_[4];
};

}  // namespace zsba

#ifndef ZSBA_SHIM_MATRIX
#define ZSBA_SHIM_MATRIX ::zsba::column
#endif
// Engine precision of a new factor_graph (set_precision changes it per graph).
#ifndef ZSBA_SHIM_PRECISION
#define ZSBA_SHIM_PRECISION ZSBA_PRECISION_FP64
#endif

namespace zsba {
namespace factor_graph {

using parameter_matrix = ZSBA_SHIM_MATRIX;

// ---- robust losses (factor_graph.hpp:851-934) ---------------------------------
class loss_function_base {
public:
    virtual ~loss_function_base() = default;
    virtual int32_t engine_kind() const = 0;
    virtual double engine_delta() const { return 0.0; }
};
class loss_trivial final : public loss_function_base {
public:
    int32_t engine_kind() const override { return ZSBA_LOSS_TRIVIAL; }
};
class loss_huber final : public loss_function_base {
public:
    explicit loss_huber(double delta) : delta_(delta) {}
    int32_t engine_kind() const override { return ZSBA_LOSS_HUBER; }
    double engine_delta() const override { return delta_; }

private:
    double delta_;
};
class loss_cauchy final : public loss_function_base {
public:
    explicit loss_cauchy(double delta) : delta_(delta) {}
    int32_t engine_kind() const override { return ZSBA_LOSS_CAUCHY; }
    double engine_delta() const override { return delta_; }

private:
    double delta_;
};
class loss_tukey final : public loss_function_base {
public:
    explicit loss_tukey(double delta) : delta_(delta) {}
    int32_t engine_kind() const override { return ZSBA_LOSS_TUKEY; }
    double engine_delta() const override { return delta_; }

private:
    double delta_;
};

// ---- vertices (factor_graph.hpp:58-146, 938-969) --------------------------------
class vertex_base {
public:
    virtual ~vertex_base() = default;
    void set_fixed(bool fixed_value = true) { fixed_ = fixed_value; }
    bool is_fixed() const { return fixed_; }
    void set_marginalised(bool marginalised_value = true) { marginalised_ = marginalised_value; }
    bool is_marginalised() const { return marginalised_; }
    int get_local_dimensions() const { return local_dimension_; }
    template <class M>
    void set_parameters(const M& m) {
        for (size_t i = 0; i < params_.size(); ++i) params_[i] = m[i][0];
    }
    parameter_matrix get_parameters() const { return parameter_matrix(params_.size(), 1, params_.data()); }
    const double* raw_parameters() const { return params_.data(); }
    double* raw_parameters() { return params_.data(); }
    virtual bool is_pose() const = 0;

protected:
    vertex_base(size_t dimension, int local_dimension) : params_(dimension, 0.0), local_dimension_(local_dimension) {}

private:
    std::vector<double> params_;
    int local_dimension_;
    bool fixed_ = false;
    bool marginalised_ = false;
};

// Parameters [tx, ty, tz, qx, qy, qz, qw] (factor_graph.hpp:941).
class vertex_pose final : public vertex_base {
public:
    vertex_pose() : vertex_base(7, 6) {}
    bool is_pose() const override { return true; }
};

class vertex_point_xyz final : public vertex_base {
public:
    vertex_point_xyz() : vertex_base(3, 3) {}
    bool is_pose() const override { return false; }
};

// ---- edges (factor_graph.hpp:148-276, 973-1049) -----------------------------------
class edge_base {
public:
    virtual ~edge_base() = default;
    bool add_vertex(vertex_base* vertex) {
        if (vertices_.size() >= 2) return false;
        vertices_.push_back(vertex);
        return true;
    }
    size_t num_vertices() const { return vertices_.size(); }
    vertex_base* get_vertex(size_t i) const { return vertices_[i]; }
    template <class M>
    void set_observation(const M& m) {
        obs_[0] = m[0][0];
        obs_[1] = m[1][0];
    }
    template <class M>
    void set_information(const M& m) {
        info_[0] = m[0][0];
        info_[1] = m[0][1];
        info_[2] = m[1][0];
        info_[3] = m[1][1];
        has_info_ = true;
    }
    void set_loss_function(loss_function_base* loss) { loss_ = loss; }
    const double* observation() const { return obs_; }
    const double* information() const { return has_info_ ? info_ : nullptr; }
    const loss_function_base* loss() const { return loss_; }
    virtual void intrinsics(double K[4]) const = 0;

private:
    std::vector<vertex_base*> vertices_;
    double obs_[2] = {0.0, 0.0};
    double info_[4] = {1.0, 0.0, 0.0, 1.0};
    bool has_info_ = false;
    loss_function_base* loss_ = nullptr;
};

template <typename camera_type>
class edge_reprojection final : public edge_base {
public:
    explicit edge_reprojection(const camera_type& camera_model) {
        camera_model.get_parameters(K_, 4);
    }
    void intrinsics(double K[4]) const override {
        for (int i = 0; i < 4; ++i) K[i] = K_[i];
    }

private:
    double K_[4];
};

// ---- the graph (factor_graph.hpp:341-629) ---------------------------------------
class factor_graph final {
public:
    explicit factor_graph(bool verbose = false) : verbose_(verbose) {}
    ~factor_graph() {
        if (handle_) zsba_destroy(handle_);
    }
    factor_graph(const factor_graph&) = delete;
    factor_graph& operator=(const factor_graph&) = delete;

    // Engine precision for solve(): ZSBA_PRECISION_FP64 (default, the reference's
    // operation order), ZSBA_PRECISION_FP64_FAST or ZSBA_PRECISION_MIXED.
    void set_precision(int32_t precision) { precision_ = precision; }
    void set_device(int32_t device) { device_ = device; }
    const std::string& last_error() const { return error_; }

    bool add_vertex(vertex_base* vertex) {
        if (!vertex || vertex_set_.count(vertex)) return false;
        vertex_set_.insert(vertex);
        (vertex->is_marginalised() ? points_ : poses_).push_back(vertex);
        structure_changed_ = true;
        return true;
    }
    bool remove_vertex(vertex_base* vertex) {
        if (!vertex_set_.count(vertex)) return false;
        vertex_set_.erase(vertex);
        erase(poses_, vertex);
        erase(points_, vertex);
        std::vector<edge_base*> keep;
        for (edge_base* e : edges_) {
            bool touches = false;
            for (size_t i = 0; i < e->num_vertices(); ++i) touches = touches || e->get_vertex(i) == vertex;
            if (touches) edge_set_.erase(e);
            else keep.push_back(e);
        }
        edges_.swap(keep);
        structure_changed_ = true;
        return true;
    }
    bool add_edge(edge_base* edge) {
        if (!edge || edge_set_.count(edge)) return false;
        edge_set_.insert(edge);
        edges_.push_back(edge);
        structure_changed_ = true;
        return true;
    }
    bool remove_edge(edge_base* edge) {
        if (!edge_set_.count(edge)) return false;
        edge_set_.erase(edge);
        erase(edges_, edge);
        structure_changed_ = true;
        return true;
    }
    std::vector<edge_base*> get_connected_edges(vertex_base* vertex) const {
        std::vector<edge_base*> out;
        for (edge_base* e : edges_)
            for (size_t i = 0; i < e->num_vertices(); ++i)
                if (e->get_vertex(i) == vertex) out.push_back(e);
        return out;
    }

    // Returns the accepted-step count; 0 for an empty graph or on failure.
    int solve(int iterations = 10, bool use_relative_convergence = false) {
        if (edges_.empty() || (poses_.empty() && points_.empty())) {
            if (verbose_) std::fprintf(stderr, "Cannot solve problem without edges or vertices\n");
            return 0;
        }
        if (!sync()) return 0;
        const int32_t accepted = zsba_solve(handle_, iterations, use_relative_convergence ? 1 : 0, precision_);
        if (accepted < 0) {
            error_ = zsba_last_error(handle_);
            return 0;
        }
        std::vector<double> P(7 * poses_.size()), X(3 * points_.size());
        if (zsba_get_poses(handle_, P.data()) < 0 || zsba_get_points(handle_, X.data()) < 0) {
            error_ = zsba_last_error(handle_);
            return 0;
        }
        for (size_t i = 0; i < poses_.size(); ++i) poses_[i]->set_parameters(parameter_matrix(7, 1, &P[7 * i]));
        for (size_t j = 0; j < points_.size(); ++j) points_[j]->set_parameters(parameter_matrix(3, 1, &X[3 * j]));
        return accepted;
    }

    double get_current_chi(bool recompute_residuals = true) {
        if (edges_.empty()) return 0.0;
        if (!sync()) return 0.0;
        return zsba_current_chi(handle_, recompute_residuals ? 1 : 0);
    }

private:
    template <class T>
    static void erase(std::vector<T*>& v, T* x) {
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i] == x) {
                v.erase(v.begin() + static_cast<std::ptrdiff_t>(i));
                return;
            }
    }

    // Sends the structure (when it changed) and every vertex's parameters and
    // fixed flag to the engine handle.
    bool sync() {
        if (structure_changed_ || !handle_) {
            if (handle_) zsba_destroy(handle_);
            zsba_options opt{};
            opt.device = device_;
            opt.verbose = verbose_ ? 1 : 0;
            handle_ = zsba_create(&opt);
            if (!handle_) {
                error_ = "zsba_create failed";
                return false;
            }
            index_.clear();
            for (size_t i = 0; i < poses_.size(); ++i) {
                if (!poses_[i]->is_pose()) return fail("a vertex that is not marginalised must be a vertex_pose");
                index_[poses_[i]] = zsba_add_pose(handle_, poses_[i]->raw_parameters(), poses_[i]->is_fixed());
            }
            for (size_t j = 0; j < points_.size(); ++j) {
                if (points_[j]->is_pose()) return fail("a marginalised vertex must be a vertex_point_xyz");
                index_[points_[j]] = zsba_add_point(handle_, points_[j]->raw_parameters(), points_[j]->is_fixed());
            }
            for (edge_base* e : edges_) {
                if (e->num_vertices() != 2) return fail("an edge needs a pose and a point");
                vertex_base* p = e->get_vertex(0);
                vertex_base* l = e->get_vertex(1);
                if (!index_.count(p) || !index_.count(l) || !p->is_pose() || l->is_pose())
                    return fail("edge vertices must be a pose then a point of this graph");
                double K[4];
                e->intrinsics(K);
                const loss_function_base* loss = e->loss();
                const int32_t kind = loss ? loss->engine_kind() : ZSBA_LOSS_NONE;
                const double delta = loss ? loss->engine_delta() : 0.0;
                if (zsba_add_observation(handle_, index_[p], index_[l], e->observation(), e->information(), K, kind,
                                         delta) < 0)
                    return fail(zsba_last_error(handle_));
            }
            structure_changed_ = false;
            return true;
        }
        for (size_t i = 0; i < poses_.size(); ++i) {
            if (zsba_set_pose(handle_, index_[poses_[i]], poses_[i]->raw_parameters()) < 0 ||
                zsba_set_pose_fixed(handle_, index_[poses_[i]], poses_[i]->is_fixed()) < 0)
                return fail(zsba_last_error(handle_));
        }
        for (size_t j = 0; j < points_.size(); ++j) {
            if (zsba_set_point(handle_, index_[points_[j]], points_[j]->raw_parameters()) < 0 ||
                zsba_set_point_fixed(handle_, index_[points_[j]], points_[j]->is_fixed()) < 0)
                return fail(zsba_last_error(handle_));
        }
        return true;
    }
    bool fail(const std::string& m) {
        error_ = m;
        structure_changed_ = true;
        return false;
    }

    bool verbose_;
    int32_t precision_ = ZSBA_SHIM_PRECISION;
    int32_t device_ = 0;
    std::string error_;
    zsba_graph* handle_ = nullptr;
    bool structure_changed_ = true;
    std::vector<vertex_base*> poses_, points_;
    std::vector<edge_base*> edges_;
    std::unordered_set<vertex_base*> vertex_set_;
    std::unordered_set<edge_base*> edge_set_;
    std::unordered_map<vertex_base*, int32_t> index_;
};

}  // namespace factor_graph
}  // namespace zsba
