// Host side of the engine: the C ABI of include/zsba.h.
//
// A handle owns a host copy of the graph (indexed by insertion id), the device
// copy in structure-of-arrays form, and one HIP stream.  zsba_solve analyses the
// structure once (observation order by point, W blocks, the lower-triangle
// blocks of the reduced camera system and the products feeding each of them),
// uploads, and then runs the Levenberg-Marquardt loop of factor_graph::solve
// (factor_graph.hpp:489-615) entirely on the device: the host enqueues attempt
// after attempt and only reads the controller's `done` flag, one attempt behind.
#include "zsba.h"
#include "common.hpp"
#include "ba_math.hpp"

#include <algorithm>
#include <tuple>
#include <mutex>
#include <map>
#include <cstdint>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

namespace zb {
void launch_end_attempt64(const dev_work64& w, hipStream_t s);
void launch_cost_stored64(const dev_problem& p, const dev_work64& w, hipStream_t s);
}  // namespace zb

using zb::dev_problem;
using zb::dev_work64;
using zb::lm_state;

namespace {

struct loss_entry {
    int32_t kind;
    double delta;
};

// Device memory chunks kept by the process after a handle releases them, per
// device, so that building a graph (a bundle adjustment per map::optimise call)
// does not call hipMalloc / hipFree once the pool holds chunks of its size.  At
// most kPoolCap bytes are kept per device; beyond that chunks are freed.
struct chunk {
    void* p;
    size_t bytes;
};
constexpr size_t kPoolCap = size_t(2) << 30;
constexpr int kPoolDevices = 64;
std::mutex pool_mutex;
std::vector<chunk> pool_free[kPoolDevices];
size_t pool_bytes[kPoolDevices] = {};

hipError_t pool_acquire(int dev, size_t bytes, chunk& out) {
    {
        std::lock_guard<std::mutex> lock(pool_mutex);
        if (dev >= 0 && dev < kPoolDevices) {
            std::vector<chunk>& v = pool_free[dev];
            size_t best = v.size();
            for (size_t i = 0; i < v.size(); ++i)
                if (v[i].bytes >= bytes && (best == v.size() || v[i].bytes < v[best].bytes)) best = i;
            if (best != v.size() && v[best].bytes <= 4 * bytes + (size_t(64) << 20)) {
                out = v[best];
                v.erase(v.begin() + static_cast<std::ptrdiff_t>(best));
                pool_bytes[dev] -= out.bytes;
                return hipSuccess;
            }
        }
    }
    out = {nullptr, bytes};
    return hipMalloc(&out.p, bytes);
}

void pool_release(int dev, const chunk& c) {
    std::lock_guard<std::mutex> lock(pool_mutex);
    if (dev < 0 || dev >= kPoolDevices || c.bytes > kPoolCap) {
        (void)hipFree(c.p);
        return;
    }
    std::vector<chunk>& v = pool_free[dev];
    while (pool_bytes[dev] + c.bytes > kPoolCap && !v.empty()) {  // drop the largest first
        size_t big = 0;
        for (size_t i = 1; i < v.size(); ++i)
            if (v[i].bytes > v[big].bytes) big = i;
        (void)hipFree(v[big].p);
        pool_bytes[dev] -= v[big].bytes;
        v.erase(v.begin() + static_cast<std::ptrdiff_t>(big));
    }
    v.push_back(c);
    pool_bytes[dev] += c.bytes;
}

// A handle's device buffers: 256-byte aligned pieces of pooled chunks (at least
// 8 MiB each), released back to the pool together.
struct dev_alloc {
    int device = 0;
    std::vector<chunk> chunks;
    size_t used = 0;   // bytes used in the last chunk
    size_t bytes = 0;  // bytes handed out
    hipError_t status = hipSuccess;
    template <class T>
    T* get(size_t n) {
        if (n == 0) n = 1;
        const size_t want = (n * sizeof(T) + 255) / 256 * 256;
        if (chunks.empty() || used + want > chunks.back().bytes) {
            chunk c;
            const size_t size = std::max(want, size_t(8) << 20);
            const hipError_t e = pool_acquire(device, size, c);
            if (e != hipSuccess) {
                status = e;
                return nullptr;
            }
            chunks.push_back(c);
            used = 0;
        }
        void* p = static_cast<char*>(chunks.back().p) + used;
        used += want;
        bytes += want;
        return static_cast<T*>(p);
    }
    void release() {
        for (const chunk& c : chunks) pool_release(device, c);
        chunks.clear();
        used = 0;
        bytes = 0;
        status = hipSuccess;
    }
};

// Streams and pinned controller records kept by the process between handles.
struct stream_slot {
    hipStream_t stream;
    void* pinned;
};
std::vector<stream_slot> stream_pool[kPoolDevices];

hipError_t stream_acquire(int dev, hipStream_t& s, void*& pinned, size_t pinned_bytes) {
    {
        std::lock_guard<std::mutex> lock(pool_mutex);
        if (dev >= 0 && dev < kPoolDevices && !stream_pool[dev].empty()) {
            s = stream_pool[dev].back().stream;
            pinned = stream_pool[dev].back().pinned;
            stream_pool[dev].pop_back();
            return hipSuccess;
        }
    }
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    e = hipHostMalloc(&pinned, pinned_bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(s);
        s = nullptr;
    }
    return e;
}

void stream_release(int dev, hipStream_t s, void* pinned) {
    std::lock_guard<std::mutex> lock(pool_mutex);
    if (dev >= 0 && dev < kPoolDevices && stream_pool[dev].size() < 64) {
        stream_pool[dev].push_back({s, pinned});
        return;
    }
    (void)hipHostFree(pinned);
    (void)hipStreamDestroy(s);
}

// Key of a block-sparse matrix block with the hash of the reference's
// sparse_block::block_key (matrix_sparse_block.hpp:47-59).  Used only to learn
// the iteration order of the reference's std::unordered_map containers, which
// sets the order of its Schur sums.
struct ref_key {
    size_t first, second;
    bool operator==(const ref_key& o) const { return first == o.first && second == o.second; }
};
struct ref_hash {
    std::size_t operator()(const ref_key& k) const {
        return ((std::hash<size_t>()(k.first) ^ (std::hash<size_t>()(k.second) << 1)) >> 1);
    }
};
using ref_map = std::unordered_map<ref_key, char, ref_hash>;

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct zsba_graph {
    zsba_options opt{};
    std::string err;

    // host graph, indexed by id
    std::vector<double> pose;
    std::vector<uint8_t> pose_fixed, pose_alive;
    std::vector<double> point;
    std::vector<uint8_t> point_fixed, point_alive;
    std::vector<int32_t> obs_pose, obs_point, obs_cam, obs_loss;
    std::vector<double> obs_uv, obs_info;
    std::vector<uint8_t> obs_alive, obs_has_info;
    std::vector<std::array<double, 4>> cams;
    std::vector<loss_entry> losses;

    bool structure_dirty = true;
    bool values_dirty = true;
    bool device_newer = false;

    // compaction of the last structure
    std::vector<int32_t> pose_slot_id, point_slot_id, obs_e_id;  // slot -> id
    std::vector<int32_t> pose_id_slot, point_id_slot;            // id -> slot (-1 dead)
    std::vector<int32_t> sorted_e;                               // s -> e

    // device
    bool hip_ready = false;
    hipStream_t stream = nullptr;
    dev_alloc mem;
    dev_problem dp;
    dev_work64 w;
    lm_state* lm_pinned = nullptr;
    double* lambda_dev = nullptr;
    uint8_t* flags_dev = nullptr;
    float* cull_mean_dev = nullptr;
    uint8_t* cull_erase_dev = nullptr;
    double* eval_dev = nullptr;
    size_t eval_capacity = 0;
    double* snap_pose = nullptr;
    double* snap_point = nullptr;
    uint8_t* frozen_dev = nullptr;  // [nl] behind-camera freeze flags (w.frozen while a solve uses them)
    bool all_points_fixed = false;  // every live point of the current structure is fixed
    bool have_snapshot = false;

    zsba_stats stats{};
    zb::kernel_timer timer;

    // Captured launch sequence of one LM attempt, per precision and freeze state
    // (the kernels' arguments), for the current structure.
    hipGraph_t att_graph[6] = {};
    hipGraphExec_t att_exec[6] = {};
    bool att_capture_failed = false;
    void drop_attempt_graphs() {
        for (int k = 0; k < 6; ++k) {
            if (att_exec[k]) (void)hipGraphExecDestroy(att_exec[k]);
            if (att_graph[k]) (void)hipGraphDestroy(att_graph[k]);
            att_exec[k] = nullptr;
            att_graph[k] = nullptr;
        }
    }

    // Exact-order state: the keys of the reference's factor_graph::h_pl container,
    // cleared and refilled by every make_hessian and kept across solves exactly as
    // that member is (factor_graph.hpp:366, 654, 704-709).
    ref_map hpl_mirror;
    std::vector<ref_key> hpl_sequence;  // first-occurrence key order of the current structure
    std::vector<int32_t> key_pair;      // pair index of hpl_sequence[i]
    bool orders_ready = false;          // set 1 computed for the current structure
    bool mirror_filled = false;         // the mirror has been filled for the current structure
    int2* d_contrib_x[2] = {nullptr, nullptr};
    int32_t* d_pose_pairs_x[2] = {nullptr, nullptr};
    int32_t* d_lm_pairs_x[2] = {nullptr, nullptr};
    std::vector<int2> contrib_fast;     // contributions grouped by block (host copy)
    std::vector<int64_t> blk_start_h;
    std::vector<int32_t> pose_pair_start_h, pose_pairs_h, lm_pair_start_h, pair_pose_h, pair_point_h;

    ~zsba_graph() {
        if (hip_ready) {
            if (stream) (void)hipStreamSynchronize(stream);
            drop_attempt_graphs();
            mem.release();
            if (eval_dev) (void)hipFree(eval_dev);
            if (stream) stream_release(opt.device, stream, lm_pinned);
        }
    }

    int32_t fail(int32_t code, const char* fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    int32_t hip_fail(hipError_t e, const char* what) {
        return fail(ZSBA_EDEVICE, "%s: %s", what, hipGetErrorString(e));
    }

    // Copies are ordered on the handle's stream and complete before returning.
    hipError_t copy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
        if (bytes == 0) return hipSuccess;
        const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, stream);
        if (e != hipSuccess) return e;
        return hipStreamSynchronize(stream);
    }

    int32_t live_poses() const { return static_cast<int32_t>(std::count(pose_alive.begin(), pose_alive.end(), 1)); }
    int32_t live_points() const { return static_cast<int32_t>(std::count(point_alive.begin(), point_alive.end(), 1)); }
    int32_t live_obs() const { return static_cast<int32_t>(std::count(obs_alive.begin(), obs_alive.end(), 1)); }

    int32_t ensure_hip() {
        if (hip_ready) return 0;
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || count == 0) return fail(ZSBA_EDEVICE, "no HIP device available");
        if (opt.device < 0 || opt.device >= count) return fail(ZSBA_EINVAL, "device %d out of range", opt.device);
        if ((e = hipSetDevice(opt.device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
        void* pinned = nullptr;
        if ((e = stream_acquire(opt.device, stream, pinned, sizeof(lm_state))) != hipSuccess)
            return hip_fail(e, "stream and pinned controller");
        lm_pinned = static_cast<lm_state*>(pinned);
        mem.device = opt.device;
        std::memset(lm_pinned, 0, sizeof(lm_state));
        hip_ready = true;
        return 0;
    }

    int32_t camera_index(const double K[4]) {
        if (!(K[0] > 0.0) || !(K[1] > 0.0)) return fail(ZSBA_EINVAL, "focal lengths must be positive (camera.hpp:78-79)");
        for (size_t i = 0; i < cams.size(); ++i)
            if (std::memcmp(cams[i].data(), K, 4 * sizeof(double)) == 0) return static_cast<int32_t>(i);
        if (cams.size() >= 65535) return fail(ZSBA_EUNSUP, "more than 65535 distinct cameras");
        cams.push_back({K[0], K[1], K[2], K[3]});
        return static_cast<int32_t>(cams.size() - 1);
    }
    int32_t loss_index(int32_t kind, double delta) {
        if (kind < ZSBA_LOSS_NONE || kind > ZSBA_LOSS_TUKEY) return fail(ZSBA_EINVAL, "unknown loss kind %d", kind);
        if (kind == ZSBA_LOSS_NONE || kind == ZSBA_LOSS_TRIVIAL) delta = 0.0;
        for (size_t i = 0; i < losses.size(); ++i)
            if (losses[i].kind == kind && std::memcmp(&losses[i].delta, &delta, sizeof(double)) == 0) return static_cast<int32_t>(i);
        if (losses.size() >= 32767) return fail(ZSBA_EUNSUP, "more than 32767 distinct losses");
        losses.push_back({kind, delta});
        return static_cast<int32_t>(losses.size() - 1);
    }

    // Pull device parameters into the host arrays if the device holds newer ones.
    int32_t sync_to_host() {
        if (!device_newer) return 0;
        std::vector<double> P(7 * pose_slot_id.size()), X(3 * point_slot_id.size());
        hipError_t e;
        if (!P.empty() && (e = copy(P.data(), w.pose, P.size() * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(e, "download poses");
        if (!X.empty() && (e = copy(X.data(), w.point, X.size() * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(e, "download points");
        for (size_t i = 0; i < pose_slot_id.size(); ++i) std::memcpy(&pose[7 * pose_slot_id[i]], &P[7 * i], 7 * sizeof(double));
        for (size_t j = 0; j < point_slot_id.size(); ++j) std::memcpy(&point[3 * point_slot_id[j]], &X[3 * j], 3 * sizeof(double));
        device_newer = false;
        return 0;
    }

    int32_t build_structure();
    int32_t upload_values();
    int32_t prepare();
    int32_t prepare_exact_orders();
    void order_set_from_mirror(std::vector<int2>& contrib, std::vector<int32_t>& pose_pairs,
                               std::vector<int32_t>& lm_pairs);
    // precision: ZSBA_PRECISION_FP64 (reference order), _FP64_FAST, _MIXED
    int32_t run(int32_t iterations, int32_t use_rel, int32_t precision);
    int32_t adjoint(const double* dL_dpoints, const double* dL_dposes, double* dL_dobs, double* dL_dK);
    int32_t gauss_newton(int32_t steps);
    int32_t hold_indefinite_points(double rel_tol, uint8_t* held);
    double adjoint_residual = 0.0;  // relative CG residual of the last adjoint
    void print_trace();
    double mixed_floor() const {
        const int e = opt.floor_exp == 0 ? kDefaultFloorExp : opt.floor_exp;
        return e < 0 ? 0.0 : std::ldexp(1.0, -e);
    }
    static constexpr int kDefaultFloorExp = 24;
    // Whether a solve in this precision applies the behind-camera freeze.
    bool freeze_for(int32_t precision) const {
        if (precision == ZSBA_PRECISION_FP64 || opt.freeze_behind < 0) return false;
        return opt.freeze_behind > 0 || precision == ZSBA_PRECISION_MIXED;
    }
    // Points flagged by the last linearisation of a solve with the freeze on.
    int32_t count_frozen() {
        if (!w.frozen || dp.nl == 0) return 0;
        std::vector<uint8_t> f(static_cast<size_t>(dp.nl));
        if (copy(f.data(), w.frozen, f.size(), hipMemcpyDeviceToHost) != hipSuccess) return -1;
        return static_cast<int32_t>(std::count(f.begin(), f.end(), 1));
    }
};

// ---- structure analysis -----------------------------------------------------
int32_t zsba_graph::build_structure() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_attempt_graphs();
    const int32_t np_all = static_cast<int32_t>(pose_alive.size());
    const int32_t nl_all = static_cast<int32_t>(point_alive.size());
    pose_slot_id.clear(); point_slot_id.clear(); obs_e_id.clear();
    pose_id_slot.assign(np_all, -1);
    point_id_slot.assign(nl_all, -1);
    for (int32_t i = 0; i < np_all; ++i)
        if (pose_alive[i]) { pose_id_slot[i] = static_cast<int32_t>(pose_slot_id.size()); pose_slot_id.push_back(i); }
    for (int32_t j = 0; j < nl_all; ++j)
        if (point_alive[j]) { point_id_slot[j] = static_cast<int32_t>(point_slot_id.size()); point_slot_id.push_back(j); }
    for (int32_t k = 0; k < static_cast<int32_t>(obs_alive.size()); ++k)
        if (obs_alive[k]) obs_e_id.push_back(k);
    const int32_t np = static_cast<int32_t>(pose_slot_id.size());
    const int32_t nl = static_cast<int32_t>(point_slot_id.size());
    const int32_t no = static_cast<int32_t>(obs_e_id.size());

    // observations of a point, insertion order (counting sort by point slot)
    std::vector<int32_t> lm_start(nl + 1, 0), e_pose(no), e_point(no);
    for (int32_t e = 0; e < no; ++e) {
        const int32_t id = obs_e_id[e];
        e_pose[e] = pose_id_slot[obs_pose[id]];
        e_point[e] = point_id_slot[obs_point[id]];
        lm_start[e_point[e] + 1]++;
    }
    for (int32_t l = 0; l < nl; ++l) lm_start[l + 1] += lm_start[l];
    sorted_e.assign(no, 0);
    std::vector<int32_t> e_sorted(no), fill(lm_start.begin(), lm_start.end() - 1);
    for (int32_t e = 0; e < no; ++e) {
        const int32_t s = fill[e_point[e]]++;
        sorted_e[s] = e;
        e_sorted[e] = s;
    }
    // observations of a pose, insertion order
    std::vector<int32_t> pose_start(np + 1, 0), pose_list(no);
    for (int32_t e = 0; e < no; ++e) pose_start[e_pose[e] + 1]++;
    for (int32_t i = 0; i < np; ++i) pose_start[i + 1] += pose_start[i];
    std::vector<int32_t> obs_pose_pos(no);
    {
        std::vector<int32_t> f(pose_start.begin(), pose_start.end() - 1);
        for (int32_t e = 0; e < no; ++e) {
            const int32_t t = f[e_pose[e]]++;
            pose_list[t] = e_sorted[e];
            obs_pose_pos[e_sorted[e]] = t;
        }
    }
    // W blocks: distinct free (pose, point) pairs, first-occurrence order per point
    std::vector<int32_t> obs_pair(no, -1), lm_pair_start(nl + 1, 0), pair_pose, pair_point, pair_first;
    std::vector<uint8_t> pfix(np), lfix(nl);
    for (int32_t i = 0; i < np; ++i) pfix[i] = pose_fixed[pose_slot_id[i]];
    for (int32_t l = 0; l < nl; ++l) lfix[l] = point_fixed[point_slot_id[l]];
    all_points_fixed = std::all_of(lfix.begin(), lfix.end(), [](uint8_t f) { return f != 0; });
    std::vector<int32_t> stamp(np, -1), stamp_pair(np, -1);
    for (int32_t l = 0; l < nl; ++l) {
        lm_pair_start[l] = static_cast<int32_t>(pair_pose.size());
        if (!lfix[l]) {
            for (int32_t s = lm_start[l]; s < lm_start[l + 1]; ++s) {
                const int32_t k = e_pose[sorted_e[s]];
                if (pfix[k]) continue;
                if (stamp[k] != l) {
                    stamp[k] = l;
                    stamp_pair[k] = static_cast<int32_t>(pair_pose.size());
                    pair_pose.push_back(k);
                    pair_point.push_back(l);
                    pair_first.push_back(s);
                }
                obs_pair[s] = stamp_pair[k];
            }
        }
    }
    lm_pair_start[nl] = static_cast<int32_t>(pair_pose.size());
    const int32_t npair = static_cast<int32_t>(pair_pose.size());

    // linearisation blocks: whole points, at most kLinBlock observations
    std::vector<int32_t> lb_start, lb_point, lp_start, le_pose, lo_start, pe_start(np + 1, 0), pe_list;
    std::vector<uint8_t> lo_list, pair_multi(npair, 0);
    {
        std::vector<int32_t> pair_count(npair, 0);
        for (int32_t s = 0; s < no; ++s)
            if (obs_pair[s] >= 0) pair_count[obs_pair[s]]++;
        for (int32_t a = 0; a < npair; ++a) pair_multi[a] = pair_count[a] > 1 ? 1 : 0;
        bool fits = true;
        for (int32_t l = 0; l < nl && fits; ++l) fits = lm_start[l + 1] - lm_start[l] <= zb::kLinBlock;
        if (fits && no > 0) {
            std::vector<int32_t> slot_of(np, -1), members;
            std::vector<std::vector<uint8_t>> offs;
            auto close_block = [&](int32_t b0, int32_t b1) {
                // entries of the block: free poses in order of first appearance
                members.clear();
                offs.clear();
                for (int32_t s = b0; s < b1; ++s) {
                    const int32_t k = e_pose[sorted_e[s]];
                    if (pfix[k]) continue;
                    if (slot_of[k] < 0) {
                        slot_of[k] = static_cast<int32_t>(members.size());
                        members.push_back(k);
                        offs.emplace_back();
                    }
                    offs[slot_of[k]].push_back(static_cast<uint8_t>(s - b0));
                }
                for (size_t m = 0; m < members.size(); ++m) {
                    le_pose.push_back(members[m]);
                    lo_start.push_back(static_cast<int32_t>(lo_list.size()));
                    lo_list.insert(lo_list.end(), offs[m].begin(), offs[m].end());
                    slot_of[members[m]] = -1;
                }
                lb_start.push_back(b1);
                lp_start.push_back(static_cast<int32_t>(le_pose.size()));
            };
            lb_start.push_back(0);
            lb_point.push_back(0);
            lp_start.push_back(0);
            int32_t b0 = 0;
            for (int32_t l = 0; l < nl; ++l) {
                if (lm_start[l + 1] - b0 > zb::kLinBlock) {
                    close_block(b0, lm_start[l]);
                    lb_point.push_back(l);
                    b0 = lm_start[l];
                }
            }
            close_block(b0, no);
            lb_point.push_back(nl);
            lo_start.push_back(static_cast<int32_t>(lo_list.size()));
            for (size_t e = 0; e < le_pose.size(); ++e) pe_start[le_pose[e] + 1]++;
            for (int32_t i = 0; i < np; ++i) pe_start[i + 1] += pe_start[i];
            pe_list.assign(le_pose.size(), 0);  // entry -> pose-major slot
            std::vector<int32_t> f(pe_start.begin(), pe_start.end() - 1);
            for (size_t e = 0; e < le_pose.size(); ++e) pe_list[e] = f[le_pose[e]]++;
        }
    }
    const int32_t n_lblk = lb_start.empty() ? 0 : static_cast<int32_t>(lb_start.size()) - 1;
    const int32_t n_lent = static_cast<int32_t>(le_pose.size());

    // key sequence of h_pl (first occurrence in insertion order, factor_graph.hpp:664-709)
    {
        hpl_sequence.clear();
        key_pair.clear();
        std::vector<uint8_t> inserted(npair, 0);
        for (int32_t e = 0; e < no; ++e) {
            const int32_t k = e_pose[e], l = e_point[e];
            if (pfix[k] || lfix[l]) continue;
            const int32_t a = obs_pair[e_sorted[e]];
            if (inserted[a]) continue;
            inserted[a] = 1;
            hpl_sequence.push_back(ref_key{static_cast<size_t>(k), static_cast<size_t>(l)});
            key_pair.push_back(a);
        }
    }
    // W blocks of each pose (for b_s)
    std::vector<int32_t> pose_pair_start(np + 1, 0), pose_pairs(npair);
    for (int32_t a = 0; a < npair; ++a) pose_pair_start[pair_pose[a] + 1]++;
    for (int32_t i = 0; i < np; ++i) pose_pair_start[i + 1] += pose_pair_start[i];
    {
        std::vector<int32_t> f(pose_pair_start.begin(), pose_pair_start.end() - 1);
        for (int32_t a = 0; a < npair; ++a) pose_pairs[f[pair_pose[a]]++] = a;
    }

    // lower-triangle blocks of the reduced system and their contributions
    std::vector<int64_t> block_key;  // row * np + col
    std::vector<int32_t> blk_count;
    const bool dense_index = static_cast<int64_t>(np) * np <= (int64_t(1) << 26);
    std::vector<int32_t> dense_map;
    std::unordered_map<int64_t, int32_t> hash_map;
    if (dense_index) dense_map.assign(static_cast<size_t>(np) * np, -1);
    auto block_of = [&](int32_t row, int32_t col) -> int32_t {
        const int64_t key = static_cast<int64_t>(row) * np + col;
        if (dense_index) {
            int32_t& b = dense_map[key];
            if (b < 0) { b = static_cast<int32_t>(block_key.size()); block_key.push_back(key); blk_count.push_back(0); }
            return b;
        }
        auto it = hash_map.find(key);
        if (it != hash_map.end()) return it->second;
        const int32_t b = static_cast<int32_t>(block_key.size());
        hash_map.emplace(key, b);
        block_key.push_back(key);
        blk_count.push_back(0);
        return b;
    };
    for (int32_t i = 0; i < np; ++i) block_of(i, i);
    for (int32_t l = 0; l < nl; ++l)
        for (int32_t a = lm_pair_start[l]; a < lm_pair_start[l + 1]; ++a)
            for (int32_t c = lm_pair_start[l]; c < lm_pair_start[l + 1]; ++c)
                if (pair_pose[a] >= pair_pose[c]) blk_count[block_of(pair_pose[a], pair_pose[c])]++;
    const int32_t nblk = static_cast<int32_t>(block_key.size());
    std::vector<int32_t> order(nblk);
    for (int32_t b = 0; b < nblk; ++b) order[b] = b;
    std::sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return block_key[x] < block_key[y]; });
    std::vector<int32_t> rank(nblk);
    for (int32_t r = 0; r < nblk; ++r) rank[order[r]] = r;
    std::vector<int64_t> blk_start(nblk + 1, 0);
    std::vector<int32_t> blk_row(nblk), blk_col(nblk), diag_blk(np);
    for (int32_t r = 0; r < nblk; ++r) {
        const int32_t b = order[r];
        blk_start[r + 1] = blk_start[r] + blk_count[b];
        blk_row[r] = static_cast<int32_t>(block_key[b] / np);
        blk_col[r] = static_cast<int32_t>(block_key[b] % np);
        if (blk_row[r] == blk_col[r]) diag_blk[blk_row[r]] = r;
    }
    const int64_t ncontrib = blk_start[nblk];
    // envelope: pose i couples to no pose left of first_pose[i]; Cholesky fill
    // stays inside it, so tiles left of it stay exactly zero
    const int32_t n_tiles = static_cast<int32_t>((6L * np + 31) / 32);
    std::vector<int32_t> tile_first(n_tiles), tile_last(n_tiles);
    int32_t panel_tiles = 0, band = 0;
    {
        std::vector<int32_t> first_pose(np);
        for (int32_t i = 0; i < np; ++i) first_pose[i] = i;
        for (int32_t r = 0; r < nblk; ++r) first_pose[blk_row[r]] = std::min(first_pose[blk_row[r]], blk_col[r]);
        for (int32_t t = 0; t < n_tiles; ++t) {
            int64_t f = INT64_MAX;
            for (int64_t r = 32L * t; r < std::min<int64_t>(32L * (t + 1), 6L * np); ++r)
                f = std::min<int64_t>(f, 6L * first_pose[r / 6]);
            tile_first[t] = static_cast<int32_t>(f / 32);
            tile_last[t] = t;
        }
        for (int32_t t = 0; t < n_tiles; ++t)
            for (int32_t k = tile_first[t]; k < t; ++k) tile_last[k] = std::max(tile_last[k], t);
        for (int32_t k = 0; k < n_tiles; ++k) panel_tiles = std::max(panel_tiles, tile_last[k] - k + 1);
        for (int64_t r = 0; r < 6L * np; ++r) band = std::max(band, static_cast<int32_t>(r - 6L * first_pose[r / 6] + 1));
    }
    std::vector<int2> contrib(static_cast<size_t>(ncontrib));
    {
        std::vector<int64_t> f(blk_start.begin(), blk_start.end() - 1);
        for (int32_t l = 0; l < nl; ++l)
            for (int32_t a = lm_pair_start[l]; a < lm_pair_start[l + 1]; ++a)
                for (int32_t c = lm_pair_start[l]; c < lm_pair_start[l + 1]; ++c)
                    if (pair_pose[a] >= pair_pose[c]) {
                        const int32_t r = rank[block_of(pair_pose[a], pair_pose[c])];
                        contrib[f[r]++] = make_int2(a, c);
                    }
    }

    // per-observation parameters
    std::vector<int32_t> obs_param(no), obs_pose_s(no), obs_point_s(no);
    std::vector<double> uv(2L * no), info;
    bool identity = true;
    for (int32_t s = 0; s < no; ++s) {
        const int32_t id = obs_e_id[sorted_e[s]];
        if (obs_has_info[id]) {
            const double* W = &obs_info[4L * id];
            if (!(W[0] == 1.0 && W[1] == 0.0 && W[2] == 0.0 && W[3] == 1.0)) identity = false;
        }
    }
    if (!identity) info.resize(4L * no);
    for (int32_t s = 0; s < no; ++s) {
        const int32_t e = sorted_e[s], id = obs_e_id[e];
        obs_param[s] = obs_cam[id] | (obs_loss[id] << 16);
        obs_pose_s[s] = e_pose[e];
        obs_point_s[s] = e_point[e];
        uv[2L * s] = obs_uv[2L * id];
        uv[2L * s + 1] = obs_uv[2L * id + 1];
        if (!identity) {
            if (obs_has_info[id]) std::memcpy(&info[4L * s], &obs_info[4L * id], 4 * sizeof(double));
            else { info[4L * s] = 1; info[4L * s + 1] = 0; info[4L * s + 2] = 0; info[4L * s + 3] = 1; }
        }
    }
    std::vector<double> cam_table(4 * std::max<size_t>(cams.size(), 1), 0.0);
    for (size_t c = 0; c < cams.size(); ++c) std::memcpy(&cam_table[4 * c], cams[c].data(), 4 * sizeof(double));
    std::vector<int32_t> lk(std::max<size_t>(losses.size(), 1), 0);
    std::vector<double> ld(std::max<size_t>(losses.size(), 1), 0.0);
    for (size_t i = 0; i < losses.size(); ++i) { lk[i] = losses[i].kind; ld[i] = losses[i].delta; }

    // allocate and upload
    (void)hipStreamSynchronize(stream);
    mem.release();
    have_snapshot = false;
    const int64_t n = 6L * np;
    dev_problem p;
    p.np = np; p.nl = nl; p.no = no; p.npair = npair; p.nblk = nblk; p.ncontrib = ncontrib;
    p.n = static_cast<int32_t>(n); p.identity_info = identity ? 1 : 0;
    // uploads are queued; the host arrays live until the synchronisation below
    auto up = [&](auto* dst, const auto& src) {
        if (!src.empty() && dst)
            (void)hipMemcpyAsync(dst, src.data(), src.size() * sizeof(src[0]), hipMemcpyHostToDevice, stream);
        return dst;
    };
    int32_t *d_obs_pose = mem.get<int32_t>(no), *d_obs_point = mem.get<int32_t>(no), *d_obs_pair = mem.get<int32_t>(no),
            *d_obs_param = mem.get<int32_t>(no), *d_obs_orig = mem.get<int32_t>(no);
    double* d_uv = mem.get<double>(2L * no);
    zb::obs_record* d_rec = mem.get<zb::obs_record>(no);
    double* d_info = identity ? nullptr : mem.get<double>(4L * no);
    int32_t *d_lm_start = mem.get<int32_t>(nl + 1), *d_pose_start = mem.get<int32_t>(np + 1), *d_pose_list = mem.get<int32_t>(no),
            *d_lm_pair_start = mem.get<int32_t>(nl + 1), *d_pair_pose = mem.get<int32_t>(npair),
            *d_pair_point = mem.get<int32_t>(npair), *d_pair_first = mem.get<int32_t>(npair),
            *d_blk_row = mem.get<int32_t>(nblk), *d_blk_col = mem.get<int32_t>(nblk), *d_diag = mem.get<int32_t>(np);
    int64_t* d_blk_start = mem.get<int64_t>(nblk + 1);
    int2* d_contrib = mem.get<int2>(static_cast<size_t>(ncontrib));
    uint8_t *d_pfix = mem.get<uint8_t>(np), *d_lfix = mem.get<uint8_t>(nl);
    double* d_cams = mem.get<double>(cam_table.size());
    int32_t* d_lk = mem.get<int32_t>(lk.size());
    double* d_ld = mem.get<double>(ld.size());
    w = dev_work64{};
    w.timer = opt.timing ? &timer : nullptr;
    timer.on = opt.timing != 0;
    timer.mask = opt.timing_classes ? static_cast<uint32_t>(opt.timing_classes) : ~0u;
    w.pose = mem.get<double>(7L * np); w.pose_backup = mem.get<double>(7L * np);
    w.point = mem.get<double>(3L * nl); w.point_backup = mem.get<double>(3L * nl);
    w.residual = mem.get<double>(2L * no); w.pose_terms = mem.get<double>(27L * no);
    w.Hpp = mem.get<double>(21L * np); w.bp = mem.get<double>(6L * np);
    w.Hll = mem.get<double>(9L * nl); w.bl = mem.get<double>(3L * nl);
    w.W = mem.get<double>(18L * npair); w.Hinv = mem.get<double>(9L * nl); w.Y = mem.get<double>(18L * npair);
    w.S = mem.get<double>(static_cast<size_t>(n * n)); w.bs = mem.get<double>(n); w.dxp = mem.get<double>(n);
    w.dxl = mem.get<double>(3L * nl);
    w.partial = mem.get<double>(3 * zb::kPartials + 8);
    w.cbuf = mem.get<double>(no);
    w.sbuf = mem.get<double>(6L * np + 3L * nl);
    w.Lt = mem.get<double>(static_cast<size_t>(n * n + n));
    const size_t band_doubles = zb::band_buffer_doubles(static_cast<int32_t>(n), band);
    w.Lband = band_doubles ? mem.get<double>(band_doubles) : nullptr;
    w.pose_terms_f = mem.get<float>(21L * no); w.pose_grad = mem.get<double>(6L * no);
    w.Hpp_f = mem.get<float>(21L * np); w.bp_m = mem.get<double>(6L * np);
    w.Hll_f = mem.get<float>(9L * nl); w.bl_m = mem.get<double>(3L * nl);
    w.W_f = mem.get<float>(18L * npair); w.Hinv_f = mem.get<float>(9L * nl); w.Y_f = mem.get<float>(18L * npair);
    w.zl = mem.get<double>(3L * nl); w.gl = mem.get<double>(3L * nl);
    w.rvec = mem.get<double>(n); w.corr = mem.get<double>(n);
    w.pose_part = mem.get<double>(static_cast<size_t>(np) * zb::kPoseChunks * 27);
    w.lin_part_h = mem.get<double>(21L * n_lent);  // room for binary64 storage
    w.lin_part_g = mem.get<double>(6L * n_lent);
    int32_t* d_lb_start = mem.get<int32_t>(lb_start.size());
    int32_t* d_lb_point = mem.get<int32_t>(lb_point.size());
    int32_t* d_lp_start = mem.get<int32_t>(lp_start.size());
    int32_t* d_lo_start = mem.get<int32_t>(lo_start.size());
    uint8_t* d_lo_list = mem.get<uint8_t>(lo_list.size());
    int32_t* d_pe_start = mem.get<int32_t>(pe_start.size());
    int32_t* d_pe_list = mem.get<int32_t>(pe_list.size());
    uint8_t* d_pair_multi = mem.get<uint8_t>(pair_multi.size());
    int32_t* d_pose_pairs = mem.get<int32_t>(npair);
    w.adj = mem.get<double>(6 * (6L * np + 3L * nl));
    w.adj_obs = mem.get<double>(2L * no);
    w.adj_frozen = mem.get<uint8_t>(nl);
    frozen_dev = mem.get<uint8_t>(nl);
    w.adj_partial = mem.get<double>(513);
    int32_t* d_tile_first = mem.get<int32_t>(n_tiles);
    int32_t* d_tile_last = mem.get<int32_t>(n_tiles);
    for (int set = 0; set < 2; ++set) {
        d_contrib_x[set] = mem.get<int2>(static_cast<size_t>(ncontrib));
        d_pose_pairs_x[set] = mem.get<int32_t>(npair);
        d_lm_pairs_x[set] = mem.get<int32_t>(npair);
    }
    int32_t* d_pose_pair_start = mem.get<int32_t>(np + 1);
    int32_t* d_obs_pose_pos = mem.get<int32_t>(no);
    w.chol_scale = mem.get<double>(4);
    w.chol_status = mem.get<int32_t>(4);
    w.lm = mem.get<lm_state>(1);
    lambda_dev = mem.get<double>(1);
    flags_dev = mem.get<uint8_t>(no);
    cull_mean_dev = mem.get<float>(nl);
    cull_erase_dev = mem.get<uint8_t>(nl);
    snap_pose = mem.get<double>(7L * np);
    snap_point = mem.get<double>(3L * nl);
    if (mem.status != hipSuccess) {
        mem.release();
        return fail(ZSBA_ENOMEM, "device allocation of %.1f MB failed", mem.bytes / 1e6);
    }
    // the band factor's entries outside the band are never written: zero once
    if (w.Lband) (void)hipMemsetAsync(w.Lband, 0, band_doubles * sizeof(double), stream);
    std::vector<int32_t> orig(no);
    for (int32_t s = 0; s < no; ++s) orig[s] = sorted_e[s];
    p.obs_pose = up(d_obs_pose, obs_pose_s); p.obs_point = up(d_obs_point, obs_point_s);
    p.obs_pair = up(d_obs_pair, obs_pair); p.obs_param = up(d_obs_param, obs_param); p.obs_orig = up(d_obs_orig, orig);
    p.obs_uv = up(d_uv, uv); p.obs_info = identity ? nullptr : up(d_info, info);
    std::vector<zb::obs_record> rec(no);
    for (int32_t s = 0; s < no; ++s) {
        const int32_t a = obs_pair[s];
        rec[s] = {obs_pose_s[s], obs_point_s[s], obs_param[s], a < 0 ? -1 : 2 * a + (pair_first[a] == s ? 1 : 0),
                  uv[2L * s], uv[2L * s + 1]};
    }
    p.obs_rec = up(d_rec, rec);
    p.lm_start = up(d_lm_start, lm_start); p.pose_start = up(d_pose_start, pose_start); p.pose_list = up(d_pose_list, pose_list);
    p.lm_pair_start = up(d_lm_pair_start, lm_pair_start); p.pair_pose = up(d_pair_pose, pair_pose);
    p.pair_point = up(d_pair_point, pair_point); p.pair_first = up(d_pair_first, pair_first);
    p.blk_row = up(d_blk_row, blk_row); p.blk_col = up(d_blk_col, blk_col); p.blk_start = up(d_blk_start, blk_start);
    p.contrib = up(d_contrib, contrib); p.diag_blk = up(d_diag, diag_blk);
    p.pose_fixed = up(d_pfix, pfix); p.point_fixed = up(d_lfix, lfix);
    p.pose_pair_start = up(d_pose_pair_start, pose_pair_start);
    p.pose_pairs = up(d_pose_pairs, pose_pairs);
    p.n_lblk = n_lblk;
    p.n_lent = n_lent;
    if (n_lblk > 0) {
        p.lb_start = up(d_lb_start, lb_start); p.lb_point = up(d_lb_point, lb_point); p.lp_start = up(d_lp_start, lp_start);
        p.lo_start = up(d_lo_start, lo_start); p.lo_list = up(d_lo_list, lo_list);
        p.pe_start = up(d_pe_start, pe_start); p.e_slot = up(d_pe_list, pe_list);
    }
    p.pair_multi = up(d_pair_multi, pair_multi);
    w.tile_first = up(d_tile_first, tile_first);
    w.tile_last = up(d_tile_last, tile_last);
    w.panel_tiles = panel_tiles;
    w.band = band;
    w.factor_path = opt.factor_path;
    p.obs_pose_pos = up(d_obs_pose_pos, obs_pose_pos);
    for (int set = 0; set < 2; ++set) {
        p.contrib_x[set] = d_contrib_x[set];
        p.pose_pairs_x[set] = d_pose_pairs_x[set];
        p.lm_pairs_x[set] = d_lm_pairs_x[set];
    }
    contrib_fast = contrib;
    blk_start_h = blk_start;
    pose_pair_start_h = pose_pair_start;
    pose_pairs_h = pose_pairs;
    lm_pair_start_h = lm_pair_start;
    pair_pose_h = pair_pose;
    pair_point_h = pair_point;
    orders_ready = false;
    mirror_filled = false;
    p.cams = up(d_cams, cam_table); p.loss_kind = up(d_lk, lk); p.loss_delta = up(d_ld, ld);
    (void)hipMemsetAsync(w.chol_status, 0, 4 * sizeof(int32_t), stream);
    // points without observations are in no linearisation block: their H_ll, b_l
    // and freeze flag stay zero
    (void)hipMemsetAsync(w.Hll_f, 0, sizeof(float) * 9 * static_cast<size_t>(nl), stream);
    (void)hipMemsetAsync(w.Hll, 0, sizeof(double) * 9 * static_cast<size_t>(nl), stream);
    (void)hipMemsetAsync(w.bl_m, 0, sizeof(double) * 3 * static_cast<size_t>(nl), stream);
    (void)hipMemsetAsync(frozen_dev, 0, static_cast<size_t>(nl), stream);
    (void)hipStreamSynchronize(stream);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "structure upload");
    dp = p;
    structure_dirty = false;
    values_dirty = true;
    return 0;
}

// Refill the mirror as make_hessian refills h_pl, then derive the orders the
// reference sums in: Y = multiply(h_pl, h_ll^-1) is a fresh map reserved to
// |h_pl| and filled in h_pl order (matrix_sparse_block_diagonal.hpp:107-122);
// h_lp = get_transpose() likewise with swapped keys (matrix_sparse_block.hpp:132-139).
// P(i,k) accumulates over Y in Y order, b_s over Y in Y order, the point
// back-substitution over h_lp in h_lp order.
void zsba_graph::order_set_from_mirror(std::vector<int2>& contrib, std::vector<int32_t>& pose_pairs,
                                       std::vector<int32_t>& lm_pairs) {
    hpl_mirror.clear();
    for (const ref_key& k : hpl_sequence)
        if (hpl_mirror.find(k) == hpl_mirror.end()) hpl_mirror[k] = 0;
    std::unordered_map<int64_t, int32_t> pair_of;
    pair_of.reserve(key_pair.size() * 2 + 1);
    const int64_t nl = static_cast<int64_t>(point_slot_id.size());
    for (size_t i = 0; i < hpl_sequence.size(); ++i)
        pair_of[static_cast<int64_t>(hpl_sequence[i].first) * nl + static_cast<int64_t>(hpl_sequence[i].second)] = key_pair[i];
    ref_map Y, LP;
    Y.reserve(hpl_mirror.size());
    LP.reserve(hpl_mirror.size());
    for (auto it = hpl_mirror.begin(); it != hpl_mirror.end(); ++it) {
        if (Y.find(it->first) == Y.end()) Y[it->first] = 0;
        LP[ref_key{it->first.second, it->first.first}] = 0;
    }
    const size_t npair = key_pair.size();
    std::vector<int32_t> rank_y(npair, 0), rank_lp(npair, 0);
    int32_t r = 0;
    for (auto it = Y.begin(); it != Y.end(); ++it, ++r)
        rank_y[pair_of[static_cast<int64_t>(it->first.first) * nl + static_cast<int64_t>(it->first.second)]] = r;
    r = 0;
    for (auto it = LP.begin(); it != LP.end(); ++it, ++r)
        rank_lp[pair_of[static_cast<int64_t>(it->first.second) * nl + static_cast<int64_t>(it->first.first)]] = r;
    contrib = contrib_fast;
    const size_t nblk = blk_start_h.size() - 1;
    for (size_t b = 0; b < nblk; ++b)
        std::stable_sort(contrib.begin() + blk_start_h[b], contrib.begin() + blk_start_h[b + 1],
                         [&](const int2& u, const int2& v) { return rank_y[u.x] < rank_y[v.x]; });
    pose_pairs = pose_pairs_h;
    for (size_t i = 0; i + 1 < pose_pair_start_h.size(); ++i)
        std::sort(pose_pairs.begin() + pose_pair_start_h[i], pose_pairs.begin() + pose_pair_start_h[i + 1],
                  [&](int32_t u, int32_t v) { return rank_y[u] < rank_y[v]; });
    lm_pairs.resize(npair);
    for (size_t a = 0; a < npair; ++a) lm_pairs[a] = static_cast<int32_t>(a);
    for (size_t l = 0; l + 1 < lm_pair_start_h.size(); ++l)
        std::sort(lm_pairs.begin() + lm_pair_start_h[l], lm_pairs.begin() + lm_pair_start_h[l + 1],
                  [&](int32_t u, int32_t v) { return rank_lp[u] < rank_lp[v]; });
}

// Orders for the first linearisation of this solve (set 0) and for the later
// ones (set 1).  After one refill of the current key sequence the mirror's
// bucket count no longer changes, so every later refill iterates identically.
int32_t zsba_graph::prepare_exact_orders() {
    std::vector<int2> c;
    std::vector<int32_t> pp, lp;
    hipError_t e;
    auto upload = [&](int set) -> hipError_t {
        hipError_t r = copy(d_contrib_x[set], c.data(), c.size() * sizeof(int2), hipMemcpyHostToDevice);
        if (r != hipSuccess) return r;
        r = copy(d_pose_pairs_x[set], pp.data(), pp.size() * sizeof(int32_t), hipMemcpyHostToDevice);
        if (r != hipSuccess) return r;
        return copy(d_lm_pairs_x[set], lp.data(), lp.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    };
    if (!orders_ready) {
        order_set_from_mirror(c, pp, lp);  // fill #1 of this structure
        if ((e = upload(0)) != hipSuccess) return hip_fail(e, "order upload");
        order_set_from_mirror(c, pp, lp);  // fill #2: steady state
        if ((e = upload(1)) != hipSuccess) return hip_fail(e, "order upload");
        orders_ready = true;
        mirror_filled = true;
    } else if (mirror_filled) {
        // set 0 of this solve equals the steady-state order
        if ((e = copy(d_contrib_x[0], d_contrib_x[1], contrib_fast.size() * sizeof(int2), hipMemcpyDeviceToDevice)) != hipSuccess)
            return hip_fail(e, "order copy");
        if ((e = copy(d_pose_pairs_x[0], d_pose_pairs_x[1], pose_pairs_h.size() * sizeof(int32_t), hipMemcpyDeviceToDevice)) != hipSuccess)
            return hip_fail(e, "order copy");
        if ((e = copy(d_lm_pairs_x[0], d_lm_pairs_x[1], pair_pose_h.size() * sizeof(int32_t), hipMemcpyDeviceToDevice)) != hipSuccess)
            return hip_fail(e, "order copy");
    }
    return 0;
}

int32_t zsba_graph::upload_values() {
    const size_t np = pose_slot_id.size(), nl = point_slot_id.size();
    std::vector<double> P(7 * np), X(3 * nl);
    for (size_t i = 0; i < np; ++i) std::memcpy(&P[7 * i], &pose[7L * pose_slot_id[i]], 7 * sizeof(double));
    for (size_t j = 0; j < nl; ++j) std::memcpy(&X[3 * j], &point[3L * point_slot_id[j]], 3 * sizeof(double));
    hipError_t e;
    if (np && (e = copy(w.pose, P.data(), P.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "upload poses");
    if (nl && (e = copy(w.point, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "upload points");
    values_dirty = false;
    device_newer = false;
    return 0;
}

int32_t zsba_graph::prepare() {
    int32_t rc = ensure_hip();
    if (rc) return rc;
    const double t0 = now_ms();
    if (structure_dirty) {
        rc = sync_to_host();
        if (rc) return rc;
        rc = build_structure();
        if (rc) return rc;
    }
    if (values_dirty) {
        rc = upload_values();
        if (rc) return rc;
    }
    stats.setup_ms = now_ms() - t0;
    return 0;
}

// The verbose LM trace of factor_graph(true) (factor_graph.hpp:526, 581, 589)
// cannot be printed from the device; the summary line is printed instead.
void zsba_graph::print_trace() {
    std::fprintf(stderr, "[zsba] iterations %d, attempts %d, accepted %d, rejected %d, failed solves %d, chi2 % 7.7f -> % 7.7f, lambda % 8.7f\n",
                 lm_pinned->executed_iterations, lm_pinned->attempts, lm_pinned->success, lm_pinned->rejected,
                 lm_pinned->solve_failures, lm_pinned->initial_chi2, lm_pinned->chi2, lm_pinned->lambda);
}

int32_t zsba_graph::run(int32_t iterations, int32_t use_rel, int32_t precision) {
    hipError_t e;
    const bool exact = precision == ZSBA_PRECISION_FP64;
    const bool mixed = precision == ZSBA_PRECISION_MIXED;
    if (exact) {
        const int32_t rc = prepare_exact_orders();
        if (rc) return rc;
    }
    timer.reset();
    w.frozen = freeze_for(precision) ? frozen_dev : nullptr;
    const double t0 = now_ms();
    std::memset(lm_pinned, 0, sizeof(lm_state));
    lm_pinned->iterations = iterations;
    lm_pinned->use_relative = use_rel ? 1 : 0;
    lm_pinned->require_descent = exact ? 0 : 1;
    if ((e = hipMemcpyAsync(w.lm, lm_pinned, sizeof(lm_state), hipMemcpyHostToDevice, stream)) != hipSuccess)
        return hip_fail(e, "upload controller");
    // every point fixed (map::optimise(n, true, ...)): the reference-order solve
    // in one launch
    const bool motion = exact && all_points_fixed && dp.np > 0 && dp.np <= zb::kMotionMaxPoses;
    if (motion) {
        zb::launch_motion_solve(dp, w, stream);
    } else {
        zb::launch_cost64(dp, w, stream, 0, false, exact);
        if (mixed) {
            zb::launch_linearise_mixed(dp, w, stream, false);
            zb::launch_init_lm_mixed(dp, w, stream);
        } else {
            zb::launch_linearise64(dp, w, stream, false, exact);
            zb::launch_init_lm64(dp, w, stream);
        }
    }
    const int32_t max_attempts = motion ? 0 : iterations * 10;
    // one attempt: every kernel gated on the controller in device memory, so the
    // same sequence serves every attempt of the solve
    auto enqueue_attempt = [&]() {
        if (mixed) {
            zb::launch_reduced_system_mixed(dp, w, stream, nullptr);
            zb::launch_cholesky_solve64(dp.n, w, stream, true, false, mixed_floor());
            zb::launch_refine_mixed(dp, w, stream, opt.refine_steps, nullptr);
            zb::launch_after_solve64(w, stream);
            zb::launch_backsub_update_mixed(dp, w, stream);
        } else {
            zb::launch_reduced_system64(dp, w, stream, nullptr, exact);
            zb::launch_cholesky_solve64(dp.n, w, stream, true, exact);
            zb::launch_after_solve64(w, stream);
            zb::launch_backsub_update64(dp, w, stream, exact);
        }
        // the decision rides on the cost's final sum, the controller reset on the
        // linearisation's bookkeeping (the same state updates, in the same order)
        zb::launch_cost64(dp, w, stream, 1, true, exact, true);
        zb::launch_rollback64(dp, w, stream);
        if (mixed) zb::launch_linearise_mixed(dp, w, stream, true, true);
        else zb::launch_linearise64(dp, w, stream, true, exact, true);
    };
    hipGraphExec_t att = nullptr;
    if (max_attempts > 0 && !w.timer && opt.attempt_graph >= 0 && !att_capture_failed) {
        const int key = 2 * (precision < 0 ? 0 : (precision > 2 ? 2 : precision)) + (w.frozen ? 1 : 0);
        if (!att_exec[key]) {
            hipGraph_t gr = nullptr;
            hipGraphExec_t ex = nullptr;
            bool ok = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (ok) {
                enqueue_attempt();
                ok = hipStreamEndCapture(stream, &gr) == hipSuccess && gr;
            }
            ok = ok && hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) == hipSuccess;
            if (ok) {
                att_graph[key] = gr;
                att_exec[key] = ex;
            } else {
                // launch one by one from now on
                if (gr) (void)hipGraphDestroy(gr);
                (void)hipGetLastError();
                att_capture_failed = true;
            }
        }
        att = att_exec[key];
    }
    hipEvent_t ev[2];
    (void)hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    int32_t* done_pinned = &lm_pinned->done;
    for (int32_t a = 0; a < max_attempts; ++a) {
        if (att) {
            if ((e = hipGraphLaunch(att, stream)) != hipSuccess) {
                (void)hipStreamSynchronize(stream);
                (void)hipEventDestroy(ev[0]);
                (void)hipEventDestroy(ev[1]);
                return hip_fail(e, "hipGraphLaunch");
            }
        } else {
            enqueue_attempt();
        }
        (void)hipMemcpyAsync(done_pinned, &w.lm->done, sizeof(int32_t), hipMemcpyDeviceToHost, stream);
        (void)hipEventRecord(ev[a & 1], stream);
        // Keep one attempt queued behind the one whose flag is read.
        if (a >= 1) {
            (void)hipEventSynchronize(ev[(a - 1) & 1]);
            if (*done_pinned) break;
        }
    }
    e = hipStreamSynchronize(stream);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (e != hipSuccess) return hip_fail(e, "solve");
    if ((e = copy(lm_pinned, w.lm, sizeof(lm_state), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "download controller");

    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "kernel launch");
    stats.solve_ms = now_ms() - t0;
    timer.collect();
    for (int k = 0; k < zb::kernel_timer::kClasses && k < ZSBA_K_COUNT; ++k) {
        stats.kernel_ms[k] = timer.ms[k];
        stats.kernel_launches[k] = timer.launches[k];
    }
    stats.iterations = lm_pinned->executed_iterations;
    stats.attempts = lm_pinned->attempts;
    stats.linearisations = lm_pinned->linearisations;
    stats.accepted = lm_pinned->success;
    stats.rejected = lm_pinned->rejected;
    stats.solve_failures = lm_pinned->solve_failures;
    stats.initial_chi2 = lm_pinned->initial_chi2;
    stats.final_chi2 = lm_pinned->chi2;
    stats.lambda = lm_pinned->lambda;
    stats.frozen_points = count_frozen();
    w.frozen = nullptr;  // the other entry points follow the reference's rule
    if (opt.verbose) print_trace();
    device_newer = true;
    return lm_pinned->success;
}

// ---- undamped Gauss-Newton polish ------------------------------------------------
// steps x (linearise, reduced system with lambda = 0 and a 1e-12 max|S_ii|
// diagonal floor for the empty rows of fixed poses, factorise, update), every
// step applied.  Converges quadratically to the stationary point the adjoint
// differentiates.
int32_t zsba_graph::gauss_newton(int32_t steps) {
    hipError_t e;
    const double zero = 0.0;
    if ((e = copy(lambda_dev, &zero, sizeof zero, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "lambda");
    for (int32_t k = 0; k < steps; ++k) {
        zb::launch_linearise64(dp, w, stream, false);
        zb::launch_reduced_system64(dp, w, stream, lambda_dev, false);
        zb::launch_cholesky_solve64(dp.n, w, stream, false, false, 1e-12);
        int32_t status = 0;
        if ((e = copy(&status, w.chol_status, sizeof status, hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(e, "status");
        if (status != 0) return fail(ZSBA_ESTATE, "Gauss-Newton step %d: the normal equations are not positive definite", k);
        zb::launch_gn_apply64(dp, w, stream);
    }
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "Gauss-Newton");
    device_newer = true;
    return 0;
}

// ---- points without a finite optimum ------------------------------------------
// The exact Hessian's 3x3 block of every free point (three products H e_k with
// e_k the k-th coordinate of every point, forward-mode duals through the
// gradient arithmetic) and the undamped Gauss-Newton block.  A point whose
// Gauss-Newton block is singular under the reference's inverse rule, or whose
// exact block is not positive definite (pivots of its symmetric part at most
// rel_tol of its trace), has no local optimum in its own coordinates at this
// state; it is set fixed.  held[id] = 1 for those points.  Returns their count.
int32_t zsba_graph::hold_indefinite_points(double rel_tol, uint8_t* held) {
    hipError_t e;
    const int64_t np = dp.np, nl = dp.nl, N = 6 * np + 3 * nl;
    if (nl == 0) return 0;
    const double zero = 0.0;
    if ((e = copy(lambda_dev, &zero, sizeof zero, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "lambda");
    zb::launch_linearise64(dp, w, stream, false);
    zb::launch_reduced_system64(dp, w, stream, lambda_dev, false);
    zb::launch_adjoint_frozen(dp, w, stream);
    double* dv = w.adj;
    double* out = dv + N;
    std::vector<double> v(N, 0.0), blocks(9 * nl, 0.0), col(N);
    for (int k = 0; k < 3; ++k) {
        for (int64_t j = 0; j < nl; ++j) {
            v[6 * np + 3 * j + 0] = 0.0;
            v[6 * np + 3 * j + 1] = 0.0;
            v[6 * np + 3 * j + 2] = 0.0;
            v[6 * np + 3 * j + k] = 1.0;
        }
        if ((e = copy(dv, v.data(), N * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "e_k");
        zb::launch_grad_tangent(dp, w, dv, nullptr, out, stream);
        if ((e = copy(col.data(), out, N * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(e, "H e_k");
        for (int64_t j = 0; j < nl; ++j)
            for (int i = 0; i < 3; ++i) blocks[9 * j + 3 * i + k] = col[6 * np + 3 * j + i];
    }
    std::vector<uint8_t> frozen(nl);
    if ((e = copy(frozen.data(), w.adj_frozen, nl, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "frozen");
    int32_t count = 0;
    for (int64_t j = 0; j < nl; ++j) {
        const int32_t id = point_slot_id[j];
        if (point_fixed[id]) continue;
        bool hold = frozen[j] != 0;
        if (!hold) {
            double a[9];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) a[3 * r + c] = 0.5 * (blocks[9 * j + 3 * r + c] + blocks[9 * j + 3 * c + r]);
            const double tol = rel_tol * (std::fabs(a[0]) + std::fabs(a[4]) + std::fabs(a[8]));
            // pivots of the symmetric 3x3 block (Gaussian elimination)
            const double p1 = a[0];
            const double p2 = p1 > tol ? a[4] - a[1] * a[3] / p1 : -1.0;
            const double m12 = a[5] - a[3] * a[2] / p1;
            const double p3 = p2 > tol ? (a[8] - a[2] * a[6] / p1) - m12 * m12 / p2 : -1.0;
            hold = !(p1 > tol && p2 > tol && p3 > tol);
        }
        if (hold) {
            point_fixed[id] = 1;
            structure_dirty = true;
            ++count;
            if (held) held[id] = 1;
        }
    }
    return count;
}

// ---- adjoint of the converged solve --------------------------------------------
// H lambda = dL/dx by conjugate gradients with the Gauss-Newton Schur factor as
// preconditioner (kernels_adjoint.hip), then dL/duv = -lambda^T dg/duv and
// dL/dK = -lambda^T dg/dK.  Returns the CG iterations.
int32_t zsba_graph::adjoint(const double* dL_dpoints, const double* dL_dposes, double* dL_dobs, double* dL_dK) {
    hipError_t e;
    const int64_t np = dp.np, nl = dp.nl, N = 6 * np + 3 * nl;
    double* x = w.adj;  // lambda
    double* r = x + N;
    double* z = r + N;
    double* pv = z + N;
    double* Ap = pv + N;
    double* b = Ap + N;
    // right-hand side in slot order (rows of fixed vertices zero)
    std::vector<double> hb(N, 0.0);
    for (int64_t i = 0; i < np; ++i) {
        const int32_t id = pose_slot_id[i];
        if (dL_dposes && !pose_fixed[id])
            for (int k = 0; k < 6; ++k) hb[6 * i + k] = dL_dposes[6L * id + k];
    }
    for (int64_t j = 0; j < nl; ++j) {
        const int32_t id = point_slot_id[j];
        if (dL_dpoints && !point_fixed[id])
            for (int k = 0; k < 3; ++k) hb[6 * np + 3 * j + k] = dL_dpoints[3L * id + k];
    }
    if ((e = copy(b, hb.data(), N * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "rhs");
    // Gauss-Newton preconditioner at x*: undamped blocks, reduced system with a
    // 1e-10 max|S_ii| diagonal floor (fixed poses have empty rows)
    const double zero = 0.0;
    if ((e = copy(lambda_dev, &zero, sizeof zero, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "lambda");
    zb::launch_linearise64(dp, w, stream, false);
    zb::launch_reduced_system64(dp, w, stream, lambda_dev, false);
    zb::launch_cholesky_solve64(dp.n, w, stream, false, false, 1e-10);
    zb::launch_adjoint_frozen(dp, w, stream);
    zb::launch_zero_frozen_rows(dp, w, b, stream);
    int32_t status = 0;
    if ((e = copy(&status, w.chol_status, sizeof status, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "status");
    if (status != 0) return fail(ZSBA_ESTATE, "the Gauss-Newton system at the current state is not positive definite");
    auto dot = [&](const double* a, const double* c, double& out) -> hipError_t {
        zb::launch_dot(a, c, N, w.adj_partial, w.adj_partial + 512, stream);
        return copy(&out, w.adj_partial + 512, sizeof(double), hipMemcpyDeviceToHost);
    };
    // CG: x = 0, r = b, z = M^-1 r, p = z
    (void)hipMemsetAsync(x, 0, N * sizeof(double), stream);
    (void)hipMemcpyAsync(r, b, N * sizeof(double), hipMemcpyDeviceToDevice, stream);
    zb::launch_adjoint_precond(dp, w, r, z, stream);
    (void)hipMemcpyAsync(pv, z, N * sizeof(double), hipMemcpyDeviceToDevice, stream);
    double rz = 0.0, bb = 0.0;
    if ((e = dot(r, z, rz)) != hipSuccess || (e = dot(b, b, bb)) != hipSuccess) return hip_fail(e, "cg");
    // Stop at a relative residual of 1e-12 or when it has not improved for 10
    // iterations (the rounding floor, about eps cond(H)); fail above 1e-8.
    int32_t it = 0;
    const int32_t max_it = 300;
    double best = 1.0;
    int32_t since_best = 0;
    adjoint_residual = 0.0;
    if (bb > 0.0) {
        for (it = 1; it <= max_it; ++it) {
            zb::launch_grad_tangent(dp, w, pv, nullptr, Ap, stream);
            double pAp = 0.0;
            if ((e = dot(pv, Ap, pAp)) != hipSuccess) return hip_fail(e, "cg");
            if (!(pAp > 0.0)) return fail(ZSBA_ESTATE, "the Hessian at the current state is not positive definite");
            const double alpha = rz / pAp;
            zb::launch_axpby(alpha, pv, 1.0, x, N, stream);
            zb::launch_axpby(-alpha, Ap, 1.0, r, N, stream);
            double rr = 0.0;
            if ((e = dot(r, r, rr)) != hipSuccess) return hip_fail(e, "cg");
            const double rel = std::sqrt(rr / bb);
            if (opt.verbose) std::fprintf(stderr, "[zsba] adjoint CG %d: relative residual %.3e\n", it, rel);
            adjoint_residual = rel;
            if (rel < best) {
                best = rel;
                since_best = 0;
            } else if (++since_best >= 10) {
                break;
            }
            if (rel <= 1e-12) break;
            zb::launch_adjoint_precond(dp, w, r, z, stream);
            double rz_new = 0.0;
            if ((e = dot(r, z, rz_new)) != hipSuccess) return hip_fail(e, "cg");
            const double beta = rz_new / rz;
            rz = rz_new;
            zb::launch_axpby(1.0, z, beta, pv, N, stream);
        }
        if (it > max_it) it = max_it;
        if (!(adjoint_residual <= 1e-8))
            return fail(ZSBA_ESTATE, "conjugate gradients stopped at relative residual %.3e after %d iterations",
                        adjoint_residual, it);
    }
    // dL/dK_j = -lambda . dg/dK_j
    if (dL_dK) {
        for (int j = 0; j < 4; ++j) {
            double dK[4] = {0.0, 0.0, 0.0, 0.0};
            dK[j] = 1.0;
            zb::launch_grad_tangent(dp, w, nullptr, dK, Ap, stream);
            double t = 0.0;
            if ((e = dot(x, Ap, t)) != hipSuccess) return hip_fail(e, "dK");
            dL_dK[j] = -t;
        }
    }
    if (dL_dobs) {
        zb::launch_obs_grad(dp, w, x, w.adj_obs, stream);
        std::vector<double> d(2L * dp.no);
        if ((e = copy(d.data(), w.adj_obs, d.size() * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(e, "dL/duv");
        for (int32_t k = 0; k < dp.no; ++k) {
            dL_dobs[2L * obs_e_id[k]] = d[2L * k];
            dL_dobs[2L * obs_e_id[k] + 1] = d[2L * k + 1];
        }
    }
    return it;
}

// ---- batched local BA -----------------------------------------------------------
// Every window whose reduced system fits one workgroup (zb::kBatchMaxN rows) is
// solved in lockstep by one sequence of launches; larger windows and the
// reference-order FP64 precision go through the per-graph solve.  A solve is one
// hipGraph launch of the start and the first `iterations` steps (an accepted
// attempt per iteration), then launches of kMoreSteps steps until every window
// has finished, each checked one launch behind through a flag in pinned memory.
// Both graphs are captured once per batch shape and replayed; the window table
// and the workgroup tables are refreshed before each solve.  The workgroup size
// of the factorisation and the Schur contributions in flight are chosen per
// batch shape by timing the candidates on the batch itself (states restored
// after each trial) and remembered for the process.
namespace {
struct tune_key {
    int32_t count, max_n, obs_bucket, mixed;
    bool operator<(const tune_key& o) const {
        return std::tie(count, max_n, obs_bucket, mixed) < std::tie(o.count, o.max_n, o.obs_bucket, o.mixed);
    }
};
std::mutex tune_mutex;
std::map<tune_key, std::pair<int32_t, int32_t>> tune_cache;  // threads, Schur contributions in flight
constexpr int32_t kMoreSteps = 2;
constexpr int kXcds = 8;
}  // namespace

struct zsba_batch {
    std::vector<zsba_graph*> graphs;
    std::string err;
    zsba_batch_info info{};
    hipStream_t stream = nullptr;
    int device = -1;
    zb::batch_window* d_wins = nullptr;
    lm_state* d_out = nullptr;
    lm_state* h_out = nullptr;
    uint64_t* d_pc = nullptr;  // phase clock: last stamp, its phase, ticks per phase
    uint64_t* h_pc = nullptr;
    int32_t* d_flag = nullptr;
    int32_t* h_flag = nullptr;  // 2 slots
    int2* d_maps = nullptr;
    size_t maps_capacity = 0;
    int32_t capacity = 0;
    std::vector<zb::batch_window> h_wins;
    std::vector<int2> h_maps;
    int32_t min_threads = 256;  // one factorisation row thread per row below a diagonal tile
    zb::batch_grid grid{};
    int32_t first_steps = 0, refine_rounds = 0, max_steps = 0;
    std::vector<int32_t> win_graph;  // window -> graph index
    hipGraphExec_t exec_first = nullptr, exec_more = nullptr;
    std::vector<int32_t> graph_key;

    ~zsba_batch() { release(); }
    void drop_graphs() {
        if (exec_first) (void)hipGraphExecDestroy(exec_first);
        if (exec_more) (void)hipGraphExecDestroy(exec_more);
        exec_first = exec_more = nullptr;
        graph_key.clear();
    }
    void release() {
        drop_graphs();
        if (stream) (void)hipStreamSynchronize(stream);
        if (d_wins) (void)hipFree(d_wins);
        if (d_out) (void)hipFree(d_out);
        if (h_out) (void)hipHostFree(h_out);
        if (d_pc) (void)hipFree(d_pc);
        if (h_pc) (void)hipHostFree(h_pc);
        if (d_flag) (void)hipFree(d_flag);
        if (h_flag) (void)hipHostFree(h_flag);
        if (d_maps) (void)hipFree(d_maps);
        if (stream) (void)hipStreamDestroy(stream);
        d_wins = nullptr;
        d_out = nullptr;
        h_out = nullptr;
        d_pc = nullptr;
        h_pc = nullptr;
        d_flag = nullptr;
        h_flag = nullptr;
        d_maps = nullptr;
        maps_capacity = 0;
        stream = nullptr;
        capacity = 0;
    }
    int32_t fail(int32_t code, const std::string& m) {
        err = m;
        return code;
    }
    int32_t hip_fail(hipError_t e, const char* what) {
        return fail(ZSBA_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }

    int32_t ensure(int32_t count, int dev) {
        hipError_t e;
        if (device != dev) {
            release();
            device = dev;
        }
        if ((e = hipSetDevice(dev)) != hipSuccess) return hip_fail(e, "hipSetDevice");
        if (!stream && (e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) != hipSuccess)
            return hip_fail(e, "hipStreamCreate");
        if (!d_pc) {
            const size_t bytes = sizeof(uint64_t) * (2 + zb::kBatchPhases);
            if ((e = hipMalloc(&d_pc, bytes)) != hipSuccess) return hip_fail(e, "hipMalloc");
            if ((e = hipHostMalloc(reinterpret_cast<void**>(&h_pc), bytes, hipHostMallocDefault)) != hipSuccess)
                return hip_fail(e, "hipHostMalloc");
            if ((e = hipMalloc(&d_flag, sizeof(int32_t))) != hipSuccess) return hip_fail(e, "hipMalloc");
            if ((e = hipHostMalloc(reinterpret_cast<void**>(&h_flag), 2 * sizeof(int32_t), hipHostMallocDefault)) !=
                hipSuccess)
                return hip_fail(e, "hipHostMalloc");
        }
        if (count > capacity) {
            drop_graphs();
            if (d_wins) (void)hipFree(d_wins);
            if (d_out) (void)hipFree(d_out);
            if (h_out) (void)hipHostFree(h_out);
            d_wins = nullptr;
            d_out = nullptr;
            h_out = nullptr;
            if ((e = hipMalloc(&d_wins, sizeof(zb::batch_window) * count)) != hipSuccess) return hip_fail(e, "hipMalloc");
            if ((e = hipMalloc(&d_out, sizeof(lm_state) * count)) != hipSuccess) return hip_fail(e, "hipMalloc");
            if ((e = hipHostMalloc(reinterpret_cast<void**>(&h_out), sizeof(lm_state) * count, hipHostMallocDefault)) !=
                hipSuccess)
                return hip_fail(e, "hipHostMalloc");
            capacity = count;
        }
        return 0;
    }

    // Workgroups of window w in table k (see zb::batch_map).
    static int64_t share(int k, const zb::batch_window& bw, bool mixed) {
        const dev_problem& p = bw.p;
        const bool by_blocks = mixed && p.n_lblk > 0;
        auto cdiv = [](int64_t a, int64_t b) { return (a + b - 1) / b; };
        auto sum_grid = [&](int64_t n) {
            const int64_t G = cdiv(n, 256);
            return G < 1 ? int64_t(1) : (G > zb::kPartials ? int64_t(zb::kPartials) : G);
        };
        switch (k) {
            case zb::BM_WIN: return 1;
            case zb::BM_LBLK: return by_blocks ? p.n_lblk : 0;
            case zb::BM_LPOSE: return by_blocks ? p.np : 0;
            case zb::BM_LPTS: return by_blocks ? 0 : cdiv(p.nl, 256);
            case zb::BM_LCHUNK: return by_blocks ? 0 : cdiv(static_cast<int64_t>(p.np) * zb::kPoseChunks, 4);
            case zb::BM_LENTRY: return by_blocks ? 0 : cdiv(27L * p.np, 256);
            case zb::BM_PTS: return cdiv(p.nl, 256);
            case zb::BM_INVZ: return cdiv(p.nl + static_cast<int64_t>(p.n) * p.n, 256);
            case zb::BM_PAIR: return cdiv(p.npair, 256);
            case zb::BM_SBLK: return p.nblk;
            case zb::BM_PTSN: return mixed ? cdiv(static_cast<int64_t>(p.nl) + p.n, 256) : 0;
            case zb::BM_APPLY: return sum_grid(static_cast<int64_t>(p.np) + p.nl);
            case zb::BM_COST: return sum_grid(p.no);
            case zb::BM_SLOT: return cdiv(static_cast<int64_t>(p.np) + p.nl, 256);
            default: return 0;
        }
    }

    // The workgroup tables.  Windows go to XCDs balanced by observation count
    // (largest first, each to the least-loaded XCD); in each table, XCD x's
    // workgroups are x, x + 8, x + 16, ... (the dispatcher's round robin), holding
    // its windows' shares in window order.  BM_WIN keeps window i at workgroup i.
    int32_t build_maps(bool mixed) {
        const int32_t nw = static_cast<int32_t>(h_wins.size());
        std::vector<int32_t> order(nw), xcd(nw, 0);
        for (int32_t i = 0; i < nw; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t a, int32_t b) { return h_wins[a].p.no > h_wins[b].p.no; });
        int64_t load[kXcds] = {};
        for (int32_t w : order) {
            int x = 0;
            for (int y = 1; y < kXcds; ++y)
                if (load[y] < load[x]) x = y;
            xcd[w] = x;
            load[x] += h_wins[w].p.no + 1;
        }
        h_maps.clear();
        size_t offset[zb::kBatchMaps];
        for (int k = 0; k < zb::kBatchMaps; ++k) {
            offset[k] = h_maps.size();
            if (k == zb::BM_WIN) {
                for (int32_t i = 0; i < nw; ++i) h_maps.push_back(make_int2(i, 0));
                grid.n_wg[k] = nw;
                continue;
            }
            std::vector<int2> list[kXcds];
            for (int32_t w = 0; w < nw; ++w) {
                const int64_t c = share(k, h_wins[w], mixed);
                for (int64_t j = 0; j < c; ++j) list[xcd[w]].push_back(make_int2(w, static_cast<int>(j)));
            }
            size_t J = 0;
            for (int x = 0; x < kXcds; ++x) J = std::max(J, list[x].size());
            if (J * kXcds > static_cast<size_t>(INT32_MAX)) return fail(ZSBA_EINVAL, "batch too large");
            for (size_t j = 0; j < J; ++j)
                for (int x = 0; x < kXcds; ++x) h_maps.push_back(j < list[x].size() ? list[x][j] : make_int2(-1, 0));
            grid.n_wg[k] = static_cast<int32_t>(J * kXcds);
        }
        hipError_t e;
        if (h_maps.size() > maps_capacity) {
            if (d_maps) (void)hipFree(d_maps);
            d_maps = nullptr;
            maps_capacity = 0;
            if ((e = hipMalloc(&d_maps, sizeof(int2) * h_maps.size())) != hipSuccess) return hip_fail(e, "hipMalloc");
            maps_capacity = h_maps.size();
        }
        if (!h_maps.empty() &&
            (e = hipMemcpyAsync(d_maps, h_maps.data(), sizeof(int2) * h_maps.size(), hipMemcpyHostToDevice, stream)) !=
                hipSuccess)
            return hip_fail(e, "upload workgroup tables");
        for (int k = 0; k < zb::kBatchMaps; ++k) grid.map[k] = d_maps + offset[k];
        return 0;
    }

    void enqueue_first(int32_t threads, int32_t inflight, bool mixed) {
        const int32_t nw = static_cast<int32_t>(h_wins.size());
        zb::launch_batch_begin(d_wins, nw, d_pc, stream);
        zb::launch_batch_steps(d_wins, grid, threads, inflight, mixed, refine_rounds, first_steps, d_pc, stream);
        zb::launch_batch_finish(d_wins, nw, d_out, d_flag, d_pc, stream);
    }
    void enqueue_more(int32_t threads, int32_t inflight, bool mixed) {
        const int32_t nw = static_cast<int32_t>(h_wins.size());
        zb::launch_batch_steps(d_wins, grid, threads, inflight, mixed, refine_rounds, kMoreSteps, d_pc, stream);
        zb::launch_batch_finish(d_wins, nw, d_out, d_flag, d_pc, stream);
    }

    int32_t capture(int32_t threads, int32_t inflight, bool mixed) {
        std::vector<int32_t> key = {threads, inflight, mixed ? 1 : 0, static_cast<int32_t>(h_wins.size()), first_steps,
                                    refine_rounds};
        for (int k = 0; k < zb::kBatchMaps; ++k) key.push_back(grid.n_wg[k]);
        if (exec_first && exec_more && key == graph_key) return 0;
        drop_graphs();
        hipError_t e;
        for (int which = 0; which < 2; ++which) {
            hipGraph_t gr = nullptr;
            if ((e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                return hip_fail(e, "hipStreamBeginCapture");
            if (which == 0) enqueue_first(threads, inflight, mixed);
            else enqueue_more(threads, inflight, mixed);
            if ((e = hipStreamEndCapture(stream, &gr)) != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            hipGraphExec_t ex = nullptr;
            e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
            (which == 0 ? exec_first : exec_more) = ex;
        }
        graph_key = key;
        return 0;
    }

    // One solve of the window table: the first launch, then more until every
    // window has finished (or the attempt limit of the per-graph solve,
    // 10 x iterations, is reached).
    int32_t run(int32_t threads, int32_t inflight, bool mixed, bool use_graphs) {
        hipError_t e;
        if (use_graphs) {
            int32_t rc = capture(threads, inflight, mixed);
            if (rc) return rc;
            if ((e = hipGraphLaunch(exec_first, stream)) != hipSuccess) return hip_fail(e, "hipGraphLaunch");
            info.graph_launches += 1;
        } else {
            enqueue_first(threads, inflight, mixed);
        }
        hipEvent_t ev[2];
        (void)hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
        (void)hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
        (void)hipMemcpyAsync(&h_flag[0], d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, stream);
        (void)hipEventRecord(ev[0], stream);
        int32_t steps = first_steps;
        for (int32_t i = 0;; ++i) {
            const bool more = steps < max_steps;
            if (more) {
                if (use_graphs) {
                    if ((e = hipGraphLaunch(exec_more, stream)) != hipSuccess) {
                        (void)hipStreamSynchronize(stream);
                        (void)hipEventDestroy(ev[0]);
                        (void)hipEventDestroy(ev[1]);
                        return hip_fail(e, "hipGraphLaunch");
                    }
                    info.graph_launches += 1;
                } else {
                    enqueue_more(threads, inflight, mixed);
                }
                (void)hipMemcpyAsync(&h_flag[(i + 1) & 1], d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, stream);
                (void)hipEventRecord(ev[(i + 1) & 1], stream);
                steps += kMoreSteps;
            }
            (void)hipEventSynchronize(ev[i & 1]);
            if (h_flag[i & 1] || !more) break;
        }
        e = hipStreamSynchronize(stream);
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        if (e != hipSuccess) return hip_fail(e, "batch solve");
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "kernel launch");
        info.steps = steps;
        return 0;
    }

    // Times every candidate (factorisation workgroup size x Schur contributions in
    // flight) once on this batch, restoring the window states after each trial,
    // and returns the fastest.
    int32_t autotune(bool mixed, int32_t& best_threads, int32_t& best_inflight) {
        hipError_t e;
        std::vector<double*> saved;
        for (int32_t k : win_graph) {
            zsba_graph* g = graphs[k];
            const size_t bytes = sizeof(double) * (7L * g->dp.np + 3L * g->dp.nl);
            double* buf = nullptr;
            if ((e = hipMalloc(&buf, bytes)) != hipSuccess) {
                for (double* s2 : saved) (void)hipFree(s2);
                return hip_fail(e, "hipMalloc");
            }
            (void)hipMemcpyAsync(buf, g->w.pose, sizeof(double) * 7L * g->dp.np, hipMemcpyDeviceToDevice, stream);
            (void)hipMemcpyAsync(buf + 7L * g->dp.np, g->w.point, sizeof(double) * 3L * g->dp.nl, hipMemcpyDeviceToDevice,
                                 stream);
            saved.push_back(buf);
        }
        (void)hipStreamSynchronize(stream);
        double best_ms = 0.0;
        best_threads = 0;
        best_inflight = 1;
        int32_t rc = 0;
        for (int32_t threads : {256, 512, 1024}) {
            if (threads < min_threads) continue;
            for (int32_t inflight : {1, 2}) {
                const double t0 = now_ms();
                if ((rc = run(threads, inflight, mixed, false))) break;
                const double ms = now_ms() - t0;
                float& slot = info.tune_ms[threads == 256 ? 0 : threads == 512 ? 1 : 2];
                if (slot == 0.f || ms < slot) slot = static_cast<float>(ms);
                if (best_threads == 0 || ms < best_ms) {
                    best_threads = threads;
                    best_inflight = inflight;
                    best_ms = ms;
                }
                for (size_t i = 0; i < win_graph.size(); ++i) {
                    zsba_graph* g = graphs[win_graph[i]];
                    (void)hipMemcpyAsync(g->w.pose, saved[i], sizeof(double) * 7L * g->dp.np, hipMemcpyDeviceToDevice,
                                         stream);
                    (void)hipMemcpyAsync(g->w.point, saved[i] + 7L * g->dp.np, sizeof(double) * 3L * g->dp.nl,
                                         hipMemcpyDeviceToDevice, stream);
                }
                (void)hipStreamSynchronize(stream);
            }
            if (rc) break;
        }
        for (double* buf : saved) (void)hipFree(buf);
        info.steps = 0;
        return rc;
    }

    int32_t solve(int32_t iterations, int32_t use_rel, int32_t precision, int32_t* accepted) {
        const int32_t count = static_cast<int32_t>(graphs.size());
        if (count > 0 && !accepted) return fail(ZSBA_EINVAL, "accepted must not be NULL");
        if (precision != ZSBA_PRECISION_FP64 && precision != ZSBA_PRECISION_MIXED && precision != ZSBA_PRECISION_FP64_FAST)
            return fail(ZSBA_EINVAL, "unknown precision " + std::to_string(precision));
        info = zsba_batch_info{};
        const double t_start = now_ms();
        h_wins.clear();
        win_graph.clear();
        int dev = -1;
        const bool mixed = precision == ZSBA_PRECISION_MIXED;
        int32_t max_refine = 0;
        for (int32_t k = 0; k < count; ++k) {
            zsba_graph* g = graphs[k];
            accepted[k] = 0;
            g->stats = zsba_stats{};
            g->stats.precision = precision;
            if (g->live_obs() == 0 || (g->live_poses() == 0 && g->live_points() == 0)) continue;
            int32_t rc = g->prepare();
            if (rc) return fail(rc, "graph " + std::to_string(k) + ": " + g->err);
            g->stats.n_obs = g->dp.no;
            g->stats.n_poses = g->dp.np;
            g->stats.n_points = g->dp.nl;
            const bool fits = g->dp.n <= zb::kBatchMaxN && g->w.panel_tiles <= zb::kBatchMaxN / 32;
            if (precision == ZSBA_PRECISION_FP64 || !fits) {
                rc = g->run(iterations, use_rel, precision);
                if (rc < 0) return fail(rc, "graph " + std::to_string(k) + ": " + g->err);
                accepted[k] = rc;
                info.individual += 1;
                continue;
            }
            if (dev >= 0 && g->opt.device != dev) return fail(ZSBA_EINVAL, "batched graphs must share one device");
            dev = g->opt.device;
            zb::batch_window bw;
            bw.p = g->dp;
            bw.w = g->w;
            bw.w.timer = nullptr;
            bw.w.frozen = g->freeze_for(precision) ? g->frozen_dev : nullptr;
            bw.iterations = iterations;
            bw.use_relative = use_rel ? 1 : 0;
            bw.refine_steps = g->opt.refine_steps;
            bw.floor_rel = mixed ? g->mixed_floor() : 0.0;
            max_refine = std::max(max_refine, bw.refine_steps);
            h_wins.push_back(bw);
            win_graph.push_back(k);
        }
        const int32_t nw = static_cast<int32_t>(h_wins.size());
        info.windows = nw;
        if (nw == 0) {
            info.solve_ms = now_ms() - t_start;
            return 0;
        }
        int32_t rc = ensure(nw, dev);
        if (rc) return rc;
        hipError_t e;
        if ((e = hipMemcpyAsync(d_wins, h_wins.data(), sizeof(zb::batch_window) * nw, hipMemcpyHostToDevice, stream)) !=
            hipSuccess)
            return hip_fail(e, "upload windows");
        if ((rc = build_maps(mixed))) return rc;
        // an accepted attempt per iteration in the first launch; at most 10
        // attempts per iteration (the controller ends an iteration after 10
        // failures), the per-graph solve's limit
        first_steps = std::max(iterations, 1);
        max_steps = 10 * std::max(iterations, 1);
        refine_rounds = mixed ? max_refine : 0;
        int32_t max_n = 0, max_panel = 1;
        int64_t obs = 0;
        for (int32_t i = 0; i < nw; ++i) {
            max_n = std::max(max_n, h_wins[i].p.n);
            max_panel = std::max(max_panel, graphs[win_graph[i]]->w.panel_tiles);
            obs += h_wins[i].p.no;
        }
        // the factorisation gives every row below a diagonal tile its own thread
        // of waves 1..: (tallest panel - 1) x 32 rows need that many threads + 64
        min_threads = (max_panel - 1) * 32 + 64 <= 256 ? 256 : 512;
        const tune_key key{nw, max_n * 64 + max_panel, static_cast<int32_t>(std::log2(static_cast<double>(obs) + 1.0)),
                           mixed ? 1 : 0};
        int32_t threads = 0, inflight = 1;
        {
            std::lock_guard<std::mutex> lock(tune_mutex);
            auto it = tune_cache.find(key);
            if (it != tune_cache.end()) {
                threads = it->second.first;
                inflight = it->second.second;
            }
        }
        if (threads == 0) {
            const double t = now_ms();
            if ((rc = autotune(mixed, threads, inflight))) return rc;
            info.autotune_wall_ms = now_ms() - t;
            std::lock_guard<std::mutex> lock(tune_mutex);
            tune_cache[key] = {threads, inflight};
        }
        info.threads = threads;
        info.schur_inflight = inflight;
        const double t0 = now_ms();
        if ((rc = run(threads, inflight, mixed, true))) return rc;
        if ((e = hipMemcpyAsync(h_out, d_out, sizeof(lm_state) * nw, hipMemcpyDeviceToHost, stream)) != hipSuccess)
            return hip_fail(e, "download controllers");
        if ((e = hipMemcpyAsync(h_pc, d_pc, sizeof(uint64_t) * (2 + zb::kBatchPhases), hipMemcpyDeviceToHost, stream)) !=
            hipSuccess)
            return hip_fail(e, "download phase clock");
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "batch solve");
        info.kernel_wall_ms = now_ms() - t0;
        for (int32_t i = 0; i < nw; ++i) {
            zsba_graph* g = graphs[win_graph[i]];
            const lm_state& s = h_out[i];
            *g->lm_pinned = s;
            g->stats.iterations = s.executed_iterations;
            g->stats.attempts = s.attempts;
            g->stats.linearisations = s.linearisations;
            g->stats.accepted = s.success;
            g->stats.rejected = s.rejected;
            g->stats.solve_failures = s.solve_failures;
            g->stats.initial_chi2 = s.initial_chi2;
            g->stats.final_chi2 = s.chi2;
            g->stats.lambda = s.lambda;
            g->stats.solve_ms = info.kernel_wall_ms;
            g->device_newer = true;
            accepted[win_graph[i]] = s.success;
        }
        // device time by phase over the whole batch (the clock runs at 100 MHz)
        for (int k = 0; k < zb::kBatchPhases; ++k) info.phase_ms[k] = static_cast<double>(h_pc[2 + k]) * 1e-5;
        info.lin_ms = info.phase_ms[zb::PH_LINEARISE];
        info.solve_ms = now_ms() - t_start;
        return 0;
    }
};

// ---- C ABI ------------------------------------------------------------------
extern "C" {

int32_t zsba_abi_version(void) { return ZSBA_ABI_VERSION; }

zsba_graph* zsba_create(const zsba_options* options) {
    zsba_graph* g = new (std::nothrow) zsba_graph();
    if (!g) return nullptr;
    if (options) g->opt = *options;
    if (g->opt.factor_path < 0 || g->opt.factor_path > 3) g->opt.factor_path = 0;
    if (g->opt.refine_steps == 0) g->opt.refine_steps = 1;
    else if (g->opt.refine_steps < 0) g->opt.refine_steps = 0;
    return g;
}

void zsba_destroy(zsba_graph* g) { delete g; }

const char* zsba_last_error(const zsba_graph* g) { return g ? g->err.c_str() : "null handle"; }

int32_t zsba_add_poses(zsba_graph* g, int32_t count, const double* params, const uint8_t* fixed) {
    if (!g || count < 0 || (count > 0 && !params)) return g ? g->fail(ZSBA_EINVAL, "bad arguments to add_poses") : ZSBA_EINVAL;
    if (g->sync_to_host()) return ZSBA_EDEVICE;
    const int32_t first = static_cast<int32_t>(g->pose_alive.size());
    g->pose.insert(g->pose.end(), params, params + 7L * count);
    for (int32_t i = 0; i < count; ++i) {
        g->pose_fixed.push_back(fixed ? (fixed[i] ? 1 : 0) : 0);
        g->pose_alive.push_back(1);
    }
    g->structure_dirty = true;
    return first;
}
int32_t zsba_add_pose(zsba_graph* g, const double params[7], int32_t fixed) {
    const uint8_t f = fixed ? 1 : 0;
    return zsba_add_poses(g, 1, params, &f);
}
int32_t zsba_add_points(zsba_graph* g, int32_t count, const double* xyz, const uint8_t* fixed) {
    if (!g || count < 0 || (count > 0 && !xyz)) return g ? g->fail(ZSBA_EINVAL, "bad arguments to add_points") : ZSBA_EINVAL;
    if (g->sync_to_host()) return ZSBA_EDEVICE;
    const int32_t first = static_cast<int32_t>(g->point_alive.size());
    g->point.insert(g->point.end(), xyz, xyz + 3L * count);
    for (int32_t i = 0; i < count; ++i) {
        g->point_fixed.push_back(fixed ? (fixed[i] ? 1 : 0) : 0);
        g->point_alive.push_back(1);
    }
    g->structure_dirty = true;
    return first;
}
int32_t zsba_add_point(zsba_graph* g, const double xyz[3], int32_t fixed) {
    const uint8_t f = fixed ? 1 : 0;
    return zsba_add_points(g, 1, xyz, &f);
}

int32_t zsba_add_observations(zsba_graph* g, int32_t count, const int32_t* pose_ids, const int32_t* point_ids,
                              const double* uv, const double* info, const double K[4], int32_t loss_kind, double loss_delta) {
    if (!g) return ZSBA_EINVAL;
    if (count < 0 || (count > 0 && (!pose_ids || !point_ids || !uv)) || !K)
        return g->fail(ZSBA_EINVAL, "bad arguments to add_observations");
    for (int32_t k = 0; k < count; ++k) {
        const int32_t ip = pose_ids[k], il = point_ids[k];
        if (ip < 0 || ip >= static_cast<int32_t>(g->pose_alive.size()) || !g->pose_alive[ip])
            return g->fail(ZSBA_EINVAL, "observation %d: unknown pose id %d", k, ip);
        if (il < 0 || il >= static_cast<int32_t>(g->point_alive.size()) || !g->point_alive[il])
            return g->fail(ZSBA_EINVAL, "observation %d: unknown point id %d", k, il);
    }
    const int32_t cam = g->camera_index(K);
    if (cam < 0) return cam;
    const int32_t li = g->loss_index(loss_kind, loss_delta);
    if (li < 0) return li;
    const int32_t first = static_cast<int32_t>(g->obs_alive.size());
    g->obs_pose.insert(g->obs_pose.end(), pose_ids, pose_ids + count);
    g->obs_point.insert(g->obs_point.end(), point_ids, point_ids + count);
    g->obs_uv.insert(g->obs_uv.end(), uv, uv + 2L * count);
    if (info) g->obs_info.insert(g->obs_info.end(), info, info + 4L * count);
    else g->obs_info.insert(g->obs_info.end(), 4L * count, 0.0);
    g->obs_has_info.insert(g->obs_has_info.end(), count, info ? 1 : 0);
    g->obs_cam.insert(g->obs_cam.end(), count, cam);
    g->obs_loss.insert(g->obs_loss.end(), count, li);
    g->obs_alive.insert(g->obs_alive.end(), count, 1);
    g->structure_dirty = true;
    return first;
}
int32_t zsba_add_observation(zsba_graph* g, int32_t pose_id, int32_t point_id, const double uv[2], const double info[4],
                             const double K[4], int32_t loss_kind, double loss_delta) {
    return zsba_add_observations(g, 1, &pose_id, &point_id, uv, info, K, loss_kind, loss_delta);
}

int32_t zsba_remove_observation(zsba_graph* g, int32_t id) {
    if (!g) return ZSBA_EINVAL;
    if (id < 0 || id >= static_cast<int32_t>(g->obs_alive.size()) || !g->obs_alive[id])
        return g->fail(ZSBA_EINVAL, "unknown observation id %d", id);
    g->obs_alive[id] = 0;
    g->structure_dirty = true;
    return 0;
}
int32_t zsba_remove_pose(zsba_graph* g, int32_t id) {
    if (!g) return ZSBA_EINVAL;
    if (id < 0 || id >= static_cast<int32_t>(g->pose_alive.size()) || !g->pose_alive[id])
        return g->fail(ZSBA_EINVAL, "unknown pose id %d", id);
    if (g->sync_to_host()) return ZSBA_EDEVICE;
    g->pose_alive[id] = 0;
    for (size_t k = 0; k < g->obs_alive.size(); ++k)
        if (g->obs_pose[k] == id) g->obs_alive[k] = 0;
    g->structure_dirty = true;
    return 0;
}
int32_t zsba_remove_point(zsba_graph* g, int32_t id) {
    if (!g) return ZSBA_EINVAL;
    if (id < 0 || id >= static_cast<int32_t>(g->point_alive.size()) || !g->point_alive[id])
        return g->fail(ZSBA_EINVAL, "unknown point id %d", id);
    if (g->sync_to_host()) return ZSBA_EDEVICE;
    g->point_alive[id] = 0;
    for (size_t k = 0; k < g->obs_alive.size(); ++k)
        if (g->obs_point[k] == id) g->obs_alive[k] = 0;
    g->structure_dirty = true;
    return 0;
}

int32_t zsba_set_pose(zsba_graph* g, int32_t id, const double params[7]) {
    if (!g || !params) return ZSBA_EINVAL;
    if (id < 0 || id >= static_cast<int32_t>(g->pose_alive.size())) return g->fail(ZSBA_EINVAL, "unknown pose id %d", id);
    if (g->sync_to_host()) return ZSBA_EDEVICE;
    std::memcpy(&g->pose[7L * id], params, 7 * sizeof(double));
    g->values_dirty = true;
    return 0;
}
int32_t zsba_set_point(zsba_graph* g, int32_t id, const double xyz[3]) {
    if (!g || !xyz) return ZSBA_EINVAL;
    if (id < 0 || id >= static_cast<int32_t>(g->point_alive.size())) return g->fail(ZSBA_EINVAL, "unknown point id %d", id);
    if (g->sync_to_host()) return ZSBA_EDEVICE;
    std::memcpy(&g->point[3L * id], xyz, 3 * sizeof(double));
    g->values_dirty = true;
    return 0;
}
int32_t zsba_set_pose_fixed(zsba_graph* g, int32_t id, int32_t fixed) {
    if (!g) return ZSBA_EINVAL;
    if (id < 0 || id >= static_cast<int32_t>(g->pose_alive.size())) return g->fail(ZSBA_EINVAL, "unknown pose id %d", id);
    const uint8_t f = fixed ? 1 : 0;
    if (g->pose_fixed[id] == f) return 0;  // unchanged: the structure stands
    if (g->sync_to_host()) return ZSBA_EDEVICE;
    g->pose_fixed[id] = f;
    g->structure_dirty = true;
    return 0;
}
int32_t zsba_set_point_fixed(zsba_graph* g, int32_t id, int32_t fixed) {
    if (!g) return ZSBA_EINVAL;
    if (id < 0 || id >= static_cast<int32_t>(g->point_alive.size())) return g->fail(ZSBA_EINVAL, "unknown point id %d", id);
    const uint8_t f = fixed ? 1 : 0;
    if (g->point_fixed[id] == f) return 0;  // unchanged: the structure stands
    if (g->sync_to_host()) return ZSBA_EDEVICE;
    g->point_fixed[id] = f;
    g->structure_dirty = true;
    return 0;
}

int32_t zsba_counts(const zsba_graph* g, int64_t counts[3]) {
    if (!g || !counts) return ZSBA_EINVAL;
    counts[0] = static_cast<int64_t>(g->pose_alive.size());
    counts[1] = static_cast<int64_t>(g->point_alive.size());
    counts[2] = static_cast<int64_t>(g->obs_alive.size());
    return 0;
}

int32_t zsba_get_poses(const zsba_graph* gc, double* out) {
    zsba_graph* g = const_cast<zsba_graph*>(gc);
    if (!g || !out) return ZSBA_EINVAL;
    if (g->sync_to_host()) return ZSBA_EDEVICE;
    for (size_t i = 0; i < g->pose_alive.size(); ++i)
        if (g->pose_alive[i]) std::memcpy(out + 7 * i, &g->pose[7 * i], 7 * sizeof(double));
    return 0;
}
int32_t zsba_get_points(const zsba_graph* gc, double* out) {
    zsba_graph* g = const_cast<zsba_graph*>(gc);
    if (!g || !out) return ZSBA_EINVAL;
    if (g->sync_to_host()) return ZSBA_EDEVICE;
    for (size_t j = 0; j < g->point_alive.size(); ++j)
        if (g->point_alive[j]) std::memcpy(out + 3 * j, &g->point[3 * j], 3 * sizeof(double));
    return 0;
}

int32_t zsba_solve(zsba_graph* g, int32_t iterations, int32_t use_relative_convergence, int32_t precision) {
    if (!g) return ZSBA_EINVAL;
    g->stats = zsba_stats{};
    g->stats.precision = precision;
    if (precision != ZSBA_PRECISION_FP64 && precision != ZSBA_PRECISION_MIXED && precision != ZSBA_PRECISION_FP64_FAST)
        return g->fail(ZSBA_EINVAL, "unknown precision %d", precision);
    // factor_graph.hpp:494-499
    if (g->live_obs() == 0 || (g->live_poses() == 0 && g->live_points() == 0)) {
        if (g->opt.verbose) std::fprintf(stderr, "Cannot solve problem without edges or vertices\n");
        return 0;
    }
    int32_t rc = g->prepare();
    if (rc) return rc;
    g->stats.n_obs = g->dp.no;
    g->stats.n_poses = g->dp.np;
    g->stats.n_points = g->dp.nl;
    return g->run(iterations, use_relative_convergence, precision);
}

double zsba_current_chi(zsba_graph* g, int32_t recompute) {
    if (!g) return NAN;
    if (g->live_obs() == 0) return 0.0;
    if (g->prepare()) return NAN;
    if (recompute) zb::launch_cost64(g->dp, g->w, g->stream, 3, false, true);
    else zb::launch_cost_stored64(g->dp, g->w, g->stream);
    double v = NAN;
    if (g->copy(&v, g->w.partial + zb::kPartials, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        g->fail(ZSBA_EDEVICE, "current_chi");
        return NAN;
    }
    return v;
}

int32_t zsba_get_inlier_flags(zsba_graph* g, uint8_t* flags) {
    if (!g || !flags) return ZSBA_EINVAL;
    if (g->live_obs() == 0) return 0;
    int32_t rc = g->prepare();
    if (rc) return rc;
    zb::launch_inliers64(g->dp, g->w, g->flags_dev, g->stream);
    std::vector<uint8_t> f(g->dp.no);
    const hipError_t e = g->copy(f.data(), g->flags_dev, f.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return g->hip_fail(e, "inlier flags");
    for (int32_t s = 0; s < g->dp.no; ++s) flags[g->obs_e_id[g->sorted_e[s]]] = f[s];
    return 0;
}

int32_t zsba_reprojection_cull(zsba_graph* g, float threshold, float* mean_error, uint8_t* erase) {
    if (!g || !mean_error || !erase) return ZSBA_EINVAL;
    if (g->live_points() == 0) return 0;
    if (g->live_obs() == 0) {  // every live point is without observations
        int32_t n = 0;
        for (size_t j = 0; j < g->point_alive.size(); ++j)
            if (g->point_alive[j]) {
                mean_error[j] = INFINITY;
                erase[j] = 1;
                ++n;
            }
        return n;
    }
    int32_t rc = g->prepare();
    if (rc) return rc;
    const int32_t nl = g->dp.nl;
    zb::launch_reprojection_cull(g->dp, g->w, threshold, g->cull_mean_dev, g->cull_erase_dev, g->stream);
    std::vector<float> m(nl);
    std::vector<uint8_t> f(nl);
    hipError_t e = g->copy(m.data(), g->cull_mean_dev, m.size() * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = g->copy(f.data(), g->cull_erase_dev, f.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return g->hip_fail(e, "reprojection cull");
    int32_t n = 0;
    for (int32_t l = 0; l < nl; ++l) {
        const int32_t id = g->point_slot_id[l];
        mean_error[id] = m[l];
        erase[id] = f[l];
        n += f[l];
    }
    return n;
}

int32_t zsba_get_residuals(zsba_graph* g, double* out) {
    if (!g || !out) return ZSBA_EINVAL;
    if (g->live_obs() == 0) return 0;
    int32_t rc = g->prepare();
    if (rc) return rc;
    zb::launch_cost64(g->dp, g->w, g->stream, 3, false);
    std::vector<double> r(2L * g->dp.no);
    const hipError_t e = g->copy(r.data(), g->w.residual, r.size() * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return g->hip_fail(e, "residuals");
    for (int32_t s = 0; s < g->dp.no; ++s) {
        const int32_t id = g->obs_e_id[g->sorted_e[s]];
        out[2L * id] = r[2L * s];
        out[2L * id + 1] = r[2L * s + 1];
    }
    return 0;
}

int32_t zsba_get_stats(const zsba_graph* g, zsba_stats* stats) {
    if (!g || !stats) return ZSBA_EINVAL;
    *stats = g->stats;
    return 0;
}

int32_t zsba_debug_eval_edges(zsba_graph* g, double* out) {
    if (!g || !out) return ZSBA_EINVAL;
    if (g->live_obs() == 0) return 0;
    int32_t rc = g->prepare();
    if (rc) return rc;
    const size_t need = 27L * g->dp.no;
    if (g->eval_capacity < need) {
        if (g->eval_dev) (void)hipFree(g->eval_dev);
        g->eval_dev = nullptr;
        g->eval_capacity = 0;
        if (hipMalloc(&g->eval_dev, need * sizeof(double)) != hipSuccess) return g->fail(ZSBA_ENOMEM, "eval buffer");
        g->eval_capacity = need;
    }
    zb::launch_eval_edges64(g->dp, g->w, g->eval_dev, g->stream);
    std::vector<double> v(need);
    const hipError_t e = g->copy(v.data(), g->eval_dev, need * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return g->hip_fail(e, "eval edges");
    for (int32_t k = 0; k < g->dp.no; ++k) std::memcpy(out + 27L * g->obs_e_id[k], &v[27L * k], 27 * sizeof(double));
    return 0;
}

int32_t zsba_debug_reduced_system(zsba_graph* g, double lambda, int32_t precision, double* S, double* b_s, double* dx_pose,
                                  double* dx_point) {
    if (!g) return ZSBA_EINVAL;
    if (precision != ZSBA_PRECISION_FP64 && precision != ZSBA_PRECISION_MIXED)
        return g->fail(ZSBA_EUNSUP, "debug system: precision %d", precision);
    const bool mixed = precision == ZSBA_PRECISION_MIXED;
    if (g->live_obs() == 0) return g->fail(ZSBA_ESTATE, "no observations");
    int32_t rc = g->prepare();
    if (rc) return rc;
    lm_state st{};
    st.lambda = lambda;
    hipError_t e;
    if ((e = hipStreamSynchronize(g->stream)) != hipSuccess) return g->hip_fail(e, "sync");
    if ((e = g->copy(g->w.lm, &st, sizeof st, hipMemcpyHostToDevice)) != hipSuccess) return g->hip_fail(e, "lm");
    if ((e = g->copy(g->lambda_dev, &lambda, sizeof lambda, hipMemcpyHostToDevice)) != hipSuccess)
        return g->hip_fail(e, "lambda");
    if (mixed) {
        zb::launch_linearise_mixed(g->dp, g->w, g->stream, false);
        zb::launch_reduced_system_mixed(g->dp, g->w, g->stream, g->lambda_dev);
    } else {
        zb::launch_linearise64(g->dp, g->w, g->stream, false);
        zb::launch_reduced_system64(g->dp, g->w, g->stream, g->lambda_dev);
    }
    const int64_t n = g->dp.n;
    if (S && (e = hipMemcpyAsync(S, g->w.S, n * n * sizeof(double), hipMemcpyDeviceToHost, g->stream)) != hipSuccess)
        return g->hip_fail(e, "S");
    if (b_s && (e = hipMemcpyAsync(b_s, g->w.bs, n * sizeof(double), hipMemcpyDeviceToHost, g->stream)) != hipSuccess)
        return g->hip_fail(e, "b_s");
    zb::launch_cholesky_solve64(g->dp.n, g->w, g->stream, false, false, mixed ? g->mixed_floor() : 0.0);
    if (mixed) {
        zb::launch_refine_mixed(g->dp, g->w, g->stream, g->opt.refine_steps, g->lambda_dev);
        zb::launch_backsub_only_mixed(g->dp, g->w, g->stream);
    } else {
        zb::launch_backsub_only64(g->dp, g->w, g->stream);
    }
    if ((e = hipStreamSynchronize(g->stream)) != hipSuccess) return g->hip_fail(e, "reduced system");
    lm_state out{};
    if ((e = g->copy(&out, g->w.lm, sizeof out, hipMemcpyDeviceToHost)) != hipSuccess) return g->hip_fail(e, "lm");
    if (out.solve_ok) {
        if (dx_pose && (e = g->copy(dx_pose, g->w.dxp, n * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
            return g->hip_fail(e, "dx_pose");
        if (dx_point) {
            std::vector<double> d(3L * g->dp.nl);
            if ((e = g->copy(d.data(), g->w.dxl, d.size() * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
                return g->hip_fail(e, "dx_point");
            std::memcpy(dx_point, d.data(), d.size() * sizeof(double));
        }
    }
    return out.solve_ok ? 1 : 0;
}

int32_t zsba_snapshot(zsba_graph* g) {
    if (!g) return ZSBA_EINVAL;
    if (g->live_obs() == 0) return g->fail(ZSBA_ESTATE, "no observations");
    int32_t rc = g->prepare();
    if (rc) return rc;
    hipError_t e;
    if ((e = g->copy(g->snap_pose, g->w.pose, 7L * g->dp.np * sizeof(double), hipMemcpyDeviceToDevice)) != hipSuccess)
        return g->hip_fail(e, "snapshot");
    if ((e = g->copy(g->snap_point, g->w.point, 3L * g->dp.nl * sizeof(double), hipMemcpyDeviceToDevice)) != hipSuccess)
        return g->hip_fail(e, "snapshot");
    g->have_snapshot = true;
    return 0;
}

int32_t zsba_restore(zsba_graph* g) {
    if (!g) return ZSBA_EINVAL;
    if (!g->have_snapshot || g->structure_dirty) return g->fail(ZSBA_ESTATE, "no snapshot of the current structure");
    if (g->values_dirty) return g->fail(ZSBA_ESTATE, "parameters were set after the snapshot");
    hipError_t e;
    if ((e = g->copy(g->w.pose, g->snap_pose, 7L * g->dp.np * sizeof(double), hipMemcpyDeviceToDevice)) != hipSuccess)
        return g->hip_fail(e, "restore");
    if ((e = g->copy(g->w.point, g->snap_point, 3L * g->dp.nl * sizeof(double), hipMemcpyDeviceToDevice)) != hipSuccess)
        return g->hip_fail(e, "restore");
    g->device_newer = true;
    return 0;
}

int32_t zsba_debug_math(int32_t fn, int32_t n, const double* in, double* out, int32_t on_device) {
    if (n < 0 || (n > 0 && (!in || !out)) || fn < 0 || fn > 3) return ZSBA_EINVAL;
    if (!on_device) {
        for (int32_t i = 0; i < n; ++i)
            out[i] = fn == 0 ? zb::cr_sin(in[i]) : fn == 1 ? zb::cr_cos(in[i]) : fn == 2 ? zb::cr_cube(in[i]) : std::sqrt(in[i]);
        return 0;
    }
    double *d_in = nullptr, *d_out = nullptr;
    if (hipMalloc(&d_in, sizeof(double) * (n + 1)) != hipSuccess) return ZSBA_ENOMEM;
    if (hipMalloc(&d_out, sizeof(double) * (n + 1)) != hipSuccess) { (void)hipFree(d_in); return ZSBA_ENOMEM; }
    int32_t rc = 0;
    if (hipMemcpy(d_in, in, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) rc = ZSBA_EDEVICE;
    if (!rc) {
        zb::launch_math_probe(fn, n, d_in, d_out, nullptr);
        if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(out, d_out, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
            rc = ZSBA_EDEVICE;
    }
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    return rc;
}

zsba_batch* zsba_batch_create(zsba_graph* const* graphs, int32_t count) {
    if (!graphs || count < 0) return nullptr;
    zsba_batch* b = new (std::nothrow) zsba_batch();
    if (!b) return nullptr;
    b->graphs.assign(graphs, graphs + count);
    for (zsba_graph* g : b->graphs) {
        if (!g) {
            delete b;
            return nullptr;
        }
    }
    return b;
}

void zsba_batch_destroy(zsba_batch* b) { delete b; }

const char* zsba_batch_last_error(const zsba_batch* b) { return b ? b->err.c_str() : "null handle"; }

int32_t zsba_batch_solve(zsba_batch* b, int32_t iterations, int32_t use_rel, int32_t precision, int32_t* accepted) {
    if (!b) return ZSBA_EINVAL;
    return b->solve(iterations, use_rel, precision, accepted);
}

int32_t zsba_batch_get_info(const zsba_batch* b, zsba_batch_info* info) {
    if (!b || !info) return ZSBA_EINVAL;
    *info = b->info;
    return 0;
}

int32_t zsba_solve_batch(zsba_graph* const* graphs, int32_t count, int32_t iterations, int32_t use_rel, int32_t precision,
                         int32_t* accepted) {
    if (!graphs || count < 0 || (count > 0 && !accepted)) return ZSBA_EINVAL;
    zsba_batch* b = zsba_batch_create(graphs, count);
    if (!b) return ZSBA_EINVAL;
    const int32_t rc = b->solve(iterations, use_rel, precision, accepted);
    if (rc < 0 && count > 0) graphs[0]->err = b->err;
    delete b;
    return rc;
}

int32_t zsba_debug_adjoint_ops(zsba_graph* g, const double* v, double* Hv, double* Minv_v) {
    if (!g || !v) return ZSBA_EINVAL;
    if (g->live_obs() == 0) return g->fail(ZSBA_ESTATE, "no observations");
    int32_t rc = g->prepare();
    if (rc) return rc;
    hipError_t e;
    const int64_t N = 6L * g->dp.np + 3L * g->dp.nl;
    double* dv = g->w.adj;
    double* out = dv + N;
    if ((e = g->copy(dv, v, N * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) return g->hip_fail(e, "v");
    {
        const double zero = 0.0;
        if ((e = g->copy(g->lambda_dev, &zero, sizeof zero, hipMemcpyHostToDevice)) != hipSuccess)
            return g->hip_fail(e, "lambda");
        zb::launch_linearise64(g->dp, g->w, g->stream, false);
        zb::launch_reduced_system64(g->dp, g->w, g->stream, g->lambda_dev, false);
        zb::launch_cholesky_solve64(g->dp.n, g->w, g->stream, false, false, 1e-10);
        zb::launch_adjoint_frozen(g->dp, g->w, g->stream);
    }
    if (Hv) {
        zb::launch_grad_tangent(g->dp, g->w, dv, nullptr, out, g->stream);
        if ((e = g->copy(Hv, out, N * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess) return g->hip_fail(e, "Hv");
    }
    if (Minv_v) {
        zb::launch_adjoint_precond(g->dp, g->w, dv, out, g->stream);
        if ((e = g->copy(Minv_v, out, N * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
            return g->hip_fail(e, "Minv v");
    }
    return 0;
}

int32_t zsba_gauss_newton(zsba_graph* g, int32_t steps) {
    if (!g || steps < 0) return ZSBA_EINVAL;
    if (g->live_obs() == 0 || steps == 0) return 0;
    int32_t rc = g->prepare();
    if (rc) return rc;
    return g->gauss_newton(steps);
}

int32_t zsba_hold_indefinite_points(zsba_graph* g, double rel_tol, uint8_t* held) {
    if (!g || !(rel_tol >= 0.0) || !(rel_tol < 1.0)) return ZSBA_EINVAL;
    if (g->live_obs() == 0) return 0;
    int32_t rc = g->prepare();
    if (rc) return rc;
    return g->hold_indefinite_points(rel_tol > 0.0 ? rel_tol : 1e-12, held);
}

int32_t zsba_adjoint(zsba_graph* g, const double* dL_dpoints, const double* dL_dposes, double* dL_dobs, double* dL_dK) {
    if (!g) return ZSBA_EINVAL;
    if (!dL_dpoints && !dL_dposes) return g->fail(ZSBA_EINVAL, "no loss gradient given");
    if (g->live_obs() == 0) return g->fail(ZSBA_ESTATE, "no observations");
    int32_t rc = g->prepare();
    if (rc) return rc;
    return g->adjoint(dL_dpoints, dL_dposes, dL_dobs, dL_dK);
}

}  // extern "C"
