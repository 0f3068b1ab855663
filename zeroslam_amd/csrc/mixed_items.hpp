// Per-item device functions of the mixed-precision solve, shared by the
// per-graph kernels (kernels_fp32.hip, kernels_fp64.hip) and the batched
// per-window solver (kernels_batch.hip).  TS is the storage type of the Hessian
// blocks (binary32 on the mixed path, binary64 for a binary64 batch); residuals,
// robust weights, gradients, the reduced system and all state updates are
// binary64.  Each function handles one point, pose, pair or reduced-system block;
// the caller maps items to threads or waves.
//
// Buffer layouts: pose-side terms of observation s in pose-major slot
// obs_pose_pos[s]: 21 H_pp terms (TS) and 6 gradient terms (binary64); H_pp 21
// lower-triangle entries per pose; H_ll and H_ll^-1 9 per point; W and Y 18 per
// (pose, point) pair, row-major 6 x 3.
#pragma once

#include "ba_math.hpp"
#include "common.hpp"

namespace zb {
namespace mixed {

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}
template <class T> __device__ __forceinline__ T wave_sum(T v);
template <> __device__ __forceinline__ float wave_sum<float>(float v) { return wave_sum_f(v); }
template <> __device__ __forceinline__ double wave_sum<double>(double v) { return wave_sum_d(v); }

// Reduce-scatter over the wave by recursive halving.  On entry slot k of every
// lane holds a partial of entry k (M slots).  The step over lane bit b pairs each
// lane with the lane across that bit: the lane with b = 0 keeps the lower half of
// the slots, its partner the upper half, and each adds the partner's partials of
// the half it keeps.  After the six steps the lane holds the wave total of entry
// `off` when `m` is 1 (every entry ends in exactly one lane; m is 0 on the lanes
// left without one).  Slots past a lane's own count only ever meet other such
// slots, so they never reach a total.  M entries cost M/2 + M/4 + ... shuffles
// (38 for M = 36) where one butterfly per entry costs 6 M.
template <class T, int M, int kBit>
struct reduce_scatter_step {
    static __device__ __forceinline__ T run(const T (&v)[M], int lane, int& off, int& m) {
        constexpr int H = (M + 1) / 2;
        const bool hi = (lane >> kBit) & 1;
        T u[H];
#pragma unroll
        for (int k = 0; k < H; ++k) {
            const T lo_v = v[k];
            const T hi_v = (H + k < M) ? v[(H + k < M) ? H + k : 0] : T(0);
            const T send = hi ? lo_v : hi_v;
            const T keep = hi ? hi_v : lo_v;
            u[k] = keep + __shfl_xor(send, 1 << kBit, 64);
        }
        if (hi) {
            off += H;
            m = m > H ? m - H : 0;
        } else {
            m = m < H ? m : H;
        }
        return reduce_scatter_step<T, H, kBit - 1>::run(u, lane, off, m);
    }
};
template <class T, int M>
struct reduce_scatter_step<T, M, -1> {
    static __device__ __forceinline__ T run(const T (&v)[M], int, int&, int&) { return v[0]; }
};

// Wave total of entry `entry` (or -1: this lane holds none) from M per-lane slots.
template <class T, int M>
__device__ __forceinline__ T wave_reduce_scatter(const T (&v)[M], int lane, int& entry) {
    static_assert(M >= 1 && M <= 64, "one entry per lane at most");
    int off = 0, m = M;
    const T total = reduce_scatter_step<T, M, 5>::run(v, lane, off, m);
    entry = m == 1 ? off : -1;
    return total;
}

__device__ __forceinline__ void obs_weight(const dev_problem& p, int s, double W[4]) {
    if (p.identity_info) {
        W[0] = 1.0; W[1] = 0.0; W[2] = 0.0; W[3] = 1.0;
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) W[i] = p.obs_info[4L * s + i];
    }
}

// Robust cost of observation s (edge_base::robust_chi2 of the recomputed
// residual, factor_graph.hpp:221-228, 990-1009); the residual is stored.
__device__ inline double cost_item(const dev_problem& p, const double* pose, const double* point, double* resid, int s) {
    const int32_t param = p.obs_param[s];
    double P[7], X[3], Pc[3], r[2];
    const double uv[2] = {p.obs_uv[2L * s], p.obs_uv[2L * s + 1]};
#pragma unroll
    for (int i = 0; i < 7; ++i) P[i] = pose[7L * p.obs_pose[s] + i];
#pragma unroll
    for (int i = 0; i < 3; ++i) X[i] = point[3L * p.obs_point[s] + i];
    pose_apply(P, X, Pc);
    residual(Pc, uv, p.cams + 4 * (param & 0xFFFF), r);
    resid[2L * s] = r[0];
    resid[2L * s + 1] = r[1];
    double W[4];
    obs_weight(p, s, W);
    const int li = param >> 16;
    return robust_chi2(p.loss_kind[li], p.loss_delta[li], r, W);
}

// Linearisation of observation s between pose parameters P and point X: the
// binary64 residual, information, robust weight and Jacobians in the reference's
// order, and the Jacobians and weighted information rounded to TS.  Shared by
// every kernel that forms Hessian terms, so that recomputed terms equal stored
// ones bit for bit.
template <class TS>
struct obs_lin {
    double r[2], W[4], Wr[4], drho, Jp[12], Jl[6];
    TS Jpf[12], Jlf[6], Wrf[4];
    bool behind;  // the camera cannot project the point (behind-camera residual)
};

template <class TS>
__device__ __forceinline__ void linearise_obs(const dev_problem& p, const double (&P)[7], const double (&X)[3], int s,
                                              const obs_record& rec, obs_lin<TS>& o) {
    const int32_t param = rec.param;
    const double* K = p.cams + 4 * (param & 0xFFFF);
    double Pc[3], q[4];
    pose_apply(P, X, Pc);
    pose_quat(P, q);
    o.behind = behind_camera(Pc);
    const double uv[2] = {rec.u, rec.v};
    residual(Pc, uv, K, o.r);
    obs_weight(p, s, o.W);
    const int li = param >> 16;
    robust_info(p.loss_kind[li], p.loss_delta[li], o.W, o.r, o.drho, o.Wr);
    jacobians(Pc, q, K, o.Jp, o.Jl);
#pragma unroll
    for (int i = 0; i < 12; ++i) o.Jpf[i] = static_cast<TS>(o.Jp[i]);
#pragma unroll
    for (int i = 0; i < 6; ++i) o.Jlf[i] = static_cast<TS>(o.Jl[i]);
#pragma unroll
    for (int i = 0; i < 4; ++i) o.Wrf[i] = static_cast<TS>(o.Wr[i]);
}

// make_hessian (factor_graph.hpp:648-732) for the observations of point l, in
// insertion order: H_ll and the W blocks accumulate in TS, b_l in binary64; with
// kPoseTerms the pose-side terms go to pose-major slots (otherwise
// reduce_pose_chunk_recompute forms them again).  Residual, robust weight and
// Jacobians are binary64 in the reference's order; the Hessian products use the
// Jacobians rounded to TS.  frozen (nullable) receives whether an observation of
// the point is behind its camera.
template <class TS, bool kPoseTerms = true>
__device__ inline void lin_point(const dev_problem& p, const double* pose, const double* point, TS* pterms, double* pgrad,
                                 TS* Hll, double* bl, TS* Wbuf, uint8_t* frozen, int l) {
    bool behind = false;
    const bool lfixed = p.point_fixed[l] != 0;
    const double X[3] = {point[3L * l], point[3L * l + 1], point[3L * l + 2]};
    TS H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double b[3] = {0, 0, 0};
    const int s0 = p.lm_start[l], s1 = p.lm_start[l + 1];
    for (int s = s0; s < s1; ++s) {
        const obs_record rec = p.obs_rec[s];
        const int ip = rec.pose;
        const bool pfixed = p.pose_fixed[ip] != 0;
        if (pfixed && lfixed) continue;
        double P[7];
#pragma unroll
        for (int i = 0; i < 7; ++i) P[i] = pose[7L * ip + i];
        obs_lin<TS> o;
        linearise_obs(p, P, X, s, rec, o);
        behind = behind || o.behind;
        if (!pfixed) {
            TS H21[21], Wpl[18];
            pose_hessian(o.Jpf, o.Jlf, o.Wrf, H21, Wpl);
            if (kPoseTerms) {
                double v[6];
                pose_grad(o.Jp, o.W, o.r, o.drho, v);
                const int64_t slot = p.obs_pose_pos[s];
#pragma unroll
                for (int i = 0; i < 21; ++i) pterms[21 * slot + i] = H21[i];
#pragma unroll
                for (int i = 0; i < 6; ++i) pgrad[6 * slot + i] = v[i];
            }
            if (!lfixed) {
                const int pr = rec.pairf >> 1;
                TS* Wp = Wbuf + 18L * pr;
                if (rec.pairf & 1) {
#pragma unroll
                    for (int i = 0; i < 18; ++i) Wp[i] = Wpl[i];
                } else {
#pragma unroll
                    for (int i = 0; i < 18; ++i) Wp[i] = Wp[i] + Wpl[i];
                }
            }
        }
        if (!lfixed) {
            TS Hl[9];
            double v[3];
            point_hessian(o.Jlf, o.Wrf, Hl);
            point_grad(o.Jl, o.W, o.r, o.drho, v);
#pragma unroll
            for (int i = 0; i < 9; ++i) H[i] = H[i] + Hl[i];
#pragma unroll
            for (int i = 0; i < 3; ++i) b[i] = b[i] - v[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) Hll[9L * l + i] = H[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) bl[3L * l + i] = b[i];
    if (frozen) frozen[l] = behind ? 1 : 0;
}

// ---- linearisation by blocks of observations (one thread each) ---------------
// Block b holds whole points (p.lb_start).  Each thread linearises one
// observation once (linearise_obs, the arithmetic lin_point uses) and stages its
// terms in LDS; then the thread of each point's first observation adds the
// point's terms in insertion order exactly as lin_point does (H_ll, b_l, the W
// blocks of pairs with several observations, the behind flag), and threads
// (entry, term) add the pose terms of each free pose the block touches, in block
// order, into per-entry partials (lin_part_h in TS, lin_part_g in binary64).
// lin_pose_finish then adds a pose's partials over its blocks.  The W terms of
// observations whose pair has several observations pass through wscratch (18
// per observation slot).
template <class TS>
struct lin_stage {
    TS h[kLinBlock][21];     // pose H_pp terms
    double g[kLinBlock][6];  // pose gradient terms
    TS ph[kLinBlock][9];     // point H_ll terms
    double pg[kLinBlock][3]; // point gradient terms
    int32_t pairf[kLinBlock];  // obs_record.pairf, or -2 for a fixed pose
    uint8_t flags[kLinBlock];  // bit 0 behind camera, bit 1 several observations in the pair
    int32_t lo_s[kLinBlock + 1];  // the block's entries: offset ranges
    uint8_t lo_l[kLinBlock];      // and offsets
};

template <class TS>
__device__ inline void lin_block(const dev_problem& p, const double* pose, const double* point, TS* Hll, double* bl,
                                 TS* Wbuf, uint8_t* frozen, TS* part_h, double* part_g, TS* wscratch, int b,
                                 lin_stage<TS>& sh) {
    const int t = threadIdx.x;
    const int s0 = p.lb_start[b], m = p.lb_start[b + 1] - s0;
    const int e0 = p.lp_start[b], ne = p.lp_start[b + 1] - e0;
    const int l0 = p.lo_start[e0];
    // the block's entry table into LDS
    for (int i = t; i <= ne; i += blockDim.x) sh.lo_s[i] = p.lo_start[e0 + i] - l0;
    for (int i = t; i < p.lo_start[e0 + ne] - l0; i += blockDim.x) sh.lo_l[i] = p.lo_list[l0 + i];
    if (t < m) {
        const int s = s0 + t;
        const obs_record rec = p.obs_rec[s];
        const bool pfixed = p.pose_fixed[rec.pose] != 0, lfixed = p.point_fixed[rec.point] != 0;
        bool behind = false, multi = false;
        if (!(pfixed && lfixed)) {
            double P[7];
#pragma unroll
            for (int i = 0; i < 7; ++i) P[i] = pose[7L * rec.pose + i];
            const double X[3] = {point[3L * rec.point], point[3L * rec.point + 1], point[3L * rec.point + 2]};
            obs_lin<TS> o;
            linearise_obs(p, P, X, s, rec, o);
            behind = o.behind;
            if (!pfixed) {
                TS H21[21], Wpl[18];
                double v[6];
                pose_hessian(o.Jpf, o.Jlf, o.Wrf, H21, Wpl);
                pose_grad(o.Jp, o.W, o.r, o.drho, v);
#pragma unroll
                for (int i = 0; i < 21; ++i) sh.h[t][i] = H21[i];
#pragma unroll
                for (int i = 0; i < 6; ++i) sh.g[t][i] = v[i];
                if (!lfixed) {
                    const int pr = rec.pairf >> 1;
                    multi = p.pair_multi[pr] != 0;
                    if (multi) {
#pragma unroll
                        for (int i = 0; i < 18; ++i) wscratch[18L * s + i] = Wpl[i];
                    } else {
                        TS* Wp = Wbuf + 18L * pr;
#pragma unroll
                        for (int i = 0; i < 18; ++i) Wp[i] = Wpl[i];
                    }
                }
            }
            if (!lfixed) {
                TS Hl[9];
                double v[3];
                point_hessian(o.Jlf, o.Wrf, Hl);
                point_grad(o.Jl, o.W, o.r, o.drho, v);
#pragma unroll
                for (int i = 0; i < 9; ++i) sh.ph[t][i] = Hl[i];
#pragma unroll
                for (int i = 0; i < 3; ++i) sh.pg[t][i] = v[i];
            }
        }
        sh.pairf[t] = pfixed ? -2 : rec.pairf;
        sh.flags[t] = static_cast<uint8_t>((behind ? 1 : 0) | (multi ? 2 : 0));
    }
    __syncthreads();
    // point sums: thread (point, term) adds the point's terms in insertion order
    // (terms 0-8 H_ll, 9-11 b_l; term 12 the behind flag and the W blocks of
    // pairs with several observations)
    const int pt0 = p.lb_point[b], npt = p.lb_point[b + 1] - pt0;
    for (int q = t; q < npt * 13; q += blockDim.x) {
        const int l = pt0 + q / 13, term = q - (q / 13) * 13;
        const bool lfixed = p.point_fixed[l] != 0;
        const int k0 = p.lm_start[l] - s0, k1 = p.lm_start[l + 1] - s0;
        if (term < 9) {
            TS acc = 0;
            if (!lfixed)
                for (int k = k0; k < k1; ++k) acc = acc + sh.ph[k][term];
            Hll[9L * l + term] = acc;
        } else if (term < 12) {
            double acc = 0.0;
            if (!lfixed)
                for (int k = k0; k < k1; ++k) acc = acc - sh.pg[k][term - 9];
            bl[3L * l + (term - 9)] = acc;
        } else {
            bool behind = false;
            for (int k = k0; k < k1; ++k) {
                const int pf = sh.pairf[k];
                const bool pfixed = pf == -2;
                if (pfixed && lfixed) continue;
                behind = behind || (sh.flags[k] & 1) != 0;
                if (!pfixed && !lfixed && (sh.flags[k] & 2)) {
                    TS* Wp = Wbuf + 18L * (pf >> 1);
                    const TS* Wo = wscratch + 18L * (s0 + k);
                    if (pf & 1) {
#pragma unroll
                        for (int i = 0; i < 18; ++i) Wp[i] = Wo[i];
                    } else {
#pragma unroll
                        for (int i = 0; i < 18; ++i) Wp[i] = Wp[i] + Wo[i];
                    }
                }
            }
            if (frozen) frozen[l] = behind ? 1 : 0;
        }
    }
    // pose partials, written at the entry's pose-major slot
    for (int q = t; q < ne * 27; q += blockDim.x) {
        const int ei = q / 27, j = q - ei * 27;
        const int o0 = sh.lo_s[ei], o1 = sh.lo_s[ei + 1];
        const int64_t slot = p.e_slot[e0 + ei];
        if (j < 21) {
            TS acc = 0;
            for (int o = o0; o < o1; ++o) acc = acc + sh.h[sh.lo_l[o]][j];
            part_h[21 * slot + j] = acc;
        } else {
            double acc = 0.0;
            for (int o = o0; o < o1; ++o) acc = acc + sh.g[sh.lo_l[o]][j - 21];
            part_g[6 * slot + (j - 21)] = acc;
        }
    }
}

// Pose i by one workgroup of kFinishThreads: thread t adds the partials of slots
// pe_start[i] + t, + kFinishThreads, ... (binary64, 27 terms), each wave adds its
// lanes by a fixed butterfly, and the waves' sums are added in wave order; H_pp
// terms are rounded to TS, b_p = -(sum of gradient terms).  red: 4 x 27 doubles.
constexpr int kFinishThreads = 256;
template <class TS>
__device__ inline void lin_pose_finish(const dev_problem& p, const TS* part_h, const double* part_g, TS* Hpp,
                                       double* bp, int i, double (*red)[27]) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int k0 = p.pe_start[i], k1 = p.pe_start[i + 1];
    double acc[27];
#pragma unroll
    for (int j = 0; j < 27; ++j) acc[j] = 0.0;
    for (int k = k0 + t; k < k1; k += kFinishThreads) {
#pragma unroll
        for (int j = 0; j < 21; ++j) acc[j] += static_cast<double>(part_h[21L * k + j]);
#pragma unroll
        for (int j = 0; j < 6; ++j) acc[21 + j] += part_g[6L * k + j];
    }
#pragma unroll
    for (int j = 0; j < 27; ++j) {
        double v = acc[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
        if (lane == 0) red[wave][j] = v;
    }
    __syncthreads();
    if (t < 27) {
        const double sum = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
        if (t < 21) Hpp[21L * i + t] = static_cast<TS>(sum);
        else bp[6L * i + (t - 21)] = -sum;
    }
}

// Chunk c of [t0, t1) split into kPoseChunks contiguous pieces.
__device__ __forceinline__ void chunk_range(int t0, int t1, int c, int& c0, int& c1) {
    const int64_t len = t1 - t0;
    c0 = t0 + static_cast<int>(len * c / kPoseChunks);
    c1 = t0 + static_cast<int>(len * (c + 1) / kPoseChunks);
}

// Wave for (pose i, chunk c): lane j < 27 sums H_pp term j (j < 21) or gradient
// term j - 21 over the chunk's slots into part[(i kPoseChunks + c) 27 + j].
template <class TS>
__device__ inline void reduce_pose_chunk(const dev_problem& p, const TS* pterms, const double* pgrad, double* part,
                                         int i, int c, int lane) {
    const bool fixed = p.pose_fixed[i] != 0;
    int t0, t1;
    chunk_range(p.pose_start[i], fixed ? p.pose_start[i] : p.pose_start[i + 1], c, t0, t1);
    constexpr int U = 8;
    double* out = part + (static_cast<int64_t>(i) * kPoseChunks + c) * 27;
    if (lane < 21) {
        const TS* src = pterms + lane;
        TS acc[4] = {0, 0, 0, 0};
        int t = t0;
        for (; t + U <= t1; t += U) {
            TS v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = src[21L * (t + u)];
#pragma unroll
            for (int u = 0; u < U; ++u) acc[u & 3] += v[u];
        }
        for (; t < t1; ++t) acc[0] += src[21L * t];
        out[lane] = static_cast<double>((acc[0] + acc[1]) + (acc[2] + acc[3]));
    } else if (lane < 27) {
        const double* src = pgrad + (lane - 21);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        int t = t0;
        for (; t + U <= t1; t += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = src[6L * (t + u)];
#pragma unroll
            for (int u = 0; u < U; ++u) acc[u & 3] += v[u];
        }
        for (; t < t1; ++t) acc[0] += src[6L * t];
        out[lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    }
}

// reduce_pose_chunk without the stored terms: the wave forms the pose terms of
// 64 slots at a time (lane = slot, linearise_obs + pose_hessian + pose_grad, the
// arithmetic lin_point stores), stages them in LDS (ht: 64 x 21 TS, gt: 64 x 6
// binary64 per wave), and lane j < 27 adds term j over the slots in
// reduce_pose_chunk's order (accumulator (t - t0) mod 4 over whole blocks of 8
// slots, the tail into accumulator 0), so part receives the same values.  The
// caller's wave must be convergent (all 64 lanes call with the same i, c).
template <class TS>
__device__ inline void reduce_pose_chunk_recompute(const dev_problem& p, const double* pose, const double* point,
                                                   double* part, int i, int c, int lane, TS (*ht)[21],
                                                   double (*gt)[6]) {
    const bool fixed = p.pose_fixed[i] != 0;
    int t0, t1;
    chunk_range(p.pose_start[i], fixed ? p.pose_start[i] : p.pose_start[i + 1], c, t0, t1);
    const int full = t0 + ((t1 - t0) & ~7);  // slots before this are in whole blocks of 8
    double P[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) P[k] = pose[7L * i + k];
    TS hacc[4] = {0, 0, 0, 0};
    double gacc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int base = t0; base < t1; base += 64) {
        const int t = base + lane;
        if (t < t1) {
            const int s = p.pose_list[t];
            const obs_record rec = p.obs_rec[s];
            const int l = rec.point;
            const double X[3] = {point[3L * l], point[3L * l + 1], point[3L * l + 2]};
            obs_lin<TS> o;
            linearise_obs(p, P, X, s, rec, o);
            TS H21[21], Wpl[18];
            double v[6];
            pose_hessian(o.Jpf, o.Jlf, o.Wrf, H21, Wpl);
            pose_grad(o.Jp, o.W, o.r, o.drho, v);
#pragma unroll
            for (int k = 0; k < 21; ++k) ht[lane][k] = H21[k];
#pragma unroll
            for (int k = 0; k < 6; ++k) gt[lane][k] = v[k];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int m = (t1 - base) < 64 ? (t1 - base) : 64;
        if (lane < 21) {
            for (int k = 0; k < m; ++k) {
                const int tt = base + k;
                hacc[tt < full ? ((tt - t0) & 3) : 0] += ht[k][lane];
            }
        } else if (lane < 27) {
            for (int k = 0; k < m; ++k) {
                const int tt = base + k;
                gacc[tt < full ? ((tt - t0) & 3) : 0] += gt[k][lane - 21];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    double* out = part + (static_cast<int64_t>(i) * kPoseChunks + c) * 27;
    if (lane < 21) out[lane] = static_cast<double>((hacc[0] + hacc[1]) + (hacc[2] + hacc[3]));
    else if (lane < 27) out[lane] = (gacc[0] + gacc[1]) + (gacc[2] + gacc[3]);
}

// Thread for (pose i, term j < 27): the chunk partials added in order; H_pp term
// (TS) or b_p = -(sum of gradient terms).
template <class TS>
__device__ inline void reduce_pose_finish(const double* part, TS* Hpp, double* bp, int i, int j) {
    const double* src = part + static_cast<int64_t>(i) * kPoseChunks * 27 + j;
    double v[kPoseChunks];
#pragma unroll
    for (int c = 0; c < kPoseChunks; ++c) v[c] = src[27 * c];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < kPoseChunks; ++c) a[c & 3] += v[c];
    const double sum = (a[0] + a[1]) + (a[2] + a[3]);
    if (j < 21) Hpp[21L * i + j] = static_cast<TS>(sum);
    else bp[6L * i + (j - 21)] = -sum;
}

// Damped point block and its inverse (factor_graph.hpp:742-765).
template <class TS>
__device__ inline void damp_invert_item(const dev_problem& p, const TS* Hll, TS* Hinv, const uint8_t* frozen,
                                        double lambda, int l) {
    const TS lam = static_cast<TS>(lambda);
    TS H[9], inv[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = Hll[9L * l + i];
    H[0] += lam; H[4] += lam; H[8] += lam;
    if (p.point_fixed[l] || (frozen && frozen[l])) {
#pragma unroll
        for (int i = 0; i < 9; ++i) inv[i] = TS(0);
    } else {
        invert3(H, inv);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) Hinv[9L * l + i] = inv[i];
}

// Y = W H_ll^-1 for pair a (matrix_sparse_block_diagonal.hpp:107-122).
template <class TS>
__device__ inline void form_Y_item(const dev_problem& p, const TS* Wbuf, const TS* Hinv, TS* Y, int a) {
    TS Wa[18], Hi[9], o[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) Wa[i] = Wbuf[18L * a + i];
    const int l = p.pair_point[a];
#pragma unroll
    for (int i = 0; i < 9; ++i) Hi[i] = Hinv[9L * l + i];
    mul<6, 3, 3>(Wa, Hi, o);
#pragma unroll
    for (int i = 0; i < 18; ++i) Y[18L * a + i] = o[i];
}

// One workgroup of kSchurWaves waves per 6x6 lower-triangle block of the
// reduced camera system: the products Y_a W_c^T accumulate in TS per lane, Y b_l
// in binary64; the block is assembled in binary64:
//   S_ik = [i == k] (H_pp,i + lambda I) - sum Y W^T,  b_s,i = b_p,i - sum Y b_l
// (factor_graph.hpp:768-781, matrix_sparse_block.hpp:158-196).
// Lane j of wave w adds contributions c0 + 64 w + j, then + 64 kSchurWaves, ...
// in order; each wave reduce-scatters its 36 (and, on a diagonal block, 6)
// sums, and the waves' totals are added in wave order in binary64.
// kInFlight: contributions per lane loaded before they are added (2 hides the
// re-reads' latency; 1 needs fewer registers); the result is the same.
constexpr int kSchurWaves = 4;
constexpr int kSchurThreads = 64 * kSchurWaves;

template <class TS>
struct schur_lds {
    TS s[kSchurWaves][36];
    double b[kSchurWaves][6];
};

template <class TS, int kInFlight = 2>
__device__ inline void schur_block(const dev_problem& p, const TS* Y, const TS* Wbuf, const TS* Hpp, const double* bp,
                                   const double* bl, double* S, double* bs, double lam, int blk, int tid,
                                   schur_lds<TS>& red) {
    const int lane = tid & 63, wave = tid >> 6;
    const int row = p.blk_row[blk], col = p.blk_col[blk];
    const bool diag = row == col;
    TS acc[36];
#pragma unroll
    for (int i = 0; i < 36; ++i) acc[i] = TS(0);
    double accb[6] = {0, 0, 0, 0, 0, 0};
    const int64_t c0 = p.blk_start[blk], c1 = p.blk_start[blk + 1];
    constexpr int kStride = kSchurThreads;
    // contributions c, c + kStride of this lane: two in flight (loads first), added in order
    auto add = [&](const TS (&Ya)[18], const TS (&Wc)[18], const double (&bv)[3]) {
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int cc = 0; cc < 6; ++cc)
                acc[r * 6 + cc] += Ya[r * 3] * Wc[cc * 3] + Ya[r * 3 + 1] * Wc[cc * 3 + 1] + Ya[r * 3 + 2] * Wc[cc * 3 + 2];
        if (diag) {
#pragma unroll
            for (int r = 0; r < 6; ++r)
                accb[r] += static_cast<double>(Ya[r * 3]) * bv[0] + static_cast<double>(Ya[r * 3 + 1]) * bv[1] +
                           static_cast<double>(Ya[r * 3 + 2]) * bv[2];
        }
    };
    auto fetch = [&](int64_t c, TS (&Ya)[18], TS (&Wc)[18], double (&bv)[3]) {
        const int2 ab = p.contrib[c];
#pragma unroll
        for (int i = 0; i < 18; ++i) Ya[i] = Y[18L * ab.x + i];
#pragma unroll
        for (int i = 0; i < 18; ++i) Wc[i] = Wbuf[18L * ab.y + i];
        bv[0] = bv[1] = bv[2] = 0.0;
        if (diag) {
            const int l = p.pair_point[ab.x];
            bv[0] = bl[3L * l];
            bv[1] = bl[3L * l + 1];
            bv[2] = bl[3L * l + 2];
        }
    };
    int64_t c = c0 + 64 * wave + lane;
    if (kInFlight == 2) {
        for (; c + kStride < c1; c += 2 * kStride) {
            TS Ya0[18], Wc0[18], Ya1[18], Wc1[18];
            double b0[3], b1[3];
            fetch(c, Ya0, Wc0, b0);
            fetch(c + kStride, Ya1, Wc1, b1);
            add(Ya0, Wc0, b0);
            add(Ya1, Wc1, b1);
        }
    }
    for (; c < c1; c += kStride) {
        TS Ya0[18], Wc0[18];
        double b0[3];
        fetch(c, Ya0, Wc0, b0);
        add(Ya0, Wc0, b0);
    }
    // Waves with no contributions hold zeros and still take part: every wave's
    // total enters the sum, so the result does not depend on the list length.
    int e;
    const TS tot = wave_reduce_scatter<TS, 36>(acc, lane, e);
    if (e >= 0) red.s[wave][e] = tot;
    if (diag) {
        int eb;
        const double totb = wave_reduce_scatter<double, 6>(accb, lane, eb);
        if (eb >= 0) red.b[wave][eb] = totb;
    }
    __syncthreads();
    if (tid < 36) {
        const int r = tid / 6, cc = tid % 6;
        double sum = static_cast<double>(red.s[0][tid]);
#pragma unroll
        for (int w = 1; w < kSchurWaves; ++w) sum += static_cast<double>(red.s[w][tid]);
        const int64_t n = p.n;
        double* Sblk = S + (6L * row) * n + 6L * col;
        if (diag) {
            if (cc <= r) {
                double base = static_cast<double>(Hpp[21L * row + r * (r + 1) / 2 + cc]);
                if (r == cc) base += lam;
                Sblk[r * n + cc] = base - sum;
            }
        } else {
            Sblk[r * n + cc] = -sum;
        }
    } else if (diag && tid >= 64 && tid < 70) {
        const int i = tid - 64;
        double sum = red.b[0][i];
#pragma unroll
        for (int w = 1; w < kSchurWaves; ++w) sum += red.b[w][i];
        bs[6L * row + i] = bp[6L * row + i] - sum;
    }
}

// ---- iterative refinement of the mixed solve in binary64 ------------------------
// The step (dx_p, dx_l) solved with the TS-block reduced system is refined
// against the binary64 normal equations of the current linearisation, applied
// without forming them: per observation the Jacobians and robust weight are
// recomputed in binary64 (the linearisation's sequence) and
//   r_p,i = b_p,i - lambda dx_p,i - sum_obs J_p^T Omega' (J_p dx_p,i + J_l dx_l,l)
//   r_l,l = b_l,l - lambda dx_l,l - sum_obs J_l^T Omega' (J_p dx_p,i + J_l dx_l,l)
// (rows of fixed vertices are zero).  The correction solves the TS-block system
// for (r_p, r_l) by the same elimination:  rs = r_p - W H_ll^-1 r_l,
// delta_p = S^-1 rs (the Cholesky factor of S), delta_l = H_ll^-1 (r_l - W^T delta_p).

// dx_l = H_ll^-1 (b_l - sum_a W_a^T dx_p(a)) for point l, into o.
template <class TS>
__device__ inline void point_step_value(const dev_problem& p, const TS* Wbuf, const TS* Hinv, const double* bl,
                                        const double* dxp, int l, double (&o)[3]) {
    double t[3] = {0.0, 0.0, 0.0};
    for (int a = p.lm_pair_start[l]; a < p.lm_pair_start[l + 1]; ++a) {
        const int k = p.pair_pose[a];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double v = 0.0;
#pragma unroll
            for (int r = 0; r < 6; ++r) v += static_cast<double>(Wbuf[18L * a + 3 * r + c]) * dxp[6L * k + r];
            t[c] += v;
        }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) v += static_cast<double>(Hinv[9L * l + 3 * r + c]) * (bl[3L * l + c] - t[c]);
        o[r] = v;
    }
}

template <class TS>
__device__ inline void point_step_item(const dev_problem& p, const TS* Wbuf, const TS* Hinv, const double* bl,
                                       const double* dxp, double* dxl, int l) {
    double o[3];
    point_step_value(p, Wbuf, Hinv, bl, dxp, l, o);
#pragma unroll
    for (int r = 0; r < 3; ++r) dxl[3L * l + r] = o[r];
}

// Point l: r_l and g_l = H_ll^-1 r_l; the pose-side products J_p^T Omega' u of
// each observation go to its pose-major slot.
template <class TS>
__device__ inline void full_residual_item(const dev_problem& p, const double* pose, const double* point, const double* dxp,
                                          const double* dxl, const double* bl, const TS* Hinv, double* pslots, double* rl,
                                          double* gl, double lam, int l) {
    const bool lfixed = p.point_fixed[l] != 0;
    const double X[3] = {point[3L * l], point[3L * l + 1], point[3L * l + 2]};
    const double dl[3] = {dxl[3L * l], dxl[3L * l + 1], dxl[3L * l + 2]};
    double acc[3] = {0.0, 0.0, 0.0};
    for (int s = p.lm_start[l]; s < p.lm_start[l + 1]; ++s) {
        const int ip = p.obs_pose[s];
        const bool pfixed = p.pose_fixed[ip] != 0;
        if (pfixed && lfixed) continue;
        const int32_t param = p.obs_param[s];
        const double* K = p.cams + 4 * (param & 0xFFFF);
        double P[7], Pc[3], q[4], r[2], Jp[12], Jl[6], W[4], Wr[4], drho;
#pragma unroll
        for (int i = 0; i < 7; ++i) P[i] = pose[7L * ip + i];
        pose_apply(P, X, Pc);
        const double uv[2] = {p.obs_uv[2L * s], p.obs_uv[2L * s + 1]};
        residual(Pc, uv, K, r);
        pose_quat(P, q);
        jacobians(Pc, q, K, Jp, Jl);
        obs_weight(p, s, W);
        const int li = param >> 16;
        robust_info(p.loss_kind[li], p.loss_delta[li], W, r, drho, Wr);
        double u[2] = {0.0, 0.0};
        if (!pfixed) {
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int k = 0; k < 6; ++k) u[m] += Jp[6 * m + k] * dxp[6L * ip + k];
        }
        if (!lfixed) {
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int k = 0; k < 3; ++k) u[m] += Jl[3 * m + k] * dl[k];
        }
        const double wu[2] = {Wr[0] * u[0] + Wr[1] * u[1], Wr[2] * u[0] + Wr[3] * u[1]};
        if (!pfixed) {
            double* out = pslots + 6L * p.obs_pose_pos[s];
#pragma unroll
            for (int k = 0; k < 6; ++k) out[k] = Jp[k] * wu[0] + Jp[6 + k] * wu[1];
        }
        if (!lfixed) {
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[k] += Jl[k] * wu[0] + Jl[3 + k] * wu[1];
        }
    }
    double rr[3] = {0.0, 0.0, 0.0};
    if (!lfixed) {
#pragma unroll
        for (int k = 0; k < 3; ++k) rr[k] = bl[3L * l + k] - lam * dl[k] - acc[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) rl[3L * l + k] = rr[k];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double o = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) o += static_cast<double>(Hinv[9L * l + 3 * r + c]) * rr[c];
        gl[3L * l + r] = o;
    }
}

// Wave for (pose i, chunk c): the chunk's share of sum_t slots_t + sum_a W_a g_l(a)
// into part[(i kPoseChunks + c) 6 + r].
template <class TS>
__device__ inline void pose_residual_chunk(const dev_problem& p, const double* pslots, const TS* Wbuf, const double* gl,
                                           double* part, int i, int c, int lane) {
    const bool fixed = p.pose_fixed[i] != 0;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (!fixed) {
        int t0, t1, a0, a1;
        chunk_range(p.pose_start[i], p.pose_start[i + 1], c, t0, t1);
        chunk_range(p.pose_pair_start[i], p.pose_pair_start[i + 1], c, a0, a1);
        for (int t = t0 + lane; t < t1; t += 64) {
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[k] += pslots[6L * t + k];
        }
        for (int t = a0 + lane; t < a1; t += 64) {
            const int a = p.pose_pairs[t];
            const int l = p.pair_point[a];
            const double g0 = gl[3L * l], g1 = gl[3L * l + 1], g2 = gl[3L * l + 2];
            const TS* Wa = Wbuf + 18L * a;
#pragma unroll
            for (int r = 0; r < 6; ++r)
                acc[r] += static_cast<double>(Wa[3 * r]) * g0 + static_cast<double>(Wa[3 * r + 1]) * g1 +
                          static_cast<double>(Wa[3 * r + 2]) * g2;
        }
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) acc[r] = wave_sum_d(acc[r]);
    if (lane < 6) {
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) a = (q == lane) ? acc[q] : a;
        part[(static_cast<int64_t>(i) * kPoseChunks + c) * 6 + lane] = a;
    }
}

// Thread for (pose i, row r < 6): rs = r_p - lambda dx_p - (chunk partials in order).
__device__ inline void pose_residual_finish(const dev_problem& p, const double* part, const double* bp,
                                            const double* dxp, double* rs, double lam, int i, int r) {
    const double* src = part + static_cast<int64_t>(i) * kPoseChunks * 6 + r;
    double v[kPoseChunks];
#pragma unroll
    for (int c = 0; c < kPoseChunks; ++c) v[c] = src[6 * c];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < kPoseChunks; ++c) a[c & 3] += v[c];
    const double sum = (a[0] + a[1]) + (a[2] + a[3]);
    rs[6L * i + r] = p.pose_fixed[i] ? 0.0 : bp[6L * i + r] - lam * dxp[6L * i + r] - sum;
}

// dx_l += H_ll^-1 (r_l - sum_a W_a^T delta_p(a)) for point l.
template <class TS>
__device__ inline void refine_point_item(const dev_problem& p, const TS* Wbuf, const TS* Hinv, const double* rl,
                                         const double* corr, double* dxl, int l) {
    double tt[3] = {0.0, 0.0, 0.0};
    for (int a = p.lm_pair_start[l]; a < p.lm_pair_start[l + 1]; ++a) {
        const int k = p.pair_pose[a];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double v = 0.0;
#pragma unroll
            for (int r = 0; r < 6; ++r) v += static_cast<double>(Wbuf[18L * a + 3 * r + c]) * corr[6L * k + r];
            tt[c] += v;
        }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double o = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) o += static_cast<double>(Hinv[9L * l + 3 * r + c]) * (rl[3L * l + c] - tt[c]);
        dxl[3L * l + r] += o;
    }
}

// ---- the refinement residual by blocks of observations -----------------------
// Round k of the refinement on linearisation block b (whole points, one thread
// per observation, kLinBlock threads).  Round 0 first forms each point's step
// dx_l = H_ll^-1 (b_l - sum_a W_a^T dx_p) (point_step_value); later rounds read
// dx_l.  Each observation's Jacobians, robust weight, u = J_p dx_p + J_l dx_l and
// its pose and point products are formed as full_residual_item forms them; the
// point sums (insertion order) give r_l and g_l = H_ll^-1 r_l exactly as there.
// Then each free pose's share of  sum_obs J_p^T Omega' u + sum_a W_a g_l(a)  over
// the block (the W term at its pair's first observation, the observations in
// block order) goes to the entry's slot part[6 e_slot + r]; refine_pose_finish
// adds a pose's slots.
struct refine_stage {
    double pv[kLinBlock][6];   // pose-side terms of each observation
    double lv[kLinBlock][3];   // point-side terms
    double dl[kLinBlock][3];   // dx_l of the block's points
    double gl[kLinBlock][3];   // g_l of the block's points
    int32_t pairf[kLinBlock];  // obs_record.pairf, or -2 for a fixed pose
    int32_t pt[kLinBlock];     // the observation's point offset in the block
    int32_t lo_s[kLinBlock + 1];
    uint8_t lo_l[kLinBlock];
};

template <class TS>
__device__ inline void refine_block(const dev_problem& p, const double* pose, const double* point, const TS* Wbuf,
                                    const TS* Hinv, const double* bl, const double* dxp, double* dxl, double* rl,
                                    double* gl, double* part, double lam, bool step_first, int b, refine_stage& sh) {
    const int t = threadIdx.x;
    const int s0 = p.lb_start[b], m = p.lb_start[b + 1] - s0;
    const int e0 = p.lp_start[b], ne = p.lp_start[b + 1] - e0;
    const int l0 = p.lo_start[e0];
    const int pt0 = p.lb_point[b], npt = p.lb_point[b + 1] - pt0;
    for (int i = t; i <= ne; i += blockDim.x) sh.lo_s[i] = p.lo_start[e0 + i] - l0;
    for (int i = t; i < p.lo_start[e0 + ne] - l0; i += blockDim.x) sh.lo_l[i] = p.lo_list[l0 + i];
    for (int q = t; q < npt; q += blockDim.x) {
        const int l = pt0 + q;
        double d[3];
        if (step_first) {
            point_step_value(p, Wbuf, Hinv, bl, dxp, l, d);
#pragma unroll
            for (int c = 0; c < 3; ++c) dxl[3L * l + c] = d[c];
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) d[c] = dxl[3L * l + c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) sh.dl[q][c] = d[c];
    }
    __syncthreads();
    if (t < m) {
        const int s = s0 + t;
        const obs_record rec = p.obs_rec[s];
        const bool pfixed = p.pose_fixed[rec.pose] != 0, lfixed = p.point_fixed[rec.point] != 0;
        const int q = rec.point - pt0;
        double pv[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, lv[3] = {0.0, 0.0, 0.0};
        if (!(pfixed && lfixed)) {
            const double* K = p.cams + 4 * (rec.param & 0xFFFF);
            double P[7], Pc[3], qt[4], r[2], Jp[12], Jl[6], W[4], Wr[4], drho;
#pragma unroll
            for (int i = 0; i < 7; ++i) P[i] = pose[7L * rec.pose + i];
            const double X[3] = {point[3L * rec.point], point[3L * rec.point + 1], point[3L * rec.point + 2]};
            pose_apply(P, X, Pc);
            const double uv[2] = {rec.u, rec.v};
            residual(Pc, uv, K, r);
            pose_quat(P, qt);
            jacobians(Pc, qt, K, Jp, Jl);
            obs_weight(p, s, W);
            const int li = rec.param >> 16;
            robust_info(p.loss_kind[li], p.loss_delta[li], W, r, drho, Wr);
            double u[2] = {0.0, 0.0};
            if (!pfixed) {
#pragma unroll
                for (int mm = 0; mm < 2; ++mm)
#pragma unroll
                    for (int k = 0; k < 6; ++k) u[mm] += Jp[6 * mm + k] * dxp[6L * rec.pose + k];
            }
            if (!lfixed) {
#pragma unroll
                for (int mm = 0; mm < 2; ++mm)
#pragma unroll
                    for (int k = 0; k < 3; ++k) u[mm] += Jl[3 * mm + k] * sh.dl[q][k];
            }
            const double wu[2] = {Wr[0] * u[0] + Wr[1] * u[1], Wr[2] * u[0] + Wr[3] * u[1]};
            if (!pfixed) {
#pragma unroll
                for (int k = 0; k < 6; ++k) pv[k] = Jp[k] * wu[0] + Jp[6 + k] * wu[1];
            }
            if (!lfixed) {
#pragma unroll
                for (int k = 0; k < 3; ++k) lv[k] = Jl[k] * wu[0] + Jl[3 + k] * wu[1];
            }
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) sh.pv[t][k] = pv[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) sh.lv[t][k] = lv[k];
        sh.pairf[t] = pfixed ? -2 : rec.pairf;
        sh.pt[t] = q;
    }
    __syncthreads();
    // point sums in insertion order: r_l, g_l
    for (int q = t; q < npt; q += blockDim.x) {
        const int l = pt0 + q;
        const bool lfixed = p.point_fixed[l] != 0;
        const int k0 = p.lm_start[l] - s0, k1 = p.lm_start[l + 1] - s0;
        double rr[3] = {0.0, 0.0, 0.0};
        if (!lfixed) {
            double acc[3] = {0.0, 0.0, 0.0};
            for (int k = k0; k < k1; ++k) {
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[c] += sh.lv[k][c];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) rr[c] = bl[3L * l + c] - lam * sh.dl[q][c] - acc[c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) rl[3L * l + c] = rr[c];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double o = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) o += static_cast<double>(Hinv[9L * l + 3 * r + c]) * rr[c];
            gl[3L * l + r] = o;
            sh.gl[q][r] = o;
        }
    }
    __syncthreads();
    // W_a g_l at the first observation of each W block
    if (t < m) {
        const int pf = sh.pairf[t];
        if (pf >= 0 && (pf & 1)) {
            const TS* Wa = Wbuf + 18L * (pf >> 1);
            const int q = sh.pt[t];
            const double g0 = sh.gl[q][0], g1 = sh.gl[q][1], g2 = sh.gl[q][2];
#pragma unroll
            for (int r = 0; r < 6; ++r)
                sh.pv[t][r] += static_cast<double>(Wa[3 * r]) * g0 + static_cast<double>(Wa[3 * r + 1]) * g1 +
                               static_cast<double>(Wa[3 * r + 2]) * g2;
        }
    }
    __syncthreads();
    // pose partials, written at the entry's pose-major slot
    for (int q = t; q < ne * 6; q += blockDim.x) {
        const int ei = q / 6, r = q - ei * 6;
        double acc = 0.0;
        for (int o = sh.lo_s[ei]; o < sh.lo_s[ei + 1]; ++o) acc += sh.pv[sh.lo_l[o]][r];
        part[6L * p.e_slot[e0 + ei] + r] = acc;
    }
}

// Pose i by one wave: lane-strided sums of its slots, a fixed butterfly, then
// rs = r_p - lambda dx_p - sum (zero for a fixed pose).
__device__ inline void refine_pose_finish(const dev_problem& p, const double* part, const double* bp,
                                          const double* dxp, double* rs, double lam, int i, int lane) {
    const bool fixed = p.pose_fixed[i] != 0;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (!fixed)
        for (int k = p.pe_start[i] + lane; k < p.pe_start[i + 1]; k += 64) {
#pragma unroll
            for (int r = 0; r < 6; ++r) acc[r] += part[6L * k + r];
        }
#pragma unroll
    for (int r = 0; r < 6; ++r) acc[r] = wave_sum_d(acc[r]);
    if (lane < 6) {
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) a = (q == lane) ? acc[q] : a;
        rs[6L * i + lane] = fixed ? 0.0 : bp[6L * i + lane] - lam * dxp[6L * i + lane] - a;
    }
}

// ---- state update with the solved step, factor_graph.hpp:556-560, 809-830 ------
// Slot t < np: pose t (T <- exp(dx) T); slot np + l: point l (X += dx).  Returns
// the slot's LM scale terms sum_i dx_i (lambda dx_i + b_i).
__device__ inline double apply_step_item(const dev_problem& p, double* pose, double* pose_backup, double* point,
                                         double* point_backup, const double* dxp, const double* dxl, const double* bp,
                                         const double* bl, double lam, int t) {
    double acc = 0.0;
    if (t < p.np) {
        double d[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) d[i] = dxp[6L * t + i];
#pragma unroll
        for (int i = 0; i < 6; ++i) acc = acc + d[i] * (lam * d[i] + bp[6L * t + i]);
        if (!p.pose_fixed[t]) {
            double P[7];
#pragma unroll
            for (int i = 0; i < 7; ++i) P[i] = pose[7L * t + i];
#pragma unroll
            for (int i = 0; i < 7; ++i) pose_backup[7L * t + i] = P[i];
            pose_plus(P, d);
#pragma unroll
            for (int i = 0; i < 7; ++i) pose[7L * t + i] = P[i];
        }
    } else if (t < p.np + p.nl) {
        const int l = t - p.np;
        double dl[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) dl[i] = dxl[3L * l + i];
#pragma unroll
        for (int i = 0; i < 3; ++i) acc = acc + dl[i] * (lam * dl[i] + bl[3L * l + i]);
        if (!p.point_fixed[l]) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                point_backup[3L * l + i] = point[3L * l + i];
                point[3L * l + i] = point[3L * l + i] + dl[i];
            }
        }
    }
    return acc;
}

__device__ inline void rollback_item(const dev_problem& p, double* pose, const double* pose_backup, double* point,
                                     const double* point_backup, int t) {
    if (t < p.np) {
        if (!p.pose_fixed[t])
            for (int i = 0; i < 7; ++i) pose[7L * t + i] = pose_backup[7L * t + i];
    } else if (t < p.np + p.nl) {
        const int l = t - p.np;
        if (!p.point_fixed[l])
            for (int i = 0; i < 3; ++i) point[3L * l + i] = point_backup[3L * l + i];
    }
}

}  // namespace mixed
}  // namespace zb
