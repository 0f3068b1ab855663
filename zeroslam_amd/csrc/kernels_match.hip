// Projection-based data association of a new frame against the map
// (slam.hpp:341-384) with the 256-bit Hamming distance of feature.hpp:1134-1188.
//
// The reference walks the landmarks in the map's iteration order and, for each
// landmark in front of the frame and inside its 10-pixel border that the frame
// does not observe yet, every matched keypoint of the frame in index order.  A
// keypoint within 2 px of the landmark's projection (squared distance not above
// 4), whose previous-frame index has not been taken, and whose descriptor is
// within 63 bits of the landmark's first observation's descriptor, becomes an
// observation; its previous-frame index is then taken.
//
// On the device the test is split in two exact halves (section "banded search"
// below gives the fast form of the first; the form described here is the one
// that handles every input):
//   1. every (landmark, keypoint) pair is tested independently against the
//      indices taken before the walk: one thread per landmark, the keypoint
//      positions staged through LDS in tiles shared by the workgroup, the
//      descriptor held in registers, Hamming distances by popcount.  A count
//      pass, a device-wide exclusive scan and a write pass lay the candidates
//      out in the walk's order (landmark order, then keypoint order);
//   2. the taking of indices during the walk: a candidate survives exactly when
//      it is the first candidate, in the walk's order, that carries its
//      previous-frame index (an earlier accepted candidate with that index would
//      have taken it, and the first one is always accepted).  One 64-bit atomic
//      minimum of (landmark, keypoint) per index finds it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "zsba.h"

namespace {

constexpr int kMatchBlock = 256;
constexpr int kScanBlock = 1024;
constexpr int kDescWords = 8;  // 256 bits

struct match_frame {
    double R[9];
    double t[3];
    double K[4];
    double u_max, v_max;  // cols - 11 and rows - 11, converted as the reference's size_t arithmetic
};

// mapped = R X + t (matrix.hpp products: sum from zero in ascending k, the
// translation added after), pinhole::project (camera.hpp:91-112), and the
// border test of slam.hpp:348-350.
__device__ __forceinline__ bool project_landmark(const match_frame& f, const double* X, double& u, double& v) {
    double m[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) s = s + f.R[3 * i + k] * X[k];
        m[i] = s + f.t[i];
    }
    if (m[2] < 1.0e-12) return false;
    const double iz = 1.0 / m[2];
    const double x = m[0] * iz, y = m[1] * iz;
    u = f.K[0] * x + f.K[2];
    v = f.K[1] * y + f.K[3];
    return !((u < 10.0) || (u > f.u_max) || (v < 10.0) || (v > f.v_max));
}

// One thread per landmark.  kWrite = false counts each landmark's candidates,
// kWrite = true writes them from the landmark's exclusive offset.  Every thread
// of the workgroup takes part in staging every keypoint tile.
template <bool kWrite>
__global__ void __launch_bounds__(kMatchBlock)
k_match_scan_keypoints(match_frame f, int32_t nl, const double* __restrict__ lm_xyz, const uint32_t* __restrict__ lm_desc,
                       const uint8_t* __restrict__ lm_observed, int32_t nk, const double* __restrict__ kp_xy,
                       const uint32_t* __restrict__ kp_desc, const int32_t* __restrict__ kp_key,
                       const uint8_t* __restrict__ key_taken, int32_t* __restrict__ count,
                       const int32_t* __restrict__ offset, int32_t* __restrict__ cand_lm, int32_t* __restrict__ cand_kp) {
    __shared__ double2 tile[kMatchBlock];
    const int32_t o = blockIdx.x * kMatchBlock + threadIdx.x;
    double u = 0.0, v = 0.0;
    bool live = false;
    uint32_t d[kDescWords];
    if (o < nl && !lm_observed[o]) {
        const double X[3] = {lm_xyz[3L * o], lm_xyz[3L * o + 1], lm_xyz[3L * o + 2]};
        live = project_landmark(f, X, u, v);
    }
#pragma unroll
    for (int w = 0; w < kDescWords; ++w) d[w] = live ? lm_desc[kDescWords * static_cast<int64_t>(o) + w] : 0u;
    int32_t n = 0;
    const int32_t base = (kWrite && live) ? offset[o] : 0;
    for (int32_t j0 = 0; j0 < nk; j0 += kMatchBlock) {
        __syncthreads();
        const int32_t j = j0 + threadIdx.x;
        if (j < nk) tile[threadIdx.x] = make_double2(kp_xy[2L * j], kp_xy[2L * j + 1]);
        __syncthreads();
        if (!live) continue;
        const int32_t jn = min(kMatchBlock, nk - j0);
        for (int32_t jj = 0; jj < jn; ++jj) {
            const double2 p = tile[jj];
            const double d0 = p.x - u, d1 = p.y - v;
            double l2 = 0.0;  // matrix::get_length_squared
            l2 = l2 + d0 * d0;
            l2 = l2 + d1 * d1;
            if (l2 > 4.0) continue;
            const int32_t k = j0 + jj;
            if (key_taken[kp_key[k]]) continue;
            uint32_t h = 0;
#pragma unroll
            for (int w = 0; w < kDescWords; ++w) h += __popc(d[w] ^ kp_desc[kDescWords * static_cast<int64_t>(k) + w]);
            if (h >= 64u) continue;
            if (kWrite) {
                cand_lm[base + n] = o;
                cand_kp[base + n] = k;
            }
            ++n;
        }
    }
    if (!kWrite && o < nl) count[o] = n;
}

// Exclusive scan of per-landmark counts: block totals, one workgroup over the
// totals, then each block's scan from its base.
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t x, int32_t& total) {
    __shared__ int32_t wsum[kScanBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int32_t inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    __syncthreads();
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    int32_t before = 0;
    total = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) {
        if (w < wv) before += wsum[w];
        total += wsum[w];
    }
    return before + inc - x;
}

__global__ void __launch_bounds__(kScanBlock) k_match_block_totals(int32_t nl, const int32_t* count, int32_t* totals) {
    const int32_t o = blockIdx.x * kScanBlock + threadIdx.x;
    int32_t total;
    (void)block_exclusive_scan(o < nl ? count[o] : 0, total);
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanBlock) k_match_scan_totals(int32_t nb, int32_t* totals, int32_t nl, int32_t* offset) {
    int32_t carry = 0;
    for (int32_t b0 = 0; b0 < nb; b0 += kScanBlock) {
        const int32_t b = b0 + threadIdx.x;
        const int32_t x = b < nb ? totals[b] : 0;
        int32_t total;
        const int32_t ex = block_exclusive_scan(x, total);
        __syncthreads();
        if (b < nb) totals[b] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) offset[nl] = carry;
}

__global__ void __launch_bounds__(kScanBlock) k_match_scan_apply(int32_t nl, const int32_t* count, const int32_t* base,
                                                                 int32_t* offset) {
    const int32_t o = blockIdx.x * kScanBlock + threadIdx.x;
    int32_t total;
    const int32_t ex = block_exclusive_scan(o < nl ? count[o] : 0, total);
    if (o < nl) offset[o] = base[blockIdx.x] + ex;
}

__global__ void __launch_bounds__(kMatchBlock) k_match_best_fill(int32_t n_keys, unsigned long long* best) {
    for (int32_t k = blockIdx.x * kMatchBlock + threadIdx.x; k < n_keys; k += gridDim.x * kMatchBlock) best[k] = ~0ull;
}

__device__ __forceinline__ unsigned long long walk_position(int32_t o, int32_t k) {
    return (static_cast<unsigned long long>(static_cast<uint32_t>(o)) << 32) | static_cast<uint32_t>(k);
}

__global__ void __launch_bounds__(kMatchBlock) k_match_first_per_key(int32_t nc, const int32_t* cand_lm, const int32_t* cand_kp,
                                                                   const int32_t* kp_key, unsigned long long* best) {
    const int32_t c = blockIdx.x * kMatchBlock + threadIdx.x;
    if (c >= nc) return;
    const int32_t k = cand_kp[c];
    atomicMin(&best[kp_key[k]], walk_position(cand_lm[c], k));
}

__global__ void __launch_bounds__(kMatchBlock) k_match_accept(int32_t nc, const int32_t* cand_lm, const int32_t* cand_kp,
                                                            const int32_t* kp_key, const unsigned long long* best,
                                                            uint8_t* accepted) {
    const int32_t c = blockIdx.x * kMatchBlock + threadIdx.x;
    if (c >= nc) return;
    const int32_t k = cand_kp[c];
    accepted[c] = best[kp_key[k]] == walk_position(cand_lm[c], k) ? 1 : 0;
}

int32_t blocks_for(int64_t n, int block) { return static_cast<int32_t>((n + block - 1) / block); }

// ---- banded search ------------------------------------------------------------
// Keypoints and landmarks are bucketed by image row in bands of kBandPx pixels.
// A keypoint passes the 2-pixel test only if its row differs from the
// landmark's by at most 2 px (the rounded sum of squares is never below its
// second term), so a landmark in band b needs the keypoints of bands b-1..b+1,
// plus those with a coordinate that is not finite: a NaN distance passes the
// test, as it does in the reference.  Rows outside the image are clamped to the
// first or last band, where they are still tested wherever they could match.
// Landmarks whose projection is not finite but passed the border test (NaN)
// and candidate lists longer than one workgroup sorts are left to the ordered
// two-pass search above, which handles every input.
constexpr int kBandPx = 8;
constexpr int kMaxBands = 4096;
constexpr int kSortMax = 4096;

__device__ __forceinline__ int row_band(double y, int nbands) {  // y finite
    const double b = floor(y / kBandPx);
    return b < 0.0 ? 0 : (b >= static_cast<double>(nbands - 1) ? nbands - 1 : static_cast<int>(b));
}

__global__ void __launch_bounds__(kMatchBlock) k_match_zero(int32_t n, int32_t* a) {
    for (int32_t i = blockIdx.x * kMatchBlock + threadIdx.x; i < n; i += gridDim.x * kMatchBlock) a[i] = 0;
}

// One workgroup: keypoints sorted by band (a counting sort; order inside a band
// does not matter), the last bin holding those with a coordinate that is not
// finite.  bin_start has nbands + 2 entries.
__global__ void __launch_bounds__(kScanBlock) k_match_bin_keypoints(int32_t nk, const double* __restrict__ kp_xy, int32_t nbands,
                                                                   int32_t* __restrict__ bin_start, double2* __restrict__ kp_sorted,
                                                                   int32_t* __restrict__ kp_sorted_id) {
    __shared__ int32_t hist[kMaxBands + 1];
    const int nbins = nbands + 1;
    for (int b = threadIdx.x; b < nbins; b += kScanBlock) hist[b] = 0;
    __syncthreads();
    auto bin_of = [&](double x, double y) { return (isfinite(x) && isfinite(y)) ? row_band(y, nbands) : nbands; };
    for (int32_t j = threadIdx.x; j < nk; j += kScanBlock) atomicAdd(&hist[bin_of(kp_xy[2L * j], kp_xy[2L * j + 1])], 1);
    __syncthreads();
    int32_t carry = 0;
    for (int b0 = 0; b0 < nbins; b0 += kScanBlock) {
        const int b = b0 + threadIdx.x;
        const int32_t x = b < nbins ? hist[b] : 0;
        int32_t total;
        const int32_t ex = block_exclusive_scan(x, total);
        __syncthreads();
        if (b < nbins) {
            hist[b] = carry + ex;
            bin_start[b] = carry + ex;
        }
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) bin_start[nbins] = carry;
    __syncthreads();
    for (int32_t j = threadIdx.x; j < nk; j += kScanBlock) {
        const double x = kp_xy[2L * j], y = kp_xy[2L * j + 1];
        const int32_t pos = atomicAdd(&hist[bin_of(x, y)], 1);
        kp_sorted[pos] = make_double2(x, y);
        kp_sorted_id[pos] = j;
    }
}

// Per landmark: the projection and border tests, then its band, or -1.  Each
// workgroup counts its landmarks per band in LDS and writes the counts to
// column blockIdx.x of the band-major table block_hist[band][workgroup], whose
// exclusive scan gives every (band, workgroup) its range of the band order.
// counters[0] counts landmarks that pass with a projection that is not finite.
__global__ void __launch_bounds__(kMatchBlock) k_match_project(match_frame f, int32_t nl, const double* __restrict__ lm_xyz,
                                                             const uint8_t* __restrict__ lm_observed, int32_t nbands,
                                                             double2* __restrict__ lm_uv, int32_t* __restrict__ lm_band,
                                                             int32_t* __restrict__ block_hist, int32_t* __restrict__ counters) {
    __shared__ int32_t hist[kMaxBands];
    for (int b = threadIdx.x; b < nbands; b += kMatchBlock) hist[b] = 0;
    __syncthreads();
    const int32_t o = blockIdx.x * kMatchBlock + threadIdx.x;
    if (o < nl) {
        double u = 0.0, v = 0.0;
        bool live = false;
        if (!lm_observed[o]) {
            const double X[3] = {lm_xyz[3L * o], lm_xyz[3L * o + 1], lm_xyz[3L * o + 2]};
            live = project_landmark(f, X, u, v);
        }
        int32_t b = -1;
        if (live) {
            if (isfinite(u) && isfinite(v)) {
                b = row_band(v, nbands);
                atomicAdd(&hist[b], 1);
            } else {
                atomicAdd(&counters[0], 1);
            }
        }
        lm_uv[o] = make_double2(u, v);
        lm_band[o] = b;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbands; b += kMatchBlock)
        block_hist[static_cast<int64_t>(b) * gridDim.x + blockIdx.x] = hist[b];
}

// Same grid as k_match_project: each landmark takes the next place in its
// (band, workgroup) range.
__global__ void __launch_bounds__(kMatchBlock) k_match_lm_scatter(int32_t nl, int32_t nbands, const int32_t* __restrict__ lm_band,
                                                                const int32_t* __restrict__ range_start,
                                                                int32_t* __restrict__ lm_sorted) {
    __shared__ int32_t cursor[kMaxBands];
    for (int b = threadIdx.x; b < nbands; b += kMatchBlock)
        cursor[b] = range_start[static_cast<int64_t>(b) * gridDim.x + blockIdx.x];
    __syncthreads();
    const int32_t o = blockIdx.x * kMatchBlock + threadIdx.x;
    if (o >= nl) return;
    const int32_t b = lm_band[o];
    if (b >= 0) lm_sorted[atomicAdd(&cursor[b], 1)] = o;
}

// One workgroup per 256 landmarks in band order: the keypoints of the bands
// from the first landmark's band - 1 to the last one's band + 1, then those not
// finite, staged through LDS.  Candidates are appended as (landmark, keypoint)
// walk positions; counters[1] counts them (only the first cap are stored).
__global__ void __launch_bounds__(kMatchBlock)
k_match_band_search(int32_t nbands, const int32_t* __restrict__ live_count, const int32_t* __restrict__ lm_sorted,
                    const double2* __restrict__ lm_uv, const int32_t* __restrict__ lm_band, const uint32_t* __restrict__ lm_desc,
                    const int32_t* __restrict__ bin_start, const double2* __restrict__ kp_sorted,
                    const int32_t* __restrict__ kp_sorted_id, const uint32_t* __restrict__ kp_desc,
                    const int32_t* __restrict__ kp_key, const uint8_t* __restrict__ key_taken, int32_t cap,
                    unsigned long long* __restrict__ cand, int32_t* __restrict__ counters) {
    __shared__ double2 tile[kMatchBlock];
    __shared__ int32_t tile_id[kMatchBlock];
    const int32_t n_live = *live_count;
    const int32_t first = blockIdx.x * kMatchBlock;
    if (first >= n_live) return;  // the whole workgroup
    const int32_t idx = first + threadIdx.x;
    const bool valid = idx < n_live;
    const int32_t last = min(n_live, first + kMatchBlock) - 1;
    const int b_lo = max(lm_band[lm_sorted[first]] - 1, 0);
    const int b_hi = min(lm_band[lm_sorted[last]] + 1, nbands - 1);
    int32_t o = 0;
    double u = 0.0, v = 0.0;
    uint32_t d[kDescWords];
    if (valid) {
        o = lm_sorted[idx];
        const double2 p = lm_uv[o];
        u = p.x;
        v = p.y;
    }
#pragma unroll
    for (int w = 0; w < kDescWords; ++w) d[w] = valid ? lm_desc[kDescWords * static_cast<int64_t>(o) + w] : 0u;
    for (int r = 0; r < 2; ++r) {
        const int32_t k0 = r == 0 ? bin_start[b_lo] : bin_start[nbands];
        const int32_t k1 = r == 0 ? bin_start[b_hi + 1] : bin_start[nbands + 1];
        for (int32_t j0 = k0; j0 < k1; j0 += kMatchBlock) {
            __syncthreads();
            const int32_t j = j0 + threadIdx.x;
            if (j < k1) {
                tile[threadIdx.x] = kp_sorted[j];
                tile_id[threadIdx.x] = kp_sorted_id[j];
            }
            __syncthreads();
            if (!valid) continue;
            const int32_t jn = min(kMatchBlock, k1 - j0);
            for (int32_t jj = 0; jj < jn; ++jj) {
                const double2 p = tile[jj];
                const double d0 = p.x - u, d1 = p.y - v;
                double l2 = 0.0;
                l2 = l2 + d0 * d0;
                l2 = l2 + d1 * d1;
                if (l2 > 4.0) continue;
                const int32_t k = tile_id[jj];
                if (key_taken[kp_key[k]]) continue;
                uint32_t h = 0;
#pragma unroll
                for (int w = 0; w < kDescWords; ++w) h += __popc(d[w] ^ kp_desc[kDescWords * static_cast<int64_t>(k) + w]);
                if (h >= 64u) continue;
                const int32_t slot = atomicAdd(&counters[1], 1);
                if (slot < cap) cand[slot] = walk_position(o, k);
            }
        }
    }
}

// One workgroup: bitonic sort of the n = counters[1] <= kSortMax walk positions.
__global__ void __launch_bounds__(kScanBlock) k_match_sort_small(unsigned long long* __restrict__ keys,
                                                                const int32_t* __restrict__ counters) {
    __shared__ unsigned long long sh[kSortMax];
    const int32_t n = min(counters[1], kSortMax);
    int32_t p = 1;
    while (p < n) p <<= 1;
    for (int32_t i = threadIdx.x; i < p; i += kScanBlock) sh[i] = i < n ? keys[i] : ~0ull;
    __syncthreads();
    for (int32_t k = 2; k <= p; k <<= 1) {
        for (int32_t j = k >> 1; j > 0; j >>= 1) {
            for (int32_t i = threadIdx.x; i < p; i += kScanBlock) {
                const int32_t l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const unsigned long long a = sh[i], b = sh[l];
                    if ((a > b) == up) {
                        sh[i] = b;
                        sh[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int32_t i = threadIdx.x; i < n; i += kScanBlock) keys[i] = sh[i];
}

__global__ void __launch_bounds__(kMatchBlock) k_match_first_per_key_packed(int32_t nc, const unsigned long long* cand,
                                                                          const int32_t* kp_key, unsigned long long* best) {
    const int32_t c = blockIdx.x * kMatchBlock + threadIdx.x;
    if (c >= nc) return;
    const unsigned long long w = cand[c];
    atomicMin(&best[kp_key[static_cast<uint32_t>(w)]], w);
}

__global__ void __launch_bounds__(kMatchBlock) k_match_accept_packed(int32_t nc, const unsigned long long* cand,
                                                                   const int32_t* kp_key, const unsigned long long* best,
                                                                   uint8_t* accepted) {
    const int32_t c = blockIdx.x * kMatchBlock + threadIdx.x;
    if (c >= nc) return;
    const unsigned long long w = cand[c];
    accepted[c] = best[kp_key[static_cast<uint32_t>(w)]] == w ? 1 : 0;
}

// ---- the search that allocates nothing ------------------------------------------
// zsba_match_by_projection_noalloc runs the banded search above on device
// inputs and outputs with every intermediate buffer carved from one static
// device array per GPU (`scratch`, below), and decides on the device what the
// host decides in zsba_match_by_projection: a call that the banded search
// cannot finish (a landmark with a projection that is not finite, more than
// kSortMax candidates, or a keypoint index outside [0, n_keys)) reports an
// error code instead of falling back to the ordered search.
constexpr int64_t kNaMaxLandmarks = int64_t(1) << 18;
constexpr int64_t kNaMaxKeypoints = int64_t(1) << 14;
constexpr int64_t kNaMaxKeys = int64_t(1) << 16;
constexpr int64_t kNaMaxBands = 256;  // images up to 2048 rows
constexpr int64_t kNaMaxGrid = (kNaMaxLandmarks + kMatchBlock - 1) / kMatchBlock;
constexpr int64_t kNaMaxCells = kNaMaxBands * kNaMaxGrid;
constexpr int64_t kNaMaxCellBlocks = (kNaMaxCells + kScanBlock - 1) / kScanBlock;

constexpr int64_t na_align(int64_t x) { return (x + 255) / 256 * 256; }
// byte offsets of the buffers in the scratch array
constexpr int64_t kNaLmUv = 0;
constexpr int64_t kNaLmBand = kNaLmUv + na_align(16 * kNaMaxLandmarks);
constexpr int64_t kNaLmSorted = kNaLmBand + na_align(4 * kNaMaxLandmarks);
constexpr int64_t kNaCellCount = kNaLmSorted + na_align(4 * kNaMaxLandmarks);
constexpr int64_t kNaCellStart = kNaCellCount + na_align(4 * kNaMaxCells);
constexpr int64_t kNaCursor = kNaCellStart + na_align(4 * (kNaMaxCells + 1));
constexpr int64_t kNaBinStart = kNaCursor + na_align(4 * kNaMaxCellBlocks);
constexpr int64_t kNaKpSorted = kNaBinStart + na_align(4 * (kNaMaxBands + 2));
constexpr int64_t kNaKpSortedId = kNaKpSorted + na_align(16 * kNaMaxKeypoints);
constexpr int64_t kNaKpKey = kNaKpSortedId + na_align(4 * kNaMaxKeypoints);
constexpr int64_t kNaCounters = kNaKpKey + na_align(4 * kNaMaxKeypoints);
constexpr int64_t kNaCand = kNaCounters + na_align(4 * 4);
constexpr int64_t kNaAccepted = kNaCand + na_align(8 * kSortMax);
constexpr int64_t kNaBest = kNaAccepted + na_align(kSortMax);
constexpr int64_t kNaScratchBytes = kNaBest + na_align(8 * kNaMaxKeys);
static_assert(kNaScratchBytes <= (int64_t(10) << 20), "the no-alloc scratch stays within 10 MiB");

// counters[0]: landmarks with a projection that is not finite; [1]: candidates;
// [2]: a keypoint index outside [0, n_keys).  Keys are copied with bad ones
// replaced by 0 so that no kernel reads outside key_taken or best.
__global__ void __launch_bounds__(kMatchBlock) k_na_keys(int32_t nk, const int32_t* __restrict__ kp_key, int32_t n_keys,
                                                       int32_t* __restrict__ key_safe, int32_t* __restrict__ counters) {
    const int32_t i = blockIdx.x * kMatchBlock + threadIdx.x;
    if (i >= nk) return;
    const int32_t k = kp_key[i];
    const bool ok = k >= 0 && k < n_keys;
    key_safe[i] = ok ? k : 0;
    if (!ok) atomicOr(&counters[2], 1);
}

__device__ __forceinline__ int32_t na_status(const int32_t* counters) {
    if (counters[2]) return ZSBA_EINVAL;
    if (counters[0] != 0 || counters[1] > kSortMax) return ZSBA_EUNSUP;
    return 0;
}

__global__ void __launch_bounds__(kMatchBlock) k_na_first_per_key(const unsigned long long* cand, const int32_t* kp_key,
                                                                unsigned long long* best, const int32_t* counters) {
    if (na_status(counters)) return;
    const int32_t c = blockIdx.x * kMatchBlock + threadIdx.x;
    if (c >= counters[1]) return;
    const unsigned long long w = cand[c];
    atomicMin(&best[kp_key[static_cast<uint32_t>(w)]], w);
}

__global__ void __launch_bounds__(kMatchBlock) k_na_accept(const unsigned long long* cand, const int32_t* kp_key,
                                                         const unsigned long long* best, uint8_t* accepted,
                                                         const int32_t* counters) {
    if (na_status(counters)) return;
    const int32_t c = blockIdx.x * kMatchBlock + threadIdx.x;
    if (c >= counters[1]) return;
    const unsigned long long w = cand[c];
    accepted[c] = best[kp_key[static_cast<uint32_t>(w)]] == w ? 1 : 0;
}

// One workgroup: the accepted candidates, in walk order, become the matches
// (the first max_matches written); count receives their number or the error.
__global__ void __launch_bounds__(kScanBlock) k_na_compact(const unsigned long long* __restrict__ cand,
                                                          const uint8_t* __restrict__ accepted,
                                                          const int32_t* __restrict__ counters, int32_t max_matches,
                                                          int32_t* __restrict__ match_landmark,
                                                          int32_t* __restrict__ match_keypoint, int32_t* __restrict__ count) {
    constexpr int kPer = kSortMax / kScanBlock;
    const int32_t status = na_status(counters);
    const int32_t nc = status ? 0 : counters[1];
    const int32_t c0 = threadIdx.x * kPer;
    int32_t mine = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) mine += (c0 + i < nc && accepted[c0 + i]) ? 1 : 0;
    int32_t total = 0;
    int32_t n = block_exclusive_scan(mine, total);
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int32_t c = c0 + i;
        if (c < nc && accepted[c]) {
            if (n < max_matches) {
                match_landmark[n] = static_cast<int32_t>(cand[c] >> 32);
                match_keypoint[n] = static_cast<int32_t>(cand[c] & 0xFFFFFFFFull);
            }
            ++n;
        }
    }
    if (threadIdx.x == 0) *count = status ? status : total;
}

}  // namespace

// Scratch of zsba_match_by_projection_noalloc: one copy per GPU, held until the
// process ends.  Host code uses only the address hipGetSymbolAddress gives for
// the current device (na_scratch), never this name.
static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex na_mutex;
unsigned char* na_ptr[kNaMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
hipError_t na_scratch(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNaMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(na_mutex);
    if (!na_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        na_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = na_ptr[dev];
    return hipSuccess;
}

}  // namespace

struct zsba_matcher {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::string err;
    // device buffers, grown on demand
    struct buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    buf lm_xyz, lm_desc, lm_obs, kp_xy, kp_desc, kp_key, key_taken, count, offset, totals, best, cand_lm, cand_kp, accepted;
    buf lm_uv, lm_band, lm_sorted, band_count, band_start, cursor, bin_start, kp_sorted, kp_sorted_id, counters, candw;
    std::vector<unsigned long long> h_w;
    std::vector<int32_t> h_lm, h_kp;
    std::vector<uint8_t> h_acc;
    double last_ms = 0.0;
    int32_t last_candidates = 0;

    ~zsba_matcher() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        for (buf* b : {&lm_xyz, &lm_desc, &lm_obs, &kp_xy, &kp_desc, &kp_key, &key_taken, &count, &offset, &totals, &best,
                       &cand_lm, &cand_kp, &accepted, &lm_uv, &lm_band, &lm_sorted, &band_count, &band_start, &cursor,
                       &bin_start, &kp_sorted, &kp_sorted_id, &counters, &candw})
            if (b->p) (void)hipFree(b->p);
        for (hipEvent_t& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int32_t fail(int32_t code, const char* fmt, ...) {
        char text[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(text, sizeof text, fmt, ap);
        va_end(ap);
        err = text;
        return code;
    }
    int32_t hip_fail(hipError_t e, const char* what) { return fail(ZSBA_EDEVICE, "%s: %s", what, hipGetErrorString(e)); }
    hipError_t reserve(buf& b, size_t bytes) {
        bytes = std::max<size_t>(bytes, 256);
        if (b.bytes >= bytes) return hipSuccess;
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        const size_t grown = bytes + bytes / 2;
        const hipError_t e = hipMalloc(&b.p, grown);
        if (e == hipSuccess) b.bytes = grown;
        return e;
    }
    hipError_t upload(buf& b, const void* src, size_t bytes) {
        hipError_t e = reserve(b, bytes);
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, stream);
        return e;
    }
    template <class T>
    T* as(buf& b) { return static_cast<T*>(b.p); }
};

extern "C" {

zsba_matcher* zsba_matcher_create(int32_t device) {
    zsba_matcher* m = new (std::nothrow) zsba_matcher();
    if (!m) return nullptr;
    m->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&m->ev[0]) != hipSuccess || hipEventCreate(&m->ev[1]) != hipSuccess) {
        delete m;
        return nullptr;
    }
    return m;
}

void zsba_matcher_destroy(zsba_matcher* m) { delete m; }

const char* zsba_matcher_last_error(const zsba_matcher* m) { return m ? m->err.c_str() : "null handle"; }

int32_t zsba_matcher_stats(const zsba_matcher* m, double* device_ms, int32_t* candidates) {
    if (!m) return ZSBA_EINVAL;
    if (device_ms) *device_ms = m->last_ms;
    if (candidates) *candidates = m->last_candidates;
    return 0;
}

int32_t zsba_match_by_projection(zsba_matcher* m, const zsba_match_frame* frame, int32_t n_landmarks,
                                 const double* landmark_xyz, const uint8_t* landmark_descriptor,
                                 const uint8_t* landmark_observed, int32_t n_keypoints, const double* keypoint_xy,
                                 const uint8_t* keypoint_descriptor, const int32_t* keypoint_key, int32_t n_keys,
                                 const uint8_t* key_taken, int32_t max_matches, int32_t* match_landmark,
                                 int32_t* match_keypoint) {
    if (!m) return ZSBA_EINVAL;
    if (!frame || n_landmarks < 0 || n_keypoints < 0 || n_keys < 0 || max_matches < 0)
        return m->fail(ZSBA_EINVAL, "bad arguments to match_by_projection");
    if ((n_landmarks > 0 && (!landmark_xyz || !landmark_descriptor || !landmark_observed)) ||
        (n_keypoints > 0 && (!keypoint_xy || !keypoint_descriptor || !keypoint_key || !key_taken)) ||
        (max_matches > 0 && (!match_landmark || !match_keypoint)))
        return m->fail(ZSBA_EINVAL, "match_by_projection: null array");
    for (int32_t i = 0; i < n_keypoints; ++i)
        if (keypoint_key[i] < 0 || keypoint_key[i] >= n_keys)
            return m->fail(ZSBA_EINVAL, "match_by_projection: keypoint %d has previous-frame index %d outside [0, %d)", i,
                           keypoint_key[i], n_keys);
    m->last_ms = 0.0;
    m->last_candidates = 0;
    if (n_landmarks == 0 || n_keypoints == 0) return 0;

    match_frame f;
    for (int i = 0; i < 9; ++i) f.R[i] = frame->rotation[i];
    for (int i = 0; i < 3; ++i) f.t[i] = frame->translation[i];
    for (int i = 0; i < 4; ++i) f.K[i] = frame->K[i];
    f.u_max = static_cast<double>(static_cast<uint64_t>(frame->cols) - 11u);
    f.v_max = static_cast<double>(static_cast<uint64_t>(frame->rows) - 11u);

    hipError_t e = hipSetDevice(m->device);
    const int64_t nl = n_landmarks, nk = n_keypoints;
    const int32_t nb = blocks_for(nl, kScanBlock);
    if (e == hipSuccess) e = m->upload(m->lm_xyz, landmark_xyz, 24 * nl);
    if (e == hipSuccess) e = m->upload(m->lm_desc, landmark_descriptor, 32 * nl);
    if (e == hipSuccess) e = m->upload(m->lm_obs, landmark_observed, nl);
    if (e == hipSuccess) e = m->upload(m->kp_xy, keypoint_xy, 16 * nk);
    if (e == hipSuccess) e = m->upload(m->kp_desc, keypoint_descriptor, 32 * nk);
    if (e == hipSuccess) e = m->upload(m->kp_key, keypoint_key, 4 * nk);
    if (e == hipSuccess) e = m->upload(m->key_taken, key_taken, static_cast<size_t>(n_keys));
    if (e == hipSuccess) e = m->reserve(m->count, 4 * nl);
    if (e == hipSuccess) e = m->reserve(m->offset, 4 * (nl + 1));
    if (e == hipSuccess) e = m->reserve(m->totals, 4 * static_cast<size_t>(nb));
    if (e == hipSuccess) e = m->reserve(m->best, 8 * static_cast<size_t>(n_keys));
    if (e != hipSuccess) return m->hip_fail(e, "match_by_projection: upload");

    const hipStream_t s = m->stream;
    (void)hipEventRecord(m->ev[0], s);
    const int32_t grid = blocks_for(nl, kMatchBlock);

    // Banded search; the ordered two-pass search below takes the inputs it leaves.
    const int64_t nbands64 = (frame->rows + kBandPx - 1) / kBandPx;
    if (nbands64 >= 1 && nbands64 <= kMaxBands && nbands64 * grid < (int64_t(1) << 30)) {
        const int32_t nbands = static_cast<int32_t>(nbands64);
        const int64_t ncell = nbands64 * grid;  // (band, workgroup) table
        const int32_t ncb = blocks_for(ncell, kScanBlock);
        e = m->reserve(m->lm_uv, 16 * nl);
        if (e == hipSuccess) e = m->reserve(m->lm_band, 4 * nl);
        if (e == hipSuccess) e = m->reserve(m->lm_sorted, 4 * nl);
        if (e == hipSuccess) e = m->reserve(m->band_count, 4 * static_cast<size_t>(ncell));
        if (e == hipSuccess) e = m->reserve(m->band_start, 4 * static_cast<size_t>(ncell + 1));
        if (e == hipSuccess) e = m->reserve(m->cursor, 4 * static_cast<size_t>(ncb));
        if (e == hipSuccess) e = m->reserve(m->bin_start, 4 * static_cast<size_t>(nbands + 2));
        if (e == hipSuccess) e = m->reserve(m->kp_sorted, 16 * nk);
        if (e == hipSuccess) e = m->reserve(m->kp_sorted_id, 4 * nk);
        if (e == hipSuccess) e = m->reserve(m->counters, 8);
        if (e == hipSuccess) e = m->reserve(m->candw, 8 * static_cast<size_t>(kSortMax));
        if (e == hipSuccess) e = m->reserve(m->accepted, kSortMax);
        if (e != hipSuccess) return m->hip_fail(e, "match_by_projection: band buffers");
        int32_t* const cell_count = m->as<int32_t>(m->band_count);
        int32_t* const cell_start = m->as<int32_t>(m->band_start);
        k_match_zero<<<1, kMatchBlock, 0, s>>>(2, m->as<int32_t>(m->counters));
        k_match_bin_keypoints<<<1, kScanBlock, 0, s>>>(n_keypoints, m->as<double>(m->kp_xy), nbands, m->as<int32_t>(m->bin_start),
                                                       m->as<double2>(m->kp_sorted), m->as<int32_t>(m->kp_sorted_id));
        k_match_project<<<grid, kMatchBlock, 0, s>>>(f, n_landmarks, m->as<double>(m->lm_xyz), m->as<uint8_t>(m->lm_obs), nbands,
                                                     m->as<double2>(m->lm_uv), m->as<int32_t>(m->lm_band), cell_count,
                                                     m->as<int32_t>(m->counters));
        const int32_t ncell32 = static_cast<int32_t>(ncell);
        k_match_block_totals<<<ncb, kScanBlock, 0, s>>>(ncell32, cell_count, m->as<int32_t>(m->cursor));
        k_match_scan_totals<<<1, kScanBlock, 0, s>>>(ncb, m->as<int32_t>(m->cursor), ncell32, cell_start);
        k_match_scan_apply<<<ncb, kScanBlock, 0, s>>>(ncell32, cell_count, m->as<int32_t>(m->cursor), cell_start);
        k_match_lm_scatter<<<grid, kMatchBlock, 0, s>>>(n_landmarks, nbands, m->as<int32_t>(m->lm_band), cell_start,
                                                        m->as<int32_t>(m->lm_sorted));
        k_match_band_search<<<grid, kMatchBlock, 0, s>>>(
            nbands, cell_start + ncell, m->as<int32_t>(m->lm_sorted), m->as<double2>(m->lm_uv),
            m->as<int32_t>(m->lm_band), m->as<uint32_t>(m->lm_desc), m->as<int32_t>(m->bin_start), m->as<double2>(m->kp_sorted),
            m->as<int32_t>(m->kp_sorted_id), m->as<uint32_t>(m->kp_desc), m->as<int32_t>(m->kp_key), m->as<uint8_t>(m->key_taken),
            kSortMax, m->as<unsigned long long>(m->candw), m->as<int32_t>(m->counters));
        int32_t cnt[2] = {0, 0};
        e = hipMemcpyAsync(cnt, m->counters.p, sizeof cnt, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return m->hip_fail(e, "match_by_projection: banded search");
        if (cnt[0] == 0 && cnt[1] >= 0 && cnt[1] <= kSortMax) {
            const int32_t nc = cnt[1];
            m->last_candidates = nc;
            if (nc > 0) {
                k_match_sort_small<<<1, kScanBlock, 0, s>>>(m->as<unsigned long long>(m->candw), m->as<int32_t>(m->counters));
                k_match_best_fill<<<std::min(blocks_for(n_keys, kMatchBlock), 1024), kMatchBlock, 0, s>>>(
                    n_keys, m->as<unsigned long long>(m->best));
                const int32_t cgrid = blocks_for(nc, kMatchBlock);
                k_match_first_per_key_packed<<<cgrid, kMatchBlock, 0, s>>>(nc, m->as<unsigned long long>(m->candw),
                                                                          m->as<int32_t>(m->kp_key),
                                                                          m->as<unsigned long long>(m->best));
                k_match_accept_packed<<<cgrid, kMatchBlock, 0, s>>>(nc, m->as<unsigned long long>(m->candw),
                                                                   m->as<int32_t>(m->kp_key), m->as<unsigned long long>(m->best),
                                                                   m->as<uint8_t>(m->accepted));
            }
            (void)hipEventRecord(m->ev[1], s);
            m->h_w.resize(nc);
            m->h_acc.resize(nc);
            e = hipSuccess;
            if (nc > 0) {
                e = hipMemcpyAsync(m->h_w.data(), m->candw.p, 8 * static_cast<size_t>(nc), hipMemcpyDeviceToHost, s);
                if (e == hipSuccess) e = hipMemcpyAsync(m->h_acc.data(), m->accepted.p, nc, hipMemcpyDeviceToHost, s);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e == hipSuccess) e = hipGetLastError();
            if (e != hipSuccess) return m->hip_fail(e, "match_by_projection");
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, m->ev[0], m->ev[1]);
            m->last_ms = ms;
            // sorted walk positions: the accepted ones are the matches, in order
            int32_t n = 0;
            for (int32_t c = 0; c < nc; ++c) {
                if (!m->h_acc[c]) continue;
                if (n < max_matches) {
                    match_landmark[n] = static_cast<int32_t>(m->h_w[c] >> 32);
                    match_keypoint[n] = static_cast<int32_t>(m->h_w[c] & 0xFFFFFFFFull);
                }
                ++n;
            }
            return n;
        }
    }
    // Ordered two-pass search: the candidate offsets are 32-bit.
    if (nl * nk > static_cast<int64_t>(INT32_MAX))
        return m->fail(ZSBA_EUNSUP, "match_by_projection: %d landmarks x %d keypoints need the ordered search, "
                       "whose candidate offsets are 32-bit", n_landmarks, n_keypoints);
    k_match_scan_keypoints<false><<<grid, kMatchBlock, 0, s>>>(
        f, n_landmarks, m->as<double>(m->lm_xyz), m->as<uint32_t>(m->lm_desc), m->as<uint8_t>(m->lm_obs), n_keypoints,
        m->as<double>(m->kp_xy), m->as<uint32_t>(m->kp_desc), m->as<int32_t>(m->kp_key), m->as<uint8_t>(m->key_taken),
        m->as<int32_t>(m->count), nullptr, nullptr, nullptr);
    k_match_block_totals<<<nb, kScanBlock, 0, s>>>(n_landmarks, m->as<int32_t>(m->count), m->as<int32_t>(m->totals));
    k_match_scan_totals<<<1, kScanBlock, 0, s>>>(nb, m->as<int32_t>(m->totals), n_landmarks, m->as<int32_t>(m->offset));
    k_match_scan_apply<<<nb, kScanBlock, 0, s>>>(n_landmarks, m->as<int32_t>(m->count), m->as<int32_t>(m->totals),
                                                  m->as<int32_t>(m->offset));
    int32_t nc = 0;
    e = hipMemcpyAsync(&nc, m->as<int32_t>(m->offset) + nl, sizeof nc, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return m->hip_fail(e, "match_by_projection: candidate count");
    m->last_candidates = nc;
    if (nc == 0) {
        (void)hipEventRecord(m->ev[1], s);
        (void)hipEventSynchronize(m->ev[1]);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, m->ev[0], m->ev[1]);
        m->last_ms = ms;
        return 0;
    }
    e = m->reserve(m->cand_lm, 4 * static_cast<size_t>(nc));
    if (e == hipSuccess) e = m->reserve(m->cand_kp, 4 * static_cast<size_t>(nc));
    if (e == hipSuccess) e = m->reserve(m->accepted, static_cast<size_t>(nc));
    if (e != hipSuccess) return m->hip_fail(e, "match_by_projection: candidate buffers");
    k_match_scan_keypoints<true><<<grid, kMatchBlock, 0, s>>>(
        f, n_landmarks, m->as<double>(m->lm_xyz), m->as<uint32_t>(m->lm_desc), m->as<uint8_t>(m->lm_obs), n_keypoints,
        m->as<double>(m->kp_xy), m->as<uint32_t>(m->kp_desc), m->as<int32_t>(m->kp_key), m->as<uint8_t>(m->key_taken),
        nullptr, m->as<int32_t>(m->offset), m->as<int32_t>(m->cand_lm), m->as<int32_t>(m->cand_kp));
    k_match_best_fill<<<std::min(blocks_for(n_keys, kMatchBlock), 1024), kMatchBlock, 0, s>>>(
        n_keys, m->as<unsigned long long>(m->best));
    const int32_t cgrid = blocks_for(nc, kMatchBlock);
    k_match_first_per_key<<<cgrid, kMatchBlock, 0, s>>>(nc, m->as<int32_t>(m->cand_lm), m->as<int32_t>(m->cand_kp),
                                                        m->as<int32_t>(m->kp_key), m->as<unsigned long long>(m->best));
    k_match_accept<<<cgrid, kMatchBlock, 0, s>>>(nc, m->as<int32_t>(m->cand_lm), m->as<int32_t>(m->cand_kp),
                                                 m->as<int32_t>(m->kp_key), m->as<unsigned long long>(m->best),
                                                 m->as<uint8_t>(m->accepted));
    (void)hipEventRecord(m->ev[1], s);
    m->h_lm.resize(nc);
    m->h_kp.resize(nc);
    m->h_acc.resize(nc);
    e = hipMemcpyAsync(m->h_lm.data(), m->cand_lm.p, 4 * static_cast<size_t>(nc), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(m->h_kp.data(), m->cand_kp.p, 4 * static_cast<size_t>(nc), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(m->h_acc.data(), m->accepted.p, static_cast<size_t>(nc), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return m->hip_fail(e, "match_by_projection");
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, m->ev[0], m->ev[1]);
    m->last_ms = ms;
    // The candidates are in the walk's order; the accepted ones are its matches.
    int32_t n = 0;
    for (int32_t c = 0; c < nc; ++c) {
        if (!m->h_acc[c]) continue;
        if (n < max_matches) {
            match_landmark[n] = m->h_lm[c];
            match_keypoint[n] = m->h_kp[c];
        }
        ++n;
    }
    return n;
}

int32_t zsba_match_by_projection_noalloc(const zsba_match_frame* frame, int32_t n_landmarks, const double* landmark_xyz,
                                         const uint8_t* landmark_descriptor, const uint8_t* landmark_observed,
                                         int32_t n_keypoints, const double* keypoint_xy,
                                         const uint8_t* keypoint_descriptor, const int32_t* keypoint_key, int32_t n_keys,
                                         const uint8_t* key_taken, int32_t max_matches, int32_t* match_landmark,
                                         int32_t* match_keypoint, int32_t* match_count, void* stream) {
    if (!frame || !match_count || n_landmarks < 0 || n_keypoints < 0 || n_keys < 0 || max_matches < 0)
        return ZSBA_EINVAL;
    if ((n_landmarks > 0 && (!landmark_xyz || !landmark_descriptor || !landmark_observed)) ||
        (n_keypoints > 0 && (!keypoint_xy || !keypoint_descriptor || !keypoint_key || !key_taken)) ||
        (max_matches > 0 && (!match_landmark || !match_keypoint)))
        return ZSBA_EINVAL;
    if (n_keypoints > 0 && n_keys == 0) return ZSBA_EINVAL;  // every previous-frame index is out of range
    if (n_landmarks > kNaMaxLandmarks || n_keypoints > kNaMaxKeypoints || n_keys > kNaMaxKeys) return ZSBA_EUNSUP;
    const int64_t nbands64 = (frame->rows + kBandPx - 1) / kBandPx;
    if (frame->rows < 1 || frame->cols < 1 || nbands64 > kNaMaxBands) return ZSBA_EUNSUP;
    unsigned char* base = nullptr;
    if (na_scratch(&base) != hipSuccess) return ZSBA_EDEVICE;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    int32_t* const counters = reinterpret_cast<int32_t*>(base + kNaCounters);
    if (n_landmarks == 0 || n_keypoints == 0) {
        // no candidates: count = 0, written by a kernel (no memset, so the call
        // can be captured into a graph)
        k_match_zero<<<1, kMatchBlock, 0, s>>>(1, match_count);
        return hipGetLastError() == hipSuccess ? 0 : ZSBA_EDEVICE;
    }
    match_frame f;
    for (int i = 0; i < 9; ++i) f.R[i] = frame->rotation[i];
    for (int i = 0; i < 3; ++i) f.t[i] = frame->translation[i];
    for (int i = 0; i < 4; ++i) f.K[i] = frame->K[i];
    f.u_max = static_cast<double>(static_cast<uint64_t>(frame->cols) - 11u);
    f.v_max = static_cast<double>(static_cast<uint64_t>(frame->rows) - 11u);
    const int32_t nbands = static_cast<int32_t>(nbands64);
    const int32_t grid = blocks_for(n_landmarks, kMatchBlock);
    const int64_t ncell = nbands64 * grid;
    const int32_t ncb = blocks_for(ncell, kScanBlock);
    // the carve-up holds these sizes (checked against the limits above)
    if (grid > kNaMaxGrid || ncell > kNaMaxCells || ncb > kNaMaxCellBlocks) return ZSBA_EUNSUP;
    double2* const lm_uv = reinterpret_cast<double2*>(base + kNaLmUv);
    int32_t* const lm_band = reinterpret_cast<int32_t*>(base + kNaLmBand);
    int32_t* const lm_sorted = reinterpret_cast<int32_t*>(base + kNaLmSorted);
    int32_t* const cell_count = reinterpret_cast<int32_t*>(base + kNaCellCount);
    int32_t* const cell_start = reinterpret_cast<int32_t*>(base + kNaCellStart);
    int32_t* const cursor = reinterpret_cast<int32_t*>(base + kNaCursor);
    int32_t* const bin_start = reinterpret_cast<int32_t*>(base + kNaBinStart);
    double2* const kp_sorted = reinterpret_cast<double2*>(base + kNaKpSorted);
    int32_t* const kp_sorted_id = reinterpret_cast<int32_t*>(base + kNaKpSortedId);
    int32_t* const kp_key = reinterpret_cast<int32_t*>(base + kNaKpKey);
    unsigned long long* const cand = reinterpret_cast<unsigned long long*>(base + kNaCand);
    uint8_t* const accepted = base + kNaAccepted;
    unsigned long long* const best = reinterpret_cast<unsigned long long*>(base + kNaBest);
    const int32_t ncell32 = static_cast<int32_t>(ncell);
    k_match_zero<<<1, kMatchBlock, 0, s>>>(4, counters);
    k_na_keys<<<blocks_for(n_keypoints, kMatchBlock), kMatchBlock, 0, s>>>(n_keypoints, keypoint_key, n_keys, kp_key,
                                                                          counters);
    k_match_bin_keypoints<<<1, kScanBlock, 0, s>>>(n_keypoints, keypoint_xy, nbands, bin_start, kp_sorted, kp_sorted_id);
    k_match_project<<<grid, kMatchBlock, 0, s>>>(f, n_landmarks, landmark_xyz, landmark_observed, nbands, lm_uv, lm_band,
                                                 cell_count, counters);
    k_match_block_totals<<<ncb, kScanBlock, 0, s>>>(ncell32, cell_count, cursor);
    k_match_scan_totals<<<1, kScanBlock, 0, s>>>(ncb, cursor, ncell32, cell_start);
    k_match_scan_apply<<<ncb, kScanBlock, 0, s>>>(ncell32, cell_count, cursor, cell_start);
    k_match_lm_scatter<<<grid, kMatchBlock, 0, s>>>(n_landmarks, nbands, lm_band, cell_start, lm_sorted);
    k_match_band_search<<<grid, kMatchBlock, 0, s>>>(
        nbands, cell_start + ncell, lm_sorted, lm_uv, lm_band, reinterpret_cast<const uint32_t*>(landmark_descriptor),
        bin_start, kp_sorted, kp_sorted_id, reinterpret_cast<const uint32_t*>(keypoint_descriptor), kp_key, key_taken,
        kSortMax, cand, counters);
    k_match_sort_small<<<1, kScanBlock, 0, s>>>(cand, counters);
    k_match_best_fill<<<std::min(blocks_for(n_keys, kMatchBlock), 1024), kMatchBlock, 0, s>>>(n_keys, best);
    const int32_t cgrid = blocks_for(kSortMax, kMatchBlock);
    k_na_first_per_key<<<cgrid, kMatchBlock, 0, s>>>(cand, kp_key, best, counters);
    k_na_accept<<<cgrid, kMatchBlock, 0, s>>>(cand, kp_key, best, accepted, counters);
    k_na_compact<<<1, kScanBlock, 0, s>>>(cand, accepted, counters, max_matches, match_landmark, match_keypoint,
                                         match_count);
    return hipGetLastError() == hipSuccess ? 0 : ZSBA_EDEVICE;
}

}  // extern "C"
