// Banded Cholesky factorisation and substitutions of the reduced camera system
// for the reassociating precisions (FP64-fast, mixed).  They replace
// matrix::decompose_cholesky / matrix::solve_cholesky
// (source/matrix_decomposition_cholesky.hpp:26-95, called at
// factor_graph.hpp:782-793) when the reduced system is a narrow band: row r of S
// is zero left of column r - band + 1 (sequential tracks couple each pose to a
// few neighbours only).  Cholesky fill stays inside that profile, so the factor
// is a band of the same width.
//
// Factorisation (k_band_factor, one workgroup: band_roles gives its waves).  The
// trailing submatrix lives in a window of W rows (W >= band + 7, a multiple of
// 16): row r sits at window position r mod W, so a row keeps its position from
// the panel that brings it in until the panel that eliminates it.  The window's
// symmetric W x W matrix is held in registers as the tiles (I, J), I >= J, of
// 16 x 16 positions, each owned by one tile wave (tile (I, J) by tile wave
// (I + J) mod the tile-wave count, which spreads every block row's tiles over
// them).  Below, "waves 1-3" stands for the tile waves and "wave 0" for the
// staging waves where entering rows are staged.  Per panel of 8 columns
// c0 .. c0 + 7 (window positions pp0 .. pp0 + 7):
//   phase A  wave 0 reads the panel's columns for every window row from LDS (PB:
//            lane l holds rows c0 + l and c0 + 64 + l) and factors them
//            right-looking: pivot d (readlane), y = 1/sqrt(d), L_rj = a_rj y,
//            the remaining panel columns updated with L_cj broadcast by readlane;
//            the forward substitution runs along on the right-hand side (RH,
//            kept by position).  The panel's L goes to LP by position, with
//            zeros at the panel's own positions (its diagonal block, with 1/L_jj
//            on the diagonal, goes to DB).  Meanwhile waves 1-3 apply the
//            previous panel's L to every tile not holding this panel's
//            positions (two v_mfma_f64_16x16x4f64 per tile: acc -= L_I L_J^T)
//            and copy the previous panel's L to the factor buffers.
//   phase B  waves 1-3 load the rows entering the window into this panel's
//            positions (staged in SR one panel ahead by wave 0; the zeros of LP
//            there keep every later update of this panel off them), then apply
//            this panel's L to the tiles holding the next panel's positions and
//            publish those columns to PB (lookahead); wave 0 stages the rows
//            entering with the next panel and loads those of the one after.
// Every tile receives every panel's update exactly once, in panel order: in
// phase B of panel k when it holds panel k + 1's positions, else in phase A of
// panel k + 1.  The sums are reassociated (four products per matrix-core step),
// so this path serves the reassociating precisions only.
//
// The factor is written twice, as rows of kRing entries: by columns
// (Lc[c kRing + d] = L_{c+d, c}) for the forward substitution and by rows
// (Lr[r kRing + d] = L_{r, r-d}) for the backward one, with 1 / L_ii at d = 0.
// Entries outside the band are zero (the buffer is cleared when it is
// allocated).
//
// Substitutions (band_sub): wave 0 runs a right-looking recurrence over a ring
// of 128 unknowns (lane l holds ring slots l and l + 64).  Step j: the lane
// holding unknown j multiplies it by 1 / L_jj (entry d = 0 of its factor row),
// readlane broadcasts the result, every slot subtracts its factor entry times it,
// and the freed slot takes the unknown 128 steps away.  The other waves stream
// the factor rows into a double-buffered LDS ring, kB rows per block, with the
// loads of up to kK blocks in flight in their registers, so wave 0 reads only
// LDS.  One barrier per block.
//
// Failure rules are the reference's: the decomposition fails when a pivot is not
// greater than 1e-14 max|S_ii| (NaN fails), the solve when some |L_ii| is not
// greater than 1e-7 max|L_ii|.  With floor_rel > 0 every diagonal entry is first
// raised by floor_rel max|S_ii| (in the window; S itself is only read).  Entries
// of S's lower triangle outside the row profile must be zero (the reduced-system
// kernels clear S before forming it).  Rows past n (up to a multiple of 8) are
// identity rows.
#include "common.hpp"

#include <cmath>
#include <cstdint>
#include <type_traits>
#include <utility>

namespace zb {
namespace {

typedef double band_d4 __attribute__((ext_vector_type(4)));

constexpr int kP = kBandPanel;  // panel columns
constexpr int kThreads = 256;   // the substitutions: wave 0 the recurrence, waves 1-3 the loads

// Wave roles of the factorisation for a window of W positions.  Up to W = 80,
// eight waves: wave 0 the panels, wave 4 (on wave 0's SIMD) half of the rows
// entering the window, waves 1-3 and 5-7 the tiles.  Wider windows keep four
// waves (wave 0 the panels and all entering rows, waves 1-3 the tiles): at two
// waves per SIMD their registers would spill.
template <int W>
struct band_roles {
    static constexpr bool kWide = W > 80;
    static constexpr int kThreads = kWide ? 256 : 512;
    static constexpr int kTileWaves = kWide ? 3 : 6;
    static constexpr int kTileLanes = 64 * kTileWaves;
    static constexpr int kStagers = kWide ? 1 : 2;
    static constexpr int kStageWave = kWide ? -1 : 4;  // the second staging wave
    // tile wave index of a wave (-1: none)
    static __device__ __forceinline__ int tile_wave(int wave) {
        return (wave == 0 || wave == kStageWave) ? -1 : ((kStageWave < 0 || wave < kStageWave) ? wave - 1 : wave - 2);
    }
};
constexpr int kLoaders = 3;     // loading waves of the substitutions
constexpr int kRing = 128;      // unknowns held by wave 0; row stride of the factor buffers
constexpr int kB = 16;          // substitution steps per block

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// f(std::integral_constant<int, i>) for i = 0 .. N - 1: every index a constant,
// so arrays indexed by it stay in registers.
template <class F, int... Is>
__device__ __forceinline__ void static_for_seq(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_seq(f, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ double lane_value(double v, int j) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), j);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), j);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// 1 / sqrt(d) within about an ulp for positive normal d: the hardware estimate
// (relative error near 2^-23), then two Newton steps y += y (1 - d y^2) / 2.
__device__ __forceinline__ double rsqrt_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = fma(-(d * y), y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}

__device__ __forceinline__ double dmax_nan(double a, double b) { return (a < b) ? b : a; }

// The lane index, hidden from loop-invariant code motion: the per-step lane masks
// of the unrolled recurrences are formed where they are used instead of being
// hoisted out and kept live in scalar registers (which spills them).
__device__ __forceinline__ int opaque_lane() {
    int lane = static_cast<int>(threadIdx.x & 63);
    asm volatile("" : "+v"(lane));
    return lane;
}

// Window size for a band: the rows a panel of kP columns reaches, in blocks of 16.
constexpr int window_for(int band) {
    return band + kP - 1 <= 64 ? 64 : band + kP - 1 <= 80 ? 80 : band + kP - 1 <= 96 ? 96
         : band + kP - 1 <= 112 ? 112 : band + kP - 1 <= 128 ? 128 : 0;
}

// Tiles (I, J), I >= J, of nb position blocks, dealt to the nt tile waves by
// (I + J) mod nt.
constexpr int tiles_of_wave(int nb, int nt, int wv) {
    int c = 0;
    for (int a = 0; a < nb; ++a)
        for (int b = 0; b <= a; ++b)
            if ((a + b) % nt == wv) ++c;
    return c;
}
constexpr int max_tiles(int nb, int nt) {
    int m = 0;
    for (int wv = 0; wv < nt; ++wv) m = tiles_of_wave(nb, nt, wv) > m ? tiles_of_wave(nb, nt, wv) : m;
    return m;
}

// Factor entries one loading lane holds per block, and blocks in flight (the
// loads outstanding stay within the 63 the wave's counter tracks).  Four blocks
// keep three blocks' loads ahead, more than an HBM round trip at the recurrence's
// pace, with half the registers and unrolled code of eight.
template <int W>
struct sub_shape {
    static constexpr int kItems = kB * W;
    static constexpr int kNI = (kItems + 64 * kLoaders - 1) / (64 * kLoaders);
    static constexpr int kK = (63 / (kNI + 1) + 1) < 4 ? (63 / (kNI + 1) + 1) : 4;
};

// LDS of the factorisation and of the substitutions (they run one after the
// other and share it), in doubles.
template <int W>
struct band_lds {
    static constexpr int kPB = 0;                          // [W][kP + 1]
    static constexpr int kLP = kPB + W * (kP + 1);         // [2][kP][W]
    static constexpr int kDB = kLP + 2 * kP * W;           // [2][kP][kP]
    static constexpr int kSR = kDB + 2 * kP * kP;          // [2][kP][W]
    static constexpr int kSRb = kSR + 2 * kP * W;          // [2][kP]
    static constexpr int kRH = kSRb + 2 * kP;              // [W]
    static constexpr int kSink = kRH + W;                  // [64]: wave 0's lanes with nothing to store
    static constexpr int kFactor = kSink + 64;
    static constexpr int kRingOff = 0;                     // [2][kB][kRing]
    static constexpr int kSlot = 2 * kB * kRing;           // [2][kRing]
    static constexpr int kOut = kSlot + 2 * kRing;         // [2][kB]: a block's unknowns, stored a block later
    static constexpr int kSub = kOut + 2 * kB;
    static constexpr int kTotal = kFactor > kSub ? kFactor : kSub;
};

// Entry (r, c), r >= c, of the system being factorised: S's lower triangle
// inside the band, the floor on the diagonal, identity rows past n.  The load
// address is always valid (element 0 when the entry is not in S).
// Both are written as selects over values computed for every lane: wave 0 runs
// them for ten items per lane per panel with no other wave on its SIMD, and
// branches around the multiply or the early returns (exec-mask updates per item)
// cost it several times the arithmetic.
__device__ __forceinline__ int64_t entry_index(int n, int wb, int r, int c) {
    // r, c >= 0: one unsigned 32 x 32 -> 64 multiply-add
    const uint64_t idx = static_cast<uint64_t>(static_cast<uint32_t>(r)) * static_cast<uint32_t>(n) +
                         static_cast<uint32_t>(c);
    const bool in = (r < n) & (r - c < wb);
    return in ? static_cast<int64_t>(idx) : 0;
}
__device__ __forceinline__ double entry_finish(double raw, int n, int wb, double lift, int r, int c) {
    const bool diag = r == c;
    const double in_band = r - c < wb ? (diag ? raw + lift : raw) : 0.0;
    return r < n ? in_band : (diag ? 1.0 : 0.0);
}

// Rows of the substitutions: n rounded up to whole blocks (factor rows past n
// are zero, so their unknowns are zero and touch nothing).
__device__ __forceinline__ int sub_rows(int n) { return (n + kB - 1) / kB * kB; }

// Substitution over factor rows F (stride kRing; entry 0 is 1 / L_jj): forward
// (kFwd: rows 0, 1, .. of the column-ordered factor) or backward (rows from the
// last, of the row-ordered factor), in blocks of kB steps.  A block's unknowns
// sit in one half of the ring (lane l: slot l in u0, slot l + 64 in u1), so each
// step is one product, one broadcast and two fused multiply-adds; the freed slots
// take their new unknowns at the end of the block.  rhs and out may be the same
// array.  Every wave of the workgroup calls it; wave 0 computes, waves
// 1 .. kLoaders load.
template <int W, bool kFwd>
__device__ __forceinline__ void band_sub(int n, const double* F, const double* rhs, double* out, double* lds,
                                         int wave, int lane) {
    using sh = sub_shape<W>;
    constexpr int kNI = sh::kNI, kK = sh::kK;
    double* ring = lds + band_lds<W>::kRingOff;
    double* rslot = lds + band_lds<W>::kSlot;
    double* xo = lds + band_lds<W>::kOut;
    const int np = sub_rows(n);
    const int nblk = np / kB;
    // row of step t of block m
    auto row_of = [&](int m, int t) __attribute__((always_inline)) { return kFwd ? m * kB + t : np - 1 - m * kB - t; };
    // entries past W stay zero
    for (int i = static_cast<int>(threadIdx.x); i < 2 * kB * kRing; i += static_cast<int>(blockDim.x)) ring[i] = 0.0;
    __syncthreads();

    const bool loader = wave >= 1 && wave <= kLoaders;
    const int li = (wave - 1) * 64 + lane;
    // loading lane's items: entry (t, d) of a block, idx = li + u 64 kLoaders
    int it_t[kNI], it_d[kNI];
    static_for<kNI>([&](auto U) __attribute__((always_inline)) {
        constexpr int u = U;
        const int idx = li + u * 64 * kLoaders;
        it_t[u] = idx < sh::kItems ? idx / W : -1;
        it_d[u] = idx < sh::kItems ? idx - (idx / W) * W : 0;
    });
    double ld[kK][kNI];
    double lr[kK];
    auto issue = [&](int m, double* dst, double& rdst) __attribute__((always_inline)) {
        const int mm = m < nblk ? m : nblk - 1;
        static_for<kNI>([&](auto U) __attribute__((always_inline)) {
            constexpr int u = U;
            const int j = row_of(mm, it_t[u] < 0 ? 0 : it_t[u]);
            dst[u] = F[j * kRing + it_d[u]];
        });
        // the refill value's validity is applied when it is stored: a select
        // here would wait for this load, and with it every load before it
        const int j = row_of(mm, lane & (kB - 1));
        const int jr = kFwd ? j + kRing : j - kRing;
        rdst = rhs[(jr >= 0 && jr < n) ? jr : 0];
    };
    auto store = [&](int m, const double* src, double rsrc) __attribute__((always_inline)) {
        double* rg = ring + (m & 1) * kB * kRing;
        static_for<kNI>([&](auto U) __attribute__((always_inline)) {
            constexpr int u = U;
            if (it_t[u] >= 0) rg[it_t[u] * kRing + it_d[u]] = src[u];
        });
        if (wave == 1 && lane < kB) {
            const int mm = m < nblk ? m : nblk - 1;
            const int j = row_of(mm, lane);
            const int jr = kFwd ? j + kRing : j - kRing;
            rslot[(m & 1) * kRing + (row_of(m, lane) & (kRing - 1))] = (jr >= 0 && jr < n) ? rsrc : 0.0;
        }
    };

    // wave 0: the ring of unknowns; slot s holds unknown j + ((s - j) & 127)
    // (forward) or j - ((j - s) & 127) (backward) at step j
    double u0 = 0.0, u1 = 0.0;
    if (wave == 0) {
        if (kFwd) {
            u0 = lane < n ? rhs[lane] : 0.0;
            u1 = lane + 64 < n ? rhs[lane + 64] : 0.0;
        } else {
            const int j0 = np - 1;
            const int c0 = j0 - ((j0 - lane) & (kRing - 1)), c1 = j0 - ((j0 - lane - 64) & (kRing - 1));
            u0 = (c0 >= 0 && c0 < n) ? rhs[c0] : 0.0;
            u1 = (c1 >= 0 && c1 < n) ? rhs[c1] : 0.0;
        }
    }
    if (loader) {
        static_for<kK>([&](auto S) __attribute__((always_inline)) { issue(S, ld[S], lr[S]); });
        store(0, ld[0], lr[0]);
    }
    __syncthreads();
    // one block: the kB steps on the unknowns in half kHi of the ring
    auto block = [&](int m, const double* l0, const double* l1, auto HI) __attribute__((always_inline)) {
        constexpr bool kHi = HI;
        double xs = 0.0;
        const int jb = row_of(m, 0);
        static_for<kB>([&](auto T) __attribute__((always_inline)) {
            constexpr int t = T;
            const int sl = (kFwd ? jb + t : jb - t) & 63;
            const double xj = lane_value(kHi ? u1 * l1[t] : u0 * l0[t], sl);
            u0 = fma(-l0[t], xj, u0);
            u1 = fma(-l1[t], xj, u1);
            xs = (lane == t) ? xj : xs;
        });
        // the block's slots (kB consecutive, from lo) take the unknowns 128 away
        const int lo = (kFwd ? jb : jb - (kB - 1)) & 63;
        const bool freed = lane >= lo && lane < lo + kB;
        const double rs = rslot[(m & 1) * kRing + (kHi ? 64 : 0) + lane];
        if (kHi) u1 = freed ? rs : u1;
        else u0 = freed ? rs : u0;
        // the block's unknowns go out through LDS: loader wave 1 stores them
        // during the next block, so wave 0 has no global store to wait on
        if (lane < kB) xo[(m & 1) * kB + lane] = xs;
    };
    auto store_out = [&](int m) __attribute__((always_inline)) {
        if (lane < kB) {
            const int j = row_of(m, lane);
            if (j < n) out[j] = xo[(m & 1) * kB + lane];
        }
    };
    for (int m0 = 0; m0 < nblk; m0 += kK) {
        static_for<kK>([&](auto S) __attribute__((always_inline)) {
            constexpr int s = S;
            const int m = m0 + s;
            if (m >= nblk) return;
            if (wave == 0) {
                const double* rg = ring + (m & 1) * kB * kRing;
                double l0[kB], l1[kB];
                const int jb = row_of(m, 0);
                static_for<kB>([&](auto T) __attribute__((always_inline)) {
                    constexpr int t = T;
                    const int j = kFwd ? jb + t : jb - t;
                    const int d0 = kFwd ? ((lane - j) & (kRing - 1)) : ((j - lane) & (kRing - 1));
                    l0[t] = rg[t * kRing + d0];
                    l1[t] = rg[t * kRing + (d0 ^ 64)];
                });
                if ((jb & 64) != 0) block(m, l0, l1, std::integral_constant<bool, true>{});
                else block(m, l0, l1, std::integral_constant<bool, false>{});
            } else if (loader) {
                constexpr int sn = (s + 1) % kK;
                if (wave == 1 && m > 0) store_out(m - 1);
                store(m + 1, ld[sn], lr[sn]);
                issue(m + kK, ld[s], lr[s]);
            }
            __syncthreads();
        });
    }
    if (wave == 1 && nblk > 0) store_out(nblk - 1);
}

template <int W>
__global__ void __launch_bounds__(band_roles<W>::kThreads) k_band_factor(int n, int wb, const double* A, const double* b, double* x,
                                                          double* Lband, double floor_rel, double* eps_out,
                                                          int32_t* status, lm_state* lm, int gated) {
    if (gated && lm->done) return;
    using roles = band_roles<W>;
    constexpr int kFactorThreads = roles::kThreads, kTileWaves = roles::kTileWaves, kTileLanes = roles::kTileLanes;
    constexpr int kStageWave = roles::kStageWave;
    constexpr int NB = W / 16;
    constexpr int kMaxT = max_tiles(NB, kTileWaves);
    constexpr int kPre = kP * W / (64 * roles::kStagers);           // entering-row items per lane of a staging wave
    constexpr int kCP = (kP * W + kTileLanes - 1) / kTileLanes;     // factor-copy items per tile lane
    using lay = band_lds<W>;
    static_assert(W % 16 == 0 && W >= 64 && W <= kRing, "window of 64 .. 128 positions in blocks of 16");
    static_assert(kP * W % (64 * roles::kStagers) == 0, "entering rows split evenly over the staging waves");
    __shared__ double smem[lay::kTotal];
    __shared__ double red[kFactorThreads];
    __shared__ int tab[kTileWaves][kMaxT];
    __shared__ int sflag;
    double* PB = smem + lay::kPB;
    double* LP = smem + lay::kLP;
    double* DB = smem + lay::kDB;
    double* SR = smem + lay::kSR;
    double* SRb = smem + lay::kSRb;
    double* RH = smem + lay::kRH;
    const int tid = threadIdx.x, lane = tid & 63;
    double* sink = smem + lay::kSink + lane;
    const int wave = uniform(tid >> 6);
    const int tw = roles::tile_wave(wave);
    const bool stager = wave == 0 || wave == kStageWave;
    const int np8 = (n + kP - 1) / kP * kP, NP = np8 / kP;
    double* Lc = Lband;
    double* Lr = Lband + static_cast<int64_t>(sub_rows(n)) * kRing;

    if (tid == 0) {
        int c[kTileWaves] = {};
        for (int wv = 0; wv < kTileWaves; ++wv)
            for (int s = 0; s < kMaxT; ++s) tab[wv][s] = -1;
        for (int a = 0; a < NB; ++a)
            for (int bb = 0; bb <= a; ++bb) {
                const int wv = (a + bb) % kTileWaves;
                tab[wv][c[wv]++] = (a << 8) | bb;
            }
    }
    // max |S_ii| (NaN never wins), the floor and the pivot threshold
    {
        double m = 0.0;
        for (int i = tid; i < n; i += kFactorThreads) m = dmax_nan(m, fabs(A[static_cast<int64_t>(i) * n + i]));
        red[tid] = m;
        __syncthreads();
        for (int s = kFactorThreads / 2; s > 0; s >>= 1) {
            if (tid < s) red[tid] = dmax_nan(red[tid], red[tid + s]);
            __syncthreads();
        }
    }
    const double mx = red[0];
    const double lift = floor_rel > 0.0 ? floor_rel * mx : 0.0;
    const double eps = (mx + lift) * 1e-14;
    if (tid == 0) {
        eps_out[0] = eps;
        status[0] = 0;
    }

    // tile waves: slots, their tiles and operand offsets; the copy items
    const int r16 = lane & 15, kq = lane >> 4;
    double acc[kMaxT][4];  // tile slot s, element v (row l / 16 + 4 v, column l % 16)
    int tI[kMaxT], tJ[kMaxT], oI[kMaxT], oJ[kMaxT];
    int c_kk[kCP], c_q[kCP];
    // wave 0: the entering-row items
    int e_i[kPre], e_q[kPre];
    double pre[kPre], pre_b = 0.0;
    static_for<kMaxT>([&](auto S) __attribute__((always_inline)) {
        constexpr int s = S;
        const int t = tw >= 0 ? uniform(tab[tw][s]) : -1;
        tI[s] = t < 0 ? -1 : (t >> 8);
        tJ[s] = t < 0 ? -1 : (t & 255);
        oI[s] = kq * W + 16 * (t < 0 ? 0 : tI[s]) + r16;
        oJ[s] = kq * W + 16 * (t < 0 ? 0 : tJ[s]) + r16;
        acc[s][0] = acc[s][1] = acc[s][2] = acc[s][3] = 0.0;
    });
    static_for<kCP>([&](auto T) __attribute__((always_inline)) {
        constexpr int t = T;
        const int idx = tw >= 0 ? tw * 64 + lane + t * kTileLanes : -1;
        c_kk[t] = (idx >= 0 && idx < kP * W) ? idx / W : -1;
        c_q[t] = (idx >= 0 && idx < kP * W) ? idx - (idx / W) * W : 0;
    });
    static_for<kPre>([&](auto T) __attribute__((always_inline)) {
        constexpr int t = T;
        // the staging waves take alternate groups of 64 items
        const int idx = lane + 64 * (roles::kStagers * t + (wave == kStageWave ? 1 : 0));
        e_i[t] = idx / W;
        e_q[t] = idx - (idx / W) * W;
    });

    // the rows entering the window with panel m: row kP m + W + i at position
    // pp0(m) + i; item (i, q): its entry at the column position q holds in panel
    // m + 1's window (npos: that panel's first position)
    auto entering = [&](int m, int t, int npos, int& r, int& c) __attribute__((always_inline)) {
        const int rn = kP * m + W + e_i[t];
        int d = e_q[t] - npos;
        d += d < 0 ? W : 0;
        const int cq = kP * (m + 1) + d;
        r = cq <= rn ? rn : cq;
        c = cq <= rn ? cq : rn;
    };
    auto prefetch = [&](int m, int npos) __attribute__((always_inline)) {
        static_for<kPre>([&](auto T) __attribute__((always_inline)) {
            constexpr int t = T;
            int r, c;
            entering(m, t, npos, r, c);
            pre[t] = A[entry_index(n, wb, r, c)];
        });
        // validity applied when staged (a select here would wait for the loads)
        const int rn = kP * m + W + (lane & (kP - 1));
        pre_b = b[rn < n ? rn : 0];
    };
    auto stage = [&](int m, int npos) __attribute__((always_inline)) {
        double* S = SR + (m & 1) * kP * W;
        static_for<kPre>([&](auto T) __attribute__((always_inline)) {
            constexpr int t = T;
            int r, c;
            entering(m, t, npos, r, c);
            S[e_i[t] * W + e_q[t]] = entry_finish(pre[t], n, wb, lift, r, c);
        });
        if (wave == 0 && lane < kP) SRb[(m & 1) * kP + lane] = (kP * m + W + lane < n) ? pre_b : 0.0;
    };
    // tile update: acc -= L_I L_J^T over the panel's kP columns (lane l: operand
    // row l % 16 at depth l / 16; result element v at row l / 16 + 4 v, column
    // l % 16).  Operands of every slot load first, then the products run.
    auto update = [&](const double* L, auto want) __attribute__((always_inline)) {
        double a0[kMaxT], b0[kMaxT], a1[kMaxT], b1[kMaxT];
        static_for<kMaxT>([&](auto S) __attribute__((always_inline)) {
            constexpr int s = S;
            a0[s] = L[oI[s]];
            b0[s] = L[oJ[s]];
            a1[s] = L[oI[s] + 4 * W];
            b1[s] = L[oJ[s] + 4 * W];
        });
        static_for<kMaxT>([&](auto S) __attribute__((always_inline)) {
            constexpr int s = S;
            if (tI[s] >= 0 && want(tI[s], tJ[s])) {
                band_d4 c = {acc[s][0], acc[s][1], acc[s][2], acc[s][3]};
                c = __builtin_amdgcn_mfma_f64_16x16x4f64(-a0[s], b0[s], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f64_16x16x4f64(-a1[s], b1[s], c, 0, 0, 0);
                acc[s][0] = c[0];
                acc[s][1] = c[1];
                acc[s][2] = c[2];
                acc[s][3] = c[3];
            }
        });
    };
    // the columns of the panel at positions x0 .. x0 + 7 (block X, half h) to
    // PB[q][kk] = entry (position q, position x0 + kk); a diagonal tile writes
    // its rows only (its columns hold the same values)
    auto publish = [&](int X, int h, int x0) __attribute__((always_inline)) {
        static_for<kMaxT>([&](auto S) __attribute__((always_inline)) {
            constexpr int s = S;
            const int I = tI[s], J = tJ[s];
            if (I < 0) return;
            const int q = 16 * J + r16;
            if (I == X) {
                static_for<4>([&](auto V) __attribute__((always_inline)) {
                    constexpr int v = V;
                    if ((v >> 1) == h) PB[q * (kP + 1) + 16 * I + kq + 4 * v - x0] = acc[s][v];
                });
            } else if (J == X && (r16 >> 3) == h) {
                static_for<4>([&](auto V) __attribute__((always_inline)) {
                    constexpr int v = V;
                    PB[(16 * I + kq + 4 * v) * (kP + 1) + q - x0] = acc[s][v];
                });
            }
        });
    };
    // the positions x0 .. x0 + 7 (block X, half h) take the entering rows:
    // S[i][q] = entry (entering row i, position q)
    auto reload = [&](const double* S, int X, int h, int x0) __attribute__((always_inline)) {
        static_for<kMaxT>([&](auto SS) __attribute__((always_inline)) {
            constexpr int s = SS;
            const int I = tI[s], J = tJ[s];
            if (I < 0) return;
            const int q = 16 * J + r16;
            // selects, not conditional stores, and addresses inside S (rows
            // taken mod 8), so the slots stay in registers
            if (I == X) {
                static_for<4>([&](auto V) __attribute__((always_inline)) {
                    constexpr int v = V;
                    const double t = S[((kq + 4 * v - 8 * h) & (kP - 1)) * W + q];
                    acc[s][v] = ((v >> 1) == h) ? t : acc[s][v];
                });
            }
            if (J == X) {
                const bool mine = (r16 >> 3) == h;
                static_for<4>([&](auto V) __attribute__((always_inline)) {
                    constexpr int v = V;
                    const double t = S[((r16 - 8 * h) & (kP - 1)) * W + 16 * I + kq + 4 * v];
                    acc[s][v] = (mine && (I != X || (v >> 1) != h)) ? t : acc[s][v];
                });
            }
        });
    };
    // panel k's L to the factor buffers, by columns and by rows: the panel's own
    // rows from its diagonal block DB (DB[kk][t] = L_{c0+t, c0+kk}, 1 / L_jj on
    // the diagonal), the others from LP by position; rows past n get zeros
    auto copy_panel = [&](int k, int pp) __attribute__((always_inline)) {
        const double* L = LP + (k & 1) * kP * W;
        const double* D = DB + (k & 1) * kP * kP;
        const int c0 = kP * k;
        static_for<kCP>([&](auto T) __attribute__((always_inline)) {
            constexpr int t = T;
            const int kk = c_kk[t], q = c_q[t];
            int d = q - pp;
            d += d < 0 ? W : 0;
            if (kk >= 0 && d >= kk) {
                const int r = c0 + d, c = c0 + kk;
                double v = d < kP ? D[kk * kP + d] : L[kk * W + q];
                v = r < n ? v : 0.0;
                Lc[static_cast<int64_t>(c) * kRing + (d - kk)] = v;
                Lr[static_cast<int64_t>(r) * kRing + (d - kk)] = v;
            }
        });
    };

    // the window of panel 0 (rows 0 .. W - 1 at positions 0 .. W - 1), its first
    // panel's columns, the right-hand side; the rows entering with panel 0, and
    // the loads of those entering with panel 1
    if (stager) {
        if (wave == 0)
            for (int q = lane; q < W; q += 64) RH[q] = q < n ? b[q] : 0.0;
        if (NP >= 2) {
            prefetch(0, kP % W);
            stage(0, kP % W);
        }
        if (NP >= 3) prefetch(1, (2 * kP) % W);
    } else {
        static_for<kMaxT>([&](auto S) __attribute__((always_inline)) {
            constexpr int s = S;
            double raw[4];
            int rr[4], cc[4];
            static_for<4>([&](auto V) __attribute__((always_inline)) {
                constexpr int v = V;
                const int p = 16 * (tI[s] < 0 ? 0 : tI[s]) + kq + 4 * v, q = 16 * (tJ[s] < 0 ? 0 : tJ[s]) + r16;
                rr[v] = p > q ? p : q;
                cc[v] = p > q ? q : p;
                raw[v] = A[entry_index(n, wb, rr[v], cc[v])];
            });
            static_for<4>([&](auto V) __attribute__((always_inline)) {
                constexpr int v = V;
                acc[s][v] = entry_finish(raw[v], n, wb, lift, rr[v], cc[v]);
            });
        });
        publish(0, 0, 0);
    }
    __syncthreads();

    bool ok = true;  // wave 0
    double lmax = 0.0, lmin = INFINITY;
    int pp0 = 0;     // window position of column c0
    for (int k = 0; k < NP; ++k) {
        const int c0 = kP * k;
        const int kb = k & 1;
        const int ppp = pp0 == 0 ? W - kP : pp0 - kP;  // the previous panel's positions
        int ppn = pp0 + kP;                            // the next panel's
        if (ppn == W) ppn = 0;
        // ---- phase A
        if (wave == 0) {
            int q0 = pp0 + lane;
            if (q0 >= W) q0 -= W;
            const bool v1 = 64 + lane < W;
            int q1 = pp0 + 64 + lane;
            if (q1 >= W) q1 -= W;
            if (!v1) q1 = q0;
            double a0[kP], a1[kP];
            static_for<kP>([&](auto KK) __attribute__((always_inline)) {
                constexpr int kk = KK;
                a0[kk] = PB[q0 * (kP + 1) + kk];
                const double t = PB[q1 * (kP + 1) + kk];
                a1[kk] = v1 ? t : 0.0;
            });
            const double rb0 = RH[q0], rb1 = RH[q1];
            // the value refilling lane j < kP's position (the row entering there)
            const double refill = SRb[kb * kP + (lane & (kP - 1))];
            double b0 = rb0, b1 = v1 ? rb1 : 0.0;
            // lane j < kP ends with the panel's column j: 1 / sqrt(d_j), d_j and
            // the forward value; entries above the diagonal (lanes < j of
            // column j) are left unreduced: nothing reads them
            double yv = 0.0, dv = 0.0, bv = 0.0;
            static_for<kP>([&](auto JJ) __attribute__((always_inline)) {
                constexpr int j = JJ;
                const double d = lane_value(a0[j], j);
                const double y = rsqrt_nr(d);
                a0[j] = a0[j] * y;  // lane j: L_jj = d y
                a1[j] = a1[j] * y;
                const double yj = lane_value(b0, j) * y;
                b0 = fma(-a0[j], yj, b0);
                b1 = fma(-a1[j], yj, b1);
                const int ln = opaque_lane();
                yv = (ln == j) ? y : yv;
                dv = (ln == j) ? d : dv;
                bv = (ln == j) ? yj : bv;
                static_for<kP>([&](auto JN) __attribute__((always_inline)) {
                    constexpr int jj = JN;
                    if constexpr (jj > j) {
                        const double t = lane_value(a0[j], jj);
                        a0[jj] = fma(-a0[j], t, a0[jj]);
                        a1[jj] = fma(-a1[j], t, a1[jj]);
                    }
                });
            });
            // the pivot rules, one column per lane
            if (lane < kP && c0 + lane < n) {
                ok &= dv > eps;
                const double l = dv * yv;
                lmax = dmax_nan(lmax, fabs(l));
                lmin = fmin(lmin, fabs(l));
            }
            // Every lane stores every value: those with nothing to store write
            // their slot of the sink, so no store waits on an exec-mask update.
            double* L = LP + kb * kP * W;
            double* D = DB + kb * kP * kP;
            static_for<kP>([&](auto KK) __attribute__((always_inline)) {
                constexpr int kk = KK;
                L[kk * W + q0] = lane < kP ? 0.0 : a0[kk];
                *(v1 ? L + kk * W + q1 : sink) = a1[kk];
                *(lane < kP ? D + kk * kP + lane : sink) = (kk == lane) ? yv : a0[kk];
            });
            *(v1 ? RH + q1 : sink) = b1;
            RH[q0] = lane >= kP ? b0 : ((k + 2 <= NP) ? refill : 0.0);
            if (lane < kP && c0 + lane < n) x[c0 + lane] = bv;
        } else if (k > 0 && tw >= 0) {
            const int Xc = pp0 >> 4;
            update(LP + (kb ^ 1) * kP * W, [&](int I, int J) __attribute__((always_inline)) { return I != Xc && J != Xc; });
            copy_panel(k - 1, ppp);
        }
        __syncthreads();
        // ---- phase B
        if (k + 1 < NP) {
            if (stager) {
                int np2 = ppn + kP;
                np2 -= np2 >= W ? W : 0;
                int np3 = np2 + kP;
                np3 -= np3 >= W ? W : 0;
                if (k + 1 <= NP - 2) stage(k + 1, np2);
                if (k + 2 <= NP - 2) prefetch(k + 2, np3);
            } else if (tw >= 0) {
                const int Xc = pp0 >> 4, hc = (pp0 >> 3) & 1;
                const int Xn = ppn >> 4, hn = (ppn >> 3) & 1;
                reload(SR + kb * kP * W, Xc, hc, pp0);
                update(LP + kb * kP * W, [&](int I, int J) __attribute__((always_inline)) { return I == Xn || J == Xn; });
                publish(Xn, hn, ppn);
            }
        }
        __syncthreads();
        pp0 = ppn;
    }
    // the last panel's L to the factor buffers; the status
    {
        int ppl = pp0 - kP;
        if (ppl < 0) ppl += W;
        if (tw >= 0) copy_panel(NP - 1, ppl);
    }
    if (wave == 0) {
        // lanes 0 .. kP - 1 hold the pivot rules of their columns
        bool all_ok = true;
        double mxl = 0.0, mnl = INFINITY;
        static_for<kP>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = J;
            all_ok &= __builtin_amdgcn_readlane(ok ? 1 : 0, j) != 0;
            mxl = dmax_nan(mxl, lane_value(lmax, j));
            mnl = fmin(mnl, lane_value(lmin, j));
        });
        if (lane == 0) sflag = !all_ok ? 1 : ((mnl > 1e-7 * mxl) ? 0 : 2);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const int st = sflag;
    if (st != 0) {
        if (tid == 0) {
            status[0] = st;
            lm->solve_ok = 0;
        }
        return;
    }
    band_sub<W, false>(n, Lr, x, x, smem, wave, lane);
    if (tid == 0) lm->solve_ok = 1;
}

// x = S^-1 b with the factor of k_band_factor (forward, then backward).
template <int W>
__global__ void __launch_bounds__(kThreads) k_band_solve(int n, const double* Lband, const double* b, double* x,
                                                         const int32_t* status, lm_state* lm, int gated) {
    if (gated && lm->done) return;
    if (*status != 0) {
        if (threadIdx.x == 0) lm->solve_ok = 0;
        return;
    }
    __shared__ double smem[band_lds<W>::kSub];
    const int lane = threadIdx.x & 63;
    const int wave = uniform(static_cast<int>(threadIdx.x) >> 6);
    const double* Lc = Lband;
    const double* Lr = Lband + static_cast<int64_t>(sub_rows(n)) * kRing;
    band_sub<W, true>(n, Lc, b, x, smem, wave, lane);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    band_sub<W, false>(n, Lr, x, x, smem, wave, lane);
    if (threadIdx.x == 0) lm->solve_ok = 1;
}

}  // namespace

static_assert(kBandMaxW + kB - 1 <= kRing, "a block's steps reach only unknowns held in the ring");
static_assert(window_for(kBandMaxW) != 0, "the widest band has a window");

int band_window(int32_t band) { return (band > 0 && band <= kBandMaxW) ? window_for(band) : 0; }

size_t band_buffer_doubles(int32_t n, int32_t band) {
    const int W = band_window(band);
    if (W == 0 || n <= 0) return 0;
    // column-ordered rows [0, np), row-ordered rows [0, np + kRing)
    const size_t np = (static_cast<size_t>(n) + kB - 1) / kB * kB;
    return static_cast<size_t>(kRing) * (2 * np + kRing);
}

bool band_path(int32_t n, const dev_work64& w, bool exact) {
    if (exact || n <= 0 || w.band <= 0 || !w.Lband) return false;
    const int W = band_window(w.band);
    if (W == 0) return false;
    if (w.factor_path == 3) return true;
    return w.factor_path == 0 && n >= kBandMinRowsPerWindow * W;
}

void launch_chol_band(int32_t n, const dev_work64& w, hipStream_t s, bool gated, double floor_rel) {
    const int g = gated ? 1 : 0;
#define ZB_BAND_FACTOR(WW)                                                                                       \
    k_band_factor<WW><<<1, band_roles<WW>::kThreads, 0, s>>>(n, w.band, w.S, w.bs, w.dxp, w.Lband, floor_rel, w.chol_scale,  \
                                             w.chol_status, w.lm, g)
    switch (band_window(w.band)) {
    case 64: ZB_BAND_FACTOR(64); break;
    case 80: ZB_BAND_FACTOR(80); break;
    case 96: ZB_BAND_FACTOR(96); break;
    case 112: ZB_BAND_FACTOR(112); break;
    case 128: ZB_BAND_FACTOR(128); break;
    default: break;
    }
#undef ZB_BAND_FACTOR
}

void launch_band_solve(int32_t n, const dev_work64& w, const double* b, double* x, hipStream_t s, bool gated) {
    const int g = gated ? 1 : 0;
#define ZB_BAND_SOLVE(WW) k_band_solve<WW><<<1, kThreads, 0, s>>>(n, w.Lband, b, x, w.chol_status, w.lm, g)
    switch (band_window(w.band)) {
    case 64: ZB_BAND_SOLVE(64); break;
    case 80: ZB_BAND_SOLVE(80); break;
    case 96: ZB_BAND_SOLVE(96); break;
    case 112: ZB_BAND_SOLVE(112); break;
    case 128: ZB_BAND_SOLVE(128); break;
    default: break;
    }
#undef ZB_BAND_SOLVE
}

}  // namespace zb
