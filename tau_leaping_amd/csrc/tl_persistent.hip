// tl_persistent.hip -- the persistent engine: one work-stealing launch runs
// whole trajectories of every (model, realization) work item.
//
// A group of lanes owns one realization at a time: GS = 64 is one wavefront
// (several independent realizations per block, no block barriers); GS equal
// to the block size makes the whole block cooperate on one realization (for
// models with tens of thousands of channels).  Per group, LDS holds the
// populations x[N], the proposed change dx[N] (integer LDS atomics, so the
// update is exact and order-free), the criticality bits (one 64-bit ballot
// per 64 channels), a staging buffer for the first random draw of 4*GS
// channels, the species sums of the tau bound (one copy per wavefront) and,
// when it fits, the propensities a[M]; otherwise every pass recomputes a_j
// from x and the channel constants (shared by all realizations, so they stay
// in cache), and no per-realization array goes to global memory.  Reductions
// (a0, a0c, tau min / argmin, cumulative searches) are wavefront shuffles and
// ballots plus a fixed-order combine across waves, so results do not depend
// on timing.
//
// The per-realization algorithm follows PerformCalc / Step
// (tau_leaping/model.cu:301-413, 640-737) with the defect fixes of the CPU
// oracle (FP64 time and cumulative sums, tau2 over a0c, exact direct events,
// leap rejection on negative populations, clamp to t_final).
#include "tl_internal.hpp"

// This file is compiled twice: as is (TL_PERSISTENT_EXT 0, the kernels for
// the reference's reaction forms and sampler, plus the host side), and from
// tl_persistent_ext.hip (TL_PERSISTENT_EXT 1: the same kernels with sources,
// channels of three reactant molecules and the exact Poisson sampler).  Keeping those
// paths out of the default instances keeps their registers out of its hot
// loops.
#ifndef TL_PERSISTENT_EXT
#define TL_PERSISTENT_EXT 0
#endif

namespace tl {

namespace {

__host__ __device__ inline size_t al16(size_t v) { return (v + 15) & ~(size_t)15; }

// Global-memory pointers.  Pointers read from memory (the work-item table)
// are generic to the compiler, which then emits flat instructions; a flat
// load also counts against the LDS wait counter, so every wait for an LDS
// result would wait for it too.  Cast to the global address space, they give
// global loads.
template <class T>
using gptr = T __attribute__((address_space(1))) *;
template <class T>
__device__ __forceinline__ gptr<T> as_global(T *p) {
    return (gptr<T>)p;
}
// Constant-address-space view of read-only model data: a load at a
// wave-uniform index becomes a scalar load (scalar cache).
template <class T>
using cptr = const T __attribute__((address_space(4))) *;

// Native vector types (the HIP vector classes do not load through
// address-space-qualified pointers).
typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));

// The device model with global-address-space pointers.
struct GModel {
    int32_t N, M, nnz;
    gptr<const double> c64;
    gptr<const float> c32;
    gptr<const int32_t> row_ptr, csr, col_ptr, csc, x0, dep_ptr, dep, aff_ptr, aff, tile_ptr;
    gptr<const int8_t> g_order;
    gptr<const v2i> rs;
    gptr<const v4i> colq;  // null when a column has more than 4 entries
    gptr<const v2i> aff_rs;
    gptr<const double> aff_c64;
    gptr<const float> aff_c32;
    gptr<const int32_t> s2;  // null unless some channel has three reactant molecules
    gptr<const int32_t> row_order;  // species by CSR row length, longest first
    // column entries of a wave-uniform channel (scalar load)
    __device__ __forceinline__ int4 colq_uniform(int32_t j) const {
        const v4i v = ((cptr<v4i>)colq)[j];
        return make_int4(v.x, v.y, v.z, v.w);
    }
    __device__ __forceinline__ int32_t aff_ptr_uniform(int32_t j) const { return ((cptr<int32_t>)aff_ptr)[j]; }
    __device__ __forceinline__ int2 rs_at(int32_t j) const {
        const v2i v = rs[j];
        return make_int2(v.x, v.y);
    }
    __device__ __forceinline__ int4 colq_at(int32_t j) const {
        const v4i v = colq[j];
        return make_int4(v.x, v.y, v.z, v.w);
    }
    template <class R>
    __device__ __forceinline__ gptr<const R> rates() const {
        if constexpr (sizeof(R) == 8)
            return c64;
        else
            return c32;
    }
};
__device__ __forceinline__ GModel global_model(const DevModel &d) {
    GModel g;
    g.N = d.N;
    g.M = d.M;
    g.nnz = d.nnz;
    g.c64 = as_global(d.c64);
    g.c32 = as_global(d.c32);
    g.row_ptr = as_global(d.row_ptr);
    g.csr = as_global(d.csr);
    g.col_ptr = as_global(d.col_ptr);
    g.csc = as_global(d.csc);
    g.x0 = as_global(d.x0);
    g.dep_ptr = as_global(d.dep_ptr);
    g.dep = as_global(d.dep);
    g.aff_ptr = as_global(d.aff_ptr);
    g.aff = as_global(d.aff);
    g.tile_ptr = as_global(d.tile_ptr);
    g.g_order = as_global(d.g_order);
    g.rs = as_global(reinterpret_cast<const v2i *>(d.rs));
    g.colq = as_global(reinterpret_cast<const v4i *>(d.colq));
    g.aff_rs = as_global(reinterpret_cast<const v2i *>(d.aff_rs));
    g.aff_c64 = as_global(d.aff_c64);
    g.s2 = as_global(d.s2);
    g.aff_c32 = as_global(d.aff_c32);
    g.row_order = as_global(d.row_order);
    return g;
}

// LDS layout of one group.  Species sums mu[N], sig[N] of the tau bound;
// block groups without a[M] in LDS also hold one propensity tile.
struct GroupLayout {
    size_t off_x, off_dx, off_snap, off_crit, off_red, off_mu, off_sig, off_wsum, off_tile, off_a, bytes;
};
__host__ __device__ inline GroupLayout group_layout(int32_t n_max, int32_t m_max, int gs, bool a_lds,
                                                    size_t real_bytes, int tile_grains) {
    GroupLayout L;
    const int nw = gs / 64;
    L.off_x = 0;
    L.off_dx = al16((size_t)n_max * 4);
    L.off_snap = L.off_dx + al16((size_t)n_max * 4);
    L.off_crit = L.off_snap + al16((size_t)n_max * 4);
    L.off_red = L.off_crit + al16((size_t)((m_max + 63) / 64) * 8);
    L.off_mu = L.off_red + al16((size_t)(4 * nw + 4) * 8);
    L.off_sig = L.off_mu + al16((size_t)n_max * 8);
    L.off_wsum = L.off_sig + al16((size_t)n_max * 8);
    L.off_tile = L.off_wsum + (gs > 64 ? al16((size_t)((m_max + 63) / 64) * 8) : 0);
    const bool tile = gs > 64 && !a_lds;
    const int32_t tile_ch = tile_grains * kTileGrain;
    L.off_a = L.off_tile + (tile ? al16((size_t)(m_max < tile_ch ? m_max : tile_ch) * real_bytes) : 0);
    L.bytes = L.off_a + (a_lds ? al16((size_t)m_max * real_bytes) : 0);
    return L;
}

// LDS of a group whose state is in global memory: reduction scratch, then the
// propensity tile.
__host__ __device__ inline size_t hbm_tile_offset(int gs) { return al16((size_t)(4 * (gs / 64) + 4) * 8); }
__host__ __device__ inline size_t hbm_lds_bytes(int32_t m_max, int gs, bool a_stored, size_t real_bytes,
                                                int tile_grains) {
    if (a_stored) return hbm_tile_offset(gs);  // a[M] is in the global slice: no tile
    const int32_t tile_ch = tile_grains * kTileGrain;
    return hbm_tile_offset(gs) + al16((size_t)(m_max < tile_ch ? m_max : tile_ch) * real_bytes);
}

// ------------------------------------------------------ uniform values --
// Values every lane of a wave holds identically are moved to scalar
// registers, so model pointers, times and step sizes do not occupy VGPRs and
// branches on them are scalar branches.
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float uni(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uni(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ unsigned long long uni(unsigned long long v) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(v & 0xffffffffull));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
// Value of lane `l` (uniform) of a wave-varying double.
__device__ __forceinline__ double readlane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Load a struct that is the same for every lane into scalar registers.
template <class T>
__device__ __forceinline__ T ld_uniform(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "dword struct");
    T out;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(p);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)src[i]);
    return out;
}

// ------------------------------------------------------- wave primitives --
// Scans and reductions over a wavefront with DPP lane moves (no LDS round
// trip): an inclusive scan shifts within rows of 16 lanes by 1, 2, 4, 8, then
// broadcasts lane 15 into the next row (row_bcast:15 on rows 1 and 3) and
// lane 31 into rows 2-3 (row_bcast:31); lane 63 then holds the total.  Lanes
// without a source read the identity.  The order of additions is fixed, so
// every run sums identically.
constexpr int kDppRowShr = 0x110;
constexpr int kDppBcast15 = 0x142;
constexpr int kDppBcast31 = 0x143;
constexpr int kDppWaveShr1 = 0x138;

template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp_i(int v, int ident) {
    return __builtin_amdgcn_update_dpp(ident, v, CTRL, ROWS, 0xf, false);
}
// Lane move for sums: lanes whose source is out of range read zero
// (bound_ctrl) and every row is written, so no identity has to be placed in
// the destination first (10 fewer VALU instructions per FP64 wave scan).
template <int CTRL>
__device__ __forceinline__ int dpp_zero_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d_zero(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_zero_i<CTRL>((int)(b & 0xffffffffll));
    const int hi = dpp_zero_i<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL, ROWS>((int)(b & 0xffffffffll), 0);
    const int hi = dpp_i<CTRL, ROWS>((int)(b >> 32), 0);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ float dpp_f(float v, float ident) {
    return __int_as_float(dpp_i<CTRL, ROWS>(__float_as_int(v), __float_as_int(ident)));
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_d_ident(double v, double ident) {
    const long long b = __double_as_longlong(v), id = __double_as_longlong(ident);
    const int lo = dpp_i<CTRL, ROWS>((int)(b & 0xffffffffll), (int)(id & 0xffffffffll));
    const int hi = dpp_i<CTRL, ROWS>((int)(b >> 32), (int)(id >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double wave_incl_scan(double v, int /*wl*/) {
    v += dpp_d_zero<kDppRowShr | 1>(v);
    v += dpp_d_zero<kDppRowShr | 2>(v);
    v += dpp_d_zero<kDppRowShr | 4>(v);
    v += dpp_d_zero<kDppRowShr | 8>(v);
    v += dpp_d<kDppBcast15, 0xa>(v);
    v += dpp_d<kDppBcast31, 0xc>(v);
    return v;
}
// Value of the previous lane (0.0 in lane 0).
__device__ __forceinline__ double wave_shift_up(double v) { return dpp_d_zero<kDppWaveShr1>(v); }

// Wave total, the same in every lane (scalar register).
__device__ __forceinline__ double wave_sum(double v) { return readlane(wave_incl_scan(v, 0), 63); }

// Wave totals of x and of y together: the upper 32 lanes of x trade places
// with the lower 32 of y (v_permlane32_swap), so lane l < 32 holds
// x_l + x_(l+32) and lane l >= 32 holds y_(l-32) + y_l; a reduction over
// each half-wave then leaves the x total in lane 31, the y total in lane 63.
__device__ __forceinline__ double wave_sum_pair_lanes(double x, double y) {
    const long long bx = __double_as_longlong(x), by = __double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)bx, (unsigned)by, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(bx >> 32), (unsigned)(by >> 32), false, false);
    const double x2 = __longlong_as_double(((long long)hi[0] << 32) | (unsigned)lo[0]);
    const double y2 = __longlong_as_double(((long long)hi[1] << 32) | (unsigned)lo[1]);
    double v = x2 + y2;
    v += dpp_d_zero<kDppRowShr | 1>(v);
    v += dpp_d_zero<kDppRowShr | 2>(v);
    v += dpp_d_zero<kDppRowShr | 4>(v);
    v += dpp_d_zero<kDppRowShr | 8>(v);
    v += dpp_d<kDppBcast15, 0xa>(v);
    return v;
}

__device__ __forceinline__ float wave_incl_scan_f(float v) {
    v += __int_as_float(dpp_zero_i<kDppRowShr | 1>(__float_as_int(v)));
    v += __int_as_float(dpp_zero_i<kDppRowShr | 2>(__float_as_int(v)));
    v += __int_as_float(dpp_zero_i<kDppRowShr | 4>(__float_as_int(v)));
    v += __int_as_float(dpp_zero_i<kDppRowShr | 8>(__float_as_int(v)));
    v += dpp_f<kDppBcast15, 0xa>(v, 0.f);
    v += dpp_f<kDppBcast31, 0xc>(v, 0.f);
    return v;
}
template <class R>
__device__ __forceinline__ R wave_total(R v) {
    if constexpr (sizeof(R) == 8)
        return wave_sum(v);
    else
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wave_incl_scan_f(v)), 63));
}

template <class R>
__device__ __forceinline__ void argmin_step(R &v, int &i, int &w, R ov, int oi, int ow) {
    if (ov < v || (ov == v && oi < i)) {
        v = ov;
        i = oi;
        w = ow;
    }
}
template <int CTRL, int ROWS, class R>
__device__ __forceinline__ void argmin_dpp(R &v, int &i, int &w) {
    constexpr int kImax = 0x7fffffff;
    R ov;
    if constexpr (sizeof(R) == 8)
        ov = dpp_d_ident<CTRL, ROWS>(v, (double)INFINITY);
    else
        ov = dpp_f<CTRL, ROWS>(v, INFINITY);
    const int oi = dpp_i<CTRL, ROWS>(i, kImax);
    const int ow = dpp_i<CTRL, ROWS>(w, 0);
    argmin_step<R>(v, i, w, ov, oi, ow);
}
// Minimum value with the smallest index among ties; result in every lane.
template <class R>
__device__ __forceinline__ void wave_argmin(R &v, int &i, int &w) {
    argmin_dpp<kDppRowShr | 1, 0xf, R>(v, i, w);
    argmin_dpp<kDppRowShr | 2, 0xf, R>(v, i, w);
    argmin_dpp<kDppRowShr | 4, 0xf, R>(v, i, w);
    argmin_dpp<kDppRowShr | 8, 0xf, R>(v, i, w);
    argmin_dpp<kDppBcast15, 0xa, R>(v, i, w);
    argmin_dpp<kDppBcast31, 0xc, R>(v, i, w);
    if constexpr (sizeof(R) == 8)
        v = readlane((double)v, 63);
    else
        v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
    i = __builtin_amdgcn_readlane(i, 63);
    w = __builtin_amdgcn_readlane(w, 63);
}

__device__ __forceinline__ int wave_max_int(int v) {
    constexpr int kImin = (int)0x80000000;
    v = max(v, dpp_i<kDppRowShr | 1>(v, kImin));
    v = max(v, dpp_i<kDppRowShr | 2>(v, kImin));
    v = max(v, dpp_i<kDppRowShr | 4>(v, kImin));
    v = max(v, dpp_i<kDppRowShr | 8>(v, kImin));
    v = max(v, dpp_i<kDppBcast15, 0xa>(v, kImin));
    v = max(v, dpp_i<kDppBcast31, 0xc>(v, kImin));
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ bool crit_bit(const uint64_t *critq, int32_t j) { return (critq[j >> 6] >> (j & 63)) & 1ull; }

// Criticality bits of channels [lo, hi) within word w (bit b = channel 64w + b).
__device__ __forceinline__ uint64_t crit_word(const uint64_t *critq, int32_t w, int32_t lo, int32_t hi) {
    const int32_t b0 = max(lo - 64 * w, 0), b1 = min(hi - 64 * w, 64);
    if (b1 <= b0) return 0ull;
    const uint64_t upto = (b1 == 64) ? ~0ull : ((1ull << b1) - 1ull);
    return critq[w] & upto & ~((1ull << b0) - 1ull);
}

// First channel j of [j0, j1) (critical ones only when masked) at which
// excl + the cumulative weight passes target, else -1; `last` is the last
// channel with positive weight.
template <class P>
__device__ __forceinline__ int32_t walk_chunk(const P &pa, const uint64_t *critq, bool masked, int32_t j0, int32_t j1,
                                              double excl, double target, int32_t &last) {
    double cum = excl;
    if (masked) {
        if (j0 < j1)
            for (int32_t w = j0 >> 6; w <= (j1 - 1) >> 6; ++w)
                for (uint64_t bits = crit_word(critq, w, j0, j1); bits; bits &= bits - 1ull) {
                    const int32_t j = 64 * w + (int32_t)__builtin_ctzll(bits);
                    const double v = (double)pa.get(j);
                    if (v > 0.0) last = j;
                    cum += v;
                    if (cum > target) return j;
                }
        return -1;
    }
    for (int32_t j = j0; j < j1; ++j) {
        const double v = (double)pa.get(j);
        if (v > 0.0) last = j;
        cum += v;
        if (cum > target) return j;
    }
    return -1;
}

// ---------------------------------------------------------------- groups --
// GS == 64: one wavefront per realization.  Only wavefront-level
// synchronisation is used, so waves of a block never wait on each other.
template <int GS>
struct Grp {
    static constexpr int NW = GS / 64;
    int lane, wl, wid;
    double *red;

    // LDS-only exchange (see simulate_item)
    __device__ __forceinline__ void lds_sync() const {
        if constexpr (GS == 64) {
            __builtin_amdgcn_wave_barrier();
            __asm__ volatile("" ::: "memory");
        } else {
            __syncthreads();
        }
    }

    __device__ __forceinline__ void sync() const {
        if constexpr (GS == 64) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        } else {
            __syncthreads();
        }
    }

    __device__ __forceinline__ double sum(double v) const {
        v = wave_sum(v);
        if constexpr (GS == 64) {
            return uni(v);
        } else {
            if (wl == 0) red[wid] = v;
            __syncthreads();
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) s += red[w];
            __syncthreads();
            return uni(s);
        }
    }

    __device__ __forceinline__ void sum2(double &a, double &b) const {
        a = wave_sum(a);
        b = wave_sum(b);
        if constexpr (GS != 64) {
            if (wl == 0) {
                red[wid] = a;
                red[NW + wid] = b;
            }
            __syncthreads();
            a = 0.0;
            b = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                a += red[w];
                b += red[NW + w];
            }
            __syncthreads();
        }
        a = uni(a);
        b = uni(b);
    }

    __device__ __forceinline__ bool any(bool p) const {
        if constexpr (GS == 64)
            return __ballot(p) != 0ull;
        else
            return __syncthreads_or(p) != 0;
    }

    __device__ __forceinline__ int max_int(int v) const {
        v = wave_max_int(v);
        if constexpr (GS != 64) {
            int *ri = reinterpret_cast<int *>(red);
            if (wl == 0) ri[wid] = v;
            __syncthreads();
            v = ri[0];
#pragma unroll
            for (int w = 1; w < NW; ++w) v = max(v, ri[w]);
            __syncthreads();
        }
        return uni(v);
    }

    template <class R>
    __device__ __forceinline__ void argmin(R &v, int &i, int &w) const {
        wave_argmin<R>(v, i, w);
        if constexpr (GS != 64) {
            R *rv = reinterpret_cast<R *>(red);
            int *ri = reinterpret_cast<int *>(red + NW);
            if (wl == 0) {
                rv[wid] = v;
                ri[2 * wid] = i;
                ri[2 * wid + 1] = w;
            }
            __syncthreads();
            v = rv[0];
            i = ri[0];
            w = ri[1];
#pragma unroll
            for (int q = 1; q < NW; ++q) {
                const R ov = rv[q];
                const int oi = ri[2 * q];
                if (ov < v || (ov == v && oi < i)) {
                    v = ov;
                    i = oi;
                    w = ri[2 * q + 1];
                }
            }
            __syncthreads();
        }
        v = uni(v);
        i = uni(i);
        w = uni(w);
    }

    // Broadcast a 64-bit value from lane 0 of the group.
    __device__ __forceinline__ unsigned long long bcast0(unsigned long long v) const {
        if constexpr (GS == 64) {
            return uni(v);
        } else {
            unsigned long long *p = reinterpret_cast<unsigned long long *>(red + 3 * NW);
            if (lane == 0) *p = v;
            __syncthreads();
            v = *p;
            __syncthreads();
            return uni(v);
        }
    }

    // Cumulative weights in channel order: lane l owns the contiguous channels
    // [l*C, (l+1)*C) and sums them; a group scan over the lane sums gives each
    // lane's inclusive / exclusive prefix.  Weight = a_j, or a_j of critical
    // channels only.  Returns the total (the same in every lane).
    struct Scan {
        double incl, excl;
        int32_t j0, j1;
    };
    template <class P>
    __device__ double scan_chunks(const P &pa, const uint64_t *critq, bool masked, int32_t M, Scan &sc) const {
        const int32_t C = (M + GS - 1) / GS;
        const int32_t j0 = min(M, lane * C), j1 = min(M, j0 + C);
        double s = 0.0;
        if (masked) {
            // critical channels only: visit the set bits of the chunk
            if (j0 < j1)
                for (int32_t w = j0 >> 6; w <= (j1 - 1) >> 6; ++w)
                    for (uint64_t bits = crit_word(critq, w, j0, j1); bits; bits &= bits - 1ull)
                        s += (double)pa.get(64 * w + (int32_t)__builtin_ctzll(bits));
        } else {
            for (int32_t j = j0; j < j1; ++j) s += (double)pa.get(j);
        }
        return scan_sums(M, s, sc);
    }

    // The group scan of scan_chunks from lane chunk sums s computed elsewhere.
    __device__ double scan_sums(int32_t M, double s, Scan &sc) const {
        const int32_t C = (M + GS - 1) / GS;
        sc.j0 = min(M, lane * C);
        sc.j1 = min(M, sc.j0 + C);
        const double incl_w = wave_incl_scan(s, wl);
        double base = 0.0, total;
        if constexpr (GS != 64) {
            if (wl == 63) red[wid] = incl_w;
            __syncthreads();
            total = 0.0;
            for (int w = 0; w < NW; ++w) {
                if (w == wid) base = total;
                total += red[w];
            }
            __syncthreads();
        } else {
            total = readlane(incl_w, 63);
        }
        sc.incl = base + incl_w;
        // runs on every lane: a lane that is masked off does not provide its
        // value to the others
        sc.excl = base + wave_shift_up(incl_w);
        return uni(total);
    }

    // First j with inclusive cumulative weight > target; -1 if no such j
    // (total <= target).  The first lane whose prefix passes the target walks
    // its own channels.
    template <class P>
    __device__ int32_t search(const P &pa, const uint64_t *critq, bool masked, int32_t M, double target) const {
        Scan sc;
        scan_chunks<P>(pa, critq, masked, M, sc);
        return pick<P>(pa, critq, masked, sc, target);
    }
    // Block groups: first critical channel at which the cumulative critical
    // propensity passes target, from the per-word sums wsum[w] (channels
    // 64w..64w+63; filled by the propensity pass).  A group scan over the
    // word sums finds the word, then one wavefront scans its 64 channels.
    // -1 when the total does not pass target.
    template <class P>
    __device__ int32_t search_critical(const P &pa, const uint64_t *critq, int32_t M, const double *wsum,
                                       double target) const {
        static_assert(GS > 64, "block groups only");
        const int32_t nwd = (M + 63) / 64;
        const int32_t C = (nwd + GS - 1) / GS;
        const int32_t w0 = min(nwd, lane * C), w1 = min(nwd, w0 + C);
        double s = 0.0;
        for (int32_t w = w0; w < w1; ++w) s += wsum[w];
        Scan sc;
        scan_sums(nwd, s, sc);
        int *slot = reinterpret_cast<int *>(red + 2 * NW);
        double *sbase = red + 2 * NW + 2;  // slot[0..2] occupy the two doubles before
        if (lane == 0) slot[0] = 1 << 30;
        __syncthreads();
        const unsigned long long hit = __ballot(sc.incl > target);
        if (hit != 0ull && wl == 0) atomicMin(&slot[0], wid * 64 + (int)__builtin_ctzll(hit));
        __syncthreads();
        const int L = slot[0];
        if (L >= (1 << 30)) {
            __syncthreads();
            return -1;
        }
        if (lane == L) {  // the word within this lane's word chunk
            double cum = sc.excl;
            int32_t W = w1 - 1;
            for (int32_t w = w0; w < w1; ++w) {
                if (cum + wsum[w] > target) {
                    W = w;
                    break;
                }
                cum += wsum[w];
            }
            slot[1] = W;
            *sbase = cum;
        }
        __syncthreads();
        if (wid == 0) {
            const int32_t W = slot[1];
            const double base = *sbase;
            const int32_t j = 64 * W + wl;
            const double v = (j < M && crit_bit(critq, j)) ? (double)pa.get(j) : 0.0;
            const double incl = base + wave_incl_scan(v, wl);
            const unsigned long long h = __ballot(incl > target);
            const unsigned long long pos = __ballot(v > 0.0);
            int32_t res;
            if (h != 0ull)
                res = 64 * W + (int32_t)__builtin_ctzll(h);
            else  // rounding at the word end: the last positive channel
                res = pos ? 64 * W + 63 - (int32_t)__builtin_clzll(pos) : -1;
            if (wl == 0) slot[2] = res;
        }
        __syncthreads();
        const int32_t out = slot[2];
        __syncthreads();
        return uni(out);
    }

    template <class P>
    __device__ int32_t pick(const P &pa, const uint64_t *critq, bool masked, const Scan &sc, double target) const {
        const int32_t j0 = sc.j0, j1 = sc.j1;
        const double incl = sc.incl, excl = sc.excl;
        const unsigned long long hit = __ballot(incl > target);
        int32_t res = -1;
        if constexpr (GS == 64) {
            if (hit == 0ull) return -1;
            const int L = (int)__builtin_ctzll(hit);
            if (wl == L) {
                int32_t last = -1;
                res = walk_chunk(pa, critq, masked, j0, j1, excl, target, last);
                if (res < 0) res = last;  // rounding at the chunk end
            }
            return uni(__builtin_amdgcn_readlane(res, L));
        } else {
            int *slot = reinterpret_cast<int *>(red + 2 * NW);
            if (lane == 0) slot[0] = 1 << 30;
            __syncthreads();
            if (hit != 0ull && wl == 0) atomicMin(&slot[0], wid * 64 + (int)__builtin_ctzll(hit));
            __syncthreads();
            const int L = slot[0];
            if (lane == L) {
                int32_t last = -1;
                res = walk_chunk(pa, critq, masked, j0, j1, excl, target, last);
                slot[1] = (res < 0) ? last : res;
            }
            __syncthreads();
            const int32_t out = (L >= (1 << 30)) ? -1 : slot[1];
            __syncthreads();
            return uni(out);
        }
    }
};

// ----------------------------------------------------------------- passes --
__device__ __forceinline__ void chan_unpack(int2 p, int &order, int32_t &s0, int32_t &s1) {
    order = (int)((uint32_t)p.x >> 29);
    s0 = p.x & 0x1fffffff;
    s1 = p.y;
}

// a_j and criticality of channel j from its packed reactants.  s2 (a
// model-wide pointer, null unless some channel is a source or has three
// reactant molecules) selects the extended formulas with a scalar branch.
template <class R>
__device__ __forceinline__ R chan_prop(int2 p, R cj, const int32_t *xs, gptr<const int32_t> s2, int32_t j, bool &cr) {
    int order;
    int32_t s0, s1;
    chan_unpack(p, order, s0, s1);
    if (s2 != nullptr) return propensity_ext<R>(order, cj, xs[s0], xs[s1], order == 4 ? xs[s2[j]] : 1, cr);
    return propensity<R>(order, cj, xs[s0], xs[s1], cr);
}

// Where the passes read propensities: an array (kept in LDS, written by the
// propensity pass and the direct-event updates), or a recomputation from the
// populations in LDS and the shared channel constants.  Both give the same
// values: a_j is a function of (c_j, x) evaluated by one formula.
template <class R>
struct PropStored {
    static constexpr bool kStored = true;
    R *a;
    __device__ __forceinline__ R get(int32_t j) const { return a[j]; }
    __device__ __forceinline__ void set(int32_t j, R v) const { a[j] = v; }
};
template <class R>
struct PropRecomputed {
    static constexpr bool kStored = false;
    gptr<const v2i> rs;
    gptr<const R> c;
    const int32_t *xs;
    gptr<const int32_t> s2;
    __device__ __forceinline__ R get(int32_t j) const {
        bool cr;
        const v2i p = rs[j];
        return chan_prop<R>(make_int2(p.x, p.y), c[j], xs, s2, j, cr);
    }
    __device__ __forceinline__ void set(int32_t, R) const {}
};

// Adds of one column entry p of a channel with propensity aj to the species
// sums (v * aj, v^2 * aj; padding entries have v = 0).
__device__ __forceinline__ void musig_add(double *mu, double *sig, int32_t p, double aj) {
    const int32_t vi = entry_value(p);
    if (vi != 0) {
        const double v = (double)vi, tmp = v * aj;
        atomicAdd(&mu[entry_index(p)], tmp);
        atomicAdd(&sig[entry_index(p)], v * tmp);
    }
}

// x[i] += v * k for one column entry p.
__device__ __forceinline__ void column_add(int32_t *x, int32_t p, int32_t k) {
    const int32_t v = entry_value(p);
    if (v != 0) atomicAdd(&x[entry_index(p)], v * k);
}

// Propensities, criticality bits, and the lane partial sums of a0 / a0c
// (one-wavefront groups).  Each non-critical channel also adds v*a_j and
// v*(v*a_j) of its column entries to the species sums mu[i], sig[i] (LDS
// atomics; within one wavefront the order of the adds is fixed, so the sums
// repeat bit for bit).
template <class R, class P>
__device__ __forceinline__ bool prop_one(const GModel &m, int32_t j, int2 rsj, R cj, int4 q, const int32_t *xs,
                                         const P &pa, double *mu, double *sig, double &s, double &sc) {
    bool cr;
    const R aj = chan_prop<R>(rsj, cj, xs, m.s2, j, cr);
    pa.set(j, aj);
    s += (double)aj;
    if (cr) sc += (double)aj;
    if (!cr && aj != (R)0) {
        if (m.colq) {
            musig_add(mu, sig, q.x, (double)aj);
            musig_add(mu, sig, q.y, (double)aj);
            musig_add(mu, sig, q.z, (double)aj);
            musig_add(mu, sig, q.w, (double)aj);
        } else {
            const int32_t e1 = m.col_ptr[j + 1];
            for (int32_t e = m.col_ptr[j]; e < e1; ++e) musig_add(mu, sig, m.csc[e], (double)aj);
        }
    }
    return cr;
}

template <class R, int GS, class P>
__device__ __forceinline__ void prop_pass(const GModel &m, gptr<const R> c, const int32_t *xs, const P &pa,
                                          uint64_t *critq, double *mu, double *sig, int lane, double &s,
                                          double &sc) {
    s = 0.0;
    sc = 0.0;
    const bool wantq = m.colq != nullptr;
    for (int32_t j0 = 0; j0 < m.M; j0 += GS) {
        const int32_t j = j0 + lane;
        bool cr = false;
        if (j < m.M) {
            const int4 q = wantq ? m.colq_at(j) : make_int4(2, 2, 2, 2);
            cr = prop_one<R>(m, j, m.rs_at(j), c[j], q, xs, pa, mu, sig, s, sc);
        }
        const unsigned long long bal = __ballot(cr);
        const int32_t tile0 = j0 + (lane & ~63);
        if ((lane & 63) == 0 && tile0 < m.M) critq[tile0 >> 6] = bal;
    }
}

// Block groups: propensities, criticality bits and lane partial sums of
// a0 / a0c, then the species sums gathered along the CSR rows, one tile of
// TILE channels at a time (tile_grains grains; the tile's propensities sit in LDS: the
// whole a[M] when it fits, else a tile buffer).  A species row belongs to L
// consecutive lanes (L a power of two, about GS / N) that stride its entries
// in the tile and combine by butterfly; the row's first lane adds the tile's
// part to mu[i], sig[i].  No atomics, and a fixed order of additions.
template <class R, int GS, bool HBM, class P>
__device__ __forceinline__ void prop_tau_tiled(const GModel &m, gptr<const R> c, const int32_t *xs, const P &pa,
                                               R *tile, int tile_grains, uint64_t *critq, double *mu, double *sig,
                                               double *wsum, const Grp<GS> &G, double &s, double &sc) {
    const int lane = G.lane;
    s = 0.0;
    sc = 0.0;
    int L = 1;
    while (L < 64 && 2 * L * m.N <= GS) L *= 2;
    const int32_t TG = tile_grains, TILE = TG * kTileGrain;
    const int32_t ng = (m.M + kTileGrain - 1) / kTileGrain;  // grains (tile_ptr has ng + 1 rows)
    const int32_t nt = (m.M + TILE - 1) / TILE;
    const int32_t rows_per_pass = GS / L;
    // (not with the state in HBM: there the permuted rows would scatter the
    // species-sum and row-offset accesses that consecutive lanes now coalesce)
    const bool sorted_rows = !HBM && m.row_order != nullptr && m.N > rows_per_pass;
    const int k = lane & (L - 1);
    for (int32_t t = 0; t < nt; ++t) {
        const int32_t t0 = t * TILE;
        const int32_t t1 = min(m.M, t0 + TILE);
        const int32_t g0 = t * TG, g1 = min(ng, g0 + TG);
        R *A;  // A[j - aoff] holds a_j for j in the tile
        int32_t aoff;
        if constexpr (P::kStored) {
            A = pa.a;
            aoff = 0;
        } else {
            A = tile;
            aoff = t0;
        }
        // two channels per lane per round, constants loaded up front
        for (int32_t j0 = t0; j0 < t1; j0 += 2 * GS) {
            const int32_t ja = j0 + lane, jb = ja + GS;
            int2 ra = make_int2(0, 0), rb = make_int2(0, 0);
            R ca = (R)0, cb = (R)0;
            if (ja < t1) {
                ra = m.rs_at(ja);
                ca = c[ja];
            }
            if (jb < t1) {
                rb = m.rs_at(jb);
                cb = c[jb];
            }
            bool cra = false, crb = false;
            R aa = (R)0, ab = (R)0;
            // the tile buffer (recomputed propensities) keeps only what the
            // gather needs: a_j of non-critical channels, 0 for critical ones
            if (ja < t1) {
                aa = chan_prop<R>(ra, ca, xs, m.s2, ja, cra);
                A[ja - aoff] = (P::kStored || !cra) ? aa : (R)0;
                s += (double)aa;
                if (cra) sc += (double)aa;
            }
            if (jb < t1) {
                ab = chan_prop<R>(rb, cb, xs, m.s2, jb, crb);
                A[jb - aoff] = (P::kStored || !crb) ? ab : (R)0;
                s += (double)ab;
                if (crb) sc += (double)ab;
            }
            const unsigned long long bala = __ballot(cra), balb = __ballot(crb);
            const int32_t ta = j0 + (lane & ~63), tb = ta + GS;
            // each wavefront holds whole 64-channel words: their critical
            // propensity sums (for the critical search), both words in one
            // reduction
            double wa = 0.0, wb = 0.0;
            if ((bala | balb) != 0ull) {
                const double v = wave_sum_pair_lanes(cra ? (double)aa : 0.0, crb ? (double)ab : 0.0);
                wa = readlane(v, 31);
                wb = readlane(v, 63);
            }
            if ((lane & 63) == 0) {
                if (ta < t1) {
                    critq[ta >> 6] = bala;
                    wsum[ta >> 6] = wa;
                }
                if (tb < t1) {
                    critq[tb >> 6] = balb;
                    wsum[tb >> 6] = wb;
                }
            }
        }
        G.sync();
        for (int32_t i0 = 0; i0 < m.N; i0 += rows_per_pass) {
            // with more than one pass, rows in order of length, longest first:
            // the rows of one pass are of similar length
            const int32_t pos = i0 + lane / L;
            const int32_t i = pos >= m.N ? m.N : (sorted_rows ? m.row_order[pos] : pos);
            double pm = 0.0, ps = 0.0;
            if (i < m.N) {
                const int32_t e0 = m.tile_ptr[(size_t)g0 * m.N + i], e1 = m.tile_ptr[(size_t)g1 * m.N + i];
                // eight entries per round with their loads issued together; the
                // additions keep the lane's entry order
                auto add = [&](int32_t p) {
                    const int32_t j = entry_index(p);
                    if (P::kStored ? !crit_bit(critq, j) : true) {  // the tile holds 0 for critical channels
                        const double v = (double)entry_value(p);
                        const double tmp = v * (double)A[j - aoff];
                        pm += tmp;
                        ps += v * tmp;
                    }
                };
                int32_t e = e0 + k;
                for (; e + 7 * L < e1; e += 8 * L) {
                    int32_t p[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) p[u] = m.csr[e + u * L];
#pragma unroll
                    for (int u = 0; u < 8; ++u) add(p[u]);
                }
                for (; e < e1; e += L) add(m.csr[e]);
            }
            for (int o = L >> 1; o > 0; o >>= 1) {
                pm += __shfl_xor(pm, o, 64);
                ps += __shfl_xor(ps, o, 64);
            }
            if (i < m.N && k == 0) {
                mu[i] = (t == 0) ? pm : mu[i] + pm;
                sig[i] = (t == 0) ? ps : sig[i] + ps;
            }
        }
        G.sync();
    }
}

// Propensities of the channels whose reactants include a species of
// channel jf's column (after jf fired once); the values equal a full
// recomputation.
template <class R, int GS, class P>
__device__ __forceinline__ void prop_update(const GModel &m, gptr<const R> c, const int32_t *xs, const P &pa,
                                            int32_t jf, int lane) {
    if constexpr (!P::kStored) return;
    if (m.aff_ptr) {
        // the listed channel's index, rs and rate load side by side
        const int32_t d1 = m.aff_ptr_uniform(jf + 1);
        for (int32_t d = m.aff_ptr_uniform(jf) + lane; d < d1; d += GS) {
            const int32_t k = m.aff[d];
            const v2i p = m.aff_rs[d];
            R ck;
            if constexpr (sizeof(R) == 8)
                ck = m.aff_c64[d];
            else
                ck = m.aff_c32[d];
            bool cr;
            pa.set(k, chan_prop<R>(make_int2(p.x, p.y), ck, xs, m.s2, k, cr));
        }
        return;
    }
    const int32_t e1 = m.col_ptr[jf + 1];
    for (int32_t e = m.col_ptr[jf]; e < e1; ++e) {
        const int32_t i = entry_index(m.csc[e]);
        const int32_t d1 = m.dep_ptr[i + 1];
        for (int32_t d = m.dep_ptr[i] + lane; d < d1; d += GS) {
            const int32_t k = m.dep[d];
            bool cr;
            pa.set(k, chan_prop<R>(m.rs_at(k), c[k], xs, m.s2, k, cr));
        }
    }
}


// Visit every channel of one leap attempt with its firing count k_j.  Lane l
// takes the four consecutive channels 4l..4l+3 of each span of 4*GS
// channels, so the Philox block holding their first draws (channels 4b..4b+3
// share block b) is its own: no staging, no barrier inside the pass.  The
// caller synchronizes before reading what f wrote.
// CRIT_A (the score kernels): f sees every channel, in channel order, with
// a_j of critical channels too (else they get 0 and a recomputed propensity
// is not evaluated for them).  Without it f acts only on k_j != 0 (integer
// adds, in any order): a lane first settles the channels whose count needs
// no search (critical ones; those whose head draw lies below the bound of a
// zero count), then draws the rest, its r-th remaining channel in round r,
// so a wave runs as many searches as its busiest lane holds instead of one
// per channel position whenever any lane has one there.  The counts are the
// same either way.  The one-wavefront FP64 kernels of the default build keep
// the draw per position: there the compaction's extra live registers cost
// more than the searches it saves (toy network 128 x 256, 20,000
// realizations: 706 -> 736 ms; with it the 1024-lane colloid kernel takes
// 933 -> 847 ms, FP32 one-wavefront 688 -> 675 ms, the PTRS builds 286 ->
// 251 ms).
// pflags: kPoissonExact / kPoissonFast, honoured in the EXT build only (the
// default build carries none of PTRS's registers).
// tile, tile_lo: block groups with recomputed propensities hold a_j of the
// non-critical channels j >= tile_lo (the last tile of the propensity pass,
// the same values) in the tile; without CRIT_A they are read there instead
// of recomputed.  tile_lo = M: no tile.
template <class R, int GS, bool CRIT_A, bool EXT, class P, class F>
__device__ __forceinline__ void for_each_firing(const GModel &m, const Grp<GS> &G, const P &pa, const uint64_t *critq,
                                                uint2 key, uint32_t epoch, uint32_t c3, R tr, int32_t jfire,
                                                int pflags, const R *tile, int32_t tile_lo, F &&f) {
    const bool exact = EXT && (pflags & kPoissonExact) != 0;
    const bool fast = EXT && (pflags & kPoissonFast) != 0;
    const uint32_t ck = c3 | kKindPoisson;
    for (int32_t c0 = 0; c0 < m.M; c0 += 4 * GS) {
        const int32_t jb = c0 + 4 * G.lane;
        if (jb >= m.M) continue;
        const uint4 h = head_block(key, (uint32_t)(jb >> 2), epoch, c3);
        const uint32_t cw = (uint32_t)(critq[jb >> 6] >> (jb & 63)) & 15u;  // jb is a multiple of 4
        // recomputed propensities: the four first (their loads overlap)
        R a4[4] = {(R)0, (R)0, (R)0, (R)0};
        if constexpr (!P::kStored) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int32_t j = jb + i;
                a4[i] = (R)0;
                if (j < m.M && (CRIT_A || !((cw >> i) & 1u)))
                    a4[i] = (!CRIT_A && j >= tile_lo) ? tile[j - tile_lo] : pa.get(j);
            }
        }
        if constexpr (!CRIT_A && !(GS == 64 && sizeof(R) == 8 && !EXT)) {
            // todo: channels that need a draw, and the critical channel that
            // fires once (its bit also in once)
            uint32_t todo = 0, once = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int32_t j = jb + i;
                if (j < m.M) {
                    if ((cw >> i) & 1u) {
                        if (j == jfire) once = 1u << i;
                    } else {
                        R aj;
                        if constexpr (P::kStored)
                            aj = pa.get(j);
                        else
                            aj = a4[i];
                        if (!poisson_zero_by_bound<R>(aj * tr, quad_pick(h, i))) todo |= 1u << i;
                    }
                }
            }
            todo |= once;
            while (todo) {
                const int i = __builtin_ctz(todo);
                todo &= todo - 1;
                const int32_t j = jb + i;
                R aj;
                if constexpr (P::kStored)
                    aj = pa.get(j);
                else
                    aj = i == 0 ? a4[0] : (i == 1 ? a4[1] : (i == 2 ? a4[2] : a4[3]));
                const bool cr = (once >> i) & 1u;
                const int32_t kj =
                    cr ? 1 : poisson_draw<R>(aj * tr, quad_pick(h, i), key, (uint32_t)j, epoch, ck, exact, fast);
                if (kj) f(j, cr, aj, kj);
            }
            continue;
        }
#pragma unroll 1
        for (int i = 0; i < 4; ++i) {
            const int32_t j = jb + i;
            if (j < m.M) {
                const bool cr = crit_bit(critq, j);
                R aj;
                if constexpr (P::kStored)
                    aj = pa.get(j);
                else
                    aj = i == 0 ? a4[0] : (i == 1 ? a4[1] : (i == 2 ? a4[2] : a4[3]));
                const int32_t kj =
                    cr ? (j == jfire ? 1 : 0)
                       : poisson_draw<R>(aj * tr, quad_pick(h, i), key, (uint32_t)j, epoch, ck, exact, fast);
                f(j, cr, aj, kj);
            }
        }
    }
}

// ------------------------------------------------------------ one item --
// Tracked species on the time grid: linear interpolation between the
// states before (snap, at t_a) and after (xs, at t_b) one main-loop step, as
// AllStat2StepStat does between logged states (statistics.cpp:15-64).
template <int GS>
__device__ __forceinline__ void grid_record(const ModelItem &mi, int64_t r, int lane, const int32_t *snap,
                                            const int32_t *xs, double t_a, double t_b, int32_t &next_g) {
    const int32_t n = mi.n_intervals, T = mi.n_track;
    const double h = mi.t_final / (double)n;
    while (next_g < n) {
        const double g = h * (double)next_g;
        if (!(g < t_b)) break;
        const double w = (t_b > t_a) ? (g - t_a) / (t_b - t_a) : 0.0;
        gptr<double> out = as_global(mi.grid + ((size_t)r * (n + 1) + next_g) * T);
        for (int32_t q = lane; q < T; q += GS) {
            const double xa = (double)snap[q], xb = (double)xs[as_global(mi.track)[q]];
            out[q] = xa + (xb - xa) * w;
        }
        next_g++;
    }
}

template <class R, int GS, bool SENS, bool EXT, bool HBM, class P>
__device__ void simulate_item(const ModelItem &mi, int64_t r, const Grp<GS> &G, int32_t *xs, int32_t *dx,
                              int32_t *snap, uint64_t *critq, double *mu, double *sig, double *wsum, R *tile, int tile_grains,
                              const P &pa) {
    GModel m_ = global_model(mi.dm);
    if constexpr (!EXT) m_.s2 = nullptr;  // the extended formulas are compiled out
    const GModel &m = m_;
    // Group barrier.  Without sensitivities the lanes of a one-wavefront group
    // only exchange data through LDS, whose operations a wavefront performs in
    // order, so a code-motion barrier suffices (no wait on outstanding global
    // loads); scores in global memory need the full fences.
    auto sync_ = [&]() {
        if constexpr (SENS)
            G.sync();
        else
            G.lds_sync();
    };
    const int lane = G.lane;
    const gptr<const R> c = m.template rates<R>();
    const uint2 key = rng_key(mi.seed, mi.rid_offset + r);
    const uint32_t c3 = mi.model_id << 2;
    const double tf = mi.t_final;
    const R eps = (R)mi.eps;
    const int pflags = EXT ? mi.poisson : 0;
    const bool exact = (pflags & kPoissonExact) != 0;
    // Likelihood-ratio score S_j = dlog p(path)/dc_j, and D_j = dt/dc_j of the
    // current time (nonzero through tau1-limited leaps; used by the clamped
    // last leap and the truncated last direct interval).
    gptr<double> S = nullptr, D = nullptr, TL = nullptr;
    if (SENS) {
        S = as_global(mi.score + (size_t)r * m.M);
        D = as_global(mi.dtime + (size_t)r * m.M);
        TL = as_global(mi.tlast + (size_t)r * m.M);
    }
    // Direct-event exposure -a_j dt / c_j is added lazily by one-wavefront
    // groups with propensities in LDS and affected-channel lists: a channel's
    // exposure is flushed when its propensity changes and at the end of the
    // block (O(affected channels) per event instead of O(M)).
    const bool lazy = SENS && GS == 64 && P::kStored && m.aff_ptr != nullptr;
    // first channel of the propensity pass's last tile (block groups with
    // recomputed propensities; M: none)
    int32_t tile_lo = m.M;
    if constexpr (GS > 64 && !P::kStored) {
        const int32_t TILE = tile_grains * kTileGrain;
        if (m.M > 0) tile_lo = ((m.M - 1) / TILE) * TILE;
    }

    for (int32_t i = lane; i < m.N; i += GS) {
        xs[i] = m.x0[i];
        dx[i] = 0;
    }
    if (GS == 64)  // scattered sums start from zero (the tiled pass assigns them)
        for (int32_t i = lane; i < m.N; i += GS) {
            mu[i] = 0.0;
            sig[i] = 0.0;
        }
    if (SENS)
        for (int32_t j = lane; j < m.M; j += GS) {
            S[j] = 0.0;
            D[j] = 0.0;
        }
    sync_();

    double t = 0.0;
    uint32_t epoch = 0;
    int64_t steps = 0, nssa = 0, nrej = 0;
    int32_t status = kOk;
    const bool grid = mi.n_intervals > 0;
    int32_t next_g = 0;
    // direct-event uniforms, 64 events at a time: lane l holds r1 and -ln(r2)
    // of the event at epoch ssa_base + l
    uint32_t ssa_base = 0;
    bool ssa_valid = false;
    double ssa_r1 = 0.0, ssa_nl = 0.0;

    while (t < tf) {
        if (mi.max_iter > 0 && steps >= mi.max_iter) {
            status = kIterCap;
            break;
        }
        const double t_before = t;
        if (grid)
            for (int32_t q = lane; q < mi.n_track; q += GS) snap[q] = xs[as_global(mi.track)[q]];
        double a0, a0c;
        if constexpr (GS == 64)
            prop_pass<R, GS>(m, c, xs, pa, critq, mu, sig, lane, a0, a0c);
        else
            prop_tau_tiled<R, GS, HBM>(m, c, xs, pa, tile, tile_grains, critq, mu, sig, wsum, G, a0, a0c);
        G.sum2(a0, a0c);
        sync_();
        if (!(a0 > 0.0)) {  // absorbing state
            t = tf;
            steps++;
            if (grid) grid_record<GS>(mi, r, lane, snap, xs, t_before, t, next_g);
            break;
        }
        // tau1: per species mean / variance bound, min and argmin
        R best = (R)kFltMax;
        int bi = -1, bw = 0;
        // tau_i from the species sums (zeroed after use for the next scatter)
        for (int32_t i = lane; i < m.N; i += GS) {
            const R mui = (R)mu[i], sgi = (R)sig[i];
            if (GS == 64) {
                mu[i] = 0.0;
                sig[i] = 0.0;
            }
            if (i < m.M) {
                int w;
                const R ti = tau_species<R>(eps, xs[i], m.g_order[i], mui, sgi, w);
                if (ti < best) {
                    best = ti;
                    bi = i;
                    bw = w;
                }
            }
        }
        G.template argmin<R>(best, bi, bw);
        double tau1 = (double)best;
        const int32_t istar = bi;
        const int which = bw;

        int retries = 0;
        for (;;) {
            if (tau1 < kSsaFactor / a0) {
                // direct branch: up to 100 exact events; after each event only
                // the channels that read a changed species are recomputed
                bool flushed = false;
                if (lazy) {
                    for (int32_t j = lane; j < m.M; j += GS) TL[j] = t;
                    __threadfence_block();
                    G.sync();
                }
                for (int s = 0; s < kSsaSteps; ++s) {
                    typename Grp<GS>::Scan sc;
                    const double a0s = G.scan_chunks(pa, critq, false, m.M, sc);
                    if (!(a0s > 0.0)) {
                        t = tf;
                        break;
                    }
                    if (!ssa_valid || epoch - ssa_base >= 64u) {
                        ssa_base = epoch;
                        ssa_valid = true;
                        double r1l, r2l;
                        step_uniforms(key, epoch + (uint32_t)G.wl, c3 | kKindSsa, r1l, r2l);
                        ssa_r1 = r1l;
                        ssa_nl = -log(r2l);
                    }
                    const int li = (int)uni(epoch - ssa_base);
                    const double r1 = uni(readlane(ssa_r1, li));
                    const double dt = uni(readlane(ssa_nl, li)) / a0s;
                    epoch++;
                    if (dt >= tf - t) {
                        if (SENS) {  // survival over rem = tf - t: d/dc_j of -a0*rem
                            const double rem = tf - t;
                            for (int32_t j = lane; j < m.M; j += GS)
                                if (c[j] != (R)0)
                                    S[j] += a0s * D[j] - (double)pa.get(j) * (lazy ? tf - TL[j] : rem) / (double)c[j];
                            flushed = true;
                        }
                        t = tf;
                        break;
                    }
                    int32_t jf = G.pick(pa, critq, false, sc, a0s * r1);
                    if (jf < 0) {
                        int last = -1;
                        for (int32_t j = lane; j < m.M; j += GS)
                            if (pa.get(j) > (R)0) last = j;
                        jf = G.max_int(last);
                    }
                    if (SENS) {
                        if (lazy) {
                            // the event's own term, then the exposure of the channels
                            // whose propensity this event changes, up to the event time
                            if (lane == 0 && c[jf] != (R)0) S[jf] += 1.0 / (double)c[jf];
                            const double te = t + dt;
                            const int32_t d1 = m.aff_ptr_uniform(jf + 1);
                            for (int32_t d = m.aff_ptr_uniform(jf) + lane; d < d1; d += GS) {
                                const int32_t k = m.aff[d];
                                if (c[k] != (R)0) S[k] -= (double)pa.get(k) * (te - TL[k]) / (double)c[k];
                                TL[k] = te;
                            }
                        } else {
                            for (int32_t j = lane; j < m.M; j += GS)
                                if (c[j] != (R)0)
                                    S[j] += ((j == jf) ? 1.0 / (double)c[j] : 0.0) - (double)pa.get(j) * dt / (double)c[j];
                        }
                    }
                    if constexpr (SENS && !P::kStored) G.sync();
                    if (m.colq) {
                        const int4 q = m.colq_uniform(jf);
                        if (lane < 4) column_add(xs, lane == 0 ? q.x : (lane == 1 ? q.y : (lane == 2 ? q.z : q.w)), 1);
                    } else {
                        const int32_t e1 = m.col_ptr[jf + 1];
                        for (int32_t e = m.col_ptr[jf] + lane; e < e1; e += GS) column_add(xs, m.csc[e], 1);
                    }
                    sync_();
                    prop_update<R, GS>(m, c, xs, pa, jf, lane);
                    t += dt;
                    nssa++;
                    sync_();
                }
                if (lazy && !flushed) {  // exposure of every channel up to the block's end
                    __threadfence_block();
                    for (int32_t j = lane; j < m.M; j += GS)
                        if (c[j] != (R)0) S[j] -= (double)pa.get(j) * (t - TL[j]) / (double)c[j];
                    G.sync();
                }
                steps++;
                break;
            }

            // ---- leap attempt ----
            double r1, r2;
            step_uniforms(key, epoch, c3 | kKindLeap, r1, r2);
            int32_t jc = m.M;
            double tau2 = (double)kFltMax;
            // The critical channel is searched for only when its waiting time
            // t2 is below tau1: otherwise tau = tau1 and nothing fires whatever
            // the search returns (tau1 <= t2 <= tau2).
            const double t2 = a0c > 0.0 ? -log(r2) / a0c : (double)kFltMax;
            if (a0c > 0.0 && t2 < tau1) {
                int32_t f;
                if constexpr (GS == 64)
                    f = G.search(pa, critq, true, m.M, a0c * r1);
                else
                    f = G.search_critical(pa, critq, m.M, wsum, a0c * r1);
                if (f >= 0) {
                    jc = f;
                    tau2 = t2;
                }
            }
            double tau = tau1 < tau2 ? tau1 : tau2;
            bool fire = (jc != m.M) && (tau2 < tau1);
            bool clamped = false;
            if (tau >= tf - t) {
                tau = tf - t;
                clamped = true;
                fire = false;
            }
            const int32_t jfire = fire ? jc : -1;
            const R tr = (R)tau;
            double qpart = 0.0;
            auto fire_one = [&](int32_t j, bool cr, R aj, int32_t kj) {
                                       if (kj) {
                                           if (m.colq) {
                                               const int4 q = m.colq_at(j);
                                               column_add(dx, q.x, kj);
                                               column_add(dx, q.y, kj);
                                               column_add(dx, q.z, kj);
                                               column_add(dx, q.w, kj);
                                           } else {
                                               const int32_t e1 = m.col_ptr[j + 1];
                                               for (int32_t e = m.col_ptr[j]; e < e1; ++e) column_add(dx, m.csc[e], kj);
                                           }
                                       }
                                       if (SENS) {
                                           const double term = score_term<R>(cr, aj, tr, tau, kj, exact);
                                           qpart += term;
                                           if (c[j] != (R)0) S[j] += term / (double)c[j];
                                       }
                                   };
            for_each_firing<R, GS, SENS, EXT>(m, G, pa, critq, key, epoch, c3, tr, jfire, pflags, tile, tile_lo,
                                              fire_one);
            sync_();  // dx complete
            bool neg = false;
            for (int32_t i = lane; i < m.N; i += GS) neg |= (xs[i] + dx[i]) < 0;
            neg = G.any(neg);
            if (neg) {
                for (int32_t i = lane; i < m.N; i += GS) dx[i] = 0;
                if (SENS) {  // take back this attempt's score terms (same draws)
                    auto undo = [&](int32_t j, bool cr, R aj, int32_t kj) {
                        if (c[j] != (R)0) S[j] -= score_term<R>(cr, aj, tr, tau, kj, exact) / (double)c[j];
                    };
                    for_each_firing<R, GS, SENS, EXT>(m, G, pa, critq, key, epoch, c3, tr, jfire, pflags, tile,
                                                      tile_lo, undo);
                }
                epoch++;
                nrej++;
                sync_();
                if (++retries > kMaxRetries) {
                    status = kRetryExhausted;
                    break;
                }
                tau1 *= 0.5;
                continue;
            }
            epoch++;
            // (score terms first: they read this leap's propensities)
            if (SENS && clamped) {
                // tau = tf - t: dtau/dc_j = -D_j
                const double Q = G.sum(qpart);
                for (int32_t j = lane; j < m.M; j += GS) S[j] -= (Q / tau) * D[j];
            } else if (SENS && jfire < 0 && istar >= 0) {
                // tau = tau1: add Q * dtau1/dc_j over the argmin species row
                const double Q = G.sum(qpart);
                double mu = 0.0, sig = 0.0;
                const int32_t e0 = m.row_ptr[istar], e1 = m.row_ptr[istar + 1];
                for (int32_t e = e0 + lane; e < e1; e += GS) {
                    const int32_t p = m.csr[e], j = entry_index(p);
                    if (crit_bit(critq, j)) continue;
                    const double v = (double)entry_value(p);
                    mu += v * (double)pa.get(j);
                    sig += v * v * (double)pa.get(j);
                }
                G.sum2(mu, sig);
                for (int32_t e = e0 + lane; e < e1; e += GS) {
                    const int32_t p = m.csr[e], j = entry_index(p);
                    if (crit_bit(critq, j) || c[j] == (R)0) continue;
                    const double v = (double)entry_value(p);
                    const double h = (double)pa.get(j) / (double)c[j];
                    const double dtau = (which == 0) ? -tau * (mu > 0.0 ? 1.0 : -1.0) * v * h / fabs(mu)
                                                     : -tau * v * v * h / sig;
                    atomicAdd((double *)&S[j], (Q / tau) * dtau);
                    atomicAdd((double *)&D[j], dtau);
                }
            }
            if constexpr (SENS && !P::kStored) G.sync();
            for (int32_t i = lane; i < m.N; i += GS) {
                xs[i] += dx[i];
                dx[i] = 0;
            }
            const double tn = clamped ? tf : t + tau;
            if (!(tn > t)) {  // time can not advance: stop instead of spinning
                status = kStalled;
                break;
            }
            t = tn;
            steps++;
            sync_();
            break;
        }
        if (status != kOk) break;
        if (grid) grid_record<GS>(mi, r, lane, snap, xs, t_before, t, next_g);
    }

    if (grid && status == kOk) {  // grid point t_final (and any left) takes the final state
        const int32_t n = mi.n_intervals, T = mi.n_track;
        for (; next_g <= n; ++next_g) {
            gptr<double> out = as_global(mi.grid + ((size_t)r * (n + 1) + next_g) * T);
            for (int32_t q = lane; q < T; q += GS) out[q] = (double)xs[as_global(mi.track)[q]];
        }
    }

    // outputs
    for (int32_t q = lane; q < mi.n_track; q += GS) as_global(mi.out_x)[r * mi.n_track + q] = xs[as_global(mi.track)[q]];
    if (lane == 0) {
        as_global(mi.out_steps)[r] = steps;
        as_global(mi.out_ssa)[r] = nssa;
        as_global(mi.out_rej)[r] = nrej;
        as_global(mi.out_status)[r] = status;
    }
    sync_();
}

// WPE: waves per SIMD the register allocation targets (more waves hide more
// latency; fewer give each wave more registers).
// HBM: the group's state (populations, proposal, snapshot, criticality words,
// species sums, word sums, and a[M] with ALDS) lives in a slice of global
// memory per block instead of LDS, for models whose state does not fit LDS;
// only the reduction scratch (and without ALDS the propensity tile) stays in
// LDS.  Block groups only:
// their barriers are full workgroup barriers, which order global memory
// between the waves of a block as they order LDS.
template <class R, int BS, int GS, bool ALDS, bool SENS, int WPE, bool HBM, bool EXT>
__global__ __launch_bounds__(BS, WPE) void k_persistent(PersistentLaunch L) {
    static_assert(!HBM || (GS == BS && GS > 64), "HBM state: block groups");
    extern __shared__ __align__(16) unsigned char lds[];
    const GroupLayout lay = group_layout(L.n_max, L.m_max, GS, ALDS, sizeof(R), L.tile_grains);
    const int tid = threadIdx.x;
    const int g = tid / GS;
    Grp<GS> G;
    G.lane = tid % GS;
    G.wl = tid & 63;
    G.wid = G.lane >> 6;
    unsigned char *base = HBM ? L.state + (size_t)blockIdx.x * L.state_stride : lds + (size_t)g * lay.bytes;
    int32_t *xs = reinterpret_cast<int32_t *>(base + lay.off_x);
    int32_t *dx = reinterpret_cast<int32_t *>(base + lay.off_dx);
    int32_t *snap = reinterpret_cast<int32_t *>(base + lay.off_snap);
    uint64_t *critq = reinterpret_cast<uint64_t *>(base + lay.off_crit);
    G.red = reinterpret_cast<double *>(HBM ? lds : base + lay.off_red);
    double *mu = reinterpret_cast<double *>(base + lay.off_mu);
    double *sig = reinterpret_cast<double *>(base + lay.off_sig);
    R *tile = reinterpret_cast<R *>(HBM ? lds + hbm_tile_offset(GS) : base + lay.off_tile);
    double *wsum = reinterpret_cast<double *>(base + lay.off_wsum);

    for (;;) {
        unsigned long long v = 0;
        if (G.lane == 0) v = atomicAdd(L.counter, 1ull);
        const int64_t item = (int64_t)G.bcast0(uni(v));
        if (item >= L.n_items) break;
        int lo = 0, hi = L.n_models - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (L.items[mid].item_base <= item) lo = mid;
            else hi = mid - 1;
        }
        const ModelItem mi = ld_uniform(&L.items[uni(lo)]);
        if constexpr (ALDS) {
            const PropStored<R> pa{reinterpret_cast<R *>(base + lay.off_a)};
            simulate_item<R, GS, SENS, EXT, HBM>(mi, item - mi.item_base, G, xs, dx, snap, critq, mu, sig, wsum, tile,
                                                 L.tile_grains, pa);
        } else {
            const GModel gm = global_model(mi.dm);
            const PropRecomputed<R> pa{gm.rs, gm.template rates<R>(), xs, EXT ? gm.s2 : nullptr};
            simulate_item<R, GS, SENS, EXT, HBM>(mi, item - mi.item_base, G, xs, dx, snap, critq, mu, sig, wsum, tile,
                                                 L.tile_grains, pa);
        }
    }
}

// ---------------------------------------------------------- dispatch ----
constexpr bool kExt = TL_PERSISTENT_EXT != 0;

template <class R, bool SENS>
const void *kernel_ptr(int bs, int gs, bool alds, int wpe, bool hbm) {
    if (hbm) {
        if (bs != 1024 || gs != 1024 || wpe != 1) return nullptr;
        return alds ? (const void *)k_persistent<R, 1024, 1024, true, SENS, 1, true, kExt>
                    : (const void *)k_persistent<R, 1024, 1024, false, SENS, 1, true, kExt>;
    }
#define TL_K(BS_, GS_, W_)                                                                             \
    if (bs == BS_ && gs == GS_ && wpe == W_)                                                           \
        return alds ? (const void *)k_persistent<R, BS_, GS_, true, SENS, W_, false, kExt>             \
                    : (const void *)k_persistent<R, BS_, GS_, false, SENS, W_, false, kExt>;
    TL_K(64, 64, 4)
    TL_K(64, 64, 5)
    TL_K(256, 64, 4)
    TL_K(256, 64, 5)
    TL_K(64, 64, 6)
    TL_K(256, 64, 6)
    TL_K(256, 256, 4)
    TL_K(256, 256, 1)
    TL_K(512, 512, 1)
    TL_K(512, 512, 4)
    TL_K(1024, 1024, 1)
#undef TL_K
    return nullptr;
}

// the kernels of this translation unit
const void *kernel_for_tu(const PersistentConfig &c, bool is64, bool sens) {
    if (is64)
        return sens ? kernel_ptr<double, true>(c.block, c.group, c.a_lds, c.wpe, c.hbm)
                    : kernel_ptr<double, false>(c.block, c.group, c.a_lds, c.wpe, c.hbm);
    return sens ? kernel_ptr<float, true>(c.block, c.group, c.a_lds, c.wpe, c.hbm)
                : kernel_ptr<float, false>(c.block, c.group, c.a_lds, c.wpe, c.hbm);
}

}  // namespace

#if TL_PERSISTENT_EXT
const void *kernel_for_ext(const PersistentConfig &c, bool is64, bool sens) { return kernel_for_tu(c, is64, sens); }
#else
const void *kernel_for_ext(const PersistentConfig &c, bool is64, bool sens);  // tl_persistent_ext.hip

static const void *kernel_for(const PersistentConfig &c, bool is64, bool sens) {
    return c.ext ? kernel_for_ext(c, is64, sens) : kernel_for_tu(c, is64, sens);
}

size_t persistent_lds_bytes(const PersistentConfig &c, bool is64, int32_t n_max, int32_t m_max) {
    if (c.hbm) return hbm_lds_bytes(m_max, c.group, c.a_lds, is64 ? 8 : 4, c.tile_grains);
    const GroupLayout L = group_layout(n_max, m_max, c.group, c.a_lds, is64 ? 8 : 4, c.tile_grains);
    return L.bytes * (size_t)(c.block / c.group);
}

size_t persistent_state_bytes(const PersistentConfig &c, bool is64, int32_t n_max, int32_t m_max) {
    if (!c.hbm) return 0;
    const size_t b = group_layout(n_max, m_max, c.group, c.a_lds, is64 ? 8 : 4, c.tile_grains).bytes;
    return (b + 255) & ~(size_t)255;
}

int persistent_occupancy(const PersistentConfig &c, bool is64, bool sens, size_t lds_bytes) {
    const void *k = kernel_for(c, is64, sens);
    if (!k) return 0;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, c.block, lds_bytes) != hipSuccess) return 0;
    return n;
}

hipError_t launch_persistent(const PersistentConfig &c, bool is64, bool sens, const PersistentLaunch &L, hipStream_t s) {
    const void *k = kernel_for(c, is64, sens);
    if (!k) return hipErrorInvalidDeviceFunction;
    void *args[] = {const_cast<PersistentLaunch *>(&L)};
    return hipLaunchKernel(k, dim3((unsigned)L.grid), dim3((unsigned)c.block), args, L.lds_bytes, s);
}

#endif  // !TL_PERSISTENT_EXT

}  // namespace tl
