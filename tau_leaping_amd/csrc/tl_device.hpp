// tl_device.hpp -- device-side building blocks shared by the stage kernels
// and the persistent engine (gfx950).
//
// Reference formulas (file:line in the reference tree):
//   propensity / criticality ... tau_leaping/functions.h:100-123 (nc = 5, model.cu:225)
//   GetG ........................ tau_leaping/CSRMatrixPacked.cu:294-309
//   tau_i ....................... tau_leaping/CSRMatrixPacked.cu:329-355
//   Poisson / Box-Muller ........ tau_leaping/model.cu:477-520, functions.h:37-54
//   packed entry ................ tau_leaping/functions.h:126-146
// The reference draws from a per-channel hybrid Tausworthe state
// (d_rng.cuh:27-35); here every draw comes from a Philox4x32-10 counter
// stream keyed per realization, so draws never depend on launch shape.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tl {

constexpr int kNc = 5;                 // criticality threshold
constexpr double kSsaFactor = 10.0;    // direct branch when tau1 < 10/a0
constexpr int kSsaSteps = 100;         // direct events per direct step
constexpr int kPoissonSwitch = 80;     // inversion below, normal approximation at or above
constexpr int kInvertCap = 300;        // inversion searches up to k = 300
constexpr int kMaxRetries = 32;        // rejected leaps in a row before giving up
constexpr float kFltMax = 3.402823466e+38f;
constexpr int32_t kTileGrain = 4096;  // channels per tile_ptr column; block-group tau tiles are 1 or 2 grains

constexpr uint32_t kKindPoisson = 0u;  // ctr[3] low bits: per-channel firing stream (draws 2, 3, ...)
constexpr uint32_t kKindLeap = 1u;     //                  leap scalars (r1, r2)
constexpr uint32_t kKindSsa = 2u;      //                  direct-event scalars
constexpr uint32_t kKindHead = 3u;     //                  first draw of channels 4b..4b+3 (block b)

// Status codes (mirror tl_api.h)
constexpr int kOk = 0;
constexpr int kRetryExhausted = 7;
constexpr int kIterCap = 8;
constexpr int kStalled = 12;

// Model constants in device memory, shared by every realization.
struct DevModel {
    int32_t N, M, nnz;
    const double *c64;       // [M] rates
    const float *c32;        // [M] rates rounded once to float
    const int8_t *order;     // [M] 1, 2, 3
    const int32_t *s0;       // [M]
    const int32_t *s1;       // [M]
    const int32_t *s2;       // [M] third reactant of Sa+Sb+Sc channels; null unless some order is 0, 4 or 5
    const int32_t *row_ptr;  // [N+1] species-major
    const int32_t *csr;      // [nnz] (channel << 4) | (v + 3)
    const int32_t *col_ptr;  // [M+1] channel-major
    const int32_t *csc;      // [nnz] (species << 4) | (v + 3)
    const int8_t *g_order;   // [N] order[i] for i < M else 0
    const int32_t *x0;       // [N]
    const int2 *rs;          // [M] (s0 | order << 29, s1): one load per channel
    const int32_t *dep_ptr;  // [N+1] species -> channels with it as a reactant
    const int32_t *dep;      // [dep_ptr[N]]
    const int4 *colq;        // [M] column entries padded with v = 0, or null (a column longer than 4)
    const int32_t *aff_ptr;  // [M+1] channel -> channels whose propensity it can change, or null
    const int32_t *aff;      // [aff_ptr[M]]
    const int2 *aff_rs;      // [aff_ptr[M]] rs of each listed channel (loads in parallel with aff)
    const double *aff_c64;   // [aff_ptr[M]] its rate
    const float *aff_c32;
    const int32_t *tile_ptr; // [n_grains + 1][N] CSR offset of each species row at channel g * kTileGrain
    const int32_t *row_order; // [N] species by CSR row length, longest first (the row gather's schedule), or
                              // null when every row has the same length
};

// ---- Philox4x32-10 ----
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 product per multiplier (v_mad_u64_u32)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k.x, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k.y,
                       (uint32_t)p0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

// key = (seed_lo, seed_hi + realization id)
__device__ __forceinline__ uint2 rng_key(uint64_t seed, int64_t rid) {
    return make_uint2((uint32_t)seed, (uint32_t)(seed >> 32) + (uint32_t)rid);
}

__device__ __forceinline__ double unif64(uint32_t b) {
    return ((double)b + 0.5) * 2.3283064365386962890625e-10;
}
__device__ __forceinline__ float unif32(uint32_t b) {
    return ((float)(b >> 9) + 0.5f) * 1.1920928955078125e-07f;
}

template <class R> struct Real;
template <> struct Real<double> {
    static __device__ __forceinline__ double unif(uint32_t b) { return unif64(b); }
    static __device__ __forceinline__ double exp_(double v) { return exp(v); }
    static __device__ __forceinline__ double log_(double v) { return log(v); }
    static __device__ __forceinline__ double sqrt_(double v) { return sqrt(v); }
    static __device__ __forceinline__ double fabs_(double v) { return fabs(v); }
    static __device__ __forceinline__ const double *rates(const DevModel &m) { return m.c64; }
};
// FP32 exp and log are evaluated in FP64 and rounded once: the device's and
// the host library's FP64 functions differ by at most an ulp of the double,
// which the rounding to float hides except within an ulp of a float rounding
// boundary, so the oracle's draws and the GPU's agree (FP32 library
// functions differ between the two by an ulp of the float).
template <> struct Real<float> {
    static __device__ __forceinline__ float unif(uint32_t b) { return unif32(b); }
    static __device__ __forceinline__ float exp_(float v) { return (float)exp((double)v); }
    static __device__ __forceinline__ float log_(float v) { return (float)log((double)v); }
    static __device__ __forceinline__ float sqrt_(float v) { return sqrtf(v); }
    static __device__ __forceinline__ float fabs_(float v) { return fabsf(v); }
    static __device__ __forceinline__ const float *rates(const DevModel &m) { return m.c32; }
};

// Propensity of one channel of order 1-3; operation order (c*x0)*x1 and
// ((c*x0)*(x0-1))/2 as in IsCriticalAndPropensity.
// Written without branches (channels of different orders share a wave); the
// roundings are those of the three formulas: c*x0, (c*x0)*f, then an exact
// halving for order 3.
template <class R>
__device__ __forceinline__ R propensity(int order, R c, int32_t x0, int32_t x1, bool &crit) {
    const bool o2 = order == 2, o3 = order == 3;
    crit = o3 ? (x0 * 2 < kNc) : ((x0 < kNc) || (o2 && x1 < kNc));
    const R t = c * (R)x0;
    const R f = (R)(o2 ? x1 : x0 - 1);
    const R t2 = (o2 || o3) ? t * f : t;
    return o3 ? t2 * (R)0.5 : t2;
}

// The forms beyond the reference's (models with m.s2 != null): order 0, a
// source, a = c and never critical; channels with three reactant molecules,
// order 4 Sa+Sb+Sc ((c*x0)*x1)*x2, order 5 2Sa+Sb (((c*x0)*(x0-1))*0.5)*x1
// (x0 the species taken twice) and order 6 3S (((c*x0)*(x0-1))*(x0-2))/6,
// critical when a reactant's coefficient times its count is below nc (the
// reference's 2S test, 2*x0 < nc, for every coefficient).  Orders 1-3 as
// propensity().
template <class R>
__device__ __forceinline__ R propensity_ext(int order, R c, int32_t x0, int32_t x1, int32_t x2, bool &crit) {
    if (order == 0) {
        crit = false;
        return c;
    }
    if (order == 4) {
        crit = (x0 < kNc) || (x1 < kNc) || (x2 < kNc);
        return ((c * (R)x0) * (R)x1) * (R)x2;
    }
    if (order == 5) {
        crit = (x0 * 2 < kNc) || (x1 < kNc);
        return (((c * (R)x0) * (R)(x0 - 1)) * (R)0.5) * (R)x1;
    }
    if (order == 6) {
        crit = x0 * 3 < kNc;
        return (((c * (R)x0) * (R)(x0 - 1)) * (R)(x0 - 2)) / (R)6;
    }
    return propensity<R>(order, c, x0, x1, crit);
}

// g of the tau bound for the order GetG reads (the channel-indexed order):
// Cao-Gillespie-Petzold's highest-order forms, with 3 for Sa+Sb+Sc,
// (3/2)(2 + 1/(x-1)) for 2Sa+Sb and 3 + 1/(x-1) + 2/(x-2) for 3S.
template <class R>
__device__ __forceinline__ R getg(int order, int32_t x) {
    if (order == 0 || order == 1) return (R)1;  // order 0: a source, the first-order bound
    if (order == 2) return (R)2;
    if (order == 3) return (R)2 + (R)1 / ((R)x - (R)1);
    if (order == 4) return (R)3;
    if (order == 5) return (R)1.5 * ((R)2 + (R)1 / ((R)x - (R)1));
    if (order == 6) return ((R)3 + (R)1 / ((R)x - (R)1)) + (R)2 / ((R)x - (R)2);
    return (R)0;
}

// tau_i from the row sums (mu, sigma^2) of species i with population xi.
// Sets which = 0 when the mean bound is the smaller one, 1 otherwise.
// In FP64, eps * x / g is divided only for the orders whose g is not 1 or 2
// (the quotient by 1 or 2 is exact, so it equals the product by 1 or 0.5):
// 2 % less time on the toy network.  The FP32 kernels keep the division (the
// changed code cost them 5 %).
template <class R>
__device__ __forceinline__ R tau_species(R eps, int32_t xi, int order_i, R mu, R sig, int &which) {
    R q;
    if constexpr (sizeof(R) == 8) {
        const R t = eps * (R)xi;
        q = order_i == 2 ? t * (R)0.5 : t;
        if (order_i > 2) q = t / getg<R>(order_i, xi);
    } else {
        q = eps * (R)xi / getg<R>(order_i, xi);
    }
    const R tmp = q > (R)1 ? q : (R)1;
    const R t1 = tmp / Real<R>::fabs_(mu);
    const R t2 = tmp * tmp / Real<R>::fabs_(sig);
    which = (t1 <= t2) ? 0 : 1;
    return t1 < t2 ? t1 : t2;
}

// Head draw of channel j at an epoch: output (j & 3) of the Philox block
// shared by channels 4b..4b+3, b = j >> 2.
__device__ __forceinline__ uint4 head_block(uint2 key, uint32_t b, uint32_t epoch, uint32_t ctr3_model) {
    return philox4x32_10(make_uint4(b, 0u, epoch, ctr3_model | kKindHead), key);
}
__device__ __forceinline__ uint32_t quad_pick(uint4 o, int s) {
    return s == 0 ? o.x : (s == 1 ? o.y : (s == 2 ? o.z : o.w));
}

// Exact Poisson sample for mean >= 10: PTRS, the transformed rejection with
// squeeze of W. Hoermann, "The transformed rejection method for generating
// Poisson random variables", Insurance: Mathematics and Economics 12 (1993)
// 39-45, in FP64 over FP64 uniforms of the channel's draws.  Attempt i takes
// draws 2i+1 and 2i+2 (draw 1 = head; draw n >= 2 as in poisson_draw):
// U = u - 1/2 from the first, V from the second.  After 64 rejected attempts
// (probability below 1e-60) it returns floor(mean), so the loop always ends.
// ln(k!) for integer k >= 0: the exact product below 10, else Stirling's
// series for ln Gamma(k + 1) to x^-7 (absolute error below 4e-13 from k = 10).
__device__ __forceinline__ double ln_factorial(double k) {
    if (k < 10.0) {
        double f = 1.0;
        for (double i = 2.0; i <= k; i += 1.0) f *= i;
        return log(f);
    }
    const double x = k + 1.0, r = 1.0 / x, r2 = r * r;
    return (x - 0.5) * log(x) - x + 0.91893853320467274178 +
           r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0 - r2 / 1680.0)));
}

__device__ __forceinline__ int32_t poisson_ptrs(double mean, uint32_t head, uint2 key, uint32_t channel,
                                                uint32_t epoch, uint32_t ctr3) {
    // the constants are recomputed per attempt (cheap on this rare path) so
    // that little stays live across the loop inside the firing pass
    uint32_t blk = 0;
    uint4 o = make_uint4(0u, 0u, 0u, 0u);
    int pos = 4;
    uint32_t ub = head;
    for (int attempt = 0; attempt < 64; ++attempt) {
        if (attempt > 0) {
            if (pos == 4) {
                o = philox4x32_10(make_uint4(blk++, channel, epoch, ctr3), key);
                pos = 0;
            }
            ub = quad_pick(o, pos++);
        }
        if (pos == 4) {
            o = philox4x32_10(make_uint4(blk++, channel, epoch, ctr3), key);
            pos = 0;
        }
        const uint32_t vb = quad_pick(o, pos++);
        const double b = 0.931 + 2.53 * sqrt(mean);
        const double a = -0.059 + 0.02483 * b;
        const double U = unif64(ub) - 0.5, V = unif64(vb);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + mean + 0.43);
        if (us >= 0.07 && V <= 0.9277 - 3.6224 / (b - 2.0)) return (int32_t)k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        const double lia = log(1.1239 + 1.1328 / (b - 3.4));
        if (log(V) + lia - log(a / (us * us) + b) <= -mean + k * log(mean) - ln_factorial(k)) return (int32_t)k;
    }
    return (int32_t)mean;
}

// Sampler flags (tl_sim_params.poisson): bit 0 = PTRS above mean 80 instead
// of the reference's normal approximation; bit 1 = PTRS from mean 10 instead
// of the inversion up to 80.
constexpr int kPoissonExact = 1;
constexpr int kPoissonFast = 2;
constexpr float kPtrsFrom = 10.f;  // PTRS from this mean (the method's range) with kPoissonFast

// Mean below which inversion draws a non-critical channel's count.
template <class R>
__device__ __forceinline__ R invert_below(bool fast) {
    return fast ? (R)kPtrsFrom : (R)kPoissonSwitch;
}

// Poisson count by inversion of the channel's first draw u: the smallest k
// with u < F(k) = sum_{i <= k} P(i), P(0) = exp(-mean), P(i) = P(i-1) mean / i
// (sequential-search inversion, L. Devroye, Non-Uniform Random Variate
// Generation, Springer 1986, sec. X.3): the Poisson law from one uniform,
// where the reference's Knuth product takes mean + 1.  The search stops at
// kInvertCap (the mass beyond is below 1e-60 for means below 80).  It returns
// 0 without evaluating exp when u < 1 - mean - margin, a bound below the
// rounded exp(-mean) (the result is the one the full search returns).
//
// In FP64 the step P(i) = (P(i-1) * mean) / i, a division by a small integer,
// takes the reciprocal from a table: with y = RN(1/i), q = RN(x * y) and the
// exactly computed remainder x - i * q (one FMA), q + (x - i * q) * y rounded
// once (a second FMA) is the correctly rounded quotient x / i (P. Markstein,
// IBM J. Res. Dev. 34 (1990) 111-119: y within half an ulp of 1/i and q
// within an ulp of x/i), so the search keeps the IEEE division's values at
// about half its instructions (toy network, FP64: 5 % less kernel time).  The
// loop counter is the same in every active lane, so the table entry is a
// scalar load, issued one step ahead.  FP32 keeps the division (its table
// load, not hidden by the cheaper FP32 step, cost as much as it saved).
struct RecipTable {
    double d[kInvertCap + 2];
};
constexpr RecipTable make_recip_table() {
    RecipTable t{};
    for (int i = 1; i <= kInvertCap + 1; ++i) t.d[i] = 1.0 / (double)i;  // IEEE division at compile time
    return t;
}
static __constant__ RecipTable kRecip = make_recip_table();

// The early exit of poisson_invert: mean < 1 and u < 1 - mean - margin, so
// the count is 0 without exp or search.
template <class R>
__device__ __forceinline__ bool poisson_zero_by_bound(R mean, uint32_t head) {
    const R u = Real<R>::unif(head);
    const R margin = sizeof(R) == 8 ? (R)9.094947017729282e-13 : (R)9.5367431640625e-07;  // 2^-40 / 2^-20
    return mean < (R)1 && u < (R)1 - mean - margin;
}

template <class R>
__device__ __forceinline__ int32_t poisson_invert(R mean, uint32_t head) {
    if (poisson_zero_by_bound<R>(mean, head)) return 0;
    const R u = Real<R>::unif(head);
    R p = Real<R>::exp_(-mean);
    R F = p;
    int32_t k = 0;
    if constexpr (sizeof(R) == 8) {
        double y = kRecip.d[1];  // RN(1/(k + 1))
        while (!(u < F) && k < kInvertCap) {
            ++k;
            const double yk = y;
            y = kRecip.d[k + 1];
            const double x = p * mean;
            const double q = x * yk;
            p = fma(fma(-(double)k, q, x), yk, q);  // = x / k, correctly rounded
            F += p;
        }
    } else {
        while (!(u < F) && k < kInvertCap) {
            ++k;
            p = (p * mean) / (R)k;
            F += p;
        }
    }
    return k;
}

// The draw of a channel whose mean is at least invert_below: PTRS below 80
// (with `fast`; and with `exact` above 80), else |mean + sqrt(mean) * z|
// truncated, z from the polar method on the pairs (head, t0), (t1, t2),
// (t3, t4), ... of the channel's draws.
template <class R>
__device__ __forceinline__ int32_t poisson_heavy(R mean, uint32_t head, uint2 key, uint32_t channel, uint32_t epoch,
                                                 uint32_t ctr3, bool exact) {
    if (mean < (R)kPoissonSwitch || exact) return poisson_ptrs((double)mean, head, key, channel, epoch, ctr3);
    uint32_t blk = 0;
    R u = 2 * Real<R>::unif(head) - 1;
    uint4 o = philox4x32_10(make_uint4(blk++, channel, epoch, ctr3), key);
    R v = 2 * Real<R>::unif(o.x) - 1;
    R s = u * u + v * v;
    if (s >= 1) {
        for (;;) {
            u = 2 * Real<R>::unif(o.y) - 1;
            v = 2 * Real<R>::unif(o.z) - 1;
            s = u * u + v * v;
            if (s < 1) break;
            const uint32_t w = o.w;
            o = philox4x32_10(make_uint4(blk++, channel, epoch, ctr3), key);
            u = 2 * Real<R>::unif(w) - 1;
            v = 2 * Real<R>::unif(o.x) - 1;
            s = u * u + v * v;
            if (s < 1) break;
        }
    }
    const R z = u * Real<R>::sqrt_(-2 * Real<R>::log_(s) / s);
    const R y = mean + Real<R>::sqrt_(mean) * z;
    return (int32_t)Real<R>::fabs_(y);
}

// Firing count of a non-critical channel: inversion below 80 (below 10 with
// `fast`, PTRS from there to 80), then the reference's normal approximation,
// or PTRS with `exact`.
template <class R>
__device__ __forceinline__ int32_t poisson_draw(R mean, uint32_t head, uint2 key, uint32_t channel,
                                                uint32_t epoch, uint32_t ctr3, bool exact, bool fast = false) {
    if (mean < invert_below<R>(fast)) return poisson_invert<R>(mean, head);
    return poisson_heavy<R>(mean, head, key, channel, epoch, ctr3, exact);
}

// Leap / direct-event scalars r1 (pick) and r2 (waiting time), FP64.
__device__ __forceinline__ void step_uniforms(uint2 key, uint32_t epoch, uint32_t ctr3_kind, double &r1,
                                              double &r2) {
    const uint4 o = philox4x32_10(make_uint4(0u, 0u, epoch, ctr3_kind), key);
    r1 = unif64(o.x);
    r2 = unif64(o.y);
}

// Likelihood-ratio score term d_j * m_j of one channel's firing count k_j at
// mean m_j = a_j * tau, d_j = d log p(k_j) / d m_j: k - m for Poisson counts
// (and every channel with the exact sampler) and critical channels; for the
// normal approximation at mean >= 80 the density of N(m, m) at k + 1/2,
// y - m + ((y - m)^2 / m - 1) / 2 with y = k + 1/2.
template <class R>
__device__ __forceinline__ double score_term(bool cr, R aj, R tr, double tau, int32_t kj, bool exact) {
    const double mj = (double)aj * tau;
    if (!exact && !cr && aj * tr >= (R)kPoissonSwitch) {
        const double d = ((double)kj + 0.5) - mj;
        return d + 0.5 * (d * d / mj - 1.0);
    }
    return (double)kj - mj;
}

// Packed entry (index << 4) | (v + 3), v in [-3, 12].
__device__ __forceinline__ int32_t entry_index(int32_t p) { return p >> 4; }
__device__ __forceinline__ int32_t entry_value(int32_t p) { return (p & 15) - 3; }

}  // namespace tl
