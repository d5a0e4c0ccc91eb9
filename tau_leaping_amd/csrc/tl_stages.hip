// tl_stages.hip -- stage kernels of the plain engine form.
//
// State is structure-of-arrays with realizations contiguous: x[i*R + r],
// a[j*R + r], so a wavefront of 64 realizations touching the same species or
// channel makes one coalesced access.  One host iteration of the plain form
// launches: propensities -> tau rows -> tau min -> select (leap scalars or a
// direct-method block) -> fire -> update -> commit.
//
// Reference stages (file:line): CalcPropensities model.cu:213-247; CompTau
// CSRMatrixPacked.cu:423-462; GetReaction2Fire model.cu:555-598 and the leap
// branch model.cu:346-390; SetNumberOfFiring model.cu:538-549; FireReactions
// CSRMatrixPacked.cu:464-503; DirectMethodStep model.cu:427-451.
#include "tl_internal.hpp"

namespace tl {

namespace {

constexpr int kT = 256;  // threads per block for the realization axis

inline unsigned blocks_for(int64_t n, int t = kT) { return (unsigned)((n + t - 1) / t); }

// Two adjacent realizations per thread (even R) load and store 2-wide.
template <class R>
struct Pair;
template <>
struct Pair<double> {
    using T = double2;
};
template <>
struct Pair<float> {
    using T = float2;
};

// ---------------------------------------------------------- propensities --
// a_j of channel j for realization r from the structure-of-arrays populations
// (extended formulas when the model has three-reactant channels).
template <class R>
__device__ __forceinline__ R soa_prop(const DevModel &m, int32_t j, R cj, const int32_t *x, int64_t Rn, int64_t r,
                                      bool &cr) {
    const int order = m.order[j];
    const int32_t x0 = x[(int64_t)m.s0[j] * Rn + r], x1 = x[(int64_t)m.s1[j] * Rn + r];
    if (m.s2) return propensity_ext<R>(order, cj, x0, x1, order == 4 ? x[(int64_t)m.s2[j] * Rn + r] : 1, cr);
    return propensity<R>(order, cj, x0, x1, cr);
}

template <class R>
__global__ __launch_bounds__(kT) void k_prop(DevModel m, int64_t Rn, int32_t nchunk, const int32_t *__restrict__ x,
                                             R *__restrict__ a, uint8_t *__restrict__ crit, double *__restrict__ a0,
                                             double *__restrict__ a0c, double *__restrict__ part,
                                             double *__restrict__ partc, const uint8_t *__restrict__ active) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (active && !active[r]) return;
    const int32_t c = blockIdx.y;
    const int32_t ch = (m.M + nchunk - 1) / nchunk;
    const int32_t j0 = c * ch, j1 = min(m.M, j0 + ch);
    const R *rc = Real<R>::rates(m);
    double s = 0.0, sc = 0.0;
    for (int32_t j = j0; j < j1; ++j) {
        bool cr;
        const R aj = soa_prop<R>(m, j, rc[j], x, Rn, r, cr);
        a[(int64_t)j * Rn + r] = aj;
        crit[(int64_t)j * Rn + r] = cr;
        s += (double)aj;
        if (cr) sc += (double)aj;
    }
    if (nchunk == 1) {
        a0[r] = s;
        a0c[r] = sc;
    } else {
        part[(int64_t)c * Rn + r] = s;
        partc[(int64_t)c * Rn + r] = sc;
    }
}

// The same for two adjacent realizations per thread (even R): 2-wide loads
// of populations and stores of propensities and flags; each realization's
// arithmetic is the single kernel's.
__device__ __forceinline__ int2 ld2(const int32_t *p) { return *reinterpret_cast<const int2 *>(p); }

template <class R>
__global__ __launch_bounds__(kT) void k_prop2(DevModel m, int64_t Rn, int32_t nchunk, const int32_t *__restrict__ x,
                                              R *__restrict__ a, uint8_t *__restrict__ crit, double *__restrict__ a0,
                                              double *__restrict__ a0c, double *__restrict__ part,
                                              double *__restrict__ partc, const uint8_t *__restrict__ active) {
    using V2 = typename Pair<R>::T;
    const int64_t r = 2 * ((int64_t)blockIdx.x * kT + threadIdx.x);
    if (r >= Rn) return;
    bool on0 = true, on1 = true;
    if (active) {
        on0 = active[r] != 0;
        on1 = active[r + 1] != 0;
        if (!on0 && !on1) return;
    }
    const int32_t c = blockIdx.y;
    const int32_t ch = (m.M + nchunk - 1) / nchunk;
    const int32_t j0 = c * ch, j1 = min(m.M, j0 + ch);
    const R *rc = Real<R>::rates(m);
    double s0 = 0.0, sc0 = 0.0, s1 = 0.0, sc1 = 0.0;
    for (int32_t j = j0; j < j1; ++j) {
        const int order = m.order[j];
        const R cj = rc[j];
        const int2 xa = ld2(x + (int64_t)m.s0[j] * Rn + r), xb = ld2(x + (int64_t)m.s1[j] * Rn + r);
        bool c0, c1;
        R p0, p1;
        if (m.s2) {
            const int2 xc = order == 4 ? ld2(x + (int64_t)m.s2[j] * Rn + r) : make_int2(1, 1);
            p0 = propensity_ext<R>(order, cj, xa.x, xb.x, xc.x, c0);
            p1 = propensity_ext<R>(order, cj, xa.y, xb.y, xc.y, c1);
        } else {
            p0 = propensity<R>(order, cj, xa.x, xb.x, c0);
            p1 = propensity<R>(order, cj, xa.y, xb.y, c1);
        }
        const int64_t o = (int64_t)j * Rn + r;
        if (on0 && on1) {
            V2 av;
            av.x = p0;
            av.y = p1;
            *reinterpret_cast<V2 *>(a + o) = av;
            *reinterpret_cast<uchar2 *>(crit + o) = make_uchar2(c0, c1);
        } else {
            if (on0) {
                a[o] = p0;
                crit[o] = c0;
            }
            if (on1) {
                a[o + 1] = p1;
                crit[o + 1] = c1;
            }
        }
        s0 += (double)p0;
        if (c0) sc0 += (double)p0;
        s1 += (double)p1;
        if (c1) sc1 += (double)p1;
    }
    double *ps = nchunk == 1 ? a0 : part + (int64_t)c * Rn;
    double *pc = nchunk == 1 ? a0c : partc + (int64_t)c * Rn;
    if (on0) {
        ps[r] = s0;
        pc[r] = sc0;
    }
    if (on1) {
        ps[r + 1] = s1;
        pc[r + 1] = sc1;
    }
}

__global__ __launch_bounds__(kT) void k_sum_parts(int64_t Rn, int32_t nchunk, const double *__restrict__ part,
                                                  const double *__restrict__ partc, double *__restrict__ a0,
                                                  double *__restrict__ a0c, const uint8_t *__restrict__ active) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (active && !active[r]) return;
    double s = 0.0, sc = 0.0;
    for (int32_t c = 0; c < nchunk; ++c) {
        s += part[(int64_t)c * Rn + r];
        sc += partc[(int64_t)c * Rn + r];
    }
    a0[r] = s;
    a0c[r] = sc;
}

// ------------------------------------------------------------------ tau --
// tau_i for species row i = blockIdx.y (+ gridDim.y strides).
template <class R>
__global__ __launch_bounds__(kT) void k_tau_rows(DevModel m, int64_t Rn, R eps, const int32_t *__restrict__ x,
                                                 const R *__restrict__ a, const uint8_t *__restrict__ crit,
                                                 R *__restrict__ tau_i, uint8_t *__restrict__ which_i,
                                                 const uint8_t *__restrict__ active) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (active && !active[r]) return;
    for (int32_t i = blockIdx.y; i < m.N; i += gridDim.y) {
        double mu = 0, sig = 0;  // FP64 sums, rounded once to R
        const int32_t e1 = m.row_ptr[i + 1];
        for (int32_t e = m.row_ptr[i]; e < e1; ++e) {
            const int32_t p = m.csr[e];
            const int32_t j = entry_index(p);
            const double v = (double)entry_value(p);
            const double tmp = v * (double)a[(int64_t)j * Rn + r] * (double)(!crit[(int64_t)j * Rn + r]);
            sig += v * tmp;
            mu += tmp;
        }
        R ti;
        int w = 0;
        if (i < m.M)
            ti = tau_species<R>(eps, x[(int64_t)i * Rn + r], m.g_order[i], (R)mu, (R)sig, w);
        else
            ti = (R)INFINITY;
        tau_i[(int64_t)i * Rn + r] = ti;
        which_i[(int64_t)i * Rn + r] = (uint8_t)w;
    }
}

// The species sums and the minimum over one contiguous chunk of rows in one
// pass (blockIdx.y = chunk), for many species: each thread keeps its chunk's
// first minimum and the bound it came from; k_tau_min_fin combines the chunks
// in order.  Nothing per species is written.
template <class R>
__global__ __launch_bounds__(kT) void k_tau_rows_part(DevModel m, int64_t Rn, R eps, int32_t nrc,
                                                      const int32_t *__restrict__ x, const R *__restrict__ a,
                                                      const uint8_t *__restrict__ crit, R *__restrict__ pbest,
                                                      int32_t *__restrict__ pidx, uint8_t *__restrict__ pwhich,
                                                      const uint8_t *__restrict__ active) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (active && !active[r]) return;
    const int32_t c = blockIdx.y;
    const int32_t ch = (m.N + nrc - 1) / nrc;
    const int32_t i0 = c * ch, i1 = min(m.N, i0 + ch);
    R best = (R)kFltMax;
    int32_t bi = -1;
    int bw = 0;
    for (int32_t i = i0; i < i1; ++i) {
        double mu = 0, sig = 0;
        const int32_t e1 = m.row_ptr[i + 1];
        for (int32_t e = m.row_ptr[i]; e < e1; ++e) {
            const int32_t p = m.csr[e];
            const int32_t j = entry_index(p);
            const double v = (double)entry_value(p);
            const double tmp = v * (double)a[(int64_t)j * Rn + r] * (double)(!crit[(int64_t)j * Rn + r]);
            sig += v * tmp;
            mu += tmp;
        }
        R ti;
        int w = 0;
        if (i < m.M)
            ti = tau_species<R>(eps, x[(int64_t)i * Rn + r], m.g_order[i], (R)mu, (R)sig, w);
        else
            ti = (R)INFINITY;
        if (ti < best) {
            best = ti;
            bi = i;
            bw = w;
        }
    }
    pbest[(int64_t)c * Rn + r] = best;
    pidx[(int64_t)c * Rn + r] = bi;
    pwhich[(int64_t)c * Rn + r] = (uint8_t)bw;
}

// The same for two adjacent realizations per thread (even R): 2-wide loads
// of propensities, flags and populations; each realization's arithmetic is
// the single kernel's.
template <class R>
__global__ __launch_bounds__(kT) void k_tau_rows_part2(DevModel m, int64_t Rn, R eps, int32_t nrc,
                                                       const int32_t *__restrict__ x, const R *__restrict__ a,
                                                       const uint8_t *__restrict__ crit, R *__restrict__ pbest,
                                                       int32_t *__restrict__ pidx, uint8_t *__restrict__ pwhich,
                                                       const uint8_t *__restrict__ active) {
    using V2 = typename Pair<R>::T;
    const int64_t r = 2 * ((int64_t)blockIdx.x * kT + threadIdx.x);
    if (r >= Rn) return;
    bool on0 = true, on1 = true;
    if (active) {
        on0 = active[r] != 0;
        on1 = active[r + 1] != 0;
        if (!on0 && !on1) return;
    }
    const int32_t c = blockIdx.y;
    const int32_t ch = (m.N + nrc - 1) / nrc;
    const int32_t i0 = c * ch, i1 = min(m.N, i0 + ch);
    R best0 = (R)kFltMax, best1 = (R)kFltMax;
    int32_t bi0 = -1, bi1 = -1;
    int bw0 = 0, bw1 = 0;
    for (int32_t i = i0; i < i1; ++i) {
        double mu0 = 0, sig0 = 0, mu1 = 0, sig1 = 0;
        const int32_t e1 = m.row_ptr[i + 1];
        for (int32_t e = m.row_ptr[i]; e < e1; ++e) {
            const int32_t p = m.csr[e];
            const int64_t o = (int64_t)entry_index(p) * Rn + r;
            const double v = (double)entry_value(p);
            const V2 av = *reinterpret_cast<const V2 *>(a + o);
            const uchar2 cv = *reinterpret_cast<const uchar2 *>(crit + o);
            const double t0 = v * (double)av.x * (double)(!cv.x);
            const double t1 = v * (double)av.y * (double)(!cv.y);
            sig0 += v * t0;
            mu0 += t0;
            sig1 += v * t1;
            mu1 += t1;
        }
        R ti0, ti1;
        int w0 = 0, w1 = 0;
        if (i < m.M) {
            const int2 xx = *reinterpret_cast<const int2 *>(x + (int64_t)i * Rn + r);
            ti0 = tau_species<R>(eps, xx.x, m.g_order[i], (R)mu0, (R)sig0, w0);
            ti1 = tau_species<R>(eps, xx.y, m.g_order[i], (R)mu1, (R)sig1, w1);
        } else {
            ti0 = ti1 = (R)INFINITY;
        }
        if (ti0 < best0) {
            best0 = ti0;
            bi0 = i;
            bw0 = w0;
        }
        if (ti1 < best1) {
            best1 = ti1;
            bi1 = i;
            bw1 = w1;
        }
    }
    const int64_t o = (int64_t)c * Rn + r;
    if (on0) {
        pbest[o] = best0;
        pidx[o] = bi0;
        pwhich[o] = (uint8_t)bw0;
    }
    if (on1) {
        pbest[o + 1] = best1;
        pidx[o + 1] = bi1;
        pwhich[o + 1] = (uint8_t)bw1;
    }
}

template <class R>
__global__ __launch_bounds__(kT) void k_tau_min(int32_t N, int64_t Rn, const R *__restrict__ tau_i,
                                                const uint8_t *__restrict__ which_i, double *__restrict__ tau1,
                                                int32_t *__restrict__ argmin, int32_t *__restrict__ which,
                                                const uint8_t *__restrict__ active) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (active && !active[r]) return;
    R best = (R)kFltMax;
    int32_t bi = -1, bw = 0;
    for (int32_t i = 0; i < N; ++i) {
        const R ti = tau_i[(int64_t)i * Rn + r];
        if (ti < best) {
            best = ti;
            bi = i;
            bw = which_i[(int64_t)i * Rn + r];
        }
    }
    tau1[r] = (double)best;
    if (argmin) argmin[r] = bi;
    if (which) which[r] = bw;
}

// The same minimum in two passes for many species: chunk c of the rows
// (blockIdx.y) writes its first minimum, then one thread per realization
// combines the chunks in order.  Strict comparisons in ascending row order
// give the sequential result exactly.
template <class R>
__global__ __launch_bounds__(kT) void k_tau_min_part(int32_t N, int64_t Rn, int32_t nrc, const R *__restrict__ tau_i,
                                                     R *__restrict__ pbest, int32_t *__restrict__ pidx,
                                                     const uint8_t *__restrict__ active) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (active && !active[r]) return;
    const int32_t c = blockIdx.y;
    const int32_t ch = (N + nrc - 1) / nrc;
    const int32_t i0 = c * ch, i1 = min(N, i0 + ch);
    R best = (R)kFltMax;
    int32_t bi = -1;
    for (int32_t i = i0; i < i1; ++i) {
        const R ti = tau_i[(int64_t)i * Rn + r];
        if (ti < best) {
            best = ti;
            bi = i;
        }
    }
    pbest[(int64_t)c * Rn + r] = best;
    pidx[(int64_t)c * Rn + r] = bi;
}

template <class R>
__global__ __launch_bounds__(kT) void k_tau_min_fin(int64_t Rn, int32_t nrc, const R *__restrict__ pbest,
                                                    const int32_t *__restrict__ pidx,
                                                    const uint8_t *__restrict__ which_i,
                                                    const uint8_t *__restrict__ pwhich, double *__restrict__ tau1,
                                                    int32_t *__restrict__ argmin, int32_t *__restrict__ which,
                                                    const uint8_t *__restrict__ active) {
    // which bound won: which_i[winner] after k_tau_min_part, the chunk's
    // pwhich after k_tau_rows_part
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (active && !active[r]) return;
    R best = (R)kFltMax;
    int32_t bi = -1, bc = -1;
    for (int32_t c = 0; c < nrc; ++c) {
        const R v = pbest[(int64_t)c * Rn + r];
        if (v < best) {
            best = v;
            bi = pidx[(int64_t)c * Rn + r];
            bc = c;
        }
    }
    tau1[r] = (double)best;
    if (argmin) argmin[r] = bi;
    if (which)
        which[r] = bi < 0 ? 0 : pwhich ? pwhich[(int64_t)bc * Rn + r] : which_i[(int64_t)bi * Rn + r];
}

// --------------------------------------------------------------- select --
// Channel and chunk walks load kWalk entries at once (one thread per
// realization: the walk is bound by load latency), then scan them in order.
constexpr int kWalk = 8;

template <class R>
__device__ void select_leap(const DevModel &m, uint2 key, uint32_t c3, uint32_t epoch, const R *a, const uint8_t *crit,
                            int64_t Rn, int64_t r, double a0c, double tau1, double t_rem, double &tau, int32_t &jc_out,
                            uint8_t &clamped, int32_t nchunk = 1, const double *partc = nullptr) {
    double r1, r2;
    step_uniforms(key, epoch, c3 | kKindLeap, r1, r2);
    int32_t jc = m.M;
    double tau2 = (double)kFltMax;
    // the critical channel is searched for only when its waiting time t2 is
    // below tau1: otherwise tau = tau1 and nothing fires whatever it is
    const double t2 = a0c > 0.0 ? -log(r2) / a0c : (double)kFltMax;
    if (a0c > 0.0 && t2 < tau1) {
        const double target = a0c * r1;
        double cum = 0.0;
        int32_t jstart = 0;
        if (partc && nchunk > 1) {
            // skip whole channel chunks by their critical sums (the propensity
            // pass's partials), then walk channels from the chunk reached
            // (eight loads in flight per batch, then the same sequential scan)
            const int32_t ch = (m.M + nchunk - 1) / nchunk;
            bool stop = false;
            for (int32_t c0 = 0; c0 + 1 < nchunk && !stop; c0 += kWalk) {
                double pv[kWalk];
#pragma unroll
                for (int u = 0; u < kWalk; ++u) pv[u] = c0 + u + 1 < nchunk ? partc[(int64_t)(c0 + u) * Rn + r] : 0.0;
#pragma unroll
                for (int u = 0; u < kWalk; ++u) {
                    if (c0 + u + 1 >= nchunk || cum + pv[u] > target) {
                        stop = true;
                        break;
                    }
                    cum += pv[u];
                    jstart += ch;
                }
            }
            jstart = min(jstart, m.M);
        }
        for (int32_t j0 = jstart; j0 < m.M && jc == m.M; j0 += kWalk) {
            R av[kWalk];
            uint8_t cv[kWalk];
#pragma unroll
            for (int u = 0; u < kWalk; ++u) {
                const bool in = j0 + u < m.M;
                cv[u] = in ? crit[(int64_t)(j0 + u) * Rn + r] : (uint8_t)0;
                av[u] = in ? a[(int64_t)(j0 + u) * Rn + r] : (R)0;
            }
#pragma unroll
            for (int u = 0; u < kWalk; ++u) {
                if (!cv[u]) continue;
                cum += (double)av[u];
                if (cum > target) {
                    jc = j0 + u;
                    break;
                }
            }
        }
        if (jc != m.M) tau2 = t2;
    }
    tau = tau1 < tau2 ? tau1 : tau2;
    bool fire = (jc != m.M) && (tau2 < tau1);
    clamped = 0;
    if (tau >= t_rem) {
        tau = t_rem;
        clamped = 1;
        fire = false;
    }
    jc_out = fire ? jc : -1;
}

template <class R>
__global__ __launch_bounds__(kT) void k_select(DevModel m, int64_t Rn, uint64_t seed, uint32_t model_id, int64_t rbase,
                                               const uint32_t *__restrict__ epoch, const R *__restrict__ a,
                                               const uint8_t *__restrict__ crit, const double *__restrict__ a0c,
                                               const double *__restrict__ tau1, const double *__restrict__ t_rem,
                                               double *__restrict__ tau, int32_t *__restrict__ jc,
                                               uint8_t *__restrict__ clamped) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    select_leap<R>(m, rng_key(seed, rbase + r), model_id << 2, epoch[r], a, crit, Rn, r, a0c[r], tau1[r], t_rem[r],
                   tau[r], jc[r], clamped[r]);
}

// ----------------------------------------------------------------- fire --
// One block covers kT realizations of channels 4q..4q+3: blockIdx.x = q*rt + tile.
// Registers for five waves per SIMD (96 VGPRs, 12 bytes of spills per lane):
// the pass is bound by its loads, which the extra waves keep in flight.  Whole
// 1e6-channel runs: five waves 6.79 s, six (92 bytes of spills) 6.91 s; on an
// earlier build four waves without spills were 2 % slower than six and eight
// (172 bytes) 29 % slower.
template <class R>
__global__ __launch_bounds__(kT, 5) void k_fire(DevModel m, int64_t Rn, uint64_t seed, uint32_t model_id, int64_t rbase,
                                                const uint32_t *__restrict__ epoch, const R *__restrict__ a,
                                                const uint8_t *__restrict__ crit, const double *__restrict__ tau,
                                                const int32_t *__restrict__ jc, int32_t *__restrict__ k,
                                                const uint8_t *__restrict__ mode, int pflags) {
    // a thread takes channels 4q..4q+3 of one realization: their first draws
    // share one Philox block (head_block), computed once while the channels'
    // propensities and criticality flags load.  blockIdx.x = realization
    // tile, q = blockIdx.y + gridDim.y * blockIdx.z (grid rows hold at most
    // 65535 quads).
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    const int32_t q = (int32_t)(blockIdx.y + gridDim.y * blockIdx.z);
    if (r >= Rn || 4 * q >= m.M) return;
    if (mode && mode[r] != 1) return;
    const int nq = min(4, m.M - 4 * q);
    R av[4];
    uint8_t cv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t o = (int64_t)(4 * q + i) * Rn + r;
        av[i] = i < nq ? a[o] : (R)0;
        cv[i] = i < nq ? crit[o] : (uint8_t)1;
    }
    const uint2 key = rng_key(seed, rbase + r);
    const uint32_t ep = epoch[r];
    const int32_t jcr = jc[r];
    const R tr = (R)tau[r];
    const uint4 h = head_block(key, (uint32_t)q, ep, model_id << 2);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= nq) break;
        const int32_t j = 4 * q + i;
        k[(int64_t)j * Rn + r] = cv[i] ? ((jcr == j) ? 1 : 0)
                                       : poisson_draw<R>(av[i] * tr, quad_pick(h, i), key, (uint32_t)j, ep,
                                                         (model_id << 2) | kKindPoisson,
                                                         (pflags & kPoissonExact) != 0, (pflags & kPoissonFast) != 0);
    }
}

// --------------------------------------------------------------- update --
// x_out_i = x_i + sum over row i of v * k_j (all N species rows).
__global__ __launch_bounds__(kT) void k_update(DevModel m, int64_t Rn, const int32_t *__restrict__ x,
                                               const int32_t *__restrict__ k, int32_t *__restrict__ x_out,
                                               uint8_t *__restrict__ neg, const uint8_t *__restrict__ mode) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (mode && mode[r] != 1) return;
    for (int32_t i = blockIdx.y; i < m.N; i += gridDim.y) {
        int32_t s = 0;
        const int32_t e1 = m.row_ptr[i + 1];
        for (int32_t e = m.row_ptr[i]; e < e1; ++e) {
            const int32_t p = m.csr[e];
            s += entry_value(p) * k[(int64_t)entry_index(p) * Rn + r];
        }
        const int32_t xo = x[(int64_t)i * Rn + r] + s;
        x_out[(int64_t)i * Rn + r] = xo;
        if (xo < 0) neg[r] = 1;
    }
}

// The same for two adjacent realizations per thread (even R): 8-byte loads
// and stores, 512 bytes per wavefront access.  A pair with one realization
// not leaping writes that one's x_out too (never read: the commit copies
// only leaping realizations) but flags only leaping ones.
__global__ __launch_bounds__(kT) void k_update2(DevModel m, int64_t Rn, const int32_t *__restrict__ x,
                                                const int32_t *__restrict__ k, int32_t *__restrict__ x_out,
                                                uint8_t *__restrict__ neg, const uint8_t *__restrict__ mode) {
    const int64_t r = 2 * ((int64_t)blockIdx.x * kT + threadIdx.x);
    if (r >= Rn) return;
    bool on0 = true, on1 = true;
    if (mode) {
        on0 = mode[r] == 1;
        on1 = mode[r + 1] == 1;
        if (!on0 && !on1) return;
    }
    bool n0 = false, n1 = false;
    for (int32_t i = blockIdx.y; i < m.N; i += gridDim.y) {
        int32_t s0 = 0, s1 = 0;
        const int32_t e1 = m.row_ptr[i + 1];
        for (int32_t e = m.row_ptr[i]; e < e1; ++e) {
            const int32_t p = m.csr[e];
            const int32_t v = entry_value(p);
            const int2 kk = *reinterpret_cast<const int2 *>(k + (int64_t)entry_index(p) * Rn + r);
            s0 += v * kk.x;
            s1 += v * kk.y;
        }
        const int2 xx = *reinterpret_cast<const int2 *>(x + (int64_t)i * Rn + r);
        const int2 xo = make_int2(xx.x + s0, xx.y + s1);
        *reinterpret_cast<int2 *>(x_out + (int64_t)i * Rn + r) = xo;
        n0 |= xo.x < 0;
        n1 |= xo.y < 0;
    }
    if (n0 && on0) neg[r] = 1;
    if (n1 && on1) neg[r + 1] = 1;
}

// ------------------------------------------------------------ direct SSA --
template <class R>
__device__ double prop_soa(const DevModel &m, const int32_t *x, R *a, uint8_t *crit, int64_t Rn, int64_t r,
                           double *a0c) {
    const R *rc = Real<R>::rates(m);
    double s = 0.0, sc = 0.0;
    for (int32_t j = 0; j < m.M; ++j) {
        bool cr;
        const R aj = soa_prop<R>(m, j, rc[j], x, Rn, r, cr);
        a[(int64_t)j * Rn + r] = aj;
        if (crit) crit[(int64_t)j * Rn + r] = cr;
        s += (double)aj;
        if (cr) sc += (double)aj;
    }
    if (a0c) *a0c = sc;
    return s;
}

// One direct-method event on column r (the fixed DirectMethodStep): returns
// false when the realization reached t_final instead.  With scores S (and
// dt/dc D), [M][R]: the event adds 1/c_jf for the fired channel and the
// exposure -a_j dt / c_j of every channel; the truncated last interval
// rem = t_final - t adds d/dc_j of -a0 rem, a0 D_j - a_j rem / c_j.
template <class R>
__device__ bool ssa_event_soa(const DevModel &m, uint2 key, uint32_t c3, uint32_t &epoch, int32_t *x, int64_t Rn,
                              int64_t r, double &t, double tf, const R *a, double a0, int32_t *fired = nullptr,
                              double *S = nullptr, const double *D = nullptr) {
    if (!(a0 > 0.0)) {
        t = tf;
        return false;
    }
    double r1, r2;
    step_uniforms(key, epoch, c3 | kKindSsa, r1, r2);
    epoch++;
    const double dt = -log(r2) / a0;
    const R *rc = Real<R>::rates(m);
    if (dt >= tf - t) {
        if (S) {
            const double rem = tf - t;
            for (int32_t j = 0; j < m.M; ++j) {
                const int64_t o = (int64_t)j * Rn + r;
                if (rc[j] != (R)0) S[o] += a0 * D[o] - (double)a[o] * rem / (double)rc[j];
            }
        }
        t = tf;
        return false;
    }
    const double target = a0 * r1;
    double cum = 0.0;
    int32_t jf = -1, last = -1;
    for (int32_t j0 = 0; j0 < m.M && jf < 0; j0 += kWalk) {
        R av[kWalk];
#pragma unroll
        for (int u = 0; u < kWalk; ++u) av[u] = j0 + u < m.M ? a[(int64_t)(j0 + u) * Rn + r] : (R)0;
#pragma unroll
        for (int u = 0; u < kWalk; ++u) {
            if (j0 + u >= m.M) break;
            const R aj = av[u];
            if (aj > 0) last = j0 + u;
            cum += (double)aj;
            if (cum > target) {
                jf = j0 + u;
                break;
            }
        }
    }
    if (jf < 0) jf = last;
    for (int32_t e = m.col_ptr[jf]; e < m.col_ptr[jf + 1]; ++e) {
        const int32_t p = m.csc[e];
        x[(int64_t)entry_index(p) * Rn + r] += entry_value(p);
    }
    if (S) {
        for (int32_t j = 0; j < m.M; ++j) {
            const int64_t o = (int64_t)j * Rn + r;
            if (rc[j] != (R)0) S[o] -= (double)a[o] * dt / (double)rc[j];
        }
        if (rc[jf] != (R)0) S[(int64_t)jf * Rn + r] += 1.0 / (double)rc[jf];
    }
    t += dt;
    if (fired) *fired = jf;
    return true;
}

// Propensities after channel jf fired: only the channels that read a species
// of jf's column change (affected-channel lists; a full pass without them),
// then a0 as the sequential FP64 sum over all channels, as prop_soa gives it.
template <class R>
__device__ double refresh_soa(const DevModel &m, const int32_t *x, R *a, int64_t Rn, int64_t r, int32_t jf) {
    if (!m.aff_ptr) return prop_soa<R>(m, x, a, nullptr, Rn, r, nullptr);
    const R *rc = Real<R>::rates(m);
    for (int32_t d = m.aff_ptr[jf]; d < m.aff_ptr[jf + 1]; ++d) {
        const int32_t k = m.aff[d];
        bool cr;
        a[(int64_t)k * Rn + r] = soa_prop<R>(m, k, rc[k], x, Rn, r, cr);
    }
    double s = 0.0;
    for (int32_t j = 0; j < m.M; ++j) s += (double)a[(int64_t)j * Rn + r];
    return s;
}

template <class R>
__global__ __launch_bounds__(kT) void k_ssa(DevModel m, int64_t Rn, uint64_t seed, uint32_t model_id, int64_t rbase,
                                            int32_t *__restrict__ x, double *__restrict__ t, uint32_t *__restrict__ epoch,
                                            double tf, int32_t n_events, int32_t *__restrict__ events,
                                            R *__restrict__ a) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    const uint2 key = rng_key(seed, rbase + r);
    double tr = t[r];
    uint32_t ep = epoch[r];
    int32_t n = 0;
    for (int32_t s = 0; s < n_events && tr < tf; ++s) {
        const double a0 = prop_soa<R>(m, x, a, nullptr, Rn, r, nullptr);
        if (!ssa_event_soa<R>(m, key, model_id << 2, ep, x, Rn, r, tr, tf, a, a0)) break;
        ++n;
    }
    t[r] = tr;
    epoch[r] = ep;
    events[r] = n;
}

// ------------------------------------------------------ plain bookkeeping --
__global__ __launch_bounds__(kT) void k_plain_init(DevModel m, int64_t Rn, int32_t *__restrict__ x, PlainState st) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    for (int32_t i = 0; i < m.N; ++i) x[(int64_t)i * Rn + r] = m.x0[i];
    st.t[r] = 0.0;
    st.override_tau1[r] = -1.0;
    st.tau1_used[r] = 0.0;
    st.tau[r] = 0.0;
    st.epoch[r] = 0;
    st.steps[r] = 0;
    st.ssa[r] = 0;
    st.rej[r] = 0;
    st.status[r] = kOk;
    st.retries[r] = 0;
    st.jc[r] = -1;
    st.active[r] = 1;
    st.mode[r] = 0;
    st.clamped[r] = 0;
    st.neg[r] = 0;
}

// Grid points of realization r passed by a step from (t_a, state xa) to
// (t_b, state xb): linear interpolation between the two states, as
// AllStat2StepStat does between logged states (statistics.cpp:15-64).
// xa(q), xb(q) give tracked species q.
template <class FA, class FB>
__device__ void grid_step(const PlainState &st, int64_t r, double tf, double t_a, double t_b, FA xa, FB xb) {
    const int32_t n = st.n_intervals, T = st.n_track;
    const double h = tf / (double)n;
    int32_t g = st.next_g[r];
    while (g < n) {
        const double tg = h * (double)g;
        if (!(tg < t_b)) break;
        const double w = (t_b > t_a) ? (tg - t_a) / (t_b - t_a) : 0.0;
        double *out = st.grid + ((size_t)r * (n + 1) + g) * T;
        for (int32_t q = 0; q < T; ++q) {
            const double va = (double)xa(q), vb = (double)xb(q);
            out[q] = va + (vb - va) * w;
        }
        ++g;
    }
    st.next_g[r] = g;
}

template <class R>
__global__ __launch_bounds__(kT) void k_plain_select(DevModel m, int64_t Rn, uint64_t seed, uint32_t model_id,
                                                     int64_t rbase, double tf, int64_t max_iter,
                                                     int32_t *__restrict__ x, R *__restrict__ a,
                                                     const uint8_t *__restrict__ crit, const double *__restrict__ a0,
                                                     const double *__restrict__ a0c, const double *__restrict__ tau1,
                                                     PlainState st, int32_t nchunk, const double *__restrict__ partc) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    st.neg[r] = 0;
    if (!st.active[r]) {
        st.mode[r] = 0;
        return;
    }
    if (max_iter > 0 && st.steps[r] >= max_iter) {
        st.status[r] = kIterCap;
        st.active[r] = 0;
        st.mode[r] = 0;
        atomicAdd(st.n_done, 1ull);
        return;
    }
    const double a0r = a0[r];
    double t = st.t[r];
    auto xtrk = [&](int32_t q) { return x[(int64_t)st.track[q] * Rn + r]; };
    if (!(a0r > 0.0)) {  // absorbing state
        if (st.grid) grid_step(st, r, tf, t, tf, xtrk, xtrk);
        st.t[r] = tf;
        st.steps[r] += 1;
        st.active[r] = 0;
        st.mode[r] = 0;
        atomicAdd(st.n_done, 1ull);
        return;
    }
    const double ov = st.override_tau1[r];
    const double tau1r = ov >= 0.0 ? ov : tau1[r];
    st.tau1_used[r] = tau1r;
    const uint2 key = rng_key(seed, rbase + r);
    const uint32_t c3 = model_id << 2;
    if (tau1r < kSsaFactor / a0r) {  // direct block
        const double t_before = t;
        if (st.grid)
            for (int32_t q = 0; q < st.n_track; ++q) st.snap[(int64_t)q * Rn + r] = xtrk(q);
        if (st.dlist) {  // k_plain_direct runs the block (mode 3: no leap stage touches r)
            st.mode[r] = 3;
            st.dlist[atomicAdd(st.dcount, 1ull)] = r;
            return;
        }
        uint32_t ep = st.epoch[r];
        double a0s = a0r;
        int64_t n = 0;
        int32_t jf = -1;
        for (int s = 0; s < kSsaSteps; ++s) {
            if (s > 0) a0s = refresh_soa<R>(m, x, a, Rn, r, jf);
            if (!ssa_event_soa<R>(m, key, c3, ep, x, Rn, r, t, tf, a, a0s, &jf, st.S, st.D)) break;
            ++n;
        }
        if (st.grid) grid_step(st, r, tf, t_before, t, [&](int32_t q) { return st.snap[(int64_t)q * Rn + r]; }, xtrk);
        st.t[r] = t;
        st.epoch[r] = ep;
        st.ssa[r] += n;
        st.steps[r] += 1;
        st.retries[r] = 0;
        st.override_tau1[r] = -1.0;
        st.mode[r] = 2;
        if (t >= tf) {
            st.active[r] = 0;
            atomicAdd(st.n_done, 1ull);
        }
        return;
    }
    double tau;
    int32_t jc;
    uint8_t cl;
    select_leap<R>(m, key, c3, st.epoch[r], a, crit, Rn, r, a0c[r], tau1r, tf - t, tau, jc, cl, nchunk, partc);
    st.tau[r] = tau;
    st.jc[r] = jc;
    st.clamped[r] = cl;
    st.mode[r] = 1;
}

// Direct-method blocks of the realizations listed by k_plain_select, one
// workgroup per realization (workgroups stride the list): the 100 events of
// DirectMethodStep (model.cu:427-451, with the fixes of ssa_event_soa), with
// the channel work of each event spread over the group.  Lane l owns the
// contiguous channels [l*C, (l+1)*C): a0 is the group scan of the lane sums
// (a fixed order, so runs repeat bit for bit), the pick walks the one chunk
// whose prefix passes a0*r1, and after an event only the channels that read
// a species of the fired column are recomputed.  Every lane computes the
// same time, epoch and draws; the fired channel comes through LDS.
constexpr int kDT = 256;

template <class R>
__global__ __launch_bounds__(kDT) void k_plain_direct(DevModel m, int64_t Rn, uint64_t seed, uint32_t model_id,
                                                      int64_t rbase, double tf, int32_t *__restrict__ x,
                                                      R *__restrict__ a, PlainState st) {
    __shared__ double wtot[kDT / 64];
    __shared__ int32_t slot[2];
    const int tid = threadIdx.x, wl = tid & 63, wid = tid >> 6;
    const unsigned long long nd = *st.dcount;
    const R *rc = Real<R>::rates(m);
    const int32_t C = (m.M + kDT - 1) / kDT;
    const int32_t j0 = min(m.M, tid * C), j1 = min(m.M, j0 + C);
    const uint32_t c3 = model_id << 2;
    for (unsigned long long li = blockIdx.x; li < nd; li += gridDim.x) {
        const int64_t r = st.dlist[li];
        const uint2 key = rng_key(seed, rbase + r);
        double t = st.t[r];
        const double t_before = t;
        uint32_t ep = st.epoch[r];
        int64_t n = 0;
        int32_t jf = -1;
        for (int s = 0; s < kSsaSteps; ++s) {
            if (s > 0) {  // propensities after jf fired
                if (m.aff_ptr) {
                    const int32_t d1 = m.aff_ptr[jf + 1];
                    for (int32_t d = m.aff_ptr[jf] + tid; d < d1; d += kDT) {
                        const int32_t k = m.aff[d];
                        bool cr;
                        a[(int64_t)k * Rn + r] = soa_prop<R>(m, k, rc[k], x, Rn, r, cr);
                    }
                } else {
                    for (int32_t j = tid; j < m.M; j += kDT) {
                        bool cr;
                        a[(int64_t)j * Rn + r] = soa_prop<R>(m, j, rc[j], x, Rn, r, cr);
                    }
                }
                __syncthreads();
            }
            // lane sum of the chunk (loads batched), then the group scan
            double ps = 0.0;
            for (int32_t jb = j0; jb < j1; jb += kWalk) {
                R av[kWalk];
#pragma unroll
                for (int u = 0; u < kWalk; ++u) av[u] = jb + u < j1 ? a[(int64_t)(jb + u) * Rn + r] : (R)0;
#pragma unroll
                for (int u = 0; u < kWalk; ++u) ps += (double)av[u];
            }
            double incl = ps;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const double y = __shfl_up(incl, o, 64);
                if (wl >= o) incl += y;
            }
            double prev = __shfl_up(incl, 1, 64);
            if (wl == 63) wtot[wid] = incl;
            if (tid == 0) {
                slot[0] = 0x7fffffff;
                slot[1] = -1;
            }
            __syncthreads();
            double base = 0.0, a0s = 0.0;
#pragma unroll
            for (int w = 0; w < kDT / 64; ++w) {
                if (w == wid) base = a0s;
                a0s += wtot[w];
            }
            const double excl = base + (wl == 0 ? 0.0 : prev);
            incl += base;
            if (!(a0s > 0.0)) {
                t = tf;
                break;
            }
            double r1, r2;
            step_uniforms(key, ep, c3 | kKindSsa, r1, r2);
            ep++;
            const double dt = -log(r2) / a0s;
            if (dt >= tf - t) {
                if (st.S) {  // survival over rem = tf - t: d/dc_j of -a0*rem
                    const double rem = tf - t;
                    for (int32_t j = tid; j < m.M; j += kDT) {
                        const int64_t o = (int64_t)j * Rn + r;
                        if (rc[j] != (R)0) st.S[o] += a0s * st.D[o] - (double)a[o] * rem / (double)rc[j];
                    }
                }
                t = tf;
                break;
            }
            const double target = a0s * r1;
            if (incl > target) atomicMin(&slot[0], tid);
            __syncthreads();
            const int L = slot[0];
            if (L == 0x7fffffff) {  // rounding at the end: the last channel with a_j > 0
                int32_t lp = -1;
                for (int32_t j = j0; j < j1; ++j)
                    if (a[(int64_t)j * Rn + r] > (R)0) lp = j;
                if (lp >= 0) atomicMax(&slot[1], lp);
            } else if (tid == L) {
                double cum = excl;
                int32_t hitj = -1, lp = -1;
                for (int32_t j = j0; j < j1 && hitj < 0; ++j) {
                    const R aj = a[(int64_t)j * Rn + r];
                    if (aj > (R)0) lp = j;
                    cum += (double)aj;
                    if (cum > target) hitj = j;
                }
                slot[1] = hitj >= 0 ? hitj : lp;  // rounding at the chunk end
            }
            __syncthreads();
            jf = slot[1];
            if (st.S) {
                for (int32_t j = tid; j < m.M; j += kDT) {
                    const int64_t o = (int64_t)j * Rn + r;
                    if (rc[j] != (R)0) {
                        st.S[o] -= (double)a[o] * dt / (double)rc[j];
                        if (j == jf) st.S[o] += 1.0 / (double)rc[j];
                    }
                }
            }
            if (tid == 0)  // entries of one column may name a species twice: one lane adds them
                for (int32_t e = m.col_ptr[jf]; e < m.col_ptr[jf + 1]; ++e) {
                    const int32_t p = m.csc[e];
                    x[(int64_t)entry_index(p) * Rn + r] += entry_value(p);
                }
            t += dt;
            ++n;
            __syncthreads();
        }
        if (tid == 0) {
            auto xtrk = [&](int32_t q) { return x[(int64_t)st.track[q] * Rn + r]; };
            if (st.grid)
                grid_step(st, r, tf, t_before, t, [&](int32_t q) { return st.snap[(int64_t)q * Rn + r]; }, xtrk);
            st.t[r] = t;
            st.epoch[r] = ep;
            st.ssa[r] += n;
            st.steps[r] += 1;
            st.retries[r] = 0;
            st.override_tau1[r] = -1.0;
            st.mode[r] = 2;
            if (t >= tf) {
                st.active[r] = 0;
                atomicAdd(st.n_done, 1ull);
            }
        }
        __syncthreads();
    }
}

// Accept or reject the leap of each realization (rows i = blockIdx.y...).
__global__ __launch_bounds__(kT) void k_plain_commit(DevModel m, int64_t Rn, int32_t *__restrict__ x,
                                                     const int32_t *__restrict__ x_new, PlainState st, double tf) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (st.mode[r] != 1) return;
    const bool neg = st.neg[r] != 0;
    if (!neg)
        for (int32_t i = blockIdx.y; i < m.N; i += gridDim.y) x[(int64_t)i * Rn + r] = x_new[(int64_t)i * Rn + r];
    if (blockIdx.y != 0) return;
    st.epoch[r] += 1;
    if (neg) {
        st.rej[r] += 1;
        const int32_t rt = st.retries[r] + 1;
        st.retries[r] = rt;
        if (rt > kMaxRetries) {
            st.status[r] = kRetryExhausted;
            st.active[r] = 0;
            atomicAdd(st.n_done, 1ull);
        } else {
            st.override_tau1[r] = st.tau1_used[r] * 0.5;
        }
    } else {
        const double t = st.clamped[r] ? tf : st.t[r] + st.tau[r];
        if (!(t > st.t[r])) {  // time can not advance: stop instead of spinning
            st.status[r] = kStalled;
            st.active[r] = 0;
            atomicAdd(st.n_done, 1ull);
            return;
        }
        st.t[r] = t;
        st.steps[r] += 1;
        st.retries[r] = 0;
        st.override_tau1[r] = -1.0;
        if (t >= tf) {
            st.active[r] = 0;
            atomicAdd(st.n_done, 1ull);
        }
    }
}

// Grid points passed by each accepted leap (runs before the commit, while x
// still holds the state before the leap and x_new the state after it).
__global__ __launch_bounds__(kT) void k_plain_grid(int64_t Rn, const int32_t *__restrict__ x,
                                                   const int32_t *__restrict__ x_new, PlainState st, double tf) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (st.mode[r] != 1 || st.neg[r]) return;
    const double t = st.t[r];
    const double tn = st.clamped[r] ? tf : t + st.tau[r];
    if (!(tn > t)) return;  // the commit ends this realization as stalled
    grid_step(st, r, tf, t, tn, [&](int32_t q) { return x[(int64_t)st.track[q] * Rn + r]; },
              [&](int32_t q) { return x_new[(int64_t)st.track[q] * Rn + r]; });
}

// Grid points at and after the final time of a realization that ended OK
// take its final state.
__global__ __launch_bounds__(kT) void k_plain_grid_final(int64_t Rn, const int32_t *__restrict__ x, PlainState st) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (st.status[r] != kOk) return;
    const int32_t n = st.n_intervals, T = st.n_track;
    for (int32_t g = st.next_g[r]; g <= n; ++g)
        for (int32_t q = 0; q < T; ++q)
            st.grid[((size_t)r * (n + 1) + g) * T + q] = (double)x[(int64_t)st.track[q] * Rn + r];
}

// Likelihood-ratio terms of an accepted leap, channel chunk blockIdx.y:
// S_j += term_j / c_j and the chunk's sum of terms into qpart (see
// score_term).
template <class R>
__global__ __launch_bounds__(kT) void k_plain_score_leap(DevModel m, int64_t Rn, int32_t nchunk,
                                                         const R *__restrict__ a, const uint8_t *__restrict__ crit,
                                                         const int32_t *__restrict__ k, double *__restrict__ qpart,
                                                         PlainState st, int exact) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (st.mode[r] != 1 || st.neg[r]) return;
    const int32_t c = blockIdx.y;
    const int32_t ch = (m.M + nchunk - 1) / nchunk;
    const int32_t j0 = c * ch, j1 = min(m.M, j0 + ch);
    const R *rc = Real<R>::rates(m);
    const double tau = st.tau[r];
    const R tr = (R)tau;
    double q = 0.0;
    for (int32_t j = j0; j < j1; ++j) {
        const int64_t o = (int64_t)j * Rn + r;
        const double term = score_term<R>(crit[o] != 0, a[o], tr, tau, k[o], exact != 0);
        q += term;
        if (rc[j] != (R)0) st.S[o] += term / (double)rc[j];
    }
    qpart[(int64_t)c * Rn + r] = q;
}

// The step-size terms of an accepted leap, with Q = the sum of its terms:
// the clamped last leap (tau = t_final - t) adds -(Q / tau) D_j; a leap of
// size tau1 adds (Q / tau) dtau1/dc_j over the arg-min species row (mean
// bound -tau sgn(mu) v h / |mu|, variance bound -tau v^2 h / sigma^2,
// h = a_j / c_j) and accumulates dtau1/dc_j into D.
template <class R>
__global__ __launch_bounds__(kT) void k_plain_score_fin(DevModel m, int64_t Rn, int32_t nchunk,
                                                        const R *__restrict__ a, const uint8_t *__restrict__ crit,
                                                        const int32_t *__restrict__ argmin,
                                                        const int32_t *__restrict__ which,
                                                        const double *__restrict__ qpart, PlainState st) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    if (st.mode[r] != 1 || st.neg[r]) return;
    double Q = 0.0;
    for (int32_t c = 0; c < nchunk; ++c) Q += qpart[(int64_t)c * Rn + r];
    const double tau = st.tau[r];
    if (st.clamped[r]) {
        for (int32_t j = 0; j < m.M; ++j) {
            const int64_t o = (int64_t)j * Rn + r;
            st.S[o] -= (Q / tau) * st.D[o];
        }
        return;
    }
    const int32_t istar = argmin[r];
    if (st.jc[r] >= 0 || istar < 0) return;
    const R *rc = Real<R>::rates(m);
    double mu = 0.0, sig = 0.0;
    const int32_t e0 = m.row_ptr[istar], e1 = m.row_ptr[istar + 1];
    for (int32_t e = e0; e < e1; ++e) {
        const int32_t p = m.csr[e];
        const int64_t o = (int64_t)entry_index(p) * Rn + r;
        if (crit[o]) continue;
        const double v = (double)entry_value(p);
        mu += v * (double)a[o];
        sig += v * v * (double)a[o];
    }
    const int w = which[r];
    for (int32_t e = e0; e < e1; ++e) {
        const int32_t p = m.csr[e], j = entry_index(p);
        const int64_t o = (int64_t)j * Rn + r;
        if (crit[o] || rc[j] == (R)0) continue;
        const double v = (double)entry_value(p);
        const double h = (double)a[o] / (double)rc[j];
        const double dtau = (w == 0) ? -tau * (mu > 0.0 ? 1.0 : -1.0) * v * h / fabs(mu) : -tau * v * v * h / sig;
        st.S[o] += (Q / tau) * dtau;
        st.D[o] += dtau;
    }
}

// out[r][j] = in[j][r] through an LDS tile of 64 x 64 (both sides coalesced).
__global__ __launch_bounds__(256) void k_transpose_f64(int64_t Rn, int32_t M, const double *__restrict__ in,
                                                       double *__restrict__ out) {
    __shared__ double tile[64][65];
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    const int32_t j0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int jj = ty; jj < 64; jj += 4) {
        const int32_t j = j0 + jj;
        const int64_t r = r0 + tx;
        tile[jj][tx] = (j < M && r < Rn) ? in[(int64_t)j * Rn + r] : 0.0;
    }
    __syncthreads();
    for (int rr = ty; rr < 64; rr += 4) {
        const int64_t r = r0 + rr;
        const int32_t j = j0 + tx;
        if (j < M && r < Rn) out[r * M + j] = tile[tx][rr];
    }
}

__global__ __launch_bounds__(kT) void k_fill_f64(double *p, double v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kT) void k_gather_track(int64_t Rn, int32_t T, const int32_t *__restrict__ track,
                                                     const int32_t *__restrict__ x, int32_t *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= Rn) return;
    for (int32_t q = 0; q < T; ++q) out[r * T + q] = x[(int64_t)track[q] * Rn + r];
}

__global__ __launch_bounds__(kT) void k_fill_u8(uint8_t *p, uint8_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ __launch_bounds__(kT) void k_fill_u64(unsigned long long *p, unsigned long long v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n) p[i] = v;
}

unsigned row_grid(int32_t n) { return (unsigned)(n < 65535 ? n : 65535); }

}  // namespace

// ----------------------------------------------------------- launchers ----
void launch_prop(bool is64, const DevModel &m, int64_t R, int32_t nchunk, const int32_t *x, void *a, uint8_t *crit,
                 double *a0, double *a0c, double *part, double *partc, const uint8_t *active, hipStream_t s) {
    // realization pairs need R even and the buffers 2-element aligned
    const size_t rs = is64 ? 8 : 4;
    const bool pairs = R % 2 == 0 && (uintptr_t)x % 8 == 0 && (uintptr_t)a % (2 * rs) == 0 && (uintptr_t)crit % 2 == 0;
    const dim3 g(blocks_for(pairs ? R / 2 : R), (unsigned)nchunk);
    if (pairs && is64)
        k_prop2<double><<<g, kT, 0, s>>>(m, R, nchunk, x, (double *)a, crit, a0, a0c, part, partc, active);
    else if (pairs)
        k_prop2<float><<<g, kT, 0, s>>>(m, R, nchunk, x, (float *)a, crit, a0, a0c, part, partc, active);
    else if (is64)
        k_prop<double><<<g, kT, 0, s>>>(m, R, nchunk, x, (double *)a, crit, a0, a0c, part, partc, active);
    else
        k_prop<float><<<g, kT, 0, s>>>(m, R, nchunk, x, (float *)a, crit, a0, a0c, part, partc, active);
    if (nchunk > 1) k_sum_parts<<<blocks_for(R), kT, 0, s>>>(R, nchunk, part, partc, a0, a0c, active);
}

namespace {
template <class R>
void tau_stage(const DevModel &m, int64_t Rn, double eps, const int32_t *x, const void *a, const uint8_t *crit,
               void *tau_i, uint8_t *which_i, double *tau1, int32_t *argmin, int32_t *which, const uint8_t *active,
               hipStream_t s, int32_t nrc, void *pbest, int32_t *pidx, uint8_t *pwhich) {
    const R *ar = (const R *)a;
    if (nrc > 1 && pbest && pidx && pwhich) {
        // realization pairs need R even and the buffers 2-element aligned
        const bool pairs = Rn % 2 == 0 && (uintptr_t)x % 8 == 0 && (uintptr_t)a % (2 * sizeof(R)) == 0 &&
                           (uintptr_t)crit % 2 == 0;
        if (pairs)
            k_tau_rows_part2<R><<<dim3(blocks_for(Rn / 2), nrc), kT, 0, s>>>(m, Rn, (R)eps, nrc, x, ar, crit,
                                                                             (R *)pbest, pidx, pwhich, active);
        else
            k_tau_rows_part<R><<<dim3(blocks_for(Rn), nrc), kT, 0, s>>>(m, Rn, (R)eps, nrc, x, ar, crit, (R *)pbest,
                                                                        pidx, pwhich, active);
        k_tau_min_fin<R><<<blocks_for(Rn), kT, 0, s>>>(Rn, nrc, (const R *)pbest, pidx, nullptr, pwhich, tau1, argmin,
                                                       which, active);
        return;
    }
    k_tau_rows<R><<<dim3(blocks_for(Rn), row_grid(m.N)), kT, 0, s>>>(m, Rn, (R)eps, x, ar, crit, (R *)tau_i, which_i,
                                                                     active);
    if (nrc > 1 && pbest && pidx) {
        k_tau_min_part<R><<<dim3(blocks_for(Rn), nrc), kT, 0, s>>>(m.N, Rn, nrc, (const R *)tau_i, (R *)pbest, pidx,
                                                                   active);
        k_tau_min_fin<R><<<blocks_for(Rn), kT, 0, s>>>(Rn, nrc, (const R *)pbest, pidx, which_i, nullptr, tau1, argmin,
                                                       which, active);
    } else {
        k_tau_min<R><<<blocks_for(Rn), kT, 0, s>>>(m.N, Rn, (const R *)tau_i, which_i, tau1, argmin, which, active);
    }
}
}  // namespace

void launch_tau(bool is64, const DevModel &m, int64_t R, double eps, const int32_t *x, const void *a,
                const uint8_t *crit, void *tau_i, uint8_t *which_i, double *tau1, int32_t *argmin, int32_t *which,
                const uint8_t *active, hipStream_t s, int32_t nrc, void *pbest, int32_t *pidx, uint8_t *pwhich) {
    if (is64)
        tau_stage<double>(m, R, eps, x, a, crit, tau_i, which_i, tau1, argmin, which, active, s, nrc, pbest, pidx,
                          pwhich);
    else
        tau_stage<float>(m, R, eps, x, a, crit, tau_i, which_i, tau1, argmin, which, active, s, nrc, pbest, pidx,
                         pwhich);
}

void launch_select(bool is64, const DevModel &m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                   const uint32_t *epoch, const void *a, const uint8_t *crit, const double *a0c, const double *tau1,
                   const double *t_rem, double *tau, int32_t *jc, uint8_t *clamped, hipStream_t s) {
    if (is64)
        k_select<double><<<blocks_for(R), kT, 0, s>>>(m, R, seed, model_id, rbase, epoch, (const double *)a, crit, a0c,
                                                      tau1, t_rem, tau, jc, clamped);
    else
        k_select<float><<<blocks_for(R), kT, 0, s>>>(m, R, seed, model_id, rbase, epoch, (const float *)a, crit, a0c,
                                                     tau1, t_rem, tau, jc, clamped);
}

void launch_fire(bool is64, const DevModel &m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                 const uint32_t *epoch, const void *a, const uint8_t *crit, const double *tau, const int32_t *jc,
                 int32_t *k, const uint8_t *mode, int pflags, hipStream_t s) {
    const int32_t nquad = (m.M + 3) / 4;
    const dim3 g(blocks_for(R), row_grid(nquad), (unsigned)((nquad + 65534) / 65535));
    if (is64)
        k_fire<double><<<g, kT, 0, s>>>(m, R, seed, model_id, rbase, epoch, (const double *)a, crit, tau, jc, k, mode,
                                        pflags);
    else
        k_fire<float><<<g, kT, 0, s>>>(m, R, seed, model_id, rbase, epoch, (const float *)a, crit, tau, jc, k, mode,
                                       pflags);
}

void launch_update(const DevModel &m, int64_t R, const int32_t *x, const int32_t *k, int32_t *x_out, uint8_t *neg,
                   const uint8_t *mode, hipStream_t s) {
    // pairs need 8-byte aligned rows: R even (the buffers are 256-byte aligned)
    const bool pairs = R % 2 == 0 && ((uintptr_t)x | (uintptr_t)k | (uintptr_t)x_out) % 8 == 0;
    if (pairs)
        k_update2<<<dim3(blocks_for(R / 2), row_grid(m.N)), kT, 0, s>>>(m, R, x, k, x_out, neg, mode);
    else
        k_update<<<dim3(blocks_for(R), row_grid(m.N)), kT, 0, s>>>(m, R, x, k, x_out, neg, mode);
}

void launch_ssa(bool is64, const DevModel &m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase, int32_t *x,
                double *t, uint32_t *epoch, double t_final, int32_t n_events, int32_t *events, void *a,
                hipStream_t s) {
    if (is64)
        k_ssa<double><<<blocks_for(R), kT, 0, s>>>(m, R, seed, model_id, rbase, x, t, epoch, t_final, n_events, events,
                                                   (double *)a);
    else
        k_ssa<float><<<blocks_for(R), kT, 0, s>>>(m, R, seed, model_id, rbase, x, t, epoch, t_final, n_events, events,
                                                  (float *)a);
}

void launch_fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s) {
    if (n > 0) k_fill_u8<<<blocks_for(n), kT, 0, s>>>(p, v, n);
}
void launch_fill_u64(unsigned long long *p, unsigned long long v, int64_t n, hipStream_t s) {
    if (n > 0) k_fill_u64<<<blocks_for(n), kT, 0, s>>>(p, v, n);
}

void launch_plain_init(const DevModel &m, int64_t R, int32_t *x, PlainState st, hipStream_t s) {
    k_plain_init<<<blocks_for(R), kT, 0, s>>>(m, R, x, st);
}

void launch_plain_select(bool is64, const DevModel &m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                         double t_final, int64_t max_iter, int32_t *x, void *a, const uint8_t *crit, const double *a0,
                         const double *a0c, const double *tau1, PlainState st, int32_t nchunk, const double *partc,
                         hipStream_t s) {
    if (is64)
        k_plain_select<double><<<blocks_for(R), kT, 0, s>>>(m, R, seed, model_id, rbase, t_final, max_iter, x,
                                                            (double *)a, crit, a0, a0c, tau1, st, nchunk, partc);
    else
        k_plain_select<float><<<blocks_for(R), kT, 0, s>>>(m, R, seed, model_id, rbase, t_final, max_iter, x,
                                                           (float *)a, crit, a0, a0c, tau1, st, nchunk, partc);
}

void launch_plain_direct(bool is64, const DevModel &m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                         double t_final, int32_t *x, void *a, PlainState st, hipStream_t s) {
    // workgroups stride the list (its length is read on the device)
    const unsigned g = (unsigned)(R < 2048 ? R : 2048);
    if (is64)
        k_plain_direct<double><<<g, kDT, 0, s>>>(m, R, seed, model_id, rbase, t_final, x, (double *)a, st);
    else
        k_plain_direct<float><<<g, kDT, 0, s>>>(m, R, seed, model_id, rbase, t_final, x, (float *)a, st);
}

void launch_plain_commit(const DevModel &m, int64_t R, int32_t *x, const int32_t *x_new, PlainState st,
                         double t_final, hipStream_t s) {
    dim3 g(blocks_for(R), row_grid(m.N));
    k_plain_commit<<<g, kT, 0, s>>>(m, R, x, x_new, st, t_final);
}

void launch_plain_grid(const DevModel &, int64_t R, const int32_t *x, const int32_t *x_new, PlainState st,
                       double t_final, hipStream_t s) {
    k_plain_grid<<<blocks_for(R), kT, 0, s>>>(R, x, x_new, st, t_final);
}

void launch_plain_grid_final(const DevModel &, int64_t R, const int32_t *x, PlainState st, double, hipStream_t s) {
    k_plain_grid_final<<<blocks_for(R), kT, 0, s>>>(R, x, st);
}

void launch_plain_score(bool is64, const DevModel &m, int64_t R, int32_t nchunk, const void *a, const uint8_t *crit,
                        const int32_t *k, const int32_t *argmin, const int32_t *which, double *qpart, PlainState st,
                        bool exact, hipStream_t s) {
    const dim3 g(blocks_for(R), (unsigned)nchunk);
    if (is64) {
        k_plain_score_leap<double><<<g, kT, 0, s>>>(m, R, nchunk, (const double *)a, crit, k, qpart, st, exact ? 1 : 0);
        k_plain_score_fin<double><<<blocks_for(R), kT, 0, s>>>(m, R, nchunk, (const double *)a, crit, argmin, which,
                                                               qpart, st);
    } else {
        k_plain_score_leap<float><<<g, kT, 0, s>>>(m, R, nchunk, (const float *)a, crit, k, qpart, st, exact ? 1 : 0);
        k_plain_score_fin<float><<<blocks_for(R), kT, 0, s>>>(m, R, nchunk, (const float *)a, crit, argmin, which,
                                                              qpart, st);
    }
}

void launch_transpose_f64(int64_t R, int32_t M, const double *in, double *out, hipStream_t s) {
    const dim3 g((unsigned)((R + 63) / 64), (unsigned)((M + 63) / 64));
    k_transpose_f64<<<g, 256, 0, s>>>(R, M, in, out);
}

void launch_fill_f64(double *p, double v, int64_t n, hipStream_t s) {
    if (n > 0) k_fill_f64<<<blocks_for(n), kT, 0, s>>>(p, v, n);
}

void launch_gather_track(int64_t R, int32_t n_track, const int32_t *track, const int32_t *x, int32_t *out,
                         hipStream_t s) {
    k_gather_track<<<blocks_for(R), kT, 0, s>>>(R, n_track, track, x, out);
}

}  // namespace tl
