/*
 * tl_oracle_real.h -- precision-generic part of the CPU oracle.
 * Included twice by tl_oracle.c with REAL = double / float and SUF = f64 / f32.
 * TEST INFRASTRUCTURE ONLY (see tl_oracle.h).
 */

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

/* Uniform in (0,1) at the working precision. */
static inline REAL FN(unif)(uint32_t b) {
#if TLO_REAL_IS_DOUBLE
    return ((double)b + 0.5) * 2.3283064365386962890625e-10;
#else
    return ((float)(b >> 9) + 0.5f) * 1.1920928955078125e-07f; /* 23 bits: exact, never 0 or 1 */
#endif
}

#if TLO_REAL_IS_DOUBLE
#define R_EXP exp
#define R_LOG log
#define R_SQRT sqrt
#define R_FABS fabs
#else
/* FP32 exp and log in FP64, rounded once (as the engine evaluates them) */
#define R_EXP(v) ((float)exp((double)(v)))
#define R_LOG(v) ((float)log((double)(v)))
#define R_SQRT sqrtf
#define R_FABS fabsf
#endif

static inline const REAL *FN(rates)(const tlo_model *m) {
#if TLO_REAL_IS_DOUBLE
    return m->rates;
#else
    return m->rates_f;
#endif
}

/* Propensity and criticality of one channel (functions.h:100-123, nc = 5 from
 * model.cu:225).  Operation order is the reference's: (c*x0)*x1 and
 * ((c*x0)*(x0-1))/2.  A source (order 0) has a = c and is never critical. */
static inline REAL FN(prop1)(int order, REAL c, int32_t x0, int32_t x1, int32_t x2, int *crit) {
    switch (order) {
    case 0:
        *crit = 0;
        return c;
    case 4: /* Sa + Sb + Sc */
        *crit = (x0 < TLO_NC) || (x1 < TLO_NC) || (x2 < TLO_NC);
        return ((c * (REAL)x0) * (REAL)x1) * (REAL)x2;
    case 5: /* 2Sa + Sb */
        *crit = (x0 * 2 < TLO_NC) || (x1 < TLO_NC);
        return (((c * (REAL)x0) * (REAL)(x0 - 1)) * (REAL)0.5) * (REAL)x1;
    case 6: /* 3S */
        *crit = x0 * 3 < TLO_NC;
        return (((c * (REAL)x0) * (REAL)(x0 - 1)) * (REAL)(x0 - 2)) / (REAL)6;
    case 1:
        *crit = x0 < TLO_NC;
        return c * (REAL)x0;
    case 2:
        *crit = (x0 < TLO_NC) || (x1 < TLO_NC);
        return c * (REAL)x0 * (REAL)x1;
    default: /* 3 */
        *crit = x0 * 2 < TLO_NC;
        return c * (REAL)x0 * (REAL)(x0 - 1) / (REAL)2;
    }
}

/* a0 and a0c accumulate in FP64 in channel order. */
static void FN(prop_one)(const tlo_model *m, const int32_t *x, ptrdiff_t xs, REAL *a, uint8_t *crit,
                         ptrdiff_t as, double *a0, double *a0c) {
    const REAL *c = FN(rates)(m);
    double s = 0.0, sc = 0.0;
    for (int32_t j = 0; j < m->M; ++j) {
        int cr;
        REAL aj = FN(prop1)(m->order[j], c[j], x[(ptrdiff_t)m->s0[j] * xs], x[(ptrdiff_t)m->s1[j] * xs],
                            x[(ptrdiff_t)m->s2[j] * xs], &cr);
        a[(ptrdiff_t)j * as] = aj;
        if (crit) crit[(ptrdiff_t)j * as] = (uint8_t)cr;
        s += (double)aj;
        if (cr) sc += (double)aj;
    }
    *a0 = s;
    if (a0c) *a0c = sc;
}

/* GetG (CSRMatrixPacked.cu:294-309). */
static inline REAL FN(getg)(int order, int32_t x) {
    switch (order) {
    case 0: /* a source: the first-order bound */
    case 1: return (REAL)1;
    case 2: return (REAL)2;
    case 3: return (REAL)2 + (REAL)1 / ((REAL)x - (REAL)1);
    case 4: return (REAL)3;
    case 5: return (REAL)1.5 * ((REAL)2 + (REAL)1 / ((REAL)x - (REAL)1));
    case 6: return ((REAL)3 + (REAL)1 / ((REAL)x - (REAL)1)) + (REAL)2 / ((REAL)x - (REAL)2);
    default: return (REAL)0;
    }
}

/* TauFn + CompTau (CSRMatrixPacked.cu:311-358, 423-462): per species row mu and
 * sigma^2 over non-critical channels, tau_i, min seeded with FLT_MAX. */
static double FN(tau_one)(const tlo_model *m, REAL eps, const int32_t *x, ptrdiff_t xs, const REAL *a,
                          const uint8_t *crit, ptrdiff_t as, int32_t *argmin, int32_t *which) {
    REAL best = (REAL)FLT_MAX;
    int32_t bi = -1, bw = 0;
    for (int32_t i = 0; i < m->N; ++i) {
        /* the row sums run in FP64 in both variants and are rounded once to
         * REAL (the FP32 variant's only departure from all-FP32 arithmetic:
         * v * a_j and v^2 * a_j are exact in FP64, so the rounded sums no
         * longer depend on the order in which a device adds them up) */
        double mu_d = 0, sig_d = 0;
        for (int32_t e = m->row_ptr[i]; e < m->row_ptr[i + 1]; ++e) {
            int32_t p = m->csr_packed[e];
            int32_t j = p >> 4;
            double v = (double)((p & 15) - 3);
            double tmp = v * (double)a[(ptrdiff_t)j * as] * (double)(!crit[(ptrdiff_t)j * as]);
            sig_d += v * tmp;
            mu_d += tmp;
        }
        if (i >= m->M) continue; /* tau_i = +inf */
        REAL mu = (REAL)mu_d, sig = (REAL)sig_d;
        int32_t xi = x[(ptrdiff_t)i * xs];
        REAL g = FN(getg)(m->g_order[i], xi);
        REAL q = eps * (REAL)xi / g;
        REAL tmp = q > (REAL)1 ? q : (REAL)1;   /* fmax, q is never NaN for i < M */
        REAL t1 = tmp / R_FABS(mu);
        REAL t2 = tmp * tmp / R_FABS(sig);
        REAL ti = t1 < t2 ? t1 : t2;
        if (ti < best) {
            best = ti;
            bi = i;
            bw = (t1 <= t2) ? 0 : 1;
        }
    }
    if (argmin) *argmin = bi;
    if (which) *which = bw;
    return (double)best;
}

/* Poisson count by inversion of one uniform u: the smallest k with
 * u < F(k) = sum_{i <= k} P(i), P(0) = exp(-mean), P(i) = P(i-1) mean / i, in
 * REAL arithmetic (sequential-search inversion, L. Devroye, Non-Uniform
 * Random Variate Generation, Springer 1986, sec. X.3).  The search stops at
 * TLO_INVERT_CAP. */
static int32_t FN(poisson_invert)(REAL mean, REAL u) {
    REAL p = R_EXP(-mean);
    REAL F = p;
    int32_t k = 0;
    while (!(u < F) && k < TLO_INVERT_CAP) {
        ++k;
        p = (p * mean) / (REAL)k;
        F += p;
    }
    return k;
}

/* Firing count of a non-critical channel (model.cu:477-520, functions.h:37-54),
 * flags TLO_POISSON_EXACT / TLO_POISSON_FAST:
 *   mean < 80 (< 10 with FAST): inversion of draw 1 (poisson_invert; the
 *     reference multiplies mean + 1 uniforms, Knuth's product, for the same
 *     law, capped at 300 draws, as the search is capped at k = 300);
 *   10 <= mean < 80 with FAST, and above 80 with EXACT: PTRS (tlo_ptrs), an
 *     exact Poisson sample from a few uniforms;
 *   otherwise: |polar Box-Muller(mean, var = mean)| truncated.
 * Draw 1 of the channel is `head` (see head_draw); draws 2, 3, ... come from
 * the channel's own stream, output ((n-2) & 3) of Philox block (n-2) >> 2 at
 * counter (block, channel, epoch, ctr3). */
static int32_t FN(poisson_stream)(REAL mean, uint32_t head, const uint32_t key[2], uint32_t channel,
                                  uint32_t epoch, uint32_t ctr3, int flags) {
    uint32_t blk[4];
    uint32_t ctr[4] = {0u, channel, epoch, ctr3};
    int have = 0, pos = 0, first = 1;
#define NEXT_U(dst)                                   \
    do {                                              \
        if (first) {                                  \
            first = 0;                                \
            (dst) = FN(unif)(head);                   \
            break;                                    \
        }                                             \
        if (pos == have) {                            \
            tlo_philox4x32_10(ctr, key, blk);         \
            ctr[0]++;                                 \
            have = 4;                                 \
            pos = 0;                                  \
        }                                             \
        (dst) = FN(unif)(blk[pos++]);                 \
    } while (0)
    const REAL invert_below = (flags & TLO_POISSON_FAST) ? (REAL)TLO_PTRS_FROM : (REAL)TLO_POISSON_SWITCH;
    if (mean < invert_below) {
        return FN(poisson_invert)(mean, FN(unif)(head));
    } else if (mean < (REAL)TLO_POISSON_SWITCH || (flags & TLO_POISSON_EXACT)) {
        return tlo_ptrs((double)mean, head, key, channel, epoch, ctr3);
    } else {
        REAL s, u, v;
        do {
            REAL ua, ub;
            NEXT_U(ua);
            NEXT_U(ub);
            u = 2 * ua - 1;
            v = 2 * ub - 1;
            s = u * u + v * v;
        } while (s >= 1);
        REAL z = u * R_SQRT(-2 * R_LOG(s) / s);
        REAL y = mean + R_SQRT(mean) * z;
        return (int32_t)R_FABS(y);
    }
#undef NEXT_U
}

int32_t FN(tlo_poisson)(REAL mean, uint64_t seed, uint32_t model_id, int64_t rid, uint32_t epoch,
                        int32_t channel, int32_t flags) {
    uint32_t key[2];
    tlo_key(seed, rid, key);
    uint32_t head = tlo_head_draw(key, model_id << 2, epoch, channel);
    return FN(poisson_stream)(mean, head, key, (uint32_t)channel, epoch, (model_id << 2) | TLO_KIND_POISSON, flags);
}

/* Critical pick (GetReaction2Fire, model.cu:555-598): upper_bound over the
 * FP64 inclusive sums of the critical propensities; tau2 = -ln(r2)/a0c. */
static void FN(select_one)(const tlo_model *m, const uint32_t key[2], uint32_t ctr3, uint32_t epoch,
                           const REAL *a, const uint8_t *crit, ptrdiff_t as, double a0c, double tau1,
                           double t_rem, double *tau_out, int32_t *jc_out, uint8_t *clamped_out) {
    uint32_t ctr[4] = {0u, 0u, epoch, ctr3 | TLO_KIND_LEAP};
    uint32_t r[4];
    tlo_philox4x32_10(ctr, key, r);
    double r1 = tlo_unif64(r[0]), r2 = tlo_unif64(r[1]);
    int32_t jc = m->M;
    double tau2 = (double)FLT_MAX;
    if (a0c > 0.0) {
        double target = a0c * r1, cum = 0.0;
        for (int32_t j = 0; j < m->M; ++j) {
            if (!crit[(ptrdiff_t)j * as]) continue;
            cum += (double)a[(ptrdiff_t)j * as];
            if (cum > target) {
                jc = j;
                break;
            }
        }
        if (jc != m->M) tau2 = -log(r2) / a0c;
    }
    double tau = tau1 < tau2 ? tau1 : tau2;
    int fire = (jc != m->M) && (tau2 < tau1);
    uint8_t clamped = 0;
    if (tau >= t_rem) {
        tau = t_rem;
        clamped = 1;
        fire = 0;
    }
    *tau_out = tau;
    *jc_out = fire ? jc : -1;
    *clamped_out = clamped;
}

/* k_j for all channels (NumberOfFiringFn, model.cu:470-474). */
static void FN(fire_one)(const tlo_model *m, const uint32_t key[2], uint32_t ctr3, uint32_t epoch,
                         const REAL *a, const uint8_t *crit, ptrdiff_t as, double tau, int32_t jc,
                         int32_t *k, ptrdiff_t ks, int flags) {
    REAL tr = (REAL)tau;
    for (int32_t j = 0; j < m->M; ++j) {
        int32_t kj;
        if (crit[(ptrdiff_t)j * as])
            kj = (j == jc) ? 1 : 0;
        else
            kj = FN(poisson_stream)(a[(ptrdiff_t)j * as] * tr, tlo_head_draw(key, ctr3, epoch, j), key, (uint32_t)j,
                                    epoch, ctr3 | TLO_KIND_POISSON, flags);
        k[(ptrdiff_t)j * ks] = kj;
    }
}

/* One direct-method event (the fixed DirectMethodStep).  Returns 1 if an event
 * fired, 0 if the realization reached t_final (absorbing state or next event
 * past t_final).  a must hold the propensities of x on entry; on return after
 * an event it is stale.  score (may be NULL) gets the LR terms; dtime holds
 * dt/dc_j of the current time (see score_leap). */
static int FN(ssa_event)(const tlo_model *m, const uint32_t key[2], uint32_t ctr3, uint32_t *epoch,
                         int32_t *x, double *t, double t_final, const REAL *a, double a0,
                         double *score, const double *dtime) {
    const REAL *c = FN(rates)(m);
    if (!(a0 > 0.0)) {
        *t = t_final;
        return 0;
    }
    uint32_t ctr[4] = {0u, 0u, *epoch, ctr3 | TLO_KIND_SSA};
    uint32_t r[4];
    tlo_philox4x32_10(ctr, key, r);
    (*epoch)++;
    double r1 = tlo_unif64(r[0]), r2 = tlo_unif64(r[1]);
    double dt = -log(r2) / a0;
    if (dt >= t_final - *t) {
        if (score) {
            /* survival over rem = t_final - t: d/dc_j of -a0*rem */
            double rem = t_final - *t;
            for (int32_t j = 0; j < m->M; ++j)
                if (c[j] != 0) score[j] += a0 * dtime[j] - (double)a[j] * rem / (double)c[j];
        }
        *t = t_final;
        return 0;
    }
    double target = a0 * r1, cum = 0.0;
    int32_t jf = -1, last_pos = -1;
    for (int32_t j = 0; j < m->M; ++j) {
        if (a[j] > 0) last_pos = j;
        cum += (double)a[j];
        if (cum > target) {
            jf = j;
            break;
        }
    }
    if (jf < 0) jf = last_pos; /* only reachable through FP64 rounding at a0*r1 ~ a0 */
    for (int32_t e = m->col_ptr[jf]; e < m->col_ptr[jf + 1]; ++e) x[m->row_ind[e]] += m->stoich[e];
    if (score) {
        for (int32_t j = 0; j < m->M; ++j)
            if (c[j] != 0) score[j] -= (double)a[j] * dt / (double)c[j];
        if (c[jf] != 0) score[jf] += 1.0 / (double)c[jf];
    }
    *t += dt;
    return 1;
}

/* ---- SoA stage entry points ---- */
void FN(tlo_propensities)(const tlo_model *m, int64_t R, const int32_t *x, REAL *a, uint8_t *crit,
                          double *a0, double *a0c) {
    for (int64_t r = 0; r < R; ++r)
        FN(prop_one)(m, x + r, R, a + r, crit + r, R, a0 + r, a0c + r);
}

void FN(tlo_tau)(const tlo_model *m, int64_t R, double eps, const int32_t *x, const REAL *a,
                 const uint8_t *crit, double *tau1, int32_t *argmin, int32_t *which) {
    for (int64_t r = 0; r < R; ++r)
        tau1[r] = FN(tau_one)(m, (REAL)eps, x + r, R, a + r, crit + r, R, argmin ? argmin + r : NULL,
                              which ? which + r : NULL);
}

void FN(tlo_select)(const tlo_model *m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                    const uint32_t *epoch, const REAL *a, const uint8_t *crit, const double *a0c,
                    const double *tau1, const double *t_rem, double *tau, int32_t *jc,
                    uint8_t *clamped) {
    for (int64_t r = 0; r < R; ++r) {
        uint32_t key[2];
        tlo_key(seed, rbase + r, key);
        FN(select_one)(m, key, model_id << 2, epoch[r], a + r, crit + r, R, a0c[r], tau1[r], t_rem[r],
                       tau + r, jc + r, clamped + r);
    }
}

void FN(tlo_fire)(const tlo_model *m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                  const uint32_t *epoch, const REAL *a, const uint8_t *crit, const double *tau,
                  const int32_t *jc, int32_t *k) {
    for (int64_t r = 0; r < R; ++r) {
        uint32_t key[2];
        tlo_key(seed, rbase + r, key);
        FN(fire_one)(m, key, model_id << 2, epoch[r], a + r, crit + r, R, tau[r], jc[r], k + r, R, 0);
    }
}

void FN(tlo_ssa)(const tlo_model *m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                 int32_t *x, double *t, uint32_t *epoch, double t_final, int32_t n_events,
                 int32_t *events) {
    int32_t *xl = (int32_t *)malloc(sizeof(int32_t) * (size_t)m->N);
    REAL *al = (REAL *)malloc(sizeof(REAL) * (size_t)m->M);
    for (int64_t r = 0; r < R; ++r) {
        uint32_t key[2];
        tlo_key(seed, rbase + r, key);
        for (int32_t i = 0; i < m->N; ++i) xl[i] = x[(ptrdiff_t)i * R + r];
        int32_t n = 0;
        for (int32_t s = 0; s < n_events && t[r] < t_final; ++s) {
            double a0;
            FN(prop_one)(m, xl, 1, al, NULL, 1, &a0, NULL);
            if (!FN(ssa_event)(m, key, model_id << 2, &epoch[r], xl, &t[r], t_final, al, a0, NULL, NULL)) break;
            ++n;
        }
        for (int32_t i = 0; i < m->N; ++i) x[(ptrdiff_t)i * R + r] = xl[i];
        events[r] = n;
    }
    free(xl);
    free(al);
}

/* ---- likelihood-ratio score of one accepted leap ----
 * term_j = d_j * m_j with d_j = dlog p(k_j)/dm_j:  Poisson / critical (and
 * every channel with the exact sampler): k - m;  normal branch (m >= 80):
 * y - m + ((y-m)^2/m - 1)/2 with y = k + 1/2.
 * tau depends on c in two cases, each adding Q * dtau/dc_j with
 * Q = sum_l term_l / tau:
 *   - tau = tau1: through the argmin species' mean / variance bound; the same
 *     dtau/dc_j accumulates in dtime (dt/dc_j of the current time);
 *   - tau = t_final - t (clamped last leap): dtau/dc_j = -dtime_j.
 * A leap set by the critical waiting time tau2 has no such term (tau2 is one
 * of the random variables). */
static void FN(score_leap)(const tlo_model *m, const REAL *a, const uint8_t *crit, const int32_t *k,
                           double tau, int tau_from_tau1, int clamped, int32_t istar, int32_t which,
                           double *score, double *dtime, int exact) {
    const REAL *c = FN(rates)(m);
    double qsum = 0.0;
    REAL tr = (REAL)tau;
    for (int32_t j = 0; j < m->M; ++j) {
        double mj = (double)a[j] * tau;
        double term;
        if (!exact && !crit[j] && (a[j] * tr) >= (REAL)TLO_POISSON_SWITCH) {
            double y = (double)k[j] + 0.5, d = y - mj;
            term = d + 0.5 * (d * d / mj - 1.0);
        } else {
            term = (double)k[j] - mj;
        }
        qsum += term;
        if (c[j] != 0) score[j] += term / (double)c[j];
    }
    if (clamped) {
        for (int32_t j = 0; j < m->M; ++j) score[j] -= (qsum / tau) * dtime[j];
        return;
    }
    if (!tau_from_tau1 || istar < 0) return;
    /* mean / variance of the argmin row recomputed in FP64 */
    double mu = 0.0, sig = 0.0;
    for (int32_t e = m->row_ptr[istar]; e < m->row_ptr[istar + 1]; ++e) {
        int32_t p = m->csr_packed[e], j = p >> 4;
        double v = (double)((p & 15) - 3);
        if (crit[j]) continue;
        mu += v * (double)a[j];
        sig += v * v * (double)a[j];
    }
    for (int32_t e = m->row_ptr[istar]; e < m->row_ptr[istar + 1]; ++e) {
        int32_t p = m->csr_packed[e], j = p >> 4;
        double v = (double)((p & 15) - 3);
        if (crit[j] || c[j] == 0) continue;
        double h = (double)a[j] / (double)c[j];
        double dtau_dc;
        if (which == 0)
            dtau_dc = -tau * (mu > 0 ? 1.0 : -1.0) * v * h / fabs(mu);
        else
            dtau_dc = -tau * v * v * h / sig;
        score[j] += (qsum / tau) * dtau_dc;
        dtime[j] += dtau_dc;
    }
}

/* ---- one realization to t_final (PerformCalc, model.cu:640-737, with the
 * fixes listed in tl_oracle.h) ---- */
typedef struct FN(work) {
    REAL *a;
    uint8_t *crit;
    int32_t *x, *xn, *xb, *k;
    double *dtime; /* [M] dt/dc_j of the current time (sensitivities) */
} FN(work);

static void FN(grid_record)(const tlo_params *p, double t_a, const int32_t *x_a, double t_b,
                            const int32_t *x_b, int32_t *next_g, double *grid) {
    if (!grid || p->n_intervals <= 0) return;
    double h = p->t_final / (double)p->n_intervals;
    while (*next_g < p->n_intervals) {
        double g = h * (double)(*next_g);
        if (!(g < t_b)) break;
        double w = (t_b > t_a) ? (g - t_a) / (t_b - t_a) : 0.0;
        for (int32_t q = 0; q < p->n_track; ++q) {
            int32_t s = p->track[q];
            grid[(size_t)(*next_g) * p->n_track + q] = (double)x_a[s] + ((double)x_b[s] - (double)x_a[s]) * w;
        }
        (*next_g)++;
    }
}

static void FN(simulate_one)(const tlo_model *m, const tlo_params *p, int64_t rid, FN(work) * w,
                             int32_t *x_final, double *score, double *grid, int64_t *steps_out,
                             int64_t *ssa_out, int64_t *rej_out, int32_t *status_out) {
    const int32_t N = m->N;
    uint32_t key[2];
    tlo_key(p->seed, rid, key);
    const uint32_t ctr3 = p->model_id << 2;
    int32_t *x = w->x;
    memcpy(x, m->x0, sizeof(int32_t) * (size_t)N);
    if (score) {
        memset(score, 0, sizeof(double) * (size_t)m->M);
        memset(w->dtime, 0, sizeof(double) * (size_t)m->M);
    }
    double t = 0.0;
    const double tf = p->t_final;
    uint32_t epoch = 0;
    int64_t steps = 0, ssa = 0, rej = 0;
    int32_t status = TLO_OK;
    int32_t next_g = 0;
    const REAL eps = (REAL)p->eps;

    while (t < tf) {
        if (p->max_iter > 0 && steps >= p->max_iter) {
            status = TLO_ERR_ITER_CAP;
            break;
        }
        double a0, a0c;
        FN(prop_one)(m, x, 1, w->a, w->crit, 1, &a0, &a0c);
        double t_before = t;
        if (grid) memcpy(w->xb, x, sizeof(int32_t) * (size_t)N); /* state before the step */
        if (!(a0 > 0.0)) { /* absorbing state */
            t = tf;
            steps++;
            FN(grid_record)(p, t_before, w->xb, t, x, &next_g, grid);
            break;
        }
        int32_t istar, which;
        double tau1 = FN(tau_one)(m, eps, x, 1, w->a, w->crit, 1, &istar, &which);
        int retries = 0;
        for (;;) {
            if (tau1 < TLO_SSA_FACTOR / a0) { /* direct branch: up to 100 events */
                double a0s = a0;
                for (int s = 0; s < TLO_SSA_STEPS; ++s) {
                    if (s > 0) FN(prop_one)(m, x, 1, w->a, NULL, 1, &a0s, NULL);
                    if (!FN(ssa_event)(m, key, ctr3, &epoch, x, &t, tf, w->a, a0s, score, w->dtime)) break;
                    ssa++;
                }
                steps++;
                break;
            }
            /* leap attempt */
            double tau;
            int32_t jc;
            uint8_t clamped;
            FN(select_one)(m, key, ctr3, epoch, w->a, w->crit, 1, a0c, tau1, tf - t, &tau, &jc, &clamped);
            FN(fire_one)(m, key, ctr3, epoch, w->a, w->crit, 1, tau, jc, w->k, 1, p->poisson);
            epoch++;
            int32_t *xn = w->xn;
            memcpy(xn, x, sizeof(int32_t) * (size_t)N);
            for (int32_t j = 0; j < m->M; ++j) {
                int32_t kj = w->k[j];
                if (!kj) continue;
                for (int32_t e = m->col_ptr[j]; e < m->col_ptr[j + 1]; ++e) xn[m->row_ind[e]] += m->stoich[e] * kj;
            }
            int neg = 0;
            for (int32_t i = 0; i < N; ++i) neg |= xn[i] < 0;
            if (neg) {
                rej++;
                if (++retries > TLO_MAX_RETRIES) {
                    status = TLO_ERR_RETRY_EXHAUSTED;
                    goto done;
                }
                tau1 *= 0.5;
                continue;
            }
            if (score)
                FN(score_leap)(m, w->a, w->crit, w->k, tau, (jc < 0 && !clamped), clamped, istar, which, score,
                               w->dtime, (p->poisson & TLO_POISSON_EXACT) != 0);
            /* swap: x <- xn */
            double tn = clamped ? tf : t + tau;
            if (!(tn > t)) {
                status = TLO_ERR_STALLED;
                goto done;
            }
            w->x = xn;
            w->xn = x;
            x = xn;
            t = tn;
            steps++;
            break;
        }
        FN(grid_record)(p, t_before, w->xb, t, x, &next_g, grid);
    }
done:
    if (grid && p->n_intervals > 0 && status == TLO_OK) {
        /* grid point t_final takes the final state (the reference drops it) */
        while (next_g <= p->n_intervals) {
            for (int32_t q = 0; q < p->n_track; ++q)
                grid[(size_t)next_g * p->n_track + q] = (double)x[p->track[q]];
            next_g++;
        }
    }
    if (x_final) memcpy(x_final, x, sizeof(int32_t) * (size_t)N);
    *steps_out = steps;
    *ssa_out = ssa;
    *rej_out = rej;
    *status_out = status;
}

#undef FN
#undef CAT
#undef CAT_
#undef R_EXP
#undef R_LOG
#undef R_SQRT
#undef R_FABS
