/*
 * tl_oracle.c -- CPU oracle for the tau-leaping hot path.
 * TEST INFRASTRUCTURE ONLY (see tl_oracle.h for scope and reference citations).
 */
#define _GNU_SOURCE
#include "tl_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ---- Philox4x32-10 ---- */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void tlo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* key = (seed_lo, seed_hi + realization id) */
static inline void tlo_key(uint64_t seed, int64_t rid, uint32_t key[2]) {
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32) + (uint32_t)rid;
}

static inline double tlo_unif64(uint32_t b) { return ((double)b + 0.5) * 2.3283064365386962890625e-10; }

/* Exact Poisson sample (the sampler option that replaces the normal
 * approximation at mean >= 80): PTRS, the transformed rejection with squeeze
 * of W. Hoermann, "The transformed rejection method for generating Poisson
 * random variables", Insurance: Mathematics and Economics 12 (1993) 39-45,
 * valid for mean >= 10, in FP64 over FP64 uniforms.  Attempt i takes the
 * channel's draws 2i+1 and 2i+2 (draw 1 = head, draws n >= 2 as in the
 * Poisson stream): U = u - 1/2 from the first, V from the second.  After
 * TLO_PTRS_CAP rejected attempts (probability below 1e-60) it returns
 * floor(mean), so the loop ends even for a non-finite mean. */
#define TLO_PTRS_CAP 64
/* ln(k!) for integer k >= 0: the exact product below 10 (k! <= 362880 is
 * exact in FP64), else Stirling's series for ln Gamma(k + 1) to x^-7
 * (absolute error below 4e-13 from k = 10). */
static double tlo_lnfact(double k) {
    if (k < 10.0) {
        double f = 1.0;
        for (double i = 2.0; i <= k; i += 1.0) f *= i;
        return log(f);
    }
    const double x = k + 1.0, r = 1.0 / x, r2 = r * r;
    return (x - 0.5) * log(x) - x + 0.91893853320467274178 +
           r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0 - r2 / 1680.0)));
}

static int32_t tlo_ptrs(double mean, uint32_t head, const uint32_t key[2], uint32_t channel, uint32_t epoch,
                        uint32_t ctr3) {
    const double smu = sqrt(mean);
    const double b = 0.931 + 2.53 * smu;
    const double a = -0.059 + 0.02483 * b;
    const double lia = log(1.1239 + 1.1328 / (b - 3.4));
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    const double lmu = log(mean);
    uint32_t ctr[4] = {0u, channel, epoch, ctr3}, blk[4];
    int pos = 4;
#define PTRS_NEXT(dst)                            \
    do {                                          \
        if (pos == 4) {                           \
            tlo_philox4x32_10(ctr, key, blk);     \
            ctr[0]++;                             \
            pos = 0;                              \
        }                                         \
        (dst) = blk[pos++];                       \
    } while (0)
    uint32_t ub = head, vb;
    for (int attempt = 0; attempt < TLO_PTRS_CAP; ++attempt) {
        if (attempt > 0) PTRS_NEXT(ub);
        PTRS_NEXT(vb);
        const double U = tlo_unif64(ub) - 0.5, V = tlo_unif64(vb);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + mean + 0.43);
        if (us >= 0.07 && V <= vr) return (int32_t)k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + lia - log(a / (us * us) + b) <= -mean + k * lmu - tlo_lnfact(k)) return (int32_t)k;
    }
#undef PTRS_NEXT
    return (int32_t)mean;
}

/* First draw of channel j at an epoch: output (j & 3) of the Philox block
 * shared by channels 4b..4b+3, b = j >> 2, at counter (b, 0, epoch,
 * ctr3 | TLO_KIND_HEAD). */
static inline uint32_t tlo_head_draw(const uint32_t key[2], uint32_t ctr3, uint32_t epoch, int32_t j) {
    uint32_t ctr[4] = {(uint32_t)j >> 2, 0u, epoch, ctr3 | TLO_KIND_HEAD};
    uint32_t o[4];
    tlo_philox4x32_10(ctr, key, o);
    return o[j & 3];
}

/* ---- model build ---- */
void tlo_model_destroy(tlo_model *m) {
    if (!m) return;
    free(m->col_ptr);
    free(m->row_ind);
    free(m->stoich);
    free(m->rates);
    free(m->rates_f);
    free(m->x0);
    free(m->order);
    free(m->s0);
    free(m->s1);
    free(m->s2);
    free(m->row_ptr);
    free(m->csr_packed);
    free(m->g_order);
    free(m);
}

/* Reaction order from the negative entries of a column, in species order
 * (OrderDetectorAndIndFiller, CSCMatrix.cu:200-231).  A column without
 * negative entries is a source (order 0; the reference returns order -1 and
 * leaves its propensity unset); Sa+Sb+Sc (4), 2Sa+Sb (5) and 3S (6) have
 * three reactant molecules (the reference treats the first two as order 2
 * and cannot pack the coefficient 3).  Other patterns are rejected. */
static int detect_order(const tlo_model *m, int32_t j) {
    int32_t nneg = 0, v[3] = {0, 0, 0}, id[3] = {-1, -1, -1};
    for (int32_t e = m->col_ptr[j]; e < m->col_ptr[j + 1]; ++e) {
        if (m->stoich[e] >= 0) continue;
        if (nneg < 3) {
            v[nneg] = m->stoich[e];
            id[nneg] = m->row_ind[e];
        }
        nneg++;
    }
    m->s2[j] = 0;
    if (nneg == 0) {
        m->order[j] = 0;
        m->s0[j] = 0;
        m->s1[j] = 0;
        return TLO_OK;
    }
    if (nneg == 1 && v[0] == -2) {
        m->order[j] = 3;
        m->s0[j] = id[0];
        m->s1[j] = id[0];
        return TLO_OK;
    }
    if (nneg == 1 && v[0] == -3) {
        m->order[j] = 6; /* 3S */
        m->s0[j] = id[0];
        m->s1[j] = id[0];
        return TLO_OK;
    }
    if (nneg == 2 && v[0] == -1 && v[1] == -1 && id[0] != id[1]) {
        m->order[j] = 2;
        m->s0[j] = id[0];
        m->s1[j] = id[1];
        return TLO_OK;
    }
    if (nneg == 1 && v[0] == -1) {
        m->order[j] = 1;
        m->s0[j] = id[0];
        m->s1[j] = id[0];
        return TLO_OK;
    }
    if (nneg == 2 && id[0] != id[1] && v[0] + v[1] == -3 && (v[0] == -2 || v[1] == -2)) {
        m->order[j] = 5; /* 2Sa + Sb, s0 = the species taken twice */
        m->s0[j] = v[0] == -2 ? id[0] : id[1];
        m->s1[j] = v[0] == -2 ? id[1] : id[0];
        return TLO_OK;
    }
    if (nneg == 3 && v[0] == -1 && v[1] == -1 && v[2] == -1 && id[0] != id[1] && id[1] != id[2] && id[0] != id[2]) {
        m->order[j] = 4; /* Sa + Sb + Sc */
        m->s0[j] = id[0];
        m->s1[j] = id[1];
        m->s2[j] = id[2];
        return TLO_OK;
    }
    return TLO_ERR_UNSUPPORTED_ORDER;
}

int tlo_model_create(int32_t N, int32_t M, const int32_t *col_ptr, const int32_t *row_ind,
                     const int8_t *stoich, const double *rates, const int32_t *x0, tlo_model **out) {
    *out = NULL;
    if (N <= 0 || M <= 0 || !col_ptr || !row_ind || !stoich || !rates || !x0) return TLO_ERR_INVALID_ARG;
    if (col_ptr[0] != 0) return TLO_ERR_INVALID_ARG;
    for (int32_t j = 0; j < M; ++j)
        if (col_ptr[j + 1] < col_ptr[j]) return TLO_ERR_INVALID_ARG;
    int32_t nnz = col_ptr[M];
    for (int32_t e = 0; e < nnz; ++e) {
        if (row_ind[e] < 0 || row_ind[e] >= N) return TLO_ERR_INVALID_ARG;
        /* 4-bit packed entries (the reference packs 3 bits, [-2, 5], functions.h:126-130) */
        if (stoich[e] < -3 || stoich[e] > 12) return TLO_ERR_STOICH_RANGE;
    }
    tlo_model *m = (tlo_model *)calloc(1, sizeof(tlo_model));
    if (!m) return TLO_ERR_ALLOC;
    m->N = N;
    m->M = M;
    m->nnz = nnz;
    m->col_ptr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(M + 1));
    m->row_ind = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    m->stoich = (int8_t *)malloc((size_t)(nnz > 0 ? nnz : 1));
    m->rates = (double *)malloc(sizeof(double) * (size_t)M);
    m->rates_f = (float *)malloc(sizeof(float) * (size_t)M);
    m->x0 = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    m->order = (int8_t *)malloc((size_t)M);
    m->s0 = (int32_t *)malloc(sizeof(int32_t) * (size_t)M);
    m->s1 = (int32_t *)malloc(sizeof(int32_t) * (size_t)M);
    m->s2 = (int32_t *)calloc((size_t)M, sizeof(int32_t));
    m->row_ptr = (int32_t *)calloc((size_t)(N + 1), sizeof(int32_t));
    m->csr_packed = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    m->g_order = (int8_t *)malloc((size_t)N);
    if (!m->col_ptr || !m->row_ind || !m->stoich || !m->rates || !m->rates_f || !m->x0 || !m->order ||
        !m->s0 || !m->s1 || !m->row_ptr || !m->csr_packed || !m->g_order) {
        tlo_model_destroy(m);
        return TLO_ERR_ALLOC;
    }
    memcpy(m->col_ptr, col_ptr, sizeof(int32_t) * (size_t)(M + 1));
    memcpy(m->row_ind, row_ind, sizeof(int32_t) * (size_t)nnz);
    memcpy(m->stoich, stoich, (size_t)nnz);
    memcpy(m->rates, rates, sizeof(double) * (size_t)M);
    for (int32_t j = 0; j < M; ++j) m->rates_f[j] = (float)rates[j];
    memcpy(m->x0, x0, sizeof(int32_t) * (size_t)N);
    for (int32_t j = 0; j < M; ++j) {
        int st = detect_order(m, j);
        if (st != TLO_OK) {
            tlo_model_destroy(m);
            return st;
        }
    }
    /* CSC -> CSR transpose (the cuSPARSE csr2csc call of CSCMatrix.cu:147-178
     * is a pure permutation); entries of a row are in channel order, and the
     * two entries of a catalyst keep their CSC order. */
    for (int32_t e = 0; e < nnz; ++e) m->row_ptr[row_ind[e] + 1]++;
    for (int32_t i = 0; i < N; ++i) m->row_ptr[i + 1] += m->row_ptr[i];
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    if (!fill) {
        tlo_model_destroy(m);
        return TLO_ERR_ALLOC;
    }
    memcpy(fill, m->row_ptr, sizeof(int32_t) * (size_t)N);
    for (int32_t j = 0; j < M; ++j)
        for (int32_t e = col_ptr[j]; e < col_ptr[j + 1]; ++e)
            m->csr_packed[fill[row_ind[e]]++] = (j << 4) + (stoich[e] + 3);
    free(fill);
    for (int32_t i = 0; i < N; ++i) m->g_order[i] = (i < M) ? m->order[i] : 0;
    *out = m;
    return TLO_OK;
}

/* Net stoichiometry for the tau bound (the product's
 * tl_model_set_net_stoichiometry; the reference has no such option, its tau
 * bound takes every entry as read, tau_leaping/CSRMatrixPacked.cu:311-358).
 * A species row lists its entries in channel order, so the two entries of a
 * catalyst are adjacent: they merge in place into their sum, and a zero sum
 * leaves the row.  The state update keeps the entries as read (same net
 * change).  Idempotent; on TLO_ERR_STOICH_RANGE the model is unchanged. */
int tlo_model_set_net_stoichiometry(tlo_model *m) {
    for (int pass = 0; pass < 2; ++pass) {
        int32_t w = 0;
        for (int32_t i = 0; i < m->N; ++i) {
            const int32_t e0 = m->row_ptr[i], e1 = m->row_ptr[i + 1];
            if (pass) m->row_ptr[i] = w;
            for (int32_t e = e0; e < e1;) {
                const int32_t j = m->csr_packed[e] >> 4;
                int32_t v = 0;
                for (; e < e1 && (m->csr_packed[e] >> 4) == j; ++e) v += (m->csr_packed[e] & 15) - 3;
                if (v < -3 || v > 12) return TLO_ERR_STOICH_RANGE; /* only in the checking pass */
                if (pass && v != 0) m->csr_packed[w++] = (j << 4) + (v + 3);
            }
        }
        if (pass) m->row_ptr[m->N] = w;
    }
    return TLO_OK;
}

/* ---- precision-generic part ---- */
#define REAL double
#define SUF f64
#define TLO_REAL_IS_DOUBLE 1
#include "tl_oracle_real.h"
#undef REAL
#undef SUF
#undef TLO_REAL_IS_DOUBLE

#define REAL float
#define SUF f32
#define TLO_REAL_IS_DOUBLE 0
#include "tl_oracle_real.h"
#undef REAL
#undef SUF
#undef TLO_REAL_IS_DOUBLE

int32_t tlo_poisson_ptrs(double mean, uint64_t seed, uint32_t model_id, int64_t rid, uint32_t epoch,
                         int32_t channel) {
    uint32_t key[2];
    tlo_key(seed, rid, key);
    const uint32_t head = tlo_head_draw(key, model_id << 2, epoch, channel);
    return tlo_ptrs(mean, head, key, (uint32_t)channel, epoch, (model_id << 2) | TLO_KIND_POISSON);
}

void tlo_poisson_ptrs_many(double mean, uint64_t seed, uint32_t model_id, int64_t rid0, int64_t n,
                           uint32_t epoch, int32_t channel, int32_t *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = tlo_poisson_ptrs(mean, seed, model_id, rid0 + i, epoch, channel);
}

void tlo_update(const tlo_model *m, int64_t R, const int32_t *x, const int32_t *k, int32_t *x_out,
                uint8_t *neg) {
    for (int64_t r = 0; r < R; ++r) {
        for (int32_t i = 0; i < m->N; ++i) x_out[(ptrdiff_t)i * R + r] = x[(ptrdiff_t)i * R + r];
        for (int32_t j = 0; j < m->M; ++j) {
            int32_t kj = k[(ptrdiff_t)j * R + r];
            if (!kj) continue;
            for (int32_t e = m->col_ptr[j]; e < m->col_ptr[j + 1]; ++e)
                x_out[(ptrdiff_t)m->row_ind[e] * R + r] += m->stoich[e] * kj;
        }
        uint8_t n = 0;
        for (int32_t i = 0; i < m->N; ++i) n |= x_out[(ptrdiff_t)i * R + r] < 0;
        neg[r] = n;
    }
}

/* ---- threaded full simulation ---- */
typedef struct sim_job {
    const tlo_model *m;
    const tlo_params *p;
    int64_t rbase, R;
    tlo_outputs *o;
    atomic_llong next;
} sim_job;

static void *sim_worker(void *arg) {
    sim_job *jb = (sim_job *)arg;
    const tlo_model *m = jb->m;
    const tlo_params *p = jb->p;
    tlo_outputs *o = jb->o;
    size_t N = (size_t)m->N, M = (size_t)m->M;
    int32_t *x = (int32_t *)malloc(sizeof(int32_t) * N), *xn = (int32_t *)malloc(sizeof(int32_t) * N);
    int32_t *xb = (int32_t *)malloc(sizeof(int32_t) * N), *k = (int32_t *)malloc(sizeof(int32_t) * M);
    int32_t *xf = (int32_t *)malloc(sizeof(int32_t) * N);
    uint8_t *crit = (uint8_t *)malloc(M);
    void *a = malloc(sizeof(double) * M);
    double *score_tmp = p->sensitivities ? (double *)malloc(sizeof(double) * M) : NULL;
    double *dtime = (double *)malloc(sizeof(double) * M);
    size_t gsz = (p->n_intervals > 0 && p->n_track > 0) ? (size_t)(p->n_intervals + 1) * (size_t)p->n_track : 0;
    for (;;) {
        int64_t r = atomic_fetch_add(&jb->next, 1);
        if (r >= jb->R) break;
        int64_t steps, ssa, rej;
        int32_t status;
        double *score = o->score ? o->score + (size_t)r * M : score_tmp;
        double *grid = (o->grid && gsz) ? o->grid + (size_t)r * gsz : NULL;
        if (p->precision == 1) {
            work_f32 w = {(float *)a, crit, x, xn, xb, k, dtime};
            simulate_one_f32(m, p, jb->rbase + r, &w, xf, score, grid, &steps, &ssa, &rej, &status);
        } else {
            work_f64 w = {(double *)a, crit, x, xn, xb, k, dtime};
            simulate_one_f64(m, p, jb->rbase + r, &w, xf, score, grid, &steps, &ssa, &rej, &status);
        }
        if (o->x_final) memcpy(o->x_final + (size_t)r * N, xf, sizeof(int32_t) * N);
        if (o->steps) o->steps[r] = steps;
        if (o->ssa_events) o->ssa_events[r] = ssa;
        if (o->rejections) o->rejections[r] = rej;
        if (o->status) o->status[r] = status;
    }
    free(x);
    free(xn);
    free(xb);
    free(k);
    free(xf);
    free(crit);
    free(a);
    free(score_tmp);
    free(dtime);
    return NULL;
}

int tlo_simulate(const tlo_model *m, const tlo_params *p, int64_t rbase, int64_t R, tlo_outputs *o) {
    int nt = p->n_threads;
    if (nt <= 0) nt = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nt < 1) nt = 1;
    if ((int64_t)nt > R) nt = (int)(R > 0 ? R : 1);
    sim_job jb;
    jb.m = m;
    jb.p = p;
    jb.rbase = rbase;
    jb.R = R;
    jb.o = o;
    atomic_init(&jb.next, 0);
    if (nt == 1) {
        sim_worker(&jb);
        return 1;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nt);
    for (int i = 0; i < nt; ++i) pthread_create(&th[i], NULL, sim_worker, &jb);
    for (int i = 0; i < nt; ++i) pthread_join(th[i], NULL);
    free(th);
    return nt;
}
