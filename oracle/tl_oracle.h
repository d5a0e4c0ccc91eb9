/*
 * tl_oracle.h -- CPU oracle for the tau-leaping hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * per-realization algorithm (CSSL-UWM/tau_leaping) with the documented defect
 * fixes applied.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline; the
 * product library (tau_leaping_amd) never links or calls it.
 *
 * Reference functions restated here (file:line in the reference tree):
 *   propensity + criticality ..... tau_leaping/functions.h:100-123,
 *                                  tau_leaping/model.cu:180-247
 *   reaction order / s0 / s1 ..... tau_leaping/CSCMatrix.cu:185-251
 *   CSC assembly from reactions .. tau_leaping/CSCHostMatrixData.cpp:19-107
 *   packed CSR entry ............. tau_leaping/functions.h:126-146
 *   tau selection (GetG, TauFn) .. tau_leaping/CSRMatrixPacked.cu:294-358,423-462
 *   step / branch ................ tau_leaping/model.cu:301-413
 *   critical pick, tau2 .......... tau_leaping/model.cu:555-598
 *   Poisson firing ............... tau_leaping/model.cu:458-549,
 *                                  tau_leaping/functions.h:37-54 (polar Box-Muller)
 *   population update ............ tau_leaping/CSRMatrixPacked.cu:464-503
 *   direct-method step ........... tau_leaping/model.cu:427-451
 *   main loop .................... tau_leaping/model.cu:640-737
 *   grid statistics .............. tau_leaping/statistics.cpp:15-87
 *
 * Deviations from the shipped reference (each fixes a defect or replaces a
 * non-reproducible part):
 *   - FP64 simulation time, leap clamped to t_final, SSA stops before an event
 *     past t_final (ref: float time, model.h:37).
 *   - a0 / a0c / cumulative searches accumulate in FP64 (ref: FP32 scans).
 *   - tau2 = ln(1/r2)/a0c (ref divides by a_k, model.cu:596).
 *   - direct step fires the selected channel once, refreshes propensities
 *     between events, dt = ln(1/r)/a0 (ref fires nothing, model.cu:442-446).
 *   - negative populations: reject the leap, halve tau1, redraw; at most 32
 *     retries (ref: no handling, model.cu:391-398).
 *   - population update covers all N species (ref loops k.size()=M rows,
 *     CSRMatrixPacked.cu:502).
 *   - RNG: Philox4x32-10 counter streams keyed per realization instead of the
 *     per-channel hybrid Tausworthe state (d_rng.cuh:27-35).
 *   - sources (no reactant, -> products) have a = c, are never critical and
 *     take g = 1 where GetG reads their order; channels with three reactant
 *     molecules, Sa+Sb+Sc (order 4: ((c*xa)*xb)*xc, g = 3), 2Sa+Sb (order
 *     5: (((c*xa)*(xa-1))*0.5)*xb, g = 1.5*(2 + 1/(x-1))) and 3S (order 6:
 *     (((c*x)*(x-1))*(x-2))/6, g = 3 + 1/(x-1) + 2/(x-2)), are critical when
 *     a reactant's coefficient times its count is below nc; other reactant
 *     patterns are rejected at model build (ref: order -1 gives an
 *     uninitialised propensity, functions.h:118-120, CSCMatrix.cu:225-231,
 *     and a third reactant overwrites the second).
 *   - packed entries hold 4 bits of stoichiometry, (index << 4) | (v + 3),
 *     v in [-3, 12] (ref: 3 bits, (index << 3) | ((v + 2) & 7), [-2, 5]),
 *     so that 3S can be written.
 *   - tau_i for species rows i >= M is +inf (ref reads order[] out of bounds,
 *     CSRMatrixPacked.cu:353).
 *   - the species sums mu_i, sigma^2_i of the tau bound accumulate in FP64 and
 *     are rounded once to the propensity type (ref: FP32 sums); v * a_j is
 *     exact in FP64, so the rounded sums do not depend on the order of the
 *     additions to within the last FP64 bit.
 *   - mean < 80: sequential-search inversion of the channel's first uniform
 *     (ref: Knuth's product of mean + 1 uniforms, capped at 300,
 *     model.cu:483-513, for the same Poisson law; the search is capped at
 *     k = 300 too).
 *   - options: an exact Poisson sampler (PTRS, Hoermann 1993) in place of
 *     the normal approximation at mean >= 80 (TLO_POISSON_EXACT), and in
 *     place of the inversion for 10 <= mean < 80 (TLO_POISSON_FAST).
 *   - the first Poisson draw of channel j at a step is output (j & 3) of the
 *     Philox block shared by channels 4*(j>>2)..4*(j>>2)+3; later draws come
 *     from the channel's own counter stream (see tlo_head_draw).
 */
#ifndef TL_ORACLE_H
#define TL_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TLO_OK = 0,
    TLO_ERR_INVALID_ARG = 9,
    TLO_ERR_UNSUPPORTED_ORDER = 3,
    TLO_ERR_STOICH_RANGE = 4,
    TLO_ERR_ALLOC = 5,
    TLO_ERR_RETRY_EXHAUSTED = 7,
    TLO_ERR_ITER_CAP = 8,
    TLO_ERR_STALLED = 12
};

/* method constants (tau_leaping/model.cu:225, model.h:34, model.cu:343,486,497) */
#define TLO_NC 5
#define TLO_SSA_FACTOR 10.0
#define TLO_SSA_STEPS 100
#define TLO_POISSON_SWITCH 80
#define TLO_INVERT_CAP 300
#define TLO_MAX_RETRIES 32
#define TLO_PTRS_FROM 10 /* PTRS from this mean (the method's range) with TLO_POISSON_FAST */

/* sampler flags (tlo_params.poisson) */
#define TLO_POISSON_EXACT 1 /* PTRS above mean 80 instead of the normal approximation */
#define TLO_POISSON_FAST 2  /* PTRS from mean 10 instead of the inversion up to 80 */

/* RNG stream kinds (ctr[3] low two bits) */
#define TLO_KIND_POISSON 0u
#define TLO_KIND_LEAP 1u
#define TLO_KIND_SSA 2u
#define TLO_KIND_HEAD 3u /* first draw of channels 4b..4b+3 */

typedef struct tlo_model {
    int32_t N, M, nnz;
    int32_t *col_ptr;    /* [M+1]  channel-major (CSC) */
    int32_t *row_ind;    /* [nnz]  species of each entry */
    int8_t *stoich;      /* [nnz]  signed stoichiometry of each entry */
    double *rates;       /* [M] */
    float *rates_f;      /* [M]    rates rounded once to float */
    int32_t *x0;         /* [N] */
    int8_t *order;       /* [M]    0: ->, 1: S->, 2: Sa+Sb->, 3: 2S->, 4: Sa+Sb+Sc->, 5: 2Sa+Sb-> (s0 = a), 6: 3S-> */
    int32_t *s0, *s1;    /* [M]    reactant species (s1 == s0 for orders 1, 3) */
    int32_t *row_ptr;    /* [N+1]  species-major (CSR) */
    int32_t *csr_packed; /* [nnz]  (channel << 4) | (v + 3), v in [-3, 12] */
    int8_t *g_order;     /* [N]    order[i] for i < M, else 0 */
    int32_t *s2;         /* [M]    third reactant of order-4 channels (else 0) */
} tlo_model;

/* Philox4x32-10 (Salmon et al., SC'11). */
void tlo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Model build from channel-major entries.  Entries of one channel must be
 * sorted by species, a species that is both reactant and product appearing as
 * two entries (reactant first), as ReadXMLToMatrixStruct builds them. */
int tlo_model_create(int32_t N, int32_t M, const int32_t *col_ptr, const int32_t *row_ind,
                     const int8_t *stoich, const double *rates, const int32_t *x0,
                     tlo_model **out);
void tlo_model_destroy(tlo_model *m);
/* tau bound over net stoichiometry: each species row keeps one entry per
 * channel holding the sum of that channel's entries for the species, zero
 * sums dropped (row_ptr / csr_packed rewritten; nnz keeps the entries as
 * read). */
int tlo_model_set_net_stoichiometry(tlo_model *m);

/* ---- stage functions over structure-of-arrays buffers ----
 * x is [N][R] int32, a is [M][R] real, crit is [M][R] uint8, per-realization
 * scalars are [R].  f64 / f32 select the propensity precision. */
void tlo_propensities_f64(const tlo_model *m, int64_t R, const int32_t *x, double *a,
                          uint8_t *crit, double *a0, double *a0c);
void tlo_propensities_f32(const tlo_model *m, int64_t R, const int32_t *x, float *a,
                          uint8_t *crit, double *a0, double *a0c);

/* tau1 per realization; argmin / which (0: mean bound, 1: variance bound)
 * may be NULL. */
void tlo_tau_f64(const tlo_model *m, int64_t R, double eps, const int32_t *x, const double *a,
                 const uint8_t *crit, double *tau1, int32_t *argmin, int32_t *which);
void tlo_tau_f32(const tlo_model *m, int64_t R, double eps, const int32_t *x, const float *a,
                 const uint8_t *crit, double *tau1, int32_t *argmin, int32_t *which);

/* Leap selection: critical pick, tau2, tau, critical firing, clamp to t_rem.
 * jc = -1 when no critical channel fires. */
void tlo_select_f64(const tlo_model *m, int64_t R, uint64_t seed, uint32_t model_id,
                    int64_t rbase, const uint32_t *epoch, const double *a, const uint8_t *crit,
                    const double *a0c, const double *tau1, const double *t_rem, double *tau,
                    int32_t *jc, uint8_t *clamped);
void tlo_select_f32(const tlo_model *m, int64_t R, uint64_t seed, uint32_t model_id,
                    int64_t rbase, const uint32_t *epoch, const float *a, const uint8_t *crit,
                    const double *a0c, const double *tau1, const double *t_rem, double *tau,
                    int32_t *jc, uint8_t *clamped);

/* Firing counts k [M][R] for one leap attempt at the given epochs. */
void tlo_fire_f64(const tlo_model *m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                  const uint32_t *epoch, const double *a, const uint8_t *crit, const double *tau,
                  const int32_t *jc, int32_t *k);
void tlo_fire_f32(const tlo_model *m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                  const uint32_t *epoch, const float *a, const uint8_t *crit, const double *tau,
                  const int32_t *jc, int32_t *k);

/* x_out = x + S k ; neg[r] = any species negative. */
void tlo_update(const tlo_model *m, int64_t R, const int32_t *x, const int32_t *k,
                int32_t *x_out, uint8_t *neg);

/* Up to n_events direct-method events per realization (x, t, epoch updated in
 * place); events[r] receives the number fired. */
void tlo_ssa_f64(const tlo_model *m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                 int32_t *x, double *t, uint32_t *epoch, double t_final, int32_t n_events,
                 int32_t *events);
void tlo_ssa_f32(const tlo_model *m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                 int32_t *x, double *t, uint32_t *epoch, double t_final, int32_t n_events,
                 int32_t *events);

/* Exact sampler (PTRS, mean >= 10) on the draws of channel j at an epoch
 * (test hook). */
int32_t tlo_poisson_ptrs(double mean, uint64_t seed, uint32_t model_id, int64_t rid, uint32_t epoch,
                         int32_t channel);
/* The same for realizations rid0 .. rid0 + n - 1. */
void tlo_poisson_ptrs_many(double mean, uint64_t seed, uint32_t model_id, int64_t rid0, int64_t n,
                           uint32_t epoch, int32_t channel, int32_t *out);

/* Single Poisson draw from the channel stream (test hook for the sampler). */
int32_t tlo_poisson_f64(double mean, uint64_t seed, uint32_t model_id, int64_t rid,
                        uint32_t epoch, int32_t channel, int32_t flags);
int32_t tlo_poisson_f32(float mean, uint64_t seed, uint32_t model_id, int64_t rid,
                        uint32_t epoch, int32_t channel, int32_t flags);

/* ---- full simulation ---- */
typedef struct tlo_params {
    double t_final;
    double eps;
    uint64_t seed;
    uint32_t model_id;
    int32_t precision;     /* 0: FP64, 1: FP32 */
    int64_t max_iter;      /* <= 0: unlimited */
    int32_t sensitivities; /* accumulate the likelihood-ratio score per channel */
    int32_t n_intervals;   /* > 0: record tracked species on the grid k*t_final/n_intervals */
    int32_t n_track;
    const int32_t *track;  /* [n_track] species indices */
    int32_t n_threads;     /* 0: all online cores */
    int32_t poisson;       /* TLO_POISSON_EXACT | TLO_POISSON_FAST */
} tlo_params;

typedef struct tlo_outputs {
    int32_t *x_final;      /* [R][N] or NULL */
    double *score;         /* [R][M] or NULL (sensitivities) */
    double *grid;          /* [R][n_intervals+1][n_track] or NULL */
    int64_t *steps;        /* [R] accepted main-loop iterations */
    int64_t *ssa_events;   /* [R] */
    int64_t *rejections;   /* [R] */
    int32_t *status;       /* [R] */
} tlo_outputs;

/* Simulate realizations rbase .. rbase+R-1 (outputs indexed 0..R-1).
 * Returns the number of threads used. */
int tlo_simulate(const tlo_model *m, const tlo_params *p, int64_t rbase, int64_t R,
                 tlo_outputs *o);

#ifdef __cplusplus
}
#endif
#endif
