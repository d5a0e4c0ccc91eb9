/*
 * tl_api.h -- C ABI of the MI355X tau-leaping engine (libtau_leaping_amd.so).
 *
 * The reference exposes its simulation path as the C++ class `Model`
 * (tau_leaping/model.h:20-98) driven by tau_leaping/tau_leaping.cpp:40-190:
 *   Model::ReadXML(fn, names)                      -> tl_model_from_stochkit
 *   ReadXMLToMatrixStruct + Model ctor             -> tl_model_from_csc
 *     (tau_leaping/CSCHostMatrixData.cpp:19-107, tau_leaping/model.cpp:17-60)
 *   Model::Clone / InitSeeds / PerformCalc over a
 *     serial realization loop (tau_leaping.cpp:127-150,
 *     model.cu:640-737, model.cu:739-775)          -> tl_simulate
 *   AllStat2StepStat / StepStat2StdDev
 *     (tau_leaping/statistics.cpp:15-87)           -> tl_result.mean / stddev
 *   ChannelsCount / ReactantsCount (model.h:44-45) -> tl_model_n_channels / _n_species
 * and internally the per-stage device operations, exposed here for parity
 * testing on caller-provided device buffers:
 *   CalcPropensities (model.cu:213-247)            -> tl_stage_propensities
 *   CSRMatrixPacked::CompTau (CSRMatrixPacked.cu:423-462) -> tl_stage_tau
 *   GetReaction2Fire + leap branch (model.cu:346-372,555-598) -> tl_stage_select
 *   SetNumberOfFiring (model.cu:538-549)           -> tl_stage_fire
 *   CSRMatrixPacked::FireReactions (CSRMatrixPacked.cu:494-503) -> tl_stage_update
 *   DirectMethodStep x100 (model.cu:343-344,427-451) -> tl_stage_ssa
 *
 * Conventions: every call returns tl_status (0 = OK) and never throws or
 * exits; tl_last_error() returns a thread-local message for the last failure.
 * Model handles are library-owned and immutable once built except through
 * tl_model_set_rates / tl_model_set_x0 / tl_model_set_net_stoichiometry.
 * Result buffers are caller-owned.
 * All simulation entry points require an AMD GPU; there is no CPU fallback.
 */
#ifndef TL_API_H
#define TL_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t tl_status;

enum {
    TL_OK = 0,
    TL_ERR_IO = 1,                /* file missing or unreadable (StochKitReader.cpp:127-131) */
    TL_ERR_PARSE = 2,             /* malformed XML / missing node / unknown rate id / count mismatch */
    TL_ERR_UNSUPPORTED_ORDER = 3, /* reactant pattern other than ->, S->, 2S->, Sa+Sb->, Sa+Sb+Sc->, 2Sa+Sb->,
                                     3S-> (the reference accepts S->, 2S->, Sa+Sb->: CSCMatrix.cu:225-231) */
    TL_ERR_STOICH_RANGE = 4,      /* stoichiometry outside [-3, 12] (4-bit packing; the reference's 3-bit
                                     packing, functions.h:126-130, holds [-2, 5]) */
    TL_ERR_ALLOC = 5,
    TL_ERR_HIP = 6,
    TL_ERR_RETRY_EXHAUSTED = 7,   /* a leap was rejected more than 32 times in a row */
    TL_ERR_ITER_CAP = 8,          /* -a maxIterations reached before t_final (model.cu:661-662) */
    TL_ERR_INVALID_ARG = 9,
    TL_ERR_NO_DEVICE = 10,
    TL_ERR_TOO_LARGE = 11,
    TL_ERR_STALLED = 12           /* an accepted leap did not advance the FP64 time */
};

enum { TL_FP64 = 0, TL_FP32 = 1 };

/* Firing counts of non-critical channels (model.cu:477-520).  NORMAL
 * (default): below mean 80 a Poisson count by inversion of the channel's
 * first uniform, the support searched outward from the mode (the reference
 * multiplies mean + 1 uniforms, Knuth's product, for the same law), and the
 * reference's truncated normal |N(m, m)| from 80.  EXACT: PTRS (Hoermann
 * 1993, an exact Poisson sample) from 80 instead.  FAST: PTRS from mean 10
 * to 80 instead of the inversion, |N(m, m)| from 80.  FAST_EXACT: PTRS from
 * mean 10. */
enum { TL_POISSON_NORMAL = 0, TL_POISSON_EXACT = 1, TL_POISSON_FAST = 2, TL_POISSON_FAST_EXACT = 3 };

/* Engine forms: PLAIN = per-model host loop of stage kernels over
 * structure-of-arrays state in HBM; PERSISTENT = one work-stealing kernel that
 * runs whole trajectories with each realization's state in LDS, or in a slice
 * of HBM per block for models too large for LDS (all models of a batch in one
 * launch); AUTO = PERSISTENT, except for a model no persistent launch shape
 * can run (PERSISTENT returns TL_ERR_TOO_LARGE) and for a model too large for
 * LDS whose persistent shape would need more than one propensity tile (16,384
 * channels for FP64 1024-lane groups) or a[M] in HBM, where PLAIN's coalesced
 * realization-major accesses were measured 1.03-7x faster: those run PLAIN.  The
 * choice is a pure function of the model sizes and the precision, and
 * tl_result.form reports it.  PLAIN holds about
 * 8 N + (5 + p) M bytes of device memory per realization (p = 8 for FP64, 4
 * for FP32), plus (p + 1) N when the species rows, cut into chunks of at
 * least 64, give fewer than 262,144 (chunk, realization) pairs: then the
 * per-species bounds are stored. */
enum { TL_MODE_AUTO = 0, TL_MODE_PLAIN = 1, TL_MODE_PERSISTENT = 2 };

/* tl_sim_params.block_size: the timed autotuner instead of the default shape */
enum { TL_BLOCK_TUNE = -1 };

typedef struct tl_model tl_model;

/* ---- models ---- */
tl_status tl_model_from_stochkit(const char *path, tl_model **out);
tl_status tl_model_from_csc(int32_t n_species, int32_t n_channels, const int32_t *col_ptr,
                            const int32_t *row_ind, const int8_t *stoich, const double *rates,
                            const int32_t *x0, tl_model **out);
void tl_model_destroy(tl_model *m);
int32_t tl_model_n_species(const tl_model *m);
int32_t tl_model_n_channels(const tl_model *m);
int32_t tl_model_nnz(const tl_model *m);
/* Species name (XML Id, or "S<i+1>" for CSC-built models); NULL if out of range. */
const char *tl_model_species_name(const tl_model *m, int32_t i);
tl_status tl_model_species_index(const tl_model *m, const char *name, int32_t *out);
tl_status tl_model_get_rates(const tl_model *m, double *rates);
tl_status tl_model_set_rates(tl_model *m, const double *rates);
tl_status tl_model_get_x0(const tl_model *m, int32_t *x0);
tl_status tl_model_set_x0(tl_model *m, const int32_t *x0);
/* tau bound from net stoichiometry (on != 0): the entries of one species in
 * one channel (a catalyst's reactant and product entries) count once, with
 * their sum, in the species sums mu_i = sum_j v_ij a_j and
 * sigma^2_i = sum_j v_ij^2 a_j; off (the default, as
 * tau_leaping/CSCHostMatrixData.cpp:74-103 builds the entries) each entry
 * counts, so a catalyst adds to sigma^2 with no change of its population.
 * Reaction orders, criticality and the state update are the same either way.
 * The reference has no such switch (the tau bound of
 * tau_leaping/CSRMatrixPacked.cu:311-358 always takes the entries as read).
 * TL_ERR_STOICH_RANGE if a net coefficient leaves [-3, 12]. */
tl_status tl_model_set_net_stoichiometry(tl_model *m, int32_t on);
int32_t tl_model_net_stoichiometry(const tl_model *m);
/* Entries of the packed layouts: tl_model_nnz, or fewer with net stoichiometry. */
int32_t tl_model_layout_nnz(const tl_model *m);
/* Derived structure (any pointer may be NULL): order[M], s0[M], s1[M],
 * species-major row_ptr[N+1] and packed entries
 * csr[tl_model_layout_nnz] = (channel<<4)|(v+3).  order: 0 source, 1 S, 2 Sa+Sb,
 * 3 2S, 4 Sa+Sb+Sc, 5 2Sa+Sb, 6 3S (the reactant side of the channel). */
tl_status tl_model_structure(const tl_model *m, int8_t *order, int32_t *s0, int32_t *s1,
                             int32_t *row_ptr, int32_t *csr);

/* ---- simulation ---- */
typedef struct tl_sim_params {
    double t_final;             /* -t */
    double eps;                 /* -e (default 0.03) */
    int64_t n_realizations;     /* -n */
    uint64_t seed;              /* RNG key; realization r uses (seed, realization_offset + r) */
    int64_t max_iter;           /* -a; <= 0: unlimited */
    int32_t precision;          /* TL_FP64 / TL_FP32 propensities; time is always FP64 */
    int32_t n_track;            /* -r; 0 with track == NULL: every species */
    const int32_t *track;       /* [n_track] species indices */
    int32_t n_intervals;        /* -i; grid t_k = k*t_final/n_intervals, k = 0..n_intervals */
    int32_t sensitivities;      /* accumulate likelihood-ratio scores, fill dmean_dc */
    int32_t device;             /* -d */
    int32_t mode;               /* TL_MODE_* */
    int32_t block_size;         /* persistent kernel lanes per realization (64..1024); 0: the default
                                   shape for the model sizes (a pure function of sizes, precision and
                                   kernel build: a seed gives the same trajectories on every run);
                                   TL_BLOCK_TUNE: time the shapes on a sample and keep the fastest */
    uint32_t model_id;          /* RNG stream id of the model within a batch */
    int64_t realization_offset; /* first realization id */
    int32_t poisson;            /* TL_POISSON_* sampler (see the enum) */
} tl_sim_params;

void tl_sim_params_default(tl_sim_params *p);

typedef struct tl_result {
    /* caller-owned outputs; NULL = not requested.  n_grid = n_intervals + 1
     * when n_intervals > 0, else 1 (the state at t_final). */
    double *mean;         /* [n_grid][n_track] */
    double *stddev;       /* [n_grid][n_track] sample standard deviation (n-1) */
    double *grid_times;   /* [n_grid] */
    int32_t *final_x;     /* [R][n_track] tracked species at t_final */
    double *dmean_dc;     /* [n_track][M] dE[X_s(t_final)]/dc_j (needs sensitivities) */
    double *dmean_dc_se;  /* [n_track][M] standard error of dmean_dc */
    int64_t *steps;       /* [R] accepted main-loop steps per realization */
    double *score;        /* [R][M] likelihood-ratio score d log p(path) / dc_j per realization
                             (needs sensitivities) */
    /* filled by the library */
    int64_t total_steps;      /* sum over realizations of accepted steps (leaps + direct blocks) */
    int64_t total_ssa_events; /* direct-method events */
    int64_t total_rejections; /* rejected leaps (negative populations) */
    int64_t n_failed;         /* realizations that ended with a non-OK status */
    int32_t first_failure;    /* status of the first failed realization */
    int32_t block_size;       /* lanes per realization the persistent kernel used */
    double sim_seconds;       /* first launch to results on the host */
    double upload_seconds;    /* one-time constant upload */
    double kernel_ms;         /* device time of the main simulation kernel(s) */
    /* caller-owned per-realization outputs (NULL = not requested) */
    int64_t *rejections;      /* [R] rejected leap attempts (negative populations) */
    int32_t *status;          /* [R] final status of each realization (TL_OK, TL_ERR_ITER_CAP, ...) */
    double *grid;             /* [R][n_grid][n_track] tracked species of each realization on the time
                                 grid (n_intervals > 0; the rows mean / stddev reduce) */
    /* filled by the library */
    int32_t form;             /* the form that ran: TL_MODE_PLAIN or TL_MODE_PERSISTENT */
} tl_result;

tl_status tl_simulate(const tl_model *m, const tl_sim_params *p, tl_result *r);

/* Many models, uneven sizes and realization counts.  params[i] / results[i]
 * belong to models[i]; precision and sensitivities must agree across the
 * batch.  PERSISTENT runs every (model, realization) as work items of
 * persistent work-stealing launches, one per size class of models, the
 * classes concurrent branches of one cached hipGraph; PLAIN runs
 * tl_simulate's plain form model after model. */
tl_status tl_batch_simulate(const tl_model *const *models, int32_t n_models,
                            const tl_sim_params *params, tl_result *results, int32_t mode);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tl_simulate's persistent form on one model, with the
 * default launch shape, writing DEVICE outputs final_x [R][n_track], steps
 * [R] and status [R]; it calls no hipMalloc / hipFree.  The kernel's other
 * device memory (work item, counter, track list, per-realization event and
 * rejection counts) is a 32 MiB array per GPU, held until the process ends:
 * at most 2,097,000 realizations and 256 tracked species per call; no
 * sensitivities, no time grid, no model whose state needs HBM (those return
 * TL_ERR_INVALID_ARG / TL_ERR_TOO_LARGE).  The model's constants must
 * already be on the device (any tl_simulate or tl_stage_* call with the
 * model there uploads them; TL_ERR_INVALID_ARG otherwise).  stream (a
 * hipStream_t, may be a capturing one): the work is enqueued on it and the
 * call returns at once; NULL: the library's stream, and the call waits.
 * Calls on one GPU share its array: overlapping calls race. */
tl_status tl_simulate_noalloc(const tl_model *m, const tl_sim_params *p, int32_t *final_x, int64_t *steps,
                              int32_t *status, void *stream);

/* ---- stages on caller-provided DEVICE buffers (structure of arrays) ----
 * x [N][R] int32, a [M][R] (double or float per precision), crit [M][R]
 * uint8, per-realization scalars [R].  Synchronous with all device work: a
 * stage waits for the work already queued on every stream of the device
 * (the caller's kernels that produce its inputs, or that still use a buffer
 * it writes), runs on the library's stream and returns when it is done. */
tl_status tl_stage_propensities(const tl_model *m, int32_t precision, int64_t R, const int32_t *x,
                                void *a, uint8_t *crit, double *a0, double *a0c);
tl_status tl_stage_tau(const tl_model *m, int32_t precision, int64_t R, double eps,
                       const int32_t *x, const void *a, const uint8_t *crit, double *tau1,
                       int32_t *argmin, int32_t *which);
tl_status tl_stage_select(const tl_model *m, int32_t precision, int64_t R, uint64_t seed,
                          uint32_t model_id, int64_t realization_offset, const uint32_t *epoch,
                          const void *a, const uint8_t *crit, const double *a0c,
                          const double *tau1, const double *t_rem, double *tau, int32_t *jc,
                          uint8_t *clamped);
tl_status tl_stage_fire(const tl_model *m, int32_t precision, int64_t R, uint64_t seed,
                        uint32_t model_id, int64_t realization_offset, const uint32_t *epoch,
                        const void *a, const uint8_t *crit, const double *tau, const int32_t *jc,
                        int32_t *k);
tl_status tl_stage_update(const tl_model *m, int64_t R, const int32_t *x, const int32_t *k,
                          int32_t *x_out, uint8_t *neg);
tl_status tl_stage_ssa(const tl_model *m, int32_t precision, int64_t R, uint64_t seed,
                       uint32_t model_id, int64_t realization_offset, int32_t *x, double *t,
                       uint32_t *epoch, double t_final, int32_t n_events, int32_t *events);

/* ---- misc ---- */
const char *tl_last_error(void);
const char *tl_version(void);
tl_status tl_device_count(int32_t *n);

#ifdef __cplusplus
}
#endif
#endif
