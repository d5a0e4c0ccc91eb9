// tl_internal.hpp -- declarations shared by the host orchestration (tl_api.cpp)
// and the HIP sources (tl_stages.hip, tl_persistent.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "tl_api.h"
#include "tl_device.hpp"

struct tl_model;

namespace tl {

tl_status set_error(tl_status st, const std::string &msg);
void release_device_copies(tl_model &m);

#define TL_HIP(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return ::tl::set_error(TL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device copy of a model's constants (one allocation per device).
struct DeviceModelCache {
    int device = -1;
    uint64_t version = 0;
    void *buf = nullptr;
    size_t bytes = 0;
    DevModel dm{};
};

// Upload (or reuse) the model's constants on the current device.
tl_status device_model(const tl_model &m, int device, DevModel &out, double *upload_seconds);

// Per-device context: the library's stream (default priority), up to three
// more for the branches of a batch graph, and the executable batch graphs
// cached by batch shape.
struct DeviceContext {
    static constexpr int kAux = 3;
    int device = -1;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    size_t lds_per_block = 0;
    hipStream_t aux[kAux] = {};
    hipEvent_t fork = nullptr, join[kAux] = {};
    std::mutex graph_mu;
    std::map<std::string, hipGraphExec_t> graphs;
};
tl_status device_context(int device, DeviceContext *&ctx);

// Grow-only device scratch owned by the context.
tl_status scratch(DeviceContext &ctx, int slot, size_t bytes, void *&ptr);

// ---------------------------------------------------------------- plain --
// Stage kernels over structure-of-arrays state ([element][R]); launched on
// ctx.stream.  is64 selects double / float propensities.
void launch_prop(bool is64, const DevModel &m, int64_t R, int32_t nchunk, const int32_t *x, void *a,
                 uint8_t *crit, double *a0, double *a0c, double *part, double *partc, const uint8_t *active,
                 hipStream_t s);
// nrc > 1 with scratch pbest [nrc][R] (propensity type) and pidx [nrc][R]:
// the species minimum in two passes over nrc row chunks.  With pwhich
// [nrc][R] as well, the species sums and each chunk's minimum are one pass
// and tau_i / which_i are not used (may be null).
void launch_tau(bool is64, const DevModel &m, int64_t R, double eps, const int32_t *x, const void *a,
                const uint8_t *crit, void *tau_i, uint8_t *which_i, double *tau1, int32_t *argmin,
                int32_t *which, const uint8_t *active, hipStream_t s, int32_t nrc = 1, void *pbest = nullptr,
                int32_t *pidx = nullptr, uint8_t *pwhich = nullptr);
void launch_select(bool is64, const DevModel &m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                   const uint32_t *epoch, const void *a, const uint8_t *crit, const double *a0c,
                   const double *tau1, const double *t_rem, double *tau, int32_t *jc, uint8_t *clamped,
                   hipStream_t s);
void launch_fire(bool is64, const DevModel &m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                 const uint32_t *epoch, const void *a, const uint8_t *crit, const double *tau, const int32_t *jc,
                 int32_t *k, const uint8_t *mode, int pflags, hipStream_t s);
void launch_update(const DevModel &m, int64_t R, const int32_t *x, const int32_t *k, int32_t *x_out, uint8_t *neg,
                   const uint8_t *mode, hipStream_t s);
void launch_ssa(bool is64, const DevModel &m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                int32_t *x, double *t, uint32_t *epoch, double t_final, int32_t n_events, int32_t *events,
                void *a, hipStream_t s);
void launch_fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s);
void launch_fill_u64(unsigned long long *p, unsigned long long v, int64_t n, hipStream_t s);

// Per-realization bookkeeping of the plain loop ([R] arrays).
struct PlainState {
    double *t, *tau1_used, *override_tau1, *tau;
    uint32_t *epoch;
    int64_t *steps, *ssa, *rej;
    int32_t *status, *retries, *jc;
    uint8_t *active, *mode, *clamped, *neg;
    unsigned long long *n_done;
    // sensitivities (null without): score S and dt/dc D, [M][R] FP64
    double *S = nullptr, *D = nullptr;
    // time grid (grid null without): the tracked species [n_track] of each
    // realization at k * t_final / n_intervals into grid [R][n_intervals+1]
    // [n_track]; next_g [R] is the next grid index, snap [n_track][R] the
    // tracked state before a direct-method block
    double *grid = nullptr;
    int32_t *next_g = nullptr, *snap = nullptr;
    const int32_t *track = nullptr;
    int32_t n_track = 0, n_intervals = 0;
    // models with many channels (dlist non-null): the select stage only lists
    // the realizations that take a direct-method block (dlist [R], dcount
    // cleared before each select), and launch_plain_direct runs each block
    // with one workgroup
    int64_t *dlist = nullptr;
    unsigned long long *dcount = nullptr;
};
// Channel count from which the plain form runs direct-method blocks one
// workgroup per realization (below it, one thread per realization).
constexpr int32_t kPlainDirectGroupM = 128;
void launch_plain_direct(bool is64, const DevModel &m, int64_t R, uint64_t seed, uint32_t model_id, int64_t rbase,
                         double t_final, int32_t *x, void *a, PlainState st, hipStream_t s);
void launch_plain_init(const DevModel &m, int64_t R, int32_t *x, PlainState st, hipStream_t s);
void launch_plain_select(bool is64, const DevModel &m, int64_t R, uint64_t seed, uint32_t model_id,
                         int64_t rbase, double t_final, int64_t max_iter, int32_t *x, void *a, const uint8_t *crit,
                         const double *a0, const double *a0c, const double *tau1, PlainState st, int32_t nchunk,
                         const double *partc, hipStream_t s);
// Grid points passed by the accepted leaps of this iteration (before the
// commit overwrites x with x_new).
void launch_plain_grid(const DevModel &m, int64_t R, const int32_t *x, const int32_t *x_new, PlainState st,
                       double t_final, hipStream_t s);
void launch_plain_commit(const DevModel &m, int64_t R, int32_t *x, const int32_t *x_new, PlainState st,
                         double t_final, hipStream_t s);
// Likelihood-ratio terms of the accepted leaps of this iteration (after the
// commit): qpart [nchunk][R] scratch; argmin / which of the tau stage.
void launch_plain_score(bool is64, const DevModel &m, int64_t R, int32_t nchunk, const void *a, const uint8_t *crit,
                        const int32_t *k, const int32_t *argmin, const int32_t *which, double *qpart, PlainState st,
                        bool exact, hipStream_t s);
// Grid points at and after each realization's final time take its final state.
void launch_plain_grid_final(const DevModel &m, int64_t R, const int32_t *x, PlainState st, double t_final,
                             hipStream_t s);
void launch_gather_track(int64_t R, int32_t n_track, const int32_t *track, const int32_t *x, int32_t *out,
                         hipStream_t s);
// out [R][M] = in [M][R]
void launch_transpose_f64(int64_t R, int32_t M, const double *in, double *out, hipStream_t s);
void launch_fill_f64(double *p, double v, int64_t n, hipStream_t s);

// ----------------------------------------------------------- persistent --
// One work item = one realization of one model.
struct ModelItem {
    DevModel dm;
    double t_final, eps;
    uint64_t seed;
    int64_t max_iter;
    int64_t item_base;   // first work-item index of this model
    int64_t R;
    int64_t rid_offset;  // realization id of item_base
    uint32_t model_id;
    int32_t n_track;
    const int32_t *track;  // [n_track]
    int32_t *out_x;        // [R][n_track]
    int64_t *out_steps, *out_ssa, *out_rej;  // [R]
    int32_t *out_status;   // [R]
    double *score;         // [R][M] (sensitivities) or null
    double *dtime;         // [R][M] dt/dc_j scratch (sensitivities) or null
    double *tlast;         // [R][M] time up to which S_j holds its direct-event exposure (sensitivities) or null
    int32_t n_intervals;   // > 0: record tracked species on the grid k*t_final/n_intervals
    double *grid;          // [R][n_intervals+1][n_track]
    int32_t poisson;       // sampler flags kPoissonExact | kPoissonFast (tl_sim_params.poisson)
};

struct PersistentConfig {
    int block;     // threads per block
    int group;     // lanes per realization (64 = one wave, == block = whole block)
    bool a_lds;    // propensities in LDS (else recomputed where they are read)
    int wpe;       // waves per SIMD the kernel is compiled for (register budget 512 / wpe)
    int tile_grains = 1;  // block groups without a in LDS: propensity tile of tile_grains * 4096 channels
    bool hbm = false;     // group state in global memory (models whose state does not fit LDS)
    bool ext = false;     // kernels with three-reactant channels and the exact sampler compiled in
    int occ = 0;          // resident blocks per CU (filled by the candidate list)
};

struct PersistentLaunch {
    const ModelItem *items;  // device array [n_models]
    int32_t n_models;
    int64_t n_items;
    int32_t n_max, m_max;
    unsigned long long *counter;
    int64_t slots;
    int grid;
    size_t lds_bytes;  // per block
    int tile_grains;   // see PersistentConfig
    unsigned char *state = nullptr;  // hbm: grid slices of state_stride bytes
    size_t state_stride = 0;
};

// LDS bytes per block for a config; 0 if it can not fit.
size_t persistent_lds_bytes(const PersistentConfig &c, bool is64, int32_t n_max, int32_t m_max);
// Global-memory state bytes per block (hbm configs; 0 otherwise).
size_t persistent_state_bytes(const PersistentConfig &c, bool is64, int32_t n_max, int32_t m_max);
// Max resident blocks per CU for the config (0 if unsupported).
int persistent_occupancy(const PersistentConfig &c, bool is64, bool sens, size_t lds_bytes);
hipError_t launch_persistent(const PersistentConfig &c, bool is64, bool sens, const PersistentLaunch &L, hipStream_t s);

// Host helpers shared by the entry points (tl_api.cpp).
tl_status check_sim_params(const tl_model *m, const tl_sim_params *p);
bool sim_needs_ext(const tl_model *m, const tl_sim_params *p);
std::vector<int32_t> sim_track_list(const tl_model *m, const tl_sim_params *p);
// The default persistent launch shape of a size class (a pure function of
// the sizes, precision and kernel build); TL_ERR_TOO_LARGE if none fits.
tl_status default_shape(DeviceContext &dc, bool is64, bool sens, bool ext, int32_t n_max, int32_t m_max,
                        PersistentConfig &cfg);
PersistentLaunch launch_plan(DeviceContext &dc, const PersistentConfig &cfg, bool is64, bool sens, ModelItem *d_items,
                             int32_t n_models, int64_t n_items, int32_t n_max, int32_t m_max,
                             unsigned long long *d_counter);
// The model's device constants if they are resident and current (no upload).
tl_status device_model_resident(const tl_model &m, int device, DevModel &out);
// ---------------------------------------------------------- reductions --
// mean / sample standard deviation (n-1) of every column of v [R][C] (either
// output may be null)
void launch_moments_i32(int64_t R, int64_t C, const int32_t *v, double *mean, double *sd, hipStream_t s);
void launch_moments_f64(int64_t R, int64_t C, const double *v, double *mean, double *sd, hipStream_t s);
// dmean[q][j] = mean_r (X_q - Xbar_q) S_j and its standard error from
// x_track [R][n_track] and score [R][M]; mean_scratch [n_track]; part
// 2 * RS * n_track * M doubles, RS = grad_slices(R, n_track, M)
int32_t grad_slices(int64_t R, int32_t n_track, int32_t M);
void launch_grad_reduce(int64_t R, int32_t n_track, int32_t M, const int32_t *x_track, const double *score,
                        double *mean_scratch, double *part, int32_t RS, double *dmean, double *dmean_se,
                        hipStream_t s);

}  // namespace tl
