// tl_api.cpp -- host orchestration behind the C ABI: device contexts, one-time
// constant upload, the plain per-model launch loop, the persistent batch
// (captured in a hipGraph), the launch-shape autotuner, moments and the
// stage entry points.
//
// Replaces the reference's driver loop (tau_leaping/tau_leaping.cpp:127-150:
// Clone n-1 times, InitSeeds, PerformCalc per realization) and the statistics
// stage (tau_leaping/statistics.cpp:15-111).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <cstdlib>
#include <vector>

#include "tl_internal.hpp"
#include "tl_model.hpp"

namespace tl {

// --------------------------------------------------------------- errors --
static thread_local std::string g_last_error;

tl_status set_error(tl_status st, const std::string &msg) {
    g_last_error = msg;
    return st;
}

// ------------------------------------------------------------- contexts --
namespace {

struct Ctx {
    DeviceContext dc;
    std::vector<std::pair<void *, size_t>> scratch;
};

std::mutex g_ctx_mu;
std::map<int, std::unique_ptr<Ctx>> g_ctx;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

tl_status device_context(int device, DeviceContext *&out) {
    std::lock_guard<std::mutex> g(g_ctx_mu);
    auto it = g_ctx.find(device);
    if (it != g_ctx.end()) {
        TL_HIP(hipSetDevice(device));
        out = &it->second->dc;
        return TL_OK;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return set_error(TL_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n)
        return set_error(TL_ERR_INVALID_ARG, "device " + std::to_string(device) + " out of range");
    TL_HIP(hipSetDevice(device));
    auto c = std::make_unique<Ctx>();
    c->dc.device = device;
    TL_HIP(hipStreamCreateWithFlags(&c->dc.stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    TL_HIP(hipGetDeviceProperties(&prop, device));
    c->dc.n_cu = prop.multiProcessorCount;
    c->dc.lds_per_block = prop.sharedMemPerBlock;
    out = &c->dc;
    g_ctx[device] = std::move(c);
    return TL_OK;
}

tl_status scratch(DeviceContext &dc, int slot, size_t bytes, void *&ptr) {
    std::lock_guard<std::mutex> g(g_ctx_mu);
    Ctx &c = *g_ctx.at(dc.device);
    if ((int)c.scratch.size() <= slot) c.scratch.resize(slot + 1, {nullptr, 0});
    auto &s = c.scratch[slot];
    if (s.second < bytes) {
        if (s.first) TL_HIP(hipFree(s.first));
        s.first = nullptr;
        s.second = 0;
        TL_HIP(hipMalloc(&s.first, bytes));
        s.second = bytes;
    }
    ptr = s.first;
    return TL_OK;
}

// -------------------------------------------------------- model upload --
void release_device_copies(tl_model &m) {
    for (auto &kv : m.dev) {
        if (kv.second->buf) {
            hipSetDevice(kv.first);
            hipFree(kv.second->buf);
        }
        delete kv.second;
    }
    m.dev.clear();
}

tl_status device_model(const tl_model &m, int device, DevModel &out, double *upload_seconds) {
    std::lock_guard<std::mutex> g(m.dev_mu);
    auto it = m.dev.find(device);
    if (it != m.dev.end() && it->second->version == m.version) {
        out = it->second->dm;
        if (upload_seconds) *upload_seconds = 0.0;
        return TL_OK;
    }
    const double t0 = now_s();
    // one allocation: c64 | c32 | s0 | s1 | row_ptr | csr | col_ptr | csc | x0 | order | g_order | rs |
    // dep_ptr | dep | colq | aff_ptr | aff | tile_ptr | aff_rs | aff_c64 | aff_c32 | s2 | row_order
    const size_t M = (size_t)m.M, N = (size_t)m.N, Z = (size_t)std::max(m.lnnz, 1);
    std::vector<size_t> off;
    size_t total = 0;
    auto add = [&](size_t bytes) {
        off.push_back(total);
        total += (bytes + 255) & ~(size_t)255;
    };
    add(M * 8);
    add(M * 4);
    add(M * 4);
    add(M * 4);
    add((N + 1) * 4);
    add(Z * 4);
    add((M + 1) * 4);
    add(Z * 4);
    add(N * 4);
    add(M);
    add(N);
    add(M * 8);  // rs: (s0 | order << 29, s1)
    const size_t ND = std::max<size_t>(m.dep.size(), 1);
    add((N + 1) * 4);  // dep_ptr
    add(ND * 4);       // dep
    add(std::max<size_t>(m.colq.size(), 1) * 4);     // colq
    add((size_t)std::max<size_t>(m.aff_ptr.size(), 1) * 4);  // aff_ptr
    add(std::max<size_t>(m.aff.size(), 1) * 4);      // aff
    add(m.tile_ptr.size() * 4);                       // tile_ptr
    const size_t NA = std::max<size_t>(m.aff.size(), 1);
    add(NA * 8);  // aff_rs
    add(NA * 8);  // aff_c64
    add(NA * 4);  // aff_c32
    add(std::max<size_t>(m.s2.size(), 1) * 4);  // s2 (three-reactant channels)
    add(N * 4);  // row_order
    std::vector<unsigned char> host(total, 0);
    bool rows_uneven = false;
    std::vector<float> c32(M);
    for (size_t j = 0; j < M; ++j) c32[j] = (float)m.rates[j];
    memcpy(&host[off[0]], m.rates.data(), M * 8);
    memcpy(&host[off[1]], c32.data(), M * 4);
    memcpy(&host[off[2]], m.s0.data(), M * 4);
    memcpy(&host[off[3]], m.s1.data(), M * 4);
    memcpy(&host[off[4]], m.row_ptr.data(), (N + 1) * 4);
    if (m.lnnz) memcpy(&host[off[5]], m.csr.data(), (size_t)m.lnnz * 4);
    memcpy(&host[off[6]], m.lcol_ptr.data(), (M + 1) * 4);
    if (m.lnnz) memcpy(&host[off[7]], m.csc.data(), (size_t)m.lnnz * 4);
    memcpy(&host[off[8]], m.x0.data(), N * 4);
    memcpy(&host[off[9]], m.order.data(), M);
    memcpy(&host[off[10]], m.g_order.data(), N);
    {
        int32_t *rs = reinterpret_cast<int32_t *>(&host[off[11]]);
        for (size_t j = 0; j < M; ++j) {
            rs[2 * j] = m.s0[j] | ((int32_t)m.order[j] << 29);
            rs[2 * j + 1] = m.s1[j];
        }
    }
    {
        // species rows longest first: the block groups' row gather gives each
        // pass's lanes rows of similar length, so a wavefront waits less on
        // its longest row (each row is still summed by the same lanes in the
        // same order)
        int32_t *ro = reinterpret_cast<int32_t *>(&host[off[22]]);
        for (size_t i = 0; i < N; ++i) ro[i] = (int32_t)i;
        std::stable_sort(ro, ro + N, [&](int32_t a, int32_t b) {
            return m.row_ptr[a + 1] - m.row_ptr[a] > m.row_ptr[b + 1] - m.row_ptr[b];
        });
        rows_uneven = N > 0 && m.row_ptr[ro[0] + 1] - m.row_ptr[ro[0]] != m.row_ptr[ro[N - 1] + 1] - m.row_ptr[ro[N - 1]];
    }
    memcpy(&host[off[12]], m.dep_ptr.data(), (N + 1) * 4);
    if (!m.dep.empty()) memcpy(&host[off[13]], m.dep.data(), m.dep.size() * 4);
    if (!m.colq.empty()) memcpy(&host[off[14]], m.colq.data(), m.colq.size() * 4);
    if (!m.aff_ptr.empty()) memcpy(&host[off[15]], m.aff_ptr.data(), m.aff_ptr.size() * 4);
    if (!m.aff.empty()) memcpy(&host[off[16]], m.aff.data(), m.aff.size() * 4);
    memcpy(&host[off[17]], m.tile_ptr.data(), m.tile_ptr.size() * 4);
    {
        int32_t *ars = reinterpret_cast<int32_t *>(&host[off[18]]);
        double *ac64 = reinterpret_cast<double *>(&host[off[19]]);
        float *ac32 = reinterpret_cast<float *>(&host[off[20]]);
        for (size_t d = 0; d < m.aff.size(); ++d) {
            const int32_t k = m.aff[d];
            ars[2 * d] = m.s0[k] | ((int32_t)m.order[k] << 29);
            ars[2 * d + 1] = m.s1[k];
            ac64[d] = m.rates[k];
            ac32[d] = (float)m.rates[k];
        }
    }

    DeviceModelCache *c;
    if (it != m.dev.end()) {
        c = it->second;
    } else {
        c = new DeviceModelCache();
        c->device = device;
        m.dev[device] = c;
    }
    if (!m.s2.empty()) memcpy(&host[off[21]], m.s2.data(), M * 4);
    if (c->bytes < total) {
        if (c->buf) TL_HIP(hipFree(c->buf));
        c->buf = nullptr;
        c->bytes = 0;
        TL_HIP(hipMalloc(&c->buf, total));
        c->bytes = total;
    }
    TL_HIP(hipMemcpy(c->buf, host.data(), total, hipMemcpyHostToDevice));
    unsigned char *b = (unsigned char *)c->buf;
    DevModel &d = c->dm;
    d.N = m.N;
    d.M = m.M;
    d.nnz = m.lnnz;
    d.c64 = (const double *)(b + off[0]);
    d.c32 = (const float *)(b + off[1]);
    d.s0 = (const int32_t *)(b + off[2]);
    d.s1 = (const int32_t *)(b + off[3]);
    d.row_ptr = (const int32_t *)(b + off[4]);
    d.csr = (const int32_t *)(b + off[5]);
    d.col_ptr = (const int32_t *)(b + off[6]);
    d.csc = (const int32_t *)(b + off[7]);
    d.x0 = (const int32_t *)(b + off[8]);
    d.order = (const int8_t *)(b + off[9]);
    d.g_order = (const int8_t *)(b + off[10]);
    d.rs = (const int2 *)(b + off[11]);
    d.dep_ptr = (const int32_t *)(b + off[12]);
    d.dep = (const int32_t *)(b + off[13]);
    d.colq = m.colq.empty() ? nullptr : (const int4 *)(b + off[14]);
    d.aff_ptr = m.aff_ptr.empty() ? nullptr : (const int32_t *)(b + off[15]);
    d.aff = m.aff_ptr.empty() ? nullptr : (const int32_t *)(b + off[16]);
    d.tile_ptr = (const int32_t *)(b + off[17]);
    d.aff_rs = (const int2 *)(b + off[18]);
    d.aff_c64 = (const double *)(b + off[19]);
    d.aff_c32 = (const float *)(b + off[20]);
    d.s2 = m.s2.empty() ? nullptr : (const int32_t *)(b + off[21]);
    d.row_order = rows_uneven ? (const int32_t *)(b + off[22]) : nullptr;
    c->version = m.version;
    out = d;
    if (upload_seconds) *upload_seconds = now_s() - t0;
    return TL_OK;
}

namespace {

// RAII device buffers for one call.
struct DevBufs {
    std::vector<void *> ptrs;
    ~DevBufs() {
        for (void *p : ptrs) hipFree(p);
    }
    template <class T>
    tl_status alloc(T *&p, size_t n) {
        void *v = nullptr;
        hipError_t e = hipMalloc(&v, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return set_error(TL_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(n * sizeof(T)) +
                                               " bytes): " + hipGetErrorString(e));
        }
        ptrs.push_back(v);
        p = (T *)v;
        return TL_OK;
    }
};

#define TL_TRY(expr)                      \
    do {                                  \
        tl_status st_ = (expr);           \
        if (st_ != TL_OK) return st_;     \
    } while (0)

// A pair of timing events, destroyed on every return path.
struct Events {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    tl_status create() {
        TL_HIP(hipEventCreate(&e0));
        TL_HIP(hipEventCreate(&e1));
        return TL_OK;
    }
    ~Events() {
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
    }
};

tl_status check_params(const tl_model *m, const tl_sim_params *p) {
    if (!m || !p) return set_error(TL_ERR_INVALID_ARG, "null model or params");
    if (!(p->t_final > 0.0) || !std::isfinite(p->t_final)) return set_error(TL_ERR_INVALID_ARG, "t_final must be > 0");
    if (!(p->eps > 0.0)) return set_error(TL_ERR_INVALID_ARG, "eps must be > 0");
    if (p->n_realizations < 1) return set_error(TL_ERR_INVALID_ARG, "n_realizations must be >= 1");
    if (p->precision != TL_FP64 && p->precision != TL_FP32) return set_error(TL_ERR_INVALID_ARG, "bad precision");
    if (p->n_track < 0 || p->n_track > m->N) return set_error(TL_ERR_INVALID_ARG, "n_track out of range");
    if (p->n_track > 0 && !p->track) return set_error(TL_ERR_INVALID_ARG, "track is NULL");
    for (int32_t q = 0; q < p->n_track; ++q)
        if (p->track[q] < 0 || p->track[q] >= m->N) return set_error(TL_ERR_INVALID_ARG, "tracked species out of range");
    if (p->n_intervals < 0) return set_error(TL_ERR_INVALID_ARG, "n_intervals must be >= 0");
    if (p->model_id >= (1u << 30)) return set_error(TL_ERR_INVALID_ARG, "model_id must be < 2^30");
    if (p->poisson < TL_POISSON_NORMAL || p->poisson > TL_POISSON_FAST_EXACT)
        return set_error(TL_ERR_INVALID_ARG, "poisson must be one of TL_POISSON_NORMAL, _EXACT, _FAST, _FAST_EXACT");
    return TL_OK;
}

// Whether a run needs the kernels with sources, three-reactant channels and
// the exact sampler compiled in (s2 is present exactly when the model has a
// source or a three-reactant channel).
bool needs_ext(const tl_model *m, const tl_sim_params *p) {
    return !m->s2.empty() || p->poisson != TL_POISSON_NORMAL;
}

std::vector<int32_t> track_list(const tl_model *m, const tl_sim_params *p) {
    std::vector<int32_t> t;
    if (p->n_track > 0)
        t.assign(p->track, p->track + p->n_track);
    else
        for (int32_t i = 0; i < m->N; ++i) t.push_back(i);
    return t;
}

void fill_summary(tl_result *res, const std::vector<int64_t> &steps, const std::vector<int64_t> &ssa,
                  const std::vector<int64_t> &rej, const std::vector<int32_t> &status) {
    res->total_steps = 0;
    res->total_ssa_events = 0;
    res->total_rejections = 0;
    res->n_failed = 0;
    res->first_failure = TL_OK;
    for (size_t r = 0; r < steps.size(); ++r) {
        res->total_steps += steps[r];
        res->total_ssa_events += ssa[r];
        res->total_rejections += rej[r];
        if (status[r] != TL_OK) {
            if (res->n_failed == 0) res->first_failure = status[r];
            res->n_failed++;
        }
    }
    if (res->steps) std::copy(steps.begin(), steps.end(), res->steps);
    if (res->rejections) std::copy(rej.begin(), rej.end(), res->rejections);
    if (res->status) std::copy(status.begin(), status.end(), res->status);
}

// The moments of a run on the device (mean and sample standard deviation of
// each tracked species, at t_final or per grid point: StepStat2StdDev,
// statistics.cpp:67-87), then the copies the caller asked for: the moments,
// final_x [R][T] and grid [R][n_grid][T].  Asynchronous on s.
tl_status finish_outputs(DevBufs &B, hipStream_t s, int64_t R, int32_t T, int32_t n_grid, double t_final,
                         const int32_t *d_outx, const double *d_grid, tl_result *res) {
    if (!res) return TL_OK;
    const int64_t C = (int64_t)n_grid * T;
    if (res->mean || res->stddev) {
        double *mom;
        TL_TRY(B.alloc(mom, (size_t)2 * C));
        if (d_grid)
            launch_moments_f64(R, C, d_grid, mom, mom + C, s);
        else
            launch_moments_i32(R, T, d_outx, mom, mom + C, s);
        if (res->mean) TL_HIP(hipMemcpyAsync(res->mean, mom, C * 8, hipMemcpyDeviceToHost, s));
        if (res->stddev) TL_HIP(hipMemcpyAsync(res->stddev, mom + C, C * 8, hipMemcpyDeviceToHost, s));
    }
    if (res->final_x) TL_HIP(hipMemcpyAsync(res->final_x, d_outx, (size_t)R * T * 4, hipMemcpyDeviceToHost, s));
    if (res->grid && d_grid) TL_HIP(hipMemcpyAsync(res->grid, d_grid, (size_t)R * C * 8, hipMemcpyDeviceToHost, s));
    if (res->grid_times) {
        if (d_grid)
            for (int32_t g = 0; g < n_grid; ++g) res->grid_times[g] = t_final / (n_grid - 1) * g;
        else
            res->grid_times[0] = t_final;
    }
    return TL_OK;
}

// ------------------------------------------------------------ plain loop --
tl_status run_plain(const tl_model *m, const tl_sim_params *p, tl_result *res) {
    DeviceContext *dc;
    TL_TRY(device_context(p->device, dc));
    hipStream_t s = dc->stream;
    DevModel dm;
    double up = 0.0;
    TL_TRY(device_model(*m, p->device, dm, &up));
    const bool is64 = p->precision == TL_FP64;
    const bool sens = p->sensitivities != 0;
    const bool grid = p->n_intervals > 0;
    const int64_t R = p->n_realizations;
    const size_t rs = is64 ? 8 : 4;
    const std::vector<int32_t> trk = track_list(m, p);
    const int32_t T = (int32_t)trk.size();
    const int32_t n_grid = grid ? p->n_intervals + 1 : 1;
    if (grid && (double)R * n_grid * T * 8.0 > 64e9) return set_error(TL_ERR_TOO_LARGE, "time-grid output exceeds 64 GB");
    // channel chunks so that the propensity pass has >= ~256k threads, and at
    // least ~sqrt(M) of them (the critical pick walks chunk sums, then one
    // chunk's channels) while the partial sums stay within 4M entries; at
    // most 65535 (one grid row each)
    const int64_t by_threads = std::max<int64_t>(1, (262144 + R - 1) / R);
    const int64_t by_walk = std::min<int64_t>((int64_t)std::ceil(std::sqrt((double)m->M)), std::max<int64_t>(1, 4194304 / R));
    int32_t nchunk = (int32_t)std::min<int64_t>(std::min<int64_t>(std::max(by_threads, by_walk), std::max(1, m->M / 64)),
                                                65535);
    DevBufs B;
    int32_t *x, *xn, *k, *d_track, *d_outx;
    void *a, *tau_i;
    uint8_t *crit, *which_i;
    double *a0, *a0c, *tau1, *part = nullptr, *partc = nullptr;
    TL_TRY(B.alloc(x, (size_t)m->N * R));
    TL_TRY(B.alloc(xn, (size_t)m->N * R));
    TL_TRY(B.alloc(k, (size_t)m->M * R));
    // species-row chunks for the tau minimum: with >= ~1M threads over at
    // least 64 rows each, the species sums and each chunk's minimum are one
    // pass and no per-species bound is stored; otherwise the bounds are
    // stored and combined over row chunks of >= ~256k threads
    const int32_t row_cap = std::min(65535, std::max(1, m->N / 64));
    const int32_t nrc_fused = (int32_t)std::min<int64_t>((1048576 + R - 1) / R, row_cap);
    const bool fused = nrc_fused > 1 && (int64_t)nrc_fused * R >= 262144;
    const int32_t nrc =
        fused ? nrc_fused : (int32_t)std::min<int64_t>(std::max<int64_t>(1, (262144 + R - 1) / R), row_cap);
    unsigned char *araw, *tauraw = nullptr;
    TL_TRY(B.alloc(araw, (size_t)m->M * R * rs));
    a = araw;
    which_i = nullptr;
    if (!fused) {
        TL_TRY(B.alloc(tauraw, (size_t)m->N * R * rs));
        TL_TRY(B.alloc(which_i, (size_t)m->N * R));
    }
    tau_i = tauraw;
    TL_TRY(B.alloc(crit, (size_t)m->M * R));
    TL_TRY(B.alloc(a0, R));
    TL_TRY(B.alloc(a0c, R));
    TL_TRY(B.alloc(tau1, R));
    if (nchunk > 1) {
        TL_TRY(B.alloc(part, (size_t)nchunk * R));
        TL_TRY(B.alloc(partc, (size_t)nchunk * R));
    }
    unsigned char *pbest = nullptr;
    int32_t *pidx = nullptr;
    uint8_t *pwhich = nullptr;
    if (nrc > 1) {
        TL_TRY(B.alloc(pbest, (size_t)nrc * R * rs));
        TL_TRY(B.alloc(pidx, (size_t)nrc * R));
        if (fused) TL_TRY(B.alloc(pwhich, (size_t)nrc * R));
    }
    TL_TRY(B.alloc(d_track, T));
    TL_TRY(B.alloc(d_outx, (size_t)T * R));
    PlainState st;
    TL_TRY(B.alloc(st.t, R));
    TL_TRY(B.alloc(st.tau1_used, R));
    TL_TRY(B.alloc(st.override_tau1, R));
    TL_TRY(B.alloc(st.tau, R));
    TL_TRY(B.alloc(st.epoch, R));
    TL_TRY(B.alloc(st.steps, R));
    TL_TRY(B.alloc(st.ssa, R));
    TL_TRY(B.alloc(st.rej, R));
    TL_TRY(B.alloc(st.status, R));
    TL_TRY(B.alloc(st.retries, R));
    TL_TRY(B.alloc(st.jc, R));
    TL_TRY(B.alloc(st.active, R));
    TL_TRY(B.alloc(st.mode, R));
    TL_TRY(B.alloc(st.clamped, R));
    TL_TRY(B.alloc(st.neg, R));
    TL_TRY(B.alloc(st.n_done, 1));
    const bool group_direct = m->M >= kPlainDirectGroupM;
    if (group_direct) {
        TL_TRY(B.alloc(st.dlist, R));
        TL_TRY(B.alloc(st.dcount, 1));
    }
    TL_HIP(hipMemcpyAsync(d_track, trk.data(), T * 4, hipMemcpyHostToDevice, s));
    // sensitivities: scores and dt/dc [M][R], the per-chunk sums of a leap's
    // terms, and the tau stage's arg-min species and bound
    double *qpart = nullptr, *d_scoreT = nullptr, *d_mean = nullptr, *d_dmean = nullptr, *d_dse = nullptr;
    int32_t *argmin = nullptr, *which = nullptr;
    if (sens) {
        TL_TRY(B.alloc(st.S, (size_t)m->M * R));
        TL_TRY(B.alloc(st.D, (size_t)m->M * R));
        TL_TRY(B.alloc(qpart, (size_t)nchunk * R));
        TL_TRY(B.alloc(argmin, R));
        TL_TRY(B.alloc(which, R));
        TL_TRY(B.alloc(d_mean, T));
        TL_TRY(B.alloc(d_dmean, (size_t)T * m->M));
        TL_TRY(B.alloc(d_dse, (size_t)T * m->M));
    }
    double *d_grid = nullptr;
    if (grid) {
        TL_TRY(B.alloc(d_grid, (size_t)R * n_grid * T));
        TL_TRY(B.alloc(st.next_g, R));
        TL_TRY(B.alloc(st.snap, (size_t)T * R));
        st.grid = d_grid;
        st.track = d_track;
        st.n_track = T;
        st.n_intervals = p->n_intervals;
    }

    Events ev;
    TL_TRY(ev.create());
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    const double t0 = now_s();
    TL_HIP(hipEventRecord(e0, s));
    launch_fill_u64(st.n_done, 0ull, 1, s);
    launch_plain_init(dm, R, x, st, s);
    if (sens) {
        launch_fill_f64(st.S, 0.0, (int64_t)m->M * R, s);
        launch_fill_f64(st.D, 0.0, (int64_t)m->M * R, s);
    }
    if (grid) launch_fill_u8((uint8_t *)st.next_g, 0, R * 4, s);
    unsigned long long *h_done = nullptr;
    TL_HIP(hipHostMalloc((void **)&h_done, sizeof(unsigned long long)));
    std::unique_ptr<unsigned long long, hipError_t (*)(void *)> done_guard(h_done, hipHostFree);
    *h_done = 0;
    const uint64_t seed = p->seed;
    const int pflags = p->poisson;
    const bool exact = (pflags & kPoissonExact) != 0;
    for (int64_t it = 0;; ++it) {
        launch_prop(is64, dm, R, nchunk, x, a, crit, a0, a0c, part, partc, st.active, s);
        launch_tau(is64, dm, R, p->eps, x, a, crit, tau_i, which_i, tau1, argmin, which, st.active, s, nrc, pbest,
                   pidx, pwhich);
        if (group_direct) launch_fill_u64(st.dcount, 0ull, 1, s);
        launch_plain_select(is64, dm, R, seed, p->model_id, p->realization_offset, p->t_final, p->max_iter, x, a, crit,
                            a0, a0c, tau1, st, nchunk, partc, s);
        if (group_direct)
            launch_plain_direct(is64, dm, R, seed, p->model_id, p->realization_offset, p->t_final, x, a, st, s);
        launch_fire(is64, dm, R, seed, p->model_id, p->realization_offset, st.epoch, a, crit, st.tau, st.jc, k, st.mode,
                    pflags, s);
        launch_update(dm, R, x, k, xn, st.neg, st.mode, s);
        if (grid) launch_plain_grid(dm, R, x, xn, st, p->t_final, s);
        launch_plain_commit(dm, R, x, xn, st, p->t_final, s);
        if (sens) launch_plain_score(is64, dm, R, nchunk, a, crit, k, argmin, which, qpart, st, exact, s);
        // a failed launch would leave the done counter short: stop on it
        TL_HIP(hipGetLastError());
        if ((it & 15) == 15) {
            TL_HIP(hipMemcpyAsync(h_done, st.n_done, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
            TL_HIP(hipStreamSynchronize(s));
            if ((int64_t)*h_done >= R) break;
        }
    }
    TL_HIP(hipEventRecord(e1, s));
    launch_gather_track(R, T, d_track, x, d_outx, s);
    if (grid) launch_plain_grid_final(dm, R, x, st, p->t_final, s);
    const bool want_grad = sens && res && (res->dmean_dc || res->dmean_dc_se);
    if (sens && res && (res->score || want_grad)) {
        TL_TRY(B.alloc(d_scoreT, (size_t)R * m->M));
        launch_transpose_f64(R, m->M, st.S, d_scoreT, s);
        if (want_grad) {
            const int32_t RS = grad_slices(R, T, m->M);
            double *gpart;
            TL_TRY(B.alloc(gpart, (size_t)2 * RS * T * m->M));
            launch_grad_reduce(R, T, m->M, d_outx, d_scoreT, d_mean, gpart, RS, d_dmean, d_dse, s);
        }
    }
    TL_TRY(finish_outputs(B, s, R, T, n_grid, p->t_final, d_outx, d_grid, res));
    std::vector<int32_t> hst(R);
    std::vector<int64_t> hsteps(R), hssa(R), hrej(R);
    TL_HIP(hipMemcpyAsync(hsteps.data(), st.steps, R * 8, hipMemcpyDeviceToHost, s));
    TL_HIP(hipMemcpyAsync(hssa.data(), st.ssa, R * 8, hipMemcpyDeviceToHost, s));
    TL_HIP(hipMemcpyAsync(hrej.data(), st.rej, R * 8, hipMemcpyDeviceToHost, s));
    TL_HIP(hipMemcpyAsync(hst.data(), st.status, R * 4, hipMemcpyDeviceToHost, s));
    if (res && sens) {
        if (res->score) TL_HIP(hipMemcpyAsync(res->score, d_scoreT, (size_t)R * m->M * 8, hipMemcpyDeviceToHost, s));
        if (res->dmean_dc)
            TL_HIP(hipMemcpyAsync(res->dmean_dc, d_dmean, (size_t)T * m->M * 8, hipMemcpyDeviceToHost, s));
        if (res->dmean_dc_se)
            TL_HIP(hipMemcpyAsync(res->dmean_dc_se, d_dse, (size_t)T * m->M * 8, hipMemcpyDeviceToHost, s));
    }
    TL_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (res) {
        res->kernel_ms = ms;
        res->upload_seconds = up;
        res->block_size = 256;
        res->form = TL_MODE_PLAIN;
        fill_summary(res, hsteps, hssa, hrej, hst);
        res->sim_seconds = now_s() - t0;
    }
    return TL_OK;
}

// ------------------------------------------------------ persistent batch --
struct Entry {
    const tl_model *m;
    const tl_sim_params *p;
    tl_result *res;
    DevModel dm;
    std::vector<int32_t> track;
    double cost;
    // device outputs
    int32_t *d_track, *d_outx, *d_status;
    int64_t *d_steps, *d_ssa, *d_rej;
    double *d_score, *d_dtime, *d_tlast, *d_grid, *d_dmean, *d_dse, *d_mean;
};

struct TuneKey {
    bool is64, sens, ext, timed;
    int32_t n_max, m_max;
    bool operator<(const TuneKey &o) const {
        return std::tie(is64, sens, ext, timed, n_max, m_max) <
               std::tie(o.is64, o.sens, o.ext, o.timed, o.n_max, o.m_max);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, PersistentConfig> g_tuned;

// ext: the kernels with three-reactant channels and the exact sampler.
std::vector<PersistentConfig> candidates(const DeviceContext &dc, bool is64, bool sens, int32_t n_max, int32_t m_max,
                                         bool ext) {
    std::vector<PersistentConfig> out;
    const int shapes[][2] = {{64, 64}, {256, 64}, {256, 256}, {512, 512}, {1024, 1024}};
    // one-wavefront groups come in register budgets of 4, 5 and 6 waves per
    // SIMD; the first listed is the default.  Measured on the 128 x 256 toy
    // network without scores: 6 waves 3783 / 3324 ms per launch (FP64 /
    // FP32), 5 waves 3907 / 3554 ms; with scores (more live values) on the
    // heterogeneous batch, 4 waves 70.5 s per batch against 71.7 s with 5
    // (same box), and 6 waves 8-27 % slower per model.
    const int wpe64[] = {6, 5, 4}, wpe64s[] = {4, 5, 6};
    // order within a shape = preference: a[M] in LDS first, except for
    // 1024-lane groups, where recomputed propensities with a large tile win
    // (the 31,250-channel colloid)
    for (auto &sh : shapes)
        for (int ai = 0; ai < 2; ++ai) {
            const int al = (sh[1] >= 1024) ? ai : 1 - ai;
            // 512-lane groups: also compiled for two blocks per CU; 256-lane
            // groups for four per CU (128 VGPRs; the unrestricted build takes
            // about 160 and fits three): measured on the heterogeneous batch's
            // 1800-2924-channel models (FP32 with scores) 15-17 % faster, on
            // the 1024-species cyclic chain 13 % (16 % with scores); five per
            // CU spilled and gained less
            const int wpe512[] = {4, 1}, wpe256[] = {4, 1};
            const int *wl = sh[1] == 64 ? (sens ? wpe64s : wpe64)
                                        : (sh[1] == 512 ? wpe512 : (sh[1] == 256 ? wpe256 : nullptr));
            const int nw = sh[1] == 64 ? 3 : (sh[1] == 512 || sh[1] == 256 ? 2 : 1);
            for (int w = 0; w < nw; ++w) {
                PersistentConfig c{sh[0], sh[1], al != 0, wl ? wl[w] : 1};
                c.ext = ext;
                // 1024-lane groups take the largest propensity tile (4, 2 or
                // 1 grains of 4096 channels) that fits: fewer tiles, fewer
                // barriers; smaller groups keep one grain (two blocks per CU)
                c.tile_grains = (sh[1] >= 1024 && al == 0) ? 4 : 1;
                size_t lds = persistent_lds_bytes(c, is64, n_max, m_max);
                while (lds > dc.lds_per_block && c.tile_grains > 1) {
                    c.tile_grains /= 2;
                    lds = persistent_lds_bytes(c, is64, n_max, m_max);
                }
                if (lds > dc.lds_per_block) continue;
                c.occ = persistent_occupancy(c, is64, sens, lds);
                if (c.occ <= 0) continue;
                out.push_back(c);
            }
        }
    if (out.empty()) {
        // nothing fits LDS: 1024-lane groups with their state in global
        // memory, either with recomputed propensities and the largest LDS
        // tile that fits (one pass over the species rows per tile), or with
        // a[M] stored in global memory too (one pass over the rows).  Measured
        // (2e4 x 1e5 and 2e5 x 1e6 toy networks): recomputed wins at 7 tiles
        // (2.8e5 vs 1.8e5 leaps/s), stored at 62 (8.7e3 vs 5.0e3); the
        // preferred one comes first.
        PersistentConfig c{1024, 1024, false, 1};
        c.hbm = true;
        c.ext = ext;
        c.tile_grains = 4;
        size_t lds = persistent_lds_bytes(c, is64, n_max, m_max);
        while (lds > dc.lds_per_block && c.tile_grains > 1) {
            c.tile_grains /= 2;
            lds = persistent_lds_bytes(c, is64, n_max, m_max);
        }
        const bool c_ok = lds <= dc.lds_per_block && persistent_occupancy(c, is64, sens, lds) > 0;
        PersistentConfig st{1024, 1024, true, 1};
        st.hbm = true;
        st.ext = ext;
        st.tile_grains = std::max(1, (m_max + kTileGrain - 1) / kTileGrain);
        lds = persistent_lds_bytes(st, is64, n_max, m_max);
        const bool st_ok = lds <= dc.lds_per_block && persistent_occupancy(st, is64, sens, lds) > 0;
        const int64_t tiles = ((int64_t)m_max + (int64_t)c.tile_grains * kTileGrain - 1) / ((int64_t)c.tile_grains * kTileGrain);
        const bool stored_first = tiles > 16;
        if (st_ok && stored_first) out.push_back(st);
        if (c_ok) out.push_back(c);
        if (st_ok && !stored_first) out.push_back(st);
    }
    return out;
}

// Default shape without timing: one wavefront per realization up to the
// one-wavefront channel limit below while the LDS (or registers) still hold
// at least 8 such groups per CU, then 256-lane blocks, and 1024 lanes from
// 8192 channels; propensities
// in LDS when they fit.  Measured: a 128 x 256 toy network runs 5x faster on
// one wavefront than on 256-lane blocks, a 300 x 900 one 1.4x; the
// 1024-species cyclic chain (37 KB of LDS per realization: 4 one-wavefront
// groups per CU) 1.9x faster on 256-lane blocks; a 960 x 3000 toy network 1.8x
// faster on 256-lane blocks than on one wavefront (2.6x on 512-lane blocks two
// per CU, but those lose on batches of such models with many direct-method
// events); the 250-species colloid (31,250 channels) 1.8x faster on 1024-lane
// blocks than on 512.
//
// The channel limit of one-wavefront groups depends on the arithmetic: on the
// heterogeneous batch's models (FP32 with scores, 2048 realizations each)
// one wavefront is 1.5x / 1.55x / 1.4x faster than 256-lane blocks at 1109 /
// 1278 / 1411 channels and 1.26x / 1.38x slower at 1674 / 1800; in FP64
// without scores it is 1.63x faster at 1278 channels and 1.21x slower at
// 1411.
int32_t one_wave_channel_limit(bool is64, bool sens) { return (is64 && !sens) ? 1344 : 1536; }

PersistentConfig heuristic(const std::vector<PersistentConfig> &cands, int32_t m_max, int32_t one_wave_limit) {
    auto pick = [&](int b, int g, int min_waves) -> const PersistentConfig * {
        for (auto &c : cands)  // preferred variants come first
            if (c.block == b && c.group == g && c.occ * (b / 64) >= min_waves) return &c;
        return nullptr;
    };
    const PersistentConfig *c = nullptr;
    if (m_max <= one_wave_limit) c = pick(64, 64, 8);
    if (!c && m_max <= one_wave_limit) c = pick(256, 64, 8);
    if (!c && m_max <= 8192) {
        // 256-lane groups in the preferred layout (a[M] in LDS first): the
        // four-per-CU register budget where LDS holds three or more such
        // groups per CU, else the unrestricted build (its extra registers are
        // free there)
        const PersistentConfig *first = pick(256, 256, 0);
        auto variant = [&](int wpe, int min_occ) -> const PersistentConfig * {
            for (auto &cc : cands)
                if (first && cc.block == 256 && cc.group == 256 && cc.a_lds == first->a_lds &&
                    cc.tile_grains == first->tile_grains && cc.wpe == wpe && cc.occ >= min_occ)
                    return &cc;
            return nullptr;
        };
        c = variant(4, 3);
        if (!c) c = variant(1, 0);
        if (!c) c = first;
    }
    if (!c) c = pick(1024, 1024, 0);
    if (!c) c = pick(512, 512, 0);
    if (!c && !cands.empty()) c = &cands.front();
    return c ? *c : PersistentConfig{0, 0, false, 1};
}

PersistentLaunch make_launch(DeviceContext &dc, const PersistentConfig &cfg, bool is64, bool sens,
                             ModelItem *d_items, int32_t n_models, int64_t n_items, int32_t n_max, int32_t m_max,
                             unsigned long long *d_counter, int64_t grid_cap) {
    const size_t lds = persistent_lds_bytes(cfg, is64, n_max, m_max);
    const int occ = std::max(1, persistent_occupancy(cfg, is64, sens, lds));
    const int ng = cfg.block / cfg.group;
    int64_t grid = (int64_t)dc.n_cu * occ;
    grid = std::min<int64_t>(grid, (n_items + ng - 1) / ng);
    if (grid_cap > 0) grid = std::min(grid, grid_cap);
    grid = std::max<int64_t>(grid, 1);
    PersistentLaunch L;
    L.items = d_items;
    L.n_models = n_models;
    L.n_items = n_items;
    L.n_max = n_max;
    L.m_max = m_max;
    L.counter = d_counter;
    L.slots = grid * ng;
    L.grid = (int)grid;
    L.lds_bytes = lds;
    L.tile_grains = cfg.tile_grains;
    L.state_stride = persistent_state_bytes(cfg, is64, n_max, m_max);
    if (cfg.hbm) {
        // at most 32 GB of group state (each block holds one slice)
        const int64_t cap = std::max<int64_t>(1, (int64_t)(32e9 / (double)L.state_stride));
        L.grid = (int)std::min<int64_t>(L.grid, cap);
        L.slots = L.grid;
    }
    return L;
}

// One persistent launch per size class.  With use_graph the counter resets
// and launches of every class are captured into one hipGraph: the classes
// are independent branches on up to four streams (the library's and three
// more, all default priority), so a class's launch fills the CUs that an
// earlier class's tail leaves idle; the first class (the most expensive) is
// dispatched first.  The executable graph is instantiated once per batch
// shape (class configs, grids, LDS) and updated in place with the new
// kernel arguments on later calls, then replayed.  TL_GRAPH=0 in the
// environment runs the same launches on the library's stream without a
// graph (for comparison).
struct ClassLaunch {
    PersistentConfig cfg;
    PersistentLaunch L;
};

std::string graph_key(bool is64, bool sens, const std::vector<ClassLaunch> &cl) {
    std::string k = std::to_string(is64) + "/" + std::to_string(sens);
    for (const auto &c : cl)
        k += "|" + std::to_string(c.cfg.block) + "," + std::to_string(c.cfg.group) + "," + std::to_string(c.cfg.a_lds) +
             "," + std::to_string(c.cfg.wpe) + "," + std::to_string(c.cfg.tile_grains) + "," +
             std::to_string(c.cfg.hbm) + "," + std::to_string(c.cfg.ext) + "," + std::to_string(c.L.grid) + "," +
             std::to_string(c.L.lds_bytes);
    return k;
}

bool graphs_enabled() {
    const char *v = getenv("TL_GRAPH");
    return !(v && v[0] == '0');
}

tl_status run_launches(DeviceContext &dc, bool is64, bool sens, const std::vector<ClassLaunch> &cl, bool use_graph,
                       float *ms_out) {
    hipStream_t s = dc.stream;
    Events ev;
    TL_TRY(ev.create());
    use_graph = use_graph && graphs_enabled();
    if (use_graph) {
        std::lock_guard<std::mutex> g(dc.graph_mu);
        const int nb = std::min<int>((int)cl.size(), 1 + DeviceContext::kAux);
        for (int b = 1; b < nb; ++b)
            if (!dc.aux[b - 1]) {
                TL_HIP(hipStreamCreateWithFlags(&dc.aux[b - 1], hipStreamNonBlocking));
                TL_HIP(hipEventCreateWithFlags(&dc.join[b - 1], hipEventDisableTiming));
            }
        if (!dc.fork) TL_HIP(hipEventCreateWithFlags(&dc.fork, hipEventDisableTiming));
        hipGraph_t gph;
        TL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t le = hipSuccess;
        auto keep = [&](hipError_t e) {
            if (le == hipSuccess) le = e;
        };
        keep(hipEventRecord(dc.fork, s));
        for (int b = 1; b < nb; ++b) keep(hipStreamWaitEvent(dc.aux[b - 1], dc.fork, 0));
        for (size_t i = 0; i < cl.size(); ++i) {
            const int b = (int)(i % nb);
            hipStream_t sb = b == 0 ? s : dc.aux[b - 1];
            launch_fill_u64(cl[i].L.counter, 0ull, 1, sb);
            keep(launch_persistent(cl[i].cfg, is64, sens, cl[i].L, sb));
        }
        for (int b = 1; b < nb; ++b) {
            keep(hipEventRecord(dc.join[b - 1], dc.aux[b - 1]));
            keep(hipStreamWaitEvent(s, dc.join[b - 1], 0));
        }
        TL_HIP(hipStreamEndCapture(s, &gph));
        std::unique_ptr<std::remove_pointer<hipGraph_t>::type, hipError_t (*)(hipGraph_t)> gguard(gph,
                                                                                                 hipGraphDestroy);
        TL_HIP(le);
        const std::string key = graph_key(is64, sens, cl);
        auto it = dc.graphs.find(key);
        hipGraphExec_t gx = nullptr;
        if (it != dc.graphs.end()) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult ur;
            if (hipGraphExecUpdate(it->second, gph, &err_node, &ur) == hipSuccess && ur == hipGraphExecUpdateSuccess) {
                gx = it->second;
            } else {
                (void)hipGetLastError();
                hipGraphExecDestroy(it->second);
                dc.graphs.erase(it);
            }
        }
        if (!gx) {
            TL_HIP(hipGraphInstantiate(&gx, gph, nullptr, nullptr, 0));
            dc.graphs[key] = gx;
        }
        TL_HIP(hipEventRecord(ev.e0, s));
        TL_HIP(hipGraphLaunch(gx, s));
        TL_HIP(hipEventRecord(ev.e1, s));
        TL_HIP(hipStreamSynchronize(s));
    } else {
        TL_HIP(hipEventRecord(ev.e0, s));
        for (const auto &c : cl) {
            launch_fill_u64(c.L.counter, 0ull, 1, s);
            TL_HIP(launch_persistent(c.cfg, is64, sens, c.L, s));
        }
        TL_HIP(hipEventRecord(ev.e1, s));
        TL_HIP(hipStreamSynchronize(s));
    }
    TL_HIP(hipGetLastError());
    float ms = 0.f;
    hipEventElapsedTime(&ms, ev.e0, ev.e1);
    if (ms_out) *ms_out = ms;
    return TL_OK;
}

tl_status launch_batch(DeviceContext &dc, const PersistentConfig &cfg, bool is64, bool sens, ModelItem *d_items,
                       int32_t n_models, int64_t n_items, int32_t n_max, int32_t m_max, unsigned long long *d_counter,
                       int64_t grid_cap, bool use_graph, float *ms_out) {
    if (cfg.hbm) return set_error(TL_ERR_INVALID_ARG, "trial launches of global-memory state shapes are not supported");
    std::vector<ClassLaunch> cl(1);
    cl[0].cfg = cfg;
    cl[0].L = make_launch(dc, cfg, is64, sens, d_items, n_models, n_items, n_max, m_max, d_counter, grid_cap);
    return run_launches(dc, is64, sens, cl, use_graph, ms_out);
}

// The launch shape of a size class.  Default (timed = false): the measured
// preferences of `heuristic`, a pure function of the sizes, precision and
// kernel build, so that a seed gives the same trajectories on every run and
// every box.  timed = true (block_size TL_BLOCK_TUNE): time each launch shape
// on a bounded sample (first few steps of a subset of the work items) and
// keep the fastest in realization-steps per second; shapes that reduce the
// species sums in another order then round differently, so the trajectories
// depend on the timing.
tl_status autotune(DeviceContext &dc, bool is64, bool sens, bool ext, bool timed, const std::vector<ModelItem> &items_in,
                   int64_t n_items, int32_t n_max, int32_t m_max, unsigned long long *d_counter, PersistentConfig &best) {
    TuneKey key{is64, sens, ext, timed, n_max, m_max};
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            best = it->second;
            return TL_OK;
        }
    }
    auto all = candidates(dc, is64, sens, n_max, m_max, ext);
    if (all.empty()) return set_error(TL_ERR_TOO_LARGE, "model too large for any persistent launch shape");
    best = heuristic(all, m_max, one_wave_channel_limit(is64, sens));
    if (best.hbm || !timed) {  // global-memory state shapes are not timed (no trial launches)
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tuned[key] = best;
        return TL_OK;
    }
    // Timed candidates: the default one-wavefront shape and the block shapes.
    // Short trials do not separate one-wavefront variants reliably (their
    // first steps are not representative); the default among them was the
    // fastest measured on full runs.
    std::vector<PersistentConfig> cands;
    for (auto &c : all)
        if (c.group != 64 || (c.block == best.block && c.group == best.group && c.a_lds == best.a_lds &&
                              c.wpe == best.wpe))
            cands.push_back(c);
    const int64_t sample = std::min<int64_t>(n_items, (int64_t)dc.n_cu * 16);
    if (cands.size() > 1 && n_items >= 4 * sample) {
        std::vector<ModelItem> trial = items_in;
        for (auto &mi : trial) mi.max_iter = (mi.max_iter > 0) ? std::min<int64_t>(mi.max_iter, 128) : 128;
        ModelItem *d_trial = nullptr;
        TL_HIP(hipMalloc(&d_trial, trial.size() * sizeof(ModelItem)));
        hipMemcpy(d_trial, trial.data(), trial.size() * sizeof(ModelItem), hipMemcpyHostToDevice);
        // the heuristic shape is kept unless another is clearly faster: the
        // trial covers only the first steps of a sample
        const PersistentConfig first = best;
        double best_rate = -1.0, first_rate = -1.0;
        for (auto &c : cands) {
            float ms = 0.f;
            tl_status st = launch_batch(dc, c, is64, sens, d_trial, (int32_t)trial.size(), sample, n_max, m_max,
                                        d_counter, 0, false, &ms);
            if (st != TL_OK) continue;
            // steps of the sampled items: items 0..sample-1 belong to the first models
            int64_t done = 0;
            int64_t left = sample;
            for (auto &mi : trial) {
                if (left <= 0) break;
                const int64_t take = std::min<int64_t>(left, mi.R);
                std::vector<int64_t> h(take);
                hipMemcpy(h.data(), mi.out_steps, take * 8, hipMemcpyDeviceToHost);
                for (auto v : h) done += v;
                left -= take;
            }
            const double rate = ms > 0 ? (double)done / ms : 0.0;
            if (c.block == first.block && c.group == first.group && c.a_lds == first.a_lds && c.wpe == first.wpe &&
                c.tile_grains == first.tile_grains)
                first_rate = rate;
            if (rate > best_rate) {
                best_rate = rate;
                best = c;
            }
        }
        hipFree(d_trial);
        if (first_rate > 0 && best_rate < 1.15 * first_rate) best = first;
    }
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tuned[key] = best;
    return TL_OK;
}

tl_status run_persistent(const tl_model *const *models, int32_t n_models, const tl_sim_params *params,
                         tl_result *results) {
    const int device = params[0].device;
    const bool is64 = params[0].precision == TL_FP64;
    const bool sens = params[0].sensitivities != 0;
    for (int32_t i = 0; i < n_models; ++i) {
        TL_TRY(check_params(models[i], &params[i]));
        if ((params[i].precision == TL_FP64) != is64 || (params[i].sensitivities != 0) != sens ||
            params[i].device != device)
            return set_error(TL_ERR_INVALID_ARG, "precision, sensitivities and device must agree across the batch");
    }
    DeviceContext *dc;
    TL_TRY(device_context(device, dc));
    hipStream_t s = dc->stream;
    DevBufs B;
    std::vector<Entry> E(n_models);
    double up_total = 0.0;
    for (int32_t i = 0; i < n_models; ++i) {
        Entry &e = E[i];
        e.m = models[i];
        e.p = &params[i];
        e.res = results ? &results[i] : nullptr;
        double up = 0.0;
        TL_TRY(device_model(*e.m, device, e.dm, &up));
        up_total += up;
        e.track = track_list(e.m, e.p);
        e.cost = (double)e.p->n_realizations * (double)(e.m->M + e.m->lnnz);
        const int64_t R = e.p->n_realizations;
        const int32_t T = (int32_t)e.track.size();
        TL_TRY(B.alloc(e.d_track, T));
        TL_TRY(B.alloc(e.d_outx, (size_t)R * T));
        TL_TRY(B.alloc(e.d_steps, R));
        TL_TRY(B.alloc(e.d_ssa, R));
        TL_TRY(B.alloc(e.d_rej, R));
        TL_TRY(B.alloc(e.d_status, R));
        e.d_score = e.d_dtime = e.d_tlast = e.d_grid = e.d_dmean = e.d_dse = e.d_mean = nullptr;
        if (sens) {
            TL_TRY(B.alloc(e.d_score, (size_t)R * e.m->M));
            TL_TRY(B.alloc(e.d_dtime, (size_t)R * e.m->M));
            TL_TRY(B.alloc(e.d_tlast, (size_t)R * e.m->M));
            TL_TRY(B.alloc(e.d_dmean, (size_t)T * e.m->M));
            TL_TRY(B.alloc(e.d_dse, (size_t)T * e.m->M));
            TL_TRY(B.alloc(e.d_mean, T));
        }
        if (e.p->n_intervals > 0) {
            const double bytes = (double)R * (e.p->n_intervals + 1) * T * 8.0;
            if (bytes > 64e9) return set_error(TL_ERR_TOO_LARGE, "time-grid output exceeds 64 GB");
            TL_TRY(B.alloc(e.d_grid, (size_t)R * (e.p->n_intervals + 1) * T));
        }
        TL_HIP(hipMemcpyAsync(e.d_track, e.track.data(), T * 4, hipMemcpyHostToDevice, s));
    }
    // Size classes: models whose one-wavefront LDS footprint is within a
    // factor of two of each other share a launch, so a small model's groups
    // are not laid out (and their occupancy set) for the largest model of the
    // batch; models past the one-wavefront channel limit form classes of
    // their own (their default shape is a block group).  An explicit block
    // size puts every model in one class.
    const int bs_req = params[0].block_size;
    const int32_t ow_limit = one_wave_channel_limit(is64, sens);
    std::map<int, std::vector<int32_t>> classes;
    for (int32_t i = 0; i < n_models; ++i) {
        int cls = 0;
        if (bs_req <= 0) {
            const size_t b = persistent_lds_bytes(PersistentConfig{64, 64, true, 5}, is64, E[i].m->N, E[i].m->M);
            size_t lim = 4096;
            while (b > lim && cls < 12) {
                lim *= 2;
                ++cls;
            }
            if (E[i].m->M > ow_limit) cls += 16;
        }
        classes[cls].push_back(i);
    }
    std::vector<ClassLaunch> launches;
    std::vector<double> class_cost;
    std::vector<PersistentConfig> entry_cfg(n_models);
    for (auto &kv : classes) {
        std::vector<int32_t> &members = kv.second;
        // work order: most expensive models first (better tail balance)
        std::stable_sort(members.begin(), members.end(), [&](int32_t a, int32_t b) { return E[a].cost > E[b].cost; });
        const int32_t nm = (int32_t)members.size();
        std::vector<ModelItem> items(nm);
        int64_t base = 0;
        int32_t n_max = 0, m_max = 0;
        bool ext = false;
        for (int32_t k = 0; k < nm; ++k) {
            Entry &e = E[members[k]];
            n_max = std::max(n_max, e.m->N);
            m_max = std::max(m_max, e.m->M);
            ext |= needs_ext(e.m, e.p);
            ModelItem &mi = items[k];
            mi.dm = e.dm;
            mi.t_final = e.p->t_final;
            mi.eps = e.p->eps;
            mi.seed = e.p->seed;
            mi.max_iter = e.p->max_iter;
            mi.item_base = base;
            mi.R = e.p->n_realizations;
            mi.rid_offset = e.p->realization_offset;
            mi.model_id = e.p->model_id;
            mi.n_track = (int32_t)e.track.size();
            mi.track = e.d_track;
            mi.out_x = e.d_outx;
            mi.out_steps = e.d_steps;
            mi.out_ssa = e.d_ssa;
            mi.out_rej = e.d_rej;
            mi.out_status = e.d_status;
            mi.score = e.d_score;
            mi.dtime = e.d_dtime;
            mi.tlast = e.d_tlast;
            mi.n_intervals = e.p->n_intervals;
            mi.grid = e.d_grid;
            mi.poisson = e.p->poisson;
            base += mi.R;
        }
        const int64_t n_items = base;
        ModelItem *d_items;
        unsigned long long *d_counter;
        TL_TRY(B.alloc(d_items, nm));
        TL_TRY(B.alloc(d_counter, 1));
        TL_HIP(hipMemcpyAsync(d_items, items.data(), nm * sizeof(ModelItem), hipMemcpyHostToDevice, s));
        TL_HIP(hipStreamSynchronize(s));

        PersistentConfig cfg;
        if (bs_req > 0) {
            auto cands = candidates(*dc, is64, sens, n_max, m_max, ext);
            const int gs = bs_req <= 64 ? 64 : bs_req;
            const int bs = bs_req <= 64 ? 64 : bs_req;
            bool found = false;
            for (auto &c : cands)
                if (c.block == bs && c.group == gs) {
                    cfg = c;
                    found = true;
                    break;
                }
            if (!found)
                for (auto &c : cands)
                    if (c.group == gs) {
                        cfg = c;
                        found = true;
                        break;
                    }
            if (!found)
                return set_error(TL_ERR_INVALID_ARG, "block_size " + std::to_string(bs_req) + " does not fit this model");
        } else {
            TL_TRY(autotune(*dc, is64, sens, ext, bs_req == TL_BLOCK_TUNE, items, n_items, n_max, m_max, d_counter,
                            cfg));
        }
        ClassLaunch c;
        c.cfg = cfg;
        c.L = make_launch(*dc, cfg, is64, sens, d_items, nm, n_items, n_max, m_max, d_counter, 0);
        if (cfg.hbm) TL_TRY(B.alloc(c.L.state, (size_t)c.L.grid * c.L.state_stride));
        launches.push_back(c);
        double cc = 0.0;
        for (int32_t k = 0; k < nm; ++k) cc += E[members[k]].cost;
        class_cost.push_back(cc);
        for (int32_t k = 0; k < nm; ++k) entry_cfg[members[k]] = cfg;
    }

    const double t0 = now_s();
    float ms = 0.f;
    {  // the most expensive class is dispatched first
        std::vector<size_t> ord(launches.size());
        for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return class_cost[a] > class_cost[b]; });
        std::vector<ClassLaunch> sorted;
        for (size_t i : ord) sorted.push_back(launches[i]);
        launches.swap(sorted);
    }
    TL_TRY(run_launches(*dc, is64, sens, launches, true, &ms));
    for (auto &e : E)
        if (sens && e.res && (e.res->dmean_dc || e.res->dmean_dc_se)) {
            const int64_t R = e.p->n_realizations;
            const int32_t T = (int32_t)e.track.size();
            const int32_t RS = grad_slices(R, T, e.m->M);
            double *gpart;
            TL_TRY(B.alloc(gpart, (size_t)2 * RS * T * e.m->M));
            launch_grad_reduce(R, T, e.m->M, e.d_outx, e.d_score, e.d_mean, gpart, RS, e.d_dmean, e.d_dse, s);
        }
    for (auto &e : E) {
        if (!e.res) continue;
        const int64_t R = e.p->n_realizations;
        const int32_t T = (int32_t)e.track.size();
        const int32_t n_grid = e.p->n_intervals > 0 ? e.p->n_intervals + 1 : 1;
        TL_TRY(finish_outputs(B, s, R, T, n_grid, e.p->t_final, e.d_outx, e.d_grid, e.res));
    }
    for (auto &e : E) {
        if (!e.res) continue;
        const int64_t R = e.p->n_realizations;
        const int32_t T = (int32_t)e.track.size();
        std::vector<int32_t> hst(R);
        std::vector<int64_t> hsteps(R), hssa(R), hrej(R);
        TL_HIP(hipMemcpyAsync(hsteps.data(), e.d_steps, R * 8, hipMemcpyDeviceToHost, s));
        TL_HIP(hipMemcpyAsync(hssa.data(), e.d_ssa, R * 8, hipMemcpyDeviceToHost, s));
        TL_HIP(hipMemcpyAsync(hrej.data(), e.d_rej, R * 8, hipMemcpyDeviceToHost, s));
        TL_HIP(hipMemcpyAsync(hst.data(), e.d_status, R * 4, hipMemcpyDeviceToHost, s));
        if (sens && e.res->dmean_dc)
            TL_HIP(hipMemcpyAsync(e.res->dmean_dc, e.d_dmean, (size_t)T * e.m->M * 8, hipMemcpyDeviceToHost, s));
        if (sens && e.res->dmean_dc_se)
            TL_HIP(hipMemcpyAsync(e.res->dmean_dc_se, e.d_dse, (size_t)T * e.m->M * 8, hipMemcpyDeviceToHost, s));
        if (sens && e.res->score)
            TL_HIP(hipMemcpyAsync(e.res->score, e.d_score, (size_t)R * e.m->M * 8, hipMemcpyDeviceToHost, s));
        TL_HIP(hipStreamSynchronize(s));
        tl_result *res = e.res;
        fill_summary(res, hsteps, hssa, hrej, hst);
        res->kernel_ms = ms;
        res->upload_seconds = up_total;
        res->form = TL_MODE_PERSISTENT;
        const PersistentConfig &ecfg = entry_cfg[&e - &E[0]];
        res->block_size = ecfg.group == 64 ? 64 : ecfg.block;
        res->sim_seconds = now_s() - t0;
    }
    return TL_OK;
}

}  // namespace

// ------------------------------------------- helpers for other entry points --
tl_status check_sim_params(const tl_model *m, const tl_sim_params *p) { return check_params(m, p); }

bool sim_needs_ext(const tl_model *m, const tl_sim_params *p) { return needs_ext(m, p); }

std::vector<int32_t> sim_track_list(const tl_model *m, const tl_sim_params *p) { return track_list(m, p); }

tl_status default_shape(DeviceContext &dc, bool is64, bool sens, bool ext, int32_t n_max, int32_t m_max,
                        PersistentConfig &cfg) {
    const auto all = candidates(dc, is64, sens, n_max, m_max, ext);
    if (all.empty()) return set_error(TL_ERR_TOO_LARGE, "model too large for any persistent launch shape");
    cfg = heuristic(all, m_max, one_wave_channel_limit(is64, sens));
    return TL_OK;
}

PersistentLaunch launch_plan(DeviceContext &dc, const PersistentConfig &cfg, bool is64, bool sens, ModelItem *d_items,
                             int32_t n_models, int64_t n_items, int32_t n_max, int32_t m_max,
                             unsigned long long *d_counter) {
    return make_launch(dc, cfg, is64, sens, d_items, n_models, n_items, n_max, m_max, d_counter, 0);
}

tl_status device_model_resident(const tl_model &m, int device, DevModel &out) {
    std::lock_guard<std::mutex> g(m.dev_mu);
    auto it = m.dev.find(device);
    if (it == m.dev.end() || it->second->version != m.version)
        return set_error(TL_ERR_INVALID_ARG,
                         "the model's constants are not on device " + std::to_string(device) +
                             " (or its rates / x0 changed since): one tl_simulate or tl_stage_* call with this model "
                             "on the device uploads them");
    out = it->second->dm;
    return TL_OK;
}

}  // namespace tl

using namespace tl;

extern "C" {

const char *tl_last_error(void) { return g_last_error.c_str(); }
const char *tl_version(void) { return "tau_leaping_amd 0.1.0 (gfx950)"; }

tl_status tl_device_count(int32_t *n) {
    if (!n) return set_error(TL_ERR_INVALID_ARG, "null");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return TL_OK;
}

void tl_sim_params_default(tl_sim_params *p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->t_final = 1.0;
    p->eps = 0.03;
    p->n_realizations = 1;
    p->seed = 1234;
    p->max_iter = 0;
    p->precision = TL_FP64;
    p->mode = TL_MODE_AUTO;
}

// AUTO mode: whether a model runs in the plain form.  Always when no
// persistent launch shape fits; and for a model whose state does not fit LDS
// when the HBM-state persistent kernel would need more than one propensity
// tile (or keep a[M] in HBM).  Measured on toy networks, R = 2048 / 8192,
// HBM-state persistent against plain: 12,000 channels (one tile) persistent
// 1.28x / 1.0x faster; 24,000 (two tiles) plain 1.03x / 1.16x faster,
// 40,000 1.09x / 1.20x, 60,000 1.14x / 1.30x, 1e5 1.8x, 2e5 2.0x, 1e6 (all
// leaps) about 7x: realizations across the lanes coalesce every gather.  A
// pure function of the sizes, the precision and the kernel build: the same
// call always runs the same form.
static tl_status auto_plain(const tl_model *m, const tl_sim_params *p, bool &plain) {
    plain = false;
    DeviceContext *dc;
    TL_TRY(device_context(p->device, dc));
    const bool is64 = p->precision == TL_FP64, ext = needs_ext(m, p);
    const auto cands = candidates(*dc, is64, p->sensitivities != 0, m->N, m->M, ext);
    if (cands.empty()) {
        plain = true;
        return TL_OK;
    }
    const PersistentConfig &f = cands.front();
    plain = f.hbm && (f.a_lds || (int64_t)m->M > (int64_t)f.tile_grains * kTileGrain);
    return TL_OK;
}

tl_status tl_simulate(const tl_model *m, const tl_sim_params *p, tl_result *r) {
    TL_TRY(check_params(m, p));
    if (p->mode == TL_MODE_PLAIN) return run_plain(m, p, r);
    if (p->mode == TL_MODE_AUTO) {
        bool plain;
        TL_TRY(auto_plain(m, p, plain));
        if (plain) return run_plain(m, p, r);
    }
    return run_persistent(&m, 1, p, r);
}

tl_status tl_batch_simulate(const tl_model *const *models, int32_t n_models, const tl_sim_params *params,
                            tl_result *results, int32_t mode) {
    if (!models || !params || n_models <= 0) return set_error(TL_ERR_INVALID_ARG, "empty batch");
    if (mode == TL_MODE_PLAIN) {
        for (int32_t i = 0; i < n_models; ++i) {
            TL_TRY(check_params(models[i], &params[i]));
            TL_TRY(run_plain(models[i], &params[i], results ? &results[i] : nullptr));
        }
        return TL_OK;
    }
    if (mode == TL_MODE_AUTO) {
        // models too large for LDS run one by one in the plain form, the rest
        // in one persistent launch
        std::vector<const tl_model *> pm;
        std::vector<tl_sim_params> pp;
        std::vector<int32_t> pi;
        for (int32_t i = 0; i < n_models; ++i) {
            TL_TRY(check_params(models[i], &params[i]));
            bool plain;
            TL_TRY(auto_plain(models[i], &params[i], plain));
            if (plain) {
                TL_TRY(run_plain(models[i], &params[i], results ? &results[i] : nullptr));
            } else {
                pm.push_back(models[i]);
                pp.push_back(params[i]);
                pi.push_back(i);
            }
        }
        if (pm.empty()) return TL_OK;
        if ((int32_t)pm.size() < n_models) {
            std::vector<tl_result> pr(pm.size());
            if (results)
                for (size_t k = 0; k < pm.size(); ++k) pr[k] = results[pi[k]];
            TL_TRY(run_persistent(pm.data(), (int32_t)pm.size(), pp.data(), results ? pr.data() : nullptr));
            if (results)
                for (size_t k = 0; k < pm.size(); ++k) results[pi[k]] = pr[k];
            return TL_OK;
        }
    }
    return run_persistent(models, n_models, params, results);
}

// ---------------------------------------------------------------- stages --
static tl_status stage_setup(const tl_model *m, int32_t precision, int64_t R, DeviceContext *&dc, DevModel &dm) {
    if (!m) return set_error(TL_ERR_INVALID_ARG, "null model");
    if (R < 1) return set_error(TL_ERR_INVALID_ARG, "R must be >= 1");
    if (precision != TL_FP64 && precision != TL_FP32) return set_error(TL_ERR_INVALID_ARG, "bad precision");
    int dev = 0;
    hipGetDevice(&dev);
    TL_TRY(device_context(dev, dc));
    // the caller's buffers may still be written (or, freshly recycled by a
    // caching allocator, read) by its own work on other streams: the stage
    // runs after all of it
    TL_HIP(hipDeviceSynchronize());
    return device_model(*m, dev, dm, nullptr);
}

static tl_status stage_finish(DeviceContext *dc) {
    TL_HIP(hipGetLastError());
    TL_HIP(hipStreamSynchronize(dc->stream));
    return TL_OK;
}

tl_status tl_stage_propensities(const tl_model *m, int32_t precision, int64_t R, const int32_t *x, void *a,
                                uint8_t *crit, double *a0, double *a0c) {
    DeviceContext *dc;
    DevModel dm;
    TL_TRY(stage_setup(m, precision, R, dc, dm));
    if (!x || !a || !crit || !a0 || !a0c) return set_error(TL_ERR_INVALID_ARG, "null buffer");
    launch_prop(precision == TL_FP64, dm, R, 1, x, a, crit, a0, a0c, nullptr, nullptr, nullptr, dc->stream);
    return stage_finish(dc);
}

tl_status tl_stage_tau(const tl_model *m, int32_t precision, int64_t R, double eps, const int32_t *x, const void *a,
                       const uint8_t *crit, double *tau1, int32_t *argmin, int32_t *which) {
    DeviceContext *dc;
    DevModel dm;
    TL_TRY(stage_setup(m, precision, R, dc, dm));
    if (!x || !a || !crit || !tau1) return set_error(TL_ERR_INVALID_ARG, "null buffer");
    void *tau_i;
    TL_TRY(scratch(*dc, 1, (size_t)m->N * R * (precision == TL_FP64 ? 8 : 4), tau_i));
    void *which_i;
    TL_TRY(scratch(*dc, 2, (size_t)m->N * R, which_i));
    launch_tau(precision == TL_FP64, dm, R, eps, x, a, crit, tau_i, (uint8_t *)which_i, tau1, argmin, which, nullptr,
               dc->stream);
    return stage_finish(dc);
}

tl_status tl_stage_select(const tl_model *m, int32_t precision, int64_t R, uint64_t seed, uint32_t model_id,
                          int64_t realization_offset, const uint32_t *epoch, const void *a, const uint8_t *crit,
                          const double *a0c, const double *tau1, const double *t_rem, double *tau, int32_t *jc,
                          uint8_t *clamped) {
    DeviceContext *dc;
    DevModel dm;
    TL_TRY(stage_setup(m, precision, R, dc, dm));
    if (!epoch || !a || !crit || !a0c || !tau1 || !t_rem || !tau || !jc || !clamped)
        return set_error(TL_ERR_INVALID_ARG, "null buffer");
    launch_select(precision == TL_FP64, dm, R, seed, model_id, realization_offset, epoch, a, crit, a0c, tau1, t_rem,
                  tau, jc, clamped, dc->stream);
    return stage_finish(dc);
}

tl_status tl_stage_fire(const tl_model *m, int32_t precision, int64_t R, uint64_t seed, uint32_t model_id,
                        int64_t realization_offset, const uint32_t *epoch, const void *a, const uint8_t *crit,
                        const double *tau, const int32_t *jc, int32_t *k) {
    DeviceContext *dc;
    DevModel dm;
    TL_TRY(stage_setup(m, precision, R, dc, dm));
    if (!epoch || !a || !crit || !tau || !jc || !k) return set_error(TL_ERR_INVALID_ARG, "null buffer");
    launch_fire(precision == TL_FP64, dm, R, seed, model_id, realization_offset, epoch, a, crit, tau, jc, k, nullptr, 0,
                dc->stream);
    return stage_finish(dc);
}

tl_status tl_stage_update(const tl_model *m, int64_t R, const int32_t *x, const int32_t *k, int32_t *x_out,
                          uint8_t *neg) {
    DeviceContext *dc;
    DevModel dm;
    TL_TRY(stage_setup(m, TL_FP64, R, dc, dm));
    if (!x || !k || !x_out || !neg) return set_error(TL_ERR_INVALID_ARG, "null buffer");
    launch_fill_u8(neg, 0, R, dc->stream);
    launch_update(dm, R, x, k, x_out, neg, nullptr, dc->stream);
    return stage_finish(dc);
}

tl_status tl_stage_ssa(const tl_model *m, int32_t precision, int64_t R, uint64_t seed, uint32_t model_id,
                       int64_t realization_offset, int32_t *x, double *t, uint32_t *epoch, double t_final,
                       int32_t n_events, int32_t *events) {
    DeviceContext *dc;
    DevModel dm;
    TL_TRY(stage_setup(m, precision, R, dc, dm));
    if (!x || !t || !epoch || !events) return set_error(TL_ERR_INVALID_ARG, "null buffer");
    void *a;
    TL_TRY(scratch(*dc, 3, (size_t)m->M * R * (precision == TL_FP64 ? 8 : 4), a));
    launch_ssa(precision == TL_FP64, dm, R, seed, model_id, realization_offset, x, t, epoch, t_final, n_events, events,
               a, dc->stream);
    return stage_finish(dc);
}

}  // extern "C"
