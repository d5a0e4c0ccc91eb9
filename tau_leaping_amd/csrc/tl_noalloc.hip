// tl_noalloc.hip -- tl_simulate_noalloc: the persistent engine for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// It runs the same persistent kernel as tl_simulate on one model, with the
// default launch shape, and allocates nothing: the work item, the work
// counter, the tracked-species list and the per-realization counts the
// kernel writes besides the caller's outputs live in `scratch`, one
// file-scope device array per GPU (held until the process ends).  Its device
// address is looked up once per device with hipGetSymbolAddress; host code
// only ever passes that address, never the array's name.  Every call on a
// GPU shares its array: calls that overlap on one GPU (two host threads, two
// streams) race and must be serialized by the caller.
#include <map>
#include <mutex>

#include "tl_internal.hpp"
#include "tl_model.hpp"

namespace tl {
namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int32_t kMaxTrack = 256;
// scratch layout: ModelItem | work counter | tracked species | ssa [R] | rej [R]
constexpr size_t kOffItem = 0, kOffCounter = 512, kOffTrack = 1024, kOffCounts = 2048;
static_assert(sizeof(ModelItem) <= kOffCounter, "work item fits its slot");
static_assert(kOffTrack + kMaxTrack * sizeof(int32_t) <= kOffCounts, "track list fits its slot");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

// The largest R whose two int64 counts per realization fit the scratch.
int64_t max_realizations() { return (int64_t)((kScratchBytes - kOffCounts - 256) / 16); }

struct TrackList {
    int32_t n;
    int32_t idx[kMaxTrack];
};

// Writes the work item and the tracked species (passed by value, so a
// captured graph holds them in its kernel node) and zeroes the work counter.
__global__ __launch_bounds__(256) void k_noalloc_setup(ModelItem mi, TrackList trk, unsigned char *base) {
    if (threadIdx.x == 0) {
        *reinterpret_cast<ModelItem *>(base + kOffItem) = mi;
        *reinterpret_cast<unsigned long long *>(base + kOffCounter) = 0ull;
    }
    int32_t *t = reinterpret_cast<int32_t *>(base + kOffTrack);
    for (int32_t i = threadIdx.x; i < trk.n; i += blockDim.x) t[i] = trk.idx[i];
}

std::mutex g_mu;
std::map<int, unsigned char *> g_scratch;  // device -> address of `scratch` there

tl_status scratch_address(int device, unsigned char *&out) {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_scratch.find(device);
    if (it != g_scratch.end()) {
        out = it->second;
        return TL_OK;
    }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the current device's copy
    if (e != hipSuccess || !p)
        return set_error(TL_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = (unsigned char *)p;
    out = (unsigned char *)p;
    return TL_OK;
}

}  // namespace
}  // namespace tl

using namespace tl;

extern "C" tl_status tl_simulate_noalloc(const tl_model *m, const tl_sim_params *p, int32_t *final_x, int64_t *steps,
                                         int32_t *status, void *stream) {
    tl_status st = check_sim_params(m, p);
    if (st != TL_OK) return st;
    if (!final_x || !steps || !status) return set_error(TL_ERR_INVALID_ARG, "null output buffer");
    if (p->sensitivities || p->n_intervals > 0)
        return set_error(TL_ERR_INVALID_ARG, "tl_simulate_noalloc runs without sensitivities and without a time grid");
    if (p->mode == TL_MODE_PLAIN) return set_error(TL_ERR_INVALID_ARG, "tl_simulate_noalloc runs the persistent form");
    if (p->block_size != 0) return set_error(TL_ERR_INVALID_ARG, "tl_simulate_noalloc takes the default launch shape");
    const int64_t R = p->n_realizations;
    if (R > max_realizations())
        return set_error(TL_ERR_TOO_LARGE, "tl_simulate_noalloc: at most " + std::to_string(max_realizations()) +
                                               " realizations per call");
    const std::vector<int32_t> trk = sim_track_list(m, p);
    if ((int32_t)trk.size() > kMaxTrack)
        return set_error(TL_ERR_TOO_LARGE, "tl_simulate_noalloc: at most 256 tracked species");
    DeviceContext *dc;
    st = device_context(p->device, dc);
    if (st != TL_OK) return st;
    DevModel dm;
    st = device_model_resident(*m, p->device, dm);
    if (st != TL_OK) return st;
    unsigned char *base;
    st = scratch_address(p->device, base);
    if (st != TL_OK) return st;
    const bool is64 = p->precision == TL_FP64, ext = sim_needs_ext(m, p);
    PersistentConfig cfg;
    st = default_shape(*dc, is64, false, ext, m->N, m->M, cfg);
    if (st != TL_OK) return st;
    if (cfg.hbm)
        return set_error(TL_ERR_TOO_LARGE, "tl_simulate_noalloc: the model's state does not fit LDS (its per-block "
                                           "state would need device memory)");

    int64_t *d_ssa = reinterpret_cast<int64_t *>(base + kOffCounts);
    int64_t *d_rej = reinterpret_cast<int64_t *>(base + kOffCounts + al256((size_t)R * 8));
    ModelItem mi{};
    mi.dm = dm;
    mi.t_final = p->t_final;
    mi.eps = p->eps;
    mi.seed = p->seed;
    mi.max_iter = p->max_iter;
    mi.item_base = 0;
    mi.R = R;
    mi.rid_offset = p->realization_offset;
    mi.model_id = p->model_id;
    mi.n_track = (int32_t)trk.size();
    mi.track = reinterpret_cast<const int32_t *>(base + kOffTrack);
    mi.out_x = final_x;
    mi.out_steps = steps;
    mi.out_ssa = d_ssa;
    mi.out_rej = d_rej;
    mi.out_status = status;
    mi.poisson = p->poisson;
    TrackList tlist;
    tlist.n = (int32_t)trk.size();
    for (int32_t q = 0; q < tlist.n; ++q) tlist.idx[q] = trk[q];

    hipStream_t s = stream ? (hipStream_t)stream : dc->stream;
    k_noalloc_setup<<<1, 256, 0, s>>>(mi, tlist, base);
    const PersistentLaunch L = launch_plan(*dc, cfg, is64, false, reinterpret_cast<ModelItem *>(base + kOffItem), 1,
                                           R, m->N, m->M, reinterpret_cast<unsigned long long *>(base + kOffCounter));
    TL_HIP(launch_persistent(cfg, is64, false, L, s));
    TL_HIP(hipGetLastError());
    if (!stream) TL_HIP(hipStreamSynchronize(s));
    return TL_OK;
}
