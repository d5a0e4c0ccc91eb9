// tl_reduce.hip -- reductions over realizations, on the device: mean and
// sample standard deviation of the tracked species (at t_final or on the
// time grid; StepStat2StdDev, tau_leaping/statistics.cpp:67-87), and the
// likelihood-ratio estimate dE[X_q(t_final)]/dc_j = mean_r (X_q - Xbar_q)
// S_j with its standard error.  Only the reduced values cross to the host.
// Every sum runs in a fixed order (per-thread strides, then a block tree or
// a slice-ordered combine), so results repeat bit for bit.
#include "tl_internal.hpp"

namespace tl {
namespace {

constexpr int kB = 256;

// Sum of v over the block, in a fixed tree order; every thread gets it.
__device__ double block_sum(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    const double s = sh[0];
    __syncthreads();
    return s;
}

// Column c (blockIdx.x + gridDim.x strides) of v [R][C]: mean, then the
// squared deviations from it (two passes, as the host formula).
template <class V>
__global__ __launch_bounds__(kB) void k_moments(int64_t Rn, int64_t C, const V *__restrict__ v,
                                                double *__restrict__ mean, double *__restrict__ sd) {
    __shared__ double sh[kB];
    for (int64_t c = blockIdx.x; c < C; c += gridDim.x) {
        double s = 0.0;
        for (int64_t r = threadIdx.x; r < Rn; r += kB) s += (double)v[r * C + c];
        const double mu = block_sum(s, sh) / (double)Rn;
        double s2 = 0.0;
        for (int64_t r = threadIdx.x; r < Rn; r += kB) {
            const double d = (double)v[r * C + c] - mu;
            s2 += d * d;
        }
        const double ss = block_sum(s2, sh);
        if (threadIdx.x == 0) {
            if (mean) mean[c] = mu;
            if (sd) sd[c] = Rn > 1 ? sqrt(ss / (double)(Rn - 1)) : 0.0;
        }
    }
}

// Partial sums of d = (X_q - Xbar_q) S_j and d^2 over realization slice
// blockIdx.z: part[(z * T + q) * M + j], and the squares RS * T * M further.
// Threads of a block take consecutive channels (coalesced reads of S [R][M]).
__global__ __launch_bounds__(kB) void k_grad_part(int64_t Rn, int32_t T, int32_t M, int32_t RS,
                                                  const int32_t *__restrict__ xt, const double *__restrict__ S,
                                                  const double *__restrict__ mean, double *__restrict__ part) {
    const int32_t j = blockIdx.x * kB + threadIdx.x;
    const int32_t sl = blockIdx.z;
    if (j >= M) return;
    const int64_t chunk = (Rn + RS - 1) / RS;
    const int64_t r0 = (int64_t)sl * chunk, r1 = r0 + chunk < Rn ? r0 + chunk : Rn;
    for (int32_t q = blockIdx.y; q < T; q += gridDim.y) {
        const double mq = mean[q];
        double acc = 0.0, acc2 = 0.0;
        for (int64_t r = r0; r < r1; ++r) {
            const double d = ((double)xt[r * T + q] - mq) * S[r * M + j];
            acc += d;
            acc2 += d * d;
        }
        const int64_t o = ((int64_t)sl * T + q) * M + j;
        part[o] = acc;
        part[(int64_t)RS * T * M + o] = acc2;
    }
}

// dmean [T][M] and its standard error from the slices, combined in order.
__global__ __launch_bounds__(kB) void k_grad_fin(int64_t Rn, int32_t T, int32_t M, int32_t RS,
                                                 const double *__restrict__ part, double *__restrict__ dmean,
                                                 double *__restrict__ se) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    const int64_t TM = (int64_t)T * M;
    if (i >= TM) return;
    double acc = 0.0, acc2 = 0.0;
    for (int32_t sl = 0; sl < RS; ++sl) {
        acc += part[(int64_t)sl * TM + i];
        acc2 += part[(int64_t)RS * TM + (int64_t)sl * TM + i];
    }
    const double mu = acc / (double)Rn;
    dmean[i] = mu;
    if (se) {
        const double var = (Rn > 1) ? (acc2 / (double)Rn - mu * mu) * (double)Rn / (double)(Rn - 1) : 0.0;
        se[i] = sqrt(fmax(var, 0.0) / (double)Rn);
    }
}

unsigned grid_cap(int64_t n) { return (unsigned)(n < 65535 ? n : 65535); }

}  // namespace

void launch_moments_i32(int64_t R, int64_t C, const int32_t *v, double *mean, double *sd, hipStream_t s) {
    k_moments<int32_t><<<grid_cap(C), kB, 0, s>>>(R, C, v, mean, sd);
}

void launch_moments_f64(int64_t R, int64_t C, const double *v, double *mean, double *sd, hipStream_t s) {
    k_moments<double><<<grid_cap(C), kB, 0, s>>>(R, C, v, mean, sd);
}

int32_t grad_slices(int64_t R, int32_t T, int32_t M) {
    const int64_t threads = (int64_t)((M + kB - 1) / kB) * kB * T;
    int64_t rs = (262144 + threads - 1) / threads;
    rs = rs < 1 ? 1 : rs;
    const int64_t by_r = (R + 63) / 64;  // at least 64 realizations per slice
    rs = rs < by_r ? rs : by_r;
    return (int32_t)(rs < 1024 ? (rs < 1 ? 1 : rs) : 1024);
}

void launch_grad_reduce(int64_t R, int32_t n_track, int32_t M, const int32_t *x_track, const double *score,
                        double *mean_scratch, double *part, int32_t RS, double *dmean, double *dmean_se,
                        hipStream_t s) {
    k_moments<int32_t><<<grid_cap(n_track), kB, 0, s>>>(R, n_track, x_track, mean_scratch, nullptr);
    const dim3 g((unsigned)((M + kB - 1) / kB), grid_cap(n_track), (unsigned)RS);
    k_grad_part<<<g, kB, 0, s>>>(R, n_track, M, RS, x_track, score, mean_scratch, part);
    const int64_t TM = (int64_t)n_track * M;
    k_grad_fin<<<(unsigned)((TM + kB - 1) / kB), kB, 0, s>>>(R, n_track, M, RS, part, dmean, dmean_se);
}

}  // namespace tl
