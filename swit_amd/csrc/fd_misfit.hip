// Misfits and adjoint sources on the device (toolbox/misfit.py), one workgroup per
// trace, FP64 accumulation, and a fixed-order FP64 sum over traces (reproducible).
//
// Trace rules as misfit.py: a trace whose observed or synthetic L2 norm is zero
// contributes nothing and gets a zero adjoint source.
//   waveform (:128-148, :263-282): adj = syn - obs (FP32, as the reference's float32
//     arrays), misfit = sum_traces sqrt(sum r^2 dt);
//   global correlation (:219-242, :338-355): o = obs/|obs|, s = syn/|syn|,
//     c = Pearson coefficient of (s, o) (np.corrcoef, means removed),
//     misfit = -sum_traces c, adj = (s c - o) / |syn|.
#include "fd_common.h"

namespace swit {

constexpr int kMisfitThreads = 256;

__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();                      // sh may still be read from a previous call
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kMisfitThreads / 64; ++w) t += sh[w];
    return t;
}

// kind 0 waveform, 1 global correlation.  obs/syn/adj: [ntraces][nt].
template <int KIND>
__global__ __launch_bounds__(kMisfitThreads) void k_misfit_trace(int nt, float dt, const float *__restrict__ obs,
                                                                 const float *__restrict__ syn,
                                                                 float *__restrict__ adj,
                                                                 double *__restrict__ tmis) {
    __shared__ double sh[kMisfitThreads / 64];
    const size_t base = (size_t)blockIdx.x * nt;
    const float *o = obs + base, *s = syn + base;
    float *a = adj + base;
    double so = 0.0, ss = 0.0, sumo = 0.0, sums = 0.0;
    for (int i = threadIdx.x; i < nt; i += kMisfitThreads) {
        const double ov = o[i], sv = s[i];
        so += ov * ov;
        ss += sv * sv;
        sumo += ov;
        sums += sv;
    }
    so = block_sum(so, sh);
    ss = block_sum(ss, sh);
    const bool live = so > 0.0 && ss > 0.0;
    if (KIND == 0) {
        double r2 = 0.0;
        for (int i = threadIdx.x; i < nt; i += kMisfitThreads) {
            const float r = s[i] - o[i];
            a[i] = live ? r : 0.0f;
            r2 += (double)r * (double)r;
        }
        r2 = block_sum(r2, sh);
        if (threadIdx.x == 0) tmis[blockIdx.x] = live ? sqrt(r2 * (double)dt) : 0.0;
        return;
    }
    sumo = block_sum(sumo, sh);
    sums = block_sum(sums, sh);
    const double no = sqrt(so), ns = sqrt(ss);
    const double mo = sumo / nt, ms = sums / nt;
    double cov = 0.0, vo = 0.0, vs = 0.0;
    for (int i = threadIdx.x; i < nt; i += kMisfitThreads) {
        const double od = (double)o[i] - mo, sd = (double)s[i] - ms;
        cov += od * sd;
        vo += od * od;
        vs += sd * sd;
    }
    cov = block_sum(cov, sh);
    vo = block_sum(vo, sh);
    vs = block_sum(vs, sh);
    // Pearson is invariant to the normalisations by |obs| and |syn|
    const double c = (live && vo > 0.0 && vs > 0.0) ? cov / sqrt(vo * vs) : 0.0;
    for (int i = threadIdx.x; i < nt; i += kMisfitThreads) {
        const double on = (double)o[i] / no, sn = (double)s[i] / ns;
        a[i] = live ? (float)((sn * c - on) / ns) : 0.0f;
    }
    if (threadIdx.x == 0) tmis[blockIdx.x] = live ? -c : 0.0;
}

// Fixed-order FP64 sum of n values into out[0]: one workgroup, each thread sums a
// strided subset in order, then a fixed tree.
__global__ __launch_bounds__(kMisfitThreads) void k_sum_fixed(int n, const double *__restrict__ v,
                                                              double *__restrict__ out) {
    __shared__ double sh[kMisfitThreads / 64];
    double t = 0.0;
    for (int i = threadIdx.x; i < n; i += kMisfitThreads) t += v[i];
    t = block_sum(t, sh);
    if (threadIdx.x == 0) out[0] = t;
}

}  // namespace swit
