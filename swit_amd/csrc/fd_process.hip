// Shot-gather processing on the device (toolbox/preprocess.py:47-200): offset and
// first-break mutes, zero-phase Butterworth filtering (scipy filtfilt), demean, linear
// detrend and Hann taper after the filter, trace and event normalisation.  Gathers are
// [ngathers][ntr][nt] float32, processed in place, in the reference's float32 flow:
//   mute     x = f32(x * m) with m the tapered mask of preprocess.py:238-262 (FP64),
//            pick = first sample with |x| > f32(0.001) * max|x| (float32, :203-209);
//   filter   odd extension of 3 * taps samples at each end in float32, direct-form II
//            transposed IIR in FP64 forward from zi * ext[0], backward from
//            zi * y[last], stored back as float32 (scipy.signal.filtfilt, padtype 'odd');
//            the recurrence keeps scipy's operation order, so the result is bit-exact;
//   detrend  x - mean, then x - (t c0 + c1) with t = f32((i+1)/nt), the least-squares
//            line fitted in FP64 and rounded to float32 (scipy.signal.detrend on float32
//            data solves it in float32), then x = f32(x * taper);
//   norms    trace: L1 / L2 / max in FP64 rounded to float32, x = x / w if w > 0;
//            event: w = sum over the gather's traces of their float32 norms, added in
//            trace order in float32 as numpy adds np.float32 scalars, x = x / w.
#include "fd_common.h"

namespace swit {
namespace pr {

constexpr int kThreads = 256;
constexpr int kFilterThreads = 64;
constexpr int kFilterBlock = 16;     // samples per block of loads in k_filtfilt
constexpr int kMaxTaps = 16;
constexpr int kMuteLength = 100;    // samples of the arrival-mute taper (preprocess.py:101)

struct ProcArgs {
    int ntr, nt;                    // traces per gather, samples
    int mute_short, mute_long;      // offset mutes on/off
    double short_dist, long_dist;   // |offset| < short_dist / > long_dist -> zero trace
    int mute_arrival;               // 0 none, 1 early (x m), 2 late (x (1 - m))
    double pick_shift;              // ceil(|window| / dt) samples
    int taps, pad;                  // filter taps (0: none); odd-extension length
    double b[kMaxTaps], a[kMaxTaps], zi[kMaxTaps];
    int norm_trace;                 // 0 none, 1 L1, 2 L2, 3 max
    int norm_event;                 // 0 none, 1 L1, 2 L2
};

__device__ __forceinline__ double wsum(double v, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) t += sh[w];
    return t;
}

__device__ __forceinline__ float wmaxf(float v, float *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    float t = sh[0];
#pragma unroll
    for (int w = 1; w < kThreads / 64; ++w) t = fmaxf(t, sh[w]);
    return t;
}

__device__ __forceinline__ int wmini(int v, int *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    int t = sh[0];
#pragma unroll
    for (int w = 1; w < kThreads / 64; ++w) t = min(t, sh[w]);
    return t;
}

// Mask value at sample i (preprocess.py:238-262, the same branch order; win holds
// sin(linspace(0, pi, 2 length))[:length]).
__device__ __forceinline__ double mute_mask(int i, int itmin, int itmax, int nt, const double *win) {
    const int L = kMuteLength;
    if (1 < itmin && itmin < itmax && itmax < nt) {
        if (i < itmin) return 0.0;
        if (i < itmax) return win[i - itmin];
        return 1.0;
    }
    if (itmin < 1 && 1 <= itmax) return i < itmax ? win[L - itmax + i] : 1.0;
    if (itmin < nt && nt < itmax) {
        if (i < itmin) return 0.0;
        return win[i - itmin];
    }
    if (itmin > nt) return 0.0;
    return 1.0;
}

// Offset and first-break mutes, one workgroup per trace.
// picks: per-trace first-break window centres (samples) given by the caller, or NULL to
// pick the first break from the data and shift it by the window (pick_shift).
__global__ __launch_bounds__(kThreads) void k_mute(ProcArgs p, const float *__restrict__ offsets,
                                                   const double *__restrict__ win, const int *__restrict__ picks,
                                                   float *__restrict__ data) {
    __shared__ float shf[kThreads / 64];
    __shared__ int shi[kThreads / 64];
    const int nt = p.nt;
    float *x = data + (size_t)blockIdx.x * nt;
    const double off = (p.mute_short || p.mute_long) ? fabs((double)offsets[blockIdx.x]) : 0.0;
    if ((p.mute_short && off < p.short_dist) || (p.mute_long && off > p.long_dist)) {
        // a zeroed trace stays +0 under the window mute (the mask is >= 0)
        for (int i = threadIdx.x; i < nt; i += kThreads) x[i] = 0.0f;
        return;
    }
    if (!p.mute_arrival) return;
    double pk;
    if (picks) {
        pk = (double)picks[blockIdx.x];
    } else {
        float m = 0.0f;
        for (int i = threadIdx.x; i < nt; i += kThreads) m = fmaxf(m, fabsf(x[i]));
        m = wmaxf(m, shf);
        const float thr = 0.001f * m;
        int first = nt;
        for (int i = threadIdx.x; i < nt; i += kThreads)
            if (fabsf(x[i]) > thr) { first = i; break; }
        first = wmini(first, shi);
        const int pick = first == nt ? 0 : first;             // argmax of an all-false mask is 0
        pk = (double)pick + p.pick_shift;
    }
    const int itmin = (int)(pk - (double)kMuteLength / 2.0);
    const int itmax = (int)((double)itmin + (double)kMuteLength);
    for (int i = threadIdx.x; i < nt; i += kThreads) {
        const double mk = mute_mask(i, itmin, itmax, nt, win);
        x[i] = (float)((double)x[i] * (p.mute_arrival == 1 ? mk : 1.0 - mk));
    }
}

// filtfilt of one trace per thread.  y: [ext][ntraces] doubles, time-major so that a
// wave's loads and stores of one time sample are contiguous.
__device__ __forceinline__ float odd_ext(const float *x, int n, int pad, int nt) {
    if (n < pad) return 2.0f * x[0] - x[pad - n];
    if (n < pad + nt) return x[n - pad];
    return 2.0f * x[nt - 1] - x[nt - 2 - (n - pad - nt)];
}

// K taps (compile time, so the delay line stays in registers).
template <int K>
__device__ __forceinline__ double df2t_step(double xn, double (&z)[K], const ProcArgs &p) {
    const double yn = z[0] + p.b[0] * xn;
#pragma unroll
    for (int k = 0; k < K - 2; ++k) z[k] = z[k + 1] + xn * p.b[k + 1] - yn * p.a[k + 1];
    z[K - 2] = xn * p.b[K - 1] - yn * p.a[K - 1];
    return yn;
}

template <int K>
__global__ __launch_bounds__(kFilterThreads) void k_filtfilt(ProcArgs p, int ntraces, float *__restrict__ data,
                                                             double *__restrict__ y) {
    const int tr = blockIdx.x * kFilterThreads + threadIdx.x;
    if (tr >= ntraces) return;
    const int nt = p.nt, pad = p.pad, L = nt + 2 * pad;
    float *x = data + (size_t)tr * nt;
    double z[K];
    z[K - 1] = 0.0;
    // forward: y = lfilter(b, a, ext, zi = zi * ext[0]).  Samples go in blocks of
    // kFilterBlock: a block's loads are issued together, then the recursion runs over
    // them, so the load latency is paid once per block rather than once per sample.
    const double x0 = (double)odd_ext(x, 0, pad, nt);
#pragma unroll
    for (int k = 0; k < K - 1; ++k) z[k] = p.zi[k] * x0;
    for (int n0 = 0; n0 < L; n0 += kFilterBlock) {
        double xb[kFilterBlock];
#pragma unroll
        for (int i = 0; i < kFilterBlock; ++i) xb[i] = n0 + i < L ? (double)odd_ext(x, n0 + i, pad, nt) : 0.0;
#pragma unroll
        for (int i = 0; i < kFilterBlock; ++i)
            if (n0 + i < L) y[(size_t)(n0 + i) * ntraces + tr] = df2t_step<K>(xb[i], z, p);
    }
    // backward over the reversed output from zi * y[L - 1]; step n of this pass is
    // time L - 1 - n of the extended trace
    const double y0 = y[(size_t)(L - 1) * ntraces + tr];
#pragma unroll
    for (int k = 0; k < K - 1; ++k) z[k] = p.zi[k] * y0;
    for (int t0 = L - 1; t0 >= 0; t0 -= kFilterBlock) {
        double yb[kFilterBlock];
#pragma unroll
        for (int i = 0; i < kFilterBlock; ++i) yb[i] = t0 - i >= 0 ? y[(size_t)(t0 - i) * ntraces + tr] : 0.0;
#pragma unroll
        for (int i = 0; i < kFilterBlock; ++i) {
            const int t = t0 - i;
            if (t < 0) break;
            const double yn = df2t_step<K>(yb[i], z, p);
            if (t >= pad && t < pad + nt) x[t - pad] = (float)yn;
        }
    }
}

// Demean, linear detrend and taper, one workgroup per trace.
__global__ __launch_bounds__(kThreads) void k_detrend_taper(ProcArgs p, const double *__restrict__ taper,
                                                            float *__restrict__ data) {
    __shared__ double sh[kThreads / 64];
    const int nt = p.nt;
    float *x = data + (size_t)blockIdx.x * nt;
    double s = 0.0;
    for (int i = threadIdx.x; i < nt; i += kThreads) s += x[i];
    s = wsum(s, sh);
    const float mean = (float)(s / nt);
    // least squares of x - mean on (t, 1), t = f32((i + 1) / nt)
    double st = 0.0, stt = 0.0, sx = 0.0, stx = 0.0;
    for (int i = threadIdx.x; i < nt; i += kThreads) {
        const float xv = x[i] - mean;
        const double t = (double)((float)(i + 1) / (float)nt);
        st += t;
        stt += t * t;
        sx += xv;
        stx += t * xv;
    }
    st = wsum(st, sh);
    stt = wsum(stt, sh);
    sx = wsum(sx, sh);
    stx = wsum(stx, sh);
    const double det = nt * stt - st * st;
    const float c0 = (float)((nt * stx - st * sx) / det);
    const float c1 = (float)((stt * sx - st * stx) / det);
    for (int i = threadIdx.x; i < nt; i += kThreads) {
        const float t = (float)(i + 1) / (float)nt;
        const float xv = (x[i] - mean) - (t * c0 + c1);
        x[i] = (float)((double)xv * taper[i]);
    }
}

// Trace normalisation, then each trace's event norm of the result into enorm.
__global__ __launch_bounds__(kThreads) void k_norm_trace(ProcArgs p, float *__restrict__ data,
                                                         float *__restrict__ enorm) {
    __shared__ double sh[kThreads / 64];
    __shared__ float shf[kThreads / 64];
    const int nt = p.nt;
    float *x = data + (size_t)blockIdx.x * nt;
    if (p.norm_trace) {
        float w;
        if (p.norm_trace == 3) {
            float m = 0.0f;
            for (int i = threadIdx.x; i < nt; i += kThreads) m = fmaxf(m, fabsf(x[i]));
            w = wmaxf(m, shf);
        } else {
            double s = 0.0;
            for (int i = threadIdx.x; i < nt; i += kThreads) s += p.norm_trace == 1 ? fabs((double)x[i])
                                                                                   : (double)x[i] * x[i];
            s = wsum(s, sh);
            w = (float)(p.norm_trace == 1 ? s : sqrt(s));
        }
        if (w > 0.0f)
            for (int i = threadIdx.x; i < nt; i += kThreads) x[i] = x[i] / w;
    }
    if (p.norm_event) {
        __syncthreads();
        double s = 0.0;
        for (int i = threadIdx.x; i < nt; i += kThreads) s += p.norm_event == 1 ? fabs((double)x[i])
                                                                                : (double)x[i] * x[i];
        s = wsum(s, sh);
        if (threadIdx.x == 0) enorm[blockIdx.x] = (float)(p.norm_event == 1 ? s : sqrt(s));
    }
}

// Event normalisation: one workgroup per gather; the gather's norm is the float32 sum
// of its traces' norms in trace order.
__global__ __launch_bounds__(kThreads) void k_norm_event(ProcArgs p, const float *__restrict__ enorm,
                                                         float *__restrict__ data) {
    __shared__ float w;
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int i = 0; i < p.ntr; ++i) s = s + enorm[(size_t)blockIdx.x * p.ntr + i];
        w = s;
    }
    __syncthreads();
    float *x = data + (size_t)blockIdx.x * p.ntr * p.nt;
    const size_t n = (size_t)p.ntr * p.nt;
    for (size_t i = threadIdx.x; i < n; i += kThreads) x[i] = x[i] / w;
}

}  // namespace pr
}  // namespace swit
