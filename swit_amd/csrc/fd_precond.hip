// Gradient preconditioning on the device (toolbox/postprocess.py:20-111): taper, the
// illumination preconditioner built from Gaussian-smoothed illuminations, gradient
// smoothing, mask and scaling, in FP64 like the reference's numpy (the in-place taper
// product keeps the float32 gradient's dtype, as there).
//
// smooth2d (toolbox/tools.py:239-275) convolves with the normalised (2 span + 1)^2
// Gaussian F[a][b] = e_a e_b / (sum e)^2, e_a = exp(-(x_a / span)^2 / 2), x_a = 2a - 2 span
// (zero padding, 'same' size), and divides by the same convolution of a field of ones.
// F is separable, so both become two 1-D passes: along z, then along x, and the ones
// field's convolution is the product of the kernel's partial sums over the valid taps.
// All arrays are [nx][nz] (z fastest).
#include "fd_common.h"

namespace swit {
namespace pc {

constexpr int kMaxThreads = 1024;

// f[a] = e_a / sum e, a = 0 .. 2 span (one thread; the sum in index order)
__global__ void k_weights(int span, double *__restrict__ f) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int n = 2 * span + 1;
    double sum = 0.0;
    for (int a = 0; a < n; ++a) {
        const double x = span > 0 ? (2.0 * a - 2.0 * span) / span : 0.0;
        f[a] = exp(-0.5 * x * x);
        sum += f[a];
    }
    for (int a = 0; a < n; ++a) f[a] /= sum;
}

// out[i][z0 + j] = sum_b in[i][z0 + j + b - span] f[b] over taps inside rows [z0, nz)
__global__ void k_smooth_z(int nx, int nz, int z0, int span, const double *__restrict__ f,
                           const double *__restrict__ in, double *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    const int nzs = nz - z0;
    if (j >= nzs || i >= nx) return;
    const double *col = in + (size_t)i * nz + z0;
    double s = 0.0;
    for (int b = max(0, span - j); b <= min(2 * span, span + nzs - 1 - j); ++b) s += col[j + b - span] * f[b];
    out[(size_t)i * nz + z0 + j] = s;
}

// out[i][z0 + j] = (sum_a in[i + a - span][z0 + j] f[a]) / (wx(i) wz(j)), with wx, wz the
// kernel's sums over the taps inside the x range and inside rows [z0, nz)
__global__ void k_smooth_x(int nx, int nz, int z0, int span, const double *__restrict__ f,
                           const double *__restrict__ in, double *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    const int nzs = nz - z0;
    if (j >= nzs || i >= nx) return;
    double s = 0.0, wx = 0.0, wz = 0.0;
    const int a0 = max(0, span - i), a1 = min(2 * span, span + nx - 1 - i);
    for (int a = a0; a <= a1; ++a) {
        s += in[(size_t)(i + a - span) * nz + z0 + j] * f[a];
        wx += f[a];
    }
    for (int b = max(0, span - j); b <= min(2 * span, span + nzs - 1 - j); ++b) wz += f[b];
    out[(size_t)i * nz + z0 + j] = s / (wx * wz);
}

__global__ void k_to_double(size_t n, const float *__restrict__ in, double *__restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        out[k] = (double)in[k];
}

// m = max over n values of v (or |v|); one workgroup
__global__ __launch_bounds__(kMaxThreads) void k_max(size_t n, const double *__restrict__ v, int absolute,
                                                     double *__restrict__ m) {
    __shared__ double sh[kMaxThreads / 64];
    double r = -INFINITY;
    for (size_t k = threadIdx.x; k < n; k += kMaxThreads) r = fmax(r, absolute ? fabs(v[k]) : v[k]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sh[0];
        for (int w = 1; w < kMaxThreads / 64; ++w) t = fmax(t, sh[w]);
        *m = t;
    }
}

// land taper before smoothing: hamming(2 ts)[ts + j] on rows j < ts, 0 below
__global__ void k_taper0(int nx, int nz, int ts, double *__restrict__ t) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= nz || i >= nx) return;
    const double pi = 3.141592653589793;
    t[(size_t)i * nz + j] = j < ts ? 0.54 - 0.46 * cos(2.0 * pi * (ts + j) / (2 * ts - 1)) : 0.0;
}

// land taper from its smoothed form t and max m: ((-(t / m) (1 - thred)) + 1)^2
__global__ void k_taper_land(size_t n, double thred, const double *__restrict__ m, double *__restrict__ t) {
    const double mx = *m;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        double v = t[k] / mx;
        v = v * (1.0 - thred);
        v = -v + 1.0;
        t[k] = v * v;
    }
}

// precond = smoothed forw / max + smoothed back / max (in c)
__global__ void k_precond_sum(size_t n, const double *__restrict__ a, const double *__restrict__ b,
                              const double *__restrict__ m, double *__restrict__ c) {
    const double ma = m[0], mb = m[1];
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        c[k] = a[k] / ma + b[k] / mb;
}

// g = float(grad * taper) (the reference's in-place product on the float32 gradient;
// taper: marine 0 on rows j < mute, 1 below; land: the array lt; none: no product),
// then g / max(precond / max precond, 1e-4)
__global__ void k_apply(int nx, int nz, int taper_kind, int mute, const float *__restrict__ grad,
                        const double *__restrict__ lt, const double *__restrict__ c, const double *__restrict__ m,
                        double *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= nz || i >= nx) return;
    const size_t k = (size_t)i * nz + j;
    float g = grad[k];
    if (taper_kind == 1) g = (float)((double)g * (j < mute ? 0.0 : 1.0));
    else if (taper_kind == 2) g = (float)((double)g * lt[k]);
    double p = c[k] / m[2];
    if (p < 1e-4) p = 1e-4;
    out[k] = (double)g / p;
}

__global__ void k_mul(size_t n, const double *__restrict__ mask, double *__restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        out[k] = out[k] * mask[k];
}

// out *= vpmax / max|out|
__global__ void k_scale(size_t n, double vpmax, const double *__restrict__ m, double *__restrict__ out) {
    const double sc = vpmax / *m;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        out[k] *= sc;
}

}  // namespace pc
}  // namespace swit
