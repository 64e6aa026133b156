// Misfits and adjoint sources on the device (toolbox/misfit.py), one workgroup per
// trace, FP64 accumulation, and a fixed-order FP64 sum over traces (reproducible).
//
// Trace rules as misfit.py: a trace whose observed or synthetic L2 norm is zero
// contributes nothing and gets a zero adjoint source.
//   waveform (:128-148, :263-282): adj = syn - obs (FP32, as the reference's float32
//     arrays), misfit = sum_traces sqrt(sum r^2 dt);
//   global correlation (:219-242, :338-355): o = obs/|obs|, s = syn/|syn|,
//     c = Pearson coefficient of (s, o) (np.corrcoef, means removed),
//     misfit = -sum_traces c, adj = (s c - o) / |syn|;
//   envelope, power 2 (:151-185, :286-308): E = x^2 + H(x)^2 with H the discrete
//     Hilbert transform of scipy.signal.hilbert, r = E_syn - E_obs,
//     misfit = dt sum_traces sum r^2, adj = 2 s r - H(2 H(s) r);
//   traveltime (:188-215, :312-334, :359-373): lag = argmax |ifft(fft(obs) conj(-fft(syn)))|
//     folded as cross_correlate_max, d = central difference of syn (FP32 as the
//     reference's float32 traces), misfit = 0.5 (lag - nt + 1)^2 dt,
//     adj = d / (sum d^2 dt) * (lag - nt + 1) dt; a trace with sum |d| = 0 is dead.
// The reference takes its FFTs of float32 traces in single precision (scipy keeps
// float32 input in complex64); here every transform runs in FP64.
#include "fd_common.h"

namespace swit {

constexpr int kMisfitThreads = 256;

// Sum over a workgroup of T threads, the same fixed order on every call; every
// thread gets the sum.  sh holds T/64 doubles.
template <int T = kMisfitThreads>
__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();                      // sh may still be read from a previous call
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < T / 64; ++w) t += sh[w];
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

// ---- FFT misfits ------------------------------------------------------------------------
// One workgroup per trace holds the trace as nt complex doubles in LDS and transforms
// it in place.  The FFT is a Stockham autosort over the radices of nt (4s first, then 2,
// 3 and the odd primes).  A stage of radix R after ns points of earlier stages maps
// butterfly b (0 <= b < n/R), jm = b mod ns, inputs x[b + q n/R] (q < R), to outputs
//   y[(b - jm) R + jm + r ns] = sum_q x[b + q n/R] W^(q jm n/(ns R)) w_R^(q r),
// W = exp(-+2 pi i / n), w_R = W^(n/R).  Radices 2, 3, 4 and 5 run as register
// butterflies (one thread per butterfly: one table twiddle, its powers by products,
// R - 1 twiddle products, then the R-point DFT with its constants); larger primes run as a gather in which each thread sums the R
// inputs of its outputs with the twiddle powers formed by repeated products.  Twiddles
// come from two LDS tables: W^e = A[e >> kTwShift] B[e & (kTwB - 1)].  Two real traces
// go through one complex transform.
constexpr int kFftMaxStages = 24;
constexpr int kFftMaxN = 9216;
constexpr int kTwShift = 7;
constexpr int kTwB = 1 << kTwShift;                          // 128
constexpr int kTwA = (kFftMaxN + kTwB - 1) / kTwB;           // 72

struct FftFactors {
    int n, nst, s;              // length, stages, twiddle split (s = kTwB)
    int radix[kFftMaxStages];
};

// Threads per workgroup: 512, so that two workgroups of up to 4096 points share a CU
// (each barrier of one overlaps the other's work).
template <int NMAX>
constexpr int fft_threads() { return 512; }

template <int NMAX>
struct FftLds {
    static constexpr int T = fft_threads<NMAX>();
    double2 x[NMAX];
    double2 ta[kTwA], tb[kTwB];      // exp(-2 pi i a kTwB / n), exp(-2 pi i b / n)
    double sh[T / 64];
    int ish[T / 64];
};

template <int NMAX>
__device__ void fft_twiddles(FftLds<NMAX> &L, const FftFactors &f) {
    constexpr int T = fft_threads<NMAX>();
    const int na = (f.n + kTwB - 1) / kTwB;
    for (int i = threadIdx.x; i < na + kTwB; i += T) {
        const int e = i < na ? i * kTwB : i - na;
        double sv, cv;
        sincospi(2.0 * (double)e / (double)f.n, &sv, &cv);
        if (i < na) L.ta[i] = make_double2(cv, -sv);
        else L.tb[e] = make_double2(cv, -sv);
    }
}

// W^e (conjugated for the inverse), 0 <= e < n
template <int NMAX, bool INV>
__device__ __forceinline__ double2 fft_w(const FftLds<NMAX> &L, int e) {
    const double2 a = L.ta[e >> kTwShift], b = L.tb[e & (kTwB - 1)];
    const double re = a.x * b.x - a.y * b.y, im = a.x * b.y + a.y * b.x;
    return make_double2(re, INV ? -im : im);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double c) { return make_double2(c * a.x, c * a.y); }
// -i a (forward sign) or +i a (inverse)
template <bool INV> __device__ __forceinline__ double2 rot(double2 a) {
    return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// R-point DFT of z into y with the forward (-) or inverse (+) sign.
template <int R, bool INV> __device__ __forceinline__ void dft_small(const double2 *z, double2 *y);
template <bool INV> __device__ __forceinline__ void dft_small2(const double2 *z, double2 *y) {
    y[0] = cadd(z[0], z[1]);
    y[1] = csub(z[0], z[1]);
}
template <bool INV> __device__ __forceinline__ void dft_small3(const double2 *z, double2 *y) {
    const double h = 0.86602540378443864676;                       // sin(2 pi / 3)
    const double2 t1 = cadd(z[1], z[2]);
    const double2 t2 = csub(z[0], cscale(t1, 0.5));
    const double2 t3 = rot<INV>(cscale(csub(z[1], z[2]), h));
    y[0] = cadd(z[0], t1);
    y[1] = cadd(t2, t3);
    y[2] = csub(t2, t3);
}
template <bool INV> __device__ __forceinline__ void dft_small4(const double2 *z, double2 *y) {
    const double2 a = cadd(z[0], z[2]), b = csub(z[0], z[2]);
    const double2 c = cadd(z[1], z[3]), d = rot<INV>(csub(z[1], z[3]));
    y[0] = cadd(a, c);
    y[1] = cadd(b, d);
    y[2] = csub(a, c);
    y[3] = csub(b, d);
}
template <bool INV> __device__ __forceinline__ void dft_small5(const double2 *z, double2 *y) {
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;   // cos(2 pi k / 5)
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;    // sin(2 pi k / 5)
    const double2 a1 = cadd(z[1], z[4]), a2 = cadd(z[2], z[3]);
    const double2 b1 = csub(z[1], z[4]), b2 = csub(z[2], z[3]);
    const double2 t1 = make_double2(z[0].x + c1 * a1.x + c2 * a2.x, z[0].y + c1 * a1.y + c2 * a2.y);
    const double2 t2 = make_double2(z[0].x + c2 * a1.x + c1 * a2.x, z[0].y + c2 * a1.y + c1 * a2.y);
    const double2 u1 = rot<INV>(make_double2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    const double2 u2 = rot<INV>(make_double2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    y[0] = cadd(z[0], cadd(a1, a2));
    y[1] = cadd(t1, u1);
    y[4] = csub(t1, u1);
    y[2] = cadd(t2, u2);
    y[3] = csub(t2, u2);
}
template <> __device__ __forceinline__ void dft_small<2, false>(const double2 *z, double2 *y) { dft_small2<false>(z, y); }
template <> __device__ __forceinline__ void dft_small<2, true>(const double2 *z, double2 *y) { dft_small2<true>(z, y); }
template <> __device__ __forceinline__ void dft_small<3, false>(const double2 *z, double2 *y) { dft_small3<false>(z, y); }
template <> __device__ __forceinline__ void dft_small<3, true>(const double2 *z, double2 *y) { dft_small3<true>(z, y); }
template <> __device__ __forceinline__ void dft_small<4, false>(const double2 *z, double2 *y) { dft_small4<false>(z, y); }
template <> __device__ __forceinline__ void dft_small<4, true>(const double2 *z, double2 *y) { dft_small4<true>(z, y); }
template <> __device__ __forceinline__ void dft_small<5, false>(const double2 *z, double2 *y) { dft_small5<false>(z, y); }
template <> __device__ __forceinline__ void dft_small<5, true>(const double2 *z, double2 *y) { dft_small5<true>(z, y); }

// W_R^m = cos(2 pi m / R) - i sin(2 pi m / R) for the composite butterflies (R = 10),
// each constant the double nearest its exact value.
template <int R> struct Wtab;
template <> struct Wtab<10> {
    static constexpr double a = 0.80901699437494742410, b = 0.30901699437494742410;   // cos 36, cos 72 deg
    static constexpr double e = 0.58778525229247312917, f = 0.95105651629515357212;   // sin 36, sin 72 deg
    static constexpr double c[10] = {1.0, a, b, -b, -a, -1.0, -a, -b, b, a};
    static constexpr double s[10] = {0.0, e, f, f, e, 0.0, -e, -f, -f, -e};
};
// a W_R^m (conjugated for the inverse); multiples of a quarter turn are exact rotations
template <int R, bool INV>
__device__ __forceinline__ double2 twiddle_const(double2 a, int m) {
    if (m == 0) return a;
    if (4 * m == R) return rot<INV>(a);
    if (2 * m == R) return make_double2(-a.x, -a.y);
    if (4 * m == 3 * R) return rot<!INV>(a);
    return cmul(a, make_double2(Wtab<R>::c[m], INV ? Wtab<R>::s[m] : -Wtab<R>::s[m]));
}

// R = A B point DFT from size-A DFTs over the B interleaved subsequences z[B n1 + n2],
// the twiddles W_R^(n2 k1), then size-B DFTs: y[k1 + A k2] (Cooley-Tukey, in registers).
template <int A, int B, bool INV>
__device__ __forceinline__ void dft_comp(const double2 *z, double2 *y) {
    double2 t[A * B];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
        double2 in[A], out[A];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) in[n1] = z[B * n1 + n2];
        dft_small<A, INV>(in, out);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) t[n2 * A + k1] = twiddle_const<A * B, INV>(out[k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        double2 in[B], out[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) in[n2] = t[n2 * A + k1];
        dft_small<B, INV>(in, out);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) y[k1 + A * k2] = out[k2];
    }
}
template <> __device__ __forceinline__ void dft_small<10, false>(const double2 *z, double2 *y) { dft_comp<5, 2, false>(z, y); }
template <> __device__ __forceinline__ void dft_small<10, true>(const double2 *z, double2 *y) { dft_comp<5, 2, true>(z, y); }

// One butterfly stage of radix R (2..5, 10), in place: all reads, barrier, all writes,
// barrier.
template <int NMAX, bool INV, int R>
__device__ void fft_stage_bfly(FftLds<NMAX> &L, int n, int ns) {
    constexpr int T = fft_threads<NMAX>();
    constexpr int BPT = (NMAX / R + T - 1) / T;      // butterflies per thread
    const int nb = n / R, mult = nb / ns;            // n / (ns R)
    double2 y[BPT][R];
    int o0[BPT];
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        const int b = threadIdx.x + k * T;
        o0[k] = -1;
        if (b < nb) {
            const int jm = b % ns;
            o0[k] = (b - jm) * R + jm;
            // twiddle powers w^q from w = W^(jm n/(ns R)) by products
            double2 z[R];
            const double2 w1 = fft_w<NMAX, INV>(L, jm * mult);
            double2 wq = w1;
            z[0] = L.x[b];
#pragma unroll
            for (int q = 1; q < R; ++q) {
                if (q >= 2) wq = cmul(wq, w1);
                z[q] = cmul(L.x[b + q * nb], wq);
            }
            dft_small<R, INV>(z, y[k]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        if (o0[k] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) L.x[o0[k] + r * ns] = y[k][r];
        }
    }
    __syncthreads();
}

// One stage of any radix as a gather over outputs, in place (reads, barrier, writes,
// barrier).
template <int NMAX, bool INV>
__device__ void fft_stage_gather(FftLds<NMAX> &L, int n, int R, int ns) {
    constexpr int T = fft_threads<NMAX>(), E = (NMAX + T - 1) / T;
    const int nr = n / R, span = ns * R, mult = n / span;
    double2 acc[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int o = threadIdx.x + k * T;
        double re = 0.0, im = 0.0;
        if (o < n) {
            const int jm = o % ns, q = (o / ns) % R;
            const int j = (o / span) * ns + jm;
            const double2 w1 = fft_w<NMAX, INV>(L, (jm + q * ns) * mult);     // step < n
            double wr = 1.0, wi = 0.0;
            for (int r = 0; r < R; ++r) {
                const double2 v = L.x[j + r * nr];
                re = fma(v.x, wr, fma(-v.y, wi, re));
                im = fma(v.x, wi, fma(v.y, wr, im));
                const double t = fma(wr, w1.x, -wi * w1.y);
                wi = fma(wr, w1.y, wi * w1.x);
                wr = t;
            }
        }
        acc[k] = make_double2(re, im);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int o = threadIdx.x + k * T;
        if (o < n) L.x[o] = acc[k];
    }
    __syncthreads();
}

// In-place DFT of L.x[0..n) (INV: inverse without the 1/n factor).  Begins and ends
// with a barrier.
template <int NMAX, bool INV>
__device__ void fft_lds(FftLds<NMAX> &L, const FftFactors &f) {
    const int n = f.n;
    int ns = 1;
    __syncthreads();
    for (int st = 0; st < f.nst; ++st) {
        const int R = f.radix[st];
        switch (R) {
        case 2: fft_stage_bfly<NMAX, INV, 2>(L, n, ns); break;
        case 3: fft_stage_bfly<NMAX, INV, 3>(L, n, ns); break;
        case 4: fft_stage_bfly<NMAX, INV, 4>(L, n, ns); break;
        case 5: fft_stage_bfly<NMAX, INV, 5>(L, n, ns); break;
        case 10: fft_stage_bfly<NMAX, INV, 10>(L, n, ns); break;
        default: fft_stage_gather<NMAX, INV>(L, n, R, ns); break;
        }
        ns *= R;
    }
}

// x <- -i sgn(k) x, sgn as scipy.signal.hilbert's h - 1: 0 at k = 0 (and k = n/2 for
// even n), +1 below n/2, -1 above.  After the inverse transform, the real and
// imaginary parts are the Hilbert transforms of the two packed real traces.
template <int NMAX>
__device__ void hilbert_multiplier(FftLds<NMAX> &L, int n) {
    constexpr int T = fft_threads<NMAX>();
    for (int k = threadIdx.x; k < n; k += T) {
        const double sg = (k == 0 || 2 * k == n) ? 0.0 : (2 * k < n ? 1.0 : -1.0);
        const double2 v = L.x[k];
        L.x[k] = make_double2(sg * v.y, -sg * v.x);
    }
}

// Envelope misfit (power 2) of one trace per workgroup.
template <int NMAX>
__global__ __launch_bounds__(fft_threads<NMAX>())
__attribute__((amdgpu_waves_per_eu(NMAX <= 4096 ? 4 : 1, 8))) void k_misfit_envelope(FftFactors f, float dt,
                                                                 const float *__restrict__ obs,
                                                                 const float *__restrict__ syn,
                                                                 float *__restrict__ adj,
                                                                 double *__restrict__ tmis) {
    constexpr int T = fft_threads<NMAX>(), E = (NMAX + T - 1) / T;
    __shared__ FftLds<NMAX> L;
    const int n = f.n;
    const size_t base = (size_t)blockIdx.x * n;
    const float *o = obs + base, *s = syn + base;
    float *a = adj + base;
    fft_twiddles(L, f);
    double so = 0.0, ss = 0.0;
    for (int i = threadIdx.x; i < n; i += T) {
        const double ov = o[i], sv = s[i];
        so += ov * ov;
        ss += sv * sv;
        L.x[i] = make_double2(sv, ov);
    }
    so = block_sum<T>(so, L.sh);
    ss = block_sum<T>(ss, L.sh);
    if (!(so > 0.0 && ss > 0.0)) {                    // uniform over the workgroup
        for (int i = threadIdx.x; i < n; i += T) a[i] = 0.0f;
        if (threadIdx.x == 0) tmis[blockIdx.x] = 0.0;
        return;
    }
    fft_lds<NMAX, false>(L, f);
    hilbert_multiplier(L, n);
    fft_lds<NMAX, true>(L, f);
    const double inv_n = 1.0 / n;
    double part[E];
    double r2 = 0.0;
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int i = threadIdx.x + k * T;
        part[k] = 0.0;
        if (i < n) {
            const double2 h = L.x[i];
            const double sh = h.x * inv_n, oh = h.y * inv_n;
            const double sv = s[i], ov = o[i];
            const double r = (sv * sv + sh * sh) - (ov * ov + oh * oh);
            r2 += r * r;
            part[k] = 2.0 * sv * r;
            L.x[i] = make_double2(2.0 * sh * r, 0.0);
        }
    }
    r2 = block_sum<T>(r2, L.sh);
    fft_lds<NMAX, false>(L, f);
    hilbert_multiplier(L, n);
    fft_lds<NMAX, true>(L, f);
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int i = threadIdx.x + k * T;
        if (i < n) a[i] = (float)(part[k] - L.x[i].x * inv_n);
    }
    if (threadIdx.x == 0) tmis[blockIdx.x] = r2 * (double)dt;
}

// Cross-correlation traveltime misfit of one trace per workgroup.
template <int NMAX>
__global__ __launch_bounds__(fft_threads<NMAX>())
__attribute__((amdgpu_waves_per_eu(NMAX <= 4096 ? 4 : 1, 8))) void k_misfit_traveltime(FftFactors f, float dt,
                                                                   const float *__restrict__ obs,
                                                                   const float *__restrict__ syn,
                                                                   float *__restrict__ adj,
                                                                   double *__restrict__ tmis) {
    constexpr int T = fft_threads<NMAX>(), E = (NMAX + T - 1) / T;
    __shared__ FftLds<NMAX> L;
    const int n = f.n;
    const size_t base = (size_t)blockIdx.x * n;
    const float *o = obs + base, *s = syn + base;
    float *a = adj + base;
    const float two_dt = 2.0f * dt;
    fft_twiddles(L, f);
    double so = 0.0, ss = 0.0, dabs = 0.0, d2 = 0.0;
    for (int i = threadIdx.x; i < n; i += T) {
        const double ov = o[i], sv = s[i];
        so += ov * ov;
        ss += sv * sv;
        L.x[i] = make_double2(ov, sv);
        if (i >= 1 && i + 1 < n) {
            const double d = (s[i + 1] - s[i - 1]) / two_dt;
            dabs += fabs(d);
            d2 += d * d;
        }
    }
    so = block_sum<T>(so, L.sh);
    ss = block_sum<T>(ss, L.sh);
    dabs = block_sum<T>(dabs, L.sh);
    d2 = block_sum<T>(d2, L.sh);
    if (!(so > 0.0 && ss > 0.0 && dabs > 0.0)) {     // uniform over the workgroup
        for (int i = threadIdx.x; i < n; i += T) a[i] = 0.0f;
        if (threadIdx.x == 0) tmis[blockIdx.x] = 0.0;
        return;
    }
    fft_lds<NMAX, false>(L, f);
    // Z = O + i S  ->  P = O * conj(-S), with O = (Z_k + conj Z_-k)/2, S = (Z_k - conj Z_-k)/(2i);
    // P_-k = conj P_k.
    double2 pk[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int i = threadIdx.x + k * T;
        pk[k] = make_double2(0.0, 0.0);
        if (2 * i <= n) {
            const double2 z1 = L.x[i], z2 = L.x[i == 0 ? 0 : n - i];
            const double orr = 0.5 * (z1.x + z2.x), oi = 0.5 * (z1.y - z2.y);
            const double sr = 0.5 * (z1.y + z2.y), si = -0.5 * (z1.x - z2.x);
            // O * conj(-S) = -(O * conj S)
            pk[k] = make_double2(-(orr * sr + oi * si), -(oi * sr - orr * si));
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int i = threadIdx.x + k * T;
        if (2 * i <= n) {
            L.x[i] = pk[k];
            if (i != 0) L.x[n - i] = make_double2(pk[k].x, -pk[k].y);
        }
    }
    fft_lds<NMAX, true>(L, f);
    // first index of the largest |cc| (numpy argmax)
    double best = -1.0;
    int bi = 0;
    for (int i = threadIdx.x; i < n; i += T) {
        const double2 c = L.x[i];
        const double m = c.x * c.x + c.y * c.y;
        if (m > best) { best = m; bi = i; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { L.sh[wave] = best; L.ish[wave] = bi; }
    __syncthreads();
    best = L.sh[0];
    bi = L.ish[0];
    for (int w = 1; w < T / 64; ++w)
        if (L.sh[w] > best || (L.sh[w] == best && L.ish[w] < bi)) { best = L.sh[w]; bi = L.ish[w]; }
    int cc = bi - 1;
    if (cc < n / 2) cc += n;
    const double lag = (double)(cc - n + 1);
    const double inv = 1.0 / (d2 * (double)dt), sc = lag * (double)dt;
    for (int i = threadIdx.x; i < n; i += T) {
        double d = 0.0;
        if (i >= 1 && i + 1 < n) d = (double)((s[i + 1] - s[i - 1]) / two_dt);
        a[i] = (float)((d * inv) * sc);
    }
    if (threadIdx.x == 0) tmis[blockIdx.x] = 0.5 * lag * lag * (double)dt;
}

}  // namespace swit
