// Batched path: every shot of the job in each launch, one fused launch per time step.
//
// Forward step kernel: a workgroup owns a TZ x TX tile for a group of shots.  Per
// shot it stages the old p, u, w of the tile plus a 3-cell halo in LDS, computes the
// new p on the tile plus the halo the velocity stencils read, then the new u, w on
// the tile, and writes the tile to the other buffer (double-buffered state: halo reads
// of a step always see the previous step).  Receiver samples and boundary strips of
// the tile are written in the same launch.
//
// Backward step kernel: a workgroup owns one tile of one shot.  It runs the
// reverse-time reconstruction of the forward wavefield on the interior window and the
// adjoint step on the padded grid through the same LDS tile, then adds the zero-lag
// imaging terms to the shot's FP32 accumulators in the reference's order; the
// cross-shot sum is FP64 at the end of the job.  (Processing several shots per
// workgroup here needs about twice the registers and runs slower.)
//
// Block -> tile order is XCD-aware: blocks are dealt round-robin to the 8 XCDs, so
// the j-th block of XCD k takes the j-th item of a contiguous range of work items;
// items walk a group of tile columns for every shot before moving on, so the group's
// coefficients and shared halos stay in that XCD's L2 (see work_item).
//
// Memory layout (device): state buffers [ns][3][nxp][ld] (p, u, w of a shot
// contiguous), reconstruction buffers [ns][3][nx][nz]; coefficients per padded cell
// CP = {den c^2 dt/dx, 1 - damp dt} and CV = {dt/(den dx), 1 - kappa_u, 1 - kappa_w}.
#include "fd_common.h"
#include "fd_kernels.h"

#include <array>
#include <vector>

namespace swit {

constexpr int kTX = 16;          // output columns of a two-step backward tile
constexpr int kHalo = 3;
constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr int kNumXcd = 8;

// Forward step kernel tile: kTZ x kFTX outputs, LDS window (kTZ+6) x (kFTX+6).
namespace fk {
constexpr int kTZ = 64;
constexpr int kFTX = 16;
constexpr int kRZ = kTZ + 2 * kHalo;
constexpr int kRX = kFTX + 2 * kHalo;
constexpr int kNR = kRZ * kRX;
constexpr int kCellsPerThread = kTZ * kFTX / kThreads;
// region of new p the velocity stencils read: (kTZ+3) x (kFTX+3) cells
constexpr int kQZ = kTZ + 3;
constexpr int kQX = kFTX + 3;
constexpr int kNQ = kQZ * kQX;
constexpr int kLPT = (kNR + kThreads - 1) / kThreads;   // staged cells per thread
constexpr int kQPT = (kNQ + kThreads - 1) / kThreads;   // region cells per thread
// LDS index of padded cell (z, x) in a window whose origin is (zo, xo).
__device__ __forceinline__ int L(int z, int x, int zo, int xo) { return (x - xo) * kRZ + (z - zo); }
}  // namespace fk

// Backward step kernel tile: kTZ x kBTX outputs.  Its LDS window is one wave tall
// (kZL = 64 rows: the outputs plus a 3-row halo above and below) and kCX columns wide;
// lane l of every wave owns window row l, so each row predicate is evaluated once per
// thread and every column loop is uniform across the wave.
namespace bk {
constexpr int kZL = 64;
constexpr int kTZ = kZL - 2 * kHalo;   // 58 output rows
constexpr int kBTX = 24;               // output columns (the default width)
constexpr int kBTXs[2] = {16, 24};     // widths the kernel is built for (autotuner candidates)
constexpr int kCX = kBTX + 2 * kHalo;  // 30 window columns
constexpr int kNL = kZL * kCX;         // floats per LDS array
// LDS index of window column c, row (lane) l.
__device__ __forceinline__ int LW(int c, int l) { return c * kZL + l; }
}  // namespace bk

// Tiling of the padded grid: tile (zt, xt) covers rows
// [max(ztop + zt*th, 1), min(ztop + (zt+1)*th - 1, nzp)] and columns
// [1 + xt*tx, min((xt+1)*tx, nxp)].
struct TileMap {
    int ztop;
    int th;         // tile height (rows)
    int tx;         // tile width (columns)
    int nzt, nxt;
    int ntiles;     // nzt * nxt
    int cg;         // tile columns per column group of the work order (see work_item)
};

struct TileBox {
    int z0, z1, x0, x1;
};

__host__ __device__ inline TileBox tile_box(const Dims &d, const TileMap &tm, int t) {
    int zt = t % tm.nzt, xt = t / tm.nzt;
    TileBox b;
    b.z0 = max(tm.ztop + zt * tm.th, 1);
    b.z1 = min(tm.ztop + (zt + 1) * tm.th - 1, d.nzp);
    b.x0 = 1 + xt * tm.tx;
    b.x1 = min((xt + 1) * tm.tx, d.nxp);
    return b;
}

// XCD-aware remap of a linear block id onto a logical work item.
__device__ __forceinline__ int xcd_logical(int b, int nb) {
    int q = nb / kNumXcd, r = nb % kNumXcd;
    int k = b % kNumXcd, j = b / kNumXcd;
    return k * q + min(k, r) + j;
}

// Logical work item -> (shot group, tile).  Items run through groups of tm.cg tile
// columns; inside a group every shot group visits all of the group's tiles before the
// next one starts.  A group's coefficients (shared by all shots) and the halos its
// tiles share stay in the XCD's L2 while the shots stream through it.
constexpr int kColGroup = 4;     // default; the autotuner may choose another
__device__ __forceinline__ void work_item(int w, int nshot, const TileMap &tm, int &s, int &t) {
    const int per = tm.cg * tm.nzt;                  // tiles of a full column group
    const int gq = w / (per * nshot);
    const int rem = w - gq * per * nshot;
    const int cols = min(tm.cg, tm.nxt - gq * tm.cg);
    const int tiles = cols * tm.nzt;
    s = rem / tiles;
    t = gq * per + rem % tiles;
}

// Split of one launch's logical work items over the XCDs: XCD k takes items
// xs[k] .. xs[k+1]-1, its j-th block the item xs[k] + j (blocks past the range exit at
// once); nsh = shots in the launch.
struct XcdSplit {
    int xs[kNumXcd + 1];
    int nsh;
};

// Strip slot of a tile: slot index and source cell.
struct StripRef {
    int slot;
    int field;      // 0 p, 1 u, 2 w
    int iz, ix;
};

// ---- 16-bit boundary strips ---------------------------------------------------------
// Strip mode SM: 0 FP32 (the reference's values), 1 FP16, 2 BF16.  In the 16-bit modes
// the values of one shot and step are scaled by 2^k, the power of two that puts the
// step's max |v| over the shot's strip (m) into [2^14, 2^15): k = min(126, 141 - e(m))
// with e(m) the biased FP32 exponent.  The scaling is exact, so the value is rounded
// once (to nearest even) into FP16 or BF16; decoding multiplies by 2^-k (exact).  Only
// m is stored per row.
__device__ __forceinline__ int strip_shift(uint32_t mbits) { return min(126, 141 - (int)((mbits >> 23) & 0xffu)); }
__device__ __forceinline__ float strip_inv(uint32_t mbits) {
    return __uint_as_float((uint32_t)(strip_shift(mbits) + 127) << 23);
}
__device__ __forceinline__ float strip_scale(uint32_t mbits) {
    return __uint_as_float((uint32_t)(127 - strip_shift(mbits)) << 23);
}

template <int SM> __device__ __forceinline__ uint16_t strip_encode(float v);
template <> __device__ __forceinline__ uint16_t strip_encode<1>(float v) {
    return __builtin_bit_cast(uint16_t, (_Float16)v);
}
template <> __device__ __forceinline__ uint16_t strip_encode<2>(float v) {
    uint32_t b = __float_as_uint(v);
    b += 0x7fffu + ((b >> 16) & 1u);          // round to nearest even (finite values)
    return (uint16_t)(b >> 16);
}
template <int SM> __device__ __forceinline__ float strip_decode(uint16_t h);
template <> __device__ __forceinline__ float strip_decode<1>(uint16_t h) {
    return (float)__builtin_bit_cast(_Float16, h);
}
template <> __device__ __forceinline__ float strip_decode<2>(uint16_t h) {
    return __uint_as_float((uint32_t)h << 16);
}

// One strip row (one shot, one step) as the backward kernel reads it.
template <int SM> struct StripRow {
    const uint16_t *b;
    float scale;
    __device__ __forceinline__ StripRow(const void *base, size_t row, int S, const uint32_t *smax)
        : b(static_cast<const uint16_t *>(base) + row * S), scale(strip_scale(smax[row])) {}
    __device__ __forceinline__ float operator[](int i) const { return strip_decode<SM>(b[i]) * scale; }
};
template <> struct StripRow<0> {
    const float *b;
    __device__ __forceinline__ StripRow(const void *base, size_t row, int S, const uint32_t *)
        : b(static_cast<const float *>(base) + row * S) {}
    __device__ __forceinline__ float operator[](int i) const { return b[i]; }
};

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Encode strip rows first_row .. first_row + gridDim.x - 1 of every shot (grid (rows, ns),
// kEncThreads threads) from the FP32 ring ring[ns][nring][S] (row r in slot r % nring):
// the row's max |v| goes to smax (float bits) and each value to 16 bits at that scale.
constexpr int kEncThreads = 256;
constexpr int kStripRing = 32;     // rows of the FP32 ring; encoded in batches of up to this many
template <int SM>
__global__ __launch_bounds__(kEncThreads) void k_strip_encode(Dims d, int first_row, int nring, int s0,
                                                              const float *__restrict__ ring,
                                                              uint32_t *__restrict__ smax,
                                                              uint16_t *__restrict__ strips16) {
    __shared__ float sh[kEncThreads / 64];
    const int s = s0 + (int)blockIdx.y, row = first_row + blockIdx.x;
    const float *v = ring + ((size_t)s * nring + row % nring) * d.S;
    float m = 0.0f;
    for (int i = threadIdx.x; i < d.S; i += kEncThreads) m = fmaxf(m, fabsf(v[i]));
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kEncThreads / 64; ++w) m = fmaxf(m, sh[w]);
    const size_t r = (size_t)s * d.nt + row;
    const uint32_t mb = __float_as_uint(m);
    if (threadIdx.x == 0) smax[r] = mb;
    const float inv = strip_inv(mb);
    uint16_t *out = strips16 + r * d.S;
    for (int i = threadIdx.x; i < d.S; i += kEncThreads) out[i] = strip_encode<SM>(v[i] * inv);
}

// Decode 16-bit strip rows to FP32: n = ns * nt * S values, row r = index / S.
template <int SM>
__global__ void k_strip_decode(size_t n, int S, const uint16_t *__restrict__ strips16,
                               const uint32_t *__restrict__ smax, float *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const size_t r = i / (size_t)S;
        out[i] = strip_decode<SM>(strips16[i]) * strip_scale(smax[r]);
    }
}

// ---- small helpers (kernels only, so captured graphs hold no memset nodes) -------
__global__ void k_fill(float *__restrict__ p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
__global__ void k_fill_d(double *__restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0;
}
// Zero the first `width` floats of each of gridDim.x blocks spaced `stride` apart.
__global__ void k_zero_rows(float *__restrict__ p, size_t stride, int width) {
    float *q = p + (size_t)blockIdx.x * stride;
    for (int i = threadIdx.x; i < width; i += blockDim.x) q[i] = 0.0f;
}
__global__ void k_zero_rows16(uint16_t *__restrict__ p, size_t stride, int width) {
    uint16_t *q = p + (size_t)blockIdx.x * stride;
    for (int i = threadIdx.x; i < width; i += blockDim.x) q[i] = 0;
}
__global__ void k_copy(const float *__restrict__ a, float *__restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

// Per-cell coefficient vectors from K, B and the damping profile.  The factors are
// the reference's (1.0 - kappa) with kappa = damp*dt (fdcore.f90:80),
// 0.5*(damp(iz,ix)+damp(iz,ix+1))*dt (:100) and 0.5*(damp(iz,ix)+damp(iz+1,ix))*dt
// (:109), each rounded exactly as there.
__global__ void k_setup_cvec(Dims d, const float *__restrict__ K, const float *__restrict__ B,
                             const float *__restrict__ prof, float2 *__restrict__ CP,
                             float4 *__restrict__ CV, float2 *__restrict__ KB) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (iz > d.nzp || ix > d.nxp) return;
    size_t k = pidx(d, iz, ix);
    float d0 = damp_at(d, prof, iz, ix);
    float dxp = ix < d.nxp ? damp_at(d, prof, iz, ix + 1) : 0.0f;
    float dzp = iz < d.nzp ? damp_at(d, prof, iz + 1, ix) : 0.0f;
    float kp = d0 * d.dt;
    float ku = 0.5f * (d0 + dxp) * d.dt;
    float kw = 0.5f * (d0 + dzp) * d.dt;
    CP[k] = make_float2(K[k], 1.0f - kp);
    CV[k] = make_float4(B[k], 1.0f - ku, 1.0f - kw, 0.0f);
    KB[k] = make_float2(K[k], B[k]);
}

// ---------------------------------------------------------------------------------
// Forward step it (fdcore.f90:75-160) for a group of G shots per workgroup.
// In/Out: [ns][3][nxp][ld].  LDS window origin (zo, xo) = (z0-3, x0-3).
// Divergence history of the exact adjoint: the divergence term of step it of shot s goes
// to p[(s * slots + it - 1 - base) * np + cell] (all steps: slots = nt, base = 0; one
// checkpoint segment after state `base`: slots = its length).  p null: not saved.
struct DSave {
    float *p;
    int slots, base;
};

// strips: null (no saving) or FP32 rows; step it's row of shot s is
// strips[(s * s_rows + s_row) * S] (the plan's strips with s_rows = nt, s_row = it - 1, or
// a ring of rows that k_strip_encode turns into 16-bit rows).
__global__ __launch_bounds__(kThreads) void k_fwd_fused(
    Dims d, int it, TileMap tm, int s0, int ns, int G, const float2 *__restrict__ CP, const float4 *__restrict__ CV,
    const ShotGeom *__restrict__ geo, const float *__restrict__ src, const float *__restrict__ In,
    float *__restrict__ Out, const int *__restrict__ rec_start, const RecRef *__restrict__ rec_list,
    float *__restrict__ seis_t, const int *__restrict__ strip_start, const StripRef *__restrict__ strip_list,
    float *__restrict__ strips, int s_rows, int s_row, DSave dsave) {
    using namespace fk;
    __shared__ float sP[kNR];
    __shared__ float sU[kNR];
    __shared__ float sW[kNR];
    __shared__ float sQ[kNR];     // new p

    // shots s0 .. s0 + ns - 1 in ceil(ns / G) groups of near-equal size (at most G), so
    // that every workgroup of the launch carries about the same work
    const int ngr = (ns + G - 1) / G;
    int grp, t;
    work_item(xcd_logical(blockIdx.x, gridDim.x), ngr, tm, grp, t);
    const TileBox tb = tile_box(d, tm, t);
    const int zo = tb.z0 - kHalo, xo = tb.x0 - kHalo;
    const int tid = threadIdx.x;
    const int fs = d.fs;
    const int np = d.nxp * d.ld;
    const int s_begin = s0 + grp * ns / ngr, s_end = s0 + (grp + 1) * ns / ngr;

    // staged fields of the first shot
    float lp[kLPT], lu[kLPT], lw[kLPT];
    auto stage_load = [&](int s) {
        const float *pin = In + (size_t)s * 3 * np;
#pragma unroll
        for (int j = 0; j < kLPT; ++j) {
            int i = tid + j * kThreads;
            int z = zo + i % kRZ, x = xo + i / kRZ;
            lp[j] = lu[j] = lw[j] = 0.f;
            if (i < kNR && z >= 1 && z <= d.nzp && x >= 1 && x <= d.nxp) {
                const unsigned k = (unsigned)((x - 1) * d.ld + (z - 1));
                lp[j] = pin[k];
                lu[j] = pin[np + k];
                lw[j] = pin[2 * np + k];
            }
        }
    };
    stage_load(s_begin);
    // coefficients of this thread's region and tile cells, shared by every shot
    float2 cp[kQPT];
    unsigned qgrid = 0, qupd = 0;
#pragma unroll
    for (int c = 0; c < kQPT; ++c) {
        int i = tid + c * kThreads;
        int z = tb.z0 - 1 + i % kQZ, x = tb.x0 - 1 + i / kQZ;
        cp[c] = make_float2(0.f, 0.f);
        if (i < kNQ && z >= 1 && z <= d.nzp && x >= 1 && x <= d.nxp) {
            qgrid |= 1u << c;
            if (z >= fs + 1 && z <= d.nzp - 2 && x >= 3 && x <= d.nxp - 2) {
                qupd |= 1u << c;
                cp[c] = CP[(unsigned)((x - 1) * d.ld + (z - 1))];
            }
        }
    }
    float4 cv[kCellsPerThread];
    unsigned tur = 0, twr = 0, tin = 0;
#pragma unroll
    for (int c = 0; c < kCellsPerThread; ++c) {
        int i = tid + c * kThreads;
        int z = tb.z0 + i % kTZ, x = tb.x0 + i / kTZ;
        cv[c] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (z <= tb.z1 && x <= tb.x1) {
            tin |= 1u << c;
            bool ur = z >= fs && z <= d.nzp - 1 && x >= 2 && x <= d.nxp - 2;
            bool wr = z >= fs && z <= d.nzp - 2 && x >= 2 && x <= d.nxp - 1;
            if (ur) tur |= 1u << c;
            if (wr) twr |= 1u << c;
            if (ur || wr) cv[c] = CV[(unsigned)((x - 1) * d.ld + (z - 1))];
        }
    }

    for (int s = s_begin; s < s_end; ++s) {
        const ShotGeom g = geo[s];
        const float srcv = d.dt * src[(size_t)s * d.nt + (it - 1)];
#pragma unroll
        for (int j = 0; j < kLPT; ++j) {
            int i = tid + j * kThreads;
            if (i < kNR) {
                sP[i] = lp[j];
                sU[i] = lu[j];
                sW[i] = lw[j];
            }
        }
        __syncthreads();
        if (s + 1 < s_end) stage_load(s + 1);   // next shot's loads overlap this shot
        // new p on rows z0-1..z0+TZ+1, cols x0-1..x0+TX+1
#pragma unroll
        for (int c = 0; c < kQPT; ++c) {
            int i = tid + c * kThreads;
            if (i >= kNQ) continue;
            int z = tb.z0 - 1 + i % kQZ, x = tb.x0 - 1 + i / kQZ;
            int l = L(z, x, zo, xo);
            float v = 0.f;
            if (qgrid >> c & 1u) {
                v = sP[l];
                if (qupd >> c & 1u) {
                    const float dv = div_term(sU[l], sU[l - kRZ], sU[l + kRZ], sU[l - 2 * kRZ], sW[l], sW[l - 1],
                                              sW[l + 1], sW[l - 2]);
                    v = cp[c].y * v - cp[c].x * dv;
                    // divergence history for the exact adjoint: tile cells, slot it-1
                    if (dsave.p && z >= tb.z0 && z <= tb.z1 && x >= tb.x0 && x <= tb.x1)
                        dsave.p[((size_t)s * dsave.slots + (it - 1 - dsave.base)) * np +
                                (unsigned)((x - 1) * d.ld + (z - 1))] = dv;
                    if (z == g.isz && x == g.isx) v = v + srcv;
                }
            }
            sQ[l] = v;
        }
        __syncthreads();
        // p mirror (fdcore.f90:92-94) on the region columns
        if (d.free_surface == 1 && fs - 1 >= tb.z0 - 1 && fs + 1 <= tb.z0 + kTZ + 1) {
            for (int x = tb.x0 - 1 + tid; x <= tb.x0 + kFTX + 1; x += kThreads)
                if (x >= 1 && x <= d.nxp) sQ[L(fs - 1, x, zo, xo)] = -sQ[L(fs + 1, x, zo, xo)];
        }
        __syncthreads();
        // new u, w on the tile (in place: each cell reads only its own old velocity)
#pragma unroll
        for (int c = 0; c < kCellsPerThread; ++c) {
            int i = tid + c * kThreads;
            int z = tb.z0 + i % kTZ, x = tb.x0 + i / kTZ;
            int l = L(z, x, zo, xo);
            if (tur >> c & 1u) sU[l] = v_update_f(sU[l], cv[c].x, cv[c].y, sQ[l], sQ[l + kRZ], sQ[l + 2 * kRZ], sQ[l - kRZ]);
            if (twr >> c & 1u) sW[l] = v_update_f(sW[l], cv[c].x, cv[c].z, sQ[l], sQ[l + 1], sQ[l + 2], sQ[l - 1]);
        }
        __syncthreads();
        // w mirror (fdcore.f90:116-120): rows fs-1 and fs share the tile that starts at fs-1
        if (d.free_surface == 1 && fs - 1 >= tb.z0 && fs <= tb.z1) {
            for (int x = tb.x0 + tid; x <= tb.x1; x += kThreads) sW[L(fs - 1, x, zo, xo)] = sW[L(fs, x, zo, xo)];
        }
        __syncthreads();
        float *pout = Out + (size_t)s * 3 * np;
#pragma unroll
        for (int c = 0; c < kCellsPerThread; ++c) {
            if (!(tin >> c & 1u)) continue;
            int i = tid + c * kThreads;
            int z = tb.z0 + i % kTZ, x = tb.x0 + i / kTZ;
            int l = L(z, x, zo, xo);
            const unsigned k = (unsigned)((x - 1) * d.ld + (z - 1));
            pout[k] = sQ[l];
            pout[np + k] = sU[l];
            pout[2 * np + k] = sW[l];
        }
        // receivers (fdcore.f90:123-128)
        {
            int r0 = rec_start[(size_t)s * (tm.ntiles + 1) + t];
            int r1 = rec_start[(size_t)s * (tm.ntiles + 1) + t + 1];
            for (int r = r0 + tid; r < r1; r += kThreads) {
                RecRef rr = rec_list[(size_t)s * d.ngmax + r];
                seis_t[((size_t)s * d.nt + (it - 1)) * d.ngmax + rr.idx] = sQ[L(rr.iz, rr.ix, zo, xo)];
            }
        }
        // strips (fdcore.f90:130-159)
        if (strips) {
            int a0 = strip_start[t], a1 = strip_start[t + 1];
            float *row = strips + ((size_t)s * s_rows + s_row) * d.S;
            for (int a = a0 + tid; a < a1; a += kThreads) {
                StripRef sr = strip_list[a];
                int l = L(sr.iz, sr.ix, zo, xo);
                row[sr.slot] = sr.field == 0 ? sQ[l] : (sr.field == 1 ? sU[l] : sW[l]);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// Backward step it (solver.f90:230-253): reconstruction (fdcore.f90:202-289), adjoint
// step (fdcore.f90:293-375) and imaging (solver.f90:237-239); one shot per workgroup.
// The reconstruction window and the adjoint window are staged together, straight from
// global memory into LDS, and their velocity and pressure updates share barriers.
// A thread owns window rows = its lane and the columns 1 + wave + 8j; its tile cells are
// the owned cells inside the tile, so each cell's {K, B} is loaded once for both the
// velocity and the pressure pass.  The damping factors (1 - kappa) are loaded only by
// tiles whose updated cells reach a damped cell; elsewhere they are exactly 1.
// RIn/ROut: [ns][3][nx][nz]; BIn/BOut: [ns][3][nxp][ld].
// acc: [ns][3][nx][nz] FP32 per-shot accumulators (kernel, forward illumination,
// backward illumination), updated in the reference's order.  TX: tile width (columns).
template <int SM, int TX>
__global__ __launch_bounds__(kThreads) void k_bwd_fused(
    Dims d, int it, TileMap tm, int s0, XcdSplit xp, const float2 *__restrict__ KB, const float2 *__restrict__ CP,
    const float4 *__restrict__ CV,
    const ShotGeom *__restrict__ geo, const float *__restrict__ src, const void *__restrict__ strips,
    const uint32_t *__restrict__ smax,
    const float *__restrict__ RIn, float *__restrict__ ROut, const float *__restrict__ BIn,
    float *__restrict__ BOut, const int *__restrict__ inj_gstart, const InjGroup *__restrict__ inj_groups,
    const RecRef *__restrict__ inj_list, const float *__restrict__ sadj_t, float *__restrict__ acc) {
    using namespace bk;
    constexpr int kCX = TX + 2 * kHalo;   // window columns
    constexpr int kNL = kZL * kCX;        // floats per LDS array
    __shared__ float rP[kNL];
    __shared__ float rU[kNL];
    __shared__ float rW[kNL];
    __shared__ float bP[kNL];     // old adjoint p, then the new one in place
    __shared__ float bU[kNL];
    __shared__ float bW[kNL];

    int s, t;
    {
        const int k = blockIdx.x % kNumXcd;
        const int w = xp.xs[k] + (int)(blockIdx.x / kNumXcd);
        if (w >= xp.xs[k + 1]) return;        // past this XCD's range: the whole workgroup
        work_item(w, xp.nsh, tm, s, t);
    }
    s += s0;                      // the launch covers shots s0 .. s0 + xp.nsh - 1
    const TileBox tb = tile_box(d, tm, t);
    const int zo = tb.z0 - kHalo, xo = tb.x0 - kHalo;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int fs = d.fs, npml = d.npml, nx = d.nx, nz = d.nz;
    const int np = d.nxp * d.ld, ni = nx * nz;
    const bool has_interior = max(tb.z0, npml + 1) <= min(tb.z1, npml + nz) &&
                              max(tb.x0, npml + 1) <= min(tb.x1, npml + nx);
    const int top0 = 3 * nx + 6 * nz;
    // updated cells span rows z0-2..z0+TZ and columns x0-2..x1+1; the velocity factors
    // also read damping at the next row / column
    const bool damped = !(tb.z0 - 2 >= npml + 1 && tb.z0 + kTZ + 1 <= npml + nz && tb.x0 - 2 >= npml + 1 &&
                          tb.x1 + 2 <= npml + nx);
    const float *Pc = reinterpret_cast<const float *>(CP);   // 1-kappa_p at 2*k+1
    const ShotGeom g = geo[s];
    const StripRow<SM> b(strips, (size_t)s * d.nt + (it - 2), d.S, smax);
    // this thread's row
    const int z = zo + lane, jz = z - npml;
    const bool zin = z >= 1 && z <= d.nzp;
    const bool rz = has_interior && jz >= 1 && jz <= nz;      // interior row
    const bool zu = z >= 2 && z <= d.nzp - 1;                  // adjoint u rows (:311)
    const bool zw = z >= 2 && z <= d.nzp - 2;                  // adjoint w rows (:320)
    const bool zp = z >= 3 && z <= d.nzp - 2;                  // adjoint p rows (:337)
    const bool lane_v = lane >= 1 && lane <= kZL - 3;          // velocity rows z0-2..z0+TZ
    const bool lane_t = lane >= kHalo && z <= tb.z1;           // tile rows
    // column sets: staged c in [0, kCX); velocities c in [1, cv_end]; tile c in [3, ct_end]
    constexpr int kJS = (kCX + kWaves - 1) / kWaves;
    constexpr int kJV = (TX + 3 + kWaves - 1) / kWaves;
    const int cv_end = tb.x1 + 1 - xo;
    const int ct_end = tb.x1 - xo;

    // ---- entry loads: both windows straight into LDS (global -> LDS transfers: lane l
    // of an instruction writes word l after the instruction's base, so no registers hold
    // the windows), cells outside the grid or the interior zeroed; then coefficients,
    // strip inserts, accumulators ----
    {
        // cells outside the grid or the interior are zeroed first: an LDS store issued
        // after pending transfers would make the compiler wait for all of them
#pragma unroll
        for (int j = 0; j < kJS; ++j) {
            const int c = wave + kWaves * j;
            if (c >= kCX) continue;
            const int x = xo + c, jx = x - npml;
            const int l = LW(c, lane);
            if (!(rz && jx >= 1 && jx <= nx)) rP[l] = rU[l] = rW[l] = 0.f;
            if (!(zin && x >= 1 && x <= d.nxp)) bP[l] = bU[l] = bW[l] = 0.f;
        }
        const float *rin = RIn + (size_t)s * 3 * ni;
        const float *bin = BIn + (size_t)s * 3 * np;
#pragma unroll
        for (int j = 0; j < kJS; ++j) {
            const int c = wave + kWaves * j;
            if (c >= kCX) continue;
            const int x = xo + c, jx = x - npml;
            if (rz && jx >= 1 && jx <= nx) {
                const unsigned k = (unsigned)((jx - 1) * nz + (jz - 1));
                __builtin_amdgcn_global_load_lds(rin + k, rP + LW(c, 0), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(rin + ni + k, rU + LW(c, 0), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(rin + 2 * ni + k, rW + LW(c, 0), 4, 0, 0);
            }
            if (zin && x >= 1 && x <= d.nxp) {
                const unsigned k = (unsigned)((x - 1) * d.ld + (z - 1));
                __builtin_amdgcn_global_load_lds(bin + k, bP + LW(c, 0), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(bin + np + k, bU + LW(c, 0), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(bin + 2 * np + k, bW + LW(c, 0), 4, 0, 0);
            }
        }
    }
    // owned cells: kb = {K, B}, fac = {1-kappa_p, 1-kappa_u, 1-kappa_w}, vsu / vsw =
    // strip values inserted into the reconstructed u / w; on tile cells tsp = strip value
    // inserted into the reconstructed p, ta* = this shot's accumulators
    float2 kb[kJV];
    float fp[kJV], fu[kJV], fw[kJV], vsu[kJV], vsw[kJV];
    float tsp[kJV], ta0[kJV], ta1[kJV], ta2[kJV];
    float *a = acc + (size_t)s * 3 * ni;
#pragma unroll
    for (int j = 0; j < kJV; ++j) {
        const int c = 1 + wave + kWaves * j;
        const int x = xo + c, jx = x - npml;
        kb[j] = make_float2(0.f, 0.f);
        fp[j] = fu[j] = fw[j] = 1.0f;
        vsu[j] = vsw[j] = tsp[j] = ta0[j] = ta1[j] = ta2[j] = 0.f;
        if (c > cv_end || !lane_v || !zin || x < 1 || x > d.nxp) continue;
        const unsigned k = (unsigned)((x - 1) * d.ld + (z - 1));
        kb[j] = KB[k];
        if (damped) {
            const float4 cv = CV[k];
            fu[j] = cv.y;
            fw[j] = cv.z;
            fp[j] = Pc[2 * k + 1];
        }
        const bool rc = rz && jx >= 1 && jx <= nx;
        if (rc) {
            if (jx == 1 || jx == nx - 1) vsu[j] = b[3 * nx + 6 * (jz - 1) + (jx == 1 ? 2 : 5)];
            if (jz == nz - 1) vsw[j] = b[3 * (jx - 1) + 2];
        }
        if (c >= kHalo && c <= ct_end && lane_t && rc) {
            const unsigned ki = (unsigned)((jx - 1) * nz + (jz - 1));
            ta0[j] = a[ki];
            ta1[j] = a[ni + ki];
            ta2[j] = a[2 * ni + ki];
            if (jz <= 3) tsp[j] = b[top0 + 6 * (jx - 1) + (jz - 1)];
            else if (jx <= 2 || jx >= nx - 1)
                tsp[j] = b[3 * nx + 6 * (jz - 1) + (jx == 1 ? 0 : jx == 2 ? 1 : jx == nx ? 3 : 4)];
            else if (jz >= nz - 1) tsp[j] = b[3 * (jx - 1) + (jz == nz ? 0 : 1)];
        }
    }
    // first adjoint-source sample of this thread's injection cell
    const int g0 = inj_gstart[(size_t)s * (tm.ntiles + 1) + t];
    const int g1 = inj_gstart[(size_t)s * (tm.ntiles + 1) + t + 1];
    const RecRef *el = inj_list + (size_t)s * d.ngmax;
    const float *sa = sadj_t + ((size_t)s * d.nt + (it - 1)) * d.ngmax;
    InjGroup ig0{0, 0};
    RecRef ir0{0, 0, 0, 0};
    float iv0 = 0.f;
    if (g0 + (int)threadIdx.x < g1) {
        ig0 = inj_groups[(size_t)s * d.ngmax + g0 + threadIdx.x];
        ir0 = el[ig0.e0];
        iv0 = sa[ir0.idx];
    }
    // the reference's s(it+1) for the reconstruction source subtraction (0 past the end)
    const float ssub = it + 1 <= d.nt ? src[(size_t)s * d.nt + it] : 0.0f;

    // ---- both windows staged: every wave's transfers complete, then the barrier ----
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // ---- velocities: reconstruction u', w' (:221-247), adjoint u', w' (:310-326) ----
#pragma unroll
    for (int j = 0; j < kJV; ++j) {
        int c = 1 + wave + kWaves * j;
        if (c > cv_end || !lane_v || !zin) continue;
        int x = xo + c, jx = x - npml;
        if (x < 1 || x > d.nxp) continue;
        int l = LW(c, lane);
        if (rz && jx >= 1 && jx <= nx) {
            if (jx == 1 || jx == nx - 1) rU[l] = vsu[j];
            else if (jx <= nx - 2 && jz >= 2 && jz <= nz - 1)
                rU[l] = v_undo(rU[l], kb[j].y, rP[l], rP[l + kZL], rP[l + 2 * kZL], rP[l - kZL]);
            if (jz == nz - 1) rW[l] = vsw[j];
            else if (jx >= 2 && jx <= nx - 1 && jz >= 2 && jz <= nz - 2)
                rW[l] = v_undo(rW[l], kb[j].y, rP[l], rP[l + 1], rP[l + 2], rP[l - 1]);
        }
        if (zu && x >= 2 && x <= d.nxp - 2)
            bU[l] = v_update_f(bU[l], kb[j].x, fu[j], bP[l], bP[l + kZL], bP[l + 2 * kZL], bP[l - kZL]);
        if (zw && x >= 2 && x <= d.nxp - 1)
            bW[l] = v_update_f(bW[l], kb[j].x, fw[j], bP[l], bP[l + 1], bP[l + 2], bP[l - 1]);
    }
    __syncthreads();
    // adjoint w mirror (:329-333), columns 2..nxp-1
    if (d.free_surface == 1) {
        const int lm = fs - 1 - zo;
        if (lane == lm && lm >= 1 && lm + 1 <= kZL - 3) {
#pragma unroll
            for (int j = 0; j < kJV; ++j) {
                int c = 1 + wave + kWaves * j;
                int x = xo + c;
                if (c <= cv_end && x >= 2 && x <= d.nxp - 1) bW[LW(c, lm)] = bW[LW(c, lm + 1)];
            }
        }
    }
    __syncthreads();
    // ---- pressures: reconstruction p' (:250-287), adjoint p' (:336-345) ------------
    // the adjoint p is updated in place: in this pass a thread reads bP only at its own cells
    float prn[kJV], pold[kJV];
    float *rout = ROut + (size_t)s * 3 * ni;
#pragma unroll
    for (int j = 0; j < kJV; ++j) {
        prn[j] = pold[j] = 0.f;
        int c = 1 + wave + kWaves * j;
        if (c < kHalo || c > ct_end || !lane_t) continue;
        int x = xo + c, jx = x - npml;
        int l = LW(c, lane);
        if (rz && jx >= 1 && jx <= nx) {
            float v;
            if (jz <= 3 || jx <= 2 || jx >= nx - 1 || jz >= nz - 1) {
                v = tsp[j];
            } else {
                v = rP[l];
                if (jz == g.isz_i && jx == g.isx_i && g.isx_i > 2 && g.isx_i < nx - 1 && g.isz_i > 3 &&
                    g.isz_i < nz - 1)
                    v = v - d.dt * ssub;
                v = p_undo(v, kb[j].x, rU[l], rU[l - kZL], rU[l + kZL], rU[l - 2 * kZL],
                           rW[l], rW[l - 1], rW[l + 1], rW[l - 2]);
            }
            const unsigned k = (unsigned)((jx - 1) * nz + (jz - 1));
            rout[k] = v;
            rout[ni + k] = rU[l];
            rout[2 * ni + k] = rW[l];
            prn[j] = v;
        }
        float v = bP[l];
        pold[j] = v;
        if (zp && x >= 3 && x <= d.nxp - 2)
            v = p_update_f(v, kb[j].y, fp[j], bU[l], bU[l - kZL], bU[l + kZL], bU[l - 2 * kZL],
                           bW[l], bW[l - 1], bW[l + 1], bW[l - 2]);
        bP[l] = v;
    }
    __syncthreads();
    // adjoint-source injection (:348-367): one thread per cell, receivers sharing a cell
    // added in receiver order
    if (g0 + (int)threadIdx.x < g1) {
        int l = LW(ir0.ix - xo, ir0.iz - zo);
        float v = bP[l] + d.dt * iv0;
        for (int e = ig0.e0 + 1; e < ig0.e1; ++e) v = v + d.dt * sa[el[e].idx];
        bP[l] = v;
    }
    for (int q = g0 + (int)threadIdx.x + kThreads; q < g1; q += kThreads) {
        InjGroup gr = inj_groups[(size_t)s * d.ngmax + q];
        RecRef r0 = el[gr.e0];
        int l = LW(r0.ix - xo, r0.iz - zo);
        float v = bP[l];
        for (int e = gr.e0; e < gr.e1; ++e) v = v + d.dt * sa[el[e].idx];
        bP[l] = v;
    }
    __syncthreads();
    // unconditional p mirror (:370-373): rows fs-1 and fs+1 are rows of one tile
    {
        const int lm = fs - 1 - zo;
        if (lane == lm && lm >= kHalo && fs + 1 <= tb.z1) {
#pragma unroll
            for (int j = 0; j < kJV; ++j) {
                int c = 1 + wave + kWaves * j;
                if (c >= kHalo && c <= ct_end) bP[LW(c, lm)] = -bP[LW(c, lm + 2)];
            }
        }
    }
    __syncthreads();
    // ---- write back; imaging dg += p_rec (p0 - p), dfw += p_rec^2, dbk += p^2 -------
    float *bout = BOut + (size_t)s * 3 * np;
#pragma unroll
    for (int j = 0; j < kJV; ++j) {
        int c = 1 + wave + kWaves * j;
        if (c < kHalo || c > ct_end || !lane_t) continue;
        int x = xo + c, jx = x - npml;
        int l = LW(c, lane);
            const unsigned k = (unsigned)((x - 1) * d.ld + (z - 1));
        float pb = bP[l];
        bout[k] = pb;
        bout[np + k] = bU[l];
        bout[2 * np + k] = bW[l];
        if (rz && jx >= 1 && jx <= nx) {
            const unsigned ki = (unsigned)((jx - 1) * nz + (jz - 1));
            float pr = prn[j];
            a[ki] = ta0[j] + pr * (pold[j] - pb);
            a[ni + ki] = ta1[j] + pr * pr;
            a[2 * ni + ki] = ta2[j] + pb * pb;
        }
    }
}

// ---------------------------------------------------------------------------------
// Two backward steps per launch: steps it and it-1 (solver.f90:230-253), each doing
// exactly the per-cell arithmetic of k_bwd_fused, so one launch here and two launches of
// k_bwd_fused give the same bits.  The window has a margin of 6 cells, 3 for each step:
// step 1 updates the window minus 3 cells per side (velocities minus their stencils'
// reach), step 2 the 52x16 tile, from step 1's values in LDS.  The three fields of both
// wavefields and the accumulators therefore cross HBM once per two steps.
// Window rows are lanes and a thread's columns are wave - 2 + 8j (j < 4), so the tile
// columns are exactly its slots j = 1, 2 and each owned cell's coefficients serve both
// steps.
// inj_*: injection groups of the tile (step 2), per-shot capacity ngmax; injx_*: those
// of the step-1 region (the tile plus 3 cells per side), per-shot capacity injx_cap.
namespace bk2 {
constexpr int kZL = 64;
constexpr int kM = 6;                  // margin: 3 cells per step
constexpr int kTZ = kZL - 2 * kM;      // 52 output rows
constexpr int kCX = kTX + 2 * kM;      // 28 window columns
constexpr int kNL = kZL * kCX;         // floats per LDS array
constexpr int kJ = 4;                  // column slots of a thread
__device__ __forceinline__ int LW(int c, int l) { return c * kZL + l; }
}  // namespace bk2

template <int SM>
__global__ __launch_bounds__(kThreads) void k_bwd_fused2(
    Dims d, int it, TileMap tm, int s0, const float2 *__restrict__ KB, const float2 *__restrict__ CP,
    const float4 *__restrict__ CV,
    const ShotGeom *__restrict__ geo, const float *__restrict__ src, const void *__restrict__ strips,
    const uint32_t *__restrict__ smax,
    const float *__restrict__ RIn, float *__restrict__ ROut, const float *__restrict__ BIn,
    float *__restrict__ BOut, const int *__restrict__ inj_gstart, const InjGroup *__restrict__ inj_groups,
    const RecRef *__restrict__ inj_list, const int *__restrict__ injx_gstart,
    const InjGroup *__restrict__ injx_groups, const RecRef *__restrict__ injx_list, int injx_cap,
    const float *__restrict__ sadj_t, float *__restrict__ acc) {
    using namespace bk2;
    __shared__ float rP[kNL];
    __shared__ float rU[kNL];
    __shared__ float rW[kNL];
    __shared__ float bP[kNL];
    __shared__ float bU[kNL];
    __shared__ float bW[kNL];

    int s, t;
    work_item(xcd_logical(blockIdx.x, gridDim.x), (int)(gridDim.x / tm.ntiles), tm, s, t);
    s += s0;                      // the launch covers shots s0 .. s0 + gridDim.x / ntiles - 1
    const TileBox tb = tile_box(d, tm, t);
    const int zo = tb.z0 - kM, xo = tb.x0 - kM;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int fs = d.fs, npml = d.npml, nx = d.nx, nz = d.nz;
    const int np = d.nxp * d.ld, ni = nx * nz;
    const int ez = tb.z1 - zo, ex = tb.x1 - xo;      // last tile row / column of the window
    const bool has_interior = max(zo, npml + 1) <= min(zo + kZL - 1, npml + nz) &&
                              max(xo, npml + 1) <= min(xo + kCX - 1, npml + nx);
    const int top0 = 3 * nx + 6 * nz;
    // updated cells span window rows 1..ez+4 and columns 1..ex+4; the velocity factors
    // also read damping at the next row / column
    const bool damped = !(zo + 1 >= npml + 1 && zo + ez + 5 <= npml + nz && xo + 1 >= npml + 1 &&
                          xo + ex + 5 <= npml + nx);
    const float *Pc = reinterpret_cast<const float *>(CP);
    const ShotGeom g = geo[s];
    const StripRow<SM> b1(strips, (size_t)s * d.nt + (it - 2), d.S, smax);
    const StripRow<SM> b2(strips, (size_t)s * d.nt + (it - 3), d.S, smax);
    const int z = zo + lane, jz = z - npml;
    const bool zin = z >= 1 && z <= d.nzp;
    const bool rz = has_interior && jz >= 1 && jz <= nz;
    const bool zu = z >= 2 && z <= d.nzp - 1;                  // adjoint u rows (:311)
    const bool zw = z >= 2 && z <= d.nzp - 2;                  // adjoint w rows (:320)
    const bool zp = z >= 3 && z <= d.nzp - 2;                  // adjoint p rows (:337)
    // rows of the two steps: velocities v*, pressures p* (p2 = the tile rows)
    const bool v1 = lane >= 1 && lane <= ez + 4 && zin, p1 = lane >= 3 && lane <= ez + 3 && zin;
    const bool v2 = lane >= 4 && lane <= ez + 1 && zin, p2 = lane >= kM && lane <= ez;
    constexpr int kJS = (kCX + kWaves - 1) / kWaves;

    // ---- entry loads: both windows straight into LDS, then coefficients, strip inserts
    // of both steps, accumulators ----
    {
        // cells outside the grid or the interior are zeroed first: an LDS store issued
        // after pending transfers would make the compiler wait for all of them
#pragma unroll
        for (int j = 0; j < kJS; ++j) {
            const int c = wave + kWaves * j;
            if (c >= kCX) continue;
            const int x = xo + c, jx = x - npml;
            const int l = LW(c, lane);
            if (!(rz && jx >= 1 && jx <= nx)) rP[l] = rU[l] = rW[l] = 0.f;
            if (!(zin && x >= 1 && x <= d.nxp)) bP[l] = bU[l] = bW[l] = 0.f;
        }
        const float *rin = RIn + (size_t)s * 3 * ni;
        const float *bin = BIn + (size_t)s * 3 * np;
#pragma unroll
        for (int j = 0; j < kJS; ++j) {
            const int c = wave + kWaves * j;
            if (c >= kCX) continue;
            const int x = xo + c, jx = x - npml;
            if (rz && jx >= 1 && jx <= nx) {
                const unsigned k = (unsigned)((jx - 1) * nz + (jz - 1));
                __builtin_amdgcn_global_load_lds(rin + k, rP + LW(c, 0), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(rin + ni + k, rU + LW(c, 0), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(rin + 2 * ni + k, rW + LW(c, 0), 4, 0, 0);
            }
            if (zin && x >= 1 && x <= d.nxp) {
                const unsigned k = (unsigned)((x - 1) * d.ld + (z - 1));
                __builtin_amdgcn_global_load_lds(bin + k, bP + LW(c, 0), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(bin + np + k, bU + LW(c, 0), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(bin + 2 * np + k, bW + LW(c, 0), 4, 0, 0);
            }
        }
    }
    // owned cells: kb = {K, B}, fac = {1-kappa_p, 1-kappa_u, 1-kappa_w}; strip values
    // inserted into the reconstructed u, w (su*, sw*) and p (sp*) at steps 1 and 2;
    // on tile cells ta = this shot's accumulators
    float2 kb[kJ];
    float fp[kJ], fu[kJ], fw[kJ], su1[kJ], sw1[kJ], su2[kJ], sw2[kJ], sp1[kJ];
    float sp2[2], ta0[2], ta1[2], ta2[2];
    float *a = acc + (size_t)s * 3 * ni;
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
        const int c = wave - 2 + kWaves * j;
        const int x = xo + c, jx = x - npml;
        kb[j] = make_float2(0.f, 0.f);
        fp[j] = fu[j] = fw[j] = 1.0f;
        su1[j] = sw1[j] = su2[j] = sw2[j] = sp1[j] = 0.f;
        if (j == 1 || j == 2) sp2[j - 1] = ta0[j - 1] = ta1[j - 1] = ta2[j - 1] = 0.f;
        if (c < 1 || c > ex + 4 || !v1 || x < 1 || x > d.nxp) continue;
        const unsigned k = (unsigned)((x - 1) * d.ld + (z - 1));
        kb[j] = KB[k];
        if (damped) {
            const float4 cv = CV[k];
            fu[j] = cv.y;
            fw[j] = cv.z;
            fp[j] = Pc[2 * k + 1];
        }
        if (!(rz && jx >= 1 && jx <= nx)) continue;
        if (jx == 1 || jx == nx - 1) {
            const int q = 3 * nx + 6 * (jz - 1) + (jx == 1 ? 2 : 5);
            su1[j] = b1[q];
            su2[j] = b2[q];
        }
        if (jz == nz - 1) {
            const int q = 3 * (jx - 1) + 2;
            sw1[j] = b1[q];
            sw2[j] = b2[q];
        }
        int q = -1;
        if (jz <= 3) q = top0 + 6 * (jx - 1) + (jz - 1);
        else if (jx <= 2 || jx >= nx - 1) q = 3 * nx + 6 * (jz - 1) + (jx == 1 ? 0 : jx == 2 ? 1 : jx == nx ? 3 : 4);
        else if (jz >= nz - 1) q = 3 * (jx - 1) + (jz == nz ? 0 : 1);
        if (q >= 0 && c >= 3 && c <= ex + 3 && p1) sp1[j] = b1[q];
        if ((j == 1 || j == 2) && c <= ex && p2) {
            if (q >= 0) sp2[j - 1] = b2[q];
            const unsigned ki = (unsigned)((jx - 1) * nz + (jz - 1));
            ta0[j - 1] = a[ki];
            ta1[j - 1] = a[ni + ki];
            ta2[j - 1] = a[2 * ni + ki];
        }
    }
    // the reference's s(it+1) of each step's source subtraction (0 past the end)
    const float ssub1 = it + 1 <= d.nt ? src[(size_t)s * d.nt + it] : 0.0f;
    const float ssub2 = src[(size_t)s * d.nt + it - 1];
    const bool src_in = g.isx_i > 2 && g.isx_i < nx - 1 && g.isz_i > 3 && g.isz_i < nz - 1;
    const int m = fs - 1 - zo;                         // window row of the mirror row fs-1

    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    // One step's velocities (reconstruction :221-247, adjoint :310-326) on columns
    // [clo, chi] of the rows vr, then the adjoint w mirror (:329-333) if rows m, m+1 are
    // velocity rows (m+1 <= vhi).
    auto velocities = [&](bool vr, int clo, int chi, int vhi, const float *su, const float *sw) {
#pragma unroll
        for (int j = 0; j < kJ; ++j) {
            const int c = wave - 2 + kWaves * j;
            if (c < clo || c > chi || !vr) continue;
            const int x = xo + c, jx = x - npml;
            if (x < 1 || x > d.nxp) continue;
            const int l = LW(c, lane);
            if (rz && jx >= 1 && jx <= nx) {
                if (jx == 1 || jx == nx - 1) rU[l] = su[j];
                else if (jx <= nx - 2 && jz >= 2 && jz <= nz - 1)
                    rU[l] = v_undo(rU[l], kb[j].y, rP[l], rP[l + kZL], rP[l + 2 * kZL], rP[l - kZL]);
                if (jz == nz - 1) rW[l] = sw[j];
                else if (jx >= 2 && jx <= nx - 1 && jz >= 2 && jz <= nz - 2)
                    rW[l] = v_undo(rW[l], kb[j].y, rP[l], rP[l + 1], rP[l + 2], rP[l - 1]);
            }
            if (zu && x >= 2 && x <= d.nxp - 2)
                bU[l] = v_update_f(bU[l], kb[j].x, fu[j], bP[l], bP[l + kZL], bP[l + 2 * kZL], bP[l - kZL]);
            if (zw && x >= 2 && x <= d.nxp - 1)
                bW[l] = v_update_f(bW[l], kb[j].x, fw[j], bP[l], bP[l + 1], bP[l + 2], bP[l - 1]);
        }
        __syncthreads();
        if (d.free_surface == 1 && lane == m && m >= 1 && m + 1 <= vhi) {
#pragma unroll
            for (int j = 0; j < kJ; ++j) {
                const int c = wave - 2 + kWaves * j;
                const int x = xo + c;
                if (c >= clo && c <= chi && x >= 2 && x <= d.nxp - 1) bW[LW(c, m)] = bW[LW(c, m + 1)];
            }
        }
        __syncthreads();
    };
    // One step's pressures (reconstruction :250-287, adjoint :336-345) on columns
    // [clo, chi] of the rows pr; both are updated in place (a thread reads rP and bP only
    // at its own cells in this pass).  Returns the reconstructed and the old adjoint p of
    // the tile slots in prn / pold.
    auto pressures = [&](bool pr, int clo, int chi, const float *sp, const float *sp_t, float ssub, float *prn,
                         float *pold) {
#pragma unroll
        for (int j = 0; j < kJ; ++j) {
            const int c = wave - 2 + kWaves * j;
            if (c < clo || c > chi || !pr) continue;
            const int x = xo + c, jx = x - npml;
            if (x < 1 || x > d.nxp) continue;
            const int l = LW(c, lane);
            if (rz && jx >= 1 && jx <= nx) {
                float v;
                if (jz <= 3 || jx <= 2 || jx >= nx - 1 || jz >= nz - 1) {
                    v = sp_t ? ((j == 1 || j == 2) ? sp_t[j - 1] : 0.f) : sp[j];
                } else {
                    v = rP[l];
                    if (jz == g.isz_i && jx == g.isx_i && src_in) v = v - d.dt * ssub;
                    v = p_undo(v, kb[j].x, rU[l], rU[l - kZL], rU[l + kZL], rU[l - 2 * kZL],
                               rW[l], rW[l - 1], rW[l + 1], rW[l - 2]);
                }
                rP[l] = v;
                if (j == 1 || j == 2) prn[j - 1] = v;
            }
            float v = bP[l];
            if (j == 1 || j == 2) pold[j - 1] = v;
            if (zp && x >= 3 && x <= d.nxp - 2)
                v = p_update_f(v, kb[j].y, fp[j], bU[l], bU[l - kZL], bU[l + kZL], bU[l - 2 * kZL],
                               bW[l], bW[l - 1], bW[l + 1], bW[l - 2]);
            bP[l] = v;
        }
        __syncthreads();
    };
    // Adjoint-source injection (:348-367) of one step's region, receivers sharing a cell
    // added in receiver order, then the p mirror (:370-373) on columns [clo, chi] when
    // row m+2 is a pressure row of the step (plo <= m+2 <= phi).
    auto inject_mirror = [&](const int *gst, const InjGroup *grp, const RecRef *lst, int cap, int row, int clo,
                             int chi, int plo, int phi) {
        const int g0 = gst[(size_t)s * (tm.ntiles + 1) + t];
        const int g1 = gst[(size_t)s * (tm.ntiles + 1) + t + 1];
        const RecRef *el = lst + (size_t)s * cap;
        const float *sa = sadj_t + ((size_t)s * d.nt + row) * d.ngmax;
        for (int q = g0 + (int)threadIdx.x; q < g1; q += kThreads) {
            const InjGroup gr = grp[(size_t)s * cap + q];
            const RecRef r0 = el[gr.e0];
            const int l = LW(r0.ix - xo, r0.iz - zo);
            float v = bP[l];
            for (int e = gr.e0; e < gr.e1; ++e) v = v + d.dt * sa[el[e].idx];
            bP[l] = v;
        }
        __syncthreads();
        if (lane == m && m >= 0 && m + 2 >= plo && m + 2 <= phi) {
#pragma unroll
            for (int j = 0; j < kJ; ++j) {
                const int c = wave - 2 + kWaves * j;
                const int x = xo + c;
                if (c >= clo && c <= chi && x >= 1 && x <= d.nxp) bP[LW(c, m)] = -bP[LW(c, m + 2)];
            }
        }
        __syncthreads();
    };

    // ---- step 1 (it): window minus 3 cells per side ----
    float pr1[2] = {0.f, 0.f}, po1[2] = {0.f, 0.f};
    velocities(v1, 1, ex + 4, ez + 4, su1, sw1);
    pressures(p1, 3, ex + 3, sp1, nullptr, ssub1, pr1, po1);
    inject_mirror(injx_gstart, injx_groups, injx_list, injx_cap, it - 1, 3, ex + 3, 3, ez + 3);
    // step 1 imaging on the tile: dg += p_rec (p0 - p), dfw += p_rec^2, dbk += p^2
#pragma unroll
    for (int j = 1; j <= 2; ++j) {
        const int c = wave - 2 + kWaves * j;
        const int jx = xo + c - npml;
        if (c > ex || !p2 || !(rz && jx >= 1 && jx <= nx)) continue;
        const float pb = bP[LW(c, lane)];
        const float pr = pr1[j - 1];
        ta0[j - 1] = ta0[j - 1] + pr * (po1[j - 1] - pb);
        ta1[j - 1] = ta1[j - 1] + pr * pr;
        ta2[j - 1] = ta2[j - 1] + pb * pb;
    }
    // ---- step 2 (it-1): the tile ----
    float pr2[2] = {0.f, 0.f}, po2[2] = {0.f, 0.f};
    velocities(v2, 4, ex + 1, ez + 1, su2, sw2);
    pressures(p2, kM, ex, nullptr, sp2, ssub2, pr2, po2);
    inject_mirror(inj_gstart, inj_groups, inj_list, d.ngmax, it - 2, kM, ex, kM, ez);

    // ---- write back; step 2 imaging ----
    float *bout = BOut + (size_t)s * 3 * np;
    float *rout = ROut + (size_t)s * 3 * ni;
#pragma unroll
    for (int j = 1; j <= 2; ++j) {
        const int c = wave - 2 + kWaves * j;
        if (c > ex || !p2) continue;
        const int x = xo + c, jx = x - npml;
        const int l = LW(c, lane);
        const unsigned k = (unsigned)((x - 1) * d.ld + (z - 1));
        const float pb = bP[l];
        bout[k] = pb;
        bout[np + k] = bU[l];
        bout[2 * np + k] = bW[l];
        if (rz && jx >= 1 && jx <= nx) {
            const unsigned ki = (unsigned)((jx - 1) * nz + (jz - 1));
            const float pr = pr2[j - 1];
            rout[ki] = pr;
            rout[ni + ki] = rU[l];
            rout[2 * ni + ki] = rW[l];
            a[ki] = ta0[j - 1] + pr * (po2[j - 1] - pb);
            a[ni + ki] = ta1[j - 1] + pr * pr;
            a[2 * ni + ki] = ta2[j - 1] + pb * pb;
        }
    }
}

// Cross-shot FP64 sum of the per-shot kernels and illuminations
// (solver.f90:263-265) and g = -2 g / c (solver.f90:270).  acc: [ns][3][nx][nz].
__global__ void k_reduce_acc(Dims d, int ns, const float *__restrict__ vp, const float *__restrict__ acc,
                             double *__restrict__ grad, double *__restrict__ fw, double *__restrict__ bk) {
    size_t ni = (size_t)d.nx * d.nz;
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ni) return;
    double g = 0.0, f = 0.0, b = 0.0;
    for (int s = 0; s < ns; ++s) {
        const float *a = acc + (size_t)s * 3 * ni;
        g += (double)a[k];
        f += (double)a[ni + k];
        b += (double)a[2 * ni + k];
    }
    grad[k] = -2.0 * g / (double)vp[k];
    fw[k] = f;
    bk[k] = b;
}

// Final interior windows of the forward state (fdcore.f90:162-165) for every shot.
// S: [ns][3][nxp][ld] -> E: [ns][3][nx][nz]
__global__ void k_end_state(Dims d, const float *__restrict__ S, float *__restrict__ E) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int s = blockIdx.z;
    if (iz > d.nz || ix > d.nx) return;
    size_t np = (size_t)d.nxp * d.ld, ni = (size_t)d.nx * d.nz;
    size_t kp = (size_t)s * 3 * np + pidx(d, iz + d.npml, ix + d.npml);
    size_t ki = (size_t)s * 3 * ni + iidx(d, iz, ix);
    E[ki] = S[kp];
    E[ni + ki] = S[np + kp];
    E[2 * ni + ki] = S[2 * np + kp];
}

// Per-plan state of the batched path.
struct BatchedState {
    int ready = 0;
    int Gf = 1;                  // at most this many shots per forward workgroup
    int ngroups_f = 0;
    TileMap fwd{}, bwd{};
    int *rec_start = nullptr;    // [ns][fwd.ntiles+1]
    RecRef *rec_list = nullptr;  // [ns][ngmax]
    int *inj_start = nullptr;    // [ns][bwd.ntiles+1] group ranges per tile
    InjGroup *inj_groups = nullptr;  // [ns][ngmax]
    RecRef *inj_list = nullptr;  // [ns][ngmax]
    // one-step backward tilings of every width in bk::kBTXs with their injection groups;
    // bwd and inj_* above are those of width index bwd_wi
    TileMap bwdw[2]{};
    int *inj_start_w[2] = {nullptr, nullptr};
    InjGroup *inj_groups_w[2] = {nullptr, nullptr};
    RecRef *inj_list_w[2] = {nullptr, nullptr};
    int bwd_wi = 1;
    // two-step backward kernel: its tiling, the injection groups of its tiles (step 2)
    // and of its tiles grown by 3 cells per side (step 1, per-shot capacity injx_cap)
    TileMap bwd2{};
    TileMap extm{};              // tiles of the fused exact-adjoint step (ex::kExTX wide)
    int bwd_steps = 2;           // backward steps per launch: 1 or 2
    int *inj2_start = nullptr;
    InjGroup *inj2_groups = nullptr;
    RecRef *inj2_list = nullptr;
    int *injx_start = nullptr;
    InjGroup *injx_groups = nullptr;
    RecRef *injx_list = nullptr;
    int injx_cap = 0;
    int *strip_start = nullptr;  // [fwd.ntiles+1]
    StripRef *strip_list = nullptr;
    float2 *KB = nullptr;        // [nxp][ld] {K, B}
    float2 *CP = nullptr;        // [nxp][ld]
    float4 *CV = nullptr;        // [nxp][ld]
    float *F[2] = {nullptr, nullptr};   // forward state [ns][3][nxp][ld], ping-pong
    float *Bk[2] = {nullptr, nullptr};  // adjoint state [ns][3][nxp][ld]
    float *R[2] = {nullptr, nullptr};   // reconstruction state [ns][3][nx][nz]
    float *end = nullptr;        // saved final forward state [ns][3][nx][nz]
    float *src = nullptr;        // plan-owned copy of the source traces
    float *src_ex = nullptr;     // the exact forward's own copy (its checkpointed sweep reruns it)
    int64_t exact_gen = 0;       // count of exact forwards; names the history dhist holds
    float *acc = nullptr;        // [ns][3][nx][nz] per-shot FP32 accumulators
    float *dhist = nullptr;      // [ns][nt or ck][nxp][ld] divergence history (exact adjoint)
    int ck = 0;                  // checkpoint interval (0: the whole history is kept)
    int nck = 0;                 // checkpoints: states after steps 1, 1+ck, 1+2ck, ...
    float *ckpt = nullptr;       // [nck][ns][3][nxp][ld]
    float *lam = nullptr;        // [ns][3][nxp][ld] adjoint variables of the exact adjoint
    float *lam2 = nullptr;       // second buffer of the fused exact step (ping-pong with lam)
    int exact_fused = 1;         // 1: fused exact step kernel; 0: the seven separate passes
    int *exinj_start = nullptr;  // receiver groups (recording cells) of the backward tiles
    InjGroup *exinj_groups = nullptr;   // grown by 3 cells, per-shot capacity exinj_cap
    RecRef *exinj_list = nullptr;
    int exinj_cap = 0;
    float *qbuf = nullptr;       // [ns][2][nxp][ld] exact-adjoint temporaries
    float *gk = nullptr;         // [ns][nxp][ld] per-shot gradient with respect to K
    double *gsum = nullptr;      // [nxp][ld] K-gradient summed over shots (FP64)
    float *gseis = nullptr;      // [ns][ngmax][nt] copy of dJ/dseis for the captured loop
    hipGraphExec_t gex_f = nullptr, gex_b = nullptr;   // exact forward / adjoint loops
    bool exact_valid = false;    // dhist holds the last forward_exact of the current model
    int strip_mode = 0;          // 0 FP32, 1 FP16, 2 BF16
    uint16_t *strips16 = nullptr;   // [ns][nt][S] 16-bit strips
    uint32_t *smax = nullptr;    // [ns][nt] per-row max |v| (float bits)
    float *ring = nullptr;       // [ns][kStripRing][S] FP32 rows awaiting their 16-bit encoding
    double *out = nullptr;       // [3][nx][nz]
    hipStream_t cap = nullptr;   // capture stream
    hipGraphExec_t gfwd[2] = {nullptr, nullptr};
    hipGraphExec_t gbwd = nullptr;
    // two streams (SWIT_OPT_STREAMS 2): each half of the shots runs its time loops on a
    // stream of its own, from graphs of its own ([half][save] and [half]); the exact
    // loops likewise ([forward/adjoint][half])
    int streams = 1;
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipGraphExec_t gfwd_h[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    hipGraphExec_t gbwd_h[2] = {nullptr, nullptr};
    hipGraphExec_t gex_h[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    // work split of the one-step backward kernel over the XCDs (1: by estimated cost,
    // SWIT_OPT_BWD_BALANCE) and the splits computed so far, keyed by {shots, column group,
    // tile width, balanced}
    int bwd_balance = 1;
    std::vector<std::pair<std::array<int, 4>, XcdSplit>> splits;
    int use_graph = 1;
    int gen = 0;                 // bumped when captured launch arguments change
    int profile = 0;             // bracket each time loop with a pair of events
    struct Timed { int which; long long launches; hipEvent_t e0, e1; };
    std::vector<Timed> pending;  // recorded, not yet read
    std::vector<hipEvent_t> ev_pool;
    double prof_ms[2] = {0, 0};  // accumulated fwd / bwd step-kernel time
    long long prof_n[2] = {0, 0};
};

}  // namespace swit

struct swit_plan;
static int batched_init(swit_plan *pl);
static void batched_release(swit_plan *pl);
static int batched_set_model(swit_plan *pl, hipStream_t st);
static int batched_forward(swit_plan *pl, const float *src, int save, hipStream_t st);
static int batched_adjoint(swit_plan *pl, const float *src, double *grad, double *fw, double *bk,
                           hipStream_t st);
