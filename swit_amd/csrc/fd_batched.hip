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
// cross-shot sum is FP64 at the end of the job.
//
// Block -> tile order is XCD-aware: blocks are dealt round-robin to the 8 XCDs, so
// the j-th block of XCD k takes the j-th item of a contiguous range of work items;
// items are ordered z-fastest within a shot group so neighbouring tiles (which share
// halos) run on the same XCD, close in time, and hit the same L2.
//
// Memory layout (device): state buffers [ns][3][nxp][ld] (p, u, w of a shot
// contiguous), reconstruction buffers [ns][3][nx][nz]; coefficients per padded cell
// CP = {den c^2 dt/dx, 1 - damp dt} and CV = {dt/(den dx), 1 - kappa_u, 1 - kappa_w}.
#include "fd_common.h"
#include "fd_kernels.h"

#include <vector>

namespace swit {

constexpr int kTZ = 64;          // tile rows (iz, the contiguous axis)
constexpr int kTX = 16;          // tile columns (ix)
constexpr int kHalo = 3;
constexpr int kRZ = kTZ + 2 * kHalo;
constexpr int kRX = kTX + 2 * kHalo;
constexpr int kNR = kRZ * kRX;
constexpr int kThreads = 512;
constexpr int kCellsPerThread = kTZ * kTX / kThreads;
constexpr int kNumXcd = 8;
// Region of new values the tile's stencils read: (TZ+3) x (TX+3) cells.
constexpr int kQZ = kTZ + 3;
constexpr int kQX = kTX + 3;
constexpr int kNQ = kQZ * kQX;

// Tiling of the padded grid: tile (zt, xt) covers rows
// [max(ztop + zt*TZ, 1), min(ztop + (zt+1)*TZ - 1, nzp)] and columns
// [1 + xt*TX, min((xt+1)*TX, nxp)].
struct TileMap {
    int ztop;
    int nzt, nxt;
    int ntiles;     // nzt * nxt
};

struct TileBox {
    int z0, z1, x0, x1;
};

__device__ __forceinline__ TileBox tile_box(const Dims &d, const TileMap &tm, int t) {
    int zt = t % tm.nzt, xt = t / tm.nzt;
    TileBox b;
    b.z0 = max(tm.ztop + zt * kTZ, 1);
    b.z1 = min(tm.ztop + (zt + 1) * kTZ - 1, d.nzp);
    b.x0 = 1 + xt * kTX;
    b.x1 = min((xt + 1) * kTX, d.nxp);
    return b;
}

// XCD-aware remap of a linear block id onto a logical work item.
__device__ __forceinline__ int xcd_logical(int b, int nb) {
    int q = nb / kNumXcd, r = nb % kNumXcd;
    int k = b % kNumXcd, j = b / kNumXcd;
    return k * q + min(k, r) + j;
}

// Strip slot of a tile: slot index and source cell.
struct StripRef {
    int slot;
    int field;      // 0 p, 1 u, 2 w
    int iz, ix;
};

// Receiver of a tile: receiver index and cell.
struct RecRef {
    int idx;
    int iz, ix;
    int pad;
};

// Injection group: the receivers of one tile that share one cell, in receiver order.
struct InjGroup {
    int e0, e1;     // entry range in the shot's injection list
};

// LDS index of padded cell (z, x) in a region whose origin is (zo, xo).
__device__ __forceinline__ int L(int z, int x, int zo, int xo) { return (x - xo) * kRZ + (z - zo); }

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
                             float4 *__restrict__ CV) {
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
}

// ---------------------------------------------------------------------------------
// Forward step it (fdcore.f90:75-160) for a group of G shots per workgroup.
// In/Out: [ns][3][nxp][ld].
__global__ __launch_bounds__(kThreads) void k_fwd_fused(
    Dims d, int it, TileMap tm, int ns, int G, const float2 *__restrict__ CP, const float4 *__restrict__ CV,
    const ShotGeom *__restrict__ geo, const float *__restrict__ src, const float *__restrict__ In,
    float *__restrict__ Out, const int *__restrict__ rec_start, const RecRef *__restrict__ rec_list,
    float *__restrict__ seis_t, const int *__restrict__ strip_start, const StripRef *__restrict__ strip_list,
    float *__restrict__ strips) {
    __shared__ float sP[kNR];
    __shared__ float sU[kNR];
    __shared__ float sW[kNR];
    __shared__ float sQ[kNR];     // new p

    const int work = xcd_logical(blockIdx.x, gridDim.x);
    const int grp = work / tm.ntiles;
    const int t = work % tm.ntiles;
    const TileBox tb = tile_box(d, tm, t);
    const int zo = tb.z0 - kHalo, xo = tb.x0 - kHalo;
    const int tid = threadIdx.x;
    const int fs = d.fs;
    const int np = d.nxp * d.ld;

    const int s_begin = grp * G, s_end = min(ns, s_begin + G);
    for (int s = s_begin; s < s_end; ++s) {
        const ShotGeom g = geo[s];
        const float srcv = d.dt * src[(size_t)s * d.nt + (it - 1)];
        const float *pin = In + (size_t)s * 3 * np;
        // stage old state (zero outside the grid)
        for (int i = tid; i < kNR; i += kThreads) {
            int z = zo + i % kRZ, x = xo + i / kRZ;
            float p = 0.f, u = 0.f, w = 0.f;
            if (z >= 1 && z <= d.nzp && x >= 1 && x <= d.nxp) {
                int k = (x - 1) * d.ld + (z - 1);
                p = pin[k];
                u = pin[np + k];
                w = pin[2 * np + k];
            }
            sP[i] = p;
            sU[i] = u;
            sW[i] = w;
        }
        __syncthreads();
        // new p on rows z0-1..z0+TZ+1, cols x0-1..x0+TX+1
        for (int i = tid; i < kNQ; i += kThreads) {
            int z = tb.z0 - 1 + i % kQZ, x = tb.x0 - 1 + i / kQZ;
            int l = L(z, x, zo, xo);
            float v = 0.f;
            if (z >= 1 && z <= d.nzp && x >= 1 && x <= d.nxp) {
                v = sP[l];
                if (z >= fs + 1 && z <= d.nzp - 2 && x >= 3 && x <= d.nxp - 2) {
                    float2 cp = CP[(x - 1) * d.ld + (z - 1)];
                    v = p_update_f(v, cp.x, cp.y, sU[l], sU[l - kRZ], sU[l + kRZ], sU[l - 2 * kRZ],
                                   sW[l], sW[l - 1], sW[l + 1], sW[l - 2]);
                    if (z == g.isz && x == g.isx) v = v + srcv;
                }
            }
            sQ[l] = v;
        }
        __syncthreads();
        // p mirror (fdcore.f90:92-94) on the region columns
        if (d.free_surface == 1 && fs - 1 >= tb.z0 - 1 && fs + 1 <= tb.z0 + kTZ + 1) {
            for (int x = tb.x0 - 1 + tid; x <= tb.x0 + kTX + 1; x += kThreads)
                if (x >= 1 && x <= d.nxp) sQ[L(fs - 1, x, zo, xo)] = -sQ[L(fs + 1, x, zo, xo)];
        }
        __syncthreads();
        // new u, w on the tile (in place: each cell reads only its own old velocity)
#pragma unroll
        for (int c = 0; c < kCellsPerThread; ++c) {
            int i = tid + c * kThreads;
            int z = tb.z0 + i % kTZ, x = tb.x0 + i / kTZ;
            if (z > tb.z1 || x > tb.x1) continue;
            int l = L(z, x, zo, xo);
            bool ur = z >= fs && z <= d.nzp - 1 && x >= 2 && x <= d.nxp - 2;
            bool wr = z >= fs && z <= d.nzp - 2 && x >= 2 && x <= d.nxp - 1;
            if (ur || wr) {
                float4 cv = CV[(x - 1) * d.ld + (z - 1)];
                if (ur) sU[l] = v_update_f(sU[l], cv.x, cv.y, sQ[l], sQ[l + kRZ], sQ[l + 2 * kRZ], sQ[l - kRZ]);
                if (wr) sW[l] = v_update_f(sW[l], cv.x, cv.z, sQ[l], sQ[l + 1], sQ[l + 2], sQ[l - 1]);
            }
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
            int i = tid + c * kThreads;
            int z = tb.z0 + i % kTZ, x = tb.x0 + i / kTZ;
            if (z > tb.z1 || x > tb.x1) continue;
            int l = L(z, x, zo, xo), k = (x - 1) * d.ld + (z - 1);
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
            float *row = strips + ((size_t)s * d.nt + (it - 1)) * d.S;
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
// RIn/ROut: [ns][3][nx][nz]; BIn/BOut: [ns][3][nxp][ld].
// acc: [ns][3][nx][nz] FP32 per-shot accumulators (kernel, forward illumination,
// backward illumination), updated in the reference's order.
__global__ __launch_bounds__(kThreads) void k_bwd_fused(
    Dims d, int it, TileMap tm, const float2 *__restrict__ CP, const float4 *__restrict__ CV,
    const ShotGeom *__restrict__ geo, const float *__restrict__ src, const float *__restrict__ strips,
    const float *__restrict__ RIn, float *__restrict__ ROut, const float *__restrict__ BIn,
    float *__restrict__ BOut, const int *__restrict__ inj_gstart, const InjGroup *__restrict__ inj_groups,
    const RecRef *__restrict__ inj_list, const float *__restrict__ sadj_t, float *__restrict__ acc) {
    __shared__ float sP[kNR];
    __shared__ float sU[kNR];
    __shared__ float sW[kNR];
    __shared__ float sQ[kNR];

    const int work = xcd_logical(blockIdx.x, gridDim.x);
    const int s = work / tm.ntiles;
    const int t = work % tm.ntiles;
    const TileBox tb = tile_box(d, tm, t);
    const int zo = tb.z0 - kHalo, xo = tb.x0 - kHalo;
    const int tid = threadIdx.x;
    const int fs = d.fs, npml = d.npml, nx = d.nx, nz = d.nz;
    const int np = d.nxp * d.ld;
    const int ni = nx * nz;
    // does this tile touch the interior window?
    const bool has_interior = max(tb.z0, npml + 1) <= min(tb.z1, npml + nz) &&
                              max(tb.x0, npml + 1) <= min(tb.x1, npml + nx);
    const int S = d.S;
    const int top0 = 3 * nx + 6 * nz;
    const float *Bc = reinterpret_cast<const float *>(CV);   // B at 4*k
    const float *Kc = reinterpret_cast<const float *>(CP);   // K at 2*k, 1-kappa_p at 2*k+1
    const ShotGeom g = geo[s];

    float prn[kCellsPerThread];     // reconstructed p at it-1 on owned cells
#pragma unroll
    for (int c = 0; c < kCellsPerThread; ++c) prn[c] = 0.f;
    // ---------------- reconstruction on the interior window -----------------
    if (has_interior) {
        const float *rin = RIn + (size_t)s * 3 * ni;
        const float *b = strips + ((size_t)s * d.nt + (it - 2)) * S;
        for (int i = tid; i < kNR; i += kThreads) {
            int jz = zo + i % kRZ - npml, jx = xo + i / kRZ - npml;
            float p = 0.f, u = 0.f, w = 0.f;
            if (jz >= 1 && jz <= nz && jx >= 1 && jx <= nx) {
                int k = (jx - 1) * nz + (jz - 1);
                p = rin[k];
                u = rin[ni + k];
                w = rin[2 * ni + k];
            }
            sP[i] = p;
            sU[i] = u;
            sW[i] = w;
        }
        __syncthreads();
        // u', w' on rows z0-2..z0+TZ, cols x0-2..x0+TX (fdcore.f90:221-247)
        for (int i = tid; i < kNQ; i += kThreads) {
            int z = tb.z0 - 2 + i % kQZ, x = tb.x0 - 2 + i / kQZ;
            int jz = z - npml, jx = x - npml;
            if (jz < 1 || jz > nz || jx < 1 || jx > nx) continue;
            int l = L(z, x, zo, xo);
            if (jx == 1 || jx == nx - 1) {
                int jj = 3 * nx + 6 * (jz - 1);
                sU[l] = jx == 1 ? b[jj + 2] : b[jj + 5];
            } else if (jx >= 2 && jx <= nx - 2 && jz >= 2 && jz <= nz - 1) {
                float alpha = Bc[4 * ((x - 1) * d.ld + (z - 1))];
                sU[l] = v_undo(sU[l], alpha, sP[l], sP[l + kRZ], sP[l + 2 * kRZ], sP[l - kRZ]);
            }
            if (jz == nz - 1) {
                sW[l] = b[3 * (jx - 1) + 2];
            } else if (jx >= 2 && jx <= nx - 1 && jz >= 2 && jz <= nz - 2) {
                float alpha = Bc[4 * ((x - 1) * d.ld + (z - 1))];
                sW[l] = v_undo(sW[l], alpha, sP[l], sP[l + 1], sP[l + 2], sP[l - 1]);
            }
        }
        __syncthreads();
        // p' on owned tile cells (fdcore.f90:250-287)
        float *rout = ROut + (size_t)s * 3 * ni;
#pragma unroll
        for (int c = 0; c < kCellsPerThread; ++c) {
            int i = tid + c * kThreads;
            int z = tb.z0 + i % kTZ, x = tb.x0 + i / kTZ;
            int jz = z - npml, jx = x - npml;
            if (z > tb.z1 || x > tb.x1 || jz < 1 || jz > nz || jx < 1 || jx > nx) continue;
            int l = L(z, x, zo, xo);
            float v;
            if (jz <= 3) {
                v = b[top0 + 6 * (jx - 1) + (jz - 1)];
            } else if (jx <= 2 || jx >= nx - 1) {
                int jj = 3 * nx + 6 * (jz - 1);
                v = jx == 1 ? b[jj] : jx == 2 ? b[jj + 1] : jx == nx ? b[jj + 3] : b[jj + 4];
            } else if (jz >= nz - 1) {
                int jj = 3 * (jx - 1);
                v = jz == nz ? b[jj] : b[jj + 1];
            } else {
                v = sP[l];
                if (jz == g.isz_i && jx == g.isx_i && g.isx_i > 2 && g.isx_i < nx - 1 && g.isz_i > 3 &&
                    g.isz_i < nz - 1) {
                    // the reference reads s(it+1), one past the end at it = nt (taken as 0)
                    float sv = it + 1 <= d.nt ? src[(size_t)s * d.nt + it] : 0.0f;
                    v = v - d.dt * sv;
                }
                float alpha = Kc[2 * ((x - 1) * d.ld + (z - 1))];
                v = p_undo(v, alpha, sU[l], sU[l - kRZ], sU[l + kRZ], sU[l - 2 * kRZ],
                           sW[l], sW[l - 1], sW[l + 1], sW[l - 2]);
            }
            int k = (jx - 1) * nz + (jz - 1);
            rout[k] = v;
            rout[ni + k] = sU[l];
            rout[2 * ni + k] = sW[l];
            prn[c] = v;
        }
        __syncthreads();
    }
    // ---------------- adjoint step on the padded grid ------------------------
    const float *bin = BIn + (size_t)s * 3 * np;
    for (int i = tid; i < kNR; i += kThreads) {
        int z = zo + i % kRZ, x = xo + i / kRZ;
        float p = 0.f, u = 0.f, w = 0.f;
        if (z >= 1 && z <= d.nzp && x >= 1 && x <= d.nxp) {
            int k = (x - 1) * d.ld + (z - 1);
            p = bin[k];
            u = bin[np + k];
            w = bin[2 * np + k];
        }
        sP[i] = p;
        sU[i] = u;
        sW[i] = w;
    }
    __syncthreads();
    // u', w' with alpha = den c^2 dt/dx on the region (fdcore.f90:310-326)
    for (int i = tid; i < kNQ; i += kThreads) {
        int z = tb.z0 - 2 + i % kQZ, x = tb.x0 - 2 + i / kQZ;
        bool ur = z >= 2 && z <= d.nzp - 1 && x >= 2 && x <= d.nxp - 2;
        bool wr = z >= 2 && z <= d.nzp - 2 && x >= 2 && x <= d.nxp - 1;
        if (!(ur || wr)) continue;
        int l = L(z, x, zo, xo);
        int k = (x - 1) * d.ld + (z - 1);
        float alpha = Kc[2 * k];
        float4 cv = CV[k];
        if (ur) sU[l] = v_update_f(sU[l], alpha, cv.y, sP[l], sP[l + kRZ], sP[l + 2 * kRZ], sP[l - kRZ]);
        if (wr) sW[l] = v_update_f(sW[l], alpha, cv.z, sP[l], sP[l + 1], sP[l + 2], sP[l - 1]);
    }
    __syncthreads();
    // w mirror (fdcore.f90:329-333), columns 2..nxp-1 of the region
    if (d.free_surface == 1 && fs - 1 >= tb.z0 - 2 && fs <= tb.z0 + kTZ) {
        for (int x = tb.x0 - 2 + tid; x <= tb.x0 + kTX; x += kThreads)
            if (x >= 2 && x <= d.nxp - 1) sW[L(fs - 1, x, zo, xo)] = sW[L(fs, x, zo, xo)];
    }
    __syncthreads();
    // p' with alpha = dt/(den dx) on the tile into sQ (fdcore.f90:336-345)
#pragma unroll
    for (int c = 0; c < kCellsPerThread; ++c) {
        int i = tid + c * kThreads;
        int z = tb.z0 + i % kTZ, x = tb.x0 + i / kTZ;
        if (z > tb.z1 || x > tb.x1) continue;
        int l = L(z, x, zo, xo);
        float v = sP[l];
        if (z >= 3 && z <= d.nzp - 2 && x >= 3 && x <= d.nxp - 2) {
            int k = (x - 1) * d.ld + (z - 1);
            float alpha = Bc[4 * k];
            float fac = Kc[2 * k + 1];
            v = p_update_f(v, alpha, fac, sU[l], sU[l - kRZ], sU[l + kRZ], sU[l - 2 * kRZ],
                           sW[l], sW[l - 1], sW[l + 1], sW[l - 2]);
        }
        sQ[l] = v;
    }
    __syncthreads();
    // adjoint-source injection (fdcore.f90:348-367): one thread per cell, receivers
    // sharing a cell added in receiver order
    {
        int g0 = inj_gstart[(size_t)s * (tm.ntiles + 1) + t];
        int g1 = inj_gstart[(size_t)s * (tm.ntiles + 1) + t + 1];
        const RecRef *el = inj_list + (size_t)s * d.ngmax;
        const float *sa = sadj_t + ((size_t)s * d.nt + (it - 1)) * d.ngmax;
        for (int q = g0 + tid; q < g1; q += kThreads) {
            InjGroup gr = inj_groups[(size_t)s * d.ngmax + q];
            RecRef r0 = el[gr.e0];
            int l = L(r0.iz, r0.ix, zo, xo);
            float v = sQ[l];
            for (int e = gr.e0; e < gr.e1; ++e) v = v + d.dt * sa[el[e].idx];
            sQ[l] = v;
        }
    }
    __syncthreads();
    // unconditional p mirror (fdcore.f90:370-373): rows fs-1 and fs+1 share a tile
    if (fs - 1 >= tb.z0 && fs + 1 <= tb.z1) {
        for (int x = tb.x0 + tid; x <= tb.x1; x += kThreads) sQ[L(fs - 1, x, zo, xo)] = -sQ[L(fs + 1, x, zo, xo)];
    }
    __syncthreads();
    // write back; imaging dg += p_rec (p0 - p), dfw += p_rec^2, dbk += p^2
    float *bout = BOut + (size_t)s * 3 * np;
    float *a = acc + (size_t)s * 3 * ni;
#pragma unroll
    for (int c = 0; c < kCellsPerThread; ++c) {
        int i = tid + c * kThreads;
        int z = tb.z0 + i % kTZ, x = tb.x0 + i / kTZ;
        if (z > tb.z1 || x > tb.x1) continue;
        int l = L(z, x, zo, xo), k = (x - 1) * d.ld + (z - 1);
        float pb = sQ[l];
        bout[k] = pb;
        bout[np + k] = sU[l];
        bout[2 * np + k] = sW[l];
        int jz = z - npml, jx = x - npml;
        if (jz >= 1 && jz <= nz && jx >= 1 && jx <= nx) {
            int ki = (jx - 1) * nz + (jz - 1);
            float pr = prn[c];
            a[ki] = a[ki] + pr * (sP[l] - pb);
            a[ni + ki] = a[ni + ki] + pr * pr;
            a[2 * ni + ki] = a[2 * ni + ki] + pb * pb;
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
    int Gf = 1;                  // shots per forward workgroup
    int ngroups_f = 0;
    TileMap fwd{}, bwd{};
    int *rec_start = nullptr;    // [ns][fwd.ntiles+1]
    RecRef *rec_list = nullptr;  // [ns][ngmax]
    int *inj_start = nullptr;    // [ns][bwd.ntiles+1] group ranges per tile
    InjGroup *inj_groups = nullptr;  // [ns][ngmax]
    RecRef *inj_list = nullptr;  // [ns][ngmax]
    int *strip_start = nullptr;  // [fwd.ntiles+1]
    StripRef *strip_list = nullptr;
    float2 *CP = nullptr;        // [nxp][ld]
    float4 *CV = nullptr;        // [nxp][ld]
    float *F[2] = {nullptr, nullptr};   // forward state [ns][3][nxp][ld], ping-pong
    float *Bk[2] = {nullptr, nullptr};  // adjoint state [ns][3][nxp][ld]
    float *R[2] = {nullptr, nullptr};   // reconstruction state [ns][3][nx][nz]
    float *end = nullptr;        // saved final forward state [ns][3][nx][nz]
    float *src = nullptr;        // plan-owned copy of the source traces
    float *acc = nullptr;        // [ns][3][nx][nz] per-shot FP32 accumulators
    double *out = nullptr;       // [3][nx][nz]
    hipStream_t cap = nullptr;   // capture stream
    hipGraphExec_t gfwd[2] = {nullptr, nullptr};
    hipGraphExec_t gbwd = nullptr;
    int use_graph = 1;
    int profile = 0;             // record events around the step kernels
    std::vector<hipEvent_t> ev;  // [2 * nt] start/stop pairs
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
