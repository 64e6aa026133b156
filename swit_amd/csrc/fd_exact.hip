// Exact discrete adjoint of the forward scheme (the batched forward's arithmetic,
// fdcore.f90:75-128): the transpose of every operation of a forward step, applied in
// reverse order, with the divergence history the forward stored.  Gives the gradient
// of any misfit of the seismograms with respect to the velocity model through
// K = den c^2 dt/dx; density and the damping profile (and so vmin) are held fixed.
//
// Forward step it (state at it-1 -> it), R_p, R_u, R_w the update regions:
//   p~ = fp p - K D(u, w)            on R_p   D = c1(u0-u-1+w0-w-1) + c2(u+1-u-2+w+1-w-2)
//   p~(src) += dt s(it);  p = p~ with p(fs-1,:) = -p~(fs+1,:)
//   u = fu u - B Gx(p) on R_u,  w~ = fw w - B Gz(p) on R_w,  w = w~ with w(fs-1,:) = w~(fs,:)
//   d(it) = p(receiver cells)
// Reverse step it, adjoint variables (lp, lu, lw) of (p, u, w) at it:
//   lp(rec) += g(it);  lw(fs) += lw(fs-1), lw(fs-1) = 0
//   qu = B lu on R_u, qw = B lw on R_w;  lu *= fu on R_u, lw *= fw on R_w
//   lp -= Gx^T qu + Gz^T qw;  lp(fs+1) -= lp(fs-1), lp(fs-1) = 0
//   on R_p: gK -= lp D(it),  r = K lp,  lp *= fp
//   lu -= Dx^T r,  lw -= Dz^T r
// with Gx^T q (j) = c1 (q(j-1) - q(j)) + c2 (q(j-2) - q(j+1)) and
//      Dx^T r (j) = c1 (r(j) - r(j+1)) + c2 (r(j-1) - r(j+2))  (and the same in z).
// Cells outside the update regions hold fields that stay zero; their adjoint values
// never feed back and are left as they come.
#include "fd_common.h"

namespace swit {
namespace ex {

__device__ __forceinline__ bool in_rp(const Dims &d, int iz, int ix) {
    return iz >= d.fs + 1 && iz <= d.nzp - 2 && ix >= 3 && ix <= d.nxp - 2;
}
__device__ __forceinline__ bool in_ru(const Dims &d, int iz, int ix) {
    return iz >= d.fs && iz <= d.nzp - 1 && ix >= 2 && ix <= d.nxp - 2;
}
__device__ __forceinline__ bool in_rw(const Dims &d, int iz, int ix) {
    return iz >= d.fs && iz <= d.nzp - 2 && ix >= 2 && ix <= d.nxp - 1;
}
// value of a [nxp][ld] field at (iz, ix), zero outside the padded grid
__device__ __forceinline__ float at(const Dims &d, const float *f, int iz, int ix) {
    return (iz >= 1 && iz <= d.nzp && ix >= 1 && ix <= d.nxp) ? f[pidx(d, iz, ix)] : 0.0f;
}

#define EX_CELL                                                       \
    const int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;         \
    const int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;         \
    const int s = blockIdx.z;                                         \
    if (iz > d.nzp || ix > d.nxp) return;                             \
    const size_t np = (size_t)d.nxp * d.ld;                           \
    const size_t k = pidx(d, iz, ix);                                 \
    (void)np;                                                         \
    (void)k

// lam: [ns][3][np]; q: [ns][2][np]; gk: [ns][np]; D: divergence history (DSave)
// One thread per recording cell (the shot's cell groups), adding its receivers'
// dJ/dseis in receiver order, as the fused step does.
__global__ void k_ex_inject(Dims d, int it, const int *__restrict__ ngrp, const InjGroup *__restrict__ grp,
                            const RecRef *__restrict__ lst, const float *__restrict__ g, float *__restrict__ lam) {
    const int s = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ngrp[s]) return;
    const size_t np = (size_t)d.nxp * d.ld;
    const InjGroup q = grp[(size_t)s * d.ngmax + t];
    const RecRef *l = lst + (size_t)s * d.ngmax;
    float *c = lam + (size_t)s * 3 * np + pidx(d, l[q.e0].iz, l[q.e0].ix);
    float acc = *c;
    for (int e = q.e0; e < q.e1; ++e) acc += g[((size_t)s * d.ngmax + l[e].idx) * d.nt + (it - 1)];
    *c = acc;
}

// grid (ceil(nxp/256), ns): one thread per column
__global__ void k_ex_mirror(Dims d, int field, float *__restrict__ lam) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int s = blockIdx.y;
    if (ix > d.nxp) return;
    const size_t np = (size_t)d.nxp * d.ld;
    float *f = lam + ((size_t)s * 3 + field) * np;
    const int fs = d.fs;
    if (field == 2) {            // w(fs-1) = w(fs)
        f[pidx(d, fs, ix)] += f[pidx(d, fs - 1, ix)];
        f[pidx(d, fs - 1, ix)] = 0.0f;
    } else {                     // p(fs-1) = -p(fs+1)
        f[pidx(d, fs + 1, ix)] -= f[pidx(d, fs - 1, ix)];
        f[pidx(d, fs - 1, ix)] = 0.0f;
    }
}

__global__ void k_ex_vel(Dims d, const float2 *__restrict__ KB, const float4 *__restrict__ CV,
                         float *__restrict__ lam, float *__restrict__ q) {
    EX_CELL;
    float *lu = lam + ((size_t)s * 3 + 1) * np, *lw = lam + ((size_t)s * 3 + 2) * np;
    float *qu = q + (size_t)s * 2 * np, *qw = qu + np;
    const bool ru = in_ru(d, iz, ix), rw = in_rw(d, iz, ix);
    const float B = KB[k].y;
    const float4 cv = CV[k];
    const float u = lu[k], w = lw[k];
    qu[k] = ru ? B * u : 0.0f;
    qw[k] = rw ? B * w : 0.0f;
    if (ru) lu[k] = cv.y * u;
    if (rw) lw[k] = cv.z * w;
}

__global__ void k_ex_p(Dims d, const float *__restrict__ q, float *__restrict__ lam) {
    EX_CELL;
    const float *qu = q + (size_t)s * 2 * np, *qw = qu + np;
    float *lp = lam + (size_t)s * 3 * np;
    const float gx = kC1 * (at(d, qu, iz, ix - 1) - at(d, qu, iz, ix)) +
                     kC2 * (at(d, qu, iz, ix - 2) - at(d, qu, iz, ix + 1));
    const float gz = kC1 * (at(d, qw, iz - 1, ix) - at(d, qw, iz, ix)) +
                     kC2 * (at(d, qw, iz - 2, ix) - at(d, qw, iz + 1, ix));
    lp[k] = lp[k] - gx - gz;
}

__global__ void k_ex_grad(Dims d, int it, const float2 *__restrict__ KB, const float2 *__restrict__ CP,
                          DSave D, float *__restrict__ lam, float *__restrict__ q, float *__restrict__ gk) {
    EX_CELL;
    float *lp = lam + (size_t)s * 3 * np;
    float *r = q + (size_t)s * 2 * np;
    if (!in_rp(d, iz, ix)) {
        r[k] = 0.0f;
        return;
    }
    const float p = lp[k];
    gk[(size_t)s * np + k] -= p * D.p[((size_t)s * D.slots + (it - 1 - D.base)) * np + k];
    r[k] = KB[k].x * p;
    lp[k] = CP[k].y * p;
}

__global__ void k_ex_vel2(Dims d, const float *__restrict__ q, float *__restrict__ lam) {
    EX_CELL;
    const float *r = q + (size_t)s * 2 * np;
    float *lu = lam + ((size_t)s * 3 + 1) * np, *lw = lam + ((size_t)s * 3 + 2) * np;
    const float dx = kC1 * (at(d, r, iz, ix) - at(d, r, iz, ix + 1)) + kC2 * (at(d, r, iz, ix - 1) - at(d, r, iz, ix + 2));
    const float dz = kC1 * (at(d, r, iz, ix) - at(d, r, iz + 1, ix)) + kC2 * (at(d, r, iz - 1, ix) - at(d, r, iz + 2, ix));
    lu[k] = lu[k] - dx;
    lw[k] = lw[k] - dz;
}

// Tile width of the fused step (its own tiling of the backward rows), window columns
// (tile + 3-cell halo on each side) and floats per LDS array.
constexpr int kExTX = 32;
constexpr int kExCX = kExTX + 2 * kHalo;
constexpr int kExNL = bk::kZL * kExCX;

// ---------------------------------------------------------------------------------
// One reverse step it of all shots, fused: the operations of k_ex_inject, k_ex_mirror
// (w), k_ex_vel, k_ex_p, k_ex_mirror (p), k_ex_grad and k_ex_vel2, in that order, with
// the same arithmetic per cell, so the same bits (receivers sharing a cell are added in
// receiver order).  One workgroup per tile (58 x kExTX, the backward rows) of one shot.
// Its window is the tile plus 3 cells per side (64 x kExCX): lane l of every wave owns window
// row l, and a thread's cells are columns wave + 8j of its row, so the z neighbours of a
// cell are the adjacent lanes' values (qw, r and the w mirror use DPP wave shifts;
// lane_up: lane l gets lane l+1's value, lane_dn: lane l-1's, 0 past the ends, called
// with every lane active).
// - stage lp of the window in LDS and lu, lw of the thread's cells in registers (only
//   their own cells read them); inject into lp; transpose the w mirror;
// - qu = B lu (LDS), qw = B lw (registers), lu *= fu, lw *= fw on the window;
// - lp -= Gx^T qu + Gz^T qw on window rows and columns 2 .. extent-2 (the region whose r
//   the tile's velocity stencils read), then transpose the p mirror (column by column,
//   in the wave that owns the column);
// - on the same region: gK -= lp D(it) (tile cells), r = K lp (LDS), lp *= fp;
// - on the tile: lu -= Dx^T r, lw -= Dz^T r; write lp, lu, lw and gK.
// The tiling keeps the mirror rows away from window edges where a value they need
// would lie outside the window: no tile starts at rows fs-1 .. fs+3.
// One launch covers shots s0 .. s0 + gridDim.x / ntiles - 1.
// LIn/LOut: [ns][3][nxp][ld] (lp, lu, lw), ping-pong; gk: [ns][nxp][ld];
// g: [ns][ngmax][nt]; inj_*: receiver groups at recording cells of the tiles grown by 3.
__device__ __forceinline__ float lane_up(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_dn(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__global__ __launch_bounds__(kThreads) void k_ex_step(
    Dims d, int it, TileMap tm, int s0, const float2 *__restrict__ KB, const float2 *__restrict__ CP,
    const float4 *__restrict__ CV, DSave D, const float *__restrict__ LIn, float *__restrict__ LOut,
    float *__restrict__ gk, const int *__restrict__ inj_gstart, const InjGroup *__restrict__ inj_groups,
    const RecRef *__restrict__ inj_list, int inj_cap, const float *__restrict__ g) {
    using namespace bk;
    __shared__ float sP[kExNL];
    __shared__ float sA[kExNL];     // qu, then r

    int s, t;
    work_item(xcd_logical(blockIdx.x, gridDim.x), (int)(gridDim.x / tm.ntiles), tm, s, t);
    s += s0;
    const TileBox tb = tile_box(d, tm, t);
    const int zo = tb.z0 - kHalo, xo = tb.x0 - kHalo;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int np = d.nxp * d.ld;
    const int z = zo + lane;
    const bool zin = z >= 1 && z <= d.nzp;
    const bool lane_r = lane >= 2 && lane <= kZL - 2;          // rows of the lp / r region
    const bool lane_t = lane >= kHalo && z <= tb.z1;           // tile rows
    const int ct_end = tb.x1 - xo;                             // last tile column
    const int m = d.fs - 1 - zo;                               // window row of fs-1
    constexpr int kJS = (kExCX + kWaves - 1) / kWaves;
    const float *lin = LIn + (size_t)s * 3 * np;
    float *lout = LOut + (size_t)s * 3 * np;
    float *gks = gk + (size_t)s * np;
    const float *Dp = D.p + ((size_t)s * D.slots + (it - 1 - D.base)) * np;

    // ---- stage lp (LDS: injection writes other threads' cells); lu, lw of this
    // thread's cells into registers (only their own cells read them) ----
    float lu[kJS], lw[kJS], qw[kJS];
#pragma unroll
    for (int j = 0; j < kJS; ++j) {
        const int c = wave + kWaves * j;
        lu[j] = lw[j] = qw[j] = 0.f;
        if (c >= kExCX) continue;
        const int x = xo + c;
        if (zin && x >= 1 && x <= d.nxp) {
            const unsigned k = (unsigned)((x - 1) * d.ld + (z - 1));
            __builtin_amdgcn_global_load_lds(lin + k, sP + LW(c, 0), 4, 0, 0);
            lu[j] = lin[np + k];
            lw[j] = lin[2 * np + k];
        } else {
            sP[LW(c, lane)] = 0.f;
        }
    }
    // coefficients of this thread's cells: B (window), K and 1-kappa_p (region), the
    // velocity factors, history and gK (tile)
    float2 kb[kJS];
    float fu[kJS], fw[kJS], fp[kJS], dv[kJS], gkv[kJS];
#pragma unroll
    for (int j = 0; j < kJS; ++j) {
        const int c = wave + kWaves * j;
        const int x = xo + c;
        kb[j] = make_float2(0.f, 0.f);
        fu[j] = fw[j] = fp[j] = 1.0f;
        dv[j] = gkv[j] = 0.f;
        if (c >= kExCX || !zin || x < 1 || x > d.nxp) continue;
        const unsigned k = (unsigned)((x - 1) * d.ld + (z - 1));
        kb[j] = KB[k];
        if (c >= 2 && c <= kExCX - 2 && lane_r) fp[j] = CP[k].y;
        if (c >= kHalo && c <= ct_end && lane_t) {
            const float4 cv = CV[k];
            fu[j] = cv.y;
            fw[j] = cv.z;
            if (in_rp(d, z, x)) {
                dv[j] = Dp[k];
                gkv[j] = gks[k];
            }
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    // ---- lp(receivers) += g(it) (k_ex_inject), one thread per cell in receiver order ----
    {
        const int g0 = inj_gstart[(size_t)s * (tm.ntiles + 1) + t];
        const int g1 = inj_gstart[(size_t)s * (tm.ntiles + 1) + t + 1];
        const RecRef *el = inj_list + (size_t)s * inj_cap;
        for (int q = g0 + (int)threadIdx.x; q < g1; q += kThreads) {
            const InjGroup gr = inj_groups[(size_t)s * inj_cap + q];
            const RecRef r0 = el[gr.e0];
            const int l = LW(r0.ix - xo, r0.iz - zo);
            float v = sP[l];
            for (int e = gr.e0; e < gr.e1; ++e) v = v + g[((size_t)s * d.ngmax + el[e].idx) * d.nt + (it - 1)];
            sP[l] = v;
        }
    }
    // ---- w mirror transposed (k_ex_mirror, w): lw(fs) += lw(fs-1), lw(fs-1) = 0; each
    // column in the wave that owns it ----
    if (d.free_surface == 1 && m >= 0 && m < kZL) {
#pragma unroll
        for (int j = 0; j < kJS; ++j) {
            const int c = wave + kWaves * j;
            const int x = xo + c;
            if (c >= kExCX || x < 1 || x > d.nxp) continue;
            const float wm = lane_dn(lw[j]);            // lane m+1 gets lane m's lw
            if (lane == m + 1) lw[j] = lw[j] + wm;
            if (lane == m) lw[j] = 0.0f;
        }
    }
    // ---- qu = B lu, qw = B lw (zero off R_u / R_w), lu *= fu, lw *= fw (k_ex_vel) ----
#pragma unroll
    for (int j = 0; j < kJS; ++j) {
        const int c = wave + kWaves * j;
        if (c >= kExCX) continue;
        const int x = xo + c;
        const int l = LW(c, lane);
        const bool g_in = zin && x >= 1 && x <= d.nxp;
        const bool ru = g_in && in_ru(d, z, x), rw = g_in && in_rw(d, z, x);
        const float u = lu[j], w = lw[j];
        sA[l] = ru ? kb[j].y * u : 0.0f;
        qw[j] = rw ? kb[j].y * w : 0.0f;
        lu[j] = ru ? fu[j] * u : u;
        lw[j] = rw ? fw[j] * w : w;
    }
    __syncthreads();
    // ---- lp -= Gx^T qu + Gz^T qw (k_ex_p) on the region, then the p mirror transposed
    // (k_ex_mirror, p): lp(fs+1) -= lp(fs-1), lp(fs-1) = 0 ----
#pragma unroll
    for (int j = 0; j < kJS; ++j) {
        const int c = wave + kWaves * j;
        const int x = xo + c;
        if (c < 2 || c > kExCX - 2 || x < 1 || x > d.nxp) continue;
        // qw of the rows above and below: the neighbouring lanes' values
        const float qd1 = lane_dn(qw[j]), qd2 = lane_dn(qd1), qu1 = lane_up(qw[j]);
        if (!lane_r || !zin) continue;
        const int l = LW(c, lane);
        const float gx = kC1 * (sA[l - kZL] - sA[l]) + kC2 * (sA[l - 2 * kZL] - sA[l + kZL]);
        const float gz = kC1 * (qd1 - qw[j]) + kC2 * (qd2 - qu1);
        sP[l] = sP[l] - gx - gz;
    }
    if (d.free_surface == 1 && m >= 0 && m < kZL) {
#pragma unroll
        for (int j = 0; j < kJS; ++j) {
            const int c = wave + kWaves * j;
            const int x = xo + c;
            if (c < 2 || c > kExCX - 2 || x < 1 || x > d.nxp) continue;
            const float pm = sP[LW(c, m)];
            if (m + 2 < kZL && lane == m + 2) sP[LW(c, m + 2)] = sP[LW(c, m + 2)] - pm;
            if (lane == m) sP[LW(c, m)] = 0.0f;
        }
    }
    __syncthreads();
    // ---- on R_p: gK -= lp D(it), r = K lp, lp *= fp; r = 0 elsewhere (k_ex_grad) ----
    float lp[kJS], rr[kJS];
#pragma unroll
    for (int j = 0; j < kJS; ++j) {
        const int c = wave + kWaves * j;
        lp[j] = rr[j] = 0.f;
        if (c >= kExCX) continue;
        const int x = xo + c;
        const int l = LW(c, lane);
        const float p = sP[l];
        lp[j] = p;
        if (c < 2 || c > kExCX - 2 || !lane_r) continue;
        float r = 0.0f;
        if (zin && x >= 1 && x <= d.nxp && in_rp(d, z, x)) {
            if (c >= kHalo && c <= ct_end && lane_t) gkv[j] = gkv[j] - p * dv[j];
            r = kb[j].x * p;
            lp[j] = fp[j] * p;
        }
        sA[l] = r;
        rr[j] = r;
    }
    __syncthreads();
    // ---- lu -= Dx^T r, lw -= Dz^T r (k_ex_vel2) on the tile; write back ----
#pragma unroll
    for (int j = 0; j < kJS; ++j) {
        const int c = wave + kWaves * j;
        if (c < kHalo || c > ct_end) continue;
        const float ru1 = lane_up(rr[j]), ru2 = lane_up(ru1), rd1 = lane_dn(rr[j]);
        if (!lane_t) continue;
        const int x = xo + c;
        const int l = LW(c, lane);
        const float dx = kC1 * (sA[l] - sA[l + kZL]) + kC2 * (sA[l - kZL] - sA[l + 2 * kZL]);
        const float dz = kC1 * (rr[j] - ru1) + kC2 * (rd1 - ru2);
        const unsigned k = (unsigned)((x - 1) * d.ld + (z - 1));
        lout[k] = lp[j];
        lout[np + k] = lu[j] - dx;
        lout[2 * np + k] = lw[j] - dz;
        if (in_rp(d, z, x)) gks[k] = gkv[j];
    }
}

// FP64 sum over shots (in shot order) of the K-gradient of every padded cell.
// gs: [nxp][ld].  grid-stride over the cells.
__global__ void k_ex_shotsum(Dims d, int ns, const float *__restrict__ gk, double *__restrict__ gs) {
    const size_t np = (size_t)d.nxp * d.ld;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (size_t)gridDim.x * blockDim.x) {
        double sum = 0.0;
        for (int s = 0; s < ns; ++s) sum += (double)gk[(size_t)s * np + k];
        gs[k] = sum;
    }
}

// dJ/dc on the interior: the shot-summed K-gradient of every padded cell that replicates
// the interior cell (io.f90:231-300 padding), summed in FP64 column by column, times
// dK/dc = 2 den c dt/dx.
__global__ void k_ex_fold(Dims d, const float *__restrict__ vp, const float *__restrict__ rho,
                          const double *__restrict__ gs, double *__restrict__ grad) {
    const int jz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int jx = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (jz > d.nz || jx > d.nx) return;
    const int npml = d.npml;
    const int z0 = jz == 1 ? 1 : npml + jz, z1 = jz == d.nz ? d.nzp : npml + jz;
    const int x0 = jx == 1 ? 1 : npml + jx, x1 = jx == d.nx ? d.nxp : npml + jx;
    double sum = 0.0;
    for (int ix = x0; ix <= x1; ++ix)
        for (int iz = z0; iz <= z1; ++iz) sum += gs[pidx(d, iz, ix)];
    const size_t ki = iidx(d, jz, jx);
    grad[ki] = sum * 2.0 * (double)rho[ki] * (double)vp[ki] * (double)d.dt / (double)d.dx;
}

#undef EX_CELL

}  // namespace ex
}  // namespace swit
