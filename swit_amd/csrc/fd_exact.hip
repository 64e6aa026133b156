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
__global__ void k_ex_inject(Dims d, int it, const ShotGeom *__restrict__ geo, const int *__restrict__ rec_cells,
                            const float *__restrict__ g, float *__restrict__ lam) {
    const int s = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= geo[s].ng) return;
    const size_t np = (size_t)d.nxp * d.ld;
    atomicAdd(&lam[(size_t)s * 3 * np + rec_cells[(size_t)s * d.ngmax + r]],
              g[((size_t)s * d.ngmax + r) * d.nt + (it - 1)]);
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

// dJ/dc on the interior: sum over shots (FP64) of the K-gradient of every padded cell
// that replicates the interior cell (io.f90:231-300 padding), times dK/dc = 2 den c dt/dx.
__global__ void k_ex_fold(Dims d, int ns, const float *__restrict__ vp, const float *__restrict__ rho,
                          const float *__restrict__ gk, double *__restrict__ grad) {
    const int jz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int jx = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (jz > d.nz || jx > d.nx) return;
    const int npml = d.npml;
    const int z0 = jz == 1 ? 1 : npml + jz, z1 = jz == d.nz ? d.nzp : npml + jz;
    const int x0 = jx == 1 ? 1 : npml + jx, x1 = jx == d.nx ? d.nxp : npml + jx;
    const size_t np = (size_t)d.nxp * d.ld;
    double sum = 0.0;
    for (int s = 0; s < ns; ++s) {
        const float *g = gk + (size_t)s * np;
        for (int ix = x0; ix <= x1; ++ix)
            for (int iz = z0; iz <= z1; ++iz) sum += (double)g[pidx(d, iz, ix)];
    }
    const size_t ki = iidx(d, jz, jx);
    grad[ki] = sum * 2.0 * (double)rho[ki] * (double)vp[ki] * (double)d.dt / (double)d.dx;
}

#undef EX_CELL

}  // namespace ex
}  // namespace swit
