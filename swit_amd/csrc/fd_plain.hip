// Per-step kernels of the plain path: one thread per cell, one launch per sweep,
// grid.z = shots in the launch.  Each kernel restates one sweep of
// fd2dmpi/fdcore.f90 with the reference's loop bounds and operation order.
#include "fd_common.h"
#include "fd_kernels.h"

namespace swit {

// ---- setup ----------------------------------------------------------------------

// Padding (io.f90:231-300 reduces to clamping the interior index) and coefficients.
__global__ void k_setup_coef(Dims d, const float *__restrict__ vp, const float *__restrict__ rho,
                             float *__restrict__ K, float *__restrict__ B) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (iz > d.nzp || ix > d.nxp) return;
    int jz = min(max(iz - d.npml, 1), d.nz);
    int jx = min(max(ix - d.npml, 1), d.nx);
    float c = vp[(size_t)(jx - 1) * d.nz + (jz - 1)];
    float den = rho[(size_t)(jx - 1) * d.nz + (jz - 1)];
    // fdcore.f90:79 alpha = den*c*c*dt/dx ; :99 alpha = dt/(den*dx)
    K[pidx(d, iz, ix)] = den * c * c * d.dt / d.dx;
    B[pidx(d, iz, ix)] = d.dt / (den * d.dx);
}

// vmin over interior rows 2..nz (math.f90:342-361 via solver.f90:73), then the
// damping profile (pml.f90:80-87).  One workgroup.  Also the model statistics the
// plan checks before any step runs: stats[0] = max vp (the stability limit of
// toolbox/base.py:57-59), stats[1] = number of vp or rho values that are not finite
// and positive.
__global__ void k_setup_pml(Dims d, const float *__restrict__ vp, const float *__restrict__ rho,
                            float *__restrict__ prof, float *__restrict__ vmin_out, float *__restrict__ stats) {
    __shared__ float red[1024], rmax[1024];
    __shared__ int rbad[1024];
    float v = 10000000000.0f, vmax = 0.0f;
    int bad = 0;
    const int n = d.nx * d.nz;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        int iz = k % d.nz + 1;
        const float c = vp[k], r = rho[k];
        if (iz > 1) v = fminf(v, c);
        vmax = fmaxf(vmax, c);
        bad += !(c > 0.0f && c <= 3.0e38f) + !(r > 0.0f && r <= 3.0e38f);
    }
    red[threadIdx.x] = v;
    rmax[threadIdx.x] = vmax;
    rbad[threadIdx.x] = bad;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[threadIdx.x] = fminf(red[threadIdx.x], red[threadIdx.x + s]);
            rmax[threadIdx.x] = fmaxf(rmax[threadIdx.x], rmax[threadIdx.x + s]);
            rbad[threadIdx.x] += rbad[threadIdx.x + s];
        }
        __syncthreads();
    }
    float vmin = red[0];
    float a = (float)(d.npml - 1) * d.dx;
    float kappa = 3.0f * vmin * kLn1000 / (2.0f * a);
    for (int i = threadIdx.x + 1; i <= d.npml; i += blockDim.x) {
        float xa = (float)(i - 1) * d.dx / a;
        prof[i] = kappa * xa * xa;
    }
    if (threadIdx.x == 0) {
        prof[0] = 0.0f;
        *vmin_out = vmin;
        stats[0] = rmax[0];
        stats[1] = (float)rbad[0];
    }
}

__global__ void k_setup_damp(Dims d, const float *__restrict__ prof, float *__restrict__ damp) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (iz > d.nzp || ix > d.nxp) return;
    damp[pidx(d, iz, ix)] = damp_at(d, prof, iz, ix);
}

// ---- forward (fdcore.f90:75-160) -----------------------------------------------

// p-sweep (:77-85), injection (:88-89) and the p mirror (:92-94).
__global__ void k_fwd_p(Dims d, int it, int shot0, const float *__restrict__ K,
                        const float *__restrict__ damp, const ShotGeom *__restrict__ geo,
                        const float *__restrict__ src, float *__restrict__ P,
                        const float *__restrict__ U, const float *__restrict__ W) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int s = shot0 + blockIdx.z;
    if (iz > d.nzp || ix > d.nxp) return;
    size_t off = (size_t)s * d.nxp * d.ld;
    float *p = P + off;
    const float *u = U + off, *w = W + off;
    if (iz >= d.fs + 1 && iz <= d.nzp - 2 && ix >= 3 && ix <= d.nxp - 2) {
        size_t k = pidx(d, iz, ix);
        float kappa = damp[k] * d.dt;
        float v = p_update(p[k], K[k], kappa,
                           u[k], u[k - d.ld], u[k + d.ld], u[k - 2 * d.ld],
                           w[k], w[k - 1], w[k + 1], w[k - 2]);
        const ShotGeom g = geo[s];
        if (iz == g.isz && ix == g.isx) v = v + d.dt * src[(size_t)s * d.nt + (it - 1)];
        p[k] = v;
        if (d.free_surface == 1 && iz == d.fs + 1) p[pidx(d, d.fs - 1, ix)] = -v;
    } else if (d.free_surface == 1 && iz == d.fs + 1) {
        // columns the sweep does not touch still get the mirror of their (zero) value
        p[pidx(d, d.fs - 1, ix)] = -p[pidx(d, iz, ix)];
    }
}

// u-sweep (:97-104), w-sweep (:106-113) and the w mirror (:116-120).
__global__ void k_fwd_v(Dims d, int shot0, const float *__restrict__ B,
                        const float *__restrict__ damp, const float *__restrict__ P,
                        float *__restrict__ U, float *__restrict__ W) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int s = shot0 + blockIdx.z;
    if (iz > d.nzp || ix > d.nxp) return;
    size_t off = (size_t)s * d.nxp * d.ld;
    const float *p = P + off;
    float *u = U + off, *w = W + off;
    size_t k = pidx(d, iz, ix);
    if (iz >= d.fs && iz <= d.nzp - 1 && ix >= 2 && ix <= d.nxp - 2) {
        float kappa = 0.5f * (damp[k] + damp[k + d.ld]) * d.dt;
        u[k] = v_update(u[k], B[k], kappa, p[k], p[k + d.ld], p[k + 2 * d.ld], p[k - d.ld]);
    }
    if (iz >= d.fs && iz <= d.nzp - 2 && ix >= 2 && ix <= d.nxp - 1) {
        float kappa = 0.5f * (damp[k] + damp[k + 1]) * d.dt;
        float v = v_update(w[k], B[k], kappa, p[k], p[k + 1], p[k + 2], p[k - 1]);
        w[k] = v;
        if (d.free_surface == 1 && iz == d.fs) w[pidx(d, d.fs - 1, ix)] = v;
    }
}

// Receiver sampling (:123-128) and strip save (:130-159) after step `it`.
// seis_t: [ns][nt][ngmax] (time-major staging); strips: [ns][nt][S].
__global__ void k_fwd_record(Dims d, int it, int shot0, const ShotGeom *__restrict__ geo,
                             const int *__restrict__ rec_fwd, const uint32_t *__restrict__ strip_map,
                             const float *__restrict__ P, const float *__restrict__ U,
                             const float *__restrict__ W, float *__restrict__ seis_t,
                             float *__restrict__ strips) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int s = shot0 + blockIdx.z;
    size_t off = (size_t)s * d.nxp * d.ld;
    const ShotGeom g = geo[s];
    if (t < d.ngmax) {
        float v = 0.0f;
        if (t < g.ng) v = P[off + rec_fwd[(size_t)s * d.ngmax + t]];
        seis_t[((size_t)s * d.nt + (it - 1)) * d.ngmax + t] = v;
        if (it == 2) seis_t[(size_t)s * d.nt * d.ngmax + t] = 0.0f;   // csg(1,:) = 0
    }
    if (strips && t < d.S) {
        uint32_t m = strip_map[t];
        uint32_t f = m >> kStripFieldShift, o = m & ((1u << kStripFieldShift) - 1u);
        const float *src = f == 0 ? P : (f == 1 ? U : W);
        strips[((size_t)s * d.nt + (it - 1)) * d.S + t] = src[off + o];
        if (it == 2) strips[(size_t)s * d.nt * d.S + t] = 0.0f;       // boundary(:,1) = 0
    }
}

// End state (:162-165): interior windows of p, u, w.  ends: [3][ns][nx][nz].
__global__ void k_fwd_end(Dims d, int shot0, const float *__restrict__ P, const float *__restrict__ U,
                          const float *__restrict__ W, float *__restrict__ pe, float *__restrict__ ue,
                          float *__restrict__ we) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int s = shot0 + blockIdx.z;
    if (iz > d.nz || ix > d.nx) return;
    size_t off = (size_t)s * d.nxp * d.ld;
    size_t kp = off + pidx(d, iz + d.npml, ix + d.npml);
    size_t ki = (size_t)s * d.nx * d.nz + iidx(d, iz, ix);
    pe[ki] = P[kp];
    ue[ki] = U[kp];
    we[ki] = W[kp];
}

// ---- reconstruction (fdcore.f90:202-289) on the interior window ---------------

// Velocity undo (:221-235) and strip inserts w(nz-1,:), u(:,1), u(:,nx-1) (:238-247).
// b: strips of the shot at column it-1.
__global__ void k_reco_v(Dims d, int it, int shot0, const float *__restrict__ B,
                         const float *__restrict__ strips, const float *__restrict__ Pr,
                         float *__restrict__ Ur, float *__restrict__ Wr) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int s = shot0 + blockIdx.z;
    if (iz > d.nz || ix > d.nx) return;
    size_t off = (size_t)s * d.nx * d.nz;
    const float *p = Pr + off;
    float *u = Ur + off, *w = Wr + off;
    const float *b = strips + ((size_t)s * d.nt + (it - 2)) * d.S;
    size_t k = iidx(d, iz, ix);
    float alpha = B[pidx(d, iz + d.npml, ix + d.npml)];
    const int nz = d.nz, nx = d.nx;
    // u: interior columns 2..nx-2 updated; columns 1 and nx-1 come from the strips
    if (ix == 1 || ix == nx - 1) {
        size_t jj = (size_t)3 * nx + (size_t)6 * (iz - 1);
        u[k] = ix == 1 ? b[jj + 2] : b[jj + 5];
    } else if (ix >= 2 && ix <= nx - 2 && iz >= 2 && iz <= nz - 1) {
        u[k] = v_undo(u[k], alpha, p[k], p[k + nz], p[k + 2 * nz], p[k - nz]);
    }
    // w: rows 2..nz-2 updated; row nz-1 comes from the strips
    if (iz == nz - 1) {
        w[k] = b[(size_t)3 * (ix - 1) + 2];
    } else if (ix >= 2 && ix <= nx - 1 && iz >= 2 && iz <= nz - 2) {
        w[k] = v_undo(w[k], alpha, p[k], p[k + 1], p[k + 2], p[k - 1]);
    }
}

// Source subtraction (:250-257), pressure undo (:260-267) and strip inserts
// (:269-287; top rows win over sides, sides over bottom, as in the reference order).
__global__ void k_reco_p(Dims d, int it, int shot0, const float *__restrict__ K,
                         const ShotGeom *__restrict__ geo, const float *__restrict__ src,
                         const float *__restrict__ strips, float *__restrict__ Pr,
                         const float *__restrict__ Ur, const float *__restrict__ Wr) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int s = shot0 + blockIdx.z;
    if (iz > d.nz || ix > d.nx) return;
    size_t off = (size_t)s * d.nx * d.nz;
    float *p = Pr + off;
    const float *u = Ur + off, *w = Wr + off;
    const float *b = strips + ((size_t)s * d.nt + (it - 2)) * d.S;
    const int nz = d.nz, nx = d.nx;
    size_t k = iidx(d, iz, ix);
    const size_t top0 = (size_t)3 * nx + (size_t)6 * nz;
    if (iz <= 3) {
        p[k] = b[top0 + (size_t)6 * (ix - 1) + (iz - 1)];
        return;
    }
    if (ix <= 2 || ix >= nx - 1) {
        size_t jj = (size_t)3 * nx + (size_t)6 * (iz - 1);
        p[k] = ix == 1 ? b[jj] : ix == 2 ? b[jj + 1] : ix == nx ? b[jj + 3] : b[jj + 4];
        return;
    }
    if (iz >= nz - 1) {
        size_t jj = (size_t)3 * (ix - 1);
        p[k] = iz == nz ? b[jj] : b[jj + 1];
        return;
    }
    // iz in 4..nz-2, ix in 3..nx-2: undo
    float v = p[k];
    const ShotGeom g = geo[s];
    if (iz == g.isz_i && ix == g.isx_i && g.isx_i > 2 && g.isx_i < nx - 1 && g.isz_i > 3 &&
        g.isz_i < nz - 1) {
        // the reference reads s(it+1), one past the end at it = nt (taken as 0)
        float sv = it + 1 <= d.nt ? src[(size_t)s * d.nt + it] : 0.0f;
        v = v - d.dt * sv;
    }
    float alpha = K[pidx(d, iz + d.npml, ix + d.npml)];
    p[k] = p_undo(v, alpha, u[k], u[k - nz], u[k + nz], u[k - 2 * nz],
                  w[k], w[k - 1], w[k + 1], w[k - 2]);
}

// ---- adjoint (fdcore.f90:293-375) ----------------------------------------------

// u/w sweeps with alpha = den c^2 dt/dx (:310-326) and the w mirror (:329-333).
__global__ void k_back_v(Dims d, int shot0, const float *__restrict__ K,
                         const float *__restrict__ damp, const float *__restrict__ P,
                         float *__restrict__ U, float *__restrict__ W) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int s = shot0 + blockIdx.z;
    if (iz > d.nzp || ix > d.nxp) return;
    size_t off = (size_t)s * d.nxp * d.ld;
    const float *p = P + off;
    float *u = U + off, *w = W + off;
    size_t k = pidx(d, iz, ix);
    float alpha = K[k];
    if (iz >= 2 && iz <= d.nzp - 1 && ix >= 2 && ix <= d.nxp - 2) {
        float kappa = 0.5f * (damp[k] + damp[k + d.ld]) * d.dt;
        u[k] = v_update(u[k], alpha, kappa, p[k], p[k + d.ld], p[k + 2 * d.ld], p[k - d.ld]);
    }
    if (iz >= 2 && iz <= d.nzp - 2 && ix >= 2 && ix <= d.nxp - 1) {
        bool mirror = d.free_surface == 1;
        // row fs-1 is overwritten by the mirror of row fs
        if (!(mirror && iz == d.fs - 1)) {
            float kappa = 0.5f * (damp[k] + damp[k + 1]) * d.dt;
            float v = v_update(w[k], alpha, kappa, p[k], p[k + 1], p[k + 2], p[k - 1]);
            w[k] = v;
            if (mirror && iz == d.fs) w[pidx(d, d.fs - 1, ix)] = v;
        }
    }
}

// p sweep with alpha = dt/(den dx) over rows 3..nzp-2 (:336-345), Pc -> Pn.
// Row fs-1 is left to the mirror kernel.
__global__ void k_back_p(Dims d, int shot0, const float *__restrict__ B,
                         const float *__restrict__ damp, const float *__restrict__ Pc,
                         float *__restrict__ Pn, const float *__restrict__ U,
                         const float *__restrict__ W) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int s = shot0 + blockIdx.z;
    if (iz > d.nzp || ix > d.nxp) return;
    if (!(iz >= 3 && iz <= d.nzp - 2 && ix >= 3 && ix <= d.nxp - 2)) return;
    if (iz == d.fs - 1) return;
    size_t off = (size_t)s * d.nxp * d.ld;
    const float *u = U + off, *w = W + off;
    size_t k = pidx(d, iz, ix);
    float kappa = damp[k] * d.dt;
    Pn[off + k] = p_update(Pc[off + k], B[k], kappa,
                           u[k], u[k - d.ld], u[k + d.ld], u[k - 2 * d.ld],
                           w[k], w[k - 1], w[k + 1], w[k - 2]);
}

// Adjoint-source injection (:348-367).  sadj_t: [ns][nt][ngmax].  One thread per
// injection cell (cell groups of the shot: ngrp[s] groups, each a range of the shot's
// receiver list sorted stably by cell), adding its receivers in receiver order as the
// reference's loop does, so shared cells give the reference's bits every run.
__global__ void k_back_inject(Dims d, int it, int shot0, const int *__restrict__ ngrp,
                              const InjGroup *__restrict__ grp, const RecRef *__restrict__ lst,
                              const float *__restrict__ sadj_t, float *__restrict__ Pn) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int s = shot0 + blockIdx.z;
    if (t >= ngrp[s]) return;
    const InjGroup g = grp[(size_t)s * d.ngmax + t];
    const RecRef *l = lst + (size_t)s * d.ngmax;
    float *c = Pn + (size_t)s * d.nxp * d.ld + pidx(d, l[g.e0].iz, l[g.e0].ix);
    const float *a = sadj_t + ((size_t)s * d.nt + (it - 1)) * d.ngmax;
    float acc = *c;
    for (int e = g.e0; e < g.e1; ++e) acc = acc + d.dt * a[l[e].idx];
    *c = acc;
}

// Unconditional p mirror (:370-373) and imaging (solver.f90:237-239):
// dg += p_rec (p0 - p), dfw += p_rec^2, dbk += p^2 over the interior.
__global__ void k_back_mirror_image(Dims d, int shot0, const float *__restrict__ Pc,
                                    float *__restrict__ Pn, const float *__restrict__ Pr,
                                    float *__restrict__ dg, float *__restrict__ dfw,
                                    float *__restrict__ dbk) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int ix = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int s = shot0 + blockIdx.z;
    if (iz > d.nzp || ix > d.nxp) return;
    size_t off = (size_t)s * d.nxp * d.ld;
    if (iz == d.fs - 1) {
        Pn[off + pidx(d, iz, ix)] = -Pn[off + pidx(d, d.fs + 1, ix)];
        return;
    }
    int jz = iz - d.npml, jx = ix - d.npml;
    if (jz < 1 || jz > d.nz || jx < 1 || jx > d.nx) return;
    size_t k = off + pidx(d, iz, ix);
    size_t ki = (size_t)s * d.nx * d.nz + iidx(d, jz, jx);
    float pr = Pr[ki];
    float p0 = Pc[k];
    float pb = Pn[k];
    dg[ki] = dg[ki] + pr * (p0 - pb);
    dfw[ki] = dfw[ki] + pr * pr;
    dbk[ki] = dbk[ki] + pb * pb;
}

// ---- reductions and layout ------------------------------------------------------

// Cross-shot FP64 sum (solver.f90:263-265) and g = -2 g / c (solver.f90:270).
__global__ void k_reduce_shots(Dims d, int ns, const float *__restrict__ vp,
                               const float *__restrict__ dg, const float *__restrict__ dfw,
                               const float *__restrict__ dbk, double *__restrict__ grad,
                               double *__restrict__ fw, double *__restrict__ bk) {
    size_t n = (size_t)d.nx * d.nz;
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double g = 0.0, f = 0.0, b = 0.0;
    for (int s = 0; s < ns; ++s) {
        g += (double)dg[(size_t)s * n + k];
        f += (double)dfw[(size_t)s * n + k];
        b += (double)dbk[(size_t)s * n + k];
    }
    grad[k] = -2.0 * g / (double)vp[k];
    fw[k] = f;
    bk[k] = b;
}

// [ns][a][b] -> [ns][b][a]
__global__ void k_transpose(int a, int b, const float *__restrict__ in, float *__restrict__ out) {
    __shared__ float tile[32][33];
    int s = blockIdx.z;
    int b0 = blockIdx.x * 32, a0 = blockIdx.y * 32;
    const float *src = in + (size_t)s * a * b;
    float *dst = out + (size_t)s * a * b;
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        int ia = a0 + j, ib = b0 + threadIdx.x;
        if (ia < a && ib < b) tile[j][threadIdx.x] = src[(size_t)ia * b + ib];
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        int ib = b0 + j, ia = a0 + threadIdx.x;
        if (ia < a && ib < b) dst[(size_t)ib * a + ia] = tile[threadIdx.x][j];
    }
}

}  // namespace swit
