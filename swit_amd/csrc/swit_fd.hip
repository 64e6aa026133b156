// C ABI and plan layer of the gfx950 2-D acoustic forward/adjoint solver
// (declared in include/swit_fd.h).  One translation unit: the kernel files are
// included below so the library is a single hipcc invocation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "swit_fd.h"
#include "fd_common.h"
#include "fd_kernels.h"
#include "fd_plain.hip"
#include "fd_batched.hip"

using namespace swit;

namespace {

thread_local char g_err[1024] = "";

int set_err(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_err(SWIT_EHIP, "%s failed: %s (%s:%d)", #expr,                   \
                           hipGetErrorString(e_), __FILE__, __LINE__);                  \
    } while (0)

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

struct swit_plan {
    Dims d{};
    int ns = 0;
    int flags = 0;
    std::vector<ShotGeom> geo_h;
    std::vector<void *> allocs;
    int64_t bytes = 0;
    // geometry
    ShotGeom *geo = nullptr;
    int *rec_fwd = nullptr;       // [ns][ngmax] padded offset of the recording cell
    int *rec_inj = nullptr;       // [ns][ngmax] padded offset of the adjoint injection cell
    uint32_t *strip_map = nullptr;
    // model
    float *K = nullptr, *B = nullptr, *damp = nullptr, *prof = nullptr, *vmin = nullptr;
    float *c_int = nullptr;
    // wavefields
    float *P = nullptr, *U = nullptr, *W = nullptr;
    float *Pb[2] = {nullptr, nullptr}, *Ub = nullptr, *Wb = nullptr;
    float *Pr = nullptr, *Ur = nullptr, *Wr = nullptr;
    float *pe = nullptr, *ue = nullptr, *we = nullptr;
    float *strips = nullptr;
    float *seis_t = nullptr, *sadj_t = nullptr;
    float *dg = nullptr, *dfw = nullptr, *dbk = nullptr;
    bool model_set = false;
    bool strips_valid = false;
    BatchedState bat;

    template <typename T>
    int alloc(T **p, size_t count) {
        void *q = nullptr;
        size_t nbytes = count * sizeof(T);
        if (nbytes == 0) nbytes = 256;
        hipError_t e = hipMalloc(&q, nbytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_err(SWIT_ENOMEM, "hipMalloc of %zu bytes failed: %s", nbytes,
                           hipGetErrorString(e));
        }
        allocs.push_back(q);
        bytes += (int64_t)nbytes;
        *p = (T *)q;
        return SWIT_OK;
    }
    void release() {
        for (void *q : allocs) (void)hipFree(q);
        allocs.clear();
    }
};

#define TRY(expr)                 \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != SWIT_OK) {     \
            return rc_;           \
        }                         \
    } while (0)

static int plan_build(swit_plan *pl, const swit_grid *g, const swit_shots *sh, int flags) {
    if (!g || !sh) return set_err(SWIT_EINVAL, "grid and shots are required");
    if (g->nx < 8 || g->nz < 8) return set_err(SWIT_EINVAL, "nx, nz must be >= 8 (got %d, %d)", g->nx, g->nz);
    if (g->npml < 3 || g->npml > 1000) return set_err(SWIT_EINVAL, "npml must be in [3, 1000] (got %d)", g->npml);
    if (g->nt < 2) return set_err(SWIT_EINVAL, "nt must be >= 2 (got %d)", g->nt);
    if (!(g->dx > 0.0f) || !(g->dt > 0.0f)) return set_err(SWIT_EINVAL, "dx and dt must be positive");
    if (sh->ns < 1 || sh->ngmax < 1) return set_err(SWIT_EINVAL, "need ns >= 1 and ngmax >= 1");
    if (!sh->ng || !sh->src_xz || !sh->rec_xz) return set_err(SWIT_EINVAL, "shot arrays are required");
    if ((flags & SWIT_STRIPS_F16) && (flags & SWIT_STRIPS_BF16))
        return set_err(SWIT_EINVAL, "choose one strip precision");

    Dims &d = pl->d;
    d.nx = g->nx;
    d.nz = g->nz;
    d.npml = g->npml;
    d.nxp = g->nx + 2 * g->npml;
    d.nzp = g->nz + 2 * g->npml;
    d.ld = (d.nzp + 31) / 32 * 32;
    d.fs = g->npml + 1;
    d.nt = g->nt;
    d.S = 9 * g->nx + 6 * g->nz;
    d.ngmax = sh->ngmax;
    d.free_surface = g->free_surface;
    d.dt = g->dt;
    d.dx = g->dx;
    pl->ns = sh->ns;
    pl->flags = flags;
    if ((int64_t)d.nxp * d.ld >= (int64_t)1 << kStripFieldShift)
        return set_err(SWIT_EINVAL, "grid too large for the strip map");

    const int ns = sh->ns, ngmax = sh->ngmax, npml = d.npml, fs = d.fs;
    pl->geo_h.resize(ns);
    std::vector<int> rf((size_t)ns * ngmax, 0), ri((size_t)ns * ngmax, 0);
    auto poff = [&](int iz, int ix) { return (ix - 1) * d.ld + (iz - 1); };
    for (int s = 0; s < ns; ++s) {
        ShotGeom &q = pl->geo_h[s];
        float xs = sh->src_xz[2 * s], zs = sh->src_xz[2 * s + 1];
        // fdcore.f90:70-72 (float32 division, truncating int)
        q.isx = npml + (int)(xs / d.dx) + 1;
        q.isz = npml + (int)(zs / d.dx) + 1;
        if (q.isz == fs) q.isz += 1;
        // fdcore.f90:216-218
        q.isx_i = (int)(xs / d.dx) + 1;
        q.isz_i = (int)(zs / d.dx) + 1;
        q.ng = sh->ng[s];
        if (q.ng < 0 || q.ng > ngmax) return set_err(SWIT_EINVAL, "shot %d: ng=%d outside [0, ngmax]", s, q.ng);
        if (q.isx < 3 || q.isx > d.nxp - 2 || q.isz < fs + 1 || q.isz > d.nzp - 2)
            return set_err(SWIT_EINVAL, "shot %d: source (%g, %g) m is outside the updated grid", s, xs, zs);
        for (int r = 0; r < q.ng; ++r) {
            float xg = sh->rec_xz[((size_t)s * ngmax + r) * 2];
            float zg = sh->rec_xz[((size_t)s * ngmax + r) * 2 + 1];
            int igx = npml + (int)(xg / d.dx) + 1;
            int igz = npml + (int)(zg / d.dx) + 1;
            // fdcore.f90:124-126 recording cell
            int gz_rec = igz == fs ? igz + 1 : igz;
            // fdcore.f90:355-359 injection cell
            int gz_inj = igz > fs ? igz : fs + 1;
            if (igx < 3 || igx > d.nxp - 2 || gz_inj < 3 || gz_inj > d.nzp - 2 || gz_rec < 1 || gz_rec > d.nzp)
                return set_err(SWIT_EINVAL, "shot %d receiver %d: (%g, %g) m is outside the updated grid",
                               s, r, xg, zg);
            rf[(size_t)s * ngmax + r] = poff(gz_rec, igx);
            ri[(size_t)s * ngmax + r] = poff(gz_inj, igx);
        }
    }
    // strip layout of fdcore.f90:130-159
    std::vector<uint32_t> sm;
    sm.reserve(d.S);
    auto put = [&](uint32_t f, int iz, int ix) { sm.push_back((f << kStripFieldShift) | (uint32_t)poff(iz, ix)); };
    const int nxp = d.nxp, nzp = d.nzp;
    for (int j = npml + 1; j <= nxp - npml; ++j) {
        put(0, nzp - npml, j); put(0, nzp - npml - 1, j); put(2, nzp - npml - 1, j);
    }
    for (int j = npml + 1; j <= nzp - npml; ++j) {
        put(0, j, npml + 1); put(0, j, npml + 2); put(1, j, npml + 1);
        put(0, j, nxp - npml); put(0, j, nxp - npml - 1); put(1, j, nxp - npml - 1);
    }
    for (int j = npml + 1; j <= nxp - npml; ++j) {
        put(0, fs, j); put(0, fs + 1, j); put(0, fs + 2, j);
        put(1, fs, j); put(1, fs + 1, j); put(2, fs, j);
    }
    if ((int)sm.size() != d.S) return set_err(SWIT_EINVAL, "internal: strip map size");

    const size_t np = (size_t)d.nxp * d.ld, ni = (size_t)d.nx * d.nz;
    TRY(pl->alloc(&pl->geo, ns));
    TRY(pl->alloc(&pl->rec_fwd, (size_t)ns * ngmax));
    TRY(pl->alloc(&pl->rec_inj, (size_t)ns * ngmax));
    TRY(pl->alloc(&pl->strip_map, d.S));
    TRY(pl->alloc(&pl->K, np));
    TRY(pl->alloc(&pl->B, np));
    TRY(pl->alloc(&pl->damp, np));
    TRY(pl->alloc(&pl->prof, 1024));
    TRY(pl->alloc(&pl->vmin, 64));
    TRY(pl->alloc(&pl->c_int, ni));
    TRY(pl->alloc(&pl->P, np * ns));
    TRY(pl->alloc(&pl->U, np * ns));
    TRY(pl->alloc(&pl->W, np * ns));
    TRY(pl->alloc(&pl->Pb[0], np * ns));
    TRY(pl->alloc(&pl->Pb[1], np * ns));
    TRY(pl->alloc(&pl->Ub, np * ns));
    TRY(pl->alloc(&pl->Wb, np * ns));
    TRY(pl->alloc(&pl->Pr, ni * ns));
    TRY(pl->alloc(&pl->Ur, ni * ns));
    TRY(pl->alloc(&pl->Wr, ni * ns));
    TRY(pl->alloc(&pl->pe, ni * ns));
    TRY(pl->alloc(&pl->ue, ni * ns));
    TRY(pl->alloc(&pl->we, ni * ns));
    TRY(pl->alloc(&pl->strips, (size_t)ns * d.nt * d.S));
    TRY(pl->alloc(&pl->seis_t, (size_t)ns * d.nt * ngmax));
    TRY(pl->alloc(&pl->sadj_t, (size_t)ns * d.nt * ngmax));
    TRY(pl->alloc(&pl->dg, ni * ns));
    TRY(pl->alloc(&pl->dfw, ni * ns));
    TRY(pl->alloc(&pl->dbk, ni * ns));
    HIP_TRY(hipMemcpy(pl->geo, pl->geo_h.data(), sizeof(ShotGeom) * ns, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pl->rec_fwd, rf.data(), sizeof(int) * rf.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pl->rec_inj, ri.data(), sizeof(int) * ri.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pl->strip_map, sm.data(), sizeof(uint32_t) * sm.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(pl->P, 0, np * ns * sizeof(float)));
    HIP_TRY(hipMemset(pl->U, 0, np * ns * sizeof(float)));
    HIP_TRY(hipMemset(pl->W, 0, np * ns * sizeof(float)));
    if (flags & SWIT_PATH_BATCHED) TRY(batched_init(pl));
    return SWIT_OK;
}

extern "C" int swit_plan_create(const swit_grid *grid, const swit_shots *shots, int flags,
                                swit_plan **out) {
    g_err[0] = 0;
    if (!out) return set_err(SWIT_EINVAL, "out is NULL");
    *out = nullptr;
    swit_plan *pl = new (std::nothrow) swit_plan();
    if (!pl) return set_err(SWIT_ENOMEM, "host allocation failed");
    int rc = plan_build(pl, grid, shots, flags);
    if (rc != SWIT_OK) {
        batched_release(pl);
        pl->release();
        delete pl;
        return rc;
    }
    *out = pl;
    return SWIT_OK;
}

extern "C" int swit_plan_destroy(swit_plan *plan) {
    if (!plan) return SWIT_OK;
    (void)hipDeviceSynchronize();
    batched_release(plan);
    plan->release();
    delete plan;
    return SWIT_OK;
}

extern "C" const char *swit_last_error(void) { return g_err; }

extern "C" int64_t swit_plan_scratch_bytes(const swit_plan *plan) { return plan ? plan->bytes : 0; }

extern "C" int swit_plan_set_model(swit_plan *pl, const float *vp, const float *rho,
                                   hipStream_t stream) {
    g_err[0] = 0;
    if (!pl || !vp || !rho) return set_err(SWIT_EINVAL, "plan, vp and rho are required");
    const Dims &d = pl->d;
    dim3 blk(64, 4);
    dim3 grd(cdiv(d.nzp, 64), cdiv(d.nxp, 4), 1);
    hipLaunchKernelGGL(k_setup_coef, grd, blk, 0, stream, d, vp, rho, pl->K, pl->B);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_setup_pml, dim3(1), dim3(1024), 0, stream, d, vp, pl->prof, pl->vmin);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_setup_damp, grd, blk, 0, stream, d, pl->prof, pl->damp);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(pl->c_int, vp, sizeof(float) * d.nx * d.nz, hipMemcpyDeviceToDevice, stream));
    pl->model_set = true;
    pl->strips_valid = false;
    if (pl->flags & SWIT_PATH_BATCHED) TRY(batched_set_model(pl, stream));
    return SWIT_OK;
}

// ---- plain path -------------------------------------------------------------------

static int plain_forward(swit_plan *pl, const float *src, int save, hipStream_t st) {
    const Dims &d = pl->d;
    const size_t np = (size_t)d.nxp * d.ld;
    HIP_TRY(hipMemsetAsync(pl->P, 0, np * pl->ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->U, 0, np * pl->ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->W, 0, np * pl->ns * sizeof(float), st));
    dim3 blk(64, 4);
    int nrec = d.S > d.ngmax ? d.S : d.ngmax;
    for (int s = 0; s < pl->ns; ++s) {
        dim3 grd(cdiv(d.nzp, 64), cdiv(d.nxp, 4), 1);
        for (int it = 2; it <= d.nt; ++it) {
            hipLaunchKernelGGL(k_fwd_p, grd, blk, 0, st, d, it, s, pl->K, pl->damp, pl->geo, src,
                               pl->P, pl->U, pl->W);
            hipLaunchKernelGGL(k_fwd_v, grd, blk, 0, st, d, s, pl->B, pl->damp, pl->P, pl->U, pl->W);
            hipLaunchKernelGGL(k_fwd_record, dim3(cdiv(nrec, 256), 1, 1), dim3(256), 0, st, d, it, s,
                               pl->geo, pl->rec_fwd, pl->strip_map, pl->P, pl->U, pl->W, pl->seis_t,
                               save ? pl->strips : nullptr);
            LAUNCH_CHECK();
        }
        if (save) {
            dim3 gi(cdiv(d.nz, 64), cdiv(d.nx, 4), 1);
            hipLaunchKernelGGL(k_fwd_end, gi, blk, 0, st, d, s, pl->P, pl->U, pl->W, pl->pe, pl->ue, pl->we);
            LAUNCH_CHECK();
        }
    }
    return SWIT_OK;
}

static int plain_adjoint(swit_plan *pl, const float *src, hipStream_t st) {
    const Dims &d = pl->d;
    const size_t np = (size_t)d.nxp * d.ld, ni = (size_t)d.nx * d.nz;
    const int ns = pl->ns;
    HIP_TRY(hipMemcpyAsync(pl->Pr, pl->pe, ni * ns * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(pl->Ur, pl->ue, ni * ns * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(pl->Wr, pl->we, ni * ns * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemsetAsync(pl->Pb[0], 0, np * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->Pb[1], 0, np * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->Ub, 0, np * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->Wb, 0, np * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->dg, 0, ni * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->dfw, 0, ni * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->dbk, 0, ni * ns * sizeof(float), st));
    dim3 blk(64, 4);
    dim3 gp(cdiv(d.nzp, 64), cdiv(d.nxp, 4), 1);
    dim3 gi(cdiv(d.nz, 64), cdiv(d.nx, 4), 1);
    for (int s = 0; s < ns; ++s) {
        int cur = 0;
        for (int it = d.nt; it >= 2; --it) {
            hipLaunchKernelGGL(k_reco_v, gi, blk, 0, st, d, it, s, pl->B, pl->strips, pl->Pr, pl->Ur, pl->Wr);
            hipLaunchKernelGGL(k_reco_p, gi, blk, 0, st, d, it, s, pl->K, pl->geo, src, pl->strips,
                               pl->Pr, pl->Ur, pl->Wr);
            hipLaunchKernelGGL(k_back_v, gp, blk, 0, st, d, s, pl->K, pl->damp, pl->Pb[cur], pl->Ub, pl->Wb);
            hipLaunchKernelGGL(k_back_p, gp, blk, 0, st, d, s, pl->B, pl->damp, pl->Pb[cur], pl->Pb[1 - cur],
                               pl->Ub, pl->Wb);
            hipLaunchKernelGGL(k_back_inject, dim3(cdiv(d.ngmax, 256), 1, 1), dim3(256), 0, st, d, it, s,
                               pl->geo, pl->rec_inj, pl->sadj_t, pl->Pb[1 - cur]);
            hipLaunchKernelGGL(k_back_mirror_image, gp, blk, 0, st, d, s, pl->Pb[cur], pl->Pb[1 - cur],
                               pl->Pr, pl->dg, pl->dfw, pl->dbk);
            LAUNCH_CHECK();
            cur = 1 - cur;
        }
    }
    return SWIT_OK;
}

// ---- public jobs -------------------------------------------------------------------

extern "C" int swit_forward(swit_plan *pl, const float *src, float *seis, int save_strips,
                            hipStream_t stream) {
    g_err[0] = 0;
    if (!pl || !src || !seis) return set_err(SWIT_EINVAL, "plan, src and seis are required");
    if (!pl->model_set) return set_err(SWIT_ESTATE, "swit_plan_set_model must be called first");
    const Dims &d = pl->d;
    pl->strips_valid = false;
    if (pl->flags & SWIT_PATH_BATCHED)
        TRY(batched_forward(pl, src, save_strips, stream));
    else
        TRY(plain_forward(pl, src, save_strips, stream));
    // [ns][nt][ngmax] -> [ns][ngmax][nt]
    hipLaunchKernelGGL(k_transpose, dim3(cdiv(d.ngmax, 32), cdiv(d.nt, 32), pl->ns), dim3(32, 8), 0,
                       stream, d.nt, d.ngmax, pl->seis_t, seis);
    LAUNCH_CHECK();
    pl->strips_valid = save_strips != 0;
    return SWIT_OK;
}

extern "C" int swit_adjoint(swit_plan *pl, const float *src, const float *adjsrc, double *grad,
                            double *illum_forw, double *illum_back, hipStream_t stream) {
    g_err[0] = 0;
    if (!pl || !src || !adjsrc || !grad || !illum_forw || !illum_back)
        return set_err(SWIT_EINVAL, "plan, src, adjsrc and the three outputs are required");
    if (!pl->strips_valid)
        return set_err(SWIT_ESTATE, "no saved strips: run swit_forward with save_strips=1 first");
    const Dims &d = pl->d;
    // [ns][ngmax][nt] -> [ns][nt][ngmax]
    hipLaunchKernelGGL(k_transpose, dim3(cdiv(d.nt, 32), cdiv(d.ngmax, 32), pl->ns), dim3(32, 8), 0,
                       stream, d.ngmax, d.nt, adjsrc, pl->sadj_t);
    LAUNCH_CHECK();
    if (pl->flags & SWIT_PATH_BATCHED) {
        TRY(batched_adjoint(pl, src, grad, illum_forw, illum_back, stream));
    } else {
        TRY(plain_adjoint(pl, src, stream));
        const size_t ni = (size_t)d.nx * d.nz;
        hipLaunchKernelGGL(k_reduce_shots, dim3(cdiv((int)ni, 256)), dim3(256), 0, stream, d, pl->ns,
                           pl->c_int, pl->dg, pl->dfw, pl->dbk, grad, illum_forw, illum_back);
        LAUNCH_CHECK();
    }
    return SWIT_OK;
}

// ---- batched path (host side) --------------------------------------------------------

static int batched_init(swit_plan *pl) {
    pl->bat.ready = 1;
    return SWIT_OK;
}

static void batched_release(swit_plan *pl) { pl->bat.ready = 0; }

static int batched_set_model(swit_plan *, hipStream_t) { return SWIT_OK; }

static int batched_forward(swit_plan *pl, const float *src, int save, hipStream_t st) {
    const Dims &d = pl->d;
    const size_t np = (size_t)d.nxp * d.ld;
    const int ns = pl->ns;
    HIP_TRY(hipMemsetAsync(pl->P, 0, np * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->U, 0, np * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->W, 0, np * ns * sizeof(float), st));
    dim3 blk(64, 4);
    dim3 grd(cdiv(d.nzp, 64), cdiv(d.nxp, 4), ns);
    int nrec = d.S > d.ngmax ? d.S : d.ngmax;
    for (int it = 2; it <= d.nt; ++it) {
        hipLaunchKernelGGL(k_fwd_p, grd, blk, 0, st, d, it, 0, pl->K, pl->damp, pl->geo, src,
                           pl->P, pl->U, pl->W);
        hipLaunchKernelGGL(k_fwd_v, grd, blk, 0, st, d, 0, pl->B, pl->damp, pl->P, pl->U, pl->W);
        hipLaunchKernelGGL(k_fwd_record, dim3(cdiv(nrec, 256), 1, ns), dim3(256), 0, st, d, it, 0,
                           pl->geo, pl->rec_fwd, pl->strip_map, pl->P, pl->U, pl->W, pl->seis_t,
                           save ? pl->strips : nullptr);
        LAUNCH_CHECK();
    }
    if (save) {
        dim3 gi(cdiv(d.nz, 64), cdiv(d.nx, 4), ns);
        hipLaunchKernelGGL(k_fwd_end, gi, blk, 0, st, d, 0, pl->P, pl->U, pl->W, pl->pe, pl->ue, pl->we);
        LAUNCH_CHECK();
    }
    return SWIT_OK;
}

static int batched_adjoint(swit_plan *pl, const float *src, double *grad, double *fw, double *bk,
                           hipStream_t st) {
    const Dims &d = pl->d;
    const size_t np = (size_t)d.nxp * d.ld, ni = (size_t)d.nx * d.nz;
    const int ns = pl->ns;
    HIP_TRY(hipMemcpyAsync(pl->Pr, pl->pe, ni * ns * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(pl->Ur, pl->ue, ni * ns * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(pl->Wr, pl->we, ni * ns * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemsetAsync(pl->Pb[0], 0, np * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->Pb[1], 0, np * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->Ub, 0, np * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->Wb, 0, np * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->dg, 0, ni * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->dfw, 0, ni * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(pl->dbk, 0, ni * ns * sizeof(float), st));
    dim3 blk(64, 4);
    dim3 gp(cdiv(d.nzp, 64), cdiv(d.nxp, 4), ns);
    dim3 gi(cdiv(d.nz, 64), cdiv(d.nx, 4), ns);
    int cur = 0;
    for (int it = d.nt; it >= 2; --it) {
        hipLaunchKernelGGL(k_reco_v, gi, blk, 0, st, d, it, 0, pl->B, pl->strips, pl->Pr, pl->Ur, pl->Wr);
        hipLaunchKernelGGL(k_reco_p, gi, blk, 0, st, d, it, 0, pl->K, pl->geo, src, pl->strips,
                           pl->Pr, pl->Ur, pl->Wr);
        hipLaunchKernelGGL(k_back_v, gp, blk, 0, st, d, 0, pl->K, pl->damp, pl->Pb[cur], pl->Ub, pl->Wb);
        hipLaunchKernelGGL(k_back_p, gp, blk, 0, st, d, 0, pl->B, pl->damp, pl->Pb[cur], pl->Pb[1 - cur],
                           pl->Ub, pl->Wb);
        hipLaunchKernelGGL(k_back_inject, dim3(cdiv(d.ngmax, 256), 1, ns), dim3(256), 0, st, d, it, 0,
                           pl->geo, pl->rec_inj, pl->sadj_t, pl->Pb[1 - cur]);
        hipLaunchKernelGGL(k_back_mirror_image, gp, blk, 0, st, d, 0, pl->Pb[cur], pl->Pb[1 - cur],
                           pl->Pr, pl->dg, pl->dfw, pl->dbk);
        LAUNCH_CHECK();
        cur = 1 - cur;
    }
    hipLaunchKernelGGL(k_reduce_shots, dim3(cdiv((int)ni, 256)), dim3(256), 0, st, d, ns,
                       pl->c_int, pl->dg, pl->dfw, pl->dbk, grad, fw, bk);
    LAUNCH_CHECK();
    return SWIT_OK;
}
