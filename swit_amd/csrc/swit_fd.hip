// C ABI and plan layer of the gfx950 2-D acoustic forward/adjoint solver
// (declared in include/swit_fd.h).  One translation unit: the kernel files are
// included below so the library is a single hipcc invocation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <utility>
#include <vector>

#include "swit_fd.h"
#include "fd_common.h"
#include "fd_kernels.h"
#include "fd_plain.hip"
#include "fd_batched.hip"
#include "fd_misfit.hip"
#include "fd_exact.hip"
#include "fd_precond.hip"
#include "fd_process.hip"

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
    std::vector<int> rec_cells_fwd, rec_cells_inj;   // host copies of rec_fwd / rec_inj
    std::vector<uint32_t> strip_map_h;
    std::vector<void *> allocs;
    int64_t bytes = 0;
    // geometry
    ShotGeom *geo = nullptr;
    int *rec_fwd = nullptr;       // [ns][ngmax] padded offset of the recording cell
    int *rec_inj = nullptr;       // [ns][ngmax] padded offset of the adjoint injection cell
    // cell groups (receivers sharing a cell, in receiver order) of the injection cells
    // (inj_*: plain adjoint) and of the recording cells (rec_*: exact adjoint's passes)
    int *inj_ngrp = nullptr, *rec_ngrp = nullptr;                 // [ns]
    InjGroup *inj_grp = nullptr, *rec_grp = nullptr;              // [ns][ngmax]
    RecRef *inj_lst = nullptr, *rec_lst = nullptr;                // [ns][ngmax]
    uint32_t *strip_map = nullptr;
    // model
    float *K = nullptr, *B = nullptr, *damp = nullptr, *prof = nullptr, *vmin = nullptr;
    float *c_int = nullptr, *rho_int = nullptr;   // interior vp and density of the model
    // wavefields
    float *P = nullptr, *U = nullptr, *W = nullptr;
    float *Pb[2] = {nullptr, nullptr}, *Ub = nullptr, *Wb = nullptr;
    float *Pr = nullptr, *Ur = nullptr, *Wr = nullptr;
    float *pe = nullptr, *ue = nullptr, *we = nullptr;
    float *strips = nullptr;
    float *seis_t = nullptr, *sadj_t = nullptr;
    float *dg = nullptr, *dfw = nullptr, *dbk = nullptr;
    float *stats_h = nullptr;     // pinned host copy of the model statistics of set_model
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
        if (stats_h) (void)hipHostFree(stats_h);
        stats_h = nullptr;
    }
};

#define TRY(expr)                 \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != SWIT_OK) {     \
            return rc_;           \
        }                         \
    } while (0)

static int build_injection(swit_plan *pl, const std::vector<int> &cells, const TileMap &tm, int margin, int cap,
                           std::vector<int> &gst, std::vector<InjGroup> &groups, std::vector<RecRef> &list);

// Cell groups over the whole padded grid (one tile): per shot, the group count, the
// groups and the receiver list they index, uploaded to the plan's device arrays.
static int upload_cell_groups(swit_plan *pl, const std::vector<int> &cells, int **ngrp, InjGroup **grp,
                              RecRef **lst) {
    const Dims &d = pl->d;
    const int ns = pl->ns;
    TileMap all{};
    all.ztop = 1;
    all.th = d.nzp;
    all.tx = d.nxp;
    all.nzt = all.nxt = all.ntiles = all.cg = 1;
    std::vector<int> gst;
    std::vector<InjGroup> groups;
    std::vector<RecRef> list;
    TRY(build_injection(pl, cells, all, 0, d.ngmax, gst, groups, list));
    std::vector<int> cnt(ns);
    for (int s = 0; s < ns; ++s) cnt[s] = gst[(size_t)s * 2 + 1];
    TRY(pl->alloc(ngrp, ns));
    TRY(pl->alloc(grp, groups.size()));
    TRY(pl->alloc(lst, list.size()));
    HIP_TRY(hipMemcpy(*ngrp, cnt.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(*grp, groups.data(), sizeof(InjGroup) * groups.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(*lst, list.data(), sizeof(RecRef) * list.size(), hipMemcpyHostToDevice));
    return SWIT_OK;
}

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
    const bool s16 = (flags & (SWIT_STRIPS_F16 | SWIT_STRIPS_BF16)) != 0;
    const bool batched = (flags & SWIT_PATH_BATCHED) != 0;
    if (s16 && !batched) return set_err(SWIT_EINVAL, "16-bit strips need the batched path");
    if ((flags & (SWIT_PLAN_EXACT | SWIT_PLAN_EXACT_CKPT)) && !batched)
        return set_err(SWIT_EINVAL, "the exact adjoint needs the batched path");
    const bool nostrips = (flags & SWIT_PLAN_NO_STRIPS) != 0;
    if (nostrips && s16) return set_err(SWIT_EINVAL, "a plan without strips takes no strip precision");

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
    pl->rec_cells_fwd = rf;
    pl->rec_cells_inj = ri;
    pl->strip_map_h = sm;

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
    HIP_TRY(hipHostMalloc((void **)&pl->stats_h, 4 * sizeof(float), hipHostMallocDefault));
    TRY(pl->alloc(&pl->c_int, ni));
    TRY(pl->alloc(&pl->rho_int, ni));
    if (!batched) {
        // per-shot wavefields of the plain path
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
        TRY(pl->alloc(&pl->dg, ni * ns));
        TRY(pl->alloc(&pl->dfw, ni * ns));
        TRY(pl->alloc(&pl->dbk, ni * ns));
    }
    if (!s16 && !nostrips) TRY(pl->alloc(&pl->strips, (size_t)ns * d.nt * d.S));
    TRY(pl->alloc(&pl->seis_t, (size_t)ns * d.nt * ngmax));
    TRY(pl->alloc(&pl->sadj_t, (size_t)ns * d.nt * ngmax));
    // Receiver columns past a shot's own count are never written by a kernel; they
    // stay zero for the plan's life, so padded traces read as dead traces.
    HIP_TRY(hipMemset(pl->seis_t, 0, sizeof(float) * ns * d.nt * ngmax));
    HIP_TRY(hipMemset(pl->sadj_t, 0, sizeof(float) * ns * d.nt * ngmax));
    HIP_TRY(hipStreamSynchronize(nullptr));
    HIP_TRY(hipMemcpy(pl->geo,pl->geo_h.data(), sizeof(ShotGeom) * ns, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pl->rec_fwd, rf.data(), sizeof(int) * rf.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pl->rec_inj, ri.data(), sizeof(int) * ri.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pl->strip_map, sm.data(), sizeof(uint32_t) * sm.size(), hipMemcpyHostToDevice));
    if (!batched) TRY(upload_cell_groups(pl, ri, &pl->inj_ngrp, &pl->inj_grp, &pl->inj_lst));
    if (flags & (SWIT_PLAN_EXACT | SWIT_PLAN_EXACT_CKPT))
        TRY(upload_cell_groups(pl, rf, &pl->rec_ngrp, &pl->rec_grp, &pl->rec_lst));
    if (batched) TRY(batched_init(pl));
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

extern "C" int swit_plan_strips(swit_plan *pl, float *out, hipStream_t stream) {
    g_err[0] = 0;
    if (!pl || !out) return set_err(SWIT_EINVAL, "plan and out are required");
    if (!pl->strips_valid) return set_err(SWIT_ESTATE, "no saved strips");
    const Dims &d = pl->d;
    const size_t n = (size_t)pl->ns * d.nt * d.S;
    const BatchedState &bs = pl->bat;
    if (bs.strip_mode == 0) {
        HIP_TRY(hipMemcpyAsync(out, pl->strips, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
        return SWIT_OK;
    }
    int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    if (bs.strip_mode == 1)
        hipLaunchKernelGGL(k_strip_decode<1>, dim3(blocks), dim3(256), 0, stream, n, d.S, bs.strips16, bs.smax, out);
    else
        hipLaunchKernelGGL(k_strip_decode<2>, dim3(blocks), dim3(256), 0, stream, n, d.S, bs.strips16, bs.smax, out);
    LAUNCH_CHECK();
    return SWIT_OK;
}

extern "C" int swit_plan_set_model(swit_plan *pl, const float *vp, const float *rho,
                                   hipStream_t stream) {
    g_err[0] = 0;
    if (!pl || !vp || !rho) return set_err(SWIT_EINVAL, "plan, vp and rho are required");
    const Dims &d = pl->d;
    dim3 blk(64, 4);
    dim3 grd(cdiv(d.nzp, 64), cdiv(d.nxp, 4), 1);
    // the model's statistics first; a refused model leaves the plan without a model
    pl->model_set = false;
    pl->strips_valid = false;
    pl->bat.exact_valid = false;
    hipLaunchKernelGGL(k_setup_pml, dim3(1), dim3(1024), 0, stream, d, vp, rho, pl->prof, pl->vmin, pl->vmin + 1);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(pl->stats_h, pl->vmin + 1, 2 * sizeof(float), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (pl->stats_h[1] != 0.0f)
        return set_err(SWIT_EINVAL, "the model holds %.0f vp or rho values that are not finite and positive",
                       (double)pl->stats_h[1]);
    // toolbox/base.py:57-59: 4th-order stability needs dt < sqrt(3/8) dx / max(vp)
    const double dt_req = std::sqrt(3.0 / 8.0) * (double)d.dx / (double)pl->stats_h[0];
    if (dt_req <= (double)d.dt)
        return set_err(SWIT_EINVAL, "modeling stability: dt = %.4f ms > dt_required = %.4f ms (max vp %.1f m/s)",
                       (double)d.dt * 1000.0, dt_req * 1000.0, (double)pl->stats_h[0]);
    hipLaunchKernelGGL(k_setup_coef, grd, blk, 0, stream, d, vp, rho, pl->K, pl->B);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_setup_damp, grd, blk, 0, stream, d, pl->prof, pl->damp);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(pl->c_int, vp, sizeof(float) * d.nx * d.nz, hipMemcpyDeviceToDevice, stream));
    HIP_TRY(hipMemcpyAsync(pl->rho_int, rho, sizeof(float) * d.nx * d.nz, hipMemcpyDeviceToDevice, stream));
    pl->model_set = true;
    pl->strips_valid = false;
    pl->bat.exact_valid = false;
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
                               pl->inj_ngrp, pl->inj_grp, pl->inj_lst, pl->sadj_t, pl->Pb[1 - cur]);
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
    if (save_strips < 0 || save_strips > 2) return set_err(SWIT_EINVAL, "save_strips must be 0, 1 or 2");
    if (save_strips == 1 && (pl->flags & SWIT_PLAN_NO_STRIPS))
        return set_err(SWIT_EINVAL, "the plan was created without strips (SWIT_PLAN_NO_STRIPS)");
    const int save = save_strips == 1;
    if (save_strips != 2) pl->strips_valid = false;
    if (pl->flags & SWIT_PATH_BATCHED)
        TRY(batched_forward(pl, src, save, stream));
    else
        TRY(plain_forward(pl, src, save, stream));
    // [ns][nt][ngmax] -> [ns][ngmax][nt]
    hipLaunchKernelGGL(k_transpose, dim3(cdiv(d.ngmax, 32), cdiv(d.nt, 32), pl->ns), dim3(32, 8), 0,
                       stream, d.nt, d.ngmax, pl->seis_t, seis);
    LAUNCH_CHECK();
    if (save_strips != 2) pl->strips_valid = save_strips == 1;
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

static TileMap make_tilemap(const Dims &d, int ztop, int th) {
    TileMap tm;
    tm.ztop = ztop;
    tm.th = th;
    tm.tx = fk::kFTX;
    tm.nzt = cdiv(d.nzp - ztop + 1, th);
    tm.nxt = cdiv(d.nxp, fk::kFTX);
    tm.ntiles = tm.nzt * tm.nxt;
    tm.cg = std::min(kColGroup, tm.nxt);
    return tm;
}

static int tile_of(const TileMap &tm, int iz, int ix) {
    int zt = (iz - tm.ztop) / tm.th, xt = (ix - 1) / tm.tx;
    return xt * tm.nzt + zt;
}

// Backward tiling: rows 2..nzp-1 are the only rows a backward step changes (adjoint
// u/w rows 2..nzp-1, p rows 3..nzp-2; the interior window lies inside), and rows
// fs-2..fs+1 must share a tile (the adjoint p stencil there reads the mirrored w(fs-1);
// the p mirror reads row fs+1 after injection), so no tile starts at fs-1..fs+1.  The
// fused exact-adjoint step uses tiles of the same rows and needs, in every tile's window
// (tile +- 3 rows), the rows its transposed mirrors read: no tile starts at fs+2 or fs+3
// either.  Pick the offset with the fewest tiles.  tx: tile width (the fused exact
// step takes wider tiles than the backward step kernel).
static TileMap backward_tilemap(const Dims &d, int tx = kTX) {
    const int th = bk::kTZ, fs = d.fs;
    TileMap best{};
    best.ntiles = 1 << 30;
    for (int shift = 0; shift < th; ++shift) {
        int ztop = 2 - shift;
        bool ok = true;
        for (int b = ztop + th; b <= d.nzp; b += th)       // tile starts after the first
            if (b >= fs - 1 && b <= fs + 3) ok = false;
        if (!ok) continue;
        TileMap tm;
        tm.ztop = ztop;
        tm.th = th;
        tm.tx = tx;
        tm.nzt = cdiv(d.nzp - 1 - ztop + 1, th);           // cover rows up to nzp-1
        tm.nxt = cdiv(d.nxp, tx);
        tm.ntiles = tm.nzt * tm.nxt;
        tm.cg = std::min(kColGroup, tm.nxt);
        if (tm.ntiles < best.ntiles) best = tm;
    }
    return best;
}

// Two-step backward tiling: 52-row tiles whose windows reach 6 rows past each side.
// Within a window the mirror row fs-1 must not sit where a step needs it while its
// source row (fs for w, fs+1 for p) lies past that step's updated rows, which is the
// case for fs-1 in [z1-1, z1+4] of a tile ending at row z1.  Rows 2..nzp-1 are
// covered; pick the offset with the fewest tiles.
static TileMap backward2_tilemap(const Dims &d) {
    const int th = bk2::kTZ, fs = d.fs;
    TileMap best{};
    best.ntiles = 1 << 30;
    for (int shift = 0; shift < th; ++shift) {
        TileMap tm;
        tm.ztop = 2 - shift;
        tm.th = th;
        tm.tx = kTX;
        tm.nzt = cdiv(d.nzp - 1 - tm.ztop + 1, th);
        tm.nxt = cdiv(d.nxp, kTX);
        tm.ntiles = tm.nzt * tm.nxt;
        tm.cg = std::min(kColGroup, tm.nxt);
        bool ok = true;
        for (int zt = 0; zt < tm.nzt; ++zt) {
            const int z1 = std::min(tm.ztop + (zt + 1) * th - 1, d.nzp);
            if (fs - 1 >= z1 - 1 && fs - 1 <= z1 + 4) ok = false;
        }
        if (ok && tm.ntiles < best.ntiles) best = tm;
    }
    return best;
}

// Injection groups per tile: for every shot, the receivers whose cell (`cells`: padded
// offsets [ns][ngmax]) lies in the tile grown by `margin` cells per side, listed tile by tile, sorted by cell within
// a tile (stably, so receivers sharing a cell stay in receiver order) and grouped by
// cell, so one thread adds every receiver of a cell in order.
// gst: [ns][ntiles+1] group ranges per tile; groups, list: [ns][cap].
static int build_injection(swit_plan *pl, const std::vector<int> &cells, const TileMap &tm, int margin, int cap,
                           std::vector<int> &gst, std::vector<InjGroup> &groups, std::vector<RecRef> &list) {
    const Dims &d = pl->d;
    const int ns = pl->ns, ngmax = d.ngmax;
    gst.assign((size_t)ns * (tm.ntiles + 1), 0);
    groups.assign((size_t)ns * cap, InjGroup{0, 0});
    list.assign((size_t)ns * cap, RecRef{0, 0, 0, 0});
    for (int s = 0; s < ns; ++s) {
        std::vector<RecRef> rec;
        for (int r = 0; r < pl->geo_h[s].ng; ++r) {
            const int ci = cells[(size_t)s * ngmax + r];
            rec.push_back(RecRef{r, ci % d.ld + 1, ci / d.ld + 1, 0});
        }
        RecRef *el = &list[(size_t)s * cap];
        InjGroup *gr = &groups[(size_t)s * cap];
        int *gs = &gst[(size_t)s * (tm.ntiles + 1)];
        int ne = 0, ng = 0;
        for (int t = 0; t < tm.ntiles; ++t) {
            const TileBox b = tile_box(d, tm, t);
            const int e0 = ne;
            for (const RecRef &r : rec) {
                if (r.iz < b.z0 - margin || r.iz > b.z1 + margin || r.ix < b.x0 - margin || r.ix > b.x1 + margin)
                    continue;
                if (ne == cap) return set_err(SWIT_EINVAL, "shot %d: more than %d injection entries", s, cap);
                el[ne++] = r;
            }
            std::stable_sort(el + e0, el + ne, [](const RecRef &x, const RecRef &y) {
                return x.ix != y.ix ? x.ix < y.ix : x.iz < y.iz;
            });
            gs[t] = ng;
            for (int e = e0; e < ne;) {
                int f = e + 1;
                while (f < ne && el[f].ix == el[e].ix && el[f].iz == el[e].iz) ++f;
                gr[ng++] = InjGroup{e, f};
                e = f;
            }
        }
        gs[tm.ntiles] = ng;
    }
    return SWIT_OK;
}

static int fill(float *p, size_t n, hipStream_t st) {
    int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, st, p, n, 0.0f);
    LAUNCH_CHECK();
    return SWIT_OK;
}

static int copyk(const float *a, float *b, size_t n, hipStream_t st) {
    int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, a, b, n);
    LAUNCH_CHECK();
    return SWIT_OK;
}

static int batched_init(swit_plan *pl) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    const int ns = pl->ns, ngmax = d.ngmax, fs = d.fs;
    bs.fwd = make_tilemap(d, fs - 1, fk::kTZ);
    for (int wi = 0; wi < 2; ++wi) bs.bwdw[wi] = backward_tilemap(d, bk::kBTXs[wi]);
    bs.bwd_wi = bk::kBTXs[0] == bk::kBTX ? 0 : 1;
    bs.bwd = bs.bwdw[bs.bwd_wi];
    bs.bwd2 = backward2_tilemap(d);
    // receivers: recording cells by forward tile
    std::vector<int> rs((size_t)ns * (bs.fwd.ntiles + 1));
    std::vector<RecRef> rl((size_t)ns * ngmax);
    for (int s = 0; s < ns; ++s) {
        std::vector<std::pair<int, RecRef>> a;
        for (int r = 0; r < pl->geo_h[s].ng; ++r) {
            int cf = pl->rec_cells_fwd[(size_t)s * ngmax + r];
            RecRef rf{r, cf % d.ld + 1, cf / d.ld + 1, 0};
            if (rf.iz < fs - 1) return set_err(SWIT_EINVAL, "shot %d receiver %d above the free surface", s, r);
            a.push_back({tile_of(bs.fwd, rf.iz, rf.ix), rf});
        }
        auto by_tile = [](const std::pair<int, RecRef> &x, const std::pair<int, RecRef> &y) { return x.first < y.first; };
        std::stable_sort(a.begin(), a.end(), by_tile);
        int *sa = &rs[(size_t)s * (bs.fwd.ntiles + 1)];
        for (int t = 0, k = 0; t <= bs.fwd.ntiles; ++t) {
            while (k < (int)a.size() && a[k].first < t) ++k;
            sa[t] = k;
        }
        for (size_t k = 0; k < a.size(); ++k) rl[(size_t)s * ngmax + k] = a[k].second;
    }
    // injection cells: tiles of the one-step kernel, tiles of the two-step kernel, and
    // those grown by 3 cells per side (a receiver lies in at most 2 x 2 of them)
    std::vector<int> is2, isx;
    std::vector<RecRef> il2, ilx;
    std::vector<InjGroup> ig2, igx;
    bs.injx_cap = 4 * ngmax;
    for (int wi = 0; wi < 2; ++wi) {
        std::vector<int> is;
        std::vector<RecRef> il;
        std::vector<InjGroup> ig;
        TRY(build_injection(pl, pl->rec_cells_inj, bs.bwdw[wi], 0, ngmax, is, ig, il));
        TRY(pl->alloc(&bs.inj_start_w[wi], is.size()));
        TRY(pl->alloc(&bs.inj_list_w[wi], il.size()));
        TRY(pl->alloc(&bs.inj_groups_w[wi], ig.size()));
        HIP_TRY(hipMemcpy(bs.inj_start_w[wi], is.data(), sizeof(int) * is.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(bs.inj_list_w[wi], il.data(), sizeof(RecRef) * il.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(bs.inj_groups_w[wi], ig.data(), sizeof(InjGroup) * ig.size(), hipMemcpyHostToDevice));
    }
    bs.inj_start = bs.inj_start_w[bs.bwd_wi];
    bs.inj_list = bs.inj_list_w[bs.bwd_wi];
    bs.inj_groups = bs.inj_groups_w[bs.bwd_wi];
    TRY(build_injection(pl, pl->rec_cells_inj, bs.bwd2, 0, ngmax, is2, ig2, il2));
    TRY(build_injection(pl, pl->rec_cells_inj, bs.bwd2, 3, bs.injx_cap, isx, igx, ilx));
    // strip slots by forward tile
    std::vector<std::pair<int, StripRef>> sv;
    for (int k = 0; k < d.S; ++k) {
        uint32_t m = pl->strip_map_h[k];
        int f = (int)(m >> kStripFieldShift), o = (int)(m & ((1u << kStripFieldShift) - 1u));
        StripRef r{k, f, o % d.ld + 1, o / d.ld + 1};
        sv.push_back({tile_of(bs.fwd, r.iz, r.ix), r});
    }
    std::stable_sort(sv.begin(), sv.end(), [](const std::pair<int, StripRef> &x, const std::pair<int, StripRef> &y) {
        return x.first < y.first;
    });
    std::vector<int> ss(bs.fwd.ntiles + 1);
    std::vector<StripRef> sl(d.S);
    for (int t = 0, k = 0; t <= bs.fwd.ntiles; ++t) {
        while (k < (int)sv.size() && sv[k].first < t) ++k;
        ss[t] = k;
    }
    for (int k = 0; k < d.S; ++k) sl[k] = sv[k].second;

    const size_t np = (size_t)d.nxp * d.ld, ni = (size_t)d.nx * d.nz;
    TRY(pl->alloc(&bs.rec_start, rs.size()));
    TRY(pl->alloc(&bs.rec_list, rl.size()));
    TRY(pl->alloc(&bs.inj2_start, is2.size()));
    TRY(pl->alloc(&bs.inj2_list, il2.size()));
    TRY(pl->alloc(&bs.inj2_groups, ig2.size()));
    TRY(pl->alloc(&bs.injx_start, isx.size()));
    TRY(pl->alloc(&bs.injx_list, ilx.size()));
    TRY(pl->alloc(&bs.injx_groups, igx.size()));
    TRY(pl->alloc(&bs.strip_start, ss.size()));
    TRY(pl->alloc(&bs.strip_list, sl.size()));
    TRY(pl->alloc(&bs.KB, np));
    TRY(pl->alloc(&bs.CP, np));
    TRY(pl->alloc(&bs.CV, np));
    for (int k = 0; k < 2; ++k) {
        TRY(pl->alloc(&bs.F[k], 3 * np * ns));
        TRY(pl->alloc(&bs.Bk[k], 3 * np * ns));
        TRY(pl->alloc(&bs.R[k], 3 * ni * ns));
    }
    TRY(pl->alloc(&bs.end, 3 * ni * ns));
    TRY(pl->alloc(&bs.src, (size_t)ns * d.nt));
    TRY(pl->alloc(&bs.acc, (size_t)ns * 3 * ni));
    TRY(pl->alloc(&bs.out, 3 * ni));
    if (pl->flags & (SWIT_PLAN_EXACT | SWIT_PLAN_EXACT_CKPT)) {
        if (pl->flags & SWIT_PLAN_EXACT_CKPT) {
            // states every ~sqrt(nt) steps, one segment of divergence history
            bs.ck = std::max(1, (int)std::lround(std::sqrt((double)d.nt)));
            bs.nck = cdiv(d.nt - 1, bs.ck);
            TRY(pl->alloc(&bs.ckpt, (size_t)bs.nck * ns * 3 * np));
            TRY(pl->alloc(&bs.dhist, (size_t)ns * bs.ck * np));
        } else {
            TRY(pl->alloc(&bs.dhist, (size_t)ns * d.nt * np));
        }
        TRY(pl->alloc(&bs.lam, (size_t)ns * 3 * np));
        TRY(pl->alloc(&bs.lam2, (size_t)ns * 3 * np));
        TRY(pl->alloc(&bs.qbuf, (size_t)ns * 2 * np));
        TRY(pl->alloc(&bs.gk, (size_t)ns * np));
        TRY(pl->alloc(&bs.gsum, np));
        std::vector<int> es;
        std::vector<RecRef> el;
        std::vector<InjGroup> eg;
        bs.exinj_cap = 4 * ngmax;
        bs.extm = backward_tilemap(d, ex::kExTX);
        TRY(build_injection(pl, pl->rec_cells_fwd, bs.extm, 3, bs.exinj_cap, es, eg, el));
        TRY(pl->alloc(&bs.exinj_start, es.size()));
        TRY(pl->alloc(&bs.exinj_groups, eg.size()));
        TRY(pl->alloc(&bs.exinj_list, el.size()));
        HIP_TRY(hipMemcpy(bs.exinj_start, es.data(), sizeof(int) * es.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(bs.exinj_groups, eg.data(), sizeof(InjGroup) * eg.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(bs.exinj_list, el.data(), sizeof(RecRef) * el.size(), hipMemcpyHostToDevice));
        TRY(pl->alloc(&bs.gseis, (size_t)ns * ngmax * d.nt));
        TRY(pl->alloc(&bs.src_ex, (size_t)ns * d.nt));
    }
    bs.strip_mode = (pl->flags & SWIT_STRIPS_F16) ? 1 : (pl->flags & SWIT_STRIPS_BF16) ? 2 : 0;
    if (bs.strip_mode && !(pl->flags & SWIT_PLAN_NO_STRIPS)) {
        TRY(pl->alloc(&bs.strips16, (size_t)ns * d.nt * d.S));
        TRY(pl->alloc(&bs.smax, (size_t)ns * d.nt));
        TRY(pl->alloc(&bs.ring, (size_t)kStripRing * ns * d.S));
    }
    HIP_TRY(hipMemcpy(bs.rec_start, rs.data(), sizeof(int) * rs.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bs.rec_list, rl.data(), sizeof(RecRef) * rl.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bs.inj2_start, is2.data(), sizeof(int) * is2.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bs.inj2_list, il2.data(), sizeof(RecRef) * il2.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bs.inj2_groups, ig2.data(), sizeof(InjGroup) * ig2.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bs.injx_start, isx.data(), sizeof(int) * isx.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bs.injx_list, ilx.data(), sizeof(RecRef) * ilx.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bs.injx_groups, igx.data(), sizeof(InjGroup) * igx.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bs.strip_start, ss.data(), sizeof(int) * ss.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bs.strip_list, sl.data(), sizeof(StripRef) * sl.size(), hipMemcpyHostToDevice));
    // every state buffer starts at zero; cells no step writes stay zero for good
    for (int k = 0; k < 2; ++k) {
        HIP_TRY(hipMemset(bs.F[k], 0, 3 * np * ns * sizeof(float)));
        HIP_TRY(hipMemset(bs.Bk[k], 0, 3 * np * ns * sizeof(float)));
    }
    bs.Gf = std::min(4, ns);
    bs.ngroups_f = cdiv(ns, bs.Gf);
    const char *eg = getenv("SWIT_NO_GRAPH");
    bs.use_graph = !(eg && atoi(eg) > 0);
    HIP_TRY(hipStreamCreateWithFlags(&bs.cap, hipStreamNonBlocking));
    bs.ready = 1;
    return SWIT_OK;
}

static void batched_release(swit_plan *pl) {
    BatchedState &bs = pl->bat;
    for (auto &g : bs.gfwd)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    if (bs.gbwd) { (void)hipGraphExecDestroy(bs.gbwd); bs.gbwd = nullptr; }
    for (auto &row : bs.gfwd_h)
        for (auto &g : row)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    for (auto &g : bs.gbwd_h)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    for (auto &row : bs.gex_h)
        for (auto &g : row)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    if (bs.gex_f) { (void)hipGraphExecDestroy(bs.gex_f); bs.gex_f = nullptr; }
    if (bs.gex_b) { (void)hipGraphExecDestroy(bs.gex_b); bs.gex_b = nullptr; }
    for (auto &t : bs.pending) { (void)hipEventDestroy(t.e0); (void)hipEventDestroy(t.e1); }
    bs.pending.clear();
    for (hipEvent_t e : bs.ev_pool) (void)hipEventDestroy(e);
    bs.ev_pool.clear();
    if (bs.cap) { (void)hipStreamDestroy(bs.cap); bs.cap = nullptr; }
    if (bs.side) { (void)hipStreamDestroy(bs.side); bs.side = nullptr; }
    if (bs.fork) { (void)hipEventDestroy(bs.fork); bs.fork = nullptr; }
    if (bs.join) { (void)hipEventDestroy(bs.join); bs.join = nullptr; }
    bs.ready = 0;
}

static int batched_set_model(swit_plan *pl, hipStream_t st) {
    const Dims &d = pl->d;
    dim3 blk(64, 4), grd(cdiv(d.nzp, 64), cdiv(d.nxp, 4), 1);
    hipLaunchKernelGGL(k_setup_cvec, grd, blk, 0, st, d, pl->K, pl->B, pl->prof, pl->bat.CP, pl->bat.CV,
                       pl->bat.KB);
    LAUNCH_CHECK();
    return SWIT_OK;
}

static int pool_event(BatchedState &bs, hipEvent_t *e) {
    if (!bs.ev_pool.empty()) {
        *e = bs.ev_pool.back();
        bs.ev_pool.pop_back();
        return SWIT_OK;
    }
    HIP_TRY(hipEventCreate(e));
    return SWIT_OK;
}

// Read every recorded loop timing (waits for its end event) into prof_ms / prof_n.
static int drain_timings(BatchedState &bs) {
    for (auto &t : bs.pending) {
        float ms = 0;
        HIP_TRY(hipEventSynchronize(t.e1));
        HIP_TRY(hipEventElapsedTime(&ms, t.e0, t.e1));
        bs.prof_ms[t.which] += ms;
        bs.prof_n[t.which] += t.launches;
        bs.ev_pool.push_back(t.e0);
        bs.ev_pool.push_back(t.e1);
    }
    bs.pending.clear();
    return SWIT_OK;
}

// Forward time loop: prologue (state clear, zero first rows), the nt-1 step kernels,
// epilogue (final state for the reconstruction).
static int forward_prologue(swit_plan *pl, int save, hipStream_t st) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    const int ns = pl->ns;
    const size_t np = (size_t)d.nxp * d.ld;
    TRY(fill(bs.F[0], 3 * np * ns, st));
    // row 0 of the seismograms and strips is zero (csg(1,:), boundary(:,1))
    hipLaunchKernelGGL(k_zero_rows, dim3(ns), dim3(256), 0, st, pl->seis_t, (size_t)d.nt * d.ngmax, d.ngmax);
    LAUNCH_CHECK();
    if (save && bs.strip_mode == 0) {
        hipLaunchKernelGGL(k_zero_rows, dim3(ns), dim3(256), 0, st, pl->strips, (size_t)d.nt * d.S, d.S);
        LAUNCH_CHECK();
    } else if (save) {
        hipLaunchKernelGGL(k_zero_rows16, dim3(ns), dim3(256), 0, st, bs.strips16, (size_t)d.nt * d.S, d.S);
        LAUNCH_CHECK();
        TRY(fill(reinterpret_cast<float *>(bs.smax), (size_t)ns * d.nt, st));   // +0.0f bits are 0
    }
    return SWIT_OK;
}

// Forward step it.  With strip saving the step's FP32 strip row goes to the plan's
// strips (FP32 plans) or to the ring that k_strip_encode empties (16-bit plans).
// A launch's shots: s0 .. s0 + n - 1 (n < 0: all of the plan's shots).
struct ShotRange {
    int s0, n;
};
static ShotRange shots_of(const swit_plan *pl, ShotRange r) { return r.n < 0 ? ShotRange{0, pl->ns} : r; }

// ---- work split over the XCDs --------------------------------------------------------
// A one-step backward launch's logical items (work_item order) go to the 8 XCDs in
// contiguous ranges: equal item counts, or (balanced) equal estimated cost, the bytes an
// item moves.  Items differ: edge tiles are narrow, damped tiles also read the damping
// factors, and tiles without interior cells skip the reconstruction and imaging.

static int span(int a0, int a1, int lo, int hi) { return std::max(0, std::min(a1, hi) - std::max(a0, lo) + 1); }

// One item (one shot, tile t): adjoint window in, {K, B} of the velocity cells, the
// damping factors of a damped tile, the adjoint tile out; the interior parts of the
// reconstruction window, its tile out and the accumulators' read-modify-write.
static double bwd_item_cost(const Dims &d, const TileMap &tm, int t) {
    const TileBox b = tile_box(d, tm, t);
    const int npml = d.npml, zo = b.z0 - kHalo, xo = b.x0 - kHalo;
    const double win = (double)span(zo, zo + bk::kZL - 1, 1, d.nzp) * span(xo, b.x1 + kHalo, 1, d.nxp);
    const double vel = (double)span(b.z0 - 2, b.z1 + 1, 1, d.nzp) * span(b.x0 - 2, b.x1 + 1, 1, d.nxp);
    const double tile = (double)(b.z1 - b.z0 + 1) * (b.x1 - b.x0 + 1);
    const double win_i = (double)span(zo, zo + bk::kZL - 1, npml + 1, npml + d.nz) *
                         span(xo, b.x1 + kHalo, npml + 1, npml + d.nx);
    const double tile_i = (double)span(b.z0, b.z1, npml + 1, npml + d.nz) * span(b.x0, b.x1, npml + 1, npml + d.nx);
    const bool damped = !(b.z0 - 2 >= npml + 1 && b.z0 + bk::kTZ + 1 <= npml + d.nz && b.x0 - 2 >= npml + 1 &&
                          b.x1 + 2 <= npml + d.nx);
    return 12 * win + 8 * vel + (damped ? 24 * vel : 0.0) + 12 * tile + 12 * win_i + 36 * tile_i;
}

// Split of a launch of `nsh` shots, cached per {shots, column group, tile width,
// balanced}.  *grid: blocks to launch (XCD k's j-th block takes item xs[k] + j).
static XcdSplit bwd_xcd_split(BatchedState &bs, const Dims &d, const TileMap &tm, int nsh, int *grid) {
    const bool balanced = bs.bwd_balance != 0;
    const std::array<int, 4> key = {nsh, tm.cg, tm.tx, balanced ? 1 : 0};
    XcdSplit xp{};
    bool found = false;
    for (const auto &e : bs.splits)
        if (e.first == key) { xp = e.second; found = true; break; }
    if (!found) {
        const int n = nsh * tm.ntiles;
        xp.nsh = nsh;
        if (!balanced) {
            const int q = n / kNumXcd, r = n % kNumXcd;
            for (int k = 0; k <= kNumXcd; ++k) xp.xs[k] = k * q + std::min(k, r);
        } else {
            std::vector<double> cost(tm.ntiles);
            for (int t = 0; t < tm.ntiles; ++t) cost[t] = bwd_item_cost(d, tm, t);
            std::vector<double> cum(n + 1, 0.0);
            for (int w = 0; w < n; ++w) {
                // the tile of logical item w, as work_item maps it
                const int per = tm.cg * tm.nzt, gq = w / (per * nsh), rem = w - gq * per * nsh;
                const int cols = std::min(tm.cg, tm.nxt - gq * tm.cg), tiles = cols * tm.nzt;
                cum[w + 1] = cum[w] + cost[gq * per + rem % tiles];
            }
            xp.xs[0] = 0;
            for (int k = 1; k < kNumXcd; ++k) {
                const double target = cum[n] * k / kNumXcd;
                int w = xp.xs[k - 1];
                while (w < n && cum[w + 1] <= target) ++w;
                if (w < n && target - cum[w] > cum[w + 1] - target) ++w;     // the nearer boundary
                xp.xs[k] = w;
            }
            xp.xs[kNumXcd] = n;
        }
        bs.splits.push_back({key, xp});
    }
    int most = 0;
    for (int k = 0; k < kNumXcd; ++k) most = std::max(most, xp.xs[k + 1] - xp.xs[k]);
    *grid = kNumXcd * most;
    return xp;
}

static void launch_fwd(swit_plan *pl, int it, const float *in, float *out, int save, hipStream_t st,
                       DSave dsave = DSave{nullptr, 0, 0}, const float *src = nullptr,
                       ShotRange range = ShotRange{0, -1}) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    const ShotRange r = shots_of(pl, range);
    float *strips = !save ? nullptr : bs.strip_mode == 0 ? pl->strips : bs.ring;
    const int s_rows = bs.strip_mode == 0 ? d.nt : kStripRing;
    const int s_row = bs.strip_mode == 0 ? it - 1 : (it - 1) % kStripRing;
    hipLaunchKernelGGL(k_fwd_fused, dim3(cdiv(r.n, bs.Gf) * bs.fwd.ntiles), dim3(kThreads), 0, st, d, it, bs.fwd,
                       r.s0, r.n, bs.Gf, bs.CP, bs.CV, pl->geo, src ? src : bs.src, in, out, bs.rec_start,
                       bs.rec_list, pl->seis_t, bs.strip_start, bs.strip_list, strips, s_rows, s_row, dsave);
}

// 16-bit plans: encode strip rows first..last (at most kStripRing, all still in the ring).
static void launch_encode(swit_plan *pl, int first, int last, hipStream_t st, ShotRange range = ShotRange{0, -1}) {
    BatchedState &bs = pl->bat;
    const ShotRange r = shots_of(pl, range);
    const dim3 grd(last - first + 1, r.n);
    if (bs.strip_mode == 1)
        hipLaunchKernelGGL(k_strip_encode<1>, grd, dim3(kEncThreads), 0, st, pl->d, first, kStripRing, r.s0,
                           bs.ring, bs.smax, bs.strips16);
    else
        hipLaunchKernelGGL(k_strip_encode<2>, grd, dim3(kEncThreads), 0, st, pl->d, first, kStripRing, r.s0,
                           bs.ring, bs.smax, bs.strips16);
}

template <int SM>
static void launch_bwd(swit_plan *pl, int it, const float *ri, float *ro, const float *bi, float *bo,
                       hipStream_t st, ShotRange r) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    const void *strips = SM == 0 ? static_cast<const void *>(pl->strips) : static_cast<const void *>(bs.strips16);
    auto kern = bs.bwd.tx == bk::kBTXs[0] ? k_bwd_fused<SM, bk::kBTXs[0]> : k_bwd_fused<SM, bk::kBTXs[1]>;
    int grid = 0;
    const XcdSplit xp = bwd_xcd_split(bs, d, bs.bwd, r.n, &grid);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), 0, st, d, it, bs.bwd, r.s0, xp, bs.KB, bs.CP, bs.CV,
                       pl->geo, bs.src, strips, bs.smax, ri, ro, bi, bo, bs.inj_start, bs.inj_groups, bs.inj_list,
                       pl->sadj_t, bs.acc);
}

// Make width index wi (of bk::kBTXs) the one-step backward kernel's tiling; the column
// group is kept where the new tiling allows it.
static void select_bwd_width(BatchedState &bs, int wi) {
    const int cg = bs.bwd.cg;
    bs.bwd_wi = wi;
    bs.bwd = bs.bwdw[wi];
    bs.bwd.cg = std::min(cg, bs.bwd.nxt);
    bs.inj_start = bs.inj_start_w[wi];
    bs.inj_list = bs.inj_list_w[wi];
    bs.inj_groups = bs.inj_groups_w[wi];
}

template <int SM>
static void launch_bwd2(swit_plan *pl, int it, const float *ri, float *ro, const float *bi, float *bo,
                        hipStream_t st, ShotRange r) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    const void *strips = SM == 0 ? static_cast<const void *>(pl->strips) : static_cast<const void *>(bs.strips16);
    hipLaunchKernelGGL(k_bwd_fused2<SM>, dim3(r.n * bs.bwd2.ntiles), dim3(kThreads), 0, st, d, it, bs.bwd2, r.s0,
                       bs.KB, bs.CP, bs.CV, pl->geo, bs.src, strips, bs.smax, ri, ro, bi, bo, bs.inj2_start,
                       bs.inj2_groups, bs.inj2_list, bs.injx_start, bs.injx_groups, bs.injx_list, bs.injx_cap,
                       pl->sadj_t, bs.acc);
}

// Backward launch that takes the reverse pass from step it down: two steps (it, it-1)
// when the plan runs the two-step kernel and it >= 3, else one.  Returns the steps done.
static int launch_backward(swit_plan *pl, int it, const float *ri, float *ro, const float *bi, float *bo,
                           hipStream_t st, ShotRange range = ShotRange{0, -1}) {
    BatchedState &bs = pl->bat;
    const ShotRange r = shots_of(pl, range);
    const bool two = bs.bwd_steps == 2 && it >= 3;
    switch (bs.strip_mode) {
    case 0: two ? launch_bwd2<0>(pl, it, ri, ro, bi, bo, st, r) : launch_bwd<0>(pl, it, ri, ro, bi, bo, st, r); break;
    case 1: two ? launch_bwd2<1>(pl, it, ri, ro, bi, bo, st, r) : launch_bwd<1>(pl, it, ri, ro, bi, bo, st, r); break;
    default: two ? launch_bwd2<2>(pl, it, ri, ro, bi, bo, st, r) : launch_bwd<2>(pl, it, ri, ro, bi, bo, st, r); break;
    }
    return two ? 2 : 1;
}

static int forward_loop(swit_plan *pl, int save, hipStream_t st, ShotRange r = ShotRange{0, -1}) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    int first = 1;                               // oldest strip row not yet encoded
    for (int it = 2; it <= d.nt; ++it) {
        const float *in = bs.F[it % 2];          // it = 2 reads F[0]
        float *out = bs.F[(it + 1) % 2];
        launch_fwd(pl, it, in, out, save, st, DSave{nullptr, 0, 0}, nullptr, r);
        LAUNCH_CHECK();
        // 16-bit plans: encode the ring once its last slot is written, and at the end
        if (save && bs.strip_mode && ((it - 1) % kStripRing == kStripRing - 1 || it == d.nt)) {
            launch_encode(pl, first, it - 1, st, r);
            LAUNCH_CHECK();
            first = it;
        }
    }
    return SWIT_OK;
}

static int forward_epilogue(swit_plan *pl, int save, hipStream_t st) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    if (save) {
        const float *fin = bs.F[(d.nt + 1) % 2];
        hipLaunchKernelGGL(k_end_state, dim3(cdiv(d.nz, 64), cdiv(d.nx, 4), pl->ns), dim3(64, 4), 0, st, d, fin,
                           bs.end);
        LAUNCH_CHECK();
    }
    return SWIT_OK;
}

static int adjoint_prologue(swit_plan *pl, hipStream_t st) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    const int ns = pl->ns;
    const size_t np = (size_t)d.nxp * d.ld, ni = (size_t)d.nx * d.nz;
    TRY(fill(bs.Bk[0], 3 * np * ns, st));
    TRY(copyk(bs.end, bs.R[0], 3 * ni * ns, st));
    TRY(fill(bs.acc, 3 * ni * ns, st));
    return SWIT_OK;
}

static int adjoint_loop(swit_plan *pl, hipStream_t st, ShotRange r = ShotRange{0, -1}) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    for (int it = d.nt, k = 0; it >= 2; k ^= 1) {
        it -= launch_backward(pl, it, bs.R[k], bs.R[k ^ 1], bs.Bk[k], bs.Bk[k ^ 1], st, r);
        LAUNCH_CHECK();
    }
    return SWIT_OK;
}

static int adjoint_epilogue(swit_plan *pl, hipStream_t st) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    const size_t ni = (size_t)d.nx * d.nz;
    hipLaunchKernelGGL(k_reduce_acc, dim3(cdiv((int)ni, 256)), dim3(256), 0, st, d, pl->ns, pl->c_int, bs.acc,
                       bs.out, bs.out + ni, bs.out + 2 * ni);
    LAUNCH_CHECK();
    return SWIT_OK;
}

// Two streams: body(k, half, stream) runs the first half of the shots (k = 0) on st and
// the second (k = 1) on the plan's second stream, forked from st and joined back into it.
template <class Body>
static int on_two_streams(swit_plan *pl, hipStream_t st, Body body) {
    BatchedState &bs = pl->bat;
    const int h = pl->ns / 2;
    const ShotRange half[2] = {{0, h}, {h, pl->ns - h}};
    if (!bs.side) {
        HIP_TRY(hipStreamCreateWithFlags(&bs.side, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&bs.fork, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&bs.join, hipEventDisableTiming));
    }
    HIP_TRY(hipEventRecord(bs.fork, st));
    HIP_TRY(hipStreamWaitEvent(bs.side, bs.fork, 0));
    for (int k = 0; k < 2; ++k) TRY(body(k, half[k], k == 0 ? st : bs.side));
    HIP_TRY(hipEventRecord(bs.join, bs.side));
    HIP_TRY(hipStreamWaitEvent(st, bs.join, 0));
    return SWIT_OK;
}

// Capture one time loop (which = 0 forward, 1 backward) into an executable graph.
static int capture(swit_plan *pl, hipGraphExec_t *exec, int which, int save, ShotRange r = ShotRange{0, -1}) {
    BatchedState &bs = pl->bat;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(bs.cap, hipStreamCaptureModeThreadLocal));
    int rc = which == 0 ? forward_loop(pl, save, bs.cap, r) : adjoint_loop(pl, bs.cap, r);
    hipError_t e = hipStreamEndCapture(bs.cap, &graph);
    if (rc != SWIT_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return set_err(SWIT_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return set_err(SWIT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    return SWIT_OK;
}

// Run one time loop on st: the captured graph, or direct launches.  With profiling on,
// a pair of events brackets the loop; the step kernels run back to back in between,
// so the pair's elapsed time / (nt - 1) is the step kernel's average duration.
static int run_loop(swit_plan *pl, int which, int save, hipStream_t st) {
    BatchedState &bs = pl->bat;
    BatchedState::Timed t{which, (long long)pl->d.nt - 1, nullptr, nullptr};
    if (bs.profile) {
        TRY(pool_event(bs, &t.e0));
        TRY(pool_event(bs, &t.e1));
        bs.pending.push_back(t);
        HIP_TRY(hipEventRecord(t.e0, st));
    }
    if (bs.streams == 2 && pl->ns >= 2) {
        TRY(on_two_streams(pl, st, [&](int k, ShotRange half, hipStream_t sk) -> int {
            if (!bs.use_graph)
                return which == 0 ? forward_loop(pl, save, sk, half) : adjoint_loop(pl, sk, half);
            hipGraphExec_t &g = which == 0 ? bs.gfwd_h[k][save ? 1 : 0] : bs.gbwd_h[k];
            if (!g) TRY(capture(pl, &g, which, save, half));
            HIP_TRY(hipGraphLaunch(g, sk));
            return SWIT_OK;
        }));
    } else if (bs.use_graph) {
        hipGraphExec_t &g = which == 0 ? bs.gfwd[save ? 1 : 0] : bs.gbwd;
        if (!g) TRY(capture(pl, &g, which, save));
        HIP_TRY(hipGraphLaunch(g, st));
    } else {
        TRY(which == 0 ? forward_loop(pl, save, st) : adjoint_loop(pl, st));
    }
    if (bs.profile) HIP_TRY(hipEventRecord(t.e1, st));
    return SWIT_OK;
}

static int batched_forward(swit_plan *pl, const float *src, int save, hipStream_t st) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    HIP_TRY(hipMemcpyAsync(bs.src, src, sizeof(float) * pl->ns * d.nt, hipMemcpyDeviceToDevice, st));
    TRY(forward_prologue(pl, save, st));
    TRY(run_loop(pl, 0, save, st));
    return forward_epilogue(pl, save, st);
}

static int batched_adjoint(swit_plan *pl, const float *src, double *grad, double *fw, double *bk,
                           hipStream_t st) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    const size_t ni = (size_t)d.nx * d.nz;
    HIP_TRY(hipMemcpyAsync(bs.src, src, sizeof(float) * pl->ns * d.nt, hipMemcpyDeviceToDevice, st));
    TRY(adjoint_prologue(pl, st));
    TRY(run_loop(pl, 1, 0, st));
    TRY(adjoint_epilogue(pl, st));
    HIP_TRY(hipMemcpyAsync(grad, bs.out, sizeof(double) * ni, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(fw, bs.out + ni, sizeof(double) * ni, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(bk, bs.out + 2 * ni, sizeof(double) * ni, hipMemcpyDeviceToDevice, st));
    return SWIT_OK;
}

// ---- options and measurement -----------------------------------------------------------

// Captured graphs hold the launch arguments: drop them when those change.
static void drop_half_graphs(BatchedState &bs) {
    for (auto &row : bs.gfwd_h)
        for (auto &g : row)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    for (auto &g : bs.gbwd_h)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    for (auto &row : bs.gex_h)
        for (auto &g : row)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

static void drop_graphs(BatchedState &bs) {
    for (auto &g : bs.gfwd)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    if (bs.gbwd) { (void)hipGraphExecDestroy(bs.gbwd); bs.gbwd = nullptr; }
    drop_half_graphs(bs);
    if (bs.gex_f) { (void)hipGraphExecDestroy(bs.gex_f); bs.gex_f = nullptr; }
    if (bs.gex_b) { (void)hipGraphExecDestroy(bs.gex_b); bs.gex_b = nullptr; }
    ++bs.gen;
}

// Average duration per time step (us) of back-to-back step-kernel launches covering
// `steps` steps (which = 0 forward with strip saving, 1 the plan's backward kernel) on the
// plan's buffers, after two warm-up launches.
static int time_steps(swit_plan *pl, int which, int steps, hipStream_t st, double *us) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_err(SWIT_EHIP, "hipEventCreate: %s", hipGetErrorString(e));
    }
    // steps per launch and launches timed
    const int spl = which == 1 && bs.bwd_steps == 2 && d.nt >= 3 ? 2 : 1;
    const int nl = cdiv(steps, spl);
    auto launch = [&](int k) {
        if (which == 0) {
            const int it = 2 + k % (d.nt - 1);
            // with strip saving, as the gradient's forward runs (a plan without strips
            // only runs forwards that save none)
            launch_fwd(pl, it, bs.F[k & 1], bs.F[(k + 1) & 1], (pl->flags & SWIT_PLAN_NO_STRIPS) ? 0 : 1, st);
        } else {
            // the plan's backward kernel; a two-step launch starts at it >= 3
            const int it = spl == 2 ? d.nt - (2 * k) % (d.nt - 2) : d.nt - k % (d.nt - 1);
            launch_backward(pl, it, bs.R[k & 1], bs.R[(k + 1) & 1], bs.Bk[k & 1], bs.Bk[(k + 1) & 1], st);
        }
    };
    int rc = SWIT_OK;
    for (int k = 0; k < 2; ++k) launch(k);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int k = 0; k < nl && e == hipSuccess; ++k) launch(k);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) rc = set_err(SWIT_EHIP, "autotune timing: %s", hipGetErrorString(e));
    *us = 1000.0 * ms / ((double)nl * spl);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

extern "C" int swit_plan_autotune(swit_plan *pl, int steps, int *choice, double *us, hipStream_t stream) {
    g_err[0] = 0;
    if (!pl || !choice || !us) return set_err(SWIT_EINVAL, "plan, choice and us are required");
    if (!(pl->flags & SWIT_PATH_BATCHED)) return set_err(SWIT_EINVAL, "the autotuner tunes the batched path");
    if (!pl->model_set) return set_err(SWIT_ESTATE, "swit_plan_set_model must be called first");
    if (steps < 1 || steps > 1000) return set_err(SWIT_EINVAL, "steps must be in [1, 1000]");
    BatchedState &bs = pl->bat;
    HIP_TRY(hipStreamSynchronize(stream));
    // the wavefield and strip buffers are scratch here: every job clears what it reads
    // first, and the saved strips are overwritten
    pl->strips_valid = false;
    bs.exact_valid = false;
    static const int kGroups[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12};
    static const int kCols[] = {1, 2, 4, 8, 16};
    constexpr int kRepeats = 3;      // each candidate keeps its best of three bursts
    double best_f = 1e30, best_b = 1e30;
    int bg = bs.Gf, bcf = bs.fwd.cg, bcb = bs.bwd.cg;
    for (int G : kGroups) {
        if (G > pl->ns) continue;
        for (int cg : kCols) {
            if (cg > bs.fwd.nxt) continue;
            bs.Gf = G;
            bs.ngroups_f = cdiv(pl->ns, G);
            bs.fwd.cg = cg;
            double t = 1e30;
            for (int r = 0; r < kRepeats; ++r) {
                double tr;
                TRY(time_steps(pl, 0, steps, stream, &tr));
                t = std::min(t, tr);
            }
            if (t < best_f) { best_f = t; bg = G; bcf = cg; }
        }
    }
    int bsteps = bs.bwd_steps, bwi = bs.bwd_wi, bbal = bs.bwd_balance;
    const int wi0 = bs.bwd_wi;
    // the backward kernels read strips: a plan without them has no backward to tune
    const int sp_max = (pl->flags & SWIT_PLAN_NO_STRIPS) ? 0 : 2;
    for (int sp = 1; sp <= sp_max; ++sp) {
        if (sp == 2 && pl->d.nt < 3) continue;
        bs.bwd_steps = sp;
        // the one-step kernel's tile width and XCD work split are candidates too; the
        // two-step kernel keeps its own tiling and split
        for (int wi = 0; wi < 2; ++wi) {
            if (sp == 2 && wi != wi0) continue;
            select_bwd_width(bs, wi);
            for (int bal = 0; bal <= (sp == 1 ? 1 : 0); ++bal) {
                if (sp == 1) bs.bwd_balance = bal;
                for (int cg : kCols) {
                    if (cg > bs.bwd.nxt || (sp == 2 && cg > bs.bwd2.nxt)) continue;
                    bs.bwd.cg = bs.bwd2.cg = cg;
                    double t = 1e30;
                    for (int r = 0; r < kRepeats; ++r) {
                        double tr;
                        TRY(time_steps(pl, 1, steps, stream, &tr));
                        t = std::min(t, tr);
                    }
                    if (t < best_b) { best_b = t; bcb = cg; bsteps = sp; bwi = wi; bbal = bs.bwd_balance; }
                }
            }
        }
    }
    bs.Gf = bg;
    bs.ngroups_f = cdiv(pl->ns, bg);
    bs.fwd.cg = bcf;
    select_bwd_width(bs, bwi);
    bs.bwd.cg = bs.bwd2.cg = bcb;
    bs.bwd_steps = bsteps;
    bs.bwd_balance = bbal;
    drop_graphs(bs);
    choice[0] = bg;
    choice[1] = bcf;
    choice[2] = bcb;
    choice[3] = bsteps;
    choice[4] = bs.bwd.tx;
    choice[5] = bbal;
    us[0] = best_f;
    us[1] = sp_max ? best_b : 0.0;
    return SWIT_OK;
}

extern "C" int swit_plan_set_option(swit_plan *pl, int option, int value) {
    g_err[0] = 0;
    if (!pl) return set_err(SWIT_EINVAL, "plan is NULL");
    BatchedState &bs = pl->bat;
    switch (option) {
    case SWIT_OPT_GRAPH:
        bs.use_graph = value != 0;
        return SWIT_OK;
    case SWIT_OPT_PROFILE:
        bs.profile = value != 0;
        if (value) {
            TRY(drain_timings(bs));
            bs.prof_ms[0] = bs.prof_ms[1] = 0;
            bs.prof_n[0] = bs.prof_n[1] = 0;
        }
        return SWIT_OK;
    case SWIT_OPT_FWD_GROUP:
        if (value < 1 || value > pl->ns) return set_err(SWIT_EINVAL, "group size %d outside [1, ns]", value);
        bs.Gf = value;
        bs.ngroups_f = cdiv(pl->ns, value);
        drop_graphs(bs);
        return SWIT_OK;
    case SWIT_OPT_FWD_COLGROUP:
    case SWIT_OPT_BWD_COLGROUP: {
        TileMap &tm = option == SWIT_OPT_FWD_COLGROUP ? bs.fwd : bs.bwd;
        if (value < 1 || value > tm.nxt) return set_err(SWIT_EINVAL, "column group %d outside [1, %d]", value, tm.nxt);
        tm.cg = value;
        if (option == SWIT_OPT_BWD_COLGROUP) bs.bwd2.cg = value;
        drop_graphs(bs);
        return SWIT_OK;
    }
    case SWIT_OPT_BWD_TILE_X: {
        int wi = -1;
        for (int k = 0; k < 2; ++k)
            if (bk::kBTXs[k] == value) wi = k;
        if (wi < 0) return set_err(SWIT_EINVAL, "backward tile width %d is not one of %d, %d", value, bk::kBTXs[0],
                                   bk::kBTXs[1]);
        select_bwd_width(bs, wi);
        drop_graphs(bs);
        return SWIT_OK;
    }
    case SWIT_OPT_EXACT_FUSED:
        bs.exact_fused = value != 0;
        if (bs.gex_b) { (void)hipGraphExecDestroy(bs.gex_b); bs.gex_b = nullptr; }
        for (auto &row : bs.gex_h)
            for (auto &g : row)
                if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
        return SWIT_OK;
    case SWIT_OPT_BWD_STEPS:
        if (value != 1 && value != 2) return set_err(SWIT_EINVAL, "backward steps per launch must be 1 or 2");
        bs.bwd_steps = value;
        drop_graphs(bs);
        return SWIT_OK;
    case SWIT_OPT_STREAMS:
        if (value != 1 && value != 2) return set_err(SWIT_EINVAL, "streams must be 1 or 2");
        bs.streams = value;
        return SWIT_OK;
    case SWIT_OPT_BWD_BALANCE:
        if (value != 0 && value != 1) return set_err(SWIT_EINVAL, "backward balance must be 0 or 1");
        bs.bwd_balance = value;
        drop_graphs(bs);
        return SWIT_OK;
    default:
        return set_err(SWIT_EINVAL, "unknown option %d", option);
    }
}

extern "C" int swit_plan_kernel_time(swit_plan *pl, int which, double *total_ms, int64_t *launches) {
    g_err[0] = 0;
    if (!pl || which < 0 || which > 1 || !total_ms || !launches) return set_err(SWIT_EINVAL, "bad arguments");
    TRY(drain_timings(pl->bat));
    *total_ms = pl->bat.prof_ms[which];
    *launches = pl->bat.prof_n[which];
    return SWIT_OK;
}

// ---- plan groups ------------------------------------------------------------------------

struct swit_group {
    std::vector<swit_plan *> plans;
    hipStream_t cap = nullptr;               // capture origin
    std::vector<hipStream_t> branch;         // one capture branch per plan
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraphExec_t gf[2] = {nullptr, nullptr};   // forward without / with strip saving
    hipGraphExec_t gb = nullptr;
    std::vector<int> gens_f[2], gens_b;          // plan generations at capture
    hipGraphExec_t gg = nullptr;                 // whole gradient (swit_group_gradient)
    std::vector<int> gens_g;
    std::vector<const void *> ptrs_g;            // the caller's arrays the gradient graph uses
    int kind_g = -1;

    void drop(hipGraphExec_t &g) {
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    }
    void release() {
        drop(gf[0]);
        drop(gf[1]);
        drop(gb);
        drop(gg);
        for (hipStream_t st : branch) (void)hipStreamDestroy(st);
        branch.clear();
        for (hipEvent_t e : join) (void)hipEventDestroy(e);
        join.clear();
        if (fork) { (void)hipEventDestroy(fork); fork = nullptr; }
        if (cap) { (void)hipStreamDestroy(cap); cap = nullptr; }
    }
    std::vector<int> gens() const {
        std::vector<int> v;
        for (swit_plan *p : plans) v.push_back(p->bat.gen);
        return v;
    }
};

// Capture every plan's whole job (which = 0 forward, 1 adjoint) on its own branch.
static int group_capture(swit_group *gr, int which, int save, hipGraphExec_t *exec) {
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(gr->cap, hipStreamCaptureModeThreadLocal));
    int rc = SWIT_OK;
    hipError_t e = hipEventRecord(gr->fork, gr->cap);
    for (size_t i = 0; i < gr->plans.size() && rc == SWIT_OK && e == hipSuccess; ++i) {
        swit_plan *pl = gr->plans[i];
        hipStream_t st = gr->branch[i];
        e = hipStreamWaitEvent(st, gr->fork, 0);
        if (e != hipSuccess) break;
        if (which == 0) {
            rc = forward_prologue(pl, save, st);
            if (rc == SWIT_OK) rc = forward_loop(pl, save, st);
            if (rc == SWIT_OK) rc = forward_epilogue(pl, save, st);
        } else {
            rc = adjoint_prologue(pl, st);
            if (rc == SWIT_OK) rc = adjoint_loop(pl, st);
            if (rc == SWIT_OK) rc = adjoint_epilogue(pl, st);
        }
        if (rc == SWIT_OK) e = hipEventRecord(gr->join[i], st);
        if (rc == SWIT_OK && e == hipSuccess) e = hipStreamWaitEvent(gr->cap, gr->join[i], 0);
    }
    hipError_t ee = hipStreamEndCapture(gr->cap, &graph);
    if (rc != SWIT_OK || e != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc != SWIT_OK) return rc;
        return set_err(SWIT_EHIP, "group capture: %s", hipGetErrorString(e != hipSuccess ? e : ee));
    }
    e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return set_err(SWIT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    return SWIT_OK;
}

extern "C" int swit_group_create(swit_plan *const *plans, int nplans, swit_group **out) {
    g_err[0] = 0;
    if (!out) return set_err(SWIT_EINVAL, "out is NULL");
    *out = nullptr;
    if (!plans || nplans < 1) return set_err(SWIT_EINVAL, "need at least one plan");
    swit_group *gr = new (std::nothrow) swit_group();
    if (!gr) return set_err(SWIT_ENOMEM, "host allocation failed");
    for (int i = 0; i < nplans; ++i) {
        swit_plan *pl = plans[i];
        if (!pl || !(pl->flags & SWIT_PATH_BATCHED)) {
            delete gr;
            return set_err(SWIT_EINVAL, "plan %d is not a batched plan", i);
        }
        for (swit_plan *q : gr->plans)
            if (q == pl) {
                delete gr;
                return set_err(SWIT_EINVAL, "plan %d appears twice", i);
            }
        gr->plans.push_back(pl);
    }
    auto fail = [&](hipError_t e) {
        gr->release();
        delete gr;
        return set_err(SWIT_EHIP, "group setup: %s", hipGetErrorString(e));
    };
    hipError_t e = hipStreamCreateWithFlags(&gr->cap, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(e);
    e = hipEventCreateWithFlags(&gr->fork, hipEventDisableTiming);
    if (e != hipSuccess) return fail(e);
    for (int i = 0; i < nplans; ++i) {
        hipStream_t st;
        hipEvent_t ev;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) return fail(e);
        gr->branch.push_back(st);
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return fail(e);
        gr->join.push_back(ev);
    }
    *out = gr;
    return SWIT_OK;
}

extern "C" int swit_group_destroy(swit_group *gr) {
    if (!gr) return SWIT_OK;
    (void)hipDeviceSynchronize();
    gr->release();
    delete gr;
    return SWIT_OK;
}

extern "C" int swit_group_forward(swit_group *gr, const float *const *src, float *const *seis, int save_strips,
                                  hipStream_t stream) {
    g_err[0] = 0;
    if (!gr || !src || !seis) return set_err(SWIT_EINVAL, "group, src and seis are required");
    if (save_strips < 0 || save_strips > 2) return set_err(SWIT_EINVAL, "save_strips must be 0, 1 or 2");
    const int save = save_strips == 1;
    const size_t n = gr->plans.size();
    for (size_t i = 0; i < n; ++i) {
        if (!src[i] || !seis[i]) return set_err(SWIT_EINVAL, "plan %zu: src and seis are required", i);
        if (!gr->plans[i]->model_set) return set_err(SWIT_ESTATE, "plan %zu: no model set", i);
        if (save && (gr->plans[i]->flags & SWIT_PLAN_NO_STRIPS))
            return set_err(SWIT_EINVAL, "plan %zu was created without strips (SWIT_PLAN_NO_STRIPS)", i);
    }
    for (size_t i = 0; i < n; ++i) {
        swit_plan *pl = gr->plans[i];
        if (save_strips != 2) pl->strips_valid = false;
        HIP_TRY(hipMemcpyAsync(pl->bat.src, src[i], sizeof(float) * pl->ns * pl->d.nt, hipMemcpyDeviceToDevice,
                               stream));
    }
    hipGraphExec_t &g = gr->gf[save];
    if (g && gr->gens_f[save] != gr->gens()) gr->drop(g);
    if (!g) {
        TRY(group_capture(gr, 0, save, &g));
        gr->gens_f[save] = gr->gens();
    }
    HIP_TRY(hipGraphLaunch(g, stream));
    for (size_t i = 0; i < n; ++i) {
        swit_plan *pl = gr->plans[i];
        const Dims &d = pl->d;
        hipLaunchKernelGGL(k_transpose, dim3(cdiv(d.ngmax, 32), cdiv(d.nt, 32), pl->ns), dim3(32, 8), 0, stream,
                           d.nt, d.ngmax, pl->seis_t, seis[i]);
        LAUNCH_CHECK();
        if (save_strips != 2) pl->strips_valid = save_strips == 1;
    }
    return SWIT_OK;
}

extern "C" int swit_group_adjoint(swit_group *gr, const float *const *src, const float *const *adjsrc,
                                  double *const *grad, double *const *illum_forw, double *const *illum_back,
                                  hipStream_t stream) {
    g_err[0] = 0;
    if (!gr || !src || !adjsrc || !grad || !illum_forw || !illum_back)
        return set_err(SWIT_EINVAL, "group, src, adjsrc and the three outputs are required");
    const size_t n = gr->plans.size();
    for (size_t i = 0; i < n; ++i) {
        if (!src[i] || !adjsrc[i] || !grad[i] || !illum_forw[i] || !illum_back[i])
            return set_err(SWIT_EINVAL, "plan %zu: missing array", i);
        if (!gr->plans[i]->strips_valid)
            return set_err(SWIT_ESTATE, "plan %zu: no saved strips: run a strip-saving forward first", i);
    }
    for (size_t i = 0; i < n; ++i) {
        swit_plan *pl = gr->plans[i];
        const Dims &d = pl->d;
        hipLaunchKernelGGL(k_transpose, dim3(cdiv(d.nt, 32), cdiv(d.ngmax, 32), pl->ns), dim3(32, 8), 0, stream,
                           d.ngmax, d.nt, adjsrc[i], pl->sadj_t);
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(pl->bat.src, src[i], sizeof(float) * pl->ns * d.nt, hipMemcpyDeviceToDevice,
                               stream));
    }
    if (gr->gb && gr->gens_b != gr->gens()) gr->drop(gr->gb);
    if (!gr->gb) {
        TRY(group_capture(gr, 1, 0, &gr->gb));
        gr->gens_b = gr->gens();
    }
    HIP_TRY(hipGraphLaunch(gr->gb, stream));
    for (size_t i = 0; i < n; ++i) {
        swit_plan *pl = gr->plans[i];
        const size_t ni = (size_t)pl->d.nx * pl->d.nz;
        const double *o = pl->bat.out;
        HIP_TRY(hipMemcpyAsync(grad[i], o, sizeof(double) * ni, hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipMemcpyAsync(illum_forw[i], o + ni, sizeof(double) * ni, hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipMemcpyAsync(illum_back[i], o + 2 * ni, sizeof(double) * ni, hipMemcpyDeviceToDevice, stream));
    }
    return SWIT_OK;
}

// ---- misfits -------------------------------------------------------------------------------

// Radices of an n-point transform: 4s first, then 2, 3 and the odd primes in order.
static FftFactors fft_factors(int n) {
    FftFactors f{};
    f.n = n;
    f.s = kTwB;
    int m = n;
    // the 10-point register butterfly first: fewer passes through LDS (4000 = 10 10 10 4
    // takes four stages instead of six)
    for (int p : {10, 4, 2, 3})
        while (m % p == 0) { f.radix[f.nst++] = p; m /= p; }
    for (int p = 5; m > 1; p += 2)
        while (m % p == 0) { f.radix[f.nst++] = p; m /= p; }
    return f;
}

template <int NMAX>
static void launch_fft_misfit(int kind, const FftFactors &f, int ntraces, float dt, const float *obs,
                              const float *syn, float *adj, double *work, hipStream_t stream) {
    if (kind == SWIT_MISFIT_ENVELOPE)
        hipLaunchKernelGGL(k_misfit_envelope<NMAX>, dim3(ntraces), dim3(fft_threads<NMAX>()), 0, stream, f, dt, obs, syn,
                           adj, work);
    else
        hipLaunchKernelGGL(k_misfit_traveltime<NMAX>, dim3(ntraces), dim3(fft_threads<NMAX>()), 0, stream, f, dt, obs, syn,
                           adj, work);
}

static int misfit_check(int kind, int ntraces, int nt) {
    if (kind != SWIT_MISFIT_WAVEFORM && kind != SWIT_MISFIT_GLOBAL_CORRELATION && kind != SWIT_MISFIT_ENVELOPE &&
        kind != SWIT_MISFIT_TRAVELTIME)
        return set_err(SWIT_EINVAL, "unknown misfit kind %d", kind);
    if (ntraces < 1 || nt < 1) return set_err(SWIT_EINVAL, "need ntraces >= 1 and nt >= 1");
    if ((kind == SWIT_MISFIT_ENVELOPE || kind == SWIT_MISFIT_TRAVELTIME) && nt > kFftMaxN)
        return set_err(SWIT_EINVAL, "envelope and traveltime misfits take nt <= %d (got %d)", kFftMaxN, nt);
    return SWIT_OK;
}

// The misfit kernels of a checked call (also captured into group gradient graphs).
static int launch_misfit(int kind, int ntraces, int nt, float dt, const float *obs, const float *syn, float *adj,
                         double *work, double *misfit, hipStream_t stream) {
    if (kind == SWIT_MISFIT_WAVEFORM) {
        hipLaunchKernelGGL(k_misfit_trace<0>, dim3(ntraces), dim3(kMisfitThreads), 0, stream, nt, dt, obs, syn, adj,
                           work);
    } else if (kind == SWIT_MISFIT_GLOBAL_CORRELATION) {
        hipLaunchKernelGGL(k_misfit_trace<1>, dim3(ntraces), dim3(kMisfitThreads), 0, stream, nt, dt, obs, syn, adj,
                           work);
    } else {
        const FftFactors f = fft_factors(nt);
        if (nt <= 2048) launch_fft_misfit<2048>(kind, f, ntraces, dt, obs, syn, adj, work, stream);
        else if (nt <= 4096) launch_fft_misfit<4096>(kind, f, ntraces, dt, obs, syn, adj, work, stream);
        else launch_fft_misfit<kFftMaxN>(kind, f, ntraces, dt, obs, syn, adj, work, stream);
    }
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(kMisfitThreads), 0, stream, ntraces, work, misfit);
    LAUNCH_CHECK();
    return SWIT_OK;
}

extern "C" int swit_misfit(int kind, int ntraces, int nt, float dt, const float *obs, const float *syn,
                           float *adj, double *work, double *misfit, hipStream_t stream) {
    g_err[0] = 0;
    if (const int rc = misfit_check(kind, ntraces, nt)) return rc;
    if (!obs || !syn || !adj || !work || !misfit) return set_err(SWIT_EINVAL, "all arrays are required");
    return launch_misfit(kind, ntraces, nt, dt, obs, syn, adj, work, misfit, stream);
}

// ---- group gradient ---------------------------------------------------------------------

// One plan's whole gradient on stream st: strip-saving forward, seismograms to seis,
// misfit and adjoint source against obs (into adj, per-trace values in work, the sum in
// misfit), the adjoint source to the time-major layout, adjoint, outputs to grad / fw / bk.
static int plan_gradient_ops(swit_plan *pl, int kind, const float *obs, float *seis, float *adj, double *work,
                             double *misfit, double *grad, double *fw, double *bk, hipStream_t st) {
    const Dims &d = pl->d;
    TRY(forward_prologue(pl, 1, st));
    TRY(forward_loop(pl, 1, st));
    TRY(forward_epilogue(pl, 1, st));
    hipLaunchKernelGGL(k_transpose, dim3(cdiv(d.ngmax, 32), cdiv(d.nt, 32), pl->ns), dim3(32, 8), 0, st, d.nt,
                       d.ngmax, pl->seis_t, seis);
    LAUNCH_CHECK();
    TRY(launch_misfit(kind, pl->ns * d.ngmax, d.nt, d.dt, obs, seis, adj, work, misfit, st));
    hipLaunchKernelGGL(k_transpose, dim3(cdiv(d.nt, 32), cdiv(d.ngmax, 32), pl->ns), dim3(32, 8), 0, st, d.ngmax,
                       d.nt, adj, pl->sadj_t);
    LAUNCH_CHECK();
    TRY(adjoint_prologue(pl, st));
    TRY(adjoint_loop(pl, st));
    TRY(adjoint_epilogue(pl, st));
    const size_t ni = (size_t)d.nx * d.nz;
    const double *o = pl->bat.out;
    HIP_TRY(hipMemcpyAsync(grad, o, sizeof(double) * ni, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(fw, o + ni, sizeof(double) * ni, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(bk, o + 2 * ni, sizeof(double) * ni, hipMemcpyDeviceToDevice, st));
    return SWIT_OK;
}

extern "C" int swit_group_gradient(swit_group *gr, int kind, const float *const *src, const float *const *obs,
                                   float *const *seis, float *const *adj, double *const *work,
                                   double *const *misfit, double *const *grad, double *const *illum_forw,
                                   double *const *illum_back, hipStream_t stream) {
    g_err[0] = 0;
    if (!gr || !src || !obs || !seis || !adj || !work || !misfit || !grad || !illum_forw || !illum_back)
        return set_err(SWIT_EINVAL, "group and every array list are required");
    const size_t n = gr->plans.size();
    std::vector<const void *> ptrs;
    for (size_t i = 0; i < n; ++i) {
        swit_plan *pl = gr->plans[i];
        if (!src[i] || !obs[i] || !seis[i] || !adj[i] || !work[i] || !misfit[i] || !grad[i] || !illum_forw[i] ||
            !illum_back[i])
            return set_err(SWIT_EINVAL, "plan %zu: missing array", i);
        if (!pl->model_set) return set_err(SWIT_ESTATE, "plan %zu: no model set", i);
        if (pl->flags & SWIT_PLAN_NO_STRIPS)
            return set_err(SWIT_EINVAL, "plan %zu was created without strips (SWIT_PLAN_NO_STRIPS)", i);
        TRY(misfit_check(kind, pl->ns * pl->d.ngmax, pl->d.nt));
        for (const void *q : {(const void *)obs[i], (const void *)seis[i], (const void *)adj[i],
                              (const void *)work[i], (const void *)misfit[i], (const void *)grad[i],
                              (const void *)illum_forw[i], (const void *)illum_back[i]})
            ptrs.push_back(q);
    }
    for (size_t i = 0; i < n; ++i) {
        swit_plan *pl = gr->plans[i];
        pl->strips_valid = false;
        HIP_TRY(hipMemcpyAsync(pl->bat.src, src[i], sizeof(float) * pl->ns * pl->d.nt, hipMemcpyDeviceToDevice,
                               stream));
    }
    // the graph holds the caller's arrays: recapture when they, the kind or a plan's
    // launch configuration change
    if (gr->gg && (gr->gens_g != gr->gens() || gr->ptrs_g != ptrs || gr->kind_g != kind)) gr->drop(gr->gg);
    if (!gr->gg) {
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(gr->cap, hipStreamCaptureModeThreadLocal));
        int rc = SWIT_OK;
        hipError_t e = hipEventRecord(gr->fork, gr->cap);
        for (size_t i = 0; i < n && rc == SWIT_OK && e == hipSuccess; ++i) {
            hipStream_t st = gr->branch[i];
            e = hipStreamWaitEvent(st, gr->fork, 0);
            if (e != hipSuccess) break;
            rc = plan_gradient_ops(gr->plans[i], kind, obs[i], seis[i], adj[i], work[i], misfit[i], grad[i],
                                   illum_forw[i], illum_back[i], st);
            if (rc == SWIT_OK) e = hipEventRecord(gr->join[i], st);
            if (rc == SWIT_OK && e == hipSuccess) e = hipStreamWaitEvent(gr->cap, gr->join[i], 0);
        }
        hipError_t ee = hipStreamEndCapture(gr->cap, &graph);
        if (rc != SWIT_OK || e != hipSuccess || ee != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (rc != SWIT_OK) return rc;
            return set_err(SWIT_EHIP, "group capture: %s", hipGetErrorString(e != hipSuccess ? e : ee));
        }
        e = hipGraphInstantiate(&gr->gg, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return set_err(SWIT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        gr->gens_g = gr->gens();
        gr->ptrs_g = ptrs;
        gr->kind_g = kind;
    }
    HIP_TRY(hipGraphLaunch(gr->gg, stream));
    for (size_t i = 0; i < n; ++i) gr->plans[i]->strips_valid = true;
    return SWIT_OK;
}

// ---- exact adjoint ---------------------------------------------------------------------------

// Forward steps it = a+1..b (state after step a in F[(a+1) % 2]) of the shots in r,
// storing the divergence history as dsave says.
static int exact_forward_steps(swit_plan *pl, int a, int b, DSave ds, ShotRange r, hipStream_t st) {
    BatchedState &bs = pl->bat;
    for (int it = a + 1; it <= b; ++it) {
        const float *in = bs.F[it % 2];
        float *out = bs.F[(it + 1) % 2];
        launch_fwd(pl, it, in, out, 0, st, ds, bs.src_ex, r);
        LAUNCH_CHECK();
    }
    return SWIT_OK;
}

// The whole forward of the shots in r: with the full divergence history, or keeping a
// checkpoint of the state every ck steps.  Every buffer is shot-major, so the shots' part
// of a checkpoint is one contiguous span.
static int exact_forward_loop(swit_plan *pl, hipStream_t st, ShotRange range = ShotRange{0, -1}) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    const ShotRange r = shots_of(pl, range);
    if (!bs.ck) return exact_forward_steps(pl, 1, d.nt, DSave{bs.dhist, d.nt, 0}, r, st);
    // checkpoints: the state after step 1 + j ck (after step 1 it is zero)
    const size_t per = (size_t)3 * d.nxp * d.ld, chunk = pl->ns * per;
    float *ck0 = bs.ckpt + r.s0 * per;
    TRY(fill(ck0, r.n * per, st));
    for (int j = 1; j < bs.nck; ++j) {
        const int a = 1 + (j - 1) * bs.ck, b = a + bs.ck;
        TRY(exact_forward_steps(pl, a, b, DSave{nullptr, 0, 0}, r, st));
        TRY(copyk(bs.F[(b + 1) % 2] + r.s0 * per, ck0 + (size_t)j * chunk, r.n * per, st));
    }
    const int a = 1 + (bs.nck - 1) * bs.ck;
    return exact_forward_steps(pl, a, d.nt, DSave{nullptr, 0, 0}, r, st);
}

// Reverse steps b..a+1 of the exact adjoint of the shots in r, reading dJ/dseis from
// bs.gseis.  The fused kernel reads lam[k] and writes lam[k^1] (lam[0] = bs.lam,
// lam[1] = bs.lam2); k carries over between segments.  The separate passes run on all
// shots (r must cover them).
static int exact_adjoint_steps(swit_plan *pl, int a, int b, DSave ds, int &k, ShotRange r, hipStream_t st) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    const int ns = pl->ns;
    if (bs.exact_fused) {
        float *lam[2] = {bs.lam, bs.lam2};
        for (int it = b; it >= a + 1; --it, k ^= 1) {
            hipLaunchKernelGGL(ex::k_ex_step, dim3(r.n * bs.extm.ntiles), dim3(kThreads), 0, st, d, it, bs.extm,
                               r.s0, bs.KB, bs.CP, bs.CV, ds, lam[k], lam[k ^ 1], bs.gk, bs.exinj_start,
                               bs.exinj_groups, bs.exinj_list, bs.exinj_cap, bs.gseis);
            LAUNCH_CHECK();
        }
        return SWIT_OK;
    }
    if (r.s0 != 0 || r.n != ns) return set_err(SWIT_EINVAL, "the separate exact passes run on all shots");
    const dim3 blk(64, 4), grd(cdiv(d.nzp, 64), cdiv(d.nxp, 4), ns);
    const dim3 gcol(cdiv(d.nxp, 256), ns), ginj(cdiv(d.ngmax, 256), ns);
    for (int it = b; it >= a + 1; --it) {
        hipLaunchKernelGGL(ex::k_ex_inject, ginj, dim3(256), 0, st, d, it, pl->rec_ngrp, pl->rec_grp, pl->rec_lst,
                           bs.gseis, bs.lam);
        if (d.free_surface == 1) hipLaunchKernelGGL(ex::k_ex_mirror, gcol, dim3(256), 0, st, d, 2, bs.lam);
        hipLaunchKernelGGL(ex::k_ex_vel, grd, blk, 0, st, d, bs.KB, bs.CV, bs.lam, bs.qbuf);
        hipLaunchKernelGGL(ex::k_ex_p, grd, blk, 0, st, d, bs.qbuf, bs.lam);
        if (d.free_surface == 1) hipLaunchKernelGGL(ex::k_ex_mirror, gcol, dim3(256), 0, st, d, 0, bs.lam);
        hipLaunchKernelGGL(ex::k_ex_grad, grd, blk, 0, st, d, it, bs.KB, bs.CP, ds, bs.lam, bs.qbuf, bs.gk);
        hipLaunchKernelGGL(ex::k_ex_vel2, grd, blk, 0, st, d, bs.qbuf, bs.lam);
        LAUNCH_CHECK();
    }
    return SWIT_OK;
}

// The whole reverse sweep of the shots in r: from the full history, or segment by segment
// from the checkpoints (each segment's forward is recomputed first, bit for bit).
static int exact_adjoint_loop(swit_plan *pl, hipStream_t st, ShotRange range = ShotRange{0, -1}) {
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    const ShotRange r = shots_of(pl, range);
    int k = 0;                      // the sweep starts from lam[0] (both buffers are zero)
    if (!bs.ck) return exact_adjoint_steps(pl, 1, d.nt, DSave{bs.dhist, d.nt, 0}, k, r, st);
    const size_t per = (size_t)3 * d.nxp * d.ld, chunk = pl->ns * per;
    for (int j = bs.nck - 1; j >= 0; --j) {
        const int a = 1 + j * bs.ck, b = std::min(a + bs.ck, d.nt);
        TRY(copyk(bs.ckpt + (size_t)j * chunk + r.s0 * per, bs.F[(a + 1) % 2] + r.s0 * per, r.n * per, st));
        TRY(exact_forward_steps(pl, a, b, DSave{bs.dhist, bs.ck, a}, r, st));
        TRY(exact_adjoint_steps(pl, a, b, DSave{bs.dhist, bs.ck, a}, k, r, st));
    }
    return SWIT_OK;
}

static int exact_capture(swit_plan *pl, int which, hipGraphExec_t *exec, ShotRange r = ShotRange{0, -1}) {
    BatchedState &bs = pl->bat;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(bs.cap, hipStreamCaptureModeThreadLocal));
    int rc = which == 0 ? exact_forward_loop(pl, bs.cap, r) : exact_adjoint_loop(pl, bs.cap, r);
    hipError_t e = hipStreamEndCapture(bs.cap, &graph);
    if (rc != SWIT_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return set_err(SWIT_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return set_err(SWIT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    return SWIT_OK;
}

// Run one exact loop (which = 0 forward, 1 adjoint) on st.  With two streams each half of
// the shots runs on a stream of its own, as in run_loop (the adjoint only with the fused
// kernel: the separate passes cover all shots).
static int exact_run(swit_plan *pl, int which, hipStream_t st) {
    BatchedState &bs = pl->bat;
    if (bs.streams == 2 && pl->ns >= 2 && (which == 0 || bs.exact_fused)) {
        return on_two_streams(pl, st, [&](int k, ShotRange half, hipStream_t sk) -> int {
            if (!bs.use_graph)
                return which == 0 ? exact_forward_loop(pl, sk, half) : exact_adjoint_loop(pl, sk, half);
            hipGraphExec_t &g = bs.gex_h[which][k];
            if (!g) TRY(exact_capture(pl, which, &g, half));
            HIP_TRY(hipGraphLaunch(g, sk));
            return SWIT_OK;
        });
    }
    if (!bs.use_graph) return which == 0 ? exact_forward_loop(pl, st) : exact_adjoint_loop(pl, st);
    hipGraphExec_t &g = which == 0 ? bs.gex_f : bs.gex_b;
    if (!g) TRY(exact_capture(pl, which, &g));
    HIP_TRY(hipGraphLaunch(g, st));
    return SWIT_OK;
}

extern "C" int swit_forward_exact(swit_plan *pl, const float *src, float *seis, hipStream_t stream) {
    g_err[0] = 0;
    if (!pl || !src || !seis) return set_err(SWIT_EINVAL, "plan, src and seis are required");
    if (!(pl->flags & (SWIT_PLAN_EXACT | SWIT_PLAN_EXACT_CKPT)))
        return set_err(SWIT_EINVAL, "plan was not created with SWIT_PLAN_EXACT or SWIT_PLAN_EXACT_CKPT");
    if (!pl->model_set) return set_err(SWIT_ESTATE, "swit_plan_set_model must be called first");
    BatchedState &bs = pl->bat;
    const Dims &d = pl->d;
    bs.exact_valid = false;
    HIP_TRY(hipMemcpyAsync(bs.src_ex, src, sizeof(float) * pl->ns * d.nt, hipMemcpyDeviceToDevice, stream));
    TRY(forward_prologue(pl, 0, stream));
    TRY(exact_run(pl, 0, stream));
    hipLaunchKernelGGL(k_transpose, dim3(cdiv(d.ngmax, 32), cdiv(d.nt, 32), pl->ns), dim3(32, 8), 0, stream, d.nt,
                       d.ngmax, pl->seis_t, seis);
    LAUNCH_CHECK();
    bs.exact_valid = true;
    ++bs.exact_gen;
    return SWIT_OK;
}

extern "C" int64_t swit_plan_exact_generation(const swit_plan *pl) {
    return pl && pl->bat.exact_valid ? pl->bat.exact_gen : 0;
}

extern "C" int swit_adjoint_exact(swit_plan *pl, const float *gseis, double *grad, hipStream_t stream) {
    g_err[0] = 0;
    if (!pl || !gseis || !grad) return set_err(SWIT_EINVAL, "plan, gseis and grad are required");
    if (!(pl->flags & (SWIT_PLAN_EXACT | SWIT_PLAN_EXACT_CKPT)))
        return set_err(SWIT_EINVAL, "plan was not created with SWIT_PLAN_EXACT or SWIT_PLAN_EXACT_CKPT");
    BatchedState &bs = pl->bat;
    if (!bs.exact_valid) return set_err(SWIT_ESTATE, "no divergence history: run swit_forward_exact first");
    const Dims &d = pl->d;
    const int ns = pl->ns;
    const size_t np = (size_t)d.nxp * d.ld;
    HIP_TRY(hipMemcpyAsync(bs.gseis, gseis, sizeof(float) * ns * d.ngmax * d.nt, hipMemcpyDeviceToDevice, stream));
    TRY(fill(bs.lam, 3 * np * ns, stream));
    TRY(fill(bs.lam2, 3 * np * ns, stream));
    TRY(fill(bs.gk, np * ns, stream));
    TRY(exact_run(pl, 1, stream));
    hipLaunchKernelGGL(ex::k_ex_shotsum, dim3((unsigned)std::min<size_t>(cdiv((int)np, 256), 4096)), dim3(256), 0,
                       stream, d, ns, bs.gk, bs.gsum);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(ex::k_ex_fold, dim3(cdiv(d.nz, 64), cdiv(d.nx, 4)), dim3(64, 4), 0, stream, d, pl->c_int,
                       pl->rho_int, bs.gsum, grad);
    LAUNCH_CHECK();
    return SWIT_OK;
}

// ---- gradient preconditioning -----------------------------------------------------------

static int precond_spans(const swit_precond *p, int *sill) {
    const int mn = std::min(p->nx, p->nz);
    *sill = mn > 40 ? 40 : mn / 2;                     // postprocess.py:62-65
    return std::max(std::max(*sill, p->grad_smooth), p->grad_mute / 2);
}

extern "C" int swit_grad_precond_work(const swit_precond *p, size_t *bytes) {
    g_err[0] = 0;
    if (!p || !bytes) return set_err(SWIT_EINVAL, "parameters and bytes are required");
    if (p->nx < 1 || p->nz < 1) return set_err(SWIT_EINVAL, "nx and nz must be positive");
    int sill;
    const int smax = precond_spans(p, &sill);
    *bytes = sizeof(double) * ((size_t)4 * p->nx * p->nz + 3 * (size_t)(2 * smax + 1) + 16);
    return SWIT_OK;
}

namespace {
int precond_check(const swit_precond *p, const float *grad, const float *forw, const float *back, const double *out) {
    if (!p || !grad || !forw || !back || !out)
        return set_err(SWIT_EINVAL, "parameters, grad, forw, back and out are required");
    if (p->nx < 1 || p->nz < 1) return set_err(SWIT_EINVAL, "nx and nz must be positive");
    if (p->grad_mute < 0 || p->grad_mute > p->nz || p->grad_smooth < 0)
        return set_err(SWIT_EINVAL, "grad_mute must be in [0, nz] and grad_smooth >= 0");
    return SWIT_OK;
}

// The preconditioner's launches on a checked parameter set and a workspace of at least
// swit_grad_precond_work bytes.  Every workspace value is written before it is read
// (fields, kernel weights, maxima), so the workspace needs no clearing.
int precond_run(const swit_precond *p, const float *grad, const float *forw, const float *back, const double *mask,
                double *out, void *work, hipStream_t stream) {
    const int nx = p->nx, nz = p->nz;
    const size_t n = (size_t)nx * nz;
    int sill;
    const int smax = precond_spans(p, &sill);
    double *A = static_cast<double *>(work), *B = A + n, *C = B + n, *T = C + n;
    double *f_ill = T + n, *f_grad = f_ill + (2 * smax + 1), *f_tap = f_grad + (2 * smax + 1);
    double *M = f_tap + (2 * smax + 1);                // maxima: forw, back, precond, taper, |grad|
    const dim3 b2(256), g2(cdiv(nz, 256), nx);
    const int ge = (int)std::min<size_t>(cdiv((int)std::min<size_t>(n, (size_t)1 << 30), 256), 2048);
    // smooth2d of the field in X (in place through T), rows z0.. nz-1
    auto smooth = [&](double *X, int z0, int span, const double *f) -> int {
        const dim3 g(cdiv(nz - z0, 256), nx);
        hipLaunchKernelGGL(pc::k_smooth_z, g, b2, 0, stream, nx, nz, z0, span, f, X, T);
        hipLaunchKernelGGL(pc::k_smooth_x, g, b2, 0, stream, nx, nz, z0, span, f, T, X);
        LAUNCH_CHECK();
        return SWIT_OK;
    };
    hipLaunchKernelGGL(pc::k_weights, dim3(1), dim3(1), 0, stream, sill, f_ill);
    // illumination preconditioner (postprocess.py:62-75)
    hipLaunchKernelGGL(pc::k_to_double, dim3(ge), dim3(256), 0, stream, n, forw, A);
    hipLaunchKernelGGL(pc::k_to_double, dim3(ge), dim3(256), 0, stream, n, back, B);
    LAUNCH_CHECK();
    TRY(smooth(A, 0, sill, f_ill));
    TRY(smooth(B, 0, sill, f_ill));
    hipLaunchKernelGGL(pc::k_max, dim3(1), dim3(pc::kMaxThreads), 0, stream, n, A, 0, M);
    hipLaunchKernelGGL(pc::k_max, dim3(1), dim3(pc::kMaxThreads), 0, stream, n, B, 0, M + 1);
    hipLaunchKernelGGL(pc::k_precond_sum, dim3(ge), dim3(256), 0, stream, n, A, B, M, C);
    hipLaunchKernelGGL(pc::k_max, dim3(1), dim3(pc::kMaxThreads), 0, stream, n, C, 0, M + 2);
    LAUNCH_CHECK();
    // taper (postprocess.py:56-58, 88-111): marine zeroes the rows above the mute, land
    // damps them with a smoothed Hamming ramp
    int kind = 0;
    if (p->grad_mute > 0) {
        if (p->marine) {
            kind = 1;
        } else {
            kind = 2;
            const int ts = p->grad_mute;
            hipLaunchKernelGGL(pc::k_weights, dim3(1), dim3(1), 0, stream, ts / 2, f_tap);
            hipLaunchKernelGGL(pc::k_taper0, g2, b2, 0, stream, nx, nz, ts, A);
            LAUNCH_CHECK();
            TRY(smooth(A, 0, ts / 2, f_tap));
            hipLaunchKernelGGL(pc::k_max, dim3(1), dim3(pc::kMaxThreads), 0, stream, n, A, 0, M + 3);
            hipLaunchKernelGGL(pc::k_taper_land, dim3(ge), dim3(256), 0, stream, n, (double)p->grad_thred, M + 3, A);
            LAUNCH_CHECK();
        }
    }
    hipLaunchKernelGGL(pc::k_apply, g2, b2, 0, stream, nx, nz, kind, p->grad_mute, grad, A, C, M, out);
    LAUNCH_CHECK();
    // gradient smoothing (postprocess.py:77-83): below the mute for marine data
    if (p->grad_smooth > 0) {
        const int z0 = p->marine ? p->grad_mute : 0;
        hipLaunchKernelGGL(pc::k_weights, dim3(1), dim3(1), 0, stream, p->grad_smooth, f_grad);
        LAUNCH_CHECK();
        if (z0 < nz) TRY(smooth(out, z0, p->grad_smooth, f_grad));
    }
    if (mask) {
        hipLaunchKernelGGL(pc::k_mul, dim3(ge), dim3(256), 0, stream, n, mask, out);
        LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(pc::k_max, dim3(1), dim3(pc::kMaxThreads), 0, stream, n, out, 1, M + 4);
    hipLaunchKernelGGL(pc::k_scale, dim3(ge), dim3(256), 0, stream, n, p->vpmax, M + 4, out);
    LAUNCH_CHECK();
    return SWIT_OK;
}

}  // namespace

// Device-resident workspace of swit_grad_precond_noalloc.  Each device that runs it holds
// its own copy, for the life of the process; calls on one device share it.
static __device__ __align__(256) unsigned char scratch[SWIT_PRECOND_SCRATCH_BYTES];

namespace {
constexpr size_t kPrecondScratchBytes = SWIT_PRECOND_SCRATCH_BYTES;
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kScratchDevices];   // device address of `scratch`, per device

// The current device's address of `scratch`, resolved once per device with that device
// current.  Host code must use this address: the array's own name is a host stand-in.
int precond_scratch(hipStream_t stream, void **ptr) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices)
        return set_err(SWIT_EINVAL, "device %d is beyond the %d devices the workspace table holds", dev,
                       kScratchDevices);
    if (stream) {
        hipDevice_t sdev = -1;
        HIP_TRY(hipStreamGetDevice(stream, &sdev));
        if (sdev != dev)
            return set_err(SWIT_EINVAL, "the stream belongs to device %d but device %d is current", (int)sdev, dev);
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *q = nullptr;
        HIP_TRY(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
        if (!q) return set_err(SWIT_EHIP, "hipGetSymbolAddress returned no address for the workspace");
        g_scratch_ptr[dev] = q;
    }
    *ptr = g_scratch_ptr[dev];
    return SWIT_OK;
}
}  // namespace

extern "C" int swit_grad_precond(const swit_precond *p, const float *grad, const float *forw, const float *back,
                                 const double *mask, double *out, void *work, size_t work_bytes,
                                 hipStream_t stream) {
    g_err[0] = 0;
    TRY(precond_check(p, grad, forw, back, out));
    if (!work) return set_err(SWIT_EINVAL, "work is required");
    size_t need;
    TRY(swit_grad_precond_work(p, &need));
    if (work_bytes < need) return set_err(SWIT_EINVAL, "work is %zu bytes, needs %zu", work_bytes, need);
    return precond_run(p, grad, forw, back, mask, out, work, stream);
}

extern "C" int swit_grad_precond_noalloc(const swit_precond *p, const float *grad, const float *forw,
                                         const float *back, const double *mask, double *out, hipStream_t stream) {
    g_err[0] = 0;
    TRY(precond_check(p, grad, forw, back, out));
    size_t need;
    TRY(swit_grad_precond_work(p, &need));
    if (need > kPrecondScratchBytes)
        return set_err(SWIT_EINVAL,
                       "a %d x %d model needs %zu bytes of workspace; swit_grad_precond_noalloc holds %zu "
                       "(use swit_grad_precond with a workspace)",
                       p->nx, p->nz, need, kPrecondScratchBytes);
    void *work = nullptr;
    TRY(precond_scratch(stream, &work));
    return precond_run(p, grad, forw, back, mask, out, work, stream);
}

// ---- shot-gather processing --------------------------------------------------------------

namespace {
constexpr int kProcChunk = 16384;   // traces per filtfilt launch (workspace 8 (nt + 6 taps) each)

int process_check(const swit_process_params *p) {
    if (!p) return set_err(SWIT_EINVAL, "parameters are required");
    if (p->ngathers < 1 || p->ntr < 1 || p->nt < 1) return set_err(SWIT_EINVAL, "need ngathers, ntr, nt >= 1");
    if (p->filter_taps) {
        if (p->filter_taps < 2 || p->filter_taps > SWIT_PROCESS_MAX_TAPS)
            return set_err(SWIT_EINVAL, "filter_taps must be 2..%d", SWIT_PROCESS_MAX_TAPS);
        if (p->a[0] != 1.0) return set_err(SWIT_EINVAL, "filter coefficients must have a[0] = 1");
        if (p->nt <= 3 * p->filter_taps)
            return set_err(SWIT_EINVAL, "the filter needs nt > %d (3 * taps)", 3 * p->filter_taps);
    }
    if (p->norm_trace < 0 || p->norm_trace > SWIT_NORM_MAX) return set_err(SWIT_EINVAL, "bad norm_trace");
    if (p->norm_event < 0 || p->norm_event > SWIT_NORM_L2) return set_err(SWIT_EINVAL, "bad norm_event");
    if (p->mute_late_arrival && !(p->dt > 0.0)) return set_err(SWIT_EINVAL, "the window mute needs dt > 0");
    return SWIT_OK;
}

size_t process_bytes(const swit_process_params *p) {
    const size_t ntraces = (size_t)p->ngathers * p->ntr;
    size_t b = 0;
    if (p->filter_taps) b += sizeof(double) * (size_t)(p->nt + 6 * p->filter_taps) * std::min<size_t>(ntraces, kProcChunk);
    if (p->norm_event) b = std::max(b, sizeof(float) * ntraces);
    return std::max<size_t>(b, 256);
}

template <int K>
void launch_filtfilt_k(const pr::ProcArgs &a, int n, float *data, double *y, hipStream_t st) {
    hipLaunchKernelGGL(pr::k_filtfilt<K>, dim3((n + pr::kFilterThreads - 1) / pr::kFilterThreads),
                       dim3(pr::kFilterThreads), 0, st, a, n, data, y);
}

template <int K = 2>
void launch_filtfilt(int taps, const pr::ProcArgs &a, int n, float *data, double *y, hipStream_t st) {
    if constexpr (K <= pr::kMaxTaps) {
        if (taps == K) return launch_filtfilt_k<K>(a, n, data, y, st);
        launch_filtfilt<K + 1>(taps, a, n, data, y, st);
    }
}
}  // namespace

extern "C" int swit_process_work(const swit_process_params *p, size_t *bytes) {
    g_err[0] = 0;
    if (const int rc = process_check(p)) return rc;
    if (!bytes) return set_err(SWIT_EINVAL, "bytes is required");
    *bytes = process_bytes(p);
    return SWIT_OK;
}

extern "C" int swit_process(const swit_process_params *p, const float *offsets, const double *mute_win,
                            const double *taper, const int32_t *picks, float *data, void *work, size_t work_bytes,
                            hipStream_t stream) {
    g_err[0] = 0;
    if (const int rc = process_check(p)) return rc;
    if (!data || !work) return set_err(SWIT_EINVAL, "data and work are required");
    if (work_bytes < process_bytes(p)) return set_err(SWIT_EINVAL, "work needs %zu bytes", process_bytes(p));
    if ((p->mute_offset_short || p->mute_offset_long) && !offsets)
        return set_err(SWIT_EINVAL, "offset mutes need offsets");
    if (p->mute_late_arrival && !mute_win) return set_err(SWIT_EINVAL, "the window mute needs mute_win");
    if (p->detrend_taper && !taper) return set_err(SWIT_EINVAL, "detrend_taper needs taper");
    pr::ProcArgs a{};
    a.ntr = p->ntr;
    a.nt = p->nt;
    a.mute_short = p->mute_offset_short;
    a.mute_long = p->mute_offset_long;
    a.short_dist = p->mute_offset_short_dis;
    a.long_dist = p->mute_offset_long_dis;
    a.mute_arrival = p->mute_late_arrival ? (p->mute_late_window > 0.0 ? 2 : 1) : 0;
    a.pick_shift = p->mute_late_arrival ? std::ceil(std::fabs(p->mute_late_window) / p->dt) : 0.0;
    a.taps = p->filter_taps;
    a.pad = 3 * p->filter_taps;
    for (int k = 0; k < p->filter_taps; ++k) {
        a.b[k] = p->b[k];
        a.a[k] = p->a[k];
        a.zi[k] = k + 1 < p->filter_taps ? p->zi[k] : 0.0;
    }
    a.norm_trace = p->norm_trace;
    a.norm_event = p->norm_event;
    const int ntraces = p->ngathers * p->ntr;
    if (a.mute_short || a.mute_long || a.mute_arrival) {
        hipLaunchKernelGGL(pr::k_mute, dim3(ntraces), dim3(pr::kThreads), 0, stream, a, offsets, mute_win,
                           p->mute_late_arrival ? picks : nullptr, data);
        LAUNCH_CHECK();
    }
    if (a.taps) {
        double *y = static_cast<double *>(work);
        for (int t0 = 0; t0 < ntraces; t0 += kProcChunk) {
            const int n = std::min(kProcChunk, ntraces - t0);
            launch_filtfilt(a.taps, a, n, data + (size_t)t0 * p->nt, y, stream);
            LAUNCH_CHECK();
        }
    }
    if (p->detrend_taper) {
        hipLaunchKernelGGL(pr::k_detrend_taper, dim3(ntraces), dim3(pr::kThreads), 0, stream, a, taper, data);
        LAUNCH_CHECK();
    }
    if (a.norm_trace || a.norm_event) {
        float *enorm = static_cast<float *>(work);
        hipLaunchKernelGGL(pr::k_norm_trace, dim3(ntraces), dim3(pr::kThreads), 0, stream, a, data, enorm);
        LAUNCH_CHECK();
        if (a.norm_event) {
            hipLaunchKernelGGL(pr::k_norm_event, dim3(p->ngathers), dim3(pr::kThreads), 0, stream, a, enorm, data);
            LAUNCH_CHECK();
        }
    }
    return SWIT_OK;
}
