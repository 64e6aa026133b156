/*
 * CPU restatement of SWIT's 2-D acoustic staggered-grid solver (fd2dmpi), generic
 * over the arithmetic type.  TEST INFRASTRUCTURE ONLY: this file is the oracle the
 * HIP kernels are checked against; no product code path may call it.
 *
 * Included twice by fd2d_oracle.c with
 *     R  = float  / double
 *     FN(name) = name##_f32 / name##_f64
 *
 * Array convention: every 2-D field is stored with iz fastest, i.e. element (iz, ix)
 * (1-based, as in the Fortran) lives at [(ix-1)*ld + (iz-1)].  This is byte-identical
 * to Fortran's column-major (iz, ix) and to numpy (nx, nz) row-major.
 *
 * Index arithmetic (source/receiver cells) is always done in float32 exactly as the
 * reference does (real(4) division, truncating int()), so the FP32 and FP64 runs see
 * the same cells.
 */

#define IDX(iz, ix, ld) ((size_t)((ix) - 1) * (size_t)(ld) + (size_t)((iz) - 1))

/* ------------------------------------------------------------------------------
 * Model padding into the PML frame.
 * Follows fd2dmpi/io.f90:231-264 (readvelfile) and :268-300 (read_densityfile):
 * interior copied to [npml+1 : npml+nz, npml+1 : npml+nx]; left/right columns
 * replicate the edge interior column over the interior rows; then top/bottom rows
 * replicate the edge padded row over the full width (corners included).
 * in:  [nx][nz]   out: [nx+2npml][nz+2npml]
 */
void FN(orc_pad)(int nx, int nz, int npml, const R *in, R *out)
{
    const int nxp = nx + 2 * npml, nzp = nz + 2 * npml;
    for (int ix = 1; ix <= nx; ++ix)
        for (int iz = 1; iz <= nz; ++iz)
            out[IDX(npml + iz, npml + ix, nzp)] = in[IDX(iz, ix, nz)];
    for (int ix = 1; ix <= npml; ++ix)
        for (int iz = npml + 1; iz <= npml + nz; ++iz) {
            out[IDX(iz, ix, nzp)] = out[IDX(iz, npml + 1, nzp)];
            out[IDX(iz, nxp - npml + ix, nzp)] = out[IDX(iz, nxp - npml, nzp)];
        }
    for (int iz = 1; iz <= npml; ++iz)
        for (int ix = 1; ix <= nxp; ++ix) {
            out[IDX(iz, ix, nzp)] = out[IDX(npml + 1, ix, nzp)];
            out[IDX(nzp - npml + iz, ix, nzp)] = out[IDX(nzp - npml, ix, nzp)];
        }
}

/* ------------------------------------------------------------------------------
 * Minimum velocity for the PML profile.
 * Follows fd2dmpi/math.f90:342-361 (min_value_2d_topo) as called at
 * fd2dmpi/solver.f90:73 with the interior window of c and fs-npml: the free-surface
 * row is npml+1 everywhere (io.f90:777-800), so interior row 1 is skipped.
 * c_pad: [nxp][nzp]
 */
R FN(orc_vmin)(int nx, int nz, int npml, const R *c_pad)
{
    const int nzp = nz + 2 * npml;
    const int fs_rel = 1;            /* fs(j) - npml for every column */
    R v = (R)10000000000.0;
    for (int i = 1; i <= nz; ++i)
        for (int j = 1; j <= nx; ++j)
            if (i > fs_rel) {
                R x = c_pad[IDX(npml + i, npml + j, nzp)];
                if (v > x) v = x;
            }
    return v;
}

/* ------------------------------------------------------------------------------
 * Damping field.
 * Follows fd2dmpi/pml.f90:74-101 (setup_pml): quadratic profile
 *   a = (npml-1) dx, kappa = 3 vmin ln(1000) / (2a), damp(i) = kappa ((i-1) dx / a)^2,
 * written into full-height side columns, then overwritten in the top/bottom
 * trapezoids ix in [npml-i+2, nxp-npml+i-1].
 * damp: [nxp][nzp]
 */
void FN(orc_setup_pml)(int nx, int nz, int npml, R dx, R vmin, R *damp)
{
    const int nxp = nx + 2 * npml, nzp = nz + 2 * npml;
    R prof[1024];
    for (size_t k = 0; k < (size_t)nxp * nzp; ++k) damp[k] = (R)0;
    R a = (R)(npml - 1) * dx;
    R kappa = (R)3.0 * vmin * LOGR((R)1000.0) / ((R)2.0 * a);
    for (int i = 1; i <= npml; ++i) {
        R xa = (R)(i - 1) * dx / a;
        prof[i] = kappa * xa * xa;
    }
    for (int i = 1; i <= npml; ++i)
        for (int iz = 1; iz <= nzp; ++iz) {
            damp[IDX(iz, npml - i + 1, nzp)] = prof[i];
            damp[IDX(iz, nxp + i - npml, nzp)] = prof[i];
        }
    for (int iz = 1; iz <= npml; ++iz)
        for (int ix = 1 + (npml - (iz - 1)); ix <= nxp - (npml - (iz - 1)); ++ix) {
            damp[IDX(npml - iz + 1, ix, nzp)] = prof[iz];
            damp[IDX(nzp + iz - npml, ix, nzp)] = prof[iz];
        }
}

/* ------------------------------------------------------------------------------
 * One shot, forward job.
 * Follows fd2dmpi/fdcore.f90:30-198 (staggered42_modeling_is):
 *   p-sweep (:77-85), inject dt*s(it) (:88-89), p mirror (:92-94), u-sweep (:97-104),
 *   w-sweep (:106-113), w mirror (:116-120), receivers (:123-128), strips (:130-159),
 *   end state (:162-176).
 * c, den, damp: [nxp][nzp] padded.  s: [nt].  csg: [ng][nt] (csg[.,0] = 0).
 * boundary: [nt][9nx+6nz] (column it-1 holds step it; column 0 stays 0) or NULL.
 * p_end/u_end/w_end: [nx][nz] or NULL.  psnap: [nt][nx][nz] interior p after each
 * step (test hook) or NULL.
 */
int FN(orc_forward_shot)(int nx, int nz, int npml, int nt, int free_surface,
                         float dx, float dt,
                         const R *c, const R *den, const R *damp,
                         float xs, float zs, int ng, const float *xg, const float *zg,
                         const R *s, R *csg, R *boundary,
                         R *p_end, R *u_end, R *w_end, R *psnap)
{
    const int nxp = nx + 2 * npml, nzp = nz + 2 * npml;
    const int fs = npml + 1;
    const size_t S = (size_t)9 * nx + (size_t)6 * nz;
    const R c1 = (R)(9.0 / 8.0), c2 = (R)(-1.0 / 24.0);
    const R rdt = (R)dt, rdx = (R)dx;
    size_t n = (size_t)nxp * nzp;
    R *p = (R *)calloc(n, sizeof(R));
    R *u = (R *)calloc(n, sizeof(R));
    R *w = (R *)calloc(n, sizeof(R));
    int *igx = (int *)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    int *igz = (int *)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    if (!p || !u || !w || !igx || !igz) { free(p); free(u); free(w); free(igx); free(igz); return 1; }

    /* fdcore.f90:70-72 */
    int isx = npml + (int)(xs / dx) + 1;
    int isz = npml + (int)(zs / dx) + 1;
    if (isz == fs) isz = isz + 1;
    /* fdcore.f90:124-126 */
    for (int ig = 0; ig < ng; ++ig) {
        igx[ig] = npml + (int)(xg[ig] / dx) + 1;
        igz[ig] = npml + (int)(zg[ig] / dx) + 1;
        if (igz[ig] == fs) igz[ig] = igz[ig] + 1;
    }
    for (int ig = 0; ig < ng; ++ig) csg[(size_t)ig * nt] = (R)0;
    if (boundary) for (size_t k = 0; k < S; ++k) boundary[k] = (R)0;
    if (psnap) for (size_t k = 0; k < (size_t)nx * nz; ++k) psnap[k] = (R)0;

#define P(iz, ix) p[IDX(iz, ix, nzp)]
#define U(iz, ix) u[IDX(iz, ix, nzp)]
#define W(iz, ix) w[IDX(iz, ix, nzp)]
#define CC(iz, ix) c[IDX(iz, ix, nzp)]
#define DEN(iz, ix) den[IDX(iz, ix, nzp)]
#define DMP(iz, ix) damp[IDX(iz, ix, nzp)]

    for (int it = 2; it <= nt; ++it) {
        for (int ix = 3; ix <= nxp - 2; ++ix)
            for (int iz = fs + 1; iz <= nzp - 2; ++iz) {
                R alpha = DEN(iz, ix) * CC(iz, ix) * CC(iz, ix) * rdt / rdx;
                R kappa = DMP(iz, ix) * rdt;
                P(iz, ix) = ((R)1.0 - kappa) * P(iz, ix) - alpha *
                    (c1 * (U(iz, ix) - U(iz, ix - 1) + W(iz, ix) - W(iz - 1, ix)) +
                     c2 * (U(iz, ix + 1) - U(iz, ix - 2) + W(iz + 1, ix) - W(iz - 2, ix)));
            }
        P(isz, isx) = P(isz, isx) + rdt * s[it - 1];
        if (free_surface == 1)
            for (int ix = 1; ix <= nxp; ++ix) P(fs - 1, ix) = -P(fs + 1, ix);
        for (int ix = 2; ix <= nxp - 2; ++ix)
            for (int iz = fs; iz <= nzp - 1; ++iz) {
                R alpha = rdt / (DEN(iz, ix) * rdx);
                R kappa = (R)0.5 * (DMP(iz, ix) + DMP(iz, ix + 1)) * rdt;
                U(iz, ix) = ((R)1.0 - kappa) * U(iz, ix) - alpha *
                    (c1 * (P(iz, ix + 1) - P(iz, ix)) + c2 * (P(iz, ix + 2) - P(iz, ix - 1)));
            }
        for (int ix = 2; ix <= nxp - 1; ++ix)
            for (int iz = fs; iz <= nzp - 2; ++iz) {
                R alpha = rdt / (DEN(iz, ix) * rdx);
                R kappa = (R)0.5 * (DMP(iz, ix) + DMP(iz + 1, ix)) * rdt;
                W(iz, ix) = ((R)1.0 - kappa) * W(iz, ix) - alpha *
                    (c1 * (P(iz + 1, ix) - P(iz, ix)) + c2 * (P(iz + 2, ix) - P(iz - 1, ix)));
            }
        if (free_surface == 1)
            for (int ix = 1; ix <= nxp; ++ix) W(fs - 1, ix) = W(fs, ix);
        for (int ig = 0; ig < ng; ++ig)
            csg[(size_t)ig * nt + (it - 1)] = P(igz[ig], igx[ig]);
        if (boundary) {
            R *b = boundary + (size_t)(it - 1) * S;
            size_t jj = 0;
            for (int j = npml + 1; j <= nxp - npml; ++j) {
                b[jj + 0] = P(nzp - npml, j);
                b[jj + 1] = P(nzp - npml - 1, j);
                b[jj + 2] = W(nzp - npml - 1, j);
                jj += 3;
            }
            for (int j = npml + 1; j <= nzp - npml; ++j) {
                b[jj + 0] = P(j, npml + 1);
                b[jj + 1] = P(j, npml + 2);
                b[jj + 2] = U(j, npml + 1);
                b[jj + 3] = P(j, nxp - npml);
                b[jj + 4] = P(j, nxp - npml - 1);
                b[jj + 5] = U(j, nxp - npml - 1);
                jj += 6;
            }
            for (int j = npml + 1; j <= nxp - npml; ++j) {
                b[jj + 0] = P(fs, j);
                b[jj + 1] = P(fs + 1, j);
                b[jj + 2] = P(fs + 2, j);
                b[jj + 3] = U(fs, j);
                b[jj + 4] = U(fs + 1, j);
                b[jj + 5] = W(fs, j);
                jj += 6;
            }
        }
        if (psnap) {
            R *q = psnap + (size_t)(it - 1) * nx * nz;
            for (int ix = 1; ix <= nx; ++ix)
                for (int iz = 1; iz <= nz; ++iz) q[IDX(iz, ix, nz)] = P(npml + iz, npml + ix);
        }
    }
    if (p_end)
        for (int ix = 1; ix <= nx; ++ix)
            for (int iz = 1; iz <= nz; ++iz) {
                p_end[IDX(iz, ix, nz)] = P(npml + iz, npml + ix);
                u_end[IDX(iz, ix, nz)] = U(npml + iz, npml + ix);
                w_end[IDX(iz, ix, nz)] = W(npml + iz, npml + ix);
            }
    free(p); free(u); free(w); free(igx); free(igz);
    return 0;
}

/* ------------------------------------------------------------------------------
 * Reverse-time reconstruction of one step on the interior window.
 * Follows fd2dmpi/fdcore.f90:202-289 (staggered42_reco_it) with fs(ix)-npml = 1.
 * ci/deni: interior [nx][nz] sections of c and den (solver.f90:234).
 * b: strip column it-1, i.e. boundary + (it-2)*S.  s: [nt]; the reference reads
 * s(it+1), which is past the end at it = nt: that value is taken as 0 here.
 */
static void FN(orc_reco_step)(int nx, int nz, int it, int nt, R rdx, R rdt,
                              int isx, int isz, const R *ci, const R *deni,
                              const R *b, const R *s, R *p, R *u, R *w)
{
    const R c1 = (R)(9.0 / 8.0), c2 = (R)(-1.0 / 24.0);
#define IP(iz, ix) p[IDX(iz, ix, nz)]
#define IU(iz, ix) u[IDX(iz, ix, nz)]
#define IW(iz, ix) w[IDX(iz, ix, nz)]
#define IC(iz, ix) ci[IDX(iz, ix, nz)]
#define ID(iz, ix) deni[IDX(iz, ix, nz)]
    for (int ix = 2; ix <= nx - 2; ++ix)
        for (int iz = 2; iz <= nz - 1; ++iz) {
            R alpha = rdt / (ID(iz, ix) * rdx);
            IU(iz, ix) = IU(iz, ix) + alpha *
                (c1 * (IP(iz, ix + 1) - IP(iz, ix)) + c2 * (IP(iz, ix + 2) - IP(iz, ix - 1)));
        }
    for (int ix = 2; ix <= nx - 1; ++ix)
        for (int iz = 2; iz <= nz - 2; ++iz) {
            R alpha = rdt / (ID(iz, ix) * rdx);
            IW(iz, ix) = IW(iz, ix) + alpha *
                (c1 * (IP(iz + 1, ix) - IP(iz, ix)) + c2 * (IP(iz + 2, ix) - IP(iz - 1, ix)));
        }
    size_t jj = 0;
    for (int j = 1; j <= nx; ++j) { IW(nz - 1, j) = b[jj + 2]; jj += 3; }
    for (int j = 1; j <= nz; ++j) { IU(j, 1) = b[jj + 2]; IU(j, nx - 1) = b[jj + 5]; jj += 6; }
    if (isx > 2 && isx < nx - 1 && isz > 3 && isz < nz - 1) {
        R sv = (it + 1 <= nt) ? s[it] : (R)0;   /* s(it+1) */
        /* isz > fs-npml = 1 always holds here */
        IP(isz, isx) = IP(isz, isx) - rdt * sv;
    }
    for (int ix = 3; ix <= nx - 2; ++ix)
        for (int iz = 3; iz <= nz - 2; ++iz) {
            R alpha = ID(iz, ix) * IC(iz, ix) * IC(iz, ix) * rdt / rdx;
            IP(iz, ix) = IP(iz, ix) + alpha *
                (c1 * (IU(iz, ix) - IU(iz, ix - 1) + IW(iz, ix) - IW(iz - 1, ix)) +
                 c2 * (IU(iz, ix + 1) - IU(iz, ix - 2) + IW(iz + 1, ix) - IW(iz - 2, ix)));
        }
    jj = 0;
    for (int j = 1; j <= nx; ++j) { IP(nz, j) = b[jj + 0]; IP(nz - 1, j) = b[jj + 1]; jj += 3; }
    for (int j = 1; j <= nz; ++j) {
        IP(j, 1) = b[jj + 0]; IP(j, 2) = b[jj + 1];
        IP(j, nx) = b[jj + 3]; IP(j, nx - 1) = b[jj + 4];
        jj += 6;
    }
    for (int j = 1; j <= nx; ++j) {
        IP(1, j) = b[jj + 0]; IP(2, j) = b[jj + 1]; IP(3, j) = b[jj + 2];
        jj += 6;
    }
#undef IP
#undef IU
#undef IW
#undef IC
#undef ID
}

/* ------------------------------------------------------------------------------
 * One adjoint step on the padded grid.
 * Follows fd2dmpi/fdcore.f90:293-375 (staggered42_back_it): u/w sweeps with
 * coefficient den*c^2*dt/dx (:310-326), w mirror (:329-333), p sweep with
 * dt/(den*dx) over rows 3..nzp-2 (:336-345), adjoint injection at the receiver cells
 * (:348-367), unconditional p mirror (:370-373).
 * sadj: [ng][nt]
 */
static void FN(orc_back_step)(int nxp, int nzp, int fs, int it, int nt, int free_surface,
                              R rdx, R rdt, const R *c, const R *den, const R *damp,
                              int ng, const int *igx, const int *igz, const R *sadj,
                              R *p, R *u, R *w)
{
    const R c1 = (R)(9.0 / 8.0), c2 = (R)(-1.0 / 24.0);
    for (int ix = 2; ix <= nxp - 2; ++ix)
        for (int iz = 2; iz <= nzp - 1; ++iz) {
            R alpha = DEN(iz, ix) * CC(iz, ix) * CC(iz, ix) * rdt / rdx;
            R kappa = (R)0.5 * (DMP(iz, ix) + DMP(iz, ix + 1)) * rdt;
            U(iz, ix) = ((R)1.0 - kappa) * U(iz, ix) - alpha *
                (c1 * (P(iz, ix + 1) - P(iz, ix)) + c2 * (P(iz, ix + 2) - P(iz, ix - 1)));
        }
    for (int ix = 2; ix <= nxp - 1; ++ix)
        for (int iz = 2; iz <= nzp - 2; ++iz) {
            R alpha = DEN(iz, ix) * CC(iz, ix) * CC(iz, ix) * rdt / rdx;
            R kappa = (R)0.5 * (DMP(iz, ix) + DMP(iz + 1, ix)) * rdt;
            W(iz, ix) = ((R)1.0 - kappa) * W(iz, ix) - alpha *
                (c1 * (P(iz + 1, ix) - P(iz, ix)) + c2 * (P(iz + 2, ix) - P(iz - 1, ix)));
        }
    if (free_surface == 1)
        for (int ix = 2; ix <= nxp - 1; ++ix) W(fs - 1, ix) = W(fs, ix);
    for (int ix = 3; ix <= nxp - 2; ++ix)
        for (int iz = 3; iz <= nzp - 2; ++iz) {
            R alpha = rdt / (DEN(iz, ix) * rdx);
            R kappa = DMP(iz, ix) * rdt;
            P(iz, ix) = ((R)1.0 - kappa) * P(iz, ix) - alpha *
                (c1 * (U(iz, ix) - U(iz, ix - 1) + W(iz, ix) - W(iz - 1, ix)) +
                 c2 * (U(iz, ix + 1) - U(iz, ix - 2) + W(iz + 1, ix) - W(iz - 2, ix)));
        }
    for (int ig = 0; ig < ng; ++ig) {
        R v = rdt * sadj[(size_t)ig * nt + (it - 1)];
        if (igz[ig] > fs) P(igz[ig], igx[ig]) = P(igz[ig], igx[ig]) + v;
        else P(fs + 1, igx[ig]) = P(fs + 1, igx[ig]) + v;
    }
    for (int ix = 1; ix <= nxp; ++ix) P(fs - 1, ix) = -P(fs + 1, ix);
}

/* ------------------------------------------------------------------------------
 * One shot, adjoint job.
 * Follows fd2dmpi/solver.f90:194-259: per-shot zeroing of state and accumulators
 * (:197-210), then for it = nt..2: p0_bk = p_bk (:232), reconstruction (:234),
 * adjoint step (:235), imaging dg += p_rec (p0_bk - p_bk), dfw += p_rec^2,
 * dbk += p_bk^2 on the interior (:237-239).
 * p_end/u_end/w_end are consumed (overwritten by the reconstruction).
 * dg/dfw/dbk: [nx][nz] outputs (zeroed here).  adj_rec: [ng][nt] adjoint pressure at
 * the receiver cells after each step (the dev solver's adjoint record,
 * dev/solver/fdsolver.f90:204-217) or NULL.  prec: [nt][nx][nz] reconstructed p
 * (time it-1 stored at slot it-2) or NULL.
 */
int FN(orc_adjoint_shot)(int nx, int nz, int npml, int nt, int free_surface,
                         float dx, float dt,
                         const R *c, const R *den, const R *damp,
                         float xs, float zs, int ng, const float *xg, const float *zg,
                         const R *s, const R *sadj, const R *boundary,
                         R *p_end, R *u_end, R *w_end,
                         R *dg, R *dfw, R *dbk, R *adj_rec, R *prec)
{
    const int nxp = nx + 2 * npml, nzp = nz + 2 * npml;
    const int fs = npml + 1;
    const size_t S = (size_t)9 * nx + (size_t)6 * nz;
    const R rdt = (R)dt, rdx = (R)dx;
    size_t n = (size_t)nxp * nzp, ni = (size_t)nx * nz;
    R *p = (R *)calloc(n, sizeof(R));
    R *u = (R *)calloc(n, sizeof(R));
    R *w = (R *)calloc(n, sizeof(R));
    R *p0 = (R *)calloc(ni, sizeof(R));
    R *ci = (R *)malloc(sizeof(R) * ni);
    R *di = (R *)malloc(sizeof(R) * ni);
    int *igx = (int *)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    int *igz = (int *)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    if (!p || !u || !w || !p0 || !ci || !di || !igx || !igz) {
        free(p); free(u); free(w); free(p0); free(ci); free(di); free(igx); free(igz);
        return 1;
    }
    for (int ix = 1; ix <= nx; ++ix)
        for (int iz = 1; iz <= nz; ++iz) {
            ci[IDX(iz, ix, nz)] = CC(npml + iz, npml + ix);
            di[IDX(iz, ix, nz)] = DEN(npml + iz, npml + ix);
        }
    /* fdcore.f90:216-218: interior source cell, no free-surface push-down */
    int isx = (int)(xs / dx) + 1;
    int isz = (int)(zs / dx) + 1;
    /* fdcore.f90:350-352: receiver cells without push-down; injection handles fs */
    for (int ig = 0; ig < ng; ++ig) {
        igx[ig] = npml + (int)(xg[ig] / dx) + 1;
        igz[ig] = npml + (int)(zg[ig] / dx) + 1;
    }
    for (size_t k = 0; k < ni; ++k) { dg[k] = (R)0; dfw[k] = (R)0; dbk[k] = (R)0; }
    if (adj_rec) for (size_t k = 0; k < (size_t)ng * nt; ++k) adj_rec[k] = (R)0;

    for (int it = nt; it >= 2; --it) {
        for (int ix = 1; ix <= nx; ++ix)
            for (int iz = 1; iz <= nz; ++iz) p0[IDX(iz, ix, nz)] = P(npml + iz, npml + ix);
        FN(orc_reco_step)(nx, nz, it, nt, rdx, rdt, isx, isz, ci, di,
                          boundary + (size_t)(it - 2) * S, s, p_end, u_end, w_end);
        FN(orc_back_step)(nxp, nzp, fs, it, nt, free_surface, rdx, rdt, c, den, damp,
                          ng, igx, igz, sadj, p, u, w);
        for (int ix = 1; ix <= nx; ++ix)
            for (int iz = 1; iz <= nz; ++iz) {
                size_t k = IDX(iz, ix, nz);
                R pr = p_end[k];
                R pb = P(npml + iz, npml + ix);
                dg[k] = dg[k] + pr * (p0[k] - pb);
                dfw[k] = dfw[k] + pr * pr;
                dbk[k] = dbk[k] + pb * pb;
            }
        if (adj_rec)
            for (int ig = 0; ig < ng; ++ig) {
                int gz = igz[ig] == fs ? igz[ig] + 1 : igz[ig];
                adj_rec[(size_t)ig * nt + (it - 1)] = P(gz, igx[ig]);
            }
        if (prec) {
            R *q = prec + (size_t)(it - 2) * ni;
            for (size_t k = 0; k < ni; ++k) q[k] = p_end[k];
        }
    }
    free(p); free(u); free(w); free(p0); free(ci); free(di); free(igx); free(igz);
    return 0;
}

/* ------------------------------------------------------------------------------
 * Final scaling of the summed kernel: g = -2 g / c on the interior.
 * Follows fd2dmpi/solver.f90:270.  c_pad: [nxp][nzp]; g: [nx][nz] in place.
 */
void FN(orc_scale_gradient)(int nx, int nz, int npml, const R *c_pad, R *g)
{
    const int nzp = nz + 2 * npml;
    for (int ix = 1; ix <= nx; ++ix)
        for (int iz = 1; iz <= nz; ++iz) {
            size_t k = IDX(iz, ix, nz);
            g[k] = (R)(-2.0) * g[k] / c_pad[IDX(npml + iz, npml + ix, nzp)];
        }
}

#undef P
#undef U
#undef W
#undef CC
#undef DEN
#undef DMP
