// Shared device-side definitions of the gfx950 2-D acoustic solver.
//
// Index convention: (iz, ix) are 1-based like the reference Fortran; a padded field
// element (iz, ix) lives at [(ix-1)*ld + (iz-1)], iz fastest (coalesced along iz).
// All per-cell arithmetic below keeps the reference's FP32 operation order; the
// library is compiled with -ffp-contract=off so no FMA changes a rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swit {

// Staggered 4th-order coefficients, fd2dmpi/global.f90:34-35.
constexpr float kC1 = 9.0f / 8.0f;
constexpr float kC2 = (float)(-1.0 / 24.0);
// ln(1000) rounded to float32 (0x1.ba18aap+2), the value of log(1000.0) in
// fd2dmpi/pml.f90:82.
constexpr float kLn1000 = 0x1.ba18aap+2f;

// Sizes every kernel needs, passed by value.
struct Dims {
    int nx, nz, npml;    // interior grid and PML width
    int nxp, nzp;        // padded grid (pml.f90:24-38)
    int ld;              // leading dimension of padded fields (>= nzp)
    int fs;              // free-surface row, npml+1 (io.f90:792-796)
    int nt;
    int S;               // strip floats per step, 9nx+6nz
    int ngmax;
    int free_surface;
    float dt, dx;
};

__device__ __forceinline__ size_t pidx(const Dims &d, int iz, int ix) {
    return (size_t)(ix - 1) * d.ld + (iz - 1);
}
__device__ __forceinline__ size_t iidx(const Dims &d, int iz, int ix) {
    return (size_t)(ix - 1) * d.nz + (iz - 1);
}

// Damping at padded cell (iz, ix): side columns over the full height, then the
// top/bottom trapezoids overwrite (fd2dmpi/pml.f90:88-99).  prof[1..npml].
__device__ __forceinline__ float damp_at(const Dims &d, const float *prof, int iz, int ix) {
    const int npml = d.npml;
    float v = 0.0f;
    if (ix <= npml) v = prof[npml - ix + 1];
    else if (ix >= d.nxp - npml + 1) v = prof[ix - (d.nxp - npml)];
    if (iz <= npml) {
        int i = npml - iz + 1;
        if (ix >= npml - i + 2 && ix <= d.nxp - npml + i - 1) v = prof[i];
    } else if (iz >= d.nzp - npml + 1) {
        int i = iz - (d.nzp - npml);
        if (ix >= npml - i + 2 && ix <= d.nxp - npml + i - 1) v = prof[i];
    }
    return v;
}

// ---- per-cell updates, reference operation order --------------------------------

// Pressure update (fdcore.f90:79-84 forward with alpha = den c^2 dt/dx; the adjoint's
// :339-344 passes alpha = dt/(den dx)).
__device__ __forceinline__ float p_update(float p, float alpha, float kappa,
                                          float u0, float um1, float up1, float um2,
                                          float w0, float wm1, float wp1, float wm2) {
    return (1.0f - kappa) * p - alpha * (kC1 * (u0 - um1 + w0 - wm1) + kC2 * (up1 - um2 + wp1 - wm2));
}

// Velocity update (fdcore.f90:99-103 / :108-112; adjoint :312-316 / :321-325):
// v = (1-kappa) v - alpha (c1 (q1 - q0) + c2 (q2 - qm1)).
__device__ __forceinline__ float v_update(float v, float alpha, float kappa,
                                          float q0, float q1, float q2, float qm1) {
    return (1.0f - kappa) * v - alpha * (kC1 * (q1 - q0) + kC2 * (q2 - qm1));
}

// The same updates with the damping factor (1 - kappa) precomputed: the reference
// rounds kappa, then 1 - kappa, then multiplies, so storing 1 - kappa keeps every bit.
// Divergence term of the pressure update: c1 (u0-um1+w0-wm1) + c2 (up1-um2+wp1-wm2).
__device__ __forceinline__ float div_term(float u0, float um1, float up1, float um2,
                                          float w0, float wm1, float wp1, float wm2) {
    return kC1 * (u0 - um1 + w0 - wm1) + kC2 * (up1 - um2 + wp1 - wm2);
}
__device__ __forceinline__ float p_update_f(float p, float alpha, float fac,
                                            float u0, float um1, float up1, float um2,
                                            float w0, float wm1, float wp1, float wm2) {
    return fac * p - alpha * div_term(u0, um1, up1, um2, w0, wm1, wp1, wm2);
}
__device__ __forceinline__ float v_update_f(float v, float alpha, float fac,
                                            float q0, float q1, float q2, float qm1) {
    return fac * v - alpha * (kC1 * (q1 - q0) + kC2 * (q2 - qm1));
}

// Reconstruction velocity undo (fdcore.f90:223-226 / :231-234).
__device__ __forceinline__ float v_undo(float v, float alpha, float q0, float q1, float q2, float qm1) {
    return v + alpha * (kC1 * (q1 - q0) + kC2 * (q2 - qm1));
}

// Reconstruction pressure undo (fdcore.f90:262-266).
__device__ __forceinline__ float p_undo(float p, float alpha,
                                        float u0, float um1, float up1, float um2,
                                        float w0, float wm1, float wp1, float wm2) {
    return p + alpha * (kC1 * (u0 - um1 + w0 - wm1) + kC2 * (up1 - um2 + wp1 - wm2));
}

// A receiver's index and cell (1-based padded row, column).
struct RecRef {
    int idx;
    int iz, ix;
    int pad;
};

// Injection group: the receivers (of one tile, or of the whole grid) that share one
// cell, in receiver order.
struct InjGroup {
    int e0, e1;     // entry range in the shot's injection list
};

}  // namespace swit
