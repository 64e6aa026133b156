"""Pins of the CPU oracle (C restatement of fd2dmpi) against what the reference's own
files and the mathematics fix, so that a plausible mistake (a dropped term, a wrong
sign, index or time level) fails one of them.

  * dot-product known answer of the reference (dev/test/02_adjoint_test.ipynb:276):
    a = 1.25501e35, ratio 1.00000, rerr 0.00000;
  * closed forms: PML padding (io.f90:251-258), vmin (math.f90:342-361), quadratic
    damping profile and its corner rule (pml.f90:80-99);
  * reciprocity of the forward operator with damping off (pressure source / pressure
    receiver: G(r,s) K_s = G(s,r) K_r for the staggered scheme with G = -D^T);
  * reverse-time reconstruction reproduces the forward pressure at every step
    (fdcore.f90:202-289) and the forward illumination equals sum_t p^2;
  * receiver / source cell rules (fdcore.f90:70-72,124-126);
  * gradient kernel (solver.f90:237-270): proportional to dt, invariant to a uniform
    density scale, same sign as and correlated with the exact discrete gradient
    (torch autograd through a vectorised restatement checked against the oracle);
    with the reconstruction exact and no boundary in reach, <g, dc> equals dt times
    the finite-difference derivative (fixes the -2 and the 1/c of solver.f90:270);
  * adjoint coefficient placement (fdcore.f90:310-345 against :77-113): in a
    heterogeneous medium the adjoint step is the forward scheme of the medium
    (c, 1/(den c^2)), so its record equals that forward run's seismograms;
  * one-step impulse response against the textbook 4th-order staggered stencil
    (c1 = 9/8, c2 = -1/24) and B = dt/(rho dx);
  * homogeneous-medium pressure against the exact 2-D Green's-function solution
    (arrival, waveform and amplitude).
"""
import numpy as np
import pytest
from scipy.integrate import cumulative_trapezoid

from oracle import oracle as O


def _src(nt, dt, f0, amp=1.0):
    return cumulative_trapezoid(O.ricker(nt, dt, f0, amp), initial=0)


# ---------------------------------------------------------------------------------
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_dot_product_known_answer(precision):
    """Reference notebook dev/test/02_adjoint_test.ipynb: homogeneous 4000 m/s, Gardner
    density, 401x401, dx 10, dt 1 ms, nt 1001, npml 60, 5 Hz Ricker x 1e20, source
    (2000, 2000) m, receivers (2000, 2000) and (3000, 2000) m.  The notebook prints
    'a = 1.25501e+35, b = 1.25501e+35, ratio = 1.00000e+00, rerr = 0.00000' (:276)."""
    nx = nz = 401
    npml, dx, dt, nt = 60, 10.0, 0.001, 1001
    vp = np.full((nx, nz), 4000.0, np.float32)
    rho = (np.power(vp.astype(np.float64), 0.25) * 310).astype(np.float32)
    src = _src(nt, dt, 5.0, 1e20)
    src32 = src.astype(np.float32)[None]
    sxz = np.array([[2000, 2000]], np.float32)
    rxz = np.array([[[2000, 2000], [3000, 2000]]], np.float32)
    m = O.Model(vp, rho, dx, npml, precision)
    f = O.forward(m, nt, dt, sxz, rxz, src32)
    obs = f["seis"][0].astype(np.float32)
    adj = np.zeros((1, 2, nt), np.float32)
    adj[0, 1] = obs[1]
    a_ = O.adjoint(m, nt, dt, sxz, rxz, src32, adj, f, record=True)
    rec = a_["adj_rec"][0].astype(np.float32)
    a = float(np.inner(obs[1], adj[0, 1]))
    b = float(np.inner(src, rec[0]))
    assert f"{a:2.5e}" == "1.25501e+35"
    assert f"{b:2.5e}" == "1.25501e+35"
    assert f"{a / b:2.5e}" == "1.00000e+00"
    assert f"{(a - b) / (a + b):2.5f}" in ("0.00000", "-0.00000")


# ---------------------------------------------------------------------------------
def test_padding_is_edge_replication():
    rng = np.random.default_rng(0)
    vp = rng.uniform(1500, 4000, (23, 17)).astype(np.float32)
    m = O.Model(vp, vp, 10.0, 7, "f64")
    np.testing.assert_array_equal(m.c, np.pad(vp.astype(np.float64), 7, mode="edge"))


def test_vmin_skips_first_interior_row():
    rng = np.random.default_rng(1)
    vp = rng.uniform(2000, 4000, (20, 15)).astype(np.float32)
    vp[5, 0] = 1000.0          # interior row 1: ignored
    vp[7, 9] = 1700.0
    m = O.Model(vp, vp, 10.0, 5, "f64")
    assert m.vmin == pytest.approx(1700.0)


@pytest.mark.parametrize("nx,nz,npml", [(30, 20, 8), (12, 40, 15), (25, 25, 3)])
def test_damping_closed_form(nx, nz, npml):
    """damp(iz, ix) = kappa0 ((i-1) dx / a)^2 with i the depth into the absorbing layer;
    corners take the larger of the row and column depth (the trapezoid overwrite of
    pml.f90:94-99), and the profile is mirror-symmetric."""
    dx = 12.5
    vp = np.full((nx, nz), 3000.0, np.float32)
    m = O.Model(vp, vp, dx, npml, "f64")
    nxp, nzp = nx + 2 * npml, nz + 2 * npml
    a = (npml - 1) * dx
    k0 = 3.0 * 3000.0 * np.log(1000.0) / (2.0 * a)
    prof = np.zeros(npml + 1)
    prof[1:] = k0 * ((np.arange(1, npml + 1) - 1) * dx / a) ** 2
    ix = np.arange(1, nxp + 1)
    iz = np.arange(1, nzp + 1)
    dxi = np.where(ix <= npml, npml - ix + 1, np.where(ix >= nxp - npml + 1, ix - (nxp - npml), 0))
    dzi = np.where(iz <= npml, npml - iz + 1, np.where(iz >= nzp - npml + 1, iz - (nzp - npml), 0))
    expect = prof[np.maximum(dxi[:, None], dzi[None, :])]
    np.testing.assert_allclose(m.damp, expect, rtol=1e-12, atol=0)
    np.testing.assert_array_equal(m.damp, m.damp[::-1, :])
    np.testing.assert_array_equal(m.damp, m.damp[:, ::-1])
    assert np.all(m.damp[npml:nxp - npml, npml:nzp - npml] == 0)


# ---------------------------------------------------------------------------------
def _hetero(nx, nz, seed):
    rng = np.random.default_rng(seed)
    from scipy.ndimage import gaussian_filter
    vp = 2500 + 600 * gaussian_filter(rng.standard_normal((nx, nz)), 3)
    vp = np.clip(vp, 1800, 3400).astype(np.float32)
    rho = (310 * np.power(vp.astype(np.float64), 0.25) * (1 + 0.1 * gaussian_filter(
        rng.standard_normal((nx, nz)), 2))).astype(np.float32)
    return vp, rho


@pytest.mark.parametrize("damped", [False, True])
def test_reciprocity(damped):
    """The scheme is a leapfrog of a skew-adjoint operator (the velocity stencil is
    minus the transpose of the divergence stencil, also across the free-surface mirrors)
    plus diagonal damping, so a pressure injection at s recorded at r obeys
    G(r,s) K_s = G(s,r) K_r with K = den c^2 dt/dx, in a heterogeneous medium."""
    nx, nz, npml, nt, dx, dt = 50, 36, 6, 260, 10.0, 0.001
    vp, rho = _hetero(nx, nz, 3)
    nxp, nzp = nx + 2 * npml, nz + 2 * npml
    m = O.Model(vp, rho, dx, npml, "f64", damp_override=None if damped else np.zeros((nxp, nzp)))
    src = _src(nt, dt, 20.0)[None]
    pa = np.array([12 * dx, 9 * dx], np.float32)
    pb = np.array([37 * dx, 21 * dx], np.float32)
    fa = O.forward(m, nt, dt, pa[None], pb[None, None], src, save_strips=False)["seis"][0, 0]
    fb = O.forward(m, nt, dt, pb[None], pa[None, None], src, save_strips=False)["seis"][0, 0]

    def K(p):
        ix, iz = int(p[0] / dx), int(p[1] / dx)
        return float(rho[ix, iz]) * float(vp[ix, iz]) ** 2
    lhs, rhs = fa * K(pa), fb * K(pb)
    assert np.abs(lhs).max() > 0
    assert np.abs(lhs - rhs).max() <= 1e-10 * np.abs(lhs).max()
    # the weighting matters: the plain traces differ by K_r / K_s
    assert np.abs(fa - fb).max() > 1e-2 * np.abs(fa).max()


# ---------------------------------------------------------------------------------
def _recon_case(src_depth_cells, nt=300):
    nx, nz, npml, dx, dt = 48, 32, 8, 10.0, 0.001
    vp, rho = _hetero(nx, nz, 4)
    m = O.Model(vp, rho, dx, npml, "f64")
    src = _src(nt, dt, 18.0)[None]
    sxz = np.array([[20 * dx, src_depth_cells * dx]], np.float32)
    rxz = np.stack([np.arange(nx) * dx, np.zeros(nx)], 1)[None].astype(np.float32)
    f = O.forward(m, nt, dt, sxz, rxz, src, psnap=True)
    adj = np.zeros((1, nx, nt))
    a = O.adjoint(m, nt, dt, sxz, rxz, src, adj, f, prec_snap=True)
    return f, a, nt


def test_reconstruction_reproduces_forward_pressure():
    """fdcore.f90:202-289 undoes one forward step from the saved strips: for a source
    in the top three interior rows the reconstructed pressure at time it-1 equals the
    forward pressure at it-1 for every step (FP64 round-off)."""
    f, a, nt = _recon_case(2)
    fwd = f["psnap"][0]          # slot k: time k+1
    rec = a["prec"][0]           # slot it-2: time it-1
    scale = np.abs(fwd).max()
    err = max(np.abs(rec[k - 1] - fwd[k - 1]).max() for k in range(1, nt))
    assert err <= 1e-9 * scale
    # forward illumination (solver.f90:238) = sum over reconstructed times 1..nt-1
    np.testing.assert_allclose(a["illum_forw"], (fwd[: nt - 1] ** 2).sum(0), rtol=1e-8,
                               atol=1e-12 * (fwd ** 2).sum(0).max())


def test_reconstruction_deep_source_reference_offset():
    """For a source below the third interior row the reference subtracts s(it+1)
    instead of s(it) (fdcore.f90:250-257): the reconstruction then departs from the
    forward pressure by about dt (s(it+1) - s(it)) at the source cell.  The oracle
    keeps that behaviour."""
    f, a, nt = _recon_case(10)
    fwd = f["psnap"][0]
    rec = a["prec"][0]
    err = max(np.abs(rec[k - 1] - fwd[k - 1]).max() for k in range(1, nt))
    assert err > 1e-4 * np.abs(fwd).max()


def test_receiver_push_down_and_truncation():
    """Receivers at z = 0 record one cell below the free surface, the same cell as
    z = dx (fdcore.f90:124-126); coordinates truncate to the cell below."""
    nx, nz, npml, nt, dx, dt = 30, 20, 5, 120, 10.0, 0.001
    vp, rho = _hetero(nx, nz, 5)
    m = O.Model(vp, rho, dx, npml, "f64")
    rxz = np.array([[[100, 0], [100, 10], [100, 9.9], [109.9, 10], [100, 20]]], np.float32)
    s = O.forward(m, nt, dt, np.array([[50, 30]], np.float32), rxz, _src(nt, dt, 25.0)[None],
                  save_strips=False)["seis"][0]
    np.testing.assert_array_equal(s[0], s[1])      # z 0 and dx share fs+1
    np.testing.assert_array_equal(s[2], s[0])      # 9.9 m truncates to row 0 -> pushed
    np.testing.assert_array_equal(s[3], s[1])      # 109.9 m truncates to column 10
    assert np.abs(s[4] - s[1]).max() > 0


# ---------------------------------------------------------------------------------
def _kernel(c, rho, dt, nt, fd_eps=None):
    """Reference kernel for 1/2 |d - d_obs|^2 (d_obs from c + 20 dc) on a 60 x 40 grid,
    optionally with the central finite difference of the misfit along dc."""
    nx, nz = c.shape
    npml, dx = 10, 10.0
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    dc = np.exp(-((X - nx // 2) ** 2 + (Z - 25) ** 2) / (2 * 3.0 ** 2))
    src = _src(nt, dt, 15.0)[None].astype(np.float32)
    sxz = np.array([[15 * dx, 5 * dx]], np.float32)
    rxz = np.stack([np.arange(nx) * dx, np.full(nx, 2 * dx)], 1)[None].astype(np.float32)

    def seis(cc, save=False):
        m = O.Model(cc, rho, dx, npml, "f64")
        return m, O.forward(m, nt, dt, sxz, rxz, src, save_strips=save)
    dobs = seis(c + 20 * dc)[1]["seis"]
    m, f = seis(c, True)
    r = f["seis"] - dobs
    g = O.adjoint(m, nt, dt, sxz, rxz, src, r, f)["grad"]
    fd = None
    if fd_eps:
        jp = 0.5 * np.sum((seis(c + fd_eps * dc)[1]["seis"] - dobs) ** 2)
        jm = 0.5 * np.sum((seis(c - fd_eps * dc)[1]["seis"] - dobs) ** 2)
        fd = (jp - jm) / (2 * fd_eps)
    return g, dc, r, fd


def test_gradient_kernel_scales_with_dt_and_ignores_density_scale():
    """The adjoint field carries the injection weight dt (fdcore.f90:353) where the
    exact adjoint carries 1, so kernel / (finite-difference derivative) is proportional
    to dt; a uniform density factor cancels between den c^2 and 1/den."""
    nx, nz = 60, 40
    c = np.full((nx, nz), 2500.0)
    rho = np.full((nx, nz), 2000.0, np.float32)
    g1, dc, _, fd1 = _kernel(c, rho, 0.001, 500, fd_eps=0.5)
    g2, _, _, fd2 = _kernel(c, rho, 0.0005, 1000, fd_eps=0.5)
    g3, _, _, _ = _kernel(c, 2 * rho, 0.001, 500)
    r1, r2 = np.sum(g1 * dc) / fd1, np.sum(g2 * dc) / fd2
    assert r1 > 0 and r2 > 0                   # same sign as the derivative
    assert abs(r2 / r1 - 0.5) < 0.01
    np.testing.assert_allclose(g3, g1, rtol=1e-9, atol=1e-12 * np.abs(g1).max())


def test_gradient_kernel_against_exact_discrete_gradient():
    """torch autograd through a vectorised FP64 restatement of the forward step gives
    the exact discrete gradient of 1/2 |d - d_obs|^2.  The restatement is first checked
    against the oracle's seismograms; the reference kernel (not an exact transpose)
    must then agree with it in sign and correlate strongly."""
    torch = pytest.importorskip("torch")
    torch.set_default_dtype(torch.float64)
    nx, nz, npml, nt, dx, dt = 60, 40, 10, 500, 10.0, 0.001
    c0 = np.full((nx, nz), 2500.0)
    rho = np.full((nx, nz), 2000.0)
    g_ref, dc, r, _ = _kernel(c0, rho.astype(np.float32), dt, nt)
    m = O.Model(c0, rho.astype(np.float32), dx, npml, "f64")
    damp = torch.tensor(m.damp.T)
    nxp, nzp, fs = nx + 2 * npml, nz + 2 * npml, npml
    c1, c2 = 9 / 8, -1 / 24
    src = _src(nt, dt, 15.0).astype(np.float32).astype(np.float64)

    def pad(a):
        return torch.nn.functional.pad(a[None, None], (npml,) * 4, mode="replicate")[0, 0]

    dt32 = float(np.float32(dt))      # the float32 DT_WORK the solver reads

    def forward(c):
        C, D = pad(c.T), pad(torch.tensor(rho.T))
        K, B = D * C * C * dt32 / dx, dt32 / (D * dx)
        p = torch.zeros(nzp, nxp)
        u = torch.zeros(nzp, nxp)
        w = torch.zeros(nzp, nxp)
        out = []
        for it in range(2, nt + 1):
            zs, ze, xs, xe = fs + 1, nzp - 2, 2, nxp - 2
            div = (c1 * (u[zs:ze, xs:xe] - u[zs:ze, xs - 1:xe - 1] + w[zs:ze, xs:xe] - w[zs - 1:ze - 1, xs:xe])
                   + c2 * (u[zs:ze, xs + 1:xe + 1] - u[zs:ze, xs - 2:xe - 2] + w[zs + 1:ze + 1, xs:xe]
                           - w[zs - 2:ze - 2, xs:xe]))
            pn = p.clone()
            pn[zs:ze, xs:xe] = (1 - damp[zs:ze, xs:xe] * dt32) * p[zs:ze, xs:xe] - K[zs:ze, xs:xe] * div
            pn[npml + 5, npml + 15] = pn[npml + 5, npml + 15] + dt32 * src[it - 1]
            pn[fs - 1, :] = -pn[fs + 1, :]
            p = pn
            zs, ze, xs, xe = fs, nzp - 1, 1, nxp - 2
            un = u.clone()
            ku = 0.5 * (damp[zs:ze, xs:xe] + damp[zs:ze, xs + 1:xe + 1]) * dt32
            un[zs:ze, xs:xe] = (1 - ku) * u[zs:ze, xs:xe] - B[zs:ze, xs:xe] * (
                c1 * (p[zs:ze, xs + 1:xe + 1] - p[zs:ze, xs:xe]) + c2 * (p[zs:ze, xs + 2:xe + 2] - p[zs:ze, xs - 1:xe - 1]))
            zs, ze, xs, xe = fs, nzp - 2, 1, nxp - 1
            wn = w.clone()
            kw = 0.5 * (damp[zs:ze, xs:xe] + damp[zs + 1:ze + 1, xs:xe]) * dt32
            wn[zs:ze, xs:xe] = (1 - kw) * w[zs:ze, xs:xe] - B[zs:ze, xs:xe] * (
                c1 * (p[zs + 1:ze + 1, xs:xe] - p[zs:ze, xs:xe]) + c2 * (p[zs + 2:ze + 2, xs:xe] - p[zs - 1:ze - 1, xs:xe]))
            wn[fs - 1, :] = wn[fs, :]
            u, w = un, wn
            out.append(p[npml + 2, npml:npml + nx])
        return torch.stack(out, 1)

    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    # the oracle reads models as float32 (the reference's vel.bin)
    dobs = forward(torch.tensor((c0 + 20 * dc).astype(np.float32).astype(np.float64))).detach()
    ct = torch.tensor(c0, requires_grad=True)
    d = forward(ct)
    # the restatement reproduces the oracle's residual (same forward operator)
    np.testing.assert_allclose((d - dobs).detach().numpy(), r[0][:, 1:], rtol=0,
                               atol=1e-6 * np.abs(r).max())
    (0.5 * ((d - dobs) ** 2).sum()).backward()
    g_ex = ct.grad.numpy()
    inner = (slice(5, -5), slice(5, -5))
    assert np.sign(np.sum(g_ref * dc)) == np.sign(np.sum(g_ex * dc))
    assert np.corrcoef(g_ref[inner].ravel(), g_ex[inner].ravel())[0, 1] > 0.9


# ---------------------------------------------------------------------------------
def test_one_step_impulse_response_textbook_stencil():
    """After one step from rest the only nonzero pressure is the injected dt s(2) at the
    source cell, and the velocities are -B G p with the 4th-order staggered first
    derivative G = [-c2, -c1, c1, c2] / dx offsets, c1 = 9/8, c2 = -1/24 (Levander
    1988): u at columns isx-2..isx+1 (w the same in z) is B P [1/24, -9/8, 9/8, -1/24]
    with B = dt / (rho dx)."""
    nx, nz, npml, dx, dt = 30, 30, 8, 10.0, 0.001
    vp = np.full((nx, nz), 2000.0, np.float32)
    rho = np.full((nx, nz), 1800.0, np.float32)
    m = O.Model(vp, rho, dx, npml, "f64")
    src = np.zeros((1, 2), np.float32)
    src[0, 1] = 3.0                                   # s(2) = 3
    sxz = np.array([[15 * dx, 14 * dx]], np.float32)  # interior cell (ix, iz) = (16, 15) (1-based)
    f = O.forward(m, 2, dt, sxz, sxz[:, None, :], src)
    P = np.float64(np.float32(dt)) * 3.0
    p, u, w = f["p_end"][0], f["u_end"][0], f["w_end"][0]
    ix, iz = 15, 14                                   # 0-based interior indices of the source
    expect = np.zeros_like(p)
    expect[ix, iz] = P
    np.testing.assert_allclose(p, expect, rtol=0, atol=1e-30)
    B = np.float64(np.float32(dt)) / (1800.0 * dx)
    pattern = B * P * np.array([1 / 24, -9 / 8, 9 / 8, -1 / 24])
    np.testing.assert_allclose(u[ix - 2:ix + 2, iz], pattern, rtol=1e-12)
    np.testing.assert_allclose(w[ix, iz - 2:iz + 2], pattern, rtol=1e-12)
    mask = np.ones_like(u, bool)
    mask[ix - 2:ix + 2, iz] = False
    assert np.all(u[mask] == 0)


def test_homogeneous_green_function():
    """Pressure at 200-300 m from a point source in a homogeneous medium equals the exact
    2-D solution of p_tt = c^2 lap p + q_t with q = s(t) dx^2 delta (the injection
    p += dt s adds s per unit time to one cell of area dx^2):
        p(r, t) = dx^2 (G * s')(r, t),  G = H(ct - r) / (2 pi c sqrt(c^2 t^2 - r^2)),
    G integrated over each sample analytically (acosh).  A wrong speed, weight or sign
    moves the arrival or the amplitude; measured agreement: correlation 0.99988,
    amplitude within 0.05 % (16 cells per shortest wavelength)."""
    n, npml, dx, c, f0, dt, nt = 301, 40, 5.0, 2000.0, 10.0, 0.0005, 1000
    vp = np.full((n, n), c, np.float32)
    rho = np.full((n, n), 1000.0, np.float32)
    m = O.Model(vp, rho, dx, npml, "f64")
    src = _src(nt, dt, f0)[None]
    sx = sz = 150 * dx
    rec = np.array([[[sx + 60 * dx, sz], [sx, sz + 40 * dx], [sx + 40 * dx, sz + 30 * dx]]], np.float32)
    seis = O.forward(m, nt, dt, np.array([[sx, sz]], np.float32), rec, src.astype(np.float32),
                     save_strips=False)["seis"][0]
    dt32 = float(np.float32(dt))
    sp = np.gradient(src[0], dt32)

    def g_int(r):
        k = np.arange(nt)
        ua = np.maximum((k - 0.5) * dt32, r / c)
        ub = np.maximum((k + 0.5) * dt32, r / c)
        F = lambda u: np.arccosh(np.maximum(c * u / r, 1.0)) / (2 * np.pi * c * c)
        return F(ub) - F(ua)
    for i, r in enumerate([60 * dx, 40 * dx, 50 * dx]):
        pan = dx * dx * np.convolve(sp, g_int(r))[:nt]
        num = seis[i]
        # pressure sits half a step from the analytic sample times (staggered in time)
        corr = max(np.dot(np.roll(num, -sh), pan) / np.linalg.norm(num) / np.linalg.norm(pan) for sh in (-1, 0))
        assert corr > 0.9995, (r, corr)
        assert abs(np.abs(num).max() / np.abs(pan).max() - 1) < 5e-3


# ---------------------------------------------------------------------------------
def test_adjoint_is_forward_of_the_swapped_medium():
    """The adjoint step updates velocities with den c^2 dt/dx and pressure with
    dt/(den dx) (fdcore.f90:310-345), the forward the other way round (:77-113).  Relabel
    the adjoint's velocity after its next velocity update as the current one: the
    adjoint is then the forward leapfrog of the medium (c, den' = 1/(den c^2)), whose
    pressure coefficient den' c^2 dt/dx = dt/(den dx) and velocity coefficient
    dt/(den' dx) = den c^2 dt/dx.  Adjoint source sadj(it) at reverse step it is
    forward source s'(nt+2-it); the adjoint pressure after step it is the forward
    record at nt+2-it.  With the source, the receivers and every wave inside the
    interior for the whole record (neither the free-surface rows, where the two
    sweeps differ, nor the absorbing layer are reached), the two agree to FP64
    round-off in a heterogeneous medium, where an adjoint without the swap (the
    original medium's forward) differs by the density contrast."""
    from scipy.ndimage import gaussian_filter
    nx = nz = 120
    npml, dx, dt, nt = 10, 10.0, 0.001, 300
    rng = np.random.default_rng(7)
    vp = np.clip(2500 + 7000 * gaussian_filter(rng.standard_normal((nx, nz)), 4), 1900, 3100)
    vp = vp.astype(np.float32)
    rho = (310 * np.power(vp.astype(np.float64), 0.25)
           * (1 + 2.0 * gaussian_filter(rng.standard_normal((nx, nz)), 3))).astype(np.float32)
    assert rho.min() > 0 and rho.max() / rho.min() > 1.5
    src = _src(nt, dt, 25.0)
    sxz = np.array([[60 * dx, 60 * dx]], np.float32)
    rxz = np.array([[[60, 60], [50, 62], [70, 55], [60, 48], [66, 71]]], np.float32) * dx
    ng = rxz.shape[1]
    m = O.Model(vp, rho, dx, npml, "f64")
    f = O.forward(m, nt, dt, sxz, rxz, src[None].astype(np.float32))
    adj = np.zeros((1, ng, nt), np.float32)
    adj[0, 0] = src
    rec = O.adjoint(m, nt, dt, sxz, rxz, src[None].astype(np.float32), adj, f, record=True)["adj_rec"][0]

    j = np.arange(1, nt)                      # forward slot j = it'-1 <-> adjoint slot nt-j
    s_rev = np.zeros(nt)
    s_rev[j] = adj[0, 0][nt - j]
    rec_rev = np.zeros_like(rec)
    rec_rev[:, j] = rec[:, nt - j]
    rho_sw = 1.0 / (rho.astype(np.float64) * vp.astype(np.float64) ** 2)
    m_sw = O.Model(vp.astype(np.float64), rho_sw, dx, npml, "f64", exact_input=True)
    fwd_sw = O.forward(m_sw, nt, dt, sxz, rxz, s_rev[None], save_strips=False)["seis"][0]
    scale = np.abs(fwd_sw).max()
    assert scale > 0
    assert np.abs(rec_rev - fwd_sw).max() <= 1e-12 * scale
    # control: the forward of the original medium (an adjoint without the swap)
    fwd_plain = O.forward(m, nt, dt, sxz, rxz, s_rev[None], save_strips=False)["seis"][0]
    assert np.abs(rec_rev - fwd_plain).max() > 1e-2 * scale


@pytest.mark.parametrize("dt", [0.001, 0.0005])
def test_gradient_kernel_absolute_scale(dt):
    """g = -2/c sum_t p_rec (p0_bk - p_bk) (solver.f90:237,270).  In a homogeneous medium
    the swapped adjoint is the exact transpose, its pressure is dt times the exact
    adjoint variable (injection weight dt, fdcore.f90:353), and -(2/c) sum p dlambda is
    the summation by parts of sum lambda (dK/dc) (-D v) with dK/dc = 2K/c.  So
    <g, dc> = dt dJ/dc for J = 1/2 |d - d_obs|^2, provided the reconstruction is exact
    (source at depth 2 cells, as in every benchmark geometry; a deeper source takes the
    s(it+1) path of fdcore.f90:250-257) and no wave that reaches the free surface or
    the absorbing layer returns to the receivers within the record.  The remaining
    difference is the half-step offset between p_rec and the adjoint difference:
    1.6 % at dt = 1 ms, 0.5 % at 0.5 ms.  A -1/c scale gives 0.5, a sign error -1."""
    nx, nz, npml, dx = 120, 100, 10, 10.0
    nt = int(round(0.3 / dt))
    c = np.full((nx, nz), 2500.0)
    rho = np.full((nx, nz), 2000.0, np.float32)
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    dc = np.exp(-((X - 60) ** 2 + (Z - 30) ** 2) / (2 * 3.0 ** 2))
    src = _src(nt, dt, 20.0)[None].astype(np.float32)
    sxz = np.array([[60 * dx, 2 * dx]], np.float32)
    xs = np.arange(45, 76)
    rxz = np.stack([xs * dx, np.full(len(xs), 45 * dx)], 1)[None].astype(np.float32)

    def seis(cc, save=False):
        m = O.Model(cc, rho, dx, npml, "f64", exact_input=True)
        return m, O.forward(m, nt, dt, sxz, rxz, src, save_strips=save)
    dobs = seis(c + 20 * dc)[1]["seis"]
    m, f = seis(c, True)
    r = f["seis"] - dobs
    g = O.adjoint(m, nt, dt, sxz, rxz, src, r, f)["grad"]
    eps = 0.5
    jp = 0.5 * np.sum((seis(c + eps * dc)[1]["seis"] - dobs) ** 2)
    jm = 0.5 * np.sum((seis(c - eps * dc)[1]["seis"] - dobs) ** 2)
    ratio = np.sum(g * dc) / ((jp - jm) / (2 * eps)) / float(np.float32(dt))
    assert abs(ratio - 1) < (0.025 if dt == 0.001 else 0.008), ratio
