"""Parameter objects for callers of the drop-in solver.

Callers written for the reference toolbox build `simulate(model, source, receiver,
system)` and `optimize(...)` and hand them to `forward` / `adjoint`; the class names,
constructor arguments and attribute names are that contract (toolbox/base.py:19-360)
and are kept.  What the constructors do is this module's own:

  * `simulate` validates the job before anything runs (array shapes, the 4th-order
    stability limit on dt, the dispersion rule of thumb as a warning, coordinates
    inside the model) and lays out the working folders under `system.homepath` that
    the solver and the inversion loop write into;
  * `model` derives the padded extents, the axes and the velocity range;
  * `optimize` validates the inversion settings against the supported choices.
"""
from __future__ import annotations

import os
import shutil
from multiprocessing import cpu_count

import numpy as np

# Folders (relative to homepath) that the solver and the inversion loop write into;
# the first group is emptied whenever a new simulation is set up (the observed data
# under data/obs survive).
_RESET = ('data/syn', 'data/tempdata', 'parfile', 'outputs', 'figures')
_TREE = {
    'data': ('obs', 'syn', 'tempdata'),
    'parfile': ('forward_parfile', 'forward_source', 'adjoint_parfile', 'adjoint_source', 'model'),
    'outputs': ('velocity', 'gradient', 'direction', 'LBFGS_memory'),
    'figures': ('model', 'waveform'),
}

# 4th-order staggered leapfrog: stable for dt * vmax / dx below sqrt(3/8)
_COURANT_LIMIT = np.sqrt(3.0 / 8.0)
# at least this many grid cells per wavelength of the peak frequency
_CELLS_PER_WAVELENGTH = 10.0


class simulate(object):
    """One modelling job: model, sources, receivers and the system settings.
    Constructing it checks the job and prepares `system.homepath`
    (toolbox/base.py:19-142)."""

    def __init__(self, model, source, receiver, system):
        self.model = model
        self.source = source
        self.receiver = receiver
        self.system = system
        self.initialize()

    def _check(self):
        m, home = self.model, self.system.homepath
        for name in ('vp', 'rho'):
            got = np.shape(getattr(m, name))
            if got != (m.nx, m.nz):
                raise ValueError('%s has shape %s; the grid is nx x nz = (%d, %d)' % (name, got, m.nx, m.nz))
        if not home.endswith('/'):
            self.system.homepath = home + '/'
        vmax, vmin = float(np.max(m.vp)), float(np.min(m.vp))
        dt_max = _COURANT_LIMIT * m.dx / vmax
        if m.dt >= dt_max:
            raise ValueError('unstable time step: dt = %.4f ms, the grid and max velocity %.1f m/s allow '
                             'less than %.4f ms' % (m.dt * 1e3, vmax, dt_max * 1e3))
        dx_max = vmin / self.source.f0 / _CELLS_PER_WAVELENGTH
        if m.dx > dx_max:
            print('Warning: dx = %.2f m is coarse for f0 = %.2f Hz and min velocity %.1f m/s '
                  '(dispersion; dx <= %.2f m or f0 <= %.2f Hz avoids it)'
                  % (m.dx, self.source.f0, vmin, dx_max, vmin / m.dx / _CELLS_PER_WAVELENGTH))
        lo, hi = float(m.xx.min()), float(m.xx.max())
        xs = np.asarray(self.source.xz)[:, 0]
        xr = np.asarray(self.receiver.xz)[..., 0]
        if min(xs.min(), xr.min()) < lo or max(xs.max(), xr.max()) > hi:
            raise ValueError('a source or receiver x coordinate lies outside [%.1f, %.1f] m' % (lo, hi))
        # one process per shot at most, and at most half the host's cores
        self.system.mpiproc = min(self.system.mpiproc, self.source.n, max(cpu_count() // 2, 1))

    def _builddir(self):
        home = self.system.homepath
        for sub in _RESET:
            if os.path.exists(home + sub):
                shutil.rmtree(home + sub)
        for top, subs in _TREE.items():
            for sub in subs:
                os.makedirs(os.path.join(home, top, sub), exist_ok=True)

    def _screenprint(self):
        m = self.model
        print('2-D acoustic modelling on one GPU: %d x %d cells of %.1f m, %d steps of %.2f ms, %d shots'
              % (m.nx, m.nz, m.dx, m.nt, m.dt * 1e3, self.source.n))

    def initialize(self):
        self._check()
        self._builddir()
        self._screenprint()


class model(object):
    """Grid, time axis and the vp / rho models (toolbox/base.py:145-179).  `fs` is
    kept as given; the solver applies the free surface either way, as the
    reference's does."""

    def __init__(self, nx, nz, dx, dt, nt, fs, pml, vp, rho):
        self.nx, self.nz, self.dx = nx, nz, dx
        self.dt, self.nt = dt, nt
        self.fs, self.pml = fs, pml
        self.nx_pml = nx + 2 * pml
        self.nz_pml = nz + pml * (2 - fs)
        self.xx = dx * np.arange(nx)
        self.zz = dx * np.arange(nz)
        self.t = dt * np.arange(nt)
        self.vp, self.rho = vp, rho
        self.vpmax, self.vpmin = vp.max(), vp.min()
        self.savesnap = 0
        self.savestep = 1


class source(object):
    """Peak frequency, shot count, positions (n, 2) in metres and wavelets
    (toolbox/base.py:182-190)."""

    def __init__(self, f0, n, xz, wavelet):
        self.f0, self.n, self.xz, self.wavelet = f0, n, xz, wavelet


class receiver(object):
    """Receivers per shot and their positions (shots, n, 2) in metres
    (toolbox/base.py:193-199)."""

    def __init__(self, n, xz):
        self.n, self.xz = n, xz


class system(object):
    """Working folder, process count and figure aspect (toolbox/base.py:202-210)."""

    def __init__(self, homepath, mpiproc, figaspect=1):
        self.homepath, self.mpiproc, self.figaspect = homepath, mpiproc, figaspect


# supported settings of `optimize`
_SCHEMES = ('NLCG', 'LBFGS')
_MISFITS = ('Waveform', 'Envelope', 'Traveltime', 'Globalcorrelation', 'RTM')
_NORMS = ('Max-Trace', 'L1-Event', 'L2-Event', 'L1-Trace', 'L2-Trace', 'None')
_FILTERS = ('Lowpass', 'Bandpass', 'Highpass', 'None')
# gradient taper floor by acquisition setting
_THRESHOLD = {'marine': 0.0, 'offshore': 0.0, 'land': 0.001, 'onshore': 0.001}


class optimize(object):
    """Inversion settings (toolbox/base.py:213-360): the argument order and attribute
    names are the toolbox's; invalid choices raise ValueError."""

    def __init__(self, misfit_type, scheme, maxiter, step_length, vpmax, vpmin, marine_or_land,
                 grad_mute, grad_smooth, fre_filter, fre_low, fre_high, mute_late_arrival,
                 mute_late_window, normalize, mute_offset_short, mute_offset_long,
                 mute_offset_short_dis, mute_offset_long_dis, grad_mask=None):
        self.iter = 0
        self.misfit_type, self.scheme, self.maxiter = misfit_type, scheme, maxiter
        self.step_length, self.vpmax, self.vpmin = step_length, vpmax, vpmin
        self.marine_or_land = marine_or_land
        self.grad_mute, self.grad_smooth, self.grad_mask = grad_mute, grad_smooth, grad_mask
        self.fre_filter, self.fre_low, self.fre_high = fre_filter, fre_low, fre_high
        self.mute_late_arrival, self.mute_late_window = mute_late_arrival, mute_late_window
        self.normalize = normalize
        self.mute_offset_short, self.mute_offset_long = mute_offset_short, mute_offset_long
        self.mute_offset_short_dis, self.mute_offset_long_dis = mute_offset_short_dis, mute_offset_long_dis
        key = str(marine_or_land).lower()
        if key not in _THRESHOLD:
            raise ValueError('marine_or_land must be Marine/Offshore or Land/Onshore, not %r' % (marine_or_land,))
        self.grad_thred = _THRESHOLD[key]
        self._check()
        print('Inversion: %s, misfit %s, %d iterations' % (self.scheme, self.misfit_type, self.maxiter))

    def _check(self):
        choices = ((self.scheme, _SCHEMES, 'scheme'), (self.misfit_type, _MISFITS, 'misfit_type'),
                   (self.fre_filter, _FILTERS, 'fre_filter'))
        for value, allowed, name in choices:
            if value not in allowed:
                raise ValueError('%s %r is not one of %s' % (name, value, ', '.join(allowed)))
        if not any(k in self.normalize for k in _NORMS):
            raise ValueError('normalize %r names none of %s' % (self.normalize, ', '.join(_NORMS)))
        for lo, hi, what in ((self.vpmin, self.vpmax, 'vpmin <= vpmax'), (self.fre_low, self.fre_high,
                             'fre_low <= fre_high'), (self.mute_offset_short_dis, self.mute_offset_long_dis,
                                                      'mute_offset_short_dis <= mute_offset_long_dis')):
            if lo > hi:
                raise ValueError('need %s (got %g > %g)' % (what, lo, hi))
        if self.misfit_type == 'RTM':
            # an image is one gradient
            self.maxiter = 1
