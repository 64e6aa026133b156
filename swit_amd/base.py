"""Simulation and optimisation parameter objects with the attribute layout of the
reference toolbox (toolbox/base.py:19-360), so callers written for it (inversion loop,
misfit, pre/postprocessing, example scripts) can build `simu` / `optim` the same way
and pass them to swit_amd.solver.forward / adjoint."""
from __future__ import annotations

import os
import shutil
from multiprocessing import cpu_count

import numpy as np


class simulate(object):
    """toolbox/base.py:19-142: holds model/source/receiver/system, checks them and
    prepares the working folders under system.homepath."""

    def __init__(self, model, source, receiver, system):
        self.model = model
        self.source = source
        self.receiver = receiver
        self.system = system
        self.initialize()

    def _check(self):
        """toolbox/base.py:33-77."""
        dx, dt, f0 = self.model.dx, self.model.dt, self.source.f0
        vpmin, vpmax = np.min(self.model.vp), np.max(self.model.vp)
        dt_req = np.sqrt(3.0 / 8.0) * dx / vpmax
        dx_req = vpmin / f0 / 10.0
        f0_req = vpmin / dx / 10.0
        if np.shape(self.model.vp) != (self.model.nx, self.model.nz):
            raise ValueError('Wrong dimensions of Vp, not consistant with nx = %d, nz = %d'
                             % (self.model.nx, self.model.nz))
        if np.shape(self.model.rho) != (self.model.nx, self.model.nz):
            raise ValueError('Wrong dimensions of Rho, not consistant with nx = %d, nz = %d'
                             % (self.model.nx, self.model.nz))
        if self.system.homepath[-1] != '/':
            self.system.homepath += '/'
        if dt_req <= dt:
            raise ValueError('modeling stability: dt = %.4f ms > dt_required = %.4f ms: '
                             % (dt * 1000, dt_req * 1000))
        if dx_req < dx:
            print('Warning: modeling dispersion, dx = %.2f m > dx_required =  %.2f m' % (dx, dx_req))
            print('Warning: modeling dispersion, f0 = %.2f Hz > f0_required = %.2f Hz' % (f0, f0_req))
        sx = np.asarray(self.source.xz)[:, 0]
        rx = np.asarray(self.receiver.xz)[..., 0]
        if (sx.min() < self.model.xx.min() or sx.max() > self.model.xx.max()
                or rx.min() < self.model.xx.min() or rx.max() > self.model.xx.max()):
            raise ValueError('source or receiver coordinates are out of range')
        self.system.mpiproc = min([self.system.mpiproc, self.source.n, max(cpu_count() // 2, 1)])

    def _builddir(self):
        """toolbox/base.py:80-119: clear data/syn, data/tempdata, parfile, outputs,
        figures and recreate the folder tree."""
        home = self.system.homepath
        for sub in ('data/syn', 'data/tempdata', 'parfile', 'outputs', 'figures'):
            if os.path.exists(home + sub):
                shutil.rmtree(home + sub)
        for sub in ('', 'data', 'data/obs/', 'data/syn/', 'data/tempdata/', 'parfile',
                    'parfile/forward_parfile/', 'parfile/forward_source/',
                    'parfile/adjoint_parfile/', 'parfile/adjoint_source/', 'parfile/model/',
                    'outputs', 'outputs/velocity/', 'outputs/gradient/', 'outputs/direction/',
                    'outputs/LBFGS_memory/', 'figures', 'figures/model/', 'figures/waveform/'):
            os.makedirs(home + sub, exist_ok=True)

    def _screenprint(self):
        print('*****************************************************')
        print('\n        Seismic Waveform Inversion Toolbox         \n')
        print('*****************************************************\n')
        print('Forward modeling : nx, nz = %d, %d' % (self.model.nx, self.model.nz))
        print('Forward modeling : dx = %.1f m' % (self.model.dx))
        print('Forward modeling : dt = %.2f ms, %d steps' % (self.model.dt * 1000, self.model.nt))
        print('Forward modeling : %d shots, batched on one GPU' % (self.source.n))

    def initialize(self):
        self._check()
        self._builddir()
        self._screenprint()


class model(object):
    """toolbox/base.py:145-179."""

    def __init__(self, nx, nz, dx, dt, nt, fs, pml, vp, rho):
        self.nx = nx
        self.nz = nz
        self.dx = dx
        self.dt = dt
        self.nt = nt
        self.fs = fs
        self.pml = pml
        self.nx_pml = self.nx + self.pml * 2
        self.nz_pml = self.nz + self.pml * (2 - self.fs)
        self.xx = np.arange(0, self.nx * self.dx, self.dx)
        self.zz = np.arange(0, self.nz * self.dx, self.dx)
        self.t = np.linspace(0, self.dt * self.nt, num=self.nt, endpoint=False)
        self.rho = rho
        self.vp = vp
        self.vpmax = vp.max()
        self.vpmin = vp.min()
        self.savesnap = 0
        self.savestep = 1
        # the solver applies the free surface whatever fs is (toolbox/base.py:176-177)


class source(object):
    """toolbox/base.py:182-190."""

    def __init__(self, f0, n, xz, wavelet):
        self.f0 = f0
        self.n = n
        self.xz = xz
        self.wavelet = wavelet


class receiver(object):
    """toolbox/base.py:193-199."""

    def __init__(self, n, xz):
        self.n = n
        self.xz = xz


class system(object):
    """toolbox/base.py:202-210."""

    def __init__(self, homepath, mpiproc, figaspect=1):
        self.homepath = homepath
        self.mpiproc = mpiproc
        self.figaspect = figaspect


class optimize(object):
    """toolbox/base.py:213-360 (parameter checks; screen print reduced to one line)."""

    def __init__(self, misfit_type, scheme, maxiter, step_length, vpmax, vpmin, marine_or_land,
                 grad_mute, grad_smooth, fre_filter, fre_low, fre_high, mute_late_arrival,
                 mute_late_window, normalize, mute_offset_short, mute_offset_long,
                 mute_offset_short_dis, mute_offset_long_dis, grad_mask=None):
        self.iter = 0
        self.misfit_type = misfit_type
        self.scheme = scheme
        self.maxiter = maxiter
        self.step_length = step_length
        self.vpmax = vpmax
        self.vpmin = vpmin
        self.marine_or_land = marine_or_land
        self.grad_mute = grad_mute
        self.grad_smooth = grad_smooth
        self.grad_mask = grad_mask
        self.fre_filter = fre_filter
        self.fre_low = fre_low
        self.fre_high = fre_high
        self.mute_late_arrival = mute_late_arrival
        self.mute_late_window = mute_late_window
        self.normalize = normalize
        self.mute_offset_short = mute_offset_short
        self.mute_offset_long = mute_offset_long
        self.mute_offset_short_dis = mute_offset_short_dis
        self.mute_offset_long_dis = mute_offset_long_dis
        if self.marine_or_land.lower() in ['marine', 'offshore']:
            self.grad_thred = 0.0
        elif self.marine_or_land.lower() in ['land', 'onshore']:
            self.grad_thred = 0.001
        else:
            raise ValueError('not supported modeling marine_or_land: %s' % (self.marine_or_land))
        self._check()
        print('Inversion: %s, misfit %s, %d iterations' % (self.scheme, self.misfit_type, self.maxiter))

    def _check(self):
        if self.scheme not in ['NLCG', 'LBFGS']:
            raise ValueError('not supported inversion scheme: %s' % self.scheme)
        if self.misfit_type not in ['Waveform', 'Envelope', 'Traveltime', 'Globalcorrelation', 'RTM']:
            raise ValueError('not supported misfit function: %s' % self.misfit_type)
        if self.misfit_type in ['RTM']:
            self.maxiter = 1
        if not any(k in self.normalize for k in ('Max-Trace', 'L1-Event', 'L2-Event', 'L1-Trace',
                                                 'L2-Trace', 'None')):
            raise ValueError('not supported normalization:', self.normalize)
        if self.fre_filter not in ['Lowpass', 'Bandpass', 'Highpass', 'None']:
            raise ValueError('not supported frequency filter: %s' % self.fre_filter)
        if self.vpmax < self.vpmin:
            raise ValueError('vpmax=%f m/s is less than vpmin=%f m/s\n' % (self.vpmax, self.vpmin))
        if self.fre_low > self.fre_high:
            raise ValueError('fre_low > fre_high')
        if self.mute_offset_short_dis > self.mute_offset_long_dis:
            raise ValueError('mute_offset_short_dis > mute_offset_long_dis')
