"""Inversion loop of the reference toolbox (toolbox/inversion.py:30-482) on top of this
package's solver: the same call order per iteration (forward 'syn' with snapshots,
process_workflow, misfit, adjoint, search direction, backtracking line search with
forward-only runs, decision), the same NLCG (Polak-Ribiere, periodic restart after 6,
descent check) and L-BFGS (memory 5, Liu-Nocedal initial scaling, descent check,
restart on failure) with float32 S / Y history, the same step-length safeguards and
model clipping.  The L-BFGS history lives in memory instead of the reference's
outputs/LBFGS_memory/ memmaps; plotting is not reproduced.  Outputs written as
toolbox/tools.py:340-359 does (gradient / direction / velocity per iteration, misfit
and line-search logs).
"""
from __future__ import annotations

import copy
import os

import numpy as np

from . import preprocess
from .tools import array2vector, savebinfloat32, vector2array


def _solver():
    from . import solver
    return solver


def _misfit(simu, misfit_type):
    from . import misfit
    return misfit.misfit(simu, misfit_type)


def inversion(simu, optim, inv_model):
    """inversion.py:30-79."""
    simu.model.vp = inv_model['vp']
    simu.model.rho = inv_model['rho']
    inv_scheme = optimize_init(simu, optim)
    inv_scheme['v_now'] = array2vector(simu.model.vp)
    inv_scheme['v_old'] = array2vector(simu.model.vp)
    solver = _solver()
    while optim.iter < optim.maxiter:
        optim.iter += 1
        print('\n-----------  iteration %d  -----------\n' % optim.iter)
        solver.forward(simu, simu_type='syn', savesnap=1)
        preprocess.process_workflow(simu, optim, simu_type='syn')
        inv_scheme['f_now'] = _misfit(simu, optim.misfit_type)
        inv_scheme['g_now'] = solver.adjoint(simu, optim)
        if optim.scheme in ['NLCG']:
            inv_scheme = NLCG(inv_scheme)
        elif optim.scheme in ['LBFGS']:
            inv_scheme = LBFGS(inv_scheme)
        inv_scheme = backtrack(simu, optim, inv_scheme)
        inv_scheme = desicion_make(simu, optim, inv_scheme)
        save_inv_scheme(simu, optim, inv_scheme)
    print('\n-----------  iteration end  -----------\n')
    return inv_scheme


def rtm(simu, optim, inv_model):
    """inversion.py:83-107 (image only; no plot)."""
    simu.model.vp = inv_model['vp']
    simu.model.rho = inv_model['rho']
    inv_scheme = optimize_init(simu, optim)
    inv_scheme['v_now'] = array2vector(simu.model.vp)
    inv_scheme['v_old'] = array2vector(simu.model.vp)
    solver = _solver()
    solver.forward(simu, simu_type='syn', savesnap=1)
    inv_scheme['g_now'] = solver.adjoint(simu, optim)
    return inv_scheme


def NLCG(inv_scheme):
    """inversion.py:112-171."""
    cg_itmax = 6
    g_now, g_old, d_old = inv_scheme['g_now'], inv_scheme['g_old'], inv_scheme['d_old']
    cg_it = inv_scheme['cg_it'] + 1
    if cg_it == 1:
        d_now = -g_now
        status = 0
    elif cg_it > cg_itmax:
        d_now = -g_now
        status = 1
        print('restarting NLCG... [periodic restart]')
    else:
        beta = pollak_ribere(g_now, g_old)
        if beta <= 0.:
            beta = 0.
            status = 1
            print('restarting NLCG... [negative beta]')
        else:
            status = 0
        d_now = -g_now + beta * d_old
    if check_descent(d_now, g_now) > 0.:
        d_now = -g_now
        status = 1
        print('restarting NLCG... [not a descent direction]')
    else:
        print('find NLCG direction')
    if status:
        cg_it = 0
    inv_scheme['d_now'] = d_now
    inv_scheme['cg_it'] = cg_it
    return inv_scheme


def lbfgs_direction(S, Y, k, g_now):
    """Two-loop recursion over the k newest pairs (column 0 newest) with the
    Liu-Nocedal initial scaling (inversion.py:222-239).  Returns z ~ H g."""
    rho = np.zeros(k)
    alpha = np.zeros(k)
    q = g_now
    for i in range(k):
        rho[i] = 1. / np.dot(Y[:, i], S[:, i])
        alpha[i] = rho[i] * np.dot(S[:, i], q)
        q = q - alpha[i] * Y[:, i]
    invH = np.dot(Y[:, 0], S[:, 0]) / np.dot(Y[:, 0], Y[:, 0])
    z = invH * q
    for i in range(k - 1, -1, -1):
        beta = rho[i] * np.dot(Y[:, i], z)
        z = z + (alpha[i] - beta) * S[:, i]
    return z


def LBFGS(inv_scheme):
    """inversion.py:174-263, with the S / Y history held in inv_scheme (float32)."""
    memory = inv_scheme['memory']
    memory_used = inv_scheme['memory_used']
    lbfgs_it = inv_scheme['lbfgs_it'] + 1
    g_now, g_old = inv_scheme['g_now'], inv_scheme['g_old']
    v_now, v_old = inv_scheme['v_now'], inv_scheme['v_old']
    S, Y = inv_scheme['S'], inv_scheme['Y']
    if lbfgs_it == 1:
        d_now = -g_now
        status = 0
    else:
        s = v_now - v_old
        y = g_now - g_old
        if memory_used == 0:
            S[:, 0] = s
            Y[:, 0] = y
            memory_used = 1
        else:
            S[:, 1:] = S[:, :-1]
            Y[:, 1:] = Y[:, :-1]
            S[:, 0] = s
            Y[:, 0] = y
            if memory_used < memory:
                memory_used += 1
        z = lbfgs_direction(S, Y, memory_used, g_now)
        if check_descent(-z, g_now) > 0:
            d_now = -g_now
            status = 1
            print('restarting LBFGS... [not a descent direction]')
        else:
            d_now = -z
            status = 0
    if status:
        lbfgs_it = 1
        memory_used = 0
        S[:] = 0.
        Y[:] = 0.
    else:
        print('find LBFGS direction')
    inv_scheme['memory_used'] = memory_used
    inv_scheme['lbfgs_it'] = lbfgs_it
    inv_scheme['d_now'] = d_now
    return inv_scheme


def backtrack(simu, optim, inv_scheme):
    """inversion.py:266-338: halve the step until the misfit decreases (at most 6
    tries); L-BFGS after its first iteration starts from a step that keeps max|d step|
    within [20, 300] m/s."""
    dec, search_max, vmax_thresh, vmin_thresh = 0.5, 6, 300, 20
    nx, nz = simu.model.nx, simu.model.nz
    f_now, v_now, d_now = inv_scheme['f_now'], inv_scheme['v_now'], inv_scheme['d_now']
    if optim.scheme in ['LBFGS']:
        if inv_scheme['lbfgs_it'] == 1:
            step = optim.step_length
        elif max(abs(d_now)) > vmax_thresh:
            step = 1.0 * vmax_thresh / max(abs(d_now))
        elif max(abs(d_now)) < vmin_thresh:
            step = 1.0 * vmin_thresh / max(abs(d_now))
        else:
            step = 1.0
    else:
        step = optim.step_length
    solver = _solver()
    ls_iter = 0
    while True:
        ls_iter += 1
        v_try = model_update(v_now, d_now * step, optim.vpmax, optim.vpmin)
        simu.model.vp = vector2array(v_try, nx, nz)
        solver.forward(simu, simu_type='syn', savesnap=0)
        preprocess.process_workflow(simu, optim, simu_type='syn')
        f_try = _misfit(simu, optim.misfit_type)
        if f_try >= f_now:
            print("line search: step = %.4f, f_now = %.4e, f_try = %.4e. Not decrease." % (step, f_now, f_try))
            width = dec
        else:
            inv_scheme = updata_inv_scheme(inv_scheme, 0, ls_iter, step, f_try, v_try)
            print("line search: step = %.4f, f_now = %.4e, f_try = %.4e. Succeed." % (step, f_now, f_try))
            return inv_scheme
        if ls_iter >= search_max:
            inv_scheme = updata_inv_scheme(inv_scheme, -1, ls_iter, step, f_try, v_try)
            print("line search fails: exceed the iteration of linesearch.")
            return inv_scheme
        step = step * width


def optimize_init(simu, optim):
    """inversion.py:341-400 (memory 5 for L-BFGS)."""
    nx, nz = simu.model.nx, simu.model.nz
    temp = np.zeros(nx * nz, dtype=np.float32)
    common = {'g_now': temp, 'g_old': temp, 'd_now': temp, 'v_old': temp, 'v_now': temp,
              'v_try': temp, 'f_old': 0., 'f_now': 0., 'f_try': 0., 'ls_iter': 0, 'ls_status': 0,
              'ls_step': 0., 'nx': nx, 'nz': nz}
    if optim.scheme in ['NLCG']:
        return dict(common, d_old=temp, cg_it=0)
    memory = 5
    return dict(common, memory=memory, memory_used=0, lbfgs_it=0,
                S=np.zeros((nx * nz, memory), dtype=np.float32),
                Y=np.zeros((nx * nz, memory), dtype=np.float32))


def desicion_make(simu, optim, inv_scheme):
    """inversion.py:403-449."""
    nx, nz = simu.model.nx, simu.model.nz
    if optim.scheme in ['NLCG']:
        if inv_scheme['ls_status'] == -1 and inv_scheme['cg_it'] > 1:
            inv_scheme['cg_it'] = 0
            inv_scheme['g_old'] = inv_scheme['g_old'] * 0.
            inv_scheme['d_old'] = inv_scheme['d_old'] * 0.
            print('line search fails: restarting the line search ...')
        else:
            inv_scheme['g_old'] = inv_scheme['g_now']
            inv_scheme['d_old'] = inv_scheme['d_now']
            inv_scheme['f_old'] = inv_scheme['f_now']
            inv_scheme['f_now'] = inv_scheme['f_try']
            inv_scheme['v_now'] = inv_scheme['v_try']
            print('accept line search result')
    elif optim.scheme in ['LBFGS']:
        if inv_scheme['ls_status'] == -1:
            old = copy.deepcopy(inv_scheme)
            inv_scheme = optimize_init(simu, optim)
            for key in ('v_now', 'f_now', 'g_now', 'd_now'):
                inv_scheme[key] = old[key]
            inv_scheme['v_old'] = old['v_now']
            print('line search fails: restarting the line search ...')
        else:
            inv_scheme['g_old'] = inv_scheme['g_now']
            inv_scheme['v_old'] = inv_scheme['v_now']
            inv_scheme['v_now'] = inv_scheme['v_try']
            inv_scheme['f_old'] = inv_scheme['f_now']
            inv_scheme['f_now'] = inv_scheme['f_try']
            print('accept line search result')
    simu.model.vp = vector2array(inv_scheme['v_now'], nx, nz)
    return inv_scheme


def updata_inv_scheme(inv_scheme, ls_status, ls_iter, ls_step, f_try, v_try):
    """inversion.py:452-462."""
    inv_scheme['ls_status'] = ls_status
    inv_scheme['ls_iter'] = ls_iter
    inv_scheme['ls_step'] = ls_step
    inv_scheme['f_try'] = f_try
    inv_scheme['v_try'] = v_try
    return inv_scheme


def model_update(vp, update, vpmax, vpmin):
    """inversion.py:465-474: step and clip to [vpmin, vpmax]."""
    vp_new = vp + update
    vp_new[vp_new > vpmax] = vpmax
    vp_new[vp_new < vpmin] = vpmin
    return vp_new


def pollak_ribere(g_now, g_old):
    """inversion.py:477-484."""
    return np.dot(g_now, g_now - g_old) / np.dot(g_old, g_old)


def check_conjugacy(g_now, g_old):
    return abs(np.dot(g_now, g_old) / np.dot(g_now, g_now))


def check_descent(d_now, g_now):
    return np.dot(d_now, g_now) / np.dot(g_now, g_now)


def _savetxt(path, value, fmt):
    """Append one value per line, '%d' or '%.4e' (toolbox/tools.py:307-319)."""
    with open(path, 'a') as f:
        f.write(('%d\n' if fmt == 'int' else '%.4e\n') % value)


def save_inv_scheme(simu, optim, inv_scheme):
    """toolbox/tools.py:340-359 (gradient, direction, velocity per iteration; misfit and
    line-search logs)."""
    out = simu.system.homepath + 'outputs'
    for sub in ('gradient', 'direction', 'velocity'):
        os.makedirs(out + '/' + sub, exist_ok=True)
    it = optim.iter
    savebinfloat32(out + '/gradient/grad-%d.bin' % it, inv_scheme['g_now'])
    savebinfloat32(out + '/direction/dirc-%d.bin' % it, inv_scheme['d_now'])
    savebinfloat32(out + '/velocity/vp-%d.bin' % it, inv_scheme['v_now'])
    if optim.iter == 1:
        _savetxt(out + '/misfit_data.dat', inv_scheme['f_old'], 'float')
    _savetxt(out + '/misfit_data.dat', inv_scheme['f_now'], 'float')
    _savetxt(out + '/line_search_step.dat', inv_scheme['ls_step'], 'float')
    _savetxt(out + '/line_search_iteration.dat', inv_scheme['ls_iter'], 'int')
