"""bench.py's contract on the CPU: the algorithmic byte counts behind the roofline (per
shot and time step, forward + backward) against the closed-form figures for three grids,
and the single-GPU rule (--gpus above 1 exits with a message, before any GPU work)."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


@pytest.mark.parametrize("nx,nz,npml,ng,expected", [
    (201, 81, 40, 201, 4_189_056),      # 201x81, 40-cell PML, 201 receivers
    (481, 141, 40, 481, 12_769_856),    # the bench workload
    (801, 267, 50, 801, 35_861_472),    # 801x267, 50-cell PML, 801 receivers
])
def test_algorithmic_bytes_per_shot_step(nx, nz, npml, ng, expected):
    from bench import algorithmic_bytes
    fwd, bwd = algorithmic_bytes(nx, nz, npml, ng)
    n_pad, n_int, S = (nx + 2 * npml) * (nz + 2 * npml), nx * nz, 9 * nx + 6 * nz
    # forward 36 B per padded cell + strip write + receiver write; backward 36 B per
    # padded cell (adjoint) + 32 (reconstruction) + 24 (imaging) per interior cell +
    # strip read + adjoint-source read
    assert fwd == 36 * n_pad + 4 * S + 4 * ng
    assert bwd == 36 * n_pad + 56 * n_int + 4 * S + 4 * ng
    assert fwd + bwd == expected


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
