"""The eigenwave3d.py driver: command line of the reference's
tests/eigenwave3d.py (:281-330), generate-only on the CPU, and a full
default run on the GPU checked against the oracle's L2 norms (1e-5)."""
import os
import re
import subprocess
import sys

import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, 'opesci-fd_amd', 'eigenwave3d.py')


def run_cli(args, cwd, timeout=600):
    return subprocess.run([sys.executable, CLI] + args, cwd=str(cwd), capture_output=True, text=True,
                          timeout=timeout)


def test_cli_help_and_generate(tmp_path):
    r = run_cli(['--help'], tmp_path)
    assert r.returncode == 0 and 'default' in r.stdout and '--spatial_order' in r.stdout
    r = run_cli(['default', '-so', '4'], tmp_path)
    assert r.returncode == 0, r.stderr
    assert 'Eigenwave3D example (mode=default)' in r.stdout and 'stress kernel AI' in r.stdout
    assert (tmp_path / 'eigenwave3d.h').exists()
    r = run_cli(['default', '-c', 'g++'], tmp_path)
    assert r.returncode != 0 and 'GPU only' in r.stderr
    r = run_cli(['default', '--pluto'], tmp_path)
    assert r.returncode != 0


@pytest.mark.gpu
def test_cli_default_execute_matches_oracle(tmp_path):
    r = run_cli(['default', '-c', 'hip', '-x'], tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    got = dict((m.group(1), float(m.group(2))) for m in re.finditer(r'(\w+)_l2: ([-0-9.e+]+)', r.stdout))
    assert len(got) == 9
    # default(): 100^3 cells, dt = 0.002, tmax = 1 (500 steps), rho = 1, vp = 1, vs = 0.5
    ref = oracle.run(order=4, n=(100, 100, 100), ntsteps=500, dt=0.002, rho=1.0, vp=1.0, vs=0.5)
    for f, name in enumerate(oracle.FIELDS):
        assert abs(got[name] - ref['l2'][f]) <= 1e-5 * ref['l2'][f] + 1e-9, (name, got[name], ref['l2'][f])
