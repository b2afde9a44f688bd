"""The Fortran host layer (iso_c_binding module + OneStep-order driver)
drives the GPU path: the driver's MoveES matches the oracle on the arrays it
wrote, and every full step deposits exactly the particles' charge."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG
from helpers import AMU, ME, QE, assert_close

DRIVER = os.path.join(PKG, "fortran", "build", "ipm_onestep_driver")


def _read_particles(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:8], dtype=np.int64)[0])
    return np.frombuffer(raw[8:8 + 64 * n], dtype=np.float64).reshape(n, 8).T.copy()


def test_fortran_module_compiled():
    assert os.path.exists(os.path.join(PKG, "fortran", "build", "ipm_hip.mod"))
    assert os.path.exists(DRIVER)


@pytest.mark.gpu
def test_fortran_driver_onestep(oracle, tmp_path):
    print(f"running {DRIVER}", flush=True)
    out = subprocess.run([DRIVER, str(tmp_path), "4"], capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:], flush=True)
    assert out.returncode == 0, out.stderr
    nz, nr, dx, dt = 33, 17, 1e-3, 5e-11
    f = np.fromfile(tmp_path / "fields.bin", dtype=np.float64)
    ez, er = f[:(nz + 1) * nr], f[(nz + 1) * nr:]
    assert er.size == nz * (nr + 1)
    pre = _read_particles(tmp_path / "pre.bin")
    post = _read_particles(tmp_path / "post.bin")
    ref = oracle.push(nz, nr, dx, dt, -QE, (ME / AMU) * AMU, ez, er, pre)
    assert_close(post, ref, rtol=1e-12, atol=1e-300, what="Fortran-driven MoveES")
    steps = [l for l in out.stdout.splitlines() if l.startswith("step")]
    assert len(steps) == 4 and "driver done" in out.stdout
    for l in steps:
        qg = float(re.search(r"q_grid\s+(\S+)", l).group(1))
        qp = float(re.search(r"q_particles\s+(\S+)", l).group(1))
        qa = float(re.search(r"q_abs\s+(\S+)", l).group(1))
        assert qa > 0 and abs(qg - qp) <= 1e-12 * qa, l
    # the solver system of the last step, assembled on the device through the
    # Fortran binding, against the oracle on the Rho and Chi the driver read
    sysf = np.fromfile(tmp_path / "system.bin", dtype=np.float64)
    nn, nc = nz * nr, (nz + 2) * (nr + 2)
    assert sysf.size == 9 * nn + nc
    grid = lambda a: a.reshape(nr, nz).T
    rho, chi = grid(sysf[:nn]), sysf[nn:nn + nc].reshape(nr + 2, nz + 2).T
    sigma, iv = grid(sysf[nn + nc:2 * nn + nc]), grid(sysf[2 * nn + nc:3 * nn + nc])
    coe = [grid(sysf[(3 + k) * nn + nc:(4 + k) * nn + nc]) for k in range(6)]
    assert np.abs(chi).max() > 0
    i, j = np.meshgrid(np.arange(1, nz), np.arange(1, nr), indexing="ij")
    eps = 1.0 + 0.01 * i + 0.1 * j
    pt = np.zeros((nz, nr), dtype=np.int32)
    pt[:2, :], pt[-2:, :] = 1, 2
    bound = np.where(pt == 1, 100.0, 0.0)
    ref = oracle.solver_coe(chi, eps, pt, with_chi=True)
    for k, name in enumerate("ABCDE"):
        assert np.array_equal(coe[k], ref[name]), name
    src = oracle.solver_source(dx, rho, iv, sigma=sigma, bound=bound, bound_mask=(pt > 0).astype(np.int32))
    assert np.array_equal(coe[5], src)
