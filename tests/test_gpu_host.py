"""The C++ host mirror (primal_dual_contacts_miror_amd/host): its
GpuPrimalDualForceSolver::fillForces, the ForceSolver plugin entry the
reference's DynamicSystem calls (src/Solver/ForceSolverBase.h:78-82), driven
through the test driver tests/host_fill_forces.cpp on the oracle's contact
list, must give the oracle's velocity increments within 1e-6 |g| dt and the
Python binding's results bit for bit, leave every contact's objId as it
found it, and write the local forces into the contact list.  The stack demo
(host/stack_demo.cpp, the mirror's DynamicSystem) must print what the Python
binding computes for its scene and stay within 1e-5 of the oracle.

The driver's input check (a truncated file is refused before any device
call) runs without a GPU.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import orc
from primal_dual_contacts_miror_amd import pdc, scenes
from primal_dual_contacts_miror_amd.params import PdcSolveStats, default_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "primal_dual_contacts_miror_amd", "bin")
DRIVER = os.path.join(BIN, "pdc_host_fill_forces")
DEMO = os.path.join(BIN, "pdc_stack_demo")


def _need(path):
    if not os.path.exists(path):
        pytest.fail("%s is not built (run __graft_entry__.build())" % path)


def _write_input(path, sc, params, state, co):
    n, C = sc.n, co.n
    _, quat, vel, ang = state
    with open(path, "wb") as f:
        f.write(bytes(params))
        f.write(np.array([n, C], np.int32).tobytes())
        for a, t in ((sc.shape_type, np.int32), (sc.is_static, np.int32), (sc.emit, np.int32),
                     (sc.dims, np.float64), (sc.mass, np.float64), (sc.inertia, np.float64),
                     (quat, np.float64), (vel, np.float64), (ang, np.float64),
                     (co.ids, np.int32), (co.T.reshape(-1, 36), np.float64), (co.depth, np.float64)):
            f.write(np.ascontiguousarray(a, dtype=t).tobytes())


def _read_output(path, n, C):
    raw = open(path, "rb").read()
    acc = np.frombuffer(raw, np.float64, 6 * n).reshape(n, 6)
    lf = np.frombuffer(raw, np.float64, 3 * C, 48 * n).reshape(C, 3)
    off = 48 * n + 24 * C
    size = ctypes.sizeof(PdcSolveStats)
    st = PdcSolveStats.from_buffer_copy(raw[off:off + size])
    kept = int(np.frombuffer(raw, np.int32, 1, off + size)[0])
    return acc, lf, st, kept


def test_driver_refuses_truncated_input(tmp_path):
    _need(DRIVER)
    sc = scenes.stack_10()
    p = str(tmp_path / "in.bin")
    with open(p, "wb") as f:
        f.write(bytes(sc.params))
        f.write(np.array([sc.n, 5], np.int32).tobytes())
    r = subprocess.run([DRIVER, p, str(tmp_path / "out.bin")], capture_output=True, timeout=60)
    assert r.returncode == 1 and not os.path.exists(str(tmp_path / "out.bin"))


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["stack_10", "pyramid_210"])
def test_host_mirror_fill_forces(require_gpu, tmp_path, which):
    _need(DRIVER)
    sc = scenes.stack_10(tolerance=1e-10) if which == "stack_10" else scenes.pyramid_210(tolerance=1e-10)
    p = sc.params
    p.pcg_tol = 1e-13
    o = orc.OracleScene(sc, p)
    o.set_threads(16)
    o.set_params(p)
    o.step(30 if which == "stack_10" else 25)
    state = o.get_state()
    n_c, _ = o.detect()
    co = o.contacts()
    assert n_c > 0
    o.set_linear_solver(orc.SOLVER_DIRECT)
    acc_o, _, st_o = o.fill_forces(co, *state[1:])
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, sc, p, state, co)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    acc_h, lf_h, st_h, kept = _read_output(fout, sc.n, co.n)
    solver = pdc.PrimalDualForceSolver(sc, p)
    acc_p, lf_p, st_p = solver.fill_forces(co, *state[1:])
    solver.ctx.close()
    gdt = np.linalg.norm(p.gravity[:]) * p.timestep
    err = np.max(np.abs(acc_h - acc_o)) / gdt
    print("%s: %d contacts, host mirror vs oracle %.3e |g|dt, Newton %d (oracle %d)"
          % (which, co.n, err, st_h.newton_iterations, st_o.newton_iterations))
    assert kept == 1
    assert err < 1e-6
    assert st_h.n_contacts == co.n and st_h.newton_iterations == st_p.newton_iterations
    assert np.array_equal(acc_h, acc_p) and np.array_equal(lf_h, lf_p)
    assert np.any(lf_h[:, 0] > 0)
    assert np.all(acc_h[np.asarray(sc.is_static, bool)] == 0)


def _demo_scene():
    """The stack of host/stack_demo.cpp: a static ground box of mass 1 and unit
    inertia that emits no samples, ten unit cubes at y = 0.5 + 0.001 + 1.001 k,
    library default parameters."""
    sc = scenes.stack_10()
    sc.params = default_params()
    sc.mass[0] = 1.0
    sc.inertia[0] = 1.0
    for k in range(10):
        sc.pos[1 + k, 1] = 0.5 + 0.001 + k * 1.001
    return sc


@pytest.mark.gpu
def test_stack_demo_comes_to_rest(require_gpu):
    """The demo (the mirror's DynamicSystem over the C ABI) prints what the
    Python binding computes on the same scene, to its printed digits; the
    stack is at rest and within 1e-5 of the oracle's stack after the same
    100 steps."""
    _need(DEMO)
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout)
    m = re.search(r"t=([\d.]+) contacts=(\d+) top cube y=([-\d.e+]+) vy=([-\d.e+]+)", r.stdout)
    assert m, r.stdout
    assert "Average #iterations" in r.stdout and "Contact Detection" in r.stdout
    sc = _demo_scene()
    ds = pdc.DynamicSystem(sc, sc.params)
    for _ in range(100):
        ds.step(1)
    pos, _, vel, _ = ds.state()
    n_c, _ = ds.ctx.detect_contacts()
    ds.ctx.close()
    assert m.group(1) == "1.000"
    assert int(m.group(2)) == n_c
    assert m.group(3) == "%.6f" % pos[10, 1] and m.group(4) == "%.3e" % vel[10, 1]
    assert np.max(np.abs(vel[1:])) < 1e-3
    o = orc.OracleScene(sc, sc.params)
    o.set_threads(16)
    o.step(100)
    pos_o = o.get_state()[0]
    print("top cube y %.9f, oracle %.9f; largest position difference %.3e"
          % (pos[10, 1], pos_o[10, 1], np.max(np.abs(pos - pos_o))))
    assert np.max(np.abs(pos - pos_o)) < 1e-5
