"""The C++ host layer (m2c_amd/host/m2c_host.hpp, reference-shaped SpaceOperator / TimeIntegrator /
LevelSetOperator / ExactRiemannSolver classes over the C ABI) driving the GPU from a C++ program,
checked bit for bit against the oracle on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from m2c_amd import build, workloads as W
from oracle import pyoracle as po
from tests import cases

pytestmark = pytest.mark.gpu


def demo(args, tmp_path):
    exe = build.DEMO
    assert os.path.exists(exe), "m2c_amd/m2c_demo is built by m2c_amd/build.py"
    out = str(tmp_path / "out.bin")
    r = subprocess.run([exe] + [str(a) for a in args] + [out], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    return np.fromfile(out, dtype=np.uint8).tobytes(), r.stdout


def test_cpp_host_layer_sod(tmp_path):
    raw, txt = demo(["sod", 200, 0.1], tmp_path)
    settings, V0, ID0 = W.sod_1d(200)
    ost, t, n, Vo, _ = cases.run_oracle(settings, V0, ID0, tmax=0.1, maxsteps=100000, vanalbada_pow=False)
    V = np.frombuffer(raw[:200 * 5 * 8], dtype=np.float64).reshape(1, 1, 200, 5)
    assert f"steps={n} " in txt
    assert np.array_equal(V, Vo)


def test_cpp_host_layer_shock_bubble(tmp_path):
    raw, txt = demo(["bubble", 0.125, 12], tmp_path)
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.125)
    ost, t, n, Vo, o, IDo, phio = cases.run_oracle(settings, V0, ID0, maxsteps=12, phi0=phi0, vanalbada_pow=False)
    N = V0.shape[0] * V0.shape[1] * V0.shape[2]
    V = np.frombuffer(raw[:N * 40], dtype=np.float64).reshape(V0.shape)
    ID = np.frombuffer(raw[N * 40:N * 44], dtype=np.int32).reshape(ID0.shape)
    phi = np.frombuffer(raw[N * 44:N * 52], dtype=np.float64).reshape(ID0.shape)
    assert np.array_equal(ID, IDo)
    assert np.array_equal(V, Vo)
    assert np.array_equal(phi, phio)


def deck_file(tmp_path, text):
    path = tmp_path / "input.st"
    path.write_text(text)
    return str(path)


def test_cpp_host_layer_runs_the_sod_deck(tmp_path):
    # Main.cpp's flow from an input deck: IoData, initial condition, march to Time.MaxTime
    raw, txt = demo(["deck", deck_file(tmp_path, W.sod_deck(200))], tmp_path)
    settings, V0, ID0 = W.sod_1d(200)
    # the deck's Output object: snapshots every 100 steps plus the forced first and last
    assert (tmp_path / "sod.pvd").exists() and (tmp_path / "sod_0000.vtr").exists()
    ost, t, n, Vo, _ = cases.run_oracle(settings, V0, ID0, tmax=0.1, maxsteps=100000, vanalbada_pow=False)
    V = np.frombuffer(raw[:200 * 5 * 8], dtype=np.float64).reshape(1, 1, 200, 5)
    assert f"steps={n} " in txt
    assert np.array_equal(V, Vo)


def test_cpp_host_layer_runs_the_shock_bubble_deck(tmp_path):
    text = W.shock_bubble_deck(0.125, reinit_freq=5) + "Time.MaxIts = 12;\n"
    raw, txt = demo(["deck", deck_file(tmp_path, text)], tmp_path)
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.125)
    ost, t, n, Vo, o, IDo, phio = cases.run_oracle(settings, V0, ID0, maxsteps=12, phi0=phi0, vanalbada_pow=False)
    N = V0.shape[0] * V0.shape[1] * V0.shape[2]
    assert "steps=12 " in txt
    V = np.frombuffer(raw[:N * 40], dtype=np.float64).reshape(V0.shape)
    ID = np.frombuffer(raw[N * 40:N * 44], dtype=np.int32).reshape(ID0.shape)
    phi = np.frombuffer(raw[N * 44:N * 52], dtype=np.float64).reshape(ID0.shape)
    assert np.array_equal(ID, IDo) and np.array_equal(V, Vo) and np.array_equal(phi, phio)


def test_cpp_host_layer_fixed_time_step_deck(tmp_path):
    # Time.TimeStep > 0 (SpaceOperator.cpp:2219-2221): a prescribed dt, the last step shortened to
    # land on MaxTime (Main.cpp:535-539)
    text = W.sod_deck(100).replace("MaxTime = 0.1;", "MaxTime = 0.01; TimeStep = 0.0011;")
    raw, txt = demo(["deck", deck_file(tmp_path, text)], tmp_path)
    settings, V0, ID0 = W.sod_1d(100)
    o = po.Oracle(cases.oracle_config(settings, vanalbada_pow=False))
    V, ID = o.ghosted(V0, ID0)
    t, n = 0.0, 0
    while t < 0.01:
        dt = 0.01 - t if t + 0.0011 > 0.01 else 0.0011
        assert o.advance(V, ID, dt) == 0
        t += dt
        n += 1
    assert f"steps={n} " in txt and n >= 10
    Vg = np.frombuffer(raw[:100 * 5 * 8], dtype=np.float64).reshape(1, 1, 100, 5)
    assert np.array_equal(Vg, V[1:-1, 1:-1, 1:-1])


def test_cpp_host_layer_onesided_riemann(tmp_path):
    # ExactRiemannSolver::ComputeOneSidedRiemannSolution of the host layer, air and water
    rng = np.random.default_rng(13)
    n = 300
    q = np.zeros((n, 12))
    d = rng.normal(size=(n, 3))
    q[:, 0:3] = d / np.linalg.norm(d, axis=1)[:, None]
    ids = rng.integers(0, 2, n)
    q[:, 3] = np.where(ids == 0, rng.uniform(0.5, 2.0, n), rng.uniform(990.0, 1050.0, n))
    q[:, 4:7] = rng.uniform(-80.0, 80.0, (n, 3))
    q[:, 7] = np.where(ids == 0, rng.uniform(5e4, 2e5, n), rng.uniform(1e5, 1e7, n))
    q[:, 8] = ids
    q[:, 9:12] = rng.uniform(-80.0, 80.0, (n, 3))
    src = tmp_path / "problems.bin"
    q.tofile(str(src))
    raw, txt = demo(["onesided", str(src)], tmp_path)
    r = np.frombuffer(raw, dtype=np.float64).reshape(n, 12)
    assert f"problems={n}" in txt
    o = po.Oracle(po.make_config(4, 4, 4, (0, 1, 0, 1, 0, 1), [po.BC_SYMMETRY] * 6,
                                 [po.sg(1.4, 0.0), po.sg(6.12, 3.43e8)]))
    for t in range(n):
        st, Vs, mid, Vsm = o.riemann_onesided(q[t, 0:3], q[t, 3:8], int(ids[t]), q[t, 9:12])
        assert r[t, 0] == (0 if st == 0 else 1) and r[t, 1] == mid
        assert np.array_equal(r[t, 2:7], Vs) and np.array_equal(r[t, 7:12], Vsm), t
