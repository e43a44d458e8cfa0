"""Outputs that need the device: internal energy and temperature (m2c_eos_derived), line plots, and a
deck run with outputs through the C++ host driver.

Temperature laws are pinned by identities the laws must satisfy rather than by restating them:
- SG (VarFcnSG.h:70-78): with p_inf = 0, method 1 (cv) gives T = p / ((gamma-1) rho cv), method 2
  (cp = gamma cv) gives the same T, and method 3 with T0 = e0/cv and rho0 > 0 reduces to e/cv;
- NASG (VarFcnNASG.h:61): p = (gamma-1) cv T / (1/rho - b) - pc, the NASG thermal equation of state;
- MG (VarFcnMG.h:172-179): T = T0 + (e - e0)/cv with e from the oracle's p -> E conversion.
The internal energy equals the oracle's conversion e = (rho E)/rho at zero velocity.
"""
import os
import subprocess

import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import build, workloads as W
from oracle import pyoracle as po
from tests import cases

pytestmark = pytest.mark.gpu


def solver_with(mats):
    s = M.Solver(M.make_config(4, 4, 4, (0, 1, 0, 1, 0, 1), [M.BC_SYMMETRY] * 6, mats))
    return s


def oracle_e(mat, rho, p):
    o = po.Oracle(cases.oracle_config(dict(nx=2, ny=2, nz=2, box=(0, 1, 0, 1, 0, 1), bc=[M.BC_SYMMETRY] * 6,
                                           mats=[mat])))
    return np.array([o.p2c(np.array([r, 0.0, 0.0, 0.0, q]))[4] / r for r, q in zip(rho, p)])


def test_sg_temperature_laws():
    rng = np.random.default_rng(0)
    rho, p = rng.uniform(0.5, 2.0, 64), rng.uniform(0.5, 2.0, 64)
    g, cv = 1.4, 717.5
    mats = [M.material(M.EOS_SG, [g, 0.0], tlaw=(cv, 0.0, 0.0, 0.0, 0.0, 0.0)),          # method 1
            M.material(M.EOS_SG, [g, 0.0], tlaw=(0.0, 0.0, 0.0, 0.0, g * cv, 0.0)),      # method 2
            M.material(M.EOS_SG, [g, 0.0], tlaw=(cv, 300.0, 300.0 * cv, 1.2, 0.0, 0.0))]  # method 3
    s = solver_with(mats)
    exact = p / ((g - 1.0) * rho * cv)
    for m in range(3):
        e, T = s.eos_derived(rho, p, m)
        assert np.max(np.abs(e - oracle_e(mats[m], rho, p)) / e) < 4e-16
        assert np.max(np.abs(T - exact) / exact) < 1e-14, m
    s.close()


def test_stiffened_method_three_reference_point():
    # at rho = rho0 the SG law gives T = T0 + (e - e0)/cv exactly as method 1 does
    g, pinf, cv = 6.12, 3.43e8, 4180.0
    mats = [M.material(M.EOS_SG, [g, pinf], tlaw=(cv, 293.0, 1.0e5, 1000.0, 0.0, 0.0)),
            M.material(M.EOS_SG, [g, pinf], tlaw=(cv, 293.0, 1.0e5, 0.0, 0.0, 0.0))]
    s = solver_with(mats)
    p = np.array([1e5, 1e7, 1e9])
    e3, T3 = s.eos_derived(np.full(3, 1000.0), p, 0)
    e1, T1 = s.eos_derived(np.full(3, 1000.0), p, 1)
    assert np.allclose(T3, T1, rtol=1e-12, atol=0)
    s.close()


def test_nasg_and_mg_temperature():
    g, pc, b, q, cv = 1.19, 7.028e8, 6.61e-4, -1177788.0, 3610.0
    mats = [M.material(M.EOS_NASG, [g, pc, b, q], tlaw=(cv, 0, 0, 0, 0, 0)),
            M.material(M.EOS_MG, [8.96e3, 3933.0, 1.99, 1.5, 1.0e3], tlaw=(390.0, 300.0, 0.0, 0.0, 0.0, 0.0))]
    s = solver_with(mats)
    rho, p = np.array([900.0, 1000.0, 1100.0]), np.array([1e5, 1e8, 1e9])
    e, T = s.eos_derived(rho, p, 0)
    assert np.max(np.abs((g - 1.0) * cv * T / (1.0 / rho - b) - pc - p) / (p + pc)) < 1e-12
    rho, p = np.array([8960.0, 9000.0, 9500.0]), np.array([1e5, 1e9, 2e10])
    e, T = s.eos_derived(rho, p, 1)
    eo = oracle_e(mats[1], rho, p)
    assert np.max(np.abs(e - eo) / np.abs(eo)) < 1e-14
    assert np.max(np.abs(T - (300.0 + (e - 1.0e3) / 390.0)) / T) < 1e-14
    e, T = s.eos_derived(rho, p, np.array([0, 1, 1]))   # per-point materials
    assert np.isfinite(T).all()
    s.close()


def test_line_plots_and_temperature_probes(tmp_path):
    d = str(tmp_path)
    text = W.shock_bubble_deck(0.0625, reinit_freq=3) + f"""
Equations.Material[0].StiffenedGasModel.SpecificHeatAtConstantVolume = 4180.0;
Equations.Material[1].StiffenedGasModel.SpecificHeatAtConstantVolume = 717.5;
under Output {{ Prefix = "{d}/"; Frequency = 3; Density = On; Pressure = On; InternalEnergyPerUnitMass = On;
  Temperature = On; LevelSet0 = On;
  LinePlot[0].FileName = "axis"; LinePlot[0].NumberOfPoints = 24; LinePlot[0].Frequency = 3;
  LinePlot[0].X0 = 0.5; LinePlot[0].Y0 = 6.5; LinePlot[0].Z0 = 6.5;
  LinePlot[0].Xmax = 23.5; LinePlot[0].Ymax = 6.5; LinePlot[0].Zmax = 6.5;
  under Probes {{ Frequency = 1; Temperature = "T.txt"; Node[0].X = 10.5; Node[0].Y = 6.5; Node[0].Z = 6.5; }}
}}
"""
    cfg, run, _ = M.config_from_deck(text)
    V, ID, phi = M.initial_condition_from_deck(text)
    s = M.Solver(cfg)
    s.upload(V, ID)
    s.upload_levelset(phi)
    o = M.Output(text)
    assert o.needs_device
    o.write(0.0, 0.0, 0, V, ID, phi, solver=s, force=True)
    for step in range(1, 7):
        st = s.run(1e300, 1, sync_every=1)
        V, ID = s.download()
        phi = s.download_levelset()
        o.write(st.t, st.dt, step, V, ID, phi, solver=s)
    o.close()
    e, T = s.eos_derived(V[..., 0], V[..., 4], ID)
    s.close()
    lines = open(os.path.join(d, "axis_0002.txt")).read().splitlines()   # frames at steps 0, 3, 6
    assert lines[0] == "## Line: (5.000000e-01, 6.500000e+00, 6.500000e+00) -> (2.350000e+01, 6.500000e+00, 6.500000e+00)"
    assert lines[1] == "## Number of points: 24 (h = 1.000000e+00)"
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[4:]])
    assert rows.shape == (24, 9)
    # the points are the cell centres of row (j, k) = (6, 6): the values are the cells' own
    f = lambda a: np.array([float("%.8e" % v) for v in a])  # noqa: E731
    assert np.array_equal(rows[:, 0], f(np.arange(24.0)))
    assert np.array_equal(rows[:, 1], f(V[6, 6, :, 0]))
    assert np.array_equal(rows[:, 5], f(V[6, 6, :, 4]))
    assert np.array_equal(rows[:, 6], f(T.reshape(ID.shape)[6, 6, :]))
    assert np.array_equal(rows[:, 7], f(ID[6, 6, :].astype(float)))
    assert np.array_equal(rows[:, 8], f(phi[6, 6, :]))
    probe = [ln for ln in open(os.path.join(d, "T.txt")) if not ln.startswith("##")]
    assert len(probe) == 7 and float(probe[-1].split()[2]) == float("%.8e" % T.reshape(ID.shape)[6, 6, 10])


def test_host_driver_writes_the_outputs_of_a_deck(tmp_path):
    d = str(tmp_path)
    text = W.shock_bubble_deck(0.125, reinit_freq=5) + f"""Time.MaxIts = 12;
under Output {{ Prefix = "{d}/"; Solution = "sb"; Frequency = 5; Density = On; Velocity = On; Pressure = On;
  MaterialID = On; LevelSet0 = On; }}
"""
    path = tmp_path / "input.st"
    path.write_text(text)
    out = str(tmp_path / "out.bin")
    r = subprocess.run([build.DEMO, "deck", str(path), out], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "steps=12 " in r.stdout
    pvd = open(os.path.join(d, "sb.pvd")).read()
    frames = [ln for ln in pvd.splitlines() if "<DataSet" in ln]
    assert [f.split('file="')[1][:-3] for f in frames] == ["sb_0000.vtr", "sb_0001.vtr", "sb_0002.vtr", "sb_0003.vtr"]
    from tests.test_output import read_vtr
    snap = read_vtr(os.path.join(d, "sb_0003.vtr"))
    raw = np.fromfile(out, dtype=np.uint8).tobytes()
    settings, V0, ID0, _ = W.shock_bubble(scale=0.125)
    N = ID0.size
    V = np.frombuffer(raw[:N * 40], dtype=np.float64).reshape(V0.shape)
    ID = np.frombuffer(raw[N * 40:N * 44], dtype=np.int32)
    phi = np.frombuffer(raw[N * 44:N * 52], dtype=np.float64)
    assert np.array_equal(snap["density"], V[..., 0].ravel()) and np.array_equal(snap["pressure"], V[..., 4].ravel())
    assert np.array_equal(snap["velocity"], V[..., 1:4].reshape(-1, 3))
    assert np.array_equal(snap["materialid"], ID.astype(float)) and np.array_equal(snap["levelset0"], phi)
