"""Input-deck reader (m2c_config_from_deck, m2c_initial_condition_from_deck): host code, no GPU.

What pins it:
- the workloads' decks read back to the very configuration and initial state (bit for bit) that the
  workload builders produce directly (sod_1d, shock_bubble);
- the syntax of parser/Parser.y and parser/Lexer.l case by case (precedence, integer division,
  signed literals, Fortran exponents, comments, `end`), and the assignment rules of
  parser/Assigner.h (integer fields refuse doubles, keyword fields refuse unknown keywords);
- IoData's defaults and aliases, field by field against the constructors in IoData.cpp;
- the level set of the shapes against closed-form signed distances (capsule, finite cylinder,
  axis-aligned and rotated boxes), the sign of phi against the material ID, and the distance
  property |phi(a) - phi(b)| <= |a - b|.
The reference's own parser is not built here (flex and bison generate it); the reader's syntax is
pinned by the cases below rather than by a reference build.
"""
import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W

MESH = """under Mesh {
  X0 = 0; Xmax = 8; Y0 = 0; Ymax = 4; Z0 = 0; Zmax = 2;
  NumberOfCellsX = 8; NumberOfCellsY = 4; NumberOfCellsZ = 2;
  BoundaryConditionX0 = Inlet; BoundaryConditionXmax = Outlet;
  BoundaryConditionY0 = SlipWall; BoundaryConditionYmax = StickWall;
  BoundaryConditionZ0 = Symmetry; BoundaryConditionZmax = Symmetry;
}
Equations.Material[0].EquationOfState = StiffenedGas;
"""


def read(extra, base=MESH):
    return M.config_from_deck(base + extra)


def err(extra, base=MESH):
    with pytest.raises(M.DeckError) as e:
        M.config_from_deck(base + extra)
    return e.value


def same_config(a, b):
    return bytes(a) == bytes(b)


# ------------------------------------------------------------------ workloads as decks

def test_sod_deck_reads_back_to_the_workload():
    text = W.sod_deck(200)
    cfg, run, ignored = M.config_from_deck(text)
    settings, V0, ID0 = W.sod_1d(200)
    assert same_config(cfg, W.config_from(settings))
    assert run.max_time == 0.1 and run.max_its == 2**31 - 1 and run.time_step == -1.0
    line = [ln.startswith("under Output") for ln in text.splitlines()].index(True) + 1
    assert ignored == [f"ignored Output (line {line})"]
    V, ID, phi = M.initial_condition_from_deck(text)
    assert phi is None
    assert np.array_equal(V, V0) and np.array_equal(ID, ID0)


@pytest.mark.parametrize("scale", [0.0625, 0.125])
def test_shock_bubble_deck_reads_back_to_the_workload(scale):
    text = W.shock_bubble_deck(scale, reinit_freq=3)
    cfg, run, ignored = M.config_from_deck(text)
    settings, V0, ID0, phi0 = W.shock_bubble(scale=scale, reinit_freq=3)
    assert same_config(cfg, W.config_from(settings))
    assert ignored == []
    V, ID, phi = M.initial_condition_from_deck(text)
    assert np.array_equal(V, V0) and np.array_equal(ID, ID0) and np.array_equal(phi, phi0)


@pytest.mark.parametrize("scale", [0.0625, 0.125])
def test_stretched_shock_bubble_deck_reads_back_to_the_workload(scale):
    # Mesh.ControlPointX/Y/Z: the cell counts, control points, cell centres of the shapes and the
    # sphere's level set are the workload's, bit for bit
    text = W.shock_bubble_deck(scale, reinit_freq=3, stretched=True)
    assert "ControlPointY[2]" in text and "NumberOfCells" not in text
    cfg, run, ignored = M.config_from_deck(text)
    settings, V0, ID0, phi0 = W.shock_bubble_stretched(scale=scale, reinit_freq=3)
    assert (cfg.nx, cfg.ny, cfg.nz) == (settings["nx"], settings["ny"], settings["nz"])
    assert same_config(cfg, W.config_from(settings))
    assert ignored == []
    V, ID, phi = M.initial_condition_from_deck(text)
    assert np.array_equal(V, V0) and np.array_equal(ID, ID0) and np.array_equal(phi, phi0)


def test_control_points_set_the_cell_count():
    base = MESH.replace("NumberOfCellsX = 8; NumberOfCellsY = 4; NumberOfCellsZ = 2;",
                        "NumberOfCellsY = 4; NumberOfCellsZ = 2;")
    cfg, _, _ = read("Mesh.ControlPointX[3].Coordinate = 2.0; Mesh.ControlPointX[3].CellWidth = 0.25;", base=base)
    assert cfg.nx == 32 and cfg.mesh_npts[0] == 1 and cfg.mesh_npts[1] == 0
    assert (cfg.mesh_pts[0][0][0], cfg.mesh_pts[0][0][1]) == (2.0, 0.25)
    # MeshGenerator.cpp:41-48: an axis needs cells or control points, not both, and points inside it
    assert err("Mesh.ControlPointX[0].CellWidth = 0.5;").code == M.ERR_ARG
    assert err("", base=base).code == M.ERR_ARG
    assert err("Mesh.ControlPointX[0].Coordinate = 9.0; Mesh.ControlPointX[0].CellWidth = 0.5;", base=base).code == M.ERR_ARG
    # a width left at its default (0) has no cell count
    assert err("Mesh.ControlPointX[0].Coordinate = 2.0;", base=base).code == M.ERR_ARG


# ------------------------------------------------------------------ syntax (Parser.y, Lexer.l)

def test_integer_and_double_arithmetic():
    cfg, _, _ = read("""Mesh.NumberOfCellsX = 17/2;            // integer division: 8
Mesh.Xmax = 7/2.0;                       // double: 3.5
Mesh.Ymax = 1 + 2*3;                     // precedence: 7
Mesh.Zmax = (1 + 2)*3 - 4/3;             // 9 - 1 = 8 (integer), assigned to a double
Mesh.X0 = -7/2;                          // C division truncates toward zero: -3
Mesh.Y0 = 2.5d-1;                        // Fortran exponent
Mesh.Z0 = .5e1 - 1.e0 - 4;               // 5 - 1 - 4 = 0
""")
    assert cfg.nx == 8 and cfg.x1 == 3.5 and cfg.y1 == 7.0 and cfg.z1 == 8.0
    assert cfg.x0 == -3.0 and cfg.y0 == 0.25 and cfg.z0 == 0.0


def test_mixed_arithmetic_promotes_left_to_right():
    cfg, _, _ = read("Mesh.Xmax = 1/2*4.0; Mesh.Ymax = 1.0*1/2; Mesh.Zmax = 3 - 1.5 - 1;")
    assert cfg.x1 == 0.0 and cfg.y1 == 0.5 and cfg.z1 == 0.5


def test_a_signed_literal_after_an_operand_is_a_syntax_error():
    # Lexer.l makes "-1" one number, so "3 -1" is two numbers in a row
    e = err("Mesh.Xmax = 3 -1;")
    assert e.code == M.ERR_ARG and "line 9" in str(e)
    cfg, _, _ = read("Mesh.Xmax = 3 - 1;")
    assert cfg.x1 == 2.0


def test_comments_strings_and_end():
    cfg, _, ignored = read("""/* a block
comment over lines */ Mesh.Xmax = 2; // trailing
under Output { Prefix = "out/run-1_a.b/"; }
end
this text is never read { ] ;
""")
    assert cfg.x1 == 2.0 and ignored == ["ignored Output (line 11)"]


@pytest.mark.parametrize("bad,where", [("Mesh.Xmax = ;", "line 9"), ("Mesh.Xmax = 2", "line 9"),
                                       ("under Mesh { Xmax = 2; ", "line 9"), ("}", "line 9"),
                                       ("Mesh.Xmax = 2 $;", "line 9"), ("/* open", "line 9"),
                                       ("Mesh.Xmax = \"a b\";", "line 9"), ("Mesh[1.5].X0 = 1;", "line 9")])
def test_syntax_errors_name_their_line(bad, where):
    e = err(bad)
    assert e.code == M.ERR_ARG and where in str(e), str(e)


# ------------------------------------------------------------------ assignment rules (Assigner.h)

def test_integer_fields_refuse_doubles_and_doubles_take_integers():
    e = err("Mesh.NumberOfCellsX = 8.0;")
    assert e.code == M.ERR_ARG and "integer" in str(e)
    cfg, _, _ = read("Equations.Material[0].StiffenedGasModel.PressureConstant = 3;")
    assert cfg.mat[0].p[1] == 3.0


def test_keywords():
    e = err("Space.NavierStokes.Flux = Rusanov;")
    assert e.code == M.ERR_ARG and "Rusanov" in str(e)
    e = err("Space.NavierStokes.Flux = 2;")
    assert e.code == M.ERR_ARG


# ------------------------------------------------------------------ objects, aliases, defaults

def test_io_data_defaults():
    cfg, run, _ = read("")
    assert cfg.flux == M.FLUX_HLLC and cfg.roe_delta == 0.2           # SchemeData
    assert cfg.rec_linear == 1 and cfg.limiter == M.LIM_NONE and cfg.gminmod_alpha == 1.2
    assert cfg.slope_zero_near_interface == 0                         # ReconstructionData
    assert cfg.ts == M.TS_RK2 and cfg.cfl == 0.5                      # ExplicitData, TsData
    assert cfg.mp_flux == 0 and cfg.mp_recon == 0                     # MultiPhaseData
    r = cfg.riemann                                                   # ExactRiemannSolverData
    assert (r.maxIts_main, r.maxIts_bracket, r.maxIts_shock, r.numSteps_rarefaction) == (200, 100, 200, 200)
    assert (r.tol_main, r.tol_shock, r.tol_rarefaction) == (1e-4, 1e-12, 1e-8)
    assert (r.min_pressure, r.failure_threshold, r.pressure_at_failure) == (-1e8, 0.2, 1e-8)
    assert cfg.ls_matid == -1 and cfg.reinit_freq == -1 and cfg.reinit_maxits == 30
    assert cfg.reinit_cfl == 0.8 and cfg.reinit_tol == 2e-4 and list(cfg.ls_bc) == [1] * 6
    m = cfg.mat[0]                                                    # MaterialModelData
    assert m.eos == M.EOS_SG and m.p[0] == 1.4 and m.p[1] == 0.0
    assert m.rhomin == 0.0 and m.pmin == -M.DBL_MAX and m.rhomax == M.DBL_MAX and m.failsafe_density == 1e-10
    assert list(cfg.inlet) == [1e-6, 0, 0, 0, 0] and list(cfg.outlet) == [1e-6, 0, 0, 0, 0]   # StateVariable
    assert list(cfg.bc) == [M.BC_INLET, M.BC_OUTLET, M.BC_SLIPWALL, M.BC_STICKWALL, M.BC_SYMMETRY, M.BC_SYMMETRY]
    assert list(run.initial_state) == [1e-6, 0, 0, 0, 0] and run.initial_material == 0


def test_config_defaults_entry_point():
    c = M.config_defaults()
    assert c.nmat == 0 and (c.nx, c.ny, c.nz) == (-1, -1, -1) and list(c.bc) == [0] * 6
    assert c.limiter == M.LIM_NONE and c.flux == M.FLUX_HLLC and c.cfl == 0.5


@pytest.mark.parametrize("eos,block,expect", [
    ("NobleAbelStiffenedGas", "NobleAbelStiffenedGasModel { SpecificHeatRatio = 1.19; PressureConstant = 7028e5;"
     " VolumeConstant = 6.61e-4; EnergyConstant = -1177788; }", (M.EOS_NASG, [1.19, 7.028e8, 6.61e-4, -1177788.0])),
    ("MieGruneisen", "MieGruneisenModel { ReferenceDensity = 2.0; }", (M.EOS_MG, [2.0, 3.933e6, 1.99, 1.5, 0.0])),
    ("JonesWilkinsLee", "JonesWilkinsLeeModel { Omega = 0.3; Rho0 = 1.6e-3; }",
     (M.EOS_JWL, [0.3, 3.712e11, 3.23e9, 4.15, 0.95, 1.6e-3])),
])
def test_equations_of_state(eos, block, expect):
    cfg, _, _ = read(f"under Equations {{ under Material[1] {{ EquationOfState = {eos}; under {block} }} }}")
    assert cfg.nmat == 2 and cfg.mat[1].eos == expect[0]
    assert list(cfg.mat[1].p[:len(expect[1])]) == expect[1]


def test_aliases_share_one_object_and_the_last_assignment_wins():
    cfg, _, _ = read("""BoundaryConditions.Inlet.Density = 2.0;
BoundaryConditions.Farfield0.Pressure = 3.0;
BoundaryConditions.Farfield.Density = 4.0;
BoundaryConditions.Farfield1.VelocityX = 5.0;
NavierStokesEquations.Material[0].StiffenedGasModel.SpecificHeatRatio = 1.3;
""")
    assert list(cfg.inlet) == [4.0, 0, 0, 0, 3.0] and cfg.outlet[1] == 5.0 and cfg.mat[0].p[0] == 1.3


def test_bare_names_resolve_through_enclosing_objects():
    cfg, _, _ = read("""under Equations {
  under Material[1] { EquationOfState = StiffenedGas; }
  Material[2].EquationOfState = StiffenedGas;
}
under Space {
  under LevelSet[0] {
    MaterialID = 1;
    under Reinitialization { Frequency = 3; MaxIts = 10; }
  }
  NavierStokes.Flux = Roe;
  under NavierStokes { under Reconstruction { Limiter = GeneralizedMinMod; } EntropyFixCoefficient = 0.1; }
}
under Mesh { Time.CFL = 0.4; }
under Time { under Explicit { Type = RungeKutta3; } MaxIts = 12; }
""")
    assert cfg.nmat == 3 and cfg.ls_matid == 1 and cfg.reinit_freq == 3 and cfg.reinit_maxits == 10
    assert cfg.flux == M.FLUX_ROE and cfg.limiter == M.LIM_GENMINMOD and cfg.roe_delta == 0.1
    assert cfg.cfl == 0.4 and cfg.ts == M.TS_RK3


def test_riemann_normal_keywords():
    # MultiPhaseData::riemann_normal: LevelSet 0, Mesh 1 (the default), Average 2 (IoData.cpp:1445-1448)
    assert read("")[0].mp_riemann_normal == 1
    for kw, code in (("LevelSet", 0), ("Mesh", 1), ("Average", 2)):
        cfg, _, _ = read(f"MultiPhase.RiemannNormal = {kw};")
        assert cfg.mp_riemann_normal == code, kw


def test_constant_reconstruction_depth():
    assert read("")[0].mp_conrec_depth == 0.0
    assert read("MultiPhase.ConstantReconstructionDepth = 0.25;")[0].mp_conrec_depth == 0.25


def test_navier_stokes_flux_keywords():
    # SchemeData::flux: Roe 0, LocalLaxFriedrichs 1, HLLC 2, Godunov 3 (IoData.cpp:1170-1172)
    for kw, code in (("Roe", M.FLUX_ROE), ("LocalLaxFriedrichs", M.FLUX_LLF), ("HLLC", M.FLUX_HLLC),
                     ("Godunov", M.FLUX_GODUNOV)):
        cfg, _, _ = read(f"Space.NavierStokes.Flux = {kw};")
        assert cfg.flux == code, kw


def test_multiphase_flux_keywords_map_to_the_library_numbering():
    # MultiPhaseData numbers Exact 0, Numerical 1, LocalLaxFriedrichs 2 (IoData.cpp:1425-1428)
    for kw, code in (("Exact", 1), ("Numerical", 0), ("LocalLaxFriedrichs", 2)):
        cfg, _, _ = read(f"MultiPhase.Flux = {kw}; MultiPhase.ReconstructionAtInterface = Linear;")
        assert cfg.mp_flux == code and cfg.mp_recon == 1


def test_exact_riemann_solution_and_level_set_options():
    cfg, _, _ = read("""under ExactRiemannSolution { MaxIts = 50; MaxItsBracketing = 20; MaxItsShock = 30;
  IntegrationSteps = 100; Tolerance = 1e-6; ToleranceShock = 1e-10; ToleranceRarefaction = 1e-7;
  MinPressure = -1e5; FailureThreshold = 0.3; PrescribedPressureUponFailure = 1e-6; }
under Space { under LevelSet[3] { MaterialID = 0; BoundaryConditionX0 = LinearExtrapolation;
  BoundaryConditionYmax = NonNegative; under Reinitialization { CFL = 0.5; ConvergenceTolerance = 1e-5; } } }
Space.NavierStokes.Reconstruction.SlopeNearInterface = Zero;
""")
    r = cfg.riemann
    assert (r.maxIts_main, r.maxIts_bracket, r.maxIts_shock, r.numSteps_rarefaction) == (50, 20, 30, 100)
    assert (r.tol_main, r.tol_shock, r.tol_rarefaction, r.min_pressure) == (1e-6, 1e-10, 1e-7, -1e5)
    assert (r.failure_threshold, r.pressure_at_failure) == (0.3, 1e-6)
    assert cfg.ls_matid == 0 and list(cfg.ls_bc) == [2, 1, 1, 3, 1, 1]
    assert cfg.reinit_cfl == 0.5 and cfg.reinit_tol == 1e-5 and cfg.slope_zero_near_interface == 1


@pytest.mark.parametrize("extra", [
    "Mesh.NumberOfCellsZ = 0;",                       # MeshGenerator.cpp:41-44
    "Mesh.BoundaryConditionZmax = None;",            # SpaceOperator.cpp:1731
    "Equations.Material[2].EquationOfState = StiffenedGas;",   # Main.cpp:112-115 (gap at 1)
    "DefaultInitialState.MaterialID = 4;",
    "GeometricEntities.Sphere[0].Side = Exterior;",  # SpaceOperator.cpp:866-869
    "GeometricEntities.Plane[0].Inclusion = Union;",
])
def test_values_the_reference_rejects(extra):
    e = err(extra)
    assert e.code == M.ERR_ARG, str(e)


def test_missing_mesh_and_materials():
    assert err("", base="").code == M.ERR_ARG
    assert err("", base=MESH.replace("Equations.Material[0].EquationOfState = StiffenedGas;", "")).code == M.ERR_ARG


@pytest.mark.parametrize("extra", [
    "Time.Type = Implicit;", "Time.LocalTimeStepping = On;", "Time.ConvergenceTolerance = 1e-6;",
    "Mesh.Type = Cylindrical;", "Mesh.BoundaryConditionX0 = Overset;",
    "Equations.Material[0].EquationOfState = ANEOSBirchMurnaghanDebye;",
    "Equations.Material[0].ViscosityModel.Type = Artificial;",
    "Equations.MaterialTransition[0].FromMaterialID = 0;",
    "Space.NavierStokes.Reconstruction.VariableType = Conservative;",
    "Space.NavierStokes.Reconstruction.Fixes.Sphere[0].Radius = 1;",
    "Space.NavierStokes.Smoothing.Type = Box;",
    "Space.LevelSet[0].Bandwidth = 6;", "Space.LevelSet[0].Solver = FiniteVolume;",
    "Space.LevelSet[0].Reinitialization.FirstLayerTreatment = ConstrainedMethod1;",
    "Space.LevelSet[0].MaterialID = 0; Space.LevelSet[1].MaterialID = 0;",
    "MultiPhase.PhaseChange = Extrapolation;",
    "MultiPhase.LatentHeatTransfer = On;", "MultiPhase.LevelSetCorrectionFrequency = 5;",
    "EmbeddedBoundaryMethod.EmbeddedSurfaces.Surface[0].MeshFile = \"a.top\";", "Laser.SourcePower = 1;",
    "GeometricEntities.Spheroid[0].Length = 1;", "InitialCondition.UserDataFile = \"ic.txt\";",
])
def test_options_not_provided(extra):
    e = err(extra)
    assert e.code == M.ERR_UNSUPPORTED, str(e)


def test_ignored_objects_are_reported():
    _, _, ignored = read("""under Output { Prefix = "r/"; Frequency = 5; under Probes { Frequency = 1; } }
TerminalVisualization.Variable = Pressure;
Mesh.Comment = 3;
""")
    assert ignored == ["ignored Mesh.Comment (line 11)", "ignored Output (line 9)",
                       "ignored TerminalVisualization (line 10)"]


def test_a_failed_read_leaves_the_outputs_alone():
    cfg = M.Config()
    cfg.nx = 77
    run = M.DeckRun()
    buf = M.C.create_string_buffer(256)
    rc = M.lib().m2c_config_from_deck(b"Mesh.Xmax = ;", M.C.byref(cfg), M.C.byref(run), buf, 256)
    assert rc == M.ERR_ARG and cfg.nx == 77 and b"line 1" in buf.value
    # a short message buffer is truncated, never overrun
    small = M.C.create_string_buffer(b"\xff" * 8, 8)
    rc = M.lib().m2c_config_from_deck(b"Mesh.Xmax = ;", M.C.byref(cfg), None, small, 5)
    assert rc == M.ERR_ARG and small.raw[4] == 0 and small.raw[5:] == b"\xff" * 3


# ------------------------------------------------------------------ initial conditions

BOX = """under Mesh {
  X0 = -1; Xmax = 1; Y0 = -1; Ymax = 1; Z0 = -1; Zmax = 1;
  NumberOfCellsX = 40; NumberOfCellsY = 40; NumberOfCellsZ = 40;
  BoundaryConditionX0 = Symmetry; BoundaryConditionXmax = Symmetry;
  BoundaryConditionY0 = Symmetry; BoundaryConditionYmax = Symmetry;
  BoundaryConditionZ0 = Symmetry; BoundaryConditionZmax = Symmetry;
}
under Equations { Material[0].EquationOfState = StiffenedGas; Material[1].EquationOfState = StiffenedGas; }
Space.LevelSet[0].MaterialID = 1;
DefaultInitialState.Density = 1.0; DefaultInitialState.Pressure = 1.0;
"""


def centres(n=40, lo=-1.0, hi=1.0):
    h = (hi - lo) / n
    c = lo + 0.5 * h + np.arange(n) * h
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    return x, y, z, h


def shape_ic(geometry):
    return M.initial_condition_from_deck(BOX + "under GeometricEntities {\n" + geometry + "\n}\n")


def check_distance_function(phi, ID, h, tol=1e-12):
    """phi < 0 exactly on the cells of material 1, and |phi(a) - phi(b)| <= |a - b| for neighbours."""
    assert np.array_equal(phi < 0, ID == 1)
    for ax in range(3):
        assert np.max(np.abs(np.diff(phi, axis=ax))) <= h * (1 + tol)


AIR = "under InitialState { MaterialID = 1; Density = 0.01; Pressure = 2.0; }"


def test_capsule_is_the_distance_to_its_axis_segment_minus_the_radius():
    V, ID, phi = shape_ic(f"""under CylinderWithSphericalCaps[0] {{
  CylinderCenter_x = 0.1; CylinderCenter_y = -0.05; Axis_x = 1; Axis_y = 1; Axis_z = 0.5;
  CylinderRadius = 0.3; CylinderHeight = 0.8; FrontSphericalCap = On; BackSphericalCap = On;
  {AIR} }}""")
    x, y, z, h = centres()
    a = np.array([1.0, 1.0, 0.5]) / 1.5
    p = np.stack([x - 0.1, y + 0.05, z], axis=-1)
    t = np.clip(p @ a, -0.4, 0.4)
    exact = np.linalg.norm(p - t[..., None] * a, axis=-1) - 0.3
    assert np.max(np.abs(phi - exact)) < 1e-13
    check_distance_function(phi, ID, h)
    assert np.all(V[ID == 1][:, 0] == 0.01) and np.all(V[ID == 0][:, 0] == 1.0)


def test_finite_cylinder_is_the_box_distance_in_the_meridian_plane():
    V, ID, phi = shape_ic(f"""under CylinderAndCone[0] {{
  BaseCenter_x = -0.5; BaseCenter_y = 0.1; BaseCenter_z = -0.2; Axis_x = 1; Axis_y = 0; Axis_z = 0;
  CylinderRadius = 0.4; CylinderHeight = 1.1; ConeHeight = 0.0;
  {AIR} }}""")
    x, y, z, h = centres()
    ax_, r = x + 0.5, np.hypot(y - 0.1, z + 0.2)
    qa = np.maximum(-ax_, ax_ - 1.1)          # outside the slab 0 < ax < 1.1
    qr = r - 0.4
    exact = np.where((qa > 0) | (qr > 0), np.hypot(np.maximum(qa, 0), np.maximum(qr, 0)), np.maximum(qa, qr))
    assert np.max(np.abs(phi - exact)) < 1e-13
    check_distance_function(phi, ID, h)


def test_cone_tip_distance_and_sign():
    # cylinder of height 0.4 and a 45-degree cone (the axis defaults to z, CylinderConeData): the
    # generator runs from (0.4, R) to the tip (0.4 + R, 0)
    V, ID, phi = shape_ic(f"""under CylinderAndCone[0] {{
  BaseCenter_x = -0.6; Axis_x = 1; Axis_z = 0; CylinderRadius = 0.3; CylinderHeight = 0.4; ConeHeight = 5.0;
  ConeOpeningAngleInDegrees = 45; {AIR} }}""")
    x, y, z, h = centres()
    check_distance_function(phi, ID, h)
    ax_, r = x + 0.6, np.hypot(y, z)
    on_axis_beyond_tip = (r < 1e-12) & (ax_ > 0.7)
    assert not np.any(on_axis_beyond_tip)   # the grid has no cell on the axis; use the nearest points
    beyond = (ax_ > 0.75) & (r < 0.05)
    tip = np.hypot(ax_ - 0.7, r)
    # past the tip, within the cone's axial shadow, the nearest boundary point is the tip itself
    assert np.max(np.abs(phi[beyond] - tip[beyond])) < 1e-13


def test_boxes_axis_aligned_and_rotated():
    x, y, z, h = centres()
    V, ID, phi = shape_ic(f"""under Parallelepiped[0] {{
  X0 = -0.45; Y0 = -0.3; Z0 = -0.25; Ax = 0.9; By = 0.5; Cz = 0.65; {AIR} }}""")
    q = np.stack([np.abs(x - 0.0) - 0.45, np.abs(y + 0.05) - 0.25, np.abs(z - 0.075) - 0.325], axis=-1)
    exact = np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(axis=-1), 0)
    assert np.max(np.abs(phi - exact)) < 1e-13
    check_distance_function(phi, ID, h)

    # a cube of side 0.8 centred at the origin, rotated 30 degrees about z then 20 about x
    a, b = np.radians(30.0), np.radians(20.0)
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    R = Rx @ Rz
    e = 0.8 * R.T     # rows: the rotated edge vectors
    O = -0.5 * (e[0] + e[1] + e[2])
    def fields(names, v):
        return " ".join(f"{n} = {float(t)!r};" for n, t in zip(names, v))
    V, ID, phi = shape_ic(f"""under Parallelepiped[0] {{
  {fields(("X0", "Y0", "Z0"), O)} {fields(("Ax", "Ay", "Az"), e[0])}
  {fields(("Bx", "By", "Bz"), e[1])} {fields(("Cx", "Cy", "Cz"), e[2])} {AIR} }}""")
    p = np.stack([x, y, z], axis=-1) @ R     # back into the cube's frame
    q = np.abs(p) - 0.4
    exact = np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(axis=-1), 0)
    assert np.max(np.abs(phi - exact)) < 1e-12
    check_distance_function(phi, ID, h)


def test_a_left_handed_box_is_rejected():
    e = pytest.raises(M.DeckError, shape_ic, f"under Parallelepiped[0] {{ Ax = 1; By = 1; Cz = -1; {AIR} }}")
    assert e.value.code == M.ERR_ARG


def test_shapes_apply_by_kind_then_index_and_the_level_set_keeps_the_nearest_surface():
    V, ID, phi = shape_ic("""under Sphere[1] { Center_x = 0.2; Radius = 0.5;
  under InitialState { MaterialID = 1; Density = 3.0; Pressure = 1.0; } }
under Sphere[0] { Center_x = -0.2; Radius = 0.5;
  under InitialState { MaterialID = 1; Density = 2.0; Pressure = 1.0; } }
under Plane[0] { Point_x = 0.0; Normal_x = 1; under InitialState { MaterialID = 0; Density = 5.0; Pressure = 1.0; } }
""")
    x, y, z, h = centres()
    d0 = np.sqrt((x + 0.2) ** 2 + y ** 2 + z ** 2) - 0.5
    d1 = np.sqrt((x - 0.2) ** 2 + y ** 2 + z ** 2) - 0.5
    # planes come first; sphere 1 is applied after sphere 0 and wins where they overlap
    rho = np.where(x > 0, 5.0, 1.0)
    rho = np.where(d0 < 0, 2.0, rho)
    rho = np.where(d1 < 0, 3.0, rho)
    assert np.array_equal(V[..., 0], rho)
    assert np.array_equal(ID, ((d0 < 0) | (d1 < 0)).astype(np.int32))
    # the plane belongs to material 0 and leaves the level set of material 1 alone
    diag = np.sqrt(12.0)
    ref = np.full_like(phi, diag)
    for d in (d0, d1):
        ref = np.where(np.abs(d) < np.abs(ref), d, ref)
    assert np.array_equal(phi, ref)


def test_host_driver_reports_a_bad_deck_before_touching_the_device(tmp_path):
    import subprocess
    from m2c_amd import build
    path = tmp_path / "bad.st"
    path.write_text(MESH + "Space.NavierStokes.Flux = HLLE;\n")
    r = subprocess.run([build.DEMO, "deck", str(path), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 3 and "line 9" in r.stderr and "HLLE" in r.stderr, r.stderr
    path.write_text(MESH + "Time.Type = Implicit;\n")
    r = subprocess.run([build.DEMO, "deck", str(path), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 3 and "status 4" in r.stderr, r.stderr


def test_viscosity_and_heat_diffusion_models():
    cfg, _, _ = read("""under Equations {
  under Material[0] { under ViscosityModel { Type = Constant; DynamicViscosity = 2.5e-3; BulkViscosity = 1e-4; }
                      under HeatDiffusionModel { Type = Constant; Diffusivity = 0.6; } }
  under Material[1] { EquationOfState = StiffenedGas; under ViscosityModel { Type = Sutherland; } }
  under Material[2] { EquationOfState = StiffenedGas; HeatDiffusionModel.Diffusivity = 3.0; }
}
""")
    m0, m1, m2 = cfg.mat[0], cfg.mat[1], cfg.mat[2]
    assert (m0.viscosity, m0.mu, m0.bulk, m0.kappa) == (M.VISC_CONSTANT, 2.5e-3, 1e-4, 0.6)
    # ViscosityModelData defaults (Sutherland: mu0 1.716e-5, T0 273, S 111)
    assert (m1.viscosity, m1.mu, m1.su_T0, m1.su_S, m1.bulk) == (M.VISC_SUTHERLAND, 1.716e-5, 273.0, 111.0, 0.0)
    # a diffusivity without Type = Constant is not used (HeatDiffusionModelData type None)
    assert m2.viscosity == M.VISC_NONE and m2.kappa == 0.0
