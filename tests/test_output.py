"""Outputs of a deck's Output object (m2c_output_*): host code, checked without a GPU.

Pinned by:
- the collection file text of Output.cpp:28-46 and :392-404, byte for byte;
- the snapshot read back by an independent VTK XML reader below: coordinates are the cell centres,
  the arrays equal the state bit for bit, in the reference's field order;
- isTimeToWrite (Utils.cpp:293-313) on exact binary times;
- probe values against closed forms: a probe at a cell centre returns the cell value exactly; a
  linear field is reproduced inside; the face-ghost rules near walls and inlets (mirror, inlet state,
  linear extrapolation, the non-negative wall cap) give values fixed by hand; edge and corner
  ghosts are replaced by the mean of the valid corners (a uniform field stays uniform).
Internal energy, temperature and line plots need the device (tests/test_gpu_output.py).
"""
import os
import re
import struct

import numpy as np
import pytest

import m2c_amd as M

MESH = """under Mesh {{
  X0 = 0; Xmax = 8; Y0 = 0; Ymax = 4; Z0 = 0; Zmax = 2;
  NumberOfCellsX = 8; NumberOfCellsY = 4; NumberOfCellsZ = 2;
  BoundaryConditionX0 = {bx0}; BoundaryConditionXmax = Outlet;
  BoundaryConditionY0 = SlipWall; BoundaryConditionYmax = StickWall;
  BoundaryConditionZ0 = Symmetry; BoundaryConditionZmax = Symmetry;
}}
Equations.Material[0].EquationOfState = StiffenedGas;
Equations.Material[1].EquationOfState = StiffenedGas;
BoundaryConditions.Inlet.Density = 7.0; BoundaryConditions.Inlet.VelocityX = 3.0;
BoundaryConditions.Inlet.Pressure = 11.0;
"""


def deck(d, out, bx0="Inlet", ls=""):
    return MESH.format(bx0=bx0) + ls + f'under Output {{ Prefix = "{d}/"; ' + out + " }\n"


def state(seed=0, shape=(2, 4, 8)):
    rng = np.random.default_rng(seed)
    V = np.zeros(shape + (5,))
    V[..., 0] = rng.uniform(1, 2, shape)
    V[..., 1:4] = rng.uniform(-1, 1, shape + (3,))
    V[..., 4] = rng.uniform(1, 2, shape)
    ID = (rng.random(shape) < 0.3).astype(np.int32)
    return V, ID


def read_vtr(path):
    """A VTK XML RectilinearGrid with raw appended data and UInt64 headers: {name: array}."""
    raw = open(path, "rb").read()
    head, _, blob = raw.partition(b'<AppendedData encoding="raw">\n_')
    head = head.decode()
    assert 'type="RectilinearGrid"' in head and 'header_type="UInt64"' in head
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]*)"', head).group(1).split()]
    out = {"_extent": ext, "_order": []}
    for m in re.finditer(r'<DataArray type="Float64" Name="([^"]+)" NumberOfComponents="(\d+)" '
                         r'format="appended" offset="(\d+)" />', head):
        name, nc, off = m.group(1), int(m.group(2)), int(m.group(3))
        nbytes = struct.unpack_from("<Q", blob, off)[0]
        a = np.frombuffer(blob, dtype="<f8", count=nbytes // 8, offset=off + 8)
        out[name] = a.reshape(-1, nc) if nc > 1 else a
        out["_order"].append(name)
    assert blob.endswith(b"\n  </AppendedData>\n</VTKFile>\n")
    return out


def test_collection_and_snapshots(tmp_path):
    d = str(tmp_path)
    o = M.Output(deck(d, "Solution = \"run\"; Frequency = 2; Density = On; Velocity = On; Pressure = On; "
                         "MaterialID = On; LevelSet0 = On;",
                      ls="Space.LevelSet[0].MaterialID = 1;\n"))
    pvd0 = open(os.path.join(d, "run.pvd")).read()
    assert pvd0 == ('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1"\nbyte_order="LittleEndian">\n'
                    '  <Collection>\n  </Collection>\n</VTKFile>\n')
    V, ID = state()
    phi = np.random.default_rng(1).normal(size=ID.shape)
    assert o.write(0.0, 0.0, 0, V, ID, phi, force=True) == 1
    assert o.write(0.5, 0.5, 1, V, ID, phi) == 0
    V2, ID2 = state(3)
    assert o.write(1.0, 0.5, 2, V2, ID2, phi) == 1
    assert o.write(1.0, 0.5, 2, V2, ID2, phi, force=True) == 0     # no duplicate at the same time
    o.close()
    assert open(os.path.join(d, "run.pvd")).read() == (
        '<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1"\nbyte_order="LittleEndian">\n'
        '  <Collection>\n  <DataSet timestep="0.000000e+00" file="run_0000.vtr"/>\n'
        '  <DataSet timestep="1.000000e+00" file="run_0001.vtr"/>\n  </Collection>\n</VTKFile>\n')
    s = read_vtr(os.path.join(d, "run_0001.vtr"))
    assert s["_extent"] == [0, 7, 0, 3, 0, 1]
    assert s["_order"] == ["Xcoord", "Ycoord", "Zcoord", "density", "velocity", "pressure", "materialid", "levelset0"]
    assert np.array_equal(s["Xcoord"], 0.5 + np.arange(8)) and np.array_equal(s["Zcoord"], [0.5, 1.5])
    assert np.array_equal(s["density"], V2[..., 0].ravel())
    assert np.array_equal(s["velocity"], V2[..., 1:4].reshape(-1, 3))
    assert np.array_equal(s["pressure"], V2[..., 4].ravel())
    assert np.array_equal(s["materialid"], ID2.ravel().astype(float))
    assert np.array_equal(s["levelset0"], phi.ravel())


def test_snapshots_are_forced_even_without_fields(tmp_path):
    # Main.cpp:444/632 force a snapshot at the start and the end; Output.Frequency defaults to 0
    d = str(tmp_path)
    o = M.Output(deck(d, ""))
    V, ID = state()
    assert o.write(0.0, 0.0, 0, V, ID, force=True) == 1
    assert o.write(0.1, 0.1, 1, V, ID) == 0
    s = read_vtr(os.path.join(d, "solution_0000.vtr"))
    assert s["_order"] == ["Xcoord", "Ycoord", "Zcoord"]


def test_time_interval():
    # floor(t/T) changes: written at t = 0.25, 0.5, 0.75, 1.0 (binary-exact times)
    import tempfile
    d = tempfile.mkdtemp()
    o = M.Output(deck(d, "TimeInterval = 0.25; Density = On;"))
    V, ID = state()
    written = [k for k in range(1, 9) if o.write(0.125 * k, 0.125, k, V, ID)]
    assert written == [2, 4, 6, 8]


PROBES = ('under Probes {{ Frequency = 1; Density = "rho.txt"; VelocityX = "u.txt"; VelocityY = "v.txt"; '
          'Pressure = "p.txt"; MaterialID = "id.txt"; LevelSet0 = "phi.txt"; {nodes} }}')


def probe_nodes(pts):
    return " ".join(f"Node[{q}].X = {x!r}; Node[{q}].Y = {y!r}; Node[{q}].Z = {z!r};" for q, (x, y, z) in enumerate(pts))


def probe_values(d, name):
    rows = [ln.split() for ln in open(os.path.join(d, name)) if not ln.startswith("##")]
    return np.array([[float(v) for v in r[2:]] for r in rows])


def test_probe_file_format(tmp_path):
    d = str(tmp_path)
    o = M.Output(deck(d, PROBES.format(nodes=probe_nodes([(1.5, 2.5, 0.5), (4.0, 1.0, 1.0)])),
                      ls="Space.LevelSet[0].MaterialID = 1;\n"))
    V, ID = state()
    phi = np.zeros(ID.shape)
    assert o.write(0.0, 0.0, 0, V, ID, phi, force=True) == 3     # the forced snapshot and the probes
    assert o.write(0.25, 0.25, 1, V, ID, phi) == 2
    o.close()
    lines = open(os.path.join(d, "rho.txt")).read().splitlines()
    assert lines[0] == "## Probe 0: 1.500000e+00, 2.500000e+00, 5.000000e-01"
    assert lines[1] == "## Probe 1: 4.000000e+00, 1.000000e+00, 1.000000e+00"
    assert lines[2] == "## Time step  |  Time  |  Solutions at probe nodes (0, 1, 2, etc.)"
    # ProbeOutput.cpp:453 (%10d for density) and :463 (%8d for the others)
    assert lines[3].startswith("         0      0.00000000e+00    ")
    assert open(os.path.join(d, "u.txt")).read().splitlines()[4].startswith("       1      2.50000000e-01    ")
    # a probe at a cell centre reads that cell exactly (printed with %16.8e)
    assert probe_values(d, "rho.txt")[0, 0] == float("%.8e" % V[0, 2, 1, 0])
    assert probe_values(d, "id.txt")[0, 0] == ID[0, 2, 1]


def test_probe_interpolation_is_exact_for_linear_fields(tmp_path):
    d = str(tmp_path)
    pts = [(2.3, 1.7, 0.9), (6.1, 2.2, 1.4), (4.9, 0.6, 1.1)]
    o = M.Output(deck(d, PROBES.format(nodes=probe_nodes(pts))))
    z, y, x = np.meshgrid([0.5, 1.5], 0.5 + np.arange(4), 0.5 + np.arange(8), indexing="ij")
    V = np.zeros((2, 4, 8, 5))
    V[..., 0] = 1.0 + 0.25 * x - 0.5 * y + 0.125 * z
    V[..., 4] = 2.0 - 0.1 * x + 0.3 * y * 1.0 + 0.7 * z
    o.write(0.0, 0.0, 0, V, np.zeros((2, 4, 8), np.int32), force=True)
    o.close()
    rho, p = probe_values(d, "rho.txt")[0], probe_values(d, "p.txt")[0]
    for q, (px, py, pz) in enumerate(pts):
        assert abs(rho[q] - (1.0 + 0.25 * px - 0.5 * py + 0.125 * pz)) < 1e-8   # %16.8e text
        assert abs(p[q] - (2.0 - 0.1 * px + 0.3 * py + 0.7 * pz)) < 1e-8


def test_stretched_mesh_coordinates_and_probes(tmp_path):
    # Mesh.ControlPointX / Y: the snapshot's coordinates are the generated cell centres, and the
    # probes' trilinear stencils sit on them (a linear field is reproduced at any point inside)
    d = str(tmp_path)
    mesh = MESH.format(bx0="Inlet").replace(
        "NumberOfCellsX = 8; NumberOfCellsY = 4;",
        "under ControlPointX[0] { Coordinate = 0; CellWidth = 0.5; } "
        "under ControlPointX[1] { Coordinate = 8; CellWidth = 1.5; } "
        "under ControlPointY[0] { Coordinate = 2; CellWidth = 0.25; }")
    pts = [(2.3, 1.7, 0.9), (6.1, 2.2, 1.4), (0.4, 0.6, 1.1)]
    o = M.Output(mesh + f'under Output {{ Prefix = "{d}/"; Density = On; '
                 + PROBES.format(nodes=probe_nodes(pts)) + " }\n")
    xs, _ = M.mesh_1d(0.0, 8.0, points=[(0.0, 0.5), (8.0, 1.5)])
    ys, _ = M.mesh_1d(0.0, 4.0, points=[(2.0, 0.25)])
    assert len(ys) == 16 and not np.allclose(np.diff(xs), np.diff(xs)[0])
    z, y, x = np.meshgrid([0.5, 1.5], ys, xs, indexing="ij")
    V = np.zeros(x.shape + (5,))
    V[..., 0] = 1.0 + 0.25 * x - 0.5 * y + 0.125 * z
    V[..., 4] = 2.0
    o.write(0.0, 0.0, 0, V, np.zeros(x.shape, np.int32), force=True)
    o.close()
    s = read_vtr(os.path.join(d, "solution_0000.vtr"))
    assert np.array_equal(s["Xcoord"], xs) and np.array_equal(s["Ycoord"], ys)
    assert np.array_equal(s["density"], V[..., 0].ravel())
    rho = probe_values(d, "rho.txt")[0]
    for q, (px, py, pz) in enumerate(pts):
        assert abs(rho[q] - (1.0 + 0.25 * px - 0.5 * py + 0.125 * pz)) < 1e-8


def test_probe_ghost_rules(tmp_path):
    d = str(tmp_path)
    # x0 inlet (7, 3, 0, 0, 11); y0 slip wall, ymax stick wall (velocity mirrored); the z faces symmetric
    pts = [(0.2, 2.0, 1.0),      # between the inlet ghost (x=-0.5) and the first cells (x=0.5)
           (4.0, 0.25, 1.0),     # between the slip-wall ghost (y=-0.5) and y=0.5
           (4.0, 3.75, 1.0),     # stick wall
           (0.25, 0.25, 0.25)]   # a corner: edge and corner ghosts are invalid
    o = M.Output(deck(d, PROBES.format(nodes=probe_nodes(pts)),
                      ls="under Space { under LevelSet[0] { MaterialID = 1; BoundaryConditionX0 = NonNegative;"
                         " BoundaryConditionY0 = LinearExtrapolation; } }\n"))
    V = np.zeros((2, 4, 8, 5))
    V[..., 0] = 2.0
    V[..., 1] = 0.5
    V[..., 2] = 0.25
    V[..., 4] = 5.0
    z, y, x = np.meshgrid([0.5, 1.5], 0.5 + np.arange(4), 0.5 + np.arange(8), indexing="ij")
    phi = x - 0.3 + 0.5 * y
    o.write(0.0, 0.0, 0, V, np.zeros((2, 4, 8), np.int32), phi, force=True)
    o.close()
    rho, u, v, p, ph = (probe_values(d, f)[0] for f in ("rho.txt", "u.txt", "v.txt", "p.txt", "phi.txt"))
    # inlet: (1 - xi) * ghost + xi * cell with xi = 0.7 along x
    assert abs(rho[0] - (0.3 * 7.0 + 0.7 * 2.0)) < 1e-8 and abs(u[0] - (0.3 * 3.0 + 0.7 * 0.5)) < 1e-8
    assert abs(p[0] - (0.3 * 11.0 + 0.7 * 5.0)) < 1e-8
    # slip wall at y0: v mirrored, u kept; xi = 0.75 along y
    assert abs(v[1] - (0.25 * -0.25 + 0.75 * 0.25)) < 1e-8 and abs(u[1] - 0.5) < 1e-8
    # stick wall at ymax: all velocity components mirrored; xi = 0.25 from y = 3.5 to the ghost at 4.5
    assert abs(u[2] - (0.75 * 0.5 + 0.25 * -0.5)) < 1e-8 and abs(v[2] - (0.75 * 0.25 + 0.25 * -0.25)) < 1e-8
    # corner (xi = 0.75 in every direction): of the corners (i, j, k) in {-1, 0}^3, the interior cell
    # and the three face ghosts are valid (density 2, 2 and 2, inlet 7); the four edge and corner
    # ghosts take their mean 13/4. By hand: c00 = 13/4, c01 = c10 = 37/16, c11 = 13/4,
    # c0 = 163/64, c1 = 193/64, result = 742/256.
    assert abs(rho[3] - 742.0 / 256.0) < 1e-8
    # the level set: linear extrapolation across y0 keeps the linear field; at x0 the non-negative
    # cap leaves the extrapolated ghost where it is above -h/2
    assert abs(ph[1] - (4.0 - 0.3 + 0.5 * 0.25)) < 1e-8
    g = max(-0.5, -0.5 - 0.3 + 0.5 * 2.0)   # ghost x = -0.5, y = 2.0: extrapolated phi = 0.2 stays
    assert abs(ph[0] - (0.3 * g + 0.7 * (0.5 - 0.3 + 0.5 * 2.0))) < 1e-8


def test_nonnegative_cap_binds(tmp_path):
    d = str(tmp_path)
    o = M.Output(deck(d, PROBES.format(nodes=probe_nodes([(0.0, 2.0, 1.0)])),
                      ls="under Space { under LevelSet[0] { MaterialID = 1; BoundaryConditionX0 = NonNegative; } }\n"))
    z, y, x = np.meshgrid([0.5, 1.5], 0.5 + np.arange(4), 0.5 + np.arange(8), indexing="ij")
    phi = x - 3.0                      # extrapolated ghost -3.5 is capped at -h/2 = -0.5
    o.write(0.0, 0.0, 0, *state(), phi=phi, force=True)
    o.close()
    assert abs(probe_values(d, "phi.txt")[0, 0] - (0.5 * -0.5 + 0.5 * -2.5)) < 1e-8


@pytest.mark.parametrize("extra,code", [
    (PROBES.format(nodes=probe_nodes([(9.0, 1.0, 1.0)])), M.ERR_ARG),           # outside the domain
    (PROBES.format(nodes="Node[1].X = 1; Node[1].Y = 1; Node[1].Z = 1;"), M.ERR_ARG),   # index gap
    ('LinePlot[0].FileName = "l"; LinePlot[0].NumberOfPoints = 1;', M.ERR_ARG),
])
def test_output_errors(tmp_path, extra, code):
    with pytest.raises(M.DeckError) as e:
        M.Output(deck(str(tmp_path), extra))
    assert e.value.code == code


def test_temperature_needs_a_temperature_law(tmp_path):
    text = deck(str(tmp_path), "Temperature = On;").replace(
        "Equations.Material[1].EquationOfState = StiffenedGas;", "Equations.Material[1].EquationOfState = JonesWilkinsLee;")
    with pytest.raises(M.DeckError) as e:
        M.Output(text)
    assert e.value.code == M.ERR_ARG and "temperature law" in str(e.value)


def test_outputs_not_written_are_listed(tmp_path):
    o = M.Output(deck(str(tmp_path), "DeltaTemperature = On; MeshInformation = \"mesh.txt\"; Density = On;"))
    assert o.not_written == ["not written: Output.DeltaTemperature", "not written: Output.MeshInformation"]
    assert not o.needs_device
    assert M.Output(deck(str(tmp_path), "InternalEnergyPerUnitMass = On;")).needs_device
