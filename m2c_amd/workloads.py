"""Synthetic workloads of the benchmark configurations (initial conditions + solver settings).

Each builder returns (settings, V, ID): `settings` is a dict of solver options understood by
m2c_amd.make_config, V the primitive state (nz, ny, nx, 5) and ID the material IDs (nz, ny, nx).

  sod_1d            400x1x1 two-state Riemann problem (gamma=1.4, van Albada, HLLC, RK2, CFL 0.5)
  periodic_modes    N^3 periodic gas: rho=p=1 plus 16 seeded Fourier modes of amplitude 1e-2 on
                    density and velocity and a Gaussian pressure pulse (amplitude 0.5, sigma 8
                    cells) at the centre; gamma=1.4, van Albada, HLLC, RK3, CFL 0.4
  shock_bubble      two-material shock-bubble: stiffened-gas water (gamma=6.12, p_inf=3.43e8 Pa,
                    rho=1000, p=1e5) around a level-set sphere of air (gamma=1.4, rho=1.2, p=1e5)
                    of radius 24 cells centred at x=160 cells on a 384x192x192 grid; a planar shock
                    in the water at x=64 cells with post-shock p=1e9 Pa; exact two-material Riemann
                    solves at interface faces, level-set reinitialisation every 5 steps, HLLC+MUSCL
                    (van Albada), RK2, CFL 0.5. `scale` shrinks every length for small test runs.
  shock_bubble_stretched
                    shock_bubble on a mesh stretched by control points (Mesh.ControlPointX/Y/Z):
                    cells of width 0.75 h across the bubble, growing linearly to 1.5 h at the box's
                    faces (h = the uniform run's width; box, materials, bubble and shock as
                    shock_bubble)
  periodic_viscous  periodic_modes with constant viscosity (mu 0.01, bulk 0.002) and heat diffusion
                    (kappa 0.02, T = p / rho)
  sphere_obstacle   a supersonic gas stream (Mach 2 along x, gamma 1.4) past a sphere bounded by an
                    embedded wall: cells inside the sphere are inactive and every mesh face between an
                    inside and an outside cell is cut, its wall records holding the sphere's outward
                    normal and the wall velocity at the crossing of the cell-centre edge
  batched_riemann   independent single-material Riemann-type instances: edge lengths drawn
                    uniformly from lo..hi per axis, left/right states log-uniform (rho in
                    [0.1, 10], p in [0.1, 100], u = 0) split by a random plane (normal uniform on
                    the sphere, through a uniform point of the box); gamma = 1.4, HLLC + van
                    Albada, RK2, CFL 0.5, slip walls
The periodic boundary and the Fourier-mode initial state have no counterpart in the reference's
input format (IoData.h:266-267 has no periodic code); they are the documented extension.
"""
import numpy as np

from . import (BC_INLET, BC_OUTLET, BC_PERIODIC, BC_SLIPWALL, BC_SYMMETRY, EOS_SG, FLUX_HLLC, LIM_VANALBADA,
               TS_RK2, TS_RK3, VISC_CONSTANT, material, mesh_1d)

FOURIER_SEED = 2024


def sod_1d(n=400):
    settings = dict(nx=n, ny=1, nz=1, box=(0.0, 1.0, 0.0, 1.0 / n, 0.0, 1.0 / n),
                    bc=[BC_INLET, BC_OUTLET, BC_SYMMETRY, BC_SYMMETRY, BC_SYMMETRY, BC_SYMMETRY],
                    mats=[material(EOS_SG, [1.4, 0.0])], inlet=(1.0, 0.0, 0.0, 0.0, 1.0),
                    outlet=(0.125, 0.0, 0.0, 0.0, 0.1), flux=FLUX_HLLC, limiter=LIM_VANALBADA, ts=TS_RK2,
                    cfl=0.5)
    x = (np.arange(n) + 0.5) / n
    V = np.zeros((1, 1, n, 5))
    V[..., 0] = np.where(x < 0.5, 1.0, 0.125)
    V[..., 4] = np.where(x < 0.5, 1.0, 0.1)
    return settings, V, np.zeros((1, 1, n), dtype=np.int32)


def fourier_modes(seed=FOURIER_SEED, nmodes=16, kmax=4):
    """Integer wave vectors in [-kmax, kmax]^3 minus the origin and uniform phases per field."""
    rng = np.random.default_rng(seed)
    ks = []
    while len(ks) < nmodes:
        k = rng.integers(-kmax, kmax + 1, size=3)
        if np.any(k != 0):
            ks.append(k)
    phases = rng.uniform(0.0, 2.0 * np.pi, size=(nmodes, 4))
    return np.array(ks), phases


def periodic_modes(n=256, ny=None, nz=None, amp=1e-2, pulse=0.5, sigma=8.0, seed=FOURIER_SEED):
    nx = n
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    settings = dict(nx=nx, ny=ny, nz=nz, box=(0.0, float(nx), 0.0, float(ny), 0.0, float(nz)),
                    bc=[BC_PERIODIC] * 6, mats=[material(EOS_SG, [1.4, 0.0])], flux=FLUX_HLLC,
                    limiter=LIM_VANALBADA, ts=TS_RK3, cfl=0.4)
    ks, ph = fourier_modes(seed)
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) / nz, (np.arange(ny) + 0.5) / ny, (np.arange(nx) + 0.5) / nx,
                          indexing="ij")
    V = np.zeros((nz, ny, nx, 5))
    V[..., 0] = 1.0
    V[..., 4] = 1.0
    for m in range(len(ks)):
        arg = 2.0 * np.pi * (ks[m, 0] * x + ks[m, 1] * y + ks[m, 2] * z)
        for f, comp in enumerate((0, 1, 2, 3)):
            V[..., comp] += amp * np.sin(arg + ph[m, f])
    cz, cy, cx = nz / 2.0, ny / 2.0, nx / 2.0
    r2 = ((x * nx - cx) ** 2 + (y * ny - cy) ** 2 + (z * nz - cz) ** 2)
    V[..., 4] += pulse * np.exp(-r2 / (2.0 * sigma * sigma))
    return settings, V, np.zeros((nz, ny, nx), dtype=np.int32)


def periodic_viscous(n=256, mu=0.01, bulk=0.002, kappa=0.02, **kw):
    """periodic_modes with a viscous, heat-conducting gas: constant viscosity mu and bulk viscosity,
    constant diffusivity kappa, temperature law T = e / cv with cv = 1 / (gamma - 1) (so T = p / rho).
    The diffusive step limit h^2 rho / (6 mu) stays far above the CFL 0.4 step (h = 1)."""
    settings, V, ID = periodic_modes(n=n, **kw)
    gas = material(EOS_SG, [1.4, 0.0], tlaw=(2.5, 0.0, 0.0, 0.0, 0.0, 0.0), viscosity=VISC_CONSTANT, mu=mu, bulk=bulk,
                   kappa=kappa)
    return dict(settings, mats=[gas]), V, ID


WATER = (6.12, 3.43e8)
AIR = (1.4, 0.0)


def sg_shock_state(gamma, pinf, rho1, p1, p2):
    """Post-shock (rho2, u2) behind a right-running shock into stiffened gas at rest (rho1, p1).

    Rankine-Hugoniot for p + p_inf, which obeys the perfect-gas relations."""
    P1, P2 = p1 + pinf, p2 + pinf
    g = gamma
    rho2 = rho1 * ((g + 1.0) * P2 + (g - 1.0) * P1) / ((g - 1.0) * P2 + (g + 1.0) * P1)
    u2 = np.sqrt((p2 - p1) * (1.0 / rho1 - 1.0 / rho2))
    return rho2, u2


def shock_bubble(nx=384, ny=192, nz=192, radius=24.0, cx=160.0, shock_x=64.0, p_shock=1e9, h=1.0,
                 reinit_freq=5, scale=1.0):
    if scale != 1.0:
        nx, ny, nz = max(int(round(nx * scale)), 4), max(int(round(ny * scale)), 4), max(int(round(nz * scale)), 4)
        radius, cx, shock_x = radius * scale, cx * scale, shock_x * scale
    rho_w, p0 = 1000.0, 1e5
    rho2, u2 = sg_shock_state(WATER[0], WATER[1], rho_w, p0, p_shock)
    post = (rho2, u2, 0.0, 0.0, p_shock)
    settings = dict(nx=nx, ny=ny, nz=nz, box=(0.0, nx * h, 0.0, ny * h, 0.0, nz * h),
                    bc=[BC_INLET, BC_OUTLET, BC_SYMMETRY, BC_SYMMETRY, BC_SYMMETRY, BC_SYMMETRY],
                    mats=[material(EOS_SG, list(WATER)), material(EOS_SG, list(AIR))],
                    inlet=post, outlet=(rho_w, 0.0, 0.0, 0.0, p0), flux=FLUX_HLLC, limiter=LIM_VANALBADA,
                    ts=TS_RK2, cfl=0.5, ls_matid=1, reinit_freq=reinit_freq)
    zc, yc, xc = np.meshgrid((np.arange(nz) + 0.5) * h, (np.arange(ny) + 0.5) * h, (np.arange(nx) + 0.5) * h,
                             indexing="ij")
    phi = np.sqrt((xc - cx * h) ** 2 + (yc - 0.5 * ny * h) ** 2 + (zc - 0.5 * nz * h) ** 2) - radius * h
    ID = np.where(phi < 0.0, 1, 0).astype(np.int32)
    V = np.zeros((nz, ny, nx, 5))
    V[..., 0] = np.where(ID == 1, 1.2, rho_w)
    V[..., 4] = p0
    shocked = xc < shock_x * h
    for c in range(5):
        V[..., c] = np.where(shocked, post[c], V[..., c])
    return settings, V, ID, phi


def shock_bubble_stretched(nx=384, ny=192, nz=192, radius=24.0, cx=160.0, shock_x=64.0, p_shock=1e9, h=1.0,
                           reinit_freq=5, scale=1.0, fine=0.75, coarse=1.5):
    """shock_bubble on a stretched mesh: width fine*h over the bubble's extent along each axis,
    linear to coarse*h at the box's faces; the cell counts are the mesh generator's."""
    settings, _, _, _ = shock_bubble(nx, ny, nz, radius, cx, shock_x, p_shock, h, reinit_freq, scale)
    if scale != 1.0:
        radius, cx, shock_x = radius * scale, cx * scale, shock_x * scale
    box = settings["box"]
    ctr = (cx * h, 0.5 * (box[2] + box[3]), 0.5 * (box[4] + box[5]))
    pts, axes = [], []
    for d in range(3):
        lo, hi = box[2 * d], box[2 * d + 1]
        p = [(lo, coarse * h), (ctr[d] - radius * h, fine * h), (ctr[d] + radius * h, fine * h), (hi, coarse * h)]
        pts.append(p)
        axes.append(mesh_1d(lo, hi, points=p)[0])
    settings.update(nx=len(axes[0]), ny=len(axes[1]), nz=len(axes[2]), control_points=pts)
    zc, yc, xc = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    phi = np.sqrt((xc - ctr[0]) ** 2 + (yc - ctr[1]) ** 2 + (zc - ctr[2]) ** 2) - radius * h
    ID = np.where(phi < 0.0, 1, 0).astype(np.int32)
    rho_w, p0 = 1000.0, 1e5
    post = settings["inlet"]
    V = np.zeros(ID.shape + (5,))
    V[..., 0] = np.where(ID == 1, 1.2, rho_w)
    V[..., 4] = p0
    shocked = xc < shock_x * h
    for c in range(5):
        V[..., c] = np.where(shocked, post[c], V[..., c])
    return settings, V, ID, phi


def _num(x):
    """A number as deck text that reads back to the same double (17 significant digits)."""
    return repr(float(x))


def _state(name, V, matid=None):
    s = f"  under {name} {{\n"
    if matid is not None:
        s += f"    MaterialID = {matid};\n"
    for key, v in zip(("Density", "VelocityX", "VelocityY", "VelocityZ", "Pressure"), V):
        s += f"    {key} = {_num(v)};\n"
    return s + "  }\n"


def sod_deck(n=400):
    """sod_1d as an M2C input deck: the same settings and (through a plane) the same initial state."""
    settings, _, _ = sod_1d(n)
    return f"""// Sod shock tube, {n} x 1 x 1 cells
under Mesh {{
  X0 = 0.0; Xmax = 1.0; Y0 = 0.0; Ymax = {_num(1.0 / n)}; Z0 = 0.0; Zmax = {_num(1.0 / n)};
  NumberOfCellsX = {n}; NumberOfCellsY = 1; NumberOfCellsZ = 1;
  BoundaryConditionX0 = Farfield0; BoundaryConditionXmax = Farfield1;
  BoundaryConditionY0 = Symmetry; BoundaryConditionYmax = Symmetry;
  BoundaryConditionZ0 = Symmetry; BoundaryConditionZmax = Symmetry;
}}
under Equations {{
  under Material[0] {{
    EquationOfState = StiffenedGas;
    under StiffenedGasModel {{ SpecificHeatRatio = 1.4; PressureConstant = 0.0; }}
  }}
}}
under BoundaryConditions {{
{_state("Inlet", settings["inlet"])}{_state("Outlet", settings["outlet"])}}}
under DefaultInitialState {{ Density = 0.125; Pressure = 0.1; }}
under InitialCondition {{
  under GeometricEntities {{
    under Plane[0] {{
      Point_x = 0.5; Normal_x = -1.0;
{_state("InitialState", (1.0, 0.0, 0.0, 0.0, 1.0))}    }}
  }}
}}
under Space {{
  under NavierStokes {{
    Flux = HLLC;
    under Reconstruction {{ Type = Linear; Limiter = VanAlbada; }}
  }}
}}
under Time {{
  Type = Explicit; MaxTime = 0.1; CFL = 0.5;
  under Explicit {{ Type = RungeKutta2; }}
}}
under Output {{ Prefix = ""; Solution = "sod"; Frequency = 100; Density = On; Velocity = On; Pressure = On; }}
"""


def shock_bubble_deck(scale=1.0, reinit_freq=5, stretched=False):
    """shock_bubble as an M2C input deck (h = 1): water, a level-set sphere of air, a planar shock;
    stretched=True: shock_bubble_stretched's mesh, given by Mesh.ControlPointX/Y/Z."""
    settings, _, _, _ = shock_bubble(scale=scale, reinit_freq=reinit_freq)
    nx, ny, nz = settings["nx"], settings["ny"], settings["nz"]
    radius, cx, shock_x = 24.0, 160.0, 64.0
    if scale != 1.0:
        radius, cx, shock_x = radius * scale, cx * scale, shock_x * scale
    cells = f"NumberOfCellsX = {nx}; NumberOfCellsY = {ny}; NumberOfCellsZ = {nz};"
    if stretched:
        pts = shock_bubble_stretched(scale=scale, reinit_freq=reinit_freq)[0]["control_points"]
        cells = "\n  ".join(f"under ControlPoint{a}[{q}] {{ Coordinate = {_num(x)}; CellWidth = {_num(w)}; }}"
                             for a, p in zip("XYZ", pts) for q, (x, w) in enumerate(p))
    return f"""// shock-bubble: water (stiffened gas) around an air bubble, planar shock at x = {_num(shock_x)}
under Mesh {{
  X0 = 0; Xmax = {nx}; Y0 = 0; Ymax = {ny}; Z0 = 0; Zmax = {nz};
  {cells}
  BoundaryConditionX0 = Inlet; BoundaryConditionXmax = Outlet;
  BoundaryConditionY0 = Symmetry; BoundaryConditionYmax = Symmetry;
  BoundaryConditionZ0 = Symmetry; BoundaryConditionZmax = Symmetry;
}}
under Equations {{
  Material[0].EquationOfState = StiffenedGas;
  Material[0].StiffenedGasModel.SpecificHeatRatio = {_num(WATER[0])};
  Material[0].StiffenedGasModel.PressureConstant = {_num(WATER[1])};
  under Material[1] {{
    EquationOfState = StiffenedGas;
    under StiffenedGasModel {{ SpecificHeatRatio = {_num(AIR[0])}; PressureConstant = {_num(AIR[1])}; }}
  }}
}}
under BoundaryConditions {{
{_state("Inlet", settings["inlet"])}{_state("Outlet", settings["outlet"])}}}
under DefaultInitialState {{ Density = 1000.0; Pressure = 1.0e5; }}
under GeometricEntities {{
  under Plane[0] {{
    Point_x = {_num(shock_x)}; Normal_x = -1;
{_state("InitialState", settings["inlet"], 0)}  }}
  under Sphere[0] {{
    Center_x = {_num(cx)}; Center_y = {_num(0.5 * ny)}; Center_z = {_num(0.5 * nz)}; Radius = {_num(radius)};
{_state("InitialState", (1.2, 0.0, 0.0, 0.0, 1.0e5), 1)}  }}
}}
under Space {{
  under NavierStokes {{
    Flux = HLLC;
    under Reconstruction {{ Type = Linear; Limiter = VanAlbada; }}
  }}
  under LevelSet[0] {{
    MaterialID = 1;
    under Reinitialization {{ Frequency = {reinit_freq}; }}
  }}
}}
under MultiPhase {{ Flux = Numerical; ReconstructionAtInterface = Constant; }}
under Time {{ Type = Explicit; CFL = 0.5; under Explicit {{ Type = RungeKutta2; }} }}
"""


BATCH_SEED = 4242


def sphere_obstacle(nx=48, ny=32, nz=32, radius=6.0, centre=None, vwall=(0.0, 0.0, 0.0), seed=5, amp=0.05):
    """A Mach-2 gas stream past a sphere of `radius` cells: (settings, V, ID, faces, walls). Material 1
    is inactive (the inside of the sphere); `faces` (n, 4) lists every axis, i, j, k (plus cell) of a
    face between an inside and an outside cell, `walls` (n, 12) its forward and backward wall records
    (unit outward normal at the crossing of the segment between the two cell centres, wall velocity).
    The free stream carries a seeded perturbation of relative amplitude `amp`."""
    cx, cy, cz = centre if centre is not None else (nx / 3.0, ny / 2.0, nz / 2.0)
    gas = material(EOS_SG, [1.4, 0.0])
    rho, p = 1.0, 1.0
    u = 2.0 * np.sqrt(1.4 * p / rho)
    settings = dict(nx=nx, ny=ny, nz=nz, box=(0.0, float(nx), 0.0, float(ny), 0.0, float(nz)),
                    bc=[BC_INLET, BC_OUTLET, BC_SLIPWALL, BC_SLIPWALL, BC_SLIPWALL, BC_SLIPWALL],
                    mats=[gas, gas], inactive_id=1, inlet=(rho, u, 0.0, 0.0, p), outlet=(rho, u, 0.0, 0.0, p),
                    flux=FLUX_HLLC, limiter=LIM_VANALBADA, ts=TS_RK2, cfl=0.5)
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    inside = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 < radius * radius
    ID = inside.astype(np.int32)
    rng = np.random.default_rng(seed)
    V = np.zeros((nz, ny, nx, 5))
    V[..., 0] = rho * (1.0 + amp * rng.uniform(-1.0, 1.0, (nz, ny, nx)))
    V[..., 1] = u * (1.0 + amp * rng.uniform(-1.0, 1.0, (nz, ny, nx)))
    V[..., 2] = amp * u * rng.uniform(-1.0, 1.0, (nz, ny, nx))
    V[..., 4] = p * (1.0 + amp * rng.uniform(-1.0, 1.0, (nz, ny, nx)))
    V[inside, 1:4] = 0.0
    faces, walls = [], []
    C = np.array([cx, cy, cz])
    for d in range(3):
        sl_m = [slice(None)] * 3
        sl_p = [slice(None)] * 3
        sl_m[2 - d] = slice(0, -1)
        sl_p[2 - d] = slice(1, None)
        cut = inside[tuple(sl_m)] != inside[tuple(sl_p)]
        for kk, jj, ii in zip(*np.nonzero(cut)):
            q = [ii, jj, kk]
            q[d] += 1                                     # the plus cell
            a = np.array([q[0] + 0.5, q[1] + 0.5, q[2] + 0.5])
            a[d] -= 1.0                                    # minus cell centre
            e = np.zeros(3)
            e[d] = 1.0
            # |a + t e - C|^2 = r^2 on t in (0, 1)
            w = a - C
            bq, cq = w[d], float(w @ w) - radius * radius
            disc = np.sqrt(max(bq * bq - cq, 0.0))
            t = -bq + disc if cq < 0.0 else -bq - disc
            pt = a + t * e
            nrm = (pt - C) / np.sqrt(float((pt - C) @ (pt - C)))
            faces.append([d, q[0], q[1], q[2]])
            walls.append(np.concatenate([nrm, vwall, nrm, vwall]))
    return settings, V, ID, np.array(faces, dtype=np.int32).reshape(-1, 4), np.array(walls).reshape(-1, 12)


# second gas of the two-material batched instances (stiffened gas, gamma 1.67, no stiffening)
GAS2 = (1.67, 0.0)


def batched_riemann(n=64, lo=48, hi=160, seed=BATCH_SEED, rho_range=(0.1, 10.0), p_range=(0.1, 100.0), select=None,
                    two_every=0, reinit_freq=5):
    """Returns a list of (settings, V, ID) or, for two-material instances, (settings, V, ID, phi),
    one per instance (with `select`, only the instances of those indices, drawn from the same random
    stream as the whole set).

    two_every = k > 0: every k-th instance (indices k-1, 2k-1, ...) is two-material: the gas of the
    left side of its split plane is a second gas (GAS2, material 1, tracked by a level set, the
    signed distance to the plane, reinitialised every `reinit_freq` steps), with exact Riemann
    solves at the interface faces; the others keep one gas (material 0, gamma 1.4).

    CFL is 0.4: the time step follows SpaceOperator.cpp:2287, dt = CFL * min over the three
    directions of dx/lambda, so an oblique wave sees the sum of the three directional Courant
    numbers; at CFL 0.5 MUSCL-RK2 turns unstable within 20 steps on oblique split planes (the
    oracle and the GPU agree on the blow-up)."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        nx, ny, nz = (int(v) for v in rng.integers(lo, hi + 1, size=3))
        normal = rng.normal(size=3)
        normal /= np.linalg.norm(normal)
        point = rng.uniform(0.0, 1.0, size=3) * np.array([nx, ny, nz])
        rho = np.exp(rng.uniform(np.log(rho_range[0]), np.log(rho_range[1]), size=2))
        p = np.exp(rng.uniform(np.log(p_range[0]), np.log(p_range[1]), size=2))
        if select is not None and q not in select:
            continue
        two = two_every > 0 and (q + 1) % two_every == 0
        settings = dict(nx=nx, ny=ny, nz=nz, box=(0.0, float(nx), 0.0, float(ny), 0.0, float(nz)),
                        bc=[BC_SYMMETRY] * 6, mats=[material(EOS_SG, [1.4, 0.0])], flux=FLUX_HLLC,
                        limiter=LIM_VANALBADA, ts=TS_RK2, cfl=0.4)
        z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
        dist = (x - point[0]) * normal[0] + (y - point[1]) * normal[1] + (z - point[2]) * normal[2]
        left = dist < 0.0
        V = np.zeros((nz, ny, nx, 5))
        V[..., 0] = np.where(left, rho[0], rho[1])
        V[..., 4] = np.where(left, p[0], p[1])
        if two:
            settings.update(mats=[material(EOS_SG, [1.4, 0.0]), material(EOS_SG, list(GAS2))], ls_matid=1,
                            reinit_freq=reinit_freq)
            out.append((settings, V, np.where(left, 1, 0).astype(np.int32), dist.copy()))
        else:
            out.append((settings, V, np.zeros((nz, ny, nx), dtype=np.int32)))
    return out


def config_from(settings, **over):
    from . import make_config
    s = dict(settings)
    s.update(over)
    return make_config(s.pop("nx"), s.pop("ny"), s.pop("nz"), s.pop("box"), s.pop("bc"), s.pop("mats"), **s)


def interface_problems(n, seed=8):
    """Pointwise water/air interface problems of the shock-bubble's kind along x (water = material 0,
    air = 1 of shock_bubble's settings; air on the left for half of them): (Vm, idm, Vp, idp)."""
    rng = np.random.default_rng(seed)
    Vm = np.zeros((n, 5))
    Vp = np.zeros((n, 5))
    Vm[:, 0] = rng.uniform(1000.0, 1250.0, n)
    Vm[:, 4] = 10.0 ** rng.uniform(5.0, 9.0, n)
    Vm[:, 1] = rng.uniform(-100.0, 450.0, n)
    Vp[:, 0] = 1.2 * 10.0 ** rng.uniform(0.0, 2.5, n)
    Vp[:, 4] = 10.0 ** rng.uniform(5.0, 9.0, n)
    Vp[:, 1] = rng.uniform(-100.0, 100.0, n)
    Vm[:, 2:4] = rng.uniform(-10.0, 10.0, (n, 2))
    Vp[:, 2:4] = rng.uniform(-10.0, 10.0, (n, 2))
    idm = np.zeros(n, np.int32)
    idp = np.ones(n, np.int32)
    swap = rng.random(n) < 0.5
    Vm[swap], Vp[swap] = Vp[swap].copy(), Vm[swap].copy()
    idm[swap], idp[swap] = 1, 0
    return Vm, idm, Vp, idp
