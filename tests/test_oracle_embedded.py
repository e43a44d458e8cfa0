"""Pins of the oracle's mesh faces cut by embedded walls (oracle/fv.c: face_flux, recon_cell),
restating SpaceOperator::ComputeAdvectionFluxes with EmbeddedBoundaryDataSets
(SpaceOperator.cpp:2438-2484, 2569-2611, 2696-2737) and Reconstructor's slope zeroing next to
intersected edges (Reconstructor.cpp:656-664):
* wiring, by hand on a constant-reconstruction grid: the residual of the two cells beside one cut
  face is assembled here from the pointwise numerical flux and the pointwise wall-face flux (both
  pinned on their own: test_oracle_fv.py, test_gpu_onesided.py), in the reference's summation
  order; the forward record serves the minus cell, the backward record the plus cell, an inactive
  side takes no flux;
* ReconstructionAtInterface LINEAR: on a state linear in x with binary-fraction steps every
  reconstructed face state is the exact midpoint, so the wall side's state is known exactly;
* mirror symmetry: a wall plane across the whole domain with a mirror-image state on either side
  gives mirror-image residuals (a flux taken as F(V*, V) on the wrong side breaks it);
* decoupling: with slopes zeroed next to cut faces, nothing on one side of a closed wall plane
  changes the residual on the other side, bit for bit; without the zeroing it does.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

GAS = po.sg(1.4, 0.0)


def make(nx=8, ny=3, nz=2, rec_linear=False, slope_zero=False, inactive_id=-1, mats=None):
    mats = mats or [GAS]
    cfg = po.make_config(nx, ny, nz, (0, nx, 0, ny, 0, nz), [po.BC_SLIPWALL] * 6, mats, rec_linear=rec_linear,
                         slope_zero=slope_zero, inactive_id=inactive_id)
    return po.Oracle(cfg)


def field(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    V = np.zeros((nz, ny, nx, 5))
    V[..., 0] = 1.0 + 0.3 * rng.random((nz, ny, nx))
    V[..., 1:4] = 0.4 * (rng.random((nz, ny, nx, 3)) - 0.5)
    V[..., 4] = 1.0 + 0.5 * rng.random((nz, ny, nx))
    return V


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.sqrt((v * v).sum())


def residual_by_hand(o, V, ID, i, j, k, override, inactive=-1):
    """R of interior cell (i,j,k) on a constant-reconstruction grid: face fluxes F(V_minus, V_plus),
    summed west, south, back (subtracted), east, north, front (added), times 1/(-vol); `override`
    maps a face name to the flux to use instead."""
    me = ID[k + 1, j + 1, i + 1]

    def F(d, a, b):
        ia, ib = a, b
        ida, idb = ID[ia[2] + 1, ia[1] + 1, ia[0] + 1], ID[ib[2] + 1, ib[1] + 1, ib[0] + 1]
        if ida == inactive or idb == inactive:
            return np.zeros(5)
        return o.flux(d, V[ia[2] + 1, ia[1] + 1, ia[0] + 1], V[ib[2] + 1, ib[1] + 1, ib[0] + 1], me)

    faces = {
        "w": lambda: F(0, (i - 1, j, k), (i, j, k)), "e": lambda: F(0, (i, j, k), (i + 1, j, k)),
        "s": lambda: F(1, (i, j - 1, k), (i, j, k)), "n": lambda: F(1, (i, j, k), (i, j + 1, k)),
        "b": lambda: F(2, (i, j, k - 1), (i, j, k)), "t": lambda: F(2, (i, j, k), (i, j, k + 1)),
    }
    Fl = {f: override[f] if f in override else faces[f]() for f in faces}
    ax, ay, az = 1.0, 1.0, 1.0
    inv = 1.0 / (-1.0)
    out = np.zeros(5)
    for p in range(5):
        s = 0.0
        s -= Fl["w"][p] * ax
        s -= Fl["s"][p] * ay
        s -= Fl["b"][p] * az
        s += Fl["e"][p] * ax
        s += Fl["n"][p] * ay
        s += Fl["t"][p] * az
        out[p] = s * inv
    return out


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_cut_face_wiring_by_hand(mode):
    nx, ny, nz = 6, 4, 3
    o = make(nx, ny, nz)
    V0 = field(nx, ny, nz, 3)
    V, ID = o.ghosted(V0)
    # one cut x face between cells (2,1,1) and (3,1,1); forward and backward records differ
    nf, vf = unit([0.9, 0.3, -0.2]), np.array([0.05, -0.02, 0.01])
    nb, vb = unit([-0.8, 0.1, 0.4]), np.array([-0.03, 0.04, 0.0])
    assert o.set_embedded_faces([[0, 3, 1, 1]], [np.concatenate([nf, vf, nb, vb])], normal_mode=mode) == 0
    st, R = o.residual(V, ID)
    assert st == 0
    s1, lf1, _ = o.wall_face_flux(0, 1, mode, nf, vf, V[2, 2, 3])     # minus cell (2,1,1), forward record
    s2, lf2, _ = o.wall_face_flux(0, -1, mode, nb, vb, V[2, 2, 4])    # plus cell (3,1,1), backward record
    assert s1 == 0 and s2 == 0
    assert np.array_equal(R[2, 2, 3], residual_by_hand(o, V, ID, 2, 1, 1, {"e": lf1}))
    assert np.array_equal(R[2, 2, 4], residual_by_hand(o, V, ID, 3, 1, 1, {"w": lf2}))
    # cells away from the cut face are untouched by it
    o2 = make(nx, ny, nz)
    _, R2 = o2.residual(V, ID)
    diff = np.any(R != R2, axis=-1)
    assert set(zip(*np.nonzero(diff))) == {(2, 2, 3), (2, 2, 4)}


def test_linear_reconstruction_at_cut_face():
    # a state linear in x with binary-fraction steps: van Albada's slope is the difference itself,
    # so every reconstructed x-face state is the exact midpoint; y and z slopes are zero
    nx, ny, nz = 8, 2, 2
    o = make(nx, ny, nz, rec_linear=True)
    V0 = np.zeros((nz, ny, nx, 5))
    x = np.arange(nx, dtype=np.float64)
    V0[..., 0] = 1.0 + 0.125 * x
    V0[..., 1] = -0.25 + 0.0625 * x
    V0[..., 4] = 2.0 + 0.25 * x
    V, ID = o.ghosted(V0)
    nf, vf = unit([0.9, 0.3, -0.2]), np.array([0.05, -0.02, 0.01])
    nb, vb = unit([-0.8, 0.1, 0.4]), np.array([-0.03, 0.04, 0.0])
    assert o.set_embedded_faces([[0, 4, 0, 0]], [np.concatenate([nf, vf, nb, vb])], normal_mode=0, recon=1) == 0
    st, R = o.residual(V, ID)
    assert st == 0

    def mid(a):   # the x-face state at a + 1/2
        return 0.5 * (V0[0, 0, a] + V0[0, 0, a + 1])

    for a in range(2, 5):
        assert np.array_equal(mid(a), V0[0, 0, a] + 0.5 * (V0[0, 0, a + 1] - V0[0, 0, a]))
    _, lf1, _ = o.wall_face_flux(0, 1, 0, nf, vf, mid(3))
    _, lf2, _ = o.wall_face_flux(0, -1, 0, nb, vb, mid(3))
    west = o.flux(0, mid(2), mid(2))
    east = o.flux(0, mid(4), mid(4))
    assert np.array_equal(R[1, 1, 4], residual_by_hand(o, V, ID, 3, 0, 0, {"w": west, "e": lf1}))
    assert np.array_equal(R[1, 1, 5], residual_by_hand(o, V, ID, 4, 0, 0, {"w": lf2, "e": east}))
    # CONSTANT takes the cell values instead
    o.set_embedded_faces([[0, 4, 0, 0]], [np.concatenate([nf, vf, nb, vb])], normal_mode=0, recon=0)
    _, Rc = o.residual(V, ID)
    _, lf1c, _ = o.wall_face_flux(0, 1, 0, nf, vf, V0[0, 0, 3])
    assert np.array_equal(Rc[1, 1, 4], residual_by_hand(o, V, ID, 3, 0, 0, {"w": west, "e": lf1c}))
    assert not np.array_equal(lf1c, lf1)


def test_inactive_side_takes_no_flux():
    nx, ny, nz = 6, 3, 2
    o = make(nx, ny, nz, inactive_id=1, mats=[GAS, po.sg(1.4, 0.0)])
    V0 = field(nx, ny, nz, 5)
    ID0 = np.zeros((nz, ny, nx), np.int32)
    ID0[..., 3:] = 1                       # cells i >= 3 are inside the solid
    V, ID = o.ghosted(V0, ID0)
    faces = [[0, 3, j, k] for k in range(nz) for j in range(ny)]
    nw = unit([1.0, 0.2, 0.0])
    walls = [np.concatenate([nw, [0.0, 0.0, 0.0], -nw, [0.0, 0.0, 0.0]])] * len(faces)
    assert o.set_embedded_faces(faces, walls, normal_mode=0) == 0
    st, R = o.residual(V, ID)
    assert st == 0
    assert np.all(R[1:-1, 1:-1, 4:-1] == 0.0)     # inactive cells: every flux zero
    for k in range(nz):
        for j in range(ny):
            _, lf1, _ = o.wall_face_flux(0, 1, 0, nw, np.zeros(3), V[k + 1, j + 1, 3])
            exp = residual_by_hand(o, V, ID, 2, j, k, {"e": lf1}, inactive=1)
            assert np.array_equal(R[k + 1, j + 1, 3], exp)


def mirror_x(A):
    """Mirror image in x of an interior (nz, ny, nx, 5) array: order reversed, x velocity negated."""
    B = A[:, :, ::-1].copy()
    B[..., 1] = -B[..., 1]
    return B


@pytest.mark.parametrize("slope_zero", [False, True])
def test_wall_plane_mirror_symmetry(slope_zero):
    nx, ny, nz = 8, 3, 2
    o = make(nx, ny, nz, rec_linear=True, slope_zero=slope_zero)
    half = field(nx // 2, ny, nz, 11)
    V0 = np.concatenate([half, mirror_x(half)], axis=2)
    V, ID = o.ghosted(V0)
    faces = [[0, nx // 2, j, k] for k in range(nz) for j in range(ny)]
    walls = [np.concatenate([[1.0, 0, 0], [0.0, 0, 0], [-1.0, 0, 0], [0.0, 0, 0]])] * len(faces)
    assert o.set_embedded_faces(faces, walls, normal_mode=1, recon=1) == 0
    st, R = o.residual(V, ID)
    assert st == 0
    Ri = R[1:-1, 1:-1, 1:-1]
    Rm = mirror_x(Ri)
    scale = np.abs(Ri).max()
    assert np.abs(Ri - Rm).max() <= 1e-13 * scale
    # the walls change the residual (the wall state differs from the neighbour's)
    o2 = make(nx, ny, nz, rec_linear=True, slope_zero=slope_zero)
    _, R2 = o2.residual(V, ID)
    assert np.abs(R2 - R).max() > 1e-6 * scale


@pytest.mark.parametrize("recon", [0, 1])
def test_closed_wall_plane_decouples_with_zero_slopes(recon):
    nx, ny, nz = 8, 3, 2
    V0 = field(nx, ny, nz, 21)
    V1 = V0.copy()
    V1[:, :, nx // 2:] = field(nx - nx // 2, ny, nz, 22)     # a different state right of the wall
    faces = [[0, nx // 2, j, k] for k in range(nz) for j in range(ny)]
    nw = unit([1.0, 0.3, -0.1])
    walls = [np.concatenate([nw, [0.02, 0, 0], -nw, [0.02, 0, 0]])] * len(faces)

    def left(slope_zero, Vi):
        o = make(nx, ny, nz, rec_linear=True, slope_zero=slope_zero)
        V, ID = o.ghosted(Vi)
        assert o.set_embedded_faces(faces, walls, normal_mode=2, recon=recon) == 0
        st, R = o.residual(V, ID)
        assert st == 0
        return R[1:-1, 1:-1, 1:nx // 2 + 1]

    assert np.array_equal(left(True, V0), left(True, V1))
    assert not np.array_equal(left(False, V0), left(False, V1))


def test_setter_rejects_boundary_and_repeated_faces():
    o = make(6, 3, 2)
    w = np.zeros(12)
    w[0] = w[6] = 1.0
    assert o.set_embedded_faces([[0, 0, 1, 1]], [w]) == -1      # the domain's west face
    assert o.set_embedded_faces([[0, 6, 1, 1]], [w]) == -1      # the east face
    assert o.set_embedded_faces([[1, 2, 3, 1]], [w]) == -1      # y = ny is a boundary face
    assert o.set_embedded_faces([[0, 2, 1, 1], [0, 2, 1, 1]], [w, w]) == -1
    assert o.set_embedded_faces([[3, 2, 1, 1]], [w]) == -1
    assert o.set_embedded_faces([[0, 2, 1, 1], [1, 2, 1, 1]], [w, w]) == 0
    assert o.set_embedded_faces([], []) == 0
