"""Hand-worked pins of the oracle's state update after the interface moves (oracle/multiphase.c:
local_update, fix_unresolved), restating MultiPhaseOperator::UpdateStateVariablesByRiemannSolutions,
LocalUpdateByRiemannSolutions and FixUnresolvedNodes (MultiPhaseOperator.cpp:857-1028, 1156-1320)
on small grids, with the expected values worked out here in plain Python from the reference's
rules and the star states of the independently pinned exact Riemann solver (orc_riemann; its pins
are in test_oracle_riemann.py):
* a cell swept by the neighbouring material takes the star state that the interface face stored
  on that side (RiemannSolutions left/right/bottom/top, SpaceOperator.cpp:2518-2523), weighted by
  the upwind component of the neighbour's unit velocity; without an upwind face it retries with
  weight 1;
* two upwind faces combine as (w1 V1 + w2 V2) * (1 / (w1 + w2)) (Vec5D::operator/=);
* a cell with no stored solution is unresolved: it takes the upwind-weighted mean of resolved
  same-material neighbours in its 3x3x3 block, else their inverse-distance mean, else keeps its
  velocity and pressure and takes the inverse-distance mean density of the nearest layer (up to 10)
  holding such cells, else the failsafe density; the reference's `reset` flag is not cleared
  between unresolved cells, so a later cell of the upwind branch adds onto its own old state.
Grids use unit spacing, so cell centres and distances are exact.
"""
import math

import numpy as np

from oracle import pyoracle as po

WATER = (6.12, 3.43e8)
AIR = (1.4, 0.0)


def make(nx, ny, nz=1):
    mats = [po.sg(*WATER), po.sg(*AIR)]
    cfg = po.make_config(nx, ny, nz, (0, nx, 0, ny, 0, nz), [po.BC_SYMMETRY] * 6, mats, ls_matid=1,
                         rec_linear=False)
    return po.Oracle(cfg)


def fields(ID, vel):
    """Water (1000 kg/m3) where ID == 0, air (1.2) where ID == 1, p = 1e5, uniform velocity."""
    nz, ny, nx = ID.shape
    V = np.zeros((nz, ny, nx, 5))
    V[..., 0] = np.where(ID == 1, 1.2, 1000.0)
    V[..., 1:4] = vel
    V[..., 4] = 1e5
    return V


def store(o, V0, ID0):
    V, ID = o.ghosted(V0, ID0)
    assert o.store_riemann(V, ID, 1) >= 0
    return V, ID


def star(o, d, Vm, idm, Vp, idp):
    n = np.zeros(3)
    n[d] = 1.0
    st, Vs, mid, Vsm, Vsp = o.riemann(n, Vm, idm, Vp, idp)
    assert st == 0
    return Vsm, Vsp


def update(o, V, ID, ID1_int):
    IDn = ID.copy()
    ID1 = ID.copy()
    ID1[1:-1, 1:-1, 1:-1] = ID1_int
    o.L.orc_apply_bc_id(o.h, po._ip(ID1))
    nun = o.update_states(IDn, ID1, V)
    return nun, V[1:-1, 1:-1, 1:-1]


def test_swept_cell_takes_upwind_star_state():
    o = make(8, 1)
    ID0 = np.zeros((1, 1, 8), np.int32)
    ID0[..., 4:] = 1
    V0 = fields(ID0, (10.0, 0.0, 0.0))
    V, ID = store(o, V0, ID0)
    Vsm, _ = star(o, 0, V0[0, 0, 3], 0, V0[0, 0, 4], 1)
    ID1 = ID0.copy()
    ID1[0, 0, 4] = 0                      # the water advanced into cell 4
    nun, Vi = update(o, V, ID, ID1)
    assert nun == 0
    assert np.array_equal(Vi[0, 0, 4], Vsm)   # weight u/|u| = 1
    assert np.array_equal(Vi[0, 0, 5], V0[0, 0, 5])


def test_downwind_cell_retries_without_upwinding():
    o = make(8, 1)
    ID0 = np.zeros((1, 1, 8), np.int32)
    ID0[..., 4:] = 1
    V0 = fields(ID0, (-10.0, 0.0, 0.0))
    V, ID = store(o, V0, ID0)
    Vsm, Vsp = star(o, 0, V0[0, 0, 3], 0, V0[0, 0, 4], 1)
    ID1 = ID0.copy()
    ID1[0, 0, 4] = 0        # flow runs the other way: no upwind face, weight 1 on the retry
    nun, Vi = update(o, V, ID, ID1)
    assert nun == 0 and np.array_equal(Vi[0, 0, 4], Vsm)
    # the plus side's record (RiemannSolutions::right of cell 3) for air advancing to the left
    o2 = make(8, 1)
    V, ID = store(o2, V0, ID0)
    ID1 = ID0.copy()
    ID1[0, 0, 3] = 1
    nun, Vi = update(o2, V, ID, ID1)
    assert nun == 0 and np.array_equal(Vi[0, 0, 3], Vsp)


def test_two_upwind_faces_are_weighted_by_velocity_components():
    o = make(6, 6)
    ID0 = np.zeros((1, 6, 6), np.int32)
    ID0[0, 3:, 3:] = 1                    # an air quadrant; cell (i=3, j=3) is its corner
    V0 = fields(ID0, (3.0, 4.0, 0.0))     # |v| = 5
    V, ID = store(o, V0, ID0)
    Vx, _ = star(o, 0, V0[0, 3, 2], 0, V0[0, 3, 3], 1)     # left face of the corner
    Vy, _ = star(o, 1, V0[0, 2, 3], 0, V0[0, 3, 3], 1)     # bottom face
    ID1 = ID0.copy()
    ID1[0, 3, 3] = 0
    nun, Vi = update(o, V, ID, ID1)
    assert nun == 0
    vl, vb = V0[0, 3, 2], V0[0, 2, 3]
    w1 = vl[1] / math.sqrt(vl[1] * vl[1] + vl[2] * vl[2] + vl[3] * vl[3])
    w2 = vb[2] / math.sqrt(vb[1] * vb[1] + vb[2] * vb[2] + vb[3] * vb[3])
    assert (w1, w2) == (0.6, 0.8)
    expect = [w1 * Vx[p] for p in range(5)]
    expect = [expect[p] + w2 * Vy[p] for p in range(5)]
    inv = 1.0 / (w1 + w2)
    expect = np.array([x * inv for x in expect])
    assert np.array_equal(Vi[0, 3, 3], expect)


def test_unresolved_cell_takes_nearest_layer_density():
    o = make(16, 1)
    ID0 = np.zeros((1, 1, 16), np.int32)
    ID0[..., :4] = 1
    V0 = fields(ID0, (10.0, 0.0, 0.0))
    V, ID = store(o, V0, ID0)
    ID1 = ID0.copy()
    ID1[0, 0, 10] = 1          # no stored solution at cell 10, no air cell within its 3x3x3 block
    nun, Vi = update(o, V, ID, ID1)
    assert nun == 1
    # nearest air cells: cell 3 at distance 7 (layer 7)
    sw = 1.0 / 7.0
    dens = (1.0 / 7.0) * 1.2
    assert Vi[0, 0, 10, 0] == dens / sw
    assert np.array_equal(Vi[0, 0, 10, 1:], V0[0, 0, 10, 1:])   # velocity and pressure kept


def test_unresolved_cell_beyond_ten_layers_takes_failsafe_density():
    o = make(32, 1)
    ID0 = np.zeros((1, 1, 32), np.int32)
    ID0[..., :4] = 1
    V0 = fields(ID0, (10.0, 0.0, 0.0))
    V, ID = store(o, V0, ID0)
    ID1 = ID0.copy()
    ID1[0, 0, 20] = 1          # 17 cells from the nearest air cell
    nun, Vi = update(o, V, ID, ID1)
    assert nun == 1
    assert Vi[0, 0, 20, 0] == 1e-10
    assert np.array_equal(Vi[0, 0, 20, 1:], V0[0, 0, 20, 1:])


def test_unresolved_cells_extrapolate_upwind_with_persistent_reset():
    o = make(8, 8)
    ID0 = np.zeros((1, 8, 8), np.int32)
    ID0[0, :, :2] = 1                     # air strip i <= 1, flow to the right
    V0 = fields(ID0, (10.0, 1.0, 0.0))
    V, ID = store(o, V0, ID0)
    ID1 = ID0.copy()
    for j in (3, 5):
        ID1[0, j, 2] = 1   # resolved by the stored left-face solution
        ID1[0, j, 3] = 1   # unresolved: its left face was water/water
    stars = {j: star(o, 0, V0[0, j, 1], 1, V0[0, j, 2], 0)[0] for j in (3, 5)}
    nun, Vi = update(o, V, ID, ID1)
    assert nun == 2
    vl = V0[0, 3, 1]
    wl = vl[1] / math.sqrt(vl[1] * vl[1] + vl[2] * vl[2] + vl[3] * vl[3])   # 10 / sqrt(101)
    for j in (3, 5):
        inv = 1.0 / wl
        assert np.array_equal(Vi[0, j, 2], np.array([(wl * x) * inv for x in stars[j]]))
    # the first unresolved cell (3,3): same-material neighbours in its 3x3 block that are resolved:
    # (2,3) only ((2,2), (2,4) are water). Weight max(0, d . v_hat), d from the neighbour to the cell.
    reset = False
    out = {}
    for j in (3, 5):
        nb = Vi[0, j, 2]
        v1 = nb[1:4].copy()
        n1 = math.sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2])
        inv = 1.0 / n1
        v1 = [v1[0] * inv, v1[1] * inv, v1[2] * inv]
        w = max(0.0, 1.0 * v1[0] + 0.0 * v1[1] + 0.0 * v1[2])
        assert w > 0
        if reset:
            v = [V0[0, j, 3][p] + w * nb[p] for p in range(5)]     # adds onto the cell's own old state
        else:
            v = [w * nb[p] for p in range(5)]
            reset = True
        iw = 1.0 / w
        out[j] = np.array([x * iw for x in v])
    assert np.array_equal(Vi[0, 3, 3], out[3])
    assert np.array_equal(Vi[0, 5, 3], out[5])
    assert not np.allclose(Vi[0, 5, 3], Vi[0, 5, 2], rtol=1e-3)   # the quirk is visible
