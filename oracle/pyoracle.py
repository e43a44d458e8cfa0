"""ctypes binding of the CPU oracle (liboracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and by the cpu_baseline
leg of bench.py as the checker. The product package never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

EOS_SG, EOS_NASG, EOS_MG, EOS_JWL, EOS_DUMMY = 0, 1, 2, 3, 5
BC_INLET, BC_OUTLET, BC_SLIPWALL, BC_STICKWALL, BC_SYMMETRY, BC_OVERSET, BC_PERIODIC = 1, 2, 3, 4, 5, 6, 7
FLUX_ROE, FLUX_LLF, FLUX_HLLC, FLUX_GODUNOV = 0, 1, 2, 3
LIM_NONE, LIM_GENMINMOD, LIM_VANALBADA = 0, 1, 2
TS_FE, TS_RK2, TS_RK3 = 0, 1, 2


class Eos(C.Structure):
    _fields_ = [("type", C.c_int)] + [(n, C.c_double) for n in (
        "rhomin", "pmin", "rhomax", "pmax", "failsafe_density",
        "gam", "pstiff", "gam1",
        "pc", "b", "q", "gam_pc",
        "rho0", "c0", "Gamma0", "s", "e0", "rho0_c0_c0", "Gamma0_over_2", "Gamma0_rho0",
        "omega", "A1", "A2", "R1", "R2", "R1rho0", "R2rho0", "omega_over_R1rho0", "omega_over_R2rho0",
        "d_p0", "d_rho0", "d_e0", "t_cv", "t_T0", "t_e0", "t_rho0", "t_cp", "t_h0")] + [
        ("visc", C.c_int), ("mu", C.c_double), ("bulk", C.c_double), ("su_T0", C.c_double), ("su_S", C.c_double),
        ("kappa", C.c_double)]


class RiemannCfg(C.Structure):
    _fields_ = [("maxIts_main", C.c_int), ("maxIts_bracket", C.c_int), ("maxIts_shock", C.c_int),
                ("numSteps_rarefaction", C.c_int)] + [(n, C.c_double) for n in (
                    "tol_main", "tol_shock", "tol_rarefaction", "min_pressure", "failure_threshold",
                    "pressure_at_failure", "surrogate_residual_tol")]


class Config(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double), ("y1", C.c_double),
                ("z0", C.c_double), ("z1", C.c_double),
                ("bc", C.c_int * 6), ("inlet", C.c_double * 5), ("outlet", C.c_double * 5),
                ("nmat", C.c_int), ("inactive_id", C.c_int), ("mat", Eos * 8),
                ("flux", C.c_int), ("roe_delta", C.c_double),
                ("rec_type", C.c_int), ("limiter", C.c_int), ("slope_zero_near_interface", C.c_int),
                ("gminmod_alpha", C.c_double), ("vanalbada_pow", C.c_int),
                ("ts", C.c_int), ("cfl", C.c_double),
                ("mp_flux", C.c_int), ("mp_recon", C.c_int),
                ("riemann", RiemannCfg), ("nthreads", C.c_int),
                ("ls_on", C.c_int), ("ls_matid", C.c_int), ("ls_bc", C.c_int * 6),
                ("reinit_freq", C.c_int), ("reinit_maxits", C.c_int),
                ("reinit_cfl", C.c_double), ("reinit_tol", C.c_double), ("mp_normal", C.c_int),
                ("conrec_depth", C.c_double),
                ("mesh_npts", C.c_int * 3), ("mesh_pts", (C.c_double * 2 * 16) * 3)]


def default_riemann():
    """ExactRiemannSolverData defaults, IoData.cpp:1345-1357."""
    r = RiemannCfg()
    r.maxIts_main, r.maxIts_bracket, r.maxIts_shock, r.numSteps_rarefaction = 200, 100, 200, 200
    r.tol_main, r.tol_shock, r.tol_rarefaction = 1e-4, 1e-12, 1e-8
    r.min_pressure, r.failure_threshold, r.pressure_at_failure = -1e8, 0.2, 1e-8
    r.surrogate_residual_tol = 1e-6
    return r


def sg(gamma, pinf=0.0, rhomin=0.0, pmin=-1.7976931348623157e308):
    e = Eos()
    e.type = EOS_SG
    e.gam, e.pstiff = gamma, pinf
    e.rhomin, e.pmin = rhomin, pmin
    e.rhomax = e.pmax = 1.7976931348623157e308
    e.failsafe_density = 1e-10
    return e


def eos(kind, **kw):
    e = Eos()
    e.type = kind
    e.rhomin, e.pmin = 0.0, -1.7976931348623157e308
    e.rhomax = e.pmax = 1.7976931348623157e308
    e.failsafe_density = 1e-10
    for k, v in kw.items():
        setattr(e, k, v)
    return e


def make_config(nx, ny, nz, box, bc, mats, inlet=(1, 0, 0, 0, 1), outlet=(1, 0, 0, 0, 1),
                flux=FLUX_HLLC, limiter=LIM_VANALBADA, ts=TS_RK2, cfl=0.5, rec_linear=True,
                vanalbada_pow=True, inactive_id=-1, nthreads=0, roe_delta=0.2, gminmod_alpha=1.2,
                slope_zero=False, mp_flux=0, mp_recon=0, ls_matid=None, ls_bc=(1, 1, 1, 1, 1, 1),
                reinit_freq=-1, reinit_maxits=30, reinit_cfl=0.8, reinit_tol=2.0e-4, mp_normal=1,
                conrec_depth=0.0, control_points=None):
    """control_points: None, or per axis None / a list of (coordinate, cell width) pairs
    (Mesh.ControlPointX/Y/Z); nx/ny/nz of a stretched axis must be its generated cell count."""
    c = Config()
    c.nx, c.ny, c.nz = nx, ny, nz
    c.x0, c.x1, c.y0, c.y1, c.z0, c.z1 = box
    for i in range(6):
        c.bc[i] = bc[i]
    for i in range(5):
        c.inlet[i] = inlet[i]
        c.outlet[i] = outlet[i]
    c.nmat = len(mats)
    for i, m in enumerate(mats):
        c.mat[i] = m
    c.inactive_id = inactive_id
    c.flux, c.roe_delta = flux, roe_delta
    c.rec_type = 1 if rec_linear else 0
    c.limiter, c.gminmod_alpha = limiter, gminmod_alpha
    c.slope_zero_near_interface = int(slope_zero)
    c.vanalbada_pow = int(vanalbada_pow)
    c.ts, c.cfl = ts, cfl
    c.mp_flux, c.mp_recon = mp_flux, mp_recon
    c.riemann = default_riemann()
    c.nthreads = nthreads
    c.ls_on = int(ls_matid is not None)
    c.ls_matid = ls_matid if ls_matid is not None else -1
    for i in range(6):
        c.ls_bc[i] = ls_bc[i]
    c.reinit_freq, c.reinit_maxits = reinit_freq, reinit_maxits
    c.reinit_cfl, c.reinit_tol = reinit_cfl, reinit_tol
    c.mp_normal = mp_normal
    c.conrec_depth = conrec_depth
    for d, pts in enumerate(control_points or (None, None, None)):
        if pts:
            c.mesh_npts[d] = len(pts)
            for q, (x, w) in enumerate(pts):
                c.mesh_pts[d][q][0], c.mesh_pts[d][q][1] = x, w
    return c


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        D = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int)
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.POINTER(Config)]
        L.orc_destroy.argtypes = [P]
        L.orc_ncells_ghosted.restype = C.c_long
        L.orc_ncells_ghosted.argtypes = [P]
        L.orc_numerical_flux.argtypes = [P, C.c_int, D, D, C.c_int, D]
        L.orc_apply_bc.argtypes = [P, D]
        L.orc_apply_bc_id.argtypes = [P, I]
        L.orc_compute_residual.argtypes = [P, D, I, D]
        L.orc_set_levelset.argtypes = [P, D]
        L.orc_interface_normal.argtypes = [P, D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, D]
        L.orc_time_step.argtypes = [P, D, I, D]
        L.orc_advance.argtypes = [P, D, I, C.c_double]
        L.orc_run.argtypes = [P, D, I, D, C.c_double, C.c_int, D, I]
        L.orc_apply_bc_phi.argtypes = [P, D]
        L.orc_ls_residual.argtypes = [P, D, D, D]
        L.orc_ls_reinitialize.argtypes = [P, D]
        L.orc_update_id.argtypes = [P, D, D, I]
        L.orc_riemann.argtypes = [P, D, D, C.c_int, D, C.c_int, D, I, D, D]
        L.orc_riemann_onesided.argtypes = [P, D, D, C.c_int, D, D, I, D]
        L.orc_onesided_normal.argtypes = [C.c_int, C.c_int, C.c_int, D, D]
        L.orc_wall_face_flux.argtypes = [P, C.c_int, C.c_int, C.c_int, D, D, D, C.c_int, D, D]
        L.orc_set_embedded_faces.argtypes = [P, C.c_int, I, D, C.c_int, C.c_int]
        L.orc_eos_p2c.argtypes = [P, C.c_int, D, D]
        L.orc_eos_c2p.argtypes = [P, C.c_int, D, D]
        L.orc_eos_c2.argtypes = [P, C.c_int, C.c_double, C.c_double]
        L.orc_eos_c2.restype = C.c_double
        L.orc_eos_point.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_double, D]
        L.orc_store_riemann.argtypes = [P, D, I, C.c_int]
        L.orc_update_states.argtypes = [P, I, I, D]
        L.orc_update_states.restype = C.c_long
        L.orc_material_interface_flux.argtypes = [P, C.c_int, D, C.c_int, D, C.c_int, D]
        L.orc_mp_counters.argtypes = [P, C.POINTER(C.c_long), C.POINTER(C.c_long), D]
        for f in ("orc_det_log", "orc_det_exp", "orc_pow_quarter", "orc_pow_fifth"):
            getattr(L, f).argtypes = [C.c_double]
            getattr(L, f).restype = C.c_double
        L.orc_counters.argtypes = [P, C.POINTER(C.c_long), C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.orc_mesh_1d.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, D, C.c_int, D, D, I]
        L.orc_edge_surface_walls.argtypes = [C.c_int, C.c_void_p, C.c_int, I, I, I, D]
        _lib = L
    return _lib


def mesh_1d(x0, xmax, n=0, control_points=None):
    """Cell centres and widths of one mesh axis (MeshGenerator.cpp:99-259): n uniform cells, or the
    stretched cells of a list of (coordinate, width) control points."""
    pts = np.ascontiguousarray(np.array(control_points or [], dtype=np.float64).reshape(-1))
    cnt = C.c_int(0)
    L = lib()
    pp = _dp(pts) if pts.size else None
    if L.orc_mesh_1d(x0, xmax, n, pts.size // 2, pp, 0, None, None, C.byref(cnt)) != 0:
        raise ValueError("mesh_1d: control points and widths in conflict, or a bad axis")
    x = np.zeros(cnt.value)
    dx = np.zeros(cnt.value)
    L.orc_mesh_1d(x0, xmax, n, pts.size // 2, pp, cnt.value, _dp(x), _dp(dx), C.byref(cnt))
    return x, dx


class Surface(C.Structure):
    _fields_ = [("nnodes", C.c_int), ("nelems", C.c_int), ("npoints", C.c_int),
                ("udot", C.POINTER(C.c_double)), ("elems", C.POINTER(C.c_int)), ("elem_norm", C.POINTER(C.c_double)),
                ("point_tid", C.POINTER(C.c_int)), ("point_xi", C.POINTER(C.c_double)),
                ("point_dist", C.POINTER(C.c_double))]


def edge_surface_walls(surfaces, xf, xb):
    """FindEdgeSurfaceIntersections restated (oracle/embedded_walls.c): surfaces as dicts with udot,
    elems, elem_norm, tid, xi, dist; xf / xb [faces][surfaces]. Returns (cut bool, walls [faces][12])."""
    keep = []
    arr = (Surface * max(len(surfaces), 1))()
    D, I = C.POINTER(C.c_double), C.POINTER(C.c_int)
    for i, sf in enumerate(surfaces):
        a = [np.ascontiguousarray(sf["udot"], dtype=np.float64), np.ascontiguousarray(sf["elems"], dtype=np.int32),
             np.ascontiguousarray(sf["elem_norm"], dtype=np.float64), np.ascontiguousarray(sf["tid"], dtype=np.int32),
             np.ascontiguousarray(sf["xi"], dtype=np.float64), np.ascontiguousarray(sf["dist"], dtype=np.float64)]
        keep.append(a)
        arr[i] = Surface(a[0].size // 3, a[1].size // 3, a[3].size, a[0].ctypes.data_as(D), a[1].ctypes.data_as(I),
                         a[2].ctypes.data_as(D), a[3].ctypes.data_as(I), a[4].ctypes.data_as(D), a[5].ctypes.data_as(D))
    xf = np.ascontiguousarray(xf, dtype=np.int32).reshape(-1, len(surfaces))
    xb = np.ascontiguousarray(xb, dtype=np.int32).reshape(-1, len(surfaces))
    n = len(xf)
    cut = np.zeros(n, np.int32)
    wall = np.zeros((n, 12))
    rc = lib().orc_edge_surface_walls(len(surfaces), C.cast(arr, C.c_void_p), n, _ip(xf), _ip(xb), _ip(cut), _dp(wall))
    if rc != 0:
        raise ValueError("orc_edge_surface_walls: index out of range or one-sided crossing")
    return cut.astype(bool), wall


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """Ghosted AoS state V[k+1, j+1, i+1, 5] and integer IDs, as the reference stores them."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.L = lib()
        self.h = self.L.orc_create(C.byref(cfg))
        self.shape = (cfg.nz + 2, cfg.ny + 2, cfg.nx + 2)

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def ghosted(self, V_int, ID_int=None):
        """Embed interior arrays (nz,ny,nx,5) / (nz,ny,nx) into ghosted ones and fill ghosts."""
        V = np.zeros(self.shape + (5,), dtype=np.float64)
        V[1:-1, 1:-1, 1:-1] = V_int
        ID = np.zeros(self.shape, dtype=np.int32)
        if ID_int is not None:
            ID[1:-1, 1:-1, 1:-1] = ID_int
        self.L.orc_apply_bc_id(self.h, _ip(ID))
        self.L.orc_apply_bc(self.h, _dp(V))
        return V, ID

    def flux(self, d, Vm, Vp, mid=0):
        F = np.zeros(5)
        Vm = np.ascontiguousarray(Vm, dtype=np.float64)
        Vp = np.ascontiguousarray(Vp, dtype=np.float64)
        self.L.orc_numerical_flux(self.h, d, _dp(Vm), _dp(Vp), mid, _dp(F))
        return F

    def residual(self, V, ID, phi=None):
        """Residual of the ghosted state; phi (ghosted) gives interface faces their level-set normal."""
        R = np.zeros_like(V)
        if phi is not None:
            phi = np.ascontiguousarray(phi, dtype=np.float64)
            self.L.orc_set_levelset(self.h, _dp(phi))
        st = self.L.orc_compute_residual(self.h, _dp(V), _ip(ID), _dp(R))
        if phi is not None:
            self.L.orc_set_levelset(self.h, None)
        return st, R

    def time_step(self, V, ID):
        dt = C.c_double()
        st = self.L.orc_time_step(self.h, _dp(V), _ip(ID), C.byref(dt))
        return st, dt.value

    def advance(self, V, ID, dt):
        return self.L.orc_advance(self.h, _dp(V), _ip(ID), dt)

    def run(self, V, ID, tmax, maxsteps, phi=None):
        t = C.c_double()
        n = C.c_int()
        pp = _dp(phi) if phi is not None else None
        st = self.L.orc_run(self.h, _dp(V), _ip(ID), pp, tmax, maxsteps, C.byref(t), C.byref(n))
        return st, t.value, n.value

    def ghosted_phi(self, phi_int):
        phi = np.zeros(self.shape, dtype=np.float64)
        phi[1:-1, 1:-1, 1:-1] = phi_int
        self.L.orc_apply_bc_phi(self.h, _dp(phi))
        return phi

    def apply_bc_phi(self, phi):
        self.L.orc_apply_bc_phi(self.h, _dp(phi))

    def ls_residual(self, V, phi):
        R = np.zeros(self.shape, dtype=np.float64)
        self.L.orc_ls_residual(self.h, _dp(V), _dp(phi), _dp(R))
        return R

    def reinitialize(self, phi):
        return self.L.orc_ls_reinitialize(self.h, _dp(phi))

    def update_id(self, phi0, phi, ID):
        self.L.orc_update_id(self.h, _dp(phi0), _dp(phi), _ip(ID))

    def riemann(self, d, Vm, idm, Vp, idp):
        """Exact two-material Riemann solution along unit normal d: (status, Vs, id, Vsm, Vsp)."""
        dd = np.ascontiguousarray(d, dtype=np.float64)
        Vm = np.ascontiguousarray(Vm, dtype=np.float64)
        Vp = np.ascontiguousarray(Vp, dtype=np.float64)
        Vs, Vsm, Vsp = np.zeros(5), np.zeros(5), np.zeros(5)
        mid = C.c_int()
        st = self.L.orc_riemann(self.h, _dp(dd), _dp(Vm), idm, _dp(Vp), idp, _dp(Vs), C.byref(mid),
                                _dp(Vsm), _dp(Vsp))
        return st, Vs, mid.value, Vsm, Vsp

    def riemann_onesided(self, d, Vm, idl, Ustar):
        """One-sided (wall) Riemann solution along unit normal d: (status, Vs, id, Vsm); Vs is zero
        when id == -1 (the wall recedes)."""
        dd = np.ascontiguousarray(d, dtype=np.float64)
        Vm = np.ascontiguousarray(Vm, dtype=np.float64)
        Us = np.ascontiguousarray(Ustar, dtype=np.float64)
        Vs, Vsm = np.zeros(5), np.zeros(5)
        mid = C.c_int()
        st = self.L.orc_riemann_onesided(self.h, _dp(dd), _dp(Vm), idl, _dp(Us), _dp(Vs), C.byref(mid), _dp(Vsm))
        return st, Vs, mid.value, Vsm

    def onesided_normal(self, d, fb, mode, nwall):
        nw = np.ascontiguousarray(nwall, dtype=np.float64)
        out = np.zeros(3)
        self.L.orc_onesided_normal(d, fb, mode, _dp(nw), _dp(out))
        return out

    def wall_face_flux(self, axis, side, mode, nwall, vwall, V, mid=0):
        """Flux on the fluid side of a wall-cut face: (status, F, Vstar)."""
        nw = np.ascontiguousarray(nwall, dtype=np.float64)
        vw = np.ascontiguousarray(vwall, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        F, Vs = np.zeros(5), np.zeros(5)
        st = self.L.orc_wall_face_flux(self.h, axis, side, mode, _dp(nw), _dp(vw), _dp(V), mid, _dp(F), _dp(Vs))
        return st, F, Vs

    def set_embedded_faces(self, faces, walls, normal_mode=1, recon=0):
        """Wall-cut faces for later residuals and runs: faces (n, 4) = axis, i, j, k of the plus cell;
        walls (n, 12) = nwallf, vwallf, nwallb, vwallb. Returns the oracle's status."""
        f = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).reshape(-1, 4))
        w = np.ascontiguousarray(np.asarray(walls, dtype=np.float64).reshape(-1, 12))
        return self.L.orc_set_embedded_faces(self.h, f.shape[0], _ip(f), _dp(w), normal_mode, recon)

    def interface_normal(self, phi, d, i, j, k, myid, nbid):
        """Interface Riemann direction at the d-face of (i,j,k) from the ghosted level set phi."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.zeros(3)
        self.L.orc_interface_normal(self.h, _dp(phi), d, i, j, k, myid, nbid, _dp(out))
        return out

    def p2c(self, V, mid=0):
        V = np.ascontiguousarray(V, dtype=np.float64)
        U = np.zeros(5)
        self.L.orc_eos_p2c(self.h, mid, _dp(V), _dp(U))
        return U

    def c2p(self, U, mid=0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        V = np.zeros(5)
        self.L.orc_eos_c2p(self.h, mid, _dp(U), _dp(V))
        return V

    def c2(self, rho, p, mid=0):
        return self.L.orc_eos_c2(self.h, mid, rho, p)

    def eos_point(self, rho, e, p, mid=0):
        """(p(rho,e), e(rho,p), dp/drho|e, Gamma, c^2(rho,e), CheckState(rho,p)) of material mid."""
        out = np.zeros(6)
        self.L.orc_eos_point(self.h, mid, rho, e, p, _dp(out))
        return out

    def store_riemann(self, V, ID, stamp=1):
        """Store the interface Riemann solutions of the ghosted V as a step's first stage does."""
        return self.L.orc_store_riemann(self.h, _dp(V), _ip(ID), stamp)

    def update_states(self, IDn, ID, V):
        """State update after the interface moved (IDn before, ID after); V in place. Returns the
        number of unresolved cells."""
        return self.L.orc_update_states(self.h, _ip(IDn), _ip(ID), _dp(V))

    def interface_flux(self, d, Vm, idm, Vp, idp):
        """Local Lax-Friedrichs flux between materials idm and idp along axis d."""
        Vm = np.ascontiguousarray(Vm, dtype=np.float64)
        Vp = np.ascontiguousarray(Vp, dtype=np.float64)
        F = np.zeros(5)
        self.L.orc_material_interface_flux(self.h, d, _dp(Vm), idm, _dp(Vp), idp, _dp(F))
        return F

    def mp_counters(self):
        """(reinitialisation iterations, unresolved cells, reinitialisation CFL in force)."""
        a, b, c = C.c_long(), C.c_long(), C.c_double()
        self.L.orc_mp_counters(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def counters(self):
        a, b, c = C.c_long(), C.c_long(), C.c_long()
        self.L.orc_counters(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value
