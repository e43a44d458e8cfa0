"""m2c_amd -- MI355X-native explicit finite-volume update of M2C.

Python front end of the C ABI declared in include/m2c_hip.h (libm2c_hip.so, built in-tree for
gfx950 by m2c_amd/build.py). The compute path is the HIP library; this module only marshals
configuration and host arrays. Importing it without the built library raises ImportError.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libm2c_hip.so")

EOS_SG, EOS_NASG, EOS_MG, EOS_JWL, EOS_DUMMY = 0, 1, 2, 3, 5
BC_INLET, BC_OUTLET, BC_SLIPWALL, BC_STICKWALL, BC_SYMMETRY, BC_OVERSET, BC_PERIODIC = 1, 2, 3, 4, 5, 6, 7
FLUX_ROE, FLUX_LLF, FLUX_HLLC, FLUX_GODUNOV = 0, 1, 2, 3
LIM_NONE, LIM_GENMINMOD, LIM_VANALBADA = 0, 1, 2
TS_FE, TS_RK2, TS_RK3 = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
# VarFcnBase pointwise operations of m2c_varfcn_batch (m2c_varfcn_op)
(VF_PRESSURE, VF_INTERNAL_ENERGY, VF_DPDRHO, VF_BIG_GAMMA, VF_SOUND_SPEED_SQUARE, VF_SOUND_SPEED, VF_ENTHALPY,
 VF_TOTAL_ENTHALPY, VF_MACH, VF_C2P, VF_P2C, VF_CHECK_STATE, VF_CLIP) = range(13)
OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_UNSUPPORTED = 0, 1, 2, 3, 4
DBL_MAX = 1.7976931348623157e308


class Material(C.Structure):
    _fields_ = [("eos", C.c_int), ("rhomin", C.c_double), ("pmin", C.c_double), ("rhomax", C.c_double),
                ("pmax", C.c_double), ("failsafe_density", C.c_double), ("p", C.c_double * 8),
                ("tlaw", C.c_double * 6), ("viscosity", C.c_int), ("mu", C.c_double), ("bulk", C.c_double),
                ("su_T0", C.c_double), ("su_S", C.c_double), ("kappa", C.c_double)]


class RiemannParams(C.Structure):
    _fields_ = [("maxIts_main", C.c_int), ("maxIts_bracket", C.c_int), ("maxIts_shock", C.c_int),
                ("numSteps_rarefaction", C.c_int)] + [(n, C.c_double) for n in (
                    "tol_main", "tol_shock", "tol_rarefaction", "min_pressure", "failure_threshold",
                    "pressure_at_failure", "surrogate_residual_tol")]


class Config(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double), ("y1", C.c_double),
                ("z0", C.c_double), ("z1", C.c_double),
                ("bc", C.c_int * 6), ("inlet", C.c_double * 5), ("outlet", C.c_double * 5),
                ("nmat", C.c_int), ("inactive_id", C.c_int), ("mat", Material * 8),
                ("flux", C.c_int), ("roe_delta", C.c_double), ("rec_linear", C.c_int),
                ("limiter", C.c_int), ("gminmod_alpha", C.c_double),
                ("slope_zero_near_interface", C.c_int), ("ts", C.c_int), ("cfl", C.c_double),
                ("mp_flux", C.c_int), ("mp_recon", C.c_int), ("riemann", RiemannParams),
                ("precision", C.c_int), ("device", C.c_int),
                ("ls_matid", C.c_int), ("ls_bc", C.c_int * 6), ("reinit_freq", C.c_int),
                ("reinit_maxits", C.c_int), ("reinit_cfl", C.c_double), ("reinit_tol", C.c_double),
                ("mp_riemann_normal", C.c_int), ("mp_conrec_depth", C.c_double),
                ("mesh_npts", C.c_int * 3), ("mesh_pts", (C.c_double * 2 * 16) * 3)]


class DeckRun(C.Structure):
    _fields_ = [("max_its", C.c_int), ("max_time", C.c_double), ("time_step", C.c_double),
                ("initial_state", C.c_double * 5), ("initial_material", C.c_int)]


class BatchTuning(C.Structure):
    _fields_ = [("kchunk", C.c_int), ("blocks", C.c_int), ("ms_per_step", C.c_double), ("candidates", C.c_int),
                ("tile_x", C.c_int), ("tile_y", C.c_int)]


class Status(C.Structure):
    _fields_ = [("error", C.c_int), ("reverted_faces", C.c_longlong), ("clipped_cells", C.c_longlong),
                ("riemann_errors", C.c_longlong), ("steps", C.c_longlong), ("t", C.c_double),
                ("dt", C.c_double), ("cfl", C.c_double), ("reinit_iterations", C.c_longlong),
                ("unresolved_cells", C.c_longlong), ("surrogate_accepted", C.c_longlong),
                ("surrogate_fallbacks", C.c_longlong)]


class TrainConfig(C.Structure):
    _fields_ = [("n_samples", C.c_longlong), ("n_holdout", C.c_longlong), ("seed", C.c_ulonglong),
                ("batch", C.c_int), ("epochs", C.c_int), ("lr", C.c_double), ("lr_min", C.c_double),
                ("newton_max", C.c_int), ("tol", C.c_double), ("verbose", C.c_int)]


class TrainReport(C.Structure):
    _fields_ = [("data_ms", C.c_double), ("train_ms", C.c_double), ("samples_per_s", C.c_double),
                ("steps", C.c_longlong), ("valid_samples", C.c_longlong), ("first_epoch_loss", C.c_double),
                ("last_epoch_loss", C.c_double), ("holdout_loss", C.c_double),
                ("holdout_accept_mlp", C.c_double), ("holdout_accept_acoustic", C.c_double),
                ("holdout_iters_mlp", C.c_double), ("holdout_iters_acoustic", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


VISC_NONE, VISC_CONSTANT, VISC_SUTHERLAND = 0, 1, 2


def material(kind, params, rhomin=0.0, pmin=-DBL_MAX, rhomax=DBL_MAX, pmax=DBL_MAX, failsafe_density=1e-10,
             tlaw=(0.0, 0.0, 0.0, 0.0, 0.0, 0.0), viscosity=VISC_NONE, mu=0.0, bulk=0.0, su_T0=273.0, su_S=111.0,
             kappa=0.0):
    """MaterialModelData (IoData.cpp:860-906): EOS kind + parameters + cut-offs, the temperature law
    (cv, T0, e0, rho0, cp, h0; outputs, Sutherland viscosity and heat diffusion), the viscosity model
    (ViscoFcn.h) and the heat diffusivity (HeatDiffusionFcn.h)."""
    m = Material()
    m.viscosity, m.mu, m.bulk, m.su_T0, m.su_S, m.kappa = viscosity, mu, bulk, su_T0, su_S, kappa
    m.eos = kind
    m.rhomin, m.pmin, m.rhomax, m.pmax, m.failsafe_density = rhomin, pmin, rhomax, pmax, failsafe_density
    for i, v in enumerate(params):
        m.p[i] = v
    for i, v in enumerate(tlaw):
        m.tlaw[i] = v
    return m


def riemann_defaults():
    """ExactRiemannSolverData defaults (IoData.cpp:1345-1357)."""
    r = RiemannParams()
    r.maxIts_main, r.maxIts_bracket, r.maxIts_shock, r.numSteps_rarefaction = 200, 100, 200, 200
    r.tol_main, r.tol_shock, r.tol_rarefaction = 1e-4, 1e-12, 1e-8
    r.min_pressure, r.failure_threshold, r.pressure_at_failure = -1e8, 0.2, 1e-8
    r.surrogate_residual_tol = 1e-6
    return r


def make_config(nx, ny, nz, box, bc, mats, inlet=(1, 0, 0, 0, 1), outlet=(1, 0, 0, 0, 1), flux=FLUX_HLLC,
                limiter=LIM_VANALBADA, ts=TS_RK2, cfl=0.5, rec_linear=True, inactive_id=-1, roe_delta=0.2,
                gminmod_alpha=1.2, slope_zero=False, mp_flux=0, mp_recon=0, precision=PREC_FP64, device=0,
                ls_matid=None, ls_bc=(1, 1, 1, 1, 1, 1), reinit_freq=-1, reinit_maxits=30, reinit_cfl=0.8,
                reinit_tol=2.0e-4, mp_normal=1, conrec_depth=0.0, control_points=None):
    """Solver options; the level-set ones follow LevelSetSchemeData / LevelSetReinitializationData
    (IoData.cpp:1205-1266): ls_matid=None disables the level set. mp_normal: MultiPhase.RiemannNormal
    (1 Mesh, 0 LevelSet, 2 Average); conrec_depth: MultiPhase.ConstantReconstructionDepth.
    control_points: per axis None (uniform) or a list of (coordinate, cell width) pairs, Mesh.
    ControlPointX/Y/Z; the axis's cell count must be the one mesh_1d returns for them."""
    c = Config()
    c.nx, c.ny, c.nz = nx, ny, nz
    c.x0, c.x1, c.y0, c.y1, c.z0, c.z1 = box
    for i in range(6):
        c.bc[i] = bc[i]
    for i in range(5):
        c.inlet[i], c.outlet[i] = inlet[i], outlet[i]
    c.nmat = len(mats)
    for i, m in enumerate(mats):
        c.mat[i] = m
    c.inactive_id = inactive_id
    c.flux, c.roe_delta = flux, roe_delta
    c.rec_linear = int(rec_linear)
    c.limiter, c.gminmod_alpha = limiter, gminmod_alpha
    c.slope_zero_near_interface = int(slope_zero)
    c.ts, c.cfl = ts, cfl
    c.mp_flux, c.mp_recon = mp_flux, mp_recon
    c.riemann = riemann_defaults()
    c.precision, c.device = precision, device
    c.ls_matid = -1 if ls_matid is None else ls_matid
    for i in range(6):
        c.ls_bc[i] = ls_bc[i]
    c.reinit_freq, c.reinit_maxits = reinit_freq, reinit_maxits
    c.reinit_cfl, c.reinit_tol = reinit_cfl, reinit_tol
    c.mp_riemann_normal = mp_normal
    c.mp_conrec_depth = conrec_depth
    for d, pts in enumerate(control_points or (None, None, None)):
        pts = list(pts or [])
        if len(pts) > 16:
            raise ValueError("at most 16 control points per axis")
        c.mesh_npts[d] = len(pts)
        for q, (x, h) in enumerate(pts):
            c.mesh_pts[d][q][0], c.mesh_pts[d][q][1] = x, h
    return c


def mesh_1d(x0, xmax, n=0, points=None):
    """Cell centres and widths of one mesh axis (m2c_mesh_1d: MeshGenerator's uniform axis of n
    cells, or its stretched axis from (coordinate, cell width) control points)."""
    L = lib()
    pts = np.ascontiguousarray(np.array(points or [], dtype=np.float64).reshape(-1))
    m = C.c_int(0)
    _chk(L.m2c_mesh_1d(x0, xmax, n, len(pts) // 2, _dp(pts) if len(pts) else None, 0, None, None, C.byref(m)),
         "m2c_mesh_1d")
    x = np.empty(m.value)
    w = np.empty(m.value)
    _chk(L.m2c_mesh_1d(x0, xmax, n, len(pts) // 2, _dp(pts) if len(pts) else None, m.value, _dp(x), _dp(w),
                       C.byref(m)), "m2c_mesh_1d")
    return x, w


class Surface(C.Structure):
    """m2c_surface: one embedded surface (triangles, normals, node velocities) and its edge crossings."""
    _fields_ = [("nnodes", C.c_int), ("nelems", C.c_int), ("npoints", C.c_int),
                ("udot", C.POINTER(C.c_double)), ("elems", C.POINTER(C.c_int)), ("elem_norm", C.POINTER(C.c_double)),
                ("point_tid", C.POINTER(C.c_int)), ("point_xi", C.POINTER(C.c_double)),
                ("point_dist", C.POINTER(C.c_double))]


def _surface_structs(surfaces, struct):
    """ctypes structs for dicts with udot [n][3], elems [m][3], elem_norm [m][3], tid [p], xi [p][3],
    dist [p]; the arrays are kept alive in the returned list."""
    keep, out = [], (struct * max(len(surfaces), 1))()
    for i, sf in enumerate(surfaces):
        a = dict(udot=np.ascontiguousarray(sf["udot"], dtype=np.float64).reshape(-1, 3),
                 elems=np.ascontiguousarray(sf["elems"], dtype=np.int32).reshape(-1, 3),
                 elem_norm=np.ascontiguousarray(sf["elem_norm"], dtype=np.float64).reshape(-1, 3),
                 tid=np.ascontiguousarray(sf["tid"], dtype=np.int32).reshape(-1),
                 xi=np.ascontiguousarray(sf["xi"], dtype=np.float64).reshape(-1, 3),
                 dist=np.ascontiguousarray(sf["dist"], dtype=np.float64).reshape(-1))
        keep.append(a)
        D, I = C.POINTER(C.c_double), C.POINTER(C.c_int)
        out[i] = struct(len(a["udot"]), len(a["elems"]), len(a["tid"]), a["udot"].ctypes.data_as(D),
                        a["elems"].ctypes.data_as(I), a["elem_norm"].ctypes.data_as(D), a["tid"].ctypes.data_as(I),
                        a["xi"].ctypes.data_as(D), a["dist"].ctypes.data_as(D))
    return out, keep


def edge_surface_walls(surfaces, xf, xb):
    """FindEdgeSurfaceIntersections over faces (m2c_edge_surface_walls): surfaces as dicts (see
    _surface_structs), xf / xb [faces][surfaces] crossing indices (-1 none). Returns (cut [faces] bool,
    walls [faces][12]: nwallf, vwallf, nwallb, vwallb; rows of uncut faces are zero)."""
    xf = np.ascontiguousarray(xf, dtype=np.int32).reshape(-1, len(surfaces))
    xb = np.ascontiguousarray(xb, dtype=np.int32).reshape(-1, len(surfaces))
    n = len(xf)
    arr, keep = _surface_structs(surfaces, Surface)
    cut = np.zeros(n, np.int32)
    wall = np.zeros((n, 12))
    _chk(lib().m2c_edge_surface_walls(len(surfaces), arr, n, xf.ctypes.data_as(C.POINTER(C.c_int)),
                                      xb.ctypes.data_as(C.POINTER(C.c_int)), cut.ctypes.data_as(C.POINTER(C.c_int)),
                                      _dp(wall)), "m2c_edge_surface_walls")
    del keep
    return cut.astype(bool), wall


_lib = None


def lib():
    """Load libm2c_hip.so (fails loudly when the HIP library has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libm2c_hip.so is not built; run m2c_amd/build.py (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        D = C.POINTER(C.c_double)
        L.m2c_ctx_create.argtypes = [C.POINTER(Config), C.POINTER(P)]
        L.m2c_ctx_destroy.argtypes = [P]
        L.m2c_upload_state.argtypes = [P, D, C.POINTER(C.c_int32)]
        L.m2c_download_state.argtypes = [P, D, C.POINTER(C.c_int32)]
        L.m2c_compute_residual.argtypes = [P, D]
        L.m2c_time_step_size.argtypes = [P, D, D]
        L.m2c_advance_dt.argtypes = [P, C.c_double]
        L.m2c_run.argtypes = [P, C.c_double, C.c_int, C.c_int]
        L.m2c_get_status.argtypes = [P, C.POINTER(Status)]
        L.m2c_reset_time.argtypes = [P, C.c_double]
        L.m2c_synchronize.argtypes = [P]
        L.m2c_flux_batch.argtypes = [P, C.c_int, C.c_int, D, D, C.c_int, D]
        L.m2c_time_steps.argtypes = [P, C.c_int, D, D]
        L.m2c_device_state.argtypes = [P, C.POINTER(D)]
        L.m2c_set_kchunk.argtypes = [P, C.c_int]
        L.m2c_stream_copy.argtypes = [P, C.c_longlong, C.c_int, D]
        I = C.POINTER(C.c_int)
        L.m2c_riemann_batch.argtypes = [P, C.c_int, D, D, I, D, I, D, I, D, D, I]
        L.m2c_varfcn_batch.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_int32), D, D, C.POINTER(C.c_int32)]
        L.m2c_interface_flux_batch.argtypes = [P, C.c_int, C.c_int, D, C.POINTER(C.c_int32), D,
                                               C.POINTER(C.c_int32), D]
        L.m2c_onesided_riemann_batch.argtypes = [P, C.c_int, D, D, I, D, D, I, D, I]
        L.m2c_wall_face_flux_batch.argtypes = [P, C.c_int, I, I, C.c_int, D, D, D, I, D, D, I]
        L.m2c_set_embedded_faces.argtypes = [P, C.c_int, I, D, C.c_int, C.c_int]
        L.m2c_upload_levelset.argtypes = [P, D]
        L.m2c_download_levelset.argtypes = [P, D]
        Pb = C.c_void_p
        L.m2c_batch_create.argtypes = [C.POINTER(Config), C.c_int, C.POINTER(Pb)]
        L.m2c_batch_destroy.argtypes = [Pb]
        L.m2c_batch_upload.argtypes = [Pb, C.c_int, D, C.POINTER(C.c_int32)]
        L.m2c_batch_download.argtypes = [Pb, C.c_int, D, C.POINTER(C.c_int32)]
        L.m2c_batch_upload_levelset.argtypes = [Pb, C.c_int, D]
        L.m2c_mesh_1d.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, D, C.c_int, D, D, C.POINTER(C.c_int)]
        L.m2c_edge_surface_walls.argtypes = [C.c_int, C.POINTER(Surface), C.c_int, C.POINTER(C.c_int),
                                             C.POINTER(C.c_int), C.POINTER(C.c_int), D]
        L.m2c_batch_download_levelset.argtypes = [Pb, C.c_int, D]
        L.m2c_batch_run.argtypes = [Pb, C.c_int, C.c_int, D]
        L.m2c_batch_status.argtypes = [Pb, C.c_int, C.POINTER(Status)]
        L.m2c_batch_autotune.argtypes = [Pb, C.c_int, C.POINTER(BatchTuning)]
        L.m2c_batch_set_tuning.argtypes = [Pb, C.c_int, C.c_int]
        L.m2c_batch_set_tile.argtypes = [Pb, C.c_int, C.c_int]
        F = C.POINTER(C.c_float)
        L.m2c_mlp_create.argtypes = [C.c_int, C.c_ulonglong, C.POINTER(P)]
        L.m2c_mlp_destroy.argtypes = [P]
        L.m2c_mlp_get_params.argtypes = [P, F]
        L.m2c_mlp_set_params.argtypes = [P, F]
        L.m2c_mlp_forward.argtypes = [P, C.c_int, F, F]
        L.m2c_mlp_gradient.argtypes = [P, C.c_int, F, F, D, F]
        L.m2c_mlp_samples.argtypes = [P, C.c_longlong, C.c_ulonglong, F, F, D]
        L.m2c_mlp_train.argtypes = [P, C.POINTER(TrainConfig), C.POINTER(TrainReport)]
        L.m2c_mlp_bench_infer.argtypes = [P, C.c_int, C.c_int, D]
        L.m2c_mlp_bench_train.argtypes = [P, C.c_int, C.c_int, D]
        L.m2c_set_surrogate.argtypes = [P, P, C.c_int]
        L.m2c_surrogate_riemann_batch.argtypes = [P, P, C.c_int, C.c_int, I, D, I, D, I, D, I, D, D, I, I]
        L.m2c_version.restype = C.c_char_p
        L.m2c_config_defaults.argtypes = [C.POINTER(Config)]
        L.m2c_config_from_deck.argtypes = [C.c_char_p, C.POINTER(Config), C.POINTER(DeckRun), C.c_char_p, C.c_int]
        L.m2c_initial_condition_from_deck.argtypes = [C.c_char_p, D, C.POINTER(C.c_int32), D, C.c_char_p, C.c_int]
        L.m2c_eos_derived.argtypes = [P, C.c_int, D, D, C.POINTER(C.c_int32), D, D]
        L.m2c_time_diffusion_pass.argtypes = [P, D]
        L.m2c_output_create.argtypes = [C.c_char_p, C.POINTER(P), C.c_char_p, C.c_int]
        L.m2c_output_write.argtypes = [P, P, C.c_double, C.c_double, C.c_int, D, C.POINTER(C.c_int32), D, C.c_int,
                                       C.POINTER(C.c_int)]
        L.m2c_output_needs_device.argtypes = [P]
        L.m2c_output_destroy.argtypes = [P]
        _lib = L
    return _lib


class M2CError(RuntimeError):
    """A library call that returned an error status: `code` is ERR_ARG, ERR_HIP, ERR_STATE or
    ERR_UNSUPPORTED."""

    def __init__(self, message, code=None):
        super().__init__(message)
        self.code = code


def _chk(rc, what):
    if rc != OK:
        raise M2CError(f"{what} failed with status {rc}", rc)


def axis_normals(dirs, n):
    """(n, 3) unit normals from mesh axes (n ints) or from normals already given as (n, 3)."""
    a = np.asarray(dirs)
    if a.ndim >= 1 and a.shape[-1] == 3 and a.dtype.kind == "f":
        return np.ascontiguousarray(np.broadcast_to(a.astype(np.float64), (n, 3)))
    ax = np.broadcast_to(a.astype(np.int64), (n,))
    if ((ax < 0) | (ax > 2)).any():
        raise ValueError("mesh axis outside 0..2")
    return np.ascontiguousarray(np.eye(3)[ax])


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """One simulation context on one GPU (SpaceOperator + TimeIntegrator drop-in)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.L = lib()
        self.h = C.c_void_p()
        _chk(self.L.m2c_ctx_create(C.byref(cfg), C.byref(self.h)), "m2c_ctx_create")
        self.shape = (cfg.nz, cfg.ny, cfg.nx)

    def close(self):
        if self.h:
            self.L.m2c_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, V, ID=None):
        V = np.ascontiguousarray(V, dtype=np.float64)
        assert V.shape == self.shape + (5,)
        idp = None
        if ID is not None:
            ID = np.ascontiguousarray(ID, dtype=np.int32)
            assert ID.shape == self.shape
            idp = ID.ctypes.data_as(C.POINTER(C.c_int32))
        _chk(self.L.m2c_upload_state(self.h, _dp(V), idp), "m2c_upload_state")

    def download(self):
        V = np.empty(self.shape + (5,), dtype=np.float64)
        ID = np.empty(self.shape, dtype=np.int32)
        _chk(self.L.m2c_download_state(self.h, _dp(V), ID.ctypes.data_as(C.POINTER(C.c_int32))),
             "m2c_download_state")
        return V, ID

    def upload_levelset(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        assert phi.shape == self.shape
        _chk(self.L.m2c_upload_levelset(self.h, _dp(phi)), "m2c_upload_levelset")

    def download_levelset(self):
        phi = np.empty(self.shape, dtype=np.float64)
        _chk(self.L.m2c_download_levelset(self.h, _dp(phi)), "m2c_download_levelset")
        return phi

    def riemann(self, dirs, Vm, idm, Vp, idp):
        """Exact two-material Riemann solutions of n problems: (status, Vs, id, Vsm, Vsp). dirs: unit
        normals (n x 3), or mesh axes (n ints: 0 x, 1 y, 2 z) turned into unit normals here."""
        Vm = np.ascontiguousarray(np.atleast_2d(Vm), dtype=np.float64)
        Vp = np.ascontiguousarray(np.atleast_2d(Vp), dtype=np.float64)
        n = Vm.shape[0]
        ip = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.int32), (n,)))
        dd = axis_normals(dirs, n)
        im, iq = ip(idm), ip(idp)
        Vs, Vsm, Vsp = np.empty_like(Vm), np.empty_like(Vm), np.empty_like(Vm)
        mid, st = np.empty(n, np.int32), np.empty(n, np.int32)
        cp = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        _chk(self.L.m2c_riemann_batch(self.h, n, _dp(dd), _dp(Vm), cp(im), _dp(Vp), cp(iq), _dp(Vs), cp(mid),
                                      _dp(Vsm), _dp(Vsp), cp(st)), "m2c_riemann_batch")
        return st, Vs, mid, Vsm, Vsp

    def varfcn(self, op, ids, x):
        """VarFcnBase pointwise operation `op` (VF_*) of materials ids on points x ((n, 2) for the
        (rho, e) / (rho, p) operations, (n, 5) for state vectors): (out, flag)."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
        n = x.shape[0]
        ids = np.ascontiguousarray(np.broadcast_to(np.asarray(ids, dtype=np.int32), (n,)))
        out = np.empty((n, 5) if op in (VF_C2P, VF_P2C, VF_CLIP) else n, dtype=np.float64)
        flag = np.empty(n, np.int32)
        ci = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        _chk(self.L.m2c_varfcn_batch(self.h, op, n, ci(ids), _dp(x), _dp(out), ci(flag)), "m2c_varfcn_batch")
        return out, flag

    def interface_flux(self, d, Vm, idm, Vp, idp):
        """Local Lax-Friedrichs flux between two materials along axis d for n face pairs (n x 5)."""
        Vm = np.ascontiguousarray(np.atleast_2d(Vm), dtype=np.float64)
        Vp = np.ascontiguousarray(np.atleast_2d(Vp), dtype=np.float64)
        n = Vm.shape[0]
        ip = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.int32), (n,)))
        im, iq = ip(idm), ip(idp)
        F = np.empty_like(Vm)
        ci = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        _chk(self.L.m2c_interface_flux_batch(self.h, d, n, _dp(Vm), ci(im), _dp(Vp), ci(iq), _dp(F)),
             "m2c_interface_flux_batch")
        return F

    def onesided_riemann(self, dirs, Vm, idl, Ustar):
        """One-sided (wall) Riemann solutions of n problems along unit normals dirs (n x 3) against walls
        moving with velocities Ustar (n x 3): (status, Vs, id, Vsm); id = -1 and Vs = 0 where the wall
        recedes."""
        Vm = np.ascontiguousarray(np.atleast_2d(Vm), dtype=np.float64)
        n = Vm.shape[0]
        d3 = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n, 3)))
        ip = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.int32), (n,)))
        dd, us, im = d3(dirs), d3(Ustar), ip(idl)
        Vs, Vsm = np.empty_like(Vm), np.empty_like(Vm)
        mid, st = np.empty(n, np.int32), np.empty(n, np.int32)
        cp = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        _chk(self.L.m2c_onesided_riemann_batch(self.h, n, _dp(dd), _dp(Vm), cp(im), _dp(us), _dp(Vs), cp(mid),
                                               _dp(Vsm), cp(st)), "m2c_onesided_riemann_batch")
        return st, Vs, mid, Vsm

    def set_embedded_faces(self, faces, walls, normal_mode=1, recon=0):
        """Mesh faces cut by embedded walls, fluxed in every later stage: faces (n, 4) = axis and the
        plus cell's i, j, k; walls (n, 12) = nwallf, vwallf, nwallb, vwallb. normal_mode 0 wall normal,
        1 mesh axis, 2 their average; recon 0 constant, 1 linear. An empty list removes the walls."""
        f = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).reshape(-1, 4))
        w = np.ascontiguousarray(np.asarray(walls, dtype=np.float64).reshape(-1, 12))
        if f.shape[0] != w.shape[0]:
            raise ValueError("faces and walls differ in length")
        _chk(self.L.m2c_set_embedded_faces(self.h, f.shape[0], f.ctypes.data_as(C.POINTER(C.c_int)), _dp(w),
                                           normal_mode, recon), "m2c_set_embedded_faces")

    def wall_face_flux(self, axis, side, normal_mode, nwall, vwall, V, ids):
        """Fluxes on the fluid side of n wall-cut faces: (status, F, Vstar). side +1: the fluid cell is
        on the minus side of the face; normal_mode 0 wall normal, 1 mesh axis, 2 their average."""
        V = np.ascontiguousarray(np.atleast_2d(V), dtype=np.float64)
        n = V.shape[0]
        d3 = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n, 3)))
        ip = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.int32), (n,)))
        ax, sd, im, nw, vw = ip(axis), ip(side), ip(ids), d3(nwall), d3(vwall)
        F, Vst = np.empty_like(V), np.empty_like(V)
        st = np.empty(n, np.int32)
        cp = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        _chk(self.L.m2c_wall_face_flux_batch(self.h, n, cp(ax), cp(sd), normal_mode, _dp(nw), _dp(vw), _dp(V),
                                             cp(im), _dp(F), _dp(Vst), cp(st)), "m2c_wall_face_flux_batch")
        return st, F, Vst

    def set_surrogate(self, net, newton_max=3):
        """Interface faces use the learned solver `net` (an Mlp; None = exact solver only)."""
        self._net = net   # kept alive while attached
        _chk(self.L.m2c_set_surrogate(self.h, net.h if net is not None else None, newton_max), "m2c_set_surrogate")

    def surrogate_riemann(self, net, dirs, Vm, idm, Vp, idp, newton_max=3):
        """riemann() through the learned solver with the exact fallback: (status, Vs, id, Vsm, Vsp, path)."""
        Vm = np.ascontiguousarray(np.atleast_2d(Vm), dtype=np.float64)
        Vp = np.ascontiguousarray(np.atleast_2d(Vp), dtype=np.float64)
        n = Vm.shape[0]
        ip = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.int32), (n,)))
        d, im, iq = ip(dirs), ip(idm), ip(idp)
        Vs, Vsm, Vsp = np.empty_like(Vm), np.empty_like(Vm), np.empty_like(Vm)
        mid, st, path = np.empty(n, np.int32), np.empty(n, np.int32), np.empty(n, np.int32)
        cp = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        _chk(self.L.m2c_surrogate_riemann_batch(self.h, net.h, newton_max, n, cp(d), _dp(Vm), cp(im), _dp(Vp),
                                                cp(iq), _dp(Vs), cp(mid), _dp(Vsm), _dp(Vsp), cp(st), cp(path)),
             "m2c_surrogate_riemann_batch")
        return st, Vs, mid, Vsm, Vsp, path

    def residual(self):
        R = np.empty(self.shape + (5,), dtype=np.float64)
        _chk(self.L.m2c_compute_residual(self.h, _dp(R)), "m2c_compute_residual")
        return R

    def time_step(self):
        dt, cfl = C.c_double(), C.c_double()
        _chk(self.L.m2c_time_step_size(self.h, C.byref(dt), C.byref(cfl)), "m2c_time_step_size")
        return dt.value

    def advance(self, dt):
        _chk(self.L.m2c_advance_dt(self.h, dt), "m2c_advance_dt")

    def run(self, tmax, maxsteps, sync_every=64):
        _chk(self.L.m2c_run(self.h, tmax, maxsteps, sync_every), "m2c_run")
        return self.status()

    def reset_time(self, t=0.0):
        _chk(self.L.m2c_reset_time(self.h, t), "m2c_reset_time")

    def status(self):
        s = Status()
        _chk(self.L.m2c_get_status(self.h, C.byref(s)), "m2c_get_status")
        return s

    def flux(self, d, Vm, Vp, mid=0):
        Vm = np.ascontiguousarray(np.atleast_2d(Vm), dtype=np.float64)
        Vp = np.ascontiguousarray(np.atleast_2d(Vp), dtype=np.float64)
        F = np.empty_like(Vm)
        _chk(self.L.m2c_flux_batch(self.h, d, Vm.shape[0], _dp(Vm), _dp(Vp), mid, _dp(F)), "m2c_flux_batch")
        return F

    def eos_derived(self, rho, p, ids):
        """Internal energy per unit mass and temperature at n points (device EOS): (e, T)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64).ravel()
        p = np.ascontiguousarray(p, dtype=np.float64).ravel()
        ids = np.asarray(ids, dtype=np.int32)
        ids = np.ascontiguousarray(ids.ravel() if ids.size == rho.size else np.broadcast_to(ids, rho.shape))
        e, T = np.empty_like(rho), np.empty_like(rho)
        _chk(self.L.m2c_eos_derived(self.h, rho.size, _dp(rho), _dp(p), ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                    _dp(e), _dp(T)), "m2c_eos_derived")
        return e, T

    def set_kchunk(self, kchunk):
        """z planes per workgroup of the fused stage kernel (0: the default launch shape)."""
        _chk(self.L.m2c_set_kchunk(self.h, kchunk), "m2c_set_kchunk")

    def time_steps(self, nsteps):
        a, b = C.c_double(), C.c_double()
        _chk(self.L.m2c_time_steps(self.h, nsteps, C.byref(a), C.byref(b)), "m2c_time_steps")
        return a.value, b.value

    def diffusion_pass_ms(self):
        """Average face pass before each stage (viscous / heat terms, Godunov fluxes; ms) of the last
        time_steps() call."""
        ms = C.c_double()
        _chk(self.L.m2c_time_diffusion_pass(self.h, C.byref(ms)), "m2c_time_diffusion_pass")
        return ms.value


def stream_copy_GBs(solver, nbytes=1 << 30, reps=20):
    """Measured copy bandwidth (read + write bytes per second, GB/s) of a plain copy kernel."""
    ms = C.c_double()
    _chk(solver.L.m2c_stream_copy(solver.h, nbytes, reps, C.byref(ms)), "m2c_stream_copy")
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9, ms.value


def version():
    return lib().m2c_version().decode()


def config_defaults():
    """IoData's defaults for every field of Config (no materials, no cells, boundaries None)."""
    c = Config()
    _chk(lib().m2c_config_defaults(C.byref(c)), "m2c_config_defaults")
    return c


class DeckError(M2CError):
    """An input deck the reader rejects: `code` is ERR_ARG (the reference rejects it too) or
    ERR_UNSUPPORTED (valid for the reference, not provided here)."""

    def __init__(self, code, message):
        super().__init__(message, code)


def config_from_deck(text, msglen=1 << 16):
    """Read an M2C input deck (its text) into (Config, DeckRun, ignored) where `ignored` lists the
    options outside the finite-volume update that the reader skipped. Raises DeckError."""
    cfg, run = Config(), DeckRun()
    buf = C.create_string_buffer(msglen)
    rc = lib().m2c_config_from_deck(text.encode(), C.byref(cfg), C.byref(run), buf, msglen)
    msg = buf.value.decode()
    if rc != OK:
        raise DeckError(rc, msg)
    return cfg, run, [line for line in msg.splitlines() if line]


def initial_condition_from_deck(text, msglen=1 << 16):
    """Initial state of a deck on its mesh: (V (nz, ny, nx, 5), ID (nz, ny, nx) int32, phi) where phi
    is the level set of the deck's level-set material, or None when it has none. Raises DeckError."""
    cfg, _, _ = config_from_deck(text, msglen)
    nz, ny, nx = cfg.nz, cfg.ny, cfg.nx
    V = np.zeros((nz, ny, nx, 5))
    ID = np.zeros((nz, ny, nx), dtype=np.int32)
    phi = np.zeros((nz, ny, nx)) if cfg.ls_matid >= 0 else None
    buf = C.create_string_buffer(msglen)
    rc = lib().m2c_initial_condition_from_deck(text.encode(), _dp(V), ID.ctypes.data_as(C.POINTER(C.c_int32)),
                                               None if phi is None else _dp(phi), buf, msglen)
    if rc != OK:
        raise DeckError(rc, buf.value.decode())
    return V, ID, phi


class Output:
    """The Output object of a deck (snapshots, probes, line plots): Output::OutputSolutions."""

    def __init__(self, text, msglen=1 << 16):
        self.L = lib()
        self.h = C.c_void_p()
        buf = C.create_string_buffer(msglen)
        rc = self.L.m2c_output_create(text.encode(), C.byref(self.h), buf, msglen)
        msg = buf.value.decode()
        if rc != OK:
            raise DeckError(rc, msg)
        self.not_written = [line for line in msg.splitlines() if line]

    @property
    def needs_device(self):
        return bool(self.L.m2c_output_needs_device(self.h))

    def write(self, time, dt, step, V, ID, phi=None, solver=None, force=False):
        """Returns the bit set of what was written (1 snapshot, 2 probes, 4 line plots)."""
        V = np.ascontiguousarray(V, dtype=np.float64)
        ID = np.ascontiguousarray(ID, dtype=np.int32)
        ph = None if phi is None else _dp(np.ascontiguousarray(phi, dtype=np.float64))
        w = C.c_int()
        _chk(self.L.m2c_output_write(self.h, solver.h if solver is not None else None, time, dt, step, _dp(V),
                                     ID.ctypes.data_as(C.POINTER(C.c_int32)), ph, int(force), C.byref(w)),
             "m2c_output_write")
        return w.value

    def close(self):
        if self.h:
            self.L.m2c_output_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_deck(path):
    """An input deck file: (Config, DeckRun, ignored options, V, ID, phi)."""
    with open(path) as f:
        text = f.read()
    cfg, run, ignored = config_from_deck(text)
    V, ID, phi = initial_condition_from_deck(text)
    return cfg, run, ignored, V, ID, phi


class Batch:
    """Many independent single- or two-material instances advanced together (one persistent launch
    per RK stage, a hipGraph per step)."""

    def __init__(self, cfgs):
        self.L = lib()
        arr = (Config * len(cfgs))(*cfgs)
        self.h = C.c_void_p()
        _chk(self.L.m2c_batch_create(arr, len(cfgs), C.byref(self.h)), "m2c_batch_create")
        self.shapes = [(c.nz, c.ny, c.nx) for c in cfgs]

    def close(self):
        if self.h:
            self.L.m2c_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, i, V, ID=None):
        V = np.ascontiguousarray(V, dtype=np.float64)
        assert V.shape == self.shapes[i] + (5,)
        idp = None
        if ID is not None:
            ID = np.ascontiguousarray(ID, dtype=np.int32)
            idp = ID.ctypes.data_as(C.POINTER(C.c_int32))
        _chk(self.L.m2c_batch_upload(self.h, i, _dp(V), idp), "m2c_batch_upload")

    def download(self, i):
        V = np.empty(self.shapes[i] + (5,), dtype=np.float64)
        ID = np.empty(self.shapes[i], dtype=np.int32)
        _chk(self.L.m2c_batch_download(self.h, i, _dp(V), ID.ctypes.data_as(C.POINTER(C.c_int32))),
             "m2c_batch_download")
        return V, ID

    def upload_levelset(self, i, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        assert phi.shape == self.shapes[i]
        _chk(self.L.m2c_batch_upload_levelset(self.h, i, _dp(phi)), "m2c_batch_upload_levelset")

    def download_levelset(self, i):
        phi = np.empty(self.shapes[i], dtype=np.float64)
        _chk(self.L.m2c_batch_download_levelset(self.h, i, _dp(phi)), "m2c_batch_download_levelset")
        return phi

    def run(self, nsteps, graph=True):
        ms = C.c_double()
        _chk(self.L.m2c_batch_run(self.h, nsteps, int(graph), C.byref(ms)), "m2c_batch_run")
        return ms.value

    def status(self, i):
        s = Status()
        _chk(self.L.m2c_batch_status(self.h, i, C.byref(s)), "m2c_batch_status")
        return s

    def autotune(self, trial_steps=3):
        t = BatchTuning()
        _chk(self.L.m2c_batch_autotune(self.h, trial_steps, C.byref(t)), "m2c_batch_autotune")
        return t

    def set_tuning(self, kchunk, blocks):
        _chk(self.L.m2c_batch_set_tuning(self.h, kchunk, blocks), "m2c_batch_set_tuning")

    def set_tile(self, tile_x, tile_y):
        _chk(self.L.m2c_batch_set_tile(self.h, tile_x, tile_y), "m2c_batch_set_tile")


NPARAM = 200963


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The learned interface solver's network (10-256x4-3 SiLU) on one GPU."""

    def __init__(self, seed=1, device=0):
        self.L = lib()
        self.h = C.c_void_p()
        _chk(self.L.m2c_mlp_create(device, seed, C.byref(self.h)), "m2c_mlp_create")

    def close(self):
        if self.h:
            self.L.m2c_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(NPARAM, np.float32)
        _chk(self.L.m2c_mlp_get_params(self.h, _fp(p)), "m2c_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (NPARAM,)
        _chk(self.L.m2c_mlp_set_params(self.h, _fp(p)), "m2c_mlp_set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((X.shape[0], 3), np.float32)
        _chk(self.L.m2c_mlp_forward(self.h, X.shape[0], _fp(X), _fp(Y)), "m2c_mlp_forward")
        return Y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(NPARAM, np.float32)
        loss = C.c_double()
        _chk(self.L.m2c_mlp_gradient(self.h, X.shape[0], _fp(X), _fp(T), C.byref(loss), _fp(g)), "m2c_mlp_gradient")
        return loss.value, g

    def samples(self, n, seed, raw=True):
        X = np.empty((n, 10), np.float32)
        T = np.empty((n, 3), np.float32)
        R = np.empty((n, 14), np.float64) if raw else None
        _chk(self.L.m2c_mlp_samples(self.h, n, seed, _fp(X), _fp(T), _dp(R) if raw else None), "m2c_mlp_samples")
        return X, T, R

    def train(self, n_samples=1 << 20, n_holdout=1 << 16, seed=7, batch=8192, epochs=4, lr=2e-3, lr_min=1e-5,
              newton_max=2, tol=1e-6, verbose=False):
        c = TrainConfig()
        c.n_samples, c.n_holdout, c.seed, c.batch, c.epochs = n_samples, n_holdout, seed, batch, epochs
        c.lr, c.lr_min, c.newton_max, c.tol, c.verbose = lr, lr_min, newton_max, tol, int(verbose)
        r = TrainReport()
        _chk(self.L.m2c_mlp_train(self.h, C.byref(c), C.byref(r)), "m2c_mlp_train")
        return r

    def bench_train(self, batch, steps=50):
        ms = C.c_double()
        _chk(self.L.m2c_mlp_bench_train(self.h, batch, steps, C.byref(ms)), "m2c_mlp_bench_train")
        return ms.value

    def bench_infer(self, n, reps=20):
        ms = C.c_double()
        _chk(self.L.m2c_mlp_bench_infer(self.h, n, reps, C.byref(ms)), "m2c_mlp_bench_infer")
        return ms.value


SURROGATE_PARAMS = os.path.join(HERE, "data", "sg_riemann_mlp.npy")


def load_surrogate(path=SURROGATE_PARAMS, device=0):
    """The learned interface solver with the parameters written by m2c_amd/train_surrogate.py."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: run m2c_amd/train_surrogate.py on the GPU")
    net = Mlp(seed=0, device=device)
    net.set_params(np.load(path, allow_pickle=False))
    return net
