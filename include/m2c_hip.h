/*
 * m2c_hip.h -- C ABI of the MI355X-native explicit finite-volume update.
 *
 * Plain C linkage, plain pointers and sizes. Every call returns an m2c_status code (0 = success)
 * instead of calling exit() as the reference does; fatal numerical conditions of the reference
 * are reported through m2c_get_status().
 *
 * Host arrays crossing this ABI use the reference's natural layout without ghosts:
 * k-j-i order, x fastest, 5 primitive variables (rho, u, v, w, p) innermost for V
 * (the interior of SpaceVariable3D, SpaceVariable.cpp:254-260). Material IDs are int32.
 * On the device the context stores structure-of-arrays fields, x fastest, no ghost frame:
 * ghost values are produced inside the kernels from the boundary rules.
 */
#ifndef M2C_HIP_H
#define M2C_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define M2C_MAX_MAT 8

enum m2c_status_code {
  M2C_OK = 0,
  M2C_ERR_ARG = 1,        /* invalid argument or configuration */
  M2C_ERR_HIP = 2,        /* HIP runtime error */
  M2C_ERR_STATE = 3,      /* non-hyperbolic state (the reference exits: SpaceOperator.cpp:624-631) */
  M2C_ERR_UNSUPPORTED = 4 /* option recognised but not provided by this library */
};

/* Equation-of-state kinds, VarFcnBase.h:40-41 */
enum m2c_eos_kind { M2C_EOS_SG = 0, M2C_EOS_NASG = 1, M2C_EOS_MG = 2, M2C_EOS_JWL = 3, M2C_EOS_DUMMY = 5 };
/* Boundary codes, IoData.h:266-267, plus PERIODIC (no reference counterpart) */
enum m2c_bc { M2C_BC_INLET = 1, M2C_BC_OUTLET = 2, M2C_BC_SLIPWALL = 3, M2C_BC_STICKWALL = 4,
              M2C_BC_SYMMETRY = 5, M2C_BC_OVERSET = 6, M2C_BC_PERIODIC = 7 };
/* NavierStokes flux, SchemeData::flux (IoData.cpp:1172): ROE=0, LLF=1, HLLC=2, GODUNOV=3. GODUNOV
 * (FluxFcnGodunov.h) solves the exact Riemann problem of every face; FP64 contexts only. */
enum m2c_flux { M2C_FLUX_ROE = 0, M2C_FLUX_LLF = 1, M2C_FLUX_HLLC = 2, M2C_FLUX_GODUNOV = 3 };
enum m2c_limiter { M2C_LIM_NONE = 0, M2C_LIM_GENMINMOD = 1, M2C_LIM_VANALBADA = 2 };
enum m2c_ts { M2C_TS_FE = 0, M2C_TS_RK2 = 1, M2C_TS_RK3 = 2 };
/* FP64: the drop-in path, bit-identical to the reference's arithmetic. MIXED: reconstruction,
 * limiters, face-state checks and fluxes in FP32; fetch with ConservativeToPrimitive (p(e)
 * inversion), residual accumulation, RK stage combination, clipping, CFL and the exact interface
 * Riemann solves in FP64. */
enum m2c_precision { M2C_PREC_FP64 = 0, M2C_PREC_MIXED = 1 };

typedef struct {
  int eos;                          /* m2c_eos_kind */
  double rhomin, pmin, rhomax, pmax, failsafe_density; /* MaterialModelData, IoData.cpp:860-870 */
  double p[8];                      /* SG: gamma, p_inf. NASG: gamma, pc, b, q.
                                       MG: rho0, c0, Gamma0, s, e0. JWL: omega, A1, A2, R1, R2, rho0.
                                       DUMMY: p0, rho0, e0. */
  double tlaw[6];                   /* temperature law (output only): cv, T0, e0, rho0, cp, h0.
                                       SG: all six (VarFcnSG.h:126-142); NASG: cv; MG: cv, T0, cp, h0
                                       (its e0 is p[4]); JWL and DUMMY have none. */
  int viscosity;                    /* ViscosityModel.Type (ViscoFcn.h): 0 none, 1 constant, 2 Sutherland */
  double mu;                        /* DynamicViscosity (Sutherland: ReferenceDynamicViscosity) */
  double bulk;                      /* BulkViscosity */
  double su_T0, su_S;               /* Sutherland: ReferenceTemperature, SutherlandConstant */
  double kappa;                     /* HeatDiffusionModel (Type Constant): Diffusivity; 0 = none */
} m2c_material;

typedef struct {
  int maxIts_main, maxIts_bracket, maxIts_shock, numSteps_rarefaction;
  double tol_main, tol_shock, tol_rarefaction, min_pressure, failure_threshold, pressure_at_failure;
  double surrogate_residual_tol;
} m2c_riemann_params;                /* ExactRiemannSolverData, IoData.cpp:1345-1357 */

#define M2C_MAX_MESH_PTS 16

typedef struct {
  int nx, ny, nz;                   /* cells along each axis (with control points: as m2c_mesh_1d) */
  double x0, x1, y0, y1, z0, z1;
  int bc[6];                        /* x0, xmax, y0, ymax, z0, zmax */
  double inlet[5], outlet[5];       /* BoundaryConditions.Inlet / Outlet (primitive) */
  int nmat;
  int inactive_id;                  /* -1 when there is no inactive (dummy) material */
  m2c_material mat[M2C_MAX_MAT];
  int flux;                         /* m2c_flux */
  double roe_delta;                 /* Harten entropy-fix coefficient */
  int rec_linear;                   /* 0: constant, 1: linear reconstruction */
  int limiter;                      /* m2c_limiter */
  double gminmod_alpha;
  int slope_zero_near_interface;
  int ts;                           /* m2c_ts */
  double cfl;
  int mp_flux, mp_recon;            /* MultiPhase.Flux (0 numerical, 1 exact, 2 LLF), Recon (0/1) */
  m2c_riemann_params riemann;
  int precision;                    /* m2c_precision */
  int device;                       /* HIP device ordinal */
  /* level set (LevelSetSchemeData / LevelSetReinitializationData, IoData.cpp:1205-1266) */
  int ls_matid;                     /* material tracked by the level set (phi < 0), -1 = none */
  int ls_bc[6];                     /* 1 zero-Neumann, 2 linear extrapolation, 3 non-negative,
                                       7 periodic (x0, xmax, y0, ymax, z0, zmax) */
  int reinit_freq;                  /* reinitialise every reinit_freq steps (<= 0: never) */
  int reinit_maxits;
  double reinit_cfl, reinit_tol;
  /* MultiPhase.RiemannNormal (MultiPhaseData, IoData.cpp:1445-1448): direction of the exact Riemann
   * problem at a material interface. 1 Mesh (the face's axis, the default); 0 LevelSet (the
   * level-set gradient at the face, SpaceOperator::CalculateGradPhiAtCellInterface); 2 Average
   * (their normalised sum). LevelSet and Average need the level set and interfaces only between
   * material 0 and the level set's material. */
  int mp_riemann_normal;
  /* MultiPhase.ConstantReconstructionDepth (MultiPhaseData::conRec_depth): cells with
   * |phi(t_n)| below it use constant reconstruction (SpaceOperator::TagNodesOutsideConRecDepth);
   * 0 = none (the default) */
  double mp_conrec_depth;
  /* Mesh.ControlPointX/Y/Z[i] (MeshData, IoData.cpp:486-495; MeshResolution1DPointData: Coordinate,
   * CellWidth): mesh_npts[d] > 0 stretches axis d as MeshGenerator::ComputeMesh1DNonUniform does
   * (MeshGenerator.cpp:163-259), and n[d] must equal the number of cells that generates
   * (m2c_mesh_1d); 0 = the uniform axis of n[d] cells (ComputeMesh1DUniform). Stretched meshes run
   * with the reference's per-cell limiter coefficients A, B, K (Reconstructor.cpp:53-121), face
   * areas and volumes (SpaceOperator.cpp:160-177), CFL widths, level-set coordinates and
   * reinitialisation steps; not with periodic axes, viscous or heat-diffusion terms, embedded walls
   * or batches (M2C_ERR_UNSUPPORTED). */
  int mesh_npts[3];
  double mesh_pts[3][M2C_MAX_MESH_PTS][2];
} m2c_config;

/* The cells of one mesh axis (MeshGenerator.cpp:99-259, the reference's steps and rounding):
 * npts = 0: n uniform cells of [x0, xmax] (ComputeMesh1DUniform); npts > 0: the cells
 * ComputeMesh1DNonUniform generates from the control points pts (npts pairs of coordinate and cell
 * width; n unused). Writes up to cap cell centres x and widths dx (either may be NULL) and the
 * number of cells to *nout; M2C_ERR_ARG where the reference stops with an error (control points
 * outside the axis, widths in conflict with the points) and for widths that are not positive and
 * finite or an axis of 1e9 cells or more (where the reference divides by zero or overflows). */
int m2c_mesh_1d(double x0, double xmax, int n, int npts, const double *pts, int cap, double *x, double *dx,
                int *nout);

typedef struct {
  int error;                        /* sticky m2c_status_code of device-detected conditions */
  long long reverted_faces;         /* reconstructed face states reverted to the cell value */
  long long clipped_cells;
  long long riemann_errors;
  long long steps;
  double t, dt, cfl;
  long long reinit_iterations;      /* pseudo-time iterations of all reinitialisations */
  long long unresolved_cells;       /* cells updated by FixUnresolvedNodes */
  long long surrogate_accepted;     /* interface solves taken from the learned solver */
  long long surrogate_fallbacks;    /* ... and those that failed its check and ran the exact solver */
} m2c_status;

typedef struct m2c_ctx m2c_ctx;

/* Context life cycle. Replaces the SpaceOperator/Reconstructor/VarFcn/FluxFcn factories of
 * Main.cpp:116-182 and the DataManagers3D allocations (SpaceVariable.cpp:49-156). */
int m2c_ctx_create(const m2c_config *cfg, m2c_ctx **out);
int m2c_ctx_destroy(m2c_ctx *ctx);

/* State transfer (V: nz*ny*nx*5 doubles, id: nz*ny*nx int32 or NULL for all-zero). */
int m2c_upload_state(m2c_ctx *ctx, const double *V, const int32_t *id);
int m2c_download_state(m2c_ctx *ctx, double *V, int32_t *id);

/* Level set of the material ls_matid (nz*ny*nx doubles, no ghosts): LevelSetOperator's Phi
 * (LevelSetOperator.h:86-90). Upload before the first step when ls_matid >= 0. */
int m2c_upload_levelset(m2c_ctx *ctx, const double *phi);
int m2c_download_levelset(m2c_ctx *ctx, double *phi);

/* SpaceOperator::ComputeResidual (SpaceOperator.h:136-140) on the current state;
 * R: nz*ny*nx*5 doubles (already divided by -volume, SpaceOperator.cpp:3394-3398). */
int m2c_compute_residual(m2c_ctx *ctx, double *R);

/* SpaceOperator::ComputeTimeStepSize (SpaceOperator.h:129-130). */
int m2c_time_step_size(m2c_ctx *ctx, double *dt, double *cfl);

/* TimeIntegratorBase::AdvanceOneTimeStep (TimeIntegrator.h:66-73) with the given dt. */
int m2c_advance_dt(m2c_ctx *ctx, double dt);

/* Main.cpp:514-621: repeat {dt = CFL dt; clamp to tmax; advance} while t < tmax and
 * steps < maxsteps. Time and dt stay on the device; the host synchronises every
 * `sync_every` steps (0 = only at the end). */
int m2c_run(m2c_ctx *ctx, double tmax, int maxsteps, int sync_every);

int m2c_get_status(m2c_ctx *ctx, m2c_status *st);
int m2c_reset_time(m2c_ctx *ctx, double t);
int m2c_synchronize(m2c_ctx *ctx);

/* FluxFcnBase::ComputeNumericalFluxAtCellInterface (FluxFcnBase.h:65-69) for n face pairs
 * (Vm, Vp: n*5 doubles, F: n*5 doubles), evaluated by the device kernels. */
int m2c_flux_batch(m2c_ctx *ctx, int dir, int n, const double *Vm, const double *Vp, int id, double *F);

/* VarFcnBase::GetInternalEnergyPerUnitMass and GetTemperature (VarFcnBase.h, VarFcnSG.h:70-78,
 * VarFcnNASG.h:61, VarFcnMG.h:172-179) for n points (rho, p, material id), evaluated on the device
 * with the context's materials: e[t] always, T[t] when T is not NULL (NaN for a material without a
 * temperature law: JWL, dummy). Used by the outputs. */
int m2c_eos_derived(m2c_ctx *ctx, int n, const double *rho, const double *p, const int32_t *id, double *e, double *T);

/* ExactRiemannSolverBase::ComputeRiemannSolution(double *dir, double *Vm, int idm, double *Vp, int idp,
 * double *Vs, int &id, double *Vsm, double *Vsp) (ExactRiemannSolverBase.h:35-39) for n independent
 * problems, evaluated by the device solver of the interface faces. dir: n*3 doubles, the unit normal
 * of problem t at dir[3t..3t+2] (the mesh axis, or a level-set normal as
 * SpaceOperator::GetNormalForBimaterialRiemann gives it); Vm, Vp, Vs, Vsm, Vsp: n*5 doubles; idm, idp,
 * id, status: n ints. status[t]: 0 success, 1 solver failure (counted as a Riemann error by the
 * reference), -1 invalid input state or material index (the reference exits), 2 rarefaction path
 * longer than the device cache. */
int m2c_riemann_batch(m2c_ctx *ctx, int n, const double *dir, const double *Vm, const int *idm, const double *Vp,
                      const int *idp, double *Vs, int *id, double *Vsm, double *Vsp, int *status);

/* m2c_riemann_batch for callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc). No counterpart in the reference. Every pointer
 * is DEVICE memory owned by the caller (same layout as m2c_riemann_batch); the call allocates
 * nothing: the exact solver's rarefaction-path cache lives in a fixed device array of the library
 * (25 MB per GPU, held until the process ends) serving up to 512 concurrent solves, and n is
 * unbounded (grid-stride). Material indices are checked on the device (status -1). Calls that
 * overlap on one GPU share that array and race; the call returns after its stream has finished.
 * M2C_ERR_ARG if the context's path cache exceeds the array's 1024 entries per side. */
int m2c_riemann_batch_noalloc(m2c_ctx *ctx, int n, const double *dir, const double *Vm, const int *idm,
                              const double *Vp, const int *idp, double *Vs, int *id, double *Vsm, double *Vsp,
                              int *status);

/* VarFcnBase's pointwise set (VarFcnBase.h:62-294) for n points, evaluated on the device with the
 * context's materials: point t uses material id[t]. The operation fixes the strides of in / out:
 *   (rho, e) -> 1: PRESSURE (GetPressure), DPDRHO (GetDpdrho), BIG_GAMMA (GetBigGamma),
 *                  SOUND_SPEED_SQUARE (ComputeSoundSpeedSquare), SOUND_SPEED (ComputeSoundSpeed);
 *   (rho, p) -> 1: INTERNAL_ENERGY (GetInternalEnergyPerUnitMass), ENTHALPY (ComputeEnthalpyPerUnitMass);
 *   V[5] -> 1:     TOTAL_ENTHALPY (ComputeTotalEnthalpyPerUnitMass), MACH (ComputeMachNumber),
 *                  CHECK_STATE (CheckState(V): 1 invalid, 0 valid);
 *   U[5] -> V[5]:  C2P (ConservativeToPrimitive); V[5] -> U[5]: P2C (PrimitiveToConservative);
 *   V[5] -> V[5]:  CLIP (ClipDensityAndPressure).
 * flag (n ints, may be NULL): CHECK_STATE's result, CLIP's "clipped", and for SOUND_SPEED and MACH 1
 * where c^2 <= 0 (resp. < 0) makes the reference exit (the output is NaN there). */
enum m2c_varfcn_op {
  M2C_VF_PRESSURE = 0, M2C_VF_INTERNAL_ENERGY = 1, M2C_VF_DPDRHO = 2, M2C_VF_BIG_GAMMA = 3,
  M2C_VF_SOUND_SPEED_SQUARE = 4, M2C_VF_SOUND_SPEED = 5, M2C_VF_ENTHALPY = 6, M2C_VF_TOTAL_ENTHALPY = 7,
  M2C_VF_MACH = 8, M2C_VF_C2P = 9, M2C_VF_P2C = 10, M2C_VF_CHECK_STATE = 11, M2C_VF_CLIP = 12
};
int m2c_varfcn_batch(m2c_ctx *ctx, int op, int n, const int32_t *id, const double *in, double *out, int32_t *flag);

/* FluxFcnBase::ComputeNumericalFluxAtMaterialInterface(int dir, double *Vm, int idm, double *Vp, int idp,
 * double *F) (FluxFcnBase.h:72-76), which the reference provides for the local Lax-Friedrichs flux
 * (FluxFcnLLF.h:112-185) and SpaceOperator calls between two materials under MultiPhase.Flux = LLF
 * (SpaceOperator.cpp:2492-2496), for n face pairs along axis dir: Vm, Vp, F n*5 doubles. */
int m2c_interface_flux_batch(m2c_ctx *ctx, int dir, int n, const double *Vm, const int32_t *idm, const double *Vp,
                             const int32_t *idp, double *F);

/* ExactRiemannSolverBase::ComputeOneSidedRiemannSolution (ExactRiemannSolverBase.h:45-49) for n
 * independent problems: the fluid state Vm[t] (material idl[t]) against a wall moving with velocity
 * Ustar[3t..3t+2], along the unit normal dir[3t..3t+2] that points from the fluid to the wall.
 * Vs[t]: the state at the wall, id[t] its material (-1 and Vs zero when the wall recedes from the
 * fluid); Vsm[t]: the fluid-side star state. status[t] as for m2c_riemann_batch. */
int m2c_onesided_riemann_batch(m2c_ctx *ctx, int n, const double *dir, const double *Vm, const int *idl,
                               const double *Ustar, double *Vs, int *id, double *Vsm, int *status);

/* The flux on the fluid side of n mesh faces cut by an embedded wall, as SpaceOperator's advection
 * loop computes it (SpaceOperator.cpp:2438-2484 for i-1/2 faces, :2569-2611 j-1/2, :2696-2737
 * k-1/2): face t is normal to axis[t]; side[t] = +1 when the fluid cell (state V[t], material id[t])
 * is on the minus side of the face, -1 on the plus side; nwall/vwall[3t..3t+2] are the wall's unit
 * normal and velocity at the intersection. The one-sided problem runs along
 * GetNormalForOneSidedRiemann's direction (SpaceOperator.cpp:3506-3533; normal_mode 0 the wall normal
 * (EMBEDDED_SURFACE), 1 the mesh axis (MESH, the default), 2 their normalised sum (AVERAGE)); its
 * star state is clipped (ClipDensityAndPressure) into Vstar[t] and F[t] is the context's numerical
 * flux along the axis, F(V, V*) for side +1 and F(V*, V) for side -1. Finding the intersections
 * (EmbeddedBoundaryDataSet) is the caller's. */
int m2c_wall_face_flux_batch(m2c_ctx *ctx, int n, const int *axis, const int *side, int normal_mode,
                             const double *nwall, const double *vwall, const double *V, const int *id, double *F,
                             double *Vstar, int *status);

/* Mesh faces cut by embedded walls, fluxed inside every later stage of m2c_step / m2c_run /
 * m2c_compute_residual as SpaceOperator::ComputeAdvectionFluxes does when it is given
 * EmbeddedBoundaryDataSets (SpaceOperator.cpp:2438-2484 i-1/2, :2569-2611 j-1/2, :2696-2737 k-1/2):
 * each active side of a cut face solves its one-sided problem against the wall and takes the flux
 * between its state and the clipped star state; an inactive side's flux is zero. With
 * Reconstruction SlopeNearInterface Zero, slopes of cells with a cut face along an axis are zero
 * along it (Reconstructor.cpp:656-664). Replaces the intersections that FindEdgeSurfaceIntersections
 * (SpaceOperator.cpp:3434-3500) reads from the EmbeddedBoundaryDataSets: face t is normal to axis
 * face[4t] with plus cell (face[4t+1], face[4t+2], face[4t+3]), and wall[12t..12t+11] holds that
 * function's nwallf, vwallf (the intersection seen from the minus cell) and nwallb, vwallb (seen
 * from the plus cell). normal_mode: EmbeddedBoundaryMethod RiemannNormal (0 EMBEDDED_SURFACE,
 * 1 MESH, 2 AVERAGE); recon: ReconstructionAtInterface (0 CONSTANT, 1 LINEAR). Faces must be
 * interior and listed once; the walls stay until replaced (n = 0 removes them). Contexts with a
 * level set, material interfaces, viscous or heat-diffusion terms or a periodic boundary return
 * M2C_ERR_UNSUPPORTED. Host arrays. */
int m2c_set_embedded_faces(m2c_ctx *ctx, int n, const int *face, const double *wall, int normal_mode, int recon);

/* One embedded surface as the embedded-boundary operator hands it to the fluid solver: the
 * triangulated surface (TriangulatedSurface: node velocities Udot, triangles elems, unit normals
 * elemNorm) and its intersection points with mesh edges (EmbeddedBoundaryDataSet.h:34-60: the
 * triangle tid, barycentric coordinates xi within it, and dist, the distance from the edge's first
 * node to the crossing). Host arrays. */
typedef struct {
  int nnodes, nelems, npoints;
  const double *udot;        /* [nnodes][3] */
  const int *elems;          /* [nelems][3] node indices */
  const double *elem_norm;   /* [nelems][3] */
  const int *point_tid;      /* [npoints] */
  const double *point_xi;    /* [npoints][3] */
  const double *point_dist;  /* [npoints] */
} m2c_surface;

/* The wall records of mesh faces, as SpaceOperator::FindEdgeSurfaceIntersections
 * (SpaceOperator.cpp:3434-3500) derives them for ComputeAdvectionFluxes: for face t and surface s,
 * xf[t*nsurf+s] and xb[t*nsurf+s] are the indices of the surface's first intersection point along the
 * face's edge seen from its minus and from its plus cell (-1: none; the EmbeddedBoundaryDataSet's
 * XForward / XBackward). Over the surfaces the forward crossing with the smallest dist and the
 * backward one with the largest are taken (the first surface wins ties); each gives the triangle's
 * normal and the barycentric interpolation of its nodes' velocities. cut[t] = 1 and wall[12t..12t+11]
 * = nwallf, vwallf, nwallb, vwallb (m2c_set_embedded_faces' layout) for a face with a crossing;
 * cut[t] = 0 and the record left alone without one. M2C_ERR_ARG for an index out of range and for a
 * face with a crossing in one direction only (where the reference asserts). Host code. */
int m2c_edge_surface_walls(int nsurf, const m2c_surface *surf, int nfaces, const int *xf, const int *xb, int *cut,
                           double *wall);

/* ---- many independent instances (batched) ------------------------------------------------
 * A batch holds n instances, each its own grid, boundaries, state and CFL-adaptive time step; one
 * flux, precision and RK2/RK3 scheme for all, and one material table (the materials an instance
 * defines are the leading entries of the first instance's); the flux HLLC, or Roe or LLF in FP64,
 * else M2C_ERR_UNSUPPORTED. An instance is single-material, or two-material: materials meeting at
 * interface faces (exact Riemann solves there), with a level set (advection, reinitialisation,
 * material-ID and state update of swept cells). Every RK stage runs as persistent launches draining
 * per-XCD tile queues: one for all instances, or, when two-material instances are present, one for
 * the single-material instances and one for the two-material ones; the latter's interface passes,
 * level-set stages and end-of-step tasks run around those launches on the batch's stream. A time
 * step is captured once in a hipGraph and replayed (with two-material instances, a
 * reinitialisation step is replayed as two graphs with the host's check of the reinitialisation
 * failsafe between them). An instance evolves bit for bit as it would in its own context (no
 * counterpart in the reference, which runs one problem per process). */
typedef struct m2c_batch m2c_batch;
typedef struct {
  int kchunk;                       /* z planes per work item */
  int blocks;                       /* resident persistent workgroups */
  double ms_per_step;               /* best time per batch step among the candidates */
  int candidates;
  int tile_x, tile_y;               /* columns per work item: 32 x 8 or 16 x 16 */
} m2c_batch_tuning;
int m2c_batch_create(const m2c_config *cfgs, int n, m2c_batch **out);
int m2c_batch_destroy(m2c_batch *b);
int m2c_batch_upload(m2c_batch *b, int i, const double *V, const int32_t *id);
int m2c_batch_download(m2c_batch *b, int i, double *V, int32_t *id);
/* level set of a two-material instance with one (nz*ny*nx doubles, as m2c_upload_levelset) */
int m2c_batch_upload_levelset(m2c_batch *b, int i, const double *phi);
int m2c_batch_download_levelset(m2c_batch *b, int i, double *phi);
/* nsteps steps of every instance; use_graph: replay the captured step graph; ms: elapsed (HIP events) */
int m2c_batch_run(m2c_batch *b, int nsteps, int use_graph, double *ms);
int m2c_batch_status(m2c_batch *b, int i, m2c_status *st);
/* times (tile shape, z-chunk depth, resident workgroups) candidates for trial_steps steps each on
 * a copy of the state, keeps the fastest; the state, clocks and counters are restored */
int m2c_batch_autotune(m2c_batch *b, int trial_steps, m2c_batch_tuning *out);
int m2c_batch_set_tuning(m2c_batch *b, int kchunk, int blocks);
/* the tile shape of the work items: 32 x 8 or 16 x 16 columns (else M2C_ERR_UNSUPPORTED) */
int m2c_batch_set_tile(m2c_batch *b, int tile_x, int tile_y);

/* Learned interface solver (no counterpart in the reference). A 10-256-256-256-256-3 SiLU network
 * predicts the star state of a two-stiffened-gas Riemann problem along a mesh axis; the predicted
 * p* is polished by at most `newton_max` Newton updates of the closed-form stiffened-gas wave curves
 * and accepted only when the velocity-continuity residual |u*_L(p) - u*_R(p)| / max(c_L, c_R) is
 * <= riemann.surrogate_residual_tol. Otherwise ExactRiemannSolverBase::ComputeRiemannSolution
 * (ExactRiemannSolverBase.h:35-39) runs as in the plain path. Network inputs: ln(rho_R/rho_L)/7,
 * ln(P_R/P_L)/9, (u_R-u_L)/(c_L+c_R), ln(c_R/c_L)/5, ln gam_L, ln gam_R, pinf_L/P_L, pinf_R/P_R,
 * ln(Z_R/Z_L)/7 and ln((p_acoustic + min pinf)/sqrt(P_L P_R))/4 with P = p + pinf, Z = rho c;
 * outputs ln((p* + min pinf)/sqrt(P_L P_R))/4, (u* - (u_L+u_R)/2)/(c_L+c_R), ln(rho*_L/rho_L)/4.
 * Parameters: W0 [256][10], b0, W1..W3 [256][256], b1..b3, W4 [3][256], b4 (FP32, flat, 200963);
 * the kernels compute in bf16 MFMA with FP32 accumulation. */
typedef struct m2c_mlp m2c_mlp;
typedef struct {
  long long n_samples;              /* training samples from the exact solver (a power of two) */
  long long n_holdout;              /* held-out samples (multiple of 64; 0 = none) */
  unsigned long long seed;
  int batch;                        /* samples per Adam step (multiple of 64, <= n_samples) */
  int epochs;
  double lr, lr_min;                /* Adam step size, cosine-decayed to lr_min over the run */
  int newton_max;                   /* Newton updates allowed after a prediction (held-out check) */
  double tol;                       /* continuity-residual acceptance (held-out check) */
  int verbose;                      /* progress lines on stderr */
} m2c_train_config;
typedef struct {
  double data_ms;                   /* exact-solver sample generation */
  double train_ms;                  /* training steps (HIP events) */
  double samples_per_s;             /* training throughput */
  long long steps, valid_samples;
  double first_epoch_loss, last_epoch_loss, holdout_loss;   /* mean squared error per sample */
  double holdout_accept_mlp;        /* share of held-out problems accepted from the network's p* */
  double holdout_accept_acoustic;   /* ... from the linearised (acoustic) p*, same Newton budget */
  double holdout_iters_mlp, holdout_iters_acoustic;   /* mean Newton updates to reach tol (cap 50) */
} m2c_train_report;
int m2c_mlp_create(int device, unsigned long long seed, m2c_mlp **out);
int m2c_mlp_destroy(m2c_mlp *net);
int m2c_mlp_nparams(void);
int m2c_mlp_get_params(m2c_mlp *net, float *params);
int m2c_mlp_set_params(m2c_mlp *net, const float *params);   /* also resets the Adam moments */
/* Y[n][3] = net(X[n][10]) (host arrays) */
int m2c_mlp_forward(m2c_mlp *net, int n, const float *X, float *Y);
/* loss = sum (y - t)^2 / n over rows and outputs (rows with NaN targets excluded) and its gradient
 * (flat parameter layout); n a multiple of 64 */
int m2c_mlp_gradient(m2c_mlp *net, int n, const float *X, const float *T, double *loss, float *grad);
/* n training samples generated by the exact solver: X [n][10], T [n][3] (NaN where it fails),
 * raw [n][14] = rho,u,p,gam,pinf of the left and right states, p*, u*, rho*_L, rho*_R (or NULL) */
int m2c_mlp_samples(m2c_mlp *net, long long n, unsigned long long seed, float *X, float *T, double *raw);
int m2c_mlp_train(m2c_mlp *net, const m2c_train_config *cfg, m2c_train_report *report);
/* average duration (ms) of one inference launch over n device-resident rows */
int m2c_mlp_bench_infer(m2c_mlp *net, int n, int reps, double *ms_per_launch);
/* average duration (ms) of one training step (forward, backward, weight gradients, Adam) with
 * batch B on device-resident synthetic rows; the parameters change */
int m2c_mlp_bench_train(m2c_mlp *net, int B, int steps, double *ms_per_step);
/* Interface faces of `ctx` use the network (NULL: the exact solver only). The network must stay
 * alive while attached. Requires two-material interfaces and MultiPhase.Flux = numerical. */
int m2c_set_surrogate(m2c_ctx *ctx, const m2c_mlp *net, int newton_max);
/* m2c_riemann_batch through the network with the exact fallback; path[t] = 1 network, 0 exact */
int m2c_surrogate_riemann_batch(m2c_ctx *ctx, const m2c_mlp *net, int newton_max, int n, const int *dir,
                                const double *Vm, const int *idm, const double *Vp, const int *idp, double *Vs,
                                int *id, double *Vsm, double *Vsp, int *status, int *path);

/* Timing helper: runs `nsteps` CFL-adaptive time steps (no tmax) and reports the average
 * duration (ms) of one fused stage-kernel launch and of one whole step, measured with HIP events
 * recorded on the context's stream around every stage launch and around the whole loop. */
int m2c_time_steps(m2c_ctx *ctx, int nsteps, double *ms_stage_kernel, double *ms_step);

/* Average duration (ms) per stage of the face passes that precede the stage kernel (viscous /
 * heat-diffusion terms, Godunov fluxes; three face kernels each) in the last m2c_time_steps call,
 * which leaves them out of ms_stage_kernel; 0 without such a pass. */
int m2c_time_diffusion_pass(m2c_ctx *ctx, double *ms);

/* Measured HBM copy bandwidth: a plain device copy kernel moves `nbytes` (read + write) `reps`
 * times between two buffers of that size; returns the average kernel duration (ms), timed with
 * HIP events on the context's stream. Also the known-size calibration kernel for traffic
 * counters. */
int m2c_stream_copy(m2c_ctx *ctx, long long nbytes, int reps, double *ms_avg);

/* Launch shape of the fused stage kernel (no reference counterpart): each 32x8 column tile marches
 * up chunks of kchunk z planes; 0 restores the default (chunks such that the grid has about 2048
 * workgroups). Results do not depend on it. */
int m2c_set_kchunk(m2c_ctx *ctx, int kchunk);

/* Raw device pointer of the primitive state (SoA, 5*nz*ny*nx doubles) for zero-copy callers. */
int m2c_device_state(m2c_ctx *ctx, double **V_soa);

const char *m2c_version(void);

/* ---- input decks ----------------------------------------------------------------------------
 * Host-only calls (no device is touched). */

/* IoData's defaults for the fields of m2c_config (the IoData.cpp constructors: MeshData :453,
 * ReconstructionData :1049, SchemeData :1154, LevelSetReinitializationData :1205, LevelSetSchemeData
 * :1245, ExactRiemannSolverData :1345, MultiPhaseData :1400, ExplicitData :1472, TsData :1493).
 * No materials and no cells: nmat = 0, nx = ny = nz = -1, boundaries None (0). */
int m2c_config_defaults(m2c_config *cfg);

/* Run controls and the uniform initial state of a deck (TsData, IoData.cpp:1493-1528; the
 * top-level DefaultInitialState, IoData.cpp:1695). */
typedef struct {
  int max_its;                      /* Time.MaxIts (INT_MAX by default) */
  double max_time;                  /* Time.MaxTime (1e10) */
  double time_step;                 /* Time.TimeStep: fixed dt when > 0 (-1: CFL-adaptive) */
  double initial_state[5];          /* DefaultInitialState: rho, u, v, w, p (1e-6, 0, 0, 0, 0) */
  int initial_material;             /* DefaultInitialState.MaterialID */
} m2c_deck_run;

/* Reads an M2C input deck (the text of the file main() passes to IoData::readCmdFile,
 * IoData.cpp:3235, through the grammar of parser/Parser.y) into *cfg and, when run is not
 * NULL, *run. Returns M2C_OK, M2C_ERR_ARG for a syntax error or a value the reference rejects, or
 * M2C_ERR_UNSUPPORTED for a valid option this library does not provide. msg (msglen bytes, may
 * be NULL) receives the error with its line, or on success the ignored options (output and other
 * objects outside the finite-volume update), one per line. *cfg and *run are written only on
 * success. */
int m2c_config_from_deck(const char *text, m2c_config *cfg, m2c_deck_run *run, char *msg, int msglen);

/* Initial state of a deck on its mesh (cell centres as m2c_mesh_1d generates them):
 * SpaceOperator::SetInitialCondition
 * (SpaceOperator.cpp:696-960) and, for the material of its level set, LevelSetOperator::
 * SetInitialCondition (LevelSetOperator.cpp:291-556). DefaultInitialState everywhere, then the
 * GeometricEntities Plane, CylinderAndCone, CylinderWithSphericalCaps, Sphere and Parallelepiped
 * (each kind in index order, later ones overriding). V: nz*ny*nx*5 doubles, id: nz*ny*nx int32 (the
 * layout of m2c_upload_state), phi: nz*ny*nx doubles or NULL (left alone when the deck has no level
 * set). Spheroid, Point, ArbitraryEnclosure and UserDataFile give M2C_ERR_UNSUPPORTED. */
int m2c_initial_condition_from_deck(const char *text, double *V, int32_t *id, double *phi, char *msg, int msglen);

/* ---- outputs ---------------------------------------------------------------------------------
 * The Output object of an input deck (Output::Output / InitializeOutput, Output.cpp:21-113):
 * creates <Prefix><Solution>.pvd (an empty collection, Output.cpp:28-46), opens the probe files
 * with their headers (ProbeOutput.cpp:13-175) and locates probe and line-plot points
 * (ProbeOutput.cpp:226-339). Host-only. msg receives the error, or the requested outputs this
 * library does not write (one per line). */
typedef struct m2c_output m2c_output;
int m2c_output_create(const char *text, m2c_output **out, char *msg, int msglen);

/* Output::OutputSolutions (Output.cpp:117-143) with the state of one step: the snapshot
 * (<Solution>_NNNN.vtr and its pvd line), the probes and the line plots whose isTimeToWrite
 * (Utils.cpp:293-313) holds. V, ID, phi: the host layout of m2c_download_state /
 * m2c_download_levelset (phi required when the deck has a level set). ctx evaluates internal
 * energy and temperature on the device (m2c_eos_derived); it may be NULL when
 * m2c_output_needs_device(o) is 0. written (may be NULL) gets bit 0 snapshot, bit 1 probes,
 * bit 2 line plots. */
int m2c_output_write(m2c_output *o, m2c_ctx *ctx, double time, double dt, int time_step, const double *V,
                     const int32_t *ID, const double *phi, int force_write, int *written);
int m2c_output_needs_device(const m2c_output *o);
int m2c_output_destroy(m2c_output *o);

#ifdef __cplusplus
}
#endif
#endif
