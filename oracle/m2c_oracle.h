/*
 * oracle/m2c_oracle.h -- CPU restatement of M2C's explicit finite-volume update.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * load liboracle.so as the checker. The product (m2c_amd/, include/) never links it.
 *
 * Data layout mirrors the reference's SpaceVariable3D (SpaceVariable.cpp:254-260): array-of-
 * structures, k-j-i, x fastest, one ghost layer at index -1 and N in every direction.
 * Arrays passed in/out of this API hold the ghosted box: (nx+2)*(ny+2)*(nz+2) cells.
 */
#ifndef M2C_ORACLE_H
#define M2C_ORACLE_H

#include "eos.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_MAT 8

/* boundary codes: IoData.h:266-267 (NONE..OVERSET), plus PERIODIC as an extension */
enum { ORC_BC_NONE = 0, ORC_BC_INLET = 1, ORC_BC_OUTLET = 2, ORC_BC_SLIPWALL = 3, ORC_BC_STICKWALL = 4,
       ORC_BC_SYMMETRY = 5, ORC_BC_OVERSET = 6, ORC_BC_PERIODIC = 7 };
enum { ORC_REC_CONSTANT = 0, ORC_REC_LINEAR = 1 };
enum { ORC_LIM_NONE = 0, ORC_LIM_GENMINMOD = 1, ORC_LIM_VANALBADA = 2 };
enum { ORC_TS_FE = 0, ORC_TS_RK2 = 1, ORC_TS_RK3 = 2 };
enum { ORC_MPFLUX_NUMERICAL = 0, ORC_MPFLUX_EXACT = 1, ORC_MPFLUX_LLF = 2 };

typedef struct {
  /* Riemann solver controls, IoData.cpp:1345-1357 (ExactRiemannSolverData) */
  int maxIts_main, maxIts_bracket, maxIts_shock, numSteps_rarefaction;
  double tol_main, tol_shock, tol_rarefaction, min_pressure, failure_threshold, pressure_at_failure;
  double surrogate_residual_tol;
} orc_riemann_cfg;

typedef struct {
  int nx, ny, nz;
  double x0, x1, y0, y1, z0, z1;
  int bc[6];                /* x0, xmax, y0, ymax, z0, zmax */
  double inlet[5], outlet[5];
  int nmat;                 /* number of entries in mat[], the last one may be the inactive dummy */
  int inactive_id;          /* -1 if no inactive material */
  orc_eos mat[ORC_MAX_MAT];
  int flux;                 /* ORC_FLUX_* */
  double roe_delta;         /* Harten entropy-fix coefficient (IoData.cpp:1158) */
  int rec_type, limiter, slope_zero_near_interface;
  double gminmod_alpha;
  int vanalbada_pow;        /* 1: pow(dq,k) exactly as Reconstructor.h:107-108; 0: dq*dq */
  int ts;                   /* ORC_TS_* */
  double cfl;
  int mp_flux, mp_recon;    /* MultiPhaseData flux / recon (IoData.cpp:1400-1416) */
  orc_riemann_cfg riemann;
  int nthreads;
  /* level set (LevelSetSchemeData / LevelSetReinitializationData, IoData.cpp:1205-1266) */
  int ls_on, ls_matid;
  int ls_bc[6];             /* 0 none, 1 zero-Neumann, 2 linear extrapolation, 3 non-negative, 7 periodic */
  int reinit_freq, reinit_maxits;
  double reinit_cfl, reinit_tol;
  int mp_normal;            /* MultiPhase.RiemannNormal: 0 LevelSet, 1 Mesh, 2 Average */
  double conrec_depth;      /* MultiPhase.ConstantReconstructionDepth (0: none) */
  /* Mesh.ControlPointX/Y/Z: mesh_npts[d] > 0 stretches axis d (MeshGenerator.cpp:163-259), with
   * n[d] the number of cells that generates; 0 = uniform */
  int mesh_npts[3];
  double mesh_pts[3][16][2];
} orc_config;

/* cells of one mesh axis (oracle/mesh.c): n uniform cells, or the stretched cells of npts control
 * points (coordinate, width); up to cap centres x / widths dx, their count in *nout; -1 on error */
int orc_mesh_1d(double x0, double xmax, int n, int npts, const double *pts, int cap, double *x, double *dx,
                int *nout);

/* one embedded surface: node velocities, triangles, unit normals, and its edge crossings (triangle,
 * barycentric coordinates, distance from the edge's first node) */
typedef struct {
  int nnodes, nelems, npoints;
  const double *udot;
  const int *elems;
  const double *elem_norm;
  const int *point_tid;
  const double *point_xi;
  const double *point_dist;
} orc_surface;

/* wall records of nfaces mesh faces from the forward / backward crossing indices xf, xb
 * ([face][surface], -1 none): cut[t] and wall[12t..] = nwallf, vwallf, nwallb, vwallb (oracle/embedded_walls.c);
 * -1 for an index out of range or a crossing seen from one side only */
int orc_edge_surface_walls(int nsurf, const orc_surface *surf, int nfaces, const int *xf, const int *xb, int *cut,
                           double *wall);

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const orc_config *cfg);
void orc_destroy(orc_ctx *c);
long orc_ncells_ghosted(const orc_ctx *c);

/* Pointwise entry points (for pointwise parity tests). */
void orc_numerical_flux(const orc_ctx *c, int dir, const double *Vm, const double *Vp, int id, double *F);

/* SpaceOperator::ApplyBoundaryConditions (SpaceOperator.cpp:1667-1745) on a ghosted V. */
void orc_apply_bc(const orc_ctx *c, double *V);
/* MultiPhaseOperator::UpdateMaterialIDAtGhostNodes (MultiPhaseOperator.cpp:796-819). */
void orc_apply_bc_id(const orc_ctx *c, int *ID);

/* SpaceOperator::ComputeResidual (SpaceOperator.cpp:3351-3409), single stage.
 * R is ghosted (only interior entries are meaningful). Returns the number of Riemann errors,
 * or -1 if a reconstructed state violates hyperbolicity (the reference exits). */
int orc_compute_residual(orc_ctx *c, const double *V, const int *ID, double *R);

/* The (ghosted) level set that orc_compute_residual's interface faces take their normal from when
 * RiemannNormal is LevelSet or Average (NULL: the mesh normal). orc_run sets Phi(n) itself. */
void orc_set_levelset(orc_ctx *c, const double *phi);

/* Direction of the interface Riemann problem at the d-face of cell (i,j,k) (between (i,j,k) - e_d,
 * material nbid, and (i,j,k), material myid): GetNormalForBimaterialRiemann with
 * CalculateGradPhiAtCellInterface (SpaceOperator.cpp:2966-3048) on the ghosted level set phi. */
void orc_interface_normal(const orc_ctx *c, const double *phi, int d, int i, int j, int k, int myid, int nbid,
                          double *dir);

/* SpaceOperator::ComputeTimeStepSize (SpaceOperator.cpp:2202-2227). Returns 0 or -1 (c^2<0). */
int orc_time_step(const orc_ctx *c, const double *V, const int *ID, double *dt);

/* One time step (TimeIntegrator.cpp FE :150-226, RK2 :291-428, RK3 :496-690) for the
 * Navier-Stokes part. V must have valid ghosts; on return V = V(n+1) with ghosts filled. */
int orc_advance(orc_ctx *c, double *V, const int *ID, double dt);

/* Main.cpp:514-621 time loop: runs until t >= tmax or nsteps steps. phi (ghosted, may be NULL
 * when the level set is off) is advanced with the flow; ID is updated by the level set. */
int orc_run(orc_ctx *c, double *V, int *ID, double *phi, double tmax, int maxsteps, double *t_out, int *steps_out);

/* LevelSetOperator::ApplyBoundaryConditions (LevelSetOperator.cpp:934-1033) on a ghosted phi. */
void orc_apply_bc_phi(const orc_ctx *c, double *phi);

/* LevelSetOperator::ComputeResidualFDM_FullDomain (LevelSetOperator.cpp:1066-1141). */
void orc_ls_residual(orc_ctx *c, const double *V, const double *phi, double *R);

/* LevelSetReinitializer::ReinitializeFullDomain (LevelSetReinitializer.cpp:74-172), FIXED first
 * layer. Returns the number of iterations, or -1 on failure. */
int orc_ls_reinitialize(orc_ctx *c, double *phi);

/* MultiPhaseOperator::UpdateMaterialIDByLevelSet (MultiPhaseOperator.cpp:127-330) */
void orc_update_id(orc_ctx *c, const double *phi0, const double *phi, int *ID);

/* Exact two-material Riemann problem (ExactRiemannSolverBase::ComputeRiemannSolution). */
int orc_riemann(const orc_ctx *c, const double *dir, const double *Vm, int idm, const double *Vp, int idp,
                double *Vs, int *id, double *Vsm, double *Vsp);

/* One-sided (wall) Riemann problem (ExactRiemannSolverBase::ComputeOneSidedRiemannSolution) of the
 * fluid state Vm (material idl) against a wall moving with velocity Ustar along the unit normal dir
 * from the fluid to the wall. id = -1 (Vs untouched) when the wall recedes. */
int orc_riemann_onesided(const orc_ctx *c, const double *dir, const double *Vm, int idl, const double *Ustar,
                         double *Vs, int *id, double *Vsm);

/* GetNormalForOneSidedRiemann (SpaceOperator.cpp:3506-3533) and the flux on the fluid side of a
 * wall-cut face (SpaceOperator.cpp:2438-2484); side +1: fluid cell on the minus side of the face. */
void orc_onesided_normal(int d, int fb, int mode, const double *nwall, double *dir);
int orc_wall_face_flux(const orc_ctx *c, int axis, int side, int mode, const double *nwall, const double *vwall,
                       const double *V, int id, double *F, double *Vstar);

/* Mesh faces cut by embedded walls, for orc_compute_residual and orc_run (SpaceOperator.cpp:
 * 2438-2484, 2569-2611, 2696-2737; FindEdgeSurfaceIntersections :3434-3500 is the caller's): face t
 * is normal to axis face[4t] with plus cell (face[4t+1], face[4t+2], face[4t+3]); wall[12t..] holds
 * nwallf, vwallf (forward intersection, seen from the minus cell) and nwallb, vwallb (backward,
 * seen from the plus cell). normal_mode is EmbeddedBoundaryMethod RiemannNormal (0 EMBEDDED_SURFACE,
 * 1 MESH, 2 AVERAGE), recon its ReconstructionAtInterface (0 CONSTANT, 1 LINEAR). Faces on the domain
 * boundary are rejected. n = 0 removes the walls. Returns 0, or -1 on a bad face. */
int orc_set_embedded_faces(orc_ctx *c, int n, const int *face, const double *wall, int normal_mode, int recon);

/* Pointwise EOS of material id: PrimitiveToConservative / ConservativeToPrimitive
 * (VarFcnBase.h:169-197) and the squared sound speed at (rho, p) (VarFcnBase.h:215-219). */
void orc_eos_p2c(const orc_ctx *c, int id, const double *V, double *U);
void orc_eos_c2p(const orc_ctx *c, int id, const double *U, double *V);
double orc_eos_c2(const orc_ctx *c, int id, double rho, double p);

/* Pointwise EOS of material id: out[0] p(rho,e), out[1] e(rho,p), out[2] dp/drho at constant e,
 * out[3] Gamma = (1/rho) dp/de, out[4] c^2(rho,e), out[5] CheckState(rho,p) (1 = invalid). */
void orc_eos_point(const orc_ctx *c, int id, double rho, double e, double p, double *out);

/* The interface Riemann solutions of V as the first stage of time step `stamp` stores them
 * (SpaceOperator.cpp:2518-2523); returns the residual's status. */
int orc_store_riemann(orc_ctx *c, const double *V, const int *ID, int stamp);
/* UpdateStateVariablesByRiemannSolutions + FixUnresolvedNodes (MultiPhaseOperator.cpp:857-911,
 * 1156-1320) on the ghosted V from the stored solutions; IDn before, ID after the interface moved.
 * Returns the number of unresolved cells. */
long orc_update_states(orc_ctx *c, const int *IDn, const int *ID, double *V);

/* FluxFcnLLF::ComputeNumericalFluxAtMaterialInterface (FluxFcnLLF.h:112-185) between materials idm, idp. */
void orc_material_interface_flux(const orc_ctx *c, int dir, const double *Vm, int idm, const double *Vp, int idp,
                                 double *F);

/* Multiphase counters since creation: reinitialisation iterations (retries included), unresolved
 * cells of the state update, and the reinitialisation CFL number in force (divided by 1.5 on each
 * retry, LevelSetReinitializer.cpp:155-168). */
void orc_mp_counters(const orc_ctx *c, long *reinit_iters, long *unresolved, double *reinit_cfl);

/* Deterministic log/exp used by the Riemann solver's step control (oracle/riemann.c). */
double orc_det_log(double x);
double orc_det_exp(double y);
double orc_pow_quarter(double x);
double orc_pow_fifth(double x);

/* Diagnostics counters of the last call (reverted face states, clipped cells). */
void orc_counters(const orc_ctx *c, long *reverted_faces, long *clipped_cells, long *riemann_errors);

#ifdef __cplusplus
}
#endif
#endif
