// m2c_host.hpp -- C++ host layer over the C ABI of include/m2c_hip.h, shaped like the reference's
// operator interfaces so that code written against them keeps its calls:
//
//   SpaceOperator           SpaceOperator.h:103-140 (ComputeResidual, ComputeTimeStepSize,
//                           ConservativeToPrimitive-free state transfer)
//   TimeIntegrator          TimeIntegrator.h:66-73 (AdvanceOneTimeStep) and the time loop of
//                           Main.cpp:514-621
//   ExactRiemannSolver      ExactRiemannSolverBase.h:35-49 (ComputeRiemannSolution,
//                           ComputeOneSidedRiemannSolution)
//   LevelSetOperator        LevelSetOperator.h:86-90 (the level set Phi of one material)
//   VarFcn                  VarFcnBase.h:62-294 (the pointwise equation-of-state set of one
//                           material, evaluated on the device)
//   FluxFcn                 FluxFcnBase.h:65-76 (ComputeNumericalFluxAtCellInterface and
//                           ComputeNumericalFluxAtMaterialInterface)
//
// Arrays are the interior of SpaceVariable3D (k-j-i, dof innermost, no ghosts). Fatal conditions
// that make the reference call exit() throw m2c_host::Error carrying the status code; a program
// that does not catch it terminates, as the reference does.
#pragma once
#include <cmath>
#include <cstdint>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/m2c_hip.h"

namespace m2c_host {

struct Error : std::runtime_error {
  int code;
  Error(const std::string &what, int c) : std::runtime_error(what + " (status " + std::to_string(c) + ")"), code(c) {}
};

inline void check(int rc, const char *what) {
  if (rc != M2C_OK) throw Error(what, rc);
}

// Solver options in IoData's vocabulary (defaults of IoData.cpp for the fields on this path).
inline m2c_config default_config(int nx, int ny, int nz, double x0, double x1, double y0, double y1, double z0,
                                 double z1) {
  m2c_config c = {};
  c.nx = nx; c.ny = ny; c.nz = nz;
  c.x0 = x0; c.x1 = x1; c.y0 = y0; c.y1 = y1; c.z0 = z0; c.z1 = z1;
  for (int s = 0; s < 6; s++) { c.bc[s] = M2C_BC_SYMMETRY; c.ls_bc[s] = 1; }
  c.inlet[0] = c.outlet[0] = 1.0;
  c.inlet[4] = c.outlet[4] = 1.0;
  c.nmat = 0;
  c.inactive_id = -1;
  c.flux = M2C_FLUX_HLLC;
  c.roe_delta = 0.2;               // IoData.cpp:1158
  c.rec_linear = 1;                // ReconstructionData: linear, no limiter (IoData.cpp:1049-1056)
  c.limiter = M2C_LIM_NONE;
  c.gminmod_alpha = 1.2;
  c.ts = M2C_TS_RK2;               // TimeIntegration default: Runge-Kutta 2
  c.cfl = 0.5;
  c.mp_flux = 0;                   // MultiPhaseData: Numerical flux, Constant reconstruction
  c.mp_recon = 0;
  m2c_riemann_params &r = c.riemann;   // ExactRiemannSolverData, IoData.cpp:1345-1357
  r.maxIts_main = 200; r.maxIts_bracket = 100; r.maxIts_shock = 200; r.numSteps_rarefaction = 200;
  r.tol_main = 1e-4; r.tol_shock = 1e-12; r.tol_rarefaction = 1e-8;
  r.min_pressure = -1e8; r.failure_threshold = 0.2; r.pressure_at_failure = 1e-8;
  r.surrogate_residual_tol = 1e-6;
  c.precision = M2C_PREC_FP64;
  c.device = 0;
  c.ls_matid = -1;
  c.reinit_freq = -1;              // LevelSetReinitializationData, IoData.cpp:1205-1212
  c.reinit_maxits = 30;
  c.reinit_cfl = 0.8;
  c.reinit_tol = 2.0e-4;
  c.mp_riemann_normal = 1;         // MultiPhaseData::riemann_normal = MESH (IoData.cpp:1408)
  c.mp_conrec_depth = 0.0;         // MultiPhaseData::conRec_depth (IoData.cpp:1416)
  return c;
}

inline int add_stiffened_gas(m2c_config &c, double gamma, double pinf, double failsafe_density = 1e-10) {
  m2c_material &m = c.mat[c.nmat];
  m = {};
  m.eos = M2C_EOS_SG;
  m.rhomin = 0.0; m.pmin = -1.7976931348623157e308;
  m.rhomax = m.pmax = 1.7976931348623157e308;
  m.failsafe_density = failsafe_density;
  m.p[0] = gamma; m.p[1] = pinf;
  return c.nmat++;
}

// IoData(argc, argv) -> readCmdFile (IoData.cpp:3214-3250): an input deck read into the solver
// options, the run controls of Time, the ignored options outside the finite-volume update, and the
// initial state of DefaultInitialState and GeometricEntities on the deck's mesh.
class IoData {
 public:
  explicit IoData(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw Error("cannot open input file " + path, M2C_ERR_ARG);
    std::stringstream ss;
    ss << f.rdbuf();
    text_ = ss.str();
    parse();
  }
  static IoData FromText(const std::string &text) { return IoData(text, 0); }

  m2c_config config;
  m2c_deck_run run;
  std::vector<std::string> ignored;

  long long cells() const { return (long long)config.nx * config.ny * config.nz; }
  bool has_levelset() const { return config.ls_matid >= 0; }
  const std::string &text() const { return text_; }
  // SpaceOperator::SetInitialCondition and LevelSetOperator::SetInitialCondition; phi is left
  // empty when the deck has no level set
  void InitialCondition(std::vector<double> &V, std::vector<int32_t> &ID, std::vector<double> &phi) const {
    V.assign(5 * (size_t)cells(), 0.0);
    ID.assign((size_t)cells(), 0);
    phi.assign(has_levelset() ? (size_t)cells() : 0, 0.0);
    char msg[1024];
    check_msg(m2c_initial_condition_from_deck(text_.c_str(), V.data(), ID.data(), has_levelset() ? phi.data() : nullptr,
                                              msg, sizeof msg),
              msg);
  }

 private:
  IoData(const std::string &text, int) : text_(text) { parse(); }
  void parse() {
    std::vector<char> msg(1 << 16);
    check_msg(m2c_config_from_deck(text_.c_str(), &config, &run, msg.data(), (int)msg.size()), msg.data());
    std::stringstream ss(msg.data());
    for (std::string line; std::getline(ss, line);)
      if (!line.empty()) ignored.push_back(line);
  }
  static void check_msg(int rc, const char *msg) {
    if (rc != M2C_OK) throw Error(std::string("input deck: ") + msg, rc);
  }
  std::string text_;
};

// Output (Output.h): snapshots, probes and line plots of a deck's Output object.
class Output {
 public:
  explicit Output(const IoData &iod) {
    std::vector<char> msg(1 << 16);
    const int rc = m2c_output_create(iod.text().c_str(), &out_, msg.data(), (int)msg.size());
    if (rc != M2C_OK) throw Error(std::string("output: ") + msg.data(), rc);
    std::stringstream ss(msg.data());
    for (std::string line; std::getline(ss, line);)
      if (!line.empty()) not_written.push_back(line);
  }
  ~Output() { m2c_output_destroy(out_); }
  Output(const Output &) = delete;
  Output &operator=(const Output &) = delete;
  // Output::OutputSolutions (Output.cpp:117-143); returns what was written (1 snapshot, 2 probes, 4 lines)
  int OutputSolutions(m2c_ctx *ctx, double t, double dt, int time_step, const double *V, const int32_t *ID,
                      const double *phi, bool force_write) {
    int w = 0;
    check(m2c_output_write(out_, ctx, t, dt, time_step, V, ID, phi, force_write ? 1 : 0, &w), "OutputSolutions");
    return w;
  }
  std::vector<std::string> not_written;

 private:
  m2c_output *out_ = nullptr;
};

// One device context: the SpaceOperator / TimeIntegrator / LevelSetOperator of a subdomain.
class Context {
 public:
  explicit Context(const m2c_config &cfg) : cfg_(cfg) { check(m2c_ctx_create(&cfg_, &ctx_), "m2c_ctx_create"); }
  ~Context() { m2c_ctx_destroy(ctx_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  m2c_ctx *get() const { return ctx_; }
  const m2c_config &config() const { return cfg_; }
  long long cells() const { return (long long)cfg_.nx * cfg_.ny * cfg_.nz; }

 private:
  m2c_config cfg_;
  m2c_ctx *ctx_ = nullptr;
};

class SpaceOperator {
 public:
  explicit SpaceOperator(Context &c) : c_(c) {}
  // state V (5 primitive dofs per cell) and material IDs
  void SetState(const double *V, const int32_t *ID) { check(m2c_upload_state(c_.get(), V, ID), "m2c_upload_state"); }
  void GetState(double *V, int32_t *ID) const { check(m2c_download_state(c_.get(), V, ID), "m2c_download_state"); }
  // SpaceOperator::ComputeResidual(V, ID, R): R = -(sum of face fluxes)/volume on the current state
  void ComputeResidual(double *R) { check(m2c_compute_residual(c_.get(), R), "ComputeResidual"); }
  // SpaceOperator::ComputeTimeStepSize(V, ID, dt, cfl)
  void ComputeTimeStepSize(double &dt, double &cfl) {
    check(m2c_time_step_size(c_.get(), &dt, &cfl), "ComputeTimeStepSize");
  }
  // The EmbeddedBoundaryDataSet argument of ComputeResidual: faces cut by embedded walls, as
  // FindEdgeSurfaceIntersections reports them (face[4t] axis, then the plus cell's i, j, k;
  // wall[12t..] nwallf, vwallf, nwallb, vwallb), kept for every later stage until replaced
  void SetEmbeddedFaces(const std::vector<int> &face, const std::vector<double> &wall, int riemann_normal = 1,
                        int recon = 0) {
    if (face.size() % 4 || wall.size() != 3 * face.size())
      throw std::runtime_error("SetEmbeddedFaces: 4 ints and 12 doubles per face");
    check(m2c_set_embedded_faces(c_.get(), (int)(face.size() / 4), face.data(), wall.data(), riemann_normal, recon),
          "m2c_set_embedded_faces");
  }
  // SpaceOperator::FindEdgeSurfaceIntersections over a list of faces: xf / xb hold, per face and
  // surface, the index of the surface's forward / backward crossing (-1 none); cut and wall (12 per
  // face, SetEmbeddedFaces' layout) are resized and filled. Host code, no device.
  static void FindEdgeSurfaceIntersections(const std::vector<m2c_surface> &surfaces, const std::vector<int> &xf,
                                           const std::vector<int> &xb, std::vector<int> &cut,
                                           std::vector<double> &wall) {
    const size_t ns = surfaces.size();
    if (ns == 0 || xf.size() != xb.size() || xf.size() % ns)
      throw std::runtime_error("FindEdgeSurfaceIntersections: xf and xb need one entry per face and surface");
    const int n = (int)(xf.size() / ns);
    cut.assign((size_t)n, 0);
    wall.assign((size_t)n * 12, 0.0);
    check(m2c_edge_surface_walls((int)ns, surfaces.data(), n, xf.data(), xb.data(), cut.data(), wall.data()),
          "m2c_edge_surface_walls");
  }

 private:
  Context &c_;
};

// MeshGenerator::ComputeMesh1DUniform / ComputeMesh1DNonUniform: the centres and widths of one axis
// (n uniform cells, or the cells of (coordinate, width) control points). Host code.
inline void ComputeMesh1D(double x0, double xmax, int n, const std::vector<std::pair<double, double>> &points,
                          std::vector<double> &x, std::vector<double> &dx) {
  std::vector<double> pts;
  for (const auto &p : points) { pts.push_back(p.first); pts.push_back(p.second); }
  int m = 0;
  check(m2c_mesh_1d(x0, xmax, n, (int)points.size(), pts.empty() ? nullptr : pts.data(), 0, nullptr, nullptr, &m),
        "m2c_mesh_1d");
  x.assign((size_t)m, 0.0);
  dx.assign((size_t)m, 0.0);
  check(m2c_mesh_1d(x0, xmax, n, (int)points.size(), pts.empty() ? nullptr : pts.data(), m, x.data(), dx.data(), &m),
        "m2c_mesh_1d");
}

class LevelSetOperator {
 public:
  explicit LevelSetOperator(Context &c) : c_(c) {}
  void SetPhi(const double *phi) { check(m2c_upload_levelset(c_.get(), phi), "m2c_upload_levelset"); }
  void GetPhi(double *phi) const { check(m2c_download_levelset(c_.get(), phi), "m2c_download_levelset"); }

 private:
  Context &c_;
};

class TimeIntegrator {
 public:
  explicit TimeIntegrator(Context &c) : c_(c) {}
  // TimeIntegratorBase::AdvanceOneTimeStep with the given dt (level set, IDs and end-of-step
  // tasks included)
  void AdvanceOneTimeStep(double dt) {
    check(m2c_advance_dt(c_.get(), dt), "AdvanceOneTimeStep");
    status_check();
  }
  // Main.cpp:514-621: CFL-adaptive steps until tmax or maxsteps (dt and t stay on the device)
  m2c_status Run(double tmax, int maxsteps, int sync_every = 64) {
    check(m2c_run(c_.get(), tmax, maxsteps, sync_every), "m2c_run");
    return status_check();
  }
  // Main.cpp:514-621 with a prescribed Time.TimeStep (SpaceOperator.cpp:2219-2221): dt each step,
  // shortened to land on tmax
  m2c_status RunFixed(double dt, double tmax, int maxsteps) {
    m2c_status s = Status();
    for (long long n = 0; s.t < tmax && n < maxsteps; n++) {
      AdvanceOneTimeStep(s.t + dt > tmax ? tmax - s.t : dt);
      s = Status();
    }
    return s;
  }
  m2c_status Status() const {
    m2c_status s;
    check(m2c_get_status(c_.get(), &s), "m2c_get_status");
    return s;
  }

 private:
  m2c_status status_check() const {
    m2c_status s = Status();
    if (s.error != M2C_OK) throw Error("non-hyperbolic state or solver failure", s.error);
    return s;
  }
  Context &c_;
};

// ExactRiemannSolverBase::ComputeRiemannSolution along any unit normal dir (the mesh axis of
// MultiPhase.RiemannNormal = Mesh, or the level-set / averaged normals of LevelSet / Average).
// Returns 0 on success, 1 on a solver failure (the reference counts it); throws on an invalid input
// state (the reference exits).
class ExactRiemannSolver {
 public:
  explicit ExactRiemannSolver(Context &c) : c_(c) {}
  int ComputeRiemannSolution(const double *dir, const double *Vm, int idm, const double *Vp, int idp, double *Vs,
                             int &id, double *Vsm, double *Vsp) {
    int st = 0;
    check(m2c_riemann_batch(c_.get(), 1, dir, Vm, &idm, Vp, &idp, Vs, &id, Vsm, Vsp, &st), "m2c_riemann_batch");
    if (st < 0) throw Error("ComputeRiemannSolution: invalid input state", M2C_ERR_STATE);
    return st == 0 ? 0 : 1;
  }

  // ExactRiemannSolverBase::ComputeOneSidedRiemannSolution: the fluid state Vm (material idl)
  // against a wall moving with velocity Ustar, along any unit normal dir from the fluid to the wall.
  // id = -1 and Vs left as it was when the wall recedes from the fluid. Returns 0 on success, 1 on a
  // solver failure (the reference counts it); throws on an invalid input state (the reference exits).
  int ComputeOneSidedRiemannSolution(const double *dir, const double *Vm, int idl, const double *Ustar, double *Vs,
                                     int &id, double *Vsm) {
    double vs[5];
    int st = 0;
    check(m2c_onesided_riemann_batch(c_.get(), 1, dir, Vm, &idl, Ustar, vs, &id, Vsm, &st),
          "m2c_onesided_riemann_batch");
    if (st < 0) throw Error("ComputeOneSidedRiemannSolution: invalid input state", M2C_ERR_STATE);
    if (id != -1)
      for (int p = 0; p < 5; p++) Vs[p] = vs[p];
    return st == 0 ? 0 : 1;
  }

 private:
  Context &c_;
};

// VarFcnBase of material `id` (VarFcnBase.h:62-294): the pointwise equation-of-state calls with the
// reference's signatures, evaluated by the device kernels on the context's material table. Each call
// is one device round trip; the *_n forms take n points at once. Where the reference exits
// (ComputeSoundSpeed with c^2 <= 0, ComputeMachNumber with c^2 < 0) these throw.
class VarFcn {
 public:
  VarFcn(Context &c, int id) : c_(c), id_(id) {
    if (id < 0 || id >= c.config().nmat) throw Error("VarFcn: material index out of range", M2C_ERR_ARG);
  }
  int GetType() const { return c_.config().mat[id_].eos; }
  double GetPressure(double rho, double e) { return scalar(M2C_VF_PRESSURE, rho, e); }
  double GetInternalEnergyPerUnitMass(double rho, double p) { return scalar(M2C_VF_INTERNAL_ENERGY, rho, p); }
  double GetDpdrho(double rho, double e) { return scalar(M2C_VF_DPDRHO, rho, e); }
  double GetBigGamma(double rho, double e) { return scalar(M2C_VF_BIG_GAMMA, rho, e); }
  double ComputeSoundSpeedSquare(double rho, double e) { return scalar(M2C_VF_SOUND_SPEED_SQUARE, rho, e); }
  double ComputeSoundSpeed(double rho, double e) {
    int32_t f = 0;
    double c = scalar(M2C_VF_SOUND_SPEED, rho, e, &f);
    if (f) throw Error("ComputeSoundSpeed: square root of a negative number", M2C_ERR_STATE);
    return c;
  }
  double ComputeEnthalpyPerUnitMass(double rho, double p) { return scalar(M2C_VF_ENTHALPY, rho, p); }
  double ComputeTotalEnthalpyPerUnitMass(const double *V) { return vec1(M2C_VF_TOTAL_ENTHALPY, V); }
  double ComputeMachNumber(const double *V) {
    int32_t f = 0;
    double m = vec1(M2C_VF_MACH, V, &f);
    if (f) throw Error("ComputeMachNumber: negative c^2", M2C_ERR_STATE);
    return m;
  }
  bool CheckState(const double *V, bool /*silence*/ = false) { return vec1(M2C_VF_CHECK_STATE, V) != 0.0; }
  bool CheckState(double rho, double p, bool /*silence*/ = false) {
    const double V[5] = {rho, 0.0, 0.0, 0.0, p};
    return vec1(M2C_VF_CHECK_STATE, V) != 0.0;
  }
  void ConservativeToPrimitive(const double *U, double *V) { vec5(M2C_VF_C2P, U, V); }
  void PrimitiveToConservative(const double *V, double *U) { vec5(M2C_VF_P2C, V, U); }
  bool ClipDensityAndPressure(double *V, double *U = nullptr) {
    int32_t f = 0;
    double out[5];
    vec5(M2C_VF_CLIP, V, out, &f);
    for (int q = 0; q < 5; q++) V[q] = out[q];
    if (f && U) PrimitiveToConservative(V, U);
    return f != 0;
  }
  // n points: in holds (rho, e) / (rho, p) pairs or 5-vectors as m2c_varfcn_batch describes
  void Evaluate_n(int op, int n, const double *in, double *out, int32_t *flag = nullptr) {
    std::vector<int32_t> ids((size_t)n, id_);
    check(m2c_varfcn_batch(c_.get(), op, n, ids.data(), in, out, flag), "m2c_varfcn_batch");
  }

 private:
  double scalar(int op, double a, double b, int32_t *flag = nullptr) {
    const double in[2] = {a, b};
    double out = 0.0;
    int32_t f = 0;
    check(m2c_varfcn_batch(c_.get(), op, 1, &id_, in, &out, &f), "m2c_varfcn_batch");
    if (flag) *flag = f;
    return out;
  }
  double vec1(int op, const double *V, int32_t *flag = nullptr) {
    double out = 0.0;
    int32_t f = 0;
    check(m2c_varfcn_batch(c_.get(), op, 1, &id_, V, &out, &f), "m2c_varfcn_batch");
    if (flag) *flag = f;
    return out;
  }
  void vec5(int op, const double *in, double *out, int32_t *flag = nullptr) {
    int32_t f = 0;
    check(m2c_varfcn_batch(c_.get(), op, 1, &id_, in, out, &f), "m2c_varfcn_batch");
    if (flag) *flag = f;
  }
  Context &c_;
  int32_t id_;
};

// FluxFcnBase (FluxFcnBase.h:65-76) of the context's numerical flux: the face flux between two
// states of one material, and the local Lax-Friedrichs flux between two materials.
class FluxFcn {
 public:
  explicit FluxFcn(Context &c) : c_(c) {}
  void ComputeNumericalFluxAtCellInterface(int dir, const double *Vm, const double *Vp, int id, double *F) {
    check(m2c_flux_batch(c_.get(), dir, 1, Vm, Vp, id, F), "m2c_flux_batch");
  }
  void ComputeNumericalFluxAtMaterialInterface(int dir, const double *Vm, int idm, const double *Vp, int idp,
                                               double *F) {
    check(m2c_interface_flux_batch(c_.get(), dir, 1, Vm, &idm, Vp, &idp, F), "m2c_interface_flux_batch");
  }

 private:
  Context &c_;
};

}  // namespace m2c_host
