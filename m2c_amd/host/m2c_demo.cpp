// m2c_demo.cpp -- a driver written against the reference-shaped host layer: the water/air
// shock-bubble (or Sod's problem) advanced on the GPU through the C ABI.
//
//   m2c_demo sod <n> <tmax> <out.bin>                    Sod 400x1x1-type problem to tmax
//   m2c_demo bubble <scale> <steps> <out.bin>            shock-bubble, RK2, reinit every 5 steps
//   m2c_demo deck <input file> <out.bin>                 any input deck: its mesh, options, initial
//                                                        state and Time.MaxTime / MaxIts / TimeStep
//   m2c_demo onesided <problems.bin> <out.bin>           one-sided (wall) Riemann problems of air
//                                                        (material 0) and water (material 1): 12
//                                                        doubles per problem (dir, V, id, wall
//                                                        velocity) in, 12 out (status, id, Vs, Vsm)
//   m2c_demo pointwise <problems.bin> <out.bin>          VarFcn, FluxFcn and ExactRiemannSolver calls on
//                                                        the materials of pointwise_materials (air SG,
//                                                        water SG, NASG, MG, JWL): 16 doubles per
//                                                        problem in, 32 out (layout at run_pointwise)
//
// Writes the primitive state (nz*ny*nx*5 doubles), then the IDs (int32) and, with a level set, the
// level set, to out.bin; prints steps and time.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "m2c_host.hpp"

using namespace m2c_host;

static int run_sod(int n, double tmax, const char *out) {
  m2c_config c = default_config(n, 1, 1, 0.0, 1.0, 0.0, 1.0 / n, 0.0, 1.0 / n);
  c.limiter = M2C_LIM_VANALBADA;
  add_stiffened_gas(c, 1.4, 0.0);
  c.bc[0] = M2C_BC_INLET;
  c.bc[1] = M2C_BC_OUTLET;
  double inlet[5] = {1.0, 0.0, 0.0, 0.0, 1.0}, outlet[5] = {0.125, 0.0, 0.0, 0.0, 0.1};
  memcpy(c.inlet, inlet, sizeof inlet);
  memcpy(c.outlet, outlet, sizeof outlet);
  Context ctx(c);
  SpaceOperator spo(ctx);
  TimeIntegrator ti(ctx);
  std::vector<double> V(5 * (size_t)n, 0.0);
  std::vector<int32_t> ID(n, 0);
  for (int i = 0; i < n; i++) {
    double x = (i + 0.5) / n;
    V[5 * i] = x < 0.5 ? 1.0 : 0.125;
    V[5 * i + 4] = x < 0.5 ? 1.0 : 0.1;
  }
  spo.SetState(V.data(), ID.data());
  m2c_status s = ti.Run(tmax, 1000000, 32);
  spo.GetState(V.data(), ID.data());
  FILE *f = fopen(out, "wb");
  if (!f) return 1;
  fwrite(V.data(), sizeof(double), V.size(), f);
  fwrite(ID.data(), sizeof(int32_t), ID.size(), f);
  fclose(f);
  printf("sod n=%d steps=%lld t=%.17g\n", n, s.steps, s.t);
  return 0;
}

static int run_bubble(double scale, int steps, const char *out) {
  // water (material 0) around a level-set sphere of air (material 1), a planar shock in the water
  int nx = (int)lround(384 * scale), ny = (int)lround(192 * scale), nz = (int)lround(192 * scale);
  double radius = 24.0 * scale, cx = 160.0 * scale, shock_x = 64.0 * scale;
  m2c_config c = default_config(nx, ny, nz, 0.0, nx, 0.0, ny, 0.0, nz);
  c.limiter = M2C_LIM_VANALBADA;
  add_stiffened_gas(c, 6.12, 3.43e8);
  add_stiffened_gas(c, 1.4, 0.0);
  c.ls_matid = 1;
  c.reinit_freq = 5;
  // post-shock water for p = 1e9 Pa from the stiffened-gas Rankine-Hugoniot relations
  const double g = 6.12, pinf = 3.43e8, rho1 = 1000.0, p1 = 1e5, p2 = 1e9;
  double P1 = p1 + pinf, P2 = p2 + pinf;
  double rho2 = rho1 * ((g + 1) * P2 + (g - 1) * P1) / ((g - 1) * P2 + (g + 1) * P1);
  double u2 = std::sqrt((p2 - p1) * (1.0 / rho1 - 1.0 / rho2));
  c.bc[0] = M2C_BC_INLET;
  c.bc[1] = M2C_BC_OUTLET;
  double inlet[5] = {rho2, u2, 0.0, 0.0, p2}, outlet[5] = {rho1, 0.0, 0.0, 0.0, p1};
  memcpy(c.inlet, inlet, sizeof inlet);
  memcpy(c.outlet, outlet, sizeof outlet);
  Context ctx(c);
  SpaceOperator spo(ctx);
  LevelSetOperator lso(ctx);
  TimeIntegrator ti(ctx);
  size_t N = (size_t)nx * ny * nz;
  std::vector<double> V(5 * N, 0.0), phi(N);
  std::vector<int32_t> ID(N);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t q = ((size_t)k * ny + j) * nx + i;
        double x = i + 0.5, y = j + 0.5, z = k + 0.5;
        phi[q] = std::sqrt((x - cx) * (x - cx) + (y - 0.5 * ny) * (y - 0.5 * ny) + (z - 0.5 * nz) * (z - 0.5 * nz)) - radius;
        ID[q] = phi[q] < 0.0 ? 1 : 0;
        double *v = &V[5 * q];
        v[0] = ID[q] ? 1.2 : rho1;
        v[4] = p1;
        if (x < shock_x) { v[0] = rho2; v[1] = u2; v[4] = p2; }
      }
  spo.SetState(V.data(), ID.data());
  lso.SetPhi(phi.data());
  m2c_status s = ti.Run(1e300, steps, 0);
  spo.GetState(V.data(), ID.data());
  lso.GetPhi(phi.data());
  FILE *f = fopen(out, "wb");
  if (!f) return 1;
  fwrite(V.data(), sizeof(double), V.size(), f);
  fwrite(ID.data(), sizeof(int32_t), ID.size(), f);
  fwrite(phi.data(), sizeof(double), phi.size(), f);
  fclose(f);
  printf("bubble %dx%dx%d steps=%lld t=%.17g reinit_iterations=%lld\n", nx, ny, nz, s.steps, s.t,
         s.reinit_iterations);
  return 0;
}

// Main.cpp's flow: read the deck, set the initial condition, march to MaxTime or MaxIts; with an
// Output object in the deck, the outputs of Main.cpp:444, :619 and :632 (forced at the start and
// the end, by Output.Frequency / TimeInterval in between)
static int run_deck(const char *path, const char *out) {
  IoData iod(path);
  for (const std::string &line : iod.ignored)
    if (line.compare(0, 14, "ignored Output") != 0) fprintf(stderr, "m2c_demo: %s\n", line.c_str());
  std::vector<double> V, phi;
  std::vector<int32_t> ID;
  iod.InitialCondition(V, ID, phi);
  Context ctx(iod.config);
  SpaceOperator spo(ctx);
  LevelSetOperator lso(ctx);
  TimeIntegrator ti(ctx);
  spo.SetState(V.data(), ID.data());
  if (iod.has_levelset()) lso.SetPhi(phi.data());
  const bool outputs = iod.text().find("Output") != std::string::npos;
  m2c_status s;
  if (!outputs) {
    s = iod.run.time_step > 0.0 ? ti.RunFixed(iod.run.time_step, iod.run.max_time, iod.run.max_its)
                                : ti.Run(iod.run.max_time, iod.run.max_its, 64);
  } else {
    Output o(iod);
    for (const std::string &line : o.not_written) fprintf(stderr, "m2c_demo: %s\n", line.c_str());
    double *ph = iod.has_levelset() ? phi.data() : nullptr;
    s = ti.Status();
    o.OutputSolutions(ctx.get(), s.t, 0.0, (int)s.steps, V.data(), ID.data(), ph, true);
    double dt = 0.0;
    while (s.t < iod.run.max_time && s.steps < iod.run.max_its) {
      if (iod.run.time_step > 0.0) {
        dt = s.t + iod.run.time_step > iod.run.max_time ? iod.run.max_time - s.t : iod.run.time_step;
        ti.AdvanceOneTimeStep(dt);
      } else {
        double cfl;           // the step's dt, as m2c_run will take it (Main.cpp:524-539)
        spo.ComputeTimeStepSize(dt, cfl);
        if (s.t + dt > iod.run.max_time) dt = iod.run.max_time - s.t;
        ti.Run(iod.run.max_time, 1, 1);
      }
      s = ti.Status();
      spo.GetState(V.data(), ID.data());
      if (ph) lso.GetPhi(ph);
      o.OutputSolutions(ctx.get(), s.t, dt, (int)s.steps, V.data(), ID.data(), ph, false);
    }
    o.OutputSolutions(ctx.get(), s.t, dt, (int)s.steps, V.data(), ID.data(), ph, true);
  }
  spo.GetState(V.data(), ID.data());
  if (iod.has_levelset()) lso.GetPhi(phi.data());
  FILE *f = fopen(out, "wb");
  if (!f) return 1;
  fwrite(V.data(), sizeof(double), V.size(), f);
  fwrite(ID.data(), sizeof(int32_t), ID.size(), f);
  fwrite(phi.data(), sizeof(double), phi.size(), f);
  fclose(f);
  printf("deck %dx%dx%d steps=%lld t=%.17g\n", iod.config.nx, iod.config.ny, iod.config.nz, s.steps, s.t);
  return 0;
}

static int run_onesided(const char *in, const char *out) {
  std::vector<double> q;
  {
    FILE *f = fopen(in, "rb");
    if (!f) return 1;
    double v;
    while (fread(&v, sizeof v, 1, f) == 1) q.push_back(v);
    fclose(f);
  }
  if (q.size() % 12) return 1;
  m2c_config c = default_config(4, 4, 4, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0);
  add_stiffened_gas(c, 1.4, 0.0);
  add_stiffened_gas(c, 6.12, 3.43e8);
  Context ctx(c);
  ExactRiemannSolver rs(ctx);
  std::vector<double> r;
  for (size_t t = 0; t < q.size(); t += 12) {
    const double *p = &q[t];
    double Vs[5] = {0, 0, 0, 0, 0}, Vsm[5];
    int id = 0;
    int st = rs.ComputeOneSidedRiemannSolution(p, p + 3, (int)p[8], p + 9, Vs, id, Vsm);
    r.push_back(st);
    r.push_back(id);
    r.insert(r.end(), Vs, Vs + 5);
    r.insert(r.end(), Vsm, Vsm + 5);
  }
  FILE *f = fopen(out, "wb");
  if (!f) return 1;
  fwrite(r.data(), sizeof(double), r.size(), f);
  fclose(f);
  printf("onesided problems=%zu\n", q.size() / 12);
  return 0;
}

// air, water (stiffened gases), NASG water, Mie-Grueneisen aluminium, JWL TNT products
static m2c_config pointwise_materials() {
  m2c_config c = default_config(4, 4, 4, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0);
  add_stiffened_gas(c, 1.4, 0.0);
  add_stiffened_gas(c, 6.12, 3.43e8);
  const int kinds[3] = {M2C_EOS_NASG, M2C_EOS_MG, M2C_EOS_JWL};
  const double params[3][6] = {{1.19, 6.217e8, 6.7212e-4, -1177788.0, 0, 0},
                               {2785.0, 5328.0, 2.0, 1.338, 0.0, 0},
                               {0.28, 3.712e11, 3.23e9, 4.15, 0.95, 1630.0}};
  for (int k = 0; k < 3; k++) {
    m2c_material &m = c.mat[c.nmat++];
    m = c.mat[0];
    m.eos = kinds[k];
    for (int q = 0; q < 8; q++) m.p[q] = q < 6 ? params[k][q] : 0.0;
  }
  return c;
}

// One problem: in = id, rho, e, p, u, v, w | dir[3], idm, idp, Vm_u, Vp_u... laid out as
//   [0] material id  [1] rho  [2] e  [3] p  [4..6] u, v, w      (VarFcn of that material)
//   [7..9] unit normal  [10] axis  [11] rho_p  [12] p_p  [13] id_p  [14..15] unused
// The Riemann problem and the fluxes pair the state (rho, u, v, w, p) of material id with
// (rho_p, -u, w, v, p_p) of material id_p. out (32 doubles):
//   [0] GetPressure(rho, e) [1] GetInternalEnergyPerUnitMass(rho, p) [2] GetDpdrho [3] GetBigGamma
//   [4] ComputeSoundSpeedSquare [5] CheckState(rho, p) [6..10] PrimitiveToConservative(V)
//   [11..15] ConservativeToPrimitive of that U [16] ComputeTotalEnthalpyPerUnitMass(V)
//   [17] status of ComputeRiemannSolution along the normal (0/1), [18] its id, [19..23] Vsm
//   [24..28] FluxFcn::ComputeNumericalFluxAtMaterialInterface along the axis (or, for equal
//   materials, ComputeNumericalFluxAtCellInterface) [29] Vsp[0] [30] Vsp[4] [31] Vs[0]
static int run_pointwise(const char *in, const char *out) {
  std::vector<double> q;
  {
    FILE *f = fopen(in, "rb");
    if (!f) return 1;
    double v;
    while (fread(&v, sizeof v, 1, f) == 1) q.push_back(v);
    fclose(f);
  }
  if (q.size() % 16) return 1;
  m2c_config c = pointwise_materials();
  Context ctx(c);
  ExactRiemannSolver rs(ctx);
  FluxFcn ff(ctx);
  std::vector<double> r;
  for (size_t t = 0; t < q.size(); t += 16) {
    const double *p = &q[t];
    const int id = (int)p[0], idp = (int)p[13], axis = (int)p[10];
    VarFcn vf(ctx, id);
    double o[32] = {0};
    const double rho = p[1], e = p[2], pr = p[3];
    o[0] = vf.GetPressure(rho, e);
    o[1] = vf.GetInternalEnergyPerUnitMass(rho, pr);
    o[2] = vf.GetDpdrho(rho, e);
    o[3] = vf.GetBigGamma(rho, e);
    o[4] = vf.ComputeSoundSpeedSquare(rho, e);
    o[5] = vf.CheckState(rho, pr) ? 1.0 : 0.0;
    const double V[5] = {rho, p[4], p[5], p[6], pr};
    vf.PrimitiveToConservative(V, o + 6);
    vf.ConservativeToPrimitive(o + 6, o + 11);
    o[16] = vf.ComputeTotalEnthalpyPerUnitMass(V);
    const double Vp[5] = {p[11], -p[4], p[6], p[5], p[12]};
    double Vs[5], Vsm[5], Vsp[5];
    int rid = 0;
    o[17] = rs.ComputeRiemannSolution(p + 7, V, id, Vp, idp, Vs, rid, Vsm, Vsp);
    o[18] = rid;
    for (int k = 0; k < 5; k++) o[19 + k] = Vsm[k];
    if (id == idp) ff.ComputeNumericalFluxAtCellInterface(axis, V, Vp, id, o + 24);
    else ff.ComputeNumericalFluxAtMaterialInterface(axis, V, id, Vp, idp, o + 24);
    o[29] = Vsp[0];
    o[30] = Vsp[4];
    o[31] = Vs[0];
    r.insert(r.end(), o, o + 32);
  }
  FILE *f = fopen(out, "wb");
  if (!f) return 1;
  fwrite(r.data(), sizeof(double), r.size(), f);
  fclose(f);
  printf("pointwise problems=%zu\n", q.size() / 16);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 4 || (strcmp(argv[1], "deck") && strcmp(argv[1], "onesided") && strcmp(argv[1], "pointwise") && argc < 5)) {
    fprintf(stderr,
            "usage: %s sod <n> <tmax> <out> | bubble <scale> <steps> <out> | deck <input> <out> | "
            "onesided <in> <out> | pointwise <in> <out>\n",
            argv[0]);
    return 2;
  }
  try {
    if (!strcmp(argv[1], "deck")) return run_deck(argv[2], argv[3]);
    if (!strcmp(argv[1], "onesided")) return run_onesided(argv[2], argv[3]);
    if (!strcmp(argv[1], "pointwise")) return run_pointwise(argv[2], argv[3]);
    if (!strcmp(argv[1], "sod")) return run_sod(atoi(argv[2]), atof(argv[3]), argv[4]);
    if (!strcmp(argv[1], "bubble")) return run_bubble(atof(argv[2]), atoi(argv[3]), argv[4]);
  } catch (const Error &e) {
    fprintf(stderr, "m2c_demo: %s\n", e.what());
    return 3;
  }
  fprintf(stderr, "unknown case %s\n", argv[1]);
  return 2;
}
