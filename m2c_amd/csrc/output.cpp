// output.cpp -- the solution outputs of an M2C input deck's Output object (host code).
//
//   snapshots   Output::WriteSolutionSnapshot (Output.cpp:147-418): <prefix><Solution>_NNNN.vtr, a
//               VTK XML rectilinear grid over the cell centres (raw appended Float64 arrays with
//               UInt64 byte counts) holding the requested fields in the reference's order, and
//               one <DataSet> line per snapshot in <prefix><Solution>.pvd (Output.cpp:28-46,
//               :392-404, same text).
//   probes      ProbeOutput (ProbeOutput.cpp:13-175, 226-339, 440-640): one text file per
//               requested variable, a line per output step with the value at every probe node,
//               trilinear in the cell-centre lattice including the face ghosts of the boundary
//               rules.
//   line plots  ProbeOutput's line form (ProbeOutput.cpp:178-214, 342-437): one text file per output
//               step with all variables at NumberOfPoints points along the segment.
// When to write follows isTimeToWrite (Utils.cpp:293-313). Temperature and internal energy come
// from the device through m2c_eos_derived (the equation of state lives in one place).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "deck.hpp"

using namespace m2c_deck;

namespace {

bool time_to_write(double time, double dt, int step, double freq_dt, int freq, double last, bool force) {
  if (force) return time > last + 0.1 * dt;
  if (freq_dt > 0) return std::floor(time / freq_dt) != std::floor((time - dt) / freq_dt);
  return freq > 0 && step % freq == 0;
}

std::string frame_name(const std::string &base, int frame, const char *ext) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "_%04d", frame);
  return base + buf + ext;
}

enum ProbeVar { P_DENSITY, P_VX, P_VY, P_VZ, P_PRESSURE, P_TEMPERATURE, P_MATERIALID, P_LEVELSET0, P_NVAR };
const char *const kProbeKeys[P_NVAR] = {"Density", "VelocityX", "VelocityY", "VelocityZ", "Pressure", "Temperature",
                                        "MaterialID", "LevelSet0"};

// trilinear stencil of one probe point (ProbeOutput::SetupInterpolation)
struct Stencil {
  int i, j, k;          // lower corner in the ghosted lattice (-1 .. N-1)
  double xi[3];
  bool valid[8];        // c000 c100 c010 c110 c001 c101 c011 c111
  int nvalid;
};

// point values the interpolation reads, the face ghosts included (boundary rules of
// SpaceOperator.cpp:1667-1745, MultiPhaseOperator.cpp:796-815, LevelSetOperator.cpp:934-1033)
struct Field {
  const m2c_config &c;
  const double *V;
  const int32_t *ID;
  const double *phi;
  const std::vector<double> *mx, *mw;   // cell centres and widths per axis (MeshGenerator's)
  // centre of index q along axis d; the ghosts one width beyond the boundary cells
  // (SpaceOperator::PopulateGhostBoundaryCoordinates, SpaceOperator.cpp:219-300)
  double xc(int d, int q) const {
    const int n = (int)mx[d].size();
    if (q < 0) return mx[d][0] - mw[d][0];
    if (q >= n) return mx[d][n - 1] + mw[d][n - 1];
    return mx[d][q];
  }
  size_t idx(int i, int j, int k) const { return ((size_t)k * c.ny + j) * c.nx + i; }
  // the outside direction of a face ghost: 2*dim + (0 low, 1 high), -1 inside
  int side(int i, int j, int k) const {
    if (i < 0) return 0;
    if (i >= c.nx) return 1;
    if (j < 0) return 2;
    if (j >= c.ny) return 3;
    if (k < 0) return 4;
    if (k >= c.nz) return 5;
    return -1;
  }
  void image(int &i, int &j, int &k) const {
    i = i < 0 ? -i - 1 : (i >= c.nx ? 2 * c.nx - i - 1 : i);
    j = j < 0 ? -j - 1 : (j >= c.ny ? 2 * c.ny - j - 1 : j);
    k = k < 0 ? -k - 1 : (k >= c.nz ? 2 * c.nz - k - 1 : k);
  }
  void state(int i, int j, int k, double *v, int &id) const {
    const int s = side(i, j, k);
    int ii = i, jj = j, kk = k;
    image(ii, jj, kk);
    id = ID[idx(ii, jj, kk)];
    if (s >= 0 && (c.bc[s] == M2C_BC_INLET || c.bc[s] == M2C_BC_OUTLET)) {
      const double *f = c.bc[s] == M2C_BC_INLET ? c.inlet : c.outlet;
      for (int p = 0; p < 5; p++) v[p] = f[p];
      return;
    }
    for (int p = 0; p < 5; p++) v[p] = V[5 * idx(ii, jj, kk) + p];
    if (s >= 0 && (c.bc[s] == M2C_BC_SYMMETRY || c.bc[s] == M2C_BC_SLIPWALL)) v[1 + s / 2] *= -1.0;
    else if (s >= 0 && c.bc[s] == M2C_BC_STICKWALL) { v[1] = -1.0 * v[1]; v[2] = -1.0 * v[2]; v[3] = -1.0 * v[3]; }
  }
  double levelset(int i, int j, int k) const {
    const int s = side(i, j, k);
    if (s < 0) return phi[idx(i, j, k)];
    const int dim = s / 2, hi = s % 2, bc = c.ls_bc[s];
    const int n[3] = {c.nx, c.ny, c.nz};
    int q[3] = {i, j, k};
    const int qg = q[dim];
    int im[3] = {i, j, k};
    im[dim] = hi ? n[dim] - 1 : 0;
    double v = phi[idx(im[0], im[1], im[2])];
    if ((bc == 2 || bc == 3) && n[dim] >= 2) {
      const int st = hi ? -1 : 1;
      int a[3] = {i, j, k}, b[3] = {i, j, k};
      a[dim] = qg + st;
      b[dim] = qg + 2 * st;
      const double r = xc(dim, qg), r1 = xc(dim, qg + st), r2 = xc(dim, qg + 2 * st);
      const double f1 = phi[idx(a[0], a[1], a[2])], f2 = phi[idx(b[0], b[1], b[2])];
      v = f1 + (f2 - f1) / (r2 - r1) * (r - r1);
    }
    if (bc == 3) {
      const double d2wall = 0.5 * std::fabs(xc(dim, im[dim]) - xc(dim, qg));
      v = std::max(-d2wall, v);
    }
    return v;
  }
};

double interpolate(const Stencil &s, double c[8]) {
  if (s.nvalid < 8) {
    const double avg = (c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7]) / s.nvalid;
    for (int q = 0; q < 8; q++)
      if (!s.valid[q]) c[q] = avg;
  }
  // MathTools/trilinear_interpolation.h:52-62
  const double *xi = s.xi;
  const double c00 = c[0] * (1.0 - xi[0]) + c[1] * xi[0];
  const double c01 = c[4] * (1.0 - xi[0]) + c[5] * xi[0];
  const double c10 = c[2] * (1.0 - xi[0]) + c[3] * xi[0];
  const double c11 = c[6] * (1.0 - xi[0]) + c[7] * xi[0];
  const double c0 = c00 * (1.0 - xi[1]) + c10 * xi[1];
  const double c1 = c01 * (1.0 - xi[1]) + c11 * xi[1];
  return c0 * (1.0 - xi[2]) + c1 * xi[2];
}

struct Line {
  double x0[3], x1[3];
  int n = 0, frequency = -100, frame = 0;
  double frequency_dt = -1.0, last = -1.0;
  std::string base;
  std::vector<Stencil> st;
};

}  // namespace

struct m2c_output {
  m2c_config cfg;
  std::vector<double> mx[3], mw[3];   // cell centres and widths of the mesh's axes
  int ls_index = -1;                 // the deck's LevelSet[i] index, -1 without a level set
  std::string prefix, base = "solution";
  int frequency = 0, frame = 0;
  double frequency_dt = -1.0, last = -1.0;
  bool f_density = false, f_velocity = false, f_pressure = false, f_materialid = false, f_energy = false,
       f_temperature = false, f_levelset[5] = {false, false, false, false, false};
  // probes
  int p_frequency = -100;
  double p_frequency_dt = -1.0, p_last = -1.0;
  std::vector<Stencil> probes;
  FILE *pfile[P_NVAR] = {};
  std::vector<Line> lines;
  bool needs_eos = false;            // temperature or internal energy is written

  ~m2c_output() {
    for (FILE *f : pfile)
      if (f) std::fclose(f);
  }
};

namespace {

// ProbeOutput::SetupInterpolation for one point; false when it lies outside the ghosted lattice
bool locate(const m2c_output &o, const double p[3], Stencil &s) {
  const m2c_config &c = o.cfg;
  const Field F{c, nullptr, nullptr, nullptr, o.mx, o.mw};
  const int n[3] = {c.nx, c.ny, c.nz};
  int ijk[3];
  for (int d = 0; d < 3; d++) {
    if (p[d] < F.xc(d, -1) || p[d] >= F.xc(d, n[d])) return false;
    int q = -1;
    while (!(p[d] < F.xc(d, q + 1))) q++;
    ijk[d] = q;
    s.xi[d] = (p[d] - F.xc(d, q)) / (F.xc(d, q + 1) - F.xc(d, q));
  }
  s.i = ijk[0]; s.j = ijk[1]; s.k = ijk[2];
  s.nvalid = 0;
  for (int q = 0; q < 8; q++) {
    const int i = s.i + (q & 1), j = s.j + ((q >> 1) & 1), k = s.k + ((q >> 2) & 1);
    // SpaceVariable.h:150-159: edge and corner ghosts are never populated
    const int out = (i < 0 || i >= c.nx) + (j < 0 || j >= c.ny) + (k < 0 || k >= c.nz);
    s.valid[q] = out <= 1;
    s.nvalid += s.valid[q];
  }
  return s.nvalid > 0;
}

std::string num_e(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%e", v);
  return b;
}

int read_output(const char *text, m2c_output &o, std::string &report) {
  Node root;
  m2c_deck_run run;
  int rc = read_config(text, root, o.cfg, run, report);
  if (rc != M2C_OK) return rc;
  if ((rc = mesh_axes(o.cfg, o.mx, o.mw)) != M2C_OK) return rc;
  report.clear();
  try {
    const Node *sp = root.find("Space");
    const Node *lsa = sp ? sp->find("LevelSet") : nullptr;
    if (lsa && !lsa->kids.empty()) o.ls_index = std::atoi(lsa->kids.begin()->first.c_str() + 1);
    const Node *out = root.find("Output");
    if (!out) return M2C_OK;
    gets(out, "Prefix", o.prefix);
    gets(out, "Solution", o.base);
    geti(out, "Frequency", o.frequency);
    getd(out, "TimeInterval", o.frequency_dt);
    auto onoff = [&](const char *key, bool &flag) {
      int v = 0;
      gett(out, key, v, {{"Off", 0}, {"On", 1}});
      flag = v == 1;
    };
    onoff("Density", o.f_density);
    onoff("Velocity", o.f_velocity);
    onoff("Pressure", o.f_pressure);
    onoff("MaterialID", o.f_materialid);
    onoff("InternalEnergyPerUnitMass", o.f_energy);
    onoff("Temperature", o.f_temperature);
    for (int l = 0; l < 5; l++) onoff(("LevelSet" + std::to_string(l)).c_str(), o.f_levelset[l]);
    for (const char *key : {"DeltaTemperature", "DeltaInternalEnergyPerUnitMass", "LaserRadiance", "ReferenceMap",
                            "MeanCharge", "HeavyParticlesDensity", "ElectronDensity"}) {
      bool f = false;
      onoff(key, f);
      if (f) report += std::string("not written: Output.") + key + "\n";
    }
    for (const char *key : {"MeshInformation", "MeshPartition", "CutPlane", "MaterialVolumes"})
      if (leaf(out, key) || out->find(key)) report += std::string("not written: Output.") + key + "\n";
    // probes
    const Node *pr = out->find("Probes");
    if (pr) {
      geti(pr, "Frequency", o.p_frequency);
      getd(pr, "TimeInterval", o.p_frequency_dt);
      const Node *nodes = pr->find("Node");
      const auto el = elements(nodes);
      for (size_t q = 0; q < el.size(); q++) {
        // ProbeOutput.cpp:24-33: node indices are 0 .. n-1
        if (!nodes->elem((long long)q)) fail(M2C_ERR_ARG, nodes->line, "probe node indices must be 0 .. n-1");
        double p[3] = {-1.0e20, -1.0e20, -1.0e20};
        getd(el[q], "X", p[0]);
        getd(el[q], "Y", p[1]);
        getd(el[q], "Z", p[2]);
        Stencil s;
        if (!locate(o, p, s)) fail(M2C_ERR_ARG, el[q]->line, "cannot locate probe node " + std::to_string(q) + " in the domain");
        o.probes.push_back(s);
      }
      std::string names[P_NVAR];
      for (int v = 0; v < P_NVAR; v++) gets(pr, kProbeKeys[v], names[v]);
      for (const char *key : {"DeltaTemperature", "LaserRadiance", "LevelSet1", "LevelSet2", "LevelSet3", "LevelSet4",
                              "IonizationResult"}) {
        std::string f;
        gets(pr, key, f);
        if (!f.empty()) report += std::string("not written: Output.Probes.") + key + "\n";
      }
      if (!o.probes.empty())
        for (int v = 0; v < P_NVAR; v++) {
          if (names[v].empty()) continue;
          const std::string path = o.prefix + names[v];
          o.pfile[v] = std::fopen(path.c_str(), "w");
          if (!o.pfile[v]) fail(M2C_ERR_ARG, 0, "cannot open " + path);
          if (v == P_TEMPERATURE) o.needs_eos = true;
          // ProbeOutput.cpp:158-166
          std::string hdr;
          for (size_t q = 0; q < el.size(); q++) {
            double p[3] = {-1.0e20, -1.0e20, -1.0e20};
            getd(el[q], "X", p[0]);
            getd(el[q], "Y", p[1]);
            getd(el[q], "Z", p[2]);
            hdr += "## Probe " + std::to_string(q) + ": " + num_e(p[0]) + ", " + num_e(p[1]) + ", " + num_e(p[2]) + "\n";
          }
          hdr += "## Time step  |  Time  |  Solutions at probe nodes (0, 1, 2, etc.)\n";
          std::fputs(hdr.c_str(), o.pfile[v]);
          std::fflush(o.pfile[v]);
        }
    }
    // line plots (Output.cpp:48-60: indices 0 .. n-1)
    const Node *lp = out->find("LinePlot");
    const auto lel = elements(lp);
    for (size_t q = 0; q < lel.size(); q++) {
      if (!lp->elem((long long)q)) fail(M2C_ERR_ARG, lp->line, "line plot indices must be 0 .. n-1");
      Line L;
      const Node *n = lel[q];
      L.x0[0] = L.x0[1] = L.x0[2] = L.x1[0] = L.x1[1] = L.x1[2] = 0.0;
      getd(n, "X0", L.x0[0]); getd(n, "Y0", L.x0[1]); getd(n, "Z0", L.x0[2]);
      getd(n, "Xmax", L.x1[0]); getd(n, "Ymax", L.x1[1]); getd(n, "Zmax", L.x1[2]);
      geti(n, "NumberOfPoints", L.n);
      geti(n, "Frequency", L.frequency);
      getd(n, "TimeInterval", L.frequency_dt);
      gets(n, "FileName", L.base);
      if (L.n == 1) fail(M2C_ERR_ARG, n->line, "must have more than 1 point for a line plot");
      for (int i = 0; i < L.n; i++) {
        double p[3];
        for (int d = 0; d < 3; d++) p[d] = L.x0[d] + i * ((L.x1[d] - L.x0[d]) / (L.n - 1));
        Stencil s;
        if (!locate(o, p, s)) fail(M2C_ERR_ARG, n->line, "cannot locate line-plot point " + std::to_string(i));
        L.st.push_back(s);
      }
      if (L.n > 1 && !L.base.empty()) o.needs_eos = true;
      o.lines.push_back(L);
    }
    if (o.f_energy || o.f_temperature) o.needs_eos = true;
  } catch (const DeckError &e) {
    report = e.msg;
    return e.code;
  }
  // materials without a temperature law (the reference exits in VarFcnBase::GetTemperature)
  bool temp = o.f_temperature || o.pfile[P_TEMPERATURE];
  for (const Line &L : o.lines) temp = temp || (L.n > 1 && !L.base.empty());
  for (int m = 0; temp && m < o.cfg.nmat; m++)
    if (o.cfg.mat[m].eos == M2C_EOS_JWL || o.cfg.mat[m].eos == M2C_EOS_DUMMY) {
      report = "temperature output requested with material " + std::to_string(m) + ", which has no temperature law";
      return M2C_ERR_ARG;
    }
  // Output.cpp:28-46: an empty collection, always (snapshots are also forced at the start and end)
  const std::string pvd = o.prefix + o.base + ".pvd";
  FILE *f = std::fopen(pvd.c_str(), "w");
  if (!f) {
    report = "cannot open " + pvd;
    return M2C_ERR_ARG;
  }
  std::fputs("<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\"\nbyte_order=\"LittleEndian\">\n"
             "  <Collection>\n  </Collection>\n</VTKFile>\n", f);
  std::fclose(f);
  return M2C_OK;
}

// e and T at every cell (or at the given points) through the device
int derived(m2c_ctx *ctx, const std::vector<double> &rho, const std::vector<double> &p, const std::vector<int32_t> &id,
            std::vector<double> &e, std::vector<double> &T) {
  e.assign(rho.size(), 0.0);
  T.assign(rho.size(), 0.0);
  if (rho.empty()) return M2C_OK;
  if (!ctx) return M2C_ERR_ARG;
  return m2c_eos_derived(ctx, (int)rho.size(), rho.data(), p.data(), id.data(), e.data(), T.data());
}

void put_array(std::string &xml, std::vector<char> &blob, const char *name, int ncomp, const double *data, size_t n) {
  xml += "        <DataArray type=\"Float64\" Name=\"" + std::string(name) + "\" NumberOfComponents=\"" +
         std::to_string(ncomp) + "\" format=\"appended\" offset=\"" + std::to_string(blob.size()) + "\" />\n";
  const unsigned long long bytes = sizeof(double) * n;
  const char *b = reinterpret_cast<const char *>(&bytes);
  blob.insert(blob.end(), b, b + sizeof bytes);
  const char *d = reinterpret_cast<const char *>(data);
  blob.insert(blob.end(), d, d + bytes);
}

int write_snapshot(m2c_output &o, m2c_ctx *ctx, double time, const double *V, const int32_t *ID, const double *phi) {
  const m2c_config &c = o.cfg;
  const size_t N = (size_t)c.nx * c.ny * c.nz;
  const Field F{c, V, ID, phi, o.mx, o.mw};
  std::vector<char> blob;
  std::string coords, fields;
  const int n[3] = {c.nx, c.ny, c.nz};
  const char *cn[3] = {"Xcoord", "Ycoord", "Zcoord"};
  for (int d = 0; d < 3; d++) {
    std::vector<double> x(n[d]);
    for (int q = 0; q < n[d]; q++) x[q] = F.xc(d, q);
    put_array(coords, blob, cn[d], 1, x.data(), x.size());
  }
  std::vector<double> s(N), v3;
  if (o.f_density) {
    for (size_t q = 0; q < N; q++) s[q] = V[5 * q];
    put_array(fields, blob, "density", 1, s.data(), N);
  }
  if (o.f_velocity) {
    v3.resize(3 * N);
    for (size_t q = 0; q < N; q++)
      for (int p = 0; p < 3; p++) v3[3 * q + p] = V[5 * q + 1 + p];
    put_array(fields, blob, "velocity", 3, v3.data(), N * 3);
  }
  if (o.f_pressure) {
    for (size_t q = 0; q < N; q++) s[q] = V[5 * q + 4];
    put_array(fields, blob, "pressure", 1, s.data(), N);
  }
  std::vector<double> e, T;
  if (o.f_energy || o.f_temperature) {
    std::vector<double> rho(N), p(N);
    std::vector<int32_t> id(ID, ID + N);
    for (size_t q = 0; q < N; q++) { rho[q] = V[5 * q]; p[q] = V[5 * q + 4]; }
    if (int rc = derived(ctx, rho, p, id, e, T)) return rc;
  }
  if (o.f_energy) put_array(fields, blob, "internal_energy", 1, e.data(), N);
  if (o.f_materialid) {
    for (size_t q = 0; q < N; q++) s[q] = ID[q];
    put_array(fields, blob, "materialid", 1, s.data(), N);
  }
  if (o.ls_index >= 0 && o.ls_index < 5 && o.f_levelset[o.ls_index] && phi) {
    char name[16];
    std::snprintf(name, sizeof name, "levelset%d", o.ls_index);
    put_array(fields, blob, name, 1, phi, N);
  }
  if (o.f_temperature) put_array(fields, blob, "temperature", 1, T.data(), N);

  const std::string fname = frame_name(o.base, o.frame, ".vtr");
  const std::string path = o.prefix + fname;
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return M2C_ERR_ARG;
  const std::string ext = "0 " + std::to_string(c.nx - 1) + " 0 " + std::to_string(c.ny - 1) + " 0 " +
                          std::to_string(c.nz - 1);
  std::string xml = "<?xml version=\"1.0\"?>\n"
                    "<VTKFile type=\"RectilinearGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
                    "header_type=\"UInt64\">\n"
                    "  <RectilinearGrid WholeExtent=\"" + ext + "\">\n"
                    "    <Piece Extent=\"" + ext + "\">\n"
                    "      <Coordinates>\n" + coords + "      </Coordinates>\n"
                    "      <PointData>\n" + fields + "      </PointData>\n"
                    "    </Piece>\n"
                    "  </RectilinearGrid>\n"
                    "  <AppendedData encoding=\"raw\">\n_";
  std::fwrite(xml.data(), 1, xml.size(), f);
  std::fwrite(blob.data(), 1, blob.size(), f);
  const char *tail = "\n  </AppendedData>\n</VTKFile>\n";
  std::fwrite(tail, 1, std::strlen(tail), f);
  std::fclose(f);

  // Output.cpp:392-404: replace the closing lines of the collection
  const std::string pvd = o.prefix + o.base + ".pvd";
  FILE *pf = std::fopen(pvd.c_str(), "r+");
  if (!pf) return M2C_ERR_ARG;
  std::fseek(pf, -27, SEEK_END);
  std::fprintf(pf, "  <DataSet timestep=\"%e\" file=\"%s\"/>\n", time, fname.c_str());
  std::fputs("  </Collection>\n</VTKFile>\n", pf);
  std::fclose(pf);
  o.frame++;
  o.last = time;
  return M2C_OK;
}

// the values of one variable at the corners of a stencil (ghosts by the boundary rules)
struct Corner {
  double V[8][5];
  int id[8];
  double phi[8];
};

Corner corners(const Field &F, const Stencil &s, bool with_phi) {
  Corner c;
  for (int q = 0; q < 8; q++) {
    const int i = s.i + (q & 1), j = s.j + ((q >> 1) & 1), k = s.k + ((q >> 2) & 1);
    if (!s.valid[q]) {
      for (int p = 0; p < 5; p++) c.V[q][p] = 0.0;
      c.id[q] = 0;
      c.phi[q] = 0.0;
      continue;
    }
    F.state(i, j, k, c.V[q], c.id[q]);
    c.phi[q] = with_phi ? F.levelset(i, j, k) : 0.0;
  }
  return c;
}

double interp_var(const Stencil &s, const Corner &c, int var, const double *Tc) {
  double v[8];
  for (int q = 0; q < 8; q++) {
    double x = 0.0;
    if (s.valid[q]) {
      if (var <= P_PRESSURE) x = c.V[q][var];
      else if (var == P_TEMPERATURE) x = Tc[q];
      else if (var == P_MATERIALID) x = c.id[q];
      else x = c.phi[q];
    }
    v[q] = x;
  }
  return interpolate(s, v);
}

// temperatures at the valid corners of every stencil, in one device call
int corner_temperatures(m2c_ctx *ctx, const std::vector<Corner> &cs, const std::vector<Stencil> &st,
                        std::vector<double> &T) {
  std::vector<double> rho, p;
  std::vector<int32_t> id;
  for (size_t s = 0; s < cs.size(); s++)
    for (int q = 0; q < 8; q++)
      if (st[s].valid[q]) {
        rho.push_back(cs[s].V[q][0]);
        p.push_back(cs[s].V[q][4]);
        id.push_back(cs[s].id[q]);
      }
  std::vector<double> e, Tv;
  if (int rc = derived(ctx, rho, p, id, e, Tv)) return rc;
  T.assign(8 * cs.size(), 0.0);
  size_t k = 0;
  for (size_t s = 0; s < cs.size(); s++)
    for (int q = 0; q < 8; q++)
      if (st[s].valid[q]) T[8 * s + q] = Tv[k++];
  return M2C_OK;
}

int write_probes(m2c_output &o, m2c_ctx *ctx, double time, int step, const double *V, const int32_t *ID,
                 const double *phi) {
  const Field F{o.cfg, V, ID, phi, o.mx, o.mw};
  std::vector<Corner> cs;
  for (const Stencil &s : o.probes) cs.push_back(corners(F, s, phi != nullptr));
  std::vector<double> T;
  if (o.pfile[P_TEMPERATURE])
    if (int rc = corner_temperatures(ctx, cs, o.probes, T)) return rc;
  for (int v = 0; v < P_NVAR; v++) {
    FILE *f = o.pfile[v];
    if (!f) continue;
    if (v == P_LEVELSET0 && !phi) continue;
    // ProbeOutput.cpp:453 (density: %10d) and :463-609 (the others: %8d)
    std::fprintf(f, v == P_DENSITY ? "%10d    %16.8e    " : "%8d    %16.8e    ", step, time);
    for (size_t q = 0; q < o.probes.size(); q++)
      std::fprintf(f, "%16.8e    ", interp_var(o.probes[q], cs[q], v, T.empty() ? nullptr : &T[8 * q]));
    std::fputs("\n", f);
    std::fflush(f);
  }
  o.p_last = time;
  return M2C_OK;
}

int write_line(m2c_output &o, Line &L, m2c_ctx *ctx, double time, int step, const double *V, const int32_t *ID,
               const double *phi) {
  if (L.base.empty()) return M2C_OK;
  const Field F{o.cfg, V, ID, phi, o.mx, o.mw};
  std::vector<Corner> cs;
  for (const Stencil &s : L.st) cs.push_back(corners(F, s, phi != nullptr));
  std::vector<double> T;
  if (int rc = corner_temperatures(ctx, cs, L.st, T)) return rc;
  const double dx = (L.x1[0] - L.x0[0]) / (L.n - 1), dy = (L.x1[1] - L.x0[1]) / (L.n - 1),
               dz = (L.x1[2] - L.x0[2]) / (L.n - 1);
  const double h = std::sqrt(dx * dx + dy * dy + dz * dz);
  const std::string path = o.prefix + frame_name(L.base, L.frame, ".txt");
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) return M2C_ERR_ARG;
  // ProbeOutput.cpp:371-385
  std::fprintf(f, "## Line: (%e, %e, %e) -> (%e, %e, %e)\n", L.x0[0], L.x0[1], L.x0[2], L.x1[0], L.x1[1], L.x1[2]);
  std::fprintf(f, "## Number of points: %d (h = %e)\n", L.n, h);
  std::fprintf(f, "## Time: %e, Time step: %d.\n", time, step);
  std::fputs("## Coordinate  |  Density  |  Velocity (Vx,Vy,Vz)  |  Pressure  |  Temperature  |  Material ID  "
             "|  LevelSet(s)\n", f);
  for (int q = 0; q < L.n; q++) {
    const Stencil &s = L.st[q];
    const double *Tq = &T[8 * q];
    std::fprintf(f, "%16.8e  %16.8e  %16.8e  %16.8e  %16.8e  %16.8e  %16.8e  %16.8e", q * h,
                 interp_var(s, cs[q], P_DENSITY, Tq), interp_var(s, cs[q], P_VX, Tq), interp_var(s, cs[q], P_VY, Tq),
                 interp_var(s, cs[q], P_VZ, Tq), interp_var(s, cs[q], P_PRESSURE, Tq),
                 interp_var(s, cs[q], P_TEMPERATURE, Tq), interp_var(s, cs[q], P_MATERIALID, Tq));
    if (phi) std::fprintf(f, "%16.8e  ", interp_var(s, cs[q], P_LEVELSET0, Tq));
    std::fputs("\n", f);
  }
  std::fclose(f);
  L.frame++;
  L.last = time;
  return M2C_OK;
}

void put_msg(const std::string &report, char *msg, int msglen) {
  if (msg && msglen > 0) {
    std::strncpy(msg, report.c_str(), (size_t)msglen - 1);
    msg[msglen - 1] = 0;
  }
}

}  // namespace

extern "C" {

int m2c_output_create(const char *text, m2c_output **out, char *msg, int msglen) {
  if (msg && msglen > 0) msg[0] = 0;
  if (!text || !out) return M2C_ERR_ARG;
  *out = nullptr;
  m2c_output *o = new m2c_output();
  std::string report;
  const int rc = read_output(text, *o, report);
  put_msg(report, msg, msglen);
  if (rc != M2C_OK) {
    delete o;
    return rc;
  }
  *out = o;
  return M2C_OK;
}

int m2c_output_needs_device(const m2c_output *o) { return o && o->needs_eos ? 1 : 0; }

int m2c_output_write(m2c_output *o, m2c_ctx *ctx, double time, double dt, int time_step, const double *V,
                     const int32_t *ID, const double *phi, int force_write, int *written) {
  if (written) *written = 0;
  if (!o || !V || !ID) return M2C_ERR_ARG;
  if (o->ls_index >= 0 && !phi) return M2C_ERR_ARG;
  int w = 0;
  if (time_to_write(time, dt, time_step, o->frequency_dt, o->frequency, o->last, force_write != 0)) {
    if (int rc = write_snapshot(*o, ctx, time, V, ID, phi)) return rc;
    w |= 1;
  }
  if (!o->probes.empty() &&
      time_to_write(time, dt, time_step, o->p_frequency_dt, o->p_frequency, o->p_last, force_write != 0)) {
    if (int rc = write_probes(*o, ctx, time, time_step, V, ID, phi)) return rc;
    w |= 2;
  }
  for (Line &L : o->lines)
    if (L.n > 1 && time_to_write(time, dt, time_step, L.frequency_dt, L.frequency, L.last, force_write != 0)) {
      if (int rc = write_line(*o, L, ctx, time, time_step, V, ID, phi)) return rc;
      w |= 4;
    }
  if (written) *written = w;
  return M2C_OK;
}

int m2c_output_destroy(m2c_output *o) {
  delete o;
  return M2C_OK;
}

}  // extern "C"
