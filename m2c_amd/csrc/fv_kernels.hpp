// fv_kernels.hpp -- the fused finite-volume stage kernel and its helpers (gfx950).
//
// One launch performs one Runge-Kutta stage of SpaceOperator::ComputeResidual followed by the
// stage combination of TimeIntegratorRK2/RK3/FE (TimeIntegrator.cpp:150-226, 291-428, 496-690):
//
//   ghost values      produced on the fly from the boundary rules (SpaceOperator.cpp:1667-1745,
//                     Reconstructor.cpp:795-875, MultiPhaseOperator.cpp:796-819)
//   MUSCL             Reconstructor.cpp:423-779, limiters Reconstructor.h:95-110
//   face-state check  Reconstructor.cpp:719-775 + SpaceOperator.cpp:3198-3347 (a face state that
//                     fails CheckState or would be clipped reverts to the cell value)
//   face flux         FluxFcn*::ComputeNumericalFluxAtCellInterface, once per face
//   accumulation      SpaceOperator.cpp:2423-2815 order (west, south, back subtracted, then east,
//                     north, front added), then *1/(-vol) as Vector5D.h:174-185
//   stage combine     U = (a*U_s + b*U_n) + c*dt*R, V = C2P(U), clip, CheckState
//
// Layout: a 32x8 block of (x,y) columns marches up a chunk of z planes. Plane k of the stencil
// field, with a 2-cell halo in x and y, lives in LDS (structure of arrays); each thread keeps
// V(k..k+2) of its own column in registers for the z direction. Slopes are computed once per
// cell and direction, each face flux once per face (faces on a tile's x/y edge are also computed
// by the neighbouring tile).
#pragma once
#include "physics.hpp"

namespace m2c {

constexpr int TX = 32, TY = 8, HALO = 2;
constexpr int LX = TX + 2 * HALO, LY = TY + 2 * HALO;
constexpr int NTHR = TX * TY;
constexpr int NSX = (TX + 2) * TY;             // x slopes: x in [-1, TX+1), y in [0, TY)
constexpr int NSY = TX * (TY + 2);             // y slopes: x in [0, TX), y in [-1, TY+1)
constexpr int NSL = NSX + NSY;
constexpr int NFX = (TX + 1) * TY;             // x faces: x in [0, TX], y in [0, TY)
constexpr int NFY = TX * (TY + 1);             // y faces: x in [0, TX), y in [0, TY]
constexpr int NHALO = 2 * HALO * TY + 2 * HALO * TX;
constexpr int SLOPE_ROUNDS = (NSL + NTHR - 1) / NTHR;

// A tile of TX_ x TY_ columns (NTHR threads, one column each) and its LDS extents; the constants
// above are the 32 x 8 tile of the single-instance kernel, the batched kernel also runs 16 x 16.
template <int TX_, int TY_> struct TileShape {
  static_assert(TX_ * TY_ == NTHR, "one thread per column");
  static constexpr int TX = TX_, TY = TY_, LX = TX_ + 2 * HALO, LY = TY_ + 2 * HALO;
  static constexpr int NSX = (TX_ + 2) * TY_, NSY = TX_ * (TY_ + 2), NSL = NSX + NSY;
  static constexpr int NFX = (TX_ + 1) * TY_, NFY = TX_ * (TY_ + 1);
  static constexpr int NHALO = 2 * HALO * TY_ + 2 * HALO * TX_;
  static_assert(NHALO <= NTHR, "one halo cell per thread");
  static constexpr int SLOPE_ROUNDS = (NSL + NTHR - 1) / NTHR;
};
using DefaultTile = TileShape<TX, TY>;
#define M2C_TILE_CONSTANTS(TS)                                                                 \
  constexpr int TX = TS::TX, TY = TS::TY, LX = TS::LX, LY = TS::LY, NSX = TS::NSX, NSL = TS::NSL, \
                NFX = TS::NFX, NFY = TS::NFY, NHALO = TS::NHALO, SLOPE_ROUNDS = TS::SLOPE_ROUNDS;   \
  (void)LX; (void)LY; (void)NSX; (void)NSL; (void)NFX; (void)NFY; (void)NHALO; (void)SLOPE_ROUNDS

enum { BC_INLET = 1, BC_OUTLET = 2, BC_SLIPWALL = 3, BC_STICKWALL = 4, BC_SYMMETRY = 5, BC_OVERSET = 6,
       BC_PERIODIC = 7 };

// status bits written by the kernels
enum { ST_BAD_STATE = 1u, ST_CLIP_INTERMEDIATE = 2u, ST_NAN_FLUX = 4u, ST_NEG_C2 = 8u, ST_RIEMANN = 16u,
       ST_RIEMANN_PATH = 32u, ST_LIST_OVERFLOW = 64u, ST_REINIT_FAIL = 128u };

struct Grid {
  int nx, ny, nz;
  long long N;                 // nx*ny*nz
  int bc[6];
  double inlet[5], outlet[5];
  double dx, dy, dz, ax, ay, az, inv_mvol;
  double A[3], B[3];
  int K[3];
  int flux, limiter, rec_linear, slope_zero;
  double alpha, roe_delta;
  double conrec_depth;         // |phi(t_n)| below this: constant reconstruction (0: none)
  int inactive_id, nmat, mp_flux, mp_recon;
  int has_interfaces;          // more than one active material
  // stretched mesh (0: the uniform values above hold): per axis d the cell widths w[d][i + 2] for
  // i in [-2, n+1] (a ghost takes its boundary cell's width, SpaceOperator.cpp:219-300) and the
  // limiter coefficients cA, cB, cK of the cells 0..n-1 (Reconstructor.cpp:53-121). Read by the
  // run-time-EOS variants of the kernels; a stretched mesh selects one of them.
  int stretched;
  const double *w[3];
  const double *cA[3], *cB[3];
  const int *cK[3];
};

struct Status {
  unsigned int bits;
  unsigned int pad;
  unsigned long long reverted, clipped, riemann_err;
};

// device-side run control of the time loop (Main.cpp:514-621)
struct Clock {
  double t, dt, tmax, cfl;
  unsigned long long minbits;  // running min of dx/(|u|+c) as ordered bits of a positive double
  int steps, maxsteps, active, pad;
};

struct StageArgs {
  const double *S;             // stencil field (V if IN_CONS==0, U_s otherwise), SoA 5*N
  const double *Vn;            // primitive state at t_n (for U_n = P2C(V_n))
  const double *Us;            // previous stage conservative state
  double *Out;                 // U_out, V_out or R (SoA 5*N)
  const int8_t *ID;
  const double *Fi;            // interface fluxes, 3 directions x 10 x Nf (may be null)
  long long Nf;                // faces per direction in Fi: (nx+1)*(ny+1)*(nz+1)
  Clock *clk;
  Status *st;
  double a, b, c;              // U = (a*U_s + b*U_n) + (c*dt)*R
  int first;                   // 1: first stage (U = U_n + dt*R)
  int kchunk;
  int ignore_active;           // 1: run even if the clock is inactive (direct API calls)
  int no_cfl;                  // 1: the final stage does not feed the CFL minimum (it is taken later)
  // viscous and heat-diffusion terms per face (null: none). Dv[d]: 10 components x faces of
  // direction d (0-4 added to the plus cell, 5-9 subtracted from the minus cell); Dh[d]: the heat
  // flux per face (energy only). Faces of direction d are indexed (k, j, i) with n_d + 1 planes
  // along d.
  const double *Dv[3] = {nullptr, nullptr, nullptr};
  const double *Dh[3] = {nullptr, nullptr, nullptr};
  // Godunov fluxes of the faces between two cells of one active material (FLUX_GODUNOV only):
  // G[d][p * n_d + f], faces indexed as Dv
  const double *G[3] = {nullptr, nullptr, nullptr};
  // the level set at t_n when a constant-reconstruction depth is set (null otherwise)
  const double *phin = nullptr;
  // faces cut by embedded walls (null: none): bit d of Wm[face key] marks the d-face whose plus
  // cell has that key; its two fluxes are read from Fi like an interface face's
  const uint8_t *Wm = nullptr;
};

// Each translation unit that launches kernels holds its own copy of these tables (internal
// linkage) and exports an upload function for it. The batched translation unit (M2C_BATCHED)
// reads the grid of the instance a block is working on from LDS instead; the multi-instance
// interface unit (M2C_IFACE_MULTI) reads grid and materials from LDS copies of its instance's.
#ifdef M2C_IFACE_MULTI
static __shared__ DevEos s_ieos[8];
#define M2C_EOS s_ieos
#else
static __constant__ DevEos c_eos[8];
#define M2C_EOS c_eos
#endif
#ifdef M2C_BATCHED
static __shared__ Grid s_bgrid;
#define M2C_G s_bgrid
#define M2C_BATCHED_LAYOUT 1
#elif defined(M2C_IFACE_MULTI)
static __shared__ Grid s_igrid;
#define M2C_G s_igrid
#else
static __constant__ Grid c_grid;
#define M2C_G c_grid
#endif

// Material constants of a cell with ID id. The stiffened-gas variants (one material, or two with
// IDs 0 and 1) read only the seven constants assigned below; they come from uniform loads of the
// table and a select by ID instead of a per-lane load of the entry, which would put a memory
// latency inside every face-state check and flux.
template <int EK, bool MP> __device__ __forceinline__ decltype(auto) mat_of(int id) {
  if constexpr (EK == EOS_SG) {
    DevEos m{};
    const DevEos &a = M2C_EOS[0];
    if constexpr (MP) {
      const DevEos &b = M2C_EOS[1];
      const bool o = id != 0;
      m.rhomin = o ? b.rhomin : a.rhomin; m.pmin = o ? b.pmin : a.pmin;
      m.rhomax = o ? b.rhomax : a.rhomax; m.pmax = o ? b.pmax : a.pmax;
      m.pstiff = o ? b.pstiff : a.pstiff; m.gam1 = o ? b.gam1 : a.gam1;
      m.gam_pstiff = o ? b.gam_pstiff : a.gam_pstiff;
    } else {
      m.rhomin = a.rhomin; m.pmin = a.pmin; m.rhomax = a.rhomax; m.pmax = a.pmax;
      m.pstiff = a.pstiff; m.gam1 = a.gam1; m.gam_pstiff = a.gam_pstiff;
    }
    m.type = EOS_SG;
    return m;
  } else {
    return (M2C_EOS[id]);
  }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// image index of a ghost cell (SpaceOperator.cpp:449-480); periodic wraps
__device__ __forceinline__ int image_idx(int i, int n, bool per) {
  if (i < 0) return per ? ((i % n) + n) % n : -i - 1;
  if (i >= n) return per ? i % n : n + (i - n) - 1;
  return i;
}

// A cell fetch in two halves: fetch_begin issues the loads of the (possibly ghost) cell (i,j,k),
// fetch_end turns what arrived into the stencil value (ConservativeToPrimitive of a conservative
// stage field, the boundary rule's sign flips), so that a prefetch can be converted once its loads
// have landed. `inner`: the caller knows (i,j,k) lies in the box (no ghost rules apply).
struct Fetched {
  double raw[5];
  int id, bc, side;
};

__device__ __forceinline__ void fetch_begin(const double *__restrict__ S, const int8_t *__restrict__ ID, int i,
                                            int j, int k, Fetched &f, bool inner = false) {
  const Grid &g = M2C_G;
  long long idx;
  int bc = 0, side = -1;
  if (inner) {
    idx = ((long long)k * g.ny + j) * g.nx + i;
  } else {
    int oi = (i < 0) ? 0 : (i >= g.nx ? 1 : -1);
    int oj = (j < 0) ? 2 : (j >= g.ny ? 3 : -1);
    int ok = (k < 0) ? 4 : (k >= g.nz ? 5 : -1);
    int nout = (oi >= 0) + (oj >= 0) + (ok >= 0);
    side = oi >= 0 ? oi : (oj >= 0 ? oj : ok);
    bc = (nout == 1) ? g.bc[side] : 0;
    int ii = clampi(image_idx(i, g.nx, oi >= 0 && g.bc[oi] == BC_PERIODIC), 0, g.nx - 1);
    int jj = clampi(image_idx(j, g.ny, oj >= 0 && g.bc[oj] == BC_PERIODIC), 0, g.ny - 1);
    int kk = clampi(image_idx(k, g.nz, ok >= 0 && g.bc[ok] == BC_PERIODIC), 0, g.nz - 1);
    idx = ((long long)kk * g.ny + jj) * g.nx + ii;
  }
  f.id = ID[idx];
  f.bc = bc;
  f.side = side;
  if (bc == BC_INLET || bc == BC_OUTLET) {
    const double *s = (bc == BC_INLET) ? g.inlet : g.outlet;
#pragma unroll
    for (int p = 0; p < 5; p++) f.raw[p] = s[p];
    return;
  }
#pragma unroll
  for (int p = 0; p < 5; p++) f.raw[p] = S[p * g.N + idx];
}

template <int IN_CONS, int EK, bool MP = (EK < 0)>
__device__ __forceinline__ void fetch_end(const Fetched &f, double *V, unsigned &flags) {
  const int bc = f.bc, side = f.side;
  if (bc == BC_INLET || bc == BC_OUTLET) {
#pragma unroll
    for (int p = 0; p < 5; p++) V[p] = f.raw[p];
    return;
  }
  if (IN_CONS) {
    const auto &m = mat_of<EK, MP>(f.id);
    eos_c2p<EK>(m, f.raw, V);
    if (eos_clip<double, EK>(m, V)) flags |= ST_CLIP_INTERMEDIATE;
  } else {
#pragma unroll
    for (int p = 0; p < 5; p++) V[p] = f.raw[p];
  }
  if (bc == BC_SYMMETRY || bc == BC_SLIPWALL) {
    if (side < 2) V[1] *= -1.0;
    else if (side < 4) V[2] *= -1.0;
    else V[3] *= -1.0;
  } else if (bc == BC_STICKWALL) {
    V[1] = -1.0 * V[1]; V[2] = -1.0 * V[2]; V[3] = -1.0 * V[3];
  }
}

// Load the stencil value of (possibly ghost) cell (i,j,k).
template <int IN_CONS, int EK, bool MP = (EK < 0)>
__device__ __forceinline__ void fetch_cell(const double *__restrict__ S, const int8_t *__restrict__ ID, int i,
                                           int j, int k, double *V, int &id, unsigned &flags, bool inner = false) {
  Fetched f;
  fetch_begin(S, ID, i, j, k, f, inner);
  fetch_end<IN_CONS, EK, MP>(f, V, flags);
  id = f.id;
}

// Limiter coefficients A, B, K of the cell with index c along DIM (Reconstructor.cpp:53-121): the
// uniform mesh's (A = B = 1), or with STR and a stretched mesh the cell's own (a cell beyond the box
// takes its boundary cell's; such a slope never reaches a face state, the boundary rules give the
// ghost's face value)
struct SlopeCoef {
  double A, B;
  int K;
};
template <int DIM, bool STR> __device__ __forceinline__ SlopeCoef slope_coef(int c) {
  const Grid &g = M2C_G;
  if (STR && g.stretched) {
    const int n = DIM == 0 ? g.nx : (DIM == 1 ? g.ny : g.nz);
    const int q = clampi(c, 0, n - 1);
    return {g.cA[DIM][q], g.cB[DIM][q], g.cK[DIM][q]};
  }
  return {g.A[DIM], g.B[DIM], g.K[DIM]};
}

template <typename TC> __device__ __forceinline__ TC limited_slope(TC dq0, TC dq1, const SlopeCoef &cf) {
  const Grid &g = M2C_G;
  switch (g.limiter) {
  case LIM_GENMINMOD: {
    TC A = TC(cf.A), B = TC(cf.B), al = TC(g.alpha);
    if (dq0 < TC(0) && dq1 < TC(0)) return smax(al * dq0, smax(B / (A + TC(1)) * (dq0 + dq1), al * dq1));
    if (dq0 > TC(0) && dq1 > TC(0)) return smin(al * dq0, smin(B / (A + TC(1)) * (dq0 + dq1), al * dq1));
    return TC(0);
  }
  case LIM_VANALBADA: {
    if (dq0 * dq1 <= TC(0)) return TC(0);
    if (dabs(dq0 * dq1) < TC(1e-16)) return TC(0);
    TC d0k = dq0, d1k = dq1;
    for (int t = 1; t < cf.K; t++) { d0k *= dq0; d1k *= dq1; }
    return TC(cf.B) * (dq1 * d0k + dq0 * d1k) / (d0k + TC(cf.A) * d1k);
  }
  default: return TC(0.5) * (dq0 + dq1);
  }
}

// SpaceOperator::TagNodesOutsideConRecDepth (SpaceOperator.cpp:2861-2920): the (possibly ghost)
// cell (i,j,k), taken at its image as the stencil fetch takes it, reconstructs constantly when
// |phi(t_n)| there is below the depth.
__device__ __forceinline__ bool conrec_cell(const double *phin, int i, int j, int k) {
  const Grid &g = M2C_G;
  if (!phin) return false;
  const int ii = clampi(image_idx(i, g.nx, (i < 0 ? g.bc[0] : g.bc[1]) == BC_PERIODIC), 0, g.nx - 1);
  const int jj = clampi(image_idx(j, g.ny, (j < 0 ? g.bc[2] : g.bc[3]) == BC_PERIODIC), 0, g.ny - 1);
  const int kk = clampi(image_idx(k, g.nz, (k < 0 ? g.bc[4] : g.bc[5]) == BC_PERIODIC), 0, g.nz - 1);
  return __builtin_fabs(phin[((long long)kk * g.ny + jj) * g.nx + ii]) < g.conrec_depth;
}

template <typename TC> __device__ __forceinline__ void zero5(TC *s) {
#pragma unroll
  for (int p = 0; p < 5; p++) s[p] = TC(0);
}

// Slopes of the cell with index c along DIM from its two neighbours (STR: stretched meshes served).
template <int DIM, typename TC, bool STR = false>
__device__ __forceinline__ void cell_slopes(const TC *vl, const TC *vc, const TC *vr, int idl, int idc, int idr,
                                            TC *sig, int c) {
  const Grid &g = M2C_G;
  if (!g.rec_linear || idc == g.inactive_id || (g.slope_zero && (idc != idl || idc != idr))) {
#pragma unroll
    for (int p = 0; p < 5; p++) sig[p] = TC(0);
    return;
  }
  const SlopeCoef cf = slope_coef<DIM, STR>(c);
  bool zl = (idl == g.inactive_id), zr = (idr == g.inactive_id);
#pragma unroll
  for (int p = 0; p < 5; p++) {
    TC dq0 = zl ? TC(0) : vc[p] - vl[p];
    TC dq1 = zr ? TC(0) : vr[p] - vc[p];
    sig[p] = limited_slope<TC>(dq0, dq1, cf);
  }
}

// Face state v +/- sigma/2 with the revert rule; returns 1 if reverted.
template <int EK, bool PLUS, typename TC, bool MP = (EK < 0)>
__device__ __forceinline__ int face_state(const TC *vc, const TC *sig, int id, TC *out) {
#pragma unroll
  for (int p = 0; p < 5; p++) out[p] = PLUS ? vc[p] + TC(0.5) * sig[p] : vc[p] - TC(0.5) * sig[p];
  const auto &m = mat_of<EK, MP>(id);
  if (id != M2C_G.inactive_id && (eos_bad<TC, EK>(m, out) || eos_would_clip<TC, EK>(m, out))) {
#pragma unroll
    for (int p = 0; p < 5; p++) out[p] = vc[p];
    return 1;
  }
  return 0;
}

// Face state of a ghost cell from the boundary rule (Reconstructor.cpp:820-869).
template <int DIM, typename TC>
__device__ __forceinline__ void ghost_face(int bc, const TC *img_face, const TC *vghost, TC *out) {
  if (bc == BC_SYMMETRY || bc == BC_SLIPWALL) {
#pragma unroll
    for (int p = 0; p < 5; p++) out[p] = (p == DIM + 1) ? -img_face[p] : img_face[p];
  } else if (bc == BC_STICKWALL) {
#pragma unroll
    for (int p = 0; p < 5; p++) out[p] = (p >= 1 && p <= 3) ? -img_face[p] : img_face[p];
  } else {
#pragma unroll
    for (int p = 0; p < 5; p++) out[p] = vghost[p];
  }
}

template <int FLUX, int EK, int DIR, typename TC = double, bool MP = (EK < 0)>
__device__ __forceinline__ void numerical_flux(const TC *Vm, const TC *Vp, int id, TC *F) {
  const auto &m = mat_of<EK, MP>(id);
  if (FLUX == FLUX_HLLC) flux_hllc<TC, EK, DIR>(m, Vm, Vp, F);
  else if (FLUX == FLUX_ROE) flux_roe<TC, EK, DIR>(m, Vm, Vp, TC(M2C_G.roe_delta), F);
  else if (FLUX == FLUX_GODUNOV) {   // the Godunov pass supplies these faces (G arrays)
#pragma unroll
    for (int p = 0; p < 5; p++) F[p] = TC(0);
  } else flux_llf2<TC, EK, EK, DIR>(m, m, Vm, Vp, F);
}

// Flux of one face between cells with IDs idm (minus) and idp (plus). Interface faces solved by
// the exact Riemann pass return 1 (their two fluxes are read from the interface-flux arrays).
template <int FLUX, int EK, int DIR, typename TC, bool MP = (EK < 0)>
__device__ __forceinline__ int face_flux(const TC *Vm, const TC *Vp, const TC *cm, const TC *cp, int idm, int idp,
                                         TC *F) {
  const Grid &g = M2C_G;
  if (idm == idp && idp != g.inactive_id) {
    numerical_flux<FLUX, EK, DIR, TC, MP>(Vm, Vp, idp, F);
    return 0;
  }
#pragma unroll
  for (int p = 0; p < 5; p++) F[p] = TC(0);
  if (MP && idm != idp && idm != g.inactive_id && idp != g.inactive_id) {
    if (g.mp_flux == 2) {   // LLF at the material interface (SpaceOperator.cpp:2493-2501)
      const TC *am = g.mp_recon ? Vm : cm, *ap = g.mp_recon ? Vp : cp;
      flux_llf2<TC, -1, -1, DIR>(M2C_EOS[idm], M2C_EOS[idp], am, ap, F);
      return 0;
    }
    return 1;
  }
  return 0;
}

__device__ __forceinline__ bool is_interface(int idm, int idp) {
  const Grid &g = M2C_G;
  return idm != idp && idm != g.inactive_id && idp != g.inactive_id && g.mp_flux != 2;
}

// face (i-1/2, j, k) of direction 0 (and likewise for y, z) is keyed by its plus cell (i,j,k)
__device__ __forceinline__ long long face_key(int i, int j, int k) {
  const Grid &g = M2C_G;
  return ((long long)k * (g.ny + 1) + j) * (g.nx + 1) + i;
}

// 1 when the d-face whose plus cell is (i, j, k) is cut by an embedded wall; faces outside the
// key range (ghost cells) never are
__device__ __forceinline__ unsigned wall_cut(const uint8_t *W, int d, int i, int j, int k) {
  const Grid &g = M2C_G;
  if (i < 0 || j < 0 || k < 0 || i > g.nx || j > g.ny || k > g.nz) return 0u;
  return ((unsigned)W[face_key(i, j, k)] >> d) & 1u;
}

__device__ __forceinline__ void warp_min_atomic(unsigned long long *dst, double v) {
  // wavefront min (64 lanes), then one atomic per wave
  unsigned long long b = __double_as_longlong(v);
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(b, off, 64);
    b = o < b ? o : b;
  }
  if ((threadIdx.x & 63) == 0) atomicMin(dst, b);
}

// x (DIR 0) or y (DIR 1) face of the tile whose plus cell is at local (x, y); the flux goes to
// sF[p*NF + fi].
template <int FLUX, int EK, int DIR, typename TC, bool MP = (EK < 0), class TS = DefaultTile>
__device__ __forceinline__ void tile_face(TC (*sV)[TS::LY][TS::LX], int (*sID)[TS::LX], const TC *sS, TC *sF, int x,
                                          int y, int fi, int i0, int j0, unsigned long long &nrev) {
  M2C_TILE_CONSTANTS(TS);
  const Grid &g = M2C_G;
  constexpr int SOFF = DIR == 0 ? 0 : NSX;
  constexpr int NF = DIR == 0 ? NFX : NFY;
  const int xm = x - (DIR == 0), ym = y - (DIR == 1);
  const int gpl = DIR == 0 ? i0 + x : j0 + y;
  const int gn = DIR == 0 ? g.nx : g.ny;
  const bool inside = (DIR == 0) ? (i0 + x <= g.nx && j0 + y < g.ny) : (i0 + x < g.nx && j0 + y <= g.ny);
  TC F[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
  if (inside) {
    int sm, sp;
    if (DIR == 0) { sm = y * (TX + 2) + (xm + 1); sp = y * (TX + 2) + (x + 1); }
    else { sm = (ym + 1) * TX + x; sp = (y + 1) * TX + x; }
    TC vcm[5], vcp[5], sgm[5], sgp[5], fm[5], fp[5];
#pragma unroll
    for (int p = 0; p < 5; p++) {
      vcm[p] = sV[p][ym + HALO][xm + HALO];
      vcp[p] = sV[p][y + HALO][x + HALO];
      sgm[p] = sS[p * NSL + SOFF + sm];
      sgp[p] = sS[p * NSL + SOFF + sp];
    }
    const int idm = sID[ym + HALO][xm + HALO], idp = sID[y + HALO][x + HALO];
    const bool per = g.bc[2 * DIR] == BC_PERIODIC;
    if (gpl == 0 && !per) {
      nrev += face_state<EK, false, TC, MP>(vcp, sgp, idp, fp);
      ghost_face<DIR, TC>(g.bc[2 * DIR], fp, vcm, fm);
    } else if (gpl == gn && !per) {
      nrev += face_state<EK, true, TC, MP>(vcm, sgm, idm, fm);
      ghost_face<DIR, TC>(g.bc[2 * DIR + 1], fm, vcp, fp);
    } else {
      nrev += face_state<EK, true, TC, MP>(vcm, sgm, idm, fm);
      nrev += face_state<EK, false, TC, MP>(vcp, sgp, idp, fp);
    }
    face_flux<FLUX, EK, DIR, TC, MP>(fm, fp, vcm, vcp, idm, idp, F);
  }
#pragma unroll
  for (int p = 0; p < 5; p++) sF[p * NF + fi] = F[p];
}

// convert a fetched cell state to the compute type of the stencil
template <typename TC> __device__ __forceinline__ void to_tc(const double *v, TC *o) {
#pragma unroll
  for (int p = 0; p < 5; p++) o[p] = TC(v[p]);
}

// MODE 0: residual only (Out = R). MODE 1: intermediate stage (Out = U). MODE 2: final stage
// (Out = V, plus the CFL reduction for the next step).
// TC is the arithmetic type of reconstruction, limiters, face-state checks and fluxes: double for
// the drop-in path; float for the mixed form, where the fetch (with ConservativeToPrimitive and
// its p(e) inversion), the residual accumulation, the stage combination, ConservativeToPrimitive
// of the result, clipping and the CFL reduction stay in double.
template <int IN_CONS, int MODE, int FLUX, int EK, typename TC = double, bool MP = (EK < 0), class TS = DefaultTile>
__device__ __forceinline__ void stage_body(const StageArgs &a, const int bx, const int by, const int bz) {
  M2C_TILE_CONSTANTS(TS);
  const Grid &g = M2C_G;
  if (!a.ignore_active && !a.clk->active) return;
  // faces cut by embedded walls and stretched meshes: the run-time equation-of-state variants only
  // (the stiffened-gas variants keep their registers for the interface code)
  constexpr bool WALLS = MP && (EK < 0);
  constexpr bool STR = EK < 0;

  __shared__ TC sV[5][LY][LX];
  __shared__ int sID[LY][LX];
  __shared__ TC sS[5 * NSL];                  // slopes
  __shared__ TC sFx[5 * NFX];                 // x-face fluxes
  __shared__ TC sFy[5 * NFY];                 // y-face fluxes
  __shared__ unsigned sFlags;

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int i0 = bx * TX, j0 = by * TY;
  const int kbeg = bz * a.kchunk;
  const int kend = min(kbeg + a.kchunk, g.nz);
  const int gi = i0 + tx, gj = j0 + ty;
  const bool own = gi < g.nx && gj < g.ny;
  // columns beyond the domain edge hold the ghost values there (fetch_cell applies the boundary
  // rules), so a partial tile's neighbours of the last cells are correct
  const int ci = min(gi, g.nx + 1), cj = min(gj, g.ny + 1);
  const bool zper = g.bc[4] == BC_PERIODIC;
  // the block's x/y window (tile and halo) lies in the box: its fetches need no ghost rules
  const bool xy_in = i0 - HALO >= 0 && i0 + TX + HALO <= g.nx && j0 - HALO >= 0 && j0 + TY + HALO <= g.ny;
  unsigned flags = 0;
  unsigned long long nrev = 0;
  double cfl_min = 1.7976931348623157e308;
  if (tid == 0) sFlags = 0;

  const double dt = a.clk->dt;
  const double cdt = a.c * dt;
  constexpr bool EXACT = sizeof(TC) == sizeof(double);

  // ---- z prologue: sigma_z(kbeg), flux F_{kbeg-1/2}
  TC w0[5], w1[5], w2[5], sz0[5], sz1[5], Flo[5];
  double w0d[5];   // own cell in double (drop-in path: the stencil value itself)
  int id0, id1, id2;
  int idlo;        // material of the cell below (k-1)
  int lo_iface = 0;
  {
    double d[5];
    TC vm2[5], vm1[5];
    int idm2, idm1;
    fetch_cell<IN_CONS, EK, MP>(a.S, a.ID, ci, cj, kbeg - 2, d, idm2, flags); to_tc(d, vm2);
    fetch_cell<IN_CONS, EK, MP>(a.S, a.ID, ci, cj, kbeg - 1, d, idm1, flags); to_tc(d, vm1);
    fetch_cell<IN_CONS, EK, MP>(a.S, a.ID, ci, cj, kbeg, w0d, id0, flags); to_tc(w0d, w0);
    fetch_cell<IN_CONS, EK, MP>(a.S, a.ID, ci, cj, kbeg + 1, d, id1, flags); to_tc(d, w1);
    TC szm[5];
    cell_slopes<2, TC, STR>(vm2, vm1, w0, idm2, idm1, id0, szm, kbeg - 1);
    cell_slopes<2, TC, STR>(vm1, w0, w1, idm1, id0, id1, sz0, kbeg);
    if constexpr (MP) {   // a level set means two materials: the multi-material variants only
      if (a.phin) {
        if (conrec_cell(a.phin, ci, cj, kbeg - 1)) zero5(szm);
        if (conrec_cell(a.phin, ci, cj, kbeg)) zero5(sz0);
      }
      // slopes along z of cells with a cut low or high z face (Reconstructor.cpp:656-664)
      if (WALLS && a.Wm && g.slope_zero) {
        if (wall_cut(a.Wm, 2, ci, cj, kbeg - 1) | wall_cut(a.Wm, 2, ci, cj, kbeg)) zero5(szm);
        if (wall_cut(a.Wm, 2, ci, cj, kbeg) | wall_cut(a.Wm, 2, ci, cj, kbeg + 1)) zero5(sz0);
      }
    }
    TC fm[5], fp[5];
    nrev += face_state<EK, false, TC, MP>(w0, sz0, id0, fp);
    if (kbeg == 0 && !zper) {
      ghost_face<2, TC>(g.bc[4], fp, vm1, fm);
    } else {
      nrev += face_state<EK, true, TC, MP>(vm1, szm, idm1, fm);
    }
    lo_iface = face_flux<FLUX, EK, 2, TC, MP>(fm, fp, vm1, w0, idm1, id0, Flo);
    idlo = idm1;
    fetch_cell<IN_CONS, EK, MP>(a.S, a.ID, ci, cj, kbeg + 2, d, id2, flags); to_tc(d, w2);
  }

  // halo cell owned by this thread (x halo: 4 columns x TY rows, then y halo: 4 rows x TX columns)
  int hx = 0, hy = 0;
  if (tid < 2 * HALO * TY) {
    int c = tid % (2 * HALO), r = tid / (2 * HALO);
    hx = (c < HALO) ? c - HALO : TX + (c - HALO);
    hy = r;
  } else if (tid < NHALO) {
    int t = tid - 2 * HALO * TY;
    int r = t / TX, c = t % TX;
    hy = (r < HALO) ? r - HALO : TY + (r - HALO);
    hx = c;
  }
  const int hi = min(i0 + hx, g.nx + 1), hj = min(j0 + hy, g.ny + 1);
  TC hv[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
  int hid = 0;
  if (tid < NHALO) {
    double d[5];
    fetch_cell<IN_CONS, EK, MP>(a.S, a.ID, hi, hj, kbeg, d, hid, flags);
    to_tc(d, hv);
  }

  for (int k = kbeg; k < kend; k++) {
    // ---- prefetch V(k+3) of the column and the halo cell of plane k+1 (consumed at the end)
    double w3d[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    TC hvn[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
    int id3 = 0, hidn = 0;
    const bool in3 = xy_in && k + 3 < g.nz, in1 = xy_in && k + 1 < g.nz;
    const bool more = k + 1 < kend;
    Fetched f3, fh;
    if (more) {
      fetch_begin(a.S, a.ID, ci, cj, k + 3, f3, in3);
      if (tid < NHALO) fetch_begin(a.S, a.ID, hi, hj, k + 1, fh, in1);
    }

    // ---- plane k into LDS: own cells and the halo cell from registers
#pragma unroll
    for (int p = 0; p < 5; p++) sV[p][ty + HALO][tx + HALO] = w0[p];
    sID[ty + HALO][tx + HALO] = id0;
    if (tid < NHALO) {
#pragma unroll
      for (int p = 0; p < 5; p++) sV[p][hy + HALO][hx + HALO] = hv[p];
      sID[hy + HALO][hx + HALO] = hid;
    }
    // ---- z slope of k+1 (registers)
    cell_slopes<2, TC, STR>(w0, w1, w2, id0, id1, id2, sz1, k + 1);
    if constexpr (MP) {
      if (a.phin && conrec_cell(a.phin, ci, cj, k + 1)) zero5(sz1);
      if (WALLS && a.Wm && g.slope_zero && (wall_cut(a.Wm, 2, ci, cj, k + 1) | wall_cut(a.Wm, 2, ci, cj, k + 2)))
        zero5(sz1);
    }
    __syncthreads();

    // ---- slopes in x and y, once per cell and direction
#pragma unroll
    for (int r = 0; r < SLOPE_ROUNDS; r++) {
      int it = tid + r * NTHR;
      if (it < NSX) {
        int x = it % (TX + 2) - 1, y = it / (TX + 2);
        TC vl[5], vc[5], vr[5], sg[5];
#pragma unroll
        for (int p = 0; p < 5; p++) {
          vl[p] = sV[p][y + HALO][x - 1 + HALO];
          vc[p] = sV[p][y + HALO][x + HALO];
          vr[p] = sV[p][y + HALO][x + 1 + HALO];
        }
        cell_slopes<0, TC, STR>(vl, vc, vr, sID[y + HALO][x - 1 + HALO], sID[y + HALO][x + HALO],
                                sID[y + HALO][x + 1 + HALO], sg, i0 + x);
        if constexpr (MP) {
          if (a.phin && conrec_cell(a.phin, i0 + x, j0 + y, k)) zero5(sg);
          if (WALLS && a.Wm && g.slope_zero &&
              (wall_cut(a.Wm, 0, i0 + x, j0 + y, k) | wall_cut(a.Wm, 0, i0 + x + 1, j0 + y, k)))
            zero5(sg);
        }
#pragma unroll
        for (int p = 0; p < 5; p++) sS[p * NSL + it] = sg[p];
      } else if (it < NSL) {
        int t = it - NSX;
        int x = t % TX, y = t / TX - 1;
        TC vl[5], vc[5], vr[5], sg[5];
#pragma unroll
        for (int p = 0; p < 5; p++) {
          vl[p] = sV[p][y - 1 + HALO][x + HALO];
          vc[p] = sV[p][y + HALO][x + HALO];
          vr[p] = sV[p][y + 1 + HALO][x + HALO];
        }
        cell_slopes<1, TC, STR>(vl, vc, vr, sID[y - 1 + HALO][x + HALO], sID[y + HALO][x + HALO],
                                sID[y + 1 + HALO][x + HALO], sg, j0 + y);
        if constexpr (MP) {
          if (a.phin && conrec_cell(a.phin, i0 + x, j0 + y, k)) zero5(sg);
          if (WALLS && a.Wm && g.slope_zero &&
              (wall_cut(a.Wm, 1, i0 + x, j0 + y, k) | wall_cut(a.Wm, 1, i0 + x, j0 + y + 1, k)))
            zero5(sg);
        }
#pragma unroll
        for (int p = 0; p < 5; p++) sS[p * NSL + it] = sg[p];
      }
    }
    __syncthreads();

    // ---- x and y face fluxes: the west/south face of every column, then the east column of
    //      x faces and the north row of y faces
    tile_face<FLUX, EK, 0, TC, MP, TS>(sV, sID, sS, sFx, tx, ty, ty * (TX + 1) + tx, i0, j0, nrev);
    tile_face<FLUX, EK, 1, TC, MP, TS>(sV, sID, sS, sFy, tx, ty, ty * TX + tx, i0, j0, nrev);
    {
      // the TY east x faces and the TX north y faces each go to one wave, two waves apart, and
      // the pair rotates with the block: the two other waves of the block skip the extra flux
      // and their SIMDs run the co-resident block's waves meanwhile
      static_assert(NTHR == 256 && TX <= 64 && TY <= 64, "one wave per edge row");
      const int wave = tid >> 6, lane = tid & 63;
      const int rot = (bx + by + bz) & 3;
      if (wave == rot && lane < TY)
        tile_face<FLUX, EK, 0, TC, MP, TS>(sV, sID, sS, sFx, TX, lane, lane * (TX + 1) + TX, i0, j0, nrev);
      if (wave == ((rot + 2) & 3) && lane < TX)
        tile_face<FLUX, EK, 1, TC, MP, TS>(sV, sID, sS, sFy, lane, TY, TY * TX + lane, i0, j0, nrev);
    }

    // ---- z face k+1/2 (registers)
    TC Fhi[5];
    int hi_iface;
    {
      TC fm[5], fp[5];
      nrev += face_state<EK, true, TC, MP>(w0, sz0, id0, fm);
      if (k + 1 == g.nz && !zper) {
        ghost_face<2, TC>(g.bc[5], fm, w1, fp);
      } else {
        nrev += face_state<EK, false, TC, MP>(w1, sz1, id1, fp);
      }
      hi_iface = face_flux<FLUX, EK, 2, TC, MP>(fm, fp, w0, w1, id0, id1, Fhi);
    }
    __syncthreads();

    // ---- accumulate the residual of cell (gi, gj, k) and finish the stage
    if (own) {
      double Fw[5], Fe[5], Fs[5], Fn[5], Fb[5], Ft[5];
      const int fw = ty * (TX + 1) + tx, fe = fw + 1;
      const int fs = ty * TX + tx, fn = fs + TX;
#pragma unroll
      for (int p = 0; p < 5; p++) {
        Fw[p] = double(sFx[p * NFX + fw]);
        Fe[p] = double(sFx[p * NFX + fe]);
        Fs[p] = double(sFy[p * NFY + fs]);
        Fn[p] = double(sFy[p * NFY + fn]);
        Fb[p] = double(Flo[p]);
        Ft[p] = double(Fhi[p]);
      }
      if (MP && g.has_interfaces && a.Fi) {
        // interface faces: the plus cell takes localflux2 of its low face, the minus cell
        // localflux1 of its high face (SpaceOperator.cpp:2548-2556)
        const int me = sID[ty + HALO][tx + HALO];
        const long long Nf = a.Nf;
        if (is_interface(sID[ty + HALO][tx - 1 + HALO], me)) {
          const double *f = a.Fi + 0 * 10 * Nf + face_key(gi, gj, k);
#pragma unroll
          for (int p = 0; p < 5; p++) Fw[p] = f[(5 + p) * Nf];
        }
        if (is_interface(me, sID[ty + HALO][tx + 1 + HALO])) {
          const double *f = a.Fi + 0 * 10 * Nf + face_key(gi + 1, gj, k);
#pragma unroll
          for (int p = 0; p < 5; p++) Fe[p] = f[p * Nf];
        }
        if (is_interface(sID[ty - 1 + HALO][tx + HALO], me)) {
          const double *f = a.Fi + 1 * 10 * Nf + face_key(gi, gj, k);
#pragma unroll
          for (int p = 0; p < 5; p++) Fs[p] = f[(5 + p) * Nf];
        }
        if (is_interface(me, sID[ty + 1 + HALO][tx + HALO])) {
          const double *f = a.Fi + 1 * 10 * Nf + face_key(gi, gj + 1, k);
#pragma unroll
          for (int p = 0; p < 5; p++) Fn[p] = f[p * Nf];
        }
        if (lo_iface) {
          const double *f = a.Fi + 2 * 10 * Nf + face_key(gi, gj, k);
#pragma unroll
          for (int p = 0; p < 5; p++) Fb[p] = f[(5 + p) * Nf];
        }
        if (hi_iface) {
          const double *f = a.Fi + 2 * 10 * Nf + face_key(gi, gj, k + 1);
#pragma unroll
          for (int p = 0; p < 5; p++) Ft[p] = f[p * Nf];
        }
      }
      if (FLUX == FLUX_GODUNOV) {
        // faces between two cells of this cell's material: the Godunov pass's fluxes
        const int me = sID[ty + HALO][tx + HALO];
        if (me != g.inactive_id) {
          const long long gxw = ((long long)k * g.ny + gj) * (g.nx + 1) + gi;
          const long long gys = ((long long)k * (g.ny + 1) + gj) * g.nx + gi;
          const long long gzb = ((long long)k * g.ny + gj) * g.nx + gi;
          const long long mx = (long long)(g.nx + 1) * g.ny * g.nz, my = (long long)g.nx * (g.ny + 1) * g.nz,
                          mz = (long long)g.nx * g.ny * (g.nz + 1), mxy = (long long)g.nx * g.ny;
          const bool w = sID[ty + HALO][tx - 1 + HALO] == me, e = sID[ty + HALO][tx + 1 + HALO] == me;
          const bool so = sID[ty - 1 + HALO][tx + HALO] == me, no = sID[ty + 1 + HALO][tx + HALO] == me;
          const bool bo = idlo == me, to = id1 == me;
#pragma unroll
          for (int p = 0; p < 5; p++) {
            if (w) Fw[p] = a.G[0][p * mx + gxw];
            if (e) Fe[p] = a.G[0][p * mx + gxw + 1];
            if (so) Fs[p] = a.G[1][p * my + gys];
            if (no) Fn[p] = a.G[1][p * my + gys + g.nx];
            if (bo) Fb[p] = a.G[2][p * mz + gzb];
            if (to) Ft[p] = a.G[2][p * mz + gzb + mxy];
          }
        }
      }
      if constexpr (WALLS) {
        if (a.Wm && a.Fi) {
          // faces cut by an embedded wall, after the Godunov fluxes: as at an interface face, the
          // plus cell takes localflux2 of its low face, the minus cell localflux1 of its high face
          // (SpaceOperator.cpp:2438-2484)
          const long long Nf = a.Nf;
          const long long kw = face_key(gi, gj, k);
          const unsigned m0 = a.Wm[kw];
          const bool cw = m0 & 1u, cs = (m0 >> 1) & 1u, cb = (m0 >> 2) & 1u;
          const bool ce = a.Wm[kw + 1] & 1u;
          const bool cn = ((unsigned)a.Wm[kw + (g.nx + 1)] >> 1) & 1u;
          const bool ct = ((unsigned)a.Wm[kw + (long long)(g.nx + 1) * (g.ny + 1)] >> 2) & 1u;
#pragma unroll
          for (int p = 0; p < 5; p++) {
            if (cw) Fw[p] = a.Fi[0 * 10 * Nf + (5 + p) * Nf + kw];
            if (ce) Fe[p] = a.Fi[0 * 10 * Nf + p * Nf + kw + 1];
            if (cs) Fs[p] = a.Fi[1 * 10 * Nf + (5 + p) * Nf + kw];
            if (cn) Fn[p] = a.Fi[1 * 10 * Nf + p * Nf + kw + (g.nx + 1)];
            if (cb) Fb[p] = a.Fi[2 * 10 * Nf + (5 + p) * Nf + kw];
            if (ct) Ft[p] = a.Fi[2 * 10 * Nf + p * Nf + kw + (long long)(g.nx + 1) * (g.ny + 1)];
          }
        }
      }
      double R[5];
      // viscous and heat-diffusion faces of this cell (SpaceOperator.cpp:3367-3371): added after the
      // advective sum, own faces first, then the east, north and front faces (ViscosityOperator.cpp:
      // 259-393, HeatDiffusionOperator.cpp:96-131)
            const bool diff = a.Dv[0] != nullptr;
      const long long fxw = ((long long)k * g.ny + gj) * (g.nx + 1) + gi;
      const long long fys = ((long long)k * (g.ny + 1) + gj) * g.nx + gi;
      const long long fzb = ((long long)k * g.ny + gj) * g.nx + gi;
      const long long nfx = (long long)(g.nx + 1) * g.ny * g.nz, nfy = (long long)g.nx * (g.ny + 1) * g.nz,
                      nfz = (long long)g.nx * g.ny * (g.nz + 1), nxy = (long long)g.nx * g.ny;
      // face areas and volume (SpaceOperator.cpp:160-177, 2554): the cell's own widths
      double ax = g.ax, ay = g.ay, az = g.az, inv_mvol = g.inv_mvol;
      if (STR && g.stretched) {
        const double wx = g.w[0][gi + 2], wy = g.w[1][gj + 2], wz = g.w[2][k + 2];
        ax = wy * wz; ay = wx * wz; az = wx * wy;
        inv_mvol = 1.0 / (-(wx * wy * wz));
      }
#pragma unroll
      for (int p = 0; p < 5; p++) {
        double s = 0.0;
        s -= Fw[p] * ax;
        s -= Fs[p] * ay;
        s -= Fb[p] * az;
        s += Fe[p] * ax;
        s += Fn[p] * ay;
        s += Ft[p] * az;
        if (diff) {
          s += a.Dv[0][p * nfx + fxw];
          s += a.Dv[1][p * nfy + fys];
          s += a.Dv[2][p * nfz + fzb];
          s -= a.Dv[0][(5 + p) * nfx + fxw + 1];
          s -= a.Dv[1][(5 + p) * nfy + fys + g.nx];
          s -= a.Dv[2][(5 + p) * nfz + fzb + nxy];
          if (p == 4) {
            s += a.Dh[0][fxw];
            s += a.Dh[1][fys];
            s += a.Dh[2][fzb];
            s -= a.Dh[0][fxw + 1];
            s -= a.Dh[1][fys + g.nx];
            s -= a.Dh[2][fzb + nxy];
          }
        }
        R[p] = s * inv_mvol;
        if (!dfinite(R[p])) flags |= ST_NAN_FLUX;
      }
      const long long idx = ((long long)k * g.ny + gj) * g.nx + gi;
      if (MODE == 0) {
#pragma unroll
        for (int p = 0; p < 5; p++) a.Out[p * g.N + idx] = R[p];
      } else {
        const auto &m = mat_of<EK, MP>(id0);
        double vn[5], un[5], U[5];
        if (IN_CONS) {
#pragma unroll
          for (int p = 0; p < 5; p++) vn[p] = a.Vn[p * g.N + idx];
        } else if constexpr (EXACT) {
#pragma unroll
          for (int p = 0; p < 5; p++) vn[p] = double(w0[p]);
        } else {
#pragma unroll
          for (int p = 0; p < 5; p++) vn[p] = w0d[p];
        }
        eos_p2c<double, EK>(m, vn, un);
        if (a.first) {
#pragma unroll
          for (int p = 0; p < 5; p++) U[p] = un[p] + dt * R[p];
        } else {
#pragma unroll
          for (int p = 0; p < 5; p++) {
            double us = a.Us[p * g.N + idx];
            double t = a.a * us + a.b * un[p];
            U[p] = t + cdt * R[p];
          }
        }
        double V[5];
        eos_c2p<EK>(m, U, V);
        bool clipped = eos_clip<double, EK>(m, V);
        if (eos_bad<double, EK>(m, V)) flags |= ST_BAD_STATE;
        if (MODE == 1) {
          if (clipped) flags |= ST_CLIP_INTERMEDIATE;
#pragma unroll
          for (int p = 0; p < 5; p++) a.Out[p * g.N + idx] = U[p];
        } else {
#pragma unroll
          for (int p = 0; p < 5; p++) a.Out[p * g.N + idx] = V[p];
          // CFL contribution of V(n+1) (SpaceOperator.cpp:2128-2198)
          if (id0 != g.inactive_id) {
            double c2 = eos_c2<double, EK>(m, V[0], eos_e<double, EK>(m, V[0], V[4]));
            if (c2 < 0) flags |= ST_NEG_C2;
            double cs = dsqrt(c2);
            double lf = dabs(V[1]) + cs, lg = dabs(V[2]) + cs, lh = dabs(V[3]) + cs;
            double hx = g.dx, hy = g.dy, hz = g.dz;
            if (STR && g.stretched) { hx = g.w[0][gi + 2]; hy = g.w[1][gj + 2]; hz = g.w[2][k + 2]; }
            double loc = smin(hx / lf, smin(hy / lg, hz / lh));
            if (loc < cfl_min) cfl_min = loc;
          }
          if (clipped) atomicAdd(&a.st->clipped, 1ull);
        }
      }
    }

    // ---- the prefetched cells, converted now that their loads have landed
    if (more) {
      fetch_end<IN_CONS, EK, MP>(f3, w3d, flags);
      id3 = f3.id;
      if (tid < NHALO) {
        double d[5];
        fetch_end<IN_CONS, EK, MP>(fh, d, flags);
        to_tc(d, hvn);
        hidn = fh.id;
      }
    }
    // ---- shift the z window
#pragma unroll
    for (int p = 0; p < 5; p++) {
      Flo[p] = Fhi[p];
      sz0[p] = sz1[p];
      w0[p] = w1[p];
      w1[p] = w2[p];
      w2[p] = TC(w3d[p]);
    }
    if constexpr (!EXACT) {
      if (!IN_CONS && MODE != 0 && k + 1 < kend) {
      // mixed form: the own cell of the next plane again in double (the stage combination needs
      // V_n exactly)
        int dummy;
        unsigned fl = 0;
        fetch_cell<IN_CONS, EK, MP>(a.S, a.ID, ci, cj, k + 1, w0d, dummy, fl);
      }
    }
    lo_iface = hi_iface;
    idlo = id0;
    id0 = id1; id1 = id2; id2 = id3;
#pragma unroll
    for (int p = 0; p < 5; p++) hv[p] = hvn[p];
    hid = hidn;
    __syncthreads();
  }
  if (MODE == 2 && !a.no_cfl) warp_min_atomic(&a.clk->minbits, cfl_min);
  if (flags) atomicOr(&sFlags, flags);
  if (nrev) atomicAdd(&a.st->reverted, nrev);
  __syncthreads();
  if (tid == 0 && sFlags) atomicOr(&a.st->bits, sFlags);
}

#ifndef M2C_BATCHED
template <int IN_CONS, int MODE, int FLUX, int EK, typename TC = double, bool MP = (EK < 0)>
#ifndef M2C_STAGE_BLOCKS
#define M2C_STAGE_BLOCKS 2   // resident 256-thread blocks per CU the register budget is sized for
#endif
__global__ __launch_bounds__(NTHR, M2C_STAGE_BLOCKS) void stage_kernel(StageArgs a) {
  stage_body<IN_CONS, MODE, FLUX, EK, TC, MP>(a, blockIdx.x, blockIdx.y, blockIdx.z);
}
#endif

}  // namespace m2c
