// multiphase.hpp -- shared types of the interface-face, level-set and end-of-step kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fv_kernels.hpp"
#include "mlp.hpp"
#include "riemann_dev.hpp"

namespace m2c {

constexpr int MP_SORT_CAP = 8192;   // ordered single-thread passes handle at most this many nodes
enum { LS_FIRST = 0, LS_AXBY = 1, LS_X = 2 };

// run-time constants of the level set and the interface solver
struct MpParams {
  int ls_matid;
  int ls_bc[6];
  int reinit_freq, reinit_maxits;
  double reinit_tol;
  double sign_factor;          // smoothing width of the sign function: 1 * min(dx, dy, dz)
  double min_dx;               // min(dx, dy, dz): pseudo-time step = cfl * min_dx
  const double *xc[3];         // cell-centre coordinates, index -1..n stored at [i+1]
  const double *wc[3];         // cell widths, index -1..n stored at [i+1] (ghosts: boundary widths)
  int stretched;               // the widths vary: the reinitialisation's steps and sign widths per cell
  const double *fdl[3], *fdr[3], *fdr1y;   // FD3 coefficients per index along each axis
  // fd_uni[d]: the coefficients of fdl[d] and fdr[d] are the same for every index 2..n-3 (cells whose
  // stencil stays in the box); they are then also held in fdl_u/fdr_u. Likewise fdr1y.
  int fd_uni[3], fd_uni_r1y;
  double fdl_u[3][4], fdr_u[3][4], fdr1y_u[4];
  int ls_neumann;              // every level-set boundary is zero-Neumann: a ghost is its clamped image
  // dx_pow2[d]: all cell-centre spacings along d (ghosts included) equal one power of two h, so
  // x / h == x * inv_dx[d] exactly
  int dx_pow2[3];
  double inv_dx[3];
  RiemannCfg rcfg;
  int path_cap;                // rarefaction path entries per side and thread
  double failsafe_density[8];
  int riemann_normal;          // MultiPhase.RiemannNormal: 0 level set, 1 mesh, 2 average
};

struct ReinitState {
  int active, detected, done, iter, retry;
  int cur;                     // fused iterations: the buffer of the pair holding phi
  int accepted, pad;           // fused iterations whose result became current
  double cfl;                  // persists between calls (LevelSetReinitializer::cfl)
  unsigned long long maxres_bits, dphi_bits;
  unsigned long long acc_res_bits, acc_dphi_bits;   // fused iterations: maxima of the running one
};

// device-resident per-step counters
struct MpState {
  unsigned long long epoch;    // stamp of the Riemann solutions of the current step
  unsigned nface, nrem, nchg, nunres;
  int nseq, pad;
  ReinitState rs;
  long long reinit_iters, unresolved_total;
  unsigned long long sur_accepted, sur_fallback;   // interface faces solved by the network / the exact solver
  unsigned nfb, pad2;          // listed faces whose network solution failed its check (this stage)
};

// device buffers and launch parameters of one context
struct MpLaunch {
  hipStream_t stream;
  long long N, Nf;
  int nx, ny, nz;
  int flux, riemann_threads;
  int ls_neumann;              // every level-set boundary zero-Neumann (the fused reinitialisation applies)
  int ls_f32;                  // mixed form: the fused reinitialisation iterates in FP32
  float *phif0, *phif1, *signf;   // its pair of level-set buffers and the sign function
  int8_t *ID, *IDnew;
  unsigned *faces;
  MpState *ms;
  double *Fi, *rsV;
  const double *phin;          // the level set at t_n (the interface normal of RiemannNormal LevelSet/Average)
  unsigned long long *rsMeta;
  double *scratch;
  Status *st;
  Clock *clk;
  unsigned *rem;
  uint8_t *tagR, *tag;
  double *sign, *phibk, *rphi1a, *rphi1b;
  unsigned *chg;
  int *cidx, *lvl;
  int8_t *cnt;
  double *Vnew;
  unsigned *unres;
  // learned interface solver (sur_on): features and predictions of the first sur_cap listed faces
  int sur_on, sur_cap, sur_nmax;
  double sur_tol;
  float *surX, *surY;
  unsigned *surFb;             // listed-face indices for the exact fallback pass
  MlpWeights sur_w;
};

// One instance of the multi-instance interface pass (several contexts' exact interface solves in
// one launch, each block serving one instance): the tables a context's own pass reads from
// constant memory, and its buffers. Stage s reads S[s] (primitive for s = 0, conservative after).
struct IfaceInst {
  Grid grid;
  MpParams mp;
  DevEos eos[8];
  const double *S[3];
  const int8_t *ID;
  unsigned *faces;
  MpState *ms;
  double *Fi, *rsV;
  unsigned long long *rsMeta;
  double *scratch;
  Status *st;
  const Clock *clk;
  const double *phin;
  int store;                   // keeps the Riemann solutions (a level set follows the interface)
  int blk0, nblk;              // the launch's blocks blk0 .. blk0 + nblk - 1 serve this instance
  int lblk0, lnblk;            // likewise in the face-list launch (m2c_mp_iface_list_blocks of its grid)
};

}  // namespace m2c

extern "C" {
// the multi-instance interface pass over ninst instances (dtab: device copy) at stage `stage`
// (in_cons: its input is conservative), nblk_total blocks of 64 threads; `store`: the step's first
// stage (instances with a level set keep their Riemann solutions)
// the interface-face list of a stage (relist: build it; else keep the step's list and reset the
// fallback count), the first half of m2c_mp_launch_iface
int m2c_mp_launch_iface_list(const m2c::MpLaunch &L, int ignore_active, int relist);
// blocks of one instance's face-list sweep (grid nx x ny x nz)
int m2c_mp_iface_list_blocks(int nx, int ny, int nz);
// the face lists of the ninst instances of dtab (relist: rebuilt; else only the fallback counts
// reset), nlist_blocks blocks in all
int m2c_mp_launch_iface_list_multi(hipStream_t s, const m2c::IfaceInst *dtab, int ninst, int nlist_blocks, int relist);
int m2c_mp_launch_iface_multi(hipStream_t s, const m2c::IfaceInst *dtab, int ninst, int nblk_total, int flux,
                              int in_cons, int stage, int store);
int m2c_mp_launch_sur_batch(hipStream_t s, int n, const int *dir, const double *Vm, const int *idm, const double *Vp,
                            const int *idp, float *X, float *Y, const m2c::MlpWeights &w, int nmax, double tol,
                            double *Vs, int *id, double *Vsm, double *Vsp, int *status, int *path, double *scratch,
                            int nthreads, unsigned *fbl, unsigned *nfb);
int m2c_mp_upload(const m2c::Grid &grid, const m2c::DevEos *eos, const m2c::MpParams &mp);
int m2c_mp_upload_async(hipStream_t s, const m2c::Grid *grid, const m2c::DevEos *eos, const m2c::MpParams *mp);
int m2c_mp_step_begin(const m2c::MpLaunch &L);
int m2c_mp_launch_riemann_batch(hipStream_t s, int n, const double *dir, const double *Vm, const int *idm,
                                const double *Vp, const int *idp, double *Vs, int *id, double *Vsm, double *Vsp,
                                int *status, double *scratch, int nthreads);
int m2c_mp_launch_onesided_batch(hipStream_t s, int flux, int face, int n, const double *dir, const double *Vm,
                                 const int *idl, const double *uw, const int *ax, const int *fb, int mode, double *Vs,
                                 int *id, double *Vsm, double *F, int *status, double *scratch, int nthreads);
int m2c_mp_launch_godunov_faces(hipStream_t s, int in_cons, const double *S, const int8_t *ID, double *const *G,
                                const long long *nf, double *scratch, int nthreads, m2c::Status *st,
                                const m2c::Clock *clk, int ignore_active, const double *crphi);
int m2c_mp_launch_walls(hipStream_t s, int flux, int in_cons, const double *S, const int8_t *ID, const unsigned *faces,
                        const double *wall, int n, int mode, int recon, double *Fi, long long Nf, double *scratch,
                        int nthreads, m2c::Status *st, const m2c::Clock *clk, int ignore_active);
int m2c_mp_launch_godunov_batch(hipStream_t s, int dir, int n, const double *Vm, const double *Vp, int id, double *F,
                                double *scratch, int nthreads);
int m2c_mp_launch_iface(const m2c::MpLaunch &L, const double *S, int in_cons, int store, int ignore_active,
                        int relist);
int m2c_mp_launch_ls(const m2c::MpLaunch &L, int in_cons, int mode, int final_, const double *S, const double *phiS,
                     const double *X, const double *Y, double *out, double a, double b, double coef,
                     const double *phin);
int m2c_mp_launch_id_remaining(const m2c::MpLaunch &L, int single);
int m2c_mp_launch_reinit(const m2c::MpLaunch &L, double *phi, int maxits, int retry);
int m2c_mp_launch_state_update(const m2c::MpLaunch &L, double *V, int rounds);
}
