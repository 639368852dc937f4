// fdtd2d.hip -- gfx950 2D stepper behind include/fdtd2d.h: TEz (Hz, Ex,
// Ey), the reference's 2D driver, and TMz (Ez, Hx, Hy).
//
// One code path steps one simulation or a batch of many.  A step is five
// launches over every simulation of the batch: the E half step over a
// packed list of 64x4-cell tiles, the dispersive cells, the H half step,
// the monitors, and a one-thread-per-simulation step advance.  Thirty-two
// steps and the DFT fold are captured once in a HIP graph and replayed; the
// step number of each simulation lives on the device, so simulations with
// different step counts share the replays (a finished one skips its tiles).
// A single simulation is a batch of one.
//
// Per simulation, D-form in the reference's order (sim2/program.f90:320-370,
// sim2 = FDTDsimulations/2D/air-nanoAntennas_Si02-Sisubstrate/):
//   E: rows j0..j1 (npmly..N_y-npmly): D += Ku curl Hz (include/
//      myfdtd_2Dfunc.f90:16-32), line sources on Dx, E = D*gax of the last
//      dielectric writer (dispersive cells: their own launch); in the y slabs
//      the split-field PML (mod2/BERpmly_HzExEy_mod.f90:96-123) with Ex
//      handed to the main field on rows 1..ja and jb..N_y (:144-178);
//      periodic x: the ghost columns of Ey follow (mympi_2Dfunc.f90).
//   H: rows j0..j1: Hz -= Ku curl E (myfdtd_2Dfunc.f90:34-44); slabs: Hzx,
//      Hzy (:125-142), Hz = Hzx + Hzy on rows 1..ja-1 and jb..N_y (:156-191);
//      periodic x: ghost columns of Hz.
// Cells are independent within each half step, so one thread per cell
// reproduces the reference's loops exactly.
//
// TMz, per simulation, follows the 3D FUSED engine's order for fields with
// no z variation (sim3/include/myfdtd_func.f90:19-64 with the z differences
// zero; mod3/BERpmly_mod.f90:117-195 on the x and y faces):
//   E: non-PML cells Dz += Ku ((Hy(i)-Hy(i-1)) + (Hx(j-1)-Hx(j))), line
//      sources on Dz, Ez = Dz*gax (dispersive cells: their own launch);
//      PML cells Ezx, Ezy split updates scaled by the cell's PML medium, the
//      source on Ezy, Ez = Ezx + Ezy; periodic x: ghost columns of Ez.
//   H: non-PML cells Hx -= Ku (Ez(j+1)-Ez(j)), Hy += Ku (Ez(i+1)-Ez(i));
//      PML cells Hx = g1y Hx + g2y (Ez(j)-Ez(j+1)), Hy = g1x Hy + g2x
//      (Ez(i+1)-Ez(i)) (their z splits stay zero); ghost columns of Hy.
// A batch mixes both polarisations: a tile belongs to one simulation, so
// the branch is uniform over a wavefront.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/fdtd2d.h"
#include "fdtd_internal.h"

namespace fdtd {
namespace d2 {

constexpr int kMaxLine = 4;     // line sources per simulation
constexpr int kMaxDft = 4;     // DFT lines per simulation
constexpr int kMaxProbe = 8;   // probes per simulation
constexpr int kMaxGroup = 4;   // dispersive coefficient groups per simulation
// tile: 64 x 16 cells, four wavefronts of 64 lanes, each marching through
// four consecutive rows (the row below / above stays in registers)
constexpr int kTX = 64, kTY = 16, kWaves = 4, kRowsPerWave = kTY / kWaves;
constexpr int kChunk = 256;    // cells per work item of the list kernels
constexpr int kBlockSteps = 32;  // steps per captured graph (one DFT block)

using cplx = std::complex<double>;

// ---- device descriptors ------------------------------------------------------
template <typename real>
struct Mat {
  real gax[2];
  int32_t act[2];  // ACT_KEEP / ACT_DIEL / ACT_ADE per component (Ex, Ey)
};
// TMz: act[0], gax[0] act on Ez; gax[1] is the PML medium (1/eps of the last
// dielectric written to the cell, 1 where none); act[1] is unused.

// Dispersive state and coefficients are FP64 in both precisions: inside a
// metal E is the small difference of the large D and I, which FP32 would
// cancel away.
template <typename real>
struct AdeDev {
  const int32_t* cells;  // ghosted cell offsets
  const uint8_t* mask;   // bit 0: Ex (TMz: Ez), bit 1: Ey
  double* state;         // [word][comp][cell]
  int32_t ncells, nstate;
  AdeCoef<double> co;
};

// Lorentz-Jp cells of one coefficient group: local component 0 = x (TEz) or
// z (TMz), 1 = y (TEz); FP64 pole state [word][comp][cell], words P_l
// (l < npols), Jp_l, [Q, G].
struct LjpDev {
  const int32_t* cells;
  const uint8_t* mask;
  double* state;
  int32_t ncells;
  LjpCoef<double> co;
};

struct DftDev {
  int32_t j, nfreq, mask, n_begin, n_end, nslots;  // slots: Ex first, then Hz
  const double* omega;
  double* acc;        // [slot][re/im][m][i]
  void* samp;         // [b][slot][i], field precision
  int32_t* sstep;     // [b]: step of sample b (0: none)
};

template <typename real>
struct Inst {
  real *ex, *ey, *hz, *dx, *dy;  // at cell (0,0); offset of (i,j) = i + px*j
  const uint8_t* matid;
  const Mat<real>* tab;
  int32_t px, nx, ny, j0, j1, ja, jb, has_pml, per, ntx, nty, tile0;
  real ku;
  double dt;
  // PML: slab s (0: rows 0..P+1, 1: rows jb-1..ny+1) of pex, pey, hzx, hzy
  // with pitch px; profiles per slab: f1y f2y g1y g2y [P+2] by slab row,
  // then f1x f2x g1x g2x [nx+2] by column
  real* pml[2][4];
  const real* prof[2];
  int32_t prow0[2], P;
  // sources: value of the step about to run, computed by the advance kernel
  int32_t nsrc;
  int32_t src_j[kMaxLine];
  double src_t0[kMaxLine], src_st[kMaxLine], src_w0[kMaxLine], src_amp[kMaxLine];
  real* src_val;
  int32_t ngroup;
  AdeDev<real> grp[kMaxGroup];
  int32_t nljp;
  LjpDev ljp[kMaxGroup];
  int32_t ndft;
  DftDev dft[kMaxDft];
  int32_t nprobe;
  int32_t probe_c[kMaxProbe];  // ghosted cell offsets
  double* probe_rec;           // [n-1][probe][3]
  int32_t probe_cap;           // steps the buffer holds
  int32_t* ctl;                // [0] steps done, [1] target
  // TMz: fields, Ez splits (PML cells only), profiles f1 f2 g1 g2 [nx+2]
  // along x then [ny+2] along y, PML depths Px (x) and P (y, above)
  int32_t tmz, Px;
  real *ez, *hx, *hy, *dz;
  real *ezx, *ezy;
  const real* tprof;
  // TFSF along y: total field on rows tf_lo..tf_hi (tf_hi 0: one-sided);
  // the incident line e, h, pe, ph, f1, f2, g1, g2 (ny+2 doubles each)
  int32_t tf_on, tf_lo, tf_hi, tf_src, tf_npml;
  double* tf_line;
  double tf_ck, tf_ku, tf_sg;
  double tf_t0, tf_st, tf_w0, tf_amp;
};

// ---- host state ------------------------------------------------------------------
struct HostMat {
  int type;  // 0 dielectric, 1 drude2cp, 2 graphene, 3 Lorentz-Jp
  fdtd2d_region reg;
  double eps_r;
  fdtd_drude2cp d2;
  fdtd_graphene gr;
  fdtd_lorentz_jp lj;
  int writes[2];
  int group;
};

// whether material m updates local component q (TEz 0 = x, 1 = y; TMz 0 =
// z) through its Lorentz-Jp poles
bool ljp_comp(const HostMat& m, bool tmz, int q) {
  if (m.type != 3) return false;
  return tmz ? (q == 0 && ((m.lj.jp_mask >> 2) & 1)) : ((m.lj.jp_mask >> q) & 1) != 0;
}

struct HostLjp {
  LjpCoef<double> co;
  int nstate = 0;
  int64_t ncells = 0;
  int32_t* d_cells = nullptr;
  uint8_t* d_mask = nullptr;
  double* d_state = nullptr;
};

struct HostGroup {
  AdeCoef<double> co;
  int nstate = 0;
  int64_t ncells = 0;
  int32_t* d_cells = nullptr;
  uint8_t* d_mask = nullptr;
  void* d_state = nullptr;
};

struct HostDft {
  int j, nfreq, mask, n_begin, n_end, nslots;
  std::vector<double> omega;
  double* d_omega = nullptr;
  double* d_acc = nullptr;
  void* d_samp = nullptr;
  int32_t* d_sstep = nullptr;
};

struct Batch;

struct Sim {
  fdtd2d_grid g{};
  int nx = 0, ny = 0;
  double ds = 0, ku = 0, dt = 0;
  int prec = FDTD_FP64;
  size_t wsize = 8;
  bool has_pml = false;
  fdtd2d_pml pml{};
  int P = 0, Px = 0, ja = 0, jb = 0, j0 = 1, j1 = 0;
  double sigmam = 0;
  std::vector<HostMat> mats;
  std::vector<std::pair<int, fdtd_pulse>> srcs;
  std::vector<HostDft> dfts;
  std::vector<std::pair<int, int>> probes;
  bool committed = false;
  int n_done = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t px = 0, ncell = 0;
  void* F[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // ex ey hz dx dy at (0,0)
  uint8_t* d_matid = nullptr;
  void* d_tab = nullptr;
  void* d_pml[2][4] = {{nullptr}};
  void* d_prof[2] = {nullptr, nullptr};
  std::vector<HostGroup> groups;
  std::vector<HostLjp> ljps;
  void* d_srcval = nullptr;
  double* d_probe = nullptr;
  int probe_cap = 0;
  std::vector<DevBuf> allocs;
  size_t dev_bytes = 0;
  Batch* solo = nullptr;  // batch of one behind fdtd2d_step
  Batch* owner = nullptr;
  // TFSF along y (one per simulation): the incident line on the device
  bool has_tf = false;
  fdtd2d_tfsf tfp{};
  double* d_tf = nullptr;  // e, h, pe, ph, f1, f2, g1, g2, ny+2 each
  // TMz: F holds Ez, Hx, Hy, Dz; Ez splits and profiles of its PML
  bool tmz = false;
  void* d_ezx = nullptr;
  void* d_ezy = nullptr;
  void* d_tprof = nullptr;
};

struct Batch {
  std::vector<Sim*> sims;
  int prec = FDTD_FP64;
  int device = 0;
  hipStream_t stream = nullptr;
  void* d_inst = nullptr;
  int32_t* d_ctl = nullptr;  // [count][2]
  int32_t* d_tile_inst = nullptr;
  int4* d_ade = nullptr;     // (q, group, first cell, 0)
  int4* d_ljp = nullptr;     // (q, Lorentz-Jp group, first cell, 0)
  int nljp = 0;
  int4* d_mon = nullptr;     // (q, line or -1 for probes, first i, 0)
  int2* d_lines = nullptr;   // (q, line)
  int32_t* d_tfl = nullptr;  // simulations with a TFSF
  int ntiles = 0, nade = 0, nmon = 0, nlines = 0, ntf = 0;
  std::vector<int32_t> ctl;  // host mirror
  // order of the simulations in the packed tile list: largest first, so
  // the small ones' tiles fill the end of each launch
  std::vector<int> order;
  std::vector<char> inst;    // host image of the descriptors
  hipGraphExec_t exec = nullptr;  // one captured 32-step block
  std::vector<DevBuf> allocs;
  // batch_free deletes both solo batches and the public fdtd2d_batch
  // (derived) through this type
  virtual ~Batch() = default;
};

int sync_sim(Sim* s);

void* alloc_in(std::vector<DevBuf>& list, size_t* total, size_t bytes) {
  if (bytes == 0) bytes = 16;
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  if (hipMemset(p, 0, bytes) != hipSuccess) {
    (void)hipFree(p);
    return nullptr;
  }
  list.push_back({p, bytes});
  if (total) *total += bytes;
  return p;
}

void* upload(std::vector<DevBuf>& list, size_t* total, const void* data, size_t bytes) {
  void* p = alloc_in(list, total, bytes);
  if (p && bytes && hipMemcpy(p, data, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  return p;
}

// ---- kernels -------------------------------------------------------------------------
__device__ __forceinline__ int tile_inst_of(const int32_t* tile_inst, int t) { return tile_inst[t]; }

template <typename real>
__device__ __forceinline__ bool in_pml_tmz(const Inst<real>& S, int i, int j) {
  return (S.Px > 0 && (i <= S.Px || i > S.nx - S.Px)) || (S.P > 0 && (j <= S.P || j > S.ny - S.P));
}

// TFSF consistency terms of row j (sim1/program.f90:236-238,260-262 with the
// far plane of the two-sided form), FP64 products rounded once:
// D(lo) -= Ku*h(lo-1), D(hi+1) += Ku*h(hi); H(lo-1) -= s*Ku*e(lo),
// H(hi) += s*Ku*e(hi+1).  False: no term on this row.
template <typename real>
__device__ __forceinline__ bool tf_d_term(const Inst<real>& S, int j, real& v) {
  if (!S.tf_on) return false;
  const double* h = S.tf_line + (S.ny + 2);
  if (j == S.tf_lo) {
    v = (real)(-S.tf_ku * h[S.tf_lo - 1]);
    return true;
  }
  if (S.tf_hi > 0 && j == S.tf_hi + 1) {
    v = (real)(S.tf_ku * h[S.tf_hi]);
    return true;
  }
  return false;
}
template <typename real>
__device__ __forceinline__ bool tf_h_term(const Inst<real>& S, int j, real& v) {
  if (!S.tf_on) return false;
  const double* e = S.tf_line;
  if (j == S.tf_lo - 1) {
    v = (real)(-(S.tf_sg * S.tf_ku) * e[S.tf_lo]);
    return true;
  }
  if (S.tf_hi > 0 && j == S.tf_hi) {
    v = (real)((S.tf_sg * S.tf_ku) * e[S.tf_hi + 1]);
    return true;
  }
  return false;
}

// TMz E half step of cell (i, j); loads first (the pointers may alias as far
// as the compiler knows)
template <typename real>
__device__ __forceinline__ void e_tmz(const Inst<real>& S, int i, int j, real& carry, bool use_carry) {
  const int px = S.px;
  const int c = i + px * j;
  bool src = false;
  real sv = 0;
  for (int s = 0; s < S.nsrc; ++s)
    if (j == S.src_j[s]) {
      src = true;
      sv = S.src_val[s];
    }
  const bool pml = in_pml_tmz(S, i, j);
  real cv = 0;
  const bool corr = tf_d_term(S, j, cv);
  const real hxc = S.hx[c], hxj = use_carry ? carry : S.hx[c - px], hyc = S.hy[c], hyi = S.hy[c - 1];
  carry = hxc;  // Hx(j) is the next row's Hx(j-1)
  const Mat<real> m = S.tab[S.matid[c]];
  real ez = 0;
  bool wrote = false;
  if (!pml) {
    real dz = S.dz[c];
    dz = dz + S.ku * ((hyc - hyi) + (-hxc + hxj));
    if (src) dz = dz + sv;
    if (corr) dz = dz + cv;
    S.dz[c] = dz;
    if (m.act[0] == ACT_DIEL) {
      ez = dz * m.gax[0];
      wrote = true;
    }
  } else {
    const real r = m.gax[1];
    const real* fx = S.tprof;
    const real* fy = S.tprof + 4 * (S.nx + 2);
    // only a part whose axis's slab holds the cell is kept: a part not
    // kept is Ez less the other part, or, neither kept, Ezx is the whole of
    // Ez and Ezy 0 (the other axis's profile there is the interior's)
    const bool kx = S.Px > 0 && (i <= S.Px || i > S.nx - S.Px);
    const bool ky = S.P > 0 && (j <= S.P || j > S.ny - S.P);
    const real sx = kx ? S.ezx[c] : (real)0, sy = ky ? S.ezy[c] : (real)0;
    const real eo = (kx && ky) ? (real)0 : S.ez[c];
    const real ox = kx ? sx : (ky ? eo - sy : eo);
    const real oy = ky ? sy : (kx ? eo - sx : (real)0);
    const real zx = fx[i] * ox + (fx[(S.nx + 2) + i] * r) * (hyc + -hyi);
    real zy = fy[j] * oy + (fy[(S.ny + 2) + j] * r) * (-hxc + hxj);
    if (src) zy = zy + sv * r;
    if (corr) zy = zy + cv * r;
    if (kx) S.ezx[c] = zx;
    if (ky) S.ezy[c] = zy;
    ez = zx + zy;
    wrote = true;
  }
  if (wrote) {
    S.ez[c] = ez;
    if (S.per && i == 1) S.ez[c + S.nx] = ez;
    if (S.per && i == S.nx) S.ez[c - S.nx] = ez;
  }
}

// TMz H half step of cell (i, j)
template <typename real>
__device__ __forceinline__ void h_tmz(const Inst<real>& S, int i, int j, real& carry, bool use_carry) {
  const int px = S.px;
  const int c = i + px * j;
  const bool pml = in_pml_tmz(S, i, j);
  const real ezc = use_carry ? carry : S.ez[c], ezi = S.ez[c + 1], ezj = S.ez[c + px];
  carry = ezj;  // Ez(j+1) is the next row's Ez(j)
  const real hxc = S.hx[c], hyc = S.hy[c];
  real hx, hy;
  if (!pml) {
    hx = hxc + S.ku * (-ezj + ezc);
    hy = hyc + S.ku * (ezi - ezc);
  } else {
    const real* fx = S.tprof;
    const real* fy = S.tprof + 4 * (S.nx + 2);
    hx = fy[2 * (S.ny + 2) + j] * hxc + fy[3 * (S.ny + 2) + j] * (ezc + -ezj);
    hy = fx[2 * (S.nx + 2) + i] * hyc + fx[3 * (S.nx + 2) + i] * (-ezc + ezi);
  }
  real cv = 0;
  if (tf_h_term(S, j, cv)) hx = hx + cv;
  S.hx[c] = hx;
  S.hy[c] = hy;
  if (S.per && i == 1) S.hy[c + S.nx] = hy;
  if (S.per && i == S.nx) S.hy[c - S.nx] = hy;
}

// TEz E half step of cell (i, j); carry: Hz of the row below when use_carry
template <typename real>
__device__ __forceinline__ void e_tez(const Inst<real>& S, int i, int j, real& carry, bool use_carry) {
  const int px = S.px;
  const int c = i + px * j;
  const bool main = j >= S.j0 && j <= S.j1;
  bool src = false;
  real sv = 0;
  for (int s = 0; s < S.nsrc; ++s)
    if (j == S.src_j[s]) {
      src = true;
      sv = S.src_val[s];
    }
  const bool slab = S.has_pml && (j <= S.ja || j >= S.jb);
  real cv = 0;
  const bool corr = tf_d_term(S, j, cv);
  // loads first: the field pointers may alias as far as the compiler knows
  const real hzc = S.hz[c];
  real dxv = S.dx[c], dyv = S.dy[c];
  const real hzi = main ? S.hz[c - 1] : real(0);
  const real hzj = main ? (use_carry ? carry : S.hz[c - px]) : real(0);
  carry = hzc;
  const Mat<real> m = S.tab[S.matid[c]];
  if (main) {
    dxv = dxv + S.ku * (hzc - hzj);  // dx + Ku*curlhy
    dyv = dyv - S.ku * (hzc - hzi);  // dy - Ku*curlhx
  }
  if (src) dxv = dxv + sv;
  if (corr) dxv = dxv + cv;
  if (main) S.dy[c] = dyv;
  if (main || src || corr) S.dx[c] = dxv;
  // dielectric writers (graphene's Ey = Dy is one with gax 1)
  if (m.act[1] == ACT_DIEL) {
    const real ey = dyv * m.gax[1];
    S.ey[c] = ey;
    if (S.per && i == 1) S.ey[c + S.nx] = ey;
    if (S.per && i == S.nx) S.ey[c - S.nx] = ey;
  }
  if (slab) {
    const int s = j <= S.ja ? 0 : 1;
    const int r = j - S.prow0[s];
    const int sc = i + px * r;
    real* pex = S.pml[s][0];
    real* pey = S.pml[s][1];
    const real* hzx = S.pml[s][2];
    const real* hzy = S.pml[s][3];
    const real* py = S.prof[s];
    const real* pxp = py + 4 * (S.P + 2);
    const real curlHzx_p = -hzx[sc] - hzy[sc];
    const real curlHzx_m = hzx[sc - 1] + hzy[sc - 1];
    real curlHzy_p = hzx[sc] + hzy[sc];
    real curlHzy_m = -hzx[sc - px] - hzy[sc - px];
    if (j == S.ja) curlHzy_p = hzc;
    if (j == S.jb) curlHzy_m = -S.hz[c - px];
    const real ex = py[r] * pex[sc] + py[(S.P + 2) + r] * (curlHzy_p + curlHzy_m);
    const real ey = pxp[i] * pey[sc] + pxp[(S.nx + 2) + i] * (curlHzx_p + curlHzx_m);
    pex[sc] = ex;
    pey[sc] = ey;
    S.ex[c] = ex;  // setbounday{Eya,Eyb}: Ex of every slab row
  } else if (m.act[0] == ACT_DIEL) {
    S.ex[c] = dxv * m.gax[0];
  }
}

// TEz H half step of cell (i, j); carry: Ex of the row above when use_carry
template <typename real>
__device__ __forceinline__ void h_tez(const Inst<real>& S, int i, int j, real& carry, bool use_carry) {
  const int px = S.px;
  const int c = i + px * j;
  const bool main = j >= S.j0 && j <= S.j1;
  const bool slab = S.has_pml && (j <= S.ja || j >= S.jb);
  real hz = 0;
  bool wrote = false;
  const real exj = S.ex[c + px];
  if (main) {
    const real exc = use_carry ? carry : S.ex[c];
    const real curlE = S.ey[c + 1] - S.ey[c] - exj + exc;
    hz = S.hz[c] - S.ku * curlE;
    wrote = true;
    real cv = 0;
    if (tf_h_term(S, j, cv)) hz = hz + cv;
  }
  carry = exj;  // Ex(j+1) is the next row's Ex(j)
  if (slab) {
    const int s = j <= S.ja ? 0 : 1;
    const int r = j - S.prow0[s];
    const int sc = i + px * r;
    const real* pex = S.pml[s][0];
    const real* pey = S.pml[s][1];
    real* hzx = S.pml[s][2];
    real* hzy = S.pml[s][3];
    const real* py = S.prof[s];
    const real* pxp = py + 4 * (S.P + 2);
    const real curlExy = pex[sc + px] - pex[sc];
    const real curlEyx = -pey[sc + 1] + pey[sc];
    const real ny_ = py[2 * (S.P + 2) + r] * hzy[sc] + py[3 * (S.P + 2) + r] * curlExy;
    const real nx_ = pxp[2 * (S.nx + 2) + i] * hzx[sc] + pxp[3 * (S.nx + 2) + i] * curlEyx;
    hzy[sc] = ny_;
    hzx[sc] = nx_;
    if (j <= S.ja - 1 || j >= S.jb) {  // setbounday{Hya,Hyb}
      hz = ny_ + nx_;
      wrote = true;
    }
  }
  if (wrote) {
    S.hz[c] = hz;
    if (S.per && i == 1) S.hz[c + S.nx] = hz;
    if (S.per && i == S.nx) S.hz[c - S.nx] = hz;
  }
}

// Logical block of this block.  Blocks are dealt to the 8 XCDs (one L2
// each) round robin; renumbered, each XCD takes runs of kXcdRun consecutive
// logical blocks, so the x neighbours of a tile (whose edge lines it reads)
// mostly run on the same XCD and come from its L2, while the runs stay
// short enough that every XCD gets a like mix of tiles (TEz/TMz, PML rows).
// The grid is a multiple of 8 * kXcdRun (set at launch).
constexpr int kXcdRun = 32;
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, k = b >> 3;
  return (k / kXcdRun) * (8 * kXcdRun) + (b & 7) * kXcdRun + k % kXcdRun;
}

// ---- plain row strips ------------------------------------------------------------
// A wavefront's kRowsPerWave rows are "plain" when none is a PML, source or
// TFSF row and (TMz) none of its lanes lies in an x slab: the cell update is
// then the interior one of e_tez / h_tez / e_tmz / h_tmz, and the loads of
// all rows are issued before the first row's stores (one memory round trip
// for the strip instead of one per row).  The arithmetic is the same,
// operation for operation.
template <typename real>
__device__ __forceinline__ bool src_row(const Inst<real>& S, int j) {
  bool src = false;
  for (int s = 0; s < S.nsrc; ++s) src |= j == S.src_j[s];
  return src;
}
template <typename real>
__device__ __forceinline__ bool strip_plain(const Inst<real>& S, int i0, int j0, bool e_half) {
  const int j1 = j0 + kRowsPerWave - 1;
  if (j1 > S.ny) return false;
  if (S.tmz) {
    if (S.P > 0 && (j0 <= S.P || j1 > S.ny - S.P)) return false;
    if (S.Px > 0 && (i0 + 1 <= S.Px || std::min(i0 + kTX, S.nx) > S.nx - S.Px)) return false;
  } else {
    if (j0 < S.j0 || j1 > S.j1) return false;
    if (S.has_pml && (j0 <= S.ja || j1 >= S.jb)) return false;
  }
  if (S.tf_on) {
    const int a = e_half ? S.tf_lo : S.tf_lo - 1;
    const int b = e_half ? S.tf_hi + 1 : S.tf_hi;
    if (a >= j0 && a <= j1) return false;
    if (S.tf_hi > 0 && b >= j0 && b <= j1) return false;
  }
  if (e_half)
    for (int j = j0; j <= j1; ++j)
      if (src_row(S, j)) return false;
  return true;
}

template <typename real>
__device__ __forceinline__ void e_tez_strip(const Inst<real>& S, int i, int j0) {
  constexpr int R = kRowsPerWave;
  const int px = S.px;
  const int c0 = i + px * j0;
  real hz[R + 1], hzi[R], dx[R], dy[R];
  uint8_t id[R];
  hz[0] = S.hz[c0 - px];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + px * r;
    hz[r + 1] = S.hz[c];
    hzi[r] = S.hz[c - 1];
    dx[r] = S.dx[c];
    dy[r] = S.dy[c];
    id[r] = S.matid[c];
  }
  Mat<real> m[R];
#pragma unroll
  for (int r = 0; r < R; ++r) m[r] = S.tab[id[r]];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + px * r;
    const real dxv = dx[r] + S.ku * (hz[r + 1] - hz[r]);
    const real dyv = dy[r] - S.ku * (hz[r + 1] - hzi[r]);
    S.dy[c] = dyv;
    S.dx[c] = dxv;
    if (m[r].act[1] == ACT_DIEL) {
      const real ey = dyv * m[r].gax[1];
      S.ey[c] = ey;
      if (S.per && i == 1) S.ey[c + S.nx] = ey;
      if (S.per && i == S.nx) S.ey[c - S.nx] = ey;
    }
    if (m[r].act[0] == ACT_DIEL) S.ex[c] = dxv * m[r].gax[0];
  }
}

template <typename real>
__device__ __forceinline__ void h_tez_strip(const Inst<real>& S, int i, int j0) {
  constexpr int R = kRowsPerWave;
  const int px = S.px;
  const int c0 = i + px * j0;
  real ex[R + 1], ey[R], eyi[R], hz[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + px * r;
    ex[r] = S.ex[c];
    ey[r] = S.ey[c];
    eyi[r] = S.ey[c + 1];
    hz[r] = S.hz[c];
  }
  ex[R] = S.ex[c0 + px * R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + px * r;
    const real curlE = eyi[r] - ey[r] - ex[r + 1] + ex[r];
    const real h = hz[r] - S.ku * curlE;
    S.hz[c] = h;
    if (S.per && i == 1) S.hz[c + S.nx] = h;
    if (S.per && i == S.nx) S.hz[c - S.nx] = h;
  }
}

template <typename real>
__device__ __forceinline__ void e_tmz_strip(const Inst<real>& S, int i, int j0) {
  constexpr int R = kRowsPerWave;
  const int px = S.px;
  const int c0 = i + px * j0;
  real hx[R + 1], hy[R], hyi[R], dz[R];
  uint8_t id[R];
  hx[0] = S.hx[c0 - px];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + px * r;
    hx[r + 1] = S.hx[c];
    hy[r] = S.hy[c];
    hyi[r] = S.hy[c - 1];
    dz[r] = S.dz[c];
    id[r] = S.matid[c];
  }
  Mat<real> m[R];
#pragma unroll
  for (int r = 0; r < R; ++r) m[r] = S.tab[id[r]];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + px * r;
    const real d = dz[r] + S.ku * ((hy[r] - hyi[r]) + (-hx[r + 1] + hx[r]));
    S.dz[c] = d;
    if (m[r].act[0] == ACT_DIEL) {
      const real ez = d * m[r].gax[0];
      S.ez[c] = ez;
      if (S.per && i == 1) S.ez[c + S.nx] = ez;
      if (S.per && i == S.nx) S.ez[c - S.nx] = ez;
    }
  }
}

template <typename real>
__device__ __forceinline__ void h_tmz_strip(const Inst<real>& S, int i, int j0) {
  constexpr int R = kRowsPerWave;
  const int px = S.px;
  const int c0 = i + px * j0;
  real ez[R + 1], ezi[R], hx[R], hy[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + px * r;
    ez[r] = S.ez[c];
    ezi[r] = S.ez[c + 1];
    hx[r] = S.hx[c];
    hy[r] = S.hy[c];
  }
  ez[R] = S.ez[c0 + px * R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = c0 + px * r;
    const real x = hx[r] + S.ku * (-ez[r + 1] + ez[r]);
    const real y = hy[r] + S.ku * (ezi[r] - ez[r]);
    S.hx[c] = x;
    S.hy[c] = y;
    if (S.per && i == 1) S.hy[c + S.nx] = y;
    if (S.per && i == S.nx) S.hy[c - S.nx] = y;
  }
}

// E and H over the packed tile list: block = one tile at a time, wavefront
// w marches rows 4w..4w+3 of the tile (j uniform per wavefront).
template <typename real>
__global__ __launch_bounds__(kTX * kWaves) void k2_e(const Inst<real>* __restrict__ I,
                                                     const int32_t* __restrict__ tile_inst, int ntiles) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (int t = xcd_block(); t < ntiles; t += gridDim.x) {
    const int q = tile_inst[t];
    const Inst<real>& S = I[q];
    if (S.ctl[0] >= S.ctl[1]) continue;  // finished: nothing to do this step
    const int lt = t - S.tile0;
    const int i = (lt % S.ntx) * kTX + lane + 1;
    const int j0 = (lt / S.ntx) * kTY + wave * kRowsPerWave + 1;
    if (i > S.nx) continue;
    if (strip_plain(S, (lt % S.ntx) * kTX, j0, true)) {
      if (S.tmz)
        e_tmz_strip(S, i, j0);
      else
        e_tez_strip(S, i, j0);
      continue;
    }
    real carry = 0;
    for (int r = 0; r < kRowsPerWave; ++r) {
      const int j = j0 + r;
      if (j > S.ny) break;
      if (S.tmz)
        e_tmz(S, i, j, carry, r > 0);
      else
        e_tez(S, i, j, carry, r > 0);
    }
  }
}

template <typename real>
__global__ __launch_bounds__(kTX * kWaves) void k2_h(const Inst<real>* __restrict__ I,
                                                     const int32_t* __restrict__ tile_inst, int ntiles) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (int t = xcd_block(); t < ntiles; t += gridDim.x) {
    const int q = tile_inst[t];
    const Inst<real>& S = I[q];
    if (S.ctl[0] >= S.ctl[1]) continue;
    const int lt = t - S.tile0;
    const int i = (lt % S.ntx) * kTX + lane + 1;
    const int j0 = (lt / S.ntx) * kTY + wave * kRowsPerWave + 1;
    if (i > S.nx) continue;
    if (strip_plain(S, (lt % S.ntx) * kTX, j0, false)) {
      if (S.tmz)
        h_tmz_strip(S, i, j0);
      else
        h_tez_strip(S, i, j0);
      continue;
    }
    real carry = 0;
    for (int r = 0; r < kRowsPerWave; ++r) {
      const int j = j0 + r;
      if (j > S.ny) break;
      if (S.tmz)
        h_tmz(S, i, j, carry, r > 0);
      else
        h_tez(S, i, j, carry, r > 0);
    }
  }
}

// Dispersive cells (mod2/gold_mod.f90:54-73, mod2/graphene_mod.f90:63-92):
// E = Re[(D - I + ha S + sum_p gb_p g_p) / (ga + Sn)], then the state.
template <typename real>
__global__ __launch_bounds__(256) void k2_ade(const Inst<real>* __restrict__ I, const int4* __restrict__ work,
                                              int nwork) {
  const int w = blockIdx.x;
  if (w >= nwork) return;
  const int4 wk = work[w];
  const Inst<real>& S = I[wk.x];
  if (S.ctl[0] >= S.ctl[1]) return;
  const AdeDev<real>& G = S.grp[wk.y];
  const int e = wk.z + (int)threadIdx.x;
  if (e >= G.ncells) return;
  const int n = G.ncells;
  const int c = G.cells[e];
  const int mask = G.mask[e];
  const AdeCoef<double>& co = G.co;
  const int ncp = co.ncp;
  const int i = c % S.px;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    if (!(mask & (1 << q))) continue;
    const double dv = S.tmz ? S.dz[c] : (q == 0 ? S.dx[c] : S.dy[c]);
    double* Ip = G.state + (int64_t)(0 * 2 + q) * n + e;
    double* Sp = G.state + (int64_t)(1 * 2 + q) * n + e;
    const double Iv = *Ip, Sv = *Sp;
    double xr = (dv - Iv) + co.ha * Sv;
    double xi = 0;
    double gr[4], gi[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      if (p >= ncp) break;
      gr[p] = G.state[(int64_t)((2 + p) * 2 + q) * n + e];
      gi[p] = G.state[(int64_t)((2 + ncp + p) * 2 + q) * n + e];
      xr = xr + (co.gbr[p] * gr[p] - co.gbi[p] * gi[p]);
      xi = xi + (co.gbr[p] * gi[p] + co.gbi[p] * gr[p]);
    }
    double s1 = 0, s2 = 0;
    double* s1p = nullptr;
    double* s2p = nullptr;
    if (co.raman) {
      s1p = G.state + (int64_t)((2 + 2 * ncp) * 2 + q) * n + e;
      s2p = G.state + (int64_t)((3 + 2 * ncp) * 2 + q) * n + e;
      s1 = *s1p;
      s2 = *s2p;
    }
    const double dr = co.gar + s1, di = co.gai;
    const double ev = (xr * dr + xi * di) / (dr * dr + di * di);
    real* E = S.tmz ? S.ez : (q == 0 ? S.ex : S.ey);
    E[c] = (real)ev;
    const bool wrap = S.per && (q == 1 || S.tmz);  // Ey (TEz) and Ez (TMz) wrap in x
    if (wrap && i == 1) E[c + S.nx] = (real)ev;
    if (wrap && i == S.nx) E[c - S.nx] = (real)ev;
    if (co.raman) {  // graphene_mod.f90:74-77 (Sn_new; shift)
      const double ns = co.r1 * s1 + co.r2 * s2 + co.r0 * (ev * ev);
      *s2p = s1;
      *s1p = ns;
    }
    *Ip = Iv + co.hb * ev;
    *Sp = co.ha * Sv + co.hb * ev;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      if (p >= ncp) break;
      G.state[(int64_t)((2 + p) * 2 + q) * n + e] = (co.gbr[p] * gr[p] - co.gbi[p] * gi[p]) + co.gcr[p] * ev;
      G.state[(int64_t)((2 + ncp + p) * 2 + q) * n + e] = (co.gbr[p] * gi[p] + co.gbi[p] * gr[p]) + co.gci[p] * ev;
    }
  }
}

// Lorentz-Jp cells, E form (mod2/filmGoldLorentzJp_mod.f90:35-75,
// siLorentzJp_mod.f90, graphene3LorentzJp1nm_mod.f90: E_q += Ku*ga*(curl_q -
// ds*JpT_q), then the poles from the new E).  E is read as the previous
// step left it (k2_e does not write these components), the update runs in
// FP64, the line source of the row adds ga*g.
template <typename real>
__global__ __launch_bounds__(256) void k2_ljp(const Inst<real>* __restrict__ I, const int4* __restrict__ work,
                                              int nwork) {
  const int w = blockIdx.x;
  if (w >= nwork) return;
  const int4 wk = work[w];
  const Inst<real>& S = I[wk.x];
  if (S.ctl[0] >= S.ctl[1]) return;
  const LjpDev& G = S.ljp[wk.y];
  const int e = wk.z + (int)threadIdx.x;
  if (e >= G.ncells) return;
  const int n = G.ncells;
  const int c = G.cells[e];
  const int mask = G.mask[e];
  const LjpCoef<double>& co = G.co;
  const int px = S.px;
  const int i = c % px, j = c / px;
  real cur[2];
  real* E[2];
  if (S.tmz) {
    cur[0] = (S.hy[c] - S.hy[c - 1]) + (-S.hx[c] + S.hx[c - px]);  // curlhyx + curlhxy
    cur[1] = 0;
    E[0] = S.ez;
    E[1] = nullptr;
  } else {
    cur[0] = S.hz[c] - S.hz[c - px];   // curlhzy
    cur[1] = -S.hz[c] + S.hz[c - 1];   // curlhzx
    E[0] = S.ex;
    E[1] = S.ey;
  }
  real sv = 0;
  bool hit = false;  // the line source drives local component 0 (Dx or Dz)
  for (int s = 0; s < S.nsrc; ++s)
    if (j == S.src_j[s]) {
      sv += S.src_val[s];
      hit = true;
    }
  const int np = co.npols;
  const int wq = 2 * np;  // word of Q
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    if (!(mask & (1 << q))) continue;
    double jt = 0;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      if (l >= np) break;
      jt = jt + G.state[(int64_t)((np + l) * 2 + q) * n + e];
    }
    double ev = (double)E[q][c] + co.kuga * ((double)cur[q] - co.ds * jt);
    if (q == 0 && hit) ev = ev + co.ga * (double)sv;
    const real er = (real)ev;
    E[q][c] = er;
    const bool wrap = S.per && (S.tmz || q == 1);  // Ey (TEz) and Ez (TMz) wrap in x
    if (wrap && i == 1) E[q][c + S.nx] = er;
    if (wrap && i == S.nx) E[q][c - S.nx] = er;
    double Qv = 0;
    if (co.nl) Qv = G.state[(int64_t)(wq * 2 + q) * n + e];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      if (l >= np) break;
      double* Pp = G.state + (int64_t)(l * 2 + q) * n + e;
      double* Jp = G.state + (int64_t)((np + l) * 2 + q) * n + e;
      double p0 = (co.xi[l] * ev) / co.c0;
      if (l == 0 && co.nl) p0 = p0 + ((co.xi3 * Qv) * ev) / co.c0;
      const double jn = co.g1[l] * *Jp + co.g2[l] * (p0 - *Pp);
      *Jp = jn;
      *Pp = *Pp + co.dt * jn;
    }
    if (co.nl) {
      double* Gp = G.state + (int64_t)((wq + 1) * 2 + q) * n + e;
      const double ze = co.z0 * ev;
      const double gn = co.ng1 * *Gp + co.ng2 * (ze * ze - Qv);
      *Gp = gn;
      G.state[(int64_t)(wq * 2 + q) * n + e] = Qv + co.dt * gn;
    }
  }
}

// Monitors after step n = ctl[0] + 1: DFT line samples into ring slot b
// (sample step recorded, 0 when the line's gate is closed) and probes.
template <typename real>
__global__ __launch_bounds__(256) void k2_mon(const Inst<real>* __restrict__ I, const int4* __restrict__ work,
                                              int nwork, int b) {
  const int w = blockIdx.x;
  if (w >= nwork) return;
  const int4 wk = work[w];
  const Inst<real>& S = I[wk.x];
  if (S.ctl[0] >= S.ctl[1]) return;
  const int n = S.ctl[0] + 1;
  if (wk.y < 0) {  // probes: (n*dt, Ex, Hz), sim2/program.f90:376,397,417; TMz (n*dt, Ez, Hx)
    const int p = threadIdx.x;
    if (p < S.nprobe && n - 1 < S.probe_cap) {
      double* r = S.probe_rec + ((int64_t)(n - 1) * S.nprobe + p) * 3;
      r[0] = (double)n * S.dt;
      r[1] = (double)(S.tmz ? S.ez : S.ex)[S.probe_c[p]];
      r[2] = (double)(S.tmz ? S.hx : S.hz)[S.probe_c[p]];
    }
    return;
  }
  const DftDev& D = S.dft[wk.y];
  const bool open = n >= D.n_begin && (D.n_end <= 0 || n <= D.n_end);
  if (wk.z == 0 && threadIdx.x == 0) D.sstep[b] = open ? n : 0;
  if (!open) return;
  const int i = wk.z + (int)threadIdx.x + 1;
  if (i > S.nx) return;
  const int c = i + S.px * D.j;
  real* samp = (real*)D.samp + (int64_t)b * D.nslots * S.nx;
  int slot = 0;
  if (S.tmz) {
    if (D.mask & FDTD2D_DFT_EZ) samp[(int64_t)(slot++) * S.nx + (i - 1)] = S.ez[c];
    if (D.mask & FDTD2D_DFT_HX) samp[(int64_t)slot * S.nx + (i - 1)] = S.hx[c];
    return;
  }
  if (D.mask & FDTD2D_DFT_EX) samp[(int64_t)(slot++) * S.nx + (i - 1)] = S.ex[c];
  if (D.mask & FDTD2D_DFT_HZ) samp[(int64_t)slot * S.nx + (i - 1)] = S.hz[c];
}

// Step advance: ctl[0] += 1 for every unfinished simulation, and the line
// sources' values for the step that follows (sim2/program.f90:337,
// t = (n-1)*dt).  inc = 0 only computes the values.
template <typename real>
__global__ void k2_adv(const Inst<real>* __restrict__ I, int count, int inc) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const Inst<real>& S = I[q];
  int nd = S.ctl[0];
  if (inc) {
    if (nd >= S.ctl[1]) return;
    S.ctl[0] = ++nd;
  }
  const double t = (double)nd * S.dt;  // the next step is nd + 1
  for (int s = 0; s < S.nsrc; ++s) {
    const double a = (t - S.src_t0[s]) / S.src_st[s];
    S.src_val[s] = (real)(S.src_amp[s] * (exp(-0.5 * (a * a)) * sin(S.src_w0[s] * t)));
  }
}

// TFSF incident line, one block per simulation with a TFSF (list[b]): E half
// step (ex_inc update, the pulse at t = (n-1)*dt, mypmly_inc%updateE with
// the hand-over of its end slabs; sim1/program.f90:222-231) and H half step
// (hz_inc update, mypmly_inc%updateH; :252-255).
template <typename real>
__global__ __launch_bounds__(256) void k2_tfsf_line(const Inst<real>* __restrict__ I, const int32_t* __restrict__ list,
                                                    int n, int h_half) {
  const int w = blockIdx.x;
  if (w >= n) return;
  const Inst<real>& S = I[list[w]];
  if (S.ctl[0] >= S.ctl[1]) return;
  const int N = S.ny, P = S.tf_npml, hi = N - P + 1;
  const int64_t m = N + 2;
  double* e = S.tf_line;
  double* h = e + m;
  double* pe = e + 2 * m;
  double* ph = e + 3 * m;
  const double* f1 = e + 4 * m;
  const double* f2 = e + 5 * m;
  const double* g1 = e + 6 * m;
  const double* g2 = e + 7 * m;
  if (!h_half) {
    const double t = (double)S.ctl[0] * S.dt;  // step ctl[0]+1
    const double a = (t - S.tf_t0) / S.tf_st;
    const double g = S.tf_amp * (exp(-0.5 * (a * a)) * sin(S.tf_w0 * t));
    for (int j = P + 1 + (int)threadIdx.x; j <= hi - 1; j += blockDim.x) {
      double v = e[j] + S.tf_ck * (h[j] - h[j - 1]);
      if (j == S.tf_src) v = v + g;
      e[j] = v;
    }
    if (P > 0) {
      for (int j = 1 + (int)threadIdx.x; j <= P; j += blockDim.x) pe[j] = f1[j] * pe[j] + f2[j] * (ph[j] - ph[j - 1]);
      for (int j = hi + (int)threadIdx.x; j <= N; j += blockDim.x) pe[j] = f1[j] * pe[j] + f2[j] * (ph[j] - ph[j - 1]);
      __syncthreads();
      if (threadIdx.x == 0) {
        pe[P + 1] = e[P + 1];
        e[P] = pe[P];
        pe[hi - 1] = e[hi - 1];
        e[hi] = pe[hi];
      }
    }
  } else {
    for (int j = P + 1 + (int)threadIdx.x; j <= hi - 1; j += blockDim.x) h[j] = h[j] + S.tf_ku * (e[j + 1] - e[j]);
    if (P > 0) {
      for (int j = 1 + (int)threadIdx.x; j <= P; j += blockDim.x) ph[j] = g1[j] * ph[j] + g2[j] * (pe[j + 1] - pe[j]);
      for (int j = hi + (int)threadIdx.x; j <= N; j += blockDim.x) ph[j] = g1[j] * ph[j] + g2[j] * (pe[j + 1] - pe[j]);
      __syncthreads();
      if (threadIdx.x == 0) {
        ph[P + 1] = h[P + 1];
        h[P] = ph[P];
        ph[hi - 1] = h[hi - 1];
        h[hi] = ph[hi];
      }
    }
  }
}

// DFT fold of one ring of kBlockSteps samples: acc += dt * sum_b
// (cos, sin)(omega t_b) F_b with t_b = (n_b - 1) dt, FP64.
constexpr int kMC = 32;  // frequencies per LDS table chunk

template <typename real>
__global__ __launch_bounds__(256) void k2_flush(const Inst<real>* __restrict__ I, const int4* __restrict__ work,
                                                int nwork) {
  __shared__ double tab[kBlockSteps * kMC * 2];
  __shared__ int32_t st[kBlockSteps];
  const int w = blockIdx.x;
  if (w >= nwork) return;
  const int4 wk = work[w];
  if (wk.y < 0) return;  // the probe work item
  const Inst<real>& S = I[wk.x];
  const DftDev& D = S.dft[wk.y];
  if (threadIdx.x < kBlockSteps) st[threadIdx.x] = D.sstep[threadIdx.x];
  __syncthreads();
  const int i = wk.z + (int)threadIdx.x + 1;
  const bool act = i <= S.nx;
  const int nx = S.nx;
  for (int slot = 0; slot < D.nslots; ++slot) {
    double f[kBlockSteps];
#pragma unroll
    for (int b = 0; b < kBlockSteps; ++b)
      f[b] = (act && st[b] > 0) ? (double)((const real*)D.samp)[((int64_t)b * D.nslots + slot) * nx + (i - 1)] : 0.0;
    double* re = D.acc + (int64_t)slot * 2 * D.nfreq * nx;
    double* im = re + (int64_t)D.nfreq * nx;
    for (int m0 = 0; m0 < D.nfreq; m0 += kMC) {
      const int mc = min(kMC, D.nfreq - m0);
      __syncthreads();
      for (int e = threadIdx.x; e < kBlockSteps * mc; e += blockDim.x) {
        const int b = e / mc, mm = e - b * mc;
        double sw = 0.0, cw = 0.0;
        if (st[b] > 0) sincos(D.omega[m0 + mm] * ((double)(st[b] - 1) * S.dt), &sw, &cw);
        tab[2 * (b * kMC + mm)] = cw;
        tab[2 * (b * kMC + mm) + 1] = sw;
      }
      __syncthreads();
      if (!act) continue;
      for (int mm = 0; mm < mc; ++mm) {
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int b = 0; b < kBlockSteps; ++b) {
          sr += tab[2 * (b * kMC + mm)] * f[b];
          si += tab[2 * (b * kMC + mm) + 1] * f[b];
        }
        const int64_t o = (int64_t)(m0 + mm) * nx + (i - 1);
        re[o] += sr * S.dt;
        im[o] += si * S.dt;
      }
    }
  }
}

// The ring is folded: clear the sample steps (a kernel, not a memset, so
// the node replays safely in the captured graph).
template <typename real>
__global__ void k2_clear(const Inst<real>* __restrict__ I, const int2* __restrict__ lines, int nlines) {
  const int l = blockIdx.x;
  if (l >= nlines) return;
  const int2 ln = lines[l];
  if (threadIdx.x < kBlockSteps) I[ln.x].dft[ln.y].sstep[threadIdx.x] = 0;
}

// ---- host: set-up ------------------------------------------------------------------
int region_bounds(const Sim* s, const fdtd2d_region& r, int lo[2], int hi[2]) {
  if (r.kind == FDTD2D_REGION_BOX) {
    lo[0] = r.lo[0]; lo[1] = r.lo[1]; hi[0] = r.hi[0]; hi[1] = r.hi[1];
  } else if (r.kind == FDTD2D_REGION_DISK) {
    if (!(r.radius > 0)) return 0;
    for (int a = 0; a < 2; ++a) {
      lo[a] = (int)std::ceil(r.center[a] - r.radius);
      hi[a] = (int)std::floor(r.center[a] + r.radius);
    }
  } else {
    return 0;
  }
  const int n[2] = {s->nx, s->ny};
  for (int a = 0; a < 2; ++a) {  // mod2/material_mod.f90:53-93 (one rank)
    lo[a] = std::max(lo[a], 1);
    hi[a] = std::min(hi[a], n[a]);
  }
  return 1;
}

bool region_inside(const fdtd2d_region& r, int i, int j) {
  if (r.kind != FDTD2D_REGION_DISK) return true;
  const double dx = i - r.center[0], dy = j - r.center[1];
  return dx * dx + dy * dy <= r.radius * r.radius;
}

void drude2cp_coef(const Sim* s, const fdtd_drude2cp& p, AdeCoef<double>& c) {
  // setexp + initialize_GoldVars, mod2/gold_mod.f90:129-168
  const double dt = s->dt;
  const cplx I(0.0, 1.0);
  std::memset(&c, 0, sizeof c);
  c.ncp = 2;
  c.ku = s->ku;
  c.ha = std::exp(-p.gamma * dt);
  c.hb = (p.omega_d * p.omega_d) * dt / p.gamma;
  cplx ga = p.eps_inf;
  for (int q = 0; q < 2; ++q) {
    const cplx gb = std::exp((-p.Gamma[q] + I * p.Omega[q]) * dt);
    const cplx gc = dt * 2.0 * I * p.A[q] * p.Omega[q] * std::exp(-I * p.phi[q]);
    c.gbr[q] = gb.real(); c.gbi[q] = gb.imag();
    c.gcr[q] = gc.real(); c.gci[q] = gc.imag();
    ga = ga - gc;
  }
  c.gar = ga.real();
  c.gai = ga.imag();
}

int graphene_coef(const Sim* s, const fdtd_graphene& p, AdeCoef<double>& c) {
  const Consts& K = consts();
  const double dt = s->dt;
  const cplx I(0.0, 1.0);
  std::memset(&c, 0, sizeof c);
  c.ncp = 4;
  c.ku = s->ku;
  if (p.model == FDTD_GRAPHENE_CP4) {
    // setexp, scaleParams, initialize_GrapheneVars: mod2/graphene_mod.f90:163-236
    const double scale = p.thickness / s->ds;
    const double wD = std::sqrt(scale) * p.omega_d;
    c.ha = std::exp(-p.gamma * dt);
    c.hb = (wD * wD) * dt / p.gamma;
    cplx ga = p.eps_inf;
    for (int q = 0; q < 4; ++q) {
      const cplx gb = std::exp((-p.Gamma[q] + I * p.Omega[q]) * dt);
      const cplx gc = dt * 2.0 * I * (scale * p.A[q]) * p.Omega[q] * std::exp(-I * p.phi[q]);
      c.gbr[q] = gb.real(); c.gbi[q] = gb.imag();
      c.gcr[q] = gc.real(); c.gci[q] = gc.imag();
      ga = ga - gc;
    }
    c.gar = ga.real();
    c.gai = ga.imag();
  } else if (p.model == FDTD_GRAPHENE_KUBO) {
    // Kubo intraband sheet spread over one cell: eps_inf - wp^2/(w^2 + i w/tau)
    if (!(p.tau > 0) || !(p.temperature > 0)) return set_error(FDTD_ERR_ARG, "tau and T must be > 0");
    const double kT = K.kb * p.temperature;
    const double ef = p.fermi_ev * K.echarge;
    const double Dw = 2.0 * K.echarge * K.echarge * kT / (K.hbar * K.hbar) * std::log(2.0 * std::cosh(ef / (2.0 * kT)));
    const double wp2 = Dw / (K.pi * K.ep0 * s->ds);
    const double gam = 1.0 / p.tau;
    c.ha = std::exp(-gam * dt);
    c.hb = wp2 * dt / gam;
    c.gar = p.eps_inf;
  } else {
    return set_error(FDTD_ERR_ARG, "bad graphene model");
  }
  // Raman, central differences: mod2/graphene_mod.f90:210-212 (always set)
  c.raman = 1;
  c.r0 = p.xi3 * ((p.wNL * dt) * (p.wNL * dt)) / (1.0 + p.dnuNL * dt);
  c.r1 = (2.0 - (p.wNL * dt) * (p.wNL * dt)) / (1.0 + p.dnuNL * dt);
  c.r2 = -(1.0 - p.dnuNL * dt) / (1.0 + p.dnuNL * dt);
  return 0;
}

// pmly%initialize on one domain (mod2/BERpmly_HzExEy_mod.f90:229-426):
// per slab, y profiles by slab row and x profiles by column.
void pml_profiles(const Sim* s, int slab, std::vector<double>& py, std::vector<double>& pxv) {
  const Consts& K = consts();
  const int nx = s->nx, ny = s->ny, P = s->P, Px = s->Px;
  const double ep = (slab == 0 ? s->pml.eps_low : s->pml.eps_high) * K.ep0;
  const double epr = slab == 0 ? s->pml.eps_low : s->pml.eps_high;
  const double dt = s->dt, ds = s->ds, order = s->pml.order, sigmam = s->sigmam;
  const double Ku = dt * K.c0 / ds;
  const int W = P + 2, X = nx + 2;
  py.assign(4 * (size_t)W, 0.0);
  pxv.assign(4 * (size_t)X, 0.0);
  // defaults (allocate_vars :284-309): f1 = g1 = 1, g2 = Ku, f2 = Ku/eps of
  // the slab's half of the domain (both slabs lie in their halves)
  for (int r = 0; r < W; ++r) { py[r] = 1.0; py[W + r] = Ku / epr; py[2 * W + r] = 1.0; py[3 * W + r] = Ku; }
  for (int i = 0; i < X; ++i) { pxv[i] = 1.0; pxv[X + i] = Ku / epr; pxv[2 * X + i] = 1.0; pxv[3 * X + i] = Ku; }
  if (slab == 0) {  // rows 1..P (:343-357); slab row r = j
    for (int j = 1; j <= P; ++j) {
      const double sE = ep * sigmam * std::pow(((double)(P - j) + 0.5) / (double)P, order);
      py[j] = std::exp(-(sE * dt / ep));
      py[W + j] = (1.0 - py[j]) / (K.z0 * sE * ds);
      if (j < P) {
        const double sH = K.mu0 * sigmam * std::pow((double)(P - j) / (double)P, order);
        py[2 * W + j] = std::exp(-(sH * dt / K.mu0));
        py[3 * W + j] = K.z0 * (1.0 - py[2 * W + j]) / (sH * ds);
      }
    }
  } else {  // rows jb..ny (:358-374); slab row r = j - (jb - 1)
    int jj = P;
    for (int j = s->jb; j <= ny; ++j) {
      const int r = j - (s->jb - 1);
      if (jj < P) {
        const double sE = ep * sigmam * std::pow((double)(P - jj) / (double)P, order);
        py[r] = std::exp(-(sE * dt / ep));
        py[W + r] = (1.0 - py[r]) / (K.z0 * sE * ds);
      }
      const double sH = K.mu0 * sigmam * std::pow(((double)(P - jj) + 0.5) / (double)P, order);
      py[2 * W + r] = std::exp(-(sH * dt / K.mu0));
      py[3 * W + r] = K.z0 * (1.0 - py[2 * W + r]) / (sH * ds);
      --jj;
    }
  }
  if (Px > 0) {  // init_varx1 / init_varx2 in the slab (:377-425)
    for (int i = 1; i <= Px; ++i) {
      const double sE = ep * sigmam * std::pow(((double)(Px - i) + 0.5) / (double)Px, order);
      pxv[i] = std::exp(-(sE * dt / ep));
      pxv[X + i] = (1.0 - pxv[i]) / (K.z0 * sE * ds);
      if (i < Px) {
        const double sH = K.mu0 * sigmam * std::pow((double)(Px - i) / (double)Px, order);
        pxv[2 * X + i] = std::exp(-(sH * dt / K.mu0));
        pxv[3 * X + i] = K.z0 * (1.0 - pxv[2 * X + i]) / (sH * ds);
      }
    }
    int ii = Px;
    for (int i = nx - Px + 1; i <= nx; ++i) {
      if (ii < Px) {
        const double sE = ep * sigmam * std::pow((double)(Px - ii) / (double)Px, order);
        pxv[i] = std::exp(-(sE * dt / ep));
        pxv[X + i] = (1.0 - pxv[i]) / (K.z0 * sE * ds);
      }
      const double sH = K.mu0 * sigmam * std::pow(((double)(Px - ii) + 0.5) / (double)Px, order);
      pxv[2 * X + i] = std::exp(-(sH * dt / K.mu0));
      pxv[3 * X + i] = K.z0 * (1.0 - pxv[2 * X + i]) / (sH * ds);
      --ii;
    }
  }
}

template <typename real>
void* upload_cast(Sim* s, const std::vector<double>& v) {
  std::vector<real> h(v.begin(), v.end());
  return upload(s->allocs, &s->dev_bytes, h.data(), h.size() * sizeof(real));
}

int commit_impl(Sim* s) {
  FDTD_HIP(hipSetDevice(s->device));
  const int nx = s->nx, ny = s->ny;
  // layout: rows of whole 128-byte lines, interior i = 1 at the start of a line
  const int64_t le = 128 / (int64_t)s->wsize;
  s->px = ((int64_t)nx + 2 + le - 1) / le * le;
  s->ncell = s->px * (ny + 2);
  if (s->ncell + le >= ((int64_t)1 << 31)) return set_error(FDTD_ERR_ARG, "grid too large");
  const size_t shift = (size_t)(le - 1) * s->wsize;
  const size_t fbytes = (size_t)(s->ncell + le) * s->wsize;
  for (int f = 0; f < (s->tmz ? 4 : 5); ++f) {
    char* p = (char*)alloc_in(s->allocs, &s->dev_bytes, fbytes);
    if (!p) return set_error(FDTD_ERR_NOMEM, "field allocation");
    s->F[f] = p + shift;
  }
  if (s->tmz) {
    // the 3D engine's ALL_FACES PML on the x and y faces (vacuum grading,
    // defaults f1 = g1 = 1, f2 = g2 = Ku)
    const Consts& K = consts();
    s->Px = s->has_pml ? s->pml.npml[0] : 0;
    s->P = s->has_pml ? s->pml.npml[1] : 0;
    s->j0 = 1;
    s->j1 = ny;
    s->ja = 0;
    s->jb = ny + 1;
    const double ku_pml = s->dt * K.c0 / s->ds;
    std::vector<double> prof;
    const int n[2] = {nx, ny}, P[2] = {s->Px, s->P};
    for (int a = 0; a < 2; ++a) {
      std::vector<double> f1(n[a] + 2, 1.0), f2(n[a] + 2, ku_pml), g1(n[a] + 2, 1.0), g2(n[a] + 2, ku_pml);
      if (P[a] > 0) {
        const double sigmam = s->pml.sigmam > 0 ? s->pml.sigmam
                                                : (s->pml.order + 1.0) * K.c0 * std::log(1.0 / s->pml.R0) /
                                                      (2.0 * P[a] * s->ds);  // sim3/program.f90:102
        pml_grade(s->dt, s->ds, s->pml.order, sigmam, P[a], K.ep0, n[a], f1.data(), f2.data(), g1.data(),
                  g2.data());
      }
      for (auto* v : {&f1, &f2, &g1, &g2}) prof.insert(prof.end(), v->begin(), v->end());
    }
    s->d_tprof = s->prec == FDTD_FP32 ? upload_cast<float>(s, prof) : upload_cast<double>(s, prof);
    if (!s->d_tprof) return set_error(FDTD_ERR_NOMEM, "PML profiles");
    if (s->has_pml) {
      char* a = (char*)alloc_in(s->allocs, &s->dev_bytes, fbytes);
      char* b = (char*)alloc_in(s->allocs, &s->dev_bytes, fbytes);
      if (!a || !b) return set_error(FDTD_ERR_NOMEM, "PML allocation");
      s->d_ezx = a + shift;
      s->d_ezy = b + shift;
    }
    // rasterised below with key[0] = last writer of Ez, key[1] = last
    // dielectric (the PML medium)
    for (auto& m : s->mats) {
      m.writes[0] = 1;
      m.writes[1] = m.type == 0 ? 1 : 0;
    }
  } else if (s->has_pml) {
    const int P = s->pml.npml[1];
    s->P = P;
    s->Px = s->pml.npml[0];
    s->ja = P;
    s->jb = ny - P + 1;
    s->j0 = P;
    s->j1 = ny - P;
    const Consts& K = consts();
    s->sigmam = s->pml.sigmam > 0 ? s->pml.sigmam
                                  : (s->pml.order + 1.0) * K.c0 * std::log(1.0 / s->pml.R0) / (2.0 * P * s->ds);
    for (int sl = 0; sl < 2; ++sl) {
      for (int q = 0; q < 4; ++q) {
        s->d_pml[sl][q] = alloc_in(s->allocs, &s->dev_bytes, (size_t)s->px * (P + 2) * s->wsize);
        if (!s->d_pml[sl][q]) return set_error(FDTD_ERR_NOMEM, "PML allocation");
      }
      std::vector<double> py, pxv;
      pml_profiles(s, sl, py, pxv);
      py.insert(py.end(), pxv.begin(), pxv.end());
      s->d_prof[sl] = s->prec == FDTD_FP32 ? upload_cast<float>(s, py) : upload_cast<double>(s, py);
      if (!s->d_prof[sl]) return set_error(FDTD_ERR_NOMEM, "PML profiles");
    }
  } else {
    s->j0 = 1;
    s->j1 = ny;
    s->ja = 0;
    s->jb = ny + 1;
  }
  // dispersive groups (identical coefficients share one)
  s->groups.clear();
  s->ljps.clear();
  for (auto& m : s->mats) {
    if (m.type != 3) continue;
    LjpCoef<double> c;
    lorentz_jp_coef(s->dt, s->ds, m.lj, c);
    if (s->tmz) {  // local component 0 is z; Ez's JpT_x (ez_uses_jx) stays 0 here
      c.jmask = (m.lj.jp_mask >> 2) & 1;
    } else {
      c.jmask = m.lj.jp_mask & 3;
      c.ezjx = 0;
    }
    int gi = -1;
    for (size_t q = 0; q < s->ljps.size(); ++q)
      if (std::memcmp(&s->ljps[q].co, &c, sizeof c) == 0) gi = (int)q;
    if (gi < 0) {
      if ((int)s->ljps.size() >= kMaxGroup) return set_error(FDTD_ERR_UNSUPPORTED, "too many Lorentz-Jp materials");
      HostLjp g;
      g.co = c;
      g.nstate = 2 * c.npols + (c.nl ? 2 : 0);
      s->ljps.push_back(g);
      gi = (int)s->ljps.size() - 1;
    }
    m.group = gi;
  }
  for (auto& m : s->mats) {
    if (m.type == 0 || m.type == 3) continue;
    AdeCoef<double> c;
    if (m.type == 1) {
      drude2cp_coef(s, m.d2, c);
    } else if (s->tmz) {  // the 3D module's graphene coefficients
      double einf = 0;
      int rc = graphene3d_coef(s->dt, s->ds, m.gr, c, &einf);
      if (rc) return rc;
      c.ku = s->ku;
    } else {
      int rc = graphene_coef(s, m.gr, c);
      if (rc) return rc;
    }
    int gi = -1;
    for (size_t q = 0; q < s->groups.size(); ++q)
      if (std::memcmp(&s->groups[q].co, &c, sizeof c) == 0) gi = (int)q;
    if (gi < 0) {
      if ((int)s->groups.size() >= kMaxGroup) return set_error(FDTD_ERR_UNSUPPORTED, "too many dispersive materials");
      HostGroup g;
      g.co = c;
      g.nstate = 2 + 2 * c.ncp + (c.raman ? 2 : 0);
      s->groups.push_back(g);
      gi = (int)s->groups.size() - 1;
    }
    m.group = gi;
  }
  // rasterise: per cell, the last writer of Ex and of Ey (call order,
  // sim2/program.f90:345-348)
  using Key = std::array<int16_t, 2>;
  std::vector<Key> combos{{-1, -1}};
  std::map<Key, int> index{{combos[0], 0}};
  std::vector<uint8_t> cid((size_t)s->ncell, 0);
  for (size_t mi = 0; mi < s->mats.size(); ++mi) {
    const HostMat& m = s->mats[mi];
    int lo[2], hi[2];
    if (!region_bounds(s, m.reg, lo, hi)) return set_error(FDTD_ERR_ARG, "bad region");
    std::vector<int> trans(256, -1);
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        if (!region_inside(m.reg, i, j)) continue;
        const size_t c = (size_t)(i + s->px * j);
        const int old = cid[c];
        if (trans[old] < 0) {
          Key key = combos[old];
          for (int q = 0; q < 2; ++q)
            if (m.writes[q]) key[q] = (int16_t)mi;
          auto it = index.find(key);
          if (it == index.end()) {
            if (combos.size() >= 256) return set_error(FDTD_ERR_UNSUPPORTED, "more than 256 material combinations");
            combos.push_back(key);
            it = index.emplace(key, (int)combos.size() - 1).first;
          }
          trans[old] = it->second;
        }
        cid[c] = (uint8_t)trans[old];
      }
  }
  std::vector<Mat<double>> tab(combos.size());
  if (s->tmz) {
    for (size_t q = 0; q < combos.size(); ++q) {
      tab[q].act[0] = tab[q].act[1] = ACT_KEEP;
      tab[q].gax[0] = tab[q].gax[1] = 1.0;
      const int w = combos[q][0], d = combos[q][1];
      if (w >= 0) {
        const HostMat& m = s->mats[w];
        tab[q].act[0] = m.type == 0 ? ACT_DIEL : ACT_ADE;
        if (m.type == 0) tab[q].gax[0] = m.eps_r == 0.0 ? 0.0 : 1.0 / m.eps_r;
        if (m.type == 3 && !ljp_comp(m, true, 0)) tab[q].act[0] = ACT_DIEL;  // E = D
      }
      if (d >= 0) tab[q].gax[1] = s->mats[d].eps_r == 0.0 ? 0.0 : 1.0 / s->mats[d].eps_r;
    }
  }
  for (size_t q = 0; q < combos.size() && !s->tmz; ++q)
    for (int c = 0; c < 2; ++c) {
      tab[q].act[c] = ACT_KEEP;
      tab[q].gax[c] = 1.0;
      const int w = combos[q][c];
      if (w < 0) continue;
      const HostMat& m = s->mats[w];
      if (m.type == 0) {
        tab[q].act[c] = ACT_DIEL;
        tab[q].gax[c] = m.eps_r == 0.0 ? 0.0 : 1.0 / m.eps_r;  // simpledielectric_mod.f90:66-75
      } else if (m.type == 3 && !ljp_comp(m, false, c)) {
        tab[q].act[c] = ACT_DIEL;  // a component outside jp_mask: E = D
        tab[q].gax[c] = 1.0;
      } else if (m.type == 2 && c == 1) {
        tab[q].act[c] = ACT_DIEL;  // graphene: Ey = Dy (graphene_mod.f90:68)
        tab[q].gax[c] = 1.0;
      } else {
        tab[q].act[c] = ACT_ADE;
      }
    }
  if (s->prec == FDTD_FP32) {
    std::vector<Mat<float>> t(tab.size());
    for (size_t q = 0; q < tab.size(); ++q)
      for (int c = 0; c < 2; ++c) { t[q].gax[c] = (float)tab[q].gax[c]; t[q].act[c] = tab[q].act[c]; }
    s->d_tab = upload(s->allocs, &s->dev_bytes, t.data(), t.size() * sizeof(Mat<float>));
  } else {
    s->d_tab = upload(s->allocs, &s->dev_bytes, tab.data(), tab.size() * sizeof(Mat<double>));
  }
  s->d_matid = (uint8_t*)upload(s->allocs, &s->dev_bytes, cid.data(), cid.size());
  if (!s->d_tab || !s->d_matid) return set_error(FDTD_ERR_NOMEM, "material table");
  // dispersive cell lists: cells must lie strictly between the slabs (the
  // slabs overwrite Ex there after the materials, which a separate launch
  // for dispersive cells would undo)
  std::vector<std::vector<int32_t>> gcells(s->groups.size()), lcells(s->ljps.size());
  std::vector<std::vector<uint8_t>> gmask(s->groups.size()), lmask(s->ljps.size());
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const int64_t c = i + s->px * j;
      const Key& key = combos[cid[(size_t)c]];
      int gm[2] = {-1, -1}, lm[2] = {-1, -1};
      if (s->tmz) {
        if (key[0] >= 0 && (s->mats[key[0]].type == 1 || s->mats[key[0]].type == 2)) gm[0] = s->mats[key[0]].group;
        if (key[0] >= 0 && ljp_comp(s->mats[key[0]], true, 0)) lm[0] = s->mats[key[0]].group;
      } else {
        for (int q = 0; q < 2; ++q) {
          if (key[q] < 0) continue;
          const HostMat& m = s->mats[key[q]];
          if ((m.type == 1 || m.type == 2) && !(m.type == 2 && q == 1)) gm[q] = m.group;
          if (ljp_comp(m, false, q)) lm[q] = m.group;
        }
      }
      if (lm[0] >= 0 || lm[1] >= 0) {
        if (s->tmz ? ((s->Px > 0 && (i <= s->Px || i > nx - s->Px)) || (s->P > 0 && (j <= s->P || j > ny - s->P)))
                   : (s->has_pml && (j <= s->ja || j >= s->jb)))
          return set_error(FDTD_ERR_UNSUPPORTED, "Lorentz-Jp cells inside the PML");
        if (s->has_tf && (j == s->tfp.lo || (s->tfp.hi > 0 && j == s->tfp.hi + 1)))
          return set_error(FDTD_ERR_UNSUPPORTED, "a TFSF row through Lorentz-Jp cells");
        for (size_t g = 0; g < s->ljps.size(); ++g) {
          uint8_t mk = 0;
          for (int q = 0; q < 2; ++q)
            if (lm[q] == (int)g) mk |= (uint8_t)(1 << q);
          if (mk) {
            lcells[g].push_back((int32_t)c);
            lmask[g].push_back(mk);
          }
        }
      }
      if (gm[0] < 0 && gm[1] < 0) continue;
      if (s->tmz && ((s->Px > 0 && (i <= s->Px || i > nx - s->Px)) || (s->P > 0 && (j <= s->P || j > ny - s->P))))
        return set_error(FDTD_ERR_UNSUPPORTED, "dispersive cells inside the PML");
      if (!s->tmz && s->has_pml && (j <= s->ja || j >= s->jb))
        return set_error(FDTD_ERR_UNSUPPORTED, "dispersive cells inside a PML slab row");
      for (size_t g = 0; g < s->groups.size(); ++g) {
        uint8_t mk = 0;
        for (int q = 0; q < 2; ++q)
          if (gm[q] == (int)g) mk |= (uint8_t)(1 << q);
        if (mk) {
          gcells[g].push_back((int32_t)c);
          gmask[g].push_back(mk);
        }
      }
    }
  for (size_t g = 0; g < s->ljps.size(); ++g) {
    HostLjp& G = s->ljps[g];
    G.ncells = (int64_t)lcells[g].size();
    if (!G.ncells) continue;
    G.d_cells = (int32_t*)upload(s->allocs, &s->dev_bytes, lcells[g].data(), lcells[g].size() * sizeof(int32_t));
    G.d_mask = (uint8_t*)upload(s->allocs, &s->dev_bytes, lmask[g].data(), lmask[g].size());
    G.d_state = (double*)alloc_in(s->allocs, &s->dev_bytes, (size_t)G.nstate * 2 * (size_t)G.ncells * sizeof(double));
    if (!G.d_cells || !G.d_mask || !G.d_state) return set_error(FDTD_ERR_NOMEM, "Lorentz-Jp state");
  }
  for (size_t g = 0; g < s->groups.size(); ++g) {
    HostGroup& G = s->groups[g];
    G.ncells = (int64_t)gcells[g].size();
    if (!G.ncells) continue;
    G.d_cells = (int32_t*)upload(s->allocs, &s->dev_bytes, gcells[g].data(), gcells[g].size() * sizeof(int32_t));
    G.d_mask = (uint8_t*)upload(s->allocs, &s->dev_bytes, gmask[g].data(), gmask[g].size());
    G.d_state = alloc_in(s->allocs, &s->dev_bytes, (size_t)G.nstate * 2 * (size_t)G.ncells * sizeof(double));
    if (!G.d_cells || !G.d_mask || !G.d_state) return set_error(FDTD_ERR_NOMEM, "dispersive state");
  }
  // TFSF incident line
  if (s->has_tf) {
    const int top = s->tfp.hi > 0 ? s->tfp.hi + 1 : s->tfp.lo;
    const int P = s->has_pml ? s->pml.npml[1] : 0;
    if (s->has_pml && (s->tfp.lo - 1 <= P || top >= ny - P + (s->tmz ? 1 : 0)))
      return set_error(FDTD_ERR_ARG, "TFSF rows inside the PML");
    const size_t m = (size_t)ny + 2;
    std::vector<double> line(8 * m, 0.0);
    tfsf_line_profiles(s->dt, s->ds, ny, s->tfp.npml, s->tfp.eps, s->tfp.order, s->tfp.R0, line.data() + 4 * m);
    s->d_tf = (double*)upload(s->allocs, &s->dev_bytes, line.data(), line.size() * sizeof(double));
    if (!s->d_tf) return set_error(FDTD_ERR_NOMEM, "TFSF line");
  }
  // monitors
  for (auto& d : s->dfts) {
    d.nslots = s->tmz ? ((d.mask & FDTD2D_DFT_EZ) ? 1 : 0) + ((d.mask & FDTD2D_DFT_HX) ? 1 : 0)
                      : ((d.mask & FDTD2D_DFT_EX) ? 1 : 0) + ((d.mask & FDTD2D_DFT_HZ) ? 1 : 0);
    d.d_omega = (double*)upload(s->allocs, &s->dev_bytes, d.omega.data(), d.omega.size() * sizeof(double));
    d.d_acc = (double*)alloc_in(s->allocs, &s->dev_bytes, sizeof(double) * 2 * (size_t)d.nslots * d.nfreq * nx);
    d.d_samp = alloc_in(s->allocs, &s->dev_bytes, s->wsize * (size_t)kBlockSteps * d.nslots * nx);
    d.d_sstep = (int32_t*)alloc_in(s->allocs, &s->dev_bytes, sizeof(int32_t) * kBlockSteps);
    if (!d.d_omega || !d.d_acc || !d.d_samp || !d.d_sstep) return set_error(FDTD_ERR_NOMEM, "DFT accumulators");
  }
  s->d_srcval = alloc_in(s->allocs, &s->dev_bytes, sizeof(double) * kMaxLine);
  if (!s->d_srcval) return set_error(FDTD_ERR_NOMEM, "source values");
  FDTD_HIP(hipDeviceSynchronize());
  s->committed = true;
  return 0;
}

int ensure_committed(Sim* s) { return s->committed ? 0 : commit_impl(s); }

// probe buffer holding steps 1..steps
int ensure_probe_cap(Sim* s, int steps) {
  if (s->probes.empty() || steps <= s->probe_cap) return 0;
  int cap = std::max(steps, 2 * s->probe_cap);
  cap = std::max(cap, 1024);
  const size_t rowb = sizeof(double) * 3 * s->probes.size();
  double* p = (double*)alloc_in(s->allocs, &s->dev_bytes, (size_t)cap * rowb);
  if (!p) return set_error(FDTD_ERR_NOMEM, "probe buffer");
  if (s->d_probe && s->n_done > 0)
    FDTD_HIP(hipMemcpy(p, s->d_probe, (size_t)s->n_done * rowb, hipMemcpyDeviceToDevice));
  s->d_probe = p;
  s->probe_cap = cap;
  return 0;
}

template <typename real>
Inst<real> make_inst(Sim* s, int32_t* ctl, int tile0) {
  Inst<real> d{};
  if (s->tmz) {
    d.tmz = 1;
    d.ez = (real*)s->F[0]; d.hx = (real*)s->F[1]; d.hy = (real*)s->F[2]; d.dz = (real*)s->F[3];
    d.ezx = (real*)s->d_ezx; d.ezy = (real*)s->d_ezy;
    d.tprof = (const real*)s->d_tprof;
    d.Px = s->Px;
  } else {
    d.ex = (real*)s->F[0]; d.ey = (real*)s->F[1]; d.hz = (real*)s->F[2]; d.dx = (real*)s->F[3]; d.dy = (real*)s->F[4];
  }
  d.matid = s->d_matid;
  d.tab = (const Mat<real>*)s->d_tab;
  d.px = (int32_t)s->px;
  d.nx = s->nx; d.ny = s->ny;
  d.j0 = s->j0; d.j1 = s->j1; d.ja = s->ja; d.jb = s->jb;
  d.has_pml = s->has_pml ? 1 : 0;
  d.per = s->g.periodic_x ? 1 : 0;
  d.ntx = (s->nx + kTX - 1) / kTX;
  d.nty = (s->ny + kTY - 1) / kTY;
  d.tile0 = tile0;
  d.ku = (real)s->ku;
  d.dt = s->dt;
  d.P = s->P;
  for (int sl = 0; sl < 2; ++sl) {
    for (int q = 0; q < 4; ++q) d.pml[sl][q] = (real*)s->d_pml[sl][q];
    d.prof[sl] = (const real*)s->d_prof[sl];
  }
  d.prow0[0] = 0;
  d.prow0[1] = s->jb - 1;
  d.nsrc = (int32_t)s->srcs.size();
  for (size_t q = 0; q < s->srcs.size(); ++q) {
    d.src_j[q] = s->srcs[q].first;
    d.src_t0[q] = s->srcs[q].second.t0;
    d.src_st[q] = s->srcs[q].second.st;
    d.src_w0[q] = s->srcs[q].second.w0;
    d.src_amp[q] = s->srcs[q].second.amplitude;
  }
  d.src_val = (real*)s->d_srcval;
  d.ngroup = (int32_t)s->groups.size();
  for (size_t g = 0; g < s->groups.size(); ++g) {
    const HostGroup& G = s->groups[g];
    AdeDev<real>& A = d.grp[g];
    A.cells = G.d_cells;
    A.mask = G.d_mask;
    A.state = (double*)G.d_state;
    A.ncells = (int32_t)G.ncells;
    A.nstate = G.nstate;
    A.co = G.co;
  }
  d.nljp = (int32_t)s->ljps.size();
  for (size_t g = 0; g < s->ljps.size(); ++g) {
    const HostLjp& G = s->ljps[g];
    d.ljp[g].cells = G.d_cells;
    d.ljp[g].mask = G.d_mask;
    d.ljp[g].state = G.d_state;
    d.ljp[g].ncells = (int32_t)G.ncells;
    d.ljp[g].co = G.co;
  }
  d.ndft = (int32_t)s->dfts.size();
  for (size_t l = 0; l < s->dfts.size(); ++l) {
    const HostDft& h = s->dfts[l];
    DftDev& D = d.dft[l];
    D.j = h.j; D.nfreq = h.nfreq; D.mask = h.mask; D.n_begin = h.n_begin; D.n_end = h.n_end; D.nslots = h.nslots;
    D.omega = h.d_omega; D.acc = h.d_acc; D.samp = h.d_samp; D.sstep = h.d_sstep;
  }
  d.nprobe = (int32_t)s->probes.size();
  for (size_t p = 0; p < s->probes.size(); ++p)
    d.probe_c[p] = (int32_t)(s->probes[p].first + s->px * s->probes[p].second);
  d.probe_rec = s->d_probe;
  d.probe_cap = s->probe_cap;
  d.ctl = ctl;
  if (s->has_tf) {
    d.tf_on = 1;
    d.tf_lo = s->tfp.lo;
    d.tf_hi = s->tfp.hi;
    d.tf_src = s->tfp.src_index;
    d.tf_npml = s->tfp.npml;
    d.tf_line = s->d_tf;
    d.tf_ck = s->ku * (1.0 / s->tfp.eps);
    d.tf_ku = s->ku;
    d.tf_sg = s->tmz ? -1.0 : 1.0;  // (Ez, Hx) along y: -1; (Ex, Hz): +1
    d.tf_t0 = s->tfp.pulse.t0;
    d.tf_st = s->tfp.pulse.st;
    d.tf_w0 = s->tfp.pulse.w0;
    d.tf_amp = s->tfp.pulse.amplitude;
  }
  return d;
}

template <typename real>
int build_inst(Batch* b) {
  std::vector<Inst<real>> v(b->sims.size());
  int tile0 = 0;
  for (int q : b->order) {
    Sim* s = b->sims[(size_t)q];
    v[(size_t)q] = make_inst<real>(s, b->d_ctl + 2 * q, tile0);
    tile0 += v[(size_t)q].ntx * v[(size_t)q].nty;
  }
  const size_t bytes = v.size() * sizeof(Inst<real>);
  b->inst.assign((const char*)v.data(), (const char*)v.data() + bytes);
  FDTD_HIP(hipMemcpy(b->d_inst, b->inst.data(), bytes, hipMemcpyHostToDevice));
  return 0;
}

int batch_build(Batch* b) {
  const int count = (int)b->sims.size();
  b->prec = b->sims[0]->prec;
  b->device = b->sims[0]->device;
  FDTD_HIP(hipSetDevice(b->device));
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_check(e, "hipStreamCreate");
  const size_t isz = b->prec == FDTD_FP32 ? sizeof(Inst<float>) : sizeof(Inst<double>);
  b->d_inst = alloc_in(b->allocs, nullptr, isz * count);
  b->d_ctl = (int32_t*)alloc_in(b->allocs, nullptr, sizeof(int32_t) * 2 * count);
  if (!b->d_inst || !b->d_ctl) return set_error(FDTD_ERR_NOMEM, "batch descriptors");
  std::vector<int32_t> tile_inst, tfl;
  std::vector<int4> ade, mon, ljp;
  std::vector<int2> lines;
  b->order.resize((size_t)count);
  for (int q = 0; q < count; ++q) b->order[(size_t)q] = q;
  std::stable_sort(b->order.begin(), b->order.end(), [&](int u, int v) {
    return (int64_t)b->sims[(size_t)u]->nx * b->sims[(size_t)u]->ny > (int64_t)b->sims[(size_t)v]->nx * b->sims[(size_t)v]->ny;
  });
  for (int q : b->order) {
    Sim* s = b->sims[q];
    const int nt = ((s->nx + kTX - 1) / kTX) * ((s->ny + kTY - 1) / kTY);
    tile_inst.insert(tile_inst.end(), (size_t)nt, q);
  }
  for (int q = 0; q < count; ++q) {
    Sim* s = b->sims[q];
    for (size_t g = 0; g < s->groups.size(); ++g)
      for (int64_t c0 = 0; c0 < s->groups[g].ncells; c0 += kChunk) ade.push_back(make_int4(q, (int)g, (int)c0, 0));
    for (size_t g = 0; g < s->ljps.size(); ++g)
      for (int64_t c0 = 0; c0 < s->ljps[g].ncells; c0 += kChunk) ljp.push_back(make_int4(q, (int)g, (int)c0, 0));
    for (size_t l = 0; l < s->dfts.size(); ++l) {
      for (int i0 = 0; i0 < s->nx; i0 += kChunk) mon.push_back(make_int4(q, (int)l, i0, 0));
      lines.push_back(make_int2(q, (int)l));
    }
    if (!s->probes.empty()) mon.push_back(make_int4(q, -1, 0, 0));
    if (s->has_tf) tfl.push_back(q);
  }
  b->ntiles = (int)tile_inst.size();
  b->nade = (int)ade.size();
  b->nmon = (int)mon.size();
  b->nlines = (int)lines.size();
  b->d_tile_inst = (int32_t*)upload(b->allocs, nullptr, tile_inst.data(), tile_inst.size() * sizeof(int32_t));
  b->d_ade = (int4*)upload(b->allocs, nullptr, ade.data(), ade.size() * sizeof(int4));
  b->d_mon = (int4*)upload(b->allocs, nullptr, mon.data(), mon.size() * sizeof(int4));
  b->d_lines = (int2*)upload(b->allocs, nullptr, lines.data(), lines.size() * sizeof(int2));
  b->ntf = (int)tfl.size();
  b->nljp = (int)ljp.size();
  b->d_ljp = (int4*)upload(b->allocs, nullptr, ljp.data(), ljp.size() * sizeof(int4));
  b->d_tfl = (int32_t*)upload(b->allocs, nullptr, tfl.data(), tfl.size() * sizeof(int32_t));
  if (!b->d_tile_inst || !b->d_ade || !b->d_mon || !b->d_lines || !b->d_tfl || !b->d_ljp)
    return set_error(FDTD_ERR_NOMEM, "batch work lists");
  b->ctl.assign(2 * (size_t)count, 0);
  for (int q = 0; q < count; ++q) b->ctl[2 * q] = b->ctl[2 * q + 1] = b->sims[q]->n_done;
  FDTD_HIP(hipMemcpy(b->d_ctl, b->ctl.data(), sizeof(int32_t) * 2 * count, hipMemcpyHostToDevice));
  return b->prec == FDTD_FP32 ? build_inst<float>(b) : build_inst<double>(b);
}

template <typename real>
void enqueue_block(Batch* b) {
  hipStream_t st = b->stream;
  const Inst<real>* I = (const Inst<real>*)b->d_inst;
  // blocks of the E and H launches; each strides over the packed tile list
  int tpb = 1;
  if (const char* e = std::getenv("FDTD2D_TILES_PER_BLOCK")) {
    const int v = std::atoi(e);
    if (v >= 1 && v <= 4096) tpb = v;
  }
  unsigned grid = (unsigned)std::min(b->ntiles, std::max(8192, (b->ntiles + tpb - 1) / tpb));
  grid = (grid + 8 * kXcdRun - 1) / (8 * kXcdRun) * (8 * kXcdRun);  // whole XCD runs (xcd_block)
  const int count = (int)b->sims.size();
  for (int s = 0; s < kBlockSteps; ++s) {
    // the line's E half step may precede k2_e, which reads only the line's h
    if (b->ntf) hipLaunchKernelGGL(k2_tfsf_line<real>, dim3((unsigned)b->ntf), dim3(256), 0, st, I, b->d_tfl, b->ntf, 0);
    hipLaunchKernelGGL(k2_e<real>, dim3(grid), dim3(kTX * kWaves), 0, st, I, b->d_tile_inst, b->ntiles);
    if (b->nade) hipLaunchKernelGGL(k2_ade<real>, dim3((unsigned)b->nade), dim3(kChunk), 0, st, I, b->d_ade, b->nade);
    if (b->nljp) hipLaunchKernelGGL(k2_ljp<real>, dim3((unsigned)b->nljp), dim3(kChunk), 0, st, I, b->d_ljp, b->nljp);
    hipLaunchKernelGGL(k2_h<real>, dim3(grid), dim3(kTX * kWaves), 0, st, I, b->d_tile_inst, b->ntiles);
    // k2_h read the line's new e; its H half step follows
    if (b->ntf) hipLaunchKernelGGL(k2_tfsf_line<real>, dim3((unsigned)b->ntf), dim3(256), 0, st, I, b->d_tfl, b->ntf, 1);
    if (b->nmon) hipLaunchKernelGGL(k2_mon<real>, dim3((unsigned)b->nmon), dim3(kChunk), 0, st, I, b->d_mon, b->nmon, s);
    hipLaunchKernelGGL(k2_adv<real>, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, st, I, count, 1);
  }
  if (b->nlines) {
    hipLaunchKernelGGL(k2_flush<real>, dim3((unsigned)b->nmon), dim3(kChunk), 0, st, I, b->d_mon, b->nmon);
    hipLaunchKernelGGL(k2_clear<real>, dim3((unsigned)b->nlines), dim3(64), 0, st, I, b->d_lines, b->nlines);
  }
}

// Captures one 32-step block and instantiates it.  The exec holds its own
// copy of the nodes, so the captured graph is destroyed at once; on any
// failure nothing is kept and the next call captures afresh.
int capture(Batch* b) {
  if (b->exec) return 0;
  FDTD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  if (b->prec == FDTD_FP32)
    enqueue_block<float>(b);
  else
    enqueue_block<double>(b);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_check(e, "hipStreamEndCapture");
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hip_check(e, "hipGraphInstantiate");
  b->exec = x;
  return 0;
}

int batch_run(Batch* b, const int32_t* nsteps) {
  FDTD_HIP(hipSetDevice(b->device));
  const int count = (int)b->sims.size();
  int maxn = 0;
  for (int q = 0; q < count; ++q) {
    if (nsteps[q] < 0) return set_error(FDTD_ERR_ARG, "negative step count");
    maxn = std::max(maxn, nsteps[q]);
  }
  if (maxn == 0) return 0;
  // the previous run (of this or another batch of the same handles) must be
  // done before a probe buffer is copied into a larger one
  for (Sim* s : b->sims) {
    int rc = sync_sim(s);
    if (rc) return rc;
  }
  for (int q = 0; q < count; ++q) {
    Sim* s = b->sims[q];
    int rc = ensure_probe_cap(s, s->n_done + nsteps[q]);
    if (rc) return rc;
    b->ctl[2 * q] = s->n_done;
    b->ctl[2 * q + 1] = s->n_done + nsteps[q];
  }
  // descriptors again: a probe buffer may have moved (the graph reads them
  // from device memory, so it stays valid)
  {
    int rc = b->prec == FDTD_FP32 ? build_inst<float>(b) : build_inst<double>(b);
    if (rc) return rc;
  }
  FDTD_HIP(hipMemcpy(b->d_ctl, b->ctl.data(), sizeof(int32_t) * 2 * count, hipMemcpyHostToDevice));
  if (b->prec == FDTD_FP32)
    hipLaunchKernelGGL(k2_adv<float>, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, b->stream,
                       (const Inst<float>*)b->d_inst, count, 0);
  else
    hipLaunchKernelGGL(k2_adv<double>, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, b->stream,
                       (const Inst<double>*)b->d_inst, count, 0);
  const int nblocks = (maxn + kBlockSteps - 1) / kBlockSteps;
  // FDTD2D_GRAPH=0: the same launches of each 32-step block are enqueued
  // directly instead of replaying the captured graph (for tools that cannot
  // follow graph launches); results are the same
  const char* ge = std::getenv("FDTD2D_GRAPH");
  if (ge && std::atoi(ge) == 0) {
    for (int r = 0; r < nblocks; ++r) {
      if (b->prec == FDTD_FP32)
        enqueue_block<float>(b);
      else
        enqueue_block<double>(b);
      FDTD_HIP(hipGetLastError());
    }
  } else {
    int rc = capture(b);
    if (rc) return rc;
    for (int r = 0; r < nblocks; ++r) FDTD_HIP(hipGraphLaunch(b->exec, b->stream));
  }
  for (int q = 0; q < count; ++q) b->sims[q]->n_done += nsteps[q];
  return 0;
}

int batch_free(Batch* b) {
  if (!b) return 0;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  for (auto& a : b->allocs) (void)hipFree(a.p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  for (Sim* s : b->sims)
    if (s->owner == b) s->owner = nullptr;
  delete b;
  return 0;
}

int sync_sim(Sim* s) {
  FDTD_HIP(hipSetDevice(s->device));
  if (s->solo) FDTD_HIP(hipStreamSynchronize(s->solo->stream));
  if (s->owner) FDTD_HIP(hipStreamSynchronize(s->owner->stream));
  FDTD_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

void* field_ptr(Sim* s, int comp) {
  if (s->tmz) {
    switch (comp) {
      case FDTD2D_EZ: return s->F[0];
      case FDTD2D_HX: return s->F[1];
      case FDTD2D_HY: return s->F[2];
      case FDTD2D_DZ: return s->F[3];
      default: return nullptr;
    }
  }
  switch (comp) {
    case FDTD2D_EX: return s->F[0];
    case FDTD2D_EY: return s->F[1];
    case FDTD2D_HZ: return s->F[2];
    case FDTD2D_DX: return s->F[3];
    case FDTD2D_DY: return s->F[4];
    default: return nullptr;
  }
}

}  // namespace d2
}  // namespace fdtd

using namespace fdtd;
using namespace fdtd::d2;

struct fdtd2d_sim : public fdtd::d2::Sim {};
struct fdtd2d_batch : public fdtd::d2::Batch {};

#define GUARD2(s)                                                \
  do {                                                           \
    if (!(s)) return set_error(FDTD_ERR_ARG, "null handle");     \
  } while (0)
#define SETUP2(s)                                                                  \
  do {                                                                             \
    if ((s)->committed) return set_error(FDTD_ERR_STATE, "set-up after commit");  \
  } while (0)

extern "C" {

int fdtd2d_create(fdtd2d_sim** out, const fdtd2d_grid* g) {
  if (!out || !g) return set_error(FDTD_ERR_ARG, "null argument");
  if (g->nx < 2 || g->ny < 2 || !(g->ds > 0) || !(g->courant > 0)) return set_error(FDTD_ERR_ARG, "bad grid");
  if (g->precision != FDTD_FP64 && g->precision != FDTD_FP32) return set_error(FDTD_ERR_ARG, "bad precision");
  if (g->polarization != FDTD2D_TEZ && g->polarization != FDTD2D_TMZ) return set_error(FDTD_ERR_ARG, "bad polarization");
  int dev = 0;
  FDTD_HIP(hipGetDevice(&dev));
  fdtd2d_sim* s = new fdtd2d_sim();
  s->g = *g;
  s->nx = g->nx;
  s->ny = g->ny;
  s->ds = g->ds;
  s->ku = g->courant;
  s->dt = s->ku * s->ds / consts().c0;  // sim2/program.f90:45
  s->prec = g->precision;
  s->wsize = g->precision == FDTD_FP32 ? 4 : 8;
  s->tmz = g->polarization == FDTD2D_TMZ;
  s->device = dev;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return hip_check(e, "hipStreamCreate");
  }
  *out = s;
  return 0;
}

int fdtd2d_destroy(fdtd2d_sim* s) {
  if (!s) return 0;
  (void)hipSetDevice(s->device);
  if (s->owner) return set_error(FDTD_ERR_STATE, "simulation belongs to a live batch");
  batch_free(s->solo);
  (void)hipStreamSynchronize(s->stream);
  for (auto& a : s->allocs) (void)hipFree(a.p);
  (void)hipStreamDestroy(s->stream);
  delete s;
  return 0;
}

int fdtd2d_set_pml(fdtd2d_sim* s, const fdtd2d_pml* p) {
  GUARD2(s);
  SETUP2(s);
  if (!p) return set_error(FDTD_ERR_ARG, "null argument");
  if (s->tmz) {
    if (p->npml[0] < 0 || p->npml[1] < 0 || 2 * p->npml[0] > s->nx || 2 * p->npml[1] > s->ny)
      return set_error(FDTD_ERR_ARG, "bad npml");
    if (p->npml[0] > 0 && s->g.periodic_x) return set_error(FDTD_ERR_ARG, "PML on a periodic axis");
    if ((p->npml[0] > 0 || p->npml[1] > 0) && (!(p->order > 0) || (!(p->sigmam > 0) && !(p->R0 > 0 && p->R0 < 1))))
      return set_error(FDTD_ERR_ARG, "bad grading");
    s->pml = *p;
    s->has_pml = p->npml[0] > 0 || p->npml[1] > 0;
    return 0;
  }
  if (p->npml[1] < 1 || p->npml[0] < 0 || 2 * p->npml[1] + 2 > s->ny || 2 * p->npml[0] > s->nx)
    return set_error(FDTD_ERR_ARG, "bad npml");
  if (p->npml[1] >= s->ny / 2 - 1) return set_error(FDTD_ERR_ARG, "y PML reaches the middle of the domain");
  if (!(p->eps_low > 0) || !(p->eps_high > 0)) return set_error(FDTD_ERR_ARG, "eps_low/eps_high must be > 0");
  if (!(p->order > 0) || (!(p->sigmam > 0) && !(p->R0 > 0 && p->R0 < 1))) return set_error(FDTD_ERR_ARG, "bad grading");
  s->pml = *p;
  s->has_pml = true;
  return 0;
}

static int add_mat(fdtd2d_sim* s, const fdtd2d_region* r, HostMat& m) {
  int lo[2], hi[2];
  if (!r || !region_bounds(s, *r, lo, hi)) return set_error(FDTD_ERR_ARG, "bad region");
  m.reg = *r;
  s->mats.push_back(m);
  return 0;
}

int fdtd2d_add_dielectric(fdtd2d_sim* s, const fdtd2d_region* r, double eps_r) {
  GUARD2(s);
  SETUP2(s);
  HostMat m{};
  m.type = 0;
  m.eps_r = eps_r;
  m.writes[0] = m.writes[1] = 1;
  return add_mat(s, r, m);
}

int fdtd2d_add_drude2cp(fdtd2d_sim* s, const fdtd2d_region* r, const fdtd_drude2cp* p) {
  GUARD2(s);
  SETUP2(s);
  if (!p || !(p->gamma > 0)) return set_error(FDTD_ERR_ARG, "bad Drude+2CP parameters");
  HostMat m{};
  m.type = 1;
  m.d2 = *p;
  m.writes[0] = m.writes[1] = 1;
  return add_mat(s, r, m);
}

int fdtd2d_add_graphene(fdtd2d_sim* s, const fdtd2d_region* r, const fdtd_graphene* p) {
  GUARD2(s);
  SETUP2(s);
  if (!p) return set_error(FDTD_ERR_ARG, "null argument");
  if (p->normal_axis != 1) return set_error(FDTD_ERR_ARG, "a 2D graphene line lies along x (normal_axis 1)");
  if (p->model != FDTD_GRAPHENE_CP4 && p->model != FDTD_GRAPHENE_KUBO) return set_error(FDTD_ERR_ARG, "bad graphene model");
  HostMat m{};
  m.type = 2;
  m.gr = *p;
  m.writes[0] = m.writes[1] = 1;  // Ex through the ADE, Ey = Dy
  return add_mat(s, r, m);
}

int fdtd2d_add_lorentz_jp(fdtd2d_sim* s, const fdtd2d_region* r, const fdtd_lorentz_jp* p) {
  GUARD2(s);
  SETUP2(s);
  if (!p || p->npols < 1 || p->npols > 4 || p->jp_mask < 1 || p->jp_mask > 7 || !(p->eps_inf > 0))
    return set_error(FDTD_ERR_ARG, "bad Lorentz-Jp parameters");
  HostMat m{};
  m.type = 3;
  m.lj = *p;
  m.writes[0] = m.writes[1] = 1;
  return add_mat(s, r, m);
}

int fdtd2d_add_line_source(fdtd2d_sim* s, int j, const fdtd_pulse* p) {
  GUARD2(s);
  SETUP2(s);
  if (!p) return set_error(FDTD_ERR_ARG, "null argument");
  if (j < 1 || j > s->ny) return set_error(FDTD_ERR_ARG, "source row out of range");
  if ((int)s->srcs.size() >= kMaxLine) return set_error(FDTD_ERR_UNSUPPORTED, "too many line sources");
  for (auto& q : s->srcs)
    if (q.first == j) return set_error(FDTD_ERR_UNSUPPORTED, "two line sources on one row");
  if (!(p->st > 0)) return set_error(FDTD_ERR_ARG, "pulse width must be > 0");
  s->srcs.push_back({j, *p});
  return 0;
}

int fdtd2d_add_dft_line(fdtd2d_sim* s, int j, int nfreq, const double* omega, int mask, int n_begin, int n_end,
                        int* id) {
  GUARD2(s);
  SETUP2(s);
  if (!omega || !id) return set_error(FDTD_ERR_ARG, "null argument");
  const int all = s->tmz ? (FDTD2D_DFT_EZ | FDTD2D_DFT_HX) : (FDTD2D_DFT_EX | FDTD2D_DFT_HZ);
  if (j < 1 || j > s->ny || nfreq < 1 || !(mask & all) || (mask & ~all)) return set_error(FDTD_ERR_ARG, "bad DFT line");
  if ((int)s->dfts.size() >= kMaxDft) return set_error(FDTD_ERR_UNSUPPORTED, "too many DFT lines");
  HostDft d{};
  d.j = j;
  d.nfreq = nfreq;
  d.mask = mask;
  d.n_begin = n_begin;
  d.n_end = n_end;
  d.omega.assign(omega, omega + nfreq);
  s->dfts.push_back(d);
  *id = (int)s->dfts.size() - 1;
  return 0;
}

int fdtd2d_add_probe(fdtd2d_sim* s, int i, int j, int* id) {
  GUARD2(s);
  SETUP2(s);
  if (!id) return set_error(FDTD_ERR_ARG, "null argument");
  if (i < 1 || i > s->nx || j < 1 || j > s->ny) return set_error(FDTD_ERR_ARG, "probe out of range");
  if ((int)s->probes.size() >= kMaxProbe) return set_error(FDTD_ERR_UNSUPPORTED, "too many probes");
  s->probes.push_back({i, j});
  *id = (int)s->probes.size() - 1;
  return 0;
}

int fdtd2d_add_tfsf(fdtd2d_sim* s, const fdtd2d_tfsf* p, int* id) {
  GUARD2(s);
  SETUP2(s);
  if (!p || !id) return set_error(FDTD_ERR_ARG, "null argument");
  if (s->has_tf) return set_error(FDTD_ERR_UNSUPPORTED, "one TFSF per simulation");
  const int N = s->ny;
  if (p->lo < 2 || p->lo > N || (p->hi != 0 && (p->hi < p->lo || p->hi >= N)) || p->src_index < 1 ||
      p->src_index > N || p->npml < 0 || 2 * p->npml + 2 >= N || !(p->eps > 0) || !(p->pulse.st > 0) ||
      p->src_index <= p->npml || p->src_index > N - p->npml ||
      (p->npml > 0 && (!(p->order > 0) || !(p->R0 > 0 && p->R0 < 1))))
    return set_error(FDTD_ERR_ARG, "bad TFSF rows");
  s->has_tf = true;
  s->tfp = *p;
  *id = 0;
  return 0;
}

int fdtd2d_get_tfsf_line(fdtd2d_sim* s, int id, double* e, double* h) {
  GUARD2(s);
  if (!e || !h || id != 0 || !s->has_tf) return set_error(FDTD_ERR_ARG, "bad argument");
  int rc = ensure_committed(s);
  if (rc) return rc;
  rc = sync_sim(s);
  if (rc) return rc;
  const size_t m = (size_t)s->ny + 2;
  FDTD_HIP(hipMemcpy(e, s->d_tf, m * sizeof(double), hipMemcpyDeviceToHost));
  FDTD_HIP(hipMemcpy(h, s->d_tf + m, m * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int fdtd2d_commit(fdtd2d_sim* s) {
  GUARD2(s);
  return ensure_committed(s);
}

int fdtd2d_step(fdtd2d_sim* s, int nsteps) {
  GUARD2(s);
  if (nsteps < 0) return set_error(FDTD_ERR_ARG, "negative step count");
  if (s->owner) return set_error(FDTD_ERR_STATE, "simulation belongs to a live batch");
  int rc = ensure_committed(s);
  if (rc) return rc;
  if (!s->solo) {
    Batch* b = new Batch();
    b->sims.push_back(s);
    rc = batch_build(b);
    if (rc) {
      batch_free(b);
      return rc;
    }
    s->solo = b;
  }
  const int32_t n = nsteps;
  return batch_run(s->solo, &n);
}

int fdtd2d_sync(fdtd2d_sim* s) {
  GUARD2(s);
  return sync_sim(s);
}

int fdtd2d_get_step(fdtd2d_sim* s, int* n) {
  GUARD2(s);
  if (!n) return set_error(FDTD_ERR_ARG, "null argument");
  *n = s->n_done;
  return 0;
}

int fdtd2d_get_dt(fdtd2d_sim* s, double* dt) {
  GUARD2(s);
  if (!dt) return set_error(FDTD_ERR_ARG, "null argument");
  *dt = s->dt;
  return 0;
}

int fdtd2d_get_field(fdtd2d_sim* s, int comp, double* dst) {
  GUARD2(s);
  if (!dst) return set_error(FDTD_ERR_ARG, "null argument");
  int rc = ensure_committed(s);
  if (rc) return rc;
  void* f = field_ptr(s, comp);
  if (!f) return set_error(FDTD_ERR_ARG, "bad component");
  rc = sync_sim(s);
  if (rc) return rc;
  std::vector<char> h((size_t)s->ncell * s->wsize);
  FDTD_HIP(hipMemcpy(h.data(), f, h.size(), hipMemcpyDeviceToHost));
  for (int j = 1; j <= s->ny; ++j)
    for (int i = 1; i <= s->nx; ++i) {
      const size_t c = (size_t)(i + s->px * j);
      dst[(size_t)(i - 1) + (size_t)s->nx * (j - 1)] =
          s->prec == FDTD_FP32 ? (double)((const float*)h.data())[c] : ((const double*)h.data())[c];
    }
  return 0;
}

int fdtd2d_set_field(fdtd2d_sim* s, int comp, const double* src) {
  GUARD2(s);
  if (!src) return set_error(FDTD_ERR_ARG, "null argument");
  int rc = ensure_committed(s);
  if (rc) return rc;
  void* f = field_ptr(s, comp);
  if (!f) return set_error(FDTD_ERR_ARG, "bad component");
  rc = sync_sim(s);
  if (rc) return rc;
  std::vector<char> h((size_t)s->ncell * s->wsize);
  FDTD_HIP(hipMemcpy(h.data(), f, h.size(), hipMemcpyDeviceToHost));
  const bool wrap = s->g.periodic_x && (comp == FDTD2D_EY || comp == FDTD2D_HZ || comp == FDTD2D_EZ || comp == FDTD2D_HY);
  for (int j = 1; j <= s->ny; ++j)
    for (int i = 1; i <= s->nx; ++i) {
      const double v = src[(size_t)(i - 1) + (size_t)s->nx * (j - 1)];
      std::vector<size_t> cs{(size_t)(i + s->px * j)};
      if (wrap && i == 1) cs.push_back((size_t)(s->nx + 1 + s->px * j));
      if (wrap && i == s->nx) cs.push_back((size_t)(s->px * j));
      for (size_t c : cs) {
        if (s->prec == FDTD_FP32)
          ((float*)h.data())[c] = (float)v;
        else
          ((double*)h.data())[c] = v;
      }
    }
  FDTD_HIP(hipMemcpy(f, h.data(), h.size(), hipMemcpyHostToDevice));
  return 0;
}

int fdtd2d_get_dft(fdtd2d_sim* s, int id, int comp, double* re, double* im) {
  GUARD2(s);
  if (!re || !im || id < 0 || id >= (int)s->dfts.size()) return set_error(FDTD_ERR_ARG, "bad argument");
  int rc = ensure_committed(s);
  if (rc) return rc;
  const HostDft& d = s->dfts[id];
  int slot = -1;
  const int e_bit = s->tmz ? FDTD2D_DFT_EZ : FDTD2D_DFT_EX, h_bit = s->tmz ? FDTD2D_DFT_HX : FDTD2D_DFT_HZ;
  const int e_comp = s->tmz ? FDTD2D_EZ : FDTD2D_EX, h_comp = s->tmz ? FDTD2D_HX : FDTD2D_HZ;
  if (comp == e_comp && (d.mask & e_bit)) slot = 0;
  if (comp == h_comp && (d.mask & h_bit)) slot = (d.mask & e_bit) ? 1 : 0;
  if (slot < 0) return set_error(FDTD_ERR_ARG, "component not recorded");
  rc = sync_sim(s);
  if (rc) return rc;
  const size_t n = (size_t)d.nfreq * s->nx;
  FDTD_HIP(hipMemcpy(re, d.d_acc + (size_t)slot * 2 * n, n * sizeof(double), hipMemcpyDeviceToHost));
  FDTD_HIP(hipMemcpy(im, d.d_acc + (size_t)slot * 2 * n + n, n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int fdtd2d_get_flux(fdtd2d_sim* s, int id, double* S) {
  GUARD2(s);
  if (!S || id < 0 || id >= (int)s->dfts.size()) return set_error(FDTD_ERR_ARG, "bad argument");
  const HostDft& d = s->dfts[id];
  const int both = s->tmz ? (FDTD2D_DFT_EZ | FDTD2D_DFT_HX) : (FDTD2D_DFT_EX | FDTD2D_DFT_HZ);
  if ((d.mask & both) != both) return set_error(FDTD_ERR_ARG, "flux needs both recorded components");
  const size_t n = (size_t)d.nfreq * s->nx;
  std::vector<double> er(n), ei(n), hr(n), hi(n);
  int rc = fdtd2d_get_dft(s, id, s->tmz ? FDTD2D_EZ : FDTD2D_EX, er.data(), ei.data());
  if (!rc) rc = fdtd2d_get_dft(s, id, s->tmz ? FDTD2D_HX : FDTD2D_HZ, hr.data(), hi.data());
  if (rc) return rc;
  if (s->tmz) {  // sim3/program.f90:613-685, the Hx*Ez terms
    for (int m = 0; m < d.nfreq; ++m) {
      double sy = 0.0;
      const size_t o = (size_t)m * s->nx;
      for (int i = 0; i < s->nx; ++i) sy = sy + ((hr[o + i] * er[o + i]) + (hi[o + i] * ei[o + i])) * s->ds;
      S[m] = sy;
    }
    return 0;
  }
  // sim2/program.f90:479-529, the reference's summation order
  for (int m = 0; m < d.nfreq; ++m) {
    double sy = 0.0;
    const size_t o = (size_t)m * s->nx;
    for (int i = 0; i < s->nx; ++i) sy = sy - ((hr[o + i] * er[o + i]) + (hi[o + i] * ei[o + i])) * s->ds;
    S[m] = sy;
  }
  return 0;
}

int fdtd2d_get_probe_count(fdtd2d_sim* s, int id, int* n) {
  GUARD2(s);
  if (!n || id < 0 || id >= (int)s->probes.size()) return set_error(FDTD_ERR_ARG, "bad argument");
  *n = s->n_done;
  return 0;
}

int fdtd2d_get_probe(fdtd2d_sim* s, int id, double* out) {
  GUARD2(s);
  if (!out || id < 0 || id >= (int)s->probes.size()) return set_error(FDTD_ERR_ARG, "bad argument");
  if (s->n_done == 0) return 0;
  int rc = sync_sim(s);
  if (rc) return rc;
  const size_t np = s->probes.size();
  std::vector<double> h((size_t)s->n_done * np * 3);
  FDTD_HIP(hipMemcpy(h.data(), s->d_probe, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int r = 0; r < s->n_done; ++r)
    for (int c = 0; c < 3; ++c) out[3 * (size_t)r + c] = h[((size_t)r * np + id) * 3 + c];
  return 0;
}

int fdtd2d_batch_create(fdtd2d_batch** out, fdtd2d_sim* const* sims, int count) {
  if (!out || !sims || count < 1) return set_error(FDTD_ERR_ARG, "bad argument");
  for (int q = 0; q < count; ++q) {
    if (!sims[q]) return set_error(FDTD_ERR_ARG, "null simulation");
    if (sims[q]->prec != sims[0]->prec) return set_error(FDTD_ERR_ARG, "one precision per batch");
    if (sims[q]->owner) return set_error(FDTD_ERR_STATE, "simulation already in a batch");
    for (int r = 0; r < q; ++r)
      if (sims[r] == sims[q]) return set_error(FDTD_ERR_ARG, "simulation listed twice");
    int rc = ensure_committed(sims[q]);
    if (rc) return rc;
    rc = sync_sim(sims[q]);
    if (rc) return rc;
  }
  fdtd2d_batch* b = new fdtd2d_batch();
  for (int q = 0; q < count; ++q) b->sims.push_back(sims[q]);
  int rc = batch_build(b);
  if (rc) {
    batch_free(b);
    return rc;
  }
  for (int q = 0; q < count; ++q) sims[q]->owner = b;
  *out = b;
  return 0;
}

int fdtd2d_batch_run(fdtd2d_batch* b, const int32_t* nsteps) {
  if (!b || !nsteps) return set_error(FDTD_ERR_ARG, "null argument");
  return batch_run(b, nsteps);
}

int fdtd2d_batch_sync(fdtd2d_batch* b) {
  if (!b) return set_error(FDTD_ERR_ARG, "null argument");
  FDTD_HIP(hipSetDevice(b->device));
  FDTD_HIP(hipStreamSynchronize(b->stream));
  return 0;
}

int fdtd2d_batch_destroy(fdtd2d_batch* b) {
  if (!b) return 0;
  return batch_free(b);
}

}  // extern "C"
