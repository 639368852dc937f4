// fdtd_internal.h -- shared state of the gfx950 FDTD library (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fdtd.h"

namespace fdtd {

// misc/constants_mod.f90:3-24, evaluated the way the reference evaluates them
struct Consts {
  double pi, ep0, mu0, c0, z0, hbar, kb, echarge;
};
const Consts& consts();

// per-component action of a material-table entry
enum : int32_t { ACT_KEEP = 0, ACT_DIEL = 1, ACT_ADE = 2 };

// Material-table entry, one per distinct per-cell combination of writers.
// coef = Ku*gax (E-form update), gax = 1/eps (source scaling and D-form),
// pgax = medium of an ALL_FACES PML cell, eps = energy weight.
template <typename real>
struct alignas(16) MatEntry {
  real coef[3];
  real pgax;
  real gax[3];
  real eps;
  int32_t act[3];
  int32_t pad;
};

// Device form of the material table: coef = (Ku*gax x,y,z, PML medium gax),
// gaxe = (gax x,y,z, energy eps), flags bit q = a dielectric writes E_q.
template <typename real>
struct alignas(4 * sizeof(real)) Vec4 {
  real v[4];
};

// Coefficients of one group of dispersive cells (Drude + up to 4 CP + Raman).
template <typename real>
struct AdeCoef {
  int32_t ncp, raman;
  real ku, ha, hb, gar, gai;
  real gbr[4], gbi[4], gcr[4], gci[4];
  real r0, r1, r2;
};

// Coefficients of one group of Lorentz-Jp cells (E form, see
// fdtd_lorentz_jp).  kuga = Ku'*ga with Ku' = dt*c0/ds; xi and xi3 are
// divided by c0 in the kernel, in the reference's operation order.
template <typename real>
struct LjpCoef {
  int32_t npols, jmask, nl, ezjx;
  real kuga, ga, ds, dt, c0, z0;
  real xi[4], g1[4], g2[4];
  real xi3, ng1, ng2;
};

// Lorentz poles in D form (mod3/si_mod.f90:234-259): per pole l,
// S_l <- ga_l*S_l + gb_l*S_l(one step older) + gc_l*E.
struct LordCoef {
  int32_t npols, pad;
  double ku;
  double ga[4], gb[4], gc[4];
};

// Up to this many plane sources per handle.
constexpr int kMaxSrc = 8;

// A plane source adds val, or, when dval is set, dscale * *dval read on the
// device (the TFSF consistency terms read the incident line that way).
template <typename real>
struct SrcArgs {
  int32_t n;
  int32_t axis[kMaxSrc], index[kMaxSrc], comp[kMaxSrc];
  real val[kMaxSrc];
  const double* dval[kMaxSrc];
  double dscale[kMaxSrc];
};

template <typename real>
__device__ __forceinline__ real src_value(const SrcArgs<real>& s, int q) {
  return s.dval[q] ? (real)(s.dscale[q] * *s.dval[q]) : s.val[q];
}

// split-field components
enum { S_EXY = 0, S_EXZ, S_EYZ, S_EYX, S_EZX, S_EZY, S_HXY, S_HXZ, S_HYZ, S_HYX, S_HZX, S_HZY, S_N };

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct HostMaterial {
  int type;  // 0 dielectric, 1 drude2cp, 2 graphene, 3 Lorentz-Jp, 4 Lorentz D form (lj's table)
  fdtd_region reg;
  double eps_r;
  fdtd_drude2cp d2;
  fdtd_graphene gr;
  fdtd_lorentz_jp lj;
  int writes[3];
  int group;  // ADE group index (dispersive only)
};

struct HostSource {
  int axis, index, comp;
  fdtd_pulse p;
};

// Samples of a DFT plane are buffered for up to kDftBlock gated steps and
// folded into the FP64 accumulators in one pass (time-blocked running DFT).
// Every fdtd_step call also folds what it buffered, so a job stepped in
// chunks folds at the same steps as one stepped at once (the same bits)
// when the chunks are multiples of kDftBlock (the pool's 64, 256).
constexpr int kDftBlock = 64;

// Planes a general workgroup of the one-pass step marches through (at most).
constexpr int kGenKc = 4;

struct DftPlane {
  int axis, index, nfreq, mask, n_begin, n_end, n1, n2;
  std::vector<double> omega;
  double* d_omega = nullptr;
  double* d_acc = nullptr;   // [comp slot][re/im][m][v][u]
  void* d_samp = nullptr;    // [b][slot][v][u], field precision
  double* d_tab = nullptr;   // [b][m][cos, sin]
  int nbuf = 0;              // buffered samples
  int n_first = 0;           // step of the first buffered sample
  int slot[6];
  int nslots = 0;
  bool fused_now = false;  // this step's sample is written by the H kernel
};

struct Probe {
  int i, j, k;
  int line = -1;  // TFSF id of an incident-line probe (j = line cell)
  std::vector<double> rec;
};

// TFSF plane wave: the incident 1D line and its PML (device, FP64):
// e, h, pe, ph, f1, f2, g1, g2, each n+2 doubles (cells 0..n+1).
struct Tfsf {
  fdtd_tfsf p{};
  int n = 0, hc = 0, sg = 1;
  double* d_line = nullptr;
  int32_t* d_lp = nullptr;  // (probe index, line cell) pairs of its probes
  int nlp = 0;
};

struct AdeGroup {
  int kind = 0;          // 0 Drude/CP/Raman (coef), 1 Lorentz-Jp (lj), 2 Lorentz D form (ld)
  AdeCoef<double> coef;  // FP64 master copy
  LjpCoef<double> lj{};
  LordCoef ld{};
  double eps_inf;
  int nstate;                 // state words per component
  int64_t ncells = 0;
  int32_t* d_cells = nullptr;  // linear cell offsets
  uint8_t* d_mask = nullptr;   // component mask per cell
  void* d_state = nullptr;     // SoA [word][cell]
};

struct Sim {
  fdtd_grid g{};
  int nx = 0, ny = 0, nz = 0;
  int64_t px = 0, pxy = 0, ncell_alloc = 0;  // ghosted layout strides
  double ds = 0, ku = 0, dt = 0, ku_pml = 0;
  int prec = FDTD_FP64;
  int engine = FDTD_ENGINE_AUTO;
  size_t wsize = 8;
  fdtd_pml pml{};
  int pml_mode = FDTD_PML_NONE;
  std::vector<HostMaterial> mats;
  std::vector<HostSource> srcs;
  std::vector<DftPlane> dfts;
  std::vector<Probe> probes;
  std::vector<Tfsf> tfsfs;
  bool committed = false;
  int n_done = 0;
  int device = 0;
  hipStream_t stream = nullptr;

  // device state
  void* E[3] = {nullptr, nullptr, nullptr};
  void* H[3] = {nullptr, nullptr, nullptr};
  void* D[3] = {nullptr, nullptr, nullptr};
  uint8_t* d_matid = nullptr;  // per ghosted cell
  void* d_mattab = nullptr;    // Vec4 coef[ntab], Vec4 gaxe[ntab], uint32 flags[ntab]
  int ntab = 0;
  std::vector<MatEntry<double>> tab;  // FP64 master copy
  std::vector<AdeGroup> groups;
  // ALL_FACES PML
  int P[3] = {0, 0, 0};
  int64_t npml_cells = 0;
  void* d_split = nullptr;  // [S_N][npml_cells]
  void* d_prof = nullptr;   // per axis: f1E, f2E, g1H, g2H each n_a+2
  // REF_Y PML
  int64_t slab_cells = 0;   // (nx+2)*(P+2)*(nz+2)
  void* d_slab = nullptr;   // [2][S_N][slab_cells]
  void* d_ref_prof = nullptr;  // [2 slabs][12 profiles][max_n+2]
  int ref_prof_stride = 0;
  // probes
  double* d_probe = nullptr;     // [row][probe][7] records not yet harvested
  int32_t* d_probe_pos = nullptr;
  int probe_cap = 0;             // rows
  int probe_pending = 0;         // rows written since the last harvest
  int probe_row0 = 0;            // first row of the current fdtd_step call
  int call_base = 0;  // steps done before the current fdtd_step call
  // reductions
  double* d_red = nullptr;
  int red_cap = 0;

  std::vector<DevBuf> allocs;
  size_t dev_bytes = 0;

  // launch shape of the cell-parallel kernels (x threads, y threads) and
  // whether fused kernels deal tiles to XCDs in contiguous runs
  int bx = 64, by = 4;
  int xcdmap = 1;
  // one-pass step (FUSED engine): E and H of a step read from one copy of
  // the fields and written to the other (E2/H2/d_split2 hold the second
  // copy).  The grid is cut into tiles of 64 x op_by cells (op_by in
  // {2, 4, 8}) and z ranges of at most op_kc planes (broken at the z PML
  // slabs).  Plain tiles (one dielectric material, no PML, no x- or
  // y-normal source plane) are marched through their planes by one
  // wavefront each, E and H in one pass; the other tiles run the fused
  // cell updates, E then H, over the same two copies.
  int onepass = 0;
  int op_by = 4, op_kc = 64;
  int tl_by = -1, tl_kc = -1;        // tile lists below built for these
  int tl_gkc = 1;                    // planes per general workgroup in them
  int32_t* d_tl_plain = nullptr;     // i0, j0, k0, k1, xlo, xhi, material id, x face (0, 1)
  int32_t* d_tl_gen = nullptr;       // workgroups: x0, x1, j0, j1, k0, width (64, 32 or 16), k1, face (0, 1: y, 2: z)
  int n_tl_plain = 0, n_tl_gen = 0;
  int n_tl_xf = 0;                   // plain tiles that reach into an x slab (x-face tiles), after the others
  int tl_bx = 2;                     // rows of an x-face tile
  int64_t n_plain_cells = 0;         // cells owned by plain tiles (x-face tiles not counted)
  int64_t n_xfi_cells = 0;           // non-PML cells owned by x-face tiles
  int64_t n_xf_cells = 0;            // x-slab face cells owned by x-face tiles
  // material runs of every (j,k) row r = j + (ny+2)*k: runs h_runoff[r] ..
  // h_runoff[r+1]-1, run q starting at cell h_runx[q] with id h_runid[q]
  std::vector<int32_t> h_runoff, h_runx;
  std::vector<uint8_t> h_runid;
  // two-kernel step: E and H tiles span zm planes (1, 2, 4 or 8), each
  // thread marching its column and carrying the neighbour plane in registers
  int zm = 1;
  // FP32 two-kernel step: the H kernel writes the monitor-plane samples
  // (FDTD_DFT_FUSE=0: the separate sample kernel does, bitwise the same)
  int dft_fuse = 1;
  // dry steps (autotuner): no sources, TFSF or monitors, no step count; run
  // only while every field is still zero, so they leave the state as it was
  bool dry = false;
  bool fields_set = false;  // set_field was called
  void* E2[3] = {nullptr, nullptr, nullptr};
  void* H2[3] = {nullptr, nullptr, nullptr};
  void* d_split2 = nullptr;
  int32_t* d_rowmat = nullptr;  // uniform material id of a (j,k) row, or -1
  int32_t* d_mixrows = nullptr;  // rows j + (ny+2)*k whose material ids differ
  int nmix = 0;

  // profiling
  bool profiling = false;
  double k_ms[FDTD_K_COUNT] = {0};
  int64_t k_launches[FDTD_K_COUNT] = {0};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> event_pool;

  // cell counts
  int64_t n_pml = 0, n_ade = 0, n_main = 0;
};

int set_error(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);
#define FDTD_HIP(call)                                 \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_check(e_, #call); \
  } while (0)

void* dev_alloc(Sim* s, size_t bytes);  // zero-filled; nullptr on failure
double* reduction_buffer(Sim* s, size_t bytes);  // reusable, grown with dev_alloc

// kernels (fdtd_kernels.hip)
int launch_step(Sim* s);           // one full time step at s->n_done + 1
int launch_dft_flush(Sim* s);      // fold every buffered DFT sample
int launch_energy(Sim* s, double* we, double* wh);
int launch_flux(Sim* s, int id, double* S);
int flux_check(Sim* s, int id);                            // the plane carries the four components
int launch_flux_into(Sim* s, int id, double* dS, double* S);  // dS: nfreq doubles of device memory
int launch_set_ghosts(Sim* s, int comp);
int launch_nonfinite(Sim* s, int* d_flag);  // d_flag = 1 if E or H holds NaN/Inf (sticky)
int build_march_tiles(Sim* s, int by, int kc);  // one-pass tile lists (fdtd_api.hip)

size_t estimate_bytes(const Sim* s);  // device bytes commit would allocate

// graphene coefficients of mod3/graphene_mod.f90 (CP4 or Kubo, exact-
// exponential Raman) for time step dt and cell size ds (fdtd_api.hip)
int graphene3d_coef(double dt, double ds, const fdtd_graphene& p, AdeCoef<double>& c, double* eps_inf);

// split-field PML grading of one axis, mod3/BERpmly_mod.f90:612-874: cells
// 1..P graded low (E at P-x+1/2, H at P-x), n-P+1..n high; others untouched
void pml_grade(double dt, double ds, double order, double sigmam, int P, double ep, int n, double* f1, double* f2,
               double* g1, double* g2);

// Lorentz-Jp coefficients (mod3/goldLorentzJp_mod.f90:37-91,141-163 and the
// family) for time step dt and cell size ds
void lorentz_jp_coef(double dt, double ds, const fdtd_lorentz_jp& p, LjpCoef<double>& c);

// end-slab profiles of a TFSF incident line of n cells (mod1/
// BERpmly_ExHz_mod.f90:156-255): f1, f2, g1, g2, n+2 doubles each, into out
void tfsf_line_profiles(double dt, double ds, int n, int npml, double eps, double order, double R0, double* out);

}  // namespace fdtd

// The opaque handle of the C ABI is the simulation state itself.
struct fdtd_sim : public fdtd::Sim {};
