// fdtd_api.hip -- C ABI of the gfx950 FDTD library (include/fdtd.h): set-up,
// geometry rasterisation, coefficient tables, device state, results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>

#include "fdtd_internal.h"

namespace fdtd {

const Consts& consts() {
  static const Consts c = [] {
    Consts k{};
    k.pi = 4.0 * std::atan(1.0);      // misc/constants_mod.f90:4
    k.ep0 = 8.85418781762039e-12;     // :6
    k.mu0 = k.pi * 4e-7;              // :7
    k.c0 = 1.0 / std::sqrt(k.ep0 * k.mu0);
    k.z0 = std::sqrt(k.mu0 / k.ep0);  // mod3/BERpmly_mod.f90:625
    const double h = 6.62607015e-34;
    k.hbar = h / (2.0 * k.pi);
    k.kb = 1.380649e-23;
    k.echarge = 1.602176634e-19;
    return k;
  }();
  return c;
}

static thread_local std::string g_err;

int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_check(hipError_t e, const char* what) {
  return set_error(FDTD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Stream-ordered on the handle's stream: allocating, clearing and (at
// destroy) freeing never wait for other handles' work, so jobs that set up
// or retire while others run do not drain the device.
void* dev_alloc(Sim* s, size_t bytes) {
  if (bytes == 0) bytes = 16;
  void* p = nullptr;
  if (hipMallocAsync(&p, bytes, s->stream) != hipSuccess) return nullptr;
  if (hipMemsetAsync(p, 0, bytes, s->stream) != hipSuccess) {
    (void)hipFreeAsync(p, s->stream);
    return nullptr;
  }
  s->allocs.push_back({p, bytes});
  s->dev_bytes += bytes;
  return p;
}

// Stream-ordered release of one dev_alloc block (base pointer).
void dev_free(Sim* s, void* p) {
  for (size_t q = 0; q < s->allocs.size(); ++q)
    if (s->allocs[q].p == p) {
      (void)hipFreeAsync(p, s->stream);
      s->dev_bytes -= s->allocs[q].bytes;
      s->allocs.erase(s->allocs.begin() + (long)q);
      return;
    }
}

double* reduction_buffer(Sim* s, size_t bytes) {
  if ((size_t)s->red_cap < bytes) {
    s->d_red = (double*)dev_alloc(s, bytes);
    s->red_cap = s->d_red ? (int)bytes : 0;
  }
  return s->d_red;
}

namespace {

using cplx = std::complex<double>;

// ---- regions: box clipping of mod3/material_mod.f90:60-119 + shapes -------
bool region_bounds(const Sim* s, const fdtd_region& r, int lo[3], int hi[3]) {
  const int n[3] = {s->nx, s->ny, s->nz};
  if (r.kind == FDTD_REGION_BOX) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = r.lo[a];
      hi[a] = r.hi[a];
    }
  } else if (r.kind == FDTD_REGION_SPHERE || r.kind == FDTD_REGION_CYLINDER) {
    if (r.kind == FDTD_REGION_CYLINDER && (r.axis < 0 || r.axis > 2)) return false;
    for (int a = 0; a < 3; ++a) {
      double ext = r.radius;
      if (r.kind == FDTD_REGION_CYLINDER && a == r.axis) ext = r.half_length;
      lo[a] = (int)std::ceil(r.center[a] - ext);
      hi[a] = (int)std::floor(r.center[a] + ext);
      if (r.lo[a] <= r.hi[a] && (r.lo[a] != 0 || r.hi[a] != 0)) {
        lo[a] = std::max(lo[a], r.lo[a]);
        hi[a] = std::min(hi[a], r.hi[a]);
      }
    }
  } else {
    return false;
  }
  for (int a = 0; a < 3; ++a) {
    lo[a] = std::max(lo[a], 1);
    hi[a] = std::min(hi[a], n[a]);
  }
  return true;
}

bool region_inside(const fdtd_region& r, int i, int j, int k) {
  const double p[3] = {(double)i, (double)j, (double)k};
  if (r.kind == FDTD_REGION_BOX) return true;
  if (r.kind == FDTD_REGION_SPHERE) {
    double d2 = 0;
    for (int a = 0; a < 3; ++a) d2 += (p[a] - r.center[a]) * (p[a] - r.center[a]);
    return d2 <= r.radius * r.radius;
  }
  double d2 = 0;
  for (int a = 0; a < 3; ++a)
    if (a != r.axis) d2 += (p[a] - r.center[a]) * (p[a] - r.center[a]);
  const double dz = std::fabs(p[r.axis] - r.center[r.axis]);
  return d2 <= r.radius * r.radius && dz <= r.half_length;
}

// ---- dispersive coefficients ---------------------------------------------
void drude2cp_coef(const Sim* s, const fdtd_drude2cp& p, AdeCoef<double>& c, double* eps_inf) {
  const double dt = s->dt;
  const cplx I(0.0, 1.0);
  std::memset(&c, 0, sizeof c);
  // setexp, mod3/gold_mod.f90:187-195
  const cplx expi1 = std::exp(-I * p.phi[0]);
  const cplx expi2 = std::exp(-I * p.phi[1]);
  const cplx exp1 = std::exp((-p.Gamma[0] + I * p.Omega[0]) * dt);
  const cplx exp2 = std::exp((-p.Gamma[1] + I * p.Omega[1]) * dt);
  // initialize_GoldVars, :175-181
  c.ncp = 2;
  c.raman = 0;
  c.ha = std::exp(-p.gamma * dt);
  c.hb = (p.omega_d * p.omega_d) * dt / p.gamma;
  const cplx gc1 = dt * 2.0 * I * p.A[0] * p.Omega[0] * expi1;
  const cplx gc2 = dt * 2.0 * I * p.A[1] * p.Omega[1] * expi2;
  const cplx ga = p.eps_inf - gc1 - gc2;
  c.gbr[0] = exp1.real(); c.gbi[0] = exp1.imag();
  c.gbr[1] = exp2.real(); c.gbi[1] = exp2.imag();
  c.gcr[0] = gc1.real(); c.gci[0] = gc1.imag();
  c.gcr[1] = gc2.real(); c.gci[1] = gc2.imag();
  c.gar = ga.real();
  c.gai = ga.imag();
  *eps_inf = p.eps_inf;
}

int graphene_coef_dt(double dt, double ds, const fdtd_graphene& p, AdeCoef<double>& c, double* eps_inf) {
  const Consts& K = consts();
  const cplx I(0.0, 1.0);
  std::memset(&c, 0, sizeof c);
  c.ncp = 4;
  if (p.model == FDTD_GRAPHENE_CP4) {
    // initialize_GrapheneVars + scaleParams, mod3/graphene_mod.f90:202-269
    const double scale = p.thickness / ds;
    const double wD = std::sqrt(scale) * p.omega_d;
    c.ha = std::exp(-p.gamma * dt);
    c.hb = (wD * wD) * dt / p.gamma;
    cplx ga = p.eps_inf;
    for (int q = 0; q < 4; ++q) {
      cplx expi = std::exp(-I * p.phi[q]);
      cplx ex = std::exp((-p.Gamma[q] + I * p.Omega[q]) * dt);
      if (p.ref_init_order) {  // locals read before setexp (:222-231)
        expi = 0.0;
        ex = 0.0;
      }
      const cplx gc = dt * 2.0 * I * (scale * p.A[q]) * p.Omega[q] * expi;
      c.gbr[q] = ex.real(); c.gbi[q] = ex.imag();
      c.gcr[q] = gc.real(); c.gci[q] = gc.imag();
      ga = ga - gc;
    }
    c.gar = ga.real();
    c.gai = ga.imag();
  } else if (p.model == FDTD_GRAPHENE_KUBO) {
    // Kubo intraband Drude weight spread over one cell (see oracle)
    if (!(p.tau > 0) || !(p.temperature > 0)) return set_error(FDTD_ERR_ARG, "tau and T must be > 0");
    const double kT = K.kb * p.temperature;
    const double ef = p.fermi_ev * K.echarge;
    const double Dw = 2.0 * K.echarge * K.echarge * kT / (K.hbar * K.hbar) * std::log(2.0 * std::cosh(ef / (2.0 * kT)));
    const double wp2 = Dw / (K.pi * K.ep0 * ds);
    const double gam = 1.0 / p.tau;
    c.ha = std::exp(-gam * dt);
    c.hb = wp2 * dt / gam;
    c.gar = p.eps_inf;
    c.gai = 0.0;
  } else {
    return set_error(FDTD_ERR_ARG, "bad graphene model");
  }
  if (p.wNL > 0) {
    // exact-exponential Raman recursion, mod3/graphene_mod.f90:216-218,256-258
    const double alfaR = 0.5 * p.dnuNL;
    const double betaR = std::sqrt((p.wNL * p.wNL) - (alfaR * alfaR));
    const double omegaR = (p.wNL * p.wNL) / betaR;
    c.raman = 1;
    c.r0 = p.xi3 * omegaR * dt * std::exp(-alfaR * dt) * std::sin(betaR * dt);
    c.r1 = 2.0 * std::exp(-alfaR * dt) * std::cos(betaR * dt);
    c.r2 = -std::exp(-2.0 * alfaR * dt);
  }
  *eps_inf = p.eps_inf;
  return 0;
}

// Lorentz-Jp poles, mod3/goldLorentzJp_mod.f90:37-91,141-163 (g1, g2 of
// initialize_GoldVars, evaluated inline by the film-gold and graphene
// modules with the same expressions); Ku' = dt*c0/ds as the modules compute.
// whether material m updates component q through a dispersive group
bool is_ade_comp(const HostMaterial& m, int q) {
  if (m.type == 0) return false;
  if (m.type == 3) return (m.lj.jp_mask >> q) & 1;
  return true;
}

// si_mod's recursion coefficients, mod3/si_mod.f90:249-257, in its
// operation order
void lord_coef_dt(double dt, const fdtd_lorentz_jp& p, LordCoef& c) {
  std::memset(&c, 0, sizeof c);
  c.npols = p.npols;
  for (int l = 0; l < p.npols; ++l) {
    const double den = 1.0 + dt * p.dnu[l];
    const double wd = dt * p.w0[l], dw = p.w0[l] * dt;
    c.ga[l] = (2.0 - wd * wd) / den;
    c.gb[l] = (dt * p.dnu[l] - 1.0) / den;
    c.gc[l] = (p.xi[l] * (dw * dw)) / den;
  }
}

void ljp_coef_dt(double dt, double ds, const fdtd_lorentz_jp& p, LjpCoef<double>& c) {
  const Consts& K = consts();
  std::memset(&c, 0, sizeof c);
  c.npols = p.npols;
  c.jmask = p.jp_mask;
  c.nl = p.nonlinear ? 1 : 0;
  c.ezjx = p.ez_uses_jx ? 1 : 0;
  c.ga = 1.0 / p.eps_inf;
  c.kuga = (dt * K.c0 / ds) * c.ga;
  c.ds = ds;
  c.dt = dt;
  c.c0 = K.c0;
  c.z0 = std::sqrt(K.mu0 / K.ep0);
  for (int l = 0; l < p.npols; ++l) {
    c.xi[l] = p.xi[l];
    c.g1[l] = (1.0 - p.dnu[l] * dt) / (1.0 + p.dnu[l] * dt);
    c.g2[l] = (p.w0[l] * p.w0[l]) * dt / (1.0 + p.dnu[l] * dt);
  }
  if (c.nl) {
    c.xi3 = p.xi3;
    c.ng1 = (1.0 - p.dnuNL * dt) / (1.0 + p.dnuNL * dt);
    c.ng2 = (p.wNL * p.wNL) * dt / (1.0 + p.dnuNL * dt);
  }
}

// Profiles of a TFSF line's PML object (pmly_ExHz%initialize,
// mod1/BERpmly_ExHz_mod.f90:156-255): the low slab of rank ycoord = 0 on
// cells 1..npml and the high slab of rank Ymax-1 on n-npml+1..n, both in
// the line's medium (eplow = ephigh = eps*ep0, sim1/program.f90:134).
// out: f1, f2, g1, g2, each n+2 doubles.
void tfsf_profiles_dt(double dt, double ds, int N, int np, double eps, double order, double R0, double* out) {
  const Consts& K = consts();
  const double z0 = std::sqrt(K.mu0 / K.ep0);
  const double Ku = dt * K.c0 / ds;
  const double ep = eps * K.ep0, epr = ep / K.ep0;
  double* f1 = out;
  double* f2 = out + (N + 2);
  double* g1 = out + 2 * (N + 2);
  double* g2 = out + 3 * (N + 2);
  for (int j = 0; j <= N + 1; ++j) {
    f1[j] = 1.0;
    f2[j] = Ku / epr;
    g1[j] = 1.0;
    g2[j] = Ku;
  }
  if (np <= 0) return;
  const double sigmam = (order + 1.0) * K.c0 * std::log(1.0 / R0) / (2.0 * np * ds);  // sim1/program.f90:132
  for (int j = 1; j <= np; ++j) {
    const double sE = ep * sigmam * std::pow(((double)(np - j) + 0.5) / (double)np, order);
    f1[j] = std::exp(-(sE * dt / ep));
    f2[j] = (1.0 - f1[j]) / (z0 * sE * ds);
    if (j < np) {
      const double sH = K.mu0 * sigmam * std::pow((double)(np - j) / (double)np, order);
      g1[j] = std::exp(-(sH * dt / K.mu0));
      g2[j] = z0 * (1.0 - g1[j]) / (sH * ds);
    }
  }
  int jj = np;
  for (int j = N - np + 1; j <= N; ++j, --jj) {
    if (jj < np) {
      const double sE = ep * sigmam * std::pow((double)(np - jj) / (double)np, order);
      f1[j] = std::exp(-(sE * dt / ep));
      f2[j] = (1.0 - f1[j]) / (z0 * sE * ds);
    }
    const double sH = K.mu0 * sigmam * std::pow(((double)(np - jj) + 0.5) / (double)np, order);
    g1[j] = std::exp(-(sH * dt / K.mu0));
    g2[j] = z0 * (1.0 - g1[j]) / (sH * ds);
  }
}

// ---- PML profiles (reference formulas, one axis at a time) ----------------
struct Prof {
  double dt, ds, order;
  double sigmam;
  int P;
  void e_low(double ep, int idx, double& f1, double& f2) const {
    const Consts& K = consts();
    const double sig = ep * sigmam * std::pow(((double)(P - idx) + 0.5) / (double)P, order);
    const double alpha = sig * dt / ep;
    f1 = std::exp(-alpha);
    f2 = (1.0 - f1) / (K.z0 * sig * ds);
  }
  void h_low(int idx, double& g1, double& g2) const {
    const Consts& K = consts();
    const double sig = K.mu0 * sigmam * std::pow(((double)(P - idx)) / (double)P, order);
    const double alpha = sig * dt / K.mu0;
    g1 = std::exp(-alpha);
    g2 = K.z0 * (1.0 - g1) / (sig * ds);
  }
  void e_high(double ep, int q, double& f1, double& f2) const {
    const Consts& K = consts();
    const double sig = ep * sigmam * std::pow(((double)(P - q)) / (double)P, order);
    const double alpha = sig * dt / ep;
    f1 = std::exp(-alpha);
    f2 = (1.0 - f1) / (K.z0 * sig * ds);
  }
  void h_high(int q, double& g1, double& g2) const {
    const Consts& K = consts();
    const double sig = K.mu0 * sigmam * std::pow(((double)(P - q) + 0.5) / (double)P, order);
    const double alpha = sig * dt / K.mu0;
    g1 = std::exp(-alpha);
    g2 = K.z0 * (1.0 - g1) / (sig * ds);
  }
  // init_varx1 then init_varx2 shape on indices 1..n
  void grade(double ep, int n, double* f1, double* f2, double* g1, double* g2) const {
    for (int x = 1; x <= P; ++x) {
      e_low(ep, x, f1[x], f2[x]);
      if (x < P) h_low(x, g1[x], g2[x]);
    }
    int q = P;
    for (int x = n - P + 1; x <= n; ++x) {
      if (q < P) e_high(ep, q, f1[x], f2[x]);
      h_high(q, g1[x], g2[x]);
      --q;
    }
  }
};

double sigmam_for(const Sim* s, int P) {
  const Consts& K = consts();
  if (s->pml.sigmam > 0) return s->pml.sigmam;
  return (s->pml.order + 1.0) * K.c0 * std::log(1.0 / s->pml.R0) / (2.0 * P * s->ds);  // sim3/program.f90:102
}

void* upload_bytes(Sim* s, const void* data, size_t bytes);

template <typename real>
void* upload(Sim* s, const std::vector<double>& v) {
  std::vector<real> h(v.begin(), v.end());
  return upload_bytes(s, h.data(), h.size() * sizeof(real));
}

// the copy is ordered after the allocation's clear on the handle's stream;
// the host buffer is a temporary, so wait for the copy
void* upload_bytes(Sim* s, const void* data, size_t bytes) {
  void* p = dev_alloc(s, bytes);
  if (!p) return nullptr;
  if (bytes && hipMemcpyAsync(p, data, bytes, hipMemcpyHostToDevice, s->stream) != hipSuccess) return nullptr;
  if (hipStreamSynchronize(s->stream) != hipSuccess) return nullptr;
  return p;
}

int setup_all_faces(Sim* s) {
  const int n[3] = {s->nx, s->ny, s->nz};
  for (int a = 0; a < 3; ++a) {
    s->P[a] = s->pml.npml[a];
    if (s->P[a] < 0) return set_error(FDTD_ERR_ARG, "negative PML thickness");
    if (s->P[a] > 0 && s->g.periodic[a]) return set_error(FDTD_ERR_ARG, "PML on a periodic axis");
    if (2 * s->P[a] > n[a]) return set_error(FDTD_ERR_ARG, "PML thicker than the grid");
  }
  if (s->P[0] + s->P[1] + s->P[2] > 0 && !(s->pml.order > 0)) return set_error(FDTD_ERR_ARG, "bad PML order");
  const int Px = s->P[0], Py = s->P[1], Pz = s->P[2];
  s->npml_cells = (int64_t)2 * Pz * s->nx * s->ny + (int64_t)(s->nz - 2 * Pz) * 2 * Py * s->nx +
                  (int64_t)(s->nz - 2 * Pz) * (s->ny - 2 * Py) * 2 * Px;
  std::vector<double> prof;
  for (int a = 0; a < 3; ++a) {
    std::vector<double> f1(n[a] + 2, 1.0), f2(n[a] + 2, s->ku_pml), g1(n[a] + 2, 1.0), g2(n[a] + 2, s->ku_pml);
    if (s->P[a] > 0) {
      Prof pr{s->dt, s->ds, s->pml.order, sigmam_for(s, s->P[a]), s->P[a]};
      pr.grade(consts().ep0, n[a], f1.data(), f2.data(), g1.data(), g2.data());
    }
    prof.insert(prof.end(), f1.begin(), f1.end());
    prof.insert(prof.end(), f2.begin(), f2.end());
    prof.insert(prof.end(), g1.begin(), g1.end());
    prof.insert(prof.end(), g2.begin(), g2.end());
  }
  s->d_prof = s->prec == FDTD_FP32 ? upload<float>(s, prof) : upload<double>(s, prof);
  s->d_split = dev_alloc(s, (size_t)S_N * (size_t)std::max<int64_t>(s->npml_cells, 1) * s->wsize);
  if (!s->d_prof || !s->d_split) return set_error(FDTD_ERR_NOMEM, "PML allocation");
  if (s->onepass) {
    s->d_split2 = dev_alloc(s, (size_t)S_N * (size_t)std::max<int64_t>(s->npml_cells, 1) * s->wsize);
    if (!s->d_split2) return set_error(FDTD_ERR_NOMEM, "PML allocation");
  }
  return 0;
}

int setup_ref_y(Sim* s) {
  const Consts& K = consts();
  const int P = s->pml.npml[1];
  if (P < 1 || 2 * P + 2 > s->ny) return set_error(FDTD_ERR_ARG, "bad y PML thickness");
  if (!(s->pml.eps_low > 0) || !(s->pml.eps_high > 0)) return set_error(FDTD_ERR_ARG, "bad PML eps");
  if (!(s->pml.order > 0)) return set_error(FDTD_ERR_ARG, "bad PML order");
  if (s->g.periodic[1]) return set_error(FDTD_ERR_ARG, "REF_Y PML needs a non-periodic y axis");
  const int npx = s->pml.npml[0], npz = s->pml.npml[2];
  if (npx < 0 || npz < 0 || npx > s->nx || npz > s->nz) return set_error(FDTD_ERR_ARG, "bad x/z grading depth");
  s->slab_cells = (int64_t)(s->nx + 2) * (P + 2) * (s->nz + 2);
  s->d_slab = dev_alloc(s, (size_t)2 * S_N * (size_t)s->slab_cells * s->wsize);
  const int ps = std::max(s->nx, std::max(s->ny, s->nz)) + 2;
  s->ref_prof_stride = ps;
  const double eplow = s->pml.eps_low * K.ep0, ephigh = s->pml.eps_high * K.ep0;
  const double sigmam = sigmam_for(s, P);
  std::vector<double> all((size_t)2 * 12 * ps, 0.0);
  for (int sl = 0; sl < 2; ++sl) {
    const bool low = sl == 0;
    double* pr = all.data() + (size_t)sl * 12 * ps;
    double *f1x = pr, *f2x = pr + ps, *f1y = pr + 2 * ps, *f2y = pr + 3 * ps, *f1z = pr + 4 * ps, *f2z = pr + 5 * ps;
    double *g1x = pr + 6 * ps, *g2x = pr + 7 * ps, *g1y = pr + 8 * ps, *g2y = pr + 9 * ps, *g1z = pr + 10 * ps,
           *g2z = pr + 11 * ps;
    // allocate_vars defaults (mod3/BERpmly_mod.f90:567-601)
    const double epr = low ? eplow / K.ep0 : ephigh / K.ep0;
    for (int x = 0; x < ps; ++x) {
      f1x[x] = f1y[x] = f1z[x] = 1.0;
      f2x[x] = s->ku_pml / epr;
      f2y[x] = s->ku_pml / epr;
      f2z[x] = s->ku_pml / epr;
      g1x[x] = g1y[x] = g1z[x] = 1.0;
      g2x[x] = g2y[x] = g2z[x] = s->ku_pml;
    }
    const double ep = low ? eplow : ephigh;
    if (npx > 0) Prof{s->dt, s->ds, s->pml.order, sigmam, npx}.grade(ep, s->nx, f1x, f2x, g1x, g2x);  // :630-656
    Prof py{s->dt, s->ds, s->pml.order, sigmam, P};                                                   // :674-714
    if (low) {
      for (int j = 1; j <= P; ++j) {
        py.e_low(eplow, j, f1y[j], f2y[j]);
        if (j < P) py.h_low(j, g1y[j], g2y[j]);
      }
    } else {
      int jj = P;
      for (int j = s->ny - P + 1; j <= s->ny; ++j) {
        if (jj < P) py.e_high(ephigh, jj, f1y[j], f2y[j]);
        py.h_high(jj, g1y[j], g2y[j]);
        --jj;
      }
    }
    if (npz > 0) Prof{s->dt, s->ds, s->pml.order, sigmam, npz}.grade(ep, s->nz, f1z, f2z, g1z, g2z);  // :719-745
  }
  s->d_ref_prof = s->prec == FDTD_FP32 ? upload<float>(s, all) : upload<double>(s, all);
  if (!s->d_slab || !s->d_ref_prof) return set_error(FDTD_ERR_NOMEM, "PML allocation");
  return 0;
}

inline bool cell_in_pml(const Sim* s, int i, int j, int k) {
  if (s->pml_mode != FDTD_PML_ALL_FACES) return false;
  const int* P = s->P;
  return i <= P[0] || i > s->nx - P[0] || j <= P[1] || j > s->ny - P[1] || k <= P[2] || k > s->nz - P[2];
}

template <typename real>
void* upload_table(Sim* s) {
  const size_t n = s->tab.size();
  std::vector<Vec4<real>> coef(n), gaxe(n);
  std::vector<uint32_t> flags(n);
  for (size_t q = 0; q < n; ++q) {
    const MatEntry<double>& m = s->tab[q];
    uint32_t fl = 0;
    for (int c = 0; c < 3; ++c) {
      coef[q].v[c] = (real)m.coef[c];
      gaxe[q].v[c] = (real)m.gax[c];
      if (m.act[c] == ACT_DIEL) fl |= 1u << c;
      if (m.act[c] == ACT_ADE) fl |= 16u << c;  // written by a dispersive kernel
    }
    coef[q].v[3] = (real)m.pgax;
    gaxe[q].v[3] = (real)m.eps;
    flags[q] = fl;
  }
  std::vector<uint8_t> blob(n * (2 * sizeof(Vec4<real>) + sizeof(uint32_t)));
  std::memcpy(blob.data(), coef.data(), n * sizeof(Vec4<real>));
  std::memcpy(blob.data() + n * sizeof(Vec4<real>), gaxe.data(), n * sizeof(Vec4<real>));
  std::memcpy(blob.data() + 2 * n * sizeof(Vec4<real>), flags.data(), n * sizeof(uint32_t));
  return upload_bytes(s, blob.data(), blob.size());
}

int commit_impl(Sim* s) {
  const Consts& K = consts();
  (void)K;
  // engine
  int eng = s->g.engine;
  if (eng == FDTD_ENGINE_AUTO) eng = s->pml_mode == FDTD_PML_REF_Y ? FDTD_ENGINE_STAGED : FDTD_ENGINE_FUSED;
  if (eng == FDTD_ENGINE_FUSED && s->pml_mode == FDTD_PML_REF_Y)
    return set_error(FDTD_ERR_UNSUPPORTED, "the REF_Y PML runs on the STAGED engine");
  if (eng == FDTD_ENGINE_STAGED && s->pml_mode == FDTD_PML_ALL_FACES)
    return set_error(FDTD_ERR_UNSUPPORTED, "the ALL_FACES PML runs on the FUSED engine");
  if (eng != FDTD_ENGINE_FUSED && eng != FDTD_ENGINE_STAGED) return set_error(FDTD_ERR_ARG, "bad engine");
  s->engine = eng;
  if ((int)s->srcs.size() > kMaxSrc) return set_error(FDTD_ERR_ARG, "too many sources");

  // launch shape: whole 64-lane wavefronts along x, 256 threads per block
  {
    const int r = (s->nx + 63) / 64 * 64;
    s->bx = std::min(256, r);
    s->by = 256 / s->bx;
    // tuning overrides: FDTD_BLOCK_X in {64,128,256,512}, FDTD_BLOCK_Y in
    // {1,2,4,8} (at most 512 threads), FDTD_XCD_MAP in {0,1}
    if (const char* e = std::getenv("FDTD_BLOCK_X")) {
      const int v = std::atoi(e);
      if (v == 64 || v == 128 || v == 256 || v == 512) {
        s->bx = v;
        s->by = std::max(1, 256 / v);
      }
    }
    if (const char* e = std::getenv("FDTD_BLOCK_Y")) {
      const int v = std::atoi(e);
      if ((v == 1 || v == 2 || v == 4 || v == 8) && v * s->bx <= 512) s->by = v;
    }
    if (const char* e = std::getenv("FDTD_XCD_MAP")) s->xcdmap = std::atoi(e) != 0;
  }
  // layout: rows are whole 128-byte lines and the first interior cell of
  // every row starts a line (the field pointer is offset by line-1 cells
  // inside its allocation), so each wavefront's 64 cells span the fewest lines
  const int64_t le = 128 / (int64_t)s->wsize;  // cells per line
  s->px = ((int64_t)s->nx + 2 + le - 1) / le * le;
  s->pxy = s->px * (s->ny + 2);
  s->ncell_alloc = s->pxy * (s->nz + 2);
  if (s->ncell_alloc + le >= ((int64_t)1 << 31)) return set_error(FDTD_ERR_ARG, "grid too large for one handle");
  const size_t fbytes = (size_t)(s->ncell_alloc + le) * s->wsize;
  const size_t shift = (size_t)(le - 1) * s->wsize;
  // one-pass step: FUSED engine without Lorentz-Jp materials (their E form
  // reads the previous E in place); FDTD_ONEPASS=0 keeps the two-kernel step
  s->onepass = 0;
  if (s->engine == FDTD_ENGINE_FUSED) {
    bool ljp = false;
    for (const auto& m : s->mats) ljp |= m.type == 3;
    // the one-pass step stays off unless asked for (FDTD_ONEPASS=1) or
    // picked by fdtd_autotune, which then allocates the second copies
    const char* env = std::getenv("FDTD_ONEPASS");
    s->onepass = !ljp && env && std::atoi(env) == 1;
    if (const char* e = std::getenv("FDTD_ONEPASS_BY")) {
      const int v = std::atoi(e);
      if (v == 2 || v == 4 || (v == 8 && s->prec == FDTD_FP32)) s->op_by = v;
    }
    if (const char* e = std::getenv("FDTD_ONEPASS_KC")) {
      const int v = std::atoi(e);
      if (v >= 2 && v <= 4096) s->op_kc = v;
    }
    if (const char* e = std::getenv("FDTD_DFT_FUSE")) s->dft_fuse = std::atoi(e) != 0;
    if (const char* e = std::getenv("FDTD_ZMARCH")) {
      const int v = std::atoi(e);
      if (v == 1 || v == 2 || v == 4 || v == 8) s->zm = v;
    }
  }
  for (int q = 0; q < 3; ++q) {
    char* e = (char*)dev_alloc(s, fbytes);
    char* h = (char*)dev_alloc(s, fbytes);
    if (!e || !h) return set_error(FDTD_ERR_NOMEM, "field allocation");
    s->E[q] = e + shift;
    s->H[q] = h + shift;
    if (s->onepass) {
      char* e2 = (char*)dev_alloc(s, fbytes);
      char* h2 = (char*)dev_alloc(s, fbytes);
      if (!e2 || !h2) return set_error(FDTD_ERR_NOMEM, "field allocation");
      s->E2[q] = e2 + shift;
      s->H2[q] = h2 + shift;
    }
    if (s->engine == FDTD_ENGINE_STAGED) {
      char* d = (char*)dev_alloc(s, fbytes);
      if (!d) return set_error(FDTD_ERR_NOMEM, "field allocation");
      s->D[q] = d + shift;
    }
  }

  // PML
  if (s->pml_mode == FDTD_PML_ALL_FACES) {
    int rc = setup_all_faces(s);
    if (rc) return rc;
  } else if (s->pml_mode == FDTD_PML_REF_Y) {
    int rc = setup_ref_y(s);
    if (rc) return rc;
  }

  // dispersive groups (identical coefficients share one group)
  s->groups.clear();
  for (auto& m : s->mats) {
    if (m.type == 0) continue;
    if (m.type == 4) {
      LordCoef c;
      lord_coef_dt(s->dt, m.lj, c);
      int gi = -1;
      for (size_t q = 0; q < s->groups.size(); ++q)
        if (s->groups[q].kind == 2 && std::memcmp(&s->groups[q].ld, &c, sizeof c) == 0) gi = (int)q;
      if (gi < 0) {
        AdeGroup g;
        g.kind = 2;
        std::memset(&g.coef, 0, sizeof g.coef);
        g.ld = c;
        g.eps_inf = 1.0;
        g.nstate = (s->engine == FDTD_ENGINE_FUSED ? 1 : 0) + 2 * c.npols;  // [D], S_l, S_l older
        s->groups.push_back(g);
        gi = (int)s->groups.size() - 1;
      }
      m.group = gi;
      continue;
    }
    if (m.type == 3) {
      LjpCoef<double> c;
      ljp_coef_dt(s->dt, s->ds, m.lj, c);
      int gi = -1;
      for (size_t q = 0; q < s->groups.size(); ++q)
        if (s->groups[q].kind == 1 && std::memcmp(&s->groups[q].lj, &c, sizeof c) == 0) gi = (int)q;
      if (gi < 0) {
        AdeGroup g;
        g.kind = 1;
        std::memset(&g.coef, 0, sizeof g.coef);
        g.lj = c;
        g.eps_inf = m.lj.eps_inf;
        g.nstate = 2 * c.npols + (c.nl ? 2 : 0);  // P_l, Jp_l, [Q, G]
        s->groups.push_back(g);
        gi = (int)s->groups.size() - 1;
      }
      m.group = gi;
      continue;
    }
    AdeCoef<double> c;
    double einf = 1.0;
    if (m.type == 1) {
      drude2cp_coef(s, m.d2, c, &einf);
    } else {
      int rc = graphene_coef_dt(s->dt, s->ds, m.gr, c, &einf);
      if (rc) return rc;
    }
    int gi = -1;
    for (size_t q = 0; q < s->groups.size(); ++q)
      if (s->groups[q].kind == 0 && std::memcmp(&s->groups[q].coef, &c, sizeof c) == 0) gi = (int)q;
    if (gi < 0) {
      AdeGroup g;
      g.coef = c;
      g.eps_inf = einf;
      g.nstate = (s->engine == FDTD_ENGINE_FUSED ? 1 : 0) + 2 + 2 * c.ncp + (c.raman ? 2 : 0);
      s->groups.push_back(g);
      gi = (int)s->groups.size() - 1;
    }
    m.group = gi;
  }

  // rasterise: per cell, the combination of last writers per component
  // (materials applied in call order, sim3/program.f90:411-414)
  using Key = std::array<int16_t, 5>;  // writer x,y,z, last dielectric, last any
  std::vector<Key> combos;
  std::map<Key, int> index;
  combos.push_back({-1, -1, -1, -1, -1});
  index[combos[0]] = 0;
  std::vector<uint8_t> cid((size_t)s->ncell_alloc, 0);
  for (size_t mi = 0; mi < s->mats.size(); ++mi) {
    const HostMaterial& m = s->mats[mi];
    int lo[3], hi[3];
    if (!region_bounds(s, m.reg, lo, hi)) return set_error(FDTD_ERR_ARG, "bad region");
    std::vector<int> trans(256, -1);
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) {
          if (!region_inside(m.reg, i, j, k)) continue;
          const size_t c = (size_t)(i + s->px * j + s->pxy * k);
          const int old = cid[c];
          if (trans[old] < 0) {
            Key key = combos[old];
            for (int q = 0; q < 3; ++q)
              if (m.writes[q]) key[q] = (int16_t)mi;
            if (m.type == 0) key[3] = (int16_t)mi;
            key[4] = (int16_t)mi;
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
  // material table
  s->tab.assign(combos.size(), MatEntry<double>{});
  for (size_t q = 0; q < combos.size(); ++q) {
    const Key& key = combos[q];
    MatEntry<double>& t = s->tab[q];
    for (int c = 0; c < 3; ++c) {
      const int w = key[c];
      t.act[c] = ACT_KEEP;
      t.gax[c] = 1.0;
      t.coef[c] = s->ku;
      if (w < 0) continue;
      const HostMaterial& m = s->mats[w];
      if (m.type == 0) {
        const double gax = (m.eps_r == 0.0) ? 0.0 : 1.0 / m.eps_r;  // simpledielectric_mod.f90:91-97
        t.act[c] = ACT_DIEL;
        t.gax[c] = gax;
        t.coef[c] = s->ku * gax;
      } else if (m.type == 3 && !(m.lj.jp_mask & (1 << c))) {
        // a Lorentz-Jp material writes its other components as vacuum
        // (mod3/grapheneLorentzJp_mod.f90:66: Ey += Ku*curl)
        t.act[c] = ACT_DIEL;
        t.gax[c] = 1.0;
        t.coef[c] = s->ku;
      } else {
        t.act[c] = ACT_ADE;
      }
    }
    t.pgax = key[3] >= 0 ? (s->mats[key[3]].eps_r == 0.0 ? 0.0 : 1.0 / s->mats[key[3]].eps_r) : 1.0;
    if (key[4] >= 0) {
      const HostMaterial& m = s->mats[key[4]];
      t.eps = m.type == 0 ? m.eps_r : s->groups[m.group].eps_inf;
    } else {
      t.eps = 1.0;
    }
    t.pad = 0;
  }
  s->ntab = (int)s->tab.size();
  s->d_mattab = s->prec == FDTD_FP32 ? upload_table<float>(s) : upload_table<double>(s);
  s->d_matid = (uint8_t*)upload_bytes(s, cid.data(), cid.size());

  // the uniform material id of every (j,k) row (-1 where the row mixes ids)
  {
    std::vector<int32_t> rowmat((size_t)(s->ny + 2) * (s->nz + 2), -1);
    for (int k = 1; k <= s->nz; ++k)
      for (int j = 1; j <= s->ny; ++j) {
        const uint8_t* row = cid.data() + (size_t)(s->px * j + s->pxy * k);
        bool uni = true;
        for (int i = 2; i <= s->nx && uni; ++i) uni = row[i] == row[1];
        if (uni) rowmat[(size_t)j + (size_t)(s->ny + 2) * k] = row[1];
      }
    s->d_rowmat = (int32_t*)upload_bytes(s, rowmat.data(), rowmat.size() * sizeof(int32_t));
    if (!s->d_rowmat) return set_error(FDTD_ERR_NOMEM, "row material table");
    // runs of equal ids along each row (the one-pass tile classifier asks
    // whether a 64-cell stretch of a row holds one material)
    s->h_runoff.assign(rowmat.size() + 1, 0);
    s->h_runx.clear();
    s->h_runid.clear();
    for (size_t r = 0; r < rowmat.size(); ++r) {
      s->h_runoff[r] = (int32_t)s->h_runx.size();
      const int j = (int)(r % (size_t)(s->ny + 2)), k = (int)(r / (size_t)(s->ny + 2));
      if (j < 1 || j > s->ny || k < 1 || k > s->nz) continue;
      const uint8_t* row = cid.data() + (size_t)(s->px * j + s->pxy * k);
      for (int i = 1; i <= s->nx; ++i)
        if (i == 1 || row[i] != row[i - 1]) {
          s->h_runx.push_back(i);
          s->h_runid.push_back(row[i]);
        }
    }
    s->h_runoff[rowmat.size()] = (int32_t)s->h_runx.size();
    std::vector<int32_t> mix;
    for (int k = 1; k <= s->nz; ++k)
      for (int j = 1; j <= s->ny; ++j)
        if (rowmat[(size_t)j + (size_t)(s->ny + 2) * k] < 0) mix.push_back(j + (s->ny + 2) * k);
    s->nmix = (int)mix.size();
    if (s->nmix) {
      s->d_mixrows = (int32_t*)upload_bytes(s, mix.data(), mix.size() * sizeof(int32_t));
      if (!s->d_mixrows) return set_error(FDTD_ERR_NOMEM, "mixed row list");
    }
  }
  if (!s->d_mattab || !s->d_matid) return set_error(FDTD_ERR_NOMEM, "material table");

  // dispersive cell lists
  std::vector<std::vector<int32_t>> gcells(s->groups.size());
  std::vector<std::vector<uint8_t>> gmask(s->groups.size());
  s->n_pml = 0;
  s->n_ade = 0;
  for (int k = 1; k <= s->nz; ++k)
    for (int j = 1; j <= s->ny; ++j)
      for (int i = 1; i <= s->nx; ++i) {
        const bool pm = cell_in_pml(s, i, j, k);
        if (pm) ++s->n_pml;
        const int64_t c = i + s->px * j + s->pxy * k;
        const Key& key = combos[cid[(size_t)c]];
        if (pm) continue;
        int gm[3] = {-1, -1, -1};
        for (int q = 0; q < 3; ++q)
          if (key[q] >= 0 && is_ade_comp(s->mats[key[q]], q)) gm[q] = s->mats[key[q]].group;
        bool any = false;
        for (size_t g = 0; g < s->groups.size(); ++g) {
          uint8_t mk = 0;
          for (int q = 0; q < 3; ++q)
            if (gm[q] == (int)g) mk |= (uint8_t)(1 << q);
          if (mk) {
            gcells[g].push_back((int32_t)c);
            gmask[g].push_back(mk);
            any = true;
          }
        }
        if (any) ++s->n_ade;
      }
  s->n_main = (int64_t)s->nx * s->ny * s->nz - s->n_pml;
  for (size_t g = 0; g < s->groups.size(); ++g) {
    AdeGroup& G = s->groups[g];
    G.ncells = (int64_t)gcells[g].size();
    if (!G.ncells) continue;
    G.d_cells = (int32_t*)upload_bytes(s, gcells[g].data(), gcells[g].size() * sizeof(int32_t));
    G.d_mask = (uint8_t*)upload_bytes(s, gmask[g].data(), gmask[g].size());
    // Lorentz-Jp pole state is FP64 in both precisions (k_ljp)
    G.d_state = dev_alloc(s, (size_t)G.nstate * 3 * (size_t)G.ncells * 8);  // FP64 state in both precisions
    if (!G.d_cells || !G.d_mask || !G.d_state) return set_error(FDTD_ERR_NOMEM, "dispersive state");
  }

  // TFSF incident lines
  {
    int used = (int)s->srcs.size();
    for (auto& t : s->tfsfs) used += t.p.hi > 0 ? 2 : 1;
    if (used > kMaxSrc) return set_error(FDTD_ERR_ARG, "TFSF corrections and sources exceed 8 source slots");
  }
  for (auto& t : s->tfsfs) {
    const int a = t.p.axis;
    const int n[3] = {s->nx, s->ny, s->nz};
    int P = 0;
    if (s->pml_mode == FDTD_PML_ALL_FACES) P = s->P[a];
    if (s->pml_mode == FDTD_PML_REF_Y) {
      if (a != 1) return set_error(FDTD_ERR_UNSUPPORTED, "with the reference y PML a TFSF wave runs along y");
      P = s->pml.npml[1];
    }
    if (t.p.lo - 1 <= P || (t.p.hi > 0 && t.p.hi + 1 > n[a] - P) || (t.p.hi == 0 && t.p.lo > n[a] - P))
      return set_error(FDTD_ERR_ARG, "TFSF planes inside the PML");
    // Lorentz-Jp cells read curl H directly: keep them off the D correction planes
    for (const auto& m : s->mats) {
      if (m.type != 3) continue;
      int lo[3], hi[3];
      if (!region_bounds(s, m.reg, lo, hi)) continue;
      const int pl[2] = {t.p.lo, t.p.hi > 0 ? t.p.hi + 1 : -1};
      for (int q = 0; q < 2; ++q)
        if (pl[q] >= lo[a] && pl[q] <= hi[a]) return set_error(FDTD_ERR_UNSUPPORTED, "TFSF plane through Lorentz-Jp cells");
    }
    std::vector<double> line((size_t)8 * (t.n + 2), 0.0);
    tfsf_profiles_dt(s->dt, s->ds, t.n, t.p.npml, t.p.eps, t.p.order, t.p.R0, line.data() + (size_t)4 * (t.n + 2));
    t.d_line = (double*)upload_bytes(s, line.data(), line.size() * sizeof(double));
    if (!t.d_line) return set_error(FDTD_ERR_NOMEM, "TFSF line");
    std::vector<int32_t> lp;
    const int tid = (int)(&t - s->tfsfs.data());
    for (size_t q = 0; q < s->probes.size(); ++q)
      if (s->probes[q].line == tid) {
        lp.push_back((int32_t)q);
        lp.push_back(s->probes[q].j);
      }
    t.nlp = (int)lp.size() / 2;
    if (t.nlp) {
      t.d_lp = (int32_t*)upload_bytes(s, lp.data(), lp.size() * sizeof(int32_t));
      if (!t.d_lp) return set_error(FDTD_ERR_NOMEM, "TFSF probes");
    }
  }

  // monitors
  for (auto& d : s->dfts) {
    d.nslots = 0;
    for (int c = 0; c < 6; ++c) d.slot[c] = (d.mask & (1 << c)) ? d.nslots++ : -1;
    d.d_omega = (double*)upload_bytes(s, d.omega.data(), d.omega.size() * sizeof(double));
    d.d_acc = (double*)dev_alloc(s, sizeof(double) * 2 * (size_t)d.nslots * d.nfreq * (size_t)d.n1 * d.n2);
    d.d_samp = dev_alloc(s, s->wsize * (size_t)kDftBlock * d.nslots * (size_t)d.n1 * d.n2);
    d.d_tab = (double*)dev_alloc(s, sizeof(double) * 2 * (size_t)kDftBlock * d.nfreq);
    if (!d.d_omega || !d.d_acc || !d.d_samp || !d.d_tab) return set_error(FDTD_ERR_NOMEM, "DFT accumulators");
    d.nbuf = 0;
  }
  if (!s->probes.empty()) {
    std::vector<int32_t> pos;
    for (auto& p : s->probes) {
      pos.push_back(p.line >= 0 ? -1 : p.i);  // k_probe skips incident-line probes
      pos.push_back(p.j);
      pos.push_back(p.k);
    }
    s->d_probe_pos = (int32_t*)upload_bytes(s, pos.data(), pos.size() * sizeof(int32_t));
    if (!s->d_probe_pos) return set_error(FDTD_ERR_NOMEM, "probe table");
  }
  FDTD_HIP(hipStreamSynchronize(s->stream));
  s->committed = true;
  return 0;
}

int ensure_committed(Sim* s) {
  if (s->committed) return 0;
  FDTD_HIP(hipSetDevice(s->device));
  return commit_impl(s);
}

int flush_profile(Sim* s) {
  if (s->pending.empty()) return 0;
  FDTD_HIP(hipStreamSynchronize(s->stream));
  for (auto& pe : s->pending) {
    float ms = 0.f;
    FDTD_HIP(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
    s->k_ms[pe.first] += ms;
    s->k_launches[pe.first] += 1;
    s->event_pool.push_back(pe.second.first);
    s->event_pool.push_back(pe.second.second);
  }
  s->pending.clear();
  return 0;
}

template <typename real>
int read_field(Sim* s, void* src, double* dst) {
  std::vector<real> h((size_t)s->ncell_alloc);
  FDTD_HIP(hipMemcpyAsync(h.data(), src, h.size() * sizeof(real), hipMemcpyDeviceToHost, s->stream));
  FDTD_HIP(hipStreamSynchronize(s->stream));
  size_t o = 0;
  for (int k = 1; k <= s->nz; ++k)
    for (int j = 1; j <= s->ny; ++j) {
      const real* row = h.data() + (size_t)(s->px * j + s->pxy * k);
      for (int i = 1; i <= s->nx; ++i) dst[o++] = (double)row[i];
    }
  return 0;
}

template <typename real>
int write_field(Sim* s, void* dstp, const double* src) {
  std::vector<real> h((size_t)s->ncell_alloc);
  FDTD_HIP(hipMemcpyAsync(h.data(), dstp, h.size() * sizeof(real), hipMemcpyDeviceToHost, s->stream));
  FDTD_HIP(hipStreamSynchronize(s->stream));
  size_t o = 0;
  for (int k = 1; k <= s->nz; ++k)
    for (int j = 1; j <= s->ny; ++j) {
      real* row = h.data() + (size_t)(s->px * j + s->pxy * k);
      for (int i = 1; i <= s->nx; ++i) row[i] = (real)src[o++];
    }
  FDTD_HIP(hipMemcpyAsync(dstp, h.data(), h.size() * sizeof(real), hipMemcpyHostToDevice, s->stream));
  FDTD_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

void* field_ptr(Sim* s, int comp) {
  if (comp < 3) return s->E[comp];
  if (comp < 6) return s->H[comp - 3];
  return s->D[comp - 6];
}

}  // namespace

// Tile lists of the one-pass step for tiles of 64 x by cells (x from 1 in
// steps of 64, y from 1 in steps of by) and z ranges of at most kc planes;
// 8 words per tile: i0, j0, k0, k1, xlo, xhi, then (plain tiles) the
// material id and 1 for an x-face tile (0 otherwise).
//
// Plain tiles cover x from the first 128-byte line after the low x PML
// slab to the last line before the cell next to the high slab, in steps of
// 64, so that no 128-byte line is written partly by a plain and partly by a
// general tile (partial-line stores cost far more than their bytes); the x
// edges outside that range are general.  Plain tiles are listed row by row
// (x fastest): the lines a tile gathers for its halo columns are its x
// neighbours' own lines, which those load at about the same time (listed
// column by column instead, the halo reads from HBM grow by 40 %).  A tile
// is plain when
// every cell whose new E it computes -- the cells of its owned x range
// [xlo, xhi] over its rows and planes, the cell after xhi, the row after
// its rows, the plane after its z range, each wrapped to cell 1 on a
// periodic axis -- is a non-PML cell of one material whose three components
// are dielectric, and no x- or y-normal source plane crosses
// them (z-normal ones are added per plane); the z ranges end one plane
// before the high z slab, so no such cell is a PML cell there.  FP32 with
// ALL_FACES PML on a non-periodic x axis and no TFSF line (FDTD_XFACE=0
// turns it off): plain tiles cover every x from 1, and those whose lanes
// reach into an x slab (x-face tiles, bx rows each, listed after the
// others) update the x-slab cells with the split-field rule of an x face
// cell -- their rows and planes, the row and plane after included, lie
// between the y and z slabs, the cell after xhi is no x-slab cell, and no
// z-normal source plane crosses them.  General
// work goes to the list as workgroups over at most kGenKc planes of a box
// (FDTD_GEN_KC overrides): 8 words each, x0, x1, j0, j1, k0, the
// workgroup's width (64, 32 or 16), k1, and the face kind (face_of).
int build_march_tiles(Sim* s, int by, int kc) {
  if (s->tl_by == by && s->tl_kc == kc) return 0;
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const bool af = s->pml_mode == FDTD_PML_ALL_FACES;
  const int Px = af ? s->P[0] : 0, Py = af ? s->P[1] : 0, Pz = af ? s->P[2] : 0;
  const int pers[3] = {s->g.periodic[0] ? 1 : 0, s->g.periodic[1] ? 1 : 0, s->g.periodic[2] ? 1 : 0};
  std::vector<std::pair<int, int>> zr;
  auto segment = [&](int a, int b) {
    for (int k = a; k <= b; k += kc) zr.push_back({k, std::min(b, k + kc - 1)});
  };
  if (Pz > 0) {
    segment(1, Pz);
    segment(Pz + 1, nz - Pz - 1);
    segment(std::max(Pz + 1, nz - Pz), nz - Pz);
    segment(nz - Pz + 1, nz);
  } else {
    segment(1, nz);
  }
  // x- and y-normal source planes, TFSF terms on D included
  std::vector<std::pair<int, int>> planes;  // (axis, index)
  for (const auto& src : s->srcs)
    if (src.axis != 2) planes.push_back({src.axis, src.index});
  for (const auto& t : s->tfsfs)
    if (t.p.axis != 2) {
      planes.push_back({t.p.axis, t.p.lo});
      if (t.p.hi > 0) planes.push_back({t.p.axis, t.p.hi + 1});
    }
  auto plain_id = [&](int id) {
    const MatEntry<double>& m = s->tab[(size_t)id];
    return m.act[0] == ACT_DIEL && m.act[1] == ACT_DIEL && m.act[2] == ACT_DIEL;
  };
  // the one material id of cells [a, b] of row (j, k), or -1
  auto seg_id = [&](int j, int k, int a, int b) {
    const size_t r = (size_t)j + (size_t)(ny + 2) * k;
    const int lo = s->h_runoff[r], hi = s->h_runoff[r + 1];
    if (lo == hi) return -1;
    const int q = (int)(std::upper_bound(s->h_runx.begin() + lo, s->h_runx.begin() + hi, a) - s->h_runx.begin()) - 1;
    if (q < lo) return -1;
    const int end = q + 1 < hi ? s->h_runx[(size_t)q + 1] - 1 : nx;
    return end >= b ? (int)s->h_runid[(size_t)q] : -1;
  };
  auto pml_y = [&](int j) { return j <= Py || j > ny - Py; };
  auto pml_z = [&](int k) { return k <= Pz || k > nz - Pz; };
  // cell v of an axis of n as a tile sees it (0: a dead ghost)
  auto wrap = [](int v, int n, int per) { return v <= n ? v : (v == n + 1 && per ? 1 : 0); };
  // plain tiles start at the first 128-byte line after the low x slab and
  // end with the last line before the cell next to the high slab (i = 1
  // starts a line), so the general x edges are whole lines -- or, in FP32
  // (x-face tiles), they cover every x from 1, and the lanes of a tile that
  // lie in an x slab take the split update of an x face cell
  const int le = (int)(128 / s->wsize);
  const char* xf_env = std::getenv("FDTD_XFACE");
  const bool xface = af && Px > 0 && !pers[0] && s->wsize == 4 && s->tfsfs.empty() &&
                     !(xf_env && std::atoi(xf_env) == 0);
  const int xs = xface ? 1 : (Px > 0 ? 1 + le * ((Px + le - 1) / le) : 1);
  const int xe = xface ? nx : (Px > 0 ? le * ((nx - Px - 1) / le) : nx);
  auto pml_x = [&](int i) { return i <= Px || i > nx - Px; };
  const int ntx = xe >= xs ? (xe - xs + 64) / 64 : 0;
  std::vector<int32_t> plain, gen, xfl;
  struct GBox {
    int x0, x1, j0, j1, k0, k1;
    int face;  // 1, 2: every cell a face cell of the y or z slabs (see face_of)
  };
  // A general box whose cells are all face cells of the y slabs (1) or of
  // the z slabs (2) -- x inside the x interior, the other axis's slabs not
  // touched -- in rows of one plain dielectric each, with no source or TFSF
  // plane through it: the general kernels update it with the slab axis
  // fixed at compile time (the same arithmetic, fewer tests and loads).
  auto face_of = [&](int x0, int x1, int j0, int j1, int k0, int k1) {
    // 64-wide workgroups only (more than 32 cells): the E update reads the
    // row's id with scalar loads, which needs one row per wavefront
    if (!af || x0 <= Px || x1 > nx - Px || x1 - x0 + 1 <= 32) return 0;
    bool allz = true, noz = true, ally = true, noy = true;
    for (int k = k0; k <= k1; ++k) {
      allz = allz && pml_z(k);
      noz = noz && !pml_z(k);
    }
    for (int j = j0; j <= j1; ++j) {
      ally = ally && pml_y(j);
      noy = noy && !pml_y(j);
    }
    const int face = (allz && noy) ? 2 : ((ally && noz) ? 1 : 0);
    if (!face) return 0;
    for (int k = k0; k <= k1; ++k)
      for (int j = j0; j <= j1; ++j) {
        const int id = seg_id(j, k, 1, nx);  // the row's one id (the kernel reads it from rowmat)
        if (id < 0 || !plain_id(id)) return 0;
      }
    auto crosses = [&](int axis, int index) {
      const int lo = axis == 0 ? x0 : (axis == 1 ? j0 : k0), hi = axis == 0 ? x1 : (axis == 1 ? j1 : k1);
      return index >= lo && index <= hi;
    };
    for (const auto& src : s->srcs)
      if (crosses(src.axis, src.index)) return 0;
    for (const auto& t : s->tfsfs)
      if (crosses(t.p.axis, t.p.lo) || (t.p.hi > 0 && crosses(t.p.axis, t.p.hi + 1))) return 0;
    return face;
  };
  std::vector<GBox> gbox;
  int64_t nplain = 0, nxfi = 0, nxf = 0;
  int ntxf = 0;
  std::vector<int> ys, zs;
  // pass 0: the plain tiles of by rows; pass 1 (x-face tiles only): the
  // tiles with lanes in an x slab, bx rows each (FDTD_XFACE_BY, 2 or 4),
  // listed after the others and marched by their own launch
  const char* xb_env = std::getenv("FDTD_XFACE_BY");
  const int bx = xb_env && std::atoi(xb_env) == 4 ? 4 : 2;
  for (int pass = 0; pass < (xface ? 2 : 1); ++pass) {
  const int bt = pass ? bx : by;
  const int ntyp = (ny + bt - 1) / bt;
  for (const auto& z : zr)
    for (int ty = 0; ty < ntyp; ++ty)
      for (int tx = -1; tx <= ntx; ++tx) {
        const int j0 = 1 + bt * ty, k0 = z.first, k1 = z.second;
        const int jend = std::min(j0 + bt - 1, ny);
        if (pass == 0 && xface && tx >= 0 && tx < ntx) {  // x-face columns: pass 1
          const int a = xs + 64 * tx, b = std::min(a + 63, xe);
          if (pml_x(a) || pml_x(b)) continue;
        }
        if (pass == 1) {
          if (tx < 0 || tx == ntx) continue;
          const int a = xs + 64 * tx, b = std::min(a + 63, xe);
          if (!(pml_x(a) || pml_x(b))) continue;
        }
        if (tx < 0 || tx == ntx) {  // the x edges: general
          const int a = tx < 0 ? 1 : std::max(xe + 1, xs), b = tx < 0 ? std::min(xs - 1, nx) : nx;
          if (a <= b) gbox.push_back({a, b, j0, jend, k0, k1, 0});
          continue;
        }
        const int i0 = xs + 64 * tx;
        const int xlo = i0, xhi = std::min(i0 + 63, xe);
        if (xlo > xhi) continue;
        bool ok = true;
        // rows and planes E is computed on: own ones, the row and the plane
        // after (wrapped on a periodic axis); none may be a PML row or plane
        ys.clear();
        zs.clear();
        for (int j = j0; j <= std::min(j0 + bt - 1, ny); ++j) ys.push_back(j);
        if (j0 + bt - 1 <= ny) {
          if (const int jh = wrap(j0 + bt, ny, pers[1])) ys.push_back(jh);
        } else if (pers[1]) {
          ys.push_back(1);  // the wrapped row after ny inside the tile
        }
        for (int k = k0; k <= k1; ++k) zs.push_back(k);
        if (const int kh = wrap(k1 + 1, nz, pers[2])) zs.push_back(kh);
        for (int j : ys) ok = ok && !pml_y(j);
        for (int k : zs) ok = ok && !pml_z(k);
        // the cell after xhi (whose new E the tile computes) is no x-slab
        // cell; a tile with lanes in an x slab (x face) takes no z-normal
        // source plane either (its split update has no source term)
        const bool xft = xface && (pml_x(xlo) || pml_x(xhi));
        if (xhi + 1 <= nx && pml_x(xhi + 1)) ok = false;
        if (xft)
          for (const auto& src : s->srcs)
            if (src.axis == 2)
              for (int k : zs) ok = ok && k != src.index;
        // one plain material over the cells computed: [xlo, xhi] and the
        // cell after xhi (cell 1 past nx on a periodic x axis) of every row
        int mid = -1;
        const int xa = std::min(xhi + 1, nx);
        const bool wrap1 = xhi + 1 > nx && pers[0];
        for (size_t a = 0; a < zs.size() && ok; ++a)
          for (size_t b = 0; b < ys.size() && ok; ++b) {
            const int rm = seg_id(ys[b], zs[a], xlo, xa);
            if (rm < 0 || !plain_id(rm) || (mid >= 0 && rm != mid)) ok = false;
            if (ok && wrap1 && seg_id(ys[b], zs[a], 1, 1) != rm) ok = false;
            mid = rm;
          }
        // x- and y-normal source planes through the cells computed (x: up
        // to the cell after xhi, wrapped to 1 past nx)
        for (const auto& pl : planes) {
          if (!ok) break;
          if (pl.first == 0) {
            const int xn = xhi + 1 <= nx ? xhi + 1 : (pers[0] ? 1 : 0);  // the cell after xhi
            ok = !(pl.second >= xlo && pl.second <= xhi) && pl.second != xn;
          } else {
            for (int j : ys) ok = ok && j != pl.second;
          }
        }
        if (ok) {
          const int32_t e[8] = {i0, j0, k0, k1, xlo, xhi, mid, xft ? 1 : 0};
          std::vector<int32_t>& to = xft ? xfl : plain;
          to.insert(to.end(), e, e + 8);
          const int64_t rk = (int64_t)(jend - j0 + 1) * (k1 - k0 + 1);
          // x-slab cells of the owned range: [xlo, Px] and [nx-Px+1, xhi]
          const int64_t nsl = xft ? (int64_t)std::max(0, std::min(xhi, Px) - xlo + 1) +
                                        std::max(0, xhi - std::max(xlo, nx - Px + 1) + 1)
                                  : 0;
          (xft ? nxfi : nplain) += (xhi - xlo + 1 - nsl) * rk;
          nxf += nsl * rk;
          ntxf += xft ? 1 : 0;
        } else {
          gbox.push_back({xlo, xhi, j0, jend, k0, k1, face_of(xlo, xhi, j0, jend, k0, k1)});
        }
      }
  }
  // general boxes: rows of the same x range and z range that follow each
  // other merge; a box is cut into workgroups of 256 threads, 64 x 4 (wide
  // boxes), 32 x 8 or 16 x 16 cells (the x edges), each over a chunk of the
  // box's planes
  std::sort(gbox.begin(), gbox.end(), [](const GBox& u, const GBox& v) {
    if (u.k0 != v.k0) return u.k0 < v.k0;
    if (u.x0 != v.x0) return u.x0 < v.x0;
    if (u.x1 != v.x1) return u.x1 < v.x1;
    if (u.face != v.face) return u.face < v.face;
    return u.j0 < v.j0;
  });
  std::vector<GBox> merged;
  for (const GBox& g : gbox) {
    if (!merged.empty()) {
      GBox& m = merged.back();
      if (m.k0 == g.k0 && m.k1 == g.k1 && m.x0 == g.x0 && m.x1 == g.x1 && m.face == g.face && m.j1 + 1 == g.j0) {
        m.j1 = g.j1;
        continue;
      }
    }
    merged.push_back(g);
  }
  const char* gkc_env = std::getenv("FDTD_GEN_KC");
  const int gkc = gkc_env && std::atoi(gkc_env) > 0 ? std::atoi(gkc_env) : kGenKc;
  s->tl_gkc = gkc;
  for (const GBox& g : merged) {
    const int w = g.x1 - g.x0 + 1;
    const int bw = w <= 16 ? 16 : (w <= 32 ? 32 : 64), bh = 256 / bw;
    for (int k = g.k0; k <= g.k1; k += gkc)
      for (int j = g.j0; j <= g.j1; j += bh)
        for (int x = g.x0; x <= g.x1; x += bw) {
          const int32_t e[8] = {x, std::min(x + bw - 1, g.x1), j, std::min(j + bh - 1, g.j1),
                                k, bw, std::min(k + gkc - 1, g.k1), g.face};
          gen.insert(gen.end(), e, e + 8);
        }
  }
  if (s->d_tl_plain) dev_free(s, s->d_tl_plain);
  if (s->d_tl_gen) dev_free(s, s->d_tl_gen);
  s->d_tl_plain = s->d_tl_gen = nullptr;
  s->n_tl_plain = (int)(plain.size() / 8);  // the x-face tiles (n_tl_xf) follow in the same list
  plain.insert(plain.end(), xfl.begin(), xfl.end());
  s->n_tl_gen = (int)(gen.size() / 8);
  if (!plain.empty()) s->d_tl_plain = (int32_t*)upload_bytes(s, plain.data(), plain.size() * sizeof(int32_t));
  if (s->n_tl_gen) s->d_tl_gen = (int32_t*)upload_bytes(s, gen.data(), gen.size() * sizeof(int32_t));
  if ((!plain.empty() && !s->d_tl_plain) || (s->n_tl_gen && !s->d_tl_gen))
    return set_error(FDTD_ERR_NOMEM, "one-pass tile lists");
  s->n_plain_cells = nplain;
  s->n_xfi_cells = nxfi;
  s->n_xf_cells = nxf;
  s->n_tl_xf = ntxf;
  s->tl_bx = bx;
  s->tl_by = by;
  s->tl_kc = kc;
  return 0;
}

void lorentz_jp_coef(double dt, double ds, const fdtd_lorentz_jp& p, LjpCoef<double>& c) { ljp_coef_dt(dt, ds, p, c); }

void tfsf_line_profiles(double dt, double ds, int n, int npml, double eps, double order, double R0, double* out) {
  tfsf_profiles_dt(dt, ds, n, npml, eps, order, R0, out);
}

int graphene3d_coef(double dt, double ds, const fdtd_graphene& p, AdeCoef<double>& c, double* eps_inf) {
  return graphene_coef_dt(dt, ds, p, c, eps_inf);
}

void pml_grade(double dt, double ds, double order, double sigmam, int P, double ep, int n, double* f1, double* f2,
               double* g1, double* g2) {
  Prof{dt, ds, order, sigmam, P}.grade(ep, n, f1, f2, g1, g2);
}

// Upper bound of the device bytes commit_impl allocates (admission control).
size_t estimate_bytes(const Sim* s) {
  const size_t w = s->prec == FDTD_FP32 ? 4 : 8;
  const int64_t le = 128 / (int64_t)w;
  const int64_t px = ((int64_t)s->nx + 2 + le - 1) / le * le;
  const int64_t ncell = px * (s->ny + 2) * (s->nz + 2);
  int eng = s->g.engine;
  if (eng == FDTD_ENGINE_AUTO) eng = s->pml_mode == FDTD_PML_REF_Y ? FDTD_ENGINE_STAGED : FDTD_ENGINE_FUSED;
  bool ljp = false;
  for (const auto& m : s->mats) ljp |= m.type == 3;
  const char* env = std::getenv("FDTD_ONEPASS");
  const bool two = eng == FDTD_ENGINE_FUSED && !ljp && env && std::atoi(env) == 1;  // one-pass copies
  size_t b = (size_t)(eng == FDTD_ENGINE_STAGED ? 9 : (two ? 12 : 6)) * (size_t)(ncell + le) * w;
  b += (size_t)ncell + (size_t)(s->ny + 2) * (s->nz + 2) * 8;  // ids, row ids, mixed rows
  const int n[3] = {s->nx, s->ny, s->nz};
  if (s->pml_mode == FDTD_PML_ALL_FACES) {
    const int Px = s->pml.npml[0], Py = s->pml.npml[1], Pz = s->pml.npml[2];
    const int64_t np = (int64_t)2 * Pz * s->nx * s->ny + (int64_t)(s->nz - 2 * Pz) * 2 * Py * s->nx +
                       (int64_t)(s->nz - 2 * Pz) * (s->ny - 2 * Py) * 2 * Px;
    b += (size_t)(two ? 2 : 1) * S_N * (size_t)std::max<int64_t>(np, 1) * w + (size_t)4 * (n[0] + n[1] + n[2] + 6) * w;
  } else if (s->pml_mode == FDTD_PML_REF_Y) {
    b += (size_t)2 * S_N * (size_t)(s->nx + 2) * (s->pml.npml[1] + 2) * (s->nz + 2) * w;
  }
  for (const auto& m : s->mats) {
    if (m.type == 0) continue;
    int lo[3], hi[3];
    if (!region_bounds(s, m.reg, lo, hi)) continue;
    int64_t vol = 1;
    for (int a = 0; a < 3; ++a) vol *= std::max(0, hi[a] - lo[a] + 1);
    const int words = m.type == 3   ? 2 * m.lj.npols + 2
                      : m.type == 4 ? 1 + 2 * m.lj.npols
                                    : 1 + 2 + 2 * (m.type == 1 ? 2 : 4) + 2;
    b += (size_t)vol * (3 * words * 8 + 5);  // dispersive state is FP64 in both precisions
  }
  for (const auto& d : s->dfts) {
    int ns = 0;
    for (int c = 0; c < 6; ++c) ns += (d.mask >> c) & 1;
    const size_t np = (size_t)d.n1 * d.n2;
    b += 16 * (size_t)ns * d.nfreq * np + (size_t)kDftBlock * ns * np * w + 16 * (size_t)kDftBlock * d.nfreq;
  }
  b += (size_t)7 * 8 * 1024 * s->probes.size();
  for (const auto& t : s->tfsfs) b += (size_t)8 * 8 * (t.n + 2);
  return b + ((size_t)1 << 20);
}

}  // namespace fdtd

using namespace fdtd;

#define GUARD(s)                                                  \
  do {                                                            \
    if (!(s)) return set_error(FDTD_ERR_ARG, "null handle");      \
  } while (0)
#define SETUP_ONLY(s)                                                             \
  do {                                                                            \
    if ((s)->committed) return set_error(FDTD_ERR_STATE, "set-up after commit"); \
  } while (0)

extern "C" {


const char* fdtd_last_error(void) { return g_err.c_str(); }

void fdtd_drude2cp_gold(fdtd_drude2cp* p) {
  std::memset(p, 0, sizeof *p);
  p->eps_inf = 1.1431;  // mod3/gold_mod.f90:28-34
  p->gamma = 1.0805e14;
  p->omega_d = 1.3202e16;
  p->A[0] = 0.26698;
  p->A[1] = 3.0834;
  p->phi[0] = -1.2371;
  p->phi[1] = -1.0968;
  p->Omega[0] = 3.8711e15;
  p->Omega[1] = 4.1684e15;
  p->Gamma[0] = 4.4642e14;
  p->Gamma[1] = 2.4555e15;
}

void fdtd_graphene_ref(fdtd_graphene* p) {
  static const double ap[4] = {13.1068, 1.8417, 0.4191, 0.7737};  // mod3/graphene_mod.f90:33-41
  static const double phip[4] = {-1.8391, 0.3505, 3.0565, -0.1199};
  static const double omp[4] = {6.1264e14, 6.6843e15, 5.6289e15, 1.118e16};
  static const double rp[4] = {1e14, 7.905e14, 6.7526e14, 1.0005e14};
  std::memset(p, 0, sizeof *p);
  p->model = FDTD_GRAPHENE_CP4;
  p->normal_axis = 1;
  p->ref_init_order = 1;
  p->thickness = 3.4e-10;
  p->eps_inf = 1.0;
  p->gamma = 9e15;
  p->omega_d = 5.2796e15;
  for (int q = 0; q < 4; ++q) {
    p->A[q] = ap[q];
    p->phi[q] = phip[q];
    p->Omega[q] = omp[q];
    p->Gamma[q] = rp[q];
  }
}

int fdtd_create(fdtd_sim** out, const fdtd_grid* g) {
  if (!out || !g) return set_error(FDTD_ERR_ARG, "null argument");
  if (g->nx < 1 || g->ny < 1 || g->nz < 1 || !(g->ds > 0) || !(g->courant > 0))
    return set_error(FDTD_ERR_ARG, "bad grid");
  if (g->precision != FDTD_FP64 && g->precision != FDTD_FP32) return set_error(FDTD_ERR_ARG, "bad precision");
  int dev = 0;
  FDTD_HIP(hipGetDevice(&dev));
  fdtd_sim* s = new fdtd_sim();
  s->g = *g;
  s->nx = g->nx;
  s->ny = g->ny;
  s->nz = g->nz;
  s->ds = g->ds;
  s->ku = g->courant;
  s->dt = s->ku * s->ds / consts().c0;      // sim3/program.f90:50
  s->ku_pml = s->dt * consts().c0 / s->ds;  // mod3/BERpmly_mod.f90:506
  s->prec = g->precision;
  s->wsize = g->precision == FDTD_FP32 ? 4 : 8;
  s->device = dev;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return hip_check(e, "hipStreamCreate");
  }
  *out = s;
  return 0;
}

int fdtd_destroy(fdtd_sim* s) {
  if (!s) return 0;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& pe : s->pending) {
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  for (auto ev : s->event_pool) (void)hipEventDestroy(ev);
  for (auto& b : s->allocs) (void)hipFreeAsync(b.p, s->stream);
  if (s->stream) {
    (void)hipStreamSynchronize(s->stream);
    (void)hipStreamDestroy(s->stream);
  }
  delete s;
  return 0;
}

int fdtd_set_pml(fdtd_sim* s, const fdtd_pml* p) {
  GUARD(s);
  SETUP_ONLY(s);
  if (!p || p->mode < FDTD_PML_NONE || p->mode > FDTD_PML_ALL_FACES) return set_error(FDTD_ERR_ARG, "bad pml");
  s->pml = *p;
  s->pml_mode = p->mode;
  return 0;
}

static int add_material(fdtd_sim* s, const fdtd_region* r, HostMaterial m) {
  int lo[3], hi[3];
  if (!r || !region_bounds(s, *r, lo, hi)) return set_error(FDTD_ERR_ARG, "bad region");
  m.reg = *r;
  s->mats.push_back(m);
  return 0;
}

int fdtd_add_dielectric(fdtd_sim* s, const fdtd_region* r, double eps_r) {
  GUARD(s);
  SETUP_ONLY(s);
  HostMaterial m{};
  m.type = 0;
  m.eps_r = eps_r;
  m.writes[0] = m.writes[1] = m.writes[2] = 1;
  return add_material(s, r, m);
}

int fdtd_add_drude2cp(fdtd_sim* s, const fdtd_region* r, const fdtd_drude2cp* p) {
  GUARD(s);
  SETUP_ONLY(s);
  if (!p || !(p->gamma > 0)) return set_error(FDTD_ERR_ARG, "bad Drude+2CP parameters");
  HostMaterial m{};
  m.type = 1;
  m.d2 = *p;
  m.writes[0] = m.writes[1] = m.writes[2] = 1;
  return add_material(s, r, m);
}

int fdtd_add_graphene(fdtd_sim* s, const fdtd_region* r, const fdtd_graphene* p) {
  GUARD(s);
  SETUP_ONLY(s);
  if (!p || p->normal_axis < 0 || p->normal_axis > 2) return set_error(FDTD_ERR_ARG, "bad graphene parameters");
  if (p->model == FDTD_GRAPHENE_CP4 && !(p->gamma > 0)) return set_error(FDTD_ERR_ARG, "bad graphene gamma");
  if (p->model != FDTD_GRAPHENE_CP4 && p->model != FDTD_GRAPHENE_KUBO)
    return set_error(FDTD_ERR_ARG, "bad graphene model");
  if (p->model == FDTD_GRAPHENE_KUBO && (!(p->tau > 0) || !(p->temperature > 0)))
    return set_error(FDTD_ERR_ARG, "tau and T must be > 0");
  HostMaterial m{};
  m.type = 2;
  m.gr = *p;
  m.writes[0] = m.writes[1] = m.writes[2] = 1;
  m.writes[p->normal_axis] = 0;  // mod3/graphene_mod.f90:59-105 leaves Ey alone
  return add_material(s, r, m);
}

int fdtd_add_lorentz_jp(fdtd_sim* s, const fdtd_region* r, const fdtd_lorentz_jp* p) {
  GUARD(s);
  SETUP_ONLY(s);
  if (!p || p->npols < 1 || p->npols > 4 || p->jp_mask < 1 || p->jp_mask > 7 || !(p->eps_inf > 0))
    return set_error(FDTD_ERR_ARG, "bad Lorentz-Jp parameters");
  for (int l = 0; l < p->npols; ++l)
    if (1.0 + p->dnu[l] * s->dt == 0.0) return set_error(FDTD_ERR_ARG, "pole with dnu*dt = -1");
  if (p->ez_uses_jx && (!(p->jp_mask & 1) || !(p->jp_mask & 4)))
    return set_error(FDTD_ERR_ARG, "ez_uses_jx needs Ex and Ez in jp_mask");
  HostMaterial m{};
  m.type = 3;
  m.lj = *p;
  m.writes[0] = m.writes[1] = m.writes[2] = 1;
  return add_material(s, r, m);
}

int fdtd_add_lorentz_d(fdtd_sim* s, const fdtd_region* r, const fdtd_lorentz_jp* p) {
  GUARD(s);
  SETUP_ONLY(s);
  if (!p || p->npols < 1 || p->npols > 4 || p->jp_mask != 7 || p->nonlinear || p->ez_uses_jx ||
      p->eps_inf != 1.0)
    return set_error(FDTD_ERR_ARG, "bad D-form Lorentz parameters (eps_inf 1, all components, linear)");
  for (int l = 0; l < p->npols; ++l)
    if (1.0 + p->dnu[l] * s->dt == 0.0) return set_error(FDTD_ERR_ARG, "pole with dnu*dt = -1");
  HostMaterial m{};
  m.type = 4;
  m.lj = *p;
  m.writes[0] = m.writes[1] = m.writes[2] = 1;
  return add_material(s, r, m);
}

int fdtd_lorentz_jp_preset(int32_t kind, fdtd_lorentz_jp* p) {
  if (!p) return set_error(FDTD_ERR_ARG, "null preset");
  struct Tab {
    int npols, mask, nl, ezjx;
    double einf, xi[4], w0[4], dnu[4];
  };
  // pole tables in units of 2*pi*c0/micro for w0 and dnu, as the modules
  // declare them (file:line in include/fdtd.h, FDTD_LJP_*)
  static const Tab tabs[7] = {
      {3, 7, 0, 0, 1.053787, {2.862890, -0.0722398, 367.7668, 0},
       {-2.445679, 1.480982, 0.3371279, 0}, {0.2170955, 0.1429645, 0.02285188, 0}},
      {3, 7, 0, 0, 1.0, {8.0, 2.85, -0.107, 0}, {3.64, 2.76, 1.73, 0}, {0.0, 0.063, 2.5, 0}},
      {4, 7, 0, 0, 1.1, {-12.74873, 0.9605442, 67.66233, 55.0525},
       {4.36014, -1.261863, 0.4063879, -2.176176}, {-5.825585, 0.2087635, 1.246245, 8.059584}},
      {2, 7, 0, 0, 3.497071, {2.301113, 400.0, 0, 0}, {-2.393546, 0.3441309, 0, 0},
       {0.2460773, 0.02533959, 0, 0}},
      {4, 5, 1, 1, 1.070682, {-0.163673, 0.843432, -0.170024, -0.121134},
       {1.642784, 0.807691, 0.934364, 0.525964}, {2.267351, 1.018179, 0.621508, 0.033048}},
      {3, 5, 1, 0, 1.023242, {4.957366, 4.455092, 1.905429, 0}, {0.573075, 1.713164, 4.598546, 0},
       {0.368926, 2.42099, 1.35581, 0}},
      {3, 5, 1, 0, 1.023242, {0.0, 0.0, 0.0, 0}, {0.573075, 1.713164, 4.598546, 0},
       {0.368926, 2.42099, 1.35581, 0}},
  };
  if (kind < 0 || kind > 6) return set_error(FDTD_ERR_ARG, "unknown Lorentz-Jp preset");
  const Consts& K = consts();
  const double unit = 2.0 * K.pi * K.c0 / 1e-6;  // 2*pi*c0/micro
  const Tab& t = tabs[kind];
  std::memset(p, 0, sizeof *p);
  p->npols = t.npols;
  p->jp_mask = t.mask;
  p->nonlinear = t.nl;
  p->ez_uses_jx = t.ezjx;
  p->eps_inf = t.einf;
  for (int l = 0; l < t.npols; ++l) {
    p->xi[l] = t.xi[l];
    p->w0[l] = unit * t.w0[l];
    p->dnu[l] = unit * t.dnu[l];
  }
  return 0;
}

int fdtd_add_plane_source(fdtd_sim* s, int axis, int index, int comp, const fdtd_pulse* p) {
  GUARD(s);
  SETUP_ONLY(s);
  const int n[3] = {s->nx, s->ny, s->nz};
  if (!p || axis < 0 || axis > 2 || comp < 0 || comp > 2 || index < 1 || index > n[axis])
    return set_error(FDTD_ERR_ARG, "bad source plane");
  if ((int)s->srcs.size() >= kMaxSrc) return set_error(FDTD_ERR_ARG, "too many sources");
  if (!(p->st != 0.0)) return set_error(FDTD_ERR_ARG, "pulse width must be non-zero");
  s->srcs.push_back({axis, index, comp, *p});
  return 0;
}

int fdtd_add_dft_plane(fdtd_sim* s, int axis, int index, int nfreq, const double* omega, int comp_mask, int n_begin,
                       int n_end, int* id) {
  GUARD(s);
  SETUP_ONLY(s);
  const int n[3] = {s->nx, s->ny, s->nz};
  if (!omega || !id || axis < 0 || axis > 2 || index < 1 || index > n[axis] || nfreq < 1 || nfreq > 4096 ||
      (comp_mask & ~63) || !comp_mask)
    return set_error(FDTD_ERR_ARG, "bad dft plane");
  DftPlane d;
  d.axis = axis;
  d.index = index;
  d.nfreq = nfreq;
  d.mask = comp_mask;
  d.n_begin = n_begin;
  d.n_end = n_end;
  d.n1 = n[axis == 0 ? 1 : 0];
  d.n2 = n[axis == 2 ? 1 : 2];
  d.omega.assign(omega, omega + nfreq);
  s->dfts.push_back(d);
  *id = (int)s->dfts.size() - 1;
  return 0;
}

int fdtd_add_probe(fdtd_sim* s, int i, int j, int k, int* id) {
  GUARD(s);
  SETUP_ONLY(s);
  if (!id || i < 1 || i > s->nx || j < 1 || j > s->ny || k < 1 || k > s->nz)
    return set_error(FDTD_ERR_ARG, "bad probe");
  if (s->probes.size() >= 64) return set_error(FDTD_ERR_ARG, "too many probes");
  Probe p;
  p.i = i;
  p.j = j;
  p.k = k;
  s->probes.push_back(p);
  *id = (int)s->probes.size() - 1;
  return 0;
}

int fdtd_add_tfsf(fdtd_sim* s, const fdtd_tfsf* p, int* id) {
  GUARD(s);
  SETUP_ONLY(s);
  if (!p || !id) return set_error(FDTD_ERR_ARG, "null argument");
  const int n[3] = {s->nx, s->ny, s->nz};
  if (p->axis < 0 || p->axis > 2 || p->comp < 0 || p->comp > 2 || p->comp == p->axis)
    return set_error(FDTD_ERR_ARG, "bad TFSF axis/component");
  const int N = n[p->axis];
  if (p->lo < 2 || p->lo > N || (p->hi != 0 && (p->hi < p->lo || p->hi >= N)) || p->npml < 0 ||
      2 * p->npml + 2 >= N || p->src_index <= p->npml || p->src_index > N - p->npml || !(p->eps > 0) ||
      (p->npml > 0 && !(p->R0 > 0 && p->R0 < 1)) || !(p->pulse.st != 0.0))
    return set_error(FDTD_ERR_ARG, "bad TFSF planes or line");
  if (s->tfsfs.size() >= 4) return set_error(FDTD_ERR_ARG, "too many TFSF waves");
  Tfsf t;
  t.p = *p;
  t.n = N;
  t.hc = 3 - p->axis - p->comp;
  // orientation: +1 when (comp, axis) is (x,y), (y,z) or (z,x)
  t.sg = ((p->comp + 1) % 3 == p->axis) ? 1 : -1;
  s->tfsfs.push_back(t);
  *id = (int)s->tfsfs.size() - 1;
  return 0;
}

int fdtd_add_tfsf_probe(fdtd_sim* s, int tid, int index, int* id) {
  GUARD(s);
  SETUP_ONLY(s);
  if (!id || tid < 0 || tid >= (int)s->tfsfs.size() || index < 0 || index > s->tfsfs[tid].n + 1)
    return set_error(FDTD_ERR_ARG, "bad TFSF probe");
  if (s->probes.size() >= 64) return set_error(FDTD_ERR_ARG, "too many probes");
  int nl = 0;
  for (const auto& q : s->probes) nl += q.line == tid;
  if (nl >= 4) return set_error(FDTD_ERR_ARG, "at most 4 probes per TFSF line");
  Probe p;
  p.i = 0;
  p.j = index;
  p.k = 0;
  p.line = tid;
  s->probes.push_back(p);
  *id = (int)s->probes.size() - 1;
  return 0;
}

int fdtd_get_tfsf_line(fdtd_sim* s, int tid, double* e, double* h) {
  GUARD(s);
  if (!e || !h || tid < 0 || tid >= (int)s->tfsfs.size()) return set_error(FDTD_ERR_ARG, "bad argument");
  int rc = ensure_committed(s);
  if (rc) return rc;
  const Tfsf& t = s->tfsfs[tid];
  const size_t n = (size_t)t.n + 2;
  FDTD_HIP(hipMemcpyAsync(e, t.d_line, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  FDTD_HIP(hipMemcpyAsync(h, t.d_line + n, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  FDTD_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

int fdtd_commit(fdtd_sim* s) {
  GUARD(s);
  return ensure_committed(s);
}

// second field and split copies of the one-pass step (zero-filled)
static int alloc_onepass_copies(Sim* s) {
  const int64_t le = 128 / (int64_t)s->wsize;
  const size_t fbytes = (size_t)(s->ncell_alloc + le) * s->wsize;
  const size_t shift = (size_t)(le - 1) * s->wsize;
  for (int q = 0; q < 3; ++q) {
    char* e2 = (char*)dev_alloc(s, fbytes);
    char* h2 = (char*)dev_alloc(s, fbytes);
    if (!e2 || !h2) return set_error(FDTD_ERR_NOMEM, "one-pass field copies");
    s->E2[q] = e2 + shift;
    s->H2[q] = h2 + shift;
  }
  if (s->d_split) {
    s->d_split2 = dev_alloc(s, (size_t)S_N * (size_t)std::max<int64_t>(s->npml_cells, 1) * s->wsize);
    if (!s->d_split2) return set_error(FDTD_ERR_NOMEM, "one-pass split copies");
  }
  return 0;
}

static void free_onepass_copies(Sim* s) {
  const size_t shift = (size_t)(128 / (int64_t)s->wsize - 1) * s->wsize;
  for (int q = 0; q < 3; ++q) {
    if (s->E2[q]) dev_free(s, (char*)s->E2[q] - shift);
    if (s->H2[q]) dev_free(s, (char*)s->H2[q] - shift);
    s->E2[q] = s->H2[q] = nullptr;
  }
  if (s->d_split2) dev_free(s, s->d_split2);
  s->d_split2 = nullptr;
}

int fdtd_autotune(fdtd_sim* s, int steps, int* out_by, int* out_kc) {
  GUARD(s);
  if (steps < 1) return set_error(FDTD_ERR_ARG, "autotune needs at least one timed step");
  int rc = ensure_committed(s);
  if (rc) return rc;
  if (s->n_done != 0 || s->fields_set)
    return set_error(FDTD_ERR_STATE, "autotune runs before the first step, on untouched fields");
  FDTD_HIP(hipSetDevice(s->device));
  // candidates: the two-kernel step (by = 0) and one-pass tiles of 64 x by
  // cells marching kc planes
  std::vector<std::pair<int, int>> cand;
  bool ljp = false;
  for (const auto& m : s->mats) ljp |= m.type == 3;
  bool added = false;  // second copies allocated here for the trial
  if (s->engine == FDTD_ENGINE_FUSED && !ljp && !s->E2[0]) {
    rc = alloc_onepass_copies(s);
    if (rc) return rc;
    added = true;
  }
  if (s->engine == FDTD_ENGINE_FUSED && s->E2[0]) {
    for (int zm : {1, 2, 4, 8}) cand.push_back({0, zm});  // two-kernel step, zm planes per tile
    for (int by : {2, 4, 8})
      for (int kc : {16, 32, 64, 128}) {
        if (by == 8 && s->prec != FDTD_FP32) continue;  // FP64 tiles stop at 4 rows
        if (kc > s->nz && kc > 16) continue;
        cand.push_back({by, kc});
      }
  }
  if (cand.empty()) {
    if (out_by) *out_by = 0;
    if (out_kc) *out_kc = 0;
    return 0;
  }
  hipEvent_t e0, e1;
  FDTD_HIP(hipEventCreate(&e0));
  FDTD_HIP(hipEventCreate(&e1));
  const bool prof = s->profiling;
  s->profiling = false;
  s->dry = true;
  float best = -1.f;
  std::pair<int, int> pick = cand[0];
  for (const auto& c : cand) {
    s->onepass = c.first > 0;
    if (c.first > 0) {
      s->op_by = c.first;
      s->op_kc = c.second;
    } else {
      s->zm = c.second;
    }
    for (int w = 0; w < 2 && rc == 0; ++w) rc = launch_step(s);
    if (rc == 0 && hipEventRecord(e0, s->stream) != hipSuccess) rc = set_error(FDTD_ERR_HIP, "event");
    for (int it = 0; it < steps && rc == 0; ++it) rc = launch_step(s);
    if (rc == 0 && hipEventRecord(e1, s->stream) != hipSuccess) rc = set_error(FDTD_ERR_HIP, "event");
    float ms = 0.f;
    if (rc == 0 && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = set_error(FDTD_ERR_HIP, "event timing");
    if (rc) break;
    if (best < 0.f || ms < best) {
      best = ms;
      pick = c;
    }
  }
  s->dry = false;
  s->profiling = prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  s->onepass = pick.first > 0;
  if (pick.first > 0) {
    s->op_by = pick.first;
    s->op_kc = pick.second;
  } else {
    s->zm = pick.second;
    if (added) free_onepass_copies(s);
  }
  if (out_by) *out_by = pick.first;
  if (out_kc) *out_kc = pick.second;
  return 0;
}

int fdtd_step(fdtd_sim* s, int nsteps) {
  GUARD(s);
  if (nsteps < 0) return set_error(FDTD_ERR_ARG, "negative step count");
  int rc = ensure_committed(s);
  if (rc) return rc;
  FDTD_HIP(hipSetDevice(s->device));
  const int np = (int)s->probes.size();
  if (np && s->probe_pending + nsteps > s->probe_cap) {
    // grow the device record buffer; records already written move along on
    // the handle's stream (the old buffer is released with the handle)
    const int cap = std::max(std::max(2 * s->probe_cap, s->probe_pending + nsteps), 1024);
    double* nb = (double*)dev_alloc(s, sizeof(double) * 7 * (size_t)np * cap);
    if (!nb) return set_error(FDTD_ERR_NOMEM, "probe buffer");
    if (s->probe_pending)
      FDTD_HIP(hipMemcpyAsync(nb, s->d_probe, sizeof(double) * 7 * (size_t)np * s->probe_pending,
                              hipMemcpyDeviceToDevice, s->stream));
    s->d_probe = nb;
    s->probe_cap = cap;
  }
  s->call_base = s->n_done;
  s->probe_row0 = s->probe_pending;
  for (int it = 0; it < nsteps; ++it) {
    rc = launch_step(s);
    if (rc) return rc;
    s->n_done += 1;
  }
  if (np) s->probe_pending += nsteps;
  // DFT accumulators are complete once this call's work has run; the call
  // itself only queues work on the handle's stream
  return launch_dft_flush(s);
}

namespace fdtd {
// Move probe records written since the last harvest to the host lists.
static int harvest_probes(Sim* s) {
  const int np = (int)s->probes.size();
  if (!np || !s->probe_pending) return 0;
  std::vector<double> h(7 * (size_t)np * s->probe_pending);
  FDTD_HIP(hipMemcpyAsync(h.data(), s->d_probe, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  FDTD_HIP(hipStreamSynchronize(s->stream));
  for (int it = 0; it < s->probe_pending; ++it)
    for (int p = 0; p < np; ++p) {
      const double* r = h.data() + 7 * ((size_t)it * np + p);
      s->probes[p].rec.insert(s->probes[p].rec.end(), r, r + 7);
    }
  s->probe_pending = 0;
  return 0;
}
}  // namespace fdtd

int fdtd_sync(fdtd_sim* s) {
  GUARD(s);
  if (!s->stream) return 0;
  FDTD_HIP(hipStreamSynchronize(s->stream));
  int rc = harvest_probes(s);
  if (rc) return rc;
  return flush_profile(s);
}

namespace fdtd {
// Checkpoint layout: header, then sections, each an int64 byte count and
// its bytes, in this order: E0..E2, H0..H2 (ghosted arrays), [D0..D2
// (STAGED)], PML splits, dispersive group states, DFT accumulators, probe
// records, TFSF lines.
struct CkHeader {
  char magic[8];
  int32_t version, nx, ny, nz, prec, engine, pml_mode, n_done;
  int32_t ngroups, ndft, nprobe, ntfsf;
};
static const char kCkMagic[8] = {'F', 'D', 'T', 'D', 'C', 'K', 'P', '1'};

struct CkSection {
  void* dev;                 // device bytes, or
  std::vector<double>* host;  // a host list (probe records)
  size_t bytes;
};

static std::vector<CkSection> ck_sections(Sim* s) {
  std::vector<CkSection> v;
  const size_t fb = (size_t)s->ncell_alloc * s->wsize;
  for (int q = 0; q < 3; ++q) v.push_back({s->E[q], nullptr, fb});
  for (int q = 0; q < 3; ++q) v.push_back({s->H[q], nullptr, fb});
  if (s->engine == FDTD_ENGINE_STAGED)
    for (int q = 0; q < 3; ++q) v.push_back({s->D[q], nullptr, fb});
  if (s->d_split) v.push_back({s->d_split, nullptr, (size_t)S_N * (size_t)std::max<int64_t>(s->npml_cells, 1) * s->wsize});
  if (s->d_slab) v.push_back({s->d_slab, nullptr, (size_t)2 * S_N * (size_t)s->slab_cells * s->wsize});
  for (auto& g : s->groups)
    if (g.ncells) v.push_back({g.d_state, nullptr, (size_t)g.nstate * 3 * (size_t)g.ncells * 8});
  for (auto& d : s->dfts) v.push_back({d.d_acc, nullptr, sizeof(double) * 2 * (size_t)d.nslots * d.nfreq * (size_t)d.n1 * d.n2});
  for (auto& p : s->probes) v.push_back({nullptr, &p.rec, p.rec.size() * sizeof(double)});
  for (auto& t : s->tfsfs) v.push_back({t.d_line, nullptr, (size_t)8 * (t.n + 2) * sizeof(double)});
  return v;
}

static CkHeader ck_header(const Sim* s) {
  CkHeader h{};
  std::memcpy(h.magic, kCkMagic, 8);
  h.version = 1;
  h.nx = s->nx;
  h.ny = s->ny;
  h.nz = s->nz;
  h.prec = s->prec;
  h.engine = s->engine;
  h.pml_mode = s->pml_mode;
  h.n_done = s->n_done;
  h.ngroups = (int32_t)s->groups.size();
  h.ndft = (int32_t)s->dfts.size();
  h.nprobe = (int32_t)s->probes.size();
  h.ntfsf = (int32_t)s->tfsfs.size();
  return h;
}
}  // namespace fdtd

int fdtd_checkpoint_size(fdtd_sim* s, size_t* bytes) {
  GUARD(s);
  if (!bytes) return set_error(FDTD_ERR_ARG, "null argument");
  if (!s->committed) return set_error(FDTD_ERR_STATE, "checkpoint of a handle that is not committed");
  FDTD_HIP(hipSetDevice(s->device));
  FDTD_HIP(hipStreamSynchronize(s->stream));
  int rc = harvest_probes(s);
  if (rc) return rc;
  size_t n = sizeof(CkHeader);
  for (const auto& c : ck_sections(s)) n += sizeof(int64_t) + c.bytes;
  *bytes = n;
  return 0;
}

int fdtd_checkpoint_save(fdtd_sim* s, void* buf, size_t bytes) {
  GUARD(s);
  size_t need = 0;
  int rc = fdtd_checkpoint_size(s, &need);  // also waits and harvests the probes
  if (rc) return rc;
  if (!buf || bytes != need) return set_error(FDTD_ERR_ARG, "checkpoint buffer of the wrong size");
  for (const auto& d : s->dfts)
    if (d.nbuf) return set_error(FDTD_ERR_STATE, "DFT samples not folded");  // fdtd_step folds at its end
  char* o = (char*)buf;
  const CkHeader h = ck_header(s);
  std::memcpy(o, &h, sizeof h);
  o += sizeof h;
  for (const auto& c : ck_sections(s)) {
    const int64_t nb = (int64_t)c.bytes;
    std::memcpy(o, &nb, sizeof nb);
    o += sizeof nb;
    if (c.host) {
      if (c.bytes) std::memcpy(o, c.host->data(), c.bytes);
    } else if (c.bytes) {
      FDTD_HIP(hipMemcpyAsync(o, c.dev, c.bytes, hipMemcpyDeviceToHost, s->stream));
    }
    o += c.bytes;
  }
  FDTD_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

int fdtd_checkpoint_load(fdtd_sim* s, const void* buf, size_t bytes) {
  GUARD(s);
  if (!buf || bytes < sizeof(CkHeader)) return set_error(FDTD_ERR_ARG, "bad checkpoint buffer");
  int rc = ensure_committed(s);
  if (rc) return rc;
  if (s->n_done != 0) return set_error(FDTD_ERR_STATE, "checkpoint load after the first step");
  CkHeader h;
  std::memcpy(&h, buf, sizeof h);
  CkHeader mine = ck_header(s);
  if (std::memcmp(h.magic, kCkMagic, 8) != 0 || h.version != 1) return set_error(FDTD_ERR_ARG, "not a checkpoint");
  if (h.nx != mine.nx || h.ny != mine.ny || h.nz != mine.nz || h.prec != mine.prec || h.engine != mine.engine ||
      h.pml_mode != mine.pml_mode || h.ngroups != mine.ngroups || h.ndft != mine.ndft ||
      h.nprobe != mine.nprobe || h.ntfsf != mine.ntfsf || h.n_done < 0)
    return set_error(FDTD_ERR_ARG, "checkpoint of a handle set up differently");
  FDTD_HIP(hipSetDevice(s->device));
  // probe sections carry the recorded rows: their sizes come from the buffer
  const char* in = (const char*)buf + sizeof h;
  const char* end = (const char*)buf + bytes;
  std::vector<CkSection> secs = ck_sections(s);
  std::vector<const char*> src(secs.size());
  std::vector<int64_t> len(secs.size());
  for (size_t q = 0; q < secs.size(); ++q) {
    if (end - in < (ptrdiff_t)sizeof(int64_t)) return set_error(FDTD_ERR_ARG, "truncated checkpoint");
    std::memcpy(&len[q], in, sizeof(int64_t));
    in += sizeof(int64_t);
    const bool probe = secs[q].host != nullptr;
    if (len[q] < 0 || end - in < len[q] || (!probe && (size_t)len[q] != secs[q].bytes) ||
        (probe && (size_t)len[q] != (size_t)h.n_done * 7 * sizeof(double)))
      return set_error(FDTD_ERR_ARG, "checkpoint of a handle set up differently");
    src[q] = in;
    in += len[q];
  }
  if (in != end) return set_error(FDTD_ERR_ARG, "checkpoint of a handle set up differently");
  FDTD_HIP(hipStreamSynchronize(s->stream));
  for (size_t q = 0; q < secs.size(); ++q) {
    if (secs[q].host) {
      secs[q].host->assign((const double*)src[q], (const double*)(src[q] + len[q]));
    } else if (len[q]) {
      FDTD_HIP(hipMemcpyAsync(secs[q].dev, src[q], (size_t)len[q], hipMemcpyHostToDevice, s->stream));
    }
  }
  FDTD_HIP(hipStreamSynchronize(s->stream));
  for (auto& d : s->dfts) d.nbuf = 0;
  s->probe_pending = 0;
  s->n_done = h.n_done;
  s->fields_set = true;  // the autotuner's dry steps need zero fields
  return 0;
}

int fdtd_get_step(fdtd_sim* s, int* n) {
  GUARD(s);
  if (!n) return set_error(FDTD_ERR_ARG, "null argument");
  *n = s->n_done;
  return 0;
}

int fdtd_get_dt(fdtd_sim* s, double* dt) {
  GUARD(s);
  if (!dt) return set_error(FDTD_ERR_ARG, "null argument");
  *dt = s->dt;
  return 0;
}

int fdtd_get_field(fdtd_sim* s, int comp, double* dst) {
  GUARD(s);
  if (!dst || comp < 0 || comp > 8) return set_error(FDTD_ERR_ARG, "bad component");
  int rc = ensure_committed(s);
  if (rc) return rc;
  void* f = field_ptr(s, comp);
  if (!f) return set_error(FDTD_ERR_UNSUPPORTED, "D is kept only by the STAGED engine");
  return s->prec == FDTD_FP32 ? read_field<float>(s, f, dst) : read_field<double>(s, f, dst);
}

int fdtd_set_field(fdtd_sim* s, int comp, const double* src) {
  GUARD(s);
  if (!src || comp < 0 || comp > 8) return set_error(FDTD_ERR_ARG, "bad component");
  int rc = ensure_committed(s);
  if (rc) return rc;
  void* f = field_ptr(s, comp);
  if (!f) return set_error(FDTD_ERR_UNSUPPORTED, "D is kept only by the STAGED engine");
  rc = s->prec == FDTD_FP32 ? write_field<float>(s, f, src) : write_field<double>(s, f, src);
  if (rc) return rc;
  s->fields_set = true;
  rc = launch_set_ghosts(s, comp);
  if (rc) return rc;
  FDTD_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

int fdtd_get_dft(fdtd_sim* s, int id, int comp, double* re, double* im) {
  GUARD(s);
  if (id < 0 || id >= (int)s->dfts.size() || comp < 0 || comp > 5) return set_error(FDTD_ERR_ARG, "bad argument");
  int rc = ensure_committed(s);
  if (rc) return rc;
  DftPlane& d = s->dfts[id];
  if (d.slot[comp] < 0) return set_error(FDTD_ERR_ARG, "component not monitored");
  const size_t np = (size_t)d.n1 * d.n2 * d.nfreq;
  const double* base = d.d_acc + (size_t)d.slot[comp] * 2 * np;
  if (re) FDTD_HIP(hipMemcpyAsync(re, base, np * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  if (im) FDTD_HIP(hipMemcpyAsync(im, base + np, np * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  FDTD_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

int fdtd_get_flux(fdtd_sim* s, int id, double* S) {
  GUARD(s);
  if (!S || id < 0 || id >= (int)s->dfts.size()) return set_error(FDTD_ERR_ARG, "bad argument");
  int rc = ensure_committed(s);
  if (rc) return rc;
  return launch_flux(s, id, S);
}

int fdtd_get_probe_count(fdtd_sim* s, int id, int* nrec) {
  GUARD(s);
  if (!nrec || id < 0 || id >= (int)s->probes.size()) return set_error(FDTD_ERR_ARG, "bad argument");
  int rc = harvest_probes(s);
  if (rc) return rc;
  *nrec = (int)(s->probes[id].rec.size() / 7);
  return 0;
}

int fdtd_get_probe(fdtd_sim* s, int id, double* out) {
  GUARD(s);
  if (!out || id < 0 || id >= (int)s->probes.size()) return set_error(FDTD_ERR_ARG, "bad argument");
  int rc = harvest_probes(s);
  if (rc) return rc;
  std::memcpy(out, s->probes[id].rec.data(), s->probes[id].rec.size() * sizeof(double));
  return 0;
}

int fdtd_energy(fdtd_sim* s, double* we, double* wh) {
  GUARD(s);
  if (!we || !wh) return set_error(FDTD_ERR_ARG, "null argument");
  int rc = ensure_committed(s);
  if (rc) return rc;
  return launch_energy(s, we, wh);
}

int fdtd_memory_bytes(fdtd_sim* s, size_t* bytes) {
  GUARD(s);
  if (!bytes) return set_error(FDTD_ERR_ARG, "null argument");
  *bytes = s->dev_bytes;
  return 0;
}

int fdtd_set_profiling(fdtd_sim* s, int on) {
  GUARD(s);
  int rc = flush_profile(s);
  if (rc) return rc;
  s->profiling = on != 0;
  return 0;
}

int fdtd_kernel_stats(fdtd_sim* s, int kernel, double* total_ms, int64_t* launches) {
  GUARD(s);
  if (kernel < 0 || kernel >= FDTD_K_COUNT) return set_error(FDTD_ERR_ARG, "bad kernel id");
  int rc = flush_profile(s);
  if (rc) return rc;
  if (total_ms) *total_ms = s->k_ms[kernel];
  if (launches) *launches = s->k_launches[kernel];
  return 0;
}

int fdtd_algorithmic_bytes(fdtd_sim* s, int kernel, double* bytes) {
  GUARD(s);
  if (!bytes || kernel < 0 || kernel >= FDTD_K_COUNT) return set_error(FDTD_ERR_ARG, "bad argument");
  int rc = ensure_committed(s);
  if (rc) return rc;
  const double w = (double)s->wsize;
  const double N = (double)s->nx * s->ny * s->nz;
  double b = 0;
  if (s->engine == FDTD_ENGINE_FUSED) {
    // per non-PML cell: read H(3) + read/write E(6).  Per PML cell: read the
    // other field (3) and write the main field (3); per component, read and
    // write the parts kept (those along the axes whose slabs hold the cell)
    // and read the old component unless both parts are kept: 13 words in a
    // cell of one slab (a face), 16 in two (an edge), 18 in three (a corner).
    double n1 = 0, n2 = 0, n3 = 0;
    if (s->pml_mode == FDTD_PML_ALL_FACES) {
      const double X = 2.0 * s->P[0], Y = 2.0 * s->P[1], Z = 2.0 * s->P[2];
      const double x = s->nx - X, y = s->ny - Y, z = s->nz - Z;
      n3 = X * Y * Z;
      n2 = X * Y * z + X * y * Z + x * Y * Z;
      n1 = X * y * z + x * Y * z + x * y * Z;
    }
    const double pml_words = 13.0 * n1 + 16.0 * n2 + 18.0 * n3;
    if (s->onepass) {
      // plain tiles: a whole step of their (non-PML) cells per launch, 18
      // words per cell (the two half steps' 9 each); general tiles: each
      // half step as above.  The tile lists are those of the last step.
      // x-face tiles: 18 words a step for their non-PML cells, 26 (13 a
      // half step) for their x-slab face cells, which leave the general
      // kernels' share
      const double Np = (double)s->n_plain_cells, Ni = (double)s->n_xfi_cells, Nx = (double)s->n_xf_cells;
      if (kernel == FDTD_K_E) b = w * 18.0 * Np;
      if (kernel == FDTD_K_XF) b = w * (18.0 * Ni + 26.0 * Nx);
      if (kernel == FDTD_K_GEN_E || kernel == FDTD_K_GEN_H)
        b = w * (9.0 * (N - Np - Ni - s->n_pml) + pml_words - 13.0 * Nx);
    } else if (kernel == FDTD_K_E || kernel == FDTD_K_H) {
      b = w * (9.0 * (N - s->n_pml) + pml_words);
    }
    if (kernel == FDTD_K_ADE)
      for (auto& g : s->groups) b += (w * 6.0 + 8.0 * 2.0 * 3.0 * g.nstate) * (double)g.ncells;
  } else {
    const int P = s->pml_mode == FDTD_PML_REF_Y ? s->pml.npml[1] : 0;
    const double Nm = (double)s->nx * (s->ny - 2 * P) * s->nz;
    const double Ns = (double)s->nx * 2 * P * s->nz;
    if (kernel == FDTD_K_E || kernel == FDTD_K_H) b = w * 9.0 * Nm;
    if (kernel == FDTD_K_MAT) b = w * 6.0 * N;
    if (kernel == FDTD_K_PML_E || kernel == FDTD_K_PML_H) b = w * 24.0 * Ns;
    if (kernel == FDTD_K_ADE)
      for (auto& g : s->groups) b += (w * 6.0 + 8.0 * 2.0 * 3.0 * g.nstate) * (double)g.ncells;
  }
  *bytes = b;
  return 0;
}

int fdtd_cell_counts(fdtd_sim* s, int64_t* n_pml, int64_t* n_ade) {
  GUARD(s);
  int rc = ensure_committed(s);
  if (rc) return rc;
  if (n_pml) *n_pml = s->n_pml;
  if (n_ade) *n_ade = s->n_ade;
  return 0;
}

}  // extern "C"
