/*
 * fdtd2d_oracle.c -- FP64 CPU restatement of the reference's 2D TEz path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library's
 * 2D stepper: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it.
 *
 * It exports the ABI of include/fdtd2d.h (single simulations) with the
 * prefix `ofdtd2d_`.  The algorithm follows the reference's 2D driver loop
 * by loop on one domain, in its stage and arithmetic order (D-form):
 *
 *   updateD / updateH ........ sim2/include/myfdtd_2Dfunc.f90:16-44
 *   box clipping ............. mod2/material_mod.f90:53-93
 *   dielectric ............... mod2/simpledielectric_mod.f90:21-79
 *   gold Drude+2CP ........... mod2/gold_mod.f90:47-168
 *   graphene CP4 + Raman ..... mod2/graphene_mod.f90:56-236
 *   split-field y PML ........ mod2/BERpmly_HzExEy_mod.f90:58-426
 *   halo -> periodic x ....... sim2/include/mympi_2Dfunc.f90 (interchange_E/H)
 *   time loop, source, DFT ... sim2/program.f90:320-438
 *   flux ..................... sim2/program.f90:479-529
 *
 * One domain stands for the reference's Xmax=1, Ymax=3 layout: both PML
 * slabs, the main-field rows npmly..N_y-npmly (mympi_2Dfunc.f90 set_index)
 * and the PML default f2 split at N_y/2 (BERpmly_HzExEy_mod.f90:284-297)
 * coincide with the multi-rank run for every PML cell.
 * Extensions: disks next to boxes; graphene KUBO (Kubo intraband Drude
 * weight, the same form as the 3D oracle's).
 *
 * Pins: the shipped 2D run's output files (whole program, PML included),
 * and golden vectors of the reference's own compiled 2D modules (graphene
 * CP4 + Raman, gold, dielectrics, the si / film gold / graphene3 1 nm
 * Lorentz-Jp modules) run by oracle/ref2d_driver.f90, to 1e-12.
 *
 * TMz (Ez, Hx, Hy; the reference's 2D driver has none) restates the 3D
 * oracle's FUSED stepping (fdtd_oracle.c, ALL_FACES PML) for fields that do
 * not vary along z: the 3D updates of sim3/include/myfdtd_func.f90:19-64
 * with every z difference zero, the split-field PML of
 * mod3/BERpmly_mod.f90:117-195 on the x and y faces (Ez = Ezx + Ezy; the
 * Hxz and Hyz splits stay zero, so Hx and Hy are their y and x splits),
 * gold on Ez (mod3/gold_mod.f90:50-88) and graphene on Ez with the 3D
 * module's coefficients (mod3/graphene_mod.f90:202-283).  It is pinned
 * against the 3D oracle run with nz = 1 and periodic z.
 *
 * TFSF along y (either polarisation) restates the 3D oracle's TFSF
 * (the 1D driver's incident line, sim1/program.f90:217-262, with its
 * mod1/BERpmly_ExHz_mod.f90 end slabs) with axis y: the same line update,
 * profiles and consistency terms, pinned against the 3D oracle at nz = 1.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* the complex quotient's real part in the form fdtd_oracle.c selects */
double ofdtd_cdiv_re(double complex num, double complex den);
#define cdiv_re ofdtd_cdiv_re

#include "../include/fdtd2d.h"

#define O2_PI (4.0 * atan(1.0))
static const double O2_EP0 = 8.85418781762039e-12;
#define O2_MU0 (O2_PI * 4e-7)
#define O2_C0 (1.0 / sqrt(O2_EP0 * O2_MU0))
static const double O2_H = 6.62607015e-34;
static const double O2_KB = 1.380649e-23;
static const double O2_Q = 1.602176634e-19;

static __thread char o2_err[512];
static int o2_fail(int code, const char* msg) {
  snprintf(o2_err, sizeof o2_err, "%s", msg);
  return code;
}
const char* ofdtd2d_last_error(void) { return o2_err; }

enum { M2_DIEL = 0, M2_DRUDE2CP = 1, M2_GRAPHENE = 2, M2_LJP = 3 };

typedef struct {
  int type;
  fdtd2d_region reg;
  int x0, x1, y0, y1; /* clipped box, 1-based inclusive; empty: x0 > x1 */
  int bx, by;
  unsigned char* inside;
  double gax; /* dielectric */
  int ncp;
  double ha, hb;
  double complex gb[4], gc[4], ga;
  double raman0, raman1, raman2;
  double *Ix, *Sx, *Iy, *Sy;
  double complex *gx[4], *gy[4];
  double *sn, *snm1, *snm2;
  /* Lorentz-Jp (the mod2 LorentzJp modules): per component x, y, z */
  fdtd_lorentz_jp lj;
  double lj_ga;
  double *P[4][3], *Jp[4][3], *JpT[3], *Q[3], *G[3];
} o2mat;

typedef struct {
  int j;
  fdtd_pulse p;
} o2src;

typedef struct {
  int j, nfreq, mask, n_begin, n_end;
  double* omega;
  double *re[2], *im[2]; /* [Ex, Hz] */
} o2dft;

typedef struct {
  int i, j, nrec, cap;
  double* rec;
} o2probe;

struct o2sim {
  fdtd2d_grid g;
  int nx, ny;
  long sx, n2;
  double ds, ku, dt;
  double *ex, *ey, *hz, *dx, *dy;
  int has_pml;
  fdtd2d_pml pml;
  int npx, npy, ja, jb;
  double sigmam;
  /* the pmly object: split fields (0:N_x+1,0:N_y+1), coefficients (N_x,N_y) */
  double *pex, *pey, *hzx, *hzy;
  double *f1x, *f2x, *f1y, *f2y, *g1x, *g2x, *g1y, *g2y;
  o2mat* mats;
  int nmats;
  o2src* srcs;
  int nsrcs;
  o2dft* dfts;
  int ndfts;
  o2probe* probes;
  int nprobes;
  int committed, n_done;
  /* TMz */
  int tmz;
  double *ez, *hx, *hy, *dz;
  double *ezx, *ezy, *rz;         /* PML splits of Ez, PML medium 1/eps */
  int tp[2];                      /* PML depth along x, y */
  double *tf1[2], *tf2[2], *tg1[2], *tg2[2]; /* profiles per axis, 0..n+1 */
  /* TFSF along y: the incident line (cells 0..ny+1) and its PML */
  int has_tf;
  fdtd2d_tfsf tfp;
  int tf_lo, tf_hi; /* line slab bounds: npml and ny-npml+1 */
  double tf_sg;
  double *le, *lh, *lpe, *lph, *lf1, *lf2, *lg1, *lg2;
  /* E before the material pass, for the E-form Lorentz-Jp update */
  double *eprev[3];
};
typedef struct o2sim o2sim;

#define F2(s, i, j) ((long)(i) + (s)->sx * (long)(j))            /* ghosted (0:N+1) */
#define C2(s, i, j) ((long)((i)-1) + (long)(s)->nx * (long)((j)-1)) /* (1:N_x,1:N_y)   */

static double* z2(long n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

int ofdtd2d_create(o2sim** out, const fdtd2d_grid* g) {
  if (!out || !g) return o2_fail(FDTD_ERR_ARG, "null argument");
  if (g->nx < 2 || g->ny < 2) return o2_fail(FDTD_ERR_ARG, "grid too small");
  if (!(g->ds > 0) || !(g->courant > 0)) return o2_fail(FDTD_ERR_ARG, "ds and courant must be > 0");
  if (g->polarization != FDTD2D_TEZ && g->polarization != FDTD2D_TMZ) return o2_fail(FDTD_ERR_ARG, "bad polarization");
  o2sim* s = (o2sim*)calloc(1, sizeof(o2sim));
  s->g = *g;
  s->nx = g->nx;
  s->ny = g->ny;
  s->sx = g->nx + 2;
  s->n2 = s->sx * (long)(g->ny + 2);
  s->ds = g->ds;
  s->ku = g->courant;
  s->dt = g->courant * g->ds / O2_C0; /* sim2/program.f90:45 */
  s->ex = z2(s->n2); s->ey = z2(s->n2); s->hz = z2(s->n2); s->dx = z2(s->n2); s->dy = z2(s->n2);
  s->tmz = g->polarization == FDTD2D_TMZ;
  if (s->tmz) { s->ez = z2(s->n2); s->hx = z2(s->n2); s->hy = z2(s->n2); s->dz = z2(s->n2); }
  *out = s;
  return 0;
}

int ofdtd2d_destroy(o2sim* s) {
  if (!s) return 0;
  free(s->ex); free(s->ey); free(s->hz); free(s->dx); free(s->dy);
  free(s->pex); free(s->pey); free(s->hzx); free(s->hzy);
  free(s->f1x); free(s->f2x); free(s->f1y); free(s->f2y);
  free(s->g1x); free(s->g2x); free(s->g1y); free(s->g2y);
  free(s->ez); free(s->hx); free(s->hy); free(s->dz); free(s->ezx); free(s->ezy); free(s->rz);
  free(s->le); free(s->lh); free(s->lpe); free(s->lph); free(s->lf1); free(s->lf2); free(s->lg1); free(s->lg2);
  for (int a = 0; a < 2; ++a) { free(s->tf1[a]); free(s->tf2[a]); free(s->tg1[a]); free(s->tg2[a]); }
  for (int q = 0; q < s->nmats; ++q) {
    o2mat* m = &s->mats[q];
    free(m->inside); free(m->Ix); free(m->Sx); free(m->Iy); free(m->Sy);
    for (int p = 0; p < 4; ++p) { free(m->gx[p]); free(m->gy[p]); }
    free(m->sn); free(m->snm1); free(m->snm2);
    for (int c = 0; c < 3; ++c) {
      for (int l = 0; l < 4; ++l) { free(m->P[l][c]); free(m->Jp[l][c]); }
      free(m->JpT[c]); free(m->Q[c]); free(m->G[c]);
    }
  }
  for (int c = 0; c < 3; ++c) free(s->eprev[c]);
  free(s->mats);
  free(s->srcs);
  for (int q = 0; q < s->ndfts; ++q) {
    free(s->dfts[q].omega);
    for (int c = 0; c < 2; ++c) { free(s->dfts[q].re[c]); free(s->dfts[q].im[c]); }
  }
  free(s->dfts);
  for (int q = 0; q < s->nprobes; ++q) free(s->probes[q].rec);
  free(s->probes);
  free(s);
  return 0;
}

int ofdtd2d_set_pml(o2sim* s, const fdtd2d_pml* p) {
  if (!s || !p) return o2_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o2_fail(FDTD_ERR_STATE, "set-up after commit");
  if (s->tmz) {
    if (p->npml[0] < 0 || p->npml[1] < 0 || 2 * p->npml[0] > s->nx || 2 * p->npml[1] > s->ny)
      return o2_fail(FDTD_ERR_ARG, "bad npml");
    if (p->npml[0] > 0 && s->g.periodic_x) return o2_fail(FDTD_ERR_ARG, "PML on a periodic axis");
    s->pml = *p;
    s->has_pml = p->npml[0] > 0 || p->npml[1] > 0;
    return 0;
  }
  if (p->npml[1] < 1 || p->npml[0] < 0 || 2 * p->npml[1] + 2 > s->ny || 2 * p->npml[0] > s->nx)
    return o2_fail(FDTD_ERR_ARG, "bad npml");
  if (!(p->eps_low > 0) || !(p->eps_high > 0)) return o2_fail(FDTD_ERR_ARG, "eps_low/eps_high must be > 0");
  s->pml = *p;
  s->has_pml = 1;
  return 0;
}

/* initialize_material_indexes (mod2/material_mod.f90:53-93), one rank */
static o2mat* new_mat(o2sim* s, const fdtd2d_region* r, int type) {
  int lo[2], hi[2];
  if (r->kind == FDTD2D_REGION_BOX) {
    lo[0] = r->lo[0]; lo[1] = r->lo[1]; hi[0] = r->hi[0]; hi[1] = r->hi[1];
  } else if (r->kind == FDTD2D_REGION_DISK) {
    if (!(r->radius > 0)) return NULL;
    for (int a = 0; a < 2; ++a) {
      lo[a] = (int)ceil(r->center[a] - r->radius);
      hi[a] = (int)floor(r->center[a] + r->radius);
    }
  } else {
    return NULL;
  }
  s->mats = (o2mat*)realloc(s->mats, sizeof(o2mat) * (size_t)(s->nmats + 1));
  o2mat* m = &s->mats[s->nmats++];
  memset(m, 0, sizeof *m);
  m->type = type;
  m->reg = *r;
  const int n[2] = {s->nx, s->ny};
  int c0[2], c1[2];
  int empty = 0;
  for (int a = 0; a < 2; ++a) {
    c0[a] = lo[a] < 1 ? 1 : lo[a];
    c1[a] = hi[a] > n[a] ? n[a] : hi[a];
    if (c0[a] > c1[a]) empty = 1;
  }
  if (empty) {
    m->x0 = 1; m->x1 = 0; m->y0 = 1; m->y1 = 0;
    m->bx = m->by = 0;
  } else {
    m->x0 = c0[0]; m->x1 = c1[0]; m->y0 = c0[1]; m->y1 = c1[1];
    m->bx = c1[0] - c0[0] + 1;
    m->by = c1[1] - c0[1] + 1;
  }
  m->inside = (unsigned char*)calloc((size_t)(m->bx * m->by > 0 ? m->bx * m->by : 1), 1);
  for (int j = m->y0; j <= m->y1; ++j)
    for (int i = m->x0; i <= m->x1; ++i) {
      int in = 1;
      if (r->kind == FDTD2D_REGION_DISK) {
        const double dx = i - r->center[0], dy = j - r->center[1];
        in = dx * dx + dy * dy <= r->radius * r->radius;
      }
      m->inside[(i - m->x0) + (long)m->bx * (j - m->y0)] = (unsigned char)in;
    }
  return m;
}

int ofdtd2d_add_dielectric(o2sim* s, const fdtd2d_region* r, double eps_r) {
  if (!s || !r) return o2_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o2_fail(FDTD_ERR_STATE, "set-up after commit");
  o2mat* m = new_mat(s, r, M2_DIEL);
  if (!m) return o2_fail(FDTD_ERR_ARG, "bad region");
  m->gax = eps_r == 0.0 ? 0.0 : 1.0 / eps_r; /* simpledielectric_mod.f90:66-75 */
  return 0;
}

static void alloc_state(o2mat* m, int ncomp) {
  const long nb = (long)m->bx * m->by;
  m->Ix = z2(nb); m->Sx = z2(nb);
  for (int p = 0; p < m->ncp; ++p) m->gx[p] = (double complex*)calloc((size_t)(nb > 0 ? nb : 1), sizeof(double complex));
  if (ncomp > 1) {
    m->Iy = z2(nb); m->Sy = z2(nb);
    for (int p = 0; p < m->ncp; ++p) m->gy[p] = (double complex*)calloc((size_t)(nb > 0 ? nb : 1), sizeof(double complex));
  }
}

int ofdtd2d_add_drude2cp(o2sim* s, const fdtd2d_region* r, const fdtd_drude2cp* p) {
  if (!s || !r || !p) return o2_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o2_fail(FDTD_ERR_STATE, "set-up after commit");
  if (!(p->gamma > 0)) return o2_fail(FDTD_ERR_ARG, "gamma must be > 0");
  o2mat* m = new_mat(s, r, M2_DRUDE2CP);
  if (!m) return o2_fail(FDTD_ERR_ARG, "bad region");
  const double dt = s->dt;
  /* setexp and initialize_GoldVars, gold_mod.f90:129-168 */
  m->ncp = 2;
  m->ha = exp(-p->gamma * dt);
  m->hb = (p->omega_d * p->omega_d) * dt / p->gamma;
  for (int q = 0; q < 2; ++q) {
    m->gb[q] = cexp((-p->Gamma[q] + I * p->Omega[q]) * dt);
    m->gc[q] = dt * 2.0 * I * p->A[q] * p->Omega[q] * cexp(-I * p->phi[q]);
  }
  m->ga = p->eps_inf - m->gc[0] - m->gc[1];
  alloc_state(m, s->tmz ? 1 : 2); /* TMz: Ez in the x slots */
  return 0;
}

/* TMz graphene: the 3D module's coefficients, mod3/graphene_mod.f90:202-283
 * (CP4 with the init-order quirk when ref_init_order is set, exact-exponential
 * Raman), or the Kubo Drude weight. */
static int graphene_tmz(o2sim* s, o2mat* m, const fdtd_graphene* p) {
  const double dt = s->dt;
  if (p->model == FDTD_GRAPHENE_CP4) {
    const double scale = p->thickness / s->ds;
    const double wD = sqrt(scale) * p->omega_d;
    m->ha = exp(-p->gamma * dt);
    m->hb = (wD * wD) * dt / p->gamma;
    for (int q = 0; q < 4; ++q) {
      double complex expi = cexp(-I * p->phi[q]);
      double complex ex = cexp((-p->Gamma[q] + I * p->Omega[q]) * dt);
      if (p->ref_init_order) { expi = 0; ex = 0; }
      m->gb[q] = ex;
      m->gc[q] = dt * 2.0 * I * (scale * p->A[q]) * p->Omega[q] * expi;
    }
    m->ga = p->eps_inf - m->gc[0] - m->gc[1] - m->gc[2] - m->gc[3];
  } else if (p->model == FDTD_GRAPHENE_KUBO) {
    if (!(p->tau > 0) || !(p->temperature > 0)) return o2_fail(FDTD_ERR_ARG, "tau and T must be > 0");
    const double hbar = O2_H / (2.0 * O2_PI);
    const double kT = O2_KB * p->temperature;
    const double ef = p->fermi_ev * O2_Q;
    const double Dw = 2.0 * O2_Q * O2_Q * kT / (hbar * hbar) * log(2.0 * cosh(ef / (2.0 * kT)));
    const double wp2 = Dw / (O2_PI * O2_EP0 * s->ds);
    const double gam = 1.0 / p->tau;
    m->ha = exp(-gam * dt);
    m->hb = wp2 * dt / gam;
    for (int q = 0; q < 4; ++q) { m->gb[q] = 0; m->gc[q] = 0; }
    m->ga = p->eps_inf;
  } else {
    return o2_fail(FDTD_ERR_ARG, "bad graphene model");
  }
  m->raman0 = m->raman1 = m->raman2 = 0.0;
  if (p->wNL > 0) { /* graphene_mod.f90:216-218,256-258 */
    const double alfaR = 0.5 * p->dnuNL;
    const double betaR = sqrt((p->wNL * p->wNL) - (alfaR * alfaR));
    const double omegaR = (p->wNL * p->wNL) / betaR;
    m->raman0 = p->xi3 * omegaR * dt * exp(-alfaR * dt) * sin(betaR * dt);
    m->raman1 = 2.0 * exp(-alfaR * dt) * cos(betaR * dt);
    m->raman2 = -exp(-2.0 * alfaR * dt);
  }
  alloc_state(m, 1);
  const long nb = (long)m->bx * m->by;
  m->sn = z2(nb); m->snm1 = z2(nb); m->snm2 = z2(nb);
  return 0;
}

int ofdtd2d_add_graphene(o2sim* s, const fdtd2d_region* r, const fdtd_graphene* p) {
  if (!s || !r || !p) return o2_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o2_fail(FDTD_ERR_STATE, "set-up after commit");
  if (p->normal_axis != 1) return o2_fail(FDTD_ERR_ARG, "a 2D graphene line lies along x (normal_axis 1)");
  o2mat* m = new_mat(s, r, M2_GRAPHENE);
  if (!m) return o2_fail(FDTD_ERR_ARG, "bad region");
  const double dt = s->dt;
  m->ncp = 4;
  if (s->tmz) return graphene_tmz(s, m, p);
  if (p->model == FDTD_GRAPHENE_CP4) {
    /* setexp, scaleParams, initialize_GrapheneVars: graphene_mod.f90:163-236 */
    const double scale = p->thickness / s->ds;
    const double wD = sqrt(scale) * p->omega_d;
    m->ha = exp(-p->gamma * dt);
    m->hb = (wD * wD) * dt / p->gamma;
    for (int q = 0; q < 4; ++q) {
      m->gb[q] = cexp((-p->Gamma[q] + I * p->Omega[q]) * dt);
      m->gc[q] = dt * 2.0 * I * (scale * p->A[q]) * p->Omega[q] * cexp(-I * p->phi[q]);
    }
    m->ga = p->eps_inf - m->gc[0] - m->gc[1] - m->gc[2] - m->gc[3];
  } else if (p->model == FDTD_GRAPHENE_KUBO) {
    /* Kubo intraband sheet conductivity spread over one cell (as in the 3D
     * oracle): eps = eps_inf - wp^2/(w^2 + i w/tau), wp^2 = Dw/(pi ep0 ds) */
    if (!(p->tau > 0) || !(p->temperature > 0)) return o2_fail(FDTD_ERR_ARG, "tau and T must be > 0");
    const double hbar = O2_H / (2.0 * O2_PI);
    const double kT = O2_KB * p->temperature;
    const double ef = p->fermi_ev * O2_Q;
    const double Dw = 2.0 * O2_Q * O2_Q * kT / (hbar * hbar) * log(2.0 * cosh(ef / (2.0 * kT)));
    const double wp2 = Dw / (O2_PI * O2_EP0 * s->ds);
    const double gam = 1.0 / p->tau;
    m->ha = exp(-gam * dt);
    m->hb = wp2 * dt / gam;
    for (int q = 0; q < 4; ++q) { m->gb[q] = 0; m->gc[q] = 0; }
    m->ga = p->eps_inf;
  } else {
    return o2_fail(FDTD_ERR_ARG, "bad graphene model");
  }
  /* Raman, central differences: graphene_mod.f90:210-212 */
  m->raman0 = p->xi3 * ((p->wNL * dt) * (p->wNL * dt)) / (1.0 + p->dnuNL * dt);
  m->raman1 = (2.0 - (p->wNL * dt) * (p->wNL * dt)) / (1.0 + p->dnuNL * dt);
  m->raman2 = -(1.0 - p->dnuNL * dt) / (1.0 + p->dnuNL * dt);
  alloc_state(m, 1);
  const long nb = (long)m->bx * m->by;
  m->sn = z2(nb); m->snm1 = z2(nb); m->snm2 = z2(nb);
  return 0;
}

/* <material>_LJp%initialize_material (mod2/filmGoldLorentzJp_mod.f90:79-122
 * and siblings): ga = 1/epinf over the region, state zero */
int ofdtd2d_add_lorentz_jp(o2sim* s, const fdtd2d_region* r, const fdtd_lorentz_jp* p) {
  if (!s || !r || !p) return o2_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o2_fail(FDTD_ERR_STATE, "set-up after commit");
  if (p->npols < 1 || p->npols > 4 || p->jp_mask < 1 || p->jp_mask > 7 || !(p->eps_inf > 0))
    return o2_fail(FDTD_ERR_ARG, "bad Lorentz-Jp parameters");
  o2mat* m = new_mat(s, r, M2_LJP);
  if (!m) return o2_fail(FDTD_ERR_ARG, "bad region");
  m->lj = *p;
  m->lj_ga = 1.0 / p->eps_inf;
  const long nb = (long)m->bx * m->by;
  for (int c = 0; c < 3; ++c) {
    for (int l = 0; l < p->npols; ++l) { m->P[l][c] = z2(nb); m->Jp[l][c] = z2(nb); }
    m->JpT[c] = z2(nb); m->Q[c] = z2(nb); m->G[c] = z2(nb);
  }
  return 0;
}

static double pulse_at(const fdtd_pulse* p, double t);

/* update_EJp of one cell, mod2/filmGoldLorentzJp_mod.f90:35-75 (and the
 * nonlinear pole of graphene3LorentzJp1nm_mod.f90:52-75) in the 3D oracle's
 * form: for each component q in jp_mask, E_q = Eprev_q + Ku*ga*(curl_q -
 * ds*JpT_q) (+ ga*g on a source row), then the poles from the new E; the
 * other components take E = D.  cur, D, E: per component x, y, z (TMz uses
 * z only, TEz x and y). */
static void ljp_cell2(o2sim* s, o2mat* m, long b, long c, int j, const int* comps, int nc, const double* cur,
                      double** D, double** E, double tnow) {
  const fdtd_lorentz_jp* p = &m->lj;
  const double dt = s->dt;
  const double z0 = sqrt(O2_MU0 / O2_EP0);
  const double Ku = s->dt * O2_C0 / s->ds;
  double e[3] = {0, 0, 0};
  for (int t = 0; t < nc; ++t) {
    const int q = comps[t];
    if (!((p->jp_mask >> q) & 1)) {
      e[q] = D[q][c] * 1.0;
      E[q][c] = e[q];
      continue;
    }
    const double jt = (q == 2 && p->ez_uses_jx) ? m->JpT[0][b] : m->JpT[q][b];
    e[q] = s->eprev[q][c] + Ku * m->lj_ga * (cur[q] - s->ds * jt);
    /* the line source drives Dx (TEz) or Dz (TMz): here ga*g */
    const int sq = s->tmz ? 2 : 0;
    double sv = 0.0;
    int hit = 0;
    for (int r = 0; r < s->nsrcs; ++r)
      if (q == sq && s->srcs[r].j == j) { sv += pulse_at(&s->srcs[r].p, tnow); hit = 1; }
    if (hit) e[q] = e[q] + m->lj_ga * sv;
    E[q][c] = e[q];
  }
  for (int t = 0; t < nc; ++t) {
    const int q = comps[t];
    if (!((p->jp_mask >> q) & 1)) continue;
    m->JpT[q][b] = 0.0;
    for (int l = 0; l < p->npols; ++l) {
      double P0 = p->xi[l] * e[q] / O2_C0;
      if (l == 0 && p->nonlinear) P0 = P0 + p->xi3 * m->Q[q][b] * e[q] / O2_C0;
      m->Jp[l][q][b] = ((1.0 - p->dnu[l] * dt) / (1.0 + p->dnu[l] * dt)) * m->Jp[l][q][b] +
                       ((p->w0[l] * p->w0[l]) * dt / (1.0 + p->dnu[l] * dt)) * (P0 - m->P[l][q][b]);
      m->P[l][q][b] = m->P[l][q][b] + dt * m->Jp[l][q][b];
      m->JpT[q][b] = m->JpT[q][b] + m->Jp[l][q][b];
    }
    if (p->nonlinear) {
      m->G[q][b] = ((1.0 - p->dnuNL * dt) / (1.0 + p->dnuNL * dt)) * m->G[q][b] +
                   ((p->wNL * p->wNL) * dt / (1.0 + p->dnuNL * dt)) * ((z0 * e[q]) * (z0 * e[q]) - m->Q[q][b]);
      m->Q[q][b] = m->Q[q][b] + dt * m->G[q][b];
    }
  }
}

static void save_eprev(o2sim* s) {
  int any = 0;
  for (int q = 0; q < s->nmats; ++q) any |= s->mats[q].type == M2_LJP;
  if (!any) return;
  double* src[3] = {s->ex, s->ey, s->ez};
  for (int c = 0; c < 3; ++c) {
    if (!src[c]) continue;
    if (!s->eprev[c]) s->eprev[c] = z2(s->n2);
    memcpy(s->eprev[c], src[c], sizeof(double) * (size_t)s->n2);
  }
}

int ofdtd2d_add_line_source(o2sim* s, int j, const fdtd_pulse* p) {
  if (!s || !p) return o2_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o2_fail(FDTD_ERR_STATE, "set-up after commit");
  if (j < 1 || j > s->ny) return o2_fail(FDTD_ERR_ARG, "source row out of range");
  s->srcs = (o2src*)realloc(s->srcs, sizeof(o2src) * (size_t)(s->nsrcs + 1));
  s->srcs[s->nsrcs].j = j;
  s->srcs[s->nsrcs].p = *p;
  s->nsrcs++;
  return 0;
}

int ofdtd2d_add_dft_line(o2sim* s, int j, int nfreq, const double* omega, int mask, int n_begin, int n_end, int* id) {
  if (!s || !omega || !id) return o2_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o2_fail(FDTD_ERR_STATE, "set-up after commit");
  const int all = s->tmz ? (FDTD2D_DFT_EZ | FDTD2D_DFT_HX) : (FDTD2D_DFT_EX | FDTD2D_DFT_HZ);
  if (j < 1 || j > s->ny || nfreq < 1 || !(mask & all) || (mask & ~all))
    return o2_fail(FDTD_ERR_ARG, "bad DFT line");
  s->dfts = (o2dft*)realloc(s->dfts, sizeof(o2dft) * (size_t)(s->ndfts + 1));
  o2dft* d = &s->dfts[s->ndfts];
  memset(d, 0, sizeof *d);
  d->j = j; d->nfreq = nfreq; d->mask = mask; d->n_begin = n_begin; d->n_end = n_end;
  d->omega = (double*)malloc(sizeof(double) * (size_t)nfreq);
  memcpy(d->omega, omega, sizeof(double) * (size_t)nfreq);
  for (int c = 0; c < 2; ++c) { d->re[c] = z2((long)s->nx * nfreq); d->im[c] = z2((long)s->nx * nfreq); }
  *id = s->ndfts++;
  return 0;
}

/* the incident line's end-slab profiles: tfsf_line_profiles of the 3D
 * oracle (mod1/BERpmly_ExHz_mod.f90:156-255, sim1/program.f90:132-134) */
static void tf_line_profiles(o2sim* s) {
  const int N = s->ny, np = s->tfp.npml;
  const double dt = s->dt, ds = s->ds;
  const double z0 = sqrt(O2_MU0 / O2_EP0);
  const double Ku = dt * O2_C0 / ds;
  const double ep = s->tfp.eps * O2_EP0, epr = ep / O2_EP0;
  const double order = s->tfp.order;
  const double sigmam = (order + 1.0) * O2_C0 * log(1.0 / s->tfp.R0) / (2.0 * np * ds);
  for (int j = 0; j <= N + 1; ++j) {
    s->lf1[j] = 1.0; s->lf2[j] = Ku / epr; s->lg1[j] = 1.0; s->lg2[j] = Ku;
  }
  if (np <= 0) return;
  for (int j = 1; j <= s->tf_lo; ++j) {
    const double sE = ep * sigmam * pow(((double)(np - j) + 0.5) / (double)np, order);
    const double aE = sE * dt / ep;
    s->lf1[j] = exp(-aE);
    s->lf2[j] = (1.0 - s->lf1[j]) / (z0 * sE * ds);
    if (j < np) {
      const double sH = O2_MU0 * sigmam * pow((double)(np - j) / (double)np, order);
      const double aH = sH * dt / O2_MU0;
      s->lg1[j] = exp(-aH);
      s->lg2[j] = z0 * (1.0 - s->lg1[j]) / (sH * ds);
    }
  }
  int jj = np;
  for (int j = s->tf_hi; j <= N; ++j) {
    if (jj < np) {
      const double sE = ep * sigmam * pow((double)(np - jj) / (double)np, order);
      const double aE = sE * dt / ep;
      s->lf1[j] = exp(-aE);
      s->lf2[j] = (1.0 - s->lf1[j]) / (z0 * sE * ds);
    }
    const double sH = O2_MU0 * sigmam * pow(((double)(np - jj) + 0.5) / (double)np, order);
    const double aH = sH * dt / O2_MU0;
    s->lg1[j] = exp(-aH);
    s->lg2[j] = z0 * (1.0 - s->lg1[j]) / (sH * ds);
    jj = jj - 1;
  }
}

int ofdtd2d_add_tfsf(o2sim* s, const fdtd2d_tfsf* p, int* id) {
  if (!s || !p || !id) return o2_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o2_fail(FDTD_ERR_STATE, "set-up after commit");
  if (s->has_tf) return o2_fail(FDTD_ERR_UNSUPPORTED, "one TFSF per simulation");
  const int N = s->ny;
  if (p->lo < 2 || p->lo > N || (p->hi != 0 && (p->hi < p->lo || p->hi >= N)) || p->src_index < 1 ||
      p->src_index > N || p->npml < 0 || 2 * p->npml + 2 >= N || !(p->eps > 0) || p->src_index <= p->npml ||
      p->src_index > N - p->npml)
    return o2_fail(FDTD_ERR_ARG, "bad TFSF rows");
  s->has_tf = 1;
  s->tfp = *p;
  s->tf_lo = p->npml;
  s->tf_hi = N - p->npml + 1;
  s->tf_sg = s->tmz ? -1.0 : 1.0; /* (Ez, Hx) along y: s = -1; (Ex, Hz): +1 */
  s->le = z2(N + 2); s->lh = z2(N + 2); s->lpe = z2(N + 2); s->lph = z2(N + 2);
  s->lf1 = z2(N + 2); s->lf2 = z2(N + 2); s->lg1 = z2(N + 2); s->lg2 = z2(N + 2);
  tf_line_profiles(s);
  *id = 0;
  return 0;
}

int ofdtd2d_get_tfsf_line(o2sim* s, int id, double* e, double* h) {
  if (!s || !e || !h || id != 0 || !s->has_tf) return o2_fail(FDTD_ERR_ARG, "bad argument");
  memcpy(e, s->le, sizeof(double) * (size_t)(s->ny + 2));
  memcpy(h, s->lh, sizeof(double) * (size_t)(s->ny + 2));
  return 0;
}

/* line E then H half step: tfsf_line_E / tfsf_line_H of the 3D oracle */
static void tf_line_E(o2sim* s, double tnow) {
  const int N = s->ny, lo = s->tf_lo, hi = s->tf_hi;
  const double ck = s->ku * (1.0 / s->tfp.eps);
  for (int j = lo + 1; j <= hi - 1; ++j) s->le[j] = s->le[j] + ck * (s->lh[j] - s->lh[j - 1]);
  const double a = (tnow - s->tfp.pulse.t0) / s->tfp.pulse.st;
  s->le[s->tfp.src_index] = s->le[s->tfp.src_index] +
                            s->tfp.pulse.amplitude * (exp(-0.5 * (a * a)) * sin(s->tfp.pulse.w0 * tnow));
  if (s->tfp.npml <= 0) return;
  for (int j = 1; j <= lo; ++j) s->lpe[j] = s->lf1[j] * s->lpe[j] + s->lf2[j] * (s->lph[j] - s->lph[j - 1]);
  s->lpe[lo + 1] = s->le[lo + 1];
  s->le[lo] = s->lpe[lo];
  for (int j = hi; j <= N; ++j) s->lpe[j] = s->lf1[j] * s->lpe[j] + s->lf2[j] * (s->lph[j] - s->lph[j - 1]);
  s->lpe[hi - 1] = s->le[hi - 1];
  s->le[hi] = s->lpe[hi];
}

static void tf_line_H(o2sim* s) {
  const int N = s->ny, lo = s->tf_lo, hi = s->tf_hi;
  for (int j = lo + 1; j <= hi - 1; ++j) s->lh[j] = s->lh[j] + s->ku * (s->le[j + 1] - s->le[j]);
  if (s->tfp.npml <= 0) return;
  for (int j = 1; j <= lo; ++j) s->lph[j] = s->lg1[j] * s->lph[j] + s->lg2[j] * (s->lpe[j + 1] - s->lpe[j]);
  s->lph[lo + 1] = s->lh[lo + 1];
  s->lh[lo] = s->lph[lo];
  for (int j = hi; j <= N; ++j) s->lph[j] = s->lg1[j] * s->lph[j] + s->lg2[j] * (s->lpe[j + 1] - s->lpe[j]);
  s->lph[hi - 1] = s->lh[hi - 1];
  s->lh[hi] = s->lph[hi];
}

/* add v to D of row j (field 0) or to H (field 1: Hz for TEz, Hx for TMz)
 * on every column; TMz PML cells take the E term on the split along y
 * scaled by their PML medium, as the 3D oracle's plane_add does */
static int tmz_in_pml(const o2sim* s, int i, int j);
static void row_add(o2sim* s, int field, int j, double v) {
  for (int i = 1; i <= s->nx; ++i) {
    const long c = F2(s, i, j);
    if (field == 1) {
      double* h = s->tmz ? s->hx : s->hz;
      h[c] = h[c] + v;
    } else if (s->tmz && tmz_in_pml(s, i, j)) {
      s->ezy[c] = s->ezy[c] + v * s->rz[c];
      s->ez[c] = s->ezx[c] + s->ezy[c];
    } else {
      double* d = s->tmz ? s->dz : s->dx;
      d[c] = d[c] + v;
    }
  }
}

static void tf_corr_D(o2sim* s) {
  row_add(s, 0, s->tfp.lo, -s->ku * s->lh[s->tfp.lo - 1]);
  if (s->tfp.hi > 0) row_add(s, 0, s->tfp.hi + 1, s->ku * s->lh[s->tfp.hi]);
}
static void tf_corr_H(o2sim* s) {
  row_add(s, 1, s->tfp.lo - 1, -(s->tf_sg * s->ku) * s->le[s->tfp.lo]);
  if (s->tfp.hi > 0) row_add(s, 1, s->tfp.hi, (s->tf_sg * s->ku) * s->le[s->tfp.hi + 1]);
}

int ofdtd2d_add_probe(o2sim* s, int i, int j, int* id) {
  if (!s || !id) return o2_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o2_fail(FDTD_ERR_STATE, "set-up after commit");
  if (i < 1 || i > s->nx || j < 1 || j > s->ny) return o2_fail(FDTD_ERR_ARG, "probe out of range");
  s->probes = (o2probe*)realloc(s->probes, sizeof(o2probe) * (size_t)(s->nprobes + 1));
  memset(&s->probes[s->nprobes], 0, sizeof(o2probe));
  s->probes[s->nprobes].i = i;
  s->probes[s->nprobes].j = j;
  *id = s->nprobes++;
  return 0;
}

/* pmly%initialize on one domain: allocate_vars (:249-312), initialize_vars
 * (:314-375), init_varx1/x2 (:377-425) */
static void pml_init(o2sim* s) {
  const int nx = s->nx, ny = s->ny, npx = s->pml.npml[0], npy = s->pml.npml[1];
  const double ep0 = O2_EP0, mu0 = O2_MU0, z0 = sqrt(O2_MU0 / O2_EP0);
  const double eplow = s->pml.eps_low * ep0, ephigh = s->pml.eps_high * ep0;
  const double dt = s->dt, ds = s->ds, order = s->pml.order;
  const double Ku = dt * O2_C0 / ds;
  const double sigmam = s->sigmam;
  s->npx = npx;
  s->npy = npy;
  s->ja = npy;
  s->jb = ny - npy + 1;
  const long nc = (long)nx * ny;
  s->pex = z2(s->n2); s->pey = z2(s->n2); s->hzx = z2(s->n2); s->hzy = z2(s->n2);
  s->f1x = z2(nc); s->f2x = z2(nc); s->f1y = z2(nc); s->f2y = z2(nc);
  s->g1x = z2(nc); s->g2x = z2(nc); s->g1y = z2(nc); s->g2y = z2(nc);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const long c = C2(s, i, j);
      /* f2 defaults split at N_y/2 (:284-297) */
      if (j <= ny / 2 - 1) { s->f2x[c] = Ku / s->pml.eps_low; s->f2y[c] = Ku / s->pml.eps_low; }
      if (j >= ny / 2) { s->f2x[c] = Ku / s->pml.eps_high; s->f2y[c] = Ku / s->pml.eps_high; }
      s->f1x[c] = 1.0; s->g1x[c] = 1.0; s->g2x[c] = Ku;
      s->f1y[c] = 1.0; s->g1y[c] = 1.0; s->g2y[c] = Ku;
    }
  const int xBDhigh = nx - npx + 1, yBDhigh = ny - npy + 1;
  /* init_varx1 / init_varx2 in both slabs (xcord = 0 = lastX) */
  if (npx > 0) {
    for (int slab = 0; slab < 2; ++slab) {
      const int yi = slab == 0 ? 1 : yBDhigh, yf = slab == 0 ? npy : ny;
      const double ep = slab == 0 ? eplow : ephigh;
      for (int j = yi; j <= yf; ++j)
        for (int i = 1; i <= npx; ++i) {
          const long c = C2(s, i, j);
          const double sE = ep * sigmam * pow(((double)(npx - i) + 0.5) / (double)npx, order);
          s->f1x[c] = exp(-(sE * dt / ep));
          s->f2x[c] = (1.0 - s->f1x[c]) / (z0 * sE * ds);
          if (i < npx) {
            const double sH = mu0 * sigmam * pow((double)(npx - i) / (double)npx, order);
            s->g1x[c] = exp(-(sH * dt / mu0));
            s->g2x[c] = z0 * (1.0 - s->g1x[c]) / (sH * ds);
          }
        }
      for (int j = yi; j <= yf; ++j) {
        int ii = npx;
        for (int i = xBDhigh; i <= nx; ++i) {
          const long c = C2(s, i, j);
          if (ii < npx) {
            const double sE = ep * sigmam * pow((double)(npx - ii) / (double)npx, order);
            s->f1x[c] = exp(-(sE * dt / ep));
            s->f2x[c] = (1.0 - s->f1x[c]) / (z0 * sE * ds);
          }
          const double sH = mu0 * sigmam * pow(((double)(npx - ii) + 0.5) / (double)npx, order);
          s->g1x[c] = exp(-(sH * dt / mu0));
          s->g2x[c] = z0 * (1.0 - s->g1x[c]) / (sH * ds);
          --ii;
        }
      }
    }
  }
  for (int j = 1; j <= npy; ++j)
    for (int i = 1; i <= nx; ++i) {
      const long c = C2(s, i, j);
      const double sE = eplow * sigmam * pow(((double)(npy - j) + 0.5) / (double)npy, order);
      s->f1y[c] = exp(-(sE * dt / eplow));
      s->f2y[c] = (1.0 - s->f1y[c]) / (z0 * sE * ds);
      if (j < npy) {
        const double sH = mu0 * sigmam * pow((double)(npy - j) / (double)npy, order);
        s->g1y[c] = exp(-(sH * dt / mu0));
        s->g2y[c] = z0 * (1.0 - s->g1y[c]) / (sH * ds);
      }
    }
  for (int i = 1; i <= nx; ++i) {
    int jj = npy;
    for (int j = yBDhigh; j <= ny; ++j) {
      const long c = C2(s, i, j);
      if (jj < npy) {
        const double sE = ephigh * sigmam * pow((double)(npy - jj) / (double)npy, order);
        s->f1y[c] = exp(-(sE * dt / ephigh));
        s->f2y[c] = (1.0 - s->f1y[c]) / (z0 * sE * ds);
      }
      const double sH = mu0 * sigmam * pow(((double)(npy - jj) + 0.5) / (double)npy, order);
      s->g1y[c] = exp(-(sH * dt / mu0));
      s->g2y[c] = z0 * (1.0 - s->g1y[c]) / (sH * ds);
      --jj;
    }
  }
}

/* TMz PML: init_all_faces of the 3D oracle for the x and y axes (profiles
 * of mod3/BERpmly_mod.f90:612-874 with vacuum grading: E at d = P-x+1/2,
 * H at d = P-x below, mirrored above), defaults f1 = g1 = 1, f2 = g2 = Ku;
 * the PML medium is 1/eps of the last dielectric written to the cell. */
static int tmz_init(o2sim* s) {
  const double z0 = sqrt(O2_MU0 / O2_EP0), ep = O2_EP0, mu0 = O2_MU0;
  const double dt = s->dt, ds = s->ds, order = s->pml.order;
  const double ku_pml = dt * O2_C0 / ds;
  const int n[2] = {s->nx, s->ny};
  for (int a = 0; a < 2; ++a) {
    const int P = s->has_pml ? s->pml.npml[a] : 0;
    s->tp[a] = P;
    s->tf1[a] = z2(n[a] + 2); s->tf2[a] = z2(n[a] + 2); s->tg1[a] = z2(n[a] + 2); s->tg2[a] = z2(n[a] + 2);
    for (int x = 0; x < n[a] + 2; ++x) { s->tf1[a][x] = 1.0; s->tf2[a][x] = ku_pml; s->tg1[a][x] = 1.0; s->tg2[a][x] = ku_pml; }
    if (P == 0) continue;
    const double sigmam = s->pml.sigmam > 0 ? s->pml.sigmam
                                            : (order + 1.0) * O2_C0 * log(1.0 / s->pml.R0) / (2.0 * P * ds);
    for (int x = 1; x <= P; ++x) {
      const double sE = ep * sigmam * pow(((double)(P - x) + 0.5) / (double)P, order);
      s->tf1[a][x] = exp(-(sE * dt / ep));
      s->tf2[a][x] = (1.0 - s->tf1[a][x]) / (z0 * sE * ds);
      if (x < P) {
        const double sH = mu0 * sigmam * pow(((double)(P - x)) / (double)P, order);
        s->tg1[a][x] = exp(-(sH * dt / mu0));
        s->tg2[a][x] = z0 * (1.0 - s->tg1[a][x]) / (sH * ds);
      }
    }
    int q = P;
    for (int x = n[a] - P + 1; x <= n[a]; ++x) {
      if (q < P) {
        const double sE = ep * sigmam * pow(((double)(P - q)) / (double)P, order);
        s->tf1[a][x] = exp(-(sE * dt / ep));
        s->tf2[a][x] = (1.0 - s->tf1[a][x]) / (z0 * sE * ds);
      }
      const double sH = mu0 * sigmam * pow(((double)(P - q) + 0.5) / (double)P, order);
      s->tg1[a][x] = exp(-(sH * dt / mu0));
      s->tg2[a][x] = z0 * (1.0 - s->tg1[a][x]) / (sH * ds);
      --q;
    }
  }
  s->ezx = z2(s->n2); s->ezy = z2(s->n2); s->rz = z2(s->n2);
  for (long c = 0; c < s->n2; ++c) s->rz[c] = 1.0;
  for (int q = 0; q < s->nmats; ++q) {
    const o2mat* m = &s->mats[q];
    for (int j = m->y0; j <= m->y1; ++j)
      for (int i = m->x0; i <= m->x1; ++i) {
        if (!m->inside[(i - m->x0) + (long)m->bx * (j - m->y0)]) continue;
        const int in_pml = (s->tp[0] > 0 && (i <= s->tp[0] || i > s->nx - s->tp[0])) ||
                           (s->tp[1] > 0 && (j <= s->tp[1] || j > s->ny - s->tp[1]));
        if (!in_pml) continue;
        if (m->type == M2_DIEL) s->rz[F2(s, i, j)] = m->gax;
      }
  }
  /* a dispersive cell in the PML would be left to the PML medium there (the
   * 3D engine's rule); refuse it, as the GPU library does */
  for (int j = 1; j <= s->ny; ++j)
    for (int i = 1; i <= s->nx; ++i) {
      const int in_pml = (s->tp[0] > 0 && (i <= s->tp[0] || i > s->nx - s->tp[0])) ||
                         (s->tp[1] > 0 && (j <= s->tp[1] || j > s->ny - s->tp[1]));
      if (!in_pml) continue;
      int last = -1;
      for (int q = 0; q < s->nmats; ++q) {
        const o2mat* m = &s->mats[q];
        if (i < m->x0 || i > m->x1 || j < m->y0 || j > m->y1) continue;
        if (m->inside[(i - m->x0) + (long)m->bx * (j - m->y0)]) last = q;
      }
      if (last >= 0 && s->mats[last].type != M2_DIEL) return o2_fail(FDTD_ERR_UNSUPPORTED, "dispersive cells inside the PML");
    }
  return 0;
}

int ofdtd2d_commit(o2sim* s) {
  if (!s) return o2_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return 0;
  if (s->has_tf && s->has_pml) {
    const int P = s->pml.npml[1];
    const int top = s->tfp.hi > 0 ? s->tfp.hi + 1 : s->tfp.lo;
    if (s->tfp.lo - 1 <= P || top >= s->ny - P + (s->tmz ? 1 : 0))
      return o2_fail(FDTD_ERR_ARG, "TFSF rows inside the PML");
  }
  if (s->tmz) {
    int rc = tmz_init(s);
    if (rc) return rc;
    s->committed = 1;
    return 0;
  }
  if (s->has_pml) {
    s->sigmam = s->pml.sigmam > 0 ? s->pml.sigmam
                                  : (s->pml.order + 1.0) * O2_C0 * log(1.0 / s->pml.R0) /
                                        (2.0 * s->pml.npml[1] * s->ds); /* sim2/program.f90:93 */
    pml_init(s);
  }
  s->committed = 1;
  return 0;
}

static double pulse_at(const fdtd_pulse* p, double t) {
  const double a = (t - p->t0) / p->st;
  return p->amplitude * (exp(-0.5 * (a * a)) * sin(p->w0 * t)); /* sim2/program.f90:337 */
}

/* materials in call order: simpledielectric :29-36, gold :54-73, graphene :63-92,
 * Lorentz-Jp update_EJp (E form) */
static void update_materials(o2sim* s, double tnow) {
  save_eprev(s);
  for (int q = 0; q < s->nmats; ++q) {
    o2mat* m = &s->mats[q];
    for (int j = m->y0; j <= m->y1; ++j)
      for (int i = m->x0; i <= m->x1; ++i) {
        const long b = (i - m->x0) + (long)m->bx * (j - m->y0);
        if (!m->inside[b]) continue;
        const long c = F2(s, i, j);
        if (m->type == M2_LJP) {
          static const int xy[2] = {0, 1};
          const double cur[3] = {s->hz[c] - s->hz[c - s->sx], -s->hz[c] + s->hz[c - 1], 0.0};
          double* D[3] = {s->dx, s->dy, NULL};
          double* E[3] = {s->ex, s->ey, NULL};
          ljp_cell2(s, m, b, c, j, xy, 2, cur, D, E, tnow);
        } else if (m->type == M2_DIEL) {
          s->ex[c] = s->dx[c] * m->gax;
          s->ey[c] = s->dy[c] * m->gax;
        } else if (m->type == M2_DRUDE2CP) {
          double complex num = s->dx[c] - m->Ix[b] + m->ha * m->Sx[b] + m->gb[0] * m->gx[0][b] + m->gb[1] * m->gx[1][b];
          const double ex = cdiv_re(num, m->ga);
          s->ex[c] = ex;
          num = s->dy[c] - m->Iy[b] + m->ha * m->Sy[b] + m->gb[0] * m->gy[0][b] + m->gb[1] * m->gy[1][b];
          const double ey = cdiv_re(num, m->ga);
          s->ey[c] = ey;
          m->Ix[b] = m->Ix[b] + m->hb * ex;
          m->Sx[b] = m->ha * m->Sx[b] + m->hb * ex;
          m->gx[0][b] = m->gb[0] * m->gx[0][b] + m->gc[0] * ex;
          m->gx[1][b] = m->gb[1] * m->gx[1][b] + m->gc[1] * ex;
          m->Iy[b] = m->Iy[b] + m->hb * ey;
          m->Sy[b] = m->ha * m->Sy[b] + m->hb * ey;
          m->gy[0][b] = m->gb[0] * m->gy[0][b] + m->gc[0] * ey;
          m->gy[1][b] = m->gb[1] * m->gy[1][b] + m->gc[1] * ey;
        } else {
          s->ey[c] = s->dy[c];
          const double complex num = s->dx[c] - m->Ix[b] + m->ha * m->Sx[b] + m->gb[0] * m->gx[0][b] +
                                     m->gb[1] * m->gx[1][b] + m->gb[2] * m->gx[2][b] + m->gb[3] * m->gx[3][b];
          const double ex = cdiv_re(num, m->ga + m->sn[b]);
          s->ex[c] = ex;
          m->sn[b] = m->raman1 * m->snm1[b] + m->raman2 * m->snm2[b] + m->raman0 * (ex * ex);
          m->snm2[b] = m->snm1[b];
          m->snm1[b] = m->sn[b];
          m->Ix[b] = m->Ix[b] + m->hb * ex;
          m->Sx[b] = m->ha * m->Sx[b] + m->hb * ex;
          for (int p = 0; p < 4; ++p) m->gx[p][b] = m->gb[p] * m->gx[p][b] + m->gc[p] * ex;
        }
      }
  }
}

/* updateEs, BERpmly_HzExEy_mod.f90:96-123 */
static void pml_updateEs(o2sim* s, int yi, int yf) {
  const int ja = s->ja, jb = s->jb;
  for (int j = yi; j <= yf; ++j)
    for (int i = 1; i <= s->nx; ++i) {
      const long c = F2(s, i, j), k = C2(s, i, j);
      const double curlHzx_p = -s->hzx[c] - s->hzy[c];
      const double curlHzx_m = s->hzx[c - 1] + s->hzy[c - 1];
      double curlHzy_p = s->hzx[c] + s->hzy[c];
      double curlHzy_m = -s->hzx[c - s->sx] - s->hzy[c - s->sx];
      if (j == ja) curlHzy_p = s->hz[c];
      if (j == jb) curlHzy_m = -s->hz[c - s->sx];
      s->pex[c] = s->f1y[k] * s->pex[c] + s->f2y[k] * (curlHzy_p + curlHzy_m);
      s->pey[c] = s->f1x[k] * s->pey[c] + s->f2x[k] * (curlHzx_p + curlHzx_m);
    }
}

/* updateHs, :125-142 */
static void pml_updateHs(o2sim* s, int yi, int yf) {
  for (int j = yi; j <= yf; ++j)
    for (int i = 1; i <= s->nx; ++i) {
      const long c = F2(s, i, j), k = C2(s, i, j);
      const double curlExy = s->pex[c + s->sx] - s->pex[c];
      const double curlEyx = -s->pey[c + 1] + s->pey[c];
      s->hzy[c] = s->g1y[k] * s->hzy[c] + s->g2y[k] * curlExy;
      s->hzx[c] = s->g1x[k] * s->hzx[c] + s->g2x[k] * curlEyx;
    }
}

static void pml_E(o2sim* s) {
  const int nx = s->nx, ny = s->ny, ja = s->ja, jb = s->jb;
  /* low slab: updateEs, setboundayEya (:144-154), setPecEa (:193-199) */
  pml_updateEs(s, 1, ja);
  for (int j = 1; j <= ja; ++j)
    for (int i = 0; i <= nx + 1; ++i) s->ex[F2(s, i, j)] = s->pex[F2(s, i, j)];
  for (int i = 0; i <= nx + 1; ++i) s->pex[F2(s, i, 0)] = 0.0;
  /* high slab: setboundayEyb (:168-178), setPecEb (:201-207) */
  pml_updateEs(s, jb, ny);
  for (int j = jb; j <= ny; ++j)
    for (int i = 0; i <= nx + 1; ++i) s->ex[F2(s, i, j)] = s->pex[F2(s, i, j)];
  for (int i = 0; i <= nx + 1; ++i) s->pex[F2(s, i, ny + 1)] = 0.0;
}

static void pml_H(o2sim* s) {
  const int nx = s->nx, ny = s->ny, ja = s->ja, jb = s->jb;
  /* setboundayHya (:156-166), setPecHa (:209-216) */
  pml_updateHs(s, 1, ja);
  for (int j = 1; j <= ja - 1; ++j)
    for (int i = 0; i <= nx + 1; ++i) s->hz[F2(s, i, j)] = s->hzy[F2(s, i, j)] + s->hzx[F2(s, i, j)];
  for (int i = 0; i <= nx + 1; ++i) { s->hzx[F2(s, i, 0)] = 0.0; s->hzy[F2(s, i, 0)] = 0.0; }
  /* setboundayHyb (:180-191), setPecHb (:218-225) */
  pml_updateHs(s, jb, ny);
  for (int j = jb; j <= ny; ++j)
    for (int i = 0; i <= nx + 1; ++i) s->hz[F2(s, i, j)] = s->hzy[F2(s, i, j)] + s->hzx[F2(s, i, j)];
  for (int i = 0; i <= nx + 1; ++i) { s->hzx[F2(s, i, ny + 1)] = 0.0; s->hzy[F2(s, i, ny + 1)] = 0.0; }
}

/* interchange with PBC_x on one rank: column 1 -> ghost N_x+1, then column
 * N_x -> ghost 0, over rows 0..N_y+1 (the `liney` type) */
static void wrap_x(o2sim* s, double* f) {
  for (int j = 0; j <= s->ny + 1; ++j) f[F2(s, s->nx + 1, j)] = f[F2(s, 1, j)];
  for (int j = 0; j <= s->ny + 1; ++j) f[F2(s, 0, j)] = f[F2(s, s->nx, j)];
}

/* mask bit of DFT slot c (0: Ex or Ez, 1: Hz or Hx) */
static int dft_bit(const o2sim* s, int c) {
  return s->tmz ? (c == 0 ? FDTD2D_DFT_EZ : FDTD2D_DFT_HX) : (c == 0 ? FDTD2D_DFT_EX : FDTD2D_DFT_HZ);
}

static void monitors(o2sim* s, int n) {
  const double t = (double)(n - 1) * s->dt;
  for (int q = 0; q < s->ndfts; ++q) {
    o2dft* d = &s->dfts[q];
    if (n < d->n_begin || (d->n_end > 0 && n > d->n_end)) continue;
    /* frequencies are independent: threads change no sum's order */
#pragma omp parallel for schedule(static)
    for (int m = 0; m < d->nfreq; ++m) {
      const double cs = cos(d->omega[m] * t), sn = sin(d->omega[m] * t);
      for (int c = 0; c < 2; ++c) {
        if (!(d->mask & dft_bit(s, c))) continue;
        const double* f = s->tmz ? (c == 0 ? s->ez : s->hx) : (c == 0 ? s->ex : s->hz);
        double* re = d->re[c] + (long)m * s->nx;
        double* im = d->im[c] + (long)m * s->nx;
        for (int i = 1; i <= s->nx; ++i) {
          const double v = f[F2(s, i, d->j)];
          re[i - 1] = re[i - 1] + cs * v * s->dt;
          im[i - 1] = im[i - 1] + sn * v * s->dt;
        }
      }
    }
  }
  for (int q = 0; q < s->nprobes; ++q) {
    o2probe* p = &s->probes[q];
    if (p->nrec == p->cap) {
      p->cap = p->cap ? 2 * p->cap : 1024;
      p->rec = (double*)realloc(p->rec, sizeof(double) * 3 * (size_t)p->cap);
    }
    const long c = F2(s, p->i, p->j);
    p->rec[3 * p->nrec + 0] = (double)n * s->dt;
    p->rec[3 * p->nrec + 1] = s->tmz ? s->ez[c] : s->ex[c];
    p->rec[3 * p->nrec + 2] = s->tmz ? s->hx[c] : s->hz[c];
    p->nrec++;
  }
}

/* TMz materials, non-PML cells, in call order: dielectric (E = D*gax), gold
 * (mod3/gold_mod.f90:61-83) and graphene (mod3/graphene_mod.f90:86-100) on Ez */
static int tmz_in_pml(const o2sim* s, int i, int j) {
  return (s->tp[0] > 0 && (i <= s->tp[0] || i > s->nx - s->tp[0])) ||
         (s->tp[1] > 0 && (j <= s->tp[1] || j > s->ny - s->tp[1]));
}

static void tmz_materials(o2sim* s, double tnow) {
  save_eprev(s);
  for (int q = 0; q < s->nmats; ++q) {
    o2mat* m = &s->mats[q];
    for (int j = m->y0; j <= m->y1; ++j)
      for (int i = m->x0; i <= m->x1; ++i) {
        const long b = (i - m->x0) + (long)m->bx * (j - m->y0);
        if (!m->inside[b] || tmz_in_pml(s, i, j)) continue;
        const long c = F2(s, i, j);
        if (m->type == M2_LJP) {
          static const int z[1] = {2};
          const double cur[3] = {0.0, 0.0, (s->hy[c] - s->hy[c - 1]) + (-s->hx[c] + s->hx[c - s->sx])};
          double* D[3] = {NULL, NULL, s->dz};
          double* E[3] = {NULL, NULL, s->ez};
          ljp_cell2(s, m, b, c, j, z, 1, cur, D, E, tnow);
        } else if (m->type == M2_DIEL) {
          s->ez[c] = s->dz[c] * m->gax;
        } else if (m->type == M2_DRUDE2CP) {
          const double complex num =
              (s->dz[c] - m->Ix[b] + m->ha * m->Sx[b]) + m->gb[0] * m->gx[0][b] + m->gb[1] * m->gx[1][b];
          const double e = cdiv_re(num, m->ga);
          s->ez[c] = e;
          m->Ix[b] = m->Ix[b] + m->hb * e;
          m->Sx[b] = m->ha * m->Sx[b] + m->hb * e;
          m->gx[0][b] = m->gb[0] * m->gx[0][b] + m->gc[0] * e;
          m->gx[1][b] = m->gb[1] * m->gx[1][b] + m->gc[1] * e;
        } else {
          const double complex num = (s->dz[c] - m->Ix[b] + m->ha * m->Sx[b]) + m->gb[0] * m->gx[0][b] +
                                     m->gb[1] * m->gx[1][b] + m->gb[2] * m->gx[2][b] + m->gb[3] * m->gx[3][b];
          const double e = cdiv_re(num, m->ga + m->sn[b]);
          s->ez[c] = e;
          m->sn[b] = m->raman1 * m->snm1[b] + m->raman2 * m->snm2[b] + m->raman0 * (e * e);
          m->snm2[b] = m->snm1[b];
          m->snm1[b] = m->sn[b];
          m->Ix[b] = m->Ix[b] + m->hb * e;
          m->Sx[b] = m->ha * m->Sx[b] + m->hb * e;
          for (int p = 0; p < 4; ++p) m->gx[p][b] = m->gb[p] * m->gx[p][b] + m->gc[p] * e;
        }
      }
  }
}

static void monitors(o2sim* s, int n);
static void wrap_x(o2sim* s, double* f);

/* One TMz step in the 3D FUSED order: D update of non-PML cells
 * (myfdtd_func.f90:19-41, dz row), split-field E in the PML
 * (BERpmly_mod.f90:117-159, Ezx and Ezy), sources, materials, periodic
 * ghosts of Ez, H update of non-PML cells (:45-64), split-field H in the
 * PML (:161-195), ghosts of Hy, monitors. */
static void tmz_step(o2sim* s, int n) {
  const int nx = s->nx, ny = s->ny;
  const long sx = s->sx;
  const double ku = s->ku, t = (double)(n - 1) * s->dt;
  double *ez = s->ez, *hx = s->hx, *hy = s->hy, *dz = s->dz;
  const double *f1x = s->tf1[0], *f2x = s->tf2[0], *f1y = s->tf1[1], *f2y = s->tf2[1];
  const double *g1x = s->tg1[0], *g2x = s->tg2[0], *g1y = s->tg1[1], *g2y = s->tg2[1];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const long c = F2(s, i, j);
      if (!tmz_in_pml(s, i, j)) {
        const double curlhyx = hy[c] - hy[c - 1];
        const double curlhxy = -hx[c] + hx[c - sx];
        dz[c] = dz[c] + ku * (curlhyx + curlhxy);
      } else {
        /* a part counts as kept only where its axis's slab holds the cell;
         * a part not kept is Ez less the other part, or, neither kept, Ezx
         * is the whole of Ez and Ezy 0 (the 3D oracle's old_parts, at
         * nz = 1); every part is stored, the kept ones read */
        const double r = s->rz[c];
        const int kx = s->tp[0] > 0 && (i <= s->tp[0] || i > nx - s->tp[0]);
        const int ky = s->tp[1] > 0 && (j <= s->tp[1] || j > ny - s->tp[1]);
        const double ox = kx ? s->ezx[c] : (ky ? ez[c] - s->ezy[c] : ez[c]);
        const double oy = ky ? s->ezy[c] : (kx ? ez[c] - s->ezx[c] : 0.0);
        s->ezx[c] = f1x[i] * ox + (f2x[i] * r) * (hy[c] + -hy[c - 1]);
        s->ezy[c] = f1y[j] * oy + (f2y[j] * r) * (-hx[c] + hx[c - sx]);
        ez[c] = s->ezx[c] + s->ezy[c];
      }
    }
  for (int q = 0; q < s->nsrcs; ++q) {
    const double g = pulse_at(&s->srcs[q].p, t);
    const int j = s->srcs[q].j;
    for (int i = 1; i <= nx; ++i) {
      const long c = F2(s, i, j);
      if (tmz_in_pml(s, i, j)) {
        s->ezy[c] = s->ezy[c] + g * s->rz[c];
        ez[c] = s->ezx[c] + s->ezy[c];
      } else {
        dz[c] = dz[c] + g;
      }
    }
  }
  if (s->has_tf) {
    tf_line_E(s, t);
    tf_corr_D(s);
  }
  tmz_materials(s, t);
  if (s->g.periodic_x) wrap_x(s, ez);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const long c = F2(s, i, j);
      if (!tmz_in_pml(s, i, j)) {
        hx[c] = hx[c] + ku * (-ez[c + sx] + ez[c]);
        hy[c] = hy[c] + ku * (ez[c + 1] - ez[c]);
      } else {
        hx[c] = g1y[j] * hx[c] + g2y[j] * (ez[c] + -ez[c + sx]);
        hy[c] = g1x[i] * hy[c] + g2x[i] * (-ez[c] + ez[c + 1]);
      }
    }
  if (s->has_tf) {
    tf_line_H(s);
    tf_corr_H(s);
  }
  if (s->g.periodic_x) wrap_x(s, hy);
  monitors(s, n);
}

int ofdtd2d_step(o2sim* s, int nsteps) {
  if (!s || nsteps < 0) return o2_fail(FDTD_ERR_ARG, "bad argument");
  if (!s->committed) {
    int rc = ofdtd2d_commit(s);
    if (rc) return rc;
  }
  if (s->tmz) {
    for (int it = 0; it < nsteps; ++it) {
      tmz_step(s, s->n_done + 1);
      s->n_done++;
    }
    return 0;
  }
  const int nx = s->nx, ny = s->ny;
  const int j0 = s->has_pml ? s->npy : 1, j1 = s->has_pml ? ny - s->npy : ny;
  const double ku = s->ku;
  for (int it = 0; it < nsteps; ++it) {
    const int n = s->n_done + 1;
    const double t = (double)(n - 1) * s->dt;
    /* updateD, myfdtd_2Dfunc.f90:16-32 (cells are independent: the loop
     * order, x innermost here, changes no result) */
    for (int j = j0; j <= j1; ++j)
      for (int i = 1; i <= nx; ++i) {
        const long c = F2(s, i, j);
        const double curlhx = s->hz[c] - s->hz[c - 1];
        const double curlhy = s->hz[c] - s->hz[c - s->sx];
        s->dx[c] = s->dx[c] + ku * curlhy;
        s->dy[c] = s->dy[c] - ku * curlhx;
      }
    /* source, sim2/program.f90:331-341 */
    for (int q = 0; q < s->nsrcs; ++q) {
      const double g = pulse_at(&s->srcs[q].p, t);
      for (int i = 1; i <= nx; ++i) s->dx[F2(s, i, s->srcs[q].j)] = s->dx[F2(s, i, s->srcs[q].j)] + g;
    }
    if (s->has_tf) {
      tf_line_E(s, t);
      tf_corr_D(s);
    }
    update_materials(s, t);
    if (s->has_pml) pml_E(s);
    if (s->g.periodic_x) wrap_x(s, s->ey);
    /* updateH, myfdtd_2Dfunc.f90:34-44 */
    for (int j = j0; j <= j1; ++j)
      for (int i = 1; i <= nx; ++i) {
        const long c = F2(s, i, j);
        const double curlE = s->ey[c + 1] - s->ey[c] - s->ex[c + s->sx] + s->ex[c];
        s->hz[c] = s->hz[c] - ku * curlE;
      }
    if (s->has_pml) pml_H(s);
    if (s->has_tf) {
      tf_line_H(s);
      tf_corr_H(s);
    }
    if (s->g.periodic_x) wrap_x(s, s->hz);
    monitors(s, n);
    s->n_done = n;
  }
  return 0;
}

int ofdtd2d_sync(o2sim* s) { return s ? 0 : FDTD_ERR_ARG; }
int ofdtd2d_get_step(o2sim* s, int* n) {
  if (!s || !n) return FDTD_ERR_ARG;
  *n = s->n_done;
  return 0;
}
int ofdtd2d_get_dt(o2sim* s, double* dt) {
  if (!s || !dt) return FDTD_ERR_ARG;
  *dt = s->dt;
  return 0;
}

static double* field2(o2sim* s, int comp) {
  if (s->tmz) {
    switch (comp) {
      case FDTD2D_EZ: return s->ez;
      case FDTD2D_HX: return s->hx;
      case FDTD2D_HY: return s->hy;
      case FDTD2D_DZ: return s->dz;
      default: return NULL;
    }
  }
  switch (comp) {
    case FDTD2D_EX: return s->ex;
    case FDTD2D_EY: return s->ey;
    case FDTD2D_HZ: return s->hz;
    case FDTD2D_DX: return s->dx;
    case FDTD2D_DY: return s->dy;
    default: return NULL;
  }
}

int ofdtd2d_get_field(o2sim* s, int comp, double* dst) {
  if (!s || !dst) return o2_fail(FDTD_ERR_ARG, "null argument");
  const double* f = field2(s, comp);
  if (!f) return o2_fail(FDTD_ERR_ARG, "bad component");
  for (int j = 1; j <= s->ny; ++j)
    for (int i = 1; i <= s->nx; ++i) dst[C2(s, i, j)] = f[F2(s, i, j)];
  return 0;
}

int ofdtd2d_set_field(o2sim* s, int comp, const double* src) {
  if (!s || !src) return o2_fail(FDTD_ERR_ARG, "null argument");
  double* f = field2(s, comp);
  if (!f) return o2_fail(FDTD_ERR_ARG, "bad component");
  if (!s->committed) {
    int rc = ofdtd2d_commit(s);
    if (rc) return rc;
  }
  for (int j = 1; j <= s->ny; ++j)
    for (int i = 1; i <= s->nx; ++i) f[F2(s, i, j)] = src[C2(s, i, j)];
  if (s->g.periodic_x && (comp == FDTD2D_EY || comp == FDTD2D_HZ || comp == FDTD2D_EZ || comp == FDTD2D_HY))
    wrap_x(s, f);
  return 0;
}

int ofdtd2d_get_dft(o2sim* s, int id, int comp, double* re, double* im) {
  if (!s || !re || !im || id < 0 || id >= s->ndfts) return o2_fail(FDTD_ERR_ARG, "bad argument");
  const o2dft* d = &s->dfts[id];
  const int c = s->tmz ? (comp == FDTD2D_EZ ? 0 : (comp == FDTD2D_HX ? 1 : -1))
                       : (comp == FDTD2D_EX ? 0 : (comp == FDTD2D_HZ ? 1 : -1));
  if (c < 0 || !(d->mask & dft_bit(s, c))) return o2_fail(FDTD_ERR_ARG, "component not recorded");
  memcpy(re, d->re[c], sizeof(double) * (size_t)s->nx * d->nfreq);
  memcpy(im, d->im[c], sizeof(double) * (size_t)s->nx * d->nfreq);
  return 0;
}

int ofdtd2d_get_flux(o2sim* s, int id, double* S) {
  if (!s || !S || id < 0 || id >= s->ndfts) return o2_fail(FDTD_ERR_ARG, "bad argument");
  const o2dft* d = &s->dfts[id];
  if ((d->mask & (dft_bit(s, 0) | dft_bit(s, 1))) != (dft_bit(s, 0) | dft_bit(s, 1)))
    return o2_fail(FDTD_ERR_ARG, "flux needs both recorded components");
  if (s->tmz) { /* sim3/program.f90:613-685: the Hx*Ez terms */
    for (int m = 0; m < d->nfreq; ++m) {
      double sy = 0.0;
      const long o = (long)m * s->nx;
      for (int i = 0; i < s->nx; ++i)
        sy = sy + ((d->re[1][o + i] * d->re[0][o + i]) + (d->im[1][o + i] * d->im[0][o + i])) * s->ds;
      S[m] = sy;
    }
    return 0;
  }
  for (int m = 0; m < d->nfreq; ++m) {
    double sy = 0.0;
    const long o = (long)m * s->nx;
    for (int i = 0; i < s->nx; ++i)
      sy = sy - ((d->re[1][o + i] * d->re[0][o + i]) + (d->im[1][o + i] * d->im[0][o + i])) * s->ds;
    S[m] = sy;
  }
  return 0;
}

int ofdtd2d_get_probe_count(o2sim* s, int id, int* n) {
  if (!s || !n || id < 0 || id >= s->nprobes) return o2_fail(FDTD_ERR_ARG, "bad argument");
  *n = s->probes[id].nrec;
  return 0;
}

int ofdtd2d_get_probe(o2sim* s, int id, double* out) {
  if (!s || !out || id < 0 || id >= s->nprobes) return o2_fail(FDTD_ERR_ARG, "bad argument");
  memcpy(out, s->probes[id].rec, sizeof(double) * 3 * (size_t)s->probes[id].nrec);
  return 0;
}
