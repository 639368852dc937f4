/*
 * fdtd_oracle.c -- FP64 CPU restatement of the reference's 3D Yee hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path never calls it.
 *
 * It exports the ABI of include/fdtd.h with the prefix `ofdtd_` instead of
 * `fdtd_`, so one host layer can drive either implementation.  The algorithm
 * follows the reference loop by loop, in the reference's stage order and
 * arithmetic order (D-form: D += Ku*curl H, then E = f(D) per material, in
 * material call order), with plain loops over (0:N+1)^3 arrays:
 *
 *   updateD / updateH ........ sim3/include/myfdtd_func.f90:19-64
 *   box clipping ............. mod3/material_mod.f90:60-119
 *   dielectric ............... mod3/simpledielectric_mod.f90:22-103
 *   gold Drude+2CP ........... mod3/gold_mod.f90:50-195
 *   graphene CP4 + Raman ..... mod3/graphene_mod.f90:59-283
 *   split-field y PML ........ mod3/BERpmly_mod.f90:72-306,485-874
 *   halo -> periodic ghosts .. sim3/include/mympi_func.f90:241-690
 *   source, DFT, flux ........ sim3/program.f90:396-406,441-540,613-685
 *   constants ................ misc/constants_mod.f90:3-24
 *
 * Extensions beyond the reference (same style, same formulas):
 *   - spheres/cylinders (staircased, cell centre inside) next to boxes;
 *   - FDTD_PML_ALL_FACES: the reference's split equations and 1D grading
 *     profiles (init_varx1/x2 shape) on every face with npml > 0, the split
 *     fields driven by the main (total) fields, main = sum of splits;
 *   - graphene KUBO: Drude weight from E_F, tau, T (Kubo intraband form);
 *   - energy sums.
 * Parity pins for each function are listed in tests/ (golden vectors from the
 * reference's own modules, closed forms and invariants).
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Real part of the dispersive update's complex quotient num / den.  Complex
 * division is left to the compiler by the reference's Fortran, and builds
 * differ in the last bit: mode 0 (default) is (a*c + b*d) / (c*c + d*d), as
 * amdflang's build of the modules evaluates it (the golden vectors written by
 * the modules compiled here pin it bit for bit); mode 1 is C's complex
 * division (libgcc's ratio form), which reproduces the 2D output the
 * reference ships bit for bit (mod3/gold_mod.f90:61-68,
 * mod3/graphene_mod.f90:70-72, mod2/gold_mod.f90, mod2/graphene_mod.f90). */
int ofdtd_cdiv_mode = 0;
void ofdtd_set_cdiv_mode(int mode) { ofdtd_cdiv_mode = mode; }
double ofdtd_cdiv_re(double complex num, double complex den) {
  if (ofdtd_cdiv_mode) return creal(num / den);
  const double a = creal(num), b = cimag(num), c = creal(den), d = cimag(den);
  return (a * c + b * d) / (c * c + d * d);
}
#define cdiv_re ofdtd_cdiv_re

#include "../include/fdtd.h"

/* ---- constants: misc/constants_mod.f90:3-24 --------------------------- */
#define O_PI (4.0 * atan(1.0))
static const double O_EP0 = 8.85418781762039e-12;
#define O_MU0 (O_PI * 4e-7)
#define O_C0 (1.0 / sqrt(O_EP0 * O_MU0))
static const double O_HBAR_H = 6.62607015e-34;
static const double O_KB = 1.380649e-23;
static const double O_ECHARGE = 1.602176634e-19;

static __thread char o_err[512];
static int o_fail(int code, const char* msg) {
  snprintf(o_err, sizeof o_err, "%s", msg);
  return code;
}
const char* ofdtd_last_error(void) { return o_err; }

enum { M_DIEL = 0, M_DRUDE2CP = 1, M_GRAPHENE = 2, M_LJP = 3, M_LORD = 4 };
enum { S_EXY = 0, S_EXZ, S_EYZ, S_EYX, S_EZX, S_EZY, S_HXY, S_HXZ, S_HYZ, S_HYX, S_HZX, S_HZY, S_N };

typedef struct {
  int type;
  fdtd_region reg;
  int x0, x1, y0, y1, z0, z1; /* clipped box, 1-based inclusive; empty: x0 > x1 */
  int bx, by, bz;             /* box extents                                   */
  unsigned char* inside;      /* per box cell: shape membership                */
  int writes[3];              /* E components this material writes             */
  double gax;                 /* dielectric                                    */
  /* dispersive coefficients */
  int ncp;
  double ha, hb;
  double complex gb[4], gc[4], ga;
  double raman0, raman1, raman2;
  double eps_inf;
  /* dispersive state over the box: I, S per comp; g[p][comp]; Raman */
  double *Ia[3], *S[3];
  double complex* g[4][3];
  double *sn[3], *snm1[3], *snm2[3];
  /* Lorentz-Jp (goldLorentzJp_mod.f90 family): poles, per comp P, Jp, JpT,
   * nonlinear Q, G over the box */
  fdtd_lorentz_jp lj;
  double lj_ga;
  double *P[4][3], *Jp[4][3], *JpT[3], *Q[3], *G[3];
  /* Lorentz poles in D form (mod3/si_mod.f90): per pole the recursion
   * coefficients and, per comp over the box, S and the S one step older */
  double ld_ga[4], ld_gb[4], ld_gc[4];
  double *Sl[4][3], *Slm[4][3];
} omat;

typedef struct {
  int axis, index, comp;
  fdtd_pulse p;
  double g_now; /* value added this step (read by the E-form Lorentz-Jp cells) */
} osrc;

typedef struct {
  int axis, index, nfreq, mask, n_begin, n_end;
  int n1, n2;
  double* omega;
  double* re[6];
  double* im[6];
} odft;

typedef struct {
  int i, j, k;
  int line; /* -1: grid probe; else TFSF id, j = line cell */
  int nrec, cap;
  double* rec;
} oprobe;

/* TFSF plane wave: the incident 1D line (sim1/program.f90:217-262) with its
 * own PML object (mod1/BERpmly_ExHz_mod.f90), cells 0..n+1 */
typedef struct {
  fdtd_tfsf p;
  int n, hc, sg;            /* line length, H component, orientation sign */
  int bdlo, bdhi;           /* yBDlow = npml, yBDhigh = n - npml + 1       */
  double *e, *h, *pe, *ph;  /* line fields and the PML object's own fields */
  double *f1, *f2, *g1, *g2;
} otfsf;

/* one y slab of the reference's pmly object (multi-rank: low slab on rank
 * ycord=0, high slab on rank ycord=Ymax-1; Xmax=Zmax=1) */
typedef struct {
  int active, low;
  double* s[S_N];                                    /* (0:N+1)^3             */
  double *f1x, *f2x, *f1y, *f2y, *f1z, *f2z;         /* (N_x,N_y,N_z)         */
  double *g1x, *g2x, *g1y, *g2y, *g1z, *g2z;
} oslab;

struct osim {
  fdtd_grid g;
  int nx, ny, nz;
  long sx, sy, n3;
  double ds, ku, dt, ku_pml;
  double *E[3], *H[3], *D[3];
  fdtd_pml pml;
  int pml_mode;
  oslab slab[2];
  /* ALL_FACES */
  double* s[S_N];
  double *f1E[3], *f2E[3], *g1H[3], *g2H[3]; /* 1D, index 0..N_a+1 (vacuum) */
  double* pml_gax;                           /* per cell (0:N+1)^3           */
  omat* mats;
  int nmats;
  osrc* srcs;
  int nsrcs;
  odft* dfts;
  int ndfts;
  oprobe* probes;
  int nprobes;
  otfsf* tfsfs;
  int ntfsfs;
  double* eps_cell; /* energy weights */
  double* Eprev[3]; /* E at the start of the material pass (Lorentz-Jp only) */
  int committed;
  int n_done;
};
typedef struct osim osim;

#define IDX(s, i, j, k) ((long)(i) + (s)->sx * ((long)(j) + (s)->sy * (long)(k)))

static double pulse_at(const fdtd_pulse* p, double t);
static int in_pml_all(const osim* s, int i, int j, int k);
static int in_slab(const osim* s, int axis, int v, int n);
#define CIDX(s, i, j, k) ((long)((i)-1) + (long)(s)->nx * ((long)((j)-1) + (long)(s)->ny * (long)((k)-1)))

static double* zalloc(long n) { return (double*)calloc((size_t)n, sizeof(double)); }

void ofdtd_drude2cp_gold(fdtd_drude2cp* p) {
  memset(p, 0, sizeof *p);
  p->eps_inf = 1.1431;
  p->gamma = 1.0805e14;
  p->omega_d = 1.3202e16;
  p->A[0] = 0.26698; p->A[1] = 3.0834;
  p->phi[0] = -1.2371; p->phi[1] = -1.0968;
  p->Omega[0] = 3.8711e15; p->Omega[1] = 4.1684e15;
  p->Gamma[0] = 4.4642e14; p->Gamma[1] = 2.4555e15;
}

void ofdtd_graphene_ref(fdtd_graphene* p) {
  static const double ap[4] = {13.1068, 1.8417, 0.4191, 0.7737};
  static const double phip[4] = {-1.8391, 0.3505, 3.0565, -0.1199};
  static const double omp[4] = {6.1264e14, 6.6843e15, 5.6289e15, 1.118e16};
  static const double rp[4] = {1e14, 7.905e14, 6.7526e14, 1.0005e14};
  memset(p, 0, sizeof *p);
  p->model = FDTD_GRAPHENE_CP4;
  p->normal_axis = 1;
  p->ref_init_order = 1;
  p->thickness = 3.4e-10;
  p->eps_inf = 1.0;
  p->gamma = 9e15;
  p->omega_d = 5.2796e15;
  for (int q = 0; q < 4; ++q) {
    p->A[q] = ap[q]; p->phi[q] = phip[q]; p->Omega[q] = omp[q]; p->Gamma[q] = rp[q];
  }
}

int ofdtd_create(osim** out, const fdtd_grid* g) {
  if (!out || !g) return o_fail(FDTD_ERR_ARG, "null argument");
  if (g->nx < 1 || g->ny < 1 || g->nz < 1 || !(g->ds > 0) || !(g->courant > 0))
    return o_fail(FDTD_ERR_ARG, "bad grid");
  osim* s = (osim*)calloc(1, sizeof(osim));
  s->g = *g;
  s->nx = g->nx; s->ny = g->ny; s->nz = g->nz;
  s->sx = g->nx + 2; s->sy = g->ny + 2;
  s->n3 = s->sx * s->sy * (long)(g->nz + 2);
  s->ds = g->ds;
  s->ku = g->courant;
  s->dt = s->ku * s->ds / O_C0;          /* sim3/program.f90:50 */
  s->ku_pml = s->dt * O_C0 / s->ds;      /* mod3/BERpmly_mod.f90:506 */
  for (int c = 0; c < 3; ++c) {
    s->E[c] = zalloc(s->n3); s->H[c] = zalloc(s->n3); s->D[c] = zalloc(s->n3);
    if (!s->E[c] || !s->H[c] || !s->D[c]) return o_fail(FDTD_ERR_NOMEM, "oracle alloc");
  }
  *out = s;
  return 0;
}

int ofdtd_destroy(osim* s) {
  if (!s) return 0;
  for (int c = 0; c < 3; ++c) { free(s->E[c]); free(s->H[c]); free(s->D[c]); }
  for (int b = 0; b < 2; ++b) {
    oslab* sl = &s->slab[b];
    for (int q = 0; q < S_N; ++q) free(sl->s[q]);
    free(sl->f1x); free(sl->f2x); free(sl->f1y); free(sl->f2y); free(sl->f1z); free(sl->f2z);
    free(sl->g1x); free(sl->g2x); free(sl->g1y); free(sl->g2y); free(sl->g1z); free(sl->g2z);
  }
  for (int q = 0; q < S_N; ++q) free(s->s[q]);
  for (int a = 0; a < 3; ++a) { free(s->f1E[a]); free(s->f2E[a]); free(s->g1H[a]); free(s->g2H[a]); }
  free(s->pml_gax);
  for (int m = 0; m < s->nmats; ++m) {
    omat* mt = &s->mats[m];
    free(mt->inside);
    for (int c = 0; c < 3; ++c) {
      free(mt->Ia[c]); free(mt->S[c]); free(mt->sn[c]); free(mt->snm1[c]); free(mt->snm2[c]);
      for (int p = 0; p < 4; ++p) { free(mt->g[p][c]); free(mt->P[p][c]); free(mt->Jp[p][c]); }
      free(mt->JpT[c]); free(mt->Q[c]); free(mt->G[c]);
    }
  }
  for (int c = 0; c < 3; ++c) free(s->Eprev[c]);
  free(s->mats);
  free(s->srcs);
  for (int d = 0; d < s->ndfts; ++d) {
    free(s->dfts[d].omega);
    for (int c = 0; c < 6; ++c) { free(s->dfts[d].re[c]); free(s->dfts[d].im[c]); }
  }
  free(s->dfts);
  for (int p = 0; p < s->nprobes; ++p) free(s->probes[p].rec);
  free(s->probes);
  for (int q = 0; q < s->ntfsfs; ++q) {
    otfsf* t = &s->tfsfs[q];
    free(t->e); free(t->h); free(t->pe); free(t->ph); free(t->f1); free(t->f2); free(t->g1); free(t->g2);
  }
  free(s->tfsfs);
  free(s->eps_cell);
  free(s);
  return 0;
}

int ofdtd_set_pml(osim* s, const fdtd_pml* p) {
  if (!s || !p) return o_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o_fail(FDTD_ERR_STATE, "set-up after commit");
  if (p->mode < 0 || p->mode > 2) return o_fail(FDTD_ERR_ARG, "bad pml mode");
  s->pml = *p;
  s->pml_mode = p->mode;
  return 0;
}

/* ---- regions ------------------------------------------------------------- */
static int region_box(const osim* s, const fdtd_region* r, int lo[3], int hi[3]) {
  int n[3] = {s->nx, s->ny, s->nz};
  if (r->kind == FDTD_REGION_BOX) {
    for (int a = 0; a < 3; ++a) { lo[a] = r->lo[a]; hi[a] = r->hi[a]; }
  } else if (r->kind == FDTD_REGION_SPHERE || r->kind == FDTD_REGION_CYLINDER) {
    for (int a = 0; a < 3; ++a) {
      double ext = r->radius;
      if (r->kind == FDTD_REGION_CYLINDER && a == r->axis) ext = r->half_length;
      lo[a] = (int)ceil(r->center[a] - ext);
      hi[a] = (int)floor(r->center[a] + ext);
      /* lo/hi of the struct clip the shape when they form a valid box */
      if (r->lo[a] <= r->hi[a] && (r->lo[a] != 0 || r->hi[a] != 0)) {
        if (lo[a] < r->lo[a]) lo[a] = r->lo[a];
        if (hi[a] > r->hi[a]) hi[a] = r->hi[a];
      }
    }
  } else {
    return -1;
  }
  /* initialize_material_indexes: clip to the (single) rank 1..n */
  for (int a = 0; a < 3; ++a) {
    if (lo[a] < 1) lo[a] = 1;
    if (hi[a] > n[a]) hi[a] = n[a];
  }
  return 0;
}

static int region_inside(const fdtd_region* r, int i, int j, int k) {
  double p[3] = {(double)i, (double)j, (double)k};
  if (r->kind == FDTD_REGION_BOX) return 1;
  if (r->kind == FDTD_REGION_SPHERE) {
    double d2 = 0;
    for (int a = 0; a < 3; ++a) d2 += (p[a] - r->center[a]) * (p[a] - r->center[a]);
    return d2 <= r->radius * r->radius;
  }
  double d2 = 0;
  for (int a = 0; a < 3; ++a)
    if (a != r->axis) d2 += (p[a] - r->center[a]) * (p[a] - r->center[a]);
  double dz = fabs(p[r->axis] - r->center[r->axis]);
  return d2 <= r->radius * r->radius && dz <= r->half_length;
}

static omat* new_mat(osim* s, const fdtd_region* r, int type) {
  int lo[3], hi[3];
  if (region_box(s, r, lo, hi)) return NULL;
  s->mats = (omat*)realloc(s->mats, sizeof(omat) * (size_t)(s->nmats + 1));
  omat* m = &s->mats[s->nmats++];
  memset(m, 0, sizeof *m);
  m->type = type;
  m->reg = *r;
  m->x0 = lo[0]; m->x1 = hi[0]; m->y0 = lo[1]; m->y1 = hi[1]; m->z0 = lo[2]; m->z1 = hi[2];
  m->bx = m->x1 >= m->x0 ? m->x1 - m->x0 + 1 : 0;
  m->by = m->y1 >= m->y0 ? m->y1 - m->y0 + 1 : 0;
  m->bz = m->z1 >= m->z0 ? m->z1 - m->z0 + 1 : 0;
  long nb = (long)m->bx * m->by * m->bz;
  m->inside = (unsigned char*)calloc((size_t)(nb > 0 ? nb : 1), 1);
  for (int k = m->z0; k <= m->z1; ++k)
    for (int j = m->y0; j <= m->y1; ++j)
      for (int i = m->x0; i <= m->x1; ++i)
        m->inside[(i - m->x0) + (long)m->bx * ((j - m->y0) + (long)m->by * (k - m->z0))] =
            (unsigned char)region_inside(r, i, j, k);
  m->writes[0] = m->writes[1] = m->writes[2] = 1;
  return m;
}

#define BIDX(m, i, j, k) \
  ((long)((i) - (m)->x0) + (long)(m)->bx * ((long)((j) - (m)->y0) + (long)(m)->by * (long)((k) - (m)->z0)))

int ofdtd_add_dielectric(osim* s, const fdtd_region* r, double eps_r) {
  if (!s || !r) return o_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o_fail(FDTD_ERR_STATE, "set-up after commit");
  omat* m = new_mat(s, r, M_DIEL);
  if (!m) return o_fail(FDTD_ERR_ARG, "bad region");
  /* initialize_ConstantDielectricVars, simpledielectric_mod.f90:91-97 */
  m->gax = (eps_r == 0.0) ? 0.0 : 1.0 / eps_r;
  m->eps_inf = eps_r;
  return 0;
}

static void alloc_state(omat* m) {
  long nb = (long)m->bx * m->by * m->bz;
  if (nb < 1) nb = 1;
  for (int c = 0; c < 3; ++c) {
    m->Ia[c] = zalloc(nb);
    m->S[c] = zalloc(nb);
    for (int p = 0; p < m->ncp; ++p) m->g[p][c] = (double complex*)calloc((size_t)nb, sizeof(double complex));
  }
}

int ofdtd_add_drude2cp(osim* s, const fdtd_region* r, const fdtd_drude2cp* p) {
  if (!s || !r || !p) return o_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o_fail(FDTD_ERR_STATE, "set-up after commit");
  if (!(p->gamma > 0)) return o_fail(FDTD_ERR_ARG, "gamma must be > 0");
  omat* m = new_mat(s, r, M_DRUDE2CP);
  if (!m) return o_fail(FDTD_ERR_ARG, "bad region");
  double dt = s->dt;
  /* setexp, gold_mod.f90:187-195 */
  double complex expi1 = cexp(-I * p->phi[0]);
  double complex expi2 = cexp(-I * p->phi[1]);
  double complex exp1 = cexp((-p->Gamma[0] + I * p->Omega[0]) * dt);
  double complex exp2 = cexp((-p->Gamma[1] + I * p->Omega[1]) * dt);
  /* initialize_GoldVars, gold_mod.f90:175-181 */
  m->ncp = 2;
  m->ha = exp(-p->gamma * dt);
  m->hb = (p->omega_d * p->omega_d) * dt / p->gamma;
  m->gb[0] = exp1;
  m->gb[1] = exp2;
  m->gc[0] = dt * 2.0 * I * p->A[0] * p->Omega[0] * expi1;
  m->gc[1] = dt * 2.0 * I * p->A[1] * p->Omega[1] * expi2;
  m->ga = p->eps_inf - m->gc[0] - m->gc[1];
  m->eps_inf = p->eps_inf;
  alloc_state(m);
  return 0;
}

int ofdtd_add_graphene(osim* s, const fdtd_region* r, const fdtd_graphene* p) {
  if (!s || !r || !p) return o_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o_fail(FDTD_ERR_STATE, "set-up after commit");
  if (p->normal_axis < 0 || p->normal_axis > 2) return o_fail(FDTD_ERR_ARG, "bad normal axis");
  omat* m = new_mat(s, r, M_GRAPHENE);
  if (!m) return o_fail(FDTD_ERR_ARG, "bad region");
  double dt = s->dt;
  m->writes[p->normal_axis] = 0; /* graphene_mod.f90:59-105 writes Ex, Ez only */
  m->ncp = 4;
  m->eps_inf = p->eps_inf;
  if (p->model == FDTD_GRAPHENE_CP4) {
    /* initialize_GrapheneVars, graphene_mod.f90:202-262 */
    double scale = p->thickness / s->ds; /* scaleParams :264-269 */
    double wD = sqrt(scale) * p->omega_d;
    m->ha = exp(-p->gamma * dt);
    m->hb = (wD * wD) * dt / p->gamma;
    m->ga = p->eps_inf;
    for (int q = 0; q < 4; ++q) {
      double complex expi = cexp(-I * p->phi[q]);
      double complex ex = cexp((-p->Gamma[q] + I * p->Omega[q]) * dt);
      if (p->ref_init_order) { expi = 0; ex = 0; } /* locals copied before setexp */
      m->gb[q] = ex;
      m->gc[q] = dt * 2.0 * I * (scale * p->A[q]) * p->Omega[q] * expi;
    }
    m->ga = p->eps_inf - m->gc[0] - m->gc[1] - m->gc[2] - m->gc[3];
  } else if (p->model == FDTD_GRAPHENE_KUBO) {
    /* Kubo intraband: sigma(w) = i*Dw/(pi*(w + i/tau)),
     * Dw = (2 e^2 kB T / hbar^2) ln(2 cosh(E_F / (2 kB T))).
     * As a sheet spread over one cell: eps(w) = eps_inf - wp^2/(w^2 + i w/tau)
     * with wp^2 = Dw / (pi * ep0 * ds): the Drude form of gold_mod's ADE. */
    if (!(p->tau > 0) || !(p->temperature > 0)) return o_fail(FDTD_ERR_ARG, "tau and T must be > 0");
    double hbar = O_HBAR_H / (2.0 * O_PI);
    double kT = O_KB * p->temperature;
    double ef = p->fermi_ev * O_ECHARGE;
    double Dw = 2.0 * O_ECHARGE * O_ECHARGE * kT / (hbar * hbar) * log(2.0 * cosh(ef / (2.0 * kT)));
    double wp2 = Dw / (O_PI * O_EP0 * s->ds);
    double gam = 1.0 / p->tau;
    m->ha = exp(-gam * dt);
    m->hb = wp2 * dt / gam;
    for (int q = 0; q < 4; ++q) { m->gb[q] = 0; m->gc[q] = 0; }
    m->ga = p->eps_inf;
  } else {
    return o_fail(FDTD_ERR_ARG, "bad graphene model");
  }
  /* Raman, exact-exponential form graphene_mod.f90:216-218,256-258 */
  if (p->wNL > 0) {
    double alfaR = 0.5 * p->dnuNL;
    double betaR = sqrt((p->wNL * p->wNL) - (alfaR * alfaR));
    double omegaR = (p->wNL * p->wNL) / betaR;
    m->raman0 = p->xi3 * omegaR * dt * exp(-alfaR * dt) * sin(betaR * dt);
    m->raman1 = 2.0 * exp(-alfaR * dt) * cos(betaR * dt);
    m->raman2 = -exp(-2.0 * alfaR * dt);
  }
  alloc_state(m);
  long nb = (long)m->bx * m->by * m->bz;
  if (nb < 1) nb = 1;
  for (int c = 0; c < 3; ++c) { m->sn[c] = zalloc(nb); m->snm1[c] = zalloc(nb); m->snm2[c] = zalloc(nb); }
  return 0;
}

/* <material>_LJp%initialize_material, mod3/goldLorentzJp_mod.f90:93-163:
 * ga = 1/epinf over the box, state zero.  Nothing else is precomputed: the
 * update evaluates g1, g2 inline as filmGoldLorentzJp_mod.f90:72-83 does
 * (the same expressions initialize_GoldVars stores, :141-150).           */
int ofdtd_add_lorentz_jp(osim* s, const fdtd_region* r, const fdtd_lorentz_jp* p) {
  if (!s || !r || !p) return o_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o_fail(FDTD_ERR_STATE, "set-up after commit");
  if (p->npols < 1 || p->npols > 4 || p->jp_mask < 1 || p->jp_mask > 7 || !(p->eps_inf > 0))
    return o_fail(FDTD_ERR_ARG, "bad Lorentz-Jp parameters");
  omat* m = new_mat(s, r, M_LJP);
  if (!m) return o_fail(FDTD_ERR_ARG, "bad region");
  m->lj = *p;
  m->lj_ga = 1.0 / p->eps_inf;
  m->eps_inf = p->eps_inf;
  long nb = (long)m->bx * m->by * m->bz;
  if (nb < 1) nb = 1;
  for (int c = 0; c < 3; ++c) {
    for (int l = 0; l < p->npols; ++l) { m->P[l][c] = zalloc(nb); m->Jp[l][c] = zalloc(nb); }
    m->JpT[c] = zalloc(nb);
    m->Q[c] = zalloc(nb);
    m->G[c] = zalloc(nb);
  }
  return 0;
}

/* pole tables of the Lorentz-Jp modules (w0, dnu in 2*pi*c0/micro) */
int ofdtd_lorentz_jp_preset(int kind, fdtd_lorentz_jp* p) {
  static const int npols[7] = {3, 3, 4, 2, 4, 3, 3};
  static const int mask[7] = {7, 7, 7, 7, 5, 5, 5};
  static const double einf[7] = {1.053787, 1.0, 1.1, 3.497071, 1.070682, 1.023242, 1.023242};
  static const double xi[7][4] = {{2.862890, -0.0722398, 367.7668, 0},       /* goldLorentzJp_mod.f90:24 */
                                  {8.0, 2.85, -0.107, 0},                    /* siLorentzJp_mod.f90:24 */
                                  {-12.74873, 0.9605442, 67.66233, 55.0525}, /* titaniumJp_mod.f90:24 */
                                  {2.301113, 400.0, 0, 0},                   /* filmGoldLorentzJp_mod.f90:23 */
                                  {-0.163673, 0.843432, -0.170024, -0.121134}, /* grapheneLorentzJp_mod.f90:29 */
                                  {4.957366, 4.455092, 1.905429, 0},         /* graphene3LorentzJp1nm_mod.f90:29 */
                                  {0, 0, 0, 0}};                             /* NodispersiveGraphene.f90 */
  static const double w0[7][4] = {{-2.445679, 1.480982, 0.3371279, 0},
                                  {3.64, 2.76, 1.73, 0},
                                  {4.36014, -1.261863, 0.4063879, -2.176176},
                                  {-2.393546, 0.3441309, 0, 0},
                                  {1.642784, 0.807691, 0.934364, 0.525964},
                                  {0.573075, 1.713164, 4.598546, 0},
                                  {0.573075, 1.713164, 4.598546, 0}};
  static const double dnu[7][4] = {{0.2170955, 0.1429645, 0.02285188, 0},
                                   {0.0, 0.063, 2.5, 0},
                                   {-5.825585, 0.2087635, 1.246245, 8.059584},
                                   {0.2460773, 0.02533959, 0, 0},
                                   {2.267351, 1.018179, 0.621508, 0.033048},
                                   {0.368926, 2.42099, 1.35581, 0},
                                   {0.368926, 2.42099, 1.35581, 0}};
  if (!p || kind < 0 || kind > 6) return o_fail(FDTD_ERR_ARG, "unknown preset");
  const double unit = 2.0 * O_PI * O_C0 / 1e-6; /* 2.d0*pi*c0/micro */
  memset(p, 0, sizeof *p);
  p->npols = npols[kind];
  p->jp_mask = mask[kind];
  p->nonlinear = kind >= 4;
  p->ez_uses_jx = kind == 4;
  p->eps_inf = einf[kind];
  for (int l = 0; l < p->npols; ++l) {
    p->xi[l] = xi[kind][l];
    p->w0[l] = unit * w0[kind][l];
    p->dnu[l] = unit * dnu[kind][l];
  }
  return 0;
}

int ofdtd_add_plane_source(osim* s, int axis, int index, int comp, const fdtd_pulse* p) {
  if (!s || !p) return o_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o_fail(FDTD_ERR_STATE, "set-up after commit");
  int n[3] = {s->nx, s->ny, s->nz};
  if (axis < 0 || axis > 2 || comp < 0 || comp > 2 || index < 1 || index > n[axis])
    return o_fail(FDTD_ERR_ARG, "bad source plane");
  s->srcs = (osrc*)realloc(s->srcs, sizeof(osrc) * (size_t)(s->nsrcs + 1));
  osrc* q = &s->srcs[s->nsrcs++];
  q->axis = axis; q->index = index; q->comp = comp; q->p = *p;
  return 0;
}

static void plane_dims(const osim* s, int axis, int* n1, int* n2) {
  int n[3] = {s->nx, s->ny, s->nz};
  int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  *n1 = n[a1]; *n2 = n[a2];
}

int ofdtd_add_dft_plane(osim* s, int axis, int index, int nfreq, const double* omega, int mask,
                        int n_begin, int n_end, int* id) {
  if (!s || !omega || !id) return o_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o_fail(FDTD_ERR_STATE, "set-up after commit");
  int n[3] = {s->nx, s->ny, s->nz};
  if (axis < 0 || axis > 2 || index < 1 || index > n[axis] || nfreq < 1 || (mask & ~63) || !mask)
    return o_fail(FDTD_ERR_ARG, "bad dft plane");
  s->dfts = (odft*)realloc(s->dfts, sizeof(odft) * (size_t)(s->ndfts + 1));
  odft* d = &s->dfts[s->ndfts];
  memset(d, 0, sizeof *d);
  d->axis = axis; d->index = index; d->nfreq = nfreq; d->mask = mask;
  d->n_begin = n_begin; d->n_end = n_end;
  plane_dims(s, axis, &d->n1, &d->n2);
  d->omega = (double*)malloc(sizeof(double) * (size_t)nfreq);
  memcpy(d->omega, omega, sizeof(double) * (size_t)nfreq);
  long np = (long)d->n1 * d->n2 * nfreq;
  for (int c = 0; c < 6; ++c)
    if (mask & (1 << c)) { d->re[c] = zalloc(np); d->im[c] = zalloc(np); }
  *id = s->ndfts++;
  return 0;
}

int ofdtd_add_probe(osim* s, int i, int j, int k, int* id) {
  if (!s || !id) return o_fail(FDTD_ERR_ARG, "null argument");
  if (i < 1 || i > s->nx || j < 1 || j > s->ny || k < 1 || k > s->nz) return o_fail(FDTD_ERR_ARG, "bad probe");
  s->probes = (oprobe*)realloc(s->probes, sizeof(oprobe) * (size_t)(s->nprobes + 1));
  oprobe* p = &s->probes[s->nprobes];
  memset(p, 0, sizeof *p);
  p->i = i; p->j = j; p->k = k;
  p->line = -1;
  *id = s->nprobes++;
  return 0;
}

/* ---- TFSF ------------------------------------------------------------------ */
/* pmly_ExHz%initialize / allocate_vars / initialize_vars,
 * mod1/BERpmly_ExHz_mod.f90:156-255, for the two end slabs of one line (the
 * low slab of rank ycoord = 0 and the high slab of rank Ymax-1), with
 * eplow = ephigh = eps*ep0 (sim1/program.f90:134). */
static void tfsf_line_profiles(const osim* s, otfsf* t) {
  const int N = t->n, np = t->p.npml;
  const double dt = s->dt, ds = s->ds;
  const double z0 = sqrt(O_MU0 / O_EP0);
  const double Ku = dt * O_C0 / ds;
  const double ep = t->p.eps * O_EP0, epr = ep / O_EP0;
  const double order = t->p.order;
  const double sigmam = (order + 1.0) * O_C0 * log(1.0 / t->p.R0) / (2.0 * np * ds); /* sim1/program.f90:132 */
  for (int j = 0; j <= N + 1; ++j) {
    t->f1[j] = 1.0; t->f2[j] = Ku / epr; t->g1[j] = 1.0; t->g2[j] = Ku;
  }
  if (np <= 0) return;
  for (int j = 1; j <= t->bdlo; ++j) {
    double sE = ep * sigmam * pow(((double)(np - j) + 0.5) / (double)np, order);
    double aE = sE * dt / ep;
    t->f1[j] = exp(-aE);
    t->f2[j] = (1.0 - t->f1[j]) / (z0 * sE * ds);
    if (j < np) {
      double sH = O_MU0 * sigmam * pow((double)(np - j) / (double)np, order);
      double aH = sH * dt / O_MU0;
      t->g1[j] = exp(-aH);
      t->g2[j] = z0 * (1.0 - t->g1[j]) / (sH * ds);
    }
  }
  int jj = np;
  for (int j = t->bdhi; j <= N; ++j) {
    if (jj < np) {
      double sE = ep * sigmam * pow((double)(np - jj) / (double)np, order);
      double aE = sE * dt / ep;
      t->f1[j] = exp(-aE);
      t->f2[j] = (1.0 - t->f1[j]) / (z0 * sE * ds);
    }
    double sH = O_MU0 * sigmam * pow(((double)(np - jj) + 0.5) / (double)np, order);
    double aH = sH * dt / O_MU0;
    t->g1[j] = exp(-aH);
    t->g2[j] = z0 * (1.0 - t->g1[j]) / (sH * ds);
    jj = jj - 1;
  }
}

int ofdtd_add_tfsf(osim* s, const fdtd_tfsf* p, int* id) {
  if (!s || !p || !id) return o_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o_fail(FDTD_ERR_STATE, "set-up after commit");
  int n[3] = {s->nx, s->ny, s->nz};
  if (p->axis < 0 || p->axis > 2 || p->comp < 0 || p->comp > 2 || p->comp == p->axis)
    return o_fail(FDTD_ERR_ARG, "bad TFSF axis/component");
  int N = n[p->axis];
  if (p->lo < 2 || p->lo > N || (p->hi != 0 && (p->hi < p->lo || p->hi >= N)) || p->src_index < 1 ||
      p->src_index > N || p->npml < 0 || 2 * p->npml + 2 >= N || !(p->eps > 0))
    return o_fail(FDTD_ERR_ARG, "bad TFSF planes");
  s->tfsfs = (otfsf*)realloc(s->tfsfs, sizeof(otfsf) * (size_t)(s->ntfsfs + 1));
  otfsf* t = &s->tfsfs[s->ntfsfs];
  memset(t, 0, sizeof *t);
  t->p = *p;
  t->n = N;
  t->hc = 3 - p->axis - p->comp;
  t->sg = ((p->comp == 0 && p->axis == 1) || (p->comp == 1 && p->axis == 2) || (p->comp == 2 && p->axis == 0)) ? 1
                                                                                                           : -1;
  t->bdlo = p->npml;
  t->bdhi = N - p->npml + 1;
  t->e = zalloc(N + 2); t->h = zalloc(N + 2); t->pe = zalloc(N + 2); t->ph = zalloc(N + 2);
  t->f1 = zalloc(N + 2); t->f2 = zalloc(N + 2); t->g1 = zalloc(N + 2); t->g2 = zalloc(N + 2);
  tfsf_line_profiles(s, t);
  *id = s->ntfsfs++;
  return 0;
}

int ofdtd_add_tfsf_probe(osim* s, int tid, int index, int* id) {
  if (!s || !id || tid < 0 || tid >= s->ntfsfs || index < 0 || index > s->tfsfs[tid].n + 1)
    return o_fail(FDTD_ERR_ARG, "bad TFSF probe");
  s->probes = (oprobe*)realloc(s->probes, sizeof(oprobe) * (size_t)(s->nprobes + 1));
  oprobe* p = &s->probes[s->nprobes];
  memset(p, 0, sizeof *p);
  p->line = tid;
  p->j = index;
  *id = s->nprobes++;
  return 0;
}

int ofdtd_get_tfsf_line(osim* s, int tid, double* e, double* h) {
  if (!s || !e || !h || tid < 0 || tid >= s->ntfsfs) return o_fail(FDTD_ERR_ARG, "bad argument");
  otfsf* t = &s->tfsfs[tid];
  memcpy(e, t->e, sizeof(double) * (size_t)(t->n + 2));
  memcpy(h, t->h, sizeof(double) * (size_t)(t->n + 2));
  return 0;
}

/* line E half step: ex_inc update on the main cells, the pulse, then
 * mypmly_inc%updateE (updateE_pmy_low/high + setboundaryEa/Eb),
 * sim1/program.f90:222-231, mod1/BERpmly_ExHz_mod.f90:37-46,60-80,110-130 */
static void tfsf_line_E(osim* s, otfsf* t, double tnow) {
  const int N = t->n, lo = t->bdlo, hi = t->bdhi;
  const double ck = s->ku * (1.0 / t->p.eps);
  for (int j = lo + 1; j <= hi - 1; ++j) t->e[j] = t->e[j] + ck * (t->h[j] - t->h[j - 1]);
  t->e[t->p.src_index] = t->e[t->p.src_index] + pulse_at(&t->p.pulse, tnow);
  if (t->p.npml <= 0) return;
  for (int j = 1; j <= lo; ++j) t->pe[j] = t->f1[j] * t->pe[j] + t->f2[j] * (t->ph[j] - t->ph[j - 1]);
  t->pe[lo + 1] = t->e[lo + 1];
  t->e[lo] = t->pe[lo];
  for (int j = hi; j <= N; ++j) t->pe[j] = t->f1[j] * t->pe[j] + t->f2[j] * (t->ph[j] - t->ph[j - 1]);
  t->pe[hi - 1] = t->e[hi - 1];
  t->e[hi] = t->pe[hi];
}

/* line H half step: hz_inc update, then mypmly_inc%updateH,
 * sim1/program.f90:252-255, mod1/BERpmly_ExHz_mod.f90:48-57,82-102,117-137 */
static void tfsf_line_H(osim* s, otfsf* t) {
  const int N = t->n, lo = t->bdlo, hi = t->bdhi;
  for (int j = lo + 1; j <= hi - 1; ++j) t->h[j] = t->h[j] + s->ku * (t->e[j + 1] - t->e[j]);
  if (t->p.npml <= 0) return;
  for (int j = 1; j <= lo; ++j) t->ph[j] = t->g1[j] * t->ph[j] + t->g2[j] * (t->pe[j + 1] - t->pe[j]);
  t->ph[lo + 1] = t->h[lo + 1];
  t->h[lo] = t->ph[lo];
  for (int j = hi; j <= N; ++j) t->ph[j] = t->g1[j] * t->ph[j] + t->g2[j] * (t->pe[j + 1] - t->pe[j]);
  t->ph[hi - 1] = t->h[hi - 1];
  t->h[hi] = t->ph[hi];
}

/* add v to component q (E: field 0, H: field 1) of every cell of plane
 * `index` normal to `axis`; in ALL_FACES PML cells the value goes into the
 * split along `axis` (E scaled by the cell's PML medium, as the sources) */
static void plane_add(osim* s, int field, int q, int axis, int index, double v) {
  static const int sidxE[3][3] = {{-1, S_EXY, S_EXZ}, {S_EYX, -1, S_EYZ}, {S_EZX, S_EZY, -1}};
  static const int pairE[3][2] = {{S_EXY, S_EXZ}, {S_EYZ, S_EYX}, {S_EZX, S_EZY}};
  static const int sidxH[3][3] = {{-1, S_HXY, S_HXZ}, {S_HYX, -1, S_HYZ}, {S_HZX, S_HZY, -1}};
  static const int pairH[3][2] = {{S_HXY, S_HXZ}, {S_HYZ, S_HYX}, {S_HZX, S_HZY}};
  int n[3] = {s->nx, s->ny, s->nz};
  int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  for (int w = 1; w <= n[a2]; ++w)
    for (int u = 1; u <= n[a1]; ++u) {
      int p[3];
      p[axis] = index; p[a1] = u; p[a2] = w;
      long c = IDX(s, p[0], p[1], p[2]);
      if (s->pml_mode == FDTD_PML_ALL_FACES && in_pml_all(s, p[0], p[1], p[2])) {
        if (field == 0) {
          int si = sidxE[q][axis];
          s->s[si][c] = s->s[si][c] + v * s->pml_gax[c];
          s->E[q][c] = s->s[pairE[q][0]][c] + s->s[pairE[q][1]][c];
        } else {
          /* the H pass reads a part only where its axis's slab holds the
           * cell (old_parts): with both parts of H_q kept, H_q is their
           * sum; otherwise H_q itself takes the value */
          int si = sidxH[q][axis];
          const int st[3] = {in_slab(s, 0, p[0], s->nx), in_slab(s, 1, p[1], s->ny), in_slab(s, 2, p[2], s->nz)};
          const int both = st[q == 0 ? 1 : 0] && st[q == 2 ? 1 : 2];
          if (st[axis]) s->s[si][c] = s->s[si][c] + v;
          s->H[q][c] = both ? s->s[pairH[q][0]][c] + s->s[pairH[q][1]][c] : s->H[q][c] + v;
        }
      } else if (field == 0) {
        s->D[q][c] = s->D[q][c] + v;
      } else {
        s->H[q][c] = s->H[q][c] + v;
      }
    }
}

/* consistency conditions, sim1/program.f90:236-238 (D) and 260-262 (H),
 * plus the far plane of the two-sided form */
static void tfsf_corr_D(osim* s, otfsf* t) {
  plane_add(s, 0, t->p.comp, t->p.axis, t->p.lo, -s->ku * t->h[t->p.lo - 1]);
  if (t->p.hi > 0) plane_add(s, 0, t->p.comp, t->p.axis, t->p.hi + 1, s->ku * t->h[t->p.hi]);
}
static void tfsf_corr_H(osim* s, otfsf* t) {
  plane_add(s, 1, t->hc, t->p.axis, t->p.lo - 1, -(t->sg * s->ku) * t->e[t->p.lo]);
  if (t->p.hi > 0) plane_add(s, 1, t->hc, t->p.axis, t->p.hi, (t->sg * s->ku) * t->e[t->p.hi + 1]);
}

/* ---- PML set-up ------------------------------------------------------------ */
static double pml_sigmam(const osim* s, int npml) {
  /* sim3/program.f90:102 */
  return (s->pml.order + 1.0) * O_C0 * log(1.0 / s->pml.R0) / (2.0 * npml * s->ds);
}

/* init_varx1/x2/z1/z2 and the y profile of initialize_vars, written for one
 * axis: low side (E at d = P-idx+0.5, H at P-idx for idx < P) and high side
 * (counter q from P down: E at P-q for q < P, H at P-q+0.5), each as
 * sigma = ep*sigmam*(d/P)^m, f1 = exp(-sigma*dt/ep), f2 = (1-f1)/(z0*sigma*ds),
 * g1 = exp(-sigma_h*dt/mu0), g2 = z0*(1-g1)/(sigma_h*ds).                     */
static void prof_E_low(const osim* s, double ep, double sigmam, int P, int idx, double* f1, double* f2) {
  double z0 = sqrt(O_MU0 / O_EP0);
  double sig = ep * sigmam * pow(((double)(P - idx) + 0.5) / (double)P, s->pml.order);
  double alpha = sig * s->dt / ep;
  *f1 = exp(-alpha);
  *f2 = (1.0 - *f1) / (z0 * sig * s->ds);
}
static void prof_H_low(const osim* s, double sigmam, int P, int idx, double* g1, double* g2) {
  double z0 = sqrt(O_MU0 / O_EP0);
  double sig = O_MU0 * sigmam * pow(((double)(P - idx)) / (double)P, s->pml.order);
  double alpha = sig * s->dt / O_MU0;
  *g1 = exp(-alpha);
  *g2 = z0 * (1.0 - *g1) / (sig * s->ds);
}
static void prof_E_high(const osim* s, double ep, double sigmam, int P, int q, double* f1, double* f2) {
  double z0 = sqrt(O_MU0 / O_EP0);
  double sig = ep * sigmam * pow(((double)(P - q)) / (double)P, s->pml.order);
  double alpha = sig * s->dt / ep;
  *f1 = exp(-alpha);
  *f2 = (1.0 - *f1) / (z0 * sig * s->ds);
}
static void prof_H_high(const osim* s, double sigmam, int P, int q, double* g1, double* g2) {
  double z0 = sqrt(O_MU0 / O_EP0);
  double sig = O_MU0 * sigmam * pow(((double)(P - q) + 0.5) / (double)P, s->pml.order);
  double alpha = sig * s->dt / O_MU0;
  *g1 = exp(-alpha);
  *g2 = z0 * (1.0 - *g1) / (sig * s->ds);
}

#define C3(s, i, j, k) ((long)((i)-1) + (long)(s)->nx * ((long)((j)-1) + (long)(s)->ny * (long)((k)-1)))

/* pmly%initialize for one slab (mod3/BERpmly_mod.f90:485-874) */
static int init_ref_slab(osim* s, oslab* sl, int low) {
  int nx = s->nx, ny = s->ny, nz = s->nz;
  int npx = s->pml.npml[0], npy = s->pml.npml[1], npz = s->pml.npml[2];
  long nc = (long)nx * ny * nz;
  double eplow = s->pml.eps_low * O_EP0, ephigh = s->pml.eps_high * O_EP0;
  double sigmam = s->pml.sigmam > 0 ? s->pml.sigmam : pml_sigmam(s, npy);
  double ku = s->ku_pml;
  int yBDlow = npy, yBDhigh = ny - npy + 1;
  int xBDlow = npx, xBDhigh = nx - npx + 1;
  int zBDlow = npz, zBDhigh = nz - npz + 1;
  sl->active = 1;
  sl->low = low;
  for (int q = 0; q < S_N; ++q) if (!(sl->s[q] = zalloc(s->n3))) return -1;
  double** arrs[12] = {&sl->f1x, &sl->f2x, &sl->f1y, &sl->f2y, &sl->f1z, &sl->f2z,
                       &sl->g1x, &sl->g2x, &sl->g1y, &sl->g2y, &sl->g1z, &sl->g2z};
  for (int q = 0; q < 12; ++q) if (!(*arrs[q] = zalloc(nc))) return -1;
  /* allocate_vars :567-601 (whole-array defaults) */
  double epr = low ? eplow / O_EP0 : ephigh / O_EP0;
  for (long c = 0; c < nc; ++c) {
    sl->f2x[c] = ku / epr; sl->f2y[c] = ku / epr; sl->f2z[c] = ku / epr;
    sl->f1x[c] = 1.0; sl->f1y[c] = 1.0; sl->f1z[c] = 1.0;
    sl->g1x[c] = 1.0; sl->g1y[c] = 1.0; sl->g1z[c] = 1.0;
    sl->g2x[c] = ku; sl->g2y[c] = ku; sl->g2z[c] = ku;
  }
  double ep = low ? eplow : ephigh;
  int ylo = low ? 1 : yBDhigh, yhi = low ? npy : ny;
  /* x profile inside the slab (:630-656); Xmax = 1: both ends */
  if (xBDlow > 0) {
    for (int k = 1; k <= nz; ++k)
      for (int j = ylo; j <= yhi; ++j) {
        for (int i = 1; i <= npx; ++i) {
          prof_E_low(s, ep, sigmam, npx, i, &sl->f1x[C3(s, i, j, k)], &sl->f2x[C3(s, i, j, k)]);
          if (i < npx) prof_H_low(s, sigmam, npx, i, &sl->g1x[C3(s, i, j, k)], &sl->g2x[C3(s, i, j, k)]);
        }
        int ii = npx;
        for (int i = xBDhigh; i <= nx; ++i) {
          if (ii < npx) prof_E_high(s, ep, sigmam, npx, ii, &sl->f1x[C3(s, i, j, k)], &sl->f2x[C3(s, i, j, k)]);
          prof_H_high(s, sigmam, npx, ii, &sl->g1x[C3(s, i, j, k)], &sl->g2x[C3(s, i, j, k)]);
          --ii;
        }
      }
  }
  /* y profile (:674-714) */
  if (yBDlow > 0) {
    if (low) {
      for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= yBDlow; ++j)
          for (int i = 1; i <= nx; ++i) {
            prof_E_low(s, eplow, sigmam, npy, j, &sl->f1y[C3(s, i, j, k)], &sl->f2y[C3(s, i, j, k)]);
            if (j < npy) prof_H_low(s, sigmam, npy, j, &sl->g1y[C3(s, i, j, k)], &sl->g2y[C3(s, i, j, k)]);
          }
    } else {
      for (int k = 1; k <= nz; ++k)
        for (int i = 1; i <= nx; ++i) {
          int jj = npy;
          for (int j = yBDhigh; j <= ny; ++j) {
            if (jj < npy) prof_E_high(s, ephigh, sigmam, npy, jj, &sl->f1y[C3(s, i, j, k)], &sl->f2y[C3(s, i, j, k)]);
            prof_H_high(s, sigmam, npy, jj, &sl->g1y[C3(s, i, j, k)], &sl->g2y[C3(s, i, j, k)]);
            --jj;
          }
        }
    }
  }
  /* z profile inside the slab (:719-745); Zmax = 1: both ends */
  if (zBDlow > 0) {
    for (int k = 1; k <= npz; ++k)
      for (int j = ylo; j <= yhi; ++j)
        for (int i = 1; i <= nx; ++i) {
          prof_E_low(s, ep, sigmam, npz, k, &sl->f1z[C3(s, i, j, k)], &sl->f2z[C3(s, i, j, k)]);
          if (k < npz) prof_H_low(s, sigmam, npz, k, &sl->g1z[C3(s, i, j, k)], &sl->g2z[C3(s, i, j, k)]);
        }
    for (int i = 1; i <= nx; ++i)
      for (int j = ylo; j <= yhi; ++j) {
        int kk = npz;
        for (int k = zBDhigh; k <= nz; ++k) {
          if (kk < npz) prof_E_high(s, ep, sigmam, npz, kk, &sl->f1z[C3(s, i, j, k)], &sl->f2z[C3(s, i, j, k)]);
          prof_H_high(s, sigmam, npz, kk, &sl->g1z[C3(s, i, j, k)], &sl->g2z[C3(s, i, j, k)]);
          --kk;
        }
      }
  }
  return 0;
}

static int in_pml_all(const osim* s, int i, int j, int k) {
  const int* P = s->pml.npml;
  return (P[0] > 0 && (i <= P[0] || i >= s->nx - P[0] + 1)) ||
         (P[1] > 0 && (j <= P[1] || j >= s->ny - P[1] + 1)) ||
         (P[2] > 0 && (k <= P[2] || k >= s->nz - P[2] + 1));
}

static int init_all_faces(osim* s) {
  int n[3] = {s->nx, s->ny, s->nz};
  for (int a = 0; a < 3; ++a) {
    int P = s->pml.npml[a];
    if (P > 0 && s->g.periodic[a]) return o_fail(FDTD_ERR_ARG, "PML on a periodic axis");
    if (P > 0 && 2 * P > n[a]) return o_fail(FDTD_ERR_ARG, "PML thicker than the grid");
    s->f1E[a] = zalloc(n[a] + 2); s->f2E[a] = zalloc(n[a] + 2);
    s->g1H[a] = zalloc(n[a] + 2); s->g2H[a] = zalloc(n[a] + 2);
    for (int x = 0; x < n[a] + 2; ++x) {
      s->f1E[a][x] = 1.0; s->f2E[a][x] = s->ku_pml; s->g1H[a][x] = 1.0; s->g2H[a][x] = s->ku_pml;
    }
    if (P == 0) continue;
    double sigmam = s->pml.sigmam > 0 ? s->pml.sigmam : pml_sigmam(s, P);
    for (int x = 1; x <= P; ++x) {
      prof_E_low(s, O_EP0, sigmam, P, x, &s->f1E[a][x], &s->f2E[a][x]);
      if (x < P) prof_H_low(s, sigmam, P, x, &s->g1H[a][x], &s->g2H[a][x]);
    }
    int q = P;
    for (int x = n[a] - P + 1; x <= n[a]; ++x) {
      if (q < P) prof_E_high(s, O_EP0, sigmam, P, q, &s->f1E[a][x], &s->f2E[a][x]);
      prof_H_high(s, sigmam, P, q, &s->g1H[a][x], &s->g2H[a][x]);
      --q;
    }
  }
  for (int q = 0; q < S_N; ++q) if (!(s->s[q] = zalloc(s->n3))) return o_fail(FDTD_ERR_NOMEM, "oracle alloc");
  /* PML medium: last dielectric written to the cell (1 where none) */
  s->pml_gax = zalloc(s->n3);
  for (long c = 0; c < s->n3; ++c) s->pml_gax[c] = 1.0;
  for (int m = 0; m < s->nmats; ++m) {
    omat* mt = &s->mats[m];
    for (int k = mt->z0; k <= mt->z1; ++k)
      for (int j = mt->y0; j <= mt->y1; ++j)
        for (int i = mt->x0; i <= mt->x1; ++i) {
          if (!mt->inside[BIDX(mt, i, j, k)] || !in_pml_all(s, i, j, k)) continue;
          if (mt->type == M_DIEL) s->pml_gax[IDX(s, i, j, k)] = mt->gax;
        }
  }
  return 0;
}

int ofdtd_commit(osim* s) {
  if (!s) return o_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return 0;
  if (s->pml_mode == FDTD_PML_REF_Y) {
    int P = s->pml.npml[1];
    if (P < 1 || 2 * P + 2 > s->ny) return o_fail(FDTD_ERR_ARG, "bad y PML thickness");
    if (!(s->pml.eps_low > 0) || !(s->pml.eps_high > 0)) return o_fail(FDTD_ERR_ARG, "bad PML eps");
    if (init_ref_slab(s, &s->slab[0], 1) || init_ref_slab(s, &s->slab[1], 0))
      return o_fail(FDTD_ERR_NOMEM, "oracle alloc");
  } else if (s->pml_mode == FDTD_PML_ALL_FACES) {
    int rc = init_all_faces(s);
    if (rc) return rc;
  }
  /* energy weights: eps of the last material written to the cell */
  s->eps_cell = zalloc(s->n3);
  for (long c = 0; c < s->n3; ++c) s->eps_cell[c] = 1.0;
  for (int m = 0; m < s->nmats; ++m) {
    omat* mt = &s->mats[m];
    for (int k = mt->z0; k <= mt->z1; ++k)
      for (int j = mt->y0; j <= mt->y1; ++j)
        for (int i = mt->x0; i <= mt->x1; ++i)
          if (mt->inside[BIDX(mt, i, j, k)]) s->eps_cell[IDX(s, i, j, k)] = mt->eps_inf;
  }
  s->committed = 1;
  return 0;
}

/* ---- stages ----------------------------------------------------------------- */
/* myfdtd_func.f90:19-41 */
static void update_D(osim* s, int x0, int x1, int y0, int y1, int z0, int z1, int skip_pml) {
  double *hx = s->H[0], *hy = s->H[1], *hz = s->H[2];
  double *dx = s->D[0], *dy = s->D[1], *dz = s->D[2];
  double Ku = s->ku;
#pragma omp parallel for schedule(static) if (s->n3 > 65536 && s->nz >= 8)
  for (int k = z0; k <= z1; ++k)
    for (int j = y0; j <= y1; ++j)
      for (int i = x0; i <= x1; ++i) {
        if (skip_pml && in_pml_all(s, i, j, k)) continue;
        long c = IDX(s, i, j, k);
        double curlhzy = hz[c] - hz[IDX(s, i, j - 1, k)];
        double curlhyz = -hy[c] + hy[IDX(s, i, j, k - 1)];
        double curlhxz = hx[c] - hx[IDX(s, i, j, k - 1)];
        double curlhzx = -hz[c] + hz[IDX(s, i - 1, j, k)];
        double curlhyx = hy[c] - hy[IDX(s, i - 1, j, k)];
        double curlhxy = -hx[c] + hx[IDX(s, i, j - 1, k)];
        dx[c] = dx[c] + Ku * (curlhzy + curlhyz);
        dy[c] = dy[c] + Ku * (curlhxz + curlhzx);
        dz[c] = dz[c] + Ku * (curlhyx + curlhxy);
      }
}

/* myfdtd_func.f90:45-64 */
static void update_H(osim* s, int x0, int x1, int y0, int y1, int z0, int z1, int skip_pml) {
  double *hx = s->H[0], *hy = s->H[1], *hz = s->H[2];
  double *ex = s->E[0], *ey = s->E[1], *ez = s->E[2];
  double Ku = s->ku;
#pragma omp parallel for schedule(static) if (s->n3 > 65536 && s->nz >= 8)
  for (int k = z0; k <= z1; ++k)
    for (int j = y0; j <= y1; ++j)
      for (int i = x0; i <= x1; ++i) {
        if (skip_pml && in_pml_all(s, i, j, k)) continue;
        long c = IDX(s, i, j, k);
        double curleyz = ey[IDX(s, i, j, k + 1)] - ey[c];
        double curlezy = -ez[IDX(s, i, j + 1, k)] + ez[c];
        double curlezx = ez[IDX(s, i + 1, j, k)] - ez[c];
        double curlexz = -ex[IDX(s, i, j, k + 1)] + ex[c];
        double curlexy = ex[IDX(s, i, j + 1, k)] - ex[c];
        double curleyx = -ey[IDX(s, i + 1, j, k)] + ey[c];
        hx[c] = hx[c] + Ku * (curleyz + curlezy);
        hy[c] = hy[c] + Ku * (curlezx + curlexz);
        hz[c] = hz[c] + Ku * (curlexy + curleyx);
      }
}

static double pulse_at(const fdtd_pulse* p, double t) {
  /* sim3/program.f90:400 */
  return p->amplitude * (exp(-0.5 * ((t - p->t0) / p->st) * ((t - p->t0) / p->st)) * sin(p->w0 * t));
}

/* sim3/program.f90:396-406, generalised to any plane/component.  In
 * ALL_FACES PML cells the source goes to the split part along the normal. */
static void apply_sources(osim* s, double t) {
  for (int q = 0; q < s->nsrcs; ++q) {
    osrc* sr = &s->srcs[q];
    double g = pulse_at(&sr->p, t);
    sr->g_now = g;
    int n[3] = {s->nx, s->ny, s->nz};
    int a1 = sr->axis == 0 ? 1 : 0, a2 = sr->axis == 2 ? 1 : 2;
    for (int v = 1; v <= n[a2]; ++v)
      for (int u = 1; u <= n[a1]; ++u) {
        int p[3];
        p[sr->axis] = sr->index; p[a1] = u; p[a2] = v;
        long c = IDX(s, p[0], p[1], p[2]);
        if (s->pml_mode == FDTD_PML_ALL_FACES && in_pml_all(s, p[0], p[1], p[2])) {
          /* split index of (comp, axis) */
          static const int sidx[3][3] = {{-1, S_EXY, S_EXZ}, {S_EYX, -1, S_EYZ}, {S_EZX, S_EZY, -1}};
          int si = sidx[sr->comp][sr->axis];
          if (si < 0) continue;
          static const int pair[3][2] = {{S_EXY, S_EXZ}, {S_EYZ, S_EYX}, {S_EZX, S_EZY}};
          s->s[si][c] = s->s[si][c] + g * s->pml_gax[c];
          s->E[sr->comp][c] = s->s[pair[sr->comp][0]][c] + s->s[pair[sr->comp][1]][c];
        } else {
          s->D[sr->comp][c] = s->D[sr->comp][c] + g;
        }
      }
  }
}

/* update_EJp of the Lorentz-Jp family for one cell,
 * mod3/goldLorentzJp_mod.f90:47-88 (and grapheneLorentzJp_mod.f90:52-100 for
 * the nonlinear term and the Ez/JpxT coupling).  The E form reads E as it
 * was before this material pass (Eprev), so an earlier writer of the same
 * cell does not feed it; components outside jp_mask are vacuum (E = D). */
static void ljp_cell(osim* s, omat* mt, long b, int i, int j, int k, long c) {
  const fdtd_lorentz_jp* p = &mt->lj;
  double **Hx = &s->H[0], **Hy = &s->H[1], **Hz = &s->H[2];
  const double dt = s->dt;
  const double z0 = sqrt(O_MU0 / O_EP0);
  const double Ku = s->dt * O_C0 / s->ds;
  double curlhzy = (*Hz)[c] - (*Hz)[IDX(s, i, j - 1, k)];
  double curlhyz = -(*Hy)[c] + (*Hy)[IDX(s, i, j, k - 1)];
  double curlhxz = (*Hx)[c] - (*Hx)[IDX(s, i, j, k - 1)];
  double curlhzx = -(*Hz)[c] + (*Hz)[IDX(s, i - 1, j, k)];
  double curlhyx = (*Hy)[c] - (*Hy)[IDX(s, i - 1, j, k)];
  double curlhxy = -(*Hx)[c] + (*Hx)[IDX(s, i, j - 1, k)];
  const double cur[3] = {curlhzy + curlhyz, curlhxz + curlhzx, curlhyx + curlhxy};
  double e[3];
  for (int q = 0; q < 3; ++q) {
    if (!((p->jp_mask >> q) & 1)) {
      e[q] = s->D[q][c] * 1.0;
      s->E[q][c] = e[q];
      continue;
    }
    double jt = (q == 2 && p->ez_uses_jx) ? mt->JpT[0][b] : mt->JpT[q][b];
    e[q] = s->Eprev[q][c] + Ku * mt->lj_ga * (cur[q] - s->ds * jt);
    /* a plane source through the cell enters as D does elsewhere: ga*g */
    double sv = 0.0;
    int hit = 0;
    for (int r = 0; r < s->nsrcs; ++r) {
      const osrc* sr = &s->srcs[r];
      const int pos = sr->axis == 0 ? i : (sr->axis == 1 ? j : k);
      if (sr->comp == q && pos == sr->index) { sv += sr->g_now; hit = 1; }
    }
    if (hit) e[q] = e[q] + mt->lj_ga * sv;
    s->E[q][c] = e[q];
  }
  for (int q = 0; q < 3; ++q) {
    if (!((p->jp_mask >> q) & 1)) continue;
    mt->JpT[q][b] = 0.0;
    for (int l = 0; l < p->npols; ++l) {
      double P0 = p->xi[l] * e[q] / O_C0;
      if (l == 0 && p->nonlinear) P0 = P0 + p->xi3 * mt->Q[q][b] * e[q] / O_C0;
      mt->Jp[l][q][b] = ((1.0 - p->dnu[l] * dt) / (1.0 + p->dnu[l] * dt)) * mt->Jp[l][q][b] +
                        ((p->w0[l] * p->w0[l]) * dt / (1.0 + p->dnu[l] * dt)) * (P0 - mt->P[l][q][b]);
      mt->P[l][q][b] = mt->P[l][q][b] + dt * mt->Jp[l][q][b];
      mt->JpT[q][b] = mt->JpT[q][b] + mt->Jp[l][q][b];
    }
    if (p->nonlinear) {
      mt->G[q][b] = ((1.0 - p->dnuNL * dt) / (1.0 + p->dnuNL * dt)) * mt->G[q][b] +
                    ((p->wNL * p->wNL) * dt / (1.0 + p->dnuNL * dt)) * ((z0 * e[q]) * (z0 * e[q]) - mt->Q[q][b]);
      mt->Q[q][b] = mt->Q[q][b] + dt * mt->G[q][b];
    }
  }
}

/* Si%initialize_material, mod3/si_mod.f90:234-259: the recursion
 * coefficients of each pole (lj holds the pole table), state zero. */
int ofdtd_add_lorentz_d(osim* s, const fdtd_region* r, const fdtd_lorentz_jp* p) {
  if (!s || !r || !p) return o_fail(FDTD_ERR_ARG, "null argument");
  if (s->committed) return o_fail(FDTD_ERR_STATE, "set-up after commit");
  if (p->npols < 1 || p->npols > 4 || p->jp_mask != 7 || p->nonlinear || p->ez_uses_jx || p->eps_inf != 1.0)
    return o_fail(FDTD_ERR_ARG, "bad D-form Lorentz parameters");
  omat* m = new_mat(s, r, M_LORD);
  if (!m) return o_fail(FDTD_ERR_ARG, "bad region");
  m->lj = *p;
  m->eps_inf = 1.0;
  double dt = s->dt;
  for (int l = 0; l < p->npols; ++l) {
    /* si_mod.f90:249-257 */
    m->ld_ga[l] = (2.0 - (dt * p->w0[l]) * (dt * p->w0[l])) / (1.0 + dt * p->dnu[l]);
    m->ld_gb[l] = (dt * p->dnu[l] - 1.0) / (1.0 + dt * p->dnu[l]);
    m->ld_gc[l] = (p->xi[l] * ((p->w0[l] * dt) * (p->w0[l] * dt))) / (1.0 + dt * p->dnu[l]);
  }
  long nb = (long)m->bx * m->by * m->bz;
  if (nb < 1) nb = 1;
  for (int c = 0; c < 3; ++c)
    for (int l = 0; l < p->npols; ++l) { m->Sl[l][c] = zalloc(nb); m->Slm[l][c] = zalloc(nb); }
  return 0;
}

/* update_E of every material in call order (sim3/program.f90:411-414) */
static void update_materials(osim* s) {
  int skip_pml = s->pml_mode == FDTD_PML_ALL_FACES;
  int any_ljp = 0;
  for (int m = 0; m < s->nmats; ++m) any_ljp |= s->mats[m].type == M_LJP;
  if (any_ljp)
    for (int q = 0; q < 3; ++q) {
      if (!s->Eprev[q]) s->Eprev[q] = zalloc(s->n3);
      memcpy(s->Eprev[q], s->E[q], sizeof(double) * (size_t)s->n3);
    }
  for (int m = 0; m < s->nmats; ++m) {
    omat* mt = &s->mats[m];
    if (mt->bx == 0 || mt->by == 0 || mt->bz == 0) continue;
#pragma omp parallel for schedule(static) if (s->n3 > 65536 && s->nz >= 8)
    for (int k = mt->z0; k <= mt->z1; ++k)
      for (int j = mt->y0; j <= mt->y1; ++j)
        for (int i = mt->x0; i <= mt->x1; ++i) {
          long b = BIDX(mt, i, j, k);
          if (!mt->inside[b]) continue;
          if (skip_pml && in_pml_all(s, i, j, k)) continue;
          long c = IDX(s, i, j, k);
          if (mt->type == M_DIEL) {
            /* simpledielectric_mod.f90:181-183 */
            for (int q = 0; q < 3; ++q) s->E[q][c] = s->D[q][c] * mt->gax;
          } else if (mt->type == M_LJP) {
            ljp_cell(s, mt, b, i, j, k, c);
          } else if (mt->type == M_LORD) {
            /* si_mod.f90:57-63: E from D and the S of the previous step,
             * then each pole's recursion, then the shift */
            for (int q = 0; q < 3; ++q) {
              double e = s->D[q][c];
              for (int l = 0; l < mt->lj.npols; ++l) e = e - mt->Sl[l][q][b];
              s->E[q][c] = e;
              for (int l = 0; l < mt->lj.npols; ++l) {
                double sn = (mt->ld_ga[l] * mt->Sl[l][q][b] + mt->ld_gb[l] * mt->Slm[l][q][b]) + mt->ld_gc[l] * e;
                mt->Slm[l][q][b] = mt->Sl[l][q][b];
                mt->Sl[l][q][b] = sn;
              }
            }
          } else if (mt->type == M_DRUDE2CP) {
            /* gold_mod.f90:61-83 */
            for (int q = 0; q < 3; ++q) {
              double complex num = (s->D[q][c] - mt->Ia[q][b] + mt->ha * mt->S[q][b]) + mt->gb[0] * mt->g[0][q][b] +
                                   mt->gb[1] * mt->g[1][q][b];
              double e = cdiv_re(num, mt->ga);
              s->E[q][c] = e;
              mt->Ia[q][b] = mt->Ia[q][b] + mt->hb * e;
              mt->S[q][b] = mt->ha * mt->S[q][b] + mt->hb * e;
              mt->g[0][q][b] = mt->gb[0] * mt->g[0][q][b] + mt->gc[0] * e;
              mt->g[1][q][b] = mt->gb[1] * mt->g[1][q][b] + mt->gc[1] * e;
            }
          } else {
            /* graphene_mod.f90:70-100 */
            for (int q = 0; q < 3; ++q) {
              if (!mt->writes[q]) continue;
              double complex num = (s->D[q][c] - mt->Ia[q][b] + mt->ha * mt->S[q][b]) + mt->gb[0] * mt->g[0][q][b] +
                                   mt->gb[1] * mt->g[1][q][b] + mt->gb[2] * mt->g[2][q][b] +
                                   mt->gb[3] * mt->g[3][q][b];
              double e = cdiv_re(num, mt->ga + mt->sn[q][b]);
              s->E[q][c] = e;
              mt->sn[q][b] = mt->raman1 * mt->snm1[q][b] + mt->raman2 * mt->snm2[q][b] + mt->raman0 * (e * e);
              mt->snm2[q][b] = mt->snm1[q][b];
              mt->snm1[q][b] = mt->sn[q][b];
              mt->Ia[q][b] = mt->Ia[q][b] + mt->hb * e;
              mt->S[q][b] = mt->ha * mt->S[q][b] + mt->hb * e;
              for (int p = 0; p < 4; ++p) mt->g[p][q][b] = mt->gb[p] * mt->g[p][q][b] + mt->gc[p] * e;
            }
          }
        }
  }
}

/* ---- reference y-slab PML ------------------------------------------------ */
/* updateEs, BERpmly_mod.f90:117-159 */
static void ref_updateEs(osim* s, oslab* sl, int yi, int yf) {
  double **S = sl->s;
#pragma omp parallel for schedule(static) if (s->n3 > 65536 && s->nz >= 8)
  for (int k = 1; k <= s->nz; ++k)
    for (int j = yi; j <= yf; ++j)
      for (int i = 1; i <= s->nx; ++i) {
        long c = IDX(s, i, j, k), cjm = IDX(s, i, j - 1, k), ckm = IDX(s, i, j, k - 1), cim = IDX(s, i - 1, j, k);
        long e = C3(s, i, j, k);
        double curlHzy_p = S[S_HZX][c] + S[S_HZY][c];
        double curlHzy_m = -S[S_HZX][cjm] - S[S_HZY][cjm];
        double curlHyz_p = -S[S_HYZ][c] - S[S_HYX][c];
        double curlHyz_m = S[S_HYZ][ckm] + S[S_HYX][ckm];
        double curlHxz_p = S[S_HXY][c] + S[S_HXZ][c];
        double curlHxz_m = -S[S_HXY][ckm] - S[S_HXZ][ckm];
        double curlHzx_p = -S[S_HZX][c] - S[S_HZY][c];
        double curlHzx_m = S[S_HZX][cim] + S[S_HZY][cim];
        double curlHyx_p = S[S_HYZ][c] + S[S_HYX][c];
        double curlHyx_m = -S[S_HYZ][cim] - S[S_HYX][cim];
        double curlHxy_p = -S[S_HXY][c] - S[S_HXZ][c];
        double curlHxy_m = S[S_HXY][cjm] + S[S_HXZ][cjm];
        S[S_EYX][c] = sl->f1x[e] * S[S_EYX][c] + sl->f2x[e] * (curlHzx_p + curlHzx_m);
        S[S_EZX][c] = sl->f1x[e] * S[S_EZX][c] + sl->f2x[e] * (curlHyx_p + curlHyx_m);
        S[S_EXY][c] = sl->f1y[e] * S[S_EXY][c] + sl->f2y[e] * (curlHzy_p + curlHzy_m);
        S[S_EZY][c] = sl->f1y[e] * S[S_EZY][c] + sl->f2y[e] * (curlHxy_p + curlHxy_m);
        S[S_EXZ][c] = sl->f1z[e] * S[S_EXZ][c] + sl->f2z[e] * (curlHyz_p + curlHyz_m);
        S[S_EYZ][c] = sl->f1z[e] * S[S_EYZ][c] + sl->f2z[e] * (curlHxz_p + curlHxz_m);
      }
}

/* updateHs, BERpmly_mod.f90:161-195 */
static void ref_updateHs(osim* s, oslab* sl, int yi, int yf) {
  double **S = sl->s;
#pragma omp parallel for schedule(static) if (s->n3 > 65536 && s->nz >= 8)
  for (int k = 1; k <= s->nz; ++k)
    for (int j = yi; j <= yf; ++j)
      for (int i = 1; i <= s->nx; ++i) {
        long c = IDX(s, i, j, k), cjp = IDX(s, i, j + 1, k), ckp = IDX(s, i, j, k + 1), cip = IDX(s, i + 1, j, k);
        long e = C3(s, i, j, k);
        double curlEzy_p = -S[S_EZX][cjp] - S[S_EZY][cjp];
        double curlEzy_m = S[S_EZX][c] + S[S_EZY][c];
        double curlEyz_p = S[S_EYZ][ckp] + S[S_EYX][ckp];
        double curlEyz_m = -S[S_EYZ][c] - S[S_EYX][c];
        double curlExz_p = -S[S_EXY][ckp] - S[S_EXZ][ckp];
        double curlExz_m = S[S_EXY][c] + S[S_EXZ][c];
        double curlEzx_p = S[S_EZX][cip] + S[S_EZY][cip];
        double curlEzx_m = -S[S_EZX][c] - S[S_EZY][c];
        double curlEyx_p = -S[S_EYZ][cip] - S[S_EYX][cip];
        double curlEyx_m = S[S_EYZ][c] + S[S_EYX][c];
        double curlExy_p = S[S_EXY][cjp] + S[S_EXZ][cjp];
        double curlExy_m = -S[S_EXY][c] - S[S_EXZ][c];
        S[S_HYX][c] = sl->g1x[e] * S[S_HYX][c] + sl->g2x[e] * (curlEzx_m + curlEzx_p);
        S[S_HZX][c] = sl->g1x[e] * S[S_HZX][c] + sl->g2x[e] * (curlEyx_m + curlEyx_p);
        S[S_HXY][c] = sl->g1y[e] * S[S_HXY][c] + sl->g2y[e] * (curlEzy_m + curlEzy_p);
        S[S_HZY][c] = sl->g1y[e] * S[S_HZY][c] + sl->g2y[e] * (curlExy_m + curlExy_p);
        S[S_HXZ][c] = sl->g1z[e] * S[S_HXZ][c] + sl->g2z[e] * (curlEyz_m + curlEyz_p);
        S[S_HYZ][c] = sl->g1z[e] * S[S_HYZ][c] + sl->g2z[e] * (curlExz_m + curlExz_p);
      }
}

/* updateE_pml / updateH_pml with setbounday{E,H}y{a,b}, BERpmly_mod.f90:72-112,200-306 */
static void ref_pml_E(osim* s) {
  int P = s->pml.npml[1], ja = P, jb = s->ny - P + 1;
  oslab *lo = &s->slab[0], *hi = &s->slab[1];
  ref_updateEs(s, lo, 1, ja);
  for (int k = 1; k <= s->nz; ++k)
    for (int i = 1; i <= s->nx; ++i) {
      lo->s[S_EXY][IDX(s, i, ja + 1, k)] = s->E[0][IDX(s, i, ja + 1, k)];
      s->E[0][IDX(s, i, ja, k)] = lo->s[S_EXY][IDX(s, i, ja, k)];
      lo->s[S_EZY][IDX(s, i, ja + 1, k)] = s->E[2][IDX(s, i, ja + 1, k)];
      s->E[2][IDX(s, i, ja, k)] = lo->s[S_EZY][IDX(s, i, ja, k)];
    }
  ref_updateEs(s, hi, jb, s->ny);
  for (int k = 1; k <= s->nz; ++k)
    for (int i = 1; i <= s->nx; ++i) {
      hi->s[S_EXY][IDX(s, i, jb - 1, k)] = s->E[0][IDX(s, i, jb - 1, k)];
      s->E[0][IDX(s, i, jb, k)] = hi->s[S_EXY][IDX(s, i, jb, k)];
      hi->s[S_EZY][IDX(s, i, jb - 1, k)] = s->E[2][IDX(s, i, jb - 1, k)];
      s->E[2][IDX(s, i, jb, k)] = hi->s[S_EZY][IDX(s, i, jb, k)];
    }
}

static void ref_pml_H(osim* s) {
  int P = s->pml.npml[1], ja = P, jb = s->ny - P + 1;
  oslab *lo = &s->slab[0], *hi = &s->slab[1];
  ref_updateHs(s, lo, 1, ja);
  for (int k = 1; k <= s->nz; ++k)
    for (int i = 1; i <= s->nx; ++i) {
      lo->s[S_HXY][IDX(s, i, ja + 1, k)] = s->H[0][IDX(s, i, ja + 1, k)];
      s->H[0][IDX(s, i, ja, k)] = lo->s[S_HXY][IDX(s, i, ja, k)];
      lo->s[S_HZY][IDX(s, i, ja + 1, k)] = s->H[2][IDX(s, i, ja + 1, k)];
      s->H[2][IDX(s, i, ja, k)] = lo->s[S_HZY][IDX(s, i, ja, k)];
    }
  ref_updateHs(s, hi, jb, s->ny);
  for (int k = 1; k <= s->nz; ++k)
    for (int i = 1; i <= s->nx; ++i) {
      hi->s[S_HXY][IDX(s, i, jb - 1, k)] = s->H[0][IDX(s, i, jb - 1, k)];
      s->H[0][IDX(s, i, jb, k)] = hi->s[S_HXY][IDX(s, i, jb, k)];
      hi->s[S_HZY][IDX(s, i, jb - 1, k)] = s->H[2][IDX(s, i, jb - 1, k)];
      s->H[2][IDX(s, i, jb, k)] = hi->s[S_HZY][IDX(s, i, jb, k)];
    }
}

/* ---- all-faces split-field PML (extension) --------------------------------- */

/* cell v of an axis of n cells lies in one of the axis's two PML slabs */
static int in_slab(const osim* s, int axis, int v, int n) {
  const int P = s->pml.npml[axis];
  return v <= P || v > n - P;
}

/* Old split parts of an ALL_FACES PML cell.  A component is the sum of two
 * parts, one per transverse axis (Ex = Exy + Exz, Ey = Eyz + Eyx, Ez = Ezx +
 * Ezy; H alike).  A part counts as kept only where its axis's slab holds the
 * cell (elsewhere that axis's profile is the interior's, f1 = 1, f2 = Ku):
 * a part not kept is the old component less its other part, or, with
 * neither part kept, the first part of the sum is the whole old component
 * and the second 0.  This array keeps every part it computes (the plane
 * sources after the update read them) but reads only the kept ones.
 * Parts in the order yx, zx, xy, zy, xz, yz from split index `yx` on (the
 * enum keeps E and H parts in the same order: xy, xz, yz, yx, zx, zy).
 * An extension of mod3/BERpmly_mod.f90:117-195's split form to six faces;
 * in exact arithmetic the step of storing every part. */
static void old_parts(const osim* s, double** S, int yx, double** F, long c, int i, int j, int k, double* o) {
  const int sx = in_slab(s, 0, i, s->nx), sy = in_slab(s, 1, j, s->ny), sz = in_slab(s, 2, k, s->nz);
  const int xy = yx - 3, xz = yx - 2, yz = yx - 1, zx = yx + 1, zy = yx + 2;
  const double f0 = F[0][c], f1 = F[1][c], f2 = F[2][c];
  o[0] = sx ? S[yx][c] : (sz ? f1 - S[yz][c] : 0.0);
  o[1] = sx ? S[zx][c] : (sy ? f2 - S[zy][c] : f2);
  o[2] = sy ? S[xy][c] : (sz ? f0 - S[xz][c] : f0);
  o[3] = sy ? S[zy][c] : (sx ? f2 - S[zx][c] : 0.0);
  o[4] = sz ? S[xz][c] : (sy ? f0 - S[xy][c] : 0.0);
  o[5] = sz ? S[yz][c] : (sx ? f1 - S[yx][c] : f1);
}

static void all_pml_E(osim* s) {
  double **S = s->s, *hx = s->H[0], *hy = s->H[1], *hz = s->H[2];
  double *f1x = s->f1E[0], *f2x = s->f2E[0], *f1y = s->f1E[1], *f2y = s->f2E[1];
  double *f1z = s->f1E[2], *f2z = s->f2E[2];
#pragma omp parallel for schedule(static) if (s->n3 > 65536 && s->nz >= 8)
  for (int k = 1; k <= s->nz; ++k)
    for (int j = 1; j <= s->ny; ++j)
      for (int i = 1; i <= s->nx; ++i) {
        if (!in_pml_all(s, i, j, k)) continue;
        long c = IDX(s, i, j, k);
        double r = s->pml_gax[c];
        double curlHzy_p = hz[c], curlHzy_m = -hz[IDX(s, i, j - 1, k)];
        double curlHyz_p = -hy[c], curlHyz_m = hy[IDX(s, i, j, k - 1)];
        double curlHxz_p = hx[c], curlHxz_m = -hx[IDX(s, i, j, k - 1)];
        double curlHzx_p = -hz[c], curlHzx_m = hz[IDX(s, i - 1, j, k)];
        double curlHyx_p = hy[c], curlHyx_m = -hy[IDX(s, i - 1, j, k)];
        double curlHxy_p = -hx[c], curlHxy_m = hx[IDX(s, i, j - 1, k)];
        double o[6];
        old_parts(s, S, S_EYX, s->E, c, i, j, k, o);
        S[S_EYX][c] = f1x[i] * o[0] + (f2x[i] * r) * (curlHzx_p + curlHzx_m);
        S[S_EZX][c] = f1x[i] * o[1] + (f2x[i] * r) * (curlHyx_p + curlHyx_m);
        S[S_EXY][c] = f1y[j] * o[2] + (f2y[j] * r) * (curlHzy_p + curlHzy_m);
        S[S_EZY][c] = f1y[j] * o[3] + (f2y[j] * r) * (curlHxy_p + curlHxy_m);
        S[S_EXZ][c] = f1z[k] * o[4] + (f2z[k] * r) * (curlHyz_p + curlHyz_m);
        S[S_EYZ][c] = f1z[k] * o[5] + (f2z[k] * r) * (curlHxz_p + curlHxz_m);
        s->E[0][c] = S[S_EXY][c] + S[S_EXZ][c];
        s->E[1][c] = S[S_EYZ][c] + S[S_EYX][c];
        s->E[2][c] = S[S_EZX][c] + S[S_EZY][c];
      }
}

static void all_pml_H(osim* s) {
  double **S = s->s, *ex = s->E[0], *ey = s->E[1], *ez = s->E[2];
  double *g1x = s->g1H[0], *g2x = s->g2H[0], *g1y = s->g1H[1], *g2y = s->g2H[1];
  double *g1z = s->g1H[2], *g2z = s->g2H[2];
#pragma omp parallel for schedule(static) if (s->n3 > 65536 && s->nz >= 8)
  for (int k = 1; k <= s->nz; ++k)
    for (int j = 1; j <= s->ny; ++j)
      for (int i = 1; i <= s->nx; ++i) {
        if (!in_pml_all(s, i, j, k)) continue;
        long c = IDX(s, i, j, k);
        double curlEzy_p = -ez[IDX(s, i, j + 1, k)], curlEzy_m = ez[c];
        double curlEyz_p = ey[IDX(s, i, j, k + 1)], curlEyz_m = -ey[c];
        double curlExz_p = -ex[IDX(s, i, j, k + 1)], curlExz_m = ex[c];
        double curlEzx_p = ez[IDX(s, i + 1, j, k)], curlEzx_m = -ez[c];
        double curlEyx_p = -ey[IDX(s, i + 1, j, k)], curlEyx_m = ey[c];
        double curlExy_p = ex[IDX(s, i, j + 1, k)], curlExy_m = -ex[c];
        double o[6];
        old_parts(s, S, S_HYX, s->H, c, i, j, k, o);
        S[S_HYX][c] = g1x[i] * o[0] + g2x[i] * (curlEzx_m + curlEzx_p);
        S[S_HZX][c] = g1x[i] * o[1] + g2x[i] * (curlEyx_m + curlEyx_p);
        S[S_HXY][c] = g1y[j] * o[2] + g2y[j] * (curlEzy_m + curlEzy_p);
        S[S_HZY][c] = g1y[j] * o[3] + g2y[j] * (curlExy_m + curlExy_p);
        S[S_HXZ][c] = g1z[k] * o[4] + g2z[k] * (curlEyz_m + curlEyz_p);
        S[S_HYZ][c] = g1z[k] * o[5] + g2z[k] * (curlExz_m + curlExz_p);
        s->H[0][c] = S[S_HXY][c] + S[S_HXZ][c];
        s->H[1][c] = S[S_HYZ][c] + S[S_HYX][c];
        s->H[2][c] = S[S_HZX][c] + S[S_HZY][c];
      }
}

/* ---- periodic ghosts: interchange_E / interchange_H on one rank ----------- */
static void copy_xplane(osim* s, double* f, int from, int to) {
  for (int k = 0; k <= s->nz + 1; ++k)
    for (int j = 0; j <= s->ny + 1; ++j) f[IDX(s, to, j, k)] = f[IDX(s, from, j, k)];
}
static void copy_yplane(osim* s, double* f, int from, int to) {
  for (int k = 0; k <= s->nz + 1; ++k)
    for (int i = 0; i <= s->nx + 1; ++i) f[IDX(s, i, to, k)] = f[IDX(s, i, from, k)];
}
static void copy_zplane(osim* s, double* f, int from, int to) {
  for (int j = 0; j <= s->ny + 1; ++j)
    for (int i = 0; i <= s->nx + 1; ++i) f[IDX(s, i, j, to)] = f[IDX(s, i, j, from)];
}

/* mympi_func.f90:241-518: phase 1 sends plane 1 to ghost N+1, phase 2 plane
 * N to ghost 0; x moves Ey,Ez, y moves Ex,Ez, z moves Ex,Ey. */
static void exchange(osim* s, double** F, int phase_first_low) {
  const int* pb = s->g.periodic;
  int nx = s->nx, ny = s->ny, nz = s->nz;
  for (int ph = 0; ph < 2; ++ph) {
    int low = (ph == 0) == phase_first_low; /* low: 1 -> N+1 */
    if (pb[0]) { copy_xplane(s, F[1], low ? 1 : nx, low ? nx + 1 : 0); copy_xplane(s, F[2], low ? 1 : nx, low ? nx + 1 : 0); }
    if (pb[1]) { copy_yplane(s, F[0], low ? 1 : ny, low ? ny + 1 : 0); copy_yplane(s, F[2], low ? 1 : ny, low ? ny + 1 : 0); }
    if (pb[2]) { copy_zplane(s, F[0], low ? 1 : nz, low ? nz + 1 : 0); copy_zplane(s, F[1], low ? 1 : nz, low ? nz + 1 : 0); }
  }
}

/* ---- monitors ------------------------------------------------------------- */
static double* field_ptr(osim* s, int comp) {
  return comp < 3 ? s->E[comp] : (comp < 6 ? s->H[comp - 3] : s->D[comp - 6]);
}

static void monitors(osim* s, int n) {
  double t = (double)(n - 1) * s->dt;
  for (int d = 0; d < s->ndfts; ++d) {
    odft* df = &s->dfts[d];
    if (n < df->n_begin) continue;
    if (df->n_end > 0 && n > df->n_end) continue;
    int a1 = df->axis == 0 ? 1 : 0, a2 = df->axis == 2 ? 1 : 2;
    for (int c = 0; c < 6; ++c) {
      if (!(df->mask & (1 << c))) continue;
      double* F = field_ptr(s, c);
      /* sim3/program.f90:456-472 */
#pragma omp parallel for schedule(static) if (s->n3 > 65536)
      for (int m = 0; m < df->nfreq; ++m) {
        double cw = cos(df->omega[m] * t), sw = sin(df->omega[m] * t);
        for (int v = 1; v <= df->n2; ++v)
          for (int u = 1; u <= df->n1; ++u) {
            int p[3];
            p[df->axis] = df->index; p[a1] = u; p[a2] = v;
            double f = F[IDX(s, p[0], p[1], p[2])];
            long o = (long)(u - 1) + (long)df->n1 * ((long)(v - 1) + (long)df->n2 * m);
            df->re[c][o] = df->re[c][o] + cw * f * s->dt;
            df->im[c][o] = df->im[c][o] + sw * f * s->dt;
          }
      }
    }
  }
  for (int p = 0; p < s->nprobes; ++p) {
    oprobe* pr = &s->probes[p];
    if (pr->nrec == pr->cap) {
      pr->cap = pr->cap ? 2 * pr->cap : 1024;
      pr->rec = (double*)realloc(pr->rec, sizeof(double) * 7 * (size_t)pr->cap);
    }
    double* r = pr->rec + 7 * (long)pr->nrec++;
    r[0] = (double)n * s->dt; /* sim3/program.f90:445 writes n*dt */
    if (pr->line >= 0) { /* sim1/program.f90:272-273: the incident line */
      const otfsf* t = &s->tfsfs[pr->line];
      for (int q = 1; q < 7; ++q) r[q] = 0.0;
      r[1 + t->p.comp] = t->e[pr->j];
      r[4 + t->hc] = t->sg * t->h[pr->j];
      continue;
    }
    long c = IDX(s, pr->i, pr->j, pr->k);
    for (int q = 0; q < 3; ++q) { r[1 + q] = s->E[q][c]; r[4 + q] = s->H[q][c]; }
  }
}

int ofdtd_step(osim* s, int nsteps) {
  if (!s || nsteps < 0) return o_fail(FDTD_ERR_ARG, "bad argument");
  if (!s->committed) { int rc = ofdtd_commit(s); if (rc) return rc; }
  int nx = s->nx, ny = s->ny, nz = s->nz;
  for (int it = 0; it < nsteps; ++it) {
    int n = s->n_done + 1;
    double t = (double)(n - 1) * s->dt; /* sim3/program.f90:384 */
    if (s->pml_mode == FDTD_PML_REF_Y) {
      int P = s->pml.npml[1];
      /* set_index: edge ranks exclude the y-PML rows (mympi_func.f90:162-169) */
      update_D(s, 1, nx, P + 1, ny - P, 1, nz, 0);
      apply_sources(s, t);
      for (int q = 0; q < s->ntfsfs; ++q) { tfsf_line_E(s, &s->tfsfs[q], t); tfsf_corr_D(s, &s->tfsfs[q]); }
      update_materials(s);
      ref_pml_E(s);
      exchange(s, s->E, 1);
      update_H(s, 1, nx, P + 1, ny - P, 1, nz, 0);
      ref_pml_H(s);
      for (int q = 0; q < s->ntfsfs; ++q) { tfsf_line_H(s, &s->tfsfs[q]); tfsf_corr_H(s, &s->tfsfs[q]); }
      exchange(s, s->H, 0);
    } else if (s->pml_mode == FDTD_PML_ALL_FACES) {
      update_D(s, 1, nx, 1, ny, 1, nz, 1);
      all_pml_E(s);
      apply_sources(s, t);
      for (int q = 0; q < s->ntfsfs; ++q) { tfsf_line_E(s, &s->tfsfs[q], t); tfsf_corr_D(s, &s->tfsfs[q]); }
      update_materials(s);
      exchange(s, s->E, 1);
      update_H(s, 1, nx, 1, ny, 1, nz, 1);
      all_pml_H(s);
      for (int q = 0; q < s->ntfsfs; ++q) { tfsf_line_H(s, &s->tfsfs[q]); tfsf_corr_H(s, &s->tfsfs[q]); }
      exchange(s, s->H, 0);
    } else {
      update_D(s, 1, nx, 1, ny, 1, nz, 0);
      apply_sources(s, t);
      for (int q = 0; q < s->ntfsfs; ++q) { tfsf_line_E(s, &s->tfsfs[q], t); tfsf_corr_D(s, &s->tfsfs[q]); }
      update_materials(s);
      exchange(s, s->E, 1);
      update_H(s, 1, nx, 1, ny, 1, nz, 0);
      for (int q = 0; q < s->ntfsfs; ++q) { tfsf_line_H(s, &s->tfsfs[q]); tfsf_corr_H(s, &s->tfsfs[q]); }
      exchange(s, s->H, 0);
    }
    monitors(s, n);
    s->n_done = n;
  }
  return 0;
}

int ofdtd_sync(osim* s) { return s ? 0 : FDTD_ERR_ARG; }
int ofdtd_get_step(osim* s, int* n) { if (!s || !n) return FDTD_ERR_ARG; *n = s->n_done; return 0; }
int ofdtd_get_dt(osim* s, double* dt) { if (!s || !dt) return FDTD_ERR_ARG; *dt = s->dt; return 0; }

int ofdtd_get_field(osim* s, int comp, double* dst) {
  if (!s || !dst || comp < 0 || comp > 8) return o_fail(FDTD_ERR_ARG, "bad argument");
  double* F = field_ptr(s, comp);
  long o = 0;
  for (int k = 1; k <= s->nz; ++k)
    for (int j = 1; j <= s->ny; ++j)
      for (int i = 1; i <= s->nx; ++i) dst[o++] = F[IDX(s, i, j, k)];
  return 0;
}

int ofdtd_set_field(osim* s, int comp, const double* src) {
  if (!s || !src || comp < 0 || comp > 8) return o_fail(FDTD_ERR_ARG, "bad argument");
  double* F = field_ptr(s, comp);
  long o = 0;
  for (int k = 1; k <= s->nz; ++k)
    for (int j = 1; j <= s->ny; ++j)
      for (int i = 1; i <= s->nx; ++i) F[IDX(s, i, j, k)] = src[o++];
  /* keep the periodic ghosts consistent with the interior */
  if (comp < 3) exchange(s, s->E, 1);
  else if (comp < 6) exchange(s, s->H, 0);
  return 0;
}

int ofdtd_get_dft(osim* s, int id, int comp, double* re, double* im) {
  if (!s || id < 0 || id >= s->ndfts || comp < 0 || comp > 5) return o_fail(FDTD_ERR_ARG, "bad argument");
  odft* d = &s->dfts[id];
  if (!(d->mask & (1 << comp))) return o_fail(FDTD_ERR_ARG, "component not monitored");
  long np = (long)d->n1 * d->n2 * d->nfreq;
  if (re) memcpy(re, d->re[comp], sizeof(double) * (size_t)np);
  if (im) memcpy(im, d->im[comp], sizeof(double) * (size_t)np);
  return 0;
}

/* sim3/program.f90:613-685 (y-normal) and its x/z-normal analogues */
int ofdtd_get_flux(osim* s, int id, double* S) {
  if (!s || !S || id < 0 || id >= s->ndfts) return o_fail(FDTD_ERR_ARG, "bad argument");
  odft* d = &s->dfts[id];
  int eA, eB, hA, hB; /* S = hA*eA - hB*eB */
  if (d->axis == 1) { hA = FDTD_HX; eA = FDTD_EZ; hB = FDTD_HZ; eB = FDTD_EX; }
  else if (d->axis == 2) { hA = FDTD_HY; eA = FDTD_EX; hB = FDTD_HX; eB = FDTD_EY; }
  else { hA = FDTD_HZ; eA = FDTD_EY; hB = FDTD_HY; eB = FDTD_EZ; }
  int need = (1 << eA) | (1 << eB) | (1 << hA) | (1 << hB);
  if ((d->mask & need) != need) return o_fail(FDTD_ERR_ARG, "flux needs the four tangential components");
  for (int m = 0; m < d->nfreq; ++m) {
    double acc = 0.0;
    for (int v = 0; v < d->n2; ++v)
      for (int u = 0; u < d->n1; ++u) {
        long o = (long)u + (long)d->n1 * ((long)v + (long)d->n2 * m);
        acc = acc + ((d->re[hA][o] * d->re[eA][o]) + (d->im[hA][o] * d->im[eA][o]) -
                     (d->re[hB][o] * d->re[eB][o]) - (d->im[hB][o] * d->im[eB][o])) * s->ds;
      }
    S[m] = acc;
  }
  return 0;
}

int ofdtd_get_probe_count(osim* s, int id, int* nrec) {
  if (!s || !nrec || id < 0 || id >= s->nprobes) return o_fail(FDTD_ERR_ARG, "bad argument");
  *nrec = s->probes[id].nrec;
  return 0;
}

int ofdtd_get_probe(osim* s, int id, double* out) {
  if (!s || !out || id < 0 || id >= s->nprobes) return o_fail(FDTD_ERR_ARG, "bad argument");
  memcpy(out, s->probes[id].rec, sizeof(double) * 7 * (size_t)s->probes[id].nrec);
  return 0;
}

int ofdtd_energy(osim* s, double* we, double* wh) {
  if (!s || !we || !wh) return o_fail(FDTD_ERR_ARG, "bad argument");
  if (!s->committed) { int rc = ofdtd_commit(s); if (rc) return rc; }
  double a = 0, b = 0;
  for (int k = 1; k <= s->nz; ++k)
    for (int j = 1; j <= s->ny; ++j)
      for (int i = 1; i <= s->nx; ++i) {
        long c = IDX(s, i, j, k);
        a += s->eps_cell[c] * (s->E[0][c] * s->E[0][c] + s->E[1][c] * s->E[1][c] + s->E[2][c] * s->E[2][c]);
        b += s->H[0][c] * s->H[0][c] + s->H[1][c] * s->H[1][c] + s->H[2][c] * s->H[2][c];
      }
  *we = a;
  *wh = b;
  return 0;
}

/* Closed-form Drude+2CP permittivity, misc/my_matfunctions_mod.f90:8-31
 * (gold_ep), parameterised.  Used by the tests as the analytic pin of the
 * ADE.  Returns re/im of eps(w). */
void ofdtd_drude2cp_eps(const fdtd_drude2cp* p, double w, double* re, double* im) {
  double complex crit = 0;
  for (int q = 0; q < 2; ++q)
    crit = crit + p->A[q] * p->Omega[q] *
                      ((cexp(I * p->phi[q]) / (p->Omega[q] - w - I * p->Gamma[q])) +
                       (cexp(-I * p->phi[q]) / (p->Omega[q] + w + I * p->Gamma[q])));
  double complex e = p->eps_inf - ((p->omega_d * p->omega_d) / (w * w + I * p->gamma * w)) + crit;
  *re = creal(e);
  *im = cimag(e);
}

/* Physical constants as the oracle computes them (misc/constants_mod.f90). */
void ofdtd_constants(double* out /* pi, ep0, mu0, c0 */) {
  out[0] = O_PI; out[1] = O_EP0; out[2] = O_MU0; out[3] = O_C0;
}
