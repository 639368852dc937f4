/*
 * fdtd.h -- C ABI of the MI355X (gfx950) 3D Yee FDTD time-stepper.
 *
 * The library replaces the 3D hot path of lsmayoral/Fortran-MPI-FDTD:
 * the Yee leapfrog, the split-field PML, the Drude+2CP gold ADE, the graphene
 * CP/Raman ADE, soft plane sources, running-DFT monitors and flux integrals.
 * Every entry point below names the reference routine(s) it stands in for
 * (paths relative to the reference checkout):
 *
 *   sim3/  = FDTDsimulations/3D/air-AuAntennasSiO2-Si/
 *   sim2/  = FDTDsimulations/2D/air-nanoAntennas_Si02-Sisubstrate/
 *   sim1/  = FDTDsimulations/1D/TFSF-air_SLG_SiO2_Sidisp/
 *   mod3/  = modules/src/3D/
 *   mod1/  = modules/src/1D/
 *   misc/  = modules/src/misc/
 *
 * Conventions shared with the reference
 *   - Grid indices are 1-based and inclusive, global (the reference's
 *     xinit..zend, mod3/material_mod.f90:8).  Interior cells are 1..n per axis.
 *   - Field arrays crossing the ABI are Fortran order (x fastest), interior
 *     cells only, FP64, caller-owned host buffers.
 *   - Units are the reference's normalised ones: D and E share one scale,
 *     H is scaled by the vacuum impedance, Ku = c0*dt/ds multiplies every curl
 *     (sim3/include/myfdtd_func.f90:32-34,58-60).
 *   - Materials are applied in call order; where regions overlap, the last
 *     one added writes E (sim3/program.f90:411-414).
 *
 * All functions return 0 on success or a negative FDTD_ERR_* code; the text
 * of the most recent error on the calling thread is fdtd_last_error().
 * A handle owns all its device memory and one HIP stream; it is not
 * thread-safe, but separate handles may be driven from separate threads.
 */
#ifndef FDTD_AMD_FDTD_H
#define FDTD_AMD_FDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum {
  FDTD_OK = 0,
  FDTD_ERR_ARG = -1,         /* invalid argument                           */
  FDTD_ERR_STATE = -2,       /* call not valid in the handle's state       */
  FDTD_ERR_NOMEM = -3,       /* device or host allocation failed           */
  FDTD_ERR_HIP = -4,         /* HIP runtime error                          */
  FDTD_ERR_UNSUPPORTED = -5, /* combination not supported by this path     */
  FDTD_ERR_NONFINITE = -6    /* a served job's fields left the finite range */
};

/* ---- enumerations ---------------------------------------------------- */
enum { FDTD_FP64 = 0, FDTD_FP32 = 1 };

/* PML modes.
 * REF_Y     : the reference's Berenger split-field PML on the two y faces,
 *             with x/z grading inside the y slabs and interface coupling by
 *             copy (mod3/BERpmly_mod.f90), multi-rank semantics (the y-low
 *             slab's default f2 uses eps_low, the y-high slab's eps_high).
 * ALL_FACES : the same split-field formulation and profiles on every face
 *             whose npml > 0; main fields hold the split sums, splits are
 *             driven by the main (total) fields of the other kind.          */
enum { FDTD_PML_NONE = 0, FDTD_PML_REF_Y = 1, FDTD_PML_ALL_FACES = 2 };

/* Field components for get/set. */
enum {
  FDTD_EX = 0, FDTD_EY = 1, FDTD_EZ = 2,
  FDTD_HX = 3, FDTD_HY = 4, FDTD_HZ = 5,
  FDTD_DX = 6, FDTD_DY = 7, FDTD_DZ = 8
};

/* Region shapes used to place materials. */
enum { FDTD_REGION_BOX = 0, FDTD_REGION_SPHERE = 1, FDTD_REGION_CYLINDER = 2 };

/* Graphene models. */
enum {
  FDTD_GRAPHENE_CP4 = 0,  /* reference: Drude + 4 critical points, thickness-scaled */
  FDTD_GRAPHENE_KUBO = 1  /* Drude weight from E_F, tau, T (Kubo intraband)        */
};

/* Stepping engines (fdtd_grid.engine). */
enum {
  FDTD_ENGINE_AUTO = 0,   /* STAGED for REF_Y, FUSED otherwise                  */
  FDTD_ENGINE_STAGED = 1, /* one kernel per reference stage, D kept everywhere  */
  FDTD_ENGINE_FUSED = 2   /* fused E / H kernels, D kept only in ADE cells      */
};

/* Kernel ids for fdtd_kernel_stats. */
enum {
  FDTD_K_E = 0,       /* E half step (fused), updateD (staged), or the
                         one-pass step over its plain tiles            */
  FDTD_K_H = 1,       /* H half step (fused) or updateH (staged)   */
  FDTD_K_ADE = 2,     /* dispersive cells (gold, graphene)         */
  FDTD_K_PML_E = 3,   /* staged split-field E                       */
  FDTD_K_PML_H = 4,   /* staged split-field H                       */
  FDTD_K_MAT = 5,     /* staged dielectric E = D*gax + source       */
  FDTD_K_MON = 6,     /* DFT monitors and probes                    */
  FDTD_K_GEN_E = 7,   /* one-pass step: E of its general tiles      */
  FDTD_K_GEN_H = 8,   /* one-pass step: H of its general tiles      */
  FDTD_K_XF = 9,      /* one-pass step (FP32): its plain tiles that reach
                         into an x PML slab (x-face tiles)           */
  FDTD_K_COUNT = 10
};

typedef struct fdtd_sim fdtd_sim;

/* Grid: sim3/program.f90:44-72 (ds, Ku, nx/ny/nz, PBC_x/y/z). */
typedef struct {
  int32_t nx, ny, nz;     /* interior cells                                   */
  int32_t precision;      /* FDTD_FP64 | FDTD_FP32                            */
  int32_t periodic[3];    /* PBC_x, PBC_y, PBC_z (sim3/program.f90:12)        */
  int32_t engine;         /* FDTD_ENGINE_*                                    */
  double ds;              /* cell size [m]                                    */
  double courant;         /* Ku; dt = Ku*ds/c0 (sim3/program.f90:48-50)      */
} fdtd_grid;

/* PML: sim3/program.f90:97-102 and pmly%initialize (mod3/BERpmly_mod.f90:485). */
typedef struct {
  int32_t mode;           /* FDTD_PML_*                                        */
  int32_t npml[3];        /* cells per axis; REF_Y: x/z grading depth in slabs */
  double order;           /* polynomial grading order m                        */
  double R0;              /* design reflection; used when sigmam <= 0          */
  double sigmam;          /* > 0: peak conductivity factor used as given       */
  double eps_low;         /* REF_Y: relative eps at the y-low slab             */
  double eps_high;        /* REF_Y: relative eps at the y-high slab            */
} fdtd_pml;

/* Region: the box semantics of material_mod (1-based inclusive global
 * indices, clipped to the grid) plus staircased spheres and cylinders in
 * cell-index coordinates (cell (i,j,k) has its centre at (i,j,k)).          */
typedef struct {
  int32_t kind;           /* FDTD_REGION_*                                     */
  int32_t axis;           /* CYLINDER axis 0..2                                */
  int32_t lo[3];          /* BOX lower corner (also clips SPHERE/CYLINDER)     */
  int32_t hi[3];          /* BOX upper corner                                  */
  double center[3];       /* SPHERE/CYLINDER centre                            */
  double radius;          /* SPHERE/CYLINDER radius [cells]                    */
  double half_length;     /* CYLINDER half length along axis [cells]           */
} fdtd_region;

/* Drude + 2 critical points (mod3/gold_mod.f90:28-34). */
typedef struct {
  double eps_inf;         /* epinf                                             */
  double gamma;           /* dnur  [1/s]                                       */
  double omega_d;         /* wD    [rad/s]                                     */
  double A[2];            /* ap                                                */
  double phi[2];          /* phip  [rad]                                       */
  double Omega[2];        /* omp   [rad/s]                                     */
  double Gamma[2];        /* rp    [1/s]                                       */
} fdtd_drude2cp;

/* Graphene sheet (mod3/graphene_mod.f90:9-55). */
typedef struct {
  int32_t model;          /* FDTD_GRAPHENE_*                                   */
  int32_t normal_axis;    /* sheet normal; that E component is not written     */
  int32_t ref_init_order; /* CP4: 1 keeps the reference's init order, in which
                             the CP coefficients are read before they are set
                             (mod3/graphene_mod.f90:222-231): gb = gc = 0      */
  int32_t reserved;
  double thickness;       /* CP4: physical thickness scaled into ds (3.4e-10)  */
  double eps_inf;         /* graphene_epinf                                    */
  double gamma;           /* graphene_dnur [1/s] (CP4)                         */
  double omega_d;         /* graphene_wD [rad/s] (CP4)                         */
  double A[4], phi[4], Omega[4], Gamma[4];   /* CP4 critical points          */
  double fermi_ev;        /* KUBO: Fermi level [eV]                            */
  double tau;             /* KUBO: relaxation time [s]                         */
  double temperature;     /* KUBO: [K]                                         */
  double wNL;             /* Raman resonance [rad/s]                           */
  double dnuNL;           /* Raman linewidth [1/s]                             */
  double xi3;             /* Raman strength                                    */
} fdtd_graphene;

/* Lorentz poles in polarisation-current form (the Lorentz-Jp family:
 * mod3/goldLorentzJp_mod.f90:37-91, siLorentzJp_mod.f90, titaniumJp_mod.f90,
 * filmGoldLorentzJp_mod.f90, grapheneLorentzJp_mod.f90,
 * graphene3LorentzJp1nm_mod.f90, NodispersiveGraphene.f90).  For every
 * component q in jp_mask, with Ku' = dt*c0/ds and ga = 1/eps_inf:
 *   E_q += Ku'*ga*(curlH_q - ds*JpT_q)          (JpT_q = sum_l Jp_q,l)
 *   per pole l: P0 = xi_l*E_q/c0 (+ xi3*Q_q*E_q/c0 for l = 1 when nonlinear)
 *               Jp_q,l = g1_l*Jp_q,l + g2_l*(P0 - P_q,l);  P_q,l += dt*Jp_q,l
 *   g1 = (1 - dnu*dt)/(1 + dnu*dt),  g2 = w0^2*dt/(1 + dnu*dt)
 *   nonlinear: G_q = g1NL*G_q + g2NL*((z0*E_q)^2 - Q_q);  Q_q += dt*G_q
 * Components outside jp_mask are written as vacuum (E_q += Ku*curlH_q), as
 * the graphene modules write Ey.  No reference driver calls update_EJp:
 * overlapping regions follow the library's rule (last writer wins per
 * component) instead of adding both updates.                                */
typedef struct {
  int32_t npols;          /* 1..4                                              */
  int32_t jp_mask;        /* bit q: E_q carries the poles (7 = all)            */
  int32_t nonlinear;      /* 1: graphene third-order term (wNL, dnuNL, xi3)    */
  int32_t ez_uses_jx;     /* 1: E_z is driven by JpT_x, as
                             mod3/grapheneLorentzJp_mod.f90:67-68 does         */
  double eps_inf;         /* epinf                                             */
  double xi[4];           /* xi1 (pole strengths)                              */
  double w0[4];           /* w0l [rad/s]                                       */
  double dnu[4];          /* dnu0l [rad/s]                                     */
  double wNL, dnuNL, xi3; /* nonlinear oscillator (set by the caller)          */
} fdtd_lorentz_jp;

enum {
  FDTD_LJP_GOLD = 0,          /* mod3/goldLorentzJp_mod.f90:22-26            */
  FDTD_LJP_SI = 1,            /* mod3/siLorentzJp_mod.f90:22-26              */
  FDTD_LJP_TITANIUM = 2,      /* mod3/titaniumJp_mod.f90:22-26               */
  FDTD_LJP_FILM_GOLD = 3,     /* mod3/filmGoldLorentzJp_mod.f90:21-25        */
  FDTD_LJP_GRAPHENE = 4,      /* mod3/grapheneLorentzJp_mod.f90:27-31        */
  FDTD_LJP_GRAPHENE3_1NM = 5, /* mod3/graphene3LorentzJp1nm_mod.f90:27-31    */
  FDTD_LJP_GRAPHENE_ND = 6    /* mod3/NodispersiveGraphene.f90               */
};

/* Gaussian-modulated soft source, sim3/program.f90:107-111,400:
 * g(t) = amplitude * exp(-0.5*((t-t0)/st)^2) * sin(w0*t), t = (n-1)*dt.     */
typedef struct {
  double t0, st, w0, amplitude;
} fdtd_pulse;

/* ---- parameter presets ------------------------------------------------ */
/* Gold of mod3/gold_mod.f90:28-34 (Vial & Laroche, Appl Phys A 103:849). */
void fdtd_drude2cp_gold(fdtd_drude2cp* p);
/* Graphene CP4 + Raman defaults of mod3/graphene_mod.f90:33-41 (the Raman
 * triple wNL, dnuNL, xi3 is left zero: the reference driver sets it).      */
void fdtd_graphene_ref(fdtd_graphene* p);
/* Pole tables of the Lorentz-Jp modules (FDTD_LJP_*; w0 and dnu converted
 * with 2*pi*c0/micro as the modules do).  The graphene presets set
 * nonlinear = 1 and leave wNL, dnuNL, xi3 zero (public fields the driver
 * sets); FDTD_LJP_GRAPHENE sets ez_uses_jx = 1.  Returns 0, or
 * FDTD_ERR_ARG for an unknown kind.                                         */
int fdtd_lorentz_jp_preset(int32_t kind, fdtd_lorentz_jp* p);

/* ---- lifetime ---------------------------------------------------------- */
/* Replaces the static field arrays and grid parameters of
 * sim3/program.f90:44-77,209-214 and the rank layout of
 * sim3/include/mympi_func.f90:137-214 (one domain, no ranks).              */
int fdtd_create(fdtd_sim** out, const fdtd_grid* grid);
int fdtd_destroy(fdtd_sim* sim);
const char* fdtd_last_error(void);

/* mypmly%initialize, mod3/BERpmly_mod.f90:485-508 (coefficients of
 * :612-874, sigmam of sim3/program.f90:102).                               */
int fdtd_set_pml(fdtd_sim* sim, const fdtd_pml* pml);

/* ConstantDielectric%initialize_material + update_E,
 * mod3/simpledielectric_mod.f90:22-103 (eps_r = 0 gives gax = 0, :91-97). */
int fdtd_add_dielectric(fdtd_sim* sim, const fdtd_region* region, double eps_r);

/* Gold%initialize_material + update_E, mod3/gold_mod.f90:50-195. */
int fdtd_add_drude2cp(fdtd_sim* sim, const fdtd_region* region,
                      const fdtd_drude2cp* params);

/* <material>_LJp%initialize_material + update_EJp of the Lorentz-Jp
 * family (mod3/goldLorentzJp_mod.f90:37-163 and its siblings).  The update
 * reads curl H directly (E form), so D is not used in these cells.          */
int fdtd_add_lorentz_jp(fdtd_sim* sim, const fdtd_region* region,
                        const fdtd_lorentz_jp* params);

/* Si%initialize_material + update_E, mod3/si_mod.f90:46-82,234-259: Lorentz
 * poles in D form, a second-order recursion per pole and component:
 *   E_q = D_q - S_1 - ... - S_npols             (the S of the previous step)
 *   S_l <- ga_l*S_l + gb_l*S_l(one step older) + gc_l*E_q
 *   ga = (2 - (dt*w0)^2)/(1 + dt*dnu), gb = (dt*dnu - 1)/(1 + dt*dnu),
 *   gc = xi*(w0*dt)^2/(1 + dt*dnu)
 * over the pole table of an fdtd_lorentz_jp (FDTD_LJP_SI holds si_mod's
 * table, :36-38: the same poles its E-form sibling siLorentzJp_mod uses).
 * The module has no eps_inf and no nonlinear term: eps_inf must be 1,
 * nonlinear 0 and jp_mask 7 (all three components).                        */
int fdtd_add_lorentz_d(fdtd_sim* sim, const fdtd_region* region,
                       const fdtd_lorentz_jp* params);

/* Graphene%initialize_material + update_E, mod3/graphene_mod.f90:59-283. */
int fdtd_add_graphene(fdtd_sim* sim, const fdtd_region* region,
                      const fdtd_graphene* params);

/* Plane source, sim3/program.f90:396-406: D_comp(plane) += g(t) after the
 * D update and before E = f(D).  axis = plane normal, index 1-based.       */
int fdtd_add_plane_source(fdtd_sim* sim, int axis, int index, int comp,
                          const fdtd_pulse* pulse);

/* Total-field/scattered-field plane wave at normal incidence,
 * sim1/program.f90:86-91,217-262 (the one-sided form; the two-sided form
 * follows the dormant sim3/include/myTFSF_incEzHx_3Dfunc.f90 and
 * sim2/include/myTFSF_2Dfunc.f90).  An auxiliary 1D line along `axis`
 * (cells 1..n_axis, eps uniform, its own split-free PML of npml cells at
 * both ends as mod1/BERpmly_ExHz_mod.f90, graded with order and R0) carries
 * the incident wave: e(j) += Ku/eps*(h(j) - h(j-1)), the pulse added to
 * e(src_index), h(j) += Ku*(e(j+1) - e(j)).  Planes lo..hi (hi = 0: to the
 * end of the grid) along `axis` are total field, the rest scattered field:
 *   after the D/E update   D_comp(lo)   -= Ku*h(lo-1),  D_comp(hi+1) += Ku*h(hi)
 *   after the H update     H_h(lo-1)    -= s*Ku*e(lo),  H_h(hi) += s*Ku*e(hi+1)
 * where h is the third axis and s = +1 when (comp, axis) is (x,y), (y,z) or
 * (z,x), else -1 (the incident H is s*h).  The correction planes lie outside
 * any PML along `axis`; transverse ALL_FACES PML cells take them into the
 * split along `axis`.  With FDTD_PML_REF_Y the axis must be y.  Each TFSF
 * uses two of the 8 source slots (one when one-sided).                     */
typedef struct {
  int32_t axis;           /* propagation direction +axis                        */
  int32_t comp;           /* E component of the incident wave (!= axis)         */
  int32_t lo;             /* first total-field plane (TFSF_bd, :87)             */
  int32_t hi;             /* last total-field plane; 0: one-sided               */
  int32_t src_index;      /* line cell receiving the pulse (spos, :86)          */
  int32_t npml;           /* line PML cells per end (npml, :37)                 */
  double eps;             /* relative permittivity of the line (air_ep, :134)   */
  double order;           /* line PML grading order (:130)                      */
  double R0;              /* line PML design reflection (:131)                  */
  fdtd_pulse pulse;       /* e(src_index) += g(t), t = (n-1)*dt (:228-229)      */
} fdtd_tfsf;

/* Adds a TFSF plane wave (id: 0, 1, ...).                                   */
int fdtd_add_tfsf(fdtd_sim* sim, const fdtd_tfsf* tfsf, int* id);

/* Probe of the incident line at cell `index`, sim1/program.f90:272-273: a
 * point probe (read with fdtd_get_probe) whose records hold (n*dt, the
 * incident E in its component's slot, the incident H in its slot, zeros). */
int fdtd_add_tfsf_probe(fdtd_sim* sim, int tfsf_id, int index, int* id);

/* The incident line after the last step: e and h, n_axis + 2 doubles each
 * (cells 0..n_axis+1).                                                       */
int fdtd_get_tfsf_line(fdtd_sim* sim, int tfsf_id, double* e, double* h);

/* Running DFT on a plane, sim3/program.f90:441-540: at step n with
 * n_begin <= n <= n_end (n_end <= 0: no end), for every selected component
 * F: re += cos(omega*t)*F*dt, im += sin(omega*t)*F*dt, t = (n-1)*dt,
 * sampled after the H update.  comp_mask bit c selects component c
 * (FDTD_EX..FDTD_HZ).  Accumulators are FP64.                              */
int fdtd_add_dft_plane(fdtd_sim* sim, int axis, int index, int nfreq,
                       const double* omega, int comp_mask, int n_begin,
                       int n_end, int* id);

/* Point probe, sim3/program.f90:445-451: one record per step after the H
 * update: (n*dt, Ex, Ey, Ez, Hx, Hy, Hz).                                  */
int fdtd_add_probe(fdtd_sim* sim, int i, int j, int k, int* id);

/* Optional: freeze the set-up and allocate device state now (otherwise done
 * by the first call that needs it).                                        */
int fdtd_commit(fdtd_sim* sim);

/* Optional, no reference counterpart: pick the FUSED engine's step form by
 * timing `steps` steps of each candidate on this handle -- the two-kernel
 * step with tiles of kc = 1, 2, 4 or 8 planes (each thread marching its
 * column) and the one-pass step, each wavefront marching a tile of 64 x by
 * cells through kc planes (by in {2, 4, 8}, FP64 {2, 4}; kc in {16, 32,
 * 64, 128}).  The timed steps run without sources on the untouched (zero)
 * fields and leave them zero; they sample the DFT planes as real steps do
 * (into a buffer slot the first real step overwrites) but fold nothing and
 * record no probes.  Call it after set-up and before the first fdtd_step
 * (and before any fdtd_set_field).  *by = 0 reports the two-kernel step
 * (*kc its planes per tile).  Other engines: returns 0 with *by = *kc = 0. */
int fdtd_autotune(fdtd_sim* sim, int steps, int* by, int* kc);

/* ---- time loop ----------------------------------------------------------- */
/* nsteps iterations of the body of sim3/program.f90:382-540:
 * updateD, source, material update_E, pmly%updateE, updateH, pmly%updateH,
 * monitors; the halo exchanges become periodic index wrap on one domain.   */
int fdtd_step(fdtd_sim* sim, int nsteps);
int fdtd_sync(fdtd_sim* sim);

/* Checkpoint and restart (no reference counterpart: a reference run
 * restarts from zero fields).  The state after the steps done so far --
 * step count, E and H (and D with the STAGED engine), the PML splits, every
 * dispersive group's state, the DFT accumulators, the probe records and
 * the TFSF lines -- goes to a host buffer of fdtd_checkpoint_size bytes
 * (committed handles; waits for the handle's queued work).
 * fdtd_checkpoint_load restores it into a handle set up with the same
 * grid, precision, engine, PML, materials, monitors, probes and TFSF waves,
 * before that handle's first step (committing it if needed); stepping on
 * gives the same bits as the uninterrupted run, in either FUSED step form.
 * Mismatched handles or buffers are refused (FDTD_ERR_ARG, FDTD_ERR_STATE
 * after a step).                                                            */
int fdtd_checkpoint_size(fdtd_sim* sim, size_t* bytes);
int fdtd_checkpoint_save(fdtd_sim* sim, void* buf, size_t bytes);
int fdtd_checkpoint_load(fdtd_sim* sim, const void* buf, size_t bytes);
int fdtd_get_step(fdtd_sim* sim, int* n_done);
int fdtd_get_dt(fdtd_sim* sim, double* dt);

/* ---- results ------------------------------------------------------------- */
/* comp FDTD_EX..FDTD_DZ; dst holds nx*ny*nz doubles, x fastest.  D is
 * available with the STAGED engine only.                                   */
int fdtd_get_field(fdtd_sim* sim, int comp, double* dst);
int fdtd_set_field(fdtd_sim* sim, int comp, const double* src);

/* DFT accumulators of plane `id`, component comp: n1*n2*nfreq doubles each,
 * in-plane index order (x-normal: j,k; y-normal: i,k; z-normal: i,j),
 * frequency slowest (the reference's real_exin(i,k,m), :127-141).          */
int fdtd_get_dft(fdtd_sim* sim, int id, int comp, double* re, double* im);

/* Flux through DFT plane `id`, sim3/program.f90:613-685, generalised to
 * the plane normal: S(m) = ds * sum (Re(Et)Re(Ht') ... ) with
 * y-normal: Hx*Ez - Hz*Ex; z-normal: Hy*Ex - Hx*Ey; x-normal: Hz*Ey - Hy*Ez
 * (real*real + imag*imag).  Needs the four tangential components.        */
int fdtd_get_flux(fdtd_sim* sim, int id, double* S);

/* No counterpart in the reference: fdtd_get_flux for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The per-frequency sums go to a
 * fixed device array of the library (FDTD_FLUX_NOALLOC_MAX_FREQ doubles,
 * 8 KiB, one per GPU, held until the process ends) instead of a buffer the
 * handle allocates; planes with more frequencies are refused
 * (FDTD_ERR_ARG), and the handle must be committed (FDTD_ERR_STATE
 * otherwise: committing allocates).  Every call on one GPU shares that
 * array, so calls that overlap on one GPU race: serialize them.  The
 * result equals fdtd_get_flux's bit for bit.                              */
#define FDTD_FLUX_NOALLOC_MAX_FREQ 1024
int fdtd_get_flux_noalloc(fdtd_sim* sim, int id, double* S);

int fdtd_get_probe_count(fdtd_sim* sim, int id, int* nrec);
int fdtd_get_probe(fdtd_sim* sim, int id, double* out /* nrec*7 */);

/* Field energy sums over interior cells: we = sum eps_r*(Ex^2+Ey^2+Ez^2)
 * (eps_inf for dispersive cells, 1 where no material), wh = sum |H|^2.
 * New: the reference computes flux only.                                   */
int fdtd_energy(fdtd_sim* sim, double* we, double* wh);

/* ---- introspection -------------------------------------------------------- */
int fdtd_memory_bytes(fdtd_sim* sim, size_t* bytes);
/* HIP-event timing of every kernel launch while enabled (adds events to the
 * handle's stream; off by default).                                        */
int fdtd_set_profiling(fdtd_sim* sim, int on);
int fdtd_kernel_stats(fdtd_sim* sim, int kernel, double* total_ms,
                      int64_t* launches);
/* Bytes moved per step as counted by the algorithm (fields, ADE state, PML
 * splits; monitors excluded) and the cell mix it was computed from.        */
int fdtd_algorithmic_bytes(fdtd_sim* sim, int kernel, double* bytes_per_launch);
int fdtd_cell_counts(fdtd_sim* sim, int64_t* n_pml, int64_t* n_ade);
/* Device bytes the handle holds (committed) or would allocate at commit. */
int fdtd_estimate_bytes(fdtd_sim* sim, size_t* bytes);

/* ---- serving: many independent simulations on one GPU ---------------------
 * No counterpart in the reference (it runs one simulation per mpirun).
 * A pool admits set-up handles in arrival order while their device bytes fit
 * the HBM budget and fewer than max_active run; every admitted handle runs
 * on its own stream and advances chunk_steps at a time (rounded up to a
 * multiple of 64, the DFT fold block, so that a job's spectra equal its run
 * alone bit for bit), one chunk queued per handle per fdtd_pool_progress
 * call, so co-resident jobs overlap on the device.  After every chunk the pool checks the job's E and H for NaN or
 * Inf (one read of the fields per chunk, on the job's stream): a job whose
 * fields are no longer finite (an unstable time step, a diverging
 * material) ends FDTD_JOB_FAILED and fdtd_pool_status returns
 * FDTD_ERR_NONFINITE for it; the other jobs are not affected.  A finished
 * job keeps its memory until fdtd_pool_release; the caller owns the handles
 * (read results, then release, then destroy).                              */
typedef struct fdtd_pool fdtd_pool;
enum {
  FDTD_JOB_WAITING = 0, FDTD_JOB_RUNNING = 1, FDTD_JOB_DONE = 2,
  FDTD_JOB_FAILED = 3, FDTD_JOB_CANCELLED = 4
};
int fdtd_pool_create(fdtd_pool** out, size_t hbm_budget, int max_active, int chunk_steps);
int fdtd_pool_destroy(fdtd_pool* pool);
/* FDTD_ERR_NOMEM when the job alone exceeds the budget. */
int fdtd_pool_submit(fdtd_pool* pool, fdtd_sim* sim, int nsteps, int* ticket);
/* One non-blocking scheduling pass: retire finished chunks, queue the next
 * chunk of every running job, admit waiting jobs. */
int fdtd_pool_progress(fdtd_pool* pool, int* n_running, int* n_waiting);
int fdtd_pool_status(fdtd_pool* pool, int ticket, int* state, double* wait_ms, double* run_ms);
/* Waiting: never runs.  Running: stops after its in-flight chunk. */
int fdtd_pool_cancel(fdtd_pool* pool, int ticket);
/* Frees the job's budget share and its chunk events; afterwards its handle
 * may be destroyed.  Destroy handles only after releasing their jobs (or
 * after fdtd_pool_destroy).                                                */
int fdtd_pool_release(fdtd_pool* pool, int ticket);
int fdtd_pool_usage(fdtd_pool* pool, size_t* used, size_t* budget);

#ifdef __cplusplus
}
#endif
#endif /* FDTD_AMD_FDTD_H */
