/*
 * fdtd2d.h -- C ABI of the MI355X (gfx950) 2D Yee stepper in TEz (Hz, Ex,
 * Ey) or TMz (Ez, Hx, Hy) polarisation: single simulations, and batches of
 * many independent simulations of either polarisation stepped together (one
 * launch per half step over all of them, the step loop replayed from a
 * captured HIP graph).
 *
 * The library replaces the 2D driver of lsmayoral/Fortran-MPI-FDTD and the
 * module API it calls (paths relative to the reference checkout):
 *
 *   sim2/  = FDTDsimulations/2D/air-nanoAntennas_Si02-Sisubstrate/
 *   mod2/  = modules/src/2D/
 *
 * Conventions (shared with fdtd.h, whose status codes, precision enum and
 * material/pulse structs this header reuses)
 *   - Indices are 1-based, inclusive and global: interior cells are i=1..nx,
 *     j=1..ny (sim2/program.f90:67, arrays (0:N_x+1,0:N_y+1)).
 *   - Normalised units: D and E share one scale, Hz is scaled by the vacuum
 *     impedance, Ku = c0*dt/ds multiplies every curl
 *     (sim2/include/myfdtd_2Dfunc.f90:16-44).
 *   - The main fields are updated on rows j = npml_y .. ny-npml_y; the y
 *     slabs below and above hold the reference's split-field PML
 *     (mod2/BERpmly_HzExEy_mod.f90) and hand Ex (rows 1..npml_y and
 *     ny-npml_y+1..ny) and Hz (rows 1..npml_y-1 and ny-npml_y+1..ny) to the
 *     main fields by copy (:144-191).  With periodic_x the main fields wrap
 *     in x through their ghost columns (sim2/include/mympi_2Dfunc.f90,
 *     interchange_E/H with PBC_x); the PML's split fields never wrap.
 *   - Materials are applied in call order; where regions overlap the last
 *     one added writes E (sim2/program.f90:345-348).
 *
 * TMz (no reference counterpart: the reference's 2D driver is TEz only) is
 * the 3D path's z-invariant reduction: a grid with nz = 1 and periodic z
 * stepped by the 3D FUSED engine (include/fdtd.h) carries Ez, Hx, Hy with
 * exactly these updates:
 *   Dz += Ku*((Hy(i) - Hy(i-1)) - (Hx(j) - Hx(j-1))), Ez = Dz/eps (or the
 *   dispersive update), Hx -= Ku*(Ez(j+1) - Ez(j)), Hy += Ku*(Ez(i+1) - Ez(i)),
 * the line source adds to Dz, and the PML is the 3D engine's split-field
 * PML on all faces (FDTD_PML_ALL_FACES): slabs of npml[0] cells at both x
 * ends (non-periodic x only) and npml[1] cells at both y ends, Ez split into
 * Ezx + Ezy, Hx graded along y and Hy along x, vacuum grading, f2 scaled by
 * 1/eps of the last dielectric written to the cell; eps_low / eps_high are
 * not used.  Gold (Drude+2CP) and graphene act on Ez; graphene takes the 3D
 * module's coefficients (exact-exponential Raman recursion,
 * mod3/graphene_mod.f90:216-218,256-258; ref_init_order honoured).
 *
 * All functions return 0 or a negative FDTD_ERR_* code; fdtd_last_error()
 * gives the text of the last error on the calling thread.
 */
#ifndef FDTD_AMD_FDTD2D_H
#define FDTD_AMD_FDTD2D_H

#include "fdtd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* polarisation of a simulation */
enum { FDTD2D_TEZ = 0, FDTD2D_TMZ = 1 };
/* field components crossing the ABI: TEz, then TMz */
enum { FDTD2D_EX = 0, FDTD2D_EY = 1, FDTD2D_HZ = 2, FDTD2D_DX = 3, FDTD2D_DY = 4,
       FDTD2D_EZ = 5, FDTD2D_HX = 6, FDTD2D_HY = 7, FDTD2D_DZ = 8 };
/* region kinds: box (the reference's xinit..yend) and disk (cell centres
 * within radius of center, a cylinder seen end-on)                       */
enum { FDTD2D_REGION_BOX = 0, FDTD2D_REGION_DISK = 1 };
/* DFT line component mask bits (TEz: Ex, Hz; TMz: Ez, Hx) */
enum { FDTD2D_DFT_EX = 1, FDTD2D_DFT_HZ = 4, FDTD2D_DFT_EZ = 8, FDTD2D_DFT_HX = 16 };

typedef struct fdtd2d_sim fdtd2d_sim;
typedef struct fdtd2d_batch fdtd2d_batch;

typedef struct fdtd2d_grid {
  int32_t nx, ny;        /* interior cells (sim2/program.f90:48-49)         */
  int32_t precision;     /* FDTD_FP64 or FDTD_FP32 fields                  */
  int32_t periodic_x;    /* PBC_x (sim2/program.f90:12)                    */
  double ds;             /* cell size [m]                                   */
  double courant;        /* Ku = c0*dt/ds (sim2/program.f90:43-45)         */
  int32_t polarization;  /* FDTD2D_TEZ (the reference's) or FDTD2D_TMZ      */
  int32_t reserved;
} fdtd2d_grid;

typedef struct fdtd2d_pml {
  int32_t npml[2];       /* npmlx, npmly (sim2/program.f90:89-90)           */
  double order;          /* grading exponent m                              */
  double R0;             /* reflection target, used when sigmam <= 0        */
  double sigmam;         /* > 0: taken as is (sim2/program.f90:93)          */
  double eps_low;        /* relative permittivity beside the y-low slab     */
  double eps_high;       /* ... beside the y-high slab (:237)               */
} fdtd2d_pml;

typedef struct fdtd2d_region {
  int32_t kind;          /* FDTD2D_REGION_*                                  */
  int32_t reserved;
  int32_t lo[2], hi[2];  /* box: xinit,yinit .. xend,yend                   */
  double center[2];      /* disk centre in cell units (cell i spans i-1/2..i+1/2) */
  double radius;         /* disk radius in cells                            */
} fdtd2d_region;

/* ---- lifetime --------------------------------------------------------- */
/* Replaces the static fields and grid parameters of sim2/program.f90:40-67
 * and the rank layout of sim2/include/mympi_2Dfunc.f90 (one domain).      */
int fdtd2d_create(fdtd2d_sim** out, const fdtd2d_grid* grid);
int fdtd2d_destroy(fdtd2d_sim* sim);

/* ---- set-up (before the first step) ------------------------------------ */
/* pmly%initialize (mod2/BERpmly_HzExEy_mod.f90:229-247, profiles :249-426):
 * y slabs of depth npml[1], x grading over npml[0] cells inside them.
 * TMz: slabs on all faces as described above (npml[0] > 0 needs a
 * non-periodic x; either depth may be 0).                                 */
int fdtd2d_set_pml(fdtd2d_sim* sim, const fdtd2d_pml* pml);
/* ConstantDielectric (mod2/simpledielectric_mod.f90:21-79): E = D/eps_r.   */
int fdtd2d_add_dielectric(fdtd2d_sim* sim, const fdtd2d_region* region, double eps_r);
/* Gold (mod2/gold_mod.f90:47-168): Drude + 2 critical points, Ex and Ey.   */
int fdtd2d_add_drude2cp(fdtd2d_sim* sim, const fdtd2d_region* region, const fdtd_drude2cp* params);
/* Graphene (mod2/graphene_mod.f90:56-236): Ey = Dy; Ex through the CP4 (or
 * Kubo intraband Drude) ADE and the Raman recursion in the 2D module's
 * central-difference form (:210-212).  model, thickness, CP table, Kubo
 * and Raman fields as in fdtd.h; ref_init_order is ignored (the 2D module
 * sets its exponentials before use).  TMz: Ez through the ADE with the 3D
 * module's coefficients.  Dispersive cells must lie outside the PML.      */
int fdtd2d_add_graphene(fdtd2d_sim* sim, const fdtd2d_region* region, const fdtd_graphene* params);
/* <material>_LJp%initialize_material + update_EJp of the 2D Lorentz-Jp
 * modules (mod2/siLorentzJp_mod.f90, filmGoldLorentzJp_mod.f90,
 * graphene3LorentzJp1nm_mod.f90; the pole tables of fdtd_lorentz_jp_preset,
 * which the 2D and 3D modules share): E form from curl H, no D --
 * E_q += Ku*ga*(curl_q - ds*JpT_q) for q in jp_mask (TEz: x, y; TMz: z),
 * then the poles; the other components take E = D.  A line source through
 * these cells adds ga*g.  Cells must lie outside the PML.                  */
int fdtd2d_add_lorentz_jp(fdtd2d_sim* sim, const fdtd2d_region* region, const fdtd_lorentz_jp* params);
/* Soft line source on Dx along row j, all x (sim2/program.f90:336-340):
 * Dx(i,j) += amplitude*exp(-((t-t0)/st)^2/2)*sin(w0*t), t = (n-1)*dt.
 * TMz: on Dz; in PML cells on the split Ezy, scaled by the PML medium.    */
int fdtd2d_add_line_source(fdtd2d_sim* sim, int j, const fdtd_pulse* pulse);
/* Running DFT of Ex and/or Hz (TMz: Ez and/or Hx) along row j
 * (sim2/program.f90:373-438): for steps n_begin <= n <= n_end,
 * re += cos(omega*t)*F*dt, im += sin(omega*t)*F*dt, t = (n-1)*dt.
 * n_end <= 0: no upper gate.                                               */
int fdtd2d_add_dft_line(fdtd2d_sim* sim, int j, int nfreq, const double* omega, int comp_mask, int n_begin,
                        int n_end, int* id);
/* Plane wave by total-field / scattered-field along y (no counterpart in
 * the reference's 2D driver): the 1D driver's scheme (sim1/program.f90:86-91,
 * 217-262; the same terms as fdtd_add_tfsf with axis 1) with an incident
 * line of ny cells along y carrying Ex/Hz (TEz) or Ez/Hx (TMz), its own
 * end-slab PML (mod1/BERpmly_ExHz_mod.f90) and the pulse on line cell
 * src_index.  Total field on rows lo..hi (hi = 0: lo..ny, one-sided); per
 * step D(lo) -= Ku*h(lo-1), D(hi+1) += Ku*h(hi) and H(lo-1) -= s*Ku*e(lo),
 * H(hi) += s*Ku*e(hi+1) with s = +1 (TEz) or -1 (TMz), on every column.
 * The four correction rows must lie outside the PML.  The dormant 2D
 * include sim2/include/myTFSF_2Dfunc.f90 is not followed: its Hz term sits
 * one row off its Dx term, which leaks the incident wave into the
 * scattered region.  One per simulation.                                  */
typedef struct fdtd2d_tfsf {
  int32_t lo, hi;        /* total-field rows                                 */
  int32_t src_index;     /* line cell of the pulse                           */
  int32_t npml;          /* line PML cells at each end (sim1: npml)          */
  double eps;            /* line medium, relative permittivity               */
  double order, R0;      /* line PML grading                                 */
  fdtd_pulse pulse;
} fdtd2d_tfsf;
int fdtd2d_add_tfsf(fdtd2d_sim* sim, const fdtd2d_tfsf* tfsf, int* id);
/* The incident line: e and h, ny+2 doubles each (cells 0..ny+1).          */
int fdtd2d_get_tfsf_line(fdtd2d_sim* sim, int id, double* e, double* h);
/* Time trace of (Ex, Hz) (TMz: (Ez, Hx)) at (i,j) after each step
 * (sim2/program.f90:376,397,417).                                          */
int fdtd2d_add_probe(fdtd2d_sim* sim, int i, int j, int* id);
int fdtd2d_commit(fdtd2d_sim* sim);

/* ---- time loop (sim2/program.f90:320-458) ------------------------------- */
int fdtd2d_step(fdtd2d_sim* sim, int nsteps);
int fdtd2d_sync(fdtd2d_sim* sim);
int fdtd2d_get_step(fdtd2d_sim* sim, int* n_done);
int fdtd2d_get_dt(fdtd2d_sim* sim, double* dt);

/* ---- results -------------------------------------------------------------- */
/* comp FDTD2D_*: nx*ny doubles, x fastest, interior cells.                 */
int fdtd2d_get_field(fdtd2d_sim* sim, int comp, double* dst);
int fdtd2d_set_field(fdtd2d_sim* sim, int comp, const double* src);
/* DFT accumulators of line `id` for comp FDTD2D_EX or FDTD2D_HZ (TMz:
 * FDTD2D_EZ or FDTD2D_HX): nx*nfreq doubles each, x fastest, frequency
 * slowest (real_exin(i,m)).                                                */
int fdtd2d_get_dft(fdtd2d_sim* sim, int id, int comp, double* re, double* im);
/* Sy(m) = -ds * sum_i (ReHz*ReEx + ImHz*ImEx) (sim2/program.f90:479-529);
 * TMz: Sy(m) = ds * sum_i (ReHx*ReEz + ImHx*ImEz), the Ez/Hx terms of the
 * 3D flux (sim3/program.f90:613-685).                                      */
int fdtd2d_get_flux(fdtd2d_sim* sim, int id, double* S);
int fdtd2d_get_probe_count(fdtd2d_sim* sim, int id, int* nsamples);
/* 3 doubles per sample: n*dt, Ex, Hz (TMz: n*dt, Ez, Hx).                  */
int fdtd2d_get_probe(fdtd2d_sim* sim, int id, double* trace);

/* ---- batches (no reference counterpart: the reference runs one simulation
 * per MPI job) ------------------------------------------------------------- */
/* Pack `count` committed FP32 or FP64 simulations of one precision (TEz and
 * TMz may mix) into one batch.  The batch borrows the handles: they stay valid
 * and readable through the calls above, but must not be stepped on their
 * own while the batch lives.                                               */
int fdtd2d_batch_create(fdtd2d_batch** out, fdtd2d_sim* const* sims, int count);
/* Advance simulation q by nsteps[q] steps (each may differ; 0 leaves it). */
int fdtd2d_batch_run(fdtd2d_batch* batch, const int32_t* nsteps);
int fdtd2d_batch_sync(fdtd2d_batch* batch);
int fdtd2d_batch_destroy(fdtd2d_batch* batch);

#ifdef __cplusplus
}
#endif

#endif
