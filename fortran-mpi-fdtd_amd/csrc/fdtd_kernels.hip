// fdtd_kernels.hip -- gfx950 kernels of the 3D Yee time step.
//
// Field layout: one array per component, x fastest, with a one-cell ghost
// ring on every axis (index 0..n+1, the reference's (0:N+1) bounds,
// sim3/program.f90:77).  Offset of (i,j,k) = i + px*j + pxy*k, px padded.
// Non-periodic ghosts stay zero (PEC); periodic axes keep the ghost planes
// that the reference's halo exchange fills (sim3/include/mympi_func.f90:
// 241-690) up to date from inside the kernels: the E writer of plane 1
// also writes ghost n+1, the H writer of plane n also writes ghost 0.
//
// Two engines:
//  FUSED  (production): E = E + Ku*gax*curl H per cell (D is kept only in
//         dispersive cells), split-field PML cells in the same kernel,
//         source folded in; one H kernel; one kernel per group of
//         dispersive cells.
//  STAGED (reference order): updateD, source, material E = D*gax,
//         dispersive ADE, y-slab PML with interface copies, updateH,
//         y-slab PML H -- one kernel per stage, D kept everywhere.
#include <hip/hip_runtime.h>

#include <cmath>

#include "fdtd_internal.h"

namespace fdtd {

template <typename real>
struct FieldArgs {
  real* e[3];
  real* h[3];
  real* d[3];
  const uint8_t* matid;
  const Vec4<real>* coef;   // Ku*gax x,y,z, PML medium gax
  const Vec4<real>* gaxe;   // gax x,y,z, energy eps
  const uint32_t* flags;    // bit q: a dielectric writes E_q
  const int32_t* rowmat;    // per ghosted (j,k) row: uniform id or -1 (dword: scalar loads)
  int32_t nx, ny, nz;
  int32_t per[3];
  int32_t px, pxy;          // offsets fit 32 bits (checked at commit)
  real ku;
  // tile grid of the cell-parallel kernels: ntx*nty*ntz tiles of
  // blockDim.x*blockDim.y cells, dealt out so each of the 8 XCDs (one L2
  // each) walks a contiguous run of `tper` tiles: neighbouring rows and
  // planes are then read from the same L2.
  int32_t ntx, nty, ntz, tper, xcdmap;
};

// Tile coordinates of this block (false: padding block, nothing to do).
template <typename real>
__device__ __forceinline__ bool tile_of(const FieldArgs<real>& a, int& tx, int& ty, int& tz) {
  const int b = blockIdx.x;
  const int t = a.xcdmap ? (b & 7) * a.tper + (b >> 3) : b;
  if (t >= a.ntx * a.nty * a.ntz) return false;
  tx = t % a.ntx;
  const int r = t / a.ntx;
  ty = r % a.nty;
  tz = r / a.nty;
  return true;
}

template <typename real>
struct PmlArgs {
  int32_t P[3];
  real* s[S_N];
  const real* f1[3];
  const real* f2[3];
  const real* g1[3];
  const real* g2[3];
  real kup;  // f2 and g2 outside an axis's slabs (f1 and g1 are 1 there)
};

template <typename real>
struct RefPmlArgs {
  int32_t P, jb;
  int64_t sx, sxy;  // slab strides: i + sx*jl + sxy*k
  real* s[2][S_N];
  const real* prof[2];  // 12 profiles per slab, stride ps
  int32_t ps;
};

// ---------------------------------------------------------------------------
// ghost maintenance for periodic axes
template <typename real>
__device__ __forceinline__ void put_e_ghost(real* f, real v, int i, int j, int k, int64_t c,
                                            const FieldArgs<real>& a) {
  if (a.per[0] && i == 1) f[c + a.nx] = v;
  if (a.per[1] && j == 1) f[c + a.px * a.ny] = v;
  if (a.per[2] && k == 1) f[c + a.pxy * a.nz] = v;
}
template <typename real>
__device__ __forceinline__ void put_h_ghost(real* f, real v, int i, int j, int k, int64_t c,
                                            const FieldArgs<real>& a) {
  if (a.per[0] && i == a.nx) f[c - a.nx] = v;
  if (a.per[1] && j == a.ny) f[c - a.px * a.ny] = v;
  if (a.per[2] && k == a.nz) f[c - a.pxy * a.nz] = v;
}

__device__ __forceinline__ bool in_pml(const int32_t* P, int i, int j, int k, int nx, int ny, int nz) {
  return i <= P[0] || i > nx - P[0] || j <= P[1] || j > ny - P[1] || k <= P[2] || k > nz - P[2];
}

// cell v of an axis of n cells lies in one of the axis's two PML slabs
__device__ __forceinline__ bool in_slab(int P, int v, int n) { return v <= P || v > n - P; }

// Split parts of an ALL_FACES PML cell before the update.  A component is
// the sum of two parts, one per transverse axis (Ex = Exy + Exz, Ey = Eyz +
// Eyx, Ez = Ezx + Ezy; H alike); only a part whose axis's slab holds the
// cell is stored (sx, sy, sz), since the other axis's profile there is that
// of the interior (f1 = 1, f2 = Ku).  A part not stored is the old
// component less its other part, or, with neither part stored, the first
// part of the sum is the whole old component and the second 0 -- in exact
// arithmetic the same step, and a face cell moves 13 words a half step
// instead of 18.  sp: stored parts in the order yx, zx, xy, zy, xz, yz
// (the x parts, the y parts, the z parts); f: the old components; o: the
// parts in that order.
template <typename real>
__device__ __forceinline__ void pml_parts(bool sx, bool sy, bool sz, const real* sp, const real* f, real* o) {
  o[0] = sx ? sp[0] : (sz ? f[1] - sp[5] : (real)0);  // yx, second of y = yz + yx
  o[1] = sx ? sp[1] : (sy ? f[2] - sp[3] : f[2]);     // zx, first of z = zx + zy
  o[2] = sy ? sp[2] : (sz ? f[0] - sp[4] : f[0]);     // xy, first of x = xy + xz
  o[3] = sy ? sp[3] : (sx ? f[2] - sp[1] : (real)0);  // zy
  o[4] = sz ? sp[4] : (sy ? f[0] - sp[2] : (real)0);  // xz
  o[5] = sz ? sp[5] : (sx ? f[1] - sp[0] : f[1]);     // yz
}

// The six split updates of an ALL_FACES PML cell (mod3/BERpmly_mod.f90:150-155
// for E, :186-191 for H) from the parts o (pml_parts order) and the
// neighbouring fields: f1/g1 and f2/g2 are the cell's profiles along each
// axis (1 and Ku outside that axis's slabs), r the medium's 1/eps.  Every
// update of a PML cell, whichever kernel makes it, goes through these, so
// all of them round alike.
template <typename real>
__device__ __forceinline__ void e_pml_split(real fx1, real fx2, real fy1, real fy2, real fz1, real fz2, real r,
                                            const real* o, real hx, real hy, real hz, real hz_jm, real hx_jm,
                                            real hy_km, real hx_km, real hz_im, real hy_im, real* n) {
  n[0] = fx1 * o[0] + (fx2 * r) * (-hz + hz_im);    // yx
  n[1] = fx1 * o[1] + (fx2 * r) * (hy + (-hy_im));  // zx
  n[2] = fy1 * o[2] + (fy2 * r) * (hz + (-hz_jm));  // xy
  n[3] = fy1 * o[3] + (fy2 * r) * (-hx + hx_jm);    // zy
  n[4] = fz1 * o[4] + (fz2 * r) * (-hy + hy_km);    // xz
  n[5] = fz1 * o[5] + (fz2 * r) * (hx + (-hx_km));  // yz
}

template <typename real>
__device__ __forceinline__ void h_pml_split(real gx1, real gx2, real gy1, real gy2, real gz1, real gz2, const real* o,
                                            real ex, real ey, real ez, real ey_kp, real ex_kp, real ez_jp, real ex_jp,
                                            real ez_ip, real ey_ip, real* n) {
  n[0] = gx1 * o[0] + gx2 * (-ez + ez_ip);    // yx
  n[1] = gx1 * o[1] + gx2 * (ey + (-ey_ip));  // zx
  n[2] = gy1 * o[2] + gy2 * (ez + (-ez_jp));  // xy
  n[3] = gy1 * o[3] + gy2 * (-ex + ex_jp);    // zy
  n[4] = gz1 * o[4] + gz2 * (-ey + ey_kp);    // xz
  n[5] = gz1 * o[5] + gz2 * (ex + (-ex_kp));  // yz
}

// Compact index of an ALL_FACES PML cell: z slabs (full xy planes), then y
// slabs of the middle planes (full x rows), then the low and then the high
// x slab of the rest, each of those row after row (Px cells a row), so the
// rows a workgroup of one x slab updates are contiguous in every split
// array.  The count of PML cells is below the ghosted cell count, which
// fits 31 bits.
__device__ __forceinline__ int pml_index(const int32_t* P, int i, int j, int k, int nx, int ny, int nz) {
  const int Px = P[0], Py = P[1], Pz = P[2];
  if (k <= Pz || k > nz - Pz) {
    const int kk = k <= Pz ? k - 1 : Pz + (k - (nz - Pz + 1));
    return kk * nx * ny + (j - 1) * nx + (i - 1);
  }
  const int baseY = 2 * Pz * nx * ny;
  const int km = k - Pz - 1;
  if (j <= Py || j > ny - Py) {
    const int jj = j <= Py ? j - 1 : Py + (j - (ny - Py + 1));
    return baseY + (km * 2 * Py + jj) * nx + (i - 1);
  }
  const int nym = ny - 2 * Py;
  const int baseX = baseY + (nz - 2 * Pz) * 2 * Py * nx;
  const int jm = j - Py - 1;
  const bool hi = i > Px;
  const int ii = hi ? i - (nx - Px + 1) : i - 1;
  return baseX + (hi ? (nz - 2 * Pz) * nym * Px : 0) + (km * nym + jm) * Px + ii;
}

// ---------------------------------------------------------------------------
// FUSED engine: E half step (myfdtd_func.f90:19-41 in E-form, material
// update_E folded in, ALL_FACES split-field PML, plane sources).
//
// Every load is issued before the first store (e_cell_load, then
// e_cell_store): the field pointers may alias as far as the compiler knows,
// so a load placed after a store would wait for it and cost one more memory
// round trip.  The two phases are apart so that a thread marching a column
// can issue the next plane's loads before this plane's arithmetic.  UNI:
// the row has one material id `rm` (wave-uniform, the table entry read with
// scalar loads); otherwise the id is gathered per cell.
//
// z march: with use_carry, Hy and Hx of plane k-1 come from cy, cx (the
// previous plane of the same thread) instead of memory.
//
// In place (the two-kernel step) EO* are E* and pw is p.  Over two copies
// (the general tiles of the one-pass step, pp): the old E comes from EO*,
// the old splits from p, the new E goes to E* and the new splits to pw, and
// components no material writes are copied over.
template <typename real>
struct ECellIn {
  int c, s;
  bool pml, sx, sy, sz, any_src;
  uint32_t fl, id;
  Vec4<real> cf;
  real hx, hy, hz, hz_jm, hx_jm, hy_km, hx_km, hz_im, hy_im;
  real e0[3], sp[6], fx1, fx2, fy1, fy2, fz1, fz2;
};

template <typename real, bool UNI, int FACE = 0>
__device__ __forceinline__ void e_cell_load(const FieldArgs<real>& a, const PmlArgs<real>& p,
                                            const int* sa, int rm, int i, int j, int k, bool use_carry, real cy,
                                            real cx, const real* H0, const real* H1, const real* H2,
                                            const real* EO0, const real* EO1, const real* EO2, ECellIn<real>& in) {
  const int c = i + a.px * j + a.pxy * k;
  in.c = c;
  in.hx = H0[c]; in.hy = H1[c]; in.hz = H2[c];
  in.hz_jm = H2[c - a.px]; in.hx_jm = H0[c - a.px];
  in.hy_km = use_carry ? cy : H1[c - a.pxy];
  in.hx_km = use_carry ? cx : H0[c - a.pxy];
  in.hz_im = H2[c - 1]; in.hy_im = H1[c - 1];
  const unsigned id = UNI ? (unsigned)__builtin_amdgcn_readfirstlane(rm) : (unsigned)a.matid[c];
  in.id = id;
  in.fl = a.flags[id];
  in.cf = a.coef[id];
  in.pml = FACE ? true : in_pml(p.P, i, j, k, a.nx, a.ny, a.nz);
  in.any_src = sa[0] >= 0 || sa[1] >= 0 || sa[2] >= 0;
  in.s = 0;
  in.sx = in.sy = in.sz = false;
  for (int q = 0; q < 6; ++q) in.sp[q] = 0;
  in.e0[0] = in.e0[1] = in.e0[2] = 0;
  in.fx1 = in.fx2 = in.fy1 = in.fy2 = in.fz1 = in.fz2 = 0;
  if (in.pml) {
    const int s = pml_index(p.P, i, j, k, a.nx, a.ny, a.nz);
    bool sx, sy, sz;
    if constexpr (FACE == 0) {
      sx = in_slab(p.P[0], i, a.nx);
      sy = in_slab(p.P[1], j, a.ny);
      sz = in_slab(p.P[2], k, a.nz);
      in.fx1 = p.f1[0][i]; in.fx2 = p.f2[0][i];
      in.fy1 = p.f1[1][j]; in.fy2 = p.f2[1][j];
      in.fz1 = p.f1[2][k]; in.fz2 = p.f2[2][k];
    } else {  // a face cell of the y (FACE 1) or z (FACE 2) slabs: the other axes' profiles are 1, Ku
      sx = false;
      sy = FACE == 1;
      sz = FACE == 2;
      in.fx1 = (real)1; in.fx2 = p.kup;
      in.fy1 = sy ? p.f1[1][j] : (real)1; in.fy2 = sy ? p.f2[1][j] : p.kup;
      in.fz1 = sz ? p.f1[2][k] : (real)1; in.fz2 = sz ? p.f2[2][k] : p.kup;
    }
    in.s = s;
    in.sx = sx;
    in.sy = sy;
    in.sz = sz;
    // only the parts along the axes whose slabs hold the cell are stored
    // (pml_parts); the old component is read where a part is not
    if (sx) { in.sp[0] = p.s[S_EYX][s]; in.sp[1] = p.s[S_EZX][s]; }
    if (sy) { in.sp[2] = p.s[S_EXY][s]; in.sp[3] = p.s[S_EZY][s]; }
    if (sz) { in.sp[4] = p.s[S_EXZ][s]; in.sp[5] = p.s[S_EYZ][s]; }
    if (!(sy && sz)) in.e0[0] = EO0[c];
    if (!(sz && sx)) in.e0[1] = EO1[c];
    if (!(sx && sy)) in.e0[2] = EO2[c];
  } else {
    in.e0[0] = EO0[c]; in.e0[1] = EO1[c]; in.e0[2] = EO2[c];
  }
}

template <typename real>
__device__ __forceinline__ void e_cell_store(const FieldArgs<real>& a, const PmlArgs<real>& pw, bool pp,
                                             const real* sv, const int* sa, int i, int j, int k,
                                             const ECellIn<real>& in, real* E0, real* E1, real* E2) {
  const int c = in.c;
  const real* e0 = in.e0;
  const real hx = in.hx, hy = in.hy, hz = in.hz, hz_jm = in.hz_jm, hx_jm = in.hx_jm, hy_km = in.hy_km,
             hx_km = in.hx_km, hz_im = in.hz_im, hy_im = in.hy_im;
  if (in.pml) {
    // mod3/BERpmly_mod.f90:150-155 with the splits driven by total H
    const real r = in.cf.v[3];
    const int s = in.s;
    real o[6];
    pml_parts(in.sx, in.sy, in.sz, in.sp, e0, o);
    real n[6];
    e_pml_split(in.fx1, in.fx2, in.fy1, in.fy2, in.fz1, in.fz2, r, o, hx, hy, hz, hz_jm, hx_jm, hy_km, hx_km, hz_im,
                hy_im, n);
    real eyx = n[0], ezx = n[1], exy = n[2], ezy = n[3], exz = n[4], eyz = n[5];
    // a source plane through the PML feeds the split along its normal
    if (sa[0] == 1) exy = exy + sv[0] * r;
    if (sa[0] == 2) exz = exz + sv[0] * r;
    if (sa[1] == 0) eyx = eyx + sv[1] * r;
    if (sa[1] == 2) eyz = eyz + sv[1] * r;
    if (sa[2] == 0) ezx = ezx + sv[2] * r;
    if (sa[2] == 1) ezy = ezy + sv[2] * r;
    if (in.sx) { pw.s[S_EYX][s] = eyx; pw.s[S_EZX][s] = ezx; }
    if (in.sy) { pw.s[S_EXY][s] = exy; pw.s[S_EZY][s] = ezy; }
    if (in.sz) { pw.s[S_EXZ][s] = exz; pw.s[S_EYZ][s] = eyz; }
    const real ex = exy + exz, ey = eyz + eyx, ez = ezx + ezy;
    E0[c] = ex;
    E1[c] = ey;
    E2[c] = ez;
    // a periodic axis without PML of its own still meets the slabs of the
    // other axes at its boundary
    put_e_ghost(E0, ex, i, j, k, c, a);
    put_e_ghost(E1, ey, i, j, k, c, a);
    put_e_ghost(E2, ez, i, j, k, c, a);
    return;
  }
  // E_q += Ku*gax_q*curl_q (+ gax_q * source) for every component a
  // dielectric writes
  const real cur[3] = {(hz - hz_jm) + (-hy + hy_km),   // curlhzy + curlhyz
                       (hx - hx_km) + (-hz + hz_im),   // curlhxz + curlhzx
                       (hy - hy_im) + (-hx + hx_jm)};  // curlhyx + curlhxy
  const uint32_t fl = in.fl;
  // the source scale: loaded here, source planes are few
  Vec4<real> gx;
  if (in.any_src) gx = a.gaxe[in.id];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!(fl & (1u << q))) {
      if (pp && !(fl & (16u << q))) {  // no writer: the value carries over
        real* Eq = q == 0 ? E0 : (q == 1 ? E1 : E2);
        Eq[c] = e0[q];
        put_e_ghost(Eq, e0[q], i, j, k, c, a);
      }
      continue;
    }
    real ev = e0[q] + in.cf.v[q] * cur[q];
    if (sa[q] >= 0) ev = ev + gx.v[q] * sv[q];
    real* Eq = q == 0 ? E0 : (q == 1 ? E1 : E2);
    Eq[c] = ev;
    put_e_ghost(Eq, ev, i, j, k, c, a);
  }
}

template <typename real, bool UNI>
__device__ __forceinline__ void fused_e_cell(const FieldArgs<real>& a, const PmlArgs<real>& p,
                                             const PmlArgs<real>& pw, bool pp, const real* sv, const int* sa, int rm,
                                             int i, int j, int k, real& cy, real& cx, bool use_carry,
                                             const real* H0, const real* H1, const real* H2, const real* EO0,
                                             const real* EO1, const real* EO2, real* E0, real* E1, real* E2) {
  ECellIn<real> in;
  e_cell_load<real, UNI>(a, p, sa, rm, i, j, k, use_carry, cy, cx, H0, H1, H2, EO0, EO1, EO2, in);
  cy = in.hy;
  cx = in.hx;
  e_cell_store(a, pw, pp, sv, sa, i, j, k, in, E0, E1, E2);
}

// plane sources through cell (i,j,k): value and normal axis per E
// component (-1: none here).  Called before any memory load is in flight (a
// loop carrying per-lane values past outstanding loads makes the compiler
// wait for them).
template <typename real>
__device__ __forceinline__ void plane_sources(const SrcArgs<real>& src, int nsrc, int i, int j, int k, real* sv,
                                              int* sa) {
  for (int q = 0; q < nsrc; ++q) {
    const int ax = src.axis[q];
    const int pos = ax == 0 ? i : (ax == 1 ? j : k);
    if (pos == src.index[q]) {
      const int cq = src.comp[q];
      const real v = src_value(src, q);
      if (cq == 0) { sv[0] += v; sa[0] = ax; }
      else if (cq == 1) { sv[1] += v; sa[1] = ax; }
      else { sv[2] += v; sa[2] = ax; }
    }
  }
}

// The z march of one (i,j) column over planes tz*KZ+1 .. tz*KZ+KZ.  The
// field arrays come in as restrict pointers: no H array is written here and
// no E array is read through another pointer, so the compiler may issue the
// loads of the next plane before the stores of this one.
template <typename real, int KZ>
__device__ __forceinline__ void e_column(const FieldArgs<real>& a, const PmlArgs<real>& p, const SrcArgs<real>& src,
                                         int nsrc, int i, int j, int tz, const real* __restrict__ H0,
                                         const real* __restrict__ H1, const real* __restrict__ H2,
                                         real* __restrict__ E0, real* __restrict__ E1, real* __restrict__ E2) {
  real cy = 0, cx = 0;
  bool have = false;
#pragma unroll
  for (int r = 0; r < KZ; ++r) {
    const int k = tz * KZ + r + 1;
    if (k > a.nz) break;
    const int rm = a.rowmat[__builtin_amdgcn_readfirstlane(j + (a.ny + 2) * k)];
    if (rm < 0) {  // a mixed row: k_fused_e_mixed updates it
      have = false;
      continue;
    }
    real sv[3] = {0, 0, 0};
    int sa[3] = {-1, -1, -1};
    plane_sources(src, nsrc, i, j, k, sv, sa);
    fused_e_cell<real, true>(a, p, p, false, sv, sa, rm, i, j, k, cy, cx, have, H0, H1, H2, E0, E1, E2, E0, E1, E2);
    have = true;
  }
}

// Rows with one material id (all but the rows that cross a particle
// surface): the table entry is a wave-uniform scalar load.  A tile spans KZ
// planes: each thread marches its (i,j) column through them, carrying H of
// the plane below in registers (KZ = 1: one plane, nothing carried).
template <typename real, int KZ>
__global__ __launch_bounds__(512) void k_fused_e(FieldArgs<real> a, PmlArgs<real> p, SrcArgs<real> src,
                                                 int nsrc) {
  int tx, ty, tz;
  if (!tile_of(a, tx, ty, tz)) return;
  const int i = tx * blockDim.x + threadIdx.x + 1;
  // blockDim.x is a multiple of 64 (set at commit): a wavefront lies in one
  // (j,k) row, so j and k are wave-uniform and row data are scalar loads
  const int j = ty * blockDim.y + __builtin_amdgcn_readfirstlane((int)threadIdx.y) + 1;
  if (i > a.nx || j > a.ny) return;
  e_column<real, KZ>(a, p, src, nsrc, i, j, tz, a.h[0], a.h[1], a.h[2], a.e[0], a.e[1], a.e[2]);
}

// Rows that mix material ids: block b updates x tile b % ntx of row
// rows[b / ntx] (row index j + (ny+2)*k), gathering table entries per cell.
template <typename real>
__global__ __launch_bounds__(256) void k_fused_e_mixed(FieldArgs<real> a, PmlArgs<real> p, SrcArgs<real> src,
                                                       int nsrc, const int32_t* rows, int ntx) {
  const int r = rows[blockIdx.x / ntx];
  const int i = (blockIdx.x % ntx) * blockDim.x + threadIdx.x + 1;
  const int j = r % (a.ny + 2), k = r / (a.ny + 2);
  if (i > a.nx) return;
  real sv[3] = {0, 0, 0};
  int sa[3] = {-1, -1, -1};
  plane_sources(src, nsrc, i, j, k, sv, sa);
  real cy = 0, cx = 0;
  fused_e_cell<real, false>(a, p, p, false, sv, sa, -1, i, j, k, cy, cx, false, a.h[0], a.h[1], a.h[2], a.e[0],
                            a.e[1], a.e[2], a.e[0], a.e[1], a.e[2]);
}

// Monitor planes sampled by the FUSED H kernel: a cell on such a plane
// writes its E and new H components to the plane's sample buffer (the same
// values and layout as k_monitors, which would read them again
// afterwards; an x-normal plane has one sample per 128-byte line).
// FP32 only: in FP64 the extra registers cut the H kernel's occupancy
// (8 -> 5 waves per SIMD) and cost more than the separate sample launch.
constexpr int kFuseDft = 4;
template <typename real>
constexpr bool kFuseDftReal = sizeof(real) == 4;
template <typename real>
struct DftFuse {
  int32_t n;
  int32_t axis[kFuseDft], index[kFuseDft], n1[kFuseDft], ns[kFuseDft];
  int64_t np[kFuseDft];
  int32_t comp[kFuseDft][6];
  real* dst[kFuseDft];
};

template <typename real>
__device__ __forceinline__ void dft_fuse_write(const DftFuse<real>& df, int i, int j, int k, real ex, real ey,
                                               real ez, real hx, real hy, real hz) {
  for (int q = 0; q < df.n; ++q) {
    const int ax = df.axis[q];
    if ((ax == 0 ? i : (ax == 1 ? j : k)) != df.index[q]) continue;
    const int u = ax == 0 ? j : i, v = ax == 2 ? j : k;
    const int64_t o = (int64_t)(u - 1) + (int64_t)(v - 1) * df.n1[q];
    for (int sl = 0; sl < df.ns[q]; ++sl) {
      const int c = df.comp[q][sl];
      const real x = c == 0 ? ex : c == 1 ? ey : c == 2 ? ez : c == 3 ? hx : c == 4 ? hy : hz;
      df.dst[q][sl * df.np[q] + o] = x;
    }
  }
}

// FUSED engine: H half step (myfdtd_func.f90:45-64 + split-field H) of one
// cell, as the E step in a load phase (h_cell_load) and an arithmetic and
// store phase (h_cell_store).  z march: with use_carry, Ex and Ey of plane
// k come from cex, cey (loaded as plane k+1 by the previous plane of the
// same thread).  In place HO* are H* and pw is p; over two copies the old
// H and splits come from HO* and p, the new ones go to H* and pw.
template <typename real>
struct HCellIn {
  int c, s;
  bool pml, sx, sy, sz;
  real ex, ey, ez, ey_kp, ex_kp, ez_jp, ex_jp, ez_ip, ey_ip;
  real h0[3], sp[6], gx1, gx2, gy1, gy2, gz1, gz2;
};

template <typename real, int FACE = 0>
__device__ __forceinline__ void h_cell_load(const FieldArgs<real>& a, const PmlArgs<real>& p, int i, int j, int k,
                                            bool use_carry, real cex, real cey, const real* E0, const real* E1,
                                            const real* E2, const real* HO0, const real* HO1, const real* HO2,
                                            HCellIn<real>& in) {
  const int c = i + a.px * j + a.pxy * k;
  in.c = c;
  in.ex = use_carry ? cex : E0[c];
  in.ey = use_carry ? cey : E1[c];
  in.ez = E2[c];
  in.ey_kp = E1[c + a.pxy]; in.ex_kp = E0[c + a.pxy];
  in.ez_jp = E2[c + a.px]; in.ex_jp = E0[c + a.px];
  in.ez_ip = E2[c + 1]; in.ey_ip = E1[c + 1];
  in.pml = FACE ? true : in_pml(p.P, i, j, k, a.nx, a.ny, a.nz);
  in.s = 0;
  in.sx = in.sy = in.sz = false;
  for (int q = 0; q < 6; ++q) in.sp[q] = 0;
  in.h0[0] = in.h0[1] = in.h0[2] = 0;
  in.gx1 = in.gx2 = in.gy1 = in.gy2 = in.gz1 = in.gz2 = 0;
  if (in.pml) {
    const int s = pml_index(p.P, i, j, k, a.nx, a.ny, a.nz);
    bool sx, sy, sz;
    if constexpr (FACE == 0) {
      sx = in_slab(p.P[0], i, a.nx);
      sy = in_slab(p.P[1], j, a.ny);
      sz = in_slab(p.P[2], k, a.nz);
      in.gx1 = p.g1[0][i]; in.gx2 = p.g2[0][i];
      in.gy1 = p.g1[1][j]; in.gy2 = p.g2[1][j];
      in.gz1 = p.g1[2][k]; in.gz2 = p.g2[2][k];
    } else {
      sx = false;
      sy = FACE == 1;
      sz = FACE == 2;
      in.gx1 = (real)1; in.gx2 = p.kup;
      in.gy1 = sy ? p.g1[1][j] : (real)1; in.gy2 = sy ? p.g2[1][j] : p.kup;
      in.gz1 = sz ? p.g1[2][k] : (real)1; in.gz2 = sz ? p.g2[2][k] : p.kup;
    }
    in.s = s;
    in.sx = sx;
    in.sy = sy;
    in.sz = sz;
    if (sx) { in.sp[0] = p.s[S_HYX][s]; in.sp[1] = p.s[S_HZX][s]; }
    if (sy) { in.sp[2] = p.s[S_HXY][s]; in.sp[3] = p.s[S_HZY][s]; }
    if (sz) { in.sp[4] = p.s[S_HXZ][s]; in.sp[5] = p.s[S_HYZ][s]; }
    if (!(sy && sz)) in.h0[0] = HO0[c];
    if (!(sz && sx)) in.h0[1] = HO1[c];
    if (!(sx && sy)) in.h0[2] = HO2[c];
  } else {
    in.h0[0] = HO0[c]; in.h0[1] = HO1[c]; in.h0[2] = HO2[c];
  }
}

template <typename real>
__device__ __forceinline__ void h_cell_store(const FieldArgs<real>& a, const PmlArgs<real>& pw,
                                             const DftFuse<real>& df, int i, int j, int k, const HCellIn<real>& in,
                                             real* H0, real* H1, real* H2) {
  const int c = in.c;
  const real ex = in.ex, ey = in.ey, ez = in.ez, ey_kp = in.ey_kp, ex_kp = in.ex_kp, ez_jp = in.ez_jp,
             ex_jp = in.ex_jp, ez_ip = in.ez_ip, ey_ip = in.ey_ip;
  if (in.pml) {
    // mod3/BERpmly_mod.f90:186-191 with the splits driven by total E
    const int s = in.s;
    real o[6];
    pml_parts(in.sx, in.sy, in.sz, in.sp, in.h0, o);
    real n[6];
    h_pml_split(in.gx1, in.gx2, in.gy1, in.gy2, in.gz1, in.gz2, o, ex, ey, ez, ey_kp, ex_kp, ez_jp, ex_jp, ez_ip, ey_ip,
                n);
    const real hyx = n[0], hzx = n[1], hxy = n[2], hzy = n[3], hxz = n[4], hyz = n[5];
    if (in.sx) { pw.s[S_HYX][s] = hyx; pw.s[S_HZX][s] = hzx; }
    if (in.sy) { pw.s[S_HXY][s] = hxy; pw.s[S_HZY][s] = hzy; }
    if (in.sz) { pw.s[S_HXZ][s] = hxz; pw.s[S_HYZ][s] = hyz; }
    const real hx = hxy + hxz, hy = hyz + hyx, hz = hzx + hzy;
    H0[c] = hx;
    H1[c] = hy;
    H2[c] = hz;
    put_h_ghost(H0, hx, i, j, k, c, a);
    put_h_ghost(H1, hy, i, j, k, c, a);
    put_h_ghost(H2, hz, i, j, k, c, a);
    if constexpr (kFuseDftReal<real>) dft_fuse_write(df, i, j, k, ex, ey, ez, hx, hy, hz);
    return;
  }
  const real curleyz = ey_kp - ey, curlezy = -ez_jp + ez;
  const real curlezx = ez_ip - ez, curlexz = -ex_kp + ex;
  const real curlexy = ex_jp - ex, curleyx = -ey_ip + ey;
  const real hx = in.h0[0] + a.ku * (curleyz + curlezy);
  const real hy = in.h0[1] + a.ku * (curlezx + curlexz);
  const real hz = in.h0[2] + a.ku * (curlexy + curleyx);
  H0[c] = hx;
  H1[c] = hy;
  H2[c] = hz;
  put_h_ghost(H0, hx, i, j, k, c, a);
  put_h_ghost(H1, hy, i, j, k, c, a);
  put_h_ghost(H2, hz, i, j, k, c, a);
  if constexpr (kFuseDftReal<real>) dft_fuse_write(df, i, j, k, ex, ey, ez, hx, hy, hz);
}

template <typename real>
__device__ __forceinline__ void fused_h_cell(const FieldArgs<real>& a, const PmlArgs<real>& p,
                                             const PmlArgs<real>& pw, const DftFuse<real>& df, int i, int j, int k,
                                             real& cex, real& cey, bool use_carry, const real* E0, const real* E1,
                                             const real* E2, const real* HO0, const real* HO1, const real* HO2,
                                             real* H0, real* H1, real* H2) {
  HCellIn<real> in;
  h_cell_load(a, p, i, j, k, use_carry, cex, cey, E0, E1, E2, HO0, HO1, HO2, in);
  cex = in.ex_kp;
  cey = in.ey_kp;
  h_cell_store(a, pw, df, i, j, k, in, H0, H1, H2);
}

// H over tiles of KZ planes, each thread marching its (i,j) column; restrict
// pointers as in e_column (no E array is written here).
template <typename real, int KZ>
__device__ __forceinline__ void h_column(const FieldArgs<real>& a, const PmlArgs<real>& p, const DftFuse<real>& df,
                                         int i, int j, int tz,
                                         const real* __restrict__ E0, const real* __restrict__ E1,
                                         const real* __restrict__ E2, real* __restrict__ H0, real* __restrict__ H1,
                                         real* __restrict__ H2) {
  real cex = 0, cey = 0;
#pragma unroll
  for (int r = 0; r < KZ; ++r) {
    const int k = tz * KZ + r + 1;
    if (k > a.nz) break;
    fused_h_cell(a, p, p, df, i, j, k, cex, cey, r > 0, E0, E1, E2, H0, H1, H2, H0, H1, H2);
  }
}

template <typename real, int KZ>
__global__ __launch_bounds__(512) void k_fused_h(FieldArgs<real> a, PmlArgs<real> p, DftFuse<real> df) {
  int tx, ty, tz;
  if (!tile_of(a, tx, ty, tz)) return;
  const int i = tx * blockDim.x + threadIdx.x + 1;
  const int j = ty * blockDim.y + __builtin_amdgcn_readfirstlane((int)threadIdx.y) + 1;  // wave-uniform row
  if (i > a.nx || j > a.ny) return;
  h_column<real, KZ>(a, p, df, i, j, tz, a.e[0], a.e[1], a.e[2], a.h[0], a.h[1], a.h[2]);
}

// ---------------------------------------------------------------------------
// One-pass step (FUSED engine): the E and the H half step of a time step
// read from one copy of the fields and written to the other (old: m.eo,
// m.ho and the splits p.s; new: a.e, a.h and pw.s), so no cell reads a value
// another workgroup has already replaced, whatever order they run in.  The
// grid is cut into tiles of 64 x R cells over z ranges of at most kc planes
// (lists built on the host, fdtd_api.hip build_march_tiles):
//
//  plain tiles (one dielectric material, no PML cell among the cells they
//  own, no x- or y-normal source plane; z-normal sources are added per
//  plane) go to k_march_plain: one wavefront marches the tile through its
//  planes and does E and H in one pass, so each field crosses HBM once per
//  step (12 words per cell where the two-kernel step moves 18).  A plain
//  tile owns the cells of its x range [xlo, xhi]; every cell whose new E
//  it computes (its own, the row, plane and column after them) is a
//  non-PML cell of its material.
//
//  x-face tiles (FP32): plain tiles whose lanes reach into an x PML slab;
//  those lanes keep the x slab's parts (yx, zx) of E and H and take the
//  split-field update of a face cell (e_pml_split / h_pml_split, as the
//  general kernels do), the other lanes the plain one.  They run as a
//  launch of their own (k_march_plain<real, R, true>), so the plain tiles'
//  kernel keeps its registers.
//
//  general tiles (PML, dispersive or mixed materials, other sources, and
//  the x-slab parts of plain tiles) go to k_gen_e and then k_gen_h, the
//  fused cell updates of k_fused_e / k_fused_h over the two copies.
//
// The arithmetic of every cell is that of the two-kernel step term for
// term, so both forms give the same bits.  Dispersive components were
// written into the new E by k_ade before these kernels.
template <typename real>
struct MarchArgs {
  const real* eo[3];      // E after the previous step
  const real* ho[3];      // H after the previous step
  // 8 words per entry; plain tiles: i0, j0, k0, k1, xlo, xhi (owned x
  // range), material id, x face; general workgroups: x0, x1, j0, j1, k0,
  // width, k1, face
  const int32_t* tiles;
  int32_t ntiles, tper;
  int32_t nsplit;  // general workgroups: workgroups per entry (its planes, k_gen_e) or 1
};


// lane l receives lane l-1 (lane 0: 0) / lane l+1 (lane 63: 0); every lane
// of the wavefront must be active
template <typename real>
__device__ __forceinline__ real lane_prev(real v) {
  if constexpr (sizeof(real) == 4) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}
template <typename real>
__device__ __forceinline__ real lane_next(real v) {
  if constexpr (sizeof(real) == 4) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}
// value of lane l (wave-uniform l)
template <typename real>
__device__ __forceinline__ real lane_at(real v, int l) {
  if constexpr (sizeof(real) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

// cell index v of an axis of n cells as a tile sees it: v <= n is the cell
// itself; v = n+1 is cell 1 on a periodic axis and a dead (zero) ghost
// otherwise
__device__ __forceinline__ int march_map(int v, int n, int per, bool& live) {
  live = v <= n || (v == n + 1 && per);
  return v <= n ? v : 1;
}

// What a wavefront marches: its columns and rows (physical cells after the
// periodic wrap) and which of them are live / owned.
template <int R>
struct MarchTile {
  int ip, i0, ixh, j0, jyh, jrow[R], jxg, jxl, k0, k1;
  bool lane_live, lane_own, xh_need, xg_live, yh_live, yh_need;
  bool row_live[R], row_own[R];
};

// Old fields of one plane of a plain tile, loaded a plane ahead of use.
template <typename real, int R>
struct PlainPlane {
  real ex[R], ey[R], ez[R], hx[R], hy[R], hz[R];
  real yl_hz, yl_hx;                       // row j0-1
  real yh_ex, yh_ez, yh_hx, yh_hy, yh_hz;  // row j0+R
  real xh_ey, xh_ez, xh_hx, xh_hy, xh_hz;  // column i0+64, lane q: row j0-1+q
  real xl_hz, xl_hy;                       // column i0-1, lane q: tile row q (q = R: row j0+R)
  // x-slab lanes of x-face tiles: the kept parts (yx, zx) of the old E and H
  real pe_yx[R], pe_zx[R], ph_yx[R], ph_zx[R];
  real yh_pe_zx;  // row j0+R: old Ezx
};

// x-slab face cells in the march (tiles at the ends of a row of tiles whose
// lanes reach into an x slab, list word 7 set): the x slabs' kept parts yx,
// zx over the two copies, the x profiles, and the compact index of the x
// slabs (pml_index: base, then Px cells a row over the rows and planes
// between the y and z slabs, the high slab hioff after the low one).
template <typename real>
struct XfArgs {
  const real* eo[2];  // old Eyx, Ezx
  const real* ho[2];  // old Hyx, Hzx
  real* en[2];        // new Eyx, Ezx
  real* hn[2];        // new Hyx, Hzx
  const real *f1, *f2, *g1, *g2;
  real kup;
  int32_t Px, Py, Pz, nym, nzm, base, hioff;
};

// a lane of an x-face tile: in an x slab (on), where its row of parts
// starts, its x profiles
template <typename real>
struct XfLane {
  bool on;
  int sb;
  real f1, f2, g1, g2;
};

// offset of row j, plane k in an x slab's parts (clamped to the slab's rows
// and planes: rows and planes a tile only loads for, never stores, stay in
// bounds)
template <typename real>
__device__ __forceinline__ int xf_row(const XfArgs<real>& x, int j, int k) {
  const int km = min(max(k - x.Pz - 1, 0), x.nzm - 1);
  const int jm = min(max(j - x.Py - 1, 0), x.nym - 1);
  return (km * x.nym + jm) * x.Px;
}

// Every address below lies inside the ghosted box (rows, columns and planes
// are mapped to 0..n+1 before use), so the loads are unconditional.  A value
// loaded for a dead lane, row or plane is read only by that position's own
// new E, which is zeroed; nothing consumes the loaded registers early, so
// the loads stay in flight until the plane that uses them.
template <typename real, int R, bool XF>
__device__ __forceinline__ void plain_load(const FieldArgs<real>& a, const MarchArgs<real>& m, const XfArgs<real>& x,
                                           const XfLane<real>& xl, const MarchTile<R>& w, int k,
                                           PlainPlane<real, R>& P) {
  const int px = a.px;
  const int kb = a.pxy * k;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = kb + w.ip + px * w.jrow[r];
    P.ex[r] = m.eo[0][c];
    P.ey[r] = m.eo[1][c];
    P.ez[r] = m.eo[2][c];
    P.hx[r] = m.ho[0][c];
    P.hy[r] = m.ho[1][c];
    P.hz[r] = m.ho[2][c];
  }
  {  // row j0-1 (a real row or the ghost row 0)
    const int c = kb + w.ip + px * (w.j0 - 1);
    P.yl_hz = m.ho[2][c];
    P.yl_hx = m.ho[0][c];
  }
  if (w.yh_need) {  // row j0+R
    const int c = kb + w.ip + px * w.jyh;
    P.yh_ex = m.eo[0][c];
    P.yh_ez = m.eo[2][c];
    P.yh_hx = m.ho[0][c];
    P.yh_hy = m.ho[1][c];
    P.yh_hz = m.ho[2][c];
  }
  if (w.xh_need) {  // column i0+64, one row per lane 0..R (lane 0: row j0-1, Hx only)
    const int c = kb + w.ixh + px * w.jxg;
    P.xh_hx = m.ho[0][c];
    P.xh_ey = m.eo[1][c];
    P.xh_ez = m.eo[2][c];
    P.xh_hy = m.ho[1][c];
    P.xh_hz = m.ho[2][c];
  }
  {  // column i0-1 (a real column or the ghost column 0), one row per lane
    const int c = kb + (w.i0 - 1) + px * w.jxl;
    P.xl_hz = m.ho[2][c];
    P.xl_hy = m.ho[1][c];
  }
  if constexpr (XF) {
    if (xl.on) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int sc = xl.sb + xf_row(x, w.jrow[r], k);
        P.pe_yx[r] = x.eo[0][sc];
        P.pe_zx[r] = x.eo[1][sc];
        P.ph_yx[r] = x.ho[0][sc];
        P.ph_zx[r] = x.ho[1][sc];
      }
      P.yh_pe_zx = w.yh_need ? x.eo[1][xl.sb + xf_row(x, w.jyh, k)] : (real)0;
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) P.pe_yx[r] = P.pe_zx[r] = P.ph_yx[r] = P.ph_zx[r] = 0;
      P.yh_pe_zx = 0;
    }
  }
}

// registers carried from plane k-1 to plane k
template <typename real, int R>
struct PlainCarry {
  real hx[R], hy[R], hz[R];  // old H of plane k-1
  real ex[R], ey[R], ez[R];  // new E of plane k-1
  real yh_hy, yh_ex, yh_ez;  // row j0+R: old Hy, new Ex, Ez of plane k-1
  real xh_hx, xh_ey, xh_ez;  // column i0+64 (lane q: row j0-1+q): old Hx, new Ey, Ez of plane k-1
  real ph_yx[R], ph_zx[R];   // x-slab lanes: old H parts of plane k-1
};

// material coefficients of a plain tile (wave-uniform) and this plane's
// z-normal source terms, as fused_e_cell applies them to a dielectric cell
template <typename real>
struct PlainCoef {
  real c[3];   // Ku*gax
  real g[3];   // gax (source scaling)
  real r;      // 1/eps (the PML update's medium)
  real sv[3];  // sources of this plane per component
  bool sa[3];
};

template <typename real>
__device__ __forceinline__ real plain_e(const PlainCoef<real>& q, int c, real e0, real cur) {
  real ev = e0 + q.c[c] * cur;
  if (q.sa[c]) ev = ev + q.g[c] * q.sv[c];
  return ev;
}

template <typename real, int R, bool XF>
__device__ __forceinline__ void plain_plane(const FieldArgs<real>& a, const MarchArgs<real>& m,
                                            const XfArgs<real>& x, const XfLane<real>& xl,
                                            const SrcArgs<real>& src, int nsrc, PlainCoef<real> q,
                                            const MarchTile<R>& w, int lane, int k,
                                            const PlainPlane<real, R>& cur, PlainPlane<real, R>& nxt,
                                            PlainCarry<real, R>& c) {
  if (k + 1 <= w.k1 + 1) {
    bool lv;
    plain_load<real, R, XF>(a, m, x, xl, w, march_map(k + 1, a.nz, a.per[2], lv), nxt);
  }
  bool live_k;
  const int kk = march_map(k, a.nz, a.per[2], live_k);
  const bool store_k = k <= w.k1;
  // this plane's z-normal sources (wave-uniform), in source order
  q.sv[0] = q.sv[1] = q.sv[2] = 0;
  q.sa[0] = q.sa[1] = q.sa[2] = false;
  for (int s = 0; s < nsrc; ++s)
    if (src.axis[s] == 2 && src.index[s] == kk) {
      const int cq = src.comp[s];
      const real v = src_value(src, s);
      if (cq == 0) { q.sv[0] += v; q.sa[0] = true; }
      else if (cq == 1) { q.sv[1] += v; q.sa[1] = true; }
      else { q.sv[2] += v; q.sa[2] = true; }
    }
  // new E of plane k on the tile's rows
  real exn[R], eyn[R], ezn[R];
  real pyx[R], pzx[R];  // x-slab lanes: the new E parts yx, zx
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const real hz_jm = r ? cur.hz[r - 1] : cur.yl_hz;
    const real hx_jm = r ? cur.hx[r - 1] : cur.yl_hx;
    real hz_im = lane_prev(cur.hz[r]);
    real hy_im = lane_prev(cur.hy[r]);
    const real lz = lane_at(cur.xl_hz, r), ly = lane_at(cur.xl_hy, r);
    if (lane == 0) {
      hz_im = lz;
      hy_im = ly;
    }
    const bool live = live_k && w.lane_live && w.row_live[r];
    const real hx = cur.hx[r], hy = cur.hy[r], hz = cur.hz[r];
    real ex = plain_e(q, 0, cur.ex[r], (hz - hz_jm) + (-hy + c.hy[r]));  // curlhzy + curlhyz
    real ey = plain_e(q, 1, cur.ey[r], (hx - c.hx[r]) + (-hz + hz_im));  // curlhxz + curlhzx
    real ez = plain_e(q, 2, cur.ez[r], (hy - hy_im) + (-hx + hx_jm));    // curlhyx + curlhxy
    pyx[r] = pzx[r] = 0;
    if constexpr (XF) {
      if (xl.on) {  // a face cell of an x slab: the split update, its x parts kept
        const real f[3] = {cur.ex[r], cur.ey[r], cur.ez[r]};
        const real sp[6] = {cur.pe_yx[r], cur.pe_zx[r], 0, 0, 0, 0};
        real o[6], n[6];
        pml_parts(true, false, false, sp, f, o);
        e_pml_split(xl.f1, xl.f2, (real)1, x.kup, (real)1, x.kup, q.r, o, hx, hy, hz, hz_jm, hx_jm, c.hy[r], c.hx[r],
                    hz_im, hy_im, n);
        ex = n[2] + n[4];
        ey = n[5] + n[0];
        ez = n[1] + n[3];
        pyx[r] = n[0];
        pzx[r] = n[1];
      }
    }
    exn[r] = live ? ex : (real)0;
    eyn[r] = live ? ey : (real)0;
    ezn[r] = live ? ez : (real)0;
  }
  // new Ex, Ez of row j0+R
  real yex = 0, yez = 0;
  if (w.yh_need) {
    real hy_im = lane_prev(cur.yh_hy);
    const real ly = lane_at(cur.xl_hy, R);
    if (lane == 0) hy_im = ly;
    const bool live = live_k && w.lane_live && w.yh_live;
    real ex = plain_e(q, 0, cur.yh_ex, (cur.yh_hz - cur.hz[R - 1]) + (-cur.yh_hy + c.yh_hy));
    real ez = plain_e(q, 2, cur.yh_ez, (cur.yh_hy - hy_im) + (-cur.yh_hx + cur.hx[R - 1]));
    if constexpr (XF) {
      if (xl.on) {  // Ex = Exy + Exz, Ez = Ezx + Ezy (the Ey inputs are not needed)
        const real f[3] = {cur.yh_ex, (real)0, cur.yh_ez};
        const real sp[6] = {(real)0, cur.yh_pe_zx, 0, 0, 0, 0};
        real o[6], n[6];
        pml_parts(true, false, false, sp, f, o);
        e_pml_split(xl.f1, xl.f2, (real)1, x.kup, (real)1, x.kup, q.r, o, cur.yh_hx, cur.yh_hy, cur.yh_hz,
                    cur.hz[R - 1], cur.hx[R - 1], c.yh_hy, (real)0, (real)0, hy_im, n);
        ex = n[2] + n[4];
        ez = n[1] + n[3];
      }
    }
    yex = live ? ex : (real)0;
    yez = live ? ez : (real)0;
  }
  // new Ey, Ez of column i0+64 (lanes 1..R: row j0-1+lane)
  real xey = 0, xez = 0;
  if (w.xh_need) {
    real hzl = 0, hyl = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const real vz = lane_at(cur.hz[r], 63), vy = lane_at(cur.hy[r], 63);
      if (lane == r + 1) {
        hzl = vz;
        hyl = vy;
      }
    }
    const real hx_jm = lane_prev(cur.xh_hx);
    const bool live = live_k && w.xg_live && lane >= 1 && lane <= R;
    const real ey = plain_e(q, 1, cur.xh_ey, (cur.xh_hx - c.xh_hx) + (-cur.xh_hz + hzl));
    const real ez = plain_e(q, 2, cur.xh_ez, (cur.xh_hy - hyl) + (-cur.xh_hx + hx_jm));
    xey = live ? ey : (real)0;
    xez = live ? ez : (real)0;
  }
  // stores of the new E of plane k
  if (store_k) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!(w.lane_own && w.row_own[r])) continue;
      const int cc = w.ip + a.px * w.jrow[r] + a.pxy * kk;
      a.e[0][cc] = exn[r];
      a.e[1][cc] = eyn[r];
      a.e[2][cc] = ezn[r];
      put_e_ghost(a.e[0], exn[r], w.ip, w.jrow[r], kk, cc, a);
      put_e_ghost(a.e[1], eyn[r], w.ip, w.jrow[r], kk, cc, a);
      put_e_ghost(a.e[2], ezn[r], w.ip, w.jrow[r], kk, cc, a);
      if constexpr (XF) {
        if (xl.on) {
          const int sc = xl.sb + xf_row(x, w.jrow[r], kk);
          x.en[0][sc] = pyx[r];
          x.en[1][sc] = pzx[r];
        }
      }
    }
  }
  // new H of plane k-1 (myfdtd_func.f90:45-64)
  if (k > w.k0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const real ez_jp = r < R - 1 ? c.ez[r + 1] : c.yh_ez;
      const real ex_jp = r < R - 1 ? c.ex[r + 1] : c.yh_ex;
      real ez_ip = lane_next(c.ez[r]);
      real ey_ip = lane_next(c.ey[r]);
      const real xz = lane_at(c.xh_ez, r + 1), xy = lane_at(c.xh_ey, r + 1);
      if (lane == 63) {
        ez_ip = xz;
        ey_ip = xy;
      }
      if (!(w.lane_own && w.row_own[r])) continue;
      real hx = c.hx[r] + a.ku * ((eyn[r] - c.ey[r]) + (-ez_jp + c.ez[r]));
      real hy = c.hy[r] + a.ku * ((ez_ip - c.ez[r]) + (-exn[r] + c.ex[r]));
      real hz = c.hz[r] + a.ku * ((ex_jp - c.ex[r]) + (-ey_ip + c.ey[r]));
      const int kh = k - 1;
      if constexpr (XF) {
        if (xl.on) {
          const real f[3] = {c.hx[r], c.hy[r], c.hz[r]};
          const real sp[6] = {c.ph_yx[r], c.ph_zx[r], 0, 0, 0, 0};
          real o[6], n[6];
          pml_parts(true, false, false, sp, f, o);
          h_pml_split(xl.g1, xl.g2, (real)1, x.kup, (real)1, x.kup, o, c.ex[r], c.ey[r], c.ez[r], eyn[r], exn[r],
                      ez_jp, ex_jp, ez_ip, ey_ip, n);
          hx = n[2] + n[4];
          hy = n[5] + n[0];
          hz = n[1] + n[3];
          const int sc = xl.sb + xf_row(x, w.jrow[r], kh);
          x.hn[0][sc] = n[0];
          x.hn[1][sc] = n[1];
        }
      }
      const int cc = w.ip + a.px * w.jrow[r] + a.pxy * kh;
      a.h[0][cc] = hx;
      a.h[1][cc] = hy;
      a.h[2][cc] = hz;
      put_h_ghost(a.h[0], hx, w.ip, w.jrow[r], kh, cc, a);
      put_h_ghost(a.h[1], hy, w.ip, w.jrow[r], kh, cc, a);
      put_h_ghost(a.h[2], hz, w.ip, w.jrow[r], kh, cc, a);
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    c.hx[r] = cur.hx[r];
    c.hy[r] = cur.hy[r];
    c.hz[r] = cur.hz[r];
    c.ex[r] = exn[r];
    c.ey[r] = eyn[r];
    c.ez[r] = ezn[r];
    if constexpr (XF) {
      c.ph_yx[r] = cur.ph_yx[r];
      c.ph_zx[r] = cur.ph_zx[r];
    }
  }
  c.yh_hy = cur.yh_hy;
  c.yh_ex = yex;
  c.yh_ez = yez;
  c.xh_hx = cur.xh_hx;
  c.xh_ey = xey;
  c.xh_ez = xez;
}

// Plain tiles, four independent wavefronts per workgroup (one tile each):
// the wavefront owns the 64 x R cells of its tile over planes k0..k1; per
// plane k it computes the new E of plane k and the new H of plane k-1,
// carrying the old H and the new E of the plane below in registers while
// the loads of plane k+1 are in flight.  Neighbours come from the row
// before / after (registers), from lanes l-1 / l+1 (DPP wave shifts), and
// from halos it loads and recomputes itself: row j0-1 (old Hz, Hx), row
// j0+R (its new Ex, Ez), column i0-1 (old Hz, Hy: lanes 0..R gather one
// row each), column i0+64 (its new Ey, Ez: lanes 1..R), plane k1+1 (its new
// E).  No LDS and no barrier.
template <typename real, int R, bool XF>
__global__ __launch_bounds__(256) void k_march_plain(FieldArgs<real> a, MarchArgs<real> m, SrcArgs<real> src,
                                                     int nsrc, XfArgs<real> x) {
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x;
  const int lb = a.xcdmap ? (b & 7) * m.tper + (b >> 3) : b;
  const int tile = lb * 4 + (int)(threadIdx.x >> 6);
  if (tile >= m.ntiles) return;  // whole wavefront
  const int32_t* T = m.tiles + 8 * tile;
  MarchTile<R> w;
  w.i0 = T[0];
  w.j0 = T[1];
  w.k0 = T[2];
  w.k1 = T[3];
  const int xlo = T[4], xhi = T[5], mid = T[6];
  PlainCoef<real> q;
  {
    const Vec4<real> cf = a.coef[mid], gx = a.gaxe[mid];
    for (int d = 0; d < 3; ++d) {
      q.c[d] = cf.v[d];
      q.g[d] = gx.v[d];
    }
    q.r = cf.v[3];
  }
  // lanes xlo..xhi are owned; lane xhi+1 supplies the new E after the
  // range (the next cell, or the wrapped cell 1)
  w.ip = march_map(w.i0 + lane, a.nx, a.per[0], w.lane_live);
  w.lane_live = w.lane_live && w.i0 + lane >= xlo && w.i0 + lane <= xhi + 1;
  w.lane_own = w.i0 + lane >= xlo && w.i0 + lane <= xhi;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    w.jrow[r] = march_map(w.j0 + r, a.ny, a.per[1], w.row_live[r]);
    w.row_own[r] = w.j0 + r <= a.ny;
  }
  // halo row j0+R: needed when the last tile row is owned
  w.yh_need = w.j0 + R - 1 <= a.ny;
  w.jyh = march_map(w.j0 + R, a.ny, a.per[1], w.yh_live);
  // halo column i0+64: needed when lane 63 is owned
  w.xh_need = xhi == w.i0 + 63;
  bool xh_live;
  w.ixh = march_map(w.i0 + 64, a.nx, a.per[0], xh_live);
  // gather rows: column i0+64 lane q -> row j0-1+q (q = 0..R; lane 0 only
  // supplies Hx), column i0-1 lane q -> tile row q (q < R) or the halo row
  {
    const int g = min(lane, R);
    bool lv;
    w.jxg = g == 0 ? w.j0 - 1 : march_map(w.j0 + g - 1, a.ny, a.per[1], lv);
    w.xg_live = xh_live && (g == 0 || lv);
    w.jxl = march_map(w.j0 + g, a.ny, a.per[1], lv);
  }
  PlainCarry<real, R> c;
  {
    // old Hx, Hy of plane k0-1 (a real plane or the ghost plane 0) for the
    // k-1 terms of plane k0
    const int kb = a.pxy * (w.k0 - 1);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int cc = kb + w.ip + a.px * w.jrow[r];
      c.hx[r] = m.ho[0][cc];
      c.hy[r] = m.ho[1][cc];
      c.hz[r] = 0;
      c.ex[r] = c.ey[r] = c.ez[r] = 0;
    }
    c.yh_hy = m.ho[1][kb + w.ip + a.px * w.jyh];
    c.yh_ex = c.yh_ez = 0;
    c.xh_hx = m.ho[0][kb + w.ixh + a.px * w.jxg];
    c.xh_ey = c.xh_ez = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) c.ph_yx[r] = c.ph_zx[r] = 0;
  }
  // x-face tiles (word 7): lanes in an x slab take the split update; their
  // column of the x slab's parts and their x profiles are fixed per lane
  XfLane<real> xl{};
  if constexpr (XF) {
    const int i = w.ip;
    xl.on = T[7] != 0 && (i <= x.Px || i > a.nx - x.Px);
    if (xl.on) {
      const bool hi = i > x.Px;  // as pml_index (the slabs never overlap: 2 Px <= nx)
      xl.sb = x.base + (hi ? x.hioff + (i - (a.nx - x.Px + 1)) : i - 1);
      xl.f1 = x.f1[i];
      xl.f2 = x.f2[i];
      xl.g1 = x.g1[i];
      xl.g2 = x.g2[i];
    }
  }
  PlainPlane<real, R> A, B;
  plain_load<real, R, XF>(a, m, x, xl, w, w.k0, A);
  for (int k = w.k0; k <= w.k1 + 1; k += 2) {
    plain_plane<real, R, XF>(a, m, x, xl, src, nsrc, q, w, lane, k, A, B, c);
    if (k + 1 <= w.k1 + 1) plain_plane<real, R, XF>(a, m, x, xl, src, nsrc, q, w, lane, k + 1, B, A, c);
  }
}

// General work: workgroup = 256 threads over cells x0..x1 x j0..j1 of
// planes k0..k1 (list entry: x0, x1, j0, j1, k0, width, k1, face), 64 x 4 cells
// (a wavefront is one row), 32 x 8 or 16 x 16 (narrow boxes: the x edges),
// the fused cell updates over the two copies.  k_gen_e reads the old H and
// E and writes the new E, one cell per thread, a workgroup per plane (the
// planes of an entry dealt to neighbouring workgroups, so a plane's
// neighbour below is still in L2);
// k_gen_h then reads the new E and the old H and writes the new H, each
// thread marching its (i, j) column up through the planes, pipelined by
// hand: the loads of plane k+1 are issued before the arithmetic and stores
// of plane k (no plane reads what the kernel writes, so the order is
// free), and E of the plane above comes from registers (10 % faster than a
// cell per thread; the E update marched so holds too many registers and
// runs slower).
template <typename real>
__device__ __forceinline__ bool gen_cell(const FieldArgs<real>& a, const MarchArgs<real>& m, int& i, int& j, int& k0,
                                         int& k1, bool& wide, int& pl, int& face) {
  const int b = blockIdx.x;
  int lb = a.xcdmap ? (b & 7) * m.tper + (b >> 3) : b;
  if (lb >= m.ntiles * m.nsplit) return false;
  pl = lb % m.nsplit;
  lb /= m.nsplit;
  const int32_t* T = m.tiles + 8 * lb;
  const int bw = T[5];
  wide = bw == 64;
  const int t = (int)threadIdx.x;
  i = T[0] + (t & (bw - 1));
  j = T[2] + (wide ? __builtin_amdgcn_readfirstlane(t >> 6) : t / bw);
  k0 = T[4];
  k1 = T[6];
  face = T[7];
  return i <= T[1] && j <= T[3];
}

// FP32: eight wavefronts per SIMD (the compiler keeps some kernel arguments
// in vector registers to fit), 8 % faster than the seven its register use
// otherwise allows; the kernel waits on memory for most of its life
template <typename real>
__global__ __launch_bounds__(256, sizeof(real) == 4 ? 8 : 5) void k_gen_e(FieldArgs<real> a, PmlArgs<real> p, PmlArgs<real> pw,
                                               MarchArgs<real> m, SrcArgs<real> src, int nsrc) {
  int i, j, k0, k1, pl, face;
  bool wide;
  if (!gen_cell(a, m, i, j, k0, k1, wide, pl, face)) return;
  const int k = k0 + pl;
  if (k > k1) return;
  real sv[3] = {0, 0, 0};
  int sa[3] = {-1, -1, -1};
  real cy = 0, cx = 0;
  const int rm = wide ? a.rowmat[__builtin_amdgcn_readfirstlane(j + (a.ny + 2) * k)] : -1;
  // face cells of one slab in rows of one dielectric, no source plane
  // (build_march_tiles); the row's id is read with scalar loads, so only
  // 64-wide rows with an id take this path
  if (face && wide && rm >= 0) {
    ECellIn<real> in;
    if (face == 2)
      e_cell_load<real, true, 2>(a, p, sa, rm, i, j, k, false, cy, cx, a.h[0], a.h[1], a.h[2], m.eo[0], m.eo[1],
                                 m.eo[2], in);
    else
      e_cell_load<real, true, 1>(a, p, sa, rm, i, j, k, false, cy, cx, a.h[0], a.h[1], a.h[2], m.eo[0], m.eo[1],
                                 m.eo[2], in);
    e_cell_store(a, pw, true, sv, sa, i, j, k, in, a.e[0], a.e[1], a.e[2]);
    return;
  }
  plane_sources(src, nsrc, i, j, k, sv, sa);
  if (rm >= 0)
    fused_e_cell<real, true>(a, p, pw, true, sv, sa, rm, i, j, k, cy, cx, false, a.h[0], a.h[1], a.h[2], m.eo[0],
                             m.eo[1], m.eo[2], a.e[0], a.e[1], a.e[2]);
  else
    fused_e_cell<real, false>(a, p, pw, true, sv, sa, -1, i, j, k, cy, cx, false, a.h[0], a.h[1], a.h[2], m.eo[0],
                              m.eo[1], m.eo[2], a.e[0], a.e[1], a.e[2]);
}

// one (i, j) column of planes k0..k1, pipelined by hand (see above)
template <typename real, int FACE>
__device__ __forceinline__ void gen_h_column(const FieldArgs<real>& a, const PmlArgs<real>& p, const PmlArgs<real>& pw,
                                             const MarchArgs<real>& m, int i, int j, int k0, int k1) {
  const DftFuse<real> df{};
  HCellIn<real> cur, nxt;
  h_cell_load<real, FACE>(a, p, i, j, k0, false, (real)0, (real)0, a.e[0], a.e[1], a.e[2], m.ho[0], m.ho[1],
                          m.ho[2], cur);
  for (int k = k0; k <= k1; ++k) {
    const int kn = min(k + 1, k1);
    h_cell_load<real, FACE>(a, p, i, j, kn, kn > k, cur.ex_kp, cur.ey_kp, a.e[0], a.e[1], a.e[2], m.ho[0],
                            m.ho[1], m.ho[2], nxt);
    h_cell_store(a, pw, df, i, j, k, cur, a.h[0], a.h[1], a.h[2]);
    cur = nxt;
  }
}

template <typename real>
__global__ __launch_bounds__(256) void k_gen_h(FieldArgs<real> a, PmlArgs<real> p, PmlArgs<real> pw,
                                               MarchArgs<real> m) {
  int i, j, k0, k1, pl, face;
  bool wide;
  if (!gen_cell(a, m, i, j, k0, k1, wide, pl, face)) return;
  if (face == 2)
    gen_h_column<real, 2>(a, p, pw, m, i, j, k0, k1);
  else if (face == 1)
    gen_h_column<real, 1>(a, p, pw, m, i, j, k0, k1);
  else
    gen_h_column<real, 0>(a, p, pw, m, i, j, k0, k1);
}

// ---------------------------------------------------------------------------
// Dispersive cells: Drude + up to 4 critical points + Raman
// (mod3/gold_mod.f90:61-83, mod3/graphene_mod.f90:70-100).  State SoA:
// word w of component q for cell e at state[(w*3+q)*n + e];
// words: [D (FUSED only)], I, S, g_p re, g_p im (p < ncp), sn, snm2 (Raman).
// The state and the update are FP64 in both precisions: inside a metal E is
// the small difference of the large D and I, which FP32 would cancel away
// (an FP32 state gave 6e-5 relative spectra error on the 16-disk case);
// E is stored in field precision.
template <typename real, bool FUSED>
__device__ __forceinline__ void ade_cell(const FieldArgs<real>& a, const AdeCoef<double>& co, const int32_t* cells,
                                         const uint8_t* masks, double* state, int64_t n, const SrcArgs<real>& src,
                                         int nsrc, int64_t e) {
  const int64_t c = cells[e];
  const int mask = masks[e];
  const int k = (int)(c / a.pxy);
  const int64_t rem = c - (int64_t)k * a.pxy;
  const int j = (int)(rem / a.px);
  const int i = (int)(rem - (int64_t)j * a.px);
  double cur[3] = {0, 0, 0};
  double sv[3] = {0, 0, 0};
  if (FUSED) {
    const double hx = a.h[0][c], hy = a.h[1][c], hz = a.h[2][c];
    cur[0] = (hz - (double)a.h[2][c - a.px]) + (-hy + (double)a.h[1][c - a.pxy]);
    cur[1] = (hx - (double)a.h[0][c - a.pxy]) + (-hz + (double)a.h[2][c - 1]);
    cur[2] = (hy - (double)a.h[1][c - 1]) + (-hx + (double)a.h[0][c - a.px]);
    for (int q = 0; q < nsrc; ++q) {
      const int pos = src.axis[q] == 0 ? i : (src.axis[q] == 1 ? j : k);
      if (pos == src.index[q]) sv[src.comp[q]] += (double)src_value(src, q);
    }
  }
  const int w0 = FUSED ? 1 : 0;  // word offset of I
  const int ncp = co.ncp;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!(mask & (1 << q))) continue;
    double dv;
    if (FUSED) {
      double* dp = state + (int64_t)(0 * 3 + q) * n + e;
      dv = *dp + co.ku * cur[q];
      dv = dv + sv[q];
      *dp = dv;
    } else {
      dv = a.d[q][c];
    }
    double* Ip = state + (int64_t)((w0 + 0) * 3 + q) * n + e;
    double* Sp = state + (int64_t)((w0 + 1) * 3 + q) * n + e;
    const double Iv = *Ip, Sv = *Sp;
    double xr = (dv - Iv) + co.ha * Sv;
    double xi = 0;
    double gr[4], gi[4];
#pragma unroll
    for (int pp = 0; pp < 4; ++pp) {
      if (pp >= ncp) break;
      gr[pp] = state[(int64_t)((w0 + 2 + pp) * 3 + q) * n + e];
      gi[pp] = state[(int64_t)((w0 + 2 + ncp + pp) * 3 + q) * n + e];
      xr = xr + (co.gbr[pp] * gr[pp] - co.gbi[pp] * gi[pp]);
      xi = xi + (co.gbr[pp] * gi[pp] + co.gbi[pp] * gr[pp]);
    }
    double snv = 0, snm2 = 0;
    double* snp = nullptr;
    double* snm2p = nullptr;
    if (co.raman) {
      snp = state + (int64_t)((w0 + 2 + 2 * ncp) * 3 + q) * n + e;
      snm2p = state + (int64_t)((w0 + 3 + 2 * ncp) * 3 + q) * n + e;
      snv = *snp;
      snm2 = *snm2p;
    }
    // E = Re[X / (ga + sn)]
    const double dr = co.gar + snv, di = co.gai;
    const double ev = (xr * dr + xi * di) / (dr * dr + di * di);
    a.e[q][c] = (real)ev;
    put_e_ghost(a.e[q], (real)ev, i, j, k, c, a);
    if (co.raman) {
      const double nsn = co.r1 * snv + co.r2 * snm2 + co.r0 * (ev * ev);
      *snm2p = snv;
      *snp = nsn;
    }
    *Ip = Iv + co.hb * ev;
    *Sp = co.ha * Sv + co.hb * ev;
#pragma unroll
    for (int pp = 0; pp < 4; ++pp) {
      if (pp >= ncp) break;
      state[(int64_t)((w0 + 2 + pp) * 3 + q) * n + e] = (co.gbr[pp] * gr[pp] - co.gbi[pp] * gi[pp]) + co.gcr[pp] * ev;
      state[(int64_t)((w0 + 2 + ncp + pp) * 3 + q) * n + e] =
          (co.gbr[pp] * gi[pp] + co.gbi[pp] * gr[pp]) + co.gci[pp] * ev;
    }
  }
}

template <typename real, bool FUSED>
__global__ __launch_bounds__(256) void k_ade(FieldArgs<real> a, AdeCoef<double> co, const int32_t* cells,
                                             const uint8_t* masks, double* state, int64_t n, SrcArgs<real> src,
                                             int nsrc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  ade_cell<real, FUSED>(a, co, cells, masks, state, n, src, nsrc, e);
}

// Several groups in one launch (the one-pass step): group q takes blocks
// blk[q] .. blk[q+1]-1.  A small group alone is a short launch that leaves
// most of the GPU idle, and every launch boundary costs a drain.
constexpr int kAdeMulti = 4;
struct AdeMulti {
  int32_t ng;
  int32_t blk[kAdeMulti + 1];
  AdeCoef<double> co[kAdeMulti];
  const int32_t* cells[kAdeMulti];
  const uint8_t* masks[kAdeMulti];
  double* state[kAdeMulti];
  int64_t n[kAdeMulti];
};

template <typename real>
__global__ __launch_bounds__(256) void k_ade_multi(FieldArgs<real> a, AdeMulti g, SrcArgs<real> src, int nsrc) {
  const int b = (int)blockIdx.x;
  int q = 0;
  while (q + 1 < g.ng && b >= g.blk[q + 1]) ++q;
  const int64_t e = (int64_t)(b - g.blk[q]) * blockDim.x + threadIdx.x;
  if (e >= g.n[q]) return;
  ade_cell<real, true>(a, g.co[q], g.cells[q], g.masks[q], g.state[q], g.n[q], src, nsrc, e);
}

// ---------------------------------------------------------------------------
// Lorentz poles in D form (mod3/si_mod.f90:46-82): E = D - S_1 - ... - S_np
// with the S of the previous step, then per pole S_l <- ga_l*S_l +
// gb_l*S_l(older) + gc_l*E, in the module's operation order.  FUSED: D is
// kept in state word 0 and advanced here from curl H (E-form engine), as
// k_ade does.  State SoA, word w of component q for cell e at
// state[(w*3+q)*n + e]; words: [D], S_l (l < np), S_l one step older.  FP64
// in both precisions; E is stored in field precision.
template <typename real, bool FUSED>
__global__ __launch_bounds__(256) void k_lord(FieldArgs<real> a, LordCoef co, const int32_t* cells,
                                              const uint8_t* masks, double* state, int64_t n, SrcArgs<real> src,
                                              int nsrc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t c = cells[e];
  const int mask = masks[e];
  const int k = (int)(c / a.pxy);
  const int64_t rem = c - (int64_t)k * a.pxy;
  const int j = (int)(rem / a.px);
  const int i = (int)(rem - (int64_t)j * a.px);
  double cur[3] = {0, 0, 0};
  double sv[3] = {0, 0, 0};
  if (FUSED) {
    const double hx = a.h[0][c], hy = a.h[1][c], hz = a.h[2][c];
    cur[0] = (hz - (double)a.h[2][c - a.px]) + (-hy + (double)a.h[1][c - a.pxy]);
    cur[1] = (hx - (double)a.h[0][c - a.pxy]) + (-hz + (double)a.h[2][c - 1]);
    cur[2] = (hy - (double)a.h[1][c - 1]) + (-hx + (double)a.h[0][c - a.px]);
    for (int q = 0; q < nsrc; ++q) {
      const int pos = src.axis[q] == 0 ? i : (src.axis[q] == 1 ? j : k);
      if (pos == src.index[q]) sv[src.comp[q]] += (double)src_value(src, q);
    }
  }
  const int w0 = FUSED ? 1 : 0;
  const int np = co.npols;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!(mask & (1 << q))) continue;
    double dv;
    if (FUSED) {
      double* dp = state + (int64_t)q * n + e;
      dv = *dp + co.ku * cur[q];
      dv = dv + sv[q];
      *dp = dv;
    } else {
      dv = a.d[q][c];
    }
    double s1[4], s2[4];
    double ev = dv;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      if (l >= np) break;
      s1[l] = state[(int64_t)((w0 + l) * 3 + q) * n + e];
      s2[l] = state[(int64_t)((w0 + np + l) * 3 + q) * n + e];
      ev = ev - s1[l];
    }
    a.e[q][c] = (real)ev;
    put_e_ghost(a.e[q], (real)ev, i, j, k, c, a);
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      if (l >= np) break;
      state[(int64_t)((w0 + l) * 3 + q) * n + e] = (co.ga[l] * s1[l] + co.gb[l] * s2[l]) + co.gc[l] * ev;
      state[(int64_t)((w0 + np + l) * 3 + q) * n + e] = s1[l];
    }
  }
}

// ---------------------------------------------------------------------------
// Lorentz-Jp cells, E form (mod3/goldLorentzJp_mod.f90:37-91 and the family:
// grapheneLorentzJp_mod.f90:42-101 adds the third-order oscillator and
// drives Ez with JpT_x).  The pole state is FP64 in both precisions (an FP32
// P would lose the small per-step increments dt*Jp of slow poles); the E
// update is evaluated in FP64 and stored in field precision.  State SoA,
// word w of component q for cell e at state[(w*3+q)*n + e]; words: P_l
// (l < npols), Jp_l, [Q, G].  JpT is the running sum of the Jp_l, summed in
// pole order as the modules accumulate it.
template <typename real>
__global__ __launch_bounds__(256) void k_ljp(FieldArgs<real> a, LjpCoef<double> co, const int32_t* cells,
                                             const uint8_t* masks, double* state, int64_t n, SrcArgs<real> src,
                                             int nsrc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t c = cells[e];
  const int mask = masks[e];
  const int k = (int)(c / a.pxy);
  const int64_t rem = c - (int64_t)k * a.pxy;
  const int j = (int)(rem / a.px);
  const int i = (int)(rem - (int64_t)j * a.px);
  const real hx = a.h[0][c], hy = a.h[1][c], hz = a.h[2][c];
  const real cur[3] = {(hz - a.h[2][c - a.px]) + (-hy + a.h[1][c - a.pxy]),   // curlhzy + curlhyz
                       (hx - a.h[0][c - a.pxy]) + (-hz + a.h[2][c - 1]),     // curlhxz + curlhzx
                       (hy - a.h[1][c - 1]) + (-hx + a.h[0][c - a.px])};     // curlhyx + curlhxy
  real sv[3] = {0, 0, 0};
  int hit = 0;  // bit q: a source plane through this cell feeds E_q
  for (int q = 0; q < nsrc; ++q) {
    const int pos = src.axis[q] == 0 ? i : (src.axis[q] == 1 ? j : k);
    if (pos == src.index[q]) {
      sv[src.comp[q]] += src_value(src, q);
      hit |= 1 << src.comp[q];
    }
  }
  const int np = co.npols;
  const int wq = 2 * np;  // word of Q
  double jt[3] = {0, 0, 0};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!((co.jmask >> q) & 1)) continue;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      if (l >= np) break;
      jt[q] = jt[q] + state[(int64_t)((np + l) * 3 + q) * n + e];
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!(mask & (1 << q))) continue;
    const double jd = (q == 2 && co.ezjx) ? jt[0] : jt[q];
    double ev = (double)a.e[q][c] + co.kuga * ((double)cur[q] - co.ds * jd);
    if ((hit >> q) & 1) ev = ev + co.ga * (double)sv[q];
    const real er = (real)ev;
    a.e[q][c] = er;
    put_e_ghost(a.e[q], er, i, j, k, c, a);
    double Qv = 0;
    if (co.nl) Qv = state[(int64_t)(wq * 3 + q) * n + e];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      if (l >= np) break;
      double* Pp = state + (int64_t)(l * 3 + q) * n + e;
      double* Jp = state + (int64_t)((np + l) * 3 + q) * n + e;
      double p0 = (co.xi[l] * ev) / co.c0;
      if (l == 0 && co.nl) p0 = p0 + ((co.xi3 * Qv) * ev) / co.c0;
      const double jn = co.g1[l] * *Jp + co.g2[l] * (p0 - *Pp);
      *Jp = jn;
      *Pp = *Pp + co.dt * jn;
    }
    if (co.nl) {
      double* Gp = state + (int64_t)((wq + 1) * 3 + q) * n + e;
      const double ze = co.z0 * ev;
      const double gn = co.ng1 * *Gp + co.ng2 * (ze * ze - Qv);
      *Gp = gn;
      state[(int64_t)(wq * 3 + q) * n + e] = Qv + co.dt * gn;
    }
  }
}

// ---------------------------------------------------------------------------
// STAGED engine
// updateD, sim3/include/myfdtd_func.f90:19-41, over rows [j0, j1]
template <typename real>
__global__ __launch_bounds__(256) void k_staged_d(FieldArgs<real> a, int j0, int j1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + j0;
  const int k = blockIdx.z + 1;
  if (i > a.nx || j > j1) return;
  const int64_t c = i + a.px * j + a.pxy * k;
  const real hx = a.h[0][c], hy = a.h[1][c], hz = a.h[2][c];
  const real curlhzy = hz - a.h[2][c - a.px];
  const real curlhyz = -hy + a.h[1][c - a.pxy];
  const real curlhxz = hx - a.h[0][c - a.pxy];
  const real curlhzx = -hz + a.h[2][c - 1];
  const real curlhyx = hy - a.h[1][c - 1];
  const real curlhxy = -hx + a.h[0][c - a.px];
  a.d[0][c] = a.d[0][c] + a.ku * (curlhzy + curlhyz);
  a.d[1][c] = a.d[1][c] + a.ku * (curlhxz + curlhzx);
  a.d[2][c] = a.d[2][c] + a.ku * (curlhyx + curlhxy);
}

// plane source, sim3/program.f90:396-406 (dval: value dscale * *dval read
// here, the TFSF consistency terms)
template <typename real>
__global__ void k_staged_src(FieldArgs<real> a, int axis, int index, int comp, real val, const double* dval,
                             double dscale) {
  if (dval) val = (real)(dscale * *dval);
  const int u = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int v = blockIdx.y + 1;
  int p[3];
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  const int n[3] = {a.nx, a.ny, a.nz};
  if (u > n[a1]) return;
  p[axis] = index;
  p[a1] = u;
  p[a2] = v;
  const int64_t c = p[0] + a.px * p[1] + a.pxy * p[2];
  a.d[comp][c] = a.d[comp][c] + val;
}

// ConstantDielectric%update_E for every cell/component whose last writer is
// a dielectric (mod3/simpledielectric_mod.f90:181-183).
template <typename real>
__global__ __launch_bounds__(256) void k_staged_mat(FieldArgs<real> a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > a.nx || j > a.ny) return;
  const int c = i + a.px * j + a.pxy * k;
  const unsigned id = a.matid[c];
  const uint32_t fl = a.flags[id];
  const Vec4<real> gx = a.gaxe[id];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!(fl & (1u << q))) continue;
    const real ev = a.d[q][c] * gx.v[q];
    a.e[q][c] = ev;
    put_e_ghost(a.e[q], ev, i, j, k, c, a);
  }
}

// updateH, sim3/include/myfdtd_func.f90:45-64, over rows [j0, j1]
template <typename real>
__global__ __launch_bounds__(256) void k_staged_h(FieldArgs<real> a, int j0, int j1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + j0;
  const int k = blockIdx.z + 1;
  if (i > a.nx || j > j1) return;
  const int64_t c = i + a.px * j + a.pxy * k;
  const real ex = a.e[0][c], ey = a.e[1][c], ez = a.e[2][c];
  const real curleyz = a.e[1][c + a.pxy] - ey;
  const real curlezy = -a.e[2][c + a.px] + ez;
  const real curlezx = a.e[2][c + 1] - ez;
  const real curlexz = -a.e[0][c + a.pxy] + ex;
  const real curlexy = a.e[0][c + a.px] - ex;
  const real curleyx = -a.e[1][c + 1] + ey;
  const real hx = a.h[0][c] + a.ku * (curleyz + curlezy);
  const real hy = a.h[1][c] + a.ku * (curlezx + curlexz);
  const real hz = a.h[2][c] + a.ku * (curlexy + curleyx);
  a.h[0][c] = hx;
  a.h[1][c] = hy;
  a.h[2][c] = hz;
  put_h_ghost(a.h[0], hx, i, j, k, c, a);
  put_h_ghost(a.h[1], hy, i, j, k, c, a);
  put_h_ghost(a.h[2], hz, i, j, k, c, a);
}

// Reference y-slab PML, E: updateEs (mod3/BERpmly_mod.f90:117-159) driven by
// split H only, then setboundayEya/Eyb (:200-225,255-278).  Slab 0 holds
// rows 0..P+1, slab 1 rows jb-1..ny+1; split ghosts stay zero.
template <typename real>
__global__ __launch_bounds__(256) void k_ref_pml_e(FieldArgs<real> a, RefPmlArgs<real> p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int jr = blockIdx.y * blockDim.y + threadIdx.y;  // 0..2P-1
  const int k = blockIdx.z + 1;
  if (i > a.nx || jr >= 2 * p.P) return;
  const int sl = jr < p.P ? 0 : 1;
  const int j = sl == 0 ? jr + 1 : p.jb + (jr - p.P);
  const int jl = sl == 0 ? j : j - (p.jb - 1);
  real* const* S = p.s[sl];
  const real* pr = p.prof[sl];
  const int ps = p.ps;
  const int64_t o = i + p.sx * jl + p.sxy * k;
  const int64_t ojm = o - p.sx, okm = o - p.sxy, oim = o - 1;
  const real curlHzy_p = S[S_HZX][o] + S[S_HZY][o];
  const real curlHzy_m = -S[S_HZX][ojm] - S[S_HZY][ojm];
  const real curlHyz_p = -S[S_HYZ][o] - S[S_HYX][o];
  const real curlHyz_m = S[S_HYZ][okm] + S[S_HYX][okm];
  const real curlHxz_p = S[S_HXY][o] + S[S_HXZ][o];
  const real curlHxz_m = -S[S_HXY][okm] - S[S_HXZ][okm];
  const real curlHzx_p = -S[S_HZX][o] - S[S_HZY][o];
  const real curlHzx_m = S[S_HZX][oim] + S[S_HZY][oim];
  const real curlHyx_p = S[S_HYZ][o] + S[S_HYX][o];
  const real curlHyx_m = -S[S_HYZ][oim] - S[S_HYX][oim];
  const real curlHxy_p = -S[S_HXY][o] - S[S_HXZ][o];
  const real curlHxy_m = S[S_HXY][ojm] + S[S_HXZ][ojm];
  const real f1x = pr[0 * ps + i], f2x = pr[1 * ps + i];
  const real f1y = pr[2 * ps + j], f2y = pr[3 * ps + j];
  const real f1z = pr[4 * ps + k], f2z = pr[5 * ps + k];
  S[S_EYX][o] = f1x * S[S_EYX][o] + f2x * (curlHzx_p + curlHzx_m);
  S[S_EZX][o] = f1x * S[S_EZX][o] + f2x * (curlHyx_p + curlHyx_m);
  const real exy = f1y * S[S_EXY][o] + f2y * (curlHzy_p + curlHzy_m);
  const real ezy = f1y * S[S_EZY][o] + f2y * (curlHxy_p + curlHxy_m);
  S[S_EXY][o] = exy;
  S[S_EZY][o] = ezy;
  S[S_EXZ][o] = f1z * S[S_EXZ][o] + f2z * (curlHyz_p + curlHyz_m);
  S[S_EYZ][o] = f1z * S[S_EYZ][o] + f2z * (curlHxz_p + curlHxz_m);
  const int64_t c = i + a.px * j + a.pxy * k;
  if (sl == 0 && j == p.P) {
    S[S_EXY][o + p.sx] = a.e[0][c + a.px];
    S[S_EZY][o + p.sx] = a.e[2][c + a.px];
    a.e[0][c] = exy;
    a.e[2][c] = ezy;
  } else if (sl == 1 && j == p.jb) {
    S[S_EXY][o - p.sx] = a.e[0][c - a.px];
    S[S_EZY][o - p.sx] = a.e[2][c - a.px];
    a.e[0][c] = exy;
    a.e[2][c] = ezy;
  }
}

// Reference y-slab PML, H: updateHs (:161-195) and setboundayHya/Hyb.
template <typename real>
__global__ __launch_bounds__(256) void k_ref_pml_h(FieldArgs<real> a, RefPmlArgs<real> p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int jr = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z + 1;
  if (i > a.nx || jr >= 2 * p.P) return;
  const int sl = jr < p.P ? 0 : 1;
  const int j = sl == 0 ? jr + 1 : p.jb + (jr - p.P);
  const int jl = sl == 0 ? j : j - (p.jb - 1);
  real* const* S = p.s[sl];
  const real* pr = p.prof[sl];
  const int ps = p.ps;
  const int64_t o = i + p.sx * jl + p.sxy * k;
  const int64_t ojp = o + p.sx, okp = o + p.sxy, oip = o + 1;
  const real curlEzy_p = -S[S_EZX][ojp] - S[S_EZY][ojp];
  const real curlEzy_m = S[S_EZX][o] + S[S_EZY][o];
  const real curlEyz_p = S[S_EYZ][okp] + S[S_EYX][okp];
  const real curlEyz_m = -S[S_EYZ][o] - S[S_EYX][o];
  const real curlExz_p = -S[S_EXY][okp] - S[S_EXZ][okp];
  const real curlExz_m = S[S_EXY][o] + S[S_EXZ][o];
  const real curlEzx_p = S[S_EZX][oip] + S[S_EZY][oip];
  const real curlEzx_m = -S[S_EZX][o] - S[S_EZY][o];
  const real curlEyx_p = -S[S_EYZ][oip] - S[S_EYX][oip];
  const real curlEyx_m = S[S_EYZ][o] + S[S_EYX][o];
  const real curlExy_p = S[S_EXY][ojp] + S[S_EXZ][ojp];
  const real curlExy_m = -S[S_EXY][o] - S[S_EXZ][o];
  const real g1x = pr[6 * ps + i], g2x = pr[7 * ps + i];
  const real g1y = pr[8 * ps + j], g2y = pr[9 * ps + j];
  const real g1z = pr[10 * ps + k], g2z = pr[11 * ps + k];
  S[S_HYX][o] = g1x * S[S_HYX][o] + g2x * (curlEzx_m + curlEzx_p);
  S[S_HZX][o] = g1x * S[S_HZX][o] + g2x * (curlEyx_m + curlEyx_p);
  const real hxy = g1y * S[S_HXY][o] + g2y * (curlEzy_m + curlEzy_p);
  const real hzy = g1y * S[S_HZY][o] + g2y * (curlExy_m + curlExy_p);
  S[S_HXY][o] = hxy;
  S[S_HZY][o] = hzy;
  S[S_HXZ][o] = g1z * S[S_HXZ][o] + g2z * (curlEyz_m + curlEyz_p);
  S[S_HYZ][o] = g1z * S[S_HYZ][o] + g2z * (curlExz_m + curlExz_p);
  const int64_t c = i + a.px * j + a.pxy * k;
  if (sl == 0 && j == p.P) {
    S[S_HXY][o + p.sx] = a.h[0][c + a.px];
    S[S_HZY][o + p.sx] = a.h[2][c + a.px];
    a.h[0][c] = hxy;
    a.h[2][c] = hzy;
  } else if (sl == 1 && j == p.jb) {
    S[S_HXY][o - p.sx] = a.h[0][c - a.px];
    S[S_HZY][o - p.sx] = a.h[2][c - a.px];
    a.h[0][c] = hxy;
    a.h[2][c] = hzy;
  }
}

// ---------------------------------------------------------------------------
// Monitors: running DFT on a plane (sim3/program.f90:456-472), time-blocked:
// every gated step copies the plane's selected components into a ring of
// kDftBlock samples; a flush adds dt * sum_b cos/sin(w t_b) F_b, t_b =
// (n_b - 1) dt, to the FP64 accumulators [slot][re|im][m][v][u].
// Every monitor of a step in one launch (each kernel boundary costs about
// 9 us on this GPU): the samples of up to kMonPlanes DFT planes, each laid
// out as samples[slot][v][u] (u, v the plane's two free axes in x, y, z
// order), then the probes.  Plane q owns blocks blk[q] .. blk[q+1]-1,
// (np+255)/256 per sampled component; the blocks from blk[nplanes] on write
// the probes.
constexpr int kMonPlanes = 8;
template <typename real>
struct MonList {
  int32_t nplanes, nprobe;
  int32_t axis[kMonPlanes], index[kMonPlanes], n1[kMonPlanes];
  int64_t np[kMonPlanes];
  int32_t comp[kMonPlanes][6];
  real* dst[kMonPlanes];
  int32_t blk[kMonPlanes + 1];
  const int32_t* ppos;
  double* pout;
  double tn;
};

template <typename real>
__global__ __launch_bounds__(256) void k_monitors(FieldArgs<real> a, MonList<real> m) {
  const int b = blockIdx.x;
  if (b >= m.blk[m.nplanes]) {  // probes (incident-line probes are written by k_tfsf_h)
    for (int p = (b - m.blk[m.nplanes]) * 256 + (int)threadIdx.x; p < m.nprobe; p += 256 * (int)gridDim.x) {
      if (m.ppos[3 * p] < 0) continue;
      const int64_t c = m.ppos[3 * p] + a.px * m.ppos[3 * p + 1] + a.pxy * m.ppos[3 * p + 2];
      double* r = m.pout + 7 * p;
      r[0] = m.tn;
      for (int q = 0; q < 3; ++q) {
        r[1 + q] = (double)a.e[q][c];
        r[4 + q] = (double)a.h[q][c];
      }
    }
    return;
  }
  int q = 0;
  while (q + 1 < m.nplanes && b >= m.blk[q + 1]) ++q;
  const int64_t np = m.np[q];
  const int nbx = (int)((np + 255) / 256);
  const int lb = b - m.blk[q];
  const int slot = lb / nbx;
  const int64_t o = (int64_t)(lb - slot * nbx) * 256 + threadIdx.x;
  if (o >= np) return;
  const int n1 = m.n1[q], axis = m.axis[q];
  const int u = (int)(o % n1) + 1, v = (int)(o / n1) + 1;
  int pc[3];
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  pc[axis] = m.index[q];
  pc[a1] = u;
  pc[a2] = v;
  const int64_t c = pc[0] + a.px * pc[1] + a.pxy * pc[2];
  const int comp = m.comp[q][slot];
  const real* F = comp < 3 ? a.e[comp] : a.h[comp - 3];
  m.dst[q][(int64_t)slot * np + o] = F[c];
}

// cos/sin table of one block of samples: tab[(b*nfreq + m)*2 + {0,1}]
__global__ void k_dft_table(const double* omega, int nfreq, int n_first, int count, double dt, double* tab) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nfreq * count) return;
  const int b = q / nfreq, m = q - b * nfreq;
  const double t = (double)(n_first + b - 1) * dt;
  double sw, cw;
  sincos(omega[m] * t, &sw, &cw);
  tab[2 * q] = cw;
  tab[2 * q + 1] = sw;
}

// The fold of a block as a matrix product on the FP64 matrix cores:
// D[m][cell] = sum_b T[m][b] F[b][cell] for the cosine and the sine rows,
// then acc += dt * D.  A workgroup stages the table of up to 64
// frequencies (all the block's samples, cosine and sine side by side, rows
// padded against bank conflicts) in LDS once and then walks its share of
// the plane; each wavefront takes 16 cells of one component slot at a
// time, reads their samples, then streams them 4 at a time through 8 16x16
// tiles (4 of cosines, 4
// of sines; v_mfma_f64_16x16x4f64 on gfx950: A[i][k] comes from lane i +
// 16k, B[k][j] from lane j + 16k, and D[i][j] lands in lane j + 16 (i % 4),
// element i / 4).  The samples are not held in registers, so a block can
// be long: each flush reads and writes every accumulator once, so 64
// samples a block move half the accumulator bytes of 32.  The block's sum
// is in another order than a running sum (products fused, four samples at
// a time), a rounding difference the spectra carry at the 1e-16 level.
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kDftTabRow = 65;  // double2 entries a sample row of the staged table (64 + 1 pad)
constexpr int kDftFoldGroups = 512;  // workgroups of a fold: two a CU

template <typename real>
__global__ __launch_bounds__(256, 2) void k_dft_accum(const real* __restrict__ samp, int64_t np, int nslots, int nfreq,
                                                   int count, const double* __restrict__ tab, double dt,
                                                   double* __restrict__ acc) {
  __shared__ double2 lt[kDftBlock * kDftTabRow];
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const int jj = lane & 15, kk = lane >> 4;
  const int64_t ntile = (np + 15) / 16, nwt = ntile * nslots;
  for (int m0 = 0; m0 < nfreq; m0 += 64) {
    __syncthreads();
    for (int q = (int)threadIdx.x; q < kDftBlock * 64; q += 256) {
      const int b = q >> 6, mm = q & 63;
      double2 v = {0.0, 0.0};
      if (b < count && m0 + mm < nfreq) v = *(const double2*)(tab + 2 * ((int64_t)b * nfreq + m0 + mm));
      lt[b * kDftTabRow + mm] = v;
    }
    __syncthreads();
    // wave tiles of 16 cells of one slot, strided over the grid (the loop
    // count is uniform per wavefront; no barrier inside)
    for (int64_t wt = (int64_t)blockIdx.x * 4 + w; wt < nwt; wt += (int64_t)gridDim.x * 4) {
      const int slot = (int)(wt / ntile);
      const int64_t cell = (wt - (int64_t)slot * ntile) * 16 + jj;
      const bool cin = cell < np;
      double f[kDftBlock / 4];  // this lane's samples: b = 4 st + kk
#pragma unroll
      for (int st = 0; st < kDftBlock / 4; ++st) {
        const int b = 4 * st + kk;
        f[st] = (cin && b < count) ? (double)samp[((int64_t)b * nslots + slot) * np + cell] : 0.0;
      }
      v4d dc[4], ds[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) dc[t] = ds[t] = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
      for (int st = 0; st < kDftBlock / 4; ++st) {  // every lane: the matrix instructions need them all
        if (4 * st >= count) break;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const double2 cs = lt[(4 * st + kk) * kDftTabRow + 16 * t + jj];
          dc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(cs.x, f[st], dc[t], 0, 0, 0);
          ds[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(cs.y, f[st], ds[t], 0, 0, 0);
        }
      }
      if (!cin) continue;  // cells past the plane's end: no accumulator
      double* re = acc + (int64_t)slot * 2 * nfreq * np;
      double* im = re + (int64_t)nfreq * np;
      // accumulators: all loads of a tile pair before its stores
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        double ar[4], ai[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int m = m0 + 16 * t + kk + 4 * v;
          const int64_t q = (int64_t)m * np + cell;
          ar[v] = m < nfreq ? re[q] : 0.0;
          ai[v] = m < nfreq ? im[q] : 0.0;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int m = m0 + 16 * t + kk + 4 * v;
          if (m >= nfreq) continue;
          const int64_t q = (int64_t)m * np + cell;
          re[q] = ar[v] + dc[t][v] * dt;
          im[q] = ai[v] + ds[t][v] * dt;
        }
      }
    }
  }
}

// point probes, sim3/program.f90:445-451
// ---------------------------------------------------------------------------
// TFSF incident line (sim1/program.f90:217-262, mod1/BERpmly_ExHz_mod.f90).
// One workgroup walks the line; the PML object's own fields pe, ph live on
// cells 1..npml and n-npml+1..n and hand over at the slab edges.
struct TfsfLine {
  double *e, *h, *pe, *ph;
  const double *f1, *f2, *g1, *g2;
  int32_t n, npml, src, lo, hi, comp, hc, axis, sg;
  double ck, ku;
};

inline TfsfLine tfsf_line(const Sim* s, const Tfsf& t) {
  TfsfLine L;
  const int64_t m = t.n + 2;
  L.e = t.d_line;
  L.h = t.d_line + m;
  L.pe = t.d_line + 2 * m;
  L.ph = t.d_line + 3 * m;
  L.f1 = t.d_line + 4 * m;
  L.f2 = t.d_line + 5 * m;
  L.g1 = t.d_line + 6 * m;
  L.g2 = t.d_line + 7 * m;
  L.n = t.n;
  L.npml = t.p.npml;
  L.src = t.p.src_index;
  L.lo = t.p.lo;
  L.hi = t.p.hi;
  L.comp = t.p.comp;
  L.hc = t.hc;
  L.axis = t.p.axis;
  L.sg = t.sg;
  L.ck = s->ku * (1.0 / t.p.eps);
  L.ku = s->ku;
  return L;
}

// E half step of the line: ex_inc update (:222-224), the pulse (:228-229),
// mypmly_inc%updateE (:231; updateE_pmy_low/high + setboundaryEa/Eb)
__global__ __launch_bounds__(256) void k_tfsf_line_e(TfsfLine L, double g) {
  const int N = L.n, P = L.npml, hi = N - P + 1;
  for (int j = P + 1 + (int)threadIdx.x; j <= hi - 1; j += blockDim.x) {
    double v = L.e[j] + L.ck * (L.h[j] - L.h[j - 1]);
    if (j == L.src) v = v + g;
    L.e[j] = v;
  }
  if (P > 0) {
    for (int j = 1 + (int)threadIdx.x; j <= P; j += blockDim.x)
      L.pe[j] = L.f1[j] * L.pe[j] + L.f2[j] * (L.ph[j] - L.ph[j - 1]);
    for (int j = hi + (int)threadIdx.x; j <= N; j += blockDim.x)
      L.pe[j] = L.f1[j] * L.pe[j] + L.f2[j] * (L.ph[j] - L.ph[j - 1]);
    __syncthreads();
    if (threadIdx.x == 0) {
      L.pe[P + 1] = L.e[P + 1];
      L.e[P] = L.pe[P];
      L.pe[hi - 1] = L.e[hi - 1];
      L.e[hi] = L.pe[hi];
    }
  }
}

// H half step of the TFSF: block 0 advances the line (hz_inc update :252-254,
// mypmly_inc%updateH :255) and records its probes; the other blocks apply
// the consistency terms on H (:260-262 and, two-sided, the far plane):
// H_h(lo-1) -= s*Ku*e(lo), H_h(hi) += s*Ku*e(hi+1).  ALL_FACES PML cells on
// a plane take the term into the split along the line's axis.
template <typename real>
__global__ __launch_bounds__(256) void k_tfsf_h(FieldArgs<real> a, PmlArgs<real> p, int all_faces, TfsfLine L,
                                                const int32_t* lp, int nlp, double* rec, double tn, int n1, int n2) {
  if (blockIdx.x == 0) {
    const int N = L.n, P = L.npml, hi = N - P + 1;
    for (int j = P + 1 + (int)threadIdx.x; j <= hi - 1; j += blockDim.x) L.h[j] = L.h[j] + L.ku * (L.e[j + 1] - L.e[j]);
    if (P > 0) {
      for (int j = 1 + (int)threadIdx.x; j <= P; j += blockDim.x)
        L.ph[j] = L.g1[j] * L.ph[j] + L.g2[j] * (L.pe[j + 1] - L.pe[j]);
      for (int j = hi + (int)threadIdx.x; j <= N; j += blockDim.x)
        L.ph[j] = L.g1[j] * L.ph[j] + L.g2[j] * (L.pe[j + 1] - L.pe[j]);
      __syncthreads();
      if (threadIdx.x == 0) {
        L.ph[P + 1] = L.h[P + 1];
        L.h[P] = L.ph[P];
        L.ph[hi - 1] = L.h[hi - 1];
        L.h[hi] = L.ph[hi];
      }
    }
    __syncthreads();
    if ((int)threadIdx.x < nlp) {  // sim1/program.f90:272-273
      double* r = rec + 7 * lp[2 * threadIdx.x];
      const int j = lp[2 * threadIdx.x + 1];
      r[0] = tn;
      for (int q = 1; q < 7; ++q) r[q] = 0.0;
      r[1 + L.comp] = L.e[j];
      r[4 + L.hc] = L.sg * L.h[j];
    }
    return;
  }
  const int64_t u = (int64_t)(blockIdx.x - 1) * blockDim.x + threadIdx.x;
  const int64_t np = (int64_t)n1 * n2;
  const int nplanes = L.hi > 0 ? 2 : 1;
  if (u >= np * nplanes) return;
  const int pl = (int)(u / np);
  const int64_t w = u - (int64_t)pl * np;
  const int index = pl == 0 ? L.lo - 1 : L.hi;
  const double ev = pl == 0 ? L.e[L.lo] : L.e[L.hi + 1];
  const double val = (pl == 0 ? -(L.sg * L.ku) : (L.sg * L.ku)) * ev;
  const int a1 = L.axis == 0 ? 1 : 0, a2 = L.axis == 2 ? 1 : 2;
  int c3[3];
  c3[L.axis] = index;
  c3[a1] = (int)(w % n1) + 1;
  c3[a2] = (int)(w / n1) + 1;
  const int i = c3[0], j = c3[1], k = c3[2];
  const int c = i + a.px * j + a.pxy * k;
  const int q = L.hc;
  real hv;
  if (all_faces && in_pml(p.P, i, j, k, a.nx, a.ny, a.nz)) {
    // the part of H_q along the axis takes the value; with both parts of
    // H_q stored, H_q = the two summed as k_fused_h sums them (Hxy+Hxz,
    // Hyz+Hyx, Hzx+Hzy), otherwise H_q takes the value itself (a part not
    // stored is H_q less the other part)
    const int sa = q == 0 ? (L.axis == 1 ? S_HXY : S_HXZ)
                          : (q == 1 ? (L.axis == 0 ? S_HYX : S_HYZ) : (L.axis == 0 ? S_HZX : S_HZY));
    const int s0 = q == 0 ? S_HXY : (q == 1 ? S_HYZ : S_HZX);
    const int s1 = q == 0 ? S_HXZ : (q == 1 ? S_HYX : S_HZY);
    const bool st[3] = {in_slab(p.P[0], i, a.nx), in_slab(p.P[1], j, a.ny), in_slab(p.P[2], k, a.nz)};
    const bool both = st[q == 0 ? 1 : 0] && st[q == 2 ? 1 : 2];
    const int sp = pml_index(p.P, i, j, k, a.nx, a.ny, a.nz);
    if (st[L.axis]) p.s[sa][sp] = p.s[sa][sp] + (real)val;
    hv = both ? p.s[s0][sp] + p.s[s1][sp] : a.h[q][c] + (real)val;
  } else {
    hv = a.h[q][c] + (real)val;
  }
  a.h[q][c] = hv;
  put_h_ghost(a.h[q], hv, i, j, k, c, a);
}

// Flux of a DFT plane, one block per frequency, fixed-order tree reduction.
__global__ __launch_bounds__(256) void k_flux(const double* acc, int64_t np, int nfreq, int shA, int seA, int shB,
                                              int seB, double ds, double* S) {
  __shared__ double red[256];
  const int m = blockIdx.x;
  auto re = [&](int slot) { return acc + (int64_t)slot * 2 * nfreq * np + (int64_t)m * np; };
  auto im = [&](int slot) { return acc + (int64_t)slot * 2 * nfreq * np + (int64_t)nfreq * np + (int64_t)m * np; };
  const double *rhA = re(shA), *ihA = im(shA), *reA = re(seA), *ieA = im(seA);
  const double *rhB = re(shB), *ihB = im(shB), *reB = re(seB), *ieB = im(seB);
  double acc_l = 0.0;
  for (int64_t o = threadIdx.x; o < np; o += blockDim.x)
    acc_l += ((rhA[o] * reA[o]) + (ihA[o] * ieA[o]) - (rhB[o] * reB[o]) - (ihB[o] * ieB[o])) * ds;
  red[threadIdx.x] = acc_l;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) S[m] = red[0];
}

// Served-job sentinel: flag = 1 when any of the six field arrays holds a
// NaN or Inf (ghosts and padding are zero).  Lanes that find one store 1
// (the same value, so the race is harmless); the flag is never cleared.
template <typename real>
__global__ __launch_bounds__(256) void k_nonfinite(const real* e0, const real* e1, const real* e2, const real* h0,
                                                   const real* h1, const real* h2, int64_t n, int* flag) {
  const real* f[6] = {e0, e1, e2, h0, h1, h2};
  bool bad = false;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int q = 0; q < 6; ++q) bad |= !isfinite(f[q][c]);
  if (bad) flag[0] = 1;
}

// Energy sums over interior cells; one partial pair per block.
template <typename real>
__global__ __launch_bounds__(256) void k_energy(FieldArgs<real> a, double* part) {
  __shared__ double r1[256], r2[256];
  const int64_t n = (int64_t)a.nx * a.ny * a.nz;
  double s1 = 0, s2 = 0;
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(o % a.nx) + 1;
    const int64_t r = o / a.nx;
    const int j = (int)(r % a.ny) + 1, k = (int)(r / a.ny) + 1;
    const int64_t c = i + a.px * j + a.pxy * k;
    const double ex = a.e[0][c], ey = a.e[1][c], ez = a.e[2][c];
    const double hx = a.h[0][c], hy = a.h[1][c], hz = a.h[2][c];
    s1 += (double)a.gaxe[a.matid[c]].v[3] * (ex * ex + ey * ey + ez * ez);
    s2 += hx * hx + hy * hy + hz * hz;
  }
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r1[threadIdx.x] += r1[threadIdx.x + w];
      r2[threadIdx.x] += r2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = r2[0];
  }
}

// periodic ghost refresh of one component after a host write
template <typename real>
__global__ void k_ghosts(FieldArgs<real> a, real* f, int axis) {
  const int n[3] = {a.nx, a.ny, a.nz};
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (u > n[a1] + 1) return;
  int p[3];
  p[a1] = u;
  p[a2] = v;
  const int64_t st[3] = {1, a.px, a.pxy};
  p[axis] = 0;
  const int64_t c0 = p[0] + a.px * p[1] + a.pxy * p[2];
  f[c0 + st[axis] * (n[axis] + 1)] = f[c0 + st[axis] * 1];
  f[c0] = f[c0 + st[axis] * n[axis]];
}

// ---------------------------------------------------------------------------
// host-side launchers
namespace {

template <typename real>
FieldArgs<real> field_args(Sim* s) {
  FieldArgs<real> a{};
  for (int q = 0; q < 3; ++q) {
    a.e[q] = (real*)s->E[q];
    a.h[q] = (real*)s->H[q];
    a.d[q] = (real*)s->D[q];
  }
  a.matid = s->d_matid;
  const size_t nt = (size_t)s->ntab;
  a.coef = (const Vec4<real>*)s->d_mattab;
  a.gaxe = a.coef + nt;
  a.flags = (const uint32_t*)(a.gaxe + nt);
  a.rowmat = s->d_rowmat;
  a.nx = s->nx;
  a.ny = s->ny;
  a.nz = s->nz;
  for (int q = 0; q < 3; ++q) a.per[q] = s->g.periodic[q] ? 1 : 0;
  a.px = (int32_t)s->px;
  a.pxy = (int32_t)s->pxy;
  a.ku = (real)s->ku;
  a.ntx = (s->nx + s->bx - 1) / s->bx;
  a.nty = (s->ny + s->by - 1) / s->by;
  a.ntz = s->nz;
  const int tiles = a.ntx * a.nty * a.ntz;
  a.tper = (tiles + 7) / 8;
  a.xcdmap = s->xcdmap;
  return a;
}

template <typename real>
PmlArgs<real> pml_args(Sim* s) {
  PmlArgs<real> p{};
  for (int q = 0; q < 3; ++q) p.P[q] = s->P[q];
  for (int q = 0; q < S_N; ++q) p.s[q] = s->d_split ? (real*)s->d_split + (int64_t)q * s->npml_cells : nullptr;
  p.kup = (real)s->ku_pml;
  if (s->d_prof) {
    const real* base = (const real*)s->d_prof;
    int64_t off = 0;
    const int n[3] = {s->nx, s->ny, s->nz};
    for (int ax = 0; ax < 3; ++ax) {
      p.f1[ax] = base + off; off += n[ax] + 2;
      p.f2[ax] = base + off; off += n[ax] + 2;
      p.g1[ax] = base + off; off += n[ax] + 2;
      p.g2[ax] = base + off; off += n[ax] + 2;
    }
  }
  return p;
}

template <typename real>
RefPmlArgs<real> ref_pml_args(Sim* s) {
  RefPmlArgs<real> p{};
  p.P = s->pml.npml[1];
  p.jb = s->ny - p.P + 1;
  p.sx = s->nx + 2;
  p.sxy = (int64_t)(s->nx + 2) * (p.P + 2);
  for (int sl = 0; sl < 2; ++sl) {
    for (int q = 0; q < S_N; ++q) p.s[sl][q] = (real*)s->d_slab + ((int64_t)sl * S_N + q) * s->slab_cells;
    p.prof[sl] = (const real*)s->d_ref_prof + (int64_t)sl * 12 * s->ref_prof_stride;
  }
  p.ps = s->ref_prof_stride;
  return p;
}

struct Timer {
  Sim* s;
  int kid;
  hipEvent_t a = nullptr, b = nullptr;
  Timer(Sim* s_, int kid_) : s(s_), kid(kid_) {
    if (!s->profiling) return;
    a = get();
    if (a && hipEventRecord(a, s->stream) != hipSuccess) a = nullptr;
  }
  hipEvent_t get() {
    if (!s->event_pool.empty()) {
      hipEvent_t e = s->event_pool.back();
      s->event_pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  ~Timer() {
    if (!s->profiling || !a) return;
    b = get();
    if (!b || hipEventRecord(b, s->stream) != hipSuccess) return;
    s->pending.push_back({kid, {a, b}});
  }
};

template <typename real>
inline dim3 tile_grid(const FieldArgs<real>& a) {
  const int tiles = a.ntx * a.nty * a.ntz;
  return dim3((unsigned)(a.xcdmap ? 8 * a.tper : tiles));
}

// 8 XCD runs of `per` blocks of 4 wavefronts (seg_work undoes the dealing)

inline dim3 cell_grid(int nx, int nrows, int nz, dim3 blk) {
  return dim3((unsigned)((nx + blk.x - 1) / blk.x), (unsigned)((nrows + blk.y - 1) / blk.y), (unsigned)nz);
}

template <typename real>
SrcArgs<real> source_args(Sim* s, int n, int* nsrc) {
  SrcArgs<real> sa{};
  sa.n = n;
  const double t = (double)(n - 1) * s->dt;
  int q = 0;
  for (const auto& src : s->srcs) {
    const fdtd_pulse& p = src.p;
    // sim3/program.f90:400
    const double g = p.amplitude * (std::exp(-0.5 * ((t - p.t0) / p.st) * ((t - p.t0) / p.st)) * std::sin(p.w0 * t));
    sa.axis[q] = src.axis;
    sa.index[q] = src.index;
    sa.comp[q] = src.comp;
    sa.val[q] = (real)g;
    sa.dval[q] = nullptr;
    ++q;
  }
  // TFSF consistency terms on D, sim1/program.f90:236-238: D(lo) -= Ku*h(lo-1)
  // and, two-sided, D(hi+1) += Ku*h(hi), h read from the line on the device
  for (const auto& t : s->tfsfs) {
    const double* h = t.d_line + (t.n + 2);
    sa.axis[q] = t.p.axis;
    sa.index[q] = t.p.lo;
    sa.comp[q] = t.p.comp;
    sa.val[q] = 0;
    sa.dval[q] = h + (t.p.lo - 1);
    sa.dscale[q] = -s->ku;
    ++q;
    if (t.p.hi > 0) {
      sa.axis[q] = t.p.axis;
      sa.index[q] = t.p.hi + 1;
      sa.comp[q] = t.p.comp;
      sa.val[q] = 0;
      sa.dval[q] = h + t.p.hi;
      sa.dscale[q] = s->ku;
      ++q;
    }
  }
  *nsrc = q;
  return sa;
}

template <typename real>
void dft_flush_plane(Sim* s, DftPlane& d) {
  if (d.nbuf == 0) return;
  const int64_t np = (int64_t)d.n1 * d.n2;
  Timer tm(s, FDTD_K_MON);
  const int nt = d.nfreq * d.nbuf;
  hipLaunchKernelGGL(k_dft_table, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s->stream, d.d_omega, d.nfreq,
                     d.n_first, d.nbuf, s->dt, d.d_tab);
  // two workgroups a CU (the staged table is 65 KB of LDS), each walking
  // wave tiles of 16 cells of one slot
  const int64_t wgs = ((np + 15) / 16 * d.nslots + 3) / 4;
  hipLaunchKernelGGL(k_dft_accum<real>, dim3((unsigned)std::min<int64_t>(wgs, kDftFoldGroups)), dim3(256), 0,
                     s->stream, (const real*)d.d_samp, np, d.nslots, d.nfreq, d.nbuf, d.d_tab, s->dt, d.d_acc);
  d.nbuf = 0;
}

template <typename real>
int step_impl(Sim* s) {
  const int n = s->n_done + 1;
  const double t = (double)(n - 1) * s->dt;
  FieldArgs<real> a = field_args<real>(s);
  int nsrc = 0;
  SrcArgs<real> src = source_args<real>(s, n, &nsrc);
  if (s->dry) nsrc = 0;
  const dim3 blk((unsigned)s->bx, (unsigned)s->by, 1);
  hipStream_t st = s->stream;
  auto ade_coef = [&](const AdeGroup& g) {  // FP64 in both precisions
    AdeCoef<double> c = g.coef;
    c.ku = s->ku;
    return c;
  };
  auto lord_coef = [&](const AdeGroup& g) {
    LordCoef c = g.ld;
    c.ku = s->ku;
    return c;
  };

  // TFSF lines: E half step after the main E stage (the D terms above read
  // h(lo-1) before this step's line H update); H half step, consistency
  // terms on H and the line probes after the main H stage
  auto tfsf_e = [&](hipStream_t sq) {
    if (s->dry) return;
    for (const auto& t : s->tfsfs) {
      const fdtd_pulse& pu = t.p.pulse;
      const double tt = (double)(n - 1) * s->dt;
      const double g = pu.amplitude * (std::exp(-0.5 * ((tt - pu.t0) / pu.st) * ((tt - pu.t0) / pu.st)) *
                                       std::sin(pu.w0 * tt));
      Timer tm(s, FDTD_K_MAT);
      hipLaunchKernelGGL(k_tfsf_line_e, dim3(1), dim3(256), 0, sq, tfsf_line(s, t), g);
    }
  };
  auto tfsf_h = [&](hipStream_t sq, PmlArgs<real> pa, int all_faces) {
    if (s->dry) return;
    const int np = (int)s->probes.size();
    double* row = np ? s->d_probe + (int64_t)7 * np * (s->probe_row0 + (n - 1 - s->call_base)) : nullptr;
    for (const auto& t : s->tfsfs) {
      const int ax = t.p.axis;
      const int nn[3] = {s->nx, s->ny, s->nz};
      const int n1 = nn[ax == 0 ? 1 : 0], n2 = nn[ax == 2 ? 1 : 2];
      const int64_t cells = (int64_t)n1 * n2 * (t.p.hi > 0 ? 2 : 1);
      Timer tm(s, FDTD_K_MAT);
      hipLaunchKernelGGL((k_tfsf_h<real>), dim3((unsigned)(1 + (cells + 255) / 256)), dim3(256), 0, sq, a, pa,
                         all_faces, tfsf_line(s, t), (const int32_t*)t.d_lp, row ? t.nlp : 0, row,
                         (double)n * s->dt, n1, n2);
    }
  };
  // FP32 two-kernel steps: this step's monitor samples are written by the
  // H kernel (not with a TFSF wave: its H correction follows it); dry steps
  // write them too, into the ring slot the next real step overwrites
  auto fused_monitors = [&]() {
    DftFuse<real> df{};
    if (kFuseDftReal<real> && s->dft_fuse && s->tfsfs.empty()) {
      for (auto& d : s->dfts) {
        if (df.n == kFuseDft) break;
        if (n < d.n_begin || (d.n_end > 0 && n > d.n_end)) continue;
        const int q = df.n++;
        int ns = 0;
        for (int c = 0; c < 6; ++c)
          if (d.mask & (1 << c)) df.comp[q][ns++] = c;
        df.axis[q] = d.axis;
        df.index[q] = d.index;
        df.n1[q] = d.n1;
        df.ns[q] = ns;
        df.np[q] = (int64_t)d.n1 * d.n2;
        df.dst[q] = (real*)d.d_samp + (int64_t)d.nbuf * ns * df.np[q];
        d.fused_now = true;
      }
    }
    return df;
  };
  if (s->engine == FDTD_ENGINE_FUSED && s->onepass) {
    // new fields go to the second copies, then the copies swap roles
    FieldArgs<real> an = a;
    for (int q = 0; q < 3; ++q) {
      an.e[q] = (real*)s->E2[q];
      an.h[q] = (real*)s->H2[q];
    }
    {
      FieldArgs<real> ad = a;  // D form from the old H into the new E
      for (int q = 0; q < 3; ++q) ad.e[q] = an.e[q];
      AdeMulti gm{};
      auto launch_multi = [&]() {
        if (!gm.ng) return;
        Timer tm(s, FDTD_K_ADE);
        hipLaunchKernelGGL(k_ade_multi<real>, dim3((unsigned)gm.blk[gm.ng]), dim3(256), 0, st, ad, gm, src, nsrc);
        gm = AdeMulti{};
      };
      for (auto& g : s->groups) {
        if (!g.ncells) continue;
        if (g.kind == 2) {
          Timer tm(s, FDTD_K_ADE);
          hipLaunchKernelGGL((k_lord<real, true>), dim3((unsigned)((g.ncells + 255) / 256)), dim3(256), 0, st, ad,
                             lord_coef(g), g.d_cells, g.d_mask, (double*)g.d_state, g.ncells, src, nsrc);
          continue;
        }
        const int q = gm.ng;
        gm.co[q] = ade_coef(g);
        gm.cells[q] = g.d_cells;
        gm.masks[q] = g.d_mask;
        gm.state[q] = (double*)g.d_state;
        gm.n[q] = g.ncells;
        gm.blk[q + 1] = gm.blk[q] + (int32_t)((g.ncells + 255) / 256);
        gm.ng = q + 1;
        if (gm.ng == kAdeMulti) launch_multi();
      }
      launch_multi();
    }
    {
      int rc = build_march_tiles(s, s->op_by, s->op_kc);
      if (rc) return rc;
    }
    PmlArgs<real> p = pml_args<real>(s);  // old splits
    PmlArgs<real> pw = p;                 // new splits
    for (int q = 0; q < S_N; ++q)
      pw.s[q] = s->d_split2 ? (real*)s->d_split2 + (int64_t)q * s->npml_cells : nullptr;
    MarchArgs<real> mg{};
    for (int q = 0; q < 3; ++q) {
      mg.eo[q] = a.e[q];
      mg.ho[q] = a.h[q];
    }
    const int R = s->op_by;
    // monitor samples: the separate sample kernel (written from the march
    // instead, they cost the plain tiles more than the launches they save)
    if (s->n_tl_plain || s->n_tl_xf) {
      // x-face tiles (FP32 only, build_march_tiles): the split parts of
      // the x slabs, old (p) and new (pw), and the x profiles
      XfArgs<real> xa{};
      if (s->n_tl_xf) {
        const int Px = p.P[0], Py = p.P[1], Pz = p.P[2];
        xa.eo[0] = p.s[S_EYX];
        xa.eo[1] = p.s[S_EZX];
        xa.ho[0] = p.s[S_HYX];
        xa.ho[1] = p.s[S_HZX];
        xa.en[0] = pw.s[S_EYX];
        xa.en[1] = pw.s[S_EZX];
        xa.hn[0] = pw.s[S_HYX];
        xa.hn[1] = pw.s[S_HZX];
        xa.f1 = p.f1[0];
        xa.f2 = p.f2[0];
        xa.g1 = p.g1[0];
        xa.g2 = p.g2[0];
        xa.kup = p.kup;
        xa.Px = Px;
        xa.Py = Py;
        xa.Pz = Pz;
        xa.nym = s->ny - 2 * Py;
        xa.nzm = s->nz - 2 * Pz;
        xa.base = 2 * Pz * s->nx * s->ny + (s->nz - 2 * Pz) * 2 * Py * s->nx;
        xa.hioff = (s->nz - 2 * Pz) * xa.nym * Px;
      }
      // the plain tiles, then the x-face tiles (their own launch, tiles of
      // tl_bx rows, the kernel with the split update)
      if (s->n_tl_plain) {
        Timer tm(s, FDTD_K_E);
        mg.tiles = s->d_tl_plain;
        mg.ntiles = s->n_tl_plain;
        const int blocks = (mg.ntiles + 3) / 4;  // four wavefronts per workgroup
        mg.tper = (blocks + 7) / 8;
        const dim3 grid((unsigned)(an.xcdmap ? 8 * mg.tper : blocks));
        if (R == 2) {
          hipLaunchKernelGGL((k_march_plain<real, 2, false>), grid, dim3(256), 0, st, an, mg, src, nsrc, xa);
        } else if (R == 4) {
          hipLaunchKernelGGL((k_march_plain<real, 4, false>), grid, dim3(256), 0, st, an, mg, src, nsrc, xa);
        } else {
          if constexpr (sizeof(real) == 4)  // FP64 tiles stop at 4 rows (8 would spill)
            hipLaunchKernelGGL((k_march_plain<real, 8, false>), grid, dim3(256), 0, st, an, mg, src, nsrc, xa);
        }
      }
      if constexpr (sizeof(real) == 4) {
        if (s->n_tl_xf) {
          Timer tm(s, FDTD_K_XF);
          MarchArgs<real> mx = mg;
          mx.tiles = s->d_tl_plain + 8 * (int64_t)s->n_tl_plain;
          mx.ntiles = s->n_tl_xf;
          const int blocks = (mx.ntiles + 3) / 4;
          mx.tper = (blocks + 7) / 8;
          const dim3 grid((unsigned)(an.xcdmap ? 8 * mx.tper : blocks));
          if (s->tl_bx == 4)
            hipLaunchKernelGGL((k_march_plain<real, 4, true>), grid, dim3(256), 0, st, an, mx, src, nsrc, xa);
          else
            hipLaunchKernelGGL((k_march_plain<real, 2, true>), grid, dim3(256), 0, st, an, mx, src, nsrc, xa);
        }
      }
    }
    if (s->n_tl_gen) {
      mg.tiles = s->d_tl_gen;
      mg.ntiles = s->n_tl_gen;  // E: a workgroup per plane of each; H: one each
      FieldArgs<real> ae = an;  // E: old H in, new E out
      for (int q = 0; q < 3; ++q) ae.h[q] = a.h[q];
      {
        MarchArgs<real> me = mg;
        me.nsplit = s->tl_gkc;
        me.tper = (me.ntiles * me.nsplit + 7) / 8;
        const dim3 ge((unsigned)(an.xcdmap ? 8 * me.tper : me.ntiles * me.nsplit));
        Timer tm(s, FDTD_K_GEN_E);
        hipLaunchKernelGGL(k_gen_e<real>, ge, dim3(256), 0, st, ae, p, pw, me, src, nsrc);
      }
      mg.nsplit = 1;
      mg.tper = (mg.ntiles + 7) / 8;
      const dim3 grid((unsigned)(an.xcdmap ? 8 * mg.tper : mg.ntiles));
      {
        Timer tm(s, FDTD_K_GEN_H);
        hipLaunchKernelGGL(k_gen_h<real>, grid, dim3(256), 0, st, an, p, pw, mg);
      }
    }
    for (int q = 0; q < 3; ++q) {
      std::swap(s->E[q], s->E2[q]);
      std::swap(s->H[q], s->H2[q]);
    }
    std::swap(s->d_split, s->d_split2);
    a = field_args<real>(s);
    tfsf_e(st);
    tfsf_h(st, pml_args<real>(s), 1);
  } else if (s->engine == FDTD_ENGINE_FUSED) {
    PmlArgs<real> p = pml_args<real>(s);
    // tiles of zm planes, each thread marching its column through them
    FieldArgs<real> az = a;
    az.ntz = (s->nz + s->zm - 1) / s->zm;
    az.tper = (az.ntx * az.nty * az.ntz + 7) / 8;
    const dim3 gz = tile_grid(az);
    {
      Timer tm(s, FDTD_K_E);
      if (s->zm == 8)
        hipLaunchKernelGGL((k_fused_e<real, 8>), gz, blk, 0, st, az, p, src, nsrc);
      else if (s->zm == 4)
        hipLaunchKernelGGL((k_fused_e<real, 4>), gz, blk, 0, st, az, p, src, nsrc);
      else if (s->zm == 2)
        hipLaunchKernelGGL((k_fused_e<real, 2>), gz, blk, 0, st, az, p, src, nsrc);
      else
        hipLaunchKernelGGL((k_fused_e<real, 1>), gz, blk, 0, st, az, p, src, nsrc);
      if (s->nmix) {
        const int bm = std::min(256, (s->nx + 63) / 64 * 64);
        const int ntx = (s->nx + bm - 1) / bm;
        hipLaunchKernelGGL(k_fused_e_mixed<real>, dim3((unsigned)(s->nmix * ntx)), dim3((unsigned)bm), 0, st, a, p,
                           src, nsrc, (const int32_t*)s->d_mixrows, ntx);
      }
    }
    for (auto& g : s->groups) {
      if (!g.ncells) continue;
      Timer tm(s, FDTD_K_ADE);
      if (g.kind == 1)
        hipLaunchKernelGGL(k_ljp<real>, dim3((unsigned)((g.ncells + 255) / 256)), dim3(256), 0, st, a, g.lj,
                           g.d_cells, g.d_mask, (double*)g.d_state, g.ncells, src, nsrc);
      else if (g.kind == 2)
        hipLaunchKernelGGL((k_lord<real, true>), dim3((unsigned)((g.ncells + 255) / 256)), dim3(256), 0, st, a,
                           lord_coef(g), g.d_cells, g.d_mask, (double*)g.d_state, g.ncells, src, nsrc);
      else
        hipLaunchKernelGGL((k_ade<real, true>), dim3((unsigned)((g.ncells + 255) / 256)), dim3(256), 0, st, a,
                           ade_coef(g), g.d_cells, g.d_mask, (double*)g.d_state, g.ncells, src, nsrc);
    }
    tfsf_e(st);
    const DftFuse<real> df = fused_monitors();
    {
      Timer tm(s, FDTD_K_H);
      if (s->zm == 8)
        hipLaunchKernelGGL((k_fused_h<real, 8>), gz, blk, 0, st, az, p, df);
      else if (s->zm == 4)
        hipLaunchKernelGGL((k_fused_h<real, 4>), gz, blk, 0, st, az, p, df);
      else if (s->zm == 2)
        hipLaunchKernelGGL((k_fused_h<real, 2>), gz, blk, 0, st, az, p, df);
      else
        hipLaunchKernelGGL((k_fused_h<real, 1>), gz, blk, 0, st, az, p, df);
    }
    tfsf_h(st, p, 1);
  } else {
    const bool refy = s->pml_mode == FDTD_PML_REF_Y;
    const int P = refy ? s->pml.npml[1] : 0;
    const int j0 = 1 + P, j1 = s->ny - P;
    {
      Timer tm(s, FDTD_K_E);
      hipLaunchKernelGGL(k_staged_d<real>, cell_grid(s->nx, j1 - j0 + 1, s->nz, blk), blk, 0, st, a, j0, j1);
    }
    for (int q = 0; q < nsrc; ++q) {
      const int ax = src.axis[q];
      const int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
      const int nn[3] = {s->nx, s->ny, s->nz};
      Timer tm(s, FDTD_K_MAT);
      hipLaunchKernelGGL(k_staged_src<real>, dim3((unsigned)((nn[a1] + 63) / 64), (unsigned)nn[a2]), dim3(64), 0, st,
                         a, ax, src.index[q], src.comp[q], src.val[q], src.dval[q], src.dscale[q]);
    }
    {
      Timer tm(s, FDTD_K_MAT);
      hipLaunchKernelGGL(k_staged_mat<real>, cell_grid(s->nx, s->ny, s->nz, blk), blk, 0, st, a);
    }
    for (auto& g : s->groups) {
      if (!g.ncells) continue;
      Timer tm(s, FDTD_K_ADE);
      if (g.kind == 1)  // E form from curl H: the plane sources enter here, not through D
        hipLaunchKernelGGL(k_ljp<real>, dim3((unsigned)((g.ncells + 255) / 256)), dim3(256), 0, st, a, g.lj,
                           g.d_cells, g.d_mask, (double*)g.d_state, g.ncells, src, nsrc);
      else if (g.kind == 2)
        hipLaunchKernelGGL((k_lord<real, false>), dim3((unsigned)((g.ncells + 255) / 256)), dim3(256), 0, st, a,
                           lord_coef(g), g.d_cells, g.d_mask, (double*)g.d_state, g.ncells, src, 0);
      else
        hipLaunchKernelGGL((k_ade<real, false>), dim3((unsigned)((g.ncells + 255) / 256)), dim3(256), 0, st, a,
                           ade_coef(g), g.d_cells, g.d_mask, (double*)g.d_state, g.ncells, src, 0);
    }
    RefPmlArgs<real> rp{};
    if (refy) {
      rp = ref_pml_args<real>(s);
      Timer tm(s, FDTD_K_PML_E);
      hipLaunchKernelGGL(k_ref_pml_e<real>, cell_grid(s->nx, 2 * P, s->nz, blk), blk, 0, st, a, rp);
    }
    tfsf_e(st);
    {
      Timer tm(s, FDTD_K_H);
      hipLaunchKernelGGL(k_staged_h<real>, cell_grid(s->nx, j1 - j0 + 1, s->nz, blk), blk, 0, st, a, j0, j1);
    }
    if (refy) {
      Timer tm(s, FDTD_K_PML_H);
      hipLaunchKernelGGL(k_ref_pml_h<real>, cell_grid(s->nx, 2 * P, s->nz, blk), blk, 0, st, a, rp);
    }
    tfsf_h(st, PmlArgs<real>{}, 0);
  }
  // monitors: buffer this step's plane samples, fold a full block.  Dry
  // steps (the autotuner's) sample into the ring slot the next real step
  // overwrites, so each step form is timed with its sampling cost, and
  // neither advance the ring, fold nor record probes.
  (void)t;
  // the planes this step samples with the monitor kernel (the others were
  // written by the H kernel), with their ring slots
  MonList<real> ml{};
  auto launch_monitors = [&]() -> int {
    int nb = 0;
    for (int q = 0; q < ml.nplanes; ++q) {
      ml.blk[q] = nb;
      int ns = 0;
      while (ns < 6 && ml.comp[q][ns] >= 0) ++ns;
      nb += (int)((ml.np[q] + 255) / 256) * ns;
    }
    ml.blk[ml.nplanes] = nb;
    const int pb = ml.nprobe > 0 ? (ml.nprobe + 255) / 256 : 0;
    if (nb + pb == 0) return 0;
    Timer tm(s, FDTD_K_MON);
    hipLaunchKernelGGL(k_monitors<real>, dim3((unsigned)(nb + pb)), dim3(256), 0, st, a, ml);
    ml = MonList<real>{};
    return 0;
  };
  auto add_plane = [&](DftPlane& d, real* dst) -> int {
    if (ml.nplanes == kMonPlanes) {
      int rc = launch_monitors();
      if (rc) return rc;
    }
    const int q = ml.nplanes++;
    int ns = 0;
    for (int c = 0; c < 6; ++c) ml.comp[q][c] = -1;
    for (int c = 0; c < 6; ++c)
      if (d.mask & (1 << c)) ml.comp[q][ns++] = c;
    ml.axis[q] = d.axis;
    ml.index[q] = d.index;
    ml.n1[q] = d.n1;
    ml.np[q] = (int64_t)d.n1 * d.n2;
    ml.dst[q] = dst;
    return 0;
  };
  if (s->dry) {
    for (auto& d : s->dfts) {
      if (!d.fused_now && !(n < d.n_begin || (d.n_end > 0 && n > d.n_end))) {
        int ns = 0;
        for (int c = 0; c < 6; ++c) ns += (d.mask >> c) & 1;
        int rc = add_plane(d, (real*)d.d_samp + (int64_t)d.nbuf * ns * d.n1 * d.n2);
        if (rc) return rc;
      }
      d.fused_now = false;
    }
    return launch_monitors();
  }
  std::vector<DftPlane*> active;
  for (auto& d : s->dfts) {
    if (n < d.n_begin) continue;
    if (d.n_end > 0 && n > d.n_end) continue;
    int ns = 0;
    for (int c = 0; c < 6; ++c) ns += (d.mask >> c) & 1;
    if (d.nbuf == 0) d.n_first = n;
    if (!d.fused_now) {
      int rc = add_plane(d, (real*)d.d_samp + (int64_t)d.nbuf * ns * d.n1 * d.n2);
      if (rc) return rc;
    }
    d.fused_now = false;
    active.push_back(&d);
  }
  if (!s->probes.empty()) {
    ml.nprobe = (int)s->probes.size();
    ml.ppos = (const int32_t*)s->d_probe_pos;
    ml.pout = s->d_probe + (int64_t)7 * ml.nprobe * (s->probe_row0 + (n - 1 - s->call_base));
    ml.tn = (double)n * s->dt;
  }
  {
    int rc = launch_monitors();
    if (rc) return rc;
  }
  // fold after this step's samples are written, when the ring is full or
  // the step is a multiple of kDftBlock: the blocks then end at the same
  // steps however the calls that make the steps are cut (a call folds what
  // it buffered at its end), so runs stepped in chunks of kDftBlock
  // multiples, or resumed from a checkpoint, equal runs stepped at once
  for (DftPlane* d : active) {
    d->nbuf += 1;
    if (d->nbuf == kDftBlock || n % kDftBlock == 0) dft_flush_plane<real>(s, *d);
  }
  return 0;
}

}  // namespace

int launch_step(Sim* s) {
  // launch errors are read back with hipGetLastError, which reports the
  // last failing call of this thread: clear what an earlier, unrelated call
  // left so it is not taken for a launch failure of this step
  (void)hipGetLastError();
  int rc = s->prec == FDTD_FP32 ? step_impl<float>(s) : step_impl<double>(s);
  if (rc) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_check(e, "kernel launch");
  return 0;
}

int launch_dft_flush(Sim* s) {
  for (auto& d : s->dfts) {
    if (s->prec == FDTD_FP32)
      dft_flush_plane<float>(s, d);
    else
      dft_flush_plane<double>(s, d);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_check(e, "DFT flush");
  return 0;
}

int flux_check(Sim* s, int id) {
  const DftPlane& d = s->dfts[id];
  const int axis = d.axis;
  const int need[4] = {axis == 1 ? FDTD_HX : (axis == 2 ? FDTD_HY : FDTD_HZ),
                       axis == 1 ? FDTD_EZ : (axis == 2 ? FDTD_EX : FDTD_EY),
                       axis == 1 ? FDTD_HZ : (axis == 2 ? FDTD_HX : FDTD_HY),
                       axis == 1 ? FDTD_EX : (axis == 2 ? FDTD_EY : FDTD_EZ)};
  for (int q = 0; q < 4; ++q)
    if (!(d.mask & (1 << need[q]))) return set_error(FDTD_ERR_ARG, "flux needs the four tangential components");
  return 0;
}

int launch_flux(Sim* s, int id, double* S) {
  int rc = flux_check(s, id);
  if (rc) return rc;
  double* dS = reduction_buffer(s, sizeof(double) * (size_t)s->dfts[id].nfreq);
  if (!dS) return set_error(FDTD_ERR_NOMEM, "flux buffer");
  return launch_flux_into(s, id, dS, S);
}

// The flux reduction of plane `id` into the device buffer dS (nfreq
// doubles, each written whole), then copied to the host array S.
int launch_flux_into(Sim* s, int id, double* dS, double* S) {
  DftPlane& d = s->dfts[id];
  int eA, eB, hA, hB;
  if (d.axis == 1) { hA = FDTD_HX; eA = FDTD_EZ; hB = FDTD_HZ; eB = FDTD_EX; }
  else if (d.axis == 2) { hA = FDTD_HY; eA = FDTD_EX; hB = FDTD_HX; eB = FDTD_EY; }
  else { hA = FDTD_HZ; eA = FDTD_EY; hB = FDTD_HY; eB = FDTD_EZ; }
  const int64_t np = (int64_t)d.n1 * d.n2;
  hipLaunchKernelGGL(k_flux, dim3((unsigned)d.nfreq), dim3(256), 0, s->stream, d.d_acc, np, d.nfreq, d.slot[hA],
                     d.slot[eA], d.slot[hB], d.slot[eB], s->ds, dS);
  FDTD_HIP(hipGetLastError());
  FDTD_HIP(hipMemcpyAsync(S, dS, sizeof(double) * (size_t)d.nfreq, hipMemcpyDeviceToHost, s->stream));
  FDTD_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

int launch_energy(Sim* s, double* we, double* wh) {
  const int nb = 1024;
  double* part = reduction_buffer(s, sizeof(double) * 2 * nb);
  if (!part) return set_error(FDTD_ERR_NOMEM, "energy buffer");
  if (s->prec == FDTD_FP32)
    hipLaunchKernelGGL(k_energy<float>, dim3(nb), dim3(256), 0, s->stream, field_args<float>(s), part);
  else
    hipLaunchKernelGGL(k_energy<double>, dim3(nb), dim3(256), 0, s->stream, field_args<double>(s), part);
  FDTD_HIP(hipGetLastError());
  std::vector<double> h(2 * nb);
  FDTD_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, s->stream));
  FDTD_HIP(hipStreamSynchronize(s->stream));
  double a = 0, b = 0;
  for (int q = 0; q < nb; ++q) {
    a += h[2 * q];
    b += h[2 * q + 1];
  }
  *we = a;
  *wh = b;
  return 0;
}

int launch_nonfinite(Sim* s, int* d_flag) {
  const int64_t n = s->ncell_alloc;
  const dim3 grid((unsigned)std::min<int64_t>((n + 255) / 256, 4096));
  if (s->prec == FDTD_FP32)
    hipLaunchKernelGGL(k_nonfinite<float>, grid, dim3(256), 0, s->stream, (const float*)s->E[0],
                       (const float*)s->E[1], (const float*)s->E[2], (const float*)s->H[0], (const float*)s->H[1],
                       (const float*)s->H[2], n, d_flag);
  else
    hipLaunchKernelGGL(k_nonfinite<double>, grid, dim3(256), 0, s->stream, (const double*)s->E[0],
                       (const double*)s->E[1], (const double*)s->E[2], (const double*)s->H[0],
                       (const double*)s->H[1], (const double*)s->H[2], n, d_flag);
  FDTD_HIP(hipGetLastError());
  return 0;
}

int launch_set_ghosts(Sim* s, int comp) {
  const int n[3] = {s->nx, s->ny, s->nz};
  for (int ax = 0; ax < 3; ++ax) {
    if (!s->g.periodic[ax]) continue;
    const int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
    dim3 grid((unsigned)((n[a1] + 2 + 63) / 64), (unsigned)(n[a2] + 2));
    void* f = comp < 3 ? s->E[comp] : (comp < 6 ? s->H[comp - 3] : s->D[comp - 6]);
    if (s->prec == FDTD_FP32)
      hipLaunchKernelGGL(k_ghosts<float>, grid, dim3(64), 0, s->stream, field_args<float>(s), (float*)f, ax);
    else
      hipLaunchKernelGGL(k_ghosts<double>, grid, dim3(64), 0, s->stream, field_args<double>(s), (double*)f, ax);
  }
  FDTD_HIP(hipGetLastError());
  return 0;
}

}  // namespace fdtd
