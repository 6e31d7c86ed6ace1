// Least-squares RTM on gfx950: the Born operator L of the reference
// (fd.cpp:81-177) evaluated from stored per-step Born-source snapshots, its
// exact transpose L^T (fd_adjoint.h) with the imaging condition fused into the
// reverse time loop, FP64 inner products on the GPU, and CGLS.
//
// The reference declares conope_MPI / cgsolver_MPI (lsrtm/inversion.h:81,87)
// but defines neither; its intended normal operator is forward_MPI followed by
// adjoint_MPI over all shots (inversion.cpp:420-423).  Here L^T is the exact
// transpose of L, so CG minimises ||L m - d||^2 with a true gradient.
//
// Snapshots: the background's Born-source coefficient S^n = vt2 lap^n 2/v^3
// of every step is computed once per plan (two background passes for 16-bit
// storage: per-step max |S| first, then the quantized store) and reused by
// every L and L^T application.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"
#include "fd_adjoint.h"
#include "prof.h"
#include "seis_internal.h"

namespace seis {

// ---------------------------------------------------------------------------
// Field slots of one shot in the LSRTM pool.  Slots below LS_NADJ hold the
// adjoint's per-level state (cleared before every adjoint): three levels of
// lam, two of the precomputed raw adjoints and the one-way terms.  The
// LS_KIN..LS_KC slots hold the band coefficients of the shot (written once
// per plan): a band cell b's products of a level are coefficient x lam(b).
enum LsSlot {
  LS_FIELD = 0,
  LS_AR = 3,
  LS_OW = 5,
  LS_NADJ = 6,
  LS_KIN = 6,
  LS_KT = 7,
  LS_KC = 8,
  LS_NSLOT = 9
};

// lam of adjoint level m of a shot: slot m % 3.
template <typename T>
__device__ __forceinline__ T* level(const DevShot<T>& s, int lev) {
  return s.fld + (LS_FIELD + lev % 3) * s.fstride;
}

// Band-term masks (band_terms) of the ring cells of one window width,
// computed once per plan.
__global__ void band_terms_kernel(Ring R, int NX, int NZ, int pml, unsigned* __restrict__ out) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R.n) return;
  int x, z;
  ring_cell(R, NX, NZ, j, &x, &z);
  out[j] = band_terms(x, z, NX, NZ, pml);
}

// Sponge weight of (x, z) from the plan's table (DevShot::wtab): the band
// index of sponge_weight, clamped so that a cell one beyond the padded grid
// (read only where its term is masked out) stays inside the table.
template <typename T>
__device__ __forceinline__ T table_weight(const T* wt, int x, int z, int nx, int nz, int pml) {
  int i = -1;
  if (x < pml) i = pml - x - 1;
  else if (x >= pml + nx) i = x - pml - nx;
  else if (z < pml) i = pml - z - 1;
  else if (z >= pml + nz) i = z - pml - nz;
  return SEIS_GL(wt)[min(i, pml) + 1];
}

// Band coefficients of every band cell of a shot (edge_dist < pml): the
// products of band_adj per unit adjoint, so that a step forms a band cell's
// products as coefficient x lam(b) without the divisions of band_adj.
template <typename T>
__global__ void band_coef_kernel(Consts<T> k, const DevShot<T>* __restrict__ shots) {
  const DevShot<T>& s = shots[blockIdx.y];
  const Ring R = shot_split(k.pml, s.NX, k.NZ);
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= R.n) return;
  int x, z;
  ring_cell(R, s.NX, k.NZ, j, &x, &z);
  const Model<T> m{s.vel, s.nx, k.nz, k.pml};
  const BandAdj<T> c = band_adj(k, m, x, z, classify(x, z, s.NX, k.NZ, k.pml), T(1));
  const long long i = fidx(k.ld, x, z);
  s.fld[LS_KIN * s.fstride + i] = c.in;
  s.fld[LS_KT * s.fstride + i] = c.t;
  s.fld[LS_KC * s.fstride + i] = c.c;
}

// Raw adjoints of level `lev` for the cells with a band reader or w < 1
// (ring P: outside [pml+1, NX-pml-1) x [pml+1, NZ-pml-1)), and the one-way
// terms of lam^{lev-1} at those cells (levels lev and lev+1 are complete now),
// both stored for the band sweep and the ring.  A band cell b's products of a
// level are ain = kin(b) lam(b), at = kt(b) lam(b), ac = kc(b) lam(b)
// (band_adj).  With mk = band_terms of the cell (s.pmask):
//   raw adjoint  w lam(c) + the products ain, at of level lev of the band
//                cells that read raw(c), one per BT_AIN_* / BT_AT_* bit;
//   one-way cur  the products ac of level lev of band cell c itself and of the
//                side cells whose inner neighbour c is (BT_AC_* bits);
//   one-way prev of level lev+1: ain(c) of a side cell c, -(1 - w) lam of the
//                inner neighbours (BT_PIN_*), and the tangent products at of
//                the side cells beside c (BT_TAN_*).
// The terms the mask leaves out are added as zero, in a fixed order.
//
// The neighbourhood of a ring-P cell (x, z) that these read: lam^lev (l*) and
// kt on the 3 x 3 block around it, kin, kc and lam^{lev+1} (p*) on the cross;
// suffixes: zm / zp = z -+ 1, xm / xp = x -+ 1, mm = (x-1, z-1), pm = (x+1,
// z-1), mp = (x-1, z+1), pp = (x+1, z+1).
template <typename T>
struct PrepIn {
  T l0, l_zm, l_zp, l_xm, l_xp, l_mm, l_pm, l_mp, l_pp;
  T kin0, kin_zm, kin_zp, kin_xm, kin_xp;
  T kt_mm, kt_pm, kt_mp, kt_pp, kt_xm, kt_xp, kt_zm, kt_zp;
  T kc0, kc_zm, kc_zp, kc_xm, kc_xp;
  T p0, p_zm, p_zp, p_xm, p_xp;
};

// The raw adjoint *ar and the one-way terms *ow of a ring-P cell from its
// neighbourhood, its band-term mask, its sponge weight w and -(1 - w) of its
// four neighbours (f_*).
template <typename T>
__device__ __forceinline__ void prep_terms(const PrepIn<T>& n, unsigned mk, T w, T f_zm, T f_zp, T f_xm, T f_xp, T* ar,
                                           T* ow) {
  // raw-adjoint terms
  const T a_zm = n.kin_zm * n.l_zm, a_zp = n.kin_zp * n.l_zp, a_xm = n.kin_xm * n.l_xm, a_xp = n.kin_xp * n.l_xp;
  const T t_mm = n.kt_mm * n.l_mm, t_pm = n.kt_pm * n.l_pm, t_mp = n.kt_mp * n.l_mp, t_pp = n.kt_pp * n.l_pp;
  const T t_xm = n.kt_xm * n.l_xm, t_xp = n.kt_xp * n.l_xp;
  // one-way cur terms
  const T c_0 = n.kc0 * n.l0, c_zm = n.kc_zm * n.l_zm, c_zp = n.kc_zp * n.l_zp, c_xm = n.kc_xm * n.l_xm,
          c_xp = n.kc_xp * n.l_xp;
  // level lev+1: one-way prev terms
  const T b_0 = n.kin0 * n.p0;
  const T u_xm = n.kt_xm * n.p_xm, u_zm = n.kt_zm * n.p_zm, u_xp = n.kt_xp * n.p_xp, u_zp = n.kt_zp * n.p_zp;
  auto on = [&](unsigned bit, T v) { return (mk & bit) ? v : T(0); };
  T r = w * n.l0;
  r += on(BT_AIN_ZM, a_zm);
  r += on(BT_AT_ZM_XM, t_mm);
  r += on(BT_AT_ZM_XP, t_pm);
  r += on(BT_AIN_ZP, a_zp);
  r += on(BT_AT_ZP_XM, t_mp);
  r += on(BT_AT_ZP_XP, t_pp);
  r += on(BT_AIN_XM, a_xm);
  r += on(BT_ATC_XM, t_xm);
  r += on(BT_AT_XM_ZM, t_mm);
  r += on(BT_AT_XM_ZP, t_mp);
  r += on(BT_AIN_XP, a_xp);
  r += on(BT_ATC_XP, t_xp);
  r += on(BT_AT_XP_ZM, t_pm);
  r += on(BT_AT_XP_ZP, t_pp);
  T oc = T(0);
  oc += on(BT_AC_SELF, c_0);
  oc += on(BT_AC_ZM, c_zm);
  oc += on(BT_AC_ZP, c_zp);
  oc += on(BT_AC_XM, c_xm);
  oc += on(BT_AC_XP, c_xp);
  T op = T(0);
  op += on(BT_AIN_SELF, b_0);
  op += on(BT_PIN_ZM, f_zm * n.p_zm);
  op += on(BT_PIN_ZP, f_zp * n.p_zp);
  op += on(BT_PIN_XM, f_xm * n.p_xm);
  op += on(BT_PIN_XP, f_xp * n.p_xp);
  op += on(BT_TAN_XM, u_xm);
  op += on(BT_TAN_ZM, u_zm);
  op += on(BT_TAN_XP, u_xp);
  op += on(BT_TAN_ZP, u_zp);
  *ar = r;
  *ow = oc + op;
}

// Cell j of ring P of shot s at reverse step `step` (ring_cell order): every
// neighbour is loaded (all lie inside the field and its margin), so that no
// load waits on a branch.
template <typename T>
__device__ __forceinline__ void adj_prep(const Consts<T>& k, const DevShot<T>& s, int step, long long j) {
  const int lev = s.nt + s.ndel + 1 - step;
  if (lev < 1) return;
  const Ring R = shot_split(k.pml + 1, s.NX, k.NZ);  // ring P of this shot's width
  if (j >= R.n) return;
  int x, z;
  ring_cell(R, s.NX, k.NZ, j, &x, &z);
  const unsigned mk = SEIS_GL(s.pmask)[j];
  const long long ld = k.ld, f = s.fstride;
  const long long i = fidx(ld, x, z);
  const auto lam1 = SEIS_GL(level(s, lev)), lam2 = SEIS_GL(level(s, lev + 1));
  const auto kin = SEIS_GL(s.fld + LS_KIN * f), kt = SEIS_GL(s.fld + LS_KT * f), kc = SEIS_GL(s.fld + LS_KC * f);
  const long long zm = i - 1, zp = i + 1, xm = i - ld, xp = i + ld;
  PrepIn<T> n;
  n.l0 = lam1[i], n.l_zm = lam1[zm], n.l_zp = lam1[zp], n.l_xm = lam1[xm], n.l_xp = lam1[xp];
  n.l_mm = lam1[xm - 1], n.l_pm = lam1[xp - 1], n.l_mp = lam1[xm + 1], n.l_pp = lam1[xp + 1];
  n.kin0 = kin[i], n.kin_zm = kin[zm], n.kin_zp = kin[zp], n.kin_xm = kin[xm], n.kin_xp = kin[xp];
  n.kt_mm = kt[xm - 1], n.kt_pm = kt[xp - 1], n.kt_mp = kt[xm + 1], n.kt_pp = kt[xp + 1];
  n.kt_xm = kt[xm], n.kt_xp = kt[xp], n.kt_zm = kt[zm], n.kt_zp = kt[zp];
  n.kc0 = kc[i], n.kc_zm = kc[zm], n.kc_zp = kc[zp], n.kc_xm = kc[xm], n.kc_xp = kc[xp];
  n.p0 = lam2[i], n.p_zm = lam2[zm], n.p_zp = lam2[zp], n.p_xm = lam2[xm], n.p_xp = lam2[xp];
  // sponge weight of the cell, -(1 - w) of its four neighbours
  const int nx = s.nx, nz = k.nz, pml = k.pml;
  auto pin = [&](int bx, int bz) { return -(T(1) - table_weight(s.wtab, bx, bz, nx, nz, pml)); };
  T ar, ow;
  prep_terms(n, mk, table_weight(s.wtab, x, z, nx, nz, pml), pin(x, z - 1), pin(x, z + 1), pin(x - 1, z),
             pin(x + 1, z), &ar, &ow);
  SEIS_GL(s.fld + (LS_AR + lev % 2) * f)[i] = ar;
  SEIS_GL(s.fld + LS_OW * f)[i] = ow;
}

// Reads a raw adjoint: precomputed near the band, the level itself elsewhere.
template <typename T>
struct RawAdj {
  const T* ar;
  const T* lam;
  long long ld;
  int NX, NZ, pml;
  __device__ __forceinline__ T operator()(int x, int z) const {
    const long long i = fidx(ld, x, z);
    return (x > pml && x < NX - 1 - pml && z > pml && z < NZ - 1 - pml) ? SEIS_GL(lam)[i] : SEIS_GL(ar)[i];
  }
};

// Adjoint of the record injection plus the imaging condition for one cell.
template <typename T>
__device__ __forceinline__ T adj_finish(const Consts<T>& k, const DevShot<T>& s, int lev, int x, int z, T val,
                                        T raw_adj) {
  const int ix = x - k.pml, iz = z - k.pml;
  if (ix >= 0 && ix < s.nx && iz >= 0 && iz < k.nz) {
    const int itp = lev - 2 - s.ndel;  // the forward step lev-2 recorded level lev-1
    if (itp >= 0 && z == s.gz[ix] + k.pml) val += s.rec[(long long)ix * s.nt + itp];
    if (lev - 1 <= s.nt + s.ndel - 1) {
      const long long j = (long long)ix * k.nz + iz;
      s.grad[j] += raw_adj * snap_load(s, k.nz, lev - 1, j);
    }
  }
  return val;
}

constexpr int AWG = 256;

// One wavefront's sweep of the adjoint step over its strip [x0, x1) of 64
// rows from z0: lam^{lev-1} = 2 a + (9-point gather of Q = a vt2 / h^2) -
// lam^{lev+1}, with a the raw adjoint of level lev, then the record sample,
// the gradient update and, near the edges, the one-way terms.  DEEP: every
// cell the strip reads (4 beyond it on each side) lies inside
// [pml+1, NX-pml-1) x [pml+1, NZ-pml-1), where the raw adjoint is the level
// itself; then no cell of the strip takes a one-way term and every one is an
// active model cell.  The caller decides it per wavefront.  HALF: 16-bit
// snapshots (BF16 or FP16), else native ones; a compile-time choice, so that
// the snapshot load is one instruction into one register.
template <typename T, bool DEEP, bool HALF>
__device__ __forceinline__ void adj_sweep(const Consts<T>& k, const DevShot<T>& s, int lev, int x0, int x1, int z0,
                                          int zb, T* la, int lane) {
  const int NX = s.NX, NZ = k.NZ, pml = k.pml;
  const long long ld = k.ld;
  const T* const lam1 = level(s, lev);
  const T* const lam2 = level(s, lev + 1);
  T* out = level(s, lev - 1);
  const Model<T> m{s.vel, s.nx, k.nz, pml};
  const RawAdj<T> A1{s.fld + (LS_AR + lev % 2) * s.fstride, lam1, ld, NX, NZ, pml};
  const RawAdj<T> A2{s.fld + (LS_AR + (lev + 1) % 2) * s.fstride, lam2, ld, NX, NZ, pml};
  auto raw1 = [&](int x, int z) { return DEEP ? SEIS_GL(lam1)[fidx(ld, x, z)] : A1(x, z); };
  auto raw2 = [&](int x, int z) { return DEEP ? SEIS_GL(lam2)[fidx(ld, x, z)] : A2(x, z); };
  auto velat = [&](int x, int zz) {
    // clamped outside DEEP: idle lanes may sit past the model
    return DEEP ? SEIS_GL(s.vel)[(long long)(x - pml) * k.nz + (zz - pml)] : m.vel(x, zz);
  };
  auto vt2of = [&](T v) {
    const T a = k.dt * v;
    return a * a;
  };
  auto vt2at = [&](int x, int zz) { return vt2of(velat(x, zz)); };
  // the shot's record, snapshot and gradient arrays, held in registers for the
  // whole sweep (the stores below would otherwise make the compiler reload
  // them from the shot table every column)
  const int* const gzp = s.gz;
  const T* const recp = s.rec;
  T* const gradp = s.grad;
  const int nxm = s.nx, ntr = s.nt, snk = s.snap_kind;
  const int itp = lev - 2 - s.ndel;                  // record sample injected now
  const bool grad_on = lev - 1 <= ntr + s.ndel - 1;  // imaging with S^{lev-1}
  // Every load of the sweep is issued for every lane and column, at an
  // in-bounds address (clamped where its value goes unused), so that no
  // loaded register is conditionally defined: the column-ahead loads then stay
  // in flight until their use.
  const long long snap0 = (long long)min(lev - 1, ntr + s.ndel - 1) * nxm * k.nz;  // step lev-1 of the snapshots
  const T* const snap_n = reinterpret_cast<const T*>(s.snap) + snap0;
  const unsigned short* const snap_h = reinterpret_cast<const unsigned short*>(s.snap) + snap0;
  const int z = z0 + lane;
  const bool act = DEEP || z < zb;
  const int zl = DEEP ? z : min(z, NZ - 1);
  const bool halo = lane < 4 || lane >= 60;
  // halo rows z0-4..z0-1 (lanes 0-3) and z0+64..z0+67 (lanes 60-63); the other
  // lanes load the nearest of them (the same cache lines)
  const int hz = lane < 32 ? z0 - 4 + min(lane, 3) : min(z0 + 64 + max(lane - 60, 0), NZ - 1);
  const int hslot = lane < 4 ? lane : lane + 8;
  // queues along x: qx = a vt2 / dx^2 (rows x-4..x+4), qz and a (rows x..x+4)
  T qx[9], qz[5], lq[5];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int r = x0 - 4 + j;
    const T l = raw1(r, zl);
    const T v2 = vt2at(r, zl);
    qx[j] = l * (v2 * k.idx2);
    if (j >= 4) {
      qz[j - 4] = l * (v2 * k.idz2);
      lq[j - 4] = l;
    }
  }
  // loads of column x+1 are issued before column x is computed, including
  // the record sample, snapshot and gradient of its model cells
  const int iz = z - pml;
  const bool zin = DEEP || (iz >= 0 && iz < k.nz && act);
  const int izc = DEEP ? iz : min(max(iz, 0), k.nz - 1);
  const int itc = max(itp, 0);
  // the snapshot of a column is loaded a column ahead and decoded when used
  const T scl = (grad_on && snk != SNAP_NATIVE) ? (T)SEIS_GL(s.scale)[lev - 1] : T(1);
  // the halo raw adjoint and velocity are combined when used
  T na, nv, nhl, nhv, nprev, nrec, nsn = T(0), ngrad;
  T now = T(0);  // outside DEEP: the one-way terms of the next cell
  unsigned short nsh = 0;
  int ngz;
  const auto OW = SEIS_GL(s.fld + LS_OW * s.fstride);
  auto fetch = [&](int xx) {
    xx = min(xx, x1 - 1);
    if (!DEEP) now = OW[fidx(ld, xx, zl)];
    na = raw1(xx + 4, zl);
    nv = velat(xx + 4, zl);
    nhl = raw1(xx, hz);
    nhv = velat(xx, hz);
    nprev = raw2(xx, zl);
    const int ixc = DEEP ? xx - pml : min(max(xx - pml, 0), nxm - 1);
    ngz = SEIS_GL(gzp)[ixc];
    nrec = SEIS_GL(recp)[(long long)ixc * ntr + itc];
    const long long j = (long long)ixc * k.nz + izc;
    if (HALF)
      nsh = SEIS_GL(snap_h)[j];
    else
      nsn = SEIS_GL(snap_n)[j];
    ngrad = SEIS_GL(gradp)[j];
  };
  fetch(x0);
  for (int x = x0; x < x1; x++) {
    const T nv2 = vt2of(nv);
    qx[8] = na * (nv2 * k.idx2);
    qz[4] = na * (nv2 * k.idz2);
    lq[4] = na;
    if (halo) la[hslot] = nhl * (vt2of(nhv) * k.idz2);
    const T prev = nprev, rec_cur = nrec, grad_cur = ngrad;
    const T sn_cur = nsn;
    const unsigned short sh_cur = nsh;
    const int gz_cur = ngz;
    const T ow_cur = now;
    fetch(x + 1);
    la[lane + 4] = qz[0];
    wave_sync();
    T sz = la[lane] * k.c[4];
    sz += la[lane + 1] * k.c[3];
    sz += la[lane + 2] * k.c[2];
    sz += la[lane + 3] * k.c[1];
    sz += la[lane + 4] * k.c[0];
    sz += la[lane + 5] * k.c[1];
    sz += la[lane + 6] * k.c[2];
    sz += la[lane + 7] * k.c[3];
    sz += la[lane + 8] * k.c[4];
    wave_sync();
    T sx = qx[0] * k.c[4];
    sx += qx[1] * k.c[3];
    sx += qx[2] * k.c[2];
    sx += qx[3] * k.c[1];
    sx += qx[4] * k.c[0];
    sx += qx[5] * k.c[1];
    sx += qx[6] * k.c[2];
    sx += qx[7] * k.c[3];
    sx += qx[8] * k.c[4];
    if (act) {
      const long long i = fidx(ld, x, z);
      T val = T(2) * lq[0] + sz + sx - prev;
      const int ed = DEEP ? pml + 1 : edge_dist(x, z, NX, NZ);
      if (!DEEP && ed <= pml) val += ow_cur;  // one-way terms (prep kernel)
      // adjoint of the record extraction and the imaging condition (as adj_finish)
      const int ix = x - pml;
      if (DEEP || (ix >= 0 && ix < nxm && zin)) {
        if (itp >= 0 && z == gz_cur + pml) val += rec_cur;
        if (grad_on) {
          const T snap_cur = !HALF ? sn_cur
                                   : (T)((snk == SNAP_BF16 ? bf16_to_f32(sh_cur) : f16_to_f32(sh_cur)) * (float)scl);
          SEIS_GL(gradp)[(long long)ix * k.nz + iz] = grad_cur + lq[0] * snap_cur;
        }
      }
      SEIS_GL(out)[i] = val;
    }
#pragma unroll
    for (int j = 0; j < 8; j++) qx[j] = qx[j + 1];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      qz[j] = qz[j + 1];
      lq[j] = lq[j + 1];
    }
  }
}

// Sweep items of the adjoint step: z chunk zc (64 rows from za + 64 zc) of
// strip xs (columns from xa + bx xs) of a shot's interior sq.
// An item is DEEP when its strip and every cell it reads (4 beyond it on each
// side) lie inside [pml+1, NX-pml-1) x [pml+1, NZ-pml-1) and its chunk is
// whole.
SEIS_HD inline bool deep_item(const Split& sq, int zc, int xs, int bx, int NX, int NZ, int pml) {
  const int z0 = sq.za + zc * 64;
  const int x0 = sq.xa + xs * bx, x1 = imin(x0 + bx, sq.xb);
  const int lo = pml + 1, hix = NX - pml - 1, hiz = NZ - pml - 1;
  return x0 - 4 >= lo && x1 + 4 <= hix && z0 - 4 >= lo && z0 + 68 <= hiz && z0 + 64 <= sq.zb;
}

SEIS_HD inline int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
SEIS_HD inline int ceil_div(int a, int b) { return -floor_div(-a, b); }

// The DEEP items of a shot form a rectangle: chunks [zc0, zc0 + nzd) of
// strips [xs0, xs0 + nxd) (empty: all zero).  Each bound of deep_item is
// monotone in zc or xs; the plan checks this form against deep_item for
// every width it launches.
struct DeepBox {
  int zc0, nzd, xs0, nxd;
};

SEIS_HD inline DeepBox deep_box(const Split& sq, int nzc, int bx, int NX, int NZ, int pml) {
  const int nxs = shot_strips(sq, bx);
  const int lo = pml + 1, hix = NX - pml - 1, hiz = NZ - pml - 1;
  const int zlo = imax(0, ceil_div(lo + 4 - sq.za, 64));
  const int zhi = imin(nzc - 1, floor_div(imin(hiz - 68, sq.zb - 64) - sq.za, 64));
  const int xlo = imax(0, ceil_div(lo + 4 - sq.xa, bx));
  // x1 = min(x0 + bx, xb): every strip's end qualifies when xb does
  const int xhi = sq.xb + 4 <= hix ? nxs - 1 : imin(nxs - 1, floor_div(hix - 4 - bx - sq.xa, bx));
  DeepBox d{zlo, zhi - zlo + 1, xlo, xhi - xlo + 1};
  if (d.nzd <= 0 || d.nxd <= 0) d = DeepBox{0, 0, 0, 0};
  return d;
}

// Item f of the items outside the deep box, in strip order: the strips before
// the box (all chunks), the box's strips (chunks above and below it), the
// strips after it.
SEIS_HD inline void frame_item(const DeepBox& d, int nzc, int f, int* zc, int* xs) {
  const int c1 = d.xs0 * nzc;
  if (f < c1) {
    *zc = f % nzc;
    *xs = f / nzc;
    return;
  }
  f -= c1;
  const int per = nzc - d.nzd;
  if (f < d.nxd * per) {
    const int r = f % per;
    *zc = r < d.zc0 ? r : r + d.nzd;
    *xs = d.xs0 + f / per;
    return;
  }
  f -= d.nxd * per;
  *zc = f % nzc;
  *xs = d.xs0 + d.nxd + f / nzc;
}

// One item's sweep, the snapshot form chosen by the shot.
template <typename T, bool DEEP>
__device__ __forceinline__ void adj_item(const Consts<T>& k, const DevShot<T>& s, const Split& sq, const Layout& F,
                                         int lev, int zc, int xs, T* la) {
  const int z0 = sq.za + zc * 64;
  const int x0 = sq.xa + xs * F.bx, x1 = min(x0 + F.bx, sq.xb);
  const int lane = threadIdx.x & 63;
  if (s.snap_kind != SNAP_NATIVE)
    adj_sweep<T, DEEP, true>(k, s, lev, x0, x1, z0, sq.zb, la, lane);
  else
    adj_sweep<T, DEEP, false>(k, s, lev, x0, x1, z0, sq.zb, la, lane);
}

// One reverse step: lam^{lev-1} for every cell and grad += raw-adjoint(lev) *
// S^{lev-1}, in two launches.
//  * adj_deep_kernel: the raw adjoints and one-way terms of level lev in ring
//    P (adj_prep, work-groups below npwg) and, beside them, the sweep of every
//    DEEP item, which reads neither (fused, the two overlap: 182 us per step
//    on the 128-shot 376 x 1151 plan against about 210 us as two launches).
//  * adj_band_kernel: the sweep of the other items of the cells at least 8
//    from the edge, reading the prep's results, and (work-groups from nfwg)
//    the outer 8 cells by the general gather (adj_ring).
// The sweeps: every stencil reader within 4 cells of a swept cell uses the
// 9-point operator, so the stencil part is the 9-point gather of Q = (raw
// adjoint) vt2 / h^2 of the readers; cells at most pml from the edge add the
// one-way terms.  The wavefront's index is taken uniform across its lanes, so
// that the strip bounds, the loop and the column addresses stay in scalar
// registers; a wavefront past the last item leaves as a whole (no workgroup
// barrier follows).
template <typename T>
__global__ __launch_bounds__(AWG) void adj_deep_kernel(Consts<T> k, const DevShot<T>* __restrict__ shots, Layout F,
                                                       int npwg, int ndwg, int step) {
  __shared__ T lds[WAVES][LROW];
  const DevShot<T>& s = shots[blockIdx.y];
  if ((int)blockIdx.x < npwg) {
    adj_prep(k, s, step, (long long)blockIdx.x * AWG + threadIdx.x);
    return;
  }
  const int lev = s.nt + s.ndel + 1 - step;
  if (lev < 1) return;
  const Split sq = shot_split(F.m, s.NX, k.NZ);  // this shot's interior
  const DeepBox d = deep_box(sq, F.nzc, F.bx, s.NX, k.NZ, k.pml);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int item = xcd_tile(blockIdx.x - npwg, ndwg) * WAVES + wave;
  if (item >= d.nzd * d.nxd) return;
  adj_item<T, true>(k, s, sq, F, lev, d.zc0 + item % d.nzd, d.xs0 + item / d.nzd, lds[wave]);
}

template <typename T>
__device__ void adj_ring(const Consts<T>& k, const DevShot<T>& s, const Layout& F, int step, long long j);

template <typename T>
__global__ __launch_bounds__(AWG) void adj_band_kernel(Consts<T> k, const DevShot<T>* __restrict__ shots, Layout F,
                                                       int nfwg, int step) {
  __shared__ T lds[WAVES][LROW];
  const DevShot<T>& s = shots[blockIdx.y];
  if ((int)blockIdx.x >= nfwg) {
    adj_ring(k, s, F, step, (long long)(blockIdx.x - nfwg) * AWG + threadIdx.x);
    return;
  }
  const int lev = s.nt + s.ndel + 1 - step;
  if (lev < 1) return;
  const Split sq = shot_split(F.m, s.NX, k.NZ);  // this shot's interior
  const DeepBox d = deep_box(sq, F.nzc, F.bx, s.NX, k.NZ, k.pml);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int item = xcd_tile(blockIdx.x, nfwg) * WAVES + wave;
  if (item >= F.nzc * shot_strips(sq, F.bx) - d.nzd * d.nxd) return;
  int zc, xs;
  frame_item(d, F.nzc, item, &zc, &xs);
  adj_item<T, false>(k, s, sq, F, lev, zc, xs, lds[wave]);
}

template <typename T>
__device__ void adj_ring(const Consts<T>& k, const DevShot<T>& s, const Layout& F, int step, long long j) {
  const int lev = s.nt + s.ndel + 1 - step;
  if (lev < 1) return;
  const int NX = s.NX, NZ = k.NZ, pml = k.pml;
  const Split sq = shot_split(F.m, NX, NZ);  // this shot's interior
  T* out = level(s, lev - 1);
  const Model<T> m{s.vel, s.nx, k.nz, pml};
  const long long ld = k.ld;
  const RawAdj<T> A1{s.fld + (LS_AR + lev % 2) * s.fstride, level(s, lev), ld, NX, NZ, pml};
  const RawAdj<T> A2{s.fld + (LS_AR + (lev + 1) % 2) * s.fstride, level(s, lev + 1), ld, NX, NZ, pml};
  if (j >= sq.n) return;
  int x, z;
  ring_cell(sq, NX, NZ, j, &x, &z);
  T ar;
  // one-way terms: stored by the prep kernel for ring P, zero elsewhere
  T val = adj_value(k, m, NX, A1, A2, x, z, &ar, SEIS_GL(s.fld)[LS_OW * s.fstride + fidx(ld, x, z)]);
  val = adj_finish(k, s, lev, x, z, val, ar);
  out[fidx(ld, x, z)] = val;
}

// Sum of the per-shot gradients in shot order into the image grid: cell j
// (column j / nz) gets the gradient of every shot whose window holds that
// column, out[j] = sum_s grad_s[j - x0_s nz].
template <typename T>
__global__ void sum_grads_kernel(const DevShot<T>* __restrict__ shots, int nshots, int nz, long long n, T* out) {
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x) {
    const int col = (int)(j / nz);
    T a = T(0);
    for (int s = 0; s < nshots; s++) {
      const int c = col - shots[s].x0;
      if (c >= 0 && c < shots[s].nx) a += shots[s].grad[j - (long long)shots[s].x0 * nz];
    }
    out[j] = a;
  }
}

// FP64 inner product <a, b> of two FP32/FP64 vectors: fixed-order block
// partials, then one block folds them (deterministic).
constexpr int RB = 256;
constexpr int RGRID = 1024;

template <typename T>
__global__ __launch_bounds__(RB) void dot_partial_kernel(const T* __restrict__ a, const T* __restrict__ b, long long n,
                                                         double* part) {
  __shared__ double sh[RB];
  double acc = 0.0;
  for (long long j = blockIdx.x * (long long)RB + threadIdx.x; j < n; j += (long long)gridDim.x * RB)
    acc += (double)a[j] * (double)b[j];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(RB) void dot_final_kernel(const double* __restrict__ part, int n, double* out) {
  __shared__ double sh[RB];
  double acc = 0.0;
  for (int j = threadIdx.x; j < n; j += RB) acc += part[j];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// y = y + alpha x (alpha in FP64, element update in FP64, stored in T)
template <typename T>
__global__ void axpy_kernel(T* y, const T* x, double alpha, long long n) {
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x)
    y[j] = (T)((double)y[j] + alpha * (double)x[j]);
}

// p = s + beta p
template <typename T>
__global__ void xpby_kernel(T* p, const T* s, double beta, long long n) {
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x)
    p[j] = (T)((double)s[j] + beta * (double)p[j]);
}

// y = a * b (record mask)
template <typename T>
__global__ void mul_kernel(T* y, const T* a, const T* b, long long n) {
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x)
    y[j] = a[j] * b[j];
}

// y = y + alpha x with x in FP64 (the CG iterate update from the FP64 direction)
template <typename T>
__global__ void axpy_from64_kernel(T* y, const double* x, double alpha, long long n) {
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x)
    y[j] = (T)((double)y[j] + alpha * x[j]);
}

// dst = a * (double)src (widening copy with a scale)
template <typename S>
__global__ void scale_to64_kernel(double* dst, const S* src, double a, long long n) {
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x)
    dst[j] = a * (double)src[j];
}

// dst = (T)src
template <typename T>
__global__ void from64_kernel(T* dst, const double* src, long long n) {
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x)
    dst[j] = (T)src[j];
}

template <typename T>
__global__ void fill_kernel(T* p, T v, long long n) {
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x)
    p[j] = v;
}

// ---------------------------------------------------------------------------
// Plan: device-resident state of one LSRTM problem.

#define LS_CHECK(x)                                                             \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));                \
      return SEIS_ERR_HIP;                                                      \
    }                                                                           \
  } while (0)

template <typename T>
struct LsrtmPlan {
  GridSpec g{};
  int nshots = 0, snap_kind = 0, NXmax = 0, NZ = 0, NTmax = 0;
  long long ld = 0, model = 0, nrec = 0;
  std::vector<ShotSpec> spec;
  std::vector<int> NXs;    // padded width of each shot's window
  std::vector<DevShot<T>> host;
  DevShot<T>* dev = nullptr;
  // The exact adjoint's view of the same fields: origins `adj_shift` elements
  // later, so that row 8, where its sweep's 64-row chunks start, begins a
  // 64-element line (the forward passes' chunks start at row max(pml, 4),
  // which field_origin aligns).  Its slots are cleared or written by the
  // adjoint alone (the band coefficients once, with this view).
  std::vector<DevShot<T>> host_adj;
  DevShot<T>* dev_adj = nullptr;
  long long adj_shift = 0;
  T* pool = nullptr;       // LS_NSLOT fields per shot
  void* snaps = nullptr;   // snapshots of all shots
  float* scales = nullptr;
  T* grads = nullptr;      // per-shot gradients
  double* part = nullptr;  // reduction scratch (RGRID partials + results)
  T* cg = nullptr;         // CG vectors: s, p, W p (model) and r, q (records)
  double* cg64 = nullptr;  // FP64 CG vectors: s, p and the orthonormal gradients
  size_t cg64_elems = 0;
  const T* mask = nullptr; // optional record mask (caller-owned, device)
  const T* weight = nullptr;  // optional image-space preconditioner W of CG (caller-owned, device)
  T* masked = nullptr;     // masked copy of an adjoint input
  unsigned* pmask = nullptr;  // band_terms of the ring cells P, one table per distinct width
  T* wtab = nullptr;          // sponge weight by band index (DevShot::wtab)
  Consts<T> k{};
  Layout F{};   // forward step split (interior sweep / band)
  Layout A{};   // adjoint step split
  Layout P{};   // cells whose raw adjoint is precomputed (margin pml + 1)
  Layout B{};   // absorbing band of the forward steps
  long long model_cells = 0;  // sum over shots of the window's model cells
  bool adj_tuned = false;
  int adj_kind = SEIS_LSRTM_ADJ_EXACT;  // seis_lsrtm_set_adjoint
  hipStream_t stream = 0;

  ~LsrtmPlan() {
    for (void* p : {(void*)dev, (void*)dev_adj, (void*)pool, snaps, (void*)scales, (void*)grads, (void*)part, (void*)cg, (void*)cg64,
                    (void*)masked, (void*)pmask, (void*)wtab})
      if (p) (void)hipFree(p);
  }

  size_t snap_elem() const { return snap_kind == SNAP_NATIVE ? sizeof(T) : 2; }

  int create(const GridSpec& grid, int ns, const ShotSpec* shots, const T* vel, const T* wav, const int* gz,
             int kind) {
    g = grid;
    nshots = ns;
    snap_kind = kind;
    if (g.pml < 4) return set_error("lsrtm: pml must be >= 4"), SEIS_ERR_ARG;
    if (g.nx <= 0 || g.nz <= 0 || !(g.dt > 0) || !(g.dx > 0) || !(g.dz > 0))
      return set_error("lsrtm: bad grid"), SEIS_ERR_ARG;
    if ((long long)(g.nx + 2 * g.pml) * (g.nz + 2 * g.pml) >= (1LL << 31))
      return set_error("lsrtm: the padded grid must have fewer than 2^31 cells"), SEIS_ERR_ARG;
    if (kind < 0 || kind > 2) return set_error("lsrtm: snapshot kind must be 0, 1 or 2"), SEIS_ERR_ARG;
    if (ns <= 0) return set_error("lsrtm: no shots"), SEIS_ERR_ARG;
    NZ = g.nz + 2 * g.pml;
    {
      const int mz = field_mz(g.pml), za = 8;  // the adjoint sweep's first row (make_layout margin below)
      adj_shift = (64 - (mz + za) % 64) % 64;
      // the column keeps at least 4 rows below the grid in the shifted view
      ld = ((long long)mz + adj_shift + NZ + 4 + 63) / 64 * 64;
    }
    model = (long long)g.nx * g.nz;
    spec.assign(shots, shots + ns);
    NXs.resize(ns);
    long long snap_total = 0, scale_total = 0, pool_total = 0;
    nrec = 0;
    model_cells = 0;
    for (int s = 0; s < ns; s++) {
      const ShotSpec& q = spec[s];
      const int nx = shot_width(g, q);
      if (q.nt <= 0 || q.ndel < 0 || q.sx < 0 || q.sx >= nx || q.sz < 0 || q.sz >= g.nz)
        return set_error("lsrtm: bad shot"), SEIS_ERR_ARG;
      if (q.nx < 0 || q.x0 < 0 || q.x0 + nx > g.nx)
        return set_error("lsrtm: a shot's window [x0, x0 + nx) must lie in the grid's columns"), SEIS_ERR_ARG;
      if (nx + 2 * g.pml < 8) return set_error("lsrtm: a shot's padded width must be at least 8"), SEIS_ERR_ARG;
      NXs[s] = nx + 2 * g.pml;
      NXmax = std::max(NXmax, NXs[s]);
      NTmax = std::max(NTmax, q.nt + q.ndel);
      snap_total += (long long)(q.nt + q.ndel) * nx * g.nz;
      scale_total += q.nt + q.ndel;
      pool_total += field_elems(NXs[s], ld) * LS_NSLOT;
      model_cells += (long long)nx * g.nz;
      nrec = std::max(nrec, q.rec + (long long)nx * q.nt);
    }
    LS_CHECK(hipMalloc(&pool, sizeof(T) * (size_t)pool_total));
    LS_CHECK(hipMalloc(&snaps, snap_elem() * (size_t)snap_total));
    LS_CHECK(hipMalloc(&scales, sizeof(float) * (size_t)scale_total));
    LS_CHECK(hipMalloc(&grads, sizeof(T) * (size_t)model_cells));
    LS_CHECK(hipMalloc(&part, sizeof(double) * (RGRID + 8)));
    LS_CHECK(hipMemset(scales, 0, sizeof(float) * (size_t)scale_total));
    k = make_consts<T>(g, ld);
    F = make_layout(NXs.data(), ns, NZ, std::max(g.pml, 4));
    // adjoint sweep: every stencil reader within 4 cells is a 9-point cell
    A = make_layout(NXs.data(), ns, NZ, 8, AWG);
    P = make_layout(NXs.data(), ns, NZ, g.pml + 1, 256);
    B = band_layout(NXs.data(), ns, NZ, g.pml);
    // band-term masks of ring P, one table per distinct window width
    std::vector<int> widths(NXs);
    std::sort(widths.begin(), widths.end());
    widths.erase(std::unique(widths.begin(), widths.end()), widths.end());
    std::vector<long long> moff(widths.size());
    long long mtot = 0;
    for (size_t w = 0; w < widths.size(); w++) {
      moff[w] = mtot;
      mtot += shot_split(g.pml + 1, widths[w], NZ).n;
    }
    {
      // sponge weights by band index, formed like sponge_weight's: entry 0
      // inside the model, entry i + 1 for band index i (i = pml: one cell
      // beyond the padded grid)
      std::vector<T> wt(g.pml + 2);
      wt[0] = sponge_weight<T>(g.pml, g.pml, 1, 1, g.pml);
      for (int i = 0; i <= g.pml; i++) wt[i + 1] = sponge_weight<T>(g.pml - 1 - i, g.pml, 1, 1, g.pml);
      LS_CHECK(hipMalloc(&wtab, sizeof(T) * wt.size()));
      LS_CHECK(hipMemcpy(wtab, wt.data(), sizeof(T) * wt.size(), hipMemcpyHostToDevice));
    }
    if (mtot > 0) {
      LS_CHECK(hipMalloc(&pmask, sizeof(unsigned) * (size_t)mtot));
      for (size_t w = 0; w < widths.size(); w++) {
        const Ring R = shot_split(g.pml + 1, widths[w], NZ);
        if (R.n > 0)
          hipLaunchKernelGGL(band_terms_kernel, dim3((unsigned)((R.n + 255) / 256)), dim3(256), 0, stream, R,
                             widths[w], NZ, g.pml, pmask + moff[w]);
      }
    }
    host.resize(ns);
    const long long origin = field_origin(ld, g.pml);
    char* sp = reinterpret_cast<char*>(snaps);
    float* sc = scales;
    T* fp = pool;
    T* gp = grads;
    for (int s = 0; s < ns; s++) {
      const ShotSpec& q = spec[s];
      const int nx = NXs[s] - 2 * g.pml;
      const long long fe = field_elems(NXs[s], ld);
      DevShot<T>& d = host[s];
      std::memset(&d, 0, sizeof(d));
      d.vel = vel + q.model;
      d.wav = wav + q.wav;
      d.gz = gz + q.gz;
      d.fld = fp + origin;
      d.fstride = fe;
      fp += fe * LS_NSLOT;
      // the forward passes propagate one field; the raw-adjoint slots hold
      // their band cells' stencil values (the adjoint clears them before use)
      d.raw = d.fld + (size_t)fe * LS_AR;
      d.snap = sp;
      d.scale = sc;
      d.grad = gp;
      gp += (long long)nx * g.nz;
      d.snap_kind = kind;
      d.nx = nx;
      d.NX = NXs[s];
      d.x0 = q.x0;
      d.nt = q.nt;
      d.ndel = q.ndel;
      d.sx = q.sx;
      d.sz = q.sz;
      const size_t wi = std::lower_bound(widths.begin(), widths.end(), NXs[s]) - widths.begin();
      d.pmask = pmask ? pmask + moff[wi] : nullptr;
      d.wtab = wtab;
      sp += snap_elem() * (size_t)(q.nt + q.ndel) * nx * g.nz;
      sc += q.nt + q.ndel;
    }
    LS_CHECK(hipMalloc(&dev, sizeof(DevShot<T>) * ns));
    LS_CHECK(hipMemcpy(dev, host.data(), sizeof(DevShot<T>) * ns, hipMemcpyHostToDevice));
    host_adj = host;
    for (DevShot<T>& d : host_adj) {
      d.fld += adj_shift;
      d.raw += adj_shift;
    }
    LS_CHECK(hipMalloc(&dev_adj, sizeof(DevShot<T>) * ns));
    LS_CHECK(hipMemcpy(dev_adj, host_adj.data(), sizeof(DevShot<T>) * ns, hipMemcpyHostToDevice));
    {
      // strip length of the forward passes: the background step with the
      // max-abs reduction at step 0 on zeroed fields (writes zeros only)
      int rc = zero_fields(0, LS_NSLOT);
      if (rc) return rc;
      auto fn = [](void* c, int bx) {
        LsrtmPlan* p = static_cast<LsrtmPlan*>(c);
        launch_step<T>(SNAP_MAX, p->k, p->dev, with_strip(p->F, bx), p->B, p->nshots, 0, p->stream);
      };
      F = with_strip(F, tuned_strip("lsrtm_forward", NXmax, NZ, nshots, stream, fn, this));
    }
    // band coefficients (the slots the forward passes leave alone)
    const long long nband = shot_split(g.pml, NXmax, NZ).n;
    if (nband > 0)
      hipLaunchKernelGGL((band_coef_kernel<T>), dim3((unsigned)((nband + 255) / 256), ns), dim3(256), 0, stream, k,
                         dev_adj);
    LS_CHECK(hipGetLastError());
    return make_snapshots();
  }

  // padded cells and model cells of the shots still running at forward step st
  double cells_fwd(int st, double* mcells = nullptr) const {
    double c = 0, m = 0;
    for (int s = 0; s < nshots; s++)
      if (st < spec[s].nt + spec[s].ndel) {
        c += (double)NXs[s] * NZ;
        m += (double)(NXs[s] - 2 * g.pml) * g.nz;
      }
    if (mcells) *mcells = m;
    return c;
  }

  int zero_fields(int first, int count) {
    for (int s = 0; s < nshots; s++) {
      const long long fe = host[s].fstride;
      T* base = host[s].fld - (field_origin(ld, g.pml));
      LS_CHECK(hipMemsetAsync(base + (size_t)fe * first, 0, sizeof(T) * (size_t)fe * count, stream));
    }
    return SEIS_OK;
  }

  int make_snapshots() {
    const double E = sizeof(T);
    const int passes = snap_kind == SNAP_NATIVE ? 1 : 2;
    for (int pass = 0; pass < passes; pass++) {
      const int mode = (passes == 2 && pass == 0) ? SNAP_MAX : SNAP_STORE;
      int rc = zero_fields(LS_FIELD, 3);
      if (rc) return rc;
      for (int st = 0; st < NTmax; st++) {
        double mc;
        const double c = cells_fwd(st, &mc);
        const double by = 4 * E * c + (mode == SNAP_STORE ? snap_elem() * mc : 0.0);
        timed_launch(stream, mode == SNAP_MAX ? "lsrtm_snap_max" : "lsrtm_snap_store", by, c,
                     [&] { launch_step<T>(mode, k, dev, F, B, nshots, st, stream); });
      }
    }
    LS_CHECK(hipGetLastError());
    LS_CHECK(hipStreamSynchronize(stream));
    return SEIS_OK;
  }

  // rec (device, packed like the shot specs) = L pert (device [nx][nz])
  int born(const T* pert, T* rec) {
    for (int s = 0; s < nshots; s++) {
      host[s].dv = pert + (long long)spec[s].x0 * g.nz;  // the shot's window of the perturbation
      host[s].rec = rec + spec[s].rec;
    }
    LS_CHECK(hipMemcpyAsync(dev, host.data(), sizeof(DevShot<T>) * nshots, hipMemcpyHostToDevice, stream));
    int rc = zero_fields(LS_FIELD, 3);
    if (rc) return rc;
    const double E = sizeof(T);
    for (int st = 0; st < NTmax; st++) {
      double mc;
      const double c = cells_fwd(st, &mc);
      const double by = 4 * E * c + (snap_elem() + E) * mc;
      timed_launch(stream, "lsrtm_born", by, c, [&] { launch_step<T>(BORN_S, k, dev, F, B, nshots, st, stream); });
    }
    if (mask) hipLaunchKernelGGL((mul_kernel<T>), dim3(1024), dim3(256), 0, stream, rec, rec, mask, nrec);
    LS_CHECK(hipGetLastError());
    LS_CHECK(hipStreamSynchronize(stream));
    return SEIS_OK;
  }

  // grad (device [nx][nz]) = the adjoint of the plan's kind applied to rec
  int adjoint(const T* rec, T* grad) {
    return adj_kind == SEIS_LSRTM_ADJ_XCORR ? adjoint_xcorr(rec, grad) : adjoint_exact(rec, grad);
  }

  // Cross-correlation adjoint: the receiver field run backwards in time with
  // the forward operator (stencil, HABC, sponge), the record injected times
  // vt2 at the receivers, and grad += u S^n / vt2 at every model cell and
  // step.  In the interior this field is vt2 times the exact adjoint field,
  // so it images like L^T there; in the absorbing band it is the forward
  // operator instead of its transpose (the reference's RTM imaging, fd.cpp:
  // 592-719, with stored snapshots in place of the rebuilt source field).
  int adjoint_xcorr(const T* rec, T* grad) {
    if (mask) {
      if (!masked) LS_CHECK(hipMalloc(&masked, sizeof(T) * (size_t)nrec));
      hipLaunchKernelGGL((mul_kernel<T>), dim3(1024), dim3(256), 0, stream, masked, rec, mask, nrec);
      rec = masked;
    }
    for (int s = 0; s < nshots; s++) host[s].rec = const_cast<T*>(rec) + spec[s].rec;
    LS_CHECK(hipMemcpyAsync(dev, host.data(), sizeof(DevShot<T>) * nshots, hipMemcpyHostToDevice, stream));
    int rc = zero_fields(LS_FIELD, 3);
    if (rc) return rc;
    LS_CHECK(hipMemsetAsync(grads, 0, sizeof(T) * (size_t)model_cells, stream));
    const double E = sizeof(T);
    for (int st = 0; st < NTmax; st++) {
      double mc;
      const double c = cells_fwd(st, &mc);  // shot s runs backward steps st < nt + ndel
      const double by = 4 * E * c + (snap_elem() + 2 * E) * mc;
      timed_launch(stream, "lsrtm_adjoint_xcorr", by, c,
                   [&] { launch_step<T>(ADJ_X, k, dev, F, B, nshots, st, stream); });
    }
    hipLaunchKernelGGL((sum_grads_kernel<T>), dim3(256), dim3(256), 0, stream, dev, nshots, g.nz, model, grad);
    LS_CHECK(hipGetLastError());
    LS_CHECK(hipStreamSynchronize(stream));
    return SEIS_OK;
  }

  // Work-group counts of the two sweep launches at strip length adj_bx: the
  // most DEEP and the most other items of a shot.  deep_box is checked
  // against deep_item on every item of every shot first.
  int adj_bx = -1, ndwg = 0, nfrwg = 0, npwg = 0;
  int adj_counts(const Layout& a) {
    if (a.bx == adj_bx) return SEIS_OK;
    int dmax = 0, fmax = 0;
    for (int s = 0; s < nshots; s++) {
      const Split sq = shot_split(a.m, NXs[s], NZ);
      const int nxs = shot_strips(sq, a.bx);
      const DeepBox d = deep_box(sq, a.nzc, a.bx, NXs[s], NZ, g.pml);
      for (int xs = 0; xs < nxs; xs++)
        for (int zc = 0; zc < a.nzc; zc++) {
          const bool in = zc >= d.zc0 && zc < d.zc0 + d.nzd && xs >= d.xs0 && xs < d.xs0 + d.nxd;
          if (in != deep_item(sq, zc, xs, a.bx, NXs[s], NZ, g.pml))
            return set_error("lsrtm: adjoint sweep items do not split into a rectangle"), SEIS_ERR_ARG;
        }
      dmax = std::max(dmax, d.nzd * d.nxd);
      fmax = std::max(fmax, a.nzc * nxs - d.nzd * d.nxd);
    }
    auto wgs = [](int items) {
      const int w = (items + WAVES - 1) / WAVES;
      return (w + NXCD - 1) / NXCD * NXCD;  // xcd_tile deals whole rounds
    };
    ndwg = wgs(dmax);
    nfrwg = wgs(fmax);
    npwg = P.nbwg;  // the ring-P cells of the widest shot, AWG per work-group
    adj_bx = a.bx;
    return SEIS_OK;
  }

  // The two launches of reverse step st at the counts of adj_counts(a): ring
  // P with the DEEP items, the other items with the outer ring.
  void launch_adj_step(const Layout& a, int st) {
    if (npwg + ndwg > 0)
      hipLaunchKernelGGL((adj_deep_kernel<T>), dim3(npwg + ndwg, nshots), dim3(AWG), 0, stream, k, dev_adj, a, npwg,
                         ndwg, st);
    if (nfrwg + a.nbwg > 0)
      hipLaunchKernelGGL((adj_band_kernel<T>), dim3(nfrwg + a.nbwg, nshots), dim3(AWG), 0, stream, k, dev_adj, a,
                         nfrwg, st);
  }

  // grad (device [nx][nz]) = L^T rec, summed over shots
  int adjoint_exact(const T* rec, T* grad) {
    if (mask) {
      if (!masked) LS_CHECK(hipMalloc(&masked, sizeof(T) * (size_t)nrec));
      hipLaunchKernelGGL((mul_kernel<T>), dim3(1024), dim3(256), 0, stream, masked, rec, mask, nrec);
      rec = masked;
    }
    for (int s = 0; s < nshots; s++) host_adj[s].rec = host[s].rec = const_cast<T*>(rec) + spec[s].rec;
    LS_CHECK(hipMemcpyAsync(dev_adj, host_adj.data(), sizeof(DevShot<T>) * nshots, hipMemcpyHostToDevice, stream));
    int rc;
    if (!adj_tuned) {
      // strip length of the adjoint sweep: its step 0 (no gradient update
      // yet); the fields it writes are cleared below
      auto fn = [](void* c, int bx) {
        LsrtmPlan* p = static_cast<LsrtmPlan*>(c);
        const Layout a = with_strip(p->A, bx);
        if (p->adj_counts(a) == SEIS_OK) p->launch_adj_step(a, 0);
      };
      A = with_strip(A, tuned_strip("lsrtm_adjoint", NXmax, NZ, nshots, stream, fn, this));
      adj_tuned = true;
    }
    rc = adj_counts(A);
    if (rc) return rc;
    rc = zero_fields(0, LS_NADJ);
    if (rc) return rc;
    LS_CHECK(hipMemsetAsync(grads, 0, sizeof(T) * (size_t)model_cells, stream));
    const double E = sizeof(T);
    for (int st = 0; st <= NTmax; st++) {
      double pc = 0, c = 0, mc = 0;
      for (int s = 0; s < nshots; s++)
        if ((spec[s].nt + spec[s].ndel + 1 - st) >= 1) {
          pc += (double)shot_split(g.pml + 1, NXs[s], NZ).n;
          c += (double)NXs[s] * NZ;
          mc += (double)(NXs[s] - 2 * g.pml) * g.nz;
        }
      // sweep: 16 B per cell plus snapshot and gradient; prep: 8 B per ring-P cell
      const double by = 4 * E * c + (snap_elem() + 2 * E) * mc + 8 * E * pc;
      timed_launch(stream, "lsrtm_adjoint", by, c, [&] { launch_adj_step(A, st); });
    }
    hipLaunchKernelGGL((sum_grads_kernel<T>), dim3(256), dim3(256), 0, stream, dev_adj, nshots, g.nz, model, grad);
    LS_CHECK(hipGetLastError());
    LS_CHECK(hipStreamSynchronize(stream));
    return SEIS_OK;
  }

  int dot(const T* a, const T* b, long long n, double* out) {
    hipLaunchKernelGGL((dot_partial_kernel<T>), dim3(RGRID), dim3(RB), 0, stream, a, b, n, part);
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(RB), 0, stream, part, RGRID, part + RGRID);
    LS_CHECK(hipMemcpyAsync(out, part + RGRID, sizeof(double), hipMemcpyDeviceToHost, stream));
    LS_CHECK(hipStreamSynchronize(stream));
    return SEIS_OK;
  }

  // CGLS on min ||L x - d||^2 from x = 0 (with the record mask, if set, in
  // L).  hist[2*i] = ||r_i||^2, hist[2*i+1] = ||s_i||^2 for i = 0..done;
  // stops after niter iterations or when ||s|| <= rtol ||s_0||.
  // With a weight W set, CGLS runs on the operator L W (right
  // preconditioning: x is the preconditioned image) and returns m = W x.
  //
  // The gradients s_i = W L^T r_i, orthogonal in exact arithmetic, are kept
  // orthogonal: each new s is Gram-Schmidt-orthogonalised in FP64 against the
  // earlier ones (stored normalised, FP64).  Without it CG loses
  // orthogonality within ~10 iterations on these ill-conditioned operators
  // and its iterates drift from the exact-arithmetic ones by far more than
  // the rounding of the operator (FP64 included); with it the FP32 plan's
  // iterates stay within the operator's rounding of the FP64 plan's.  The
  // directions p, the gradients and the inner products are FP64; the
  // operator applications run in the plan's precision.
  int cgls(const T* d, int niter, double rtol, T* x_out, double* hist, int* done, T* iterates = nullptr) {
    const size_t need64 = (size_t)model * (2 + (niter + 1));
    if (!cg) LS_CHECK(hipMalloc(&cg, sizeof(T) * (size_t)(3 * model + 2 * nrec)));
    if (cg64_elems < need64) {
      if (cg64) (void)hipFree(cg64);
      cg64 = nullptr;
      LS_CHECK(hipMalloc(&cg64, sizeof(double) * need64));
      cg64_elems = need64;
    }
    T *x = x_out, *s = cg, *pt = cg + model, *r = cg + 2 * model, *q = cg + 2 * model + nrec;
    T* wp = cg + 2 * model + 2 * nrec;  // W p
    double *s64 = cg64, *p64 = cg64 + model, *U = cg64 + 2 * model;  // U: orthonormal gradients
    const int G = 1024, B = 256;
    // the operator's adjoint side: s = W (L^T r)
    auto adjoint_w = [&](const T* rr, T* ss) {
      int e = adjoint(rr, ss);
      if (!e && weight) hipLaunchKernelGGL((mul_kernel<T>), dim3(G), dim3(B), 0, stream, ss, ss, weight, model);
      return e;
    };
    // s64 = s orthogonalised against the k stored gradients, stored as U[k];
    // returns ||s64||^2
    auto orthogonalise = [&](int k, double* norm2) {
      hipLaunchKernelGGL((scale_to64_kernel<T>), dim3(G), dim3(B), 0, stream, s64, s, 1.0, model);
      for (int j = 0; j < k; j++) {
        double c;
        int e = dot64(s64, U + (size_t)j * model, model, &c);
        if (e) return e;
        hipLaunchKernelGGL((axpy_kernel<double>), dim3(G), dim3(B), 0, stream, s64, U + (size_t)j * model, -c,
                           model);
      }
      int e = dot64(s64, s64, model, norm2);
      if (e) return e;
      if (*norm2 > 0)
        hipLaunchKernelGGL((scale_to64_kernel<double>), dim3(G), dim3(B), 0, stream, U + (size_t)k * model, s64,
                           1.0 / std::sqrt(*norm2), model);
      return SEIS_OK;
    };
    hipLaunchKernelGGL((fill_kernel<T>), dim3(G), dim3(B), 0, stream, x, T(0), model);
    LS_CHECK(hipMemcpyAsync(r, d, sizeof(T) * (size_t)nrec, hipMemcpyDeviceToDevice, stream));
    int rc = adjoint_w(r, s);
    if (rc) return rc;
    double gamma, rr;
    if ((rc = orthogonalise(0, &gamma)) || (rc = dot(r, r, nrec, &rr))) return rc;
    LS_CHECK(hipMemcpyAsync(p64, s64, sizeof(double) * (size_t)model, hipMemcpyDeviceToDevice, stream));
    if (hist) {
      hist[0] = rr;
      hist[1] = gamma;
    }
    if (done) *done = 0;
    const double gamma0 = gamma;
    for (int it = 0; it < niter; it++) {
      if (!(gamma > 0) || (rtol > 0 && gamma <= rtol * rtol * gamma0)) break;
      hipLaunchKernelGGL((from64_kernel<T>), dim3(G), dim3(B), 0, stream, pt, p64, model);
      if (weight) {
        hipLaunchKernelGGL((mul_kernel<T>), dim3(G), dim3(B), 0, stream, wp, pt, weight, model);
        if ((rc = born(wp, q))) return rc;
      } else if ((rc = born(pt, q))) {
        return rc;
      }
      double delta;
      if ((rc = dot(q, q, nrec, &delta))) return rc;
      if (!(delta > 0)) break;
      const double alpha = gamma / delta;
      hipLaunchKernelGGL((axpy_from64_kernel<T>), dim3(G), dim3(B), 0, stream, x, p64, alpha, model);
      hipLaunchKernelGGL((axpy_kernel<T>), dim3(G), dim3(B), 0, stream, r, q, -alpha, nrec);
      if (iterates) {  // the image after this iteration (W x with a weight)
        T* xi = iterates + (size_t)it * model;
        if (weight)
          hipLaunchKernelGGL((mul_kernel<T>), dim3(G), dim3(B), 0, stream, xi, x, weight, model);
        else
          LS_CHECK(hipMemcpyAsync(xi, x, sizeof(T) * (size_t)model, hipMemcpyDeviceToDevice, stream));
      }
      if ((rc = adjoint_w(r, s))) return rc;
      double gnew;
      if ((rc = orthogonalise(it + 1, &gnew)) || (rc = dot(r, r, nrec, &rr))) return rc;
      const double beta = gnew / gamma;
      gamma = gnew;
      hipLaunchKernelGGL((xpby_kernel<double>), dim3(G), dim3(B), 0, stream, p64, s64, beta, model);
      if (hist) {
        hist[2 * (it + 1)] = rr;
        hist[2 * (it + 1) + 1] = gamma;
      }
      if (done) *done = it + 1;
    }
    if (weight) hipLaunchKernelGGL((mul_kernel<T>), dim3(G), dim3(B), 0, stream, x, x, weight, model);
    LS_CHECK(hipGetLastError());
    LS_CHECK(hipStreamSynchronize(stream));
    return SEIS_OK;
  }

  int dot64(const double* a, const double* b, long long n, double* out) {
    hipLaunchKernelGGL((dot_partial_kernel<double>), dim3(RGRID), dim3(RB), 0, stream, a, b, n, part);
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(RB), 0, stream, part, RGRID, part + RGRID);
    LS_CHECK(hipMemcpyAsync(out, part + RGRID, sizeof(double), hipMemcpyDeviceToHost, stream));
    LS_CHECK(hipStreamSynchronize(stream));
    return SEIS_OK;
  }
};

}  // namespace seis

using namespace seis;

struct seis_lsrtm {
  int fp64 = 0;
  LsrtmPlan<float>* f = nullptr;
  LsrtmPlan<double>* d = nullptr;
  ~seis_lsrtm() {
    delete f;
    delete d;
  }
};

extern "C" {

int seis_lsrtm_create(seis_lsrtm** plan, const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                      const float* wav, const int* gz, int snapshot_kind) {
  if (!plan || !grid || !shots || !vel || !wav || !gz) return set_error("lsrtm: null argument"), SEIS_ERR_ARG;
  seis_lsrtm* p = new seis_lsrtm();
  p->f = new LsrtmPlan<float>();
  const int rc = p->f->create(*grid, nshots, shots, vel, wav, gz, snapshot_kind);
  if (rc) {
    delete p;
    *plan = nullptr;
    return rc;
  }
  *plan = p;
  return SEIS_OK;
}

int seis_lsrtm_create_f64(seis_lsrtm** plan, const seis_grid* grid, int nshots, const seis_shot* shots,
                          const double* vel, const double* wav, const int* gz) {
  if (!plan || !grid || !shots || !vel || !wav || !gz) return set_error("lsrtm: null argument"), SEIS_ERR_ARG;
  seis_lsrtm* p = new seis_lsrtm();
  p->fp64 = 1;
  p->d = new LsrtmPlan<double>();
  const int rc = p->d->create(*grid, nshots, shots, vel, wav, gz, SNAP_NATIVE);
  if (rc) {
    delete p;
    *plan = nullptr;
    return rc;
  }
  *plan = p;
  return SEIS_OK;
}

int seis_lsrtm_destroy(seis_lsrtm* plan) {
  delete plan;
  return SEIS_OK;
}

int seis_lsrtm_born(seis_lsrtm* plan, const void* pert, void* rec) {
  if (!plan || !pert || !rec) return set_error("lsrtm: null argument"), SEIS_ERR_ARG;
  return plan->fp64 ? plan->d->born((const double*)pert, (double*)rec) : plan->f->born((const float*)pert, (float*)rec);
}

int seis_lsrtm_adjoint(seis_lsrtm* plan, const void* rec, void* grad) {
  if (!plan || !rec || !grad) return set_error("lsrtm: null argument"), SEIS_ERR_ARG;
  return plan->fp64 ? plan->d->adjoint((const double*)rec, (double*)grad)
                    : plan->f->adjoint((const float*)rec, (float*)grad);
}

int seis_lsrtm_dot(seis_lsrtm* plan, const void* a, const void* b, long long n, double* out) {
  if (!plan || !a || !b || !out) return set_error("lsrtm: null argument"), SEIS_ERR_ARG;
  return plan->fp64 ? plan->d->dot((const double*)a, (const double*)b, n, out)
                    : plan->f->dot((const float*)a, (const float*)b, n, out);
}

int seis_lsrtm_set_adjoint(seis_lsrtm* plan, int kind) {
  if (!plan) return set_error("null plan"), SEIS_ERR_ARG;
  if (kind != SEIS_LSRTM_ADJ_EXACT && kind != SEIS_LSRTM_ADJ_XCORR)
    return set_error("lsrtm: adjoint kind must be SEIS_LSRTM_ADJ_EXACT or SEIS_LSRTM_ADJ_XCORR"), SEIS_ERR_ARG;
  if (plan->fp64)
    plan->d->adj_kind = kind;
  else
    plan->f->adj_kind = kind;
  return SEIS_OK;
}

int seis_lsrtm_set_weight(seis_lsrtm* plan, const void* weight) {
  if (!plan) return set_error("null plan"), SEIS_ERR_ARG;
  if (plan->fp64)
    plan->d->weight = static_cast<const double*>(weight);
  else
    plan->f->weight = static_cast<const float*>(weight);
  return SEIS_OK;
}

int seis_lsrtm_set_mask(seis_lsrtm* plan, const void* mask) {
  if (!plan) return set_error("lsrtm: null plan"), SEIS_ERR_ARG;
  if (plan->fp64) plan->d->mask = (const double*)mask;
  else plan->f->mask = (const float*)mask;
  return SEIS_OK;
}

int seis_lsrtm_cg(seis_lsrtm* plan, const void* dobs, int niter, double rtol, void* image, double* hist,
                  int* iters) {
  return seis_lsrtm_cg_history(plan, dobs, niter, rtol, image, hist, iters, nullptr);
}

int seis_lsrtm_cg_history(seis_lsrtm* plan, const void* dobs, int niter, double rtol, void* image, double* hist,
                          int* iters, void* iterates) {
  if (!plan || !dobs || !image || niter < 0) return set_error("lsrtm: bad argument"), SEIS_ERR_ARG;
  return plan->fp64 ? plan->d->cgls((const double*)dobs, niter, rtol, (double*)image, hist, iters, (double*)iterates)
                    : plan->f->cgls((const float*)dobs, niter, rtol, (float*)image, hist, iters, (float*)iterates);
}

}  // extern "C"
