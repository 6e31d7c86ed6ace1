// HIP engine for the 2D acoustic finite-difference path on gfx950:
// forward modeling, Born modeling and reference-semantics RTM, batched over
// shots.  One launch per time step advances every shot of the batch; the
// stencil, absorbing boundary, sponge blend, source/receiver injection,
// receiver extraction, Born source, boundary-strip save/restore and the
// pointwise imaging term are fused into that launch (see fd_cell.h).  The
// Hilbert imaging term of the reference RTM runs as a second, column-wise
// kernel per reverse step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"
#include "fd_cell.h"
#include "prof.h"
#include "seis_internal.h"

namespace seis {

// Born source added to the scattered field (fd.cpp:134-139): vt2 * lap * 2/v^3 * dv,
// lap = (u_new + u_prev - 2 u_cur) / dt^2 of the background.  dv is zero in the
// padding (setDv), where the new background equals its stencil value.
template <typename T>
struct BornInject {
  const Consts<T>* k;
  const Model<T>* m;
  const T* dv;
  const T* Cb;
  const T* Pb;
  PointSource<T> src;
  int NX;
  long long own;  // field index whose background value is already known
  T own_nb;
  __device__ __forceinline__ T add(T val, int x, int z, T vt2, long long i) const {
    const int pml = k->pml;
    const int ix = x - pml, iz = z - pml;
    if (ix < 0 || ix >= m->nx || iz < 0 || iz >= m->nz) return val;
    const T d = dv[(long long)ix * m->nz + iz];
    if (d == T(0)) return val;
    const T nb = (i == own) ? own_nb : raw_value(*k, *m, NX, Cb, Pb, src, x, z);
    const T lap = (nb + Pb[i] - T(2) * Cb[i]) * k->idt2;
    const T v = m->vel(x, z);
    const T v3 = T(2) / (v * v * v);
    return val + vt2 * lap * v3 * d;
  }
  __device__ __forceinline__ T prev(const T* P, long long i, int, int) const { return P[i]; }
};

// Born source read back from the snapshots: S(step) * p at model cells.
template <typename T>
struct SnapInject {
  const DevShot<T>* s;
  int nz, pml, it;
  __device__ __forceinline__ T add(T val, int x, int z, T, long long) const {
    const int ix = x - pml, iz = z - pml;
    if (ix < 0 || ix >= s->nx || iz < 0 || iz >= nz) return val;
    const long long j = (long long)ix * nz + iz;
    const T d = s->dv[j];
    if (d == T(0)) return val;
    return val + snap_load(*s, nz, it, j) * d;
  }
  __device__ __forceinline__ T prev(const T* P, long long i, int, int) const { return P[i]; }
};

// The source of step `it` of a shot: the point source, or the plane-wave
// line source (DevShot::line_src); inactive at it = 0.
template <typename T>
__device__ __forceinline__ PointSource<T> step_source(const Consts<T>& k, const DevShot<T>& s, long long src, int it) {
  PointSource<T> ps{-1, T(0)};
  if (it - 1 < 0) return ps;
  if (s.line_src) {
    ps.line = s.wav + (it - 1);
    ps.lstride = s.nt + s.ndel;
    ps.lrow = s.sz + k.pml;
    ps.lx0 = k.pml;
    ps.lnx = s.nx;
  } else {
    ps.src = src;
    ps.amp = s.wav[it - 1];
  }
  return ps;
}

// Generic update of one cell: any cell of the padded grid, including the
// absorbing band and the outer ring, through the pure per-cell functions of
// fd_cell.h (neighbouring stencil values are recomputed, not read).
template <typename T, int MODE>
__device__ __forceinline__ void cell_update(const Consts<T>& k, const DevShot<T>& s, const StepCtx<T>& c, int x,
                                            int z) {
  const int NX = s.NX;
  const long long ld = k.ld;
  const long long i = fidx(ld, x, z);
  const int it = c.it, itp = c.itp;
  const T* P = c.P;
  const T* C = c.C;
  T* N = c.N;
  const Model<T>& m = c.m;
  if (MODE == SNAP_MAX || MODE == SNAP_STORE) {
    // band cells lie outside the model (pml >= 4 is required): background only
    const PointSource<T> ps = step_source(k, s, c.src, it);
    N[i] = final_value(k, m, NX, C, P, ps, x, z);
  } else if (MODE == BORN_S) {
    SnapInject<T> si{&s, k.nz, k.pml, it};
    const T u = final_value(k, m, NX, C, P, si, x, z);
    N[i] = u;
    const int ix = x - k.pml;
    if (itp >= 0 && ix >= 0 && ix < s.nx && z == s.gz[ix] + k.pml) s.rec[(long long)ix * s.nt + itp] = u;
  } else if (MODE == MODEL || MODE == RTM_FWD || MODE == ILLUM) {
    const PointSource<T> ps = step_source(k, s, c.src, it);
    const T u = final_value(k, m, NX, C, P, ps, x, z);
    N[i] = u;
    if (MODE == ILLUM) {
      const int ix = x - k.pml, iz = z - k.pml;
      if (it > s.ndel && ix >= 0 && ix < s.nx && iz >= 0 && iz < k.nz) s.img[i] += u * u;
    } else if (MODE == MODEL) {
      const int ix = x - k.pml;
      if (itp >= 0 && ix >= 0 && ix < s.nx && z == s.gz[ix] + k.pml) s.rec[(long long)ix * s.nt + itp] = u;
    } else if (itp >= 0) {
      int sl[4];
      const int n = strip_slots(x, z, s.nx, k.nz, k.pml, sl);
      T* st = s.strips + (long long)itp * 8 * (s.nx + k.nz);
      for (int q = 0; q < n; q++) st[sl[q]] = u;
    }
  } else if (MODE == PW_BACK) {
    // receiver field of adjointPlane, backward in time: record sample
    // it - ndel injected at the receivers; the strips of the new level are
    // the ones saved at step it - 1 (fd.cpp:829-846)
    const ReceiverInject<T> ri{s.gz, s.rec, s.nx, s.nt, it - s.ndel, k.pml};
    const T u = final_value(k, m, NX, C, P, ri, x, z);
    N[i] = u;
    if (it >= 1) {
      int sl[4];
      const int n = strip_slots(x, z, s.nx, k.nz, k.pml, sl);
      T* st = s.strips + (long long)(it - 1) * 8 * (s.nx + k.nz);
      for (int q = 0; q < n; q++) st[sl[q]] = u;
    }
  } else if (MODE == ADJ_X) {
    // cross-correlation adjoint: record sample it - ndel injected (times
    // vt2), then grad += u S^it / vt2 at model cells
    const ReceiverInject<T> ri{s.gz, s.rec, s.nx, s.nt, it - s.ndel, k.pml};
    const T u = final_value(k, m, NX, C, P, ri, x, z);
    N[i] = u;
    const int ix = x - k.pml, iz = z - k.pml;
    if (ix >= 0 && ix < s.nx && iz >= 0 && iz < k.nz) {
      const long long j = (long long)ix * k.nz + iz;
      const T a = k.dt * m.vel(x, z);
      const T vt2 = a * a;
      s.grad[j] = s.grad[j] + u * (snap_load(s, k.nz, it, j) / vt2);
    }
  } else if (MODE == PW_FWD) {
    // line-source field and U = lap 2/v^3; the receiver field rebuilt
    // forward with record sample it - ndel - 1 removed from its previous
    // level (fd.cpp:851-867)
    const PointSource<T> ps = step_source(k, s, c.src, it);
    const T u = final_value(k, m, NX, C, P, ps, x, z);
    const T lap = (u + P[i] - T(2) * C[i]) * k.idt2;
    const T v = m.vel(x, z);
    c.aux[i] = lap * (T(2) / (v * v * v));
    N[i] = u;
    const ReceiverRemoval<T> rr{s.gz, s.rec, s.nx, s.nt, it - s.ndel - 1, k.pml, k.dt, &m};
    N[3 * s.fstride + i] = final_value(k, m, NX, C + 3 * s.fstride, P + 3 * s.fstride, rr, x, z);
  } else if (MODE == BORN) {
    const PointSource<T> ps = step_source(k, s, c.src, it);
    T nb_raw;
    const T ub = final_value(k, m, NX, C, P, ps, x, z, &nb_raw);
    const T* Ps = P + 3 * s.fstride;
    const T* Cs = C + 3 * s.fstride;
    T* Ns = N + 3 * s.fstride;
    BornInject<T> bi{&k, &m, s.dv, C, P, ps, NX, i, nb_raw};
    const T us = final_value(k, m, NX, Cs, Ps, bi, x, z);
    N[i] = ub;
    Ns[i] = us;
    const int ix = x - k.pml;
    if (itp >= 0 && ix >= 0 && ix < s.nx && z == s.gz[ix] + k.pml) s.rec[(long long)ix * s.nt + itp] = us;
  } else {  // RTM_BWD
    const bool on = it - 1 > 0;
    const T v = m.vel(x, z);
    T sub = T(0);
    if (on) {
      const T as = k.dt * m.vel(s.sx + k.pml, s.sz + k.pml);
      sub = (as * as) * s.wav[it - 1];
    }
    SourceRemoval<T> sr{c.src, sub, on};
    // With SEIS_RTM_ISOLATE_BAND the reconstructed source field is held at
    // zero outside the model area: run backwards, the absorbing band grows
    // without bound, and the strips already shield the interior from it.
    const bool band = x < k.pml || x >= k.pml + s.nx || z < k.pml || z >= k.pml + k.nz;
    const bool iso = band && (s.rtm_flags & SEIS_RTM_ISOLATE_BAND);
    const T u = iso ? T(0) : final_value(k, m, NX, C, P, sr, x, z);
    const T lap = (sr.prev(P, i, x, z) + u - T(2) * C[i]) * k.idt2;
    const T* Pr = P + 3 * s.fstride;
    const T* Cr = C + 3 * s.fstride;
    T* Nr = N + 3 * s.fstride;
    ReceiverInject<T> ri{s.gz, s.rec, s.nx, s.nt, itp, k.pml};
    const T ur = final_value(k, m, NX, Cr, Pr, ri, x, z);
    const T v3 = T(2) / (v * v * v);
    c.aux[i] = lap * v3;  // U of the imaging condition; the image pass reads it
    Nr[i] = ur;
    // the current level of the next reverse step gets the saved strips
    // restored (fd.cpp:663); lap above used the value before restoration
    T out = u;
    if (itp - 1 >= 0) {
      const int sl = first_strip_slot(x, z, s.nx, k.nz, k.pml);
      if (sl >= 0) out = s.strips[(long long)(itp - 1) * 8 * (s.nx + k.nz) + sl];
    }
    N[i] = out;
    if (edge_dist(x, z, NX, k.NZ) == k.pml) s.raw[i] = u;  // stencil value for the band (see interior_sweep)
  }
}

// Band cells (closer than pml to the outer edge, outside the model): the
// step kernel stores their stencil values ("u3" of processBoundary) and the
// band kernel turns them into new values once every stencil value of the
// level exists.  The sources, receivers, Born and snapshot terms all lie
// inside the model, so only the plain stencil (or ring formula) is needed
// here; SourceRemoval only changes the previous level at the source cell.
template <typename T, int MODE>
__device__ __forceinline__ void band_raw(const Consts<T>& k, const DevShot<T>& s, const StepCtx<T>& c, int x,
                                         int z) {
  const int NX = s.NX;
  const long long i = fidx(k.ld, x, z);
  const Model<T>& m = c.m;
  T* RA = s.raw;
  T* RB = s.raw + s.fstride;
  if (MODE == RTM_BWD) {
    if (!(s.rtm_flags & SEIS_RTM_ISOLATE_BAND)) {
      T sub = T(0);
      const bool on = c.it - 1 > 0;
      if (on) {
        const T as = k.dt * m.vel(s.sx + k.pml, s.sz + k.pml);
        sub = (as * as) * s.wav[c.it - 1];
      }
      RA[i] = raw_value(k, m, NX, c.C, c.P, SourceRemoval<T>{c.src, sub, on}, x, z);
    }
    RB[i] = raw_value(k, m, NX, c.C + 3 * s.fstride, c.P + 3 * s.fstride, NoInject{}, x, z);
  } else {
    RA[i] = raw_value(k, m, NX, c.C, c.P, NoInject{}, x, z);
    if (MODE == BORN || MODE == PW_FWD)
      RB[i] = raw_value(k, m, NX, c.C + 3 * s.fstride, c.P + 3 * s.fstride, NoInject{}, x, z);
  }
}

// Stencil value of a neighbour at the new level: stored for band cells; for
// the others the new value itself (outside the band it is the stencil value).
// `lim` is pml, or pml + 1 for the reconstructed RTM source field, whose new
// level holds restored strips in the first model layer while the band reads
// that layer's stencil values (kept in R by the step kernel).
template <typename T>
struct RawNeighbour {
  const T* R;
  const T* N;
  long long ld;
  int NX, NZ, lim;
  __device__ __forceinline__ T operator()(int xx, int zz) const {
    const long long q = fidx(ld, xx, zz);
    return edge_dist(xx, zz, NX, NZ) < lim ? SEIS_GL(R)[q] : SEIS_GL(N)[q];
  }
};

// One-way update and blend of every band cell of every shot (after the step
// kernel of the same step).
template <typename T, int MODE>
__global__ __launch_bounds__(WG) void band_kernel(Consts<T> k, const DevShot<T>* __restrict__ shots, int step) {
  const DevShot<T>& s = shots[blockIdx.y];
  const int it = step_it<T>(MODE, s, step);
  if (it < 0) return;
  const int NX = s.NX, NZ = k.NZ, pml = k.pml;
  const Ring B = shot_split(pml, NX, NZ);  // this shot's absorbing band
  const long long j = (long long)blockIdx.x * WG + threadIdx.x;
  if (j >= B.n) return;
  int x, z;
  ring_cell(B, NX, NZ, j, &x, &z);
  const long long ld = k.ld;
  const long long i = fidx(ld, x, z);
  const int r0 = step % 3, r1 = (step + 1) % 3, r2 = (step + 2) % 3;
  const T* P = s.fld + r0 * s.fstride;
  const T* C = s.fld + r1 * s.fstride;
  T* N = s.fld + r2 * s.fstride;
  const Model<T> m{s.vel, s.nx, k.nz, pml};
  const long long f3 = 3 * s.fstride;
  const T* RA = s.raw;
  const T* RB = s.raw + s.fstride;
  if (MODE == RTM_BWD) {
    if (s.rtm_flags & SEIS_RTM_ISOLATE_BAND) {
      // the source field stays zero throughout the band (the source cell
      // lies in the model), so its new value and U there are +0 exactly
      N[i] = T(0);
      s.aux[(step % NAUX) * s.fstride + i] = T(0);
    } else {
      T sub = T(0);
      const bool on = it - 1 > 0;
      if (on) {
        const T as = k.dt * m.vel(s.sx + pml, s.sz + pml);
        sub = (as * as) * s.wav[it - 1];
      }
      const SourceRemoval<T> sr{fidx(ld, s.sx + pml, s.sz + pml), sub, on};
      const T u = band_blend(k, m, NX, C, P, sr, x, z, RA[i], RawNeighbour<T>{RA, N, ld, NX, NZ, pml + 1});
      const T lap = (sr.prev(P, i, x, z) + u - T(2) * C[i]) * k.idt2;
      const T v = m.vel(x, z);
      s.aux[(step % NAUX) * s.fstride + i] = lap * (T(2) / (v * v * v));
      N[i] = u;
    }
    N[f3 + i] = band_blend(k, m, NX, C + f3, P + f3, NoInject{}, x, z, RB[i],
                           RawNeighbour<T>{RB, N + f3, ld, NX, NZ, pml});
  } else if (MODE == PW_FWD) {
    const T u = band_blend(k, m, NX, C, P, NoInject{}, x, z, RA[i], RawNeighbour<T>{RA, N, ld, NX, NZ, pml});
    const T lap = (u + P[i] - T(2) * C[i]) * k.idt2;
    const T v = m.vel(x, z);
    s.aux[(step % NAUX) * s.fstride + i] = lap * (T(2) / (v * v * v));
    N[i] = u;
    const ReceiverRemoval<T> rr{s.gz, s.rec, s.nx, s.nt, it - s.ndel - 1, pml, k.dt, &m};
    N[f3 + i] = band_blend(k, m, NX, C + f3, P + f3, rr, x, z, RB[i], RawNeighbour<T>{RB, N + f3, ld, NX, NZ, pml});
  } else {
    N[i] = band_blend(k, m, NX, C, P, NoInject{}, x, z, RA[i], RawNeighbour<T>{RA, N, ld, NX, NZ, pml});
    if (MODE == BORN)
      N[f3 + i] = band_blend(k, m, NX, C + f3, P + f3, NoInject{}, x, z, RB[i],
                             RawNeighbour<T>{RB, N + f3, ld, NX, NZ, pml});
  }
}

// Stencil sums of the 9-point operator in the reference's order (fd.cpp:1263-1269).
template <typename T>
__device__ __forceinline__ T d2_z(const Consts<T>& k, const T* r) {
  T u = r[0] * k.c[4];
  u = u + r[1] * k.c[3];
  u = u + r[2] * k.c[2];
  u = u + r[3] * k.c[1];
  u = u + r[4] * k.c[0];
  u = u + r[5] * k.c[1];
  u = u + r[6] * k.c[2];
  u = u + r[7] * k.c[3];
  u = u + r[8] * k.c[4];
  return u;
}

// Interior sweep of one wavefront: z = z0 + lane (cells with z >= zend idle),
// over columns [x0, x1), from x0 up, or from x1-1 down: DIR 0 up, 1 down
// (compile-time), 2 down when rev_rt (run-time).  The current level of each field is held in a 9-deep
// register queue in x order (qa[j] = column x-4+j) whichever way the sweep
// runs, so the stencil sums keep their order; the centre row is exchanged
// through LDS for the z derivative.  Each cell's value is final (no absorbing
// band in here).  Neighbouring strips sweep in opposite directions, so the 8
// columns two strips share are read by both at about the same time (both at
// the start or both at the end of their sweeps) and the second read finds
// them in L2.
template <typename T, int MODE, int DIR>
__device__ __forceinline__ void interior_sweep(const Consts<T>& k, const DevShot<T>& s, const StepCtx<T>& c, int z0,
                                               int zend, int x0, int x1, bool rev_rt, T* lds) {
  const bool rev = DIR == 2 ? rev_rt : DIR == 1;
  constexpr bool TWO = (MODE == BORN || MODE == RTM_BWD || MODE == PW_FWD);
  const int lane = threadIdx.x & 63;
  const int z = z0 + lane;
  const bool act = z < zend;
  const int NZ = k.NZ, pml = k.pml;
  const int zl = min(z, NZ - 1);
  const long long ld = k.ld;
  const auto CA = SEIS_GL(c.C);
  const auto PA = SEIS_GL(c.P);
  const auto NA = SEIS_GL(c.N);
  const auto CB = SEIS_GL(c.C + 3 * s.fstride);
  const auto PB = SEIS_GL(c.P + 3 * s.fstride);
  const auto NB = SEIS_GL(c.N + 3 * s.fstride);
  const auto VEL = SEIS_GL(s.vel);
  const auto GZA = SEIS_GL(s.gz);
  const auto REC = SEIS_GL(s.rec);
  T* la = lds;
  T* lb = lds + LROW;
  const bool halo = lane < 4 || lane >= 60;
  // halo rows z0-4..z0-1 (lanes 0-3) and z0+64..z0+67 (lanes 60-63); every
  // other lane loads the nearest of them (the same cache lines), so that the
  // halo loads are not lane-conditional
  const int hz = lane < 32 ? z0 - 4 + min(lane, 3) : min(z0 + 64 + max(lane - 60, 0), NZ - 1);
  const int hslot = lane < 4 ? lane : lane + 8;
  const int iz = min(z - pml, k.nz - 1);  // clamped: idle lanes past the model load in bounds
  const int sxp = s.sx + pml, szp = s.sz + pml;
  // first column and the column the sweep reaches next; the queue slot that
  // the fetch fills (8 going up, 0 going down) and the fetched column's
  // offset from the swept one
  const int xs0 = rev ? x1 - 1 : x0;
  const int dx = rev ? -1 : 1;
  const long long ahead = (rev ? -4 : 4) * ld;
  T qa[9], qb[9];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int slot = rev ? j + 1 : j;
    qa[slot] = CA[(long long)(xs0 - 4 + slot) * ld + zl];
    if (TWO) qb[slot] = CB[(long long)(xs0 - 4 + slot) * ld + zl];
  }
  // The loads of the next column are issued before column x is computed (one
  // column of software pipelining; the last column re-reads itself).
  // The receiver depth of the column (and, in the reverse pass, its record
  // sample) is fetched the same way.
  constexpr bool GZ = (MODE == MODEL || MODE == BORN || MODE == BORN_S || MODE == RTM_BWD || MODE == PW_BACK ||
                       MODE == PW_FWD || MODE == ADJ_X);
  // record sample read per column: injected (RTM_BWD, PW_BACK) or removed
  // from the previous level (PW_FWD)
  constexpr bool RECIN = (MODE == RTM_BWD || MODE == PW_BACK || MODE == PW_FWD || MODE == ADJ_X);
  const int rec_t = MODE == PW_FWD ? c.it - s.ndel - 1 : c.itp;
  T nca, ncb = T(0), npa, npb = T(0), nv, nha = T(0), nhb = T(0), nrec = T(0), nss = T(0), ndv = T(0);
  SnapWord<T> nsn{T(0), 0};
  int ngz = -1;
  // RTM_BWD: the saved boundary-strip value of the source field at this
  // lane's cell of the next column (slot -1: not a strip cell), fetched with
  // the other loads so that no load is issued and awaited inside a column
  const bool restore = MODE == RTM_BWD && c.itp - 1 >= 0;
  const auto STR = SEIS_GL(s.strips + (restore ? (long long)(c.itp - 1) * 8 * (s.nx + k.nz) : 0));
  T nstr = T(0);
  int nsl = -1;
  // first_strip_slot split into this lane's row part and the column's part
  // (uniform across the wavefront): a top or bottom strip row gives slot
  // ix 4 + zst, any other model row the side strips' column base + its row
  const int izs = zl - pml;
  const bool zst_in = act && izs >= 0 && izs < k.nz;
  const int zst = !zst_in ? -2 : izs < 4 ? izs : izs >= k.nz - 4 ? 4 * s.nx + (k.nz - 1 - izs) : -1;
  const bool acc = MODE == ILLUM && c.it > s.ndel && act;  // accumulate u^2 (interior cells are model cells)
  const auto SSA = SEIS_GL(s.img);
  auto fetch = [&](int xx) {
    xx = min(max(xx, x0), x1 - 1);
    const long long r = (long long)xx * ld;
    nca = CA[r + ahead + zl];
    if (TWO) ncb = CB[r + ahead + zl];
    npa = PA[r + zl];
    if (TWO) npb = PB[r + zl];
    nv = VEL[(long long)(xx - pml) * k.nz + iz];
    nha = CA[r + hz];
    if (TWO) nhb = CB[r + hz];
    if (GZ) ngz = GZA[xx - pml];
    if (RECIN && rec_t >= 0) nrec = REC[(long long)(xx - pml) * s.nt + rec_t];
    if (MODE == ILLUM) nss = SSA[r + zl];  // used where acc
    // (loads at the clamped model row of every lane; used by active lanes)
    if (MODE == BORN_S || MODE == BORN) {
      // perturbation (and, BORN_S, Born-source snapshot) of this model cell
      const long long j = (long long)(xx - pml) * k.nz + iz;
      ndv = SEIS_GL(s.dv)[j];
      if (MODE == BORN_S) nsn = snap_fetch(s, k.nz, c.it, j);
    }
    if (restore) {
      const int ix = xx - pml;
      const int side = ix < 0 || ix >= s.nx ? -1
                       : ix < 4            ? 8 * s.nx + ix * k.nz
                       : ix >= s.nx - 4    ? 8 * s.nx + 4 * k.nz + (s.nx - 1 - ix) * k.nz
                                           : -1;
      nsl = (zst == -2 || ix < 0 || ix >= s.nx) ? -1 : zst >= 0 ? ix * 4 + zst : side >= 0 ? side + izs : -1;
      nstr = STR[max(nsl, 0)];
    }
    if (MODE == ADJ_X) {
      // gradient and Born-source snapshot of this model cell
      const long long j = (long long)(xx - pml) * k.nz + iz;
      nss = SEIS_GL(s.grad)[j];
      nsn = snap_fetch(s, k.nz, c.it, j);
    }
  };
  fetch(xs0);
  // per-step scalars
  const float snap_sc =
      ((MODE == BORN_S || MODE == ADJ_X) && s.snap_kind != SNAP_NATIVE) ? SEIS_GL(s.scale)[c.it] : 0.f;
  constexpr bool PSRC = (MODE != RTM_BWD && MODE != BORN_S && MODE != PW_BACK && MODE != ADJ_X);
  T amp = T(0);
  if (PSRC && c.it - 1 >= 0 && !s.line_src) amp = s.wav[c.it - 1];
  // the plane-wave line source: this lane's row carries it
  const bool lrow = PSRC && s.line_src && c.it - 1 >= 0 && z == szp;
  const long long lstride = s.nt + s.ndel;
  T sub = T(0);
  if (MODE == RTM_BWD && c.it - 1 > 0) {
    const T as = k.dt * c.m.vel(sxp, szp);
    sub = (as * as) * s.wav[c.it - 1];
  }
  // SNAP_MAX: this lane's max |S| over the sweep; one atomic per wavefront
  // after it (non-negative floats order like their bit patterns)
  float smax = 0.f;
  auto column = [&](const int x) {
    const long long row = (long long)x * ld;
    if (rev) {
      qa[0] = nca;
      if (TWO) qb[0] = ncb;
    } else {
      qa[8] = nca;
      if (TWO) qb[8] = ncb;
    }
    if (halo) {
      la[hslot] = nha;
      if (TWO) lb[hslot] = nhb;
    }
    la[lane + 4] = qa[4];
    if (TWO) lb[lane + 4] = qb[4];
    const int ix = x - pml;
    const T v = nv;
    const T pa_cur = npa, pb_cur = npb, rec_cur = nrec, ss_cur = nss, dv_cur = ndv;
    const SnapWord<T> sn_cur = nsn;
    const int gzr = ngz + pml;  // receiver row of this column
    const T str_cur = nstr;
    const int sl_cur = nsl;
    fetch(x + dx);
    const T av = k.dt * v;
    const T vt2 = av * av;
    const T vvzz = vt2 * k.idz2;
    const T vvxx = vt2 * k.idx2;
    const long long i = row + zl;
    wave_sync();
    const T uzz = d2_z(k, la + lane);
    T uxx = qa[0] * k.c[4];
    uxx = uxx + qa[1] * k.c[3];
    uxx = uxx + qa[2] * k.c[2];
    uxx = uxx + qa[3] * k.c[1];
    uxx = uxx + qa[4] * k.c[0];
    uxx = uxx + qa[5] * k.c[1];
    uxx = uxx + qa[6] * k.c[2];
    uxx = uxx + qa[7] * k.c[3];
    uxx = uxx + qa[8] * k.c[4];
    T pa = pa_cur;
    if (MODE == RTM_BWD && sub != T(0) && x == sxp && z == szp) pa = pa - sub;
    T ua = vvzz * uzz + vvxx * uxx + T(2) * qa[4] - pa;
    if (lrow) {
      if (ix >= 0 && ix < s.nx) ua = ua + vt2 * SEIS_GL(s.wav)[(long long)ix * lstride + c.it - 1];
    } else if (PSRC && !s.line_src && x == sxp && z == szp && c.it - 1 >= 0) {
      ua = ua + vt2 * amp;
    }
    if (MODE == SNAP_MAX || MODE == SNAP_STORE) {
      // Born source coefficient of this step (fd.cpp:134-139): vt2 * lap * 2/v^3
      const T lap = (ua + pa - T(2) * qa[4]) * k.idt2;
      const T S = vt2 * lap * (T(2) / (v * v * v));
      if (MODE == SNAP_MAX) {
        smax = fmaxf(smax, act ? fabsf((float)S) : 0.f);
      } else if (act) {
        snap_store(s, k.nz, c.it, (long long)ix * k.nz + iz, S);
      }
    }
    T ub = T(0);
    if (TWO) {
      const T uzzb = d2_z(k, lb + lane);
      T uxxb = qb[0] * k.c[4];
      uxxb = uxxb + qb[1] * k.c[3];
      uxxb = uxxb + qb[2] * k.c[2];
      uxxb = uxxb + qb[3] * k.c[1];
      uxxb = uxxb + qb[4] * k.c[0];
      uxxb = uxxb + qb[5] * k.c[1];
      uxxb = uxxb + qb[6] * k.c[2];
      uxxb = uxxb + qb[7] * k.c[3];
      uxxb = uxxb + qb[8] * k.c[4];
      T pb = pb_cur;
      if (MODE == PW_FWD && rec_t >= 0 && z == gzr) pb = pb - rec_cur * vt2;
      ub = vvzz * uzzb + vvxx * uxxb + T(2) * qb[4] - pb;
    }
    wave_sync();
    if (act) {
      if (MODE == SNAP_MAX || MODE == SNAP_STORE) {
        NA[i] = ua;
      } else if (MODE == BORN_S) {
        if (dv_cur != T(0)) ua = ua + snap_decode_scaled(s, snap_sc, sn_cur) * dv_cur;
        NA[i] = ua;
        if (c.itp >= 0 && z == gzr) REC[(long long)ix * s.nt + c.itp] = ua;
      } else if (MODE == MODEL || MODE == RTM_FWD || MODE == ILLUM) {
        NA[i] = ua;
        if (MODE == ILLUM) {
          if (acc) SSA[i] = ss_cur + ua * ua;
        } else if (MODE == MODEL) {
          if (c.itp >= 0 && z == gzr) REC[(long long)ix * s.nt + c.itp] = ua;
        } else if (c.itp >= 0) {
          store_strip_cell(SEIS_GL(s.strips + (long long)c.itp * 8 * (s.nx + k.nz)), ix, z - pml, s.nx, k.nz, ua);
        }
      } else if (MODE == BORN) {
        const T d = dv_cur;
        if (d != T(0)) {
          const T lap = (ua + pa - T(2) * qa[4]) * k.idt2;
          const T v3 = T(2) / (v * v * v);
          ub = ub + vt2 * lap * v3 * d;
        }
        NA[i] = ua;
        NB[i] = ub;
        if (c.itp >= 0 && z == gzr) REC[(long long)ix * s.nt + c.itp] = ub;
      } else if (MODE == PW_BACK) {
        if (rec_t >= 0 && z == gzr) ua = ua + rec_cur * vt2;
        NA[i] = ua;
        if (c.it >= 1)
          store_strip_cell(SEIS_GL(s.strips + (long long)(c.it - 1) * 8 * (s.nx + k.nz)), ix, z - pml, s.nx, k.nz,
                           ua);
      } else if (MODE == PW_FWD) {
        const T lap = (ua + pa - T(2) * qa[4]) * k.idt2;
        const T v3 = T(2) / (v * v * v);
        SEIS_GL(c.aux)[i] = lap * v3;
        NA[i] = ua;
        NB[i] = ub;
      } else if (MODE == ADJ_X) {
        if (rec_t >= 0 && z == gzr) ua = ua + rec_cur * vt2;
        NA[i] = ua;
        SEIS_GL(s.grad)[(long long)ix * k.nz + iz] = ss_cur + ua * (snap_decode_scaled(s, snap_sc, sn_cur) / vt2);
      } else {  // RTM_BWD: source field ua, receiver field ub
        if (z == gzr) ub = ub + rec_cur * vt2;
        const T lap = (pa + ua - T(2) * qa[4]) * k.idt2;
        const T v3 = T(2) / (v * v * v);
        SEIS_GL(c.aux)[i] = lap * v3;
        NB[i] = ub;
        NA[i] = (restore && sl_cur >= 0) ? str_cur : ua;
        // the band's one-way update reads the stencil value of this layer,
        // not the restored strip
        if (edge_dist(x, z, s.NX, NZ) == pml) SEIS_GL(s.raw)[i] = ua;
      }
    }
    if (rev) {
#pragma unroll
      for (int j = 8; j > 0; j--) {
        qa[j] = qa[j - 1];
        if (TWO) qb[j] = qb[j - 1];
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        qa[j] = qa[j + 1];
        if (TWO) qb[j] = qb[j + 1];
      }
    }
  };
  for (int x = xs0, n = x1 - x0; n > 0; x += dx, n--) column(x);
  if (MODE == SNAP_MAX) {
    for (int o = 32; o >= 1; o >>= 1) smax = fmaxf(smax, __shfl_xor(smax, o, 64));
    if (lane == 0) atomicMax(reinterpret_cast<unsigned int*>(s.scale) + c.it, __float_as_uint(smax));
  }
}

// One work-group of a step launch: block bx of shot `shot` (sweep strips
// below L.nfwg, ring cells above).
template <typename T, int MODE>
__device__ __forceinline__ void step_block(const Consts<T>& k, const DevShot<T>* __restrict__ shots, const Layout& L,
                                           int step, int bx, int shot, T (*lds)[2 * LROW]) {
  const DevShot<T>& s = shots[shot];
  StepCtx<T> c;
  c.it = step_it<T>(MODE, s, step);
  if (c.it < 0) return;
  c.itp = c.it - s.ndel;
  const int r0 = step % 3, r1 = (step + 1) % 3, r2 = (step + 2) % 3;
  c.P = s.fld + r0 * s.fstride;
  c.C = s.fld + r1 * s.fstride;
  c.N = s.fld + r2 * s.fstride;
  c.m = Model<T>{s.vel, s.nx, k.nz, k.pml};
  c.src = fidx(k.ld, s.sx + k.pml, s.sz + k.pml);
  c.aux = s.aux ? s.aux + (step % NAUX) * s.fstride : nullptr;
  const int NX = s.NX, NZ = k.NZ;
  const Split q = shot_split(L.m, NX, NZ);  // this shot's interior (widths differ within a batch)
  if (bx < L.nfwg) {
    // the wavefront's index, uniform across its lanes, keeps the strip
    // bounds, the sweep's loop and its column addresses in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = xcd_tile(bx, L.nfwg) * WAVES + wave;
    if (item >= L.nzc * shot_strips(q, L.bx)) return;  // whole wavefront leaves; no workgroup barrier follows
    const int zc = item % L.nzc, xs = item / L.nzc;
    const int z0 = q.za + zc * 64;
    const int x0 = q.xa + xs * L.bx;
    const int x1 = min(x0 + L.bx, q.xb);
    // the direction is a compile-time choice: a run-time one turns every
    // shift of the register queues into selects
    // The snapshot modes keep the direction a run-time choice: with one loop
    // per direction their sweeps measured slower (Born from snapshots on
    // the 128-shot 376 x 1151 plan: 186.5 against 172.8 us per step), the
    // RTM sweeps faster.
    constexpr bool RT_DIR = MODE == BORN_S || MODE == ADJ_X || MODE == SNAP_MAX || MODE == SNAP_STORE;
    if (RT_DIR)
      interior_sweep<T, MODE, 2>(k, s, c, z0, q.zb, x0, x1, (xs & 1) != 0, lds[wave]);
    else if (xs & 1)
      interior_sweep<T, MODE, 1>(k, s, c, z0, q.zb, x0, x1, true, lds[wave]);
    else
      interior_sweep<T, MODE, 0>(k, s, c, z0, q.zb, x0, x1, false, lds[wave]);
    return;
  }
  const long long j = (long long)(bx - L.nfwg) * WG + threadIdx.x;
  if (j >= q.n) return;
  int x, z;
  ring_cell(q, NX, NZ, j, &x, &z);
  if (edge_dist(x, z, NX, NZ) < k.pml)
    band_raw<T, MODE>(k, s, c, x, z);
  else
    cell_update<T, MODE>(k, s, c, x, z);
}

template <typename T, int MODE>
__global__ __launch_bounds__(WG) void step_kernel(Consts<T> k, const DevShot<T>* __restrict__ shots, Layout L,
                                                 int step) {
  __shared__ T lds[WAVES][2 * LROW];
  step_block<T, MODE>(k, shots, L, step, blockIdx.x, blockIdx.y, lds);
}

// The imaging condition of one reverse step for one interior column x
// (fd.cpp:681-684): img += U R, then img += 1/2 (U R - H[U] H[R]) (FaQi,
// fd.cpp:1272-1298), where U = lap 2/v^3 of the source field, R the new level
// of the receiver field and H the 61-tap Hamming-tapered Hilbert FIR
// (hilbert.h:42-69) applied over the full padded column.  The reference adds
// U R over the whole padded grid but returns only the interior, which is all
// that is accumulated here; the two additions stay separate roundings.
// Even taps are zero and are skipped (adding 0 * x leaves the sum unchanged).
constexpr int HILBERT_HALF = 30;  // FaQi's filter half length (fd.cpp:1276)
struct HilbertTaps {
  float h[2 * HILBERT_HALF + 1];
};

// The imaging pass works on tiles of FQ_TILE interior rows of one column.
// Output row z uses the inputs z-29, z-27, ..., z+29, all of the other
// parity, so each tile's U and R (rows z0-29 .. z0+FQ_TILE+28) are split by
// parity into LDS arrays: output n of parity p (z = z0 + p + 2n) reads
// A_p[n .. n+29] with A_p[k] = in[z0 - 29 + p + 2k].  A thread forms 4
// outputs of one parity from A_p[4t .. 4t+32], read as 9 aligned float4; each
// input is added into every accumulator whose window holds it, in increasing
// order of the input row, so every sum is the reference's sequence of
// roundings (hilbert.h:42-54; the terms with a zero tap or an input outside
// the column are left out: each would add a zero product to a partial sum
// that is never -0, which leaves it unchanged).
constexpr int FQ_WG = 256;
constexpr int FQ_TILE = 1024;             // output rows per work-group (512 of each parity)
constexpr int FQ_PAR = FQ_TILE / 2 + 32;  // inputs per parity array (multiple of 4)

// Reverse steps `step` and, when `pair`, step + 1 applied to one tile of one
// interior column (the image is read and written once for both).  U of step
// st is in aux field st % NAUX; R of step st is in field slot 3 + (st + 2) % 3.
// `plane`: step `step` of adjointPlane's forward pass instead (it = step+2,
// imaged when it > ndel, FaQi term only; never paired).
// Grid: (row tiles, interior columns of the widest shot, shots).
// LDS of one imaging work-group: the parity-split inputs A[step][U, R][parity][k]
// and H[U] H[R] per step and output.
struct ImageLds {
  float A[2][2][2][FQ_PAR];
  float HP[2][FQ_TILE];
};

// One imaging work-group: row tile `tile` of interior column `col` of shot `shot`.
template <int HALF>
__device__ __forceinline__ void image_block(const Consts<float>& k, const DevShot<float>* __restrict__ shots, int step,
                                            const HilbertTaps& taps, int pair, int plane, int tile, int col, int shot,
                                            ImageLds& sm) {
  static_assert(HALF == 30, "the tile arithmetic assumes FaQi's 61-tap filter");
  float (&A)[2][2][2][FQ_PAR] = sm.A;
  float (&HP)[2][FQ_TILE] = sm.HP;
  const DevShot<float>& s = shots[shot];
  const int x = k.pml + col;
  if (x >= s.NX - k.pml) return;
  const int it = plane ? step + 2 : s.nt + s.ndel - 1 - step;
  if (plane ? (it >= s.nt + s.ndel || it - s.ndel <= 0) : it < s.ndel) return;
  const bool two = !plane && pair && it - 1 >= s.ndel;
  const int NZ = k.NZ;
  const int zhi = NZ - k.pml;
  const int z0 = k.pml + tile * FQ_TILE;
  if (z0 >= zhi) return;
  const long long base = fidx(k.ld, x, 0);
  const float* U0 = s.aux + (step % NAUX) * s.fstride + base;
  const float* R0 = s.fld + (3 + (step + 2) % 3) * s.fstride + base;
  const float* U1 = s.aux + ((step + 1) % NAUX) * s.fstride + base;
  const float* R1 = s.fld + (3 + (step + 3) % 3) * s.fstride + base;
  const int nst = two ? 2 : 1;
  // rows z0-29+q, q < 2 FQ_PAR, into A[st][f][q & 1][q >> 1] (coalesced
  // reads).  Every thread issues all of its loads before it stores any: each
  // at a row clamped into the column (zero outside it), the second step's
  // from the first step's fields when there is no second step (unused then).
  constexpr int NLOAD = 2 * 2 * 2 * FQ_PAR / FQ_WG;
  static_assert(NLOAD * FQ_WG == 2 * 2 * 2 * FQ_PAR, "the tile's inputs split evenly over the work-group");
  const float* const U1s = two ? U1 : U0;
  const float* const R1s = two ? R1 : R0;
  float vals[NLOAD];
#pragma unroll
  for (int l = 0; l < NLOAD; l++) {
    const int w = threadIdx.x + l * FQ_WG;
    const int q = w % (2 * FQ_PAR), sf = w / (2 * FQ_PAR);
    const float* src = sf == 0 ? U0 : sf == 1 ? R0 : sf == 2 ? U1s : R1s;
    vals[l] = SEIS_GL(src)[min(max(z0 - 29 + q, 0), NZ - 1)];
  }
#pragma unroll
  for (int l = 0; l < NLOAD; l++) {
    const int w = threadIdx.x + l * FQ_WG;
    const int q = w % (2 * FQ_PAR), sf = w / (2 * FQ_PAR);
    const int r = z0 - 29 + q;
    A[sf >> 1][sf & 1][q & 1][q >> 1] = (r >= 0 && r < NZ) ? vals[l] : 0.f;
  }
  __syncthreads();
  const int p = threadIdx.x >> 7;   // parity of this thread's outputs
  const int t = threadIdx.x & 127;  // outputs n = 4t .. 4t + 3
  for (int st = 0; st < nst; st++) {
    const float4* u4 = reinterpret_cast<const float4*>(&A[st][0][p][4 * t]);
    const float4* r4 = reinterpret_cast<const float4*>(&A[st][1][p][4 * t]);
    float uh[4] = {0.f, 0.f, 0.f, 0.f}, rh[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int v = 0; v < 9; v++) {
      const float4 qu = u4[v], qr = r4[v];
      const float xu[4] = {qu.x, qu.y, qu.z, qu.w};
      const float xr[4] = {qr.x, qr.y, qr.z, qr.w};
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int w = 4 * v + e;  // input A_p[4t + w]
#pragma unroll
        for (int m = 0; m < 4; m++) {
          const int u = w - m;  // its index j in output (4t + m)'s window: tap h[59 - 2j]
          if (u >= 0 && u < HALF) {
            const float h = taps.h[2 * HALF - 1 - 2 * u];
            uh[m] = uh[m] + xu[e] * h;
            rh[m] = rh[m] + xr[e] * h;
          }
        }
      }
    }
#pragma unroll
    for (int m = 0; m < 4; m++) HP[st][p + 2 * (4 * t + m)] = uh[m] * rh[m];
  }
  __syncthreads();
  // img += U R (unless `plane`), then img += 1/2 (U R - H[U] H[R]), step by step
  const float scal = 2.0f / 4.0f;
  const auto IMG = SEIS_GL(s.img + base);
  // the tile's image rows are all loaded first (clamped into the interior)
  constexpr int NOUT = FQ_TILE / FQ_WG;
  float imgv[NOUT];
#pragma unroll
  for (int i = 0; i < NOUT; i++) imgv[i] = IMG[min(z0 + (int)threadIdx.x + i * FQ_WG, zhi - 1)];
#pragma unroll
  for (int i = 0; i < NOUT; i++) {
    const int q = threadIdx.x + i * FQ_WG;
    const int z = z0 + q;
    if (z >= zhi) break;
    const int pp = (q + 29) & 1, kk = (q + 29) >> 1;  // row z in the parity arrays
    float img = imgv[i];
    const float ur0 = A[0][0][pp][kk] * A[0][1][pp][kk];
    if (!plane) img += ur0;
    img += scal * (ur0 - HP[0][q]);
    if (two) {
      const float ur1 = A[1][0][pp][kk] * A[1][1][pp][kk];
      img += ur1;
      img += scal * (ur1 - HP[1][q]);
    }
    IMG[z] = img;
  }
}

// Grid: (row tiles, interior columns c0 .. c0 + gridDim.y - 1, shots).
template <int HALF>
__global__ __launch_bounds__(FQ_WG) void image_kernel(Consts<float> k, const DevShot<float>* __restrict__ shots,
                                                      int step, HilbertTaps taps, int pair, int plane, int c0) {
  __shared__ __align__(16) ImageLds sm;
  image_block<HALF>(k, shots, step, taps, pair, plane, blockIdx.x, c0 + blockIdx.y, blockIdx.z, sm);
}

// Reverse step `step` of every shot (the step kernel's work-groups) fused
// with the imaging pass of reverse steps pfirst and, when `pair`, pfirst + 1
// over interior columns c0 .. c0 + nimg / ntile - 1 (the last nimg
// work-groups: tile t % ntile of column c0 + t / ntile), where pfirst + pair
// = step - 1.  The two touch disjoint buffers: the pass reads U of those
// steps (NAUX = 4, so step writes another U buffer) and their receiver levels
// (step reads them as its previous and current level and writes the third
// slot), and writes only the image.  The imaging pass is bound by its
// floating-point work and the step by memory, so in one launch each fills the
// other's idle time:
//  * the step's work-groups (sweep, then ring cells) come first, so that the
//    sweep keeps its XCD mapping, then the imaging pass's;
//  * the kernel is held to at most 80 VGPRs (6 waves per SIMD, as many as its
//    LDS allows) so that the step keeps the waves it needs to hide memory
//    latency: at 96 VGPRs (5 waves), where the compiler put it unbounded, the
//    launch took 751 us, at 6 waves 693 us on the 26-shot survey launch at
//    strip 16;
//  * the step's waves issue ahead of the imaging waves that share their SIMD
//    (686 us).
// (Also measured there: the imaging work-groups spread evenly through the
// sweep's on every XCD, each part keeping its own XCD mapping, so that every
// XCD images the columns it sweeps: 790-816 us, the step then has too few
// waves on each CU.)
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(6, 8))) void rtm_reverse_image_kernel(
    Consts<float> k, const DevShot<float>* __restrict__ shots, Layout L, int step, HilbertTaps taps, int ntile,
    int nimg, int pfirst, int pair, int c0) {
  static_assert(FQ_WG == WG, "one work-group size for both parts");
  static_assert(sizeof(ImageLds) >= sizeof(float) * WAVES * 2 * LROW, "the step's LDS rows fit");
  __shared__ __align__(16) unsigned char sm[sizeof(ImageLds)];
  const int nstep = (int)gridDim.x - nimg;  // the step's work-groups
  const int b = blockIdx.x;
  if (b >= nstep) {
    const int t = b - nstep;
    image_block<HILBERT_HALF>(k, shots, pfirst, taps, pair, 0, t % ntile, c0 + t / ntile, blockIdx.y,
                              *reinterpret_cast<ImageLds*>(sm));
    return;
  }
  __builtin_amdgcn_s_setprio(3);
  step_block<float, RTM_BWD>(k, shots, L, step, b, blockIdx.y, reinterpret_cast<float (*)[2 * LROW]>(sm));
}

// Row tiles of the imaging pass for padded depth NZ.
inline int image_tiles(int NZ, int pml) { return (NZ - 2 * pml + FQ_TILE - 1) / FQ_TILE; }

// Writes the strips of time sample `itp` into a field (fd.cpp:1204-1223).
template <typename T>
__global__ void restore_strips_kernel(Consts<T> k, const DevShot<T>* __restrict__ shots, int slot, int from_end) {
  const DevShot<T>& s = shots[blockIdx.y];
  const int nx = s.nx, nz = k.nz, pml = k.pml;
  const int n = 8 * (nx + nz);
  const int itp = from_end ? s.nt - 1 : 0;
  T* F = s.fld + slot * s.fstride;
  const T* st = s.strips + (long long)itp * n;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    int x, z;
    if (j < 4 * nx) { x = j / 4 + pml; z = pml + j % 4; }
    else if (j < 8 * nx) { int q = j - 4 * nx; x = q / 4 + pml; z = nz + pml - 1 - q % 4; }
    else if (j < 8 * nx + 4 * nz) { int q = j - 8 * nx; x = pml + q / nz; z = q % nz + pml; }
    else { int q = j - 8 * nx - 4 * nz; x = nx + pml - 1 - q / nz; z = q % nz + pml; }
    F[fidx(k.ld, x, z)] = st[j];
  }
}

// adjointPlane's forward pass: the strips saved at time it = step + 2 go
// into the receiver field's new level (after the imaging of the step,
// fd.cpp:878).
template <typename T>
__global__ void restore_plane_strips_kernel(Consts<T> k, const DevShot<T>* __restrict__ shots, int step) {
  const DevShot<T>& s = shots[blockIdx.y];
  const int it = step + 2;
  if (it >= s.nt + s.ndel) return;
  const int nx = s.nx, nz = k.nz, pml = k.pml;
  const int n = 8 * (nx + nz);
  T* F = s.fld + (3 + (step + 2) % 3) * s.fstride;
  const T* st = s.strips + (long long)it * n;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    int x, z;
    if (j < 4 * nx) { x = j / 4 + pml; z = pml + j % 4; }
    else if (j < 8 * nx) { int q = j - 4 * nx; x = q / 4 + pml; z = nz + pml - 1 - q % 4; }
    else if (j < 8 * nx + 4 * nz) { int q = j - 8 * nx; x = pml + q / nz; z = q % nz + pml; }
    else { int q = j - 8 * nx - 4 * nz; x = nx + pml - 1 - q / nz; z = q % nz + pml; }
    F[fidx(k.ld, x, z)] = st[j];
  }
}

// Interior crop of a padded image: out[ix][iz] = img[ix+pml][iz+pml] into
// the shot's own destination.
template <typename T>
__global__ void crop_kernel(Consts<T> k, const DevShot<T>* __restrict__ shots) {
  const DevShot<T>& s = shots[blockIdx.y];
  const long long n = (long long)s.nx * k.nz;
  T* o = s.out;
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(j / k.nz), iz = (int)(j % k.nz);
    o[j] = s.img[fidx(k.ld, ix + k.pml, iz + k.pml)];
  }
}

// ---------------------------------------------------------------------------
// Host side.

template <typename T>
void launch_step(int mode, const Consts<T>& k, const DevShot<T>* dev, const Layout& L, const Layout& B, int nshots,
                 int step, hipStream_t stream) {
  dim3 grid(L.nfwg + L.nbwg, nshots), block(WG);
  dim3 bgrid((unsigned)B.nbwg, nshots);
  switch (mode) {
#define SEIS_CASE(M)                                                                            \
  case M:                                                                                       \
    hipLaunchKernelGGL((step_kernel<T, M>), grid, block, 0, stream, k, dev, L, step);           \
    if (B.nbwg > 0) hipLaunchKernelGGL((band_kernel<T, M>), bgrid, block, 0, stream, k, dev, step); \
    break;
    SEIS_CASE(MODEL)
    SEIS_CASE(BORN)
    SEIS_CASE(RTM_FWD)
    SEIS_CASE(RTM_BWD)
    SEIS_CASE(SNAP_MAX)
    SEIS_CASE(SNAP_STORE)
    SEIS_CASE(BORN_S)
    SEIS_CASE(ILLUM)
    SEIS_CASE(PW_BACK)
    SEIS_CASE(PW_FWD)
    SEIS_CASE(ADJ_X)
#undef SEIS_CASE
  }
}

template void launch_step<float>(int, const Consts<float>&, const DevShot<float>*, const Layout&, const Layout&,
                                 int, int, hipStream_t);
template void launch_step<double>(int, const Consts<double>&, const DevShot<double>*, const Layout&, const Layout&,
                                  int, int, hipStream_t);

// Zeroes the first `nfields` field slots of every shot (a fill kernel, so
// that a caller may capture the time loop in a graph).
template <typename T>
__global__ void zero_fields_kernel(const DevShot<T>* __restrict__ shots, long long origin, int nfields) {
  const DevShot<T>& s = shots[blockIdx.y];
  T* base = s.fld - origin;
  const long long n = s.fstride * nfields;
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x)
    base[j] = T(0);
}

#define HIPCHECK(x)                                                            \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));               \
      return SEIS_ERR_HIP;                                                     \
    }                                                                          \
  } while (0)

int check_grid(const GridSpec& g) {
  if (g.nx <= 0 || g.nz <= 0 || g.pml < 0) return set_error("grid: nx, nz must be > 0 and pml >= 0"), SEIS_ERR_ARG;
  if (g.nz + 2 * g.pml < 8 || g.nx + 2 * g.pml < 8)
    return set_error("grid: padded grid must be at least 8 x 8 cells"), SEIS_ERR_ARG;
  if (!(g.dt > 0) || !(g.dx > 0) || !(g.dz > 0)) return set_error("grid: dt, dx, dz must be > 0"), SEIS_ERR_ARG;
  if ((long long)(g.nx + 2 * g.pml) * (g.nz + 2 * g.pml) >= (1LL << 31))
    return set_error("grid: the padded grid must have fewer than 2^31 cells"), SEIS_ERR_ARG;
  return SEIS_OK;
}

// ---------------------------------------------------------------------------
// Batch memory: every shot of a launch needs its field slots (and, for RTM,
// its boundary strips) for the whole run, so a batch is cut into consecutive
// chunks whose working set fits a budget; the chunks run one after another.
// Results are per shot and do not depend on the cut.

namespace {
std::mutex g_limit_mu;
size_t g_limit_bytes = 0;  // seis_set_batch_limits: 0 = free memory less a reserve
int g_limit_shots = 0;     // 0 = no cap on shots per launch
}  // namespace

int shot_width(const GridSpec& g, const ShotSpec& q) { return q.nx > 0 ? q.nx : g.nx; }

// Field slots per shot of a run mode (see run_chunk).
int fields_per_shot(int mode) {
  const bool ilm = mode == RUN_ILLUM, rtm = mode == RUN_RTM;
  const int nslots = (mode == RUN_MODEL || ilm) ? 3 : 6;
  const int nraw = (mode == RUN_MODEL || ilm) ? 1 : 2;
  if (mode == RUN_RTM_PLANE) return 6 + 2 + 1 + NAUX;
  return nslots + nraw + (rtm ? 1 + NAUX : ilm ? 1 : 0);
}

// Device bytes one shot holds during a run of `mode` (fields, strips, and
// its descriptor), each part a multiple of 256 bytes.
size_t shot_bytes(int mode, const GridSpec& g, const ShotSpec& q, size_t elem, size_t desc) {
  const int NX = shot_width(g, q) + 2 * g.pml, NZ = g.nz + 2 * g.pml;
  const long long ld = field_ld(NZ, g.pml);
  size_t b = (size_t)field_elems(NX, ld) * fields_per_shot(mode) * elem;
  long long strip_t = mode == RUN_RTM ? q.nt : mode == RUN_RTM_PLANE ? q.nt + q.ndel : 0;
  size_t st = (size_t)strip_t * 8 * (shot_width(g, q) + g.nz) * elem;
  return b + (st + 255) / 256 * 256 + desc;
}

int plan_chunks(const std::vector<size_t>& bytes, size_t budget, int max_shots, std::vector<int>& starts) {
  starts.clear();
  const size_t overhead = 256;  // alignment of the descriptor array
  size_t used = 0;
  int count = 0;
  for (size_t i = 0; i < bytes.size(); i++) {
    if (bytes[i] + overhead > budget) {
      set_error("batch: one shot needs " + std::to_string(bytes[i] + overhead) + " bytes of device memory, more "
                "than the " + std::to_string(budget) + " bytes available to a batch");
      return SEIS_ERR_ARG;
    }
    if (starts.empty() || used + bytes[i] > budget || (max_shots > 0 && count >= max_shots)) {
      starts.push_back((int)i);
      used = overhead;
      count = 0;
    }
    used += bytes[i];
    count++;
  }
  starts.push_back((int)bytes.size());
  return SEIS_OK;
}

// Working-set budget of a batch call: the caller's cap, or the free device
// memory less a reserve of max(1 GiB, 1/32 of it).
int batch_budget(size_t* out) {
  size_t fr = 0, tot = 0;
  HIPCHECK(hipMemGetInfo(&fr, &tot));
  const size_t reserve = std::max<size_t>(1ull << 30, fr / 32);
  size_t b = fr > reserve ? fr - reserve : 0;
  std::lock_guard<std::mutex> lk(g_limit_mu);
  if (g_limit_bytes) b = std::min(b, g_limit_bytes);
  *out = b;
  return SEIS_OK;
}

int batch_max_shots() {
  std::lock_guard<std::mutex> lk(g_limit_mu);
  return g_limit_shots;
}

// Per-shot checks shared by the batch runners.
int check_shots(int mode, const GridSpec& g, int nshots, const ShotSpec* shots, int line_src) {
  for (int s = 0; s < nshots; s++) {
    const ShotSpec& q = shots[s];
    if (q.nt <= 0 || q.ndel < 0) return set_error("shot: nt must be > 0 and ndel >= 0"), SEIS_ERR_ARG;
    if (q.nx < 0) return set_error("shot: nx must be >= 0 (0: the grid's nx)"), SEIS_ERR_ARG;
    const int nx = shot_width(g, q);
    GridSpec gs = g;
    gs.nx = nx;
    int rc = check_grid(gs);
    if (rc) return rc;
    if (line_src && nx != g.nx) return set_error("plane-wave shots span the grid's nx"), SEIS_ERR_ARG;
    if ((!line_src && (q.sx < 0 || q.sx >= nx)) || q.sz < 0 || q.sz >= g.nz)
      return set_error("shot: source cell outside the model"), SEIS_ERR_ARG;
    if ((mode == RUN_RTM || mode == RUN_RTM_PLANE) && (nx < 4 || g.nz < 4))
      return set_error("rtm: the 4-layer boundary strips need nx >= 4 and nz >= 4"), SEIS_ERR_ARG;
  }
  return SEIS_OK;
}

// Device memory of one chunk: allocated here, or a region of a caller's
// workspace.
struct ChunkMem {
  void* p = nullptr;
  bool owned = false;
  ~ChunkMem() {
    if (owned && p) (void)hipFree(p);
  }
};

// Launch with per-kernel timing, or plainly (the workspace path allocates
// nothing, including profiler events).
template <class F>
void launch(bool quiet, hipStream_t s, const char* name, double bytes, double cells, F&& fn) {
  if (quiet)
    fn();
  else
    timed_launch(s, name, bytes, cells, fn);
}

// One chunk of a batch: shots [0, n) of `shots`, their image destinations in
// `outs`, everything in `mem` (chunk_bytes).
template <typename T>
int run_chunk(int mode, const GridSpec& g, int n, const ShotSpec* shots, T* const* outs, const T* vel, const T* dv,
              const T* wav, const int* gz, T* rec, void* mem, hipStream_t stream, int rtm_flags, int line_src,
              bool quiet) {
  const int NZ = g.nz + 2 * g.pml;
  const long long ld = field_ld(NZ, g.pml);
  const bool ilm = (mode == RUN_ILLUM);
  const int nslots = (mode == RUN_MODEL || ilm) ? 3 : 6;
  const bool rtm = (mode == RUN_RTM);
  const int nraw = (mode == RUN_MODEL || ilm) ? 1 : 2;
  const int nf = fields_per_shot(mode);
  int NTmax = 0, maxbwd = 0, nxmax = 0;
  std::vector<int> NXs(n);
  for (int s = 0; s < n; s++) {
    NTmax = std::max(NTmax, shots[s].nt + shots[s].ndel);
    maxbwd = std::max(maxbwd, shots[s].nt);
    NXs[s] = shot_width(g, shots[s]) + 2 * g.pml;
    nxmax = std::max(nxmax, shot_width(g, shots[s]));
  }
  // descriptors first, then per shot: nf fields (nslots levels, the band
  // stencil values of one or two fields, then the image and NAUX U buffers
  // for RTM or the illumination for ILLUM), then the strips
  std::vector<DevShot<T>> host(n);
  DevShot<T>* dev = reinterpret_cast<DevShot<T>*>(mem);
  char* p = reinterpret_cast<char*>(mem) + (sizeof(DevShot<T>) * n + 255) / 256 * 256;
  const long long origin = field_origin(ld, g.pml);
  for (int s = 0; s < n; s++) {
    const ShotSpec& q = shots[s];
    const int nx = shot_width(g, q);
    const long long fe = field_elems(NXs[s], ld);
    T* base = reinterpret_cast<T*>(p);
    DevShot<T>& d = host[s];
    d = DevShot<T>{};
    d.vel = vel + q.model;
    d.dv = dv ? dv + q.model : nullptr;
    d.wav = wav + q.wav;
    d.gz = gz ? gz + q.gz : nullptr;  // illumination has no receivers
    d.rec = rec ? rec + q.rec : nullptr;
    d.fld = base + origin;
    d.fstride = fe;
    d.raw = base + (size_t)fe * nslots + origin;
    d.img = (rtm || ilm) ? base + (size_t)fe * (nslots + nraw) + origin : nullptr;
    d.aux = rtm ? base + (size_t)fe * (nslots + nraw + 1) + origin : nullptr;
    d.out = outs ? outs[s] : nullptr;
    const size_t st = rtm ? (size_t)q.nt * 8 * (nx + g.nz) * sizeof(T) : 0;
    d.strips = rtm ? reinterpret_cast<T*>(p + (size_t)fe * nf * sizeof(T)) : nullptr;
    p += (size_t)fe * nf * sizeof(T) + (st + 255) / 256 * 256;
    d.nx = nx;
    d.NX = NXs[s];
    d.nt = q.nt;
    d.ndel = q.ndel;
    d.sx = q.sx;
    d.sz = q.sz;
    d.rtm_flags = rtm_flags;
    d.line_src = line_src;
  }
  HIPCHECK(hipMemcpyAsync(dev, host.data(), sizeof(DevShot<T>) * n, hipMemcpyHostToDevice, stream));
  // fields and images start at zero; every strip sample is written by the
  // forward pass before the reverse pass reads it
  hipLaunchKernelGGL((zero_fields_kernel<T>), dim3(512, n), dim3(256), 0, stream, dev, origin, nf);
  const Consts<T> k = make_consts<T>(g, ld);
  Layout L = make_layout(NXs.data(), n, NZ, std::max(g.pml, 4));
  const Layout B = band_layout(NXs.data(), n, NZ, g.pml);
  const int NXmax = L.NXmax;
  {
    // strip length: timed on the zeroed fields with this pass's step kernel
    // at step 0 (no injection yet, so the fields stay zero; records and strip
    // samples written here are rewritten by the real step 0).  RTM times a
    // forward and a reverse step together (the reverse, with two fields, is
    // most of the time); the reverse trial writes only the new level of each
    // field (strip samples restored from the not yet written strip store), U
    // and the band's stencil values, all cleared before the reverse pass or
    // written in full by the forward pass before they are read.
    struct Trial {
      int mode, mode2;
      const Consts<T>* k;
      const DevShot<T>* dev;
      Layout L, B;
      int nshots;
      hipStream_t s;
    };
    // (an ILLUM trial at step 0 accumulates nothing: step 0 never exceeds ndel)
    const int proxy = mode == RUN_MODEL ? MODEL : mode == RUN_BORN ? BORN : ilm ? ILLUM : RTM_FWD;
    Trial tr{proxy, mode == RUN_RTM ? RTM_BWD : -1, &k, dev, L, B, n, stream};
    auto fn = [](void* c, int bx) {
      const Trial* t = static_cast<const Trial*>(c);
      launch_step<T>(t->mode, *t->k, t->dev, with_strip(t->L, bx), t->B, t->nshots, 0, t->s);
      if (t->mode2 >= 0) launch_step<T>(t->mode2, *t->k, t->dev, with_strip(t->L, bx), t->B, t->nshots, 0, t->s);
    };
    const char* kind = mode == RUN_MODEL ? "forward" : mode == RUN_BORN ? "born" : ilm ? "illum" : "rtm";
    L = with_strip(L, quiet ? strip_lookup(kind, NXmax, NZ, n) : tuned_strip(kind, NXmax, NZ, n, stream, fn, &tr));
  }
  const double E = sizeof(T);
  // cells of the shots still running at forward step st / reverse step st
  auto cells_fwd = [&](int st) {
    double c = 0;
    for (int s = 0; s < n; s++)
      if (st < shots[s].nt + shots[s].ndel) c += (double)NXs[s] * NZ;
    return c;
  };
  auto cells_bwd = [&](int st, double* strip_bytes, double* faqi, double* model_cells) {
    double c = 0, sb = 0, fq = 0, mc = 0;
    for (int s = 0; s < n; s++)
      if (st < shots[s].nt) {
        const int nx = NXs[s] - 2 * g.pml;
        c += (double)NXs[s] * NZ;
        sb += 8.0 * (nx + g.nz) * E;
        fq += (double)nx * NZ;
        mc += (double)nx * g.nz;
      }
    if (strip_bytes) *strip_bytes = sb;
    if (faqi) *faqi = fq;
    if (model_cells) *model_cells = mc;
    return c;
  };
  // Algorithmic HBM bytes per cell: MODEL reads vel, C, P and writes N (16 B
  // in FP32); BORN adds the scattered C, P, N and dv (32 B for 2 updates);
  // RTM reverse reads vel, C, P, Cr, Pr and writes N, Nr and lap*2/v^3
  // (32 B for 2 updates); the image pass reads U and R over the interior
  // columns and read-modify-writes the interior image.
  if (mode == RUN_MODEL) {
    for (int st = 0; st < NTmax; st++) {
      const double a = cells_fwd(st);
      launch(quiet, stream, "step_model", 4 * E * a, a, [&] { launch_step<T>(MODEL, k, dev, L, B, n, st, stream); });
    }
  } else if (ilm) {
    // MODEL's bytes plus the read-modify-write of u^2 at the model cells
    for (int st = 0; st < NTmax; st++) {
      double a = 0, mc = 0;
      for (int s = 0; s < n; s++)
        if (st < shots[s].nt + shots[s].ndel) {
          a += (double)NXs[s] * NZ;
          mc += (double)(NXs[s] - 2 * g.pml) * g.nz;
        }
      launch(quiet, stream, "step_illum", 4 * E * a + 2 * E * mc, a,
             [&] { launch_step<T>(ILLUM, k, dev, L, B, n, st, stream); });
    }
    hipLaunchKernelGGL((crop_kernel<T>), dim3(64, n), dim3(256), 0, stream, k, dev);
  } else if (mode == RUN_BORN) {
    for (int st = 0; st < NTmax; st++) {
      const double a = cells_fwd(st);
      launch(quiet, stream, "step_born", 8 * E * a, 2 * a, [&] { launch_step<T>(BORN, k, dev, L, B, n, st, stream); });
    }
  } else {
    for (int st = 0; st < NTmax; st++) {
      double a = 0, sb = 0;
      for (int s = 0; s < n; s++)
        if (st < shots[s].nt + shots[s].ndel) {
          a += (double)NXs[s] * NZ;
          sb += 8.0 * (NXs[s] - 2 * g.pml + g.nz) * E;
        }
      launch(quiet, stream, "step_rtm_forward", 4 * E * a + sb, a,
             [&] { launch_step<T>(RTM_FWD, k, dev, L, B, n, st, stream); });
    }
    // reset the six wavefields, then restore the last strips into the current level
    hipLaunchKernelGGL((zero_fields_kernel<T>), dim3(512, n), dim3(256), 0, stream, dev, origin, nslots);
    hipLaunchKernelGGL((restore_strips_kernel<T>), dim3(16, n), dim3(256), 0, stream, k, dev, 1, 1);
    constexpr int half = HILBERT_HALF;
    HilbertTaps taps{};
    hilbert_taps_host<float>(half, taps.h);
    // The imaging pass runs inside the reverse launches
    // (rtm_reverse_image_kernel), split by interior columns into two halves:
    // launch st >= 2 images reverse steps st-2 and st-1 on half st % 2, and
    // launch 1 images step 0 alone on half 1, each while those steps' U
    // buffers and receiver levels are intact.  So half 0 takes the steps in
    // pairs (0, 1), (2, 3), ... and half 1 takes step 0, then (1, 2), (3, 4),
    // ...; every column is imaged step by step in the reference's order, its
    // image read and written once per pair, and every launch carries the same
    // share of the pass.  The steps a half has left after the loop (at most
    // two) are imaged by image_kernel.  (With the whole pass of a pair in
    // every second launch, the launches alternated between bound by memory
    // and bound by the pass's floating-point work: 0.590 against 0.565 ms per
    // reverse step on the 26-shot survey launch.  As a separate launch the
    // pass cost 0.40 ms per pair there; on a second stream beside the next
    // reverse step it measured no faster than that.)
    const int ntile = image_tiles(NZ, g.pml);
    const int hc = (nxmax + 1) / 2;  // columns of half 0; half 1 has the rest
    auto half_c0 = [&](int h) { return h == 0 ? 0 : hc; };
    auto half_cols = [&](int h) { return h == 0 ? hc : nxmax - hc; };
    // U and R of each imaged step over the half's interior columns, and the
    // half's image once
    auto image_bytes = [&](int first, int pair, int h) {
      double b = 0;
      for (int s = 0; s < n; s++) {
        const int w = NXs[s] - 2 * g.pml;
        const int cols = std::max(0, std::min(w, half_c0(h) + half_cols(h)) - half_c0(h));
        const int steps = (first < shots[s].nt) + (pair && first + 1 < shots[s].nt);
        if (steps > 0) b += 2 * E * (double)cols * (steps * (double)NZ + g.nz);
      }
      return b;
    };
    for (int st = 0; st < maxbwd; st++) {
      double sb;
      const double a = cells_bwd(st, &sb, nullptr, nullptr);
      const int h = st % 2, pfirst = st >= 2 ? st - 2 : 0, pair = st >= 2;
      const bool fused = st >= 1 && half_cols(h) > 0;
      launch(quiet, stream, "step_rtm_reverse", 8 * E * a + sb + (fused ? image_bytes(pfirst, pair, h) : 0.0), 2 * a,
             [&] {
               if (!fused) {
                 launch_step<T>(RTM_BWD, k, dev, L, B, n, st, stream);
                 return;
               }
               const int nimg = ntile * half_cols(h);
               hipLaunchKernelGGL(rtm_reverse_image_kernel, dim3(nimg + L.nfwg + L.nbwg, n), dim3(WG), 0, stream, k,
                                  dev, L, st, taps, ntile, nimg, pfirst, pair, half_c0(h));
               if (B.nbwg > 0)
                 hipLaunchKernelGGL((band_kernel<T, RTM_BWD>), dim3((unsigned)B.nbwg, n), dim3(WG), 0, stream, k, dev,
                                    st);
             });
    }
    for (int h = 0; h < 2; h++) {
      if (half_cols(h) == 0) continue;
      // steps [0, done) of half h are imaged: by the last launch of parity h
      // in the loop (from launch 2 for half 0, from launch 1 for half 1)
      int last = maxbwd - 1;
      if (last >= 0 && last % 2 != h) last--;
      const int done = last >= (h == 0 ? 2 : 1) ? last : 0;
      const int left = maxbwd - done;  // 0, 1 or 2 steps
      if (left <= 0) continue;
      launch(quiet, stream, "rtm_image", image_bytes(done, left == 2, h), 0, [&] {
        hipLaunchKernelGGL((image_kernel<half>), dim3((unsigned)ntile, (unsigned)half_cols(h), (unsigned)n),
                           dim3(FQ_WG), 0, stream, k, dev, done, taps, left == 2 ? 1 : 0, 0, half_c0(h));
      });
    }
    hipLaunchKernelGGL((crop_kernel<T>), dim3(64, n), dim3(256), 0, stream, k, dev);
    HIPCHECK(hipStreamSynchronize(stream));
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(stream));
  return SEIS_OK;
}

// Image destinations of the shots of a batch: consecutive [nx_s][nz] blocks.
template <typename T>
std::vector<T*> image_outs(const GridSpec& g, int nshots, const ShotSpec* shots, T* img) {
  std::vector<T*> o(nshots, nullptr);
  if (!img) return o;
  size_t off = 0;
  for (int s = 0; s < nshots; s++) {
    o[s] = img + off;
    off += (size_t)shot_width(g, shots[s]) * g.nz;
  }
  return o;
}

template <typename T>
int run_batch(int mode, const GridSpec& g, int nshots, const ShotSpec* shots, const T* vel, const T* dv,
              const T* wav, const int* gz, T* rec, T* img, hipStream_t stream, int rtm_flags, int line_src,
              const Workspace* ws) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (nshots <= 0) return SEIS_OK;
  if ((rc = check_shots(mode, g, nshots, shots, line_src))) return rc;
  std::vector<size_t> bytes(nshots);
  for (int s = 0; s < nshots; s++) bytes[s] = shot_bytes(mode, g, shots[s], sizeof(T), sizeof(DevShot<T>));
  size_t budget;
  if (ws) budget = ws->bytes;
  else if ((rc = batch_budget(&budget))) return rc;
  std::vector<int> starts;
  if ((rc = plan_chunks(bytes, budget, batch_max_shots(), starts))) return rc;
  const std::vector<T*> outs = image_outs(g, nshots, shots, img);
  for (size_t c = 0; c + 1 < starts.size(); c++) {
    const int s0 = starts[c], n = starts[c + 1] - s0;
    size_t need = 256;
    for (int s = s0; s < s0 + n; s++) need += bytes[s];
    ChunkMem m;
    if (ws) {
      m.p = ws->ptr;
    } else {
      HIPCHECK(hipMalloc(&m.p, need));
      m.owned = true;
    }
    if ((rc = run_chunk<T>(mode, g, n, shots + s0, outs.data() + s0, vel, dv, wav, gz, rec, m.p, stream, rtm_flags,
                           line_src, ws != nullptr)))
      return rc;
  }
  return SEIS_OK;
}

template int run_batch<float>(int, const GridSpec&, int, const ShotSpec*, const float*, const float*, const float*,
                              const int*, float*, float*, hipStream_t, int, int, const Workspace*);

// adjointPlane for one chunk of shots (all of the grid's width).
template <typename T>
int run_rtm_plane_chunk(const GridSpec& g, int n, const ShotSpec* shots, T* const* outs, const T* vel, const T* sou,
                        const int* gz, const T* rec, void* mem, hipStream_t stream) {
  const int NZ = g.nz + 2 * g.pml;
  const int NX = g.nx + 2 * g.pml;
  const long long ld = field_ld(NZ, g.pml);
  const long long fe = field_elems(NX, ld);
  int NTmax = 0;
  for (int s = 0; s < n; s++) NTmax = std::max(NTmax, shots[s].nt + shots[s].ndel);
  // per shot: source field and receiver field (3 levels each), their band
  // stencil values, the image and the U buffers; then strips of every step
  const int nslots = 6, nraw = 2;
  const int nf = fields_per_shot(RUN_RTM_PLANE);
  std::vector<DevShot<T>> host(n);
  DevShot<T>* dev = reinterpret_cast<DevShot<T>*>(mem);
  char* p = reinterpret_cast<char*>(mem) + (sizeof(DevShot<T>) * n + 255) / 256 * 256;
  const long long origin = field_origin(ld, g.pml);
  std::vector<T*> bases(n);
  for (int s = 0; s < n; s++) {
    const ShotSpec& q = shots[s];
    T* base = reinterpret_cast<T*>(p);
    bases[s] = base;
    DevShot<T>& d = host[s];
    d = DevShot<T>{};
    d.vel = vel + q.model;
    d.wav = sou + q.wav;
    d.gz = gz + q.gz;
    d.rec = const_cast<T*>(rec) + q.rec;
    d.fld = base + origin;
    d.fstride = fe;
    d.raw = base + (size_t)fe * nslots + origin;
    d.img = base + (size_t)fe * (nslots + nraw) + origin;
    d.aux = base + (size_t)fe * (nslots + nraw + 1) + origin;
    d.out = outs[s];
    const size_t st = (size_t)(q.nt + q.ndel) * 8 * (g.nx + g.nz) * sizeof(T);
    d.strips = reinterpret_cast<T*>(p + (size_t)fe * nf * sizeof(T));
    p += (size_t)fe * nf * sizeof(T) + (st + 255) / 256 * 256;
    d.nx = g.nx;
    d.NX = NX;
    d.nt = q.nt;
    d.ndel = q.ndel;
    d.sx = 0;
    d.sz = q.sz;
    d.line_src = 1;
  }
  HIPCHECK(hipMemcpyAsync(dev, host.data(), sizeof(DevShot<T>) * n, hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL((zero_fields_kernel<T>), dim3(512, n), dim3(256), 0, stream, dev, origin, nf);
  // strips too: the strips saved at the last step are those of the zero
  // level the backward pass starts from
  for (int s = 0; s < n; s++)
    HIPCHECK(hipMemsetAsync(host[s].strips, 0,
                            sizeof(T) * (size_t)(shots[s].nt + shots[s].ndel) * 8 * (g.nx + g.nz), stream));
  const Consts<T> k = make_consts<T>(g, ld);
  std::vector<int> NXs(n, NX);
  Layout L = make_layout(NXs.data(), n, NZ, std::max(g.pml, 4));
  const Layout B = band_layout(NXs.data(), n, NZ, g.pml);
  {
    // strip length: a backward-pass step 0 is a trial that writes only what
    // the real step 0 writes again (its new level, band values, strips)
    struct Trial {
      const Consts<T>* k;
      const DevShot<T>* dev;
      Layout L, B;
      int nshots;
      hipStream_t s;
    };
    Trial tr{&k, dev, L, B, n, stream};
    auto fn = [](void* c, int bx) {
      const Trial* t = static_cast<const Trial*>(c);
      launch_step<T>(PW_BACK, *t->k, t->dev, with_strip(t->L, bx), t->B, t->nshots, 0, t->s);
    };
    L = with_strip(L, tuned_strip("rtm_plane", NX, NZ, n, stream, fn, &tr));
  }
  const double cells = (double)NX * NZ;
  const double E = sizeof(T);
  const double strip = 8.0 * (g.nx + g.nz) * E;
  auto active = [&](int mode, int st) {
    int c = 0;
    for (int s = 0; s < n; s++) c += step_it<T>(mode, host[s], st) >= 0;
    return c;
  };
  // pass 1: receiver field backward in time
  for (int st = 0; st < NTmax; st++) {
    const int a = active(PW_BACK, st);
    timed_launch(stream, "step_plane_back", 4 * E * a * cells + strip * a, a * cells, [&] {
      launch_step<T>(PW_BACK, k, dev, L, B, n, st, stream);
    });
  }
  // its last two levels start the forward rebuild (fd.cpp:848-852): the
  // level of time 0 as previous, of time 1 as current; the source field
  // starts from zero
  for (int s = 0; s < n; s++) {
    const int NT = shots[s].nt + shots[s].ndel;
    T* base = bases[s];
    HIPCHECK(hipMemcpyAsync(base + (size_t)fe * 3, base + (size_t)fe * ((NT + 1) % 3), sizeof(T) * fe,
                            hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(base + (size_t)fe * 4, base + (size_t)fe * (NT % 3), sizeof(T) * fe,
                            hipMemcpyDeviceToDevice, stream));
  }
  hipLaunchKernelGGL((zero_fields_kernel<T>), dim3(512, n), dim3(256), 0, stream, dev, origin, 3);
  // pass 2: line-source field forward, receiver field rebuilt forward,
  // FaQi imaging, strips restored into the receiver field's new level
  constexpr int half = HILBERT_HALF;
  HilbertTaps taps{};
  hilbert_taps_host<float>(half, taps.h);
  const dim3 igrid((unsigned)image_tiles(NZ, g.pml), (unsigned)g.nx, (unsigned)n);
  const double faqi_cells = (double)g.nx * NZ;
  for (int st = 0; st + 2 < NTmax; st++) {
    const int a = active(PW_FWD, st);
    timed_launch(stream, "step_plane_fwd", 8 * E * a * cells, 2 * a * cells, [&] {
      launch_step<T>(PW_FWD, k, dev, L, B, n, st, stream);
    });
    timed_launch(stream, "plane_image", 2 * E * a * faqi_cells + 2 * E * a * (double)g.nx * g.nz, 0, [&] {
      hipLaunchKernelGGL((image_kernel<half>), igrid, dim3(FQ_WG), 0, stream, k, dev, st, taps, 0, 1, 0);
    });
    hipLaunchKernelGGL((restore_plane_strips_kernel<T>), dim3(16, n), dim3(256), 0, stream, k, dev, st);
  }
  hipLaunchKernelGGL((crop_kernel<T>), dim3(64, n), dim3(256), 0, stream, k, dev);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(stream));
  return SEIS_OK;
}

template <typename T>
int run_rtm_plane(const GridSpec& g, int nshots, const ShotSpec* shots, const T* vel, const T* sou, const int* gz,
                  const T* rec, T* img, hipStream_t stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  if (nshots <= 0) return SEIS_OK;
  if ((rc = check_shots(RUN_RTM_PLANE, g, nshots, shots, 1))) return rc;
  std::vector<size_t> bytes(nshots);
  for (int s = 0; s < nshots; s++) bytes[s] = shot_bytes(RUN_RTM_PLANE, g, shots[s], sizeof(T), sizeof(DevShot<T>));
  size_t budget;
  if ((rc = batch_budget(&budget))) return rc;
  std::vector<int> starts;
  if ((rc = plan_chunks(bytes, budget, batch_max_shots(), starts))) return rc;
  const std::vector<T*> outs = image_outs(g, nshots, shots, img);
  for (size_t c = 0; c + 1 < starts.size(); c++) {
    const int s0 = starts[c], n = starts[c + 1] - s0;
    size_t need = 256;
    for (int s = s0; s < s0 + n; s++) need += bytes[s];
    ChunkMem m;
    HIPCHECK(hipMalloc(&m.p, need));
    m.owned = true;
    if ((rc = run_rtm_plane_chunk<T>(g, n, shots + s0, outs.data() + s0, vel, sou, gz, rec, m.p, stream))) return rc;
  }
  return SEIS_OK;
}

template int run_rtm_plane<float>(const GridSpec&, int, const ShotSpec*, const float*, const float*, const int*,
                                  const float*, float*, hipStream_t);

// Device scratch of seis_modeling_batch_static: fields and shot descriptors
// of a modeling batch, for callers that cannot allocate device memory.  One
// copy per GPU, held until the process ends.  In host code `scratch` names a
// host-side stand-in; its device address comes from hipGetSymbolAddress.
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

// Device address of the current device's scratch, looked up once per device
// (with that device current) on its first use.
int static_scratch(void** out) {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error("static scratch: device index out of range"), SEIS_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SEIS_OK;
}

}  // namespace seis

using namespace seis;

extern "C" {

int seis_modeling_batch_static(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                               const float* wav, const int* gz, float* rec) {
  if (!grid || nshots < 0 || (nshots > 0 && !shots)) return set_error("bad grid or shot list"), SEIS_ERR_ARG;
  int rc = check_grid(*grid);
  if (rc) return rc;
  if ((rc = check_shots(RUN_MODEL, *grid, nshots, shots, 0))) return rc;
  for (int s = 0; s < nshots; s++) {
    const size_t b = shot_bytes(RUN_MODEL, *grid, shots[s], sizeof(float), sizeof(DevShot<float>)) + 256;
    if (b > kStaticScratchBytes)
      return set_error("seis_modeling_batch_static: shot " + std::to_string(s) + " needs " + std::to_string(b) +
                       " bytes of scratch, more than the " + std::to_string(kStaticScratchBytes) +
                       " bytes of the static scratch"),
             SEIS_ERR_ARG;
  }
  void* p = nullptr;
  if ((rc = static_scratch(&p))) return rc;
  const Workspace ws{p, kStaticScratchBytes};
  return run_batch<float>(RUN_MODEL, *grid, nshots, shots, vel, nullptr, wav, gz, rec, nullptr, 0, 0, 0, &ws);
}

int seis_set_batch_limits(size_t bytes, int shots) {
  if (shots < 0) return set_error("seis_set_batch_limits: shots must be >= 0"), SEIS_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_limit_mu);
  g_limit_bytes = bytes;
  g_limit_shots = shots;
  return SEIS_OK;
}

int seis_batch_plan(int mode, const seis_grid* grid, int nshots, const seis_shot* shots, size_t budget,
                    int max_shots, int* starts, int* nchunks) {
  if (!grid || (nshots > 0 && !shots) || !nchunks || nshots < 0 || max_shots < 0)
    return set_error("seis_batch_plan: bad argument"), SEIS_ERR_ARG;
  if (mode < SEIS_PLAN_MODEL || mode > SEIS_PLAN_RTM_PLANE) return set_error("seis_batch_plan: unknown mode"), SEIS_ERR_ARG;
  const int rmode = mode == SEIS_PLAN_MODEL ? RUN_MODEL : mode == SEIS_PLAN_BORN ? RUN_BORN
                    : mode == SEIS_PLAN_RTM ? RUN_RTM : mode == SEIS_PLAN_ILLUM ? RUN_ILLUM : RUN_RTM_PLANE;
  int rc = check_grid(*grid);
  if (rc) return rc;
  if ((rc = check_shots(rmode, *grid, nshots, shots, rmode == RUN_RTM_PLANE))) return rc;
  std::vector<size_t> bytes(nshots);
  for (int s = 0; s < nshots; s++) bytes[s] = shot_bytes(rmode, *grid, shots[s], sizeof(float), sizeof(DevShot<float>));
  std::vector<int> st;
  if ((rc = plan_chunks(bytes, budget, max_shots, st))) return rc;
  if (starts) std::copy(st.begin(), st.end(), starts);
  *nchunks = (int)st.size() - 1;
  return SEIS_OK;
}

}  // extern "C"
