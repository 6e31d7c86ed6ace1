// Per-cell arithmetic of one leapfrog step with the hybrid absorbing
// boundary, written so that the new value of every padded cell is a pure
// function of the current and previous levels (plus static coefficients and
// the step's injections).  A kernel can therefore compute any cell -- and, for
// the absorbing band, the stencil value of its neighbours -- without any
// ordering between cells inside a step: the stencil (fd.cpp:1226-1271), the
// edge ring and one-way band of processBoundary (fd.cpp:1053-1180) and the
// sponge blend (fd.cpp:55-56) collapse into one gather per cell.
//
// Every expression keeps the reference's operation order; the library is
// compiled without floating-point contraction, so the float instance
// reproduces the reference's FP32 results bit for bit.
//
// Grid conventions: padded grid NX x NZ (NX = nx + 2 pml, NZ = nz + 2 pml),
// x = column (slow), z = row (fast).  Field storage has a leading dimension
// `ld` >= NZ and a zeroed margin around the padded grid (see Field).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#define SEIS_HD __host__ __device__

namespace seis {

// Global-address-space view of a device pointer.  Every array these kernels
// touch is device (global) memory; reaching it through a generic pointer
// compiles to FLAT instructions, whose completion the compiler can only
// await with a full wait, which serialises each load with everything issued
// before it.  Through the global view loads become global_load and waits
// count only the loads they depend on.  On the host (CPU builds of these
// headers) the view is the pointer itself.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> gl(T* p) {
  return (gptr<T>)p;
}
#define SEIS_GL(p) ::seis::gl(p)
#else
#define SEIS_GL(p) (p)
#endif

SEIS_HD inline int imin(int a, int b) { return a < b ? a : b; }
SEIS_HD inline int imax(int a, int b) { return a > b ? a : b; }

// Step constants shared by all shots of a batch, formed on the host in T with
// the reference's expression order.
template <typename T>
struct Consts {
  int nz, pml, NZ;
  long long ld;       // z stride of field storage
  T dt, dx, dz;
  T idx2, idz2;       // 1/(dx*dx), 1/(dz*dz)                 fd.cpp:21-22
  T idt2;             // 1/dt/dt                               fd.cpp:89
  T cx, cz;           // (1/dx)(1/dx) dt*dt, (1/dz)(1/dz) dt*dt fd.cpp:1057-1059
  T a1z, a1x;         // 1/dz/2/dt, 1/dx/2/dt                  fd.cpp:1098,1124
  T a2b;              // 1/2/dt/dt (then / v)                  fd.cpp:1099
  T b1, b2;           // 1/dz, 1/dx                            fd.cpp:1150-1151
  T sq2;              // sqrt(2) in T                          fd.cpp:1152
  T c[5];             // -205/72, 8/5, -1/5, 8/315, -1/560     fd.cpp:1245-1253
};

// Shot-local model access.  The padded velocity of setVel
// (InitInversion.cpp:92-113) is the model with both indices clamped, so the
// kernel reads the unpadded model directly.
template <typename T>
struct Model {
  const T* v;   // [nx][nz]
  int nx, nz, pml;
  SEIS_HD inline T vel(int x, int z) const {
    int xi = imin(imax(x - pml, 0), nx - 1);
    int zi = imin(imax(z - pml, 0), nz - 1);
    return SEIS_GL(v)[(long long)xi * nz + zi];
  }
};

// Linear index of (x, z) in field storage.
SEIS_HD inline long long fidx(long long ld, int x, int z) { return (long long)x * ld + z; }

// Sponge weight of setAbs (InitInversion.cpp:11-40): 1 inside, linear ramp
// in the band, x ramp wins in the corners.
template <typename T>
SEIS_HD inline T sponge_weight(int x, int z, int nx, int nz, int pml) {
  int i;
  if (x < pml) i = pml - x - 1;
  else if (x >= pml + nx) i = x - pml - nx;
  else if (z < pml) i = pml - z - 1;
  else if (z >= pml + nz) i = z - pml - nz;
  else return T(1);
  return T(1) - T(i + 1) * T(1) / T(pml);
}

// The injection hooks add the step's sources to the raw stencil value of a
// cell and may correct the previous level (reverse-time source removal).
struct NoInject {
  template <typename T>
  SEIS_HD inline T add(T val, int, int, T, long long) const { return val; }
  template <typename T>
  SEIS_HD inline T prev(const T* P, long long i, int, int) const { return SEIS_GL(P)[i]; }
};

// Source term u += vt2 * amp: at one cell with amp = wav[it-1]
// (fd.cpp:49-51), or, for the plane-wave engine, at every model column ix of
// one row with that column's own sample sou[ix][it-1] (fd.cpp:223-226).
template <typename T>
struct PointSource {
  long long src;  // field index of the source cell, -1 when inactive this step
  T amp;          // wav[it-1]
  const T* line = nullptr;  // line source: column ix's sample at line[ix * lstride]
  long long lstride = 0;
  int lrow = -1, lx0 = 0, lnx = 0;  // padded row, first model column, model columns
  SEIS_HD inline T add(T val, int x, int z, T vt2, long long i) const {
    if (line) {
      const int ix = x - lx0;
      return (z == lrow && ix >= 0 && ix < lnx) ? val + vt2 * SEIS_GL(line)[ix * lstride] : val;
    }
    return i == src ? val + vt2 * amp : val;
  }
  template <typename U>
  SEIS_HD inline U prev(const U* P, long long i, int, int) const { return SEIS_GL(P)[i]; }
};

// Reverse-time source removal from the previous level (fd.cpp:659-661):
// P[src] -= vt2[src] * wav[it-1] when it-1 > 0.
template <typename T>
struct SourceRemoval {
  long long src;
  T sub;  // vt2[src] * wav[it-1], 0 when inactive
  bool on;
  SEIS_HD inline T add(T val, int, int, T, long long) const { return val; }
  SEIS_HD inline T prev(const T* P, long long i, int, int) const {
    return (on && i == src) ? SEIS_GL(P)[i] - sub : SEIS_GL(P)[i];
  }
};

// Receiver-side injection u += rec[ix][itp] * vt2 at (ix+pml, gz[ix]+pml)
// for every model column (fd.cpp:673-677); nothing for itp < 0.
template <typename T>
struct ReceiverInject {
  const int* gz;      // [nx] receiver depth per column (model cells)
  const T* rec;       // [nx][nt]
  int nx, nt, itp, pml;
  SEIS_HD inline T add(T val, int x, int z, T vt2, long long) const {
    int ix = x - pml;
    if (itp >= 0 && ix >= 0 && ix < nx && z == SEIS_GL(gz)[ix] + pml)
      return val + SEIS_GL(rec)[(long long)ix * nt + itp] * vt2;
    return val;
  }
  template <typename U>
  SEIS_HD inline U prev(const U* P, long long i, int, int) const { return SEIS_GL(P)[i]; }
};

// Record removal from the previous level, P -= rec[ix][itp] * vt2 at the
// receiver cells (adjointPlane's forward rebuild of the receiver field,
// fd.cpp:862-866); nothing for itp < 0.
template <typename T>
struct ReceiverRemoval {
  const int* gz;
  const T* rec;
  int nx, nt, itp, pml;
  T dt;
  const Model<T>* m;
  SEIS_HD inline T add(T val, int, int, T, long long) const { return val; }
  SEIS_HD inline T prev(const T* P, long long i, int x, int z) const {
    const int ix = x - pml;
    if (itp >= 0 && ix >= 0 && ix < nx && z == SEIS_GL(gz)[ix] + pml) {
      const T a = dt * m->vel(x, z);
      return SEIS_GL(P)[i] - SEIS_GL(rec)[(long long)ix * nt + itp] * (a * a);
    }
    return SEIS_GL(P)[i];
  }
};

// ---------------------------------------------------------------------------
// Stencil value ("u3" of processBoundary) of cell (x, z): stencil + injection
// inside, edge formula on the outer ring.
template <typename T, class Inj>
SEIS_HD inline T raw_value(const Consts<T>& k, const Model<T>& m, int NX, const T* C, const T* P,
                                       const Inj& inj, int x, int z) {
  const long long ld = k.ld;
  const int NZ = k.NZ;
  const long long i = fidx(ld, x, z);
  const bool xe = (x == 0) || (x == NX - 1);
  const bool ze = (z == 0) || (z == NZ - 1);
  if (xe || ze) {
    // ring, fd.cpp:1060-1091
    auto vv = [&](int xx, int zz) { T v = m.vel(xx, zz); return v * v; };
    auto cu = [&](int xx, int zz) { return SEIS_GL(C)[fidx(ld, xx, zz)]; };
    T r;
    const T c0 = cu(x, z);
    const T p0 = inj.prev(P, i, x, z);
    if (xe && ze) {
      r = T(2) * c0 - p0;
    } else if (ze) {
      r = vv(x, z) * k.cx * (cu(x + 1, z) + cu(x - 1, z) - T(2) * c0) + T(2) * c0 - p0;
    } else {
      const T f = (x == 0) ? k.cz : k.cx;  // right edge uses the x factor, fd.cpp:1066
      r = vv(x, z) * f * (cu(x, z + 1) + cu(x, z - 1) - T(2) * c0) + T(2) * c0 - p0;
    }
    if (z == 0) {
      T d1 = vv(x, 1) * k.cz * (cu(x, 2) + cu(x, 0) - T(2) * cu(x, 1));
      T d2 = vv(x, 2) * k.cz * (cu(x, 3) + cu(x, 1) - T(2) * cu(x, 2));
      r += d1 * T(2) - d2;
    } else if (z == NZ - 1) {
      T d1 = vv(x, NZ - 2) * k.cz * (cu(x, NZ - 3) + cu(x, NZ - 1) - T(2) * cu(x, NZ - 2));
      T d2 = vv(x, NZ - 3) * k.cz * (cu(x, NZ - 4) + cu(x, NZ - 2) - T(2) * cu(x, NZ - 3));
      r += d1 * T(2) - d2;
    }
    if (x == 0) {
      T d1 = vv(1, z) * k.cx * (cu(2, z) + cu(0, z) - T(2) * cu(1, z));
      T d2 = vv(2, z) * k.cx * (cu(3, z) + cu(1, z) - T(2) * cu(2, z));
      r += d1 * T(2) - d2;
    } else if (x == NX - 1) {
      T d1 = vv(NX - 2, z) * k.cx * (cu(NX - 3, z) + cu(NX - 1, z) - T(2) * cu(NX - 2, z));
      T d2 = vv(NX - 3, z) * k.cx * (cu(NX - 4, z) + cu(NX - 2, z) - T(2) * cu(NX - 3, z));
      r += d1 * T(2) - d2;
    }
    return r;
  }
  const T v = m.vel(x, z);
  const T a = k.dt * v;
  const T vt2 = a * a;
  const T vvzz = vt2 * k.idz2;
  const T vvxx = vt2 * k.idx2;
  const T c0 = SEIS_GL(C)[i];
  T uzz, uxx;
  if (x >= 4 && x <= NX - 5 && z >= 4 && z <= NZ - 5) {
    const auto cz = SEIS_GL(C) + i;
    uzz = cz[-4] * k.c[4];
    uzz = uzz + cz[-3] * k.c[3];
    uzz = uzz + cz[-2] * k.c[2];
    uzz = uzz + cz[-1] * k.c[1];
    uzz = uzz + cz[0] * k.c[0];
    uzz = uzz + cz[1] * k.c[1];
    uzz = uzz + cz[2] * k.c[2];
    uzz = uzz + cz[3] * k.c[3];
    uzz = uzz + cz[4] * k.c[4];
    uxx = cz[-4 * ld] * k.c[4];
    uxx = uxx + cz[-3 * ld] * k.c[3];
    uxx = uxx + cz[-2 * ld] * k.c[2];
    uxx = uxx + cz[-1 * ld] * k.c[1];
    uxx = uxx + cz[0] * k.c[0];
    uxx = uxx + cz[1 * ld] * k.c[1];
    uxx = uxx + cz[2 * ld] * k.c[2];
    uxx = uxx + cz[3 * ld] * k.c[3];
    uxx = uxx + cz[4 * ld] * k.c[4];
  } else {
    uzz = SEIS_GL(C)[i - 1] + SEIS_GL(C)[i + 1] - T(2) * c0;
    uxx = SEIS_GL(C)[i + ld] + SEIS_GL(C)[i - ld] - T(2) * c0;
  }
  T val = vvzz * uzz + vvxx * uxx + T(2) * c0 - inj.prev(P, i, x, z);
  return inj.add(val, x, z, vt2, i);
}

// Band position of a cell: the distance to the nearest outer edge; cells
// with dist < pml are updated by the one-way layers of processBoundary.
SEIS_HD inline int edge_dist(int x, int z, int NX, int NZ) {
  return imin(imin(x, NX - 1 - x), imin(z, NZ - 1 - z));
}

// New value of band cell (x, z) (edge_dist < pml) from its own stencil value
// r: the one-way update of processBoundary (fd.cpp:1092-1178) followed by the
// sponge blend (fd.cpp:55-56).  `rawn(xx, zz)` gives the stencil value of a
// neighbour at the same new time level.
template <typename T, class Inj, class RawN>
SEIS_HD inline T band_blend(const Consts<T>& k, const Model<T>& m, int NX, const T* C, const T* P, const Inj& inj,
                            int x, int z, T r, const RawN& rawn) {
  const int pml = k.pml, NZ = k.NZ;
  const int dl = x, dr = NX - 1 - x, dtp = z, db = NZ - 1 - z;
  const int mn = imin(imin(dl, dr), imin(dtp, db));
  const long long ld = k.ld;
  const long long i = fidx(ld, x, z);
  auto Cu = [&](int xx, int zz) { return SEIS_GL(C)[fidx(ld, xx, zz)]; };
  auto Pv = [&](int xx, int zz) { return inj.prev(P, fidx(ld, xx, zz), xx, zz); };
  const T v = m.vel(x, z);
  T ow;
  if ((dtp == mn || db == mn) && (dl == mn || dr == mn)) {
    // diagonal corner cell, fd.cpp:1145-1178
    const int sz = (dtp == mn) ? 1 : -1;
    const int sx = (dl == mn) ? 1 : -1;
    T b3;
    if (sx == 1 && sz == 1)
      b3 = (T)(std::sqrt(2.0) / (double)v / (double)k.dt);  // fd.cpp:1176 forms sqrt(2) in double
    else
      b3 = k.sq2 / v / k.dt;
    T t = -k.b1 * rawn(x, z + sz) - k.b2 * rawn(x + sx, z) + b3 * (-SEIS_GL(C)[i]);
    ow = -t / (k.b1 + k.b2 + b3);
  } else if (dtp == mn || db == mn) {
    // top / bottom band: one-way update along z, fd.cpp:1092-1118
    const int s = (dtp == mn) ? 1 : -1;
    const T nin = rawn(x, z + s), nip = rawn(x + 1, z + s), nim = rawn(x - 1, z + s);
    const T a1 = k.a1z;
    const T a2 = k.a2b / v;
    const T a3 = v / T(4) / k.dx / k.dx;
    const T p0 = Pv(x, z), pin = Pv(x, z + s);
    T t = -a1 * nin - a1 * (p0 - pin) + a2 * (p0 - T(2) * SEIS_GL(C)[i] + nin + pin - T(2) * Cu(x, z + s)) -
          a3 * (nip + nim - T(2) * nin + Pv(x + 1, z) + Pv(x - 1, z) - T(2) * p0);
    ow = -t / (a1 + a2);
  } else {
    // left / right band: one-way update along x, fd.cpp:1119-1144
    const int s = (dl == mn) ? 1 : -1;
    const T nin = rawn(x + s, z), nip = rawn(x + s, z + 1), nim = rawn(x + s, z - 1);
    const T a1 = k.a1x;
    const T a2 = k.a2b / v;
    const T a3 = v / T(4) / k.dz / k.dz;
    const T p0 = Pv(x, z), pin = Pv(x + s, z);
    T t = -a1 * nin - a1 * (p0 - pin) + a2 * (p0 - T(2) * SEIS_GL(C)[i] + nin + pin - T(2) * Cu(x + s, z)) -
          a3 * (nip + nim - T(2) * nin + Pv(x, z + 1) + Pv(x, z - 1) - T(2) * p0);
    ow = -t / (a1 + a2);
  }
  const T w = sponge_weight<T>(x, z, m.nx, m.nz, pml);
  return r * w + ow * (T(1) - w);
}

// New value of cell (x, z) after the absorbing boundary and the sponge blend,
// recomputing the neighbours' stencil values (a pure function of cur and
// prev).  `raw_out` receives the stencil value of the cell itself.
template <typename T, class Inj>
SEIS_HD inline T final_value(const Consts<T>& k, const Model<T>& m, int NX, const T* C, const T* P,
                                         const Inj& inj, int x, int z, T* raw_out = nullptr) {
  const T r = raw_value(k, m, NX, C, P, inj, x, z);
  if (raw_out) *raw_out = r;
  if (edge_dist(x, z, NX, k.NZ) >= k.pml) return r;
  auto rawn = [&](int xx, int zz) { return raw_value(k, m, NX, C, P, inj, xx, zz); };
  return band_blend(k, m, NX, C, P, inj, x, z, r, rawn);
}

}  // namespace seis
