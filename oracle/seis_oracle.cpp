// CPU oracle for the 2D acoustic finite-difference engine.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU, the algorithm of
// the reference engine (bbsun/Seis src/util/fd.cpp, src/util/InitInversion.cpp,
// src/util/wavelet.cpp, src/filter/hilbert.h) so that the HIP engine in
// seis_amd/ can be checked against it.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it, and only as a checker or as the
// timed CPU baseline; the product path never calls it.
//
// Every routine is templated on the arithmetic type: the float instance keeps
// the reference's operation order (so it reproduces the reference's FP32
// results), the double instance is the FP64 oracle used to state the error of
// every FP32/BF16/FP16 output.
//
// Layout (reference memory.h:29-44): a 2D array a[i2][i1] is one contiguous
// block with i1 fastest.  Wavefields and models are [x][z] (z fastest), shot
// records are [x][t] (t fastest).  Padded sizes are NZ = nz + 2*pml and
// NX = nx + 2*pml (fd.cpp:17-18).
//
// Pinning status of each function (what checks it in tests/):
//   sponge_weights, pad_velocity, pad_perturbation : brute force on tiny grids
//                                       and the golden vectors of the compiled reference
//   ricker, hilbert_taps              : closed forms + golden vectors
//   harmonic                          : golden vectors of the compiled reference;
//                                       kx*x = 0 and pi/2 reduce it to in and -H[in]
//   plane_prepare                     : ray parameter 0 reduces it to stacking the
//                                       wavelet and the traces by column (closed form);
//                                       its harmonic is pinned as above
//   stencil_step                      : polynomial exactness + golden vectors
//   habc_update, modeling, born, rtm  : golden vectors of the compiled reference
//   born_adjoint (exact transpose)    : dot-product test (an invariant), and
//                                       linearization of modeling
//   born_adjoint_xcorr                : equals born_adjoint where no wave reaches
//                                       the band (FP64), and the dot test with born there
// (The CG solver has no CPU restatement: the GPU CGLS is checked against CGLS
// written out in numpy on the explicit matrix of the operator, in tests/.)

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

int g_threads = 1;  // plays the role of the reference's OMP_CORE (global.cpp:2)

inline void set_omp() {
#ifdef _OPENMP
  omp_set_dynamic(0);
  omp_set_num_threads(g_threads);
#endif
}

// Padded grid geometry.
struct Geom {
  int nx, nz, pml, NX, NZ;
  Geom(int nx_, int nz_, int pml_)
      : nx(nx_), nz(nz_), pml(pml_), NX(nx_ + 2 * pml_), NZ(nz_ + 2 * pml_) {}
  size_t cells() const { return (size_t)NX * NZ; }
  size_t at(int ix, int iz) const { return (size_t)ix * NZ + iz; }
};

// Sponge blend weights, InitInversion.cpp:11-40 (setAbs).  w = 1 in the
// interior, a linear ramp in the absorbing band; the x ramp is written for
// every z, the z ramp only for interior x, so x wins in the corners.
template <typename T>
void sponge_weights(const Geom& g, T* w) {
  std::vector<T> ramp(g.pml);
  for (int i = 0; i < g.pml; i++) ramp[i] = T(1) - T(i + 1) * T(1) / T(g.pml);
  for (size_t i = 0; i < g.cells(); i++) w[i] = T(1);
  for (int iz = 0; iz < g.NZ; iz++) {
    for (int ix = 0; ix < g.pml; ix++) w[g.at(ix, iz)] = ramp[g.pml - ix - 1];
    for (int ix = g.pml + g.nx; ix < g.NX; ix++) w[g.at(ix, iz)] = ramp[ix - g.pml - g.nx];
  }
  for (int ix = g.pml; ix < g.pml + g.nx; ix++) {
    for (int iz = 0; iz < g.pml; iz++) w[g.at(ix, iz)] = ramp[g.pml - iz - 1];
    for (int iz = g.pml + g.nz; iz < g.NZ; iz++) w[g.at(ix, iz)] = ramp[iz - g.pml - g.nz];
  }
}

// Edge-replicated padded velocity, InitInversion.cpp:92-113 (setVel):
// copy the model, then replicate the first/last interior row in z for every
// padded column, then the first/last interior column in x.
template <typename T>
void pad_velocity(const Geom& g, const T* v, T* vp) {
  for (size_t i = 0; i < g.cells(); i++) vp[i] = T(0);
  for (int ix = 0; ix < g.nx; ix++)
    for (int iz = 0; iz < g.nz; iz++) vp[g.at(ix + g.pml, iz + g.pml)] = v[(size_t)ix * g.nz + iz];
  for (int iz = 0; iz < g.pml; iz++)
    for (int ix = 0; ix < g.NX; ix++) {
      vp[g.at(ix, iz)] = vp[g.at(ix, g.pml)];
      vp[g.at(ix, g.NZ - g.pml + iz)] = vp[g.at(ix, g.NZ - g.pml - 1)];
    }
  for (int ix = 0; ix < g.pml; ix++)
    for (int iz = 0; iz < g.NZ; iz++) {
      vp[g.at(ix, iz)] = vp[g.at(g.pml, iz)];
      vp[g.at(g.NX - g.pml + ix, iz)] = vp[g.at(g.NX - g.pml - 1, iz)];
    }
}

// Zero-padded perturbation, InitInversion.cpp:116-126 (setDv).
template <typename T>
void pad_perturbation(const Geom& g, const T* dv, T* dp) {
  for (size_t i = 0; i < g.cells(); i++) dp[i] = T(0);
  for (int ix = 0; ix < g.nx; ix++)
    for (int iz = 0; iz < g.nz; iz++) dp[g.at(ix + g.pml, iz + g.pml)] = dv[(size_t)ix * g.nz + iz];
}

// Ricker wavelet, wavelet.cpp:7-18.  pi is the reference's literal 3.1415926;
// the factor pi^2 f^2 is formed in double and stored to T.
template <typename T>
void ricker(T dt, T fr, int num_delay, int n, T* w) {
  T a = (T)(3.1415926 * 3.1415926 * fr * fr);
  for (int i = 0; i < n; i++) {
    T t = T(i - num_delay) * T(1) * dt;
    T t2 = t * t;
    w[i] = (T(1) - T(2) * a * t2) * std::exp(-a * t2);
  }
}

// Hamming-tapered Hilbert FIR of 2*half+1 taps, hilbert.h:59-70.  The taper
// is evaluated in double and stored to T; the tap is (taper * (i%2)) in T,
// then scaled by 2/(pi*i) in double and stored to T.
template <typename T>
void hilbert_taps(int half, T* h) {
  h[half] = T(0);
  for (int i = 1; i <= half; i++) {
    T taper = (T)(0.54 + 0.46 * std::cos(3.1415926 * i / half));
    T hi = (T)((double)(taper * T(i % 2)) * 2.0 / (3.1415926 * i));
    h[half + i] = hi;
    h[half - i] = -hi;
  }
}

// Time-domain Hilbert filter of one trace, hilbert.h:42-54: out[j] sums
// in[i] * h[j - i + half] over i in increasing order (truncated at the ends).
template <typename T>
void hilbert_apply(const T* h, int half, const T* in, T* out, int n) {
  for (int j = 0; j < n; j++) out[j] = T(0);
  const int lh = 2 * half + 1;
  for (int i = 0; i < n; i++)
    for (int m = 0; m < lh; m++) {
      int j = i + m - half;
      if (j >= 0 && j < n) out[j] += in[i] * h[m];
    }
}

// Plane-wave phase shift, InitInversion.cpp:141-151: H[in] with the 201-tap
// filter (Hilbert<float> hl(100)), then out = cos(kx x) in - sin(kx x) H[in],
// cos and sin of the T product kx*x in T (cosf / sinf for float).
template <typename T>
void harmonic(const T* in, T* out, int n, T kx, T x) {
  const T ang = kx * x;
  const T cn = std::cos(ang), sn = std::sin(ang);
  std::vector<T> h(201);
  hilbert_taps<T>(100, h.data());
  hilbert_apply<T>(h.data(), 100, in, out, n);
  for (int it = 0; it < n; it++) out[it] = cn * in[it] - sn * out[it];
}

// planeWavePrepare, inversion.cpp:860-900, for ray parameters
// p_k = pmin + k*dp (k < np): per shot s in order, the source term
// harmonic(wav, p_k, sx[s]) is added into sou_k[isx[s]] (:878-881) and the
// phase-shifted traces are placed by column (swapReord: zeroed, later traces
// overwrite earlier ones, :626-643) and added into rec_k (:887-893).
// data holds the gathers [sum ng][nt] shot after shot; col[] the column of
// every trace or -1.  sou [np][nx][ntw], rec [np][nx][nt].
template <typename T>
void plane_prepare(int nx, int nt, int ntw, int ns, int np, T pmin, T dp, const T* sx, const int* isx,
                   const int* ng, const int* col, const T* wav, const T* data, T* sou, T* rec) {
  std::vector<T> tw(ntw), tr((size_t)nx * nt), csgn;
  for (int k = 0; k < np; k++) {
    T* so = sou + (size_t)k * nx * ntw;
    T* re = rec + (size_t)k * nx * nt;
    std::fill(so, so + (size_t)nx * ntw, T(0));
    std::fill(re, re + (size_t)nx * nt, T(0));
    const T p = pmin + T(k) * dp;
    size_t r0 = 0;
    for (int s = 0; s < ns; s++) {
      harmonic<T>(wav, tw.data(), ntw, p, sx[s]);
      T* dst = so + (size_t)isx[s] * ntw;
      for (int it = 0; it < ntw; it++) dst[it] = dst[it] + tw[it];
      csgn.assign((size_t)ng[s] * nt, T(0));
      for (int g = 0; g < ng[s]; g++)
        harmonic<T>(data + (r0 + g) * nt, csgn.data() + (size_t)g * nt, nt, p, sx[s]);
      std::fill(tr.begin(), tr.end(), T(0));
      for (int g = 0; g < ng[s]; g++)
        if (col[r0 + g] >= 0)
          std::memcpy(tr.data() + (size_t)col[r0 + g] * nt, csgn.data() + (size_t)g * nt, sizeof(T) * nt);
      for (size_t i = 0; i < tr.size(); i++) re[i] = tr[i] + re[i];
      r0 += ng[s];
    }
  }
}

// One leapfrog step of the interior stencil, fd.cpp:1226-1271.
// nxt = vvzz*d2z(cur) + vvxx*d2x(cur) + 2*cur - prv.  A 3-point pass over
// [1, N-2]^2 is overwritten by the 9-point (8th-order) pass over [4, N-5]^2.
// Parallel over x only, with the thread count of g_threads, like the reference.
template <typename T>
void stencil_step(const Geom& g, const T* prv, const T* cur, T* nxt, const T* vvzz, const T* vvxx) {
  const int NX = g.NX, NZ = g.NZ;
  set_omp();
#pragma omp parallel for
  for (int ix = 1; ix <= NX - 2; ix++)
    for (int iz = 1; iz <= NZ - 2; iz++) {
      size_t c = (size_t)ix * NZ + iz;
      T uzz = cur[c - 1] + cur[c + 1] - T(2) * cur[c];
      T uxx = cur[c + NZ] + cur[c - NZ] - T(2) * cur[c];
      nxt[c] = vvzz[c] * uzz + vvxx[c] * uxx + T(2) * cur[c] - prv[c];
    }
  const T k0 = T(-205) / T(72), k1 = T(8) / T(5), k2 = T(-1) / T(5), k3 = T(8) / T(315),
          k4 = T(-1) / T(560);
  const T cf[9] = {k4, k3, k2, k1, k0, k1, k2, k3, k4};
  set_omp();
#pragma omp parallel for
  for (int ix = 4; ix <= NX - 5; ix++)
    for (int iz = 4; iz <= NZ - 5; iz++) {
      size_t c = (size_t)ix * NZ + iz;
      T uzz = cur[c - 4] * cf[0];
      for (int k = 1; k < 9; k++) uzz = uzz + cur[c - 4 + k] * cf[k];
      T uxx = cur[c - 4 * (size_t)NZ] * cf[0];
      for (int k = 1; k < 9; k++) uxx = uxx + cur[c + (size_t)(k - 4) * NZ] * cf[k];
      nxt[c] = vvzz[c] * uzz + vvxx[c] * uxx + T(2) * cur[c] - prv[c];
    }
}

// Hybrid absorbing boundary, fd.cpp:1053-1180 (processBoundary).
//   nxt : new field (stencil output), its outer ring is written here
//   cur : current field, prv : previous field
//   ow  : receives the one-way-wave values of every band cell
// Ring: 1D second difference along the edge plus a linearly extrapolated
// normal second difference (the right edge scales its z difference by the x
// factor, as the reference does at fd.cpp:1066).  Band: one-way (paraxial)
// update per layer, reading the new field one cell further in.  Corners: a
// diagonal one-way condition.
template <typename T>
void habc_update(const Geom& g, T dt, T dx, T dz, const T* vel, T* ow, T* nxt, const T* cur,
                 const T* prv) {
  const int NX = g.NX, NZ = g.NZ, pml = g.pml, nx = g.nx, nz = g.nz;
  auto A = [&](int ix, int iz) -> size_t { return (size_t)ix * NZ + iz; };
  const T dt2 = dt * dt;
  const T cx = (T(1) / dx) * (T(1) / dx) * dt2;
  const T cz = (T(1) / dz) * (T(1) / dz) * dt2;
  // ring: edge-parallel second difference
  for (int ix = 1; ix <= NX - 2; ix++) {
    for (int iz : {0, NZ - 1}) {
      size_t c = A(ix, iz);
      nxt[c] = vel[c] * vel[c] * cx * (cur[A(ix + 1, iz)] + cur[A(ix - 1, iz)] - T(2) * cur[c]) +
               T(2) * cur[c] - prv[c];
    }
  }
  for (int iz = 1; iz <= NZ - 2; iz++) {
    size_t c = A(0, iz);
    nxt[c] = vel[c] * vel[c] * cz * (cur[c + 1] + cur[c - 1] - T(2) * cur[c]) + T(2) * cur[c] - prv[c];
    c = A(NX - 1, iz);
    nxt[c] = vel[c] * vel[c] * cx * (cur[c + 1] + cur[c - 1] - T(2) * cur[c]) + T(2) * cur[c] - prv[c];
  }
  for (int ix : {0, NX - 1})
    for (int iz : {0, NZ - 1}) nxt[A(ix, iz)] = T(2) * cur[A(ix, iz)] - prv[A(ix, iz)];
  // ring: extrapolated normal second difference (z, then x)
  for (int ix = 0; ix < NX; ix++) {
    T d1 = vel[A(ix, 1)] * vel[A(ix, 1)] * cz * (cur[A(ix, 2)] + cur[A(ix, 0)] - T(2) * cur[A(ix, 1)]);
    T d2 = vel[A(ix, 2)] * vel[A(ix, 2)] * cz * (cur[A(ix, 3)] + cur[A(ix, 1)] - T(2) * cur[A(ix, 2)]);
    nxt[A(ix, 0)] += d1 * T(2) - d2;
    d1 = vel[A(ix, NZ - 2)] * vel[A(ix, NZ - 2)] * cz *
         (cur[A(ix, NZ - 3)] + cur[A(ix, NZ - 1)] - T(2) * cur[A(ix, NZ - 2)]);
    d2 = vel[A(ix, NZ - 3)] * vel[A(ix, NZ - 3)] * cz *
         (cur[A(ix, NZ - 4)] + cur[A(ix, NZ - 2)] - T(2) * cur[A(ix, NZ - 3)]);
    nxt[A(ix, NZ - 1)] += d1 * T(2) - d2;
  }
  for (int iz = 0; iz < NZ; iz++) {
    T d1 = vel[A(1, iz)] * vel[A(1, iz)] * cx * (cur[A(2, iz)] + cur[A(0, iz)] - T(2) * cur[A(1, iz)]);
    T d2 = vel[A(2, iz)] * vel[A(2, iz)] * cx * (cur[A(3, iz)] + cur[A(1, iz)] - T(2) * cur[A(2, iz)]);
    nxt[A(0, iz)] += d1 * T(2) - d2;
    d1 = vel[A(NX - 2, iz)] * vel[A(NX - 2, iz)] * cx *
         (cur[A(NX - 3, iz)] + cur[A(NX - 1, iz)] - T(2) * cur[A(NX - 2, iz)]);
    d2 = vel[A(NX - 3, iz)] * vel[A(NX - 3, iz)] * cx *
         (cur[A(NX - 4, iz)] + cur[A(NX - 2, iz)] - T(2) * cur[A(NX - 3, iz)]);
    nxt[A(NX - 1, iz)] += d1 * T(2) - d2;
  }
  // band: one-way update per layer; layer k = 1 is next to the interior
  for (int k = 1; k <= pml; k++) {
    const int lo = pml + 1 - k;  // first cell of the layer's edge-parallel span
    // z boundaries (top, bottom); the inner neighbour is at K+s
    for (int side = 0; side < 2; side++) {
      const int K = side == 0 ? pml - k : nz + k + pml - 1;
      const int s = side == 0 ? 1 : -1;
      const int hi = pml + nx + k - 2;
      for (int ix = lo; ix <= hi; ix++) {
        T v = vel[A(ix, K)];
        T a1 = T(1) / dz / T(2) / dt;
        T a2 = T(1) / T(2) / dt / dt / v;
        T a3 = v / T(4) / dx / dx;
        T t = -a1 * nxt[A(ix, K + s)] - a1 * (prv[A(ix, K)] - prv[A(ix, K + s)]) +
              a2 * (prv[A(ix, K)] - T(2) * cur[A(ix, K)] + nxt[A(ix, K + s)] + prv[A(ix, K + s)] -
                    T(2) * cur[A(ix, K + s)]) -
              a3 * (nxt[A(ix + 1, K + s)] + nxt[A(ix - 1, K + s)] - T(2) * nxt[A(ix, K + s)] +
                    prv[A(ix + 1, K)] + prv[A(ix - 1, K)] - T(2) * prv[A(ix, K)]);
        ow[A(ix, K)] = -t / (a1 + a2);
      }
    }
    // x boundaries (right, left); the inner neighbour is at K+s
    for (int side = 0; side < 2; side++) {
      const int K = side == 0 ? nx + k + pml - 1 : pml - k;
      const int s = side == 0 ? -1 : 1;
      const int hi = pml + nz + k - 2;
      for (int iz = lo; iz <= hi; iz++) {
        T v = vel[A(K, iz)];
        T a1 = T(1) / dx / T(2) / dt;
        T a2 = T(1) / T(2) / dt / dt / v;
        T a3 = v / T(4) / dz / dz;
        T t = -a1 * nxt[A(K + s, iz)] - a1 * (prv[A(K, iz)] - prv[A(K + s, iz)]) +
              a2 * (prv[A(K, iz)] - T(2) * cur[A(K, iz)] + nxt[A(K + s, iz)] + prv[A(K + s, iz)] -
                    T(2) * cur[A(K + s, iz)]) -
              a3 * (nxt[A(K + s, iz + 1)] + nxt[A(K + s, iz - 1)] - T(2) * nxt[A(K + s, iz)] +
                    prv[A(K, iz + 1)] + prv[A(K, iz - 1)] - T(2) * prv[A(K, iz)]);
        ow[A(K, iz)] = -t / (a1 + a2);
      }
    }
    // corners: bottom-right, bottom-left, top-right, top-left (reference order)
    const T b1 = T(1) / dz, b2 = T(1) / dx;
    struct Corner { int xc, zc, sx, sz; bool dbl; };
    const Corner cs[4] = {
        {nx + pml + k - 1, nz + pml + k - 1, -1, -1, false},
        {pml - k, nz + pml + k - 1, +1, -1, false},
        {nx + pml + k - 1, pml - k, -1, +1, false},
        {pml - k, pml - k, +1, +1, true},  // fd.cpp:1176 forms sqrt(2) in double
    };
    for (const Corner& q : cs) {
      T v = vel[A(q.xc, q.zc)];
      T b3 = q.dbl ? (T)(std::sqrt(2.0) / (double)v / (double)dt) : std::sqrt(T(2)) / v / dt;
      T t = -b1 * nxt[A(q.xc, q.zc + q.sz)] - b2 * nxt[A(q.xc + q.sx, q.zc)] + b3 * (-cur[A(q.xc, q.zc)]);
      ow[A(q.xc, q.zc)] = -t / (b1 + b2 + b3);
    }
  }
}

// Sponge blend over the padded grid, fd.cpp:55-56: u = u*w + ow*(1-w).
template <typename T>
void blend(const Geom& g, T* u, const T* ow, const T* w) {
  for (size_t i = 0; i < g.cells(); i++) u[i] = u[i] * w[i] + ow[i] * (T(1) - w[i]);
}

// Boundary strips (fd.cpp:1181-1225, SaveAtBoundary): `layer` rows/columns
// just inside the interior edges, over the interior span.  Storage per time
// sample: top[nx][layer], bottom[nx][layer], left[layer][nz], right[layer][nz].
template <typename T>
struct Strips {
  int layer, nt, nx, nz;
  std::vector<T> top, bot, lft, rgt;
  Strips(int layer_, int nt_, int nx_, int nz_)
      : layer(layer_), nt(nt_), nx(nx_), nz(nz_), top((size_t)nt_ * nx_ * layer_),
        bot((size_t)nt_ * nx_ * layer_), lft((size_t)nt_ * nz_ * layer_), rgt((size_t)nt_ * nz_ * layer_) {}
  void move(const Geom& g, T* u, int it, bool save) {
    size_t base_x = (size_t)it * nx * layer, base_z = (size_t)it * nz * layer;
    for (int ix = 0; ix < nx; ix++)
      for (int il = 0; il < layer; il++) {
        size_t cu = g.at(ix + g.pml, g.pml + il), cd = g.at(ix + g.pml, nz + g.pml - 1 - il);
        size_t s = base_x + (size_t)ix * layer + il;
        if (save) { top[s] = u[cu]; bot[s] = u[cd]; }
        else { u[cu] = top[s]; u[cd] = bot[s]; }
      }
    for (int il = 0; il < layer; il++)
      for (int iz = 0; iz < nz; iz++) {
        size_t cl = g.at(g.pml + il, iz + g.pml), cr = g.at(nx + g.pml - 1 - il, iz + g.pml);
        size_t s = base_z + (size_t)il * nz + iz;
        if (save) { lft[s] = u[cl]; rgt[s] = u[cr]; }
        else { u[cl] = lft[s]; u[cr] = rgt[s]; }
      }
  }
};

// Per-shot coefficient arrays (fd.cpp:34-40, 115-116).
template <typename T>
struct Coefs {
  std::vector<T> w, vel, vt2, vvzz, vvxx, v3;
  Coefs(const Geom& g, T dt, T dx, T dz, const T* v, bool need_v3)
      : w(g.cells()), vel(g.cells()), vt2(g.cells()), vvzz(g.cells()), vvxx(g.cells()) {
    sponge_weights(g, w.data());
    pad_velocity(g, v, vel.data());
    const T idx2 = T(1) / (dx * dx), idz2 = T(1) / (dz * dz);
    for (size_t i = 0; i < g.cells(); i++) {
      T a = dt * vel[i];
      vt2[i] = a * a;
      vvzz[i] = vt2[i] * idz2;
      vvxx[i] = vt2[i] * idx2;
    }
    if (need_v3) {
      v3.resize(g.cells());
      for (size_t i = 0; i < g.cells(); i++) v3[i] = T(2) / (vel[i] * vel[i] * vel[i]);
    }
  }
};

// Forward shot modeling, fd.cpp:15-80.  rec is [nx][nt].
template <typename T>
void modeling(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, const int* gz,
              const T* wav, const T* v, T* rec) {
  Geom g(nx, nz, pml);
  Coefs<T> k(g, dt, dx, dz, v, false);
  std::vector<T> a(g.cells()), b(g.cells()), c(g.cells()), ow(g.cells());
  T *u0 = a.data(), *u1 = b.data(), *u2 = c.data();
  const size_t src = g.at(sx + pml, sz + pml);
  for (size_t i = 0; i < (size_t)nx * nt; i++) rec[i] = T(0);
  const int NT = nt + ndel;
  for (int it = 0; it < NT; it++) {
    stencil_step(g, u0, u1, u2, k.vvzz.data(), k.vvxx.data());
    if (it - 1 >= 0) u2[src] += k.vt2[src] * wav[it - 1];
    habc_update(g, dt, dx, dz, k.vel.data(), ow.data(), u2, u1, u0);
    blend(g, u2, ow.data(), k.w.data());
    const int itp = it - ndel;
    if (itp >= 0)
      for (int ix = 0; ix < nx; ix++) rec[(size_t)ix * nt + itp] = u2[g.at(ix + pml, gz[ix] + pml)];
    T* t = u0; u0 = u1; u1 = u2; u2 = t;
  }
}

// Source illumination, fd.cpp:275-339: the modeling loop without receivers;
// after the blend of every step it > ndel (strictly), ss += u2 * u2 at every
// cell.  The reference sums over the padded grid and returns the interior;
// only the interior is accumulated here.  ss is [nx][nz].
template <typename T>
void illum(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, const T* wav, const T* v,
           T* ss) {
  Geom g(nx, nz, pml);
  Coefs<T> k(g, dt, dx, dz, v, false);
  std::vector<T> a(g.cells()), b(g.cells()), c(g.cells()), ow(g.cells());
  T *u0 = a.data(), *u1 = b.data(), *u2 = c.data();
  const size_t src = g.at(sx + pml, sz + pml);
  for (size_t i = 0; i < (size_t)nx * nz; i++) ss[i] = T(0);
  const int NT = nt + ndel;
  for (int it = 0; it < NT; it++) {
    stencil_step(g, u0, u1, u2, k.vvzz.data(), k.vvxx.data());
    if (it - 1 >= 0) u2[src] += k.vt2[src] * wav[it - 1];
    habc_update(g, dt, dx, dz, k.vel.data(), ow.data(), u2, u1, u0);
    blend(g, u2, ow.data(), k.w.data());
    if (it > ndel)
      for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
          const T u = u2[g.at(ix + pml, iz + pml)];
          ss[(size_t)ix * nz + iz] += u * u;
        }
    T* t = u0; u0 = u1; u1 = u2; u2 = t;
  }
}

// Born (linearized) modeling, fd.cpp:81-177: the background field drives a
// scattered field through vt2 * lap(u) * 2/v^3 * dv, lap taken from the
// background's three latest levels; the scattered field is recorded.
template <typename T>
void born(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, const int* gz,
          const T* wav, const T* v, const T* dv, T* rec) {
  Geom g(nx, nz, pml);
  Coefs<T> k(g, dt, dx, dz, v, true);
  std::vector<T> dvp(g.cells());
  pad_perturbation(g, dv, dvp.data());
  const T idt2 = T(1) / dt / dt;
  std::vector<T> a(g.cells()), b(g.cells()), c(g.cells()), ow(g.cells());
  std::vector<T> ab(g.cells()), bb(g.cells()), cb(g.cells()), owb(g.cells()), lap(g.cells());
  T *u0 = a.data(), *u1 = b.data(), *u2 = c.data();
  T *s0 = ab.data(), *s1 = bb.data(), *s2 = cb.data();
  const size_t src = g.at(sx + pml, sz + pml);
  for (size_t i = 0; i < (size_t)nx * nt; i++) rec[i] = T(0);
  const int NT = nt + ndel;
  for (int it = 0; it < NT; it++) {
    stencil_step(g, u0, u1, u2, k.vvzz.data(), k.vvxx.data());
    if (it - 1 >= 0) u2[src] += k.vt2[src] * wav[it - 1];
    habc_update(g, dt, dx, dz, k.vel.data(), ow.data(), u2, u1, u0);
    blend(g, u2, ow.data(), k.w.data());
    for (size_t i = 0; i < g.cells(); i++) lap[i] = (u2[i] + u0[i] - T(2) * u1[i]) * idt2;
    stencil_step(g, s0, s1, s2, k.vvzz.data(), k.vvxx.data());
    for (size_t i = 0; i < g.cells(); i++) s2[i] += k.vt2[i] * lap[i] * k.v3[i] * dvp[i];
    habc_update(g, dt, dx, dz, k.vel.data(), owb.data(), s2, s1, s0);
    blend(g, s2, owb.data(), k.w.data());
    const int itp = it - ndel;
    if (itp >= 0)
      for (int ix = 0; ix < nx; ix++) rec[(size_t)ix * nt + itp] = s2[g.at(ix + pml, gz[ix] + pml)];
    T* t = u0; u0 = u1; u1 = u2; u2 = t;
    t = s0; s0 = s1; s1 = s2; s2 = t;
  }
}

// RTM in the reference's semantics, fd.cpp:592-719 (adjoint):
//  1. forward pass saving `layer` boundary strips of the source field,
//  2. all fields reset to zero,
//  3. backward pass: subtract the source (wav[it-1] when it-1 > 0) from the
//     previous level, restore the strips into the current level, step the
//     source field backward; back-propagate the record injected at the
//     receivers; image += lap*2/v^3 * R over the padded grid, plus the
//     Hilbert term 0.5*(U R - H[U] H[R]) over the interior (FaQi).
// img is [nx][nz].
// `isolate`: the engine's SEIS_RTM_ISOLATE_BAND variant (not in the
// reference), which holds the reconstructed source field at zero outside the
// model area: each of its backward levels is zeroed there once it is formed.
template <typename T>
void rtm(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, const int* gz,
         const T* wav, const T* v, const T* rec, T* img, bool isolate = false) {
  const int layer = 4, half = 30;
  Geom g(nx, nz, pml);
  Coefs<T> k(g, dt, dx, dz, v, true);
  const T idt2 = T(1) / dt / dt;
  std::vector<T> a(g.cells()), b(g.cells()), c(g.cells()), ow(g.cells());
  std::vector<T> ab(g.cells()), bb(g.cells()), cb(g.cells()), owb(g.cells()), lap(g.cells()),
      imgp(g.cells());
  Strips<T> st(layer, nt, nx, nz);
  T *u0 = a.data(), *u1 = b.data(), *u2 = c.data();
  T *r0 = ab.data(), *r1 = bb.data(), *r2 = cb.data();
  const size_t src = g.at(sx + pml, sz + pml);
  const int NT = nt + ndel;
  for (int it = 0; it < NT; it++) {
    stencil_step(g, u0, u1, u2, k.vvzz.data(), k.vvxx.data());
    if (it - 1 >= 0) u2[src] += k.vt2[src] * wav[it - 1];
    habc_update(g, dt, dx, dz, k.vel.data(), ow.data(), u2, u1, u0);
    blend(g, u2, ow.data(), k.w.data());
    if (it - ndel >= 0) st.move(g, u2, it - ndel, true);
    T* t = u0; u0 = u1; u1 = u2; u2 = t;
  }
  std::fill(a.begin(), a.end(), T(0));
  std::fill(b.begin(), b.end(), T(0));
  std::fill(c.begin(), c.end(), T(0));
  std::vector<T> h(2 * half + 1), uh(g.NZ), rh(g.NZ);
  hilbert_taps(half, h.data());
  const T scal = T(2) / T(4);
  for (int it = NT - 1; it >= ndel; it--) {
    if (it - 1 > 0) u2[src] -= k.vt2[src] * wav[it - 1];
    const int itp = it - ndel;
    st.move(g, u1, itp, false);
    // source field one step backward: new level in u0, previous level u2
    stencil_step(g, u2, u1, u0, k.vvzz.data(), k.vvxx.data());
    habc_update(g, dt, dx, dz, k.vel.data(), ow.data(), u0, u1, u2);
    blend(g, u0, ow.data(), k.w.data());
    if (isolate)
      for (int ix = 0; ix < g.NX; ix++)
        for (int iz = 0; iz < g.NZ; iz++)
          if (ix < pml || ix >= pml + nx || iz < pml || iz >= pml + nz) u0[g.at(ix, iz)] = T(0);
    for (size_t i = 0; i < g.cells(); i++) lap[i] = (u2[i] + u0[i] - T(2) * u1[i]) * idt2;
    // receiver field
    stencil_step(g, r2, r1, r0, k.vvzz.data(), k.vvxx.data());
    for (int ix = 0; ix < nx; ix++) {
      size_t cc = g.at(ix + pml, gz[ix] + pml);
      r0[cc] += rec[(size_t)ix * nt + itp] * k.vt2[cc];
    }
    habc_update(g, dt, dx, dz, k.vel.data(), owb.data(), r0, r1, r2);
    blend(g, r0, owb.data(), k.w.data());
    for (size_t i = 0; i < g.cells(); i++) imgp[i] += lap[i] * k.v3[i] * r0[i];
    for (size_t i = 0; i < g.cells(); i++) lap[i] = lap[i] * k.v3[i];
    // FaQi (fd.cpp:1272-1298): per interior column, over the full padded column
    set_omp();
#pragma omp parallel
    {
      std::vector<T> uhl(g.NZ), rhl(g.NZ);
#pragma omp for
      for (int ix = pml; ix < g.NX - pml; ix++) {
        const T* U = lap.data() + (size_t)ix * g.NZ;
        const T* R = r0 + (size_t)ix * g.NZ;
        hilbert_apply(h.data(), half, U, uhl.data(), g.NZ);
        hilbert_apply(h.data(), half, R, rhl.data(), g.NZ);
        T* out = imgp.data() + (size_t)ix * g.NZ;
        for (int iz = pml; iz < g.NZ - pml; iz++) out[iz] += scal * (U[iz] * R[iz] - uhl[iz] * rhl[iz]);
      }
    }
    T* t = u2; u2 = u1; u1 = u0; u0 = t;
    t = r2; r2 = r1; r1 = r0; r0 = t;
  }
  for (int ix = 0; ix < nx; ix++)
    for (int iz = 0; iz < nz; iz++) img[(size_t)ix * nz + iz] = imgp[g.at(ix + pml, iz + pml)];
}


// ---------------------------------------------------------------------------
// Plane-wave engine (fd.h:6-7, 11).  The source is a line at row sz: every
// model column ix gets vt2 * sou[ix][it-1] (sou is [nx][nt+ndel]); receivers
// sit at row gz in every column.

template <typename T>
void add_line_source(const Geom& g, const Coefs<T>& k, T* u, int sz, const T* sou, int NT, int it) {
  if (it - 1 < 0) return;
  for (int ix = 0; ix < g.nx; ix++) {
    const size_t c = g.at(ix + g.pml, sz + g.pml);
    u[c] += k.vt2[c] * sou[(size_t)ix * NT + it - 1];
  }
}

// forwordPlane, fd.cpp:178-274: Born modeling driven by the line source.
template <typename T>
void born_plane(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz, const T* v,
                const T* sou, const T* dv, T* rec) {
  Geom g(nx, nz, pml);
  Coefs<T> k(g, dt, dx, dz, v, true);
  std::vector<T> dvp(g.cells());
  pad_perturbation(g, dv, dvp.data());
  const T idt2 = T(1) / dt / dt;
  std::vector<T> a(g.cells()), b(g.cells()), c(g.cells()), ow(g.cells());
  std::vector<T> ab(g.cells()), bb(g.cells()), cb(g.cells()), owb(g.cells()), lap(g.cells());
  T *u0 = a.data(), *u1 = b.data(), *u2 = c.data();
  T *s0 = ab.data(), *s1 = bb.data(), *s2 = cb.data();
  for (size_t i = 0; i < (size_t)nx * nt; i++) rec[i] = T(0);
  const int NT = nt + ndel;
  for (int it = 0; it < NT; it++) {
    stencil_step(g, u0, u1, u2, k.vvzz.data(), k.vvxx.data());
    add_line_source(g, k, u2, sz, sou, NT, it);
    habc_update(g, dt, dx, dz, k.vel.data(), ow.data(), u2, u1, u0);
    blend(g, u2, ow.data(), k.w.data());
    for (size_t i = 0; i < g.cells(); i++) lap[i] = (u2[i] + u0[i] - T(2) * u1[i]) * idt2;
    stencil_step(g, s0, s1, s2, k.vvzz.data(), k.vvxx.data());
    for (size_t i = 0; i < g.cells(); i++) s2[i] += k.vt2[i] * lap[i] * k.v3[i] * dvp[i];
    habc_update(g, dt, dx, dz, k.vel.data(), owb.data(), s2, s1, s0);
    blend(g, s2, owb.data(), k.w.data());
    const int itp = it - ndel;
    if (itp >= 0)
      for (int ix = 0; ix < nx; ix++) rec[(size_t)ix * nt + itp] = s2[g.at(ix + pml, gz + pml)];
    T* t = u0; u0 = u1; u1 = u2; u2 = t;
    t = s0; s0 = s1; s1 = s2; s2 = t;
  }
}

// illumPlane, fd.cpp:720-786: illum with the line source.
template <typename T>
void illum_plane(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sz, const T* v, const T* sou,
                 T* ss) {
  Geom g(nx, nz, pml);
  Coefs<T> k(g, dt, dx, dz, v, false);
  std::vector<T> a(g.cells()), b(g.cells()), c(g.cells()), ow(g.cells());
  T *u0 = a.data(), *u1 = b.data(), *u2 = c.data();
  for (size_t i = 0; i < (size_t)nx * nz; i++) ss[i] = T(0);
  const int NT = nt + ndel;
  for (int it = 0; it < NT; it++) {
    stencil_step(g, u0, u1, u2, k.vvzz.data(), k.vvxx.data());
    add_line_source(g, k, u2, sz, sou, NT, it);
    habc_update(g, dt, dx, dz, k.vel.data(), ow.data(), u2, u1, u0);
    blend(g, u2, ow.data(), k.w.data());
    if (it > ndel)
      for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
          const T u = u2[g.at(ix + pml, iz + pml)];
          ss[(size_t)ix * nz + iz] += u * u;
        }
    T* t = u0; u0 = u1; u1 = u2; u2 = t;
  }
}

// adjointPlane, fd.cpp:787-921: the receiver field is propagated backward in
// time first (record injected at row gz; the 4-layer strips of each step's
// current level saved at that step), then rebuilt forward in time from its
// last two levels (the record subtracted from the previous level at row gz,
// the strips restored into each new current level) beside the line-source
// field, imaging 1/2 (U R - H[U] H[R]) (FaQi alone) with U = lap 2/v^3.  The
// reference's two `if (itp>=0);` guards end at their semicolons, so it also
// injects samples before the record (reads outside rec) when ndel > 0; those
// samples are skipped here, which is the same for ndel = 0.  img is [nx][nz].
template <typename T>
void adjoint_plane(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz, const T* v,
                   const T* sou, const T* rec, T* img) {
  const int layer = 4, half = 30;
  Geom g(nx, nz, pml);
  Coefs<T> k(g, dt, dx, dz, v, true);
  const T idt2 = T(1) / dt / dt;
  const int NT = nt + ndel;
  Strips<T> st(layer, NT, nx, nz);
  std::vector<T> a(g.cells()), b(g.cells()), c(g.cells()), ow(g.cells());
  std::vector<T> ab(g.cells()), bb(g.cells()), cb(g.cells()), owb(g.cells()), lap(g.cells()),
      imgp(g.cells());
  T *u0 = a.data(), *u1 = b.data(), *u2 = c.data();
  T *r0 = ab.data(), *r1 = bb.data(), *r2 = cb.data();
  auto inject = [&](T* r, int itp, T sign) {
    if (itp < 0) return;
    for (int ix = 0; ix < nx; ix++) {
      const size_t cc = g.at(ix + pml, gz + pml);
      if (sign > 0)
        r[cc] += rec[(size_t)ix * nt + itp] * k.vt2[cc];
      else
        r[cc] -= rec[(size_t)ix * nt + itp] * k.vt2[cc];
    }
  };
  for (int it = NT - 1; it >= 0; it--) {
    stencil_step(g, r2, r1, r0, k.vvzz.data(), k.vvxx.data());
    inject(r0, it - ndel, T(1));
    habc_update(g, dt, dx, dz, k.vel.data(), owb.data(), r0, r1, r2);
    blend(g, r0, owb.data(), k.w.data());
    st.move(g, r1, it, true);
    T* t = r2; r2 = r1; r1 = r0; r0 = t;
  }
  {
    T* t = r0; r0 = r1; r1 = r2; r2 = t;
    std::fill(r2, r2 + g.cells(), T(0));
  }
  std::vector<T> h(2 * half + 1);
  hilbert_taps(half, h.data());
  const T scal = T(2) / T(4);
  for (int it = 2; it < NT; it++) {
    stencil_step(g, u0, u1, u2, k.vvzz.data(), k.vvxx.data());
    add_line_source(g, k, u2, sz, sou, NT, it);
    habc_update(g, dt, dx, dz, k.vel.data(), ow.data(), u2, u1, u0);
    blend(g, u2, ow.data(), k.w.data());
    for (size_t i = 0; i < g.cells(); i++) lap[i] = (u2[i] + u0[i] - T(2) * u1[i]) * idt2;
    inject(r0, it - ndel - 1, T(-1));
    stencil_step(g, r0, r1, r2, k.vvzz.data(), k.vvxx.data());
    habc_update(g, dt, dx, dz, k.vel.data(), owb.data(), r2, r1, r0);
    blend(g, r2, owb.data(), k.w.data());
    if (it - ndel > 0) {
      for (size_t i = 0; i < g.cells(); i++) lap[i] = lap[i] * k.v3[i];
      set_omp();
#pragma omp parallel
      {
        std::vector<T> uhl(g.NZ), rhl(g.NZ);
#pragma omp for
        for (int ix = pml; ix < g.NX - pml; ix++) {
          const T* U = lap.data() + (size_t)ix * g.NZ;
          const T* R = r2 + (size_t)ix * g.NZ;
          hilbert_apply(h.data(), half, U, uhl.data(), g.NZ);
          hilbert_apply(h.data(), half, R, rhl.data(), g.NZ);
          T* out = imgp.data() + (size_t)ix * g.NZ;
          for (int iz = pml; iz < g.NZ - pml; iz++) out[iz] += scal * (U[iz] * R[iz] - uhl[iz] * rhl[iz]);
        }
      }
    }
    T* t = u0; u0 = u1; u1 = u2; u2 = t;
    t = r0; r0 = r1; r1 = r2; r2 = t;
    st.move(g, r1, it, false);
  }
  for (int ix = 0; ix < nx; ix++)
    for (int iz = 0; iz < nz; iz++) img[(size_t)ix * nz + iz] = imgp[g.at(ix + pml, iz + pml)];
}

// ---------------------------------------------------------------------------
// Exact transpose of the Born operator.  The Born map of born() is linear in
// dv: the scattered field s obeys s^{n+1} = Step(s^n, s^{n-1}) + injection of
// q^n = vt2 * lap^n * 2/v^3 * dv before the absorbing boundary, and the record
// reads s^{n+1} at the receivers.  Its transpose is obtained by running every
// statement of the step backwards in scatter form (y = a x  ->  x_adj += a y_adj).
// This has no counterpart in the reference (its `adjoint` is the imaging
// operator of fd.cpp:592-719, not a transpose); the dot-product test pins it.

// Transpose of stencil_step: a_raw on non-ring cells -> a_cur, a_prv.
template <typename T>
void stencil_step_adj(const Geom& g, const T* a_raw, T* a_cur, T* a_prv, const T* vvzz, const T* vvxx) {
  const int NX = g.NX, NZ = g.NZ;
  const T k0 = T(-205) / T(72), k1 = T(8) / T(5), k2 = T(-1) / T(5), k3 = T(8) / T(315), k4 = T(-1) / T(560);
  const T cf[9] = {k4, k3, k2, k1, k0, k1, k2, k3, k4};
  for (int ix = 1; ix <= NX - 2; ix++)
    for (int iz = 1; iz <= NZ - 2; iz++) {
      const size_t c = (size_t)ix * NZ + iz;
      const T a = a_raw[c];
      a_prv[c] -= a;
      a_cur[c] += T(2) * a;
      const T az = a * vvzz[c], ax = a * vvxx[c];
      if (ix >= 4 && ix <= NX - 5 && iz >= 4 && iz <= NZ - 5) {
        for (int k = 0; k < 9; k++) {
          a_cur[c + (k - 4)] += az * cf[k];
          a_cur[c + (ptrdiff_t)(k - 4) * NZ] += ax * cf[k];
        }
      } else {
        a_cur[c - 1] += az;
        a_cur[c + 1] += az;
        a_cur[c] -= T(2) * az;
        a_cur[c + NZ] += ax;
        a_cur[c - NZ] += ax;
        a_cur[c] -= T(2) * ax;
      }
    }
}

// Transpose of habc_update.  a_ow holds the adjoint of the one-way values of
// the band cells; a_raw the adjoint of the stencil/ring values.  On return the
// one-way and ring dependencies are moved into a_raw (in-neighbours), a_cur
// and a_prv, and a_raw on the ring is cleared (the ring overwrites the stencil).
template <typename T>
void habc_update_adj(const Geom& g, T dt, T dx, T dz, const T* vel, const T* a_ow, T* a_raw, T* a_cur, T* a_prv) {
  const int NX = g.NX, NZ = g.NZ, pml = g.pml, nx = g.nx, nz = g.nz;
  auto A = [&](int ix, int iz) -> size_t { return (size_t)ix * NZ + iz; };
  // band, one-way layers
  for (int k = 1; k <= pml; k++) {
    const int lo = pml + 1 - k;
    for (int side = 0; side < 2; side++) {
      const int K = side == 0 ? pml - k : nz + k + pml - 1;
      const int s = side == 0 ? 1 : -1;
      const int hi = pml + nx + k - 2;
      for (int ix = lo; ix <= hi; ix++) {
        T v = vel[A(ix, K)];
        T a1 = T(1) / dz / T(2) / dt, a2 = T(1) / T(2) / dt / dt / v, a3 = v / T(4) / dx / dx;
        T at = -a_ow[A(ix, K)] / (a1 + a2);
        a_raw[A(ix, K + s)] += at * (-a1 + a2 + T(2) * a3);
        a_raw[A(ix + 1, K + s)] -= at * a3;
        a_raw[A(ix - 1, K + s)] -= at * a3;
        a_prv[A(ix, K)] += at * (-a1 + a2 + T(2) * a3);
        a_prv[A(ix, K + s)] += at * (a1 + a2);
        a_prv[A(ix + 1, K)] -= at * a3;
        a_prv[A(ix - 1, K)] -= at * a3;
        a_cur[A(ix, K)] -= T(2) * a2 * at;
        a_cur[A(ix, K + s)] -= T(2) * a2 * at;
      }
    }
    for (int side = 0; side < 2; side++) {
      const int K = side == 0 ? nx + k + pml - 1 : pml - k;
      const int s = side == 0 ? -1 : 1;
      const int hi = pml + nz + k - 2;
      for (int iz = lo; iz <= hi; iz++) {
        T v = vel[A(K, iz)];
        T a1 = T(1) / dx / T(2) / dt, a2 = T(1) / T(2) / dt / dt / v, a3 = v / T(4) / dz / dz;
        T at = -a_ow[A(K, iz)] / (a1 + a2);
        a_raw[A(K + s, iz)] += at * (-a1 + a2 + T(2) * a3);
        a_raw[A(K + s, iz + 1)] -= at * a3;
        a_raw[A(K + s, iz - 1)] -= at * a3;
        a_prv[A(K, iz)] += at * (-a1 + a2 + T(2) * a3);
        a_prv[A(K + s, iz)] += at * (a1 + a2);
        a_prv[A(K, iz + 1)] -= at * a3;
        a_prv[A(K, iz - 1)] -= at * a3;
        a_cur[A(K, iz)] -= T(2) * a2 * at;
        a_cur[A(K + s, iz)] -= T(2) * a2 * at;
      }
    }
    const T b1 = T(1) / dz, b2 = T(1) / dx;
    struct Corner { int xc, zc, sx, sz; bool dbl; };
    const Corner cs[4] = {
        {nx + pml + k - 1, nz + pml + k - 1, -1, -1, false},
        {pml - k, nz + pml + k - 1, +1, -1, false},
        {nx + pml + k - 1, pml - k, -1, +1, false},
        {pml - k, pml - k, +1, +1, true},
    };
    for (const Corner& q : cs) {
      T v = vel[A(q.xc, q.zc)];
      T b3 = q.dbl ? (T)(std::sqrt(2.0) / (double)v / (double)dt) : std::sqrt(T(2)) / v / dt;
      T at = -a_ow[A(q.xc, q.zc)] / (b1 + b2 + b3);
      a_raw[A(q.xc, q.zc + q.sz)] -= b1 * at;
      a_raw[A(q.xc + q.sx, q.zc)] -= b2 * at;
      a_cur[A(q.xc, q.zc)] -= b3 * at;
    }
  }
  const T dt2 = dt * dt;
  const T cx = (T(1) / dx) * (T(1) / dx) * dt2;
  const T cz = (T(1) / dz) * (T(1) / dz) * dt2;
  auto vv = [&](int ix, int iz) { T v = vel[A(ix, iz)]; return v * v; };
  // x extrapolation (applied last in the forward ring update)
  for (int iz = 0; iz < NZ; iz++) {
    for (int e = 0; e < 2; e++) {
      const int x0 = e == 0 ? 0 : NX - 1, d = e == 0 ? 1 : -1;
      const T a = a_raw[A(x0, iz)];
      const T e1 = T(2) * a * vv(x0 + d, iz) * cx;
      a_cur[A(x0 + 2 * d, iz)] += e1;
      a_cur[A(x0, iz)] += e1;
      a_cur[A(x0 + d, iz)] -= T(2) * e1;
      const T e2 = -a * vv(x0 + 2 * d, iz) * cx;
      a_cur[A(x0 + 3 * d, iz)] += e2;
      a_cur[A(x0 + d, iz)] += e2;
      a_cur[A(x0 + 2 * d, iz)] -= T(2) * e2;
    }
  }
  // z extrapolation
  for (int ix = 0; ix < NX; ix++) {
    for (int e = 0; e < 2; e++) {
      const int z0 = e == 0 ? 0 : NZ - 1, d = e == 0 ? 1 : -1;
      const T a = a_raw[A(ix, z0)];
      const T e1 = T(2) * a * vv(ix, z0 + d) * cz;
      a_cur[A(ix, z0 + 2 * d)] += e1;
      a_cur[A(ix, z0)] += e1;
      a_cur[A(ix, z0 + d)] -= T(2) * e1;
      const T e2 = -a * vv(ix, z0 + 2 * d) * cz;
      a_cur[A(ix, z0 + 3 * d)] += e2;
      a_cur[A(ix, z0 + d)] += e2;
      a_cur[A(ix, z0 + 2 * d)] -= T(2) * e2;
    }
  }
  // base values of the ring (assignments)
  for (int ix = 0; ix < NX; ix++)
    for (int iz = 0; iz < NZ; iz++) {
      const bool xe = ix == 0 || ix == NX - 1, ze = iz == 0 || iz == NZ - 1;
      if (!xe && !ze) continue;
      const T a = a_raw[A(ix, iz)];
      a_prv[A(ix, iz)] -= a;
      a_cur[A(ix, iz)] += T(2) * a;
      if (xe && ze) continue;
      if (ze) {
        const T e = a * vv(ix, iz) * cx;
        a_cur[A(ix + 1, iz)] += e;
        a_cur[A(ix - 1, iz)] += e;
        a_cur[A(ix, iz)] -= T(2) * e;
      } else {
        const T e = a * vv(ix, iz) * (ix == 0 ? cz : cx);
        a_cur[A(ix, iz + 1)] += e;
        a_cur[A(ix, iz - 1)] += e;
        a_cur[A(ix, iz)] -= T(2) * e;
      }
    }
  for (int ix = 0; ix < NX; ix++) { a_raw[A(ix, 0)] = T(0); a_raw[A(ix, NZ - 1)] = T(0); }
  for (int iz = 0; iz < NZ; iz++) { a_raw[A(0, iz)] = T(0); a_raw[A(NX - 1, iz)] = T(0); }
}

// The per-step Born source coefficient vt2 * lap^n * 2/v^3 of the background
// (fd.cpp:134-139), for n = 0..NT-1, in [NT][NX][NZ].
template <typename T>
void born_source_history(const Geom& g, T dt, T dx, T dz, const Coefs<T>& k, int ndel, int nt, int sx, int sz,
                         const T* wav, std::vector<T>& bs) {
  const size_t C = g.cells();
  const int NT = nt + ndel;
  bs.assign((size_t)NT * C, T(0));
  const T idt2 = T(1) / dt / dt;
  std::vector<T> a(C), b(C), c(C), ow(C);
  T *u0 = a.data(), *u1 = b.data(), *u2 = c.data();
  const size_t src = g.at(sx + g.pml, sz + g.pml);
  for (int it = 0; it < NT; it++) {
    stencil_step(g, u0, u1, u2, k.vvzz.data(), k.vvxx.data());
    if (it - 1 >= 0) u2[src] += k.vt2[src] * wav[it - 1];
    habc_update(g, dt, dx, dz, k.vel.data(), ow.data(), u2, u1, u0);
    blend(g, u2, ow.data(), k.w.data());
    T* q = bs.data() + (size_t)it * C;
    for (size_t i = 0; i < C; i++) q[i] = k.vt2[i] * ((u2[i] + u0[i] - T(2) * u1[i]) * idt2) * k.v3[i];
    T* t = u0; u0 = u1; u1 = u2; u2 = t;
  }
}

// grad[nx][nz] = L^T rec for the Born operator L of born().
template <typename T>
void born_adjoint(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, const int* gz,
                  const T* wav, const T* v, const T* rec, T* grad) {
  Geom g(nx, nz, pml);
  Coefs<T> k(g, dt, dx, dz, v, true);
  const size_t C = g.cells();
  const int NT = nt + ndel;
  std::vector<T> bs;
  born_source_history(g, dt, dx, dz, k, ndel, nt, sx, sz, wav, bs);
  std::vector<T> A0(C), A1(C), A2(C), a_raw(C), a_ow(C), gp(C);
  T *an1 = A0.data(), *an = A1.data(), *anm1 = A2.data();
  for (int n = NT - 1; n >= 0; n--) {
    if (n - ndel >= 0)
      for (int ix = 0; ix < nx; ix++) an1[g.at(ix + pml, gz[ix] + pml)] += rec[(size_t)ix * nt + n - ndel];
    for (size_t i = 0; i < C; i++) {
      a_raw[i] = an1[i] * k.w[i];
      a_ow[i] = an1[i] * (T(1) - k.w[i]);
    }
    habc_update_adj(g, dt, dx, dz, k.vel.data(), a_ow.data(), a_raw.data(), an, anm1);
    const T* q = bs.data() + (size_t)n * C;
    for (int ix = pml; ix < pml + nx; ix++)
      for (int iz = pml; iz < pml + nz; iz++) gp[g.at(ix, iz)] += a_raw[g.at(ix, iz)] * q[g.at(ix, iz)];
    stencil_step_adj(g, a_raw.data(), an, anm1, k.vvzz.data(), k.vvxx.data());
    std::fill(an1, an1 + C, T(0));
    T* t = an1; an1 = an; an = anm1; anm1 = t;
  }
  for (int ix = 0; ix < nx; ix++)
    for (int iz = 0; iz < nz; iz++) grad[(size_t)ix * nz + iz] = gp[g.at(ix + pml, iz + pml)];
}

// Cross-correlation adjoint of born(): the receiver field run backwards in
// time with the forward step (stencil, record injected as rec * vt2 after the
// stencil as fd.cpp:673-677 injects it, HABC, blend) and grad += u S^n / vt2
// over the model cells, S^n the Born source coefficient of step n.  In the
// interior u is vt2 times the field of born_adjoint, so the two agree where
// no wave reaches the absorbing band.
template <typename T>
void born_adjoint_xcorr(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, const int* gz,
                        const T* wav, const T* v, const T* rec, T* grad) {
  Geom g(nx, nz, pml);
  Coefs<T> k(g, dt, dx, dz, v, true);
  const size_t C = g.cells();
  const int NT = nt + ndel;
  std::vector<T> bs;
  born_source_history(g, dt, dx, dz, k, ndel, nt, sx, sz, wav, bs);
  std::vector<T> a(C), b(C), c(C), ow(C), gp(C);
  T *u0 = a.data(), *u1 = b.data(), *u2 = c.data();
  for (int n = NT - 1; n >= 0; n--) {
    stencil_step(g, u0, u1, u2, k.vvzz.data(), k.vvxx.data());
    if (n - ndel >= 0)
      for (int ix = 0; ix < nx; ix++) {
        const size_t r = g.at(ix + pml, gz[ix] + pml);
        u2[r] = u2[r] + rec[(size_t)ix * nt + n - ndel] * k.vt2[r];
      }
    habc_update(g, dt, dx, dz, k.vel.data(), ow.data(), u2, u1, u0);
    blend(g, u2, ow.data(), k.w.data());
    const T* q = bs.data() + (size_t)n * C;
    for (int ix = pml; ix < pml + nx; ix++)
      for (int iz = pml; iz < pml + nz; iz++) {
        const size_t i = g.at(ix, iz);
        gp[i] = gp[i] + u2[i] * (q[i] / k.vt2[i]);
      }
    T* t = u0; u0 = u1; u1 = u2; u2 = t;
  }
  for (int ix = 0; ix < nx; ix++)
    for (int iz = 0; iz < nz; iz++) grad[(size_t)ix * nz + iz] = gp[g.at(ix + pml, iz + pml)];
}

}  // namespace

extern "C" {

void oracle_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int oracle_get_threads(void) { return g_threads; }

#define ORACLE_TYPED(SUF, T)                                                                          \
  void oracle_ricker_##SUF(T dt, T fr, int num_delay, int n, T* w) { ricker<T>(dt, fr, num_delay, n, w); } \
  void oracle_hilbert_taps_##SUF(int half, T* h) { hilbert_taps<T>(half, h); }                        \
  void oracle_hilbert_apply_##SUF(int half, const T* in, T* out, int n) {                             \
    std::vector<T> h(2 * half + 1);                                                                   \
    hilbert_taps<T>(half, h.data());                                                                  \
    hilbert_apply<T>(h.data(), half, in, out, n);                                                     \
  }                                                                                                   \
  void oracle_harmonic_##SUF(const T* in, T* out, int n, T kx, T x) { harmonic<T>(in, out, n, kx, x); }  \
  void oracle_plane_prepare_##SUF(int nx, int nt, int ntw, int ns, int np, T pmin, T dp, const T* sx,        \
                                  const int* isx, const int* ng, const int* col, const T* wav,              \
                                  const T* data, T* sou, T* rec) {                                          \
    plane_prepare<T>(nx, nt, ntw, ns, np, pmin, dp, sx, isx, ng, col, wav, data, sou, rec);                \
  }                                                                                                   \
  void oracle_sponge_weights_##SUF(int nz, int nx, int pml, T* w) { sponge_weights<T>(Geom(nx, nz, pml), w); } \
  void oracle_pad_velocity_##SUF(int nz, int nx, int pml, const T* v, T* vp) {                        \
    pad_velocity<T>(Geom(nx, nz, pml), v, vp);                                                        \
  }                                                                                                   \
  void oracle_stencil_step_##SUF(int nz, int nx, int pml, T dt, T dx, T dz, const T* vpad,          \
                                 const T* prv, const T* cur, T* nxt) {                               \
    Geom g(nx, nz, pml);                                                                              \
    std::vector<T> vvzz(g.cells()), vvxx(g.cells());                                                  \
    const T idx2 = T(1) / (dx * dx), idz2 = T(1) / (dz * dz);                                         \
    for (size_t i = 0; i < g.cells(); i++) {                                                          \
      T a = dt * vpad[i];                                                                             \
      T vt2 = a * a;                                                                                  \
      vvzz[i] = vt2 * idz2;                                                                           \
      vvxx[i] = vt2 * idx2;                                                                           \
    }                                                                                                 \
    stencil_step<T>(g, prv, cur, nxt, vvzz.data(), vvxx.data());                                      \
  }                                                                                                   \
  int oracle_modeling_##SUF(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx,      \
                            int sz, const int* gz, const T* wav, const T* vel, T* rec) {             \
    modeling<T>(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, rec);                        \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_illum_##SUF(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, \
                         const T* wav, const T* vel, T* ss) {                                         \
    illum<T>(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, wav, vel, ss);                               \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_born_plane_##SUF(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sz,     \
                              int gz, const T* vel, const T* sou, const T* dv, T* rec) {              \
    born_plane<T>(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, vel, sou, dv, rec);                      \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_illum_plane_##SUF(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sz,    \
                               const T* vel, const T* sou, T* ss) {                                   \
    illum_plane<T>(dt, dx, dz, nt, ndel, nx, nz, pml, sz, vel, sou, ss);                              \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_adjoint_plane_##SUF(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sz,  \
                                 int gz, const T* vel, const T* sou, const T* rec, T* img) {          \
    adjoint_plane<T>(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, vel, sou, rec, img);                  \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_born_##SUF(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,  \
                        const int* gz, const T* wav, const T* vel, const T* dv, T* rec) {            \
    born<T>(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, dv, rec);                        \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_rtm_##SUF(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,   \
                       const int* gz, const T* wav, const T* vel, const T* rec, T* img) {            \
    rtm<T>(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, rec, img);                        \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_rtm_isolated_##SUF(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx,  \
                                int sz, const int* gz, const T* wav, const T* vel, const T* rec,      \
                                T* img) {                                                             \
    rtm<T>(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, rec, img, true);                  \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_born_adjoint_xcorr_##SUF(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx, \
                                      int sz, const int* gz, const T* wav, const T* vel, const T* rec,     \
                                      T* grad) {                                                        \
    born_adjoint_xcorr<T>(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, rec, grad);          \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_born_adjoint_##SUF(T dt, T dx, T dz, int nt, int ndel, int nx, int nz, int pml, int sx,  \
                                int sz, const int* gz, const T* wav, const T* vel, const T* rec,     \
                                T* grad) {                                                            \
    born_adjoint<T>(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, rec, grad);              \
    return 0;                                                                                         \
  }

ORACLE_TYPED(f32, float)
ORACLE_TYPED(f64, double)

}  // extern "C"
