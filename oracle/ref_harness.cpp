// C-callable harness around the compiled reference engine.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) against
// the reference's own sources, which stay where they lie under the reference
// tree; the result lands in oracle/_ref/ and never enters version control.
// It exists to generate golden vectors (tests/golden/make_golden.py) and to
// serve as the timed "reference" CPU baseline.  It converts flat arrays
// to the reference's row-pointer tables (memory.h:29-44), calls
// modeling / forward / adjoint / illum and the plane-wave engine (fd.h:3-11)
// and copies the newly allocated
// result back, freeing it the reference's way (memory.h:103-110).

#include <cstring>
#include <cstddef>
#include <math.h>  // hilbert.h uses cos() without including it (fd.cpp includes math.h first)
#include "memory.h"
#include "fd.h"
#include "wavelet.h"
#include "global.h"
#include "InitInversion.h"
#include "../filter/hilbert.h"
#include "io.h"
#include "parser.h"
#include "coord.h"
#include "../lsrtm/datastr.h"
#include <cstdio>
#include <string>

namespace {

// View a flat [n2][n1] block as the reference's float** table.
struct Rows {
  float** p;
  Rows(const float* base, size_t n1, size_t n2) {
    p = new float*[n2];
    for (size_t i = 0; i < n2; i++) p[i] = const_cast<float*>(base) + i * n1;
  }
  ~Rows() { delete[] p; }
};

void take(float** out, float* dst, size_t n) {
  if (out) {
    std::memcpy(dst, out[0], n * sizeof(float));
    MyAlloc<float>::free(out);
  }
}

}  // namespace

extern "C" {

void ref_set_threads(int n) { OMP_CORE = n > 0 ? n : 1; }

void ref_ricker(float dt, float fr, int num_delay, int n, float* out) {
  float* w = rickerWavelet(dt, fr, num_delay, n);
  std::memcpy(out, w, sizeof(float) * n);
  MyAlloc<float>::free(w);
}

void ref_hilbert_apply(int half, const float* in, float* out, int n) {
  Hilbert<float> h(half);
  h.apply(const_cast<float*>(in), out, n);
}

// harmonic (InitInversion.cpp:141-151): cos(kx x) in - sin(kx x) H[in] with
// the 201-tap Hilbert filter, the plane-wave phase shift of planeWavePrepare.
void ref_harmonic(const float* in, float* out, int n, float kx, float x) {
  harmonic(const_cast<float*>(in), out, n, kx, x);
}

// writeSu (io.cpp:8-58): one trace of n1 samples (n2 == 0) or n2 traces.
void ref_write_su(const char* path, int n1, int n2, const float* data) {
  if (n2 == 0) {
    writeSu(std::string(path), n1, const_cast<float*>(data));
    return;
  }
  Rows r(data, n1, n2);
  writeSu(std::string(path), n1, n2, r.p);
}

void ref_set_abs(int nz, int nx, int pml, float* w) {
  take(setAbs(nz, nx, pml), w, (size_t)(nz + 2 * pml) * (nx + 2 * pml));
}

void ref_set_vel(int nz, int nx, int pml, const float* v, float* vp) {
  Rows r(v, nz, nx);
  take(setVel(r.p, nz, nx, pml), vp, (size_t)(nz + 2 * pml) * (nx + 2 * pml));
}

int ref_modeling(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
                 const int* gz, const float* wav, const float* vel, float* rec) {
  Rows v(vel, nz, nx);
  take(modeling(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, const_cast<int*>(gz), const_cast<float*>(wav), v.p),
       rec, (size_t)nt * nx);
  return 0;
}

int ref_illum(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
              const float* wav, const float* vel, float* ss) {
  Rows v(vel, nz, nx);
  take(illum(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, const_cast<float*>(wav), v.p), ss, (size_t)nx * nz);
  return 0;
}

// Plane-wave engine (fd.h:6-7, 11): sou is [nx][nt+ndel].
int ref_born_plane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                   const float* vel, const float* sou, const float* dv, float* rec) {
  Rows v(vel, nz, nx), s(sou, nt + ndel, nx), d(dv, nz, nx);
  take(forwordPlane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, v.p, s.p, d.p), rec, (size_t)nt * nx);
  return 0;
}

int ref_illum_plane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                    const float* vel, const float* sou, float* ss) {
  Rows v(vel, nz, nx), s(sou, nt + ndel, nx);
  take(illumPlane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, v.p, s.p), ss, (size_t)nx * nz);
  return 0;
}

int ref_adjoint_plane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                      const float* vel, const float* sou, const float* rec, float* img) {
  Rows v(vel, nz, nx), s(sou, nt + ndel, nx), r(rec, nt, nx);
  take(adjointPlane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, v.p, s.p, r.p), img, (size_t)nx * nz);
  return 0;
}

int ref_born(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
             const int* gz, const float* wav, const float* vel, const float* dv, float* rec) {
  Rows v(vel, nz, nx), d(dv, nz, nx);
  take(forward(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, const_cast<int*>(gz), const_cast<float*>(wav), v.p,
               d.p),
       rec, (size_t)nt * nx);
  return 0;
}

int ref_rtm(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
            const int* gz, const float* wav, const float* vel, const float* rec, float* img) {
  Rows v(vel, nz, nx), r(rec, nt, nx);
  take(adjoint(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, const_cast<int*>(gz), const_cast<float*>(wav), v.p,
               r.p),
       img, (size_t)nz * nx);
  return 0;
}

// Inversion::readParamFile's sequence of parser calls (inversion.cpp:38-79)
// on a paramLSRTM (datastr.h:51-75), printed as "key = value" lines in the
// order of the driver's seis_read_params (floats with 9 significant digits).
// Only files whose required keys are present: the parser exits otherwise.
void ref_read_params(const char* file, char* out, int len) {
  ParserFromFile pff{std::string(file)};
  paramLSRTM param;
  pff.getInt(param.nthread);
  pff.getInt(param.ns);
  pff.getInt(param.ngmax);
  pff.getInt(param.nx);
  pff.getInt(param.nz);
  pff.getInt(param.nt);
  pff.getInt(param.npml);
  pff.getInt(param.delay);
  pff.getInt(param.delaycal);
  pff.getInt(param.mask);
  pff.getInt(param.maxiter);
  pff.getInt(param.lpad);
  pff.getInt(param.rpad);
  pff.getInt(param.planeTag);
  pff.getInt(param.sbegin);
  pff.getInt(param.send);
  pff.getFloat(param.dx);
  pff.getFloat(param.dz);
  pff.getFloat(param.dt);
  pff.getFloat(param.fr);
  pff.getFloat(param.velfx);
  pff.getString(param.precsg);
  pff.getString(param.wdir);
  pff.getString(param.v0file);
  pff.getString(param.vfile);
  pff.getString(param.coordfile);
  if (param.mask.val) {
    param.maskfile.opt = 0;
    pff.getString(param.maskfile);
  }
  if (param.planeTag.val) {
    pff.getInt(param.np);
    pff.getFloat(param.dp);
    pff.getFloat(param.pmin);
  }
  std::string o;
  char b[512];
  for (paramInt* q : {&param.nthread, &param.ns, &param.ngmax, &param.nx, &param.nz, &param.nt, &param.npml,
                      &param.delay, &param.delaycal, &param.mask, &param.maxiter, &param.lpad, &param.rpad,
                      &param.planeTag, &param.sbegin, &param.send, &param.np}) {
    std::snprintf(b, sizeof(b), "%s = %d\n", q->key.c_str(), q->val);
    o += b;
  }
  for (paramFloat* q : {&param.dx, &param.dz, &param.dt, &param.fr, &param.velfx, &param.dp, &param.pmin}) {
    std::snprintf(b, sizeof(b), "%s = %.9g\n", q->key.c_str(), (double)q->val);
    o += b;
  }
  for (paramString* q : {&param.precsg, &param.wdir, &param.v0file, &param.vfile, &param.coordfile, &param.maskfile})
    o += q->key + " = " + q->val + "\n";
  std::snprintf(out, len, "%s", o.c_str());
}

// Coords (coord.cpp:22-176): shots, receivers per shot, source (x, z) per
// shot, receiver (x, z) per trace packed shot after shot.  Returns the number
// of shots; fills the arrays when they hold cap_shots / cap_traces entries.
int ref_coords(const char* file, int cap_shots, int cap_traces, int* ng, float* sxz, float* gxz) {
  const std::string f(file);
  const int ns = Coords::countShots(f);
  int* n = nullptr;
  Coords::countReceivers(f, ns, n);
  float** sc = nullptr;
  Coords::getSources(f, ns, sc);
  float*** gc = nullptr;
  Coords::getReceivers(f, ns, n, gc);
  int total = 0;
  for (int s = 0; s < ns; s++) total += n[s];
  if (ns <= cap_shots && total <= cap_traces) {
    int t = 0;
    for (int s = 0; s < ns; s++) {
      ng[s] = n[s];
      sxz[2 * s] = sc[s][0];
      sxz[2 * s + 1] = sc[s][1];
      for (int g = 0; g < n[s]; g++, t++) {
        gxz[2 * t] = gc[s][0][g];
        gxz[2 * t + 1] = gc[s][1][g];
      }
    }
  }
  MyAlloc<int>::free(n);
  MyAlloc<float>::free(sc);
  MyAlloc<float>::free(gc);
  return ns;
}

}  // extern "C"
