// fd.h-compatible entry points (include/seis_fd.h) over the C ABI.
#include "../../../include/seis_fd.h"

#include <cstdio>
#include <cstring>
#include <initializer_list>

#include "../../../include/seis.h"

namespace {

// n2 rows of n1 floats in the reference's two-block layout, zeroed.
float** alloc2(size_t n1, size_t n2) {
  if (n1 == 0 || n2 == 0) return nullptr;
  float** buf = new float*[n2];
  buf[0] = new float[n1 * n2];
  std::memset(buf[0], 0, sizeof(float) * n1 * n2);
  for (size_t i = 1; i < n2; i++) buf[i] = buf[0] + i * n1;
  return buf;
}

void free2(float** buf) {
  if (!buf) return;
  delete[] buf[0];
  delete[] buf;
}

// The reference's 2D inputs are row tables into one block; this gathers rows
// that may not be contiguous.
struct Flat {
  float* p = nullptr;
  bool own = false;
  Flat(float** a, size_t n1, size_t n2) {
    bool contiguous = true;
    for (size_t i = 1; i < n2 && contiguous; i++) contiguous = (a[i] == a[0] + i * n1);
    if (contiguous) {
      p = a[0];
    } else {
      p = new float[n1 * n2];
      own = true;
      for (size_t i = 0; i < n2; i++) std::memcpy(p + i * n1, a[i], sizeof(float) * n1);
    }
  }
  ~Flat() {
    if (own) delete[] p;
  }
};

// Sizes and row tables are checked before any input row is read or an output
// allocated: a negative size would become a huge size_t in Flat and alloc2.
bool bad_args(const char* fn, int nt, int ndel, int nx, int nz, std::initializer_list<float**> rows) {
  const char* why = nullptr;
  if (nt <= 0 || nx <= 0 || nz <= 0) why = "nt, nx and nz must be positive";
  else if (ndel < 0) why = "ndel must be non-negative";
  for (float** r : rows)
    if (!why && !r) why = "null input array";
  if (!why) return false;
  std::fprintf(stderr, "seis: %s: %s\n", fn, why);
  return true;
}

float** finish(int rc, float** out) {
  if (rc == SEIS_OK) return out;
  std::fprintf(stderr, "seis: %s\n", seis_last_error());
  free2(out);
  return nullptr;
}

}  // namespace

float** modeling(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, int* gz,
                 float* wav, float** vel) {
  if (bad_args("modeling", nt, ndel, nx, nz, {vel})) return nullptr;
  Flat v(vel, nz, nx);
  float** rec = alloc2(nt, nx);
  if (!rec) return nullptr;
  return finish(seis_modeling(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, v.p, rec[0]), rec);
}

float** forward(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, int* gz,
                float* wav, float** vel, float** dv) {
  if (bad_args("forward", nt, ndel, nx, nz, {vel, dv})) return nullptr;
  Flat v(vel, nz, nx), d(dv, nz, nx);
  float** rec = alloc2(nt, nx);
  if (!rec) return nullptr;
  return finish(seis_born(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, v.p, d.p, rec[0]), rec);
}

float** adjoint(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, int* gz,
                float* wav, float** vel, float** rec) {
  if (bad_args("adjoint", nt, ndel, nx, nz, {vel, rec})) return nullptr;
  Flat v(vel, nz, nx), r(rec, nt, nx);
  float** img = alloc2(nz, nx);
  if (!img) return nullptr;
  return finish(seis_rtm(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, v.p, r.p, img[0]), img);
}

float** illum(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, float* wav,
              float** vel) {
  if (bad_args("illum", nt, ndel, nx, nz, {vel})) return nullptr;
  Flat v(vel, nz, nx);
  float** ss = alloc2(nz, nx);
  if (!ss) return nullptr;
  return finish(seis_illum(dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, wav, v.p, ss[0]), ss);
}

float** forwordPlane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                     float** vel, float** sou, float** dv) {
  if (bad_args("forwordPlane", nt, ndel, nx, nz, {vel, sou, dv})) return nullptr;
  Flat v(vel, nz, nx), s(sou, nt + ndel, nx), d(dv, nz, nx);
  float** rec = alloc2(nt, nx);
  if (!rec) return nullptr;
  return finish(seis_born_plane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, v.p, s.p, d.p, rec[0]), rec);
}

float** adjointPlane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                     float** vel, float** sou, float** rec) {
  if (bad_args("adjointPlane", nt, ndel, nx, nz, {vel, sou, rec})) return nullptr;
  Flat v(vel, nz, nx), s(sou, nt + ndel, nx), r(rec, nt, nx);
  float** img = alloc2(nz, nx);
  if (!img) return nullptr;
  return finish(seis_rtm_plane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, v.p, s.p, r.p, img[0]), img);
}

float** illumPlane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                   float** vel, float** sou) {
  if (bad_args("illumPlane", nt, ndel, nx, nz, {vel, sou})) return nullptr;
  Flat v(vel, nz, nx), s(sou, nt + ndel, nx);
  float** ss = alloc2(nz, nx);
  if (!ss) return nullptr;
  return finish(seis_illum_plane(dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, v.p, s.p, ss[0]), ss);
}

extern "C" void seis_fd_free(float** buf) { free2(buf); }
