// seis_fd.h -- C++ drop-in for the reference's engine header (bbsun/Seis
// src/util/fd.h:3-11; the up/down, rtm_true and boundary helpers are not
// provided).  Same names, argument order and meaning; the result is
// returned as a newly allocated 2D array laid out like MyAlloc<float>::alc
// (memory.h:29-44: a table of n2 row pointers into one zeroed block of n1*n2
// floats), so the reference's MyAlloc<float>::free (memory.h:103-110: delete[]
// buf[0]; delete[] buf;) releases it.  The work runs on the GPU through the C
// ABI of seis.h.  On an engine error these print the message and return NULL
// (the reference exits with status 0 on errors instead, debug.cpp:5-12).
#ifndef SEIS_FD_H
#define SEIS_FD_H

// rec[nx][nt] = forward shot record (fd.cpp:15-80)
float** modeling(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, int* gz,
                 float* wav, float** vel);
// rec[nx][nt] = Born record of perturbation dv (fd.cpp:81-177)
float** forward(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, int* gz,
                float* wav, float** vel, float** dv);
// img[nx][nz] = RTM image of record rec (fd.cpp:592-719)
float** adjoint(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, int* gz,
                float* wav, float** vel, float** rec);

// ss[nx][nz] = source illumination, sum of u^2 over steps it > ndel
// (fd.cpp:275-339)
float** illum(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz, float* wav,
              float** vel);

// Plane-wave engine (fd.cpp:178-274, 720-921): sou[nx][nt+ndel] fired at row
// sz of every column, receivers at row gz.
// rec[nx][nt] = Born record
float** forwordPlane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                     float** vel, float** sou, float** dv);
// img[nx][nz] = plane-wave image of record rec
float** adjointPlane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                     float** vel, float** sou, float** rec);
// ss[nx][nz] = illumination of the line source
float** illumPlane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                   float** vel, float** sou);

// Releases a result of the functions above exactly as the reference's
// MyAlloc<float>::free does (delete[] buf[0]; delete[] buf;), for callers that
// are not C++ (no counterpart in the reference; NULL is ignored).
extern "C" void seis_fd_free(float** buf);

#endif
