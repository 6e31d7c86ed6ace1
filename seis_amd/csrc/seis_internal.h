// Internal declarations shared by the engine's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/seis.h"

namespace seis {

using GridSpec = seis_grid;
using ShotSpec = seis_shot;

enum RunMode { RUN_MODEL = 0, RUN_BORN = 1, RUN_RTM = 2, RUN_ILLUM = 3, RUN_RTM_PLANE = 4 };

// Caller-provided device memory for a batch run: the run takes its fields,
// strips and shot descriptors from it instead of allocating, cutting the
// batch into chunks that fit (see seis_modeling_batch_static).
struct Workspace {
  void* ptr;
  size_t bytes;
};

// Records the calling thread's last error message.
void set_error(const std::string& msg);

// Batched modeling / Born / RTM / illumination.  line_src: wav holds per
// shot the plane-wave source sou [nx][nt+ndel] fired at row sz of every
// column (forwordPlane, illumPlane) instead of a point-source wavelet.
template <typename T>
int run_batch(int mode, const GridSpec& g, int nshots, const ShotSpec* shots, const T* vel, const T* dv,
              const T* wav, const int* gz, T* rec, T* img, hipStream_t stream, int rtm_flags = 0, int line_src = 0,
              const Workspace* ws = nullptr);

// Padded-grid validity of a grid (or of one shot's window of it).
int check_grid(const GridSpec& g);

// Columns of a shot's grid: its own window width, or the grid's nx.
int shot_width(const GridSpec& g, const ShotSpec& q);

// Batched plane-wave imaging (adjointPlane, fd.cpp:787-921): per shot the
// source sou [nx][nt+ndel] at offset shots[].wav fired at row sz, the record
// rec [nx][nt] at offset shots[].rec, receiver depths gz [nx] at offset
// shots[].gz; img gets nshots images [nx][nz].
template <typename T>
int run_rtm_plane(const GridSpec& g, int nshots, const ShotSpec* shots, const T* vel, const T* sou, const int* gz,
                  const T* rec, T* img, hipStream_t stream);

// harmonic of one trace of n samples (device buffers).
int run_harmonic(const float* in, float* out, int n, float kx, float x, hipStream_t stream);

// planeWavePrepare for np ray parameters; arguments as seis_plane_prepare.
int run_plane_prepare(int nx, int nt, int ntw, int ns, int np, float pmin, float dp, const float* sx,
                      const int* isx, const int* ng, const int* col, const float* wav, const float* data,
                      float* sou, float* rec, hipStream_t stream);

}  // namespace seis
