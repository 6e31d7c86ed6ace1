// Internal declarations shared by the engine's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/seis.h"

namespace seis {

using GridSpec = seis_grid;
using ShotSpec = seis_shot;

enum RunMode { RUN_MODEL = 0, RUN_BORN = 1, RUN_RTM = 2, RUN_ILLUM = 3 };

// Records the calling thread's last error message.
void set_error(const std::string& msg);

// Batched modeling / Born / RTM / illumination.  line_src: wav holds per
// shot the plane-wave source sou [nx][nt+ndel] fired at row sz of every
// column (forwordPlane, illumPlane) instead of a point-source wavelet.
template <typename T>
int run_batch(int mode, const GridSpec& g, int nshots, const ShotSpec* shots, const T* vel, const T* dv,
              const T* wav, const int* gz, T* rec, T* img, hipStream_t stream, int rtm_flags = 0, int line_src = 0);

// Batched plane-wave imaging (adjointPlane, fd.cpp:787-921): per shot the
// source sou [nx][nt+ndel] at offset shots[].wav fired at row sz, the record
// rec [nx][nt] at offset shots[].rec, receiver depths gz [nx] at offset
// shots[].gz; img gets nshots images [nx][nz].
template <typename T>
int run_rtm_plane(const GridSpec& g, int nshots, const ShotSpec* shots, const T* vel, const T* sou, const int* gz,
                  const T* rec, T* img, hipStream_t stream);

}  // namespace seis
