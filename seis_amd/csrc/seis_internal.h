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

template <typename T>
int run_batch(int mode, const GridSpec& g, int nshots, const ShotSpec* shots, const T* vel, const T* dv,
              const T* wav, const int* gz, T* rec, T* img, hipStream_t stream, int rtm_flags = 0);

}  // namespace seis
