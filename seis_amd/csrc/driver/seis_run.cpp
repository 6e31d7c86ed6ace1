// seis_run: command-line driver over a reference-style parameter file.
//   seis_run <param file> modeling [model file]   shot records  <precsg><is>.dat
//   seis_run <param file> born <perturbation file>  Born records <wdir>BORN<is>.dat
//   seis_run <param file> rtm [image file]        RTM image of the shot records
//   seis_run <param file> rtm_born [image file]   RTM image of the Born records
//   seis_run <param file> lsrtm [image file]      least-squares RTM (maxiter CG steps)
//   seis_run <param file> lsrtm_xcorr [file]      the same with the cross-correlation adjoint
//   seis_run <param file> lsrtm_illum [file]      the same preconditioned by the source illumination
//   seis_run <param file> illum [image file]      source illumination (per shot <wdir>ILLUM<is>.dat, summed)
//   seis_run <param file> plane_prepare           plane-wave sources / records <wdir>PLANE_S_NEW<k>.dat, PLANE_R_NEW<k>.dat
//   seis_run <param file> rtm_plane [image file]  plane-wave image (per ray parameter <wdir>IMAGE<k>.dat, summed)
//   seis_run <param file> rtm_plane_born [file]   the same of the <wdir>BORN<k> records
//   seis_run <param file> illum_plane [file]      plane-wave illumination (<wdir>ILLUM_PLANE<k>.dat, summed)
// Exit status is 0 on success and 1 on error (the reference exits 0 on errors).
#include <cstdio>
#include <cstring>
#include <string>

#include "../../../include/seis.h"

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <param file> <modeling|born|rtm|rtm_born|lsrtm|lsrtm_xcorr|lsrtm_illum|illum|plane_prepare|rtm_plane|rtm_plane_born|illum_plane> [file]\n", argv[0]);
    return 1;
  }
  const std::string mode = argv[2];
  const char* arg = argc > 3 ? argv[3] : nullptr;
  const bool takes_model = (mode == "modeling" || mode == "born");
  const int rc = seis_run(argv[1], argv[2], takes_model ? arg : nullptr, takes_model ? nullptr : arg);
  if (rc != SEIS_OK) {
    std::fprintf(stderr, "seis_run: %s\n", seis_last_error());
    return 1;
  }
  return 0;
}
