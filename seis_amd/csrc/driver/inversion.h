// Inversion driver: the reference's Inversion class (lsrtm/inversion.h:16-208)
// with its parameter file, coordinates file and per-shot file layout, running
// the shots as GPU batches instead of an MPI master/worker farm
// (inversion.cpp:141-206, 365-551, 753-807).  One process, one GPU; nprocs = 1
// works (the reference needs >= 2 ranks because rank 0 only dispatches).
#pragma once
#include <string>
#include <vector>

#include "params.h"
#include "survey.h"

namespace seis {

enum { RTM_IMG = 1, LSRTM_STEP = 2 };  // inversion.h:14-15

class Inversion {
 public:
  Inversion(int rank = 0, int nprocs = 1);
  // Reads the parameter file (inversion.cpp:38-88).
  bool readParamFile(const std::string& file);
  // Reads the coordinates file and checks ns / ngmax against the parameters
  // (inversion.cpp:89-140).
  bool getConfig();
  // Shot records of model v [nx][nz] -> <precsg><is>.dat, [ng][nt] per shot
  // (modeling_MPI, inversion.cpp:141-206).
  bool modeling(const float* v);
  // Born records of perturbation dv through the v0file background ->
  // <wdir>BORN<is>.dat (forward_MPI, inversion.cpp:475-551).
  bool forward(const float* dv);
  // Migrated image [nx][nz] of the CSG records (RTM_IMG) or of the BORN
  // records (LSRTM_STEP): per-shot images of shots sbegin..send go to
  // <wdir>CSG_IMAGE<is>.dat, then all ns per-shot images are summed
  // (adjoint_MPI, inversion.cpp:365-474).
  bool adjoint(float* img, int migTag);
  // Source illumination [nx][nz] of every shot's window through the v0file
  // background -> <wdir>ILLUM<is>.dat, summed over all ns shots in shot order
  // (illum_MPI, inversion.cpp:207-290).
  bool illum(float* ss);
  // Plane-wave synthesis (planeWavePrepare, inversion.cpp:808-912): for each
  // ray parameter p_k = pmin + k dp (k < np) the phase-shifted sum over all
  // shots of the source wavelet at the source columns and of the CSG<is>
  // gathers by receiver column -> <wdir>PLANE_S_NEW<k>.dat [nx][nt+delaycal]
  // and <wdir>PLANE_R_NEW<k>.dat [nx][nt].  Needs planeTag != 0.
  bool planeWavePrepare();
  // Plane-wave illumination (illumPlane_MPI, inversion.cpp:291-364): per ray
  // parameter <wdir>ILLUM_PLANE<k>.dat [nx][nz] of the v0file background,
  // summed over k in order into ss [nx][nz].
  bool illumPlane(float* ss);
  // Plane-wave migration (adjointPlane_MPI, inversion.cpp:1019-1103): per ray
  // parameter the image of PLANE_S_NEW<k> / PLANE_R_NEW<k> (RTM_IMG) or of
  // BORN<k> (LSRTM_STEP), times the mask when mask != 0 ->
  // <wdir>IMAGE<k>.dat, summed over k in order into img [nx][nz].
  bool adjointPlane(float* img, int migTag);
  // Least-squares RTM of the CSG records by CG on the exact Born operator
  // (the declared conope_MPI / cgsolver_MPI, inversion.h:81,87): maxIter
  // iterations or until ||L^T r|| / ||L^T d|| < reError.  img [nx][nz].
  // adjoint_kind: SEIS_LSRTM_ADJ_EXACT (the transpose) or
  // SEIS_LSRTM_ADJ_XCORR (the reference's cross-correlation imaging).
  // weight: optional image-space preconditioner [nx][nz] (host), see
  // seis_lsrtm_set_weight.
  bool cgsolver(int maxIter, float reError, float* img, std::vector<double>* rnorm = nullptr,
                int adjoint_kind = 0, const float* weight = nullptr);

  const std::string& error() const { return err_; }
  Params param;
  Survey survey;

 private:
  bool read_model(const std::string& file, std::vector<float>& v);
  bool plane_batch(int kind, std::vector<float>& out, int migTag);
  int rank_, nprocs_;
  std::string err_;
};

}  // namespace seis
