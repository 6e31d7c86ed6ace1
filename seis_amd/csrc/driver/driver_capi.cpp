// C ABI of the parameter-file driver (include/seis.h: seis_run).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/seis.h"
#include "../seis_internal.h"
#include "../wavelet.h"
#include "inversion.h"

using namespace seis;

extern "C" {

int seis_ricker(float dt, float fr, int num_delay, int n, float* out) {
  if (!out || n < 0) return set_error("ricker: bad argument"), SEIS_ERR_ARG;
  ricker_host(dt, fr, num_delay, n, out);
  return SEIS_OK;
}

int seis_read_params(const char* param_file, char* out, int out_len) {
  if (!param_file) return set_error("seis_read_params: null file"), SEIS_ERR_ARG;
  Params p;
  std::string err;
  if (!read_params(param_file, p, err)) return set_error(err), SEIS_ERR_ARG;
  const std::string text = format_params(p);
  if (out && out_len > 0) {
    const size_t n = std::min(text.size(), (size_t)out_len - 1);
    std::memcpy(out, text.data(), n);
    out[n] = 0;
  }
  return SEIS_OK;
}

int seis_read_coords(const char* coord_file, int max_shots, int max_traces, int* ns, int* ntraces, int* ng,
                     float* sxz, float* gxz) {
  if (!coord_file || !ns || !ntraces) return set_error("seis_read_coords: null argument"), SEIS_ERR_ARG;
  Survey s;
  std::string err;
  if (!read_coords(coord_file, s, err)) return set_error(err), SEIS_ERR_ARG;
  int total = 0;
  for (int is = 0; is < s.ns(); is++) total += s.ng(is);
  *ns = s.ns();
  *ntraces = total;
  if (s.ns() > max_shots || total > max_traces || !ng || !sxz || !gxz) return SEIS_OK;  // sizes only
  int t = 0;
  for (int is = 0; is < s.ns(); is++) {
    ng[is] = s.ng(is);
    sxz[2 * is] = s.sx[is];
    sxz[2 * is + 1] = s.sz[is];
    for (int g = 0; g < s.ng(is); g++, t++) {
      gxz[2 * t] = s.gx[is][g];
      gxz[2 * t + 1] = s.gz[is][g];
    }
  }
  return SEIS_OK;
}

int seis_run(const char* param_file, const char* mode, const char* model_file, const char* out_file) {
  if (!param_file || !mode) return set_error("seis_run: null argument"), SEIS_ERR_ARG;
  Inversion inv;
  if (!inv.readParamFile(param_file) || !inv.getConfig()) return set_error(inv.error()), SEIS_ERR_ARG;
  const Params& p = inv.param;
  const size_t n = (size_t)p.nx.val * p.nz.val;
  std::string err;
  const std::string m(mode);
  if (m == "modeling" || m == "born") {
    std::string f = model_file ? model_file : (m == "modeling" ? p.vfile.val : "");
    if (f.empty()) return set_error("seis_run: no model file"), SEIS_ERR_ARG;
    std::vector<float> v(n);
    if (!read_floats(f, n, v.data(), err)) return set_error(err), SEIS_ERR_ARG;
    const bool ok = m == "modeling" ? inv.modeling(v.data()) : inv.forward(v.data());
    if (!ok) return set_error(inv.error()), SEIS_ERR_HIP;
    return SEIS_OK;
  }
  if (m == "plane_prepare") {
    if (!inv.planeWavePrepare()) return set_error(inv.error()), SEIS_ERR_HIP;
    return SEIS_OK;
  }
  std::vector<float> img(n);
  if (m == "rtm" || m == "rtm_born") {
    if (!inv.adjoint(img.data(), m == "rtm" ? RTM_IMG : LSRTM_STEP)) return set_error(inv.error()), SEIS_ERR_HIP;
  } else if (m == "lsrtm" || m == "lsrtm_xcorr") {
    const int kind = m == "lsrtm" ? SEIS_LSRTM_ADJ_EXACT : SEIS_LSRTM_ADJ_XCORR;
    if (!inv.cgsolver(p.maxiter.val, 0.f, img.data(), nullptr, kind)) return set_error(inv.error()), SEIS_ERR_HIP;
  } else if (m == "lsrtm_illum") {
    // illumination-compensated LSRTM: the source illumination of every shot
    // (illum_MPI) as the preconditioner W = sqrt(max / (ss + 1e-3 max))
    std::vector<float> ss(n), w(n);
    if (!inv.illum(ss.data())) return set_error(inv.error()), SEIS_ERR_HIP;
    float mx = 0.f;
    for (float v : ss) mx = std::max(mx, v);
    if (!(mx > 0.f)) return set_error("seis_run: lsrtm_illum: the survey has no illumination"), SEIS_ERR_ARG;
    for (size_t i = 0; i < n; i++) w[i] = std::sqrt(mx / (ss[i] + 1e-3f * mx));
    if (!inv.cgsolver(p.maxiter.val, 0.f, img.data(), nullptr, SEIS_LSRTM_ADJ_EXACT, w.data()))
      return set_error(inv.error()), SEIS_ERR_HIP;
  } else if (m == "illum") {
    if (!inv.illum(img.data())) return set_error(inv.error()), SEIS_ERR_HIP;
  } else if (m == "rtm_plane" || m == "rtm_plane_born") {
    if (!p.planeTag.val) return set_error("seis_run: plane-wave modes need planeTag != 0"), SEIS_ERR_ARG;
    if (!inv.adjointPlane(img.data(), m == "rtm_plane" ? RTM_IMG : LSRTM_STEP))
      return set_error(inv.error()), SEIS_ERR_HIP;
  } else if (m == "illum_plane") {
    if (!p.planeTag.val) return set_error("seis_run: plane-wave modes need planeTag != 0"), SEIS_ERR_ARG;
    if (!inv.illumPlane(img.data())) return set_error(inv.error()), SEIS_ERR_HIP;
  } else {
    return set_error("seis_run: mode must be modeling, born, rtm, rtm_born, lsrtm, lsrtm_xcorr, lsrtm_illum, "
                     "illum, plane_prepare, rtm_plane, rtm_plane_born or illum_plane"),
           SEIS_ERR_ARG;
  }
  const char* name = m == "lsrtm" || m == "lsrtm_xcorr" || m == "lsrtm_illum" ? "LSRTM_IMAGE.dat"
                     : m == "illum"       ? "ILLUM.dat"
                     : m == "illum_plane" ? "ILLUM_PLANE.dat"
                     : m == "rtm_plane" || m == "rtm_plane_born" ? "PLANE_IMAGE.dat"
                                                                 : "RTM_IMAGE.dat";
  const std::string f = out_file ? out_file : p.wdir.val + name;
  if (!write_floats(f, n, img.data(), err)) return set_error(err), SEIS_ERR_ARG;
  return SEIS_OK;
}

}  // extern "C"
