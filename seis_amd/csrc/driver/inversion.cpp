#include "inversion.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>

#include "../../../include/seis.h"
#include "../wavelet.h"

namespace seis {

namespace {

// Device buffer owned by the driver.
struct Dev {
  void* p = nullptr;
  ~Dev() {
    if (p) seis_free(p);
  }
  bool alloc(size_t bytes) { return seis_malloc(&p, bytes) == SEIS_OK; }
  bool put(const void* src, size_t bytes) { return alloc(bytes) && seis_memcpy_htod(p, src, bytes) == SEIS_OK; }
  bool get(void* dst, size_t bytes) const { return seis_memcpy_dtoh(dst, p, bytes) == SEIS_OK; }
};

// Everything one shot needs on its model window.
struct ShotJob {
  int is;
  Window w;
  std::vector<int> igz;
  int sx, sz;
};

std::string gpu_error(const char* what) { return std::string(what) + ": " + seis_last_error(); }

}  // namespace

Inversion::Inversion(int rank, int nprocs) : rank_(rank), nprocs_(nprocs) {}

bool Inversion::readParamFile(const std::string& file) { return read_params(file, param, err_); }

bool Inversion::getConfig() {
  if (!read_coords(param.coordfile.val, survey, err_)) return false;
  if (survey.ns() != param.ns.val) {
    std::ostringstream o;
    o << "error : ns = " << param.ns.val << " from parameter files, " << survey.ns() << " from scanning "
      << param.coordfile.val;
    err_ = o.str();
    return false;
  }
  if (survey.ngmax() != param.ngmax.val) {
    std::ostringstream o;
    o << "error : ngmax = " << param.ngmax.val << " from parameter files, " << survey.ngmax() << " from scanning "
      << param.coordfile.val;
    err_ = o.str();
    return false;
  }
  return true;
}

bool Inversion::read_model(const std::string& file, std::vector<float>& v) {
  v.assign((size_t)param.nx.val * param.nz.val, 0.f);
  return read_floats(file, v.size(), v.data(), err_);
}

namespace {

// Shots staged on the host per engine call: their windowed models and
// records are copied in and out per group, so host memory stays bounded on
// survey-scale runs (the engine cuts each call further by device memory).
constexpr size_t kStageBytes = size_t(4) << 30;

// Runs one of the engine's batched calls over the shots' own windows (shots
// of every width share the launches).  `kind`: 0 modeling, 1 Born, 2 RTM,
// 3 illumination.  models / perts are global [nx][nz]; records are per shot
// [nxw][nt] on input (RTM) or output (modeling, Born); images per shot
// [nxw][nz] (RTM, illumination).
bool run_windows(const Params& p, int kind, std::vector<ShotJob>& jobs, const float* model, const float* pert,
                 std::map<int, std::vector<float>>& recs, std::map<int, std::vector<float>>& imgs, std::string& err) {
  const int nz = p.nz.val, nt = p.nt.val, ndel = p.delaycal.val;
  const int NT = nt + ndel;
  std::vector<float> wav(NT);
  ricker_host(p.dt.val, p.fr.val, p.delay.val + ndel, NT, wav.data());
  size_t j0 = 0;
  while (j0 < jobs.size()) {
    // next group: as many shots as fit the staging bound (at least one)
    size_t j1 = j0, stage = 0;
    while (j1 < jobs.size()) {
      const size_t nxw = (size_t)jobs[j1].w.nx;
      const size_t b = sizeof(float) * nxw * ((kind == 1 ? 2 : 1) * (size_t)nz + nt + (kind >= 2 ? nz : 0));
      if (j1 > j0 && stage + b > kStageBytes) break;
      stage += b;
      j1++;
    }
    const int n = (int)(j1 - j0);
    size_t M = 0, R = 0;
    std::vector<size_t> moff(n), roff(n), goff(n);
    for (int j = 0; j < n; j++) {
      const size_t nxw = (size_t)jobs[j0 + j].w.nx;
      moff[j] = M;
      roff[j] = R;
      goff[j] = M / nz;  // gz rows: one entry per column, like the models
      M += nxw * nz;
      R += nxw * nt;
    }
    std::vector<float> vm(M), dm(kind == 1 ? M : 0), wv((size_t)NT * n), rc(R, 0.f);
    std::vector<int> gz(M / nz);
    std::vector<seis_shot> sh(n);
    for (int j = 0; j < n; j++) {
      const ShotJob& q = jobs[j0 + j];
      window_model(q.w, nz, model, vm.data() + moff[j]);
      if (kind == 1) window_model(q.w, nz, pert, dm.data() + moff[j]);
      std::copy(wav.begin(), wav.end(), wv.begin() + (size_t)NT * j);
      std::copy(q.igz.begin(), q.igz.end(), gz.begin() + goff[j]);
      if (kind == 2) std::copy(recs[q.is].begin(), recs[q.is].end(), rc.begin() + roff[j]);
      sh[j] = seis_shot{q.sx, q.sz, nt, ndel, (long long)moff[j], (long long)NT * j, (long long)goff[j],
                        (long long)roff[j], q.w.nx, 0};
    }
    seis_grid g{p.nx.val, nz, p.npml.val, p.dt.val, p.dx.val, p.dz.val};
    Dev dv, dd, dw, dg, dr, di;
    if (!dv.put(vm.data(), sizeof(float) * vm.size()) || !dw.put(wv.data(), sizeof(float) * wv.size()) ||
        !dg.put(gz.data(), sizeof(int) * gz.size()) || !dr.put(rc.data(), sizeof(float) * rc.size()) ||
        (kind == 1 && !dd.put(dm.data(), sizeof(float) * dm.size())) ||
        (kind >= 2 && !di.alloc(sizeof(float) * M))) {
      err = gpu_error("device buffers");
      return false;
    }
    int rcode;
    if (kind == 0)
      rcode = seis_modeling_batch(&g, n, sh.data(), (float*)dv.p, (float*)dw.p, (int*)dg.p, (float*)dr.p);
    else if (kind == 1)
      rcode = seis_born_batch(&g, n, sh.data(), (float*)dv.p, (float*)dd.p, (float*)dw.p, (int*)dg.p, (float*)dr.p);
    else if (kind == 2)
      rcode = seis_rtm_batch(&g, n, sh.data(), (float*)dv.p, (float*)dw.p, (int*)dg.p, (float*)dr.p, (float*)di.p);
    else
      rcode = seis_illum_batch(&g, n, sh.data(), (float*)dv.p, (float*)dw.p, (float*)di.p);
    if (rcode) {
      err = gpu_error("engine batch");
      return false;
    }
    if (kind >= 2) {
      std::vector<float> im(M);
      if (!di.get(im.data(), sizeof(float) * im.size())) return err = gpu_error("download"), false;
      for (int j = 0; j < n; j++)
        imgs[jobs[j0 + j].is].assign(im.begin() + moff[j], im.begin() + moff[j] + (size_t)jobs[j0 + j].w.nx * nz);
    } else {
      if (!dr.get(rc.data(), sizeof(float) * rc.size())) return err = gpu_error("download"), false;
      for (int j = 0; j < n; j++)
        recs[jobs[j0 + j].is].assign(rc.begin() + roff[j], rc.begin() + roff[j] + (size_t)jobs[j0 + j].w.nx * nt);
    }
    if (kind == 2)
      for (int j = 0; j < n; j++) recs.erase(jobs[j0 + j].is);  // records imaged: release them
    j0 = j1;
  }
  return true;
}

// Calls fn(group) for consecutive groups of jobs whose host staging fits
// kStageBytes (at least one shot each), so that records, windowed models and
// images are read, run and written group by group.
template <class F>
bool for_stage_groups(const Params& p, int kind, std::vector<ShotJob>& jobs, F&& fn) {
  const size_t nz = (size_t)p.nz.val, nt = (size_t)p.nt.val;
  size_t j0 = 0;
  while (j0 < jobs.size()) {
    size_t j1 = j0, stage = 0;
    while (j1 < jobs.size()) {
      const size_t b = sizeof(float) * (size_t)jobs[j1].w.nx * ((kind == 1 ? 2 : 1) * nz + nt + (kind >= 2 ? nz : 0));
      if (j1 > j0 && stage + b > kStageBytes) break;
      stage += b;
      j1++;
    }
    std::vector<ShotJob> group(jobs.begin() + j0, jobs.begin() + j1);
    if (!fn(group)) return false;
    j0 = j1;
  }
  return true;
}

bool make_job(const Params& p, const Survey& s, int is, ShotJob& q, std::string& err) {
  q.is = is;
  q.w = shot_window(p, s, is);
  if (q.w.nx <= 0) return err = "empty shot window", false;
  if (!receiver_depths(p, s, is, q.w.fx, q.w.nx, q.igz, err)) return false;
  source_cell(p, s, is, q.w.fx, &q.sx, &q.sz);
  if (q.sx < 0 || q.sx >= q.w.nx || q.sz < 0 || q.sz >= p.nz.val) return err = "source outside its window", false;
  return true;
}

// Records of the engine (per window column) -> muted gathers written to file.
bool write_gathers(const Params& p, const Survey& s, std::vector<ShotJob>& jobs,
                   std::map<int, std::vector<float>>& recs, const float* vmute, bool born, std::string& err) {
  const int nt = p.nt.val;
  for (ShotJob& q : jobs) {
    std::vector<float> csg((size_t)s.ng(q.is) * nt);
    if (!reorder(p, s, q.is, nt, csg.data(), recs[q.is].data(), q.w.fx, q.w.nx, true, err)) return false;
    if (!remove_direct(p, s, q.is, csg.data(), vmute, err)) return false;
    const std::string f = born ? born_name(p, q.is) : csg_name(p, q.is);
    if (!write_floats(f, csg.size(), csg.data(), err)) return false;
  }
  return true;
}

}  // namespace

bool Inversion::modeling(const float* v) {
  std::vector<ShotJob> jobs(survey.ns());
  for (int is = 0; is < survey.ns(); is++)
    if (!make_job(param, survey, is, jobs[is], err_)) return false;
  return for_stage_groups(param, 0, jobs, [&](std::vector<ShotJob>& group) {
    std::map<int, std::vector<float>> recs, imgs;
    return run_windows(param, 0, group, v, nullptr, recs, imgs, err_) &&
           write_gathers(param, survey, group, recs, v, false, err_);
  });
}

bool Inversion::forward(const float* dv) {
  std::vector<float> v0;
  if (!read_model(param.v0file.val, v0)) return false;
  std::vector<ShotJob> jobs(survey.ns());
  for (int is = 0; is < survey.ns(); is++)
    if (!make_job(param, survey, is, jobs[is], err_)) return false;
  return for_stage_groups(param, 1, jobs, [&](std::vector<ShotJob>& group) {
    std::map<int, std::vector<float>> recs, imgs;
    return run_windows(param, 1, group, v0.data(), dv, recs, imgs, err_) &&
           write_gathers(param, survey, group, recs, v0.data(), true, err_);
  });
}

bool Inversion::adjoint(float* img, int migTag) {
  std::vector<float> v0;
  if (!read_model(param.v0file.val, v0)) return false;
  const int nx = param.nx.val, nz = param.nz.val, nt = param.nt.val, ns = survey.ns();
  std::vector<ShotJob> jobs;
  for (int is = std::max(0, param.sbegin.val); is <= std::min(ns - 1, param.send.val); is++) {
    ShotJob q;
    if (!make_job(param, survey, is, q, err_)) return false;
    jobs.push_back(q);
  }
  std::vector<float> mask;
  if (param.mask.val) {
    mask.resize((size_t)nx * nz);
    if (!read_floats(param.maskfile.val, mask.size(), mask.data(), err_)) return false;
  }
  std::vector<float> g((size_t)nx * nz);
  const bool ok = for_stage_groups(param, 2, jobs, [&](std::vector<ShotJob>& group) {
    std::map<int, std::vector<float>> recs, imgs;
    for (const ShotJob& q : group) {
      std::vector<float> csg((size_t)survey.ng(q.is) * nt);
      const std::string f = migTag == RTM_IMG ? csg_name(param, q.is) : born_name(param, q.is);
      if (!read_floats(f, csg.size(), csg.data(), err_)) return false;
      if (!remove_direct(param, survey, q.is, csg.data(), v0.data(), err_)) return false;
      std::vector<float>& r = recs[q.is];
      r.assign((size_t)q.w.nx * nt, 0.f);
      if (!reorder(param, survey, q.is, nt, csg.data(), r.data(), q.w.fx, q.w.nx, false, err_)) return false;
    }
    if (!run_windows(param, 2, group, v0.data(), nullptr, recs, imgs, err_)) return false;
    for (ShotJob& q : group) {
      window_to_global(param, imgs[q.is].data(), g.data(), q.w.fx, param.velfx.val, q.w.nx, nx, nz, false);
      if (param.mask.val)
        for (size_t i = 0; i < g.size(); i++) g[i] = g[i] * mask[i];
      if (!write_floats(shot_image_name(param, q.is), g.size(), g.data(), err_)) return false;
    }
    return true;
  });
  if (!ok) return false;
  // the image is the sum of every shot's file, including shots imaged by
  // earlier runs (inversion.cpp:460-471)
  std::fill(img, img + (size_t)nx * nz, 0.f);
  for (int is = 0; is < ns; is++) {
    if (!read_floats(shot_image_name(param, is), g.size(), g.data(), err_)) return false;
    for (size_t i = 0; i < g.size(); i++) img[i] = g[i] + img[i];
  }
  return true;
}

bool Inversion::illum(float* ss) {
  std::vector<float> v0;
  if (!read_model(param.v0file.val, v0)) return false;
  const int nx = param.nx.val, nz = param.nz.val, ns = survey.ns();
  std::vector<ShotJob> jobs(ns);
  for (int is = 0; is < ns; is++)
    if (!make_job(param, survey, is, jobs[is], err_)) return false;
  std::fill(ss, ss + (size_t)nx * nz, 0.f);
  std::vector<float> g((size_t)nx * nz);
  return for_stage_groups(param, 3, jobs, [&](std::vector<ShotJob>& group) {
    std::map<int, std::vector<float>> recs, imgs;
    if (!run_windows(param, 3, group, v0.data(), nullptr, recs, imgs, err_)) return false;
    for (ShotJob& q : group) {
      window_to_global(param, imgs[q.is].data(), g.data(), q.w.fx, param.velfx.val, q.w.nx, nx, nz, false);
      if (!write_floats(illum_name(param, q.is), g.size(), g.data(), err_)) return false;
      for (size_t i = 0; i < g.size(); i++) ss[i] = g[i] + ss[i];
    }
    return true;
  });
}

bool Inversion::planeWavePrepare() {
  if (!param.planeTag.val) return err_ = "in param planeTag must be true(none zero) to use subroutine planeWavePrepare", false;
  const int nx = param.nx.val, nt = param.nt.val, np = param.np.val, ns = survey.ns();
  const int NT = nt + param.delaycal.val;
  const float dx = param.dx.val, velfx = param.velfx.val;
  if (np <= 0) return true;
  std::vector<float> wav(NT);
  ricker_host(param.dt.val, param.fr.val, param.delay.val + param.delaycal.val, NT, wav.data());
  // source columns: every shot must lie in [velfx - 0.0001 dx, max sx]
  // (inversion.cpp:845-850, 872-880)
  float sxmax = ns > 0 ? survey.sx[0] : 0.f;
  for (int is = 1; is < ns; is++) sxmax = sxmax < survey.sx[is] ? survey.sx[is] : sxmax;
  const float velmin = velfx - 0.0001 * dx, velmax = sxmax;
  std::vector<int> isx(ns), ng(ns), col;
  size_t ntr = 0;
  for (int is = 0; is < ns; is++) {
    const float sx = survey.sx[is];
    if (!((sx - velmin) >= 0 && (velmax - sx) >= 0)) return err_ = "sx must be in the range [velmin,velmax]", false;
    isx[is] = (int)((sx - velfx) / dx + 0.001f);
    if (isx[is] < 0 || isx[is] >= nx) return err_ = "isx must be in the range [0 nx) in planeWavePrepare", false;
    std::vector<int> c;
    if (!receiver_columns(param, survey, is, velfx, nx, c, err_)) return false;
    ng[is] = (int)c.size();
    col.insert(col.end(), c.begin(), c.end());
    ntr += c.size();
  }
  std::vector<float> data(ntr * nt);
  for (int is = 0, r = 0; is < ns; r += ng[is], is++)
    if (!read_floats(csg_name(param, is), (size_t)ng[is] * nt, data.data() + (size_t)r * nt, err_)) return false;
  // All ray parameters in one call when their outputs fit in 2 GiB of device
  // memory, else one call per parameter: p_k is pmin + k dp in float either
  // way (inversion.cpp:866), so the files are the same.
  const size_t per = (size_t)nx * (NT + nt) * sizeof(float);
  const int chunk = per * np <= ((size_t)2 << 30) ? np : 1;
  Dev dd, dw, dsou, drec;
  if (!dd.put(data.data(), sizeof(float) * std::max<size_t>(data.size(), 1)) ||
      !dw.put(wav.data(), sizeof(float) * wav.size()) || !dsou.alloc((size_t)chunk * nx * NT * sizeof(float)) ||
      !drec.alloc((size_t)chunk * nx * nt * sizeof(float)))
    return err_ = gpu_error("device buffers"), false;
  std::vector<float> sou((size_t)chunk * nx * NT), rec((size_t)chunk * nx * nt);
  const std::vector<float> sxs(survey.sx.begin(), survey.sx.end());
  for (int k0 = 0; k0 < np; k0 += chunk) {
    const float pk = param.pmin.val + (float)k0 * param.dp.val;
    if (seis_plane_prepare(nx, nt, NT, ns, chunk, pk, param.dp.val, sxs.data(), isx.data(), ng.data(), col.data(),
                           (const float*)dw.p, (const float*)dd.p, (float*)dsou.p, (float*)drec.p))
      return err_ = gpu_error("plane prepare"), false;
    if (!dsou.get(sou.data(), sizeof(float) * sou.size()) || !drec.get(rec.data(), sizeof(float) * rec.size()))
      return err_ = gpu_error("download"), false;
    for (int k = 0; k < chunk; k++)
      if (!write_floats(plane_source_name(param, k0 + k), (size_t)nx * NT, sou.data() + (size_t)k * nx * NT, err_) ||
          !write_floats(plane_record_name(param, k0 + k), (size_t)nx * nt, rec.data() + (size_t)k * nx * nt, err_))
        return false;
  }
  return true;
}

// kind 0: illumination, 1: migration.  Reads the per-parameter files, runs
// every ray parameter in one batch on the whole model, returns the outputs
// [np][nx][nz].
bool Inversion::plane_batch(int kind, std::vector<float>& out, int migTag) {
  std::vector<float> v0;
  if (!read_model(param.v0file.val, v0)) return false;
  const int nx = param.nx.val, nz = param.nz.val, nt = param.nt.val, np = param.np.val;
  const int ndel = param.delaycal.val, NT = nt + ndel;
  if (survey.ns() < 1 || survey.ng(0) < 1) return err_ = "plane-wave modes need shot 0 and its first receiver", false;
  // source and receiver rows of shot 0 (inversion.cpp:331-332, 1064-1065)
  const int sz = (int)((survey.sz[0] - 0.0f) / param.dz.val + 0.0001f);
  const int gz = (int)((survey.gz[0][0] - 0.0f) / param.dz.val + 0.0001f);
  if (sz < 0 || sz >= nz || gz < 0 || gz >= nz) return err_ = "plane-wave source or receiver row outside the model", false;
  out.assign((size_t)np * nx * nz, 0.f);
  if (np <= 0) return true;
  std::vector<float> sou((size_t)np * nx * NT), rec(kind == 1 ? (size_t)np * nx * nt : 0);
  std::vector<seis_shot> sh(np);
  for (int k = 0; k < np; k++) {
    if (!read_floats(plane_source_name(param, k), (size_t)nx * NT, sou.data() + (size_t)k * nx * NT, err_)) return false;
    if (kind == 1) {
      const std::string f = migTag == RTM_IMG ? plane_record_name(param, k) : born_name(param, k);
      if (!read_floats(f, (size_t)nx * nt, rec.data() + (size_t)k * nx * nt, err_)) return false;
    }
    sh[k] = seis_shot{0, sz, nt, ndel, 0, (long long)k * nx * NT, 0, (long long)k * nx * nt};
  }
  const std::vector<int> gzs(nx, gz);
  seis_grid g{nx, nz, param.npml.val, param.dt.val, param.dx.val, param.dz.val};
  Dev dv, ds, dg, dr, di;
  if (!dv.put(v0.data(), sizeof(float) * v0.size()) || !ds.put(sou.data(), sizeof(float) * sou.size()) ||
      !dg.put(gzs.data(), sizeof(int) * nx) || (kind == 1 && !dr.put(rec.data(), sizeof(float) * rec.size())) ||
      !di.alloc(sizeof(float) * out.size()))
    return err_ = gpu_error("device buffers"), false;
  const int rc = kind == 0 ? seis_illum_plane_batch(&g, np, sh.data(), (float*)dv.p, (float*)ds.p, (float*)di.p)
                           : seis_rtm_plane_batch(&g, np, sh.data(), (float*)dv.p, (float*)ds.p, (int*)dg.p,
                                                  (float*)dr.p, (float*)di.p);
  if (rc) return err_ = gpu_error("plane-wave batch"), false;
  if (!di.get(out.data(), sizeof(float) * out.size())) return err_ = gpu_error("download"), false;
  return true;
}

bool Inversion::illumPlane(float* ss) {
  std::vector<float> all;
  if (!plane_batch(0, all, RTM_IMG)) return false;
  const size_t nm = (size_t)param.nx.val * param.nz.val;
  for (int k = 0; k < param.np.val; k++)
    if (!write_floats(plane_illum_name(param, k), nm, all.data() + k * nm, err_)) return false;
  // the sum reads every parameter's file back (inversion.cpp:352-360)
  std::fill(ss, ss + nm, 0.f);
  std::vector<float> t(nm);
  for (int k = 0; k < param.np.val; k++) {
    if (!read_floats(plane_illum_name(param, k), nm, t.data(), err_)) return false;
    for (size_t i = 0; i < nm; i++) ss[i] = t[i] + ss[i];
  }
  return true;
}

bool Inversion::adjointPlane(float* img, int migTag) {
  std::vector<float> all;
  if (!plane_batch(1, all, migTag)) return false;
  const size_t nm = (size_t)param.nx.val * param.nz.val;
  std::vector<float> mask;
  if (param.mask.val) {
    mask.resize(nm);
    if (!read_floats(param.maskfile.val, nm, mask.data(), err_)) return false;
  }
  for (int k = 0; k < param.np.val; k++) {
    float* m = all.data() + k * nm;
    if (param.mask.val)
      for (size_t i = 0; i < nm; i++) m[i] = m[i] * mask[i];
    if (!write_floats(plane_image_name(param, k), nm, m, err_)) return false;
  }
  std::fill(img, img + nm, 0.f);
  std::vector<float> t(nm);
  for (int k = 0; k < param.np.val; k++) {
    if (!read_floats(plane_image_name(param, k), nm, t.data(), err_)) return false;
    for (size_t i = 0; i < nm; i++) img[i] = t[i] + img[i];
  }
  return true;
}

bool Inversion::cgsolver(int maxIter, float reError, float* img, std::vector<double>* rnorm, int adjoint_kind,
                         const float* weight) {
  std::vector<float> v0;
  if (!read_model(param.v0file.val, v0)) return false;
  const int nx = param.nx.val, nz = param.nz.val, nt = param.nt.val, ns = survey.ns();
  const int ndel = param.delaycal.val, NT = nt + ndel;
  // every shot on its window (getVel's aperture, inversion.cpp:552-598,
  // clipped to the model's columns: the plan's shots are views of the one
  // background model); data and mask per window column
  const float dx = param.dx.val, velfx = param.velfx.val;
  std::vector<float> wav(NT), d, mask;
  ricker_host(param.dt.val, param.fr.val, param.delay.val + ndel, NT, wav.data());
  std::vector<float> wv((size_t)NT * ns);
  std::vector<int> gz;
  std::vector<seis_shot> sh(ns);
  for (int is = 0; is < ns; is++) {
    const Window w = shot_window(param, survey, is);
    if (w.nx <= 0) return err_ = "empty shot window", false;
    const int a = (int)std::floor((w.fx - velfx) / dx + 0.5f);
    const int c0 = std::max(0, a), c1 = std::min(nx - 1, a + w.nx - 1);
    if (c1 < c0) return err_ = "shot window outside the model", false;
    const int nxw = c1 - c0 + 1;
    const float fx = velfx + c0 * dx;
    std::vector<int> igz;
    if (!receiver_depths(param, survey, is, fx, nxw, igz, err_)) return false;
    int sx, sz;
    source_cell(param, survey, is, fx, &sx, &sz);
    if (sx < 0 || sx >= nxw || sz < 0 || sz >= nz) return err_ = "source outside its window", false;
    std::vector<float> csg((size_t)survey.ng(is) * nt), ones(csg.size(), 1.f);
    if (!read_floats(csg_name(param, is), csg.size(), csg.data(), err_)) return false;
    if (!remove_direct(param, survey, is, csg.data(), v0.data(), err_)) return false;
    if (!remove_direct(param, survey, is, ones.data(), v0.data(), err_)) return false;
    const size_t r0 = d.size();
    d.resize(r0 + (size_t)nxw * nt, 0.f);
    mask.resize(r0 + (size_t)nxw * nt, 0.f);
    if (!reorder(param, survey, is, nt, csg.data(), d.data() + r0, fx, nxw, false, err_)) return false;
    if (!reorder(param, survey, is, nt, ones.data(), mask.data() + r0, fx, nxw, false, err_)) return false;
    std::copy(wav.begin(), wav.end(), wv.begin() + (size_t)NT * is);
    const size_t g0 = gz.size();
    gz.insert(gz.end(), igz.begin(), igz.end());
    sh[is] = seis_shot{sx, sz, nt, ndel, (long long)c0 * nz, (long long)NT * is, (long long)g0, (long long)r0, nxw,
                       c0};
  }
  Dev dv, dw, dg, dd, dm, di, dwt;
  if (weight && !dwt.put(weight, sizeof(float) * (size_t)nx * nz)) return err_ = gpu_error("device buffers"), false;
  if (!dv.put(v0.data(), sizeof(float) * v0.size()) || !dw.put(wv.data(), sizeof(float) * wv.size()) ||
      !dg.put(gz.data(), sizeof(int) * gz.size()) || !dd.put(d.data(), sizeof(float) * d.size()) ||
      !dm.put(mask.data(), sizeof(float) * mask.size()) || !di.alloc(sizeof(float) * (size_t)nx * nz))
    return err_ = gpu_error("device buffers"), false;
  seis_grid g{nx, nz, param.npml.val, param.dt.val, param.dx.val, param.dz.val};
  seis_lsrtm* plan = nullptr;
  if (seis_lsrtm_create(&plan, &g, ns, sh.data(), (float*)dv.p, (float*)dw.p, (int*)dg.p, 1))
    return err_ = gpu_error("lsrtm plan"), false;
  std::vector<double> hist(2 * (maxIter + 1), 0.0);
  int done = 0;
  int rc = seis_lsrtm_set_mask(plan, dm.p);
  if (!rc) rc = seis_lsrtm_set_adjoint(plan, adjoint_kind);
  if (!rc && weight) rc = seis_lsrtm_set_weight(plan, dwt.p);
  if (!rc) rc = seis_lsrtm_cg(plan, dd.p, maxIter, (double)reError, di.p, hist.data(), &done);
  seis_lsrtm_destroy(plan);
  if (rc) return err_ = gpu_error("lsrtm cg"), false;
  if (!di.get(img, sizeof(float) * (size_t)nx * nz)) return err_ = gpu_error("download"), false;
  if (rnorm) {
    rnorm->clear();
    for (int i = 0; i <= done; i++) rnorm->push_back(hist[2 * i]);
  }
  return true;
}

}  // namespace seis
