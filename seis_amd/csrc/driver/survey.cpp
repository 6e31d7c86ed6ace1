#include "survey.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>

namespace seis {

int Survey::ngmax() const {
  int m = 0;
  for (const auto& g : gx) m = std::max(m, (int)g.size());
  return m;
}

bool read_coords(const std::string& path, Survey& s, std::string& err) {
  // The reference reads the file line by line with getline until eof and
  // scans each of the four fields on its own (coord.cpp:22-176); a field that
  // does not scan keeps the previous line's value.  A line whose source x
  // differs from the previous line's starts a shot, every other line adds a
  // receiver to the current one.  Its counts have two quirks kept here for
  // identical geometry: the last shot's count is one less than its lines
  // (a file that ends with a newline reads one empty line more, which repeats
  // the last receiver, so the count comes out right; without that newline the
  // last receiver is dropped), and a first shot of a single line keeps the
  // count 0 when another shot follows (countReceivers sets a shot's count
  // only from its second line on).
  std::ifstream in(path, std::ios::binary);
  if (!in) {
    err = "Unable to open " + path;
    return false;
  }
  std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  if (text.find_first_not_of(" \t\r\n") == std::string::npos) {
    err = "no traces in " + path;
    return false;
  }
  std::vector<std::string> lines;
  size_t at = 0;
  while (at <= text.size()) {  // getline until eof: a final newline yields one empty line more
    const size_t e = text.find('\n', at);
    if (e == std::string::npos) {
      if (at < text.size() || lines.empty()) lines.push_back(text.substr(at));
      else lines.push_back("");
      break;
    }
    lines.push_back(text.substr(at, e - at));
    at = e + 1;
    if (at == text.size()) {
      lines.push_back("");
      break;
    }
  }
  s = Survey();
  float f[4] = {0.f, 0.f, 0.f, 0.f};
  auto scan = [&](const std::string& l) {
    // each field on its own, as sscanf(buffer, "%*s%*s%f", ...) does
    const char* fmt[4] = {"%f", "%*s%f", "%*s%*s%f", "%*s%*s%*s%f"};
    for (int k = 0; k < 4; k++) std::sscanf(l.c_str(), fmt[k], &f[k]);
  };
  std::vector<int> ng;
  std::vector<std::vector<float>> gx, gz;
  scan(lines[0]);
  float hs0 = f[0];
  s.sx.push_back(f[0]);
  s.sz.push_back(f[1]);
  gx.push_back({f[2]});
  gz.push_back({f[3]});
  ng.push_back(0);
  int cs = 0, ig = 1;
  float hg0 = f[2], vg0 = f[3];
  for (size_t i = 1; i < lines.size(); i++) {
    scan(lines[i]);
    if (f[0] == hs0) {
      ig++;
    } else {
      cs++;
      ig = 1;
      s.sx.push_back(f[0]);
      s.sz.push_back(f[1]);
      gx.emplace_back();
      gz.emplace_back();
      ng.push_back(0);
    }
    ng[cs] = ig;
    gx[cs].push_back(f[2]);
    gz[cs].push_back(f[3]);
    hs0 = f[0];
    hg0 = f[2];
    vg0 = f[3];
  }
  if (hg0 == f[2] && vg0 == f[3]) ng[cs] = ig - 1;
  s.gx.resize(gx.size());
  s.gz.resize(gz.size());
  for (size_t is = 0; is < gx.size(); is++) {
    const size_t n = (size_t)std::max(ng[is], 0);
    s.gx[is].assign(gx[is].begin(), gx[is].begin() + std::min(n, gx[is].size()));
    s.gz[is].assign(gz[is].begin(), gz[is].begin() + std::min(n, gz[is].size()));
  }
  if (s.sx.empty()) {
    err = "no traces in " + path;
    return false;
  }
  return true;
}

bool read_floats(const std::string& path, size_t n, float* out, std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "can not open file : " + path;
    return false;
  }
  f.read(reinterpret_cast<char*>(out), (std::streamsize)(sizeof(float) * n));
  if ((size_t)f.gcount() != sizeof(float) * n) {
    err = "short read from " + path;
    return false;
  }
  return true;
}

bool write_floats(const std::string& path, size_t n, const float* data, std::string& err) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) {
    err = "can not open file: " + path;
    return false;
  }
  f.write(reinterpret_cast<const char*>(data), (std::streamsize)(sizeof(float) * n));
  return (bool)f;
}

namespace {
// Byte offsets of the SU trace-header fields this writer sets.
enum SuOff { TRACL = 0, TRACR = 4, FLDR = 8, EP = 16, CDP = 20, CDPT = 24, NS = 114, DT = 116, D1 = 180, F1 = 184,
             D2 = 188, F2 = 192 };
template <typename V>
void put(unsigned char* h, int off, V v) { std::memcpy(h + off, &v, sizeof(V)); }
}  // namespace

bool write_su(const std::string& path, int n1, int n2, const float* data, std::string& err) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) {
    err = "can not open file: " + path;
    return false;
  }
  unsigned char h[240];
  for (int i = 0; i < n2; i++) {
    std::memset(h, 0, sizeof(h));
    put<int>(h, TRACL, i + 1);
    put<int>(h, TRACR, i + 1);
    put<int>(h, FLDR, 1);
    put<int>(h, EP, 1);
    put<int>(h, CDP, i + 1);
    put<int>(h, CDPT, i + 1);
    put<unsigned short>(h, NS, (unsigned short)n1);
    put<unsigned short>(h, DT, (unsigned short)1000);
    put<float>(h, D1, 1.f);
    put<float>(h, F1, 0.f);
    put<float>(h, D2, 1.f);
    put<float>(h, F2, 1.f);
    f.write(reinterpret_cast<const char*>(h), 240);
    f.write(reinterpret_cast<const char*>(data + (size_t)i * n1), (std::streamsize)(sizeof(float) * n1));
  }
  return (bool)f;
}

// The expressions below keep the reference's float / double promotions.

Window shot_window(const Params& p, const Survey& s, int is) {
  const float dx = p.dx.val;
  const float sx = s.sx[is];
  const std::vector<float>& g = s.gx[is];
  const float gx_min = *std::min_element(g.begin(), g.end());
  const float gx_max = *std::max_element(g.begin(), g.end());
  const float x_min = std::min(sx, gx_min);
  const float x_max = std::max(sx, gx_max);
  const float x_min_pad = x_min - p.lpad.val * dx;
  const float x_max_pad = x_max + p.rpad.val * dx;
  const float velfx = p.velfx.val;
  const float x_min_vel = velfx + dx;
  const float x_max_vel = velfx + (p.nx.val - 1) * dx - dx;
  Window w;
  w.fx = x_min_pad;
  w.nx = (int)((x_max_pad - x_min_pad) / dx + 1);
  w.col.resize(std::max(w.nx, 0));
  for (int ix = 0; ix < w.nx; ix++) {
    const float xcor = w.fx + ix * dx;
    int xloc;
    if (xcor >= x_min_vel && xcor <= x_max_vel) xloc = (int)((xcor - velfx) / dx);
    else if (xcor < x_min_vel) xloc = (int)((x_min_vel - velfx) / dx);
    else xloc = (int)((x_max_vel - velfx) / dx);
    w.col[ix] = std::min(std::max(xloc, 0), p.nx.val - 1);
  }
  return w;
}

void window_model(const Window& w, int nz, const float* global, float* out) {
  for (int ix = 0; ix < w.nx; ix++)
    std::memcpy(out + (size_t)ix * nz, global + (size_t)w.col[ix] * nz, sizeof(float) * nz);
}

bool receiver_depths(const Params& p, const Survey& s, int is, float fx, int nx, std::vector<int>& igz,
                     std::string& err) {
  const float dx = p.dx.val, dz = p.dz.val;
  const float x_min = fx - 0.001 * dx;
  const float x_max = fx + (nx - 1) * dx + 0.001 * dx;
  igz.assign(nx, 0);
  for (int ig = 0; ig < s.ng(is); ig++) {
    const float x = s.gx[is][ig], z = s.gz[is][ig];
    if (x >= x_min && x <= x_max) {
      const int xloc = (int)((x - x_min) / dx + 0.1f);
      const int zloc = (int)((z - 0.0f) / dz + 0.1f);
      if (xloc < 0 || xloc >= nx || zloc < 0 || zloc >= p.nz.val) {
        err = "receiver outside the model window";
        return false;
      }
      igz[xloc] = zloc;
    }
  }
  return true;
}

void source_cell(const Params& p, const Survey& s, int is, float fx, int* sx, int* sz) {
  *sx = (int)((s.sx[is] - fx) / p.dx.val + 0.0001f);
  *sz = (int)((s.sz[is] - 0.0) / p.dz.val + 0.0001f);
}

bool receiver_columns(const Params& p, const Survey& s, int is, float fx, int nx, std::vector<int>& col,
                      std::string& err) {
  const int ng = s.ng(is);
  col.assign(ng, -1);
  const float dx = p.dx.val;
  const float x_min = fx - 0.001 * dx;
  const float x_max = fx + (nx - 1) * dx + 0.001 * dx;
  for (int ig = 0; ig < ng; ig++) {
    const float gx = s.gx[is][ig];
    if (x_min <= gx && gx <= x_max) {
      const int xloc = (int)((gx - x_min) / dx + 0.1f);
      if (xloc < 0 || xloc >= nx) {
        err = "receiver column outside the window";
        return false;
      }
      col[ig] = xloc;
    }
  }
  return true;
}

bool reorder(const Params& p, const Survey& s, int is, int nt, float* csg, float* rec, float fx, int nx,
             bool to_gather, std::string& err) {
  const int ng = s.ng(is);
  std::vector<int> col;
  if (!receiver_columns(p, s, is, fx, nx, col, err)) return false;
  if (to_gather) std::fill(csg, csg + (size_t)ng * nt, 0.f);
  else std::fill(rec, rec + (size_t)nx * nt, 0.f);
  for (int ig = 0; ig < ng; ig++) {
    if (col[ig] < 0) continue;
    if (to_gather) std::memcpy(csg + (size_t)ig * nt, rec + (size_t)col[ig] * nt, sizeof(float) * nt);
    else std::memcpy(rec + (size_t)col[ig] * nt, csg + (size_t)ig * nt, sizeof(float) * nt);
  }
  return true;
}

void window_to_global(const Params& p, const float* m, float* d, float mfx, float dfx, int nxm, int nxd, int nz,
                      bool add) {
  const int layer = 4;
  if (!add) std::fill(d, d + (size_t)nxd * nz, 0.f);
  const float dx = p.dx.val;
  const float x_min = mfx - 0.001 * dx;
  const float x_max = mfx + (nxm - 1) * dx + 0.001 * dx;
  for (int ix = layer; ix < nxd - layer; ix++) {
    const float xcor = dfx + (ix - 1) * dx;
    if (xcor >= x_min && xcor <= x_max) {
      const int xloc = (int)((xcor - x_min) / dx + 0.0001f);
      if (xloc >= 4 && xloc <= nxm - 4 && xloc < nxm)
        for (int iz = layer; iz < nz - layer; iz++) d[(size_t)ix * nz + iz] += m[(size_t)xloc * nz + iz];
    }
  }
}

bool remove_direct(const Params& p, const Survey& s, int is, float* csg, const float* v, std::string& err) {
  const int ngg = s.ng(is), nt = p.nt.val, nx = p.nx.val, nz = p.nz.val;
  const float sx = s.sx[is], sz = s.sz[is];
  const float dx = p.dx.val, dz = p.dz.val, dt = p.dt.val, fr = p.fr.val, velfx = p.velfx.val;
  const int delay = p.delay.val;
  const int isx = (int)((sx - velfx) / dx + 0.0001f);
  auto vel = [&](int i, int k) { return v[(size_t)i * nz + k]; };
  for (int ig = 0; ig < ngg; ig++) {
    float t = 0.0f;
    const float gx = s.gx[is][ig], gz = s.gz[is][ig];
    const int igx = (int)((gx - velfx) / dx + 0.0001f);
    const int igz = (int)((gz - 0.0f) / dz + 0.0001f);
    int adddelay = (int)std::fabs((gz - sz) / dz + 0.0001f);
    adddelay += 10;
    if (igz < 0 || igz >= nz) {
      err = "igz in removeDirect must be within range [0 nz)";
      return false;
    }
    // reference's branch structure, including its use of igx in the fallbacks
    if (igx <= isx) {
      for (int i = igx; i < isx; i++) {
        if (i >= 0 && i < nx) t += dx / vel(i, igz);
        else if (igx < 0) t += dx / vel(0, igz);
        else if (igx >= nx) t += dx / vel(nx - 1, igz);
      }
    } else {
      for (int i = igx; i > isx; i--) {
        if (i >= 0 && i < nx) t += dx / vel(i, igz);
        else if (igx < 0) t += dx / vel(0, igz);
        else if (igx >= nx) t += dx / vel(nx - 1, igz);
      }
    }
    int it = (int)((t / dt + delay + adddelay + 1.0f / fr / dt * 0.5f * 5.0) + 0.0001f);
    it = it <= (nt - 1) ? it : nt - 1;
    it = it >= 1 ? it : 1;
    std::fill(csg + (size_t)ig * nt, csg + (size_t)ig * nt + it, 0.f);
  }
  return true;
}

namespace {
std::string join(const std::string& a, const char* b, int is) {
  std::ostringstream o;
  o << a << b << is << ".dat";
  return o.str();
}
}  // namespace

std::string csg_name(const Params& p, int is) { return join(p.precsg.val, "", is); }
std::string born_name(const Params& p, int is) { return join(p.wdir.val, "BORN", is); }
std::string shot_image_name(const Params& p, int is) { return join(p.wdir.val, "CSG_IMAGE", is); }
std::string illum_name(const Params& p, int is) { return join(p.wdir.val, "ILLUM", is); }
std::string plane_source_name(const Params& p, int ip) { return join(p.wdir.val, "PLANE_S_NEW", ip); }
std::string plane_record_name(const Params& p, int ip) { return join(p.wdir.val, "PLANE_R_NEW", ip); }
std::string plane_image_name(const Params& p, int ip) { return join(p.wdir.val, "IMAGE", ip); }
std::string plane_illum_name(const Params& p, int ip) { return join(p.wdir.val, "ILLUM_PLANE", ip); }

}  // namespace seis
