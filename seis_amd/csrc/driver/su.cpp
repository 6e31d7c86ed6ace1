// SU interop of the reference: the SU writer writeSu (util/io.cpp:8-58) and
// the SU-to-gathers utility get_info (zzz/get_info/get_info.c:32-214), which
// splits an SU file into the per-shot <dir>CSG<i>.dat files, the coordinates
// file the driver reads (sx sz gx gz per trace) and a statistics table.
//
// The SU trace header is the standard 240-byte SEG-Y trace header with the
// CWP extensions, native byte order (mysu.h).  Only the fields these two
// programs read or write are addressed, by their byte offsets.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/seis.h"
#include "../seis_internal.h"

namespace {

using seis::set_error;

// byte offsets in the SU trace header (mysu.h)
constexpr int kHeader = 240;
constexpr int OFF_TRACL = 0, OFF_TRACR = 4, OFF_FLDR = 8, OFF_EP = 16, OFF_CDP = 20, OFF_CDPT = 24;
constexpr int OFF_GELEV = 40, OFF_SELEV = 44, OFF_SX = 72, OFF_SY = 76, OFF_GX = 80, OFF_GY = 84;
constexpr int OFF_NS = 114, OFF_DT = 116, OFF_D1 = 180, OFF_F1 = 184, OFF_D2 = 188, OFF_F2 = 192;

template <typename T>
void put(unsigned char* h, int off, T v) {
  std::memcpy(h + off, &v, sizeof(T));
}

template <typename T>
T get(const unsigned char* h, int off) {
  T v;
  std::memcpy(&v, h + off, sizeof(T));
  return v;
}

// glibc's rand() in a process that never called srand (seed 1): the TYPE_3
// additive feedback generator r[i] = r[i-31] + r[i-3] (mod 2^32) over a
// state seeded by the Park-Miller generator 16807 r mod (2^31 - 1), with the
// first 310 outputs discarded; rand() returns r[i] >> 1.  get_info's
// source-position jitter comes from this sequence.
class GlibcRand {
 public:
  GlibcRand() {
    std::vector<uint32_t> r(344);
    r[0] = 1;
    for (int i = 1; i < 31; i++) {
      const int64_t v = (16807LL * (int64_t)(int32_t)r[i - 1]) % 2147483647LL;
      r[i] = (uint32_t)(v < 0 ? v + 2147483647LL : v);
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (int i = 34; i < 344; i++) r[i] = r[i - 31] + r[i - 3];
    for (int i = 0; i < 34; i++) ring_[(310 + i) % 34] = r[310 + i];
    i_ = 344;
  }
  int next() {
    const uint32_t v = ring_[(i_ - 31) % 34] + ring_[(i_ - 3) % 34];
    ring_[i_ % 34] = v;
    i_++;
    return (int)(v >> 1);
  }

 private:
  uint32_t ring_[34];
  long long i_;
};

struct File {
  FILE* f = nullptr;
  File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// The coordinates line of one trace (get_info.c:109-112, 146-147).
void coord_line(FILE* f, float sxf, float szf, double gx, float gzf) {
  std::fprintf(f, "%12.2f %12.2f %12.2f %12.2f\n", (double)sxf, (double)szf, gx, (double)gzf);
}

// Source x of a new shot: sx * scalco plus a jitter in [0, 2) from rand()
// (get_info.c:109, 152), in double, stored as float.
float jittered(int sx, double scalco, GlibcRand& rng) {
  return (float)(sx * scalco + rng.next() % 1000 * 1.0 / 1000 * 2);
}

}  // namespace

extern "C" {

int seis_write_su(const char* path, int n1, int n2, const float* data) {
  if (!path || !data || n1 <= 0 || n2 < 0) return set_error("seis_write_su: bad argument"), SEIS_ERR_ARG;
  File fw(path, "wb");
  if (!fw.f) return set_error(std::string("seis_write_su: cannot open ") + path), SEIS_ERR_ARG;
  unsigned char h[kHeader];
  std::memset(h, 0, sizeof(h));
  put<float>(h, OFF_D1, 1.f);
  put<float>(h, OFF_D2, 1.f);
  put<float>(h, OFF_F1, 0.f);
  put<float>(h, OFF_F2, 1.f);
  put<uint16_t>(h, OFF_NS, (uint16_t)n1);
  put<int32_t>(h, OFF_EP, 1);
  put<int32_t>(h, OFF_FLDR, 1);
  const int ntr = n2 == 0 ? 1 : n2;
  // the one-trace form stores dt = 1000000 into the 16-bit field
  // (io.cpp:19), the n2-trace form 1000 (io.cpp:43)
  put<uint16_t>(h, OFF_DT, (uint16_t)(n2 == 0 ? 1000000 : 1000));
  for (int i = 0; i < ntr; i++) {
    put<int32_t>(h, OFF_TRACL, i + 1);
    put<int32_t>(h, OFF_TRACR, i + 1);
    put<int32_t>(h, OFF_CDP, i + 1);
    put<int32_t>(h, OFF_CDPT, i + 1);
    if (std::fwrite(h, 1, kHeader, fw.f) != (size_t)kHeader ||
        std::fwrite(data + (size_t)i * n1, sizeof(float), (size_t)n1, fw.f) != (size_t)n1)
      return set_error("seis_write_su: write failed"), SEIS_ERR_ARG;
  }
  return SEIS_OK;
}

int seis_su_split(const char* su_file, const char* info_file, const char* coord_file, const char* dir,
                  const char* key) {
  if (!su_file || !info_file || !coord_file || !dir || !key) return set_error("seis_su_split: null argument"), SEIS_ERR_ARG;
  File in(su_file, "rb"), info(info_file, "w"), cord(coord_file, "w");
  if (!in.f || !info.f || !cord.f) return set_error("seis_su_split: cannot open a file"), SEIS_ERR_ARG;
  // the shot key is an int header field: fldr by default (get_info.c:46-54)
  const std::string k(key);
  const int key_off = k == "sx" ? OFF_SX : k == "ep" ? OFF_EP : k == "cdp" ? OFF_CDP : OFF_FLDR;
  unsigned char tr[kHeader];
  std::memset(tr, 0, sizeof(tr));
  if (std::fread(tr, 1, kHeader, in.f) != (size_t)kHeader) return set_error("seis_su_split: no trace"), SEIS_ERR_ARG;
  const int dt = get<uint16_t>(tr, OFF_DT), nt = get<uint16_t>(tr, OFF_NS);
  const long nseek = (long)nt * 4 + kHeader;
  std::fseek(in.f, 0L, SEEK_END);
  const long fsize = std::ftell(in.f);
  const long traceall = fsize / nseek;
  auto shot = [&] { return get<int32_t>(tr, key_off); };
  // reads the header at byte pos as fread would: whatever lies there
  // overwrites the start of the previous header (nothing at the end)
  auto read_header = [&](long pos) {
    std::fseek(in.f, pos, SEEK_SET);
    const size_t got = std::fread(tr, 1, kHeader, in.f);
    (void)got;
  };
  std::vector<unsigned char> buf((size_t)nt * 4);
  auto copy_data = [&](FILE* out) {
    const size_t got = std::fread(buf.data(), 1, buf.size(), in.f);
    (void)got;
    std::fwrite(buf.data(), 1, buf.size(), out);
  };
  // extent of a receiver and its source (get_info.c:66-80)
  auto span = [&](int a, int b, int* lo, int* hi) {
    if (a < b) {
      *lo = a;
      *hi = b;
    } else {
      *lo = b;
      *hi = a;
    }
  };
  GlibcRand rng;
  int olds = shot();
  int xmin, xmax, ymin, ymax;
  span(get<int32_t>(tr, OFF_GX), get<int32_t>(tr, OFF_SX), &xmin, &xmax);
  span(get<int32_t>(tr, OFF_GY), get<int32_t>(tr, OFF_SY), &ymin, &ymax);
  int sx = get<int32_t>(tr, OFF_SX), sy = get<int32_t>(tr, OFF_SY);
  const double scalco = 1;  // get_info.c:84 overrides the header's scalco
  int xminall = xmin, xmaxall = xmax, yminall = ymin, ymaxall = ymax;
  static const char* info1 = "    CSP(SP)     NTR       NT        DT     XMINALL   XMAXALL   YMINALL   YMAXALL";
  static const char* info2 =
      "        SP       NTR     BEGIN       END        SX       SY       XMIN      XMAX      YMIN      YMAX";
  std::fprintf(info.f, "%s\n", info1);
  std::fprintf(info.f, "%10d%10d%10d%10d%10d%10d%10d%10d\n", 0, 0, 0, 0, 0, 0, 0, 0);
  std::fprintf(info.f, "%s\n", info2);
  int ishot = 0;
  auto csg_path = [&](int i) { return std::string(dir) + "CSG" + std::to_string(i) + ".dat"; };
  FILE* csg = std::fopen(csg_path(ishot++).c_str(), "w");
  if (!csg) return set_error("seis_su_split: cannot open a gather file"), SEIS_ERR_ARG;
  (void)rng.next();  // get_info.c:106 prints one rand() value first
  float sxf = jittered(sx, scalco, rng);
  float szf = (float)get<int32_t>(tr, OFF_SELEV);
  float gzf = (float)get<int32_t>(tr, OFF_GELEV);
  coord_line(cord.f, sxf, szf, get<int32_t>(tr, OFF_GX) * scalco, gzf);
  long pos = 0;
  read_header(pos);
  copy_data(csg);
  long kk = 1;
  int ntr = 1, maxntr = 1, starttrace = 1, shotnumber = 0;
  do {
    pos += nseek;
    read_header(pos);
    const bool more = kk <= traceall - 1;
    if (more) {
      if (shot() != olds) {
        std::fclose(csg);
        csg = std::fopen(csg_path(ishot++).c_str(), "w");
        if (!csg) return set_error("seis_su_split: cannot open a gather file"), SEIS_ERR_ARG;
      }
      copy_data(csg);
    }
    ntr++;
    if (more && shot() == olds) {
      gzf = (float)get<int32_t>(tr, OFF_GELEV);
      coord_line(cord.f, sxf, szf, get<int32_t>(tr, OFF_GX) * scalco, gzf);
    }
    if (shot() != olds) {
      shotnumber++;
      olds = shot();
      sxf = jittered(get<int32_t>(tr, OFF_SX), scalco, rng);
      szf = (float)get<int32_t>(tr, OFF_SELEV);
      gzf = (float)get<int32_t>(tr, OFF_GELEV);
      coord_line(cord.f, sxf, szf, get<int32_t>(tr, OFF_GX) * scalco, gzf);
      if (maxntr < ntr - 1) maxntr = ntr - 1;
      std::fprintf(info.f, "%10d%10d%10d%10d%10d%10d%10d%10d%10d%10d\n", shotnumber, ntr - 1, starttrace,
                   starttrace + ntr - 2, sx, sy, xmin, xmax, ymin, ymax);
      starttrace = (int)kk + 1;
      sx = get<int32_t>(tr, OFF_SX);
      sy = get<int32_t>(tr, OFF_SY);
      ntr = 1;
      span(get<int32_t>(tr, OFF_GX), get<int32_t>(tr, OFF_SX), &xmin, &xmax);
      span(get<int32_t>(tr, OFF_GY), get<int32_t>(tr, OFF_SY), &ymin, &ymax);
    }
    kk++;
    const int gx = get<int32_t>(tr, OFF_GX), gy = get<int32_t>(tr, OFF_GY);
    if (gx < xmin) xmin = gx;
    if (gx > xmax) xmax = gx;
    if (gy < ymin) ymin = gy;
    if (gy > ymax) ymax = gy;
    if (xminall > xmin) xminall = xmin;
    if (xmaxall < xmax) xmaxall = xmax;
    if (yminall > ymin) yminall = ymin;
    if (ymaxall < ymax) ymaxall = ymax;
  } while (kk <= traceall);
  shotnumber++;
  std::fclose(csg);
  if (maxntr < ntr - 1) maxntr = ntr - 1;
  std::fprintf(info.f, "%10d%10d%10d%10d%10d%10d%10d%10d%10d%10d\n", shotnumber, ntr - 1, starttrace,
               starttrace + ntr - 2, sx, sy, xmin, xmax, ymin, ymax);
  std::rewind(info.f);
  std::fprintf(info.f, "%s\n", info1);
  std::fprintf(info.f, "%10d%10d%10d%10d%10d%10d%10d%10d\n", shotnumber, maxntr, nt, dt, xminall, xmaxall, yminall,
               ymaxall);
  return SEIS_OK;
}

}  // extern "C"
