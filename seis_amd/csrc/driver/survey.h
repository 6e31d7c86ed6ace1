// Acquisition geometry, binary/SU I/O and the per-shot glue of the
// inversion driver (reference lsrtm/inversion.cpp:552-752, util/coord.cpp,
// util/io.h, util/io.cpp).
#pragma once
#include <string>
#include <vector>

#include "params.h"

namespace seis {

// Coordinates file: one trace per line "sx sz gx gz" (metres); a shot is a
// run of consecutive lines with the same sx, read with the reference's
// counting rules (coord.cpp:22-176; see read_coords).
struct Survey {
  std::vector<float> sx, sz;               // per shot
  std::vector<std::vector<float>> gx, gz;  // per shot, per receiver
  int ns() const { return (int)sx.size(); }
  int ng(int is) const { return (int)gx[is].size(); }
  int ngmax() const;
};
bool read_coords(const std::string& path, Survey& s, std::string& err);

// Raw native-endian arrays (io.h:8-70).
bool read_floats(const std::string& path, size_t n, float* out, std::string& err);
bool write_floats(const std::string& path, size_t n, const float* data, std::string& err);
// SU file: per trace a 240-byte header (ns = n1, dt = 1000 us, d1 = d2 = 1,
// f2 = 1, ep = fldr = 1, tracl = tracr = cdp = cdpt = trace number) followed by
// n1 float32 samples (io.cpp:34-58).
bool write_su(const std::string& path, int n1, int n2, const float* data, std::string& err);

// Per-shot window of the global model (getVel, inversion.cpp:552-598).
struct Window {
  float fx;               // x of the window's first column
  int nx;                 // columns
  std::vector<int> col;   // global model column of every window column
};
Window shot_window(const Params& p, const Survey& s, int is);
// Copies the window's columns out of a global [nx][nz] model.
void window_model(const Window& w, int nz, const float* global, float* out);

// Receiver depth (cells) of every window column; columns without a
// receiver get 0 (getIgz, inversion.cpp:600-625).
bool receiver_depths(const Params& p, const Survey& s, int is, float fx, int nx, std::vector<int>& igz,
                     std::string& err);

// Source cell of a shot in its window (inversion.cpp:186-187).
void source_cell(const Params& p, const Survey& s, int is, float fx, int* sx, int* sz);

// Window column of every receiver of shot is, as swapReord places them
// (inversion.cpp:626-643): -1 outside [fx, fx + (nx-1) dx] (with a 0.001 dx
// margin); false when a column falls outside [0, nx).
bool receiver_columns(const Params& p, const Survey& s, int is, float fx, int nx, std::vector<int>& col,
                      std::string& err);

// Gather <-> per-column record (swapReord, inversion.cpp:626-643).
//   to_gather: csg[ig][t] = rec[column of receiver ig][t]
//   otherwise: rec[column][t] = csg[ig][t] (zero for columns without a receiver)
bool reorder(const Params& p, const Survey& s, int is, int nt, float* csg, float* rec, float fx, int nx,
             bool to_gather, std::string& err);

// Window image -> global image (swapModel, inversion.cpp:644-665): global
// column ix in [4, nx-4) takes window column ((fx_g + (ix-1) dx) - fx_w)/dx
// when that lies in [4, nxw-4]; rows [4, nz-4).  Adds when `add`.
void window_to_global(const Params& p, const float* m, float* d, float mfx, float dfx, int nxm, int nxd, int nz,
                      bool add);

// Mutes the direct arrival of a gather in place (removeDirect,
// inversion.cpp:699-752); v is the global [nx][nz] model.
bool remove_direct(const Params& p, const Survey& s, int is, float* csg, const float* v, std::string& err);

// File names of the driver (inversion.cpp:667-697).
std::string csg_name(const Params& p, int is);         // <precsg><is>.dat
std::string born_name(const Params& p, int is);        // <wdir>BORN<is>.dat
std::string shot_image_name(const Params& p, int is);  // <wdir>CSG_IMAGE<is>.dat
std::string illum_name(const Params& p, int is);       // <wdir>ILLUM<is>.dat
// Plane-wave files, per ray parameter (inversion.cpp:335, 896-897, 1068).
std::string plane_source_name(const Params& p, int ip);  // <wdir>PLANE_S_NEW<ip>.dat
std::string plane_record_name(const Params& p, int ip);  // <wdir>PLANE_R_NEW<ip>.dat
std::string plane_image_name(const Params& p, int ip);   // <wdir>IMAGE<ip>.dat
std::string plane_illum_name(const Params& p, int ip);   // <wdir>ILLUM_PLANE<ip>.dat

}  // namespace seis
