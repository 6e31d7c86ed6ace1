// C ABI of the engine (include/seis.h): argument checking, host<->device
// staging for the single-shot entry points, and dispatch to the batched
// runners.  No CPU fallback exists: without a GPU every compute entry point
// returns SEIS_ERR_NODEV or SEIS_ERR_HIP.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "seis_internal.h"

namespace seis {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

namespace {

int hip_fail(const char* what, hipError_t e) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorNoDevice ? SEIS_ERR_NODEV : SEIS_ERR_HIP;
}

// Device buffer owned for the duration of a single-shot call.
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

int upload(DevBuf& b, const void* src, size_t bytes) {
  hipError_t e = hipMalloc(&b.p, bytes ? bytes : 1);
  if (e != hipSuccess) return hip_fail("hipMalloc", e);
  if (bytes) {
    e = hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail("hipMemcpy", e);
  }
  return SEIS_OK;
}

int alloc(DevBuf& b, size_t bytes) {
  hipError_t e = hipMalloc(&b.p, bytes ? bytes : 1);
  if (e != hipSuccess) return hip_fail("hipMalloc", e);
  e = hipMemset(b.p, 0, bytes ? bytes : 1);
  if (e != hipSuccess) return hip_fail("hipMemset", e);
  return SEIS_OK;
}

int check_single(int nt, int ndel, int nx, int nz, int pml, int sx, int sz, const int* gz, const void* wav,
                 const void* vel, const void* out) {
  if (!gz || !wav || !vel || !out) return set_error("null array argument"), SEIS_ERR_ARG;
  if (nt <= 0 || ndel < 0 || nx <= 0 || nz <= 0 || pml < 0)
    return set_error("nt, nx, nz must be > 0; ndel, pml >= 0"), SEIS_ERR_ARG;
  if (sx < 0 || sx >= nx || sz < 0 || sz >= nz) return set_error("source cell outside the model"), SEIS_ERR_ARG;
  for (int i = 0; i < nx; i++)
    if (gz[i] < 0 || gz[i] >= nz) return set_error("receiver depth gz outside the model"), SEIS_ERR_ARG;
  return SEIS_OK;
}

int check_batch(const seis_grid* g, int nshots, const seis_shot* shots) {
  if (!g) return set_error("null grid"), SEIS_ERR_ARG;
  if (nshots < 0 || (nshots > 0 && !shots)) return set_error("bad shot list"), SEIS_ERR_ARG;
  return SEIS_OK;
}

int single(int mode, float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
           const int* gz, const float* wav, const float* vel, const float* extra, float* out) {
  int rc = check_single(nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, out);
  if (rc) return rc;
  if (mode != RUN_MODEL && !extra) return set_error("null array argument"), SEIS_ERR_ARG;
  const size_t nm = (size_t)nx * nz, nr = (size_t)nx * nt;
  DevBuf dgz, dwav, dvel, dextra, drec, dimg;
  if ((rc = upload(dgz, gz, sizeof(int) * nx))) return rc;
  if ((rc = upload(dwav, wav, sizeof(float) * (nt + ndel)))) return rc;
  if ((rc = upload(dvel, vel, sizeof(float) * nm))) return rc;
  if (mode == RUN_BORN && (rc = upload(dextra, extra, sizeof(float) * nm))) return rc;
  if (mode == RUN_RTM) {
    if ((rc = upload(drec, extra, sizeof(float) * nr))) return rc;
    if ((rc = alloc(dimg, sizeof(float) * nm))) return rc;
  } else if ((rc = alloc(drec, sizeof(float) * nr))) {
    return rc;
  }
  seis_grid g{nx, nz, pml, dt, dx, dz};
  seis_shot s{sx, sz, nt, ndel, 0, 0, 0, 0};
  rc = run_batch<float>(mode, g, 1, &s, (const float*)dvel.p, (const float*)dextra.p, (const float*)dwav.p,
                        (const int*)dgz.p, (float*)drec.p, (float*)dimg.p, 0);
  if (rc) return rc;
  hipError_t e = mode == RUN_RTM ? hipMemcpy(out, dimg.p, sizeof(float) * nm, hipMemcpyDeviceToHost)
                                 : hipMemcpy(out, drec.p, sizeof(float) * nr, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail("hipMemcpy", e);
  return SEIS_OK;
}

// Plane-wave engine: one shot, host buffers.
int plane_single(int kind, float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                 const float* vel, const float* sou, const float* extra, float* out) {
  if (!vel || !sou || !out || (kind != 1 && !extra)) return set_error("null array argument"), SEIS_ERR_ARG;
  if (nt <= 0 || ndel < 0 || nx <= 0 || nz <= 0 || pml < 0)
    return set_error("nt, nx, nz must be > 0; ndel, pml >= 0"), SEIS_ERR_ARG;
  if (sz < 0 || sz >= nz) return set_error("source row outside the model"), SEIS_ERR_ARG;
  if (kind != 1 && (gz < 0 || gz >= nz)) return set_error("receiver row outside the model"), SEIS_ERR_ARG;
  const size_t nm = (size_t)nx * nz, nr = (size_t)nx * nt, ns = (size_t)nx * (nt + ndel);
  const std::vector<int> gzs(nx, kind != 1 ? gz : 0);
  DevBuf dgz, dsou, dvel, dextra, dout;
  int rc;
  if ((rc = upload(dgz, gzs.data(), sizeof(int) * nx))) return rc;
  if ((rc = upload(dsou, sou, sizeof(float) * ns))) return rc;
  if ((rc = upload(dvel, vel, sizeof(float) * nm))) return rc;
  const size_t nout = kind == 0 ? nr : nm;
  if (kind == 0 && (rc = upload(dextra, extra, sizeof(float) * nm))) return rc;  // dv
  if (kind == 2 && (rc = upload(dextra, extra, sizeof(float) * nr))) return rc;  // rec
  if ((rc = alloc(dout, sizeof(float) * nout))) return rc;
  seis_grid g{nx, nz, pml, dt, dx, dz};
  seis_shot s{0, sz, nt, ndel, 0, 0, 0, 0};
  if (kind == 0)
    rc = run_batch<float>(RUN_BORN, g, 1, &s, (const float*)dvel.p, (const float*)dextra.p, (const float*)dsou.p,
                          (const int*)dgz.p, (float*)dout.p, nullptr, 0, 0, 1);
  else if (kind == 1)
    rc = run_batch<float>(RUN_ILLUM, g, 1, &s, (const float*)dvel.p, nullptr, (const float*)dsou.p, nullptr, nullptr,
                          (float*)dout.p, 0, 0, 1);
  else
    rc = run_rtm_plane<float>(g, 1, &s, (const float*)dvel.p, (const float*)dsou.p, (const int*)dgz.p,
                              (const float*)dextra.p, (float*)dout.p, 0);
  if (rc) return rc;
  hipError_t e = hipMemcpy(out, dout.p, sizeof(float) * nout, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SEIS_OK : hip_fail("hipMemcpy", e);
}

}  // namespace
}  // namespace seis

using namespace seis;

extern "C" {

const char* seis_last_error(void) { return g_last_error.c_str(); }

int seis_device_count(int* count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (count) *count = (e == hipSuccess) ? n : 0;
  if (e != hipSuccess) return hip_fail("hipGetDeviceCount", e);
  return SEIS_OK;
}

int seis_set_device(int device) {
  hipError_t e = hipSetDevice(device);
  return e == hipSuccess ? SEIS_OK : hip_fail("hipSetDevice", e);
}

int seis_malloc(void** ptr, size_t bytes) {
  if (!ptr) return set_error("null pointer"), SEIS_ERR_ARG;
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  return e == hipSuccess ? SEIS_OK : hip_fail("hipMalloc", e);
}

int seis_free(void* ptr) {
  hipError_t e = hipFree(ptr);
  return e == hipSuccess ? SEIS_OK : hip_fail("hipFree", e);
}

int seis_memcpy_htod(void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? SEIS_OK : hip_fail("hipMemcpy", e);
}

int seis_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SEIS_OK : hip_fail("hipMemcpy", e);
}

int seis_memset(void* dst, int value, size_t bytes) {
  hipError_t e = hipMemset(dst, value, bytes);
  return e == hipSuccess ? SEIS_OK : hip_fail("hipMemset", e);
}

}  // extern "C"

namespace {
// Plain streaming copy: 16 B per lane per iteration, grid-stride.
__global__ __launch_bounds__(256) void copy_f32x4_kernel(float4* __restrict__ dst, const float4* __restrict__ src,
                                                         long long n4) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n4; i += (long long)gridDim.x * 256)
    dst[i] = src[i];
}
}  // namespace

extern "C" {

int seis_copy_f32(float* dst, const float* src, long long n) {
  if (n < 0 || ((n & 3) != 0)) return set_error("seis_copy_f32: n must be a non-negative multiple of 4"), SEIS_ERR_ARG;
  if (((uintptr_t)dst | (uintptr_t)src) & 15) return set_error("seis_copy_f32: pointers must be 16-byte aligned"), SEIS_ERR_ARG;
  if (n == 0) return SEIS_OK;
  hipLaunchKernelGGL(copy_f32x4_kernel, dim3(4096), dim3(256), 0, 0, reinterpret_cast<float4*>(dst),
                     reinterpret_cast<const float4*>(src), n / 4);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SEIS_OK : hip_fail("copy_f32x4_kernel", e);
}

int seis_synchronize(void) {
  hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? SEIS_OK : hip_fail("hipDeviceSynchronize", e);
}

int seis_modeling(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
                  const int* gz, const float* wav, const float* vel, float* rec) {
  return single(RUN_MODEL, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, nullptr, rec);
}

int seis_born(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
              const int* gz, const float* wav, const float* vel, const float* dv, float* rec) {
  return single(RUN_BORN, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, dv, rec);
}

int seis_rtm(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
             const int* gz, const float* wav, const float* vel, const float* rec, float* img) {
  return single(RUN_RTM, dt, dx, dz, nt, ndel, nx, nz, pml, sx, sz, gz, wav, vel, rec, img);
}

int seis_illum(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
               const float* wav, const float* vel, float* ss) {
  if (!wav || !vel || !ss) return set_error("null array argument"), SEIS_ERR_ARG;
  if (nt <= 0 || ndel < 0 || nx <= 0 || nz <= 0 || pml < 0)
    return set_error("nt, nx, nz must be > 0; ndel, pml >= 0"), SEIS_ERR_ARG;
  if (sx < 0 || sx >= nx || sz < 0 || sz >= nz) return set_error("source cell outside the model"), SEIS_ERR_ARG;
  const size_t nm = (size_t)nx * nz;
  DevBuf dwav, dvel, dss;
  int rc;
  if ((rc = upload(dwav, wav, sizeof(float) * (nt + ndel)))) return rc;
  if ((rc = upload(dvel, vel, sizeof(float) * nm))) return rc;
  if ((rc = alloc(dss, sizeof(float) * nm))) return rc;
  seis_grid g{nx, nz, pml, dt, dx, dz};
  seis_shot s{sx, sz, nt, ndel, 0, 0, 0, 0};
  rc = run_batch<float>(RUN_ILLUM, g, 1, &s, (const float*)dvel.p, nullptr, (const float*)dwav.p, nullptr, nullptr,
                        (float*)dss.p, 0);
  if (rc) return rc;
  hipError_t e = hipMemcpy(ss, dss.p, sizeof(float) * nm, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SEIS_OK : hip_fail("hipMemcpy", e);
}

int seis_illum_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel, const float* wav,
                     float* ss) {
  int rc = check_batch(grid, nshots, shots);
  if (rc) return rc;
  return run_batch<float>(RUN_ILLUM, *grid, nshots, shots, vel, nullptr, wav, nullptr, nullptr, ss, 0);
}


int seis_born_plane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                    const float* vel, const float* sou, const float* dv, float* rec) {
  return plane_single(0, dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, vel, sou, dv, rec);
}

int seis_illum_plane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                     const float* vel, const float* sou, float* ss) {
  return plane_single(1, dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, vel, sou, nullptr, ss);
}

int seis_rtm_plane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                   const float* vel, const float* sou, const float* rec, float* img) {
  return plane_single(2, dt, dx, dz, nt, ndel, nx, nz, pml, sz, gz, vel, sou, rec, img);
}

int seis_born_plane_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                          const float* dv, const float* sou, const int* gz, float* rec) {
  int rc = check_batch(grid, nshots, shots);
  if (rc) return rc;
  return run_batch<float>(RUN_BORN, *grid, nshots, shots, vel, dv, sou, gz, rec, nullptr, 0, 0, 1);
}

int seis_illum_plane_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                           const float* sou, float* ss) {
  int rc = check_batch(grid, nshots, shots);
  if (rc) return rc;
  return run_batch<float>(RUN_ILLUM, *grid, nshots, shots, vel, nullptr, sou, nullptr, nullptr, ss, 0, 0, 1);
}

int seis_rtm_plane_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                         const float* sou, const int* gz, const float* rec, float* img) {
  int rc = check_batch(grid, nshots, shots);
  if (rc) return rc;
  return run_rtm_plane<float>(*grid, nshots, shots, vel, sou, gz, rec, img, 0);
}

int seis_modeling_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                        const float* wav, const int* gz, float* rec) {
  int rc = check_batch(grid, nshots, shots);
  if (rc) return rc;
  return run_batch<float>(RUN_MODEL, *grid, nshots, shots, vel, nullptr, wav, gz, rec, nullptr, 0);
}

int seis_born_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                    const float* dv, const float* wav, const int* gz, float* rec) {
  int rc = check_batch(grid, nshots, shots);
  if (rc) return rc;
  return run_batch<float>(RUN_BORN, *grid, nshots, shots, vel, dv, wav, gz, rec, nullptr, 0);
}

int seis_rtm_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                   const float* wav, const int* gz, const float* rec, float* img) {
  int rc = check_batch(grid, nshots, shots);
  if (rc) return rc;
  return run_batch<float>(RUN_RTM, *grid, nshots, shots, vel, nullptr, wav, gz, const_cast<float*>(rec), img, 0);
}

int seis_rtm_batch_ex(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                      const float* wav, const int* gz, const float* rec, float* img, int flags) {
  int rc = check_batch(grid, nshots, shots);
  if (rc) return rc;
  if (flags & ~SEIS_RTM_ISOLATE_BAND) return set_error("rtm: unknown option bits"), SEIS_ERR_ARG;
  return run_batch<float>(RUN_RTM, *grid, nshots, shots, vel, nullptr, wav, gz, const_cast<float*>(rec), img, 0,
                          flags);
}

int seis_harmonic(const float* in, float* out, int n, float kx, float x) {
  if (!in || !out) return set_error("null array argument"), SEIS_ERR_ARG;
  if (n <= 0) return set_error("harmonic: n must be > 0"), SEIS_ERR_ARG;
  DevBuf din, dout;
  int rc;
  if ((rc = upload(din, in, sizeof(float) * n))) return rc;
  if ((rc = alloc(dout, sizeof(float) * n))) return rc;
  if ((rc = run_harmonic((const float*)din.p, (float*)dout.p, n, kx, x, 0))) return rc;
  hipError_t e = hipMemcpy(out, dout.p, sizeof(float) * n, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SEIS_OK : hip_fail("hipMemcpy", e);
}

int seis_plane_prepare(int nx, int nt, int ntw, int ns, int np, float pmin, float dp, const float* sx,
                       const int* isx, const int* ng, const int* col, const float* wav, const float* data,
                       float* sou, float* rec) {
  if (nx <= 0 || nt <= 0 || ntw <= 0 || ns < 0 || np < 0)
    return set_error("plane prepare: nx, nt, ntw must be > 0; ns, np >= 0"), SEIS_ERR_ARG;
  if (ns > 0 && (!sx || !isx || !ng || !col)) return set_error("null array argument"), SEIS_ERR_ARG;
  long long ntr = 0;
  for (int s = 0; s < ns; s++) {
    if (isx[s] < 0 || isx[s] >= nx) return set_error("plane prepare: source column outside [0, nx)"), SEIS_ERR_ARG;
    if (ng[s] < 0) return set_error("plane prepare: negative trace count"), SEIS_ERR_ARG;
    ntr += ng[s];
  }
  if (ntr > 0x7fffffffLL) return set_error("plane prepare: too many traces"), SEIS_ERR_ARG;
  for (long long i = 0; i < ntr; i++)
    if (col[i] < -1 || col[i] >= nx) return set_error("plane prepare: trace column outside [-1, nx)"), SEIS_ERR_ARG;
  if (np == 0) return SEIS_OK;
  if ((ntr > 0 && !data) || !wav || !sou || !rec) return set_error("null array argument"), SEIS_ERR_ARG;
  return run_plane_prepare(nx, nt, ntw, ns, np, pmin, dp, sx, isx, ng, col, wav, data, sou, rec, 0);
}

}  // extern "C"
