// Plane-wave synthesis of shot gathers (planeWavePrepare,
// lsrtm/inversion.cpp:808-912): for every ray parameter p_k = pmin + k dp,
//
//   sou_k[isx_s] += harmonic(wav, p_k, sx_s)          per shot s, in shot order
//   rec_k[col]   += harmonic(trace, p_k, sx_s)        per shot, one trace per column
//
// with harmonic(in, p, x) = cos(p x) in - sin(p x) H[in] (InitInversion.cpp:
// 141-151) and H the 201-tap Hamming-tapered Hilbert filter (hilbert.h).
//
// The Hilbert transform of a trace does not depend on p, so the GPU filters
// every trace and the wavelet once (hilbert_kernel), then one stacking pass
// (plane_stack_kernel) forms cos*d - sin*H[d] for a block of ray parameters
// per thread and sums the shots.  The cosines and sines are taken on the host
// with the C library's cosf / sinf of the float product p*sx, as the
// reference does, so every output is bit-identical to the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.h"
#include "prof.h"
#include "seis_internal.h"

namespace seis {

namespace {

constexpr int HARM_HALF = 100;             // Hilbert<float> hl(100), InitInversion.cpp:145
constexpr int HARM_ODD = HARM_HALF / 2;    // nonzero taps on each side (odd offsets)
constexpr int HWG = 256;                   // hilbert_kernel work-group
constexpr int HTILE = 1024;                // outputs per work-group (512 of each parity)
constexpr int HPAR = HTILE / 2 + 104;      // inputs per parity array (multiple of 4)

// Taps at odd offsets d = 1, 3, ..., 99: h[L + d]; h[L - d] = -h[L + d].
struct HarmTaps {
  float h[HARM_ODD];
};

// H[in] of each row of in [rows][n] -> out [rows][n].  out[j] is the
// reference's sum of in[i] * h[j - i + L] over i in increasing order
// (hilbert.h:42-54).  Terms whose tap is zero (even j - i) and terms outside
// the trace are left out: each would add a zero product to a partial sum that
// is never -0, which leaves it unchanged (for finite samples).
//
// Output j of a tile at j0 uses the inputs j - 99, j - 97, ..., j + 99: all of
// one parity.  The tile's inputs are split by parity into two LDS arrays so
// that output n of parity p (j = j0 + p + 2n) reads A_p[n .. n + 99] with
//   A_0[k] = in[j0 - 99 + 2k],  A_1[k] = in[j0 - 98 + 2k].
// A thread computes the outputs n = 4t .. 4t + 3 of one parity from
// A_p[4t .. 4t + 102], read as 26 aligned float4 (conflict-free), each input
// added into every accumulator whose window holds it, in increasing order.
__global__ __launch_bounds__(HWG) void hilbert_kernel(HarmTaps taps, const float* __restrict__ in,
                                                      float* __restrict__ out, int n) {
  __shared__ __align__(16) float A[2][HPAR];
  __shared__ float O[HTILE];
  const long long row = blockIdx.y;
  const int j0 = blockIdx.x * HTILE;
  const float* src = in + row * n;
  for (int k = threadIdx.x; k < 2 * HPAR; k += HWG) {
    const int p = k / HPAR, kk = k - p * HPAR;
    const int t = j0 - 99 + p + 2 * kk;
    A[p][kk] = (t >= 0 && t < n) ? src[t] : 0.f;
  }
  __syncthreads();
  const int p = threadIdx.x >> 7;   // parity of this thread's outputs
  const int t = threadIdx.x & 127;  // outputs n = 4t .. 4t + 3
  const float4* a4 = reinterpret_cast<const float4*>(&A[p][4 * t]);
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int v = 0; v < 26; v++) {
    const float4 q = a4[v];
    const float x[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int w = 4 * v + e;  // input A_p[4t + w]
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const int u = w - m;  // its index in output (4t + m)'s window
        if (u >= 0 && u < 100) {
          // h[L + 99 - 2u]: taps.h[49 - u] for u < 50, -taps.h[u - 50] after
          if (u < 50)
            acc[m] = acc[m] + x[e] * taps.h[49 - u];
          else
            acc[m] = acc[m] + x[e] * (-taps.h[u - 50]);
        }
      }
    }
  }
#pragma unroll
  for (int m = 0; m < 4; m++) O[p + 2 * (4 * t + m)] = acc[m];
  __syncthreads();
  for (int k = threadIdx.x; k < HTILE; k += HWG) {
    const int j = j0 + k;
    if (j < n) out[row * n + j] = O[k];
  }
}

constexpr int SWG = 256;  // plane_stack_kernel work-group
constexpr int KB = 8;     // ray parameters per thread

// out[k][x][t] for k in [k0, k0 + KB): the sum over shots s (in order) of
// cs[k][s].cos * d[t] - cs[k][s].sin * H[d][t], d the row map[s][x] of in (no
// term where map[s][x] < 0).  span[x] bounds the shots that hold a row in
// column x ([first, last + 1), empty for none), so the loop visits only those,
// still in shot order.  One work-group per (t tile, column x, k block).
// With `single` the output is the one shot's term itself (harmonic), not a
// sum started from zero (which turns a -0 term into +0).
__global__ __launch_bounds__(SWG) void plane_stack_kernel(const float* __restrict__ in, const float* __restrict__ hin,
                                                          int n, const int* __restrict__ map,
                                                          const int2* __restrict__ span, int nx,
                                                          const float2* __restrict__ cs, int ns, int np, int single,
                                                          float* __restrict__ out) {
  const int t = blockIdx.x * SWG + threadIdx.x;
  const int x = blockIdx.y;
  const int k0 = blockIdx.z * KB;
  const int kn = min(KB, np - k0);
  float acc[KB];
#pragma unroll
  for (int b = 0; b < KB; b++) acc[b] = 0.f;
  if (t < n) {
    const int2 sp = span[x];
    for (int s = sp.x; s < sp.y; s++) {
      const int r = map[(long long)s * nx + x];
      if (r < 0) continue;
      const float d = in[(long long)r * n + t];
      const float hd = hin[(long long)r * n + t];
#pragma unroll
      for (int b = 0; b < KB; b++) {
        if (b < kn) {
          const float2 c = cs[(long long)(k0 + b) * ns + s];
          const float v = c.x * d - c.y * hd;
          acc[b] = single ? v : v + acc[b];
        }
      }
    }
#pragma unroll
    for (int b = 0; b < KB; b++)
      if (b < kn) out[((long long)(k0 + b) * nx + x) * n + t] = acc[b];
  }
}

HarmTaps harm_taps() {
  std::vector<float> h(2 * HARM_HALF + 1);
  hilbert_taps_host<float>(HARM_HALF, h.data());
  HarmTaps t{};
  for (int k = 0; k < HARM_ODD; k++) t.h[k] = h[HARM_HALF + 2 * k + 1];
  return t;
}

int hip_err(const char* what, hipError_t e) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorNoDevice ? SEIS_ERR_NODEV : SEIS_ERR_HIP;
}

struct Scratch {
  void* p = nullptr;
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
};

int launch_hilbert(const float* in, float* out, int n, long long rows, hipStream_t st) {
  if (rows <= 0 || n <= 0) return SEIS_OK;
  const HarmTaps taps = harm_taps();
  for (long long r0 = 0; r0 < rows; r0 += 65535) {
    const int nr = (int)std::min<long long>(65535, rows - r0);
    dim3 grid((n + HTILE - 1) / HTILE, nr);
    timed_launch(st, "plane_hilbert", 8.0 * n * nr, 0, [&] {
      hilbert_kernel<<<grid, HWG, 0, st>>>(taps, in + r0 * n, out + r0 * n, n);
    });
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SEIS_OK : hip_err("hilbert_kernel", e);
}

// `placed`: (shot, column) pairs of the map that hold a row.  Algorithmic
// bytes: every placed row's samples and their H read once per block of KB
// ray parameters, every output sample written once.
int launch_stack(const float* in, const float* hin, int n, const int* map, const int2* span, int ns, int nx,
                 const float2* cs, int np, float* out, hipStream_t st, double placed, int single = 0) {
  if (np <= 0 || nx <= 0 || n <= 0) return SEIS_OK;
  const int kblocks = (np + KB - 1) / KB;
  dim3 grid((n + SWG - 1) / SWG, nx, kblocks);
  timed_launch(st, "plane_stack", 8.0 * n * placed * kblocks + 4.0 * n * nx * np, 0, [&] {
    plane_stack_kernel<<<grid, SWG, 0, st>>>(in, hin, n, map, span, nx, cs, ns, np, single, out);
  });
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SEIS_OK : hip_err("plane_stack_kernel", e);
}

// [first, last + 1) of the shots s with map[s][x] >= 0, per column x.
std::vector<int2> column_spans(const std::vector<int>& map, int ns, int nx) {
  std::vector<int2> sp(nx, int2{0, 0});
  for (int x = 0; x < nx; x++)
    for (int s = 0; s < ns; s++)
      if (map[(size_t)s * nx + x] >= 0) {
        if (sp[x].y == 0) sp[x].x = s;
        sp[x].y = s + 1;
      }
  return sp;
}

}  // namespace

int run_harmonic(const float* in, float* out, int n, float kx, float x, hipStream_t st) {
  Scratch h, cs, map, span;
  hipError_t e = hipMalloc(&h.p, sizeof(float) * n);
  if (e != hipSuccess) return hip_err("hipMalloc", e);
  const float ang = kx * x;
  const float2 c{cosf(ang), sinf(ang)};
  const int zero = 0;
  const int2 one{0, 1};
  if ((e = hipMalloc(&cs.p, sizeof(float2))) != hipSuccess || (e = hipMalloc(&map.p, sizeof(int))) != hipSuccess ||
      (e = hipMalloc(&span.p, sizeof(int2))) != hipSuccess)
    return hip_err("hipMalloc", e);
  if ((e = hipMemcpyAsync(cs.p, &c, sizeof(c), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(map.p, &zero, sizeof(int), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(span.p, &one, sizeof(int2), hipMemcpyHostToDevice, st)) != hipSuccess)
    return hip_err("hipMemcpyAsync", e);
  int rc = launch_hilbert(in, (float*)h.p, n, 1, st);
  if (!rc)
    rc = launch_stack(in, (const float*)h.p, n, (const int*)map.p, (const int2*)span.p, 1, 1, (const float2*)cs.p, 1,
                      out, st, 1.0, 1);
  if (rc) return rc;
  e = hipStreamSynchronize(st);
  return e == hipSuccess ? SEIS_OK : hip_err("plane harmonic", e);
}

int run_plane_prepare(int nx, int nt, int ntw, int ns, int np, float pmin, float dp, const float* sx,
                      const int* isx, const int* ng, const int* col, const float* wav, const float* data,
                      float* sou, float* rec, hipStream_t st) {
  long long ntr = 0;
  for (int s = 0; s < ns; s++) ntr += ng[s];
  // per-shot row maps: the source wavelet (row 0) at column isx, and for the
  // record the last trace of the shot that lands in each column (swapReord
  // copies the traces in order, inversion.cpp:626-643)
  std::vector<int> smap((size_t)ns * nx, -1), rmap((size_t)ns * nx, -1);
  long long r0 = 0;
  for (int s = 0; s < ns; s++) {
    smap[(size_t)s * nx + isx[s]] = 0;
    for (int g = 0; g < ng[s]; g++)
      if (col[r0 + g] >= 0) rmap[(size_t)s * nx + col[r0 + g]] = (int)(r0 + g);
    r0 += ng[s];
  }
  double placed = 0;
  for (int v : rmap) placed += v >= 0;
  const std::vector<int2> sspan = column_spans(smap, ns, nx), rspan = column_spans(rmap, ns, nx);
  // cos / sin of p_k * sx_s; p_k = pmin + k * dp in float (inversion.cpp:866)
  std::vector<float2> cs((size_t)np * ns);
  for (int k = 0; k < np; k++) {
    const float p = pmin + (float)k * dp;
    for (int s = 0; s < ns; s++) {
      const float ang = p * sx[s];
      cs[(size_t)k * ns + s] = float2{cosf(ang), sinf(ang)};
    }
  }
  Scratch hw, hd, dcs, dsm, drm, dss, drs;
  hipError_t e;
  if ((e = hipMalloc(&hw.p, sizeof(float) * std::max(ntw, 1))) != hipSuccess ||
      (e = hipMalloc(&hd.p, sizeof(float) * std::max<long long>(ntr * nt, 1))) != hipSuccess ||
      (e = hipMalloc(&dcs.p, sizeof(float2) * std::max<size_t>(cs.size(), 1))) != hipSuccess ||
      (e = hipMalloc(&dsm.p, sizeof(int) * smap.size())) != hipSuccess ||
      (e = hipMalloc(&drm.p, sizeof(int) * rmap.size())) != hipSuccess ||
      (e = hipMalloc(&dss.p, sizeof(int2) * nx)) != hipSuccess || (e = hipMalloc(&drs.p, sizeof(int2) * nx)) != hipSuccess)
    return hip_err("hipMalloc", e);
  if ((e = hipMemcpyAsync(dcs.p, cs.data(), sizeof(float2) * cs.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(dsm.p, smap.data(), sizeof(int) * smap.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(drm.p, rmap.data(), sizeof(int) * rmap.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(dss.p, sspan.data(), sizeof(int2) * nx, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(drs.p, rspan.data(), sizeof(int2) * nx, hipMemcpyHostToDevice, st)) != hipSuccess)
    return hip_err("hipMemcpyAsync", e);
  int rc = launch_hilbert(wav, (float*)hw.p, ntw, 1, st);
  if (!rc) rc = launch_hilbert(data, (float*)hd.p, nt, ntr, st);
  if (!rc)
    rc = launch_stack(wav, (const float*)hw.p, ntw, (const int*)dsm.p, (const int2*)dss.p, ns, nx,
                      (const float2*)dcs.p, np, sou, st, ns);
  if (!rc)
    rc = launch_stack(data, (const float*)hd.p, nt, (const int*)drm.p, (const int2*)drs.p, ns, nx,
                      (const float2*)dcs.p, np, rec, st, placed);
  if (rc) return rc;
  e = hipStreamSynchronize(st);
  return e == hipSuccess ? SEIS_OK : hip_err("plane prepare", e);
}

}  // namespace seis
