// Shared device-side structures of the engine's kernels.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fd_cell.h"
#include "seis_internal.h"

namespace seis {

// ---------------------------------------------------------------------------
// Device-side shot descriptor.
template <typename T>
struct DevShot {
  const T* vel;    // [nx][nz]
  const T* dv;     // [nx][nz] (Born) or null
  const T* wav;    // [nt+ndel]
  const int* gz;   // [nx]
  T* rec;          // [nx][nt]
  T* fld;          // slot 0 grid origin; slot j at fld + j*fstride
  T* strips;       // [nt][8*(nx+nz)] (RTM)
  T* img;          // padded image (RTM) or illumination (ILLUM) accumulator, grid origin
  T* aux;          // NAUX padded scratch fields, grid origin, field j at
                   // aux + j*fstride (RTM: lap*2/v^3 of reverse step st in
                   // field st % NAUX)
  void* snap;      // Born-source snapshots [nt+ndel][nx][nz] (LSRTM)
  float* scale;    // per-snapshot scale [nt+ndel] (LSRTM, 16-bit snapshots)
  T* grad;         // per-shot image gradient [nx][nz] (LSRTM)
  long long fstride;
  int nx, NX, nt, ndel, sx, sz;
  int snap_kind;   // SNAP_NATIVE, SNAP_BF16 or SNAP_FP16
  int rtm_flags;   // SEIS_RTM_* bits (RTM reverse pass)
  int line_src;    // 1: plane-wave line source at row sz of every model
                   // column, wav = sou [nx][nt+ndel]; 0: point source
  T* raw;          // stencil values of the band cells of the new level, grid
                   // origin; second field at raw + fstride
  T* out;          // [nx][nz] destination of the cropped image (RTM, ILLUM)
  const unsigned* pmask;  // LSRTM: band_terms of the cells of ring P of this shot's width
  const T* wtab;          // LSRTM: sponge weight by band index (sponge_weight's i):
                          // wtab[0] = 1 (inside the model), wtab[i + 1], i <= pml
  int x0;          // LSRTM: first column of the shot's window in the image grid
};

// U buffers of the RTM reverse pass: the imaging pass applies two
// consecutive steps at once, in the launch of the step after them, which
// writes a third buffer (four, so that it never writes one being read).
constexpr int NAUX = 4;

// Step kernel modes.  MODEL / BORN / RTM_* follow the reference's modeling,
// forward and adjoint (fd.cpp:15-177, 592-719); SNAP_* run the background of
// the Born operator and record its per-step source coefficient; BORN_S
// propagates the scattered field from those records; ILLUM is the source
// illumination of illum (fd.cpp:275-339): modeling without receivers that
// accumulates u^2 of the model cells into DevShot::img.
// PW_BACK and PW_FWD are the two passes of the plane-wave imaging
// adjointPlane (fd.cpp:787-921): the receiver field backward in time saving
// its strips, then the line-source field forward beside the receiver field
// rebuilt forward in time.
enum Mode {
  MODEL = 0,
  BORN = 1,
  RTM_FWD = 2,
  RTM_BWD = 3,
  SNAP_MAX = 4,
  SNAP_STORE = 5,
  BORN_S = 6,
  ILLUM = 7,
  PW_BACK = 8,
  PW_FWD = 9,
  ADJ_X = 10  // LSRTM cross-correlation adjoint: receiver field backward, imaged against the snapshots
};

// Time index of launch step `step` of a shot in `mode`, or -1 when the shot
// has no such step: forward loops run it = step; the RTM reverse pass
// it = nt+ndel-1-step down to ndel; the plane-wave passes it = nt+ndel-1-step
// down to 0, then it = step+2 up to nt+ndel-1 (fd.cpp:829, 853).
template <typename T>
SEIS_HD inline int step_it(int mode, const DevShot<T>& s, int step) {
  const int NT = s.nt + s.ndel;
  int it;
  if (mode == RTM_BWD) {
    it = NT - 1 - step;
    return it >= s.ndel ? it : -1;
  }
  if (mode == PW_BACK || mode == ADJ_X) {
    it = NT - 1 - step;
    return it >= 0 ? it : -1;
  }
  it = mode == PW_FWD ? step + 2 : step;
  return it < NT ? it : -1;
}

// Storage of the Born-source snapshots vt2 * lap * 2/v^3 (model cells only).
// 16-bit forms store value / scale[step], with scale[step] the step's max |value|.
enum SnapKind { SNAP_NATIVE = 0, SNAP_BF16 = 1, SNAP_FP16 = 2 };

SEIS_HD inline unsigned short f32_to_bf16(float f) {
  unsigned int u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (unsigned short)(u >> 16);
}

SEIS_HD inline float bf16_to_f32(unsigned short h) {
  unsigned int u = (unsigned int)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

SEIS_HD inline unsigned short f32_to_f16(float f) {
  __half h = __float2half_rn(f);
  unsigned short u;
  __builtin_memcpy(&u, &h, 2);
  return u;
}

SEIS_HD inline float f16_to_f32(unsigned short u) {
  __half h;
  __builtin_memcpy(&h, &u, 2);
  return __half2float(h);
}

// Stored snapshot word of model cell j = ix*nz + iz at step it: the value
// (native storage) or its 16-bit pattern, loaded apart from its decoding so
// that a sweep can issue the load a column ahead.
template <typename T>
struct SnapWord {
  T v;
  unsigned short h;
};

template <typename T>
__device__ __forceinline__ SnapWord<T> snap_fetch(const DevShot<T>& s, int nz, int it, long long j) {
  const long long o = (long long)it * s.nx * nz + j;
  SnapWord<T> w{T(0), 0};
  if (s.snap_kind == SNAP_NATIVE)
    w.v = SEIS_GL(reinterpret_cast<const T*>(s.snap))[o];
  else
    w.h = SEIS_GL(reinterpret_cast<const unsigned short*>(s.snap))[o];
  return w;
}

// Decoded snapshot word with the step's scale `sc` (s.scale[it]; unused for
// native storage).  A sweep loads the scale once, before its loop.
template <typename T>
__device__ __forceinline__ T snap_decode_scaled(const DevShot<T>& s, float sc, SnapWord<T> w) {
  if (s.snap_kind == SNAP_NATIVE) return w.v;
  const float f = s.snap_kind == SNAP_BF16 ? bf16_to_f32(w.h) : f16_to_f32(w.h);
  return (T)(f * sc);
}

template <typename T>
__device__ __forceinline__ T snap_decode(const DevShot<T>& s, int it, SnapWord<T> w) {
  return snap_decode_scaled(s, s.snap_kind == SNAP_NATIVE ? 0.f : SEIS_GL(s.scale)[it], w);
}

// Snapshot value of model cell j = ix*nz + iz at step it.
template <typename T>
__device__ __forceinline__ T snap_load(const DevShot<T>& s, int nz, int it, long long j) {
  return snap_decode(s, it, snap_fetch(s, nz, it, j));
}

template <typename T>
__device__ __forceinline__ void snap_store(const DevShot<T>& s, int nz, int it, long long j, T v) {
  const long long o = (long long)it * s.nx * nz + j;
  if (s.snap_kind == SNAP_NATIVE) {
    SEIS_GL(reinterpret_cast<T*>(s.snap))[o] = v;
    return;
  }
  const float sc = SEIS_GL(s.scale)[it];
  const float q = sc > 0.f ? (float)v / sc : 0.f;
  SEIS_GL(reinterpret_cast<unsigned short*>(s.snap))[o] = s.snap_kind == SNAP_BF16 ? f32_to_bf16(q) : f32_to_f16(q);
}

// Indices of (x, z) in the boundary-strip store of one time sample
// (fd.cpp:1181-1225, layer = 4): top [nx][4], bottom [nx][4], left [4][nz],
// right [4][nz].  On grids narrower than 8 cells a cell can sit in several
// strips; every slot is returned so that save and restore stay consistent.
__device__ __forceinline__ int strip_slots(int x, int z, int nx, int nz, int pml, int* s) {
  const int ix = x - pml, iz = z - pml;
  if (ix < 0 || ix >= nx || iz < 0 || iz >= nz) return 0;
  int n = 0;
  if (iz < 4) s[n++] = ix * 4 + iz;
  if (iz >= nz - 4) s[n++] = nx * 4 + ix * 4 + (nz - 1 - iz);
  if (ix < 4) s[n++] = 8 * nx + ix * nz + iz;
  if (ix >= nx - 4) s[n++] = 8 * nx + 4 * nz + (nx - 1 - ix) * nz + iz;
  return n;
}

// Stores v at every slot strip_slots lists for model cell (ix, iz), without
// the slot list: one predicated store per strip, the side strips' conditions
// depending on the column only.
template <typename T, class Ptr>
__device__ __forceinline__ void store_strip_cell(Ptr st, int ix, int iz, int nx, int nz, T v) {
  if (ix < 0 || ix >= nx || iz < 0 || iz >= nz) return;
  if (iz < 4) st[ix * 4 + iz] = v;
  if (iz >= nz - 4) st[nx * 4 + ix * 4 + (nz - 1 - iz)] = v;
  if (ix < 4) st[8 * nx + ix * nz + iz] = v;
  if (ix >= nx - 4) st[8 * nx + 4 * nz + (nx - 1 - ix) * nz + iz] = v;
}

// The first slot strip_slots would return, or -1 (no private array).
__device__ __forceinline__ int first_strip_slot(int x, int z, int nx, int nz, int pml) {
  const int ix = x - pml, iz = z - pml;
  if (ix < 0 || ix >= nx || iz < 0 || iz >= nz) return -1;
  if (iz < 4) return ix * 4 + iz;
  if (iz >= nz - 4) return nx * 4 + ix * 4 + (nz - 1 - iz);
  if (ix < 4) return 8 * nx + ix * nz + iz;
  if (ix >= nx - 4) return 8 * nx + 4 * nz + (nx - 1 - ix) * nz + iz;
  return -1;
}

// Per-step, per-shot state shared by both paths of the step kernel.
template <typename T>
struct StepCtx {
  const T* P;       // previous level (slot 0 of the first field)
  const T* C;       // current level
  T* N;             // new level
  Model<T> m;
  int it, itp;
  long long src;    // field index of the source cell
  T* aux;           // RTM reverse: U of this step (one of NAUX buffers)
};

// Wavefront-local LDS exchange barrier (the hardware orders a wave's LDS
// operations; the fences keep the compiler from reordering across it).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int WAVES = 4;     // wavefronts per workgroup
constexpr int WG = 64 * WAVES;
constexpr int LROW = 72;     // one LDS row: 64 cells + 4 halo cells on each side

// Split of one shot's padded grid NX x NZ for a step kernel: the interior
// rectangle [xa, xb) x [za, zb), at least `m` cells from every edge, holds
// only cells with w = 1 and the 9-point stencil and is swept by wavefronts
// (64 z cells each) streaming along x; every other cell is a "band" cell
// updated by cell_update.  Shots of one launch share NZ and m, their widths
// NX may differ (per-shot windows), so each kernel forms its shot's split.
struct Split {
  int xa, xb, za, zb;
  long long n;       // band cells (outside the rectangle)
};

SEIS_HD inline Split shot_split(int m, int NX, int NZ) {
  Split q{m, NX - m, m, NZ - m, 0};
  if (q.xb <= q.xa || q.zb <= q.za) {  // no interior rectangle: every cell is a band cell
    q.xa = q.xb = 0;
    q.za = 0;
    q.zb = NZ;
  }
  q.n = (long long)NX * NZ - (long long)(q.xb - q.xa) * (q.zb - q.za);
  return q;
}

// Launch shape of one step over a batch: the sweep part (blockIdx.x <
// nfwg) covers the widest shot's interior, the band part (nbwg workgroups)
// the shot with the most band cells; workgroups beyond a narrower shot's
// extent exit at once.
struct Layout {
  int m;              // interior margin (shot_split)
  int NZ;             // padded depth shared by the shots
  int nzc, bx;        // z chunks of an interior, strip length of the sweep
  int NXmax;          // widest padded grid of the launch
  int nfwg;           // sweep workgroups per shot
  long long nbandmax; // most band cells of a shot
  int nbwg;           // band workgroups per shot
};

// x strips of a shot's interior at strip length bx.
SEIS_HD inline int shot_strips(const Split& q, int bx) { return q.xb > q.xa ? (q.xb - q.xa + bx - 1) / bx : 0; }

template <typename T>
Consts<T> make_consts(const GridSpec& g, long long ld) {
  Consts<T> k;
  k.nz = g.nz;
  k.pml = g.pml;
  k.NZ = g.nz + 2 * g.pml;
  k.ld = ld;
  k.dt = (T)g.dt;
  k.dx = (T)g.dx;
  k.dz = (T)g.dz;
  k.idx2 = T(1) / (k.dx * k.dx);
  k.idz2 = T(1) / (k.dz * k.dz);
  k.idt2 = T(1) / k.dt / k.dt;
  const T dt2 = k.dt * k.dt;
  k.cx = (T(1) / k.dx) * (T(1) / k.dx) * dt2;
  k.cz = (T(1) / k.dz) * (T(1) / k.dz) * dt2;
  k.a1z = T(1) / k.dz / T(2) / k.dt;
  k.a1x = T(1) / k.dx / T(2) / k.dt;
  k.a2b = T(1) / T(2) / k.dt / k.dt;
  k.b1 = T(1) / k.dz;
  k.b2 = T(1) / k.dx;
  k.sq2 = std::sqrt(T(2));
  k.c[0] = T(-205) / T(72);
  k.c[1] = T(8) / T(5);
  k.c[2] = T(-1) / T(5);
  k.c[3] = T(8) / T(315);
  k.c[4] = T(-1) / T(560);
  return k;
}

template <typename T>
inline void hilbert_taps_host(int half, T* h) {
  h[half] = T(0);
  for (int i = 1; i <= half; i++) {
    T taper = (T)(0.54 + 0.46 * std::cos(3.1415926 * i / half));
    T hi = (T)((double)(taper * T(i % 2)) * 2.0 / (3.1415926 * i));
    h[half + i] = hi;
    h[half - i] = -hi;
  }
}

// Field storage geometry: a margin of MX columns and field_mz(pml) rows of
// zeros before the padded grid (at least 4 rows after it), z stride rounded
// to 64 elements.  The top margin puts row max(pml, 4), where the interior
// sweep's 64-row segments start, on a 256-byte boundary (3% faster sweeps).
constexpr int MX = 4;

inline int field_mz(int pml) {
  const int m = std::max(pml, 4) % 64;
  int mz = (64 - m) % 64;
  if (mz < 4) mz += 64;
  return mz;
}

inline long long field_ld(int NZ, int pml) { return ((long long)(field_mz(pml) + NZ + 4) + 63) / 64 * 64; }

// Offset of cell (0, 0) of the padded grid in a field slot.
inline long long field_origin(long long ld, int pml) { return (long long)MX * ld + field_mz(pml); }

inline long long field_elems(int NX, long long ld) { return (long long)(NX + 2 * MX) * ld; }

// XCDs of the MI355X.  The dispatcher deals the workgroups of a launch
// round-robin over them, so along one row of the grid, blocks b and b' run on
// the same XCD (and share its L2) when b = b' mod NXCD.
constexpr int NXCD = 8;

// Interior tile of block b of the sweep (nfwg a multiple of NXCD): each XCD
// takes one contiguous run of tiles, so neighbouring strips of a z chunk,
// which read 8 columns in common, meet in the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nfwg) { return (b % NXCD) * (nfwg / NXCD) + b / NXCD; }

// Strip length bx of the interior sweep: the derived counts (the sweep's
// workgroup count is padded to a multiple of NXCD; the padding blocks exit).
inline Layout with_strip(Layout L, int bx) {
  L.bx = bx;
  const Split q = shot_split(L.m, L.NXmax, L.NZ);
  L.nzc = q.xb > q.xa ? (q.zb - q.za + 63) / 64 : 0;
  const int nxs = shot_strips(q, bx);
  L.nfwg = (L.nzc * nxs + WAVES - 1) / WAVES;
  L.nfwg = (L.nfwg + NXCD - 1) / NXCD * NXCD;
  return L;
}

// Layout of a launch over shots of padded widths NXs[0..n) and depth NZ with
// interior margin m; wg threads per band workgroup.
inline Layout make_layout(const int* NXs, int n, int NZ, int m, int wg = WG, int bx = 16) {
  Layout L{};
  L.m = m;
  L.NZ = NZ;
  L.NXmax = 0;
  L.nbandmax = 0;
  for (int i = 0; i < n; i++) {
    L.NXmax = std::max(L.NXmax, NXs[i]);
    L.nbandmax = std::max(L.nbandmax, shot_split(m, NXs[i], NZ).n);
  }
  L.nbwg = (int)((L.nbandmax + wg - 1) / wg);
  return with_strip(L, bx);
}

// Single-width form.
inline Layout make_layout(int NX, int NZ, int pml, int bx = 16) {
  return make_layout(&NX, 1, NZ, std::max(pml, 4), WG, bx);
}

// Strip-length autotuning (tune.cpp).  `trial(bx)` enqueues a few launches
// with strip length bx on `stream`; the first time a configuration `key` is
// seen each candidate is timed and the fastest kept for the process.  With
// tuning off (seis_set_tuning) the default strip length is used.
int tuned_strip(const char* kind, int NX, int NZ, int nshots, hipStream_t stream, void (*trial)(void*, int),
                void* ctx);

// The strip length the tuner chose for a configuration, or the default one,
// without timing anything.
int strip_lookup(const char* kind, int NX, int NZ, int nshots);

// Enumeration of the cells outside the interior rectangle of a Split: the
// rows above and below it over the full width, then the columns left and
// right of it (z fastest within each part).
using Ring = Split;

// j < R.n < 2^31 (the padded grid is checked to hold fewer cells): 32-bit
// unsigned divisions, far cheaper on the GPU than 64-bit ones.
__device__ __forceinline__ void ring_cell(const Ring& R, int NX, int NZ, long long jl, int* x, int* z) {
  const unsigned j = (unsigned)jl;
  const unsigned top = (unsigned)NX * (unsigned)R.za, bot = (unsigned)NX * (unsigned)(NZ - R.zb);
  const unsigned hz = (unsigned)(R.zb - R.za);
  const unsigned lft = (unsigned)R.xa * hz;
  if (j < top) {
    *x = (int)(j / (unsigned)R.za);
    *z = (int)(j % (unsigned)R.za);
  } else if (j < top + bot) {
    const unsigned q = j - top, h = (unsigned)(NZ - R.zb);
    *x = (int)(q / h);
    *z = R.zb + (int)(q % h);
  } else if (j < top + bot + lft) {
    const unsigned q = j - top - bot;
    *x = (int)(q / hz);
    *z = R.za + (int)(q % hz);
  } else {
    const unsigned q = j - top - bot - lft;
    *x = R.xb + (int)(q / hz);
    *z = R.za + (int)(q % hz);
  }
}

// Launches one step of `mode` for every shot of a batch (engine.hip): the
// step kernel (every cell; band cells leave their stencil value in
// DevShot::raw) followed by the band kernel (one-way update and blend of the
// band cells from those values).
template <typename T>
void launch_step(int mode, const Consts<T>& k, const DevShot<T>* dev, const Layout& L, const Layout& B, int nshots,
                 int step, hipStream_t stream);

// Band shape of a launch: the cells closer than pml to an edge of each shot
// (shot_split with margin pml), one thread each.
inline Layout band_layout(const int* NXs, int n, int NZ, int pml) { return make_layout(NXs, n, NZ, pml, WG); }

}  // namespace seis
