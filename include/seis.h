/*
 * seis.h -- C ABI of the MI355X (gfx950) 2D acoustic finite-difference engine.
 *
 * Drop-in boundary for the reference's engine entry points (bbsun/Seis
 * src/util/fd.h:3-5): forward shot modeling, Born modeling and RTM, with the
 * reference's arguments, array layouts and results, plus batched variants that
 * take device-resident buffers and run many shots per launch.
 *
 * Layouts (reference memory.h:29-44, fd.cpp:17-18): a 2D array a[i2][i1] is
 * contiguous with i1 fastest.  Models and images are [nx][nz] (z fastest),
 * records [nx][nt] (t fastest), the wavelet has nt + ndel samples.  Grid
 * indices (sx, sz, gz[]) are cells of the unpadded model; the absorbing band
 * adds pml cells on every side.
 *
 * Every call returns SEIS_OK (0) or an error code; seis_last_error() gives the
 * message of the calling thread's last failure.  The reference exits with
 * status 0 on errors (debug.cpp:5-12); this ABI never exits.
 */
#ifndef SEIS_H
#define SEIS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEIS_OK 0
#define SEIS_ERR_ARG 1    /* invalid argument */
#define SEIS_ERR_HIP 2    /* HIP runtime failure */
#define SEIS_ERR_NODEV 3  /* no usable GPU */

/* Model grid shared by every shot of a batch: nx columns, nz rows, pml
 * absorbing cells per side, time step dt and spacings dx, dz (fd.h:3). */
typedef struct seis_grid {
  int nx, nz, pml;
  float dt, dx, dz;
} seis_grid;

/* One shot of a batch.  sx, sz: source cell (fd.h:3 sx, sz); nt: record
 * length; ndel: pre-roll steps before recording starts (fd.h:3 ndel).
 * nx: columns of this shot's own grid, its model window (the per-shot
 * window of getVel, inversion.cpp:552-598); 0 means the grid's nx.  Shots
 * of different widths share one launch; all share the grid's nz, pml, dt,
 * dx, dz.  Below, nx is the shot's width.
 * model / wav / gz / rec: element offsets of this shot's [nx][nz] model
 * (and perturbation), wavelet (nt + ndel samples), receiver-depth row (nx
 * entries, one receiver per column as in fd.cpp:58-63) and [nx][nt] record
 * inside the batch arrays.
 * x0: LSRTM plans only (seis_lsrtm_create): the shot's window is columns
 * [x0, x0 + nx) of the plan's image grid (grid->nx columns); its
 * perturbation is read from and its gradient added to those columns.  The
 * other calls ignore it. */
typedef struct seis_shot {
  int sx, sz, nt, ndel;
  long long model, wav, gz, rec;
  int nx, x0;
} seis_shot;

/* Message of the calling thread's last error ("" if none). */
const char* seis_last_error(void);

/* Device management and device buffers (for device-resident batches). */
int seis_device_count(int* count);
int seis_set_device(int device);
int seis_malloc(void** ptr, size_t bytes);
int seis_free(void* ptr);
int seis_memcpy_htod(void* dst, const void* src, size_t bytes);
int seis_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int seis_memset(void* dst, int value, size_t bytes);
/* dst[i] = src[i] for n floats (device pointers, 16-byte aligned, n % 4 == 0)
 * on the default stream: a plain streaming copy, used to calibrate HBM
 * traffic counters.  No counterpart in the reference. */
int seis_copy_f32(float* dst, const float* src, long long n);
int seis_synchronize(void);

/* Forward shot modeling; replaces `modeling` (fd.h:3, fd.cpp:15-80).
 * Host buffers: gz[nx], wav[nt+ndel], vel[nx][nz] in; rec[nx][nt] out
 * (caller-allocated; the reference returns a new MyAlloc array instead). */
int seis_modeling(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
                  const int* gz, const float* wav, const float* vel, float* rec);

/* Born modeling; replaces `forward` (fd.h:4, fd.cpp:81-177).
 * dv[nx][nz] is the velocity perturbation; rec[nx][nt] the scattered record. */
int seis_born(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
              const int* gz, const float* wav, const float* vel, const float* dv, float* rec);

/* RTM image with the reference's imaging condition; replaces `adjoint`
 * (fd.h:5, fd.cpp:592-719).  rec[nx][nt] in, img[nx][nz] out. */
int seis_rtm(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
             const int* gz, const float* wav, const float* vel, const float* rec, float* img);

/* Source illumination; replaces `illum` (fd.h:10, fd.cpp:275-339): the
 * modeling loop without receivers, ss[ix][iz] = sum over steps it > ndel of
 * u^2 after the absorbing boundary.  wav[nt+ndel], vel[nx][nz] in;
 * ss[nx][nz] out. */
int seis_illum(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sx, int sz,
               const float* wav, const float* vel, float* ss);

/* Plane-wave engine (fd.h:6-7, 11).  The source is the line of row sz: every
 * model column ix fires its own time function sou[ix][it-1] (sou is
 * [nx][nt+ndel]); receivers sit at row gz of every column.
 *   seis_born_plane  replaces forwordPlane (fd.cpp:178-274): rec[nx][nt];
 *   seis_illum_plane replaces illumPlane   (fd.cpp:720-786): ss[nx][nz]
 *                    (gz is not used, as in the reference);
 *   seis_rtm_plane   replaces adjointPlane (fd.cpp:787-921): img[nx][nz].
 *                    The reference's two `if (itp>=0);` guards there end at
 *                    their semicolons, so with ndel > 0 it also injects
 *                    samples from before the record (reading outside it);
 *                    those samples are skipped here (identical for ndel = 0). */
int seis_born_plane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                    const float* vel, const float* sou, const float* dv, float* rec);
int seis_illum_plane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                     const float* vel, const float* sou, float* ss);
int seis_rtm_plane(float dt, float dx, float dz, int nt, int ndel, int nx, int nz, int pml, int sz, int gz,
                   const float* vel, const float* sou, const float* rec, float* img);

/* Ricker wavelet of the reference (wavelet.cpp:7-18): out[i] for i < n,
 * peak frequency fr, peak at sample num_delay. */
int seis_ricker(float dt, float fr, int num_delay, int n, float* out);

/* Parameter-file driver: the reference's Inversion flow (lsrtm/inversion.cpp)
 * over one GPU.  param_file uses the reference's grammar and keys
 * (param_*.txt2, datastr.h:51-75); mode is
 *   "modeling"  records of model_file (default: vfile) -> <precsg><is>.dat
 *   "born"      Born records of perturbation model_file -> <wdir>BORN<is>.dat
 *   "rtm"       image of the <precsg> records       (adjoint_MPI, RTM_IMG)
 *   "rtm_born"  image of the <wdir>BORN records     (adjoint_MPI, LSRTM_STEP)
 *   "lsrtm"     least-squares RTM, maxiter CG iterations
 *   "lsrtm_xcorr" the same with the cross-correlation adjoint
 *               (SEIS_LSRTM_ADJ_XCORR)
 *   "lsrtm_illum" lsrtm preconditioned by the source illumination of all
 *               shots, W = sqrt(max / (ss + 1e-3 max)) (seis_lsrtm_set_weight);
 *               the per-shot ILLUM<is>.dat are written on the way
 *   "illum"     source illumination of every shot -> <wdir>ILLUM<is>.dat,
 *               and their sum (illum_MPI)
 * and, with planeTag != 0, per ray parameter p_k = pmin + k dp (k < np):
 *   "plane_prepare"  plane-wave sources and records of the <precsg> gathers
 *                    -> <wdir>PLANE_S_NEW<k>.dat, <wdir>PLANE_R_NEW<k>.dat
 *                    (planeWavePrepare); writes no image
 *   "rtm_plane"      image of each PLANE_R_NEW<k> -> <wdir>IMAGE<k>.dat, and
 *                    their sum (adjointPlane_MPI, RTM_IMG)
 *   "rtm_plane_born" the same of the <wdir>BORN<k> records (LSRTM_STEP)
 *   "illum_plane"    plane-wave illumination -> <wdir>ILLUM_PLANE<k>.dat, and
 *                    their sum (illumPlane_MPI)
 * Images are written to out_file (default <wdir>RTM_IMAGE.dat,
 * <wdir>LSRTM_IMAGE.dat, <wdir>ILLUM.dat, <wdir>PLANE_IMAGE.dat or
 * <wdir>ILLUM_PLANE.dat) as float32 [nx][nz]. */
int seis_run(const char* param_file, const char* mode, const char* model_file, const char* out_file);

/* SU interop.  seis_write_su writes n2 traces of n1 float samples
 * (data[n2][n1]) as the reference's writeSu does (io.cpp:34-58: d1 = d2 = 1,
 * f2 = 1, dt = 1000, fldr = ep = 1, tracl = tracr = cdp = cdpt = trace
 * number); n2 = 0 writes the single trace data[n1] of its one-trace form
 * (io.cpp:8-33, whose dt of 1000000 lands in the 16-bit field as 16960).
 * seis_su_split does what the reference's get_info utility does
 * (zzz/get_info/get_info.c:32-214): traces of su_file grouped by the header
 * key ("fldr" default, "ep", "cdp" or "sx") go to <dir>CSG<i>.dat, the
 * coordinates file gets "sx sz gx gz" per trace (sx with get_info's rand()
 * jitter in [0, 2), sz = selev, gz = gelev) and info_file the per-shot
 * statistics table. */
int seis_write_su(const char* path, int n1, int n2, const float* data);
int seis_su_split(const char* su_file, const char* info_file, const char* coord_file, const char* dir,
                  const char* key);

/* Parses a parameter file like Inversion::readParamFile (inversion.cpp:38-79,
 * grammar of parser.cpp:16-168) and writes "key = value" lines into out
 * (out_len bytes, NUL-terminated; floats with 9 significant digits, np / dp /
 * pmin only read with planeTag != 0, maskfile only required with mask != 0);
 * errors name the missing key. */
int seis_read_params(const char* param_file, char* out, int out_len);

/* Reads a coordinates file as Coords does (coord.cpp:22-176): one trace per
 * line "sx sz gx gz", a shot is a run of lines with the same sx.  *ns and
 * *ntraces get the shot and trace counts; when they fit max_shots /
 * max_traces, ng[ns] the traces per shot, sxz[ns][2] each shot's (sx, sz)
 * and gxz[ntraces][2] each trace's (gx, gz), shot after shot. */
int seis_read_coords(const char* coord_file, int max_shots, int max_traces, int* ns, int* ntraces, int* ng,
                     float* sxz, float* gxz);

/* Batched variants: all array arguments are DEVICE pointers (seis_malloc);
 * shots[] is a host array.  Shots run concurrently in the same launches;
 * each produces exactly what the single-shot call produces for it.
 * Replace the per-shot worker loops of modeling_MPI / forward_MPI /
 * adjoint_MPI (lsrtm/inversion.cpp:141-206, 475-551, 365-474). */
int seis_modeling_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                        const float* wav, const int* gz, float* rec);
int seis_born_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                    const float* dv, const float* wav, const int* gz, float* rec);
/* seis_modeling_batch without device allocation, for callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.
 * Same arguments and results as seis_modeling_batch; the wavefields and shot
 * descriptors live in a static device scratch of 64 MiB per GPU (held until
 * the process ends), which holds 4 padded fields per shot, e.g. 39 shots
 * of 300 x 200 cells with pml 14 at once; larger batches run in
 * consecutive chunks of the scratch.  A shot whose fields alone exceed the
 * scratch is SEIS_ERR_ARG.  Every call on a GPU shares that GPU's scratch:
 * calls that overlap on one GPU (other host threads) race.  The strip length
 * is the one the tuner chose earlier for the configuration, else 16; no
 * profiler events are recorded. */
int seis_modeling_batch_static(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                               const float* wav, const int* gz, float* rec);

/* img: nshots consecutive [nx][nz] images, one per shot in shot order, each
 * of its shot's width (the per-shot CSG_IMAGE<is> of inversion.cpp:434-439). */
int seis_rtm_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                   const float* wav, const int* gz, const float* rec, float* img);

/* ss: nshots consecutive [nx][nz] illuminations, one per shot of its width (the
 * per-shot ILLUM<is> of illum_MPI, inversion.cpp:207-290, which sums them);
 * the shots' gz and rec offsets are not used. */
int seis_illum_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                     const float* wav, float* ss);

/* Batch memory.  Every shot of a launch holds its wavefields (and, for RTM,
 * its boundary strips: nt * 8 * (nx + nz) floats) for the whole run, so each
 * batch call cuts its shot list into consecutive chunks whose working set
 * fits a budget and runs them one after another; every shot's output is the
 * same for any cut.  The budget is the free device memory less a reserve of
 * max(1 GiB, 1/32 of it), capped by seis_set_batch_limits(bytes, shots):
 * bytes > 0 caps the working set of a chunk, shots > 0 the shots per launch
 * (0, 0 restores the defaults).  A shot larger than the budget is
 * SEIS_ERR_ARG.  The reference images one shot per MPI rank at a time
 * (masterRun, inversion.cpp:753-807); no counterpart of these two calls.
 * seis_batch_plan computes the cut a batch call would make for `budget`
 * bytes and `max_shots` (0: no cap), without a GPU: chunk c holds shots
 * [starts[c], starts[c+1]) for c < *nchunks (starts: nshots + 1 ints). */
#define SEIS_PLAN_MODEL 0      /* seis_modeling_batch */
#define SEIS_PLAN_BORN 1       /* seis_born_batch, seis_born_plane_batch */
#define SEIS_PLAN_RTM 2        /* seis_rtm_batch(_ex) */
#define SEIS_PLAN_ILLUM 3      /* seis_illum_batch, seis_illum_plane_batch */
#define SEIS_PLAN_RTM_PLANE 4  /* seis_rtm_plane_batch */
int seis_set_batch_limits(size_t bytes, int shots);
int seis_batch_plan(int mode, const seis_grid* grid, int nshots, const seis_shot* shots, size_t budget,
                    int max_shots, int* starts, int* nchunks);

/* Plane-wave batches: shots[].wav is the offset of each shot's sou
 * [nx][nt+ndel], shots[].sz its source row, shots[].gz the offset of its
 * receiver depths gz[nx] (one row for the reference's plane-wave functions);
 * shots[].sx is not used, and shots[].nx must be 0 or the grid's nx (a plane
 * wave spans the model).  Outputs per shot as in the single-shot calls. */
int seis_born_plane_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                          const float* dv, const float* sou, const int* gz, float* rec);
int seis_illum_plane_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                           const float* sou, float* ss);
int seis_rtm_plane_batch(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                         const float* sou, const int* gz, const float* rec, float* img);

/* Plane-wave phase shift of one trace, harmonic (InitInversion.cpp:141-151,
 * declared InitInversion.h:8): out[t] = cos(kx x) in[t] - sin(kx x) H[in][t],
 * H the 201-tap Hamming-tapered Hilbert filter (hilbert.h, half length 100),
 * cos and sin taken in float of the float product kx*x.  Host buffers of n
 * samples. */
int seis_harmonic(const float* in, float* out, int n, float kx, float x);

/* Plane-wave synthesis of shot gathers, planeWavePrepare (inversion.cpp:
 * 808-912), for all ray parameters p_k = pmin + k*dp (k < np) in one call.
 * Host arrays:   sx[ns]   source x of each shot (metres);
 *                isx[ns]  source column, in [0, nx);
 *                ng[ns]   traces of each shot;
 *                col[sum ng] model column of each trace, -1 for none (a
 *                         shot's later trace wins a shared column, as in
 *                         swapReord, inversion.cpp:626-643).
 * Device arrays: wav[ntw] the source wavelet; data[sum ng][nt] the gathers,
 *                shot after shot; sou[np][nx][ntw] and rec[np][nx][nt] out:
 *   sou_k[isx_s] = sum over shots s of harmonic(wav, p_k, sx_s),
 *   rec_k[col]   = sum over shots of harmonic(trace, p_k, sx_s),
 * the shots summed in order, bit-identical to the reference's PLANE_S_NEW /
 * PLANE_R_NEW files. */
int seis_plane_prepare(int nx, int nt, int ntw, int ns, int np, float pmin, float dp, const float* sx,
                       const int* isx, const int* ng, const int* col, const float* wav, const float* data,
                       float* sou, float* rec);

/* seis_rtm_batch with option bits; flags = 0 is seis_rtm_batch.
 * SEIS_RTM_ISOLATE_BAND: the reference reconstructs the source wavefield
 * backwards in time with the damped operator (fd.cpp:657-667).  Run that way
 * the absorbing band grows without bound on long records in fast media
 * (e.g. 751 x 2301 cells at 4 m, v up to 5.4 km/s, dt 0.4 ms: image values
 * 1e9 by nt = 2500 and inf/NaN by nt = 5000, in the reference as here).  The
 * 4-layer strips shield the interior wavefield from the band; the band
 * reaches the image only through lap at the strip cells and through the
 * column Hilbert filter of FaQi (fd.cpp:683-684).  With this bit the source
 * field is held at zero outside the model area; the interior wavefields are
 * unchanged and the image differs only within 34 cells of the model edge.
 * No counterpart in the reference. */
#define SEIS_RTM_ISOLATE_BAND 1
int seis_rtm_batch_ex(const seis_grid* grid, int nshots, const seis_shot* shots, const float* vel,
                      const float* wav, const int* gz, const float* rec, float* img, int flags);

/* Least-squares RTM.  The reference declares the LSRTM operator and solver
 * (conope_MPI, cgsolver_MPI at lsrtm/inversion.h:81,87) but defines neither;
 * its intended operator is Born modeling followed by migration of the Born
 * records over all shots (inversion.cpp:420-423).  A plan holds, for a batch
 * of shots sharing one grid, the per-step Born-source snapshots of the
 * background (vt2 * lap * 2/v^3, fd.cpp:134-139) and applies
 *   L   : perturbation [nx][nz] -> packed Born records (as seis_born_batch),
 *   L^T : packed records -> image [nx][nz], the exact transpose of L summed
 *         over shots (not the reference's imaging operator),
 * with FP64 inner products on the GPU.  All arrays are DEVICE pointers.
 * snapshot_kind: 0 = FP32, 1 = BF16, 2 = FP16 (16-bit forms scaled per
 * snapshot by its max |value|).  pml must be >= 4. */
typedef struct seis_lsrtm seis_lsrtm;
int seis_lsrtm_create(seis_lsrtm** plan, const seis_grid* grid, int nshots, const seis_shot* shots,
                      const float* vel, const float* wav, const int* gz, int snapshot_kind);
/* FP64 plan (FP64 propagation and snapshots), for the dot-product test. */
int seis_lsrtm_create_f64(seis_lsrtm** plan, const seis_grid* grid, int nshots, const seis_shot* shots,
                          const double* vel, const double* wav, const int* gz);
int seis_lsrtm_destroy(seis_lsrtm* plan);
/* rec = L pert  (float, or double for an FP64 plan) */
int seis_lsrtm_born(seis_lsrtm* plan, const void* pert, void* rec);
/* grad = L^T rec (or the cross-correlation adjoint, see seis_lsrtm_set_adjoint) */
int seis_lsrtm_adjoint(seis_lsrtm* plan, const void* rec, void* grad);
/* Which adjoint seis_lsrtm_adjoint and seis_lsrtm_cg apply:
 *   SEIS_LSRTM_ADJ_EXACT (default): the exact transpose of the Born operator
 *     (the dot test holds to rounding);
 *   SEIS_LSRTM_ADJ_XCORR: the reference's imaging condition on the stored
 *     snapshots, as its intended normal operator pairs forward_MPI with the
 *     RTM adjoint_MPI (inversion.cpp:420-423). The receiver field is run
 *     backwards with the forward operator and the record injected times vt2,
 *     and grad += u S^n / vt2. It equals L^T in the interior and differs in
 *     the absorbing band; it costs about one Born step per step. */
#define SEIS_LSRTM_ADJ_EXACT 0
#define SEIS_LSRTM_ADJ_XCORR 1
int seis_lsrtm_set_adjoint(seis_lsrtm* plan, int kind);
/* *out = <a, b> over n elements, accumulated in FP64 on the GPU */
int seis_lsrtm_dot(seis_lsrtm* plan, const void* a, const void* b, long long n, double* out);
/* Optional record mask (device, packed like the records; NULL clears): L
 * becomes mask * L and L^T becomes L^T * mask, e.g. receiver columns times
 * the direct-arrival mute of removeDirect (inversion.cpp:699-752).  The
 * caller keeps the mask alive while the plan uses it. */
int seis_lsrtm_set_mask(seis_lsrtm* plan, const void* mask);
/* Optional image-space preconditioner of seis_lsrtm_cg (device [nx][nz] in
 * the plan's precision; NULL clears): CGLS runs on L W with W = diag(weight)
 * and returns the image m = W x.  Weighting by the inverse square root of
 * the source illumination (illum_MPI, inversion.cpp:207-290; fd.cpp:275-339)
 * is the illumination compensation for which the reference computes it.
 * The caller keeps the weight alive while the plan uses it. */
int seis_lsrtm_set_weight(seis_lsrtm* plan, const void* weight);
/* CGLS for min ||L x - dobs||^2 from x = 0: at most niter iterations, stops
 * when ||L^T r|| <= rtol ||L^T dobs|| (rtol <= 0: never); image = x.
 * hist (host, 2*(niter+1) doubles, may be NULL): ||r_i||^2, ||L^T r_i||^2;
 * *iters (may be NULL) = iterations done. */
int seis_lsrtm_cg(seis_lsrtm* plan, const void* dobs, int niter, double rtol, void* image, double* hist,
                  int* iters);
/* seis_lsrtm_cg that also writes the image after every iteration i (1..
 * iterations done) into iterates + (i-1) nx nz (device, niter images in the
 * plan's precision; NULL writes none), e.g. to follow the image error per
 * iteration.  No counterpart in the reference. */
int seis_lsrtm_cg_history(seis_lsrtm* plan, const void* dobs, int niter, double rtol, void* image, double* hist,
                          int* iters, void* iterates);

/* Strip-length autotuning of the interior sweeps (no counterpart in the
 * reference).  On by default: the first run of a configuration (kind of
 * pass, padded grid, shots per launch, device) times a few trial launches of
 * each candidate strip length (8, 16, 32, 64 columns; 1, 2, 4, 8, 16 when the
 * padded grid times the shots is under 4 Mi cells) on its zeroed fields and
 * keeps the fastest for the process.  Results are identical for every strip
 * length.  seis_set_tuning(0) turns it off (strip length 16) and forgets the
 * table; another value than 0 or 1 is SEIS_ERR_ARG.  seis_tuned_strip
 * reports the choice for a configuration (0 when none was made).  kind:
 * "forward", "born", "rtm", "illum", "rtm_plane", "lsrtm_forward",
 * "lsrtm_adjoint". */
int seis_set_tuning(int on);
int seis_tuned_strip(const char* kind, int nx_padded, int nz_padded, int nshots, int* bx);

/* Kernel timing: instrumentation with no counterpart in the reference.
 * Between seis_profile_start and seis_profile_stop every engine launch is
 * bracketed by HIP events on its own stream; stop returns, per kernel, the
 * number of launches, the summed event durations, and the algorithmic HBM
 * bytes and wavefield cell-updates those launches performed. */
typedef struct seis_kernel_stat {
  char name[40];
  int launches;
  double total_ms;
  double alg_bytes;
  double cell_updates;
} seis_kernel_stat;
int seis_profile_start(void);
int seis_profile_stop(seis_kernel_stat* stats, int max_stats, int* n_stats);

#ifdef __cplusplus
}
#endif

#endif /* SEIS_H */
