/*
 * rtm_hip.h -- C ABI of the MI355X (gfx950) RTM library, librtm_hip.so.
 *
 * The reference (caixh90/RTM_GPU) has no library boundary: its 17 __global__
 * kernels are launched from main() in the same translation unit, with raw
 * device pointers and scalars as arguments (kernel.cu:18-517, launched at
 * kernel.cu:808-929).  These entry points replace those launches and the
 * device-memory handling around them (kernel.cu:758-786, 805-806, 822-825,
 * 878-885, 909, 932-933).  The host driver that calls them keeps the
 * reference's shot loop (kernel.cu:791-991).
 *
 * Conventions
 *  - All host arrays are in the reference layout: row-major [z][x] with x
 *    contiguous, NZ x NX padded grid (kernel.cu:611-612).
 *  - Indices are 0-based padded grid indices, i.e. after the 1-based to
 *    padded conversion of kernel.cu:607-610 and kernel.cu:794-795.
 *  - Every entry point returns 0 on success or a negative RTM_E* code; the
 *    message is available from rtm_last_error().  No C++ exception crosses
 *    the ABI.  A context is used by one host thread at a time; distinct
 *    contexts (one per pthread) may run concurrently.
 *  - The library owns all device memory of a context; callers own host
 *    buffers.  Device work is ordered on the context's own HIP stream.
 */
#ifndef RTM_HIP_H
#define RTM_HIP_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  RTM_OK = 0,
  RTM_E_ARG = -1,     /* bad argument or call order */
  RTM_E_HIP = -2,     /* HIP runtime error */
  RTM_E_NOMEM = -3,   /* device allocation failed */
  RTM_E_STATE = -4,   /* entry point called in the wrong state */
  RTM_E_BLOWUP = -5   /* a wavefield holds a NaN or an infinity (see check_every) */
};

enum { RTM_STRIP_F32 = 0, RTM_STRIP_F16 = 1, RTM_STRIP_BF16 = 2 };

enum {
  RTM_FIELD_U = 0,      /* forward / reconstructed source field, newest step */
  RTM_FIELD_U_PREV = 1, /* same field, previous step */
  RTM_FIELD_Q = 2,      /* receiver (back-propagated) field, newest step */
  RTM_FIELD_I1 = 3,     /* cross-correlation image (Drel1) */
  RTM_FIELD_I2 = 4,     /* source illumination (Drel2) */
  RTM_FIELD_SU = 5,     /* running sum of u (DBWADD) */
  RTM_FIELD_SQ = 6      /* running sum of q (DFWADD) */
  /* Born pass: RTM_FIELD_Q holds the scattered field dU; adjoint pass:
   * RTM_FIELD_Q holds mu = P*lambda and RTM_FIELD_I1 the accumulator G. */
};

/* Kernel kinds for rtm_kernel_time().  A frame kind times the frame launches
 * of one step together; on migration and adjoint steps without a checkpoint
 * snapshot (and with separate launches) RTM_K_RCV_FRAME / RTM_K_ADJ_FRAME
 * include the strip restore of the reconstructed field, which
 * RTM_K_STRIP_RESTORE then does not count. */
enum {
  RTM_K_FWD_INTERIOR = 0, RTM_K_FWD_FRAME = 1, RTM_K_SRC_INTERIOR = 2, RTM_K_STRIP_RESTORE = 3,
  RTM_K_RCV_INTERIOR = 4, RTM_K_RCV_FRAME = 5, RTM_K_BORN_INTERIOR = 6, RTM_K_BORN_FRAME = 7,
  RTM_K_ADJ_INTERIOR = 8, RTM_K_ADJ_FRAME = 9,
  /* the migration's backward step outside checkpoint steps: the reconstruction
   * with the source-side image terms, the receiver step with the rest */
  RTM_K_SRCI_INTERIOR = 10, RTM_K_RCVI_INTERIOR = 11, RTM_K_NKINDS = 12
};

typedef struct rtm_params {
  int nz, nx;            /* padded grid NZ, NX (kernel.cu:611-612) */
  int n2;                /* hybrid ABC width N2 */
  int nfdmax;            /* maximum operator half-length; also strip depth */
  int nt, nt2;           /* NT and the wavelet gate NT2 (kernel.cu:613,624) */
  int ilste;             /* 0 adaptive LS coefficients, 1 Taylor (kernel.cu:744-753) */
  int icompen;           /* 1 compensated imaging (Rel_Compen), 0 Rel_NonCompen */
  int s_l, s_r, s_z, ds, n; /* receiver line (kernel.cu:607-610) */
  float tao2, h2, taoh, taoh2, hzx2_1; /* step scalars (kernel.cu:614-617,625-626) */
  int max_shots;         /* shots per batch this context can hold */
  int strip_format;      /* RTM_STRIP_F32 (reference), RTM_STRIP_F16, RTM_STRIP_BF16 */
  int record;            /* 1: allocate receiver records for forward modelling */
  int zseg;              /* rows per wave in the interior kernels (0 = auto) */
  int snap_every;        /* checkpoint snapshots (not in the reference; 0 = none): the
                            forward pass keeps the whole field u(j) and u(j+1) for
                            j = m*snap_every, 2 <= j, j+1 <= nt-3, in strip_format with a
                            per-snapshot power-of-two scale; the backward pass replaces
                            the reconstructed u(k) by its snapshot */
  int stack_nx;          /* columns of the stacked image (0 = nx - 2*n2); a ragged batch
                            (rtm_set_shot_geometry) stacks each shot at its own offset */
  int check_every;       /* not in the reference: every check_every steps (and at the
                            last step) the newest wavefields are checked for NaN and
                            infinity, and the *_advance call (or rtm_migrate) that ran
                            such a step returns RTM_E_BLOWUP; 0 = every 100 steps,
                            < 0 = never.  A time step beyond the stability limit of the
                            model (rtmh_stability) blows up this way. */
  int fuse_frame;        /* not in the reference: 1 runs a forward step's and an
                            ordinary receiver step's interior and frame work in one
                            launch with their blocks interleaved (the two write disjoint
                            cells); 0 (default) keeps two launches per kind.  Results
                            are the same bit for bit; on the reverse-VSP batch the
                            fused receiver step is 2% faster and the fused forward
                            step 11% slower, the whole job within 2% either way. */
  int frame_split;       /* not in the reference: 1 runs the edge cells of the top and
                            bottom bands of the forward and receiver frames in a
                            four-column-per-lane kernel beside the one-column kernel
                            (n2 >= 12, stencil half-lengths up to 6); 0 (default) runs
                            the whole frame in the one-column kernel.  Results are the
                            same bit for bit; the split pays only on large batches (its
                            launch takes about as long for one shot as for 128), so
                            rtm_autotune times both and keeps the faster. */
} rtm_params;

/* One shot of a ragged batch (uneven apertures; not in the reference, whose
 * runs share one grid): the shot's grid is nx columns wide (nx <= the
 * context's nx, same nz and n2), its receivers are s_l, s_l+ds, .., s_r on
 * row s_z of the context (n <= the context's n), and its interior column 0
 * lands in column x_img of the stacked image. */
typedef struct rtm_shot_geom {
  int nx;
  int r_u, r_x;          /* source, 0-based padded, inside the shot's grid */
  int s_l, s_r, n;       /* receivers, 0-based padded */
  int x_img;
} rtm_shot_geom;

typedef struct rtm_ctx rtm_ctx;

/* Allocate a context on the current HIP device (replaces kernel.cu:758-779). */
int rtm_create(const rtm_params* p, rtm_ctx** out);
int rtm_destroy(rtm_ctx* ctx);
const char* rtm_last_error(const rtm_ctx* ctx);
/* Device bytes held by the context. */
size_t rtm_device_bytes(const rtm_ctx* ctx);

/* Upload the model (replaces kernel.cu:781-786).  v, r1: [nz*nx] as built by
 * velocity() (GPU_velocity_real.cpp:6-118); w: [n2+1] (kernel.cu:687-691);
 * c[nc], index[nvel+1]: funMandC's table (LSMOrCon_rec_2D.cpp:22-73), or
 * c[nfdmax+1] from order() with index = NULL when ilste = 1; vmin, dv as
 * snapped at kernel.cu:715-720. */
int rtm_set_model(rtm_ctx* ctx, const float* v, const float* r1, const float* w, const float* c,
                  int nc, const int* index, int nvel, float vmin, float dv);
/* Wavelet table wav[j] = f(j*tao), j = 0..nt (kernel.cu:812, 889, 1261). */
int rtm_set_wavelet(rtm_ctx* ctx, const float* wav);
/* Sources of the batch: shot i injects at (r_u[i], r_x[i]) (kernel.cu:794-795). */
int rtm_set_shots(rtm_ctx* ctx, int nshot, const int* r_u, const int* r_x);
/* A ragged batch: shot i has its own grid width, receivers, source and image
 * offset.  No counterpart in the reference (a drop-in caller keeps
 * rtm_set_shots).  All shots run in the same launches.  Model windows set by
 * rtm_set_shot_model for an earlier batch are dropped: each shot of the new
 * batch needs its window again before a pass begins (RTM_E_STATE otherwise). */
int rtm_set_shot_geometry(rtm_ctx* ctx, int nshot, const rtm_shot_geom* geom);
/* Model window of one shot of a ragged batch: v, r1 [nz*geom.nx] as velocity()
 * builds them for that window.  Call after rtm_set_model (which fixes the
 * coefficient table, w, vmin and dv) and rtm_set_shot_geometry; once any shot
 * has a window, every shot of the batch must have one.  rtm_set_model returns
 * the context to one model for all shots.  Shots whose windows are equal
 * (same width and bytes, compared through two 64-bit fingerprints) share one
 * stored window, so a batch of many shots over a few models reads each
 * model's arrays once per step from cache rather than once per shot. */
int rtm_set_shot_model(rtm_ctx* ctx, int shot, const float* v, const float* r1);
/* Observed gathers of the batch, seis[shot][n][nt] as read at kernel.cu:827-856.
 * Uploaded once per batch (the reference copies one time slice per step,
 * kernel.cu:905-909). */
int rtm_set_gathers(rtm_ctx* ctx, const float* seis);

/* Use the receiver records of the last forward pass (rtm_forward_begin with a
 * context created with record = 1) as the observed gathers of the batch,
 * without a host round trip.  The reference reads gathers from disk instead
 * (kernel.cu:827-838); this serves forward modelling followed by migration. */
int rtm_gathers_from_records(rtm_ctx* ctx);

/* Forward pass (kernel.cu:798-821): begin sets u(0)=0, u(1)=f(0)/2 at the
 * source and strip slots 0,1; advance runs up to nsteps of k = 2..nt-1.
 * save_strips = 0 skips the boundary strips (pure modelling). */
int rtm_forward_begin(rtm_ctx* ctx, int save_strips);
int rtm_forward_advance(rtm_ctx* ctx, int nsteps);
/* Backward pass with imaging (kernel.cu:822-931). */
int rtm_backward_begin(rtm_ctx* ctx);
int rtm_backward_advance(rtm_ctx* ctx, int nsteps);
/* Forward then backward for the whole batch. */
int rtm_migrate(rtm_ctx* ctx);
/* on = 1: rtm_migrate captures its launch sequence into a hipGraph on first
 * use and replays it afterwards (until the model, wavelet or shots change).
 * For small grids, where launch latency dominates. */
int rtm_set_graph(rtm_ctx* ctx, int on);
/* Block until the context's stream is idle. */
int rtm_sync(rtm_ctx* ctx);

/* Read back (debug / parity): a field of one shot into out[nz*nx]. */
int rtm_read_field(rtm_ctx* ctx, int shot, int which, float* out);
/* Receiver records of the batch, out[shot][n][nt] (forward modelling). */
int rtm_read_records(rtm_ctx* ctx, float* out);
/* Strip slot k of one shot, decoded to float, layout lf[mnz][nfd],
 * rt[mnz][nfd], up[nfd][mnx], dw[nfd][mnx] (Dlf/Drt/Dup/Ddw, kernel.cu:199-205). */
int rtm_read_strip(rtm_ctx* ctx, int shot, int k, float* out);

/* Per-shot image post (kernel.cu:935-990): up = -lap(I1)*v^2/vmax^2,
 * down = I2 + whitecoe*max|I2|, both interior [mnz][mnx] row-major (mnx of
 * the context's grid; a narrower shot of a ragged batch fills its first
 * columns), for every shot of the batch, written to up[shot][..],
 * down[shot][..] when non-NULL, and added in shot order to the context's
 * stack (kernel.cu:1000-1040) at the shot's image column. */
int rtm_image_post(rtm_ctx* ctx, float vmax, float whitecoe, float* up, float* down);
/* Stack result (kernel.cu:1042-1059): img = sum(up)/nrec, illum = sum(down)/nrec,
 * img /= illum when inorm = 1.  [mnz][stack_nx] (stack_nx = mnx unless set).
 * Resets the stack. */
int rtm_stack_finish(rtm_ctx* ctx, int nrec, int inorm, float* img, float* illum);

/* Least-squares migration operators (not in the reference, which migrates
 * only).  L is Born (linearised) modelling: the scattered field dU follows
 * the background forward step with the same update and hybrid ABC, driven on
 * the interior by the reflectivity m:
 *   dU(k) = update(dU(k-1), dU(k-2)) + m*((u(k) - 2u(k-1)) + u(k-2)),
 * recorded at the receivers.  The adjoint L^T back-propagates residuals r
 * with the background reconstructed from the strips (and snapshots) as in
 * the migration; the adjoint field is carried as mu = P*lambda with
 * P = v^2*tao2*h2 per cell, so it takes the plain update and ABC with
 * P*r(k) added at the receivers, and the gradient is
 * (sum_k ((u(k+2) - 2u(k+1)) + u(k)) * mu(k+2)) / P.  With per-cell LS rows
 * (ilste = 0) the adjoint field takes the transposed stencil on the interior
 * (each tap with the coefficient of the neighbour's own row); one Taylor row is
 * symmetric.  The pair passes the dot-product test to the accuracy of the
 * hybrid ABC, which is not self-adjoint. */

/* Reflectivity m on the stacked image grid [mnz][stack_nx]; each shot reads
 * its window at x_img.  Needs the shots to be set. */
int rtm_set_reflectivity(rtm_ctx* ctx, const float* m);
/* Born pass of the batch: begin starts the background (with strips, so that
 * rtm_adjoint_begin may follow) and zeroes dU and the records; advance runs
 * up to nsteps of k = 2..nt-1.  Needs a context with record = 1; the Born
 * data are read with rtm_read_records. */
int rtm_born_begin(rtm_ctx* ctx);
int rtm_born_advance(rtm_ctx* ctx, int nsteps);
/* Adjoint pass after a complete forward or Born pass with strips, with the
 * residuals r[shot][n][nt] given by rtm_set_gathers: begin zeroes mu and the
 * gradient accumulator and takes mu(nt-1), mu(nt-2) (two updates); advance
 * runs up to nsteps of k = nt-3..0. */
int rtm_adjoint_begin(rtm_ctx* ctx);
int rtm_adjoint_advance(rtm_ctx* ctx, int nsteps);
/* Gradient of the batch after a complete adjoint pass: grad[mnz][stack_nx] =
 * sum over shots, in shot order, of each shot's G/P at its image window. */
int rtm_adjoint_gradient(rtm_ctx* ctx, float* grad);

/* Data-space vectors of conjugate-gradient iterations stay on the device:
 * the gathers (residuals read by the adjoint pass) and the records (Born
 * data of the last Born pass), both [shot][nt][n] over the batch. */
enum { RTM_DATA_GATHERS = 0, RTM_DATA_RECORDS = 1 };
/* Over shots [shot0, shot0 + nshot) of the batch (nshot = 0: through the
 * last shot), so that independent problems batched together keep their own
 * step lengths: *out = sum of x*y in FP64, with a fixed reduction order
 * (repeatable bit for bit). */
int rtm_data_dot(rtm_ctx* ctx, int x, int y, int shot0, int nshot, double* out);
/* gathers = gathers + a*records over shots [shot0, shot0 + nshot) (e.g. the
 * residual update r -= alpha*q). */
int rtm_data_axpy(rtm_ctx* ctx, float a, int shot0, int nshot);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops): the FP64 dot
 * products of rtm_data_dot on caller-owned device vectors, with no allocation,
 * no copy to the host and no synchronisation.  out[v] (device memory) receives
 * sum_i x[v*stride + i] * y[v*stride + i], i < n, for v < nvec, in the same
 * fixed reduction order as rtm_data_dot, on stream (a hipStream_t; NULL = the
 * null stream).  The partial sums live in a static device array of
 * RTM_DOT_NOALLOC_MAX_VEC * 1024 doubles (8 KiB per vector) per GPU, held until
 * the process ends; nvec beyond RTM_DOT_NOALLOC_MAX_VEC is rejected.  Calls
 * that overlap on one GPU race on that array.  Errors are reported by
 * rtm_last_error(NULL) on the calling thread. */
/* No counterpart in the reference; a measurement aid.  dst[i] = src[i] for
 * i < n (device memory, 16-byte aligned, n a multiple of 4) with a plain
 * float4 streaming kernel on stream (NULL = the null stream): the bandwidth a
 * pure streaming kernel reaches on this device. */
int rtm_stream_copy(const float* src, float* dst, long long n, void* stream);

#define RTM_DOT_NOALLOC_MAX_VEC 64
int rtm_data_dot_noalloc(const float* x, const float* y, long long n, long long stride, int nvec,
                         double* out, void* stream);

/* Autotuner (not in the reference): for each interior kernel of the
 * migration (forward, source reconstruction, receiver step + imaging) times
 * that kernel inside forward + backward step pairs of the current batch (the
 * other kernels interleaved, the buffers rotating as in a migration) for row
 * segments of 1, 2, 3, 4, 6 and 9 times its register-queue length, in three
 * interleaved rounds of max(2, nsteps/2) timed pairs per candidate, one
 * kernel at a time, and keeps the fastest per kernel for this context (Born
 * modelling and the adjoint take the receiver step's); zseg[3], when not
 * NULL, receives the three choices in rows.  Needs
 * model, wavelet and shots; leaves the passes to be begun anew and resets the
 * rtm_kernel_time counters.  Rows per wave trade the vertical halo re-read
 * (2*NFD rows per segment) against the number of waves in flight.  Then the
 * frame kernels' row runs (rtm_set_frame_runs) are timed the same way, top and
 * bottom bands first, on the sum of both frame kernels' times, and last the
 * frame split (rtm_params.frame_split) on and off. */
int rtm_autotune(rtm_ctx* ctx, int nsteps, int* zseg);
/* The autotuner's three choices set by hand: rows per wave of the forward,
 * the source-reconstruction and the receiver-step kernels (Born and the
 * adjoint take the receiver step's). */
int rtm_set_zseg(rtm_ctx* ctx, const int* zseg);
/* Rows per frame-kernel wave: the top and bottom bands run in pieces of
 * rows_tb rows (ceil(n2/3)..n2; where the four-column kernel takes the middle
 * of those bands, rows_tb applies to their ends), the side bands in pieces of
 * rows_side rows (from a quarter of the default up to nz).  rtm_autotune picks
 * both; rtm_frame_runs reads the current pair. */
int rtm_set_frame_runs(rtm_ctx* ctx, int rows_tb, int rows_side);
int rtm_frame_runs(rtm_ctx* ctx, int* rows_tb, int* rows_side);

/* Kernel timing with HIP events on the context stream: mask is a bit set of
 * kernel kinds (1 << RTM_K_*), 0 = off, -1 = all.  Resets the counters. */
int rtm_set_timing(rtm_ctx* ctx, int mask);
/* Accumulated device time (ms) and launch count of one kernel kind. */
int rtm_kernel_time(rtm_ctx* ctx, int kind, double* total_ms, long long* launches);

#ifdef __cplusplus
}
#endif
#endif
