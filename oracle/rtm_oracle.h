/*
 * rtm_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference RTM path (caixh90/RTM_GPU) used as the
 * checker for the HIP library and as the pthread CPU baseline timed beside it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library.  The product (librtm_hip.so, librtm_host.so) never links it.
 *
 * Every function cites the reference file:line it restates.  Two arithmetic
 * variants of the wave engine exist:
 *   prec = 32  "FP32 replica": float storage and arithmetic with the
 *              reference's double promotions (kernel.cu:58,106,275,312) and
 *              FMA contraction at exactly three places: the per-tap stencil
 *              sum, the tap accumulation and the FP32 final combine.
 *   prec = 64  "FP64 oracle": same float inputs, all arithmetic in double.
 */
#ifndef RTM_ORACLE_H
#define RTM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* kernel.cu:602-628 -- derived grid sizes, indices and scalars.
 * ints_out  = {NZ, NX, NT, NT2, s_l, s_r, s_z, r_x}  (0-based padded indices)
 * float_out = {taoh, tao2, h2, taoh2, hzx, hzx2_1} */
void orc_derive(int mod_nz, int mod_nx, int n2, int nt1, float tao1, float tao,
                float h, float hz, float f0, int s_l1, int s_z1, int n, int ds,
                int r_x1, int* ints_out, float* floats_out);
/* kernel.cu:794-795 -- source row from a depth in metres. */
int orc_source_row(float depth_m, float hz, int n2);
/* kernel.cu:1261-1266 -- Ricker wavelet. */
float orc_ricker(float t, float f0);
/* GPU_velocity_real.cpp:6-118 -- pad a column-major raw model, optional flip,
 * r_1 = sqrt((v*tao/h)^2/2). */
void orc_pad_model(const float* raw, int mod_nz, int mod_nx, int n2, int ifv,
                   float tao, float h, float* v, float* r1);
/* kernel.cu:704-738 -- velocity range snapped to dv and per-bin need flags. */
void orc_vrange(const float* v, int nxz, float dv, float* vmin, float* vmax, int* nvel);
void orc_bins_needed(const float* v, int nxz, float vmin, float dv, int nvel, int* need);
/* kernel.cu:687-691 -- blend weights w[i] = i/N2, i = 0..N2. */
void orc_blend_weights(int n2, float* w);

/* LSMOrCon_rec_2D.cpp:22-73 (funMandC) and callees.  Returns NC or -1 if
 * c_cap is too small.  M[nvel], Index[nvel+1], c[NC]. */
int orc_design_ls(int nthita, int nfdmax, int nfdmin, int nvel, double tao, double h,
                  double df, double eps, double fmax, double vmin, double dv,
                  const int* need, double hzx, int* M, int* Index, float* c, int c_cap);
/* LSMOrCon_rec_2D.cpp:293-346 -- operator length for one velocity. */
int orc_lenfd(int nfre, int nfdmin, int nfdmax, double vel, double tao, double h,
              double df, int nthita, double eps, double hzx);
/* LSMOrCon_rec_2D.cpp:236-287 -- LS coefficients c[0..M] (double). */
void orc_lscoef(double r, double bb, int M, double hzx, double* c);
/* LSMOrCon_rec_2D.cpp:526-551 -- Taylor coefficients for order 2*nfdmax. */
void orc_taylor(int nfdmax, float* c);
/* Dispersion error max over the (f, theta) grid for a given coefficient set,
 * same formula as LSMOrCon_rec_2D.cpp:321-336 (used by invariant tests). */
double orc_disp_maxerr(const double* c, int M, double vel, double tao, double h,
                       double df, int nfre, int nthita, double hzx);

/* ---- wave engine (one shot) ------------------------------------------- */
typedef struct orc_cfg {
  int nz, nx, n2, nfdmax, nt, nt2;
  int ilste, icompen;
  int s_l, s_r, s_z, ds, n;     /* receivers, 0-based padded */
  int r_u, r_x;                 /* source, 0-based padded */
  float tao2, h2, taoh, taoh2, hzx2_1, vmin, dv;
  int record;                   /* record receivers in the forward pass */
  int nthreads;                 /* row-parallel threads per step */
  int strip_slots;              /* strip history slots kept (0 = nt); slot k%strip_slots */
  int snap_every;               /* checkpoint snapshots of u(j), u(j+1) for j = m*snap_every,
                                   2 <= j, j+1 <= nt-3; backward replaces the reconstructed
                                   u(k) by its snapshot (0 = none; not in the reference) */
} orc_cfg;

typedef struct orc_engine orc_engine;
orc_engine* orc_create(const orc_cfg* cfg, int prec);
void orc_destroy(orc_engine* e);
/* model: v, r1 [nz*nx], w [n2+1], c [nc], index [nvel+1] (NULL for Taylor),
 * wavelet table wav[j] = f(j*tao), j = 0..nt. */
void orc_set_model(orc_engine* e, const float* v, const float* r1, const float* w,
                   const float* c, int nc, const int* index, int nvel, const float* wav);
/* Forward: kernel.cu:798-821.  begin sets u(0), u(1) and strip slots 0,1;
 * advance runs nsteps of the k loop (k = 2..nt-1).  Returns the next k. */
int orc_forward_begin(orc_engine* e);
int orc_forward_advance(orc_engine* e, int nsteps);
/* Backward: kernel.cu:822-931.  seis [n][nt] (trace-major, reference layout). */
int orc_backward_begin(orc_engine* e, const float* seis);
int orc_backward_advance(orc_engine* e, int nsteps);
/* Linearised (Born) modelling and its adjoint (not in the reference).
 * m: reflectivity [nz*nx] (used on the interior).  born: the background
 * forward step, then the scattered field dU(k) = update(dU(k-1), dU(k-2)) +
 * m*((u(k)-2u(k-1))+u(k-2)) with the same ABC; with record = 1 the record
 * holds dU at the receivers.  adjoint (after a plain forward pass with
 * strips): the background is reconstructed backward from the strips (and
 * snapshots) as in the migration; the adjoint field is carried as
 * mu = P*lambda, P = v^2*tao2*h2 per cell, so it takes the plain update
 * (FP32 combine) and ABC with P*resid[s][k] added at the receivers after the
 * update; mu(nt-1), mu(nt-2) come from two updates in adjoint_begin, and
 * G += ((u(k+2)-2u(k+1))+u(k)) * mu(k+2) on the interior.  The gradient is
 * G / P (field 9). */
void orc_set_reflectivity(orc_engine* e, const float* m);
int orc_born_begin(orc_engine* e);
int orc_born_advance(orc_engine* e, int nsteps);
int orc_adjoint_begin(orc_engine* e, const float* resid);
int orc_adjoint_advance(orc_engine* e, int nsteps);
/* which: 0 = current forward/source field u (u(k) most recent),
 *        1 = previous forward/source field,
 *        2 = receiver field q (most recent), 3 = I1, 4 = I2, 5 = sum U, 6 = sum Q,
 *        7 = scattered field dU (most recent), 8 = adjoint accumulator G,
 *        9 = gradient G / P (interior; 0 on the frame).
 * Output is float [nz*nx] (images: interior valid, frame 0). */
void orc_read_field(const orc_engine* e, int which, float* out);
void orc_read_field64(const orc_engine* e, int which, double* out);
/* Receiver record rec[s][k], s < n, k < nt (forward pass with record=1). */
void orc_read_record(const orc_engine* e, float* out);
/* Strip slot k in the engine's strip layout (lf, rt, up, dw). */
void orc_read_strip(const orc_engine* e, int k, float* out);

/* kernel.cu:935-990 -- per-shot image post.  I1, I2 full-grid [nz*nx];
 * up, down interior [mod_nz*mod_nx] row-major [z][x]. */
void orc_image_post(const float* I1, const float* I2, const float* v, int nz, int nx,
                    int n2, float vmax, float whitecoe, float* up, float* down);
/* kernel.cu:1000-1059 -- stack of per-shot up/down images and normalisation. */
void orc_stack(const float* ups, const float* downs, int nshot, int nI, int inorm,
               float* img, float* illum);

/* CPU baseline: run nshot independent shots on nthreads pthreads (one shot
 * per thread at a time), forward nsteps_f steps then backward nsteps_b steps.
 * Returns wall seconds (steady clock). */
double orc_time_shots(const orc_cfg* cfg, const float* v, const float* r1, const float* w,
                      const float* c, int nc, const int* index, int nvel, const float* wav,
                      const float* seis, int nshot, int nthreads, int nsteps_f, int nsteps_b);
/* CPU baseline, latency form: one shot whose every step is split by rows over
 * nthreads pthreads (a barrier per step).  Returns wall seconds. */
double orc_time_rows(const orc_cfg* cfg, const float* v, const float* r1, const float* w,
                     const float* c, int nc, const int* index, int nvel, const float* wav,
                     const float* seis, int nthreads, int nsteps_f, int nsteps_b);

#ifdef __cplusplus
}
#endif
#endif
