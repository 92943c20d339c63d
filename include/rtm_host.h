/*
 * rtm_host.h -- host-side C API of librtm_host.so: the driver surface of the
 * reference (caixh90/RTM_GPU) above the device library librtm_hip.so.
 *
 * It keeps the reference's parameter layout, setup arithmetic, model and
 * gather I/O, coefficient design, shot loop, per-shot image files, stacking
 * and SEG-Y output (kernel.cu:525-1213), with a pthread shot dispatcher in
 * place of the reference's serial loop (kernel.cu:791).  All functions return
 * 0 on success or a negative code; rtmh_last_error() gives the message of the
 * calling thread's last failure.
 */
#ifndef RTM_HOST_H
#define RTM_HOST_H
#ifdef __cplusplus
extern "C" {
#endif

/* The 27 label/value pairs of 2D_Real_RVSP_RTM.txt (kernel.cu:542-600), in
 * file order, followed by the 12 values of the OutPara file (kernel.cu:602-604). */
typedef struct rtmh_config {
  int nfdmax, nfdmin, n2;
  float f0, fmax, df;
  int nthita;
  float eps, dv;
  int ilste, ifv;
  float whitecoe, hz, tao;
  int inorm, icompen, nsmooth;
  float wthite_phase, angle;
  int nx_bg, nx_ed, nz_bg, nz_ed;
  char out_seis[512], out_vp[512], out_dpr[512], out_para[512], result[512];
  /* OutPara: h tao1 mod_NZ mod_NX NT1 s_l s_z n ds r_x nrec dr (1-based geometry) */
  float h, tao1;
  int mod_nz, mod_nx, nt1, s_l1, s_z1, n, ds, r_x1, nrec, dr;
} rtmh_config;

/* Quantities derived at kernel.cu:607-628. */
typedef struct rtmh_derived {
  int nz, nx, nt, nt2;
  int s_l, s_r, s_z, r_x;       /* 0-based padded */
  float taoh, tao2, h2, taoh2, hzx, hzx2_1;
} rtmh_derived;

const char* rtmh_last_error(void);

/* Parameter files. */
int rtmh_read_config(const char* rtm_txt, rtmh_config* cfg);   /* both files */
int rtmh_read_depths(const char* path, int nrec, float* depths); /* kernel.cu:695-700 */

/* Setup arithmetic (kernel.cu:607-628, 794-795, 1261-1266, 687-691). */
void rtmh_derive(const rtmh_config* cfg, rtmh_derived* d);
int rtmh_source_row(float depth_m, float hz, int n2);
float rtmh_ricker(float t, float f0);
void rtmh_wavelet_table(int nt, float tao, float f0, float* wav /* nt+1 */);
void rtmh_blend_weights(int n2, float* w /* n2+1 */);

/* Model (GPU_velocity_real.cpp:6-118; kernel.cu:704-738). raw is the
 * column-major interior model (x outer) as stored on disk. */
int rtmh_read_model(const char* path, int mod_nz, int mod_nx, float* raw);
void rtmh_pad_model(const float* raw, int mod_nz, int mod_nx, int n2, int ifv, float tao,
                    float h, float* v, float* r1);
void rtmh_vrange(const float* v, long long n, float dv, float* vmin, float* vmax, int* nvel);
void rtmh_bins_needed(const float* v, long long n, float vmin, float dv, int nvel, int* need);

/* Coefficient design (LSMOrCon_rec_2D.cpp:22-551).  rtmh_design_ls returns NC
 * (size of c) or a negative code; c_cap bounds c.  nthreads parallelises the
 * per-bin coefficient solves (results do not depend on it). */
int rtmh_design_ls(int nthita, int nfdmax, int nfdmin, int nvel, double tao, double h, double df,
                   double eps, double fmax, double vmin, double dv, const int* need, double hzx,
                   int* M, int* index, float* c, int c_cap, int nthreads);
void rtmh_taylor(int nfdmax, float* c /* nfdmax+1 */);
/* Stability of the time stepping (not in the reference, which runs any time
 * step): the largest value over the cells v[nxz] of
 * v^2*tao2*h2*(1 + hzx2_1)*max_k S(k)/4, S(k) = -(c0 + 2 sum_l c_l cos(l k)) the
 * symbol of the cell's coefficient row (c, index[nvel+1] as for rtm_set_model;
 * index = NULL: the Taylor row c[nfdmax+1]).  The interior update is stable
 * when it is <= 1; rtmh_run refuses a model whose number exceeds 1. */
double rtmh_stability(const float* v, long long nxz, const float* c, const int* index, int nvel,
                      float vmin, float dv, int nfdmax, float tao2, float h2, float hzx2_1);

/* SEG-Y (SGYWrite.cpp:3-55 with the template of SGY_Model.sgy; segy.cpp codecs). */
void rtmh_float_to_ibm(const float* in, unsigned char* out_be, long long n);
void rtmh_ibm_to_float(const unsigned char* in_be, float* out, long long n);
/* Writes ntr traces of ns samples (data[trace][sample]) copying the textual,
 * binary and first trace header of template_path; ns, dt and five trace keys
 * are patched exactly as WriteSGY does. */
int rtmh_write_segy(const char* template_path, const char* out_path, const float* data, int ntr,
                    int ns, int dt_units, const float* sx, const float* sy, float rx, float ry,
                    const float* dsr);
/* Reads a SEG-Y file: *ntr, *ns, *format from the headers; data may be NULL to
 * query sizes only, else data[trace][sample] with capacity cap floats. */
int rtmh_read_segy(const char* path, int* ntr, int* ns, int* format, float* data, long long cap);

/* Per-shot post-processing helpers on the host (kernel.cu:1000-1059). */
void rtmh_stack(const float* ups, const float* downs, int nshot, long long nI, int inorm,
                float* img, float* illum);

/* Post-stack output chain (kernel.cu:1110-1179).  Arrays are trace-major
 * ([column][sample]); columns x_st..x_end of v are converted, d is indexed
 * from its first trace.  Each returns the output length per trace (> 0) or a
 * negative code; out (ntr*len floats, ntr = x_end-x_st+1) may be NULL to query
 * the length, cap is its capacity in floats.
 *   rtmh_d2t: depth -> two-way time, time step dt (D2T,
 *             DisToTimeAndTimeToDis1D.cpp:115-173; called with dz = hz, dt = tao).
 *   rtmh_t2d: time -> depth, time step dtt and depth step dzz over a velocity of
 *             nzv samples per column (T2D, DisToTimeAndTimeToDis1D.cpp:31-112;
 *             called with dtt = tao, dzz = hz). */
int rtmh_d2t(const float* v, const float* d, int nx, int nz, int x_st, int x_end, float dz,
             float dt, float* out, long long cap);
int rtmh_t2d(const float* v, const float* d, int nx, int nt, int nzv, int x_st, int x_end,
             float dtt, float dzz, float* out, long long cap);
/* Constant-phase rotation of ntr traces of nt samples by angle degrees
 * (phase_correction, phase_correction_ricker_decon.cpp:153-218). */
void rtmh_phase_rotate(const float* din, float* dout, int ntr, int nt, float angle);

/* The driver: runs the reference's main() flow from a 2D_Real_RVSP_RTM.txt
 * file.  workers = pthread workers, each with its own librtm_hip context and
 * stream; batch = shots per device batch.  Shots whose per-shot image files
 * already exist are skipped (resume).  Returns 0 or a negative code. */
typedef struct rtmh_run_opts {
  int workers;         /* >= 1 */
  int batch;           /* shots per batch, >= 1 */
  int retries;         /* re-queue a failed batch this many times */
  int strip_format;    /* RTM_STRIP_F32 / F16 / BF16 */
  int resume;          /* 1: skip shots with existing per-shot images */
  int verbose;
  const char* sgy_template; /* path of SGY_Model.sgy (NULL: ./SGY_Model.sgy) */
  /* Called by a worker before each attempt of a batch, with the batch's
   * 0-based source indices and the attempt number (0 first); a non-zero
   * return fails that attempt, which is then re-queued like a failed device
   * batch (up to `retries` times).  NULL: none.  For progress reports,
   * cancellation and tests of the re-queue path; it may be called from
   * several worker threads at once. */
  int (*before_batch)(const int* shots, int nshots, int attempt, void* user);
  void* user;
} rtmh_run_opts;
int rtmh_run(const char* rtm_txt, const rtmh_run_opts* opts);

#ifdef __cplusplus
}
#endif
#endif
