// rtm_kernels.h -- device-side data structures shared by the gfx950 kernels
// (rtm_kernels.hip) and the C-ABI glue (rtm_capi.hip).  Internal header.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtm {

// Padded-grid geometry in device memory.  Cell (z, x) of shot s in a field
// array lives at s*fstride + z*pitch + x0 + x.  x0 is chosen so that the
// first interior column n2 starts a 128-byte cache line (so does every
// 256-column tile of the interior kernels), and the row has >= 4 guard
// columns on each side so every float4 row access stays inside its row.
struct Grid {
  int nz, nx, pitch, x0, n2, mnz, mnx;  // nx, mnx: the widest shot of the batch
  int nfd;          // nfdmax of the model (strip depth)
  long long fstride;
};

// Model operands.  ctab rows are [c0' as double | c_1 .. c_NFD | zero pad],
// c0' = (1.0 + hzx2_1) * c_0 in double (the promotion of kernel.cu:58).
struct Model {
  long long mstride;   // 0: one model for all shots; fstride: windows ShotGeo::win apart
  const float* v;
  const float* r1;
  const float* pm;     // v*v*tao2*h2 per cell (FP32, left to right), padded layout
  const unsigned* ko;  // byte offset of the cell's coefficient row in ctab, padded layout
  const float* ctab;
  const float* w;
  const float* refl;   // reflectivity (Born modelling), padded layout, per shot at rstride
  long long rstride;   // 0: one reflectivity field for all shots
  // Neighbour-tap planes of the transposed stencil (per-cell rows): plane
  // 4*(l-1)+d holds, per cell, the tap-l coefficient of the neighbour's own row,
  // d = 0 above (z-l), 1 below (z+l), 2 left (x-l), 3 right (x+l); one set of
  // 4*NFD planes (fstride each) per model window at tpstride (0: shared).
  const float* tp;
  long long tpstride;
  int cs, nrow;
  float hzx2_1, tao2, h2, taoh, taoh2;
};

struct Rcv {
  int s_l, s_r, s_z, ds, n;  // n: receivers per shot of the data/record layout (largest shot)
};

// Geometry of one shot of a batch.  A uniform batch repeats the context's
// grid and receiver line; a ragged batch gives each shot its own aperture
// window: a narrower grid (nx <= Grid::nx, same nz) with its own receivers,
// and the column of the global image its interior starts at.
struct ShotGeo {
  int nx, mnx;         // padded / interior width
  int s_l, s_r, n;     // receivers on row s_z, spacing ds (common)
  int r_u, r_x;        // source
  int x_img;           // image column of interior column 0
  int win;             // model window of the shot (Model::mstride apart; 0 with one model)
};

// Strip storage: fp32, or fp16/bf16 with a per-(shot, slot) scale.
struct Strips {
  void* base;          // [shot][slot][ss] elements
  float* scale;        // [shot][slot] (fp16/bf16)
  unsigned* amax;      // [shot][slot] max |value| bits (fp16/bf16)
  float* stage;        // [shot][2][ss] fp32 staging (fp16/bf16)
  int fmt, nslot;
  long long ss;
};

// Checkpoint snapshots of whole fields, compact [nz][nx] per (shot, slot).
struct Snaps {
  void* base;          // [shot][slot][nz*nx] elements (fp32 / fp16 / bf16)
  float* scale;        // [shot][slot] (fp16/bf16)
  unsigned* amax;      // [shot][slot] max |value| bits (fp16/bf16)
  int fmt, nsnap;
};

// Interior epilogues: forward (FP32 / FP64 combine), source reconstruction,
// receiver step + imaging, Born scattered field (FP32 / FP64 combine), the
// adjoint of Born modelling, and the split imaging of the migration: the
// reconstruction with the source-side image terms (sum u, I2 += u*u) and the
// receiver step with the rest (sum q, I1; compensated imaging only).
enum { EPI_FWD = 0, EPI_SRC = 1, EPI_RCV = 2, EPI_FWD64 = 3, EPI_BORN = 4, EPI_BORN64 = 5,
       EPI_ADJ = 6, EPI_SRCI = 7, EPI_RCVI = 8, EPI_ADJT = 9 };
enum { FR_FWD = 0, FR_RCV = 1, FR_BORN = 2, FR_ADJ = 3 };

struct IntArgs {
  Grid g;
  Model m;
  Rcv r;
  const float* u1;
  const float* u0;
  float* u2;
  const ShotGeo* geo;  // per shot
  float wav;           // source amplitude this step (0 past the gate)
  int k, nt;
  int nxt, zs;         // interior x tiles, rows per wave (requested)
  int nshot;           // shots of the launch (the fastest-varying block index)
  const float* data;   // [shot][nt][n] (receiver mode)
  const float* uk;     // u(k) for imaging (receiver mode)
  float *su, *sq, *i1, *i2;
  int icompen;
  // Born / adjoint: background fields ua, ub, uc with D2u = (ua - 2ub) + uc
  // (Born: u(k), u(k-1), u(k-2); adjoint: u(k+2), u(k+1), u(k)); the adjoint
  // accumulates into i1 when image != 0 and reads data at step k.
  const float *ua, *ub, *uc;
  int image;
};

struct FrameArgs {
  Grid g;
  Model m;
  Rcv r;
  const float* u1;
  const float* u0;
  float* u2;
  // one record per wave: (first output row, column of lane 0, band | rows << 4
  // | ncol << 20, shot); band 0 top, 1 bottom, 2 left, 3 right, 4 both sides
  // in one wave; ncol (top/bottom) limits the written lanes to 1..ncol (0: 62).
  // k_frame4 records: (first output row, first column of the 256, band, shot).
  // (frame_records)
  const int4* recs;
  const ShotGeo* geo;
  float wav;
  int k, nt;
  int f64comb;
  int nshot;           // shots of the launch
  int nrec;            // records of the launch
  const float* data;
  Strips st;
  int save_strips;
  const float *ua, *ub, *uc;  // Born background fields (as in IntArgs)
  int adjt;            // adjoint with per-cell rows: interior cells take the transposed stencil
  // receiver and adjoint frames only: rbps > 0 appends rbps blocks per shot
  // that restore the strips of slot k into the rings of field rf
  // (k_strip_restore's work)
  float* rf;
  int rbps;
};

// Launchers (rtm_kernels.hip).  nfd_inst is the compile-time stencil
// half-length the call is dispatched to (>= model nfdmax, <= n2).
// Rows of the interior kernel's register queue; a wave's segment is a
// multiple of it.
int interior_queue_rows(int epi, int nfd);
hipError_t launch_interior(int epi, int nfd_inst, IntArgs a, int nshot, hipStream_t s);
// One launch of a time step's interior kernel (epi) and frame kernel (mode),
// their blocks interleaved (k_step); only (EPI_FWD, FR_FWD) and
// (EPI_RCVI, FR_RCV) are fused (step_fusable).
bool step_fusable(int epi, int mode);
hipError_t launch_step(int epi, int mode, int nfd, IntArgs ia, FrameArgs fa, int nshot, int nrec,
                       hipStream_t s);
hipError_t launch_frame(int mode, int nfd_inst, FrameArgs a, int nshot, int nrecords,
                        hipStream_t s);
// Frame records of one shot of grid width nx: top and bottom bands in runs of
// seg_tb rows, the side bands in runs of seg rows.  FRAME_ALL: every frame
// cell for k_frame (62 written columns per record); FRAME_SCALAR: the k_frame
// part of a split frame (the n2+1 columns at each end of the top and bottom
// bands, and the side bands); FRAME_VEC: the k_frame4 part (the edge cells of
// the top and bottom bands, 252 columns per record).  A split frame is every
// frame cell once; when the shot's bands hold no float4 record, FRAME_SCALAR
// equals FRAME_ALL and FRAME_VEC is empty.
enum { FRAME_ALL = 0, FRAME_SCALAR = 1, FRAME_VEC = 2 };
void frame_records(const Grid& g, int nx, int shot, int seg_tb, int seg, int kind, int4* out,
                   int* count);
// The float4 top/bottom frame kernel (forward and receiver fields) exists for
// these compile-time stencil half-lengths.
bool frame4_supported(int nfd_inst);
hipError_t launch_frame4(int mode, int nfd_inst, FrameArgs a, int nshot, int nrecords,
                         hipStream_t s);
int frame_segment_rows(int nfd);
hipError_t launch_fill(float* p, long long n, float val, hipStream_t s);
// Neighbour-tap planes (Model::tp) of one model window from its coefficient
// offsets ko (padded layout) and the coefficient table: 4*nfd planes at tp.
hipError_t launch_tap_planes(const Grid& g, const unsigned* ko, const float* ctab, int nfd,
                             float* tp, hipStream_t s);
// Finite-value watch: reset *flag to 0xffffffff; store k in *flag (if still
// 0xffffffff) when any of the n floats at f (n a multiple of 4, 16-byte
// aligned) is a NaN or an infinity.
hipError_t launch_flag_reset(unsigned* flag, hipStream_t s);
hipError_t launch_finite_check(const float* f, long long n, int k, unsigned* flag, hipStream_t s);

// Per-shot kernels take the batch's geometry array; launches cover shots
// 0..nshot-1 (blockIdx.y).
hipError_t launch_fwd_begin(const Grid& g, float* f0, float* f1, const ShotGeo* geo, float spike,
                            int nshot, hipStream_t s);
hipError_t launch_strip_save(const Grid& g, const float* f, Strips st, const ShotGeo* geo, int k,
                             int nshot, hipStream_t s);
hipError_t launch_strip_restore(const Grid& g, float* f, Strips st, const ShotGeo* geo, int k,
                                int nshot, hipStream_t s);
// Restore blocks per shot when the receiver frame launch carries the restore.
int strip_restore_blocks(const Grid& g);
hipError_t launch_strip_finish(const Grid& g, Strips st, const ShotGeo* geo, int k, int nshot,
                               hipStream_t s);
hipError_t launch_snap_save(const Grid& g, const float* f, Snaps sn, const ShotGeo* geo, int slot,
                            int nshot, hipStream_t s);
hipError_t launch_snap_restore(const Grid& g, float* f, Snaps sn, const ShotGeo* geo, int slot,
                               int nshot, hipStream_t s);
hipError_t launch_record(const Grid& g, const Rcv& r, const ShotGeo* geo, const float* f,
                         float* rec, int k, int nt, int nshot, hipStream_t s);
hipError_t launch_bwd_begin(const Grid& g, const float* s0, const float* s1, float* q0,
                            float* q1, float* su, float* sq, float* i1, float* i2,
                            const ShotGeo* geo, float spike, int icompen, int nshot, hipStream_t s);
hipError_t launch_maxabs(const Grid& g, const ShotGeo* geo, const float* i2, unsigned* out,
                         int nshot, hipStream_t s);
// up/down: [shot][mnz][g.mnx] (each shot fills its first mnx columns)
hipError_t launch_post(const Grid& g, const ShotGeo* geo, const float* v, long long mstride,
                       const float* i1, const float* i2, const unsigned* amax, float vmax,
                       float whitecoe, float* up, float* down, int nshot, hipStream_t s);
// s1/s2: the stacked image, mnz x gw; shot s adds into columns x_img..x_img+mnx-1,
// in shot order per cell.
hipError_t launch_stack_add(const Grid& g, const ShotGeo* geo, int gw, const float* up,
                            const float* down, float* s1, float* s2, int nshot, hipStream_t s);
// Reflectivity image [mnz][gw] into padded fields at dst + shot*rstride (each
// shot's window at x_img; zero off the interior); rstride = 0: shot 0 only.
hipError_t launch_refl_scatter(const Grid& g, const ShotGeo* geo, int gw, const float* img,
                               float* dst, long long rstride, int nshot, hipStream_t s);
// grad[mnz][gw] = sum over shots, in shot order, of G/pm at each shot's window.
hipError_t launch_grad_stack(const Grid& g, const ShotGeo* geo, int gw, const float* G,
                             const float* pm, long long mstride, int nshot, float* grad,
                             hipStream_t s);
// FP64 dot products of nvec pairs of float vectors of n elements, vector v at
// a + v*stride, b + v*stride: part holds nvec*dot64_part_size() doubles of
// partial sums, out[v] receives the dot product of pair v.  The reduction
// order is fixed and the same for every vector.
int dot64_part_size();
hipError_t launch_dot64(const float* a, const float* b, long long n, long long stride, int nvec,
                        double* part, double* out, hipStream_t s);
// y = y + a*x over n elements
hipError_t launch_axpy(float* y, const float* x, float a, long long n, hipStream_t s);
// out[c][r] = in[r][c] (rows x cols, row-major)
hipError_t launch_transpose(const float* in, float* out, int rows, int cols, hipStream_t s);
hipError_t launch_stack_finish(int nI, const float* s1, const float* s2, int nrec, int inorm,
                               float* img, float* illum, hipStream_t s);

}  // namespace rtm
