// rtm_capi.hip -- the C ABI of librtm_hip.so (declared in include/rtm_hip.h).
//
// Owns the device memory of a context and sequences the kernels of
// rtm_kernels.hip on the context's stream.  Replaces the device-side part of
// the reference's main(): allocation and uploads (kernel.cu:758-786), the
// forward time loop (kernel.cu:798-825), the backward loop with imaging
// (kernel.cu:859-933).  Deliver/Deliver_EFF copies are replaced by rotating
// buffer indices; the per-step host-to-device data copy (kernel.cu:905-909)
// by one upload of the whole gather per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/rtm_hip.h"
#include "rtm_kernels.h"

using namespace rtm;

struct EventPair {
  hipEvent_t a, b;
  int kind;
};

struct rtm_ctx {
  rtm_params p{};
  Grid g{};
  Model m{};
  Rcv r{};
  Strips st{};
  Snaps sn{};
  std::vector<int> snap_slot;  // per step k: snapshot slot or -1
  int nfd = 0;           // compile-time stencil half-length: longest operator of the table
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  size_t bytes = 0;
  std::vector<void*> allocs;
  // device buffers
  float *d_v = nullptr, *d_r1 = nullptr, *d_ctab = nullptr, *d_w = nullptr, *d_pm = nullptr;
  unsigned* d_ko = nullptr;
  size_t ctab_n = 0;
  float* F[3] = {nullptr, nullptr, nullptr};
  float* Q[3] = {nullptr, nullptr, nullptr};
  float *SU = nullptr, *SQ = nullptr, *I1 = nullptr, *I2 = nullptr;
  float *d_data = nullptr, *d_rec = nullptr;
  ShotGeo* d_geo = nullptr;    // per shot of the batch
  std::vector<ShotGeo> geo;
  int gw = 0;                  // columns of the stacked image
  int4* d_tiles = nullptr;     // frame (tile, shot) records of the batch
  int ntiles = 0, tile_cap = 0;  // FRAME_ALL records; capacity of d_tiles
  int nsplit = 0, nvec = 0;      // FRAME_SCALAR and FRAME_VEC records after them
  int fseg_tb = 0, fseg_side = 0;  // rows per frame record: top/bottom bands, side bands
  // model windows of a ragged batch (allocated on first use): slot w holds
  // one distinct window; shots whose windows are equal share a slot
  float *d_mv = nullptr, *d_mr1 = nullptr, *d_mpm = nullptr;
  unsigned* d_mko = nullptr;
  std::vector<char> shot_model;
  std::vector<int> win_ref;                 // shots using each slot
  std::vector<int> win_nx;                  // width of each slot's window
  std::vector<unsigned long long> win_fp;   // two 64-bit content fingerprints per slot
  float tab_vmin = 0.f, tab_dv = 1.f;  // bin parameters of the current table
  int tab_nrow = 1, tab_cs = 4;
  bool tab_ls = true;
  float *d_up = nullptr, *d_down = nullptr, *d_s1 = nullptr, *d_s2 = nullptr;
  unsigned* d_imax = nullptr;
  // host-side state
  std::vector<float> wav;
  int nshot = 0;
  int f0 = 0, f1 = 1, f2 = 2, q0 = 0, q1 = 1, q2 = 2;
  int kf = -1, kb = -2;  // next forward / backward step; -1/-2 = not begun
  // least-squares migration operators
  bool born = false;     // the forward pass is a Born pass (records hold dU)
  bool adj = false;      // the backward pass is the adjoint of Born modelling
  float* d_refl_img = nullptr;  // reflectivity on the stacked image grid [mnz][gw]
  float* d_refl = nullptr;      // padded reflectivity fields
  size_t refl_cap = 0;
  bool refl_ok = false;
  float* d_grad = nullptr;      // [mnz][gw]
  float* d_tp = nullptr;        // neighbour-tap planes of the transposed stencil
  size_t tp_cap = 0;
  bool tp_ok = false;           // planes match the current model windows
  float* d_stage = nullptr;     // one shot's trace-major gather [n][nt] (host transfers)
  double* d_dot = nullptr;      // FP64 dot-product scratch
  // finite-value watch: every check_every steps the newest fields are checked
  unsigned* d_flag = nullptr;   // 0xffffffff, or the step of the first non-finite check
  unsigned* h_flag = nullptr;   // pinned host copy
  int check_every = 100;
  bool checked = false;         // checks launched since the flag was last read
  bool capturing = false;       // inside rtm_migrate's graph capture
  bool model_ok = false, wav_ok = false, shots_ok = false, data_ok = false, save_strips = true;
  int zs = 32, nxt = 1;
  int zs_epi[16] = {0};  // rows per wave per interior kernel kind (0: zs)
  // captured migration (hipGraph)
  bool graph_on = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_f[3] = {0, 1, 2}, g_q[3] = {0, 1, 2}, g_kf = 0, g_kb = 0;
  // timing
  int timing = 0;  // bit set of kernel kinds
  std::vector<EventPair> pending;
  std::vector<hipEvent_t> pool;
  double t_ms[RTM_K_NKINDS] = {0};
  long long t_n[RTM_K_NKINDS] = {0};
};

namespace {

int fail(rtm_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(c, call)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail((c), RTM_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                  __FILE__, __LINE__);                                                      \
  } while (0)

int dalloc(rtm_ctx* c, void** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 16;
  hipError_t e = hipMalloc(p, n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, RTM_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(e));
  }
  c->allocs.push_back(*p);
  c->bytes += n;
  return RTM_OK;
}
template <typename T>
int dalloc_t(rtm_ctx* c, T** p, size_t count) {
  return dalloc(c, reinterpret_cast<void**>(p), count * sizeof(T));
}
// Release one allocation made by dalloc before the context is destroyed.
template <typename T>
void dfree_t(rtm_ctx* c, T** p, size_t count) {
  if (!*p) return;
  for (size_t i = 0; i < c->allocs.size(); ++i)
    if (c->allocs[i] == static_cast<void*>(*p)) {
      c->allocs.erase(c->allocs.begin() + i);
      c->bytes -= count * sizeof(T);
      break;
    }
  (void)hipFree(*p);
  *p = nullptr;
}

// IEEE binary16 -> float on the host.
float half_to_float(unsigned short h) {
  const unsigned s = (h >> 15) & 1u, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  float f;
  if (e == 0) f = ldexpf((float)m, -24);
  else if (e == 31) f = m ? NAN : INFINITY;
  else f = ldexpf((float)(m | 0x400u), (int)e - 25);
  return s ? -f : f;
}

// ---- timing -------------------------------------------------------------------
hipEvent_t get_event(rtm_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
int drain_timing(rtm_ctx* c) {
  if (c->pending.empty()) return RTM_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& ep : c->pending) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, ep.a, ep.b));
    c->t_ms[ep.kind] += ms;
    c->t_n[ep.kind] += 1;
    c->pool.push_back(ep.a);
    c->pool.push_back(ep.b);
  }
  c->pending.clear();
  return RTM_OK;
}
struct Timed {
  rtm_ctx* c;
  int kind;
  hipEvent_t a = nullptr, b = nullptr;
  bool on;
  Timed(rtm_ctx* c_, int k) : c(c_), kind(k), on((c_->timing >> k) & 1) {
    if (on) {
      a = get_event(c);
      b = get_event(c);
      if (a && b) (void)hipEventRecord(a, c->stream);
    }
  }
  ~Timed() {
    if (on && a && b) {
      (void)hipEventRecord(b, c->stream);
      c->pending.push_back({a, b, kind});
      if (c->pending.size() >= 4096) drain_timing(c);
    }
  }
};

// ---- launch helpers -----------------------------------------------------------
void fill_int_args(rtm_ctx* c, IntArgs& a) {
  a.g = c->g;
  a.m = c->m;
  a.r = c->r;
  a.geo = c->d_geo;
  a.nt = c->p.nt;
  a.nxt = c->nxt;
  a.zs = c->zs;
  a.data = c->d_data;
  a.uk = nullptr;
  a.su = c->SU; a.sq = c->SQ; a.i1 = c->I1; a.i2 = c->I2;
  a.icompen = c->p.icompen;
  a.ua = a.ub = a.uc = nullptr;
  a.image = 0;
}
// Rows per wave of one interior kernel kind (autotuned per kind, else zs).
int zs_for(const rtm_ctx* c, int epi) {
  return (epi >= 0 && epi < 16 && c->zs_epi[epi] > 0) ? c->zs_epi[epi] : c->zs;
}

void fill_frame_args(rtm_ctx* c, FrameArgs& a) {
  a.g = c->g;
  a.m = c->m;
  a.r = c->r;
  a.recs = c->d_tiles;
  a.geo = c->d_geo;
  a.nt = c->p.nt;
  a.f64comb = c->p.ilste != 0;
  a.data = c->d_data;
  a.st = c->st;
  a.save_strips = 0;
  a.rf = nullptr;
  a.rbps = 0;
  a.ua = a.ub = a.uc = nullptr;
  a.adjt = 0;
}

// ---- finite-value watch -------------------------------------------------------
int watch_reset(rtm_ctx* c) {
  HIPCHK(c, launch_flag_reset(c->d_flag, c->stream));
  c->checked = false;
  return RTM_OK;
}
bool watch_due(const rtm_ctx* c, int k, int last) {
  return c->check_every > 0 && (k % c->check_every == 0 || k == last);
}
int watch(rtm_ctx* c, const float* f, int k) {
  HIPCHK(c, launch_finite_check(f, (long long)c->nshot * c->g.fstride, k, c->d_flag, c->stream));
  c->checked = true;
  return RTM_OK;
}
// Read the flag after the checks of an *_advance call (not while capturing:
// rtm_migrate reads it after the graph launch).
int watch_read(rtm_ctx* c, const char* pass) {
  if (!c->checked || c->capturing) return RTM_OK;
  HIPCHK(c, hipMemcpyAsync(c->h_flag, c->d_flag, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->checked = false;
  if (*c->h_flag != 0xffffffffu)
    return fail(c, RTM_E_BLOWUP,
                "%s pass: a wavefield holds NaN or infinity at step %u (checked every %d steps); "
                "the time step may exceed the stability limit of the model, or the inputs are "
                "not finite", pass, *c->h_flag, c->check_every);
  return RTM_OK;
}

// Interior and frame of a step in one launch (k_step) when the context asks
// for it and the pair is one of the fused kinds.
static bool fused(const rtm_ctx* c, int epi, int mode) {
  return c->p.fuse_frame != 0 && step_fusable(epi, mode);
}

// The frame of a forward or receiver step: split between k_frame (the band
// ends at the corners, the side bands) and k_frame4 (the edge cells of the top
// and bottom bands) when rtm_params.frame_split asks for it and the batch has
// float4 records, else k_frame on the FRAME_ALL records.  The two launches
// write disjoint cells.
hipError_t launch_frame_split(rtm_ctx* c, int mode, FrameArgs f) {
  if (c->p.frame_split && c->nvec > 0 && frame4_supported(c->nfd)) {
    FrameArgs fs = f;
    fs.recs = c->d_tiles + c->ntiles;
    const hipError_t e = launch_frame(mode, c->nfd, fs, c->nshot, c->nsplit, c->stream);
    if (e != hipSuccess) return e;
    fs.recs = c->d_tiles + c->ntiles + c->nsplit;
    return launch_frame4(mode, c->nfd, fs, c->nshot, c->nvec, c->stream);
  }
  return launch_frame(mode, c->nfd, f, c->nshot, c->ntiles, c->stream);
}

int forward_step(rtm_ctx* c, int k) {
  const float wv = (k < c->p.nt2) ? c->wav[k - 1] : 0.0f;  // kernel.cu:812-813
  const bool quant = c->save_strips && c->st.fmt != RTM_STRIP_F32;
  if (quant && k >= 3) HIPCHK(c, launch_strip_finish(c->g, c->st, c->d_geo, k - 1, c->nshot, c->stream));
  {
    // Add (FP32 combine) or Add_Con (FP64 combine) (kernel.cu:46-114) on the
    // interior, Add + Hybrid1/2/3 on the frame (kernel.cu:116-208)
    const int epi = c->p.ilste ? EPI_FWD64 : EPI_FWD;
    IntArgs a;
    fill_int_args(c, a);
    a.u1 = c->F[c->f1]; a.u0 = c->F[c->f0]; a.u2 = c->F[c->f2];
    a.wav = wv; a.k = k;
    a.zs = zs_for(c, epi);
    FrameArgs f;
    fill_frame_args(c, f);
    f.u1 = c->F[c->f1]; f.u0 = c->F[c->f0]; f.u2 = c->F[c->f2];
    f.wav = wv; f.k = k;
    f.save_strips = c->save_strips ? 1 : 0;
    if (fused(c, epi, FR_FWD)) {
      Timed t(c, RTM_K_FWD_INTERIOR);  // the fused launch counts as the interior kind
      HIPCHK(c, launch_step(epi, FR_FWD, c->nfd, a, f, c->nshot, c->ntiles, c->stream));
    } else {
      {
        Timed t(c, RTM_K_FWD_INTERIOR);
        HIPCHK(c, launch_interior(epi, c->nfd, a, c->nshot, c->stream));
      }
      Timed t(c, RTM_K_FWD_FRAME);
      HIPCHK(c, launch_frame_split(c, FR_FWD, f));
    }
  }
  if (quant && k == c->p.nt - 1) HIPCHK(c, launch_strip_finish(c->g, c->st, c->d_geo, k, c->nshot, c->stream));
  if (c->save_strips && c->snap_slot[k] >= 0)  // checkpoint snapshot of u(k)
    HIPCHK(c, launch_snap_save(c->g, c->F[c->f2], c->sn, c->d_geo, c->snap_slot[k], c->nshot, c->stream));
  if (c->p.record && !c->born)  // receiver recording of u(k)
    HIPCHK(c, launch_record(c->g, c->r, c->d_geo, c->F[c->f2], c->d_rec, k, c->p.nt, c->nshot, c->stream));
  if (watch_due(c, k, c->p.nt - 1)) {
    int rc = watch(c, c->F[c->f2], k);
    if (rc != RTM_OK) return rc;
  }
  int t = c->f0; c->f0 = c->f1; c->f1 = c->f2; c->f2 = t;  // Deliver (kernel.cu:820)
  return RTM_OK;
}

int backward_step(rtm_ctx* c, int k) {
  const float wv = (k < c->p.nt2) ? c->wav[k + 1] : 0.0f;  // kernel.cu:889-890
  // Rel_* (kernel.cu:489-517) is split between the two interior kernels: the
  // terms of u(k) alone (sum u += u, I2 += u*u) go with its reconstruction,
  // the rest with the receiver step.  At a checkpoint step the snapshot
  // replaces u(k) after the reconstruction, and the non-compensated image
  // (I1 += q*u) needs u(k) in the receiver step anyway: there the receiver
  // step does all.
  const bool split = c->snap_slot[k] < 0 && c->p.icompen == 1;
  const int esrc = split ? EPI_SRCI : EPI_SRC, ercv = split ? EPI_RCVI : EPI_RCV;
  {  // BKAdd_EFF (kernel.cu:896): u(k) on the interior
    IntArgs a;
    fill_int_args(c, a);
    a.u1 = c->F[c->f1]; a.u0 = c->F[c->f0]; a.u2 = c->F[c->f2];
    a.wav = wv; a.k = k;
    Timed t(c, split ? RTM_K_SRCI_INTERIOR : RTM_K_SRC_INTERIOR);
    a.zs = zs_for(c, esrc);
    HIPCHK(c, launch_interior(esrc, c->nfd, a, c->nshot, c->stream));
  }
  // strips of slot k into the rings of u(k): BKEqual of step k-1 (kernel.cu:892).
  // Nothing in the receiver step reads those rings, so on ordinary steps the
  // receiver frame launch carries the restore; a checkpoint step restores
  // before its snapshot replaces u(k).
  const bool restore_in_frame = c->snap_slot[k] < 0 && !fused(c, ercv, FR_RCV);
  if (!restore_in_frame) {
    Timed t(c, RTM_K_STRIP_RESTORE);
    HIPCHK(c, launch_strip_restore(c->g, c->F[c->f2], c->st, c->d_geo, k, c->nshot, c->stream));
  }
  if (c->snap_slot[k] >= 0)  // checkpoint: the reconstructed u(k) is replaced by its snapshot
    HIPCHK(c, launch_snap_restore(c->g, c->F[c->f2], c->sn, c->d_geo, c->snap_slot[k], c->nshot, c->stream));
  { int t = c->f0; c->f0 = c->f1; c->f1 = c->f2; c->f2 = t; }  // Deliver_EFF (kernel.cu:903)
  {  // BKAdd (kernel.cu:913) + Rel_* (kernel.cu:923-930) on the interior, BKAdd
     // on the frame + BKHybrid1/2 (kernel.cu:913-920)
    IntArgs a;
    fill_int_args(c, a);
    a.u1 = c->Q[c->q1]; a.u0 = c->Q[c->q0]; a.u2 = c->Q[c->q2];
    a.uk = c->F[c->f1];
    a.wav = 0.f; a.k = k;
    a.zs = zs_for(c, ercv);
    FrameArgs f;
    fill_frame_args(c, f);
    f.u1 = c->Q[c->q1]; f.u0 = c->Q[c->q0]; f.u2 = c->Q[c->q2];
    f.wav = 0.f; f.k = k;
    if (restore_in_frame) {  // u(k) is F[f1] after the rotation
      f.rf = c->F[c->f1];
      f.rbps = strip_restore_blocks(c->g);
    }
    const int kind = split ? RTM_K_RCVI_INTERIOR : RTM_K_RCV_INTERIOR;
    if (fused(c, ercv, FR_RCV)) {
      Timed t(c, kind);  // the fused launch counts as the interior kind
      HIPCHK(c, launch_step(ercv, FR_RCV, c->nfd, a, f, c->nshot, c->ntiles, c->stream));
    } else {
      {
        Timed t(c, kind);
        HIPCHK(c, launch_interior(ercv, c->nfd, a, c->nshot, c->stream));
      }
      Timed t(c, RTM_K_RCV_FRAME);
      HIPCHK(c, launch_frame_split(c, FR_RCV, f));
    }
  }
  if (watch_due(c, k, 0)) {  // the reconstructed u(k) and the new receiver field
    int rc = watch(c, c->F[c->f1], k);
    if (rc == RTM_OK) rc = watch(c, c->Q[c->q2], k);
    if (rc != RTM_OK) return rc;
  }
  { int t = c->q0; c->q0 = c->q1; c->q1 = c->q2; c->q2 = t; }  // Deliver (kernel.cu:921)
  return RTM_OK;
}

// Born step k: the background step, then the scattered field from dU(k-1),
// dU(k-2) and D2u = (u(k) - 2u(k-1)) + u(k-2) (the background rotation
// leaves u(k), u(k-1), u(k-2) in f1, f0, f2), recorded at the receivers.
int born_step(rtm_ctx* c, int k) {
  int rc = forward_step(c, k);
  if (rc != RTM_OK) return rc;
  {
    IntArgs a;
    fill_int_args(c, a);
    a.u1 = c->Q[c->q1]; a.u0 = c->Q[c->q0]; a.u2 = c->Q[c->q2];
    a.ua = c->F[c->f1]; a.ub = c->F[c->f0]; a.uc = c->F[c->f2];
    a.wav = 0.f; a.k = k;
    Timed t(c, RTM_K_BORN_INTERIOR);
    a.zs = zs_for(c, c->p.ilste ? EPI_BORN64 : EPI_BORN);
    HIPCHK(c, launch_interior(c->p.ilste ? EPI_BORN64 : EPI_BORN, c->nfd, a, c->nshot, c->stream));
  }
  {
    FrameArgs a;
    fill_frame_args(c, a);
    a.u1 = c->Q[c->q1]; a.u0 = c->Q[c->q0]; a.u2 = c->Q[c->q2];
    a.ua = c->F[c->f1]; a.ub = c->F[c->f0]; a.uc = c->F[c->f2];
    a.wav = 0.f; a.k = k;
    Timed t(c, RTM_K_BORN_FRAME);
    HIPCHK(c, launch_frame(FR_BORN, c->nfd, a, c->nshot, c->ntiles, c->stream));
  }
  HIPCHK(c, launch_record(c->g, c->r, c->d_geo, c->Q[c->q2], c->d_rec, k, c->p.nt, c->nshot, c->stream));
  if (watch_due(c, k, c->p.nt - 1)) {
    rc = watch(c, c->Q[c->q2], k);
    if (rc != RTM_OK) return rc;
  }
  int t = c->q0; c->q0 = c->q1; c->q1 = c->q2; c->q2 = t;
  return RTM_OK;
}

// One update of the adjoint field: mu(k) from mu(k+1), mu(k+2) with P*r(k)
// at the receivers; with image, G += D2u * mu(k+2), D2u = (u(k+2) - 2u(k+1))
// + u(k) (the background rotation leaves u(k), u(k+1), u(k+2) in f1, f0, f2).
// restore: when not null, the frame launch also restores the strips of slot k
// into the rings of that field (the background u(k) of adjoint_step).
int adjoint_update(rtm_ctx* c, int k, bool image, float* restore = nullptr) {
  {
    IntArgs a;
    fill_int_args(c, a);
    a.u1 = c->Q[c->q1]; a.u0 = c->Q[c->q0]; a.u2 = c->Q[c->q2];
    a.ua = c->F[c->f2]; a.ub = c->F[c->f0]; a.uc = c->F[c->f1];
    a.image = image ? 1 : 0;
    a.wav = 0.f; a.k = k;
    Timed t(c, RTM_K_ADJ_INTERIOR);
    // per-cell (LS) rows: the transposed stencil, so the pair is exactly adjoint
    // on the interior; one Taylor row is symmetric
    const int epi = c->p.ilste == 0 ? EPI_ADJT : EPI_ADJ;
    a.zs = zs_for(c, EPI_ADJ);
    HIPCHK(c, launch_interior(epi, c->nfd, a, c->nshot, c->stream));
  }
  {
    FrameArgs a;
    fill_frame_args(c, a);
    a.u1 = c->Q[c->q1]; a.u0 = c->Q[c->q0]; a.u2 = c->Q[c->q2];
    a.wav = 0.f; a.k = k;
    a.adjt = c->p.ilste == 0;
    if (restore) {
      a.rf = restore;
      a.rbps = strip_restore_blocks(c->g);
    }
    Timed t(c, RTM_K_ADJ_FRAME);
    HIPCHK(c, launch_frame(FR_ADJ, c->nfd, a, c->nshot, c->ntiles, c->stream));
  }
  if (image && watch_due(c, k, 0)) {
    int rc = watch(c, c->Q[c->q2], k);
    if (rc != RTM_OK) return rc;
  }
  int t = c->q0; c->q0 = c->q1; c->q1 = c->q2; c->q2 = t;
  return RTM_OK;
}

// Adjoint step k: the background u(k) reconstructed as in backward_step,
// then the adjoint update with imaging.
int adjoint_step(rtm_ctx* c, int k) {
  const float wv = (k < c->p.nt2) ? c->wav[k + 1] : 0.0f;
  {
    IntArgs a;
    fill_int_args(c, a);
    a.u1 = c->F[c->f1]; a.u0 = c->F[c->f0]; a.u2 = c->F[c->f2];
    a.wav = wv; a.k = k;
    Timed t(c, RTM_K_SRC_INTERIOR);
    a.zs = zs_for(c, EPI_SRC);
    HIPCHK(c, launch_interior(EPI_SRC, c->nfd, a, c->nshot, c->stream));
  }
  // the strip restore of u(k): in the adjoint frame launch (nothing in the
  // adjoint update reads the rings of u(k)), except before a snapshot
  const bool restore_in_frame = c->snap_slot[k] < 0;
  if (!restore_in_frame) {
    Timed t(c, RTM_K_STRIP_RESTORE);
    HIPCHK(c, launch_strip_restore(c->g, c->F[c->f2], c->st, c->d_geo, k, c->nshot, c->stream));
  }
  if (c->snap_slot[k] >= 0)
    HIPCHK(c, launch_snap_restore(c->g, c->F[c->f2], c->sn, c->d_geo, c->snap_slot[k], c->nshot, c->stream));
  { int t = c->f0; c->f0 = c->f1; c->f1 = c->f2; c->f2 = t; }
  return adjoint_update(c, k, true, restore_in_frame ? c->F[c->f1] : nullptr);
}

// Frame records of a grid nx columns wide, tagged with the shot: top and
// bottom bands in runs of seg_tb rows, side bands in runs of seg (kind:
// FRAME_ALL, FRAME_SCALAR or FRAME_VEC).
void build_tiles(const Grid& g, int nx, int shot, int seg_tb, int seg, int kind,
                 std::vector<int4>& t) {
  int n = 0;
  frame_records(g, nx, shot, seg_tb, seg, kind, nullptr, &n);
  const size_t at = t.size();
  t.resize(at + n);
  frame_records(g, nx, shot, seg_tb, seg, kind, t.data() + at, &n);
}

int choose_zs(rtm_ctx* c, int nshot);
void drop_graph(rtm_ctx* c);

// The batch's frame records for the context's run lengths, three lists one
// after the other: every frame cell for k_frame (FRAME_ALL: the Born and
// adjoint frames, the one-launch step), then the split frame of the forward
// and receiver fields (FRAME_SCALAR for k_frame, FRAME_VEC for k_frame4).
// Each list is tile-major with the shot fastest, so that one XCD holds many
// shots of the same tiles at a time.
int upload_frame_records(rtm_ctx* c) {
  const int ns = (int)c->geo.size();
  std::vector<int4> rec;
  int count[3] = {0, 0, 0};
  for (int kind = FRAME_ALL; kind <= FRAME_VEC; ++kind) {
    std::vector<std::vector<int4>> per(ns);
    size_t most = 0;
    for (int s = 0; s < ns; ++s) {
      build_tiles(c->g, c->geo[s].nx, s, c->fseg_tb, c->fseg_side, kind, per[s]);
      most = std::max(most, per[s].size());
    }
    const size_t at = rec.size();
    for (size_t t = 0; t < most; ++t)
      for (int s = 0; s < ns; ++s)
        if (t < per[s].size()) rec.push_back(per[s][t]);
    count[kind] = (int)(rec.size() - at);
  }
  if ((long long)rec.size() > c->tile_cap) return fail(c, RTM_E_ARG, "frame records exceed capacity");
  HIPCHK(c, hipMemcpyAsync(c->d_tiles, rec.data(), rec.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ntiles = count[FRAME_ALL];
  c->nsplit = count[FRAME_SCALAR];
  c->nvec = count[FRAME_VEC];
  drop_graph(c);
  return RTM_OK;
}

// Upload the batch geometry and its frame records.
int upload_geometry(rtm_ctx* c) {
  const int ns = (int)c->geo.size();
  HIPCHK(c, hipMemcpyAsync(c->d_geo, c->geo.data(), ns * sizeof(ShotGeo), hipMemcpyHostToDevice, c->stream));
  int rc = upload_frame_records(c);
  if (rc != RTM_OK) return rc;
  c->nshot = ns;
  choose_zs(c, ns);
  drop_graph(c);
  c->shots_ok = true;
  c->data_ok = false;
  c->kf = -1;
  c->kb = -2;
  return RTM_OK;
}

// Pad a model window (v, r1 of nz x nx) into the padded layout and derive the
// per-cell operands: p = v*v*tao2*h2 (kernel.cu:73) and the byte offset of the
// velocity bin's coefficient row (kernel.cu:55-56: FP32 quotient, double +
// 0.5, truncation, clamped to the table).  Guard columns repeat the edge.
int upload_model_window(rtm_ctx* c, int nx, const float* v, const float* r1, float* dv_, float* dr_,
                        float* dp_, unsigned* dk_) {
  const Grid& g = c->g;
  std::vector<float> pv((size_t)g.fstride, 0.f), pr((size_t)g.fstride, 0.f), pp((size_t)g.fstride);
  std::vector<unsigned> pk((size_t)g.fstride);
  for (int z = 0; z < g.nz; ++z)
    for (int x = 0; x < g.pitch; ++x) {
      const int xs = std::min(std::max(x - g.x0, 0), nx - 1);
      pv[(size_t)z * g.pitch + x] = v[(size_t)z * nx + xs];
      pr[(size_t)z * g.pitch + x] = r1[(size_t)z * nx + xs];
    }
  for (size_t i = 0; i < pv.size(); ++i) {
    const float vv = pv[i];
    pp[i] = vv * vv * c->p.tao2 * c->p.h2;
    int kb = 0;
    if (c->tab_ls) {
      const float qf = (vv - c->tab_vmin) / c->tab_dv;
      kb = (int)((double)qf + 0.5);
      kb = std::min(std::max(kb, 0), c->tab_nrow - 1);
    }
    pk[i] = (unsigned)kb * (unsigned)c->tab_cs * 4u;
  }
  const size_t b = (size_t)g.fstride * sizeof(float);
  HIPCHK(c, hipMemcpyAsync(dv_, pv.data(), b, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(dr_, pr.data(), b, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(dp_, pp.data(), b, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(dk_, pk.data(), b, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RTM_OK;
}

int choose_zs(rtm_ctx* c, int nshot) {
  const Grid& g = c->g;
  int zs = c->p.zseg > 0 ? c->p.zseg : 32;
  c->nxt = (g.mnx + 255) / 256;
  if (c->p.zseg <= 0) {
    while (zs > 8) {
      long long nzg = (g.mnz + 4LL * zs - 1) / (4LL * zs);
      if ((long long)c->nxt * nzg * nshot >= 1024) break;
      zs /= 2;
    }
  }
  c->zs = zs;
  return RTM_OK;
}

void drop_graph(rtm_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
}

}  // namespace

// ===========================================================================
extern "C" {

int rtm_create(const rtm_params* p, rtm_ctx** out) {
  if (!p || !out) return RTM_E_ARG;
  *out = nullptr;
  rtm_ctx* c = new (std::nothrow) rtm_ctx();
  if (!c) return RTM_E_NOMEM;
  c->p = *p;
  const int mnz = p->nz - 2 * p->n2, mnx = p->nx - 2 * p->n2;
  int rc = RTM_OK;
  if (p->n2 < 1 || mnz < 2 || mnx < 2) rc = fail(c, RTM_E_ARG, "grid too small: nz=%d nx=%d n2=%d", p->nz, p->nx, p->n2);
  else if (p->nfdmax < 1 || p->nfdmax > 12) rc = fail(c, RTM_E_ARG, "nfdmax=%d outside 1..12", p->nfdmax);
  else if (p->nfdmax > p->n2) rc = fail(c, RTM_E_ARG, "nfdmax=%d exceeds n2=%d", p->nfdmax, p->n2);
  else if (p->nt < 3) rc = fail(c, RTM_E_ARG, "nt=%d < 3", p->nt);
  else if (p->max_shots < 1) rc = fail(c, RTM_E_ARG, "max_shots < 1");
  else if (p->n < 1 || p->ds < 1) rc = fail(c, RTM_E_ARG, "receivers n=%d ds=%d", p->n, p->ds);
  else if (p->s_z < 0 || p->s_z >= p->nz || p->s_l < 0 || p->s_r >= p->nx || p->s_r != (p->n - 1) * p->ds + p->s_l)
    rc = fail(c, RTM_E_ARG, "receiver line outside grid");
  else if (p->strip_format < 0 || p->strip_format > 2) rc = fail(c, RTM_E_ARG, "strip_format");
  else if (p->snap_every < 0) rc = fail(c, RTM_E_ARG, "snap_every=%d < 0", p->snap_every);
  else if (p->stack_nx < 0) rc = fail(c, RTM_E_ARG, "stack_nx=%d < 0", p->stack_nx);
  c->check_every = p->check_every == 0 ? 100 : p->check_every;
  if (rc != RTM_OK) {
    int code = rc;
    delete c;
    return code;
  }
  c->nfd = p->nfdmax;
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) { delete c; return RTM_E_HIP; }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return RTM_E_HIP; }

  Grid& g = c->g;
  g.nz = p->nz; g.nx = p->nx; g.n2 = p->n2; g.mnz = mnz; g.mnx = mnx; g.nfd = p->nfdmax;
  const int H = (c->nfd + 3) & ~3;
  int x0 = std::max(4, H);
  while ((x0 + g.n2) % 32 != 0) ++x0;  // interior rows start on a 128-byte line
  g.x0 = x0;
  g.pitch = ((x0 + g.nx + 4) + 63) / 64 * 64;
  g.fstride = (long long)g.nz * g.pitch;
  if (g.fstride >= (1LL << 28)) {  // buffer offsets of one shot's field stay below 2^30 bytes
    int code = fail(c, RTM_E_ARG, "grid of %lld padded cells exceeds 2^28", g.fstride);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return code;
  }
  c->r = Rcv{p->s_l, p->s_r, p->s_z, p->ds, p->n};

  const size_t fb = (size_t)p->max_shots * g.fstride;
  rc = RTM_OK;
  for (int i = 0; i < 3 && rc == RTM_OK; ++i) rc = dalloc_t(c, &c->F[i], fb);
  for (int i = 0; i < 3 && rc == RTM_OK; ++i) rc = dalloc_t(c, &c->Q[i], fb);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->SU, fb);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->SQ, fb);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->I1, fb);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->I2, fb);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_v, (size_t)g.fstride);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_r1, (size_t)g.fstride);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_pm, (size_t)g.fstride);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_ko, (size_t)g.fstride);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_w, (size_t)g.n2 + 1);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_geo, (size_t)p->max_shots);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_data, (size_t)p->max_shots * p->nt * p->n);
  if (rc == RTM_OK && p->record) rc = dalloc_t(c, &c->d_rec, (size_t)p->max_shots * p->nt * p->n);
  Strips& st = c->st;
  st.fmt = p->strip_format;
  st.nslot = p->nt;
  st.ss = 2LL * g.nfd * (g.mnz + g.mnx);
  const size_t esz = st.fmt == RTM_STRIP_F32 ? 4 : 2;
  if (rc == RTM_OK) rc = dalloc(c, &st.base, (size_t)p->max_shots * st.nslot * st.ss * esz);
  if (rc == RTM_OK && st.fmt != RTM_STRIP_F32) {
    rc = dalloc_t(c, &st.scale, (size_t)p->max_shots * st.nslot);
    if (rc == RTM_OK) rc = dalloc_t(c, &st.amax, (size_t)p->max_shots * st.nslot);
    if (rc == RTM_OK) rc = dalloc_t(c, &st.stage, (size_t)p->max_shots * 2 * st.ss);
  }
  // checkpoint snapshots: u(j), u(j+1) for j = m*snap_every, 2 <= j, j+1 <= nt-3
  c->snap_slot.assign(p->nt, -1);
  Snaps& sn = c->sn;
  sn.fmt = p->strip_format;
  sn.nsnap = 0;
  if (p->snap_every > 0)
    for (int j = p->snap_every; j + 1 <= p->nt - 3; j += p->snap_every) {
      if (j < 2) continue;
      c->snap_slot[j] = sn.nsnap++;
      c->snap_slot[j + 1] = sn.nsnap++;
    }
  if (rc == RTM_OK && sn.nsnap > 0) {
    const size_t cells = (size_t)g.nz * g.nx;
    rc = dalloc(c, &sn.base, (size_t)p->max_shots * sn.nsnap * cells * esz);
    if (rc == RTM_OK && sn.fmt != RTM_STRIP_F32) {
      rc = dalloc_t(c, &sn.scale, (size_t)p->max_shots * sn.nsnap);
      if (rc == RTM_OK) rc = dalloc_t(c, &sn.amax, (size_t)p->max_shots * sn.nsnap);
    }
  }
  const size_t nI = (size_t)g.mnz * g.mnx;
  c->gw = p->stack_nx > 0 ? p->stack_nx : g.mnx;
  const size_t nS = (size_t)g.mnz * c->gw;
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_up, (size_t)p->max_shots * nI);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_down, (size_t)p->max_shots * nI);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_s1, std::max(nS, nI));
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_s2, std::max(nS, nI));
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_imax, (size_t)p->max_shots);
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_flag, 1);
  if (rc == RTM_OK && hipHostMalloc(reinterpret_cast<void**>(&c->h_flag), sizeof(unsigned)) != hipSuccess) {
    c->h_flag = nullptr;
    rc = fail(c, RTM_E_NOMEM, "hipHostMalloc of the watch flag failed");
  }
  {
    c->fseg_tb = g.n2;
    c->fseg_side = frame_segment_rows(g.nfd);
    // the most records one shot of any width takes at the shortest runs the
    // autotuner tries, over the three lists (a narrower shot may take more
    // than the widest: its side bands may not pack into one wave)
    const int min_tb = std::max(1, (g.n2 + 2) / 3), min_side = std::max(1, c->fseg_side / 4);
    int most = 0;
    for (int nx = 1; nx <= g.nx; ++nx) {
      int tot = 0;
      for (int kind = FRAME_ALL; kind <= FRAME_VEC; ++kind) {
        int n = 0;
        frame_records(g, nx, 0, min_tb, min_side, kind, nullptr, &n);
        tot += n;
      }
      most = std::max(most, tot);
    }
    c->tile_cap = most * p->max_shots;
  }
  if (rc == RTM_OK) rc = dalloc_t(c, &c->d_tiles, (size_t)c->tile_cap);
  if (rc != RTM_OK) {
    int code = rc;
    rtm_destroy(c);
    return code;
  }
  // Zero everything once so guard columns hold finite values.
  for (int i = 0; i < 3; ++i) {
    HIPCHK(c, launch_fill(c->F[i], (long long)fb, 0.f, c->stream));
    HIPCHK(c, launch_fill(c->Q[i], (long long)fb, 0.f, c->stream));
  }
  HIPCHK(c, launch_fill(c->SU, (long long)fb, 0.f, c->stream));
  HIPCHK(c, launch_fill(c->SQ, (long long)fb, 0.f, c->stream));
  HIPCHK(c, launch_fill(c->I1, (long long)fb, 0.f, c->stream));
  HIPCHK(c, launch_fill(c->I2, (long long)fb, 0.f, c->stream));
  HIPCHK(c, launch_fill(c->d_s1, (long long)std::max(nS, nI), 0.f, c->stream));
  HIPCHK(c, launch_fill(c->d_s2, (long long)std::max(nS, nI), 0.f, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *out = c;
  return RTM_OK;
}

int rtm_destroy(rtm_ctx* c) {
  if (!c) return RTM_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  for (auto& ep : c->pending) { (void)hipEventDestroy(ep.a); (void)hipEventDestroy(ep.b); }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->h_flag) (void)hipHostFree(c->h_flag);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return RTM_OK;
}

// Message of the last failed call that takes no context (per host thread).
static thread_local std::string t_free_err = "null context";

const char* rtm_last_error(const rtm_ctx* c) { return c ? c->err.c_str() : t_free_err.c_str(); }
size_t rtm_device_bytes(const rtm_ctx* c) { return c ? c->bytes : 0; }

int rtm_set_model(rtm_ctx* c, const float* v, const float* r1, const float* w, const float* coef,
                  int nc, const int* index, int nvel, float vmin, float dv) {
  if (!c || !v || !r1 || !w || !coef) return RTM_E_ARG;
  const Grid& g = c->g;
  const bool ls = (c->p.ilste == 0);
  if (ls && (!index || nvel < 1)) return fail(c, RTM_E_ARG, "LS model needs index[nvel+1]");
  if (!ls && nc < c->p.nfdmax + 1) return fail(c, RTM_E_ARG, "Taylor table needs nfdmax+1 coefficients");
  if (!(dv > 0.f)) return fail(c, RTM_E_ARG, "dv must be > 0");
  const int nrow = ls ? nvel : 1;
  // The kernels run the longest operator of the table: every shorter row is
  // zero-padded, and w1 + 0*s == w1, so the taps beyond a bin's M change
  // nothing.  The strips keep their nfdmax rings (reference layout).
  int nfd = c->p.nfdmax;
  if (ls) {
    nfd = 1;
    for (int k = 0; k < nvel; ++k) nfd = std::max(nfd, index[k + 1] - index[k] - 1);
  }
  if (nfd > c->p.nfdmax) return fail(c, RTM_E_ARG, "table operator length %d > nfdmax %d", nfd, c->p.nfdmax);
  const int CS = ((2 + nfd) + 3) & ~3;
  std::vector<float> tab((size_t)nrow * CS, 0.f);
  const double h1 = 1.0 + (double)c->p.hzx2_1;
  for (int k = 0; k < nrow; ++k) {
    int top = ls ? index[k] : 0;
    int M = ls ? index[k + 1] - index[k] - 1 : c->p.nfdmax;
    if (M > nfd) return fail(c, RTM_E_ARG, "bin %d has M=%d > nfdmax=%d", k, M, nfd);
    if (ls && (top < 0 || index[k + 1] > nc)) return fail(c, RTM_E_ARG, "index[%d] outside c[%d]", k, nc);
    double c0d = 0.0;
    if (M >= 0) c0d = h1 * (double)coef[top];  // (1.0+hzx2_1)*Dc[M_top] (kernel.cu:58)
    memcpy(&tab[(size_t)k * CS], &c0d, sizeof(double));
    for (int l = 1; l <= M; ++l) tab[(size_t)k * CS + 1 + l] = coef[top + l];
  }
  int rc = RTM_OK;
  if (!c->d_ctab || c->ctab_n < tab.size()) {
    rc = dalloc_t(c, &c->d_ctab, tab.size());
    if (rc != RTM_OK) return rc;
    c->ctab_n = tab.size();
  }
  HIPCHK(c, hipMemcpyAsync(c->d_ctab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
  c->tab_vmin = vmin; c->tab_dv = dv; c->tab_nrow = nrow; c->tab_cs = CS; c->tab_ls = ls;
  rc = upload_model_window(c, g.nx, v, r1, c->d_v, c->d_r1, c->d_pm, c->d_ko);
  if (rc != RTM_OK) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_w, w, (g.n2 + 1) * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Model& m = c->m;
  m.mstride = 0;  // one model for every shot until rtm_set_shot_model
  m.v = c->d_v; m.r1 = c->d_r1; m.ctab = c->d_ctab; m.w = c->d_w;
  m.pm = c->d_pm; m.ko = c->d_ko;
  m.cs = CS; m.nrow = nrow;
  c->shot_model.assign(c->p.max_shots, 0);
  c->win_ref.assign(c->p.max_shots, 0);
  c->win_nx.assign(c->p.max_shots, 0);
  c->win_fp.assign(2 * (size_t)c->p.max_shots, 0);
  m.hzx2_1 = c->p.hzx2_1; m.tao2 = c->p.tao2; m.h2 = c->p.h2;
  m.taoh = c->p.taoh; m.taoh2 = c->p.taoh2;
  c->nfd = nfd;
  c->tp_ok = false;
  drop_graph(c);
  c->model_ok = true;
  return RTM_OK;
}

int rtm_set_wavelet(rtm_ctx* c, const float* wav) {
  if (!c || !wav) return RTM_E_ARG;
  c->wav.assign(wav, wav + c->p.nt + 1);
  drop_graph(c);
  c->wav_ok = true;
  return RTM_OK;
}

int rtm_set_shots(rtm_ctx* c, int nshot, const int* r_u, const int* r_x) {
  if (!c || !r_u || !r_x) return RTM_E_ARG;
  if (nshot < 1 || nshot > c->p.max_shots) return fail(c, RTM_E_ARG, "nshot=%d outside 1..%d", nshot, c->p.max_shots);
  const Grid& g = c->g;
  std::vector<ShotGeo> geo(nshot);
  for (int i = 0; i < nshot; ++i) {
    if (r_u[i] < 0 || r_u[i] >= g.nz || r_x[i] < 0 || r_x[i] >= g.nx)
      return fail(c, RTM_E_ARG, "shot %d source (%d,%d) outside grid", i, r_u[i], r_x[i]);
    geo[i] = ShotGeo{g.nx, g.mnx, c->p.s_l, c->p.s_r, c->p.n, r_u[i], r_x[i], 0, 0};
  }
  if (c->m.mstride != 0) return fail(c, RTM_E_STATE, "per-shot model windows need rtm_set_shot_geometry");
  c->geo = geo;
  return upload_geometry(c);
}

int rtm_set_shot_geometry(rtm_ctx* c, int nshot, const rtm_shot_geom* sg) {
  if (!c || !sg) return RTM_E_ARG;
  if (nshot < 1 || nshot > c->p.max_shots) return fail(c, RTM_E_ARG, "nshot=%d outside 1..%d", nshot, c->p.max_shots);
  const Grid& g = c->g;
  std::vector<ShotGeo> geo(nshot);
  for (int i = 0; i < nshot; ++i) {
    const rtm_shot_geom& q = sg[i];
    const int mnx = q.nx - 2 * g.n2;
    if (q.nx > g.nx || mnx < 2)
      return fail(c, RTM_E_ARG, "shot %d: width %d outside %d..%d", i, q.nx, 2 * g.n2 + 2, g.nx);
    if (q.r_u < 0 || q.r_u >= g.nz || q.r_x < 0 || q.r_x >= q.nx)
      return fail(c, RTM_E_ARG, "shot %d source (%d,%d) outside its grid", i, q.r_u, q.r_x);
    if (q.n < 1 || q.n > c->p.n || q.s_l < 0 || q.s_r != q.s_l + (q.n - 1) * c->p.ds || q.s_r >= q.nx)
      return fail(c, RTM_E_ARG, "shot %d receivers s_l=%d s_r=%d n=%d", i, q.s_l, q.s_r, q.n);
    if (q.x_img < 0 || q.x_img + mnx > c->gw)
      return fail(c, RTM_E_ARG, "shot %d image columns %d..%d outside 0..%d", i, q.x_img, q.x_img + mnx, c->gw);
    geo[i] = ShotGeo{q.nx, mnx, q.s_l, q.s_r, q.n, q.r_u, q.r_x, q.x_img, 0};
  }
  c->geo = geo;
  // Model windows were laid out for the previous batch's widths: once the
  // context holds per-shot windows, every shot of the new batch needs its
  // own again (rtm_forward_begin / rtm_born_begin check this).
  std::fill(c->shot_model.begin(), c->shot_model.end(), 0);
  c->win_ref.assign(c->p.max_shots, 0);
  c->win_nx.assign(c->p.max_shots, 0);
  c->win_fp.assign(2 * (size_t)c->p.max_shots, 0);
  c->tp_ok = false;
  return upload_geometry(c);
}

// Two independent 64-bit fingerprints of a window's bytes (width included);
// windows whose fingerprints and widths agree share one slot.
static void window_fingerprint(int nx, const float* v, const float* r1, size_t n,
                               unsigned long long* f) {
  unsigned long long a = 0xcbf29ce484222325ULL ^ (unsigned)nx, b = 0x9e3779b97f4a7c15ULL + (unsigned)nx;
  for (const float* p : {v, r1})
    for (size_t i = 0; i < n; ++i) {
      unsigned w;
      memcpy(&w, p + i, 4);
      a = (a ^ w) * 0x100000001b3ULL;
      b = (b + w + 0x9e3779b97f4a7c15ULL) * 0xbf58476d1ce4e5b9ULL;
      b ^= b >> 31;
    }
  f[0] = a;
  f[1] = b;
}

int rtm_set_shot_model(rtm_ctx* c, int shot, const float* v, const float* r1) {
  if (!c || !v || !r1) return RTM_E_ARG;
  if (!c->model_ok) return fail(c, RTM_E_STATE, "rtm_set_model first (coefficient table)");
  if (!c->shots_ok || shot < 0 || shot >= c->nshot) return fail(c, RTM_E_ARG, "shot %d of the batch", shot);
  const Grid& g = c->g;
  if (!c->d_mv) {
    const size_t n = (size_t)c->p.max_shots * g.fstride;
    int rc = dalloc_t(c, &c->d_mv, n);
    if (rc == RTM_OK) rc = dalloc_t(c, &c->d_mr1, n);
    if (rc == RTM_OK) rc = dalloc_t(c, &c->d_mpm, n);
    if (rc == RTM_OK) rc = dalloc_t(c, &c->d_mko, n);
    if (rc != RTM_OK) return rc;
  }
  if (c->win_ref.size() != (size_t)c->p.max_shots) {
    c->win_ref.assign(c->p.max_shots, 0);
    c->win_nx.assign(c->p.max_shots, 0);
    c->win_fp.assign(2 * (size_t)c->p.max_shots, 0);
  }
  const int nx = c->geo[shot].nx;
  unsigned long long fp[2];
  window_fingerprint(nx, v, r1, (size_t)g.nz * nx, fp);
  if (c->shot_model[shot]) c->win_ref[c->geo[shot].win] -= 1;  // the shot's previous window
  int w = -1;
  for (int i = 0; i < c->p.max_shots && w < 0; ++i)
    if (c->win_ref[i] > 0 && c->win_nx[i] == nx && c->win_fp[2 * i] == fp[0] && c->win_fp[2 * i + 1] == fp[1])
      w = i;
  if (w < 0) {  // a new window: the first slot no shot uses
    for (int i = 0; i < c->p.max_shots && w < 0; ++i)
      if (c->win_ref[i] == 0) w = i;
    if (w < 0) return fail(c, RTM_E_STATE, "no free model window slot");
    const size_t o = (size_t)w * g.fstride;
    int rc = upload_model_window(c, nx, v, r1, c->d_mv + o, c->d_mr1 + o, c->d_mpm + o, c->d_mko + o);
    if (rc != RTM_OK) {
      c->shot_model[shot] = 0;
      return rc;
    }
    c->win_nx[w] = nx;
    c->win_fp[2 * w] = fp[0];
    c->win_fp[2 * w + 1] = fp[1];
    c->tp_ok = false;
    }
  c->win_ref[w] += 1;
  c->geo[shot].win = w;
  HIPCHK(c, hipMemcpyAsync(c->d_geo + shot, &c->geo[shot], sizeof(ShotGeo), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Model& m = c->m;
  m.mstride = g.fstride;
  m.v = c->d_mv; m.r1 = c->d_mr1; m.pm = c->d_mpm; m.ko = c->d_mko;
  c->shot_model[shot] = 1;
  drop_graph(c);
  return RTM_OK;
}

// One shot's trace-major staging buffer for host transfers of gathers and
// records (the device layout is time-major); transposed on the GPU.
static int stage_buffer(rtm_ctx* c) {
  if (c->d_stage) return RTM_OK;
  return dalloc_t(c, &c->d_stage, (size_t)c->p.n * c->p.nt);
}

int rtm_set_gathers(rtm_ctx* c, const float* seis) {
  if (!c || !seis) return RTM_E_ARG;
  if (!c->shots_ok) return fail(c, RTM_E_STATE, "rtm_set_shots first");
  const int n = c->p.n, nt = c->p.nt;
  int rc = stage_buffer(c);
  if (rc != RTM_OK) return rc;
  for (int s = 0; s < c->nshot; ++s) {  // [n][nt] on the host -> [nt][n] on the device
    HIPCHK(c, hipMemcpyAsync(c->d_stage, seis + (size_t)s * n * nt, sizeof(float) * n * nt,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_transpose(c->d_stage, c->d_data + (size_t)s * nt * n, n, nt, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->data_ok = true;
  return RTM_OK;
}

int rtm_gathers_from_records(rtm_ctx* c) {
  if (!c) return RTM_E_ARG;
  if (!c->p.record || !c->d_rec) return fail(c, RTM_E_STATE, "context created without records");
  if (c->kf != c->p.nt) return fail(c, RTM_E_STATE, "no complete forward pass recorded");
  // records and gathers share the time-major [shot][nt][n] layout
  HIPCHK(c, hipMemcpyAsync(c->d_data, c->d_rec, sizeof(float) * c->nshot * c->p.nt * c->p.n,
                           hipMemcpyDeviceToDevice, c->stream));
  c->data_ok = true;
  return RTM_OK;
}

int rtm_forward_begin(rtm_ctx* c, int save_strips) {
  if (!c) return RTM_E_ARG;
  if (!c->model_ok || !c->wav_ok || !c->shots_ok) return fail(c, RTM_E_STATE, "model, wavelet and shots must be set");
  if (c->m.mstride != 0)
    for (int s = 0; s < c->nshot; ++s)
      if (!c->shot_model[s]) return fail(c, RTM_E_STATE, "shot %d has no model window", s);
  c->f0 = 0; c->f1 = 1; c->f2 = 2;
  c->save_strips = save_strips != 0;
  c->born = false;
  const float spike = (float)((double)c->wav[0] / 2.0);  // f(0.0,f0)/2.0 (kernel.cu:803)
  int wrc = watch_reset(c);
  if (wrc != RTM_OK) return wrc;
  HIPCHK(c, launch_fwd_begin(c->g, c->F[0], c->F[1], c->d_geo, spike, c->nshot, c->stream));
  if (c->save_strips) {  // Equal (kernel.cu:808)
    if (c->st.fmt != RTM_STRIP_F32)
      HIPCHK(c, launch_fill(reinterpret_cast<float*>(c->st.amax), (long long)c->nshot * c->st.nslot, 0.f, c->stream));
    HIPCHK(c, launch_strip_save(c->g, c->F[0], c->st, c->d_geo, 0, c->nshot, c->stream));
    HIPCHK(c, launch_strip_save(c->g, c->F[1], c->st, c->d_geo, 1, c->nshot, c->stream));
    if (c->st.fmt != RTM_STRIP_F32) {
      HIPCHK(c, launch_strip_finish(c->g, c->st, c->d_geo, 0, c->nshot, c->stream));
      HIPCHK(c, launch_strip_finish(c->g, c->st, c->d_geo, 1, c->nshot, c->stream));
    }
  }
  if (c->p.record) {
    HIPCHK(c, launch_record(c->g, c->r, c->d_geo, c->F[0], c->d_rec, 0, c->p.nt, c->nshot, c->stream));
    HIPCHK(c, launch_record(c->g, c->r, c->d_geo, c->F[1], c->d_rec, 1, c->p.nt, c->nshot, c->stream));
  }
  c->kf = 2;
  c->kb = -2;
  return RTM_OK;
}

int rtm_forward_advance(rtm_ctx* c, int nsteps) {
  if (!c) return RTM_E_ARG;
  if (c->kf < 2) return fail(c, RTM_E_STATE, "rtm_forward_begin first");
  for (int i = 0; i < nsteps && c->kf < c->p.nt; ++i, ++c->kf) {
    int rc = forward_step(c, c->kf);
    if (rc != RTM_OK) return rc;
  }
  return watch_read(c, "forward");
}

int rtm_backward_begin(rtm_ctx* c) {
  if (!c) return RTM_E_ARG;
  if (c->kf != c->p.nt) return fail(c, RTM_E_STATE, "forward pass incomplete (k=%d of %d)", c->kf, c->p.nt);
  if (!c->save_strips) return fail(c, RTM_E_STATE, "forward pass ran without strips");
  if (!c->data_ok) return fail(c, RTM_E_STATE, "rtm_set_gathers first");
  // DBW1 = u(NT-2), DBW0 = u(NT-1) (kernel.cu:822-825): swap roles of f0/f1.
  std::swap(c->f0, c->f1);
  c->q0 = 0; c->q1 = 1; c->q2 = 2;
  c->adj = false;
  const float spike = (float)((double)c->wav[0] / 2.0);
  int wrc = watch_reset(c);
  if (wrc != RTM_OK) return wrc;
  HIPCHK(c, launch_bwd_begin(c->g, c->F[c->f0], c->F[c->f1], c->Q[0], c->Q[1], c->SU, c->SQ,
                             c->I1, c->I2, c->d_geo, spike, c->p.icompen, c->nshot, c->stream));
  // BKEqual of the first backward step (kernel.cu:892, k = NT-3): slot NT-2.
  HIPCHK(c, launch_strip_restore(c->g, c->F[c->f1], c->st, c->d_geo, c->p.nt - 2, c->nshot, c->stream));
  c->kb = c->p.nt - 3;
  return RTM_OK;
}

int rtm_backward_advance(rtm_ctx* c, int nsteps) {
  if (!c) return RTM_E_ARG;
  if (c->kb < -1) return fail(c, RTM_E_STATE, "rtm_backward_begin first");
  for (int i = 0; i < nsteps && c->kb >= 0; ++i, --c->kb) {
    int rc = backward_step(c, c->kb);
    if (rc != RTM_OK) return rc;
  }
  return watch_read(c, "backward");
}

int rtm_set_reflectivity(rtm_ctx* c, const float* m) {
  if (!c || !m) return RTM_E_ARG;
  const size_t n = (size_t)c->g.mnz * c->gw;
  if (!c->d_refl_img) {
    int rc = dalloc_t(c, &c->d_refl_img, n);
    if (rc != RTM_OK) return rc;
  }
  HIPCHK(c, hipMemcpy(c->d_refl_img, m, n * sizeof(float), hipMemcpyHostToDevice));
  c->refl_ok = true;
  return RTM_OK;
}

int rtm_born_begin(rtm_ctx* c) {
  if (!c) return RTM_E_ARG;
  if (!c->p.record || !c->d_rec) return fail(c, RTM_E_STATE, "Born modelling needs a context with record = 1");
  if (!c->refl_ok) return fail(c, RTM_E_STATE, "rtm_set_reflectivity first");
  for (int s = 0; s < c->nshot; ++s)
    if (c->geo[s].x_img + c->geo[s].mnx > c->gw)
      return fail(c, RTM_E_ARG, "shot %d image columns %d..%d outside the image width %d", s,
                  c->geo[s].x_img, c->geo[s].x_img + c->geo[s].mnx, c->gw);
  int rc = rtm_forward_begin(c, 1);
  if (rc != RTM_OK) return rc;
  // each shot reads its window of the reflectivity; a uniform batch shares one field
  const Grid& g = c->g;
  bool uniform = true;
  for (int s = 0; s < c->nshot; ++s) uniform = uniform && c->geo[s].nx == g.nx && c->geo[s].x_img == 0;
  const long long rstride = uniform ? 0 : g.fstride;
  const size_t need = (size_t)(uniform ? 1 : c->nshot) * g.fstride;
  if (c->refl_cap < need) {  // a ragged batch needs one field per shot: replace the shared one
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree_t(c, &c->d_refl, c->refl_cap);
    c->refl_cap = 0;
    const size_t cap = uniform ? need : (size_t)c->p.max_shots * g.fstride;
    rc = dalloc_t(c, &c->d_refl, cap);
    if (rc != RTM_OK) return rc;
    c->refl_cap = cap;
  }
  HIPCHK(c, launch_refl_scatter(g, c->d_geo, c->gw, c->d_refl_img, c->d_refl, rstride, c->nshot, c->stream));
  c->m.refl = c->d_refl;
  c->m.rstride = rstride;
  const long long fb = (long long)c->nshot * g.fstride;
  for (int i = 0; i < 3; ++i) HIPCHK(c, launch_fill(c->Q[i], fb, 0.f, c->stream));
  HIPCHK(c, launch_fill(c->d_rec, (long long)c->nshot * c->p.nt * c->p.n, 0.f, c->stream));
  c->q0 = 0; c->q1 = 1; c->q2 = 2;
  c->born = true;
  return RTM_OK;
}

int rtm_born_advance(rtm_ctx* c, int nsteps) {
  if (!c) return RTM_E_ARG;
  if (!c->born || c->kf < 2) return fail(c, RTM_E_STATE, "rtm_born_begin first");
  for (int i = 0; i < nsteps && c->kf < c->p.nt; ++i, ++c->kf) {
    int rc = born_step(c, c->kf);
    if (rc != RTM_OK) return rc;
  }
  return watch_read(c, "Born");
}

int rtm_adjoint_begin(rtm_ctx* c) {
  if (!c) return RTM_E_ARG;
  if (c->kf != c->p.nt) return fail(c, RTM_E_STATE, "forward pass incomplete (k=%d of %d)", c->kf, c->p.nt);
  if (!c->save_strips) return fail(c, RTM_E_STATE, "forward pass ran without strips");
  if (!c->data_ok) return fail(c, RTM_E_STATE, "rtm_set_gathers first (residuals)");
  if (c->p.ilste == 0 && !c->tp_ok) {  // neighbour-tap planes of the current model windows
    const Grid& g = c->g;
    int nwin = 1;  // slots 0 .. nwin-1 (one model: slot 0)
    if (c->m.mstride)
      for (int s = 0; s < c->nshot; ++s) nwin = std::max(nwin, c->geo[s].win + 1);
    const size_t per = (size_t)4 * c->nfd * g.fstride;
    if (per * sizeof(float) >= (1ull << 31))
      return fail(c, RTM_E_ARG, "grid too large for the transposed adjoint stencil");
    if (c->tp_cap < per * nwin) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      dfree_t(c, &c->d_tp, c->tp_cap);
      c->tp_cap = 0;
      int rc = dalloc_t(c, &c->d_tp, per * nwin);
      if (rc != RTM_OK) return rc;
      c->tp_cap = per * nwin;
    }
    for (int w = 0; w < nwin; ++w)
      if (!c->m.mstride || c->win_ref[w] > 0)
        HIPCHK(c, launch_tap_planes(g, c->m.ko + (size_t)w * c->m.mstride, c->m.ctab, c->nfd,
                                    c->d_tp + (size_t)w * per, c->stream));
    c->m.tp = c->d_tp;
    c->m.tpstride = c->m.mstride ? (long long)per : 0;
    c->tp_ok = true;
  }
  std::swap(c->f0, c->f1);  // f1 = u(nt-2), f0 = u(nt-1), as for the migration
  int wrc = watch_reset(c);
  if (wrc != RTM_OK) return wrc;
  const long long fb = (long long)c->nshot * c->g.fstride;
  for (int i = 0; i < 3; ++i) HIPCHK(c, launch_fill(c->Q[i], fb, 0.f, c->stream));
  HIPCHK(c, launch_fill(c->I1, fb, 0.f, c->stream));
  HIPCHK(c, launch_strip_restore(c->g, c->F[c->f1], c->st, c->d_geo, c->p.nt - 2, c->nshot, c->stream));
  c->q0 = 0; c->q1 = 1; c->q2 = 2;
  c->adj = true;
  c->kb = -2;
  c->kf = -1;  // the reconstruction consumes the forward end state
  for (int k = c->p.nt - 1; k >= c->p.nt - 2; --k) {  // mu(nt-1), mu(nt-2)
    int rc = adjoint_update(c, k, false);
    if (rc != RTM_OK) return rc;
  }
  c->kb = c->p.nt - 3;
  return RTM_OK;
}

int rtm_adjoint_advance(rtm_ctx* c, int nsteps) {
  if (!c) return RTM_E_ARG;
  if (!c->adj || c->kb < -1) return fail(c, RTM_E_STATE, "rtm_adjoint_begin first");
  for (int i = 0; i < nsteps && c->kb >= 0; ++i, --c->kb) {
    int rc = adjoint_step(c, c->kb);
    if (rc != RTM_OK) return rc;
  }
  return watch_read(c, "adjoint");
}

int rtm_adjoint_gradient(rtm_ctx* c, float* grad) {
  if (!c || !grad) return RTM_E_ARG;
  if (!c->adj || c->kb != -1) return fail(c, RTM_E_STATE, "adjoint pass incomplete");
  const size_t n = (size_t)c->g.mnz * c->gw;
  if (!c->d_grad) {
    int rc = dalloc_t(c, &c->d_grad, n);
    if (rc != RTM_OK) return rc;
  }
  HIPCHK(c, launch_grad_stack(c->g, c->d_geo, c->gw, c->I1, c->m.pm, c->m.mstride, c->nshot,
                              c->d_grad, c->stream));
  HIPCHK(c, hipMemcpyAsync(grad, c->d_grad, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RTM_OK;
}

// Shots [shot0, shot0 + nshot) of the batch (nshot = 0: through the last).
static int data_range(rtm_ctx* c, int shot0, int nshot, long long* off, long long* n) {
  if (shot0 < 0 || shot0 >= c->nshot || nshot < 0 || shot0 + nshot > c->nshot)
    return fail(c, RTM_E_ARG, "shots %d+%d of %d", shot0, nshot, c->nshot);
  if (nshot == 0) nshot = c->nshot - shot0;
  const long long per = (long long)c->p.nt * c->p.n;
  *off = shot0 * per;
  *n = nshot * per;
  return RTM_OK;
}

int rtm_data_dot(rtm_ctx* c, int x, int y, int shot0, int nshot, double* out) {
  if (!c || !out) return RTM_E_ARG;
  const float* v[2] = {c->d_data, c->d_rec};
  if (x < 0 || x > 1 || y < 0 || y > 1) return fail(c, RTM_E_ARG, "data vector %d, %d", x, y);
  if ((x == RTM_DATA_RECORDS || y == RTM_DATA_RECORDS) && !c->d_rec)
    return fail(c, RTM_E_STATE, "context created without records");
  if ((x == RTM_DATA_GATHERS || y == RTM_DATA_GATHERS) && !c->data_ok)
    return fail(c, RTM_E_STATE, "rtm_set_gathers first");
  long long off = 0, n = 0;
  int rc = data_range(c, shot0, nshot, &off, &n);
  if (rc != RTM_OK) return rc;
  if (!c->d_dot) {
    rc = dalloc_t(c, &c->d_dot, (size_t)dot64_part_size() + 1);
    if (rc != RTM_OK) return rc;
  }
  double* res = c->d_dot + dot64_part_size();
  HIPCHK(c, launch_dot64(v[x] + off, v[y] + off, n, 0, 1, c->d_dot, res, c->stream));
  HIPCHK(c, hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RTM_OK;
}

int rtm_data_axpy(rtm_ctx* c, float a, int shot0, int nshot) {
  if (!c) return RTM_E_ARG;
  if (!c->d_rec) return fail(c, RTM_E_STATE, "context created without records");
  if (!c->data_ok) return fail(c, RTM_E_STATE, "rtm_set_gathers first");
  long long off = 0, n = 0;
  int rc = data_range(c, shot0, nshot, &off, &n);
  if (rc != RTM_OK) return rc;
  HIPCHK(c, launch_axpy(c->d_data + off, c->d_rec + off, a, n, c->stream));
  return RTM_OK;
}

// ---- streaming copy (bandwidth reference) ----------------------------------
// A plain float4 copy, four independent 16-byte loads in flight per thread,
// enough blocks for every CU: the HBM rate a pure streaming kernel reaches on
// this device, next to which the stencil kernels' rates read.
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ src,
                                                     float4* __restrict__ dst, long long n4) {
  const long long stride = (long long)gridDim.x * 1024;
  for (long long i = (long long)blockIdx.x * 1024 + threadIdx.x; i < n4; i += stride) {
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long k = i + u * 256;
      v[u] = k < n4 ? src[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long k = i + u * 256;
      if (k < n4) dst[k] = v[u];
    }
  }
}

int rtm_stream_copy(const float* src, float* dst, long long n, void* stream) {
  if (!src || !dst || n < 0 || (n & 3) || ((reinterpret_cast<size_t>(src) |
                                             reinterpret_cast<size_t>(dst)) & 15)) {
    t_free_err = "rtm_stream_copy: null, misaligned or n not a multiple of 4";
    return RTM_E_ARG;
  }
  const long long n4 = n / 4;
  long long blocks = (n4 + 1023) / 1024;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(256), 0,
                     static_cast<hipStream_t>(stream), reinterpret_cast<const float4*>(src),
                     reinterpret_cast<float4*>(dst), n4);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    t_free_err = std::string("rtm_stream_copy: ") + hipGetErrorString(e);
    return RTM_E_HIP;
  }
  return RTM_OK;
}

// ---- FP64 dot products without allocation ---------------------------------
// The partial sums of rtm_data_dot_noalloc live in a file-scope device array,
// one copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol):
// the device address of each GPU's copy comes from hipGetSymbolAddress on the
// first call on that device and is kept per device.
static __device__ __align__(256) double scratch[RTM_DOT_NOALLOC_MAX_VEC * 1024];
static_assert(RTM_DOT_NOALLOC_MAX_VEC >= 1, "scratch size");
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int rtm_data_dot_noalloc(const float* x, const float* y, long long n, long long stride, int nvec,
                         double* out, void* stream) {
  if (!x || !y || !out) {
    t_free_err = "rtm_data_dot_noalloc: null vector or output";
    return RTM_E_ARG;
  }
  if (n < 0 || nvec < 1 || (nvec > 1 && stride < n)) {
    char b[160];
    snprintf(b, sizeof(b), "rtm_data_dot_noalloc: n = %lld, nvec = %d, stride = %lld", n, nvec,
             stride);
    t_free_err = b;
    return RTM_E_ARG;
  }
  if (nvec > RTM_DOT_NOALLOC_MAX_VEC) {
    char b[160];
    snprintf(b, sizeof(b), "rtm_data_dot_noalloc: %d vectors exceed the scratch maximum of %d",
             nvec, RTM_DOT_NOALLOC_MAX_VEC);
    t_free_err = b;
    return RTM_E_ARG;
  }
  if (dot64_part_size() * RTM_DOT_NOALLOC_MAX_VEC > (int)(sizeof(scratch) / sizeof(double))) {
    t_free_err = "rtm_data_dot_noalloc: scratch smaller than the partial sums";
    return RTM_E_STATE;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    t_free_err = std::string("rtm_data_dot_noalloc: hipGetDevice: ") + hipGetErrorString(e);
    return RTM_E_HIP;
  }
  double* part = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void* p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        t_free_err = std::string("rtm_data_dot_noalloc: hipGetSymbolAddress: ") +
                     hipGetErrorString(e);
        return RTM_E_HIP;
      }
      g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    part = g_scratch_ptr[dev];
  }
  e = launch_dot64(x, y, n, stride, nvec, part, out, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) {
    t_free_err = std::string("rtm_data_dot_noalloc: launch: ") + hipGetErrorString(e);
    return RTM_E_HIP;
  }
  return RTM_OK;
}

static int migrate_eager(rtm_ctx* c) {
  int rc = rtm_forward_begin(c, 1);
  if (rc == RTM_OK) rc = rtm_forward_advance(c, c->p.nt);
  if (rc == RTM_OK) rc = rtm_backward_begin(c);
  if (rc == RTM_OK) rc = rtm_backward_advance(c, c->p.nt);
  return rc;
}

int rtm_migrate(rtm_ctx* c) {
  if (!c) return RTM_E_ARG;
  if (!c->graph_on) return migrate_eager(c);
  if (!c->model_ok || !c->wav_ok || !c->shots_ok || !c->data_ok)
    return fail(c, RTM_E_STATE, "model, wavelet, shots and gathers must be set");
  if (!c->gexec) {
    // Capture the whole launch sequence once; the host-side buffer rotation
    // is deterministic, so the replay leaves the same state as the capture.
    const int saved_timing = c->timing;
    c->timing = 0;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int rc = migrate_eager(c);
    c->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->timing = saved_timing;
    if (rc != RTM_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(c, RTM_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    c->graph = g;
    e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(c);
      return fail(c, RTM_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    c->g_f[0] = c->f0; c->g_f[1] = c->f1; c->g_f[2] = c->f2;
    c->g_q[0] = c->q0; c->g_q[1] = c->q1; c->g_q[2] = c->q2;
    c->g_kf = c->kf;
    c->g_kb = c->kb;
  }
  HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
  c->f0 = c->g_f[0]; c->f1 = c->g_f[1]; c->f2 = c->g_f[2];
  c->q0 = c->g_q[0]; c->q1 = c->g_q[1]; c->q2 = c->g_q[2];
  c->kf = c->g_kf;
  c->kb = c->g_kb;
  c->save_strips = true;
  c->checked = c->check_every > 0;  // the graph holds the checks of both passes
  return watch_read(c, "migration");
}

int rtm_set_graph(rtm_ctx* c, int on) {
  if (!c) return RTM_E_ARG;
  drop_graph(c);
  c->graph_on = on != 0;
  return RTM_OK;
}

int rtm_set_zseg(rtm_ctx* c, const int* zseg) {
  if (!c || !zseg) return RTM_E_ARG;
  for (int i = 0; i < 3; ++i)
    if (zseg[i] < 1 || zseg[i] > 4096) return fail(c, RTM_E_ARG, "rows per wave %d", zseg[i]);
  const int fwd = c->p.ilste ? EPI_FWD64 : EPI_FWD;
  c->zs_epi[fwd] = zseg[0];
  c->zs_epi[EPI_SRC] = c->zs_epi[EPI_SRCI] = zseg[1];
  c->zs_epi[EPI_RCV] = c->zs_epi[EPI_RCVI] = zseg[2];
  c->zs_epi[EPI_BORN] = c->zs_epi[EPI_BORN64] = c->zs_epi[EPI_ADJ] = zseg[2];
  c->zs = zseg[0];
  drop_graph(c);
  return RTM_OK;
}

int rtm_set_frame_runs(rtm_ctx* c, int rows_tb, int rows_side) {
  if (!c) return RTM_E_ARG;
  const Grid& g = c->g;
  const int min_tb = std::max(1, (g.n2 + 2) / 3), min_side = std::max(1, frame_segment_rows(g.nfd) / 4);
  if (rows_tb < min_tb || rows_tb > g.n2 || rows_side < min_side || rows_side > g.nz)
    return fail(c, RTM_E_ARG, "frame runs %d / %d outside %d..%d / %d..%d", rows_tb, rows_side,
                min_tb, g.n2, min_side, g.nz);
  c->fseg_tb = rows_tb;
  c->fseg_side = rows_side;
  return c->geo.empty() ? RTM_OK : upload_frame_records(c);
}

int rtm_frame_runs(rtm_ctx* c, int* rows_tb, int* rows_side) {
  if (!c || !rows_tb || !rows_side) return RTM_E_ARG;
  *rows_tb = c->fseg_tb;
  *rows_side = c->fseg_side;
  return RTM_OK;
}

int rtm_autotune(rtm_ctx* c, int nsteps, int* zseg) {
  if (!c || nsteps < 1) return RTM_E_ARG;
  if (!c->model_ok || !c->wav_ok || !c->shots_ok)
    return fail(c, RTM_E_STATE, "model, wavelet and shots must be set");
  // candidate heights: 1, 2, 3, 4, 6 and 9 register-queue lengths of the
  // kernel (each kernel rounds a request to a multiple of its queue)
  static const int kMult[] = {1, 2, 3, 4, 6, 9};
  constexpr int kNC = 6, kRounds = 3;
  // Each interior kernel kind of the migration is timed where it runs: inside
  // forward and backward step pairs on the batch's own buffers (the frame,
  // strip and imaging kernels interleaved, the buffers rotating), one kind at a
  // time, the others held at their current choice.  Timed alone, back to back
  // on the same buffers, the receiver step preferred 8-row segments; inside
  // the step sequence 32 rows ran 13% faster.  The fields written here are
  // scratch (the passes are begun anew afterwards); the timing counters of
  // rtm_kernel_time are reset.
  const int saved_timing = c->timing, saved_check = c->check_every;
  c->check_every = -1;
  int rc = rtm_forward_begin(c, 1);
  if (rc == RTM_OK && !c->data_ok)
    rc = launch_fill(c->d_data, (long long)c->nshot * c->p.nt * c->p.n, 0.f, c->stream) == hipSuccess
             ? RTM_OK : fail(c, RTM_E_HIP, "autotune: fill");
  const int fwd = c->p.ilste ? EPI_FWD64 : EPI_FWD;
  const bool split = c->p.icompen == 1;
  const int kinds[3] = {fwd, split ? EPI_SRCI : EPI_SRC, split ? EPI_RCVI : EPI_RCV};
  const int timers[3] = {RTM_K_FWD_INTERIOR, split ? RTM_K_SRCI_INTERIOR : RTM_K_SRC_INTERIOR,
                         split ? RTM_K_RCVI_INTERIOR : RTM_K_RCV_INTERIOR};
  for (int ki = 0; ki < 3; ++ki)
    if (c->zs_epi[kinds[ki]] <= 0) c->zs_epi[kinds[ki]] = 32;
  const int nt = c->p.nt;
  int kf = 2, kb = nt - 3;
  auto pairs = [&](int n) -> int {
    for (int i = 0; i < n; ++i) {
      int r = forward_step(c, kf);
      if (r == RTM_OK) r = backward_step(c, kb);
      if (r != RTM_OK) return r;
      kf = kf + 1 < nt ? kf + 1 : 2;
      kb = kb > 0 ? kb - 1 : nt - 3;
    }
    return RTM_OK;
  };
  // One measurement: an untimed pair, then per-launch time over `per` pairs.
  // The candidates are measured in kRounds interleaved rounds, alternating in
  // direction, and compared on their summed times, so that drift of the
  // device (clocks, caches, the other kernels) falls on all of them alike.
  const int per = std::max(2, nsteps / 2);
  auto measure = [&](unsigned mask, double* ms) -> int {
    c->timing = 0;
    int r = pairs(1);
    if (r != RTM_OK) return r;
    c->timing = (int)mask;
    for (int k = 0; k < RTM_K_NKINDS; ++k)
      if (mask & (1u << k)) { c->t_ms[k] = 0.0; c->t_n[k] = 0; }
    r = pairs(per);
    if (r == RTM_OK) r = drain_timing(c);
    if (r != RTM_OK) return r;
    *ms = 0.0;
    for (int k = 0; k < RTM_K_NKINDS; ++k)
      if (mask & (1u << k)) *ms += c->t_n[k] ? c->t_ms[k] / c->t_n[k] : 1e300;
    return RTM_OK;
  };
  for (int ki = 0; ki < 3 && rc == RTM_OK; ++ki) {
    const int epi = kinds[ki];
    const int ql = interior_queue_rows(epi, c->nfd);
    double tot[kNC] = {0.0};
    for (int round = 0; round < kRounds && rc == RTM_OK; ++round)
      for (int i = 0; i < kNC && rc == RTM_OK; ++i) {
        const int ci = (round & 1) ? kNC - 1 - i : i;
        c->zs_epi[epi] = kMult[ci] * ql;
        double ms = 0.0;
        rc = measure(1u << timers[ki], &ms);
        tot[ci] += ms;
      }
    int bi = 0;
    for (int i = 1; i < kNC; ++i)
      if (tot[i] < tot[bi]) bi = i;
    c->zs_epi[epi] = kMult[bi] * ql;
  }
  // Frame record runs, timed on both frame kernels of the step pair: the top
  // and bottom bands in runs of n2, n2/2 or n2/3 rows (with float4 records:
  // their k_frame ends), then the side bands in runs of the default, half or a
  // quarter of it.  Shorter runs put more waves in flight at the cost of
  // recomputing two pre-blend rows per run.
  if (rc == RTM_OK) {
    const Grid& g = c->g;
    const int seg0 = frame_segment_rows(g.nfd);
    const int cand[2][3] = {{g.n2, (g.n2 + 1) / 2, (g.n2 + 2) / 3},
                            {seg0, std::max(1, seg0 / 2), std::max(1, seg0 / 4)}};
    // the frame kernels' time: their own kinds, or the fused launches that
    // carry them (counted under the interior kinds)
    const unsigned fmask =
        (fused(c, fwd, FR_FWD) ? 1u << RTM_K_FWD_INTERIOR : 1u << RTM_K_FWD_FRAME) |
        (fused(c, kinds[2], FR_RCV) ? 1u << timers[2] : 1u << RTM_K_RCV_FRAME);
    for (int which = 0; which < 2 && rc == RTM_OK; ++which) {
      int& run = which == 0 ? c->fseg_tb : c->fseg_side;
      double tot[3] = {0.0, 0.0, 0.0};
      for (int round = 0; round < kRounds && rc == RTM_OK; ++round)
        for (int i = 0; i < 3 && rc == RTM_OK; ++i) {
          const int ci = (round & 1) ? 2 - i : i;
          run = cand[which][ci];
          rc = upload_frame_records(c);
          double ms = 0.0;
          if (rc == RTM_OK) rc = measure(fmask, &ms);
          tot[ci] += ms;
        }
      int bi = 0;
      for (int i = 1; i < 3; ++i)
        if (tot[i] < tot[bi]) bi = i;
      run = cand[which][bi];
    }
    if (rc == RTM_OK) rc = upload_frame_records(c);
    // the frame split on and off (the four-column launch pays only when the
    // batch fills the GPU with frame waves)
    if (rc == RTM_OK && c->nvec > 0 && frame4_supported(c->nfd) && !fused(c, fwd, FR_FWD)) {
      double tot[2] = {0.0, 0.0};
      for (int round = 0; round < kRounds && rc == RTM_OK; ++round)
        for (int i = 0; i < 2 && rc == RTM_OK; ++i) {
          const int on = (round & 1) ? 1 - i : i;
          c->p.frame_split = on;
          double ms = 0.0;
          rc = measure(fmask, &ms);
          tot[on] += ms;
        }
      c->p.frame_split = tot[1] < tot[0] ? 1 : 0;
    }
  }
  c->timing = saved_timing;
  c->check_every = saved_check;
  for (int i = 0; i < RTM_K_NKINDS; ++i) { c->t_ms[i] = 0.0; c->t_n[i] = 0; }
  if (split) {  // checkpoint steps run the unsplit kernels of the same shape
    c->zs_epi[EPI_SRC] = c->zs_epi[EPI_SRCI];
    c->zs_epi[EPI_RCV] = c->zs_epi[EPI_RCVI];
  }
  c->zs_epi[EPI_BORN] = c->zs_epi[EPI_BORN64] = c->zs_epi[EPI_ADJ] = c->zs_epi[kinds[2]];
  c->zs = c->zs_epi[fwd];
  c->kf = -1;
  c->kb = -2;
  drop_graph(c);
  if (zseg) {  // the kernels of an ordinary migration step
    zseg[0] = c->zs_epi[fwd];
    zseg[1] = c->zs_epi[kinds[1]];
    zseg[2] = c->zs_epi[kinds[2]];
  }
  return rc;
}

int rtm_sync(rtm_ctx* c) {
  if (!c) return RTM_E_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RTM_OK;
}

int rtm_read_field(rtm_ctx* c, int shot, int which, float* out) {
  if (!c || !out) return RTM_E_ARG;
  if (shot < 0 || shot >= c->nshot) return fail(c, RTM_E_ARG, "shot %d", shot);
  const float* base = nullptr;
  switch (which) {
    case RTM_FIELD_U: base = c->F[c->f1]; break;
    case RTM_FIELD_U_PREV: base = c->F[c->f0]; break;
    case RTM_FIELD_Q: base = c->Q[c->q1]; break;
    case RTM_FIELD_I1: base = c->I1; break;
    case RTM_FIELD_I2: base = c->I2; break;
    case RTM_FIELD_SU: base = c->SU; break;
    case RTM_FIELD_SQ: base = c->SQ; break;
    default: return fail(c, RTM_E_ARG, "field %d", which);
  }
  const Grid& g = c->g;
  const int nx = c->geo[shot].nx;  // the shot's grid fills the first nx columns
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy2D(out, g.nx * sizeof(float), base + (size_t)shot * g.fstride + g.x0,
                        g.pitch * sizeof(float), nx * sizeof(float), g.nz, hipMemcpyDeviceToHost));
  for (int z = 0; z < g.nz; ++z)
    for (int x = 0; x < g.nx; ++x)
      if (x >= nx || (which >= RTM_FIELD_I1 &&  // images are defined on the interior only
                      (z < g.n2 || z >= g.nz - g.n2 || x < g.n2 || x >= nx - g.n2)))
        out[(size_t)z * g.nx + x] = 0.f;
  return RTM_OK;
}

int rtm_read_records(rtm_ctx* c, float* out) {
  if (!c || !out) return RTM_E_ARG;
  if (!c->p.record) return fail(c, RTM_E_STATE, "context created without records");
  const int n = c->p.n, nt = c->p.nt;
  int rc = stage_buffer(c);
  if (rc != RTM_OK) return rc;
  for (int s = 0; s < c->nshot; ++s) {  // [nt][n] on the device -> [n][nt] on the host
    HIPCHK(c, launch_transpose(c->d_rec + (size_t)s * nt * n, c->d_stage, nt, n, c->stream));
    HIPCHK(c, hipMemcpyAsync(out + (size_t)s * n * nt, c->d_stage, sizeof(float) * n * nt,
                             hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RTM_OK;
}

int rtm_read_strip(rtm_ctx* c, int shot, int k, float* out) {
  if (!c || !out) return RTM_E_ARG;
  if (shot < 0 || shot >= c->nshot || k < 0 || k >= c->p.nt) return fail(c, RTM_E_ARG, "shot/k");
  const Strips& st = c->st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const long long slot = (long long)shot * st.nslot + (k % st.nslot);
  if (st.fmt == RTM_STRIP_F32) {
    HIPCHK(c, hipMemcpy(out, (const float*)st.base + slot * st.ss, st.ss * sizeof(float), hipMemcpyDeviceToHost));
    return RTM_OK;
  }
  std::vector<unsigned short> h(st.ss);
  float sc = 1.f;
  HIPCHK(c, hipMemcpy(h.data(), (const unsigned short*)st.base + slot * st.ss, st.ss * 2, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(&sc, st.scale + slot, sizeof(float), hipMemcpyDeviceToHost));
  for (long long e = 0; e < st.ss; ++e) {
    unsigned short b = h[e];
    float f;
    if (st.fmt == RTM_STRIP_BF16) {
      unsigned u = ((unsigned)b) << 16;
      memcpy(&f, &u, 4);
    } else {
      f = half_to_float(b);
    }
    out[e] = f * sc;
  }
  return RTM_OK;
}

int rtm_image_post(rtm_ctx* c, float vmax, float whitecoe, float* up, float* down) {
  if (!c) return RTM_E_ARG;
  if (c->kb != -1 || c->adj) return fail(c, RTM_E_STATE, "backward pass incomplete");
  const Grid& g = c->g;
  const int nI = g.mnz * g.mnx;  // per shot: [mnz][mnx of the widest shot]
  HIPCHK(c, launch_fill(reinterpret_cast<float*>(c->d_imax), c->nshot, 0.f, c->stream));
  HIPCHK(c, launch_maxabs(g, c->d_geo, c->I2, c->d_imax, c->nshot, c->stream));
  HIPCHK(c, launch_post(g, c->d_geo, c->m.v, c->m.mstride, c->I1, c->I2, c->d_imax, vmax, whitecoe,
                        c->d_up, c->d_down, c->nshot, c->stream));
  HIPCHK(c, launch_stack_add(g, c->d_geo, c->gw, c->d_up, c->d_down, c->d_s1, c->d_s2, c->nshot,
                             c->stream));
  if (up) HIPCHK(c, hipMemcpyAsync(up, c->d_up, sizeof(float) * nI * c->nshot, hipMemcpyDeviceToHost, c->stream));
  if (down) HIPCHK(c, hipMemcpyAsync(down, c->d_down, sizeof(float) * nI * c->nshot, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RTM_OK;
}

int rtm_stack_finish(rtm_ctx* c, int nrec, int inorm, float* img, float* illum) {
  if (!c || !img || nrec < 1) return RTM_E_ARG;
  const int nI = c->g.mnz * c->gw;  // the stacked image
  // d_up/d_down of shot 0 are reused as the result buffers.
  HIPCHK(c, launch_stack_finish(nI, c->d_s1, c->d_s2, nrec, inorm, c->d_up, c->d_down, c->stream));
  HIPCHK(c, hipMemcpyAsync(img, c->d_up, sizeof(float) * nI, hipMemcpyDeviceToHost, c->stream));
  if (illum) HIPCHK(c, hipMemcpyAsync(illum, c->d_down, sizeof(float) * nI, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, launch_fill(c->d_s1, nI, 0.f, c->stream));
  HIPCHK(c, launch_fill(c->d_s2, nI, 0.f, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RTM_OK;
}

int rtm_set_timing(rtm_ctx* c, int mask) {
  if (!c) return RTM_E_ARG;
  int rc = drain_timing(c);
  if (rc != RTM_OK) return rc;
  c->timing = mask;
  for (int i = 0; i < RTM_K_NKINDS; ++i) { c->t_ms[i] = 0; c->t_n[i] = 0; }
  return RTM_OK;
}

int rtm_kernel_time(rtm_ctx* c, int kind, double* total_ms, long long* launches) {
  if (!c || kind < 0 || kind >= RTM_K_NKINDS) return RTM_E_ARG;
  int rc = drain_timing(c);
  if (rc != RTM_OK) return rc;
  if (total_ms) *total_ms = c->t_ms[kind];
  if (launches) *launches = c->t_n[kind];
  return RTM_OK;
}

}  // extern "C"
