// rtm_cnn.hip -- a small image-domain deblurring network for migrated images
// (not in the reference): 3x3 convolutions with bias and ReLU, a residual
// connection from input channel 0, trained with Adam on mean-squared error.
// Layers may be dilated (tap offsets times the layer's dilation), which
// widens the receptive field at the same cost.
// Convolutions run on gfx950's matrix cores: operands in BF16, accumulation
// in FP32 (v_mfma_f32_16x16x32_bf16, K = 32 per instruction); master
// weights, Adam state, loss and gradient reductions stay FP32 / FP64.
//
// Layouts.  Activations are NHWC in BF16 with the channel count padded to a
// multiple of 8 (Cp), so the eight consecutive K values a lane feeds the
// MFMA are one 16-byte load.  Convolution as an implicit GEMM: M = 16
// consecutive pixels of one row, N = 16 output channels, K = 9 taps x Cp
// input channels (k = tap*Cp + ci, padded to a multiple of 32).  The MFMA
// lane layout: A[m][k] at lane m + 16*(k/8), element k%8; B[k][n] at lane
// n + 16*(k/8), element k%8; D[m][n] at lane n + 16*(m/4), element m%4.
//
//   forward      z = conv(a, W) + b; hidden layers store relu(z) in BF16,
//                the last layer y = z + x0 in FP32
//   input grad   dz_prev = conv(dz, W flipped and transposed) * (a > 0)
//   weight grad  dW[tap][co][ci] = sum_pix dz[pix][co] * a[pix + tap][ci]:
//                M = co, N = ci, K = pixels, staged transposed through LDS;
//                each workgroup accumulates a range of rows and writes one
//                partial, two reduction kernels add the partials in a fixed
//                order (FP64), so the result does not depend on scheduling.
// Channel counts above 16 run in blocks of 16 (two blocks of output channels
// per convolution wave; four 16 x 16 blocks per weight-gradient workgroup).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "../../include/rtm_cnn.h"

namespace {

typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

constexpr int MAXC = 32;                    // channels per layer (two blocks of 16)
constexpr int MAXL = 16;                    // layers
constexpr int WG_STRIDE = 9 * MAXC * MAXC + MAXC;  // weight-grad partials per workgroup
constexpr int WG_GROUPS = 1024;             // most workgroups of one weight-gradient launch
constexpr int WG_LEVEL1 = 32;               // groups of the first reduction level
// Byte offsets into an activation (MAX_PIXELS x 32 channels x 2 bytes) stay
// below 2^31, the range of a buffer load's offset.
constexpr long long MAX_PIXELS = 1LL << 25;

__device__ __forceinline__ bf16 to_bf16(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float(((unsigned)h) << 16); }


// Buffer resources: a load whose offset is out of range (kDrop) returns zero,
// which is the convolution's zero padding without a branch around the load.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr unsigned kDrop = 0x80000000u;
__device__ __forceinline__ Rsrc rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ u32x4 bld16(Rsrc r, unsigned vo) {
  return __builtin_amdgcn_raw_buffer_load_b128(r, (int)vo, 0, 0);
}

// mode 0: hidden layer (bias, ReLU, BF16 NHWC out); mode 1: last layer (bias
// + residual, FP32 out channel 0); mode 2: input gradient (mask by act > 0).
struct ConvArgs {
  const bf16* in;     // NHWC, cin_p channels
  const bf16* wt;     // [MAXC][kpad]
  const float* bias;  // [MAXC] (modes 0, 1)
  bf16* out;          // NHWC, cout_p channels (modes 0, 2)
  float* out32;       // [n][h][w] (mode 1)
  const float* resid; // [n][h][w] (mode 1)
  const bf16* mask;   // NHWC, cout_p channels (mode 2)
  int n, h, w, cin_p, kpad, cout_p, mode;
  int ncob;           // blocks of 16 output channels
  int dil;            // dilation: tap (ky, kx) reads (y + (ky-1)*dil, x + (kx-1)*dil)
};

// Weight-stationary convolution: a wave holds its B operands (all NK K-steps
// of NCOB blocks of 16 output channels) in registers and walks 16-pixel row
// tiles, so a tile costs NK 16-byte input loads per lane and NK * NCOB MFMAs.
// K-step ks covers k = 32 ks .. 32 ks + 31; lane group kg feeds k = 32 ks +
// 8 kg .. +7, eight consecutive channels of one tap (cin_p is a multiple of 8).
template <int NK, int NCOB>
__global__ __launch_bounds__(256) void k_conv(const ConvArgs a) {
  const int lane = threadIdx.x & 63;
  const int m = lane & 15, kg = lane >> 4;
  b8 wv[NCOB][NK];
#pragma unroll
  for (int cb = 0; cb < NCOB; ++cb)
#pragma unroll
    for (int ks = 0; ks < NK; ++ks)
      wv[cb][ks] = *reinterpret_cast<const b8*>(a.wt + (cb * 16 + m) * a.kpad + ks * 32 + 8 * kg);
  int tdy[NK], tdx[NK], tci[NK];
  bool tok[NK];
#pragma unroll
  for (int ks = 0; ks < NK; ++ks) {
    const int k = ks * 32 + 8 * kg, tap = k / a.cin_p;
    tok[ks] = tap < 9;
    tdy[ks] = (tap / 3 - 1) * a.dil;
    tdx[ks] = (tap % 3 - 1) * a.dil;
    tci[ks] = k - tap * a.cin_p;
  }
  const unsigned hw = (unsigned)(a.h * a.w);
  const Rsrc RI = rsrc(a.in, (unsigned)a.n * hw * (unsigned)a.cin_p * 2u);
  const unsigned segs = (unsigned)(a.w + 15) >> 4;
  const unsigned tiles = (unsigned)(a.n * a.h) * segs;
  for (unsigned t = blockIdx.x * 4u + (threadIdx.x >> 6); t < tiles; t += gridDim.x * 4u) {
    const unsigned row = t / segs, seg = t - row * segs;
    const int y = (int)(row % (unsigned)a.h);
    const unsigned nn = row / (unsigned)a.h;
    const int x = (int)seg * 16 + m;
    b8 av[NK];
#pragma unroll
    for (int ks = 0; ks < NK; ++ks) {
      const int yy = y + tdy[ks], xx = x + tdx[ks];
      const bool ok = tok[ks] && x < a.w && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
      const unsigned vo =
          ok ? ((nn * hw + (unsigned)(yy * a.w + xx)) * (unsigned)a.cin_p + (unsigned)tci[ks]) * 2u
             : kDrop;
      av[ks] = __builtin_bit_cast(b8, bld16(RI, vo));
    }
    f4 acc[NCOB];
#pragma unroll
    for (int cb = 0; cb < NCOB; ++cb) acc[cb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < NK; ++ks)
#pragma unroll
      for (int cb = 0; cb < NCOB; ++cb)
        acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[ks], wv[cb][ks], acc[cb], 0, 0, 0);
    // D[pixel][channel]: lane & 15 = channel in the block, pixels 4 kg + j
#pragma unroll
    for (int cb = 0; cb < NCOB; ++cb) {
      const int co = cb * 16 + m;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int px = (int)seg * 16 + 4 * kg + j;
        if (px >= a.w) continue;
        const long long p = (long long)row * a.w + px;
        if (a.mode == 0) {
          if (co < a.cout_p) a.out[p * a.cout_p + co] = to_bf16(fmaxf(acc[cb][j] + a.bias[co], 0.0f));
        } else if (a.mode == 1) {
          if (co == 0) a.out32[p] = (acc[cb][j] + a.bias[0]) + a.resid[p];
        } else {
          if (co < a.cout_p) {
            const bool on = from_bf16(a.mask[p * a.cout_p + co]) > 0.0f;
            a.out[p * a.cout_p + co] = to_bf16(on ? acc[cb][j] : 0.0f);
          }
        }
      }
    }
  }
}

// Weight / bias gradient partials of a range of image rows per workgroup.
struct WgradArgs {
  const bf16* dz;   // NHWC, dz_p channels
  const bf16* act;  // NHWC, act_p channels (the layer input)
  float* part;      // [groups][WG_STRIDE]
  int n, h, w, dz_p, act_p, dil;
  int rows_per;     // image rows (over n*h) of one workgroup
};

constexpr int SROW = 40;  // LDS row stride (bf16) of a staged 32-pixel chunk

// dW[tap][co][ci] = sum_pix dz[pix][co] act[pix + tap][ci] as MFMAs with
// M = co, N = ci, K = 32 pixels of one row.  Per chunk the workgroup stages
// dz (32 px x dz_p) and the nine tap-shifted act chunks (32 px x act_p)
// transposed into LDS, so each lane's eight K values are one 16-byte LDS
// read; wave w takes the 16 x 16 block (co block w & 1, ci block w >> 1) and
// accumulates over all chunks of its rows in registers.  One partial per
// workgroup; the reduction adds them in a fixed order.
__global__ __launch_bounds__(256) void k_wgrad(const WgradArgs a) {
  __shared__ __attribute__((aligned(16))) bf16 sdz[MAXC * SROW];
  __shared__ __attribute__((aligned(16))) bf16 sact[9 * MAXC * SROW];
  const int tid = threadIdx.x, lane = tid & 63, wvi = tid >> 6;
  const int cob = wvi & 1, cib = wvi >> 1;
  const int m = lane & 15, kg = lane >> 4;
  const int R = a.n * a.h;
  const int r0 = blockIdx.x * a.rows_per, r1 = min(R, r0 + a.rows_per);
  for (int i = tid; i < MAXC * SROW; i += 256) sdz[i] = 0;  // padded channels stay zero
  for (int i = tid; i < 9 * MAXC * SROW; i += 256) sact[i] = 0;
  __syncthreads();
  const bool active = cob * 16 < a.dz_p && cib * 16 < a.act_p;
  f4 acc[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.0f;
  const int cvz = a.dz_p >> 3, cva = a.act_p >> 3;
  const Rsrc RZ = rsrc(a.dz, (unsigned)R * (unsigned)a.w * (unsigned)a.dz_p * 2u);
  const Rsrc RA = rsrc(a.act, (unsigned)R * (unsigned)a.w * (unsigned)a.act_p * 2u);
  for (int r = r0; r < r1; ++r) {
    const int y = r % a.h, nn = r / a.h;
    for (int x0 = 0; x0 < a.w; x0 += 32) {
      for (int i = tid; i < 32 * cvz; i += 256) {
        const int px = i / cvz, cg = i - px * cvz;
        const int xx = x0 + px;
        const unsigned vo = xx < a.w ? ((unsigned)(r * a.w + xx) * (unsigned)a.dz_p + 8u * cg) * 2u
                                     : kDrop;
        const u32x4 v = bld16(RZ, vo);
#pragma unroll
        for (int j = 0; j < 8; ++j)
          sdz[(8 * cg + j) * SROW + px] = (bf16)(v[j >> 1] >> (16 * (j & 1)));
      }
      for (int i = tid; i < 9 * 32 * cva; i += 256) {
        const int tap = i / (32 * cva), q = i - tap * 32 * cva, px = q / cva, cg = q - px * cva;
        const int yy = y + (tap / 3 - 1) * a.dil, xx = x0 + px + (tap % 3 - 1) * a.dil;
        const bool ok = yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
        const unsigned vo =
            ok ? ((unsigned)((nn * a.h + yy) * a.w + xx) * (unsigned)a.act_p + 8u * cg) * 2u : kDrop;
        const u32x4 v = bld16(RA, vo);
#pragma unroll
        for (int j = 0; j < 8; ++j)
          sact[(tap * MAXC + 8 * cg + j) * SROW + px] = (bf16)(v[j >> 1] >> (16 * (j & 1)));
      }
      __syncthreads();
      if (active) {
        const u32x4 araw = *reinterpret_cast<const u32x4*>(&sdz[(cob * 16 + m) * SROW + 8 * kg]);
        const b8 av = __builtin_bit_cast(b8, araw);
#pragma unroll
        for (int t = 0; t < 9; ++t) {
          const b8 bv = *reinterpret_cast<const b8*>(&sact[(t * MAXC + cib * 16 + m) * SROW + 8 * kg]);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
        }
        if (cib == 0)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            bsum += __uint_as_float(araw[j] << 16) + __uint_as_float(araw[j] & 0xffff0000u);
      }
      __syncthreads();
    }
  }
  if (!active) return;
  float* P = a.part + (long long)blockIdx.x * WG_STRIDE;
  const int ci = cib * 16 + m;
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j)  // D[m = co][n = ci]: co = cob*16 + 4 kg + j
      P[(t * MAXC + cob * 16 + 4 * kg + j) * MAXC + ci] = acc[t][j];
  // bias: lanes m, m+16, m+32, m+48 hold pixels of channel cob*16 + m
  bsum += __shfl_xor(bsum, 16);
  bsum += __shfl_xor(bsum, 32);
  if (kg == 0 && cib == 0) P[9 * MAXC * MAXC + cob * 16 + m] = bsum;
}

// Deterministic two-level sum of the workgroup partials into the layer's
// gradient (W[co][ci][ky][kx] then b[co]): level 1 adds group g's share of
// the partials in order (FP64), level 2 adds the groups in order.
__device__ __forceinline__ int wgrad_src(int e, int cin, int cout) {
  const int nw = cout * cin * 9;
  if (e < nw) {
    const int co = e / (cin * 9), r = e % (cin * 9), ci = r / 9, t = r % 9;
    return (t * MAXC + co) * MAXC + ci;
  }
  return 9 * MAXC * MAXC + (e - nw);
}

__global__ void k_wgrad_reduce1(const float* part, int P, int G, int cin, int cout, double* part2) {
  const int ne = cout * cin * 9 + cout;
  const int e = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (e >= ne) return;
  const int src = wgrad_src(e, cin, cout);
  const int p0 = (int)((long long)P * g / G), p1 = (int)((long long)P * (g + 1) / G);
  double s = 0.0;
  for (int p = p0; p < p1; ++p) s += (double)part[(long long)p * WG_STRIDE + src];
  part2[(long long)g * ne + e] = s;
}

__global__ void k_wgrad_reduce2(const double* part2, int G, int ne, float* grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double s = 0.0;
  for (int g = 0; g < G; ++g) s += part2[(long long)g * ne + e];
  grad[e] = (float)s;
}

// NCHW FP32 input -> NHWC BF16 (cp channels) and the FP32 residual channel 0.
__global__ void k_pack_input(const float* x, long long npix, int hw, int cin, int cp, bf16* act,
                             float* resid) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (long long)gridDim.x * blockDim.x) {
    const long long nn = p / hw, q = p % hw;
    for (int c = 0; c < cp; ++c)
      act[p * cp + c] = c < cin ? to_bf16(x[(nn * cin + c) * hw + q]) : (bf16)0;
    resid[p] = x[nn * cin * hw + q];
  }
}

// Loss gradient of mean((y - t)^2): dz = 2(y - t)/npix in BF16 channel 0 of
// a cp-channel NHWC array; diff = y - t for the loss.
__global__ void k_loss_grad(const float* y, const float* t, long long npix, int cp, float scale,
                            bf16* dz, float* diff) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (long long)gridDim.x * blockDim.x) {
    const float d = y[p] - t[p];
    diff[p] = d;
    for (int c = 0; c < cp; ++c) dz[p * cp + c] = c == 0 ? to_bf16(scale * d) : (bf16)0;
  }
}

// One training batch from the device-resident set: sample i of the batch is
// set entry order[b0 + i] (inputs c0 channels, then the target).
__global__ void k_gather_batch(const float* dsx, const float* dst, const int* order, int b0,
                               int batch, long long img, int c0, float* x32, float* t32) {
  const long long per = img * (c0 + 1);
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)batch * per;
       e += (long long)gridDim.x * blockDim.x) {
    const long long i = e / per, r = e - i * per;
    const long long k = order[b0 + i];
    if (r < img * c0) x32[i * img * c0 + r] = dsx[k * img * c0 + r];
    else t32[i * img + (r - img * c0)] = dst[k * img + (r - img * c0)];
  }
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
  m[i] = mi;
  v[i] = vi;
  p[i] = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// FP32 params of one layer -> BF16 forward weights [16][kf] (k = tap*cin_p +
// ci), backward-data weights [16][kb] (row ci, k = tap'*cout_p + co with the
// flipped tap 8 - tap) and the padded bias.
__global__ void k_pack_weights(const float* prm, int cin, int cout, int cin_p, int cout_p, int kf,
                               int kb, bf16* wf, bf16* wb, float* bias) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < MAXC * kf) {
    const int co = e / kf, k = e % kf, tap = k / cin_p, ci = k % cin_p;
    float v = 0.0f;
    if (co < cout && ci < cin && tap < 9) v = prm[(co * cin + ci) * 9 + tap];
    wf[e] = to_bf16(v);
  }
  if (e < MAXC * kb) {
    const int ci = e / kb, k = e % kb, tp = k / cout_p, co = k % cout_p;
    float v = 0.0f;
    if (ci < cin && co < cout && tp < 9) v = prm[(co * cin + ci) * 9 + (8 - tp)];
    wb[e] = to_bf16(v);
  }
  if (e < MAXC) bias[e] = e < cout ? prm[cout * cin * 9 + e] : 0.0f;
}

// FP64 sum of squares (fixed order: block partials, then one block).
__global__ __launch_bounds__(256) void k_sumsq(const float* d, long long n, double* part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    s += (double)d[i] * (double)d[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

inline int pad8(int c) { return (c + 7) & ~7; }
inline int pad32(int k) { return (k + 31) & ~31; }
inline unsigned nblk(long long n, int per = 256, long long cap = 8192) {
  long long b = (n + per - 1) / per;
  if (b < 1) b = 1;
  return (unsigned)(b > cap ? cap : b);
}

}  // namespace

struct rtm_cnn {
  int L = 0;
  int c[MAXL + 1] = {0};
  int cp[MAXL + 1] = {0};
  int kf[MAXL] = {0}, kb[MAXL] = {0};
  int dil[MAXL] = {0};
  int off[MAXL + 1] = {0};  // parameter offsets per layer
  int np = 0;
  long long cap_pix = 0, cap_rows = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<void*> allocs;
  float *prm = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
  bf16* wf[MAXL] = {nullptr};
  bf16* wb[MAXL] = {nullptr};
  float* bias[MAXL] = {nullptr};
  bf16* act[MAXL] = {nullptr};  // act[l]: input of layer l, NHWC cp[l]
  bf16* dz[MAXL] = {nullptr};   // dz[l]: gradient at the output of layer l, NHWC cp[l+1]
  float *x32 = nullptr, *resid = nullptr, *y32 = nullptr, *t32 = nullptr, *diff = nullptr;
  float* part = nullptr;    // weight-gradient partials [WG_GROUPS][WG_STRIDE]
  double* part2 = nullptr;  // first reduction level [WG_LEVEL1][WG_STRIDE]
  double* sq = nullptr;
  int adam_t = 0;
  bool grad_ok = false;
  // device-resident training set
  float *ds_x = nullptr, *ds_t = nullptr;
  int* ds_order = nullptr;
  long long ds_cap = 0;
  int ds_n = 0;
};

namespace {

int cfail(rtm_cnn* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define CNNCHK(c, expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return cfail((c), RTM_CNN_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int calloc_dev(rtm_cnn* c, T** p, size_t count) {
  void* q = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e != hipSuccess) return cfail(c, RTM_CNN_E_NOMEM, "hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
  c->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return RTM_CNN_OK;
}

// Release one allocation made by calloc_dev before the network is destroyed.
template <typename T>
void cfree_dev(rtm_cnn* c, T** p) {
  if (!*p) return;
  for (size_t i = 0; i < c->allocs.size(); ++i)
    if (c->allocs[i] == static_cast<void*>(*p)) {
      c->allocs.erase(c->allocs.begin() + i);
      break;
    }
  (void)hipFree(*p);
  *p = nullptr;
}

int check_shape(rtm_cnn* c, int n, int h, int w) {
  if (n < 1 || h < 1 || w < 1) return cfail(c, RTM_CNN_E_ARG, "shape %d x %d x %d", n, h, w);
  if ((long long)n * h * w > c->cap_pix) return cfail(c, RTM_CNN_E_ARG, "%lld pixels > capacity %lld", (long long)n * h * w, c->cap_pix);
  if ((long long)n * h > c->cap_rows) return cfail(c, RTM_CNN_E_ARG, "%lld rows > capacity %lld", (long long)n * h, c->cap_rows);
  return RTM_CNN_OK;
}

int pack_weights(rtm_cnn* c) {
  for (int l = 0; l < c->L; ++l) {
    const int ne = MAXC * (c->kf[l] > c->kb[l] ? c->kf[l] : c->kb[l]);
    hipLaunchKernelGGL(k_pack_weights, dim3(nblk(ne, 256, 1 << 20)), dim3(256), 0, c->stream,
                       c->prm + c->off[l], c->c[l], c->c[l + 1], c->cp[l], c->cp[l + 1], c->kf[l],
                       c->kb[l], c->wf[l], c->wb[l], c->bias[l]);
    CNNCHK(c, hipGetLastError());
  }
  return RTM_CNN_OK;
}

template <int NK, int NCOB>
void conv_go(const ConvArgs& a, unsigned blocks, hipStream_t s) {
  hipLaunchKernelGGL((k_conv<NK, NCOB>), dim3(blocks), dim3(256), 0, s, a);
}

int launch_conv(rtm_cnn* c, const ConvArgs& a) {
  const long long tiles = (long long)a.n * a.h * ((a.w + 15) / 16);
  const unsigned blocks = (unsigned)std::min<long long>((tiles + 3) / 4, 2048);
  const int nk = a.kpad / 32;
  switch (nk * 4 + a.ncob) {
    case 3 * 4 + 1: conv_go<3, 1>(a, blocks, c->stream); break;
    case 3 * 4 + 2: conv_go<3, 2>(a, blocks, c->stream); break;
    case 5 * 4 + 1: conv_go<5, 1>(a, blocks, c->stream); break;
    case 5 * 4 + 2: conv_go<5, 2>(a, blocks, c->stream); break;
    case 7 * 4 + 1: conv_go<7, 1>(a, blocks, c->stream); break;
    case 7 * 4 + 2: conv_go<7, 2>(a, blocks, c->stream); break;
    case 9 * 4 + 1: conv_go<9, 1>(a, blocks, c->stream); break;
    case 9 * 4 + 2: conv_go<9, 2>(a, blocks, c->stream); break;
    default: return cfail(c, RTM_CNN_E_ARG, "no convolution for K %d, %d channel blocks", a.kpad, a.ncob);
  }
  CNNCHK(c, hipGetLastError());
  return RTM_CNN_OK;
}

// Forward of the batch already packed in act[0] / resid; y32 = output.
int forward_dev(rtm_cnn* c, int n, int h, int w) {
  for (int l = 0; l < c->L; ++l) {
    ConvArgs a{};
    a.in = c->act[l];
    a.wt = c->wf[l];
    a.bias = c->bias[l];
    a.n = n; a.h = h; a.w = w;
    a.cin_p = c->cp[l];
    a.kpad = c->kf[l];
    a.cout_p = c->cp[l + 1];
    a.dil = c->dil[l];
    a.ncob = (l + 1 < c->L) ? (c->cp[l + 1] + 15) / 16 : 1;
    if (l + 1 < c->L) {
      a.mode = 0;
      a.out = c->act[l + 1];
    } else {
      a.mode = 1;
      a.out32 = c->y32;
      a.resid = c->resid;
    }
    int rc = launch_conv(c, a);
    if (rc != RTM_CNN_OK) return rc;
  }
  return RTM_CNN_OK;
}

int pack_input(rtm_cnn* c, const float* x_dev, int n, int h, int w) {
  const long long npix = (long long)n * h * w;
  hipLaunchKernelGGL(k_pack_input, dim3(nblk(npix)), dim3(256), 0, c->stream, x_dev, npix, h * w,
                     c->c[0], c->cp[0], c->act[0], c->resid);
  CNNCHK(c, hipGetLastError());
  return RTM_CNN_OK;
}

// Loss and gradients for the batch in act[0] / resid with targets t_dev.
int loss_grad_dev(rtm_cnn* c, const float* t_dev, int n, int h, int w, double* loss) {
  int rc = forward_dev(c, n, h, w);
  if (rc != RTM_CNN_OK) return rc;
  const long long npix = (long long)n * h * w;
  const int L = c->L;
  hipLaunchKernelGGL(k_loss_grad, dim3(nblk(npix)), dim3(256), 0, c->stream, c->y32, t_dev, npix,
                     c->cp[L], (float)(2.0 / (double)npix), c->dz[L - 1], c->diff);
  CNNCHK(c, hipGetLastError());
  for (int l = L - 1; l >= 0; --l) {
    const int rows = n * h;
    const int rows_per = (rows + WG_GROUPS - 1) / WG_GROUPS;
    const int P = (rows + rows_per - 1) / rows_per;
    WgradArgs g{c->dz[l], c->act[l], c->part, n, h, w, c->cp[l + 1], c->cp[l], c->dil[l], rows_per};
    hipLaunchKernelGGL(k_wgrad, dim3((unsigned)P), dim3(256), 0, c->stream, g);
    CNNCHK(c, hipGetLastError());
    const int ne = c->c[l + 1] * c->c[l] * 9 + c->c[l + 1];
    const int G = std::min(WG_LEVEL1, P);
    hipLaunchKernelGGL(k_wgrad_reduce1, dim3((ne + 255) / 256, G), dim3(256), 0, c->stream, c->part,
                       P, G, c->c[l], c->c[l + 1], c->part2);
    CNNCHK(c, hipGetLastError());
    hipLaunchKernelGGL(k_wgrad_reduce2, dim3((ne + 255) / 256), dim3(256), 0, c->stream, c->part2,
                       G, ne, c->grad + c->off[l]);
    CNNCHK(c, hipGetLastError());
    if (l > 0) {  // dz[l-1] = conv(dz[l], W^T flipped) * (act[l] > 0)
      ConvArgs a{};
      a.in = c->dz[l];
      a.wt = c->wb[l];
      a.out = c->dz[l - 1];
      a.mask = c->act[l];
      a.n = n; a.h = h; a.w = w;
      a.cin_p = c->cp[l + 1];
      a.kpad = c->kb[l];
      a.cout_p = c->cp[l];
      a.mode = 2;
      a.dil = c->dil[l];
      a.ncob = (c->cp[l] + 15) / 16;
      rc = launch_conv(c, a);
      if (rc != RTM_CNN_OK) return rc;
    }
  }
  if (loss) {
    hipLaunchKernelGGL(k_sumsq, dim3(256), dim3(256), 0, c->stream, c->diff, npix, c->sq);
    CNNCHK(c, hipGetLastError());
    std::vector<double> part(256);
    CNNCHK(c, hipMemcpyAsync(part.data(), c->sq, 256 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CNNCHK(c, hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (double v : part) s += v;
    *loss = s / (double)npix;
  }
  c->grad_ok = true;
  return RTM_CNN_OK;
}

int adam_dev(rtm_cnn* c, float lr, float b1, float b2, float eps) {
  c->adam_t += 1;
  const double c1 = 1.0 / (1.0 - pow((double)b1, c->adam_t));
  const double c2 = 1.0 / (1.0 - pow((double)b2, c->adam_t));
  hipLaunchKernelGGL(k_adam, dim3((c->np + 255) / 256), dim3(256), 0, c->stream, c->prm, c->grad,
                     c->am, c->av, c->np, lr, b1, b2, eps, (float)c1, (float)c2);
  CNNCHK(c, hipGetLastError());
  return pack_weights(c);
}

}  // namespace

extern "C" {

int rtm_cnn_create(int nlayer, const int* channels, const int* dilation, long long max_pixels,
                   long long max_rows, rtm_cnn** out) {
  if (!channels || !out) return RTM_CNN_E_ARG;
  *out = nullptr;
  if (nlayer < 1 || nlayer > MAXL) return RTM_CNN_E_ARG;
  for (int l = 0; l <= nlayer; ++l)
    if (channels[l] < 1 || channels[l] > MAXC) return RTM_CNN_E_ARG;
  if (dilation)
    for (int l = 0; l < nlayer; ++l)
      if (dilation[l] < 1 || dilation[l] > 64) return RTM_CNN_E_ARG;
  if (channels[nlayer] != 1 || max_pixels < 1 || max_rows < 1) return RTM_CNN_E_ARG;
  if (max_pixels > MAX_PIXELS || max_rows > max_pixels) return RTM_CNN_E_ARG;
  rtm_cnn* c = new (std::nothrow) rtm_cnn();
  if (!c) return RTM_CNN_E_NOMEM;
  c->L = nlayer;
  for (int l = 0; l <= nlayer; ++l) {
    c->c[l] = channels[l];
    c->cp[l] = pad8(channels[l]);
  }
  for (int l = 0; l < nlayer; ++l) {
    c->dil[l] = dilation ? dilation[l] : 1;
    c->kf[l] = pad32(9 * c->cp[l]);
    c->kb[l] = pad32(9 * c->cp[l + 1]);
    c->off[l + 1] = c->off[l] + c->c[l + 1] * c->c[l] * 9 + c->c[l + 1];
  }
  c->np = c->off[nlayer];
  c->cap_pix = max_pixels;
  c->cap_rows = max_rows;
  int rc = RTM_CNN_OK;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = RTM_CNN_E_HIP;
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->prm, c->np);
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->grad, c->np);
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->am, c->np);
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->av, c->np);
  for (int l = 0; l < nlayer && rc == RTM_CNN_OK; ++l) {
    rc = calloc_dev(c, &c->wf[l], (size_t)MAXC * c->kf[l]);
    if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->wb[l], (size_t)MAXC * c->kb[l]);
    if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->bias[l], MAXC);
    if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->act[l], (size_t)max_pixels * c->cp[l]);
    if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->dz[l], (size_t)max_pixels * c->cp[l + 1]);
  }
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->x32, (size_t)max_pixels * c->c[0]);
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->resid, (size_t)max_pixels);
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->y32, (size_t)max_pixels);
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->t32, (size_t)max_pixels);
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->diff, (size_t)max_pixels);
  if (rc == RTM_CNN_OK)
    rc = calloc_dev(c, &c->part, (size_t)std::min<long long>(max_rows, WG_GROUPS) * WG_STRIDE);
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->part2, (size_t)WG_LEVEL1 * WG_STRIDE);
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->sq, 256);
  if (rc == RTM_CNN_OK) {
    if (hipMemsetAsync(c->am, 0, c->np * sizeof(float), c->stream) != hipSuccess ||
        hipMemsetAsync(c->av, 0, c->np * sizeof(float), c->stream) != hipSuccess ||
        hipMemsetAsync(c->prm, 0, c->np * sizeof(float), c->stream) != hipSuccess)
      rc = RTM_CNN_E_HIP;
  }
  if (rc == RTM_CNN_OK) rc = pack_weights(c);
  if (rc == RTM_CNN_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = RTM_CNN_E_HIP;
  if (rc != RTM_CNN_OK) {
    rtm_cnn_destroy(c);
    return rc;
  }
  *out = c;
  return RTM_CNN_OK;
}

int rtm_cnn_destroy(rtm_cnn* c) {
  if (!c) return RTM_CNN_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return RTM_CNN_OK;
}

const char* rtm_cnn_last_error(const rtm_cnn* c) { return c ? c->err.c_str() : "null network"; }

int rtm_cnn_num_params(const rtm_cnn* c) { return c ? c->np : 0; }

int rtm_cnn_set_params(rtm_cnn* c, const float* p) {
  if (!c || !p) return RTM_CNN_E_ARG;
  CNNCHK(c, hipMemcpyAsync(c->prm, p, c->np * sizeof(float), hipMemcpyHostToDevice, c->stream));
  CNNCHK(c, hipMemsetAsync(c->am, 0, c->np * sizeof(float), c->stream));
  CNNCHK(c, hipMemsetAsync(c->av, 0, c->np * sizeof(float), c->stream));
  c->adam_t = 0;
  int rc = pack_weights(c);
  if (rc != RTM_CNN_OK) return rc;
  CNNCHK(c, hipStreamSynchronize(c->stream));
  return RTM_CNN_OK;
}

int rtm_cnn_get_params(rtm_cnn* c, float* p) {
  if (!c || !p) return RTM_CNN_E_ARG;
  CNNCHK(c, hipMemcpyAsync(p, c->prm, c->np * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  CNNCHK(c, hipStreamSynchronize(c->stream));
  return RTM_CNN_OK;
}

int rtm_cnn_forward(rtm_cnn* c, const float* x, int n, int h, int w, float* y) {
  if (!c || !x || !y) return RTM_CNN_E_ARG;
  int rc = check_shape(c, n, h, w);
  if (rc != RTM_CNN_OK) return rc;
  const size_t npix = (size_t)n * h * w;
  CNNCHK(c, hipMemcpyAsync(c->x32, x, npix * c->c[0] * sizeof(float), hipMemcpyHostToDevice, c->stream));
  rc = pack_input(c, c->x32, n, h, w);
  if (rc == RTM_CNN_OK) rc = forward_dev(c, n, h, w);
  if (rc != RTM_CNN_OK) return rc;
  CNNCHK(c, hipMemcpyAsync(y, c->y32, npix * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  CNNCHK(c, hipStreamSynchronize(c->stream));
  return RTM_CNN_OK;
}

int rtm_cnn_loss_grad(rtm_cnn* c, const float* x, const float* t, int n, int h, int w, double* loss,
                      float* grad) {
  if (!c || !x || !t) return RTM_CNN_E_ARG;
  int rc = check_shape(c, n, h, w);
  if (rc != RTM_CNN_OK) return rc;
  const size_t npix = (size_t)n * h * w;
  CNNCHK(c, hipMemcpyAsync(c->x32, x, npix * c->c[0] * sizeof(float), hipMemcpyHostToDevice, c->stream));
  CNNCHK(c, hipMemcpyAsync(c->t32, t, npix * sizeof(float), hipMemcpyHostToDevice, c->stream));
  rc = pack_input(c, c->x32, n, h, w);
  if (rc == RTM_CNN_OK) rc = loss_grad_dev(c, c->t32, n, h, w, loss);
  if (rc != RTM_CNN_OK) return rc;
  if (grad) CNNCHK(c, hipMemcpyAsync(grad, c->grad, c->np * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  CNNCHK(c, hipStreamSynchronize(c->stream));
  return RTM_CNN_OK;
}

int rtm_cnn_adam_step(rtm_cnn* c, float lr, float b1, float b2, float eps) {
  if (!c) return RTM_CNN_E_ARG;
  if (!c->grad_ok) return cfail(c, RTM_CNN_E_STATE, "rtm_cnn_loss_grad first");
  int rc = adam_dev(c, lr, b1, b2, eps);
  if (rc != RTM_CNN_OK) return rc;
  CNNCHK(c, hipStreamSynchronize(c->stream));
  return RTM_CNN_OK;
}

int rtm_cnn_fit(rtm_cnn* c, const float* x, const float* t, int n, int h, int w, int epochs,
                int batch, float lr, unsigned seed, double* loss_per_epoch) {
  if (!c || !x || !t || epochs < 0 || batch < 1) return RTM_CNN_E_ARG;
  if (batch > n) batch = n;
  int rc = check_shape(c, batch, h, w);
  if (rc != RTM_CNN_OK) return rc;
  const size_t img = (size_t)h * w;
  const size_t need = (size_t)n * img;
  if ((long long)need > c->ds_cap || n > c->ds_n) {  // a larger training set replaces the old copy
    CNNCHK(c, hipStreamSynchronize(c->stream));
    cfree_dev(c, &c->ds_x);
    cfree_dev(c, &c->ds_t);
    cfree_dev(c, &c->ds_order);
    c->ds_cap = 0;
    c->ds_n = 0;
    rc = calloc_dev(c, &c->ds_x, need * c->c[0]);
    if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->ds_t, need);
    if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->ds_order, (size_t)n);
    if (rc != RTM_CNN_OK) return rc;
    c->ds_cap = (long long)need;
    c->ds_n = n;
  }
  CNNCHK(c, hipMemcpyAsync(c->ds_x, x, need * c->c[0] * sizeof(float), hipMemcpyHostToDevice, c->stream));
  CNNCHK(c, hipMemcpyAsync(c->ds_t, t, need * sizeof(float), hipMemcpyHostToDevice, c->stream));
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  unsigned long long s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
  auto rnd = [&s]() {  // 64-bit LCG, high bits
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (unsigned)(s >> 33);
  };
  for (int ep = 0; ep < epochs; ++ep) {
    for (int i = n - 1; i > 0; --i) {  // Fisher-Yates
      const int j = (int)(rnd() % (unsigned)(i + 1));
      const int tmp = order[i]; order[i] = order[j]; order[j] = tmp;
    }
    // the previous epoch ended with a stream synchronisation (the loss read-back)
    CNNCHK(c, hipMemcpyAsync(c->ds_order, order.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    double sum = 0.0;
    int nb = 0;
    for (int b0 = 0; b0 + batch <= n; b0 += batch) {
      hipLaunchKernelGGL(k_gather_batch, dim3(nblk((long long)batch * img * (c->c[0] + 1))), dim3(256), 0,
                         c->stream, c->ds_x, c->ds_t, c->ds_order, b0, batch, (long long)img, c->c[0],
                         c->x32, c->t32);
      CNNCHK(c, hipGetLastError());
      rc = pack_input(c, c->x32, batch, h, w);
      double loss = 0.0;
      if (rc == RTM_CNN_OK) rc = loss_grad_dev(c, c->t32, batch, h, w, &loss);
      if (rc == RTM_CNN_OK) rc = adam_dev(c, lr, 0.9f, 0.999f, 1e-8f);
      if (rc != RTM_CNN_OK) return rc;
      sum += loss;
      ++nb;
    }
    if (loss_per_epoch) loss_per_epoch[ep] = nb ? sum / nb : 0.0;
  }
  CNNCHK(c, hipStreamSynchronize(c->stream));
  return RTM_CNN_OK;
}

}  // extern "C"
