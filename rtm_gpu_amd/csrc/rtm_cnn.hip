// rtm_cnn.hip -- a small image-domain deblurring network for migrated images
// (not in the reference): 3x3 convolutions with bias and ReLU, a residual
// connection from input channel 0, trained with Adam on mean-squared error.
// Layers may be dilated (tap offsets times the layer's dilation), which
// widens the receptive field at the same cost.
// Convolutions run on the matrix cores: operands in BF16, accumulation in
// FP32 (v_mfma_f32_16x16x16_bf16); master weights, Adam state, loss and
// gradient reductions stay FP32 / FP64.
//
// Layouts.  Activations are NHWC in BF16 with the channel count padded to a
// multiple of 4 (Cp), so the four consecutive K values a lane feeds the MFMA
// are one 8-byte load.  Convolution as an implicit GEMM per wave: M = 16
// consecutive pixels of one row, N = 16 output channels, K = 9 taps x Cp
// input channels (k = tap*Cp + ci, padded to a multiple of 16).  The MFMA
// lane layout: A[m][k] at lane m + 16*(k/4), element k%4; B[k][n] at lane
// n + 16*(k/4); D[m][n] at lane n + 16*(m/4), element m%4.
//
//   forward      z = conv(a, W) + b; hidden layers store relu(z) in BF16,
//                the last layer y = z + x0 in FP32
//   input grad   dz_prev = conv(dz, W flipped and transposed) * (a > 0)
//   weight grad  dW[tap][co][ci] = sum_pix dz[pix][co] * a[pix + tap][ci]:
//                M = co, N = ci, K = pixels; one wave per image row and
//                16 x 16 channel block writes its partial sums, a second
//                kernel adds them in row order (FP64), so the result does not
//                depend on scheduling.
// Channel counts above 16 run in blocks of 16 output channels (one wave per
// block) and, for the weight gradient, blocks of 16 input channels.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "../../include/rtm_cnn.h"

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

constexpr int MAXC = 32;                    // channels per layer (two blocks of 16)
constexpr int MAXL = 16;                    // layers
constexpr int WG_STRIDE = 9 * MAXC * MAXC + MAXC;  // weight-grad partials per row

__device__ __forceinline__ bf16 to_bf16(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float(((unsigned)h) << 16); }

__device__ __forceinline__ s4 load4(const bf16* p) {  // 8-byte aligned
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  s4 r;
  r[0] = (short)(v.x & 0xffffu); r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffffu); r[3] = (short)(v.y >> 16);
  return r;
}

// mode 0: hidden layer (bias, ReLU, BF16 NHWC out); mode 1: last layer (bias
// + residual, FP32 out channel 0); mode 2: input gradient (mask by act > 0).
struct ConvArgs {
  const bf16* in;     // NHWC, cin_p channels
  const bf16* wt;     // [16][kpad]
  const float* bias;  // [16] (modes 0, 1)
  bf16* out;          // NHWC, cout_p channels (modes 0, 2)
  float* out32;       // [n][h][w] (mode 1)
  const float* resid; // [n][h][w] (mode 1)
  const bf16* mask;   // NHWC, cout_p channels (mode 2)
  int n, h, w, cin_p, kpad, cout_p, mode;
  int ncob;           // blocks of 16 output channels
  int dil;            // dilation: tap (ky, kx) reads (y + (ky-1)*dil, x + (kx-1)*dil)
};

__global__ __launch_bounds__(256) void k_conv(const ConvArgs a) {
  const int lane = threadIdx.x & 63;
  const int segs = (a.w + 15) >> 4;
  const long long wave_all = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long per = (long long)a.n * a.h * segs;
  if (wave_all >= per * a.ncob) return;
  const int cob = (int)(wave_all / per);  // block of 16 output channels
  const long long wave = wave_all - cob * per;
  const int seg = (int)(wave % segs);
  const long long row = wave / segs;  // nn*h + y
  const int y = (int)(row % a.h);
  const long long nn = row / a.h;
  const int m = lane & 15, kg = lane >> 4;
  const int x = seg * 16 + m;
  const bf16* wrow = a.wt + (long long)(cob * 16 + m) * a.kpad;  // output channel of lane & 15
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < a.kpad; k0 += 16) {
    const int k = k0 + 4 * kg;
    const int tap = k / a.cin_p, ci = k - tap * a.cin_p;
    s4 av = {0, 0, 0, 0};
    if (tap < 9) {
      const int yy = y + (tap / 3 - 1) * a.dil, xx = x + (tap % 3 - 1) * a.dil;
      if (x < a.w && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w)
        av = load4(a.in + ((nn * a.h + yy) * a.w + xx) * a.cin_p + ci);
    }
    const s4 bv = load4(wrow + k);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
  const int co = cob * 16 + (lane & 15);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int px = seg * 16 + 4 * kg + j;
    if (px >= a.w) continue;
    const long long p = (nn * a.h + y) * a.w + px;
    if (a.mode == 0) {
      if (co < a.cout_p) a.out[p * a.cout_p + co] = to_bf16(fmaxf(acc[j] + a.bias[co], 0.0f));
    } else if (a.mode == 1) {
      if (co == 0) a.out32[p] = (acc[j] + a.bias[0]) + a.resid[p];
    } else {
      if (co < a.cout_p) {
        const bool on = from_bf16(a.mask[p * a.cout_p + co]) > 0.0f;
        a.out[p * a.cout_p + co] = to_bf16(on ? acc[j] : 0.0f);
      }
    }
  }
}

// Weight / bias gradient partials of one image row per wave.
struct WgradArgs {
  const bf16* dz;   // NHWC, dz_p channels
  const bf16* act;  // NHWC, act_p channels (the layer input)
  float* part;      // [rows][WG_STRIDE]
  int n, h, w, dz_p, act_p, dil;
};

__global__ __launch_bounds__(256) void k_wgrad(const WgradArgs a) {
  const int lane = threadIdx.x & 63;
  const long long rows = (long long)a.n * a.h;
  const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int ncob = (a.dz_p + 15) >> 4, ncib = (a.act_p + 15) >> 4;
  if (wave >= rows * ncob * ncib) return;
  const long long row = wave % rows;
  const int blk = (int)(wave / rows);
  const int cob = blk % ncob, cib = blk / ncob;  // 16 x 16 block of (co, ci)
  const int y = (int)(row % a.h);
  const long long nn = row / a.h;
  const int c16 = lane & 15, kg = lane >> 4;
  const int coA = cob * 16 + c16, ciB = cib * 16 + c16;
  f4 acc[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.0f;
  for (int x0 = 0; x0 < a.w; x0 += 16) {
    // A[m = co][k = pixel]: lane co = c16 holds pixels 4*kg + j
    s4 av;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int px = x0 + 4 * kg + j;
      bf16 v = 0;
      if (px < a.w && coA < a.dz_p) v = a.dz[((nn * a.h + y) * a.w + px) * a.dz_p + coA];
      av[j] = (short)v;
      bsum += from_bf16(v);
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int yy = y + (t / 3 - 1) * a.dil;
      // B[k = pixel][n = ci]: lane ci = c16 holds pixels 4*kg + j, shifted by the tap
      s4 bv;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int px = x0 + 4 * kg + j;
        const int xx = px + (t % 3 - 1) * a.dil;
        bf16 v = 0;
        if (px < a.w && ciB < a.act_p && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w)
          v = a.act[((nn * a.h + yy) * a.w + xx) * a.act_p + ciB];
        bv[j] = (short)v;
      }
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
    }
  }
  float* P = a.part + row * WG_STRIDE;
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j)  // D[m = co][n = ci]
      P[(t * MAXC + cob * 16 + 4 * kg + j) * MAXC + ciB] = acc[t][j];
  // bias: lanes c16, c16+16, c16+32, c16+48 hold pixels of the same channel
  bsum += __shfl_xor(bsum, 16);
  bsum += __shfl_xor(bsum, 32);
  if (kg == 0 && cib == 0) P[9 * MAXC * MAXC + coA] = bsum;
}

// grad[layer params] from the row partials: W[co][ci][ky][kx] then b[co].
__global__ void k_wgrad_reduce(const float* part, long long rows, int cin, int cout, float* grad) {
  const int nw = cout * cin * 9;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nw + cout) return;
  int src;
  if (e < nw) {
    const int co = e / (cin * 9), r = e % (cin * 9), ci = r / 9, t = r % 9;
    src = (t * MAXC + co) * MAXC + ci;
  } else {
    src = 9 * MAXC * MAXC + (e - nw);
  }
  double s = 0.0;
  for (long long r = 0; r < rows; ++r) s += (double)part[r * WG_STRIDE + src];
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

inline int pad4(int c) { return (c + 3) & ~3; }
inline int pad16(int k) { return (k + 15) & ~15; }
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
  float* part = nullptr;
  double* sq = nullptr;
  int adam_t = 0;
  bool grad_ok = false;
  // device-resident training set
  float *ds_x = nullptr, *ds_t = nullptr;
  long long ds_cap = 0;
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

int launch_conv(rtm_cnn* c, const ConvArgs& a) {
  const long long waves = (long long)a.n * a.h * ((a.w + 15) / 16) * a.ncob;
  hipLaunchKernelGGL(k_conv, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, c->stream, a);
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
    WgradArgs g{c->dz[l], c->act[l], c->part, n, h, w, c->cp[l + 1], c->cp[l], c->dil[l]};
    const long long rows = (long long)n * h;
    const long long nwave = rows * ((c->cp[l + 1] + 15) / 16) * ((c->cp[l] + 15) / 16);
    hipLaunchKernelGGL(k_wgrad, dim3((unsigned)((nwave + 3) / 4)), dim3(256), 0, c->stream, g);
    CNNCHK(c, hipGetLastError());
    const int ne = c->c[l + 1] * c->c[l] * 9 + c->c[l + 1];
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((ne + 63) / 64), dim3(64), 0, c->stream, c->part, rows,
                       c->c[l], c->c[l + 1], c->grad + c->off[l]);
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
  rtm_cnn* c = new (std::nothrow) rtm_cnn();
  if (!c) return RTM_CNN_E_NOMEM;
  c->L = nlayer;
  for (int l = 0; l <= nlayer; ++l) {
    c->c[l] = channels[l];
    c->cp[l] = pad4(channels[l]);
  }
  for (int l = 0; l < nlayer; ++l) {
    c->dil[l] = dilation ? dilation[l] : 1;
    c->kf[l] = pad16(9 * c->cp[l]);
    c->kb[l] = pad16(9 * c->cp[l + 1]);
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
  if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->part, (size_t)max_rows * WG_STRIDE);
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
  if ((long long)need > c->ds_cap) {  // a larger training set replaces the old copy
    CNNCHK(c, hipStreamSynchronize(c->stream));
    cfree_dev(c, &c->ds_x);
    cfree_dev(c, &c->ds_t);
    c->ds_cap = 0;
    rc = calloc_dev(c, &c->ds_x, need * c->c[0]);
    if (rc == RTM_CNN_OK) rc = calloc_dev(c, &c->ds_t, need);
    if (rc != RTM_CNN_OK) return rc;
    c->ds_cap = (long long)need;
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
    double sum = 0.0;
    int nb = 0;
    for (int b0 = 0; b0 + batch <= n; b0 += batch) {
      for (int i = 0; i < batch; ++i) {
        const int k = order[b0 + i];
        CNNCHK(c, hipMemcpyAsync(c->x32 + (size_t)i * img * c->c[0], c->ds_x + (size_t)k * img * c->c[0],
                                 img * c->c[0] * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        CNNCHK(c, hipMemcpyAsync(c->t32 + (size_t)i * img, c->ds_t + (size_t)k * img,
                                 img * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
      }
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
