// rtm_setup.cpp -- parameter files, setup arithmetic, model padding and
// host-side stacking of the RTM driver (librtm_host.so).
//
// The arithmetic reproduces the reference's types and evaluation order (FP32
// where it declares float, FP64 where it writes double literals or calls
// libm), because grid sizes, placements, bins and the wavelet must agree
// bit for bit.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/rtm_host.h"
#include "rtm_host_internal.h"

namespace rtmh {

thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// Non-empty lines of a text file, trailing CR/LF removed.
static bool read_lines(const char* path, std::vector<std::string>& out) {
  FILE* f = fopen(path, "r");
  if (!f) return false;
  char buf[4096];
  while (fgets(buf, sizeof(buf), f)) {
    size_t n = strlen(buf);
    while (n && (buf[n - 1] == '\n' || buf[n - 1] == '\r')) buf[--n] = 0;
    const char* p = buf;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p) out.emplace_back(p);
  }
  fclose(f);
  return true;
}

static std::string first_token(const std::string& s) {
  size_t b = s.find_first_not_of(" \t");
  if (b == std::string::npos) return "";
  size_t e = s.find_first_of(" \t", b);
  return s.substr(b, e == std::string::npos ? std::string::npos : e - b);
}

}  // namespace rtmh

using namespace rtmh;

extern "C" {

const char* rtmh_last_error(void) { return g_err.c_str(); }

// kernel.cu:542-604: 27 (label line, value line) pairs, then the OutPara file.
int rtmh_read_config(const char* path, rtmh_config* c) {
  if (!path || !c) return set_error(RTMH_E_ARG, "null argument");
  std::vector<std::string> L;
  if (!read_lines(path, L)) return set_error(RTMH_E_IO, "cannot open %s", path);
  if (L.size() < 54) return set_error(RTMH_E_FORMAT, "%s: expected 27 label/value pairs, got %zu lines", path, L.size());
  memset(c, 0, sizeof(*c));
  int i = 1;  // value lines are the odd ones
  auto I = [&](int& v) { v = atoi(L[i].c_str()); i += 2; };
  auto F = [&](float& v) { v = (float)strtod(L[i].c_str(), nullptr); i += 2; };
  auto S = [&](char* v) { std::string t = first_token(L[i]); snprintf(v, 512, "%s", t.c_str()); i += 2; };
  I(c->nfdmax); I(c->nfdmin); I(c->n2); F(c->f0); F(c->fmax); F(c->df); I(c->nthita); F(c->eps);
  F(c->dv); I(c->ilste); I(c->ifv); F(c->whitecoe); F(c->hz); F(c->tao); I(c->inorm);
  I(c->icompen); I(c->nsmooth); F(c->wthite_phase); F(c->angle); I(c->nx_bg); I(c->nx_ed);
  I(c->nz_bg); I(c->nz_ed);
  S(c->out_seis); S(c->out_vp); S(c->out_dpr); S(c->out_para); S(c->result);
  FILE* f = fopen(c->out_para, "r");
  if (!f) return set_error(RTMH_E_IO, "cannot open OutPara file %s", c->out_para);
  int n = fscanf(f, "%f %f %d %d %d %d %d %d %d %d %d %d", &c->h, &c->tao1, &c->mod_nz,
                 &c->mod_nx, &c->nt1, &c->s_l1, &c->s_z1, &c->n, &c->ds, &c->r_x1, &c->nrec,
                 &c->dr);
  fclose(f);
  if (n != 12) return set_error(RTMH_E_FORMAT, "%s: expected 12 values, read %d", c->out_para, n);
  return RTMH_OK;
}

int rtmh_read_depths(const char* path, int nrec, float* d) {  // kernel.cu:695-700
  FILE* f = fopen(path, "r");
  if (!f) return set_error(RTMH_E_IO, "cannot open depth file %s", path);
  for (int i = 0; i < nrec; ++i)
    if (fscanf(f, "%f", &d[i]) != 1) {
      fclose(f);
      return set_error(RTMH_E_FORMAT, "%s: %d of %d depths", path, i, nrec);
    }
  fclose(f);
  return RTMH_OK;
}

void rtmh_derive(const rtmh_config* c, rtmh_derived* d) {  // kernel.cu:607-628
  const int n2 = c->n2;
  d->s_l = c->s_l1 + n2 - 1;
  d->r_x = c->r_x1 + n2 - 1;
  d->s_z = c->s_z1 + n2 - 1;
  d->s_r = (c->n - 1) * c->ds + d->s_l;
  d->nz = c->mod_nz + 2 * n2;
  d->nx = c->mod_nx + 2 * n2;
  const float tao = c->tao, h = c->h;
  d->nt = (int)((c->nt1 - 1) * c->tao1 / tao + 1.5);  // float quotient, double + 1.5
  d->taoh = tao / h;
  d->tao2 = (float)((double)tao * (double)tao);        // pow(tao,2): exact square in double
  d->h2 = (float)(1 / ((double)h * (double)h));        // 1/pow(h,2)
  d->taoh2 = d->tao2 * d->h2 / 2;
  d->nt2 = (int)(2.0 / (double)(c->f0 * tao)) + 1;    // FP32 product f0*tao
  d->hzx = c->hz / h;
  d->hzx2_1 = 1 / (d->hzx * d->hzx);
}

int rtmh_source_row(float depth_m, float hz, int n2) {  // kernel.cu:794-795
  const int N = (int)depth_m;
  const float a = std::fabs((float)N / hz);
  return (int)((a + (float)n2) - 1.0f);
}

float rtmh_ricker(float t1, float f0) {  // kernel.cu:1261-1266
  const float t00 = 1 / f0;
  const double a = kRefPi * f0 * (t1 - t00);
  const double a2 = a * a;  // pow(a, 2)
  return (float)((1 - 2 * a2) * std::exp(-a2));
}

void rtmh_wavelet_table(int nt, float tao, float f0, float* wav) {
  for (int j = 0; j <= nt; ++j) wav[j] = rtmh_ricker((float)j * tao, f0);  // f((k-1)*tao,f0)
}

void rtmh_blend_weights(int n2, float* w) {  // kernel.cu:687-691
  for (int i = 0; i <= n2; ++i) w[i] = (float)((1.0 * i) / (1.0 * n2));
}

int rtmh_read_model(const char* path, int mod_nz, int mod_nx, float* raw) {
  FILE* f = fopen(path, "rb");
  if (!f) return set_error(RTMH_E_IO, "cannot open model %s", path);
  size_t n = (size_t)mod_nz * mod_nx;
  size_t got = fread(raw, sizeof(float), n, f);
  fclose(f);
  if (got != n) return set_error(RTMH_E_FORMAT, "%s: %zu of %zu floats", path, got, n);
  return RTMH_OK;
}

// GPU_velocity_real.cpp:6-118: the raw model is column-major (x outer); the
// frame repeats the nearest interior edge value; ifv = 1 mirrors x.
void rtmh_pad_model(const float* raw, int mod_nz, int mod_nx, int n2, int ifv, float tao, float h,
                    float* v, float* r1) {
  const int nz = mod_nz + 2 * n2, nx = mod_nx + 2 * n2;
  auto at = [&](int z, int x) -> float& { return v[(size_t)z * nx + x]; };
  for (int x = 0; x < mod_nx; ++x)
    for (int z = 0; z < mod_nz; ++z) at(z + n2, x + n2) = raw[(size_t)x * mod_nz + z];
  for (int z = 0; z < nz; ++z) {
    const int zs = z < n2 ? n2 : (z >= nz - n2 ? nz - n2 - 1 : z);
    for (int x = 0; x < nx; ++x) {
      const int xs = x < n2 ? n2 : (x >= nx - n2 ? nx - n2 - 1 : x);
      if (zs != z || xs != x) at(z, x) = at(zs, xs);
    }
  }
  if (ifv == 1)
    for (int z = 0; z < nz; ++z)
      for (int x = 0; x < nx / 2; ++x) std::swap(at(z, x), at(z, nx - 1 - x));
  for (size_t i = 0; i < (size_t)nz * nx; ++i) {
    const float r = v[i] * tao / h;
    const float r2 = (float)((double)r * (double)r / 2);  // pow(r,2)/2
    r1[i] = (float)std::sqrt((double)r2);
  }
}

void rtmh_vrange(const float* v, long long n, float dv, float* vmin_o, float* vmax_o, int* nvel) {
  float vmin = v[0], vmax = v[0];  // kernel.cu:704-721
  for (long long i = 0; i < n; ++i) {
    vmin = v[i] < vmin ? v[i] : vmin;
    vmax = v[i] > vmax ? v[i] : vmax;
  }
  float vel = (float)(int)(vmin / dv) * dv;
  vmin = vel > vmin ? vel - dv : vel;
  vel = (float)(int)(vmax / dv) * dv;
  vmax = vel < vmax ? vel + dv : vel;
  *vmin_o = vmin;
  *vmax_o = vmax;
  *nvel = (int)((double)((vmax - vmin) / dv) + 1.5);
}

void rtmh_bins_needed(const float* v, long long n, float vmin, float dv, int nvel, int* need) {
  for (int i = 0; i < nvel; ++i) need[i] = 0;
  for (long long i = 0; i < n; ++i) {
    const int k = (int)((double)((v[i] - vmin) / dv) + 0.5);  // kernel.cu:735
    if (k >= 0 && k < nvel) need[k] = 1;
  }
}

// kernel.cu:1000-1059: shot-ordered float sums, /nrec, optional up/down.
void rtmh_stack(const float* ups, const float* downs, int nshot, long long nI, int inorm,
                float* img, float* ill) {
  std::vector<float> a(nI, 0.f), b(nI, 0.f);
  for (int m = 0; m < nshot; ++m) {
    const float* u = ups + (size_t)m * nI;
    const float* d = downs + (size_t)m * nI;
    for (long long i = 0; i < nI; ++i) {
      a[i] += u[i];
      b[i] += d[i];
    }
  }
  for (long long i = 0; i < nI; ++i) {
    float x = a[i] / nshot, y = b[i] / nshot;
    if (inorm == 1) x = x / y;
    img[i] = x;
    ill[i] = y;
  }
}

}  // extern "C"
