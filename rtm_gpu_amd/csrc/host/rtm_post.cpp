// rtm_post.cpp -- post-stack output chain of the driver (kernel.cu:1110-1179):
// depth-to-time conversion of the stacked image, constant-phase rotation of
// the time image, and time-to-depth conversion back.
//
// The floating-point operation order of the reference helpers is kept
// (DisToTimeAndTimeToDis1D.cpp, phase_correction_ricker_decon.cpp) so that the
// three output files match the reference's bit for bit; the tests compare
// them against those helpers compiled from the reference sources.
#include <algorithm>
#include <cmath>
#include <vector>

#include "rtm_host.h"
#include "rtm_host_internal.h"

namespace {

// Linear resampling of one trace known at increasing abscissae xs[0..nin)
// onto the regular grid j*step, j < nout (DisToTimeAndTimeToDis1D.cpp:8-28).
// Sample 0 takes ys[0]; samples j with floor(xs[i-1]/step) < j <= floor(xs[i]/step)
// are interpolated between points i-1 and i.  The grid point j == nout that
// the last abscissa can reach is dropped, as there.
void resample_trace(float* out, const float* xs, const float* ys, int nout, int nin, float step) {
  out[0] = ys[0];
  int done = 0;  // last grid index written
  for (int i = 1; i < nin; ++i) {
    const int upto = (int)(xs[i] / step);
    for (int j = done + 1; j <= upto; ++j) {
      if (j >= nout) continue;
      const float frac = (j * step - xs[i - 1]) / (xs[i] - xs[i - 1]);
      out[j] = ys[i - 1] + frac * (ys[i] - ys[i - 1]);
    }
    done = upto;
  }
}

// Resample every trace of `d` (trace-major, nin samples) at the abscissae t0
// onto nout regular samples of `step`; traces are 0-based in d (the reference
// passes &D[i*Nz] with i counted from 0 whatever the first column).
void resample_all(float* out, const std::vector<float>& t0, const float* d, int ntr, int nout,
                  int nin, float step) {
  std::fill(out, out + (size_t)ntr * nout, 0.0f);
  for (int i = 0; i < ntr; ++i)
    resample_trace(out + (size_t)i * nout, &t0[(size_t)i * nin], d + (size_t)i * nin, nout, nin, step);
}

// Radix-2 decimation-in-time DFT of n = 2^k points in the form of the
// reference's kbfft (phase_correction_ricker_decon.cpp:331-384): the input
// (re, im) is permuted into (fr, fi) by bit reversal, the twiddles
// w_i = exp(-+ 2 pi i / n) are built in (re, im) by a three-multiplication
// complex recurrence from w_1 (2 pi truncated to 6.283185306), and the
// butterflies use the same three-multiplication product.  inverse scales by
// 1/n and conjugates w_1.  With amp_phase the arrays (re, im) are overwritten
// by |F|/(n/2) and a phase (+-90 when |Re| < 1e-6 |Im|, else atan(Im/Re) in
// radians), which the caller may carry into a later transform.
void fft_r2(double* re, double* im, int n, int k, double* fr, double* fi, bool inverse,
            bool amp_phase) {
  for (int it = 0; it < n; ++it) {
    int rev = 0;
    for (int b = 0, m = it; b < k; ++b, m >>= 1) rev = (rev << 1) | (m & 1);
    fr[it] = re[rev];
    fi[it] = im[rev];
  }
  // twiddle table in (re, im)
  re[0] = 1.0;
  im[0] = 0.0;
  const double ang = 6.283185306 / (1.0 * n);
  re[1] = std::cos(ang);
  im[1] = inverse ? std::sin(ang) : -std::sin(ang);
  for (int i = 2; i < n; ++i) {
    const double a = re[i - 1] * re[1];
    const double b = im[i - 1] * im[1];
    const double c = (re[i - 1] + im[i - 1]) * (re[1] + im[1]);
    re[i] = a - b;
    im[i] = c - a - b;
  }
  for (int it = 0; it + 1 < n; it += 2) {  // length-2 stage
    const double xr = fr[it], xi = fi[it];
    fr[it] = xr + fr[it + 1];
    fi[it] = xi + fi[it + 1];
    fr[it + 1] = xr - fr[it + 1];
    fi[it + 1] = xi - fi[it + 1];
  }
  int stride = n / 2;  // twiddle index stride of the current stage
  for (int len = 4; len <= n; len *= 2) {
    stride /= 2;
    const int half = len / 2;
    for (int base = 0; base <= n - len; base += len)
      for (int j = 0; j < half; ++j) {
        const double wr = re[stride * j], wi = im[stride * j];
        double& hr = fr[base + j + half];
        double& hi = fi[base + j + half];
        const double a = wr * hr;
        const double b = wi * hi;
        double c = wr + wi;
        c = c * (hr + hi);
        const double tr = a - b, ti = c - a - b;
        hr = fr[base + j] - tr;
        hi = fi[base + j] - ti;
        fr[base + j] = fr[base + j] + tr;
        fi[base + j] = fi[base + j] + ti;
      }
  }
  if (inverse)
    for (int i = 0; i < n; ++i) {
      fr[i] = fr[i] / (1.0 * n);
      fi[i] = fi[i] / (1.0 * n);
    }
  if (amp_phase)
    for (int i = 0; i < n; ++i) {
      re[i] = std::sqrt(fr[i] * fr[i] + fi[i] * fi[i]);
      re[i] = re[i] / (n / 2);
      if (std::fabs(fr[i]) < 0.000001 * std::fabs(fi[i]))
        im[i] = (fi[i] * fr[i] > 0) ? 90.0 : -90.0;
      else
        im[i] = std::atan(fi[i] / fr[i]);
    }
}

}  // namespace

extern "C" {

// kernel.cu:1133 -> DisToTimeAndTimeToDis1D.cpp:115-173.  Two-way time of
// depth sample j of column i: sum over j' = 1..j of 2*dz/v (FP32 running sum).
int rtmh_d2t(const float* v, const float* d, int nx, int nz, int x_st, int x_end, float dz,
             float dt, float* out, long long cap) {
  if (!v || !d || nz < 1 || x_st < 0 || x_end < x_st || x_end >= nx || !(dt > 0.f))
    return rtmh::set_error(rtmh::RTMH_E_ARG, "rtmh_d2t: bad arguments");
  const int ntr = x_end - x_st + 1;
  std::vector<float> t0((size_t)ntr * nz, 0.0f);
  float tmax = 0.0f;
  for (int i = x_st; i <= x_end; ++i) {
    float t = 0.0f;
    for (int j = 1; j < nz; ++j) {
      t += 2 * dz / v[(size_t)i * nz + j];
      t0[(size_t)(i - x_st) * nz + j] = t;
    }
    if (tmax < t) tmax = t;
  }
  const int nt = (int)(tmax / dt);
  if (nt < 1) return rtmh::set_error(rtmh::RTMH_E_ARG, "rtmh_d2t: empty time axis");
  if (out) {
    if (cap < (long long)ntr * nt) return rtmh::set_error(rtmh::RTMH_E_ARG, "rtmh_d2t: out too small");
    resample_all(out, t0, d, ntr, nt, nz, dt);
  }
  return nt;
}

// kernel.cu:1175 -> DisToTimeAndTimeToDis1D.cpp:31-112.  Depth of time sample
// j of column i: sum over j' = 1..j of dtt * v(depth index) / 2, the depth
// index being (int)((float)(int)depth / dzz) clamped to nzv-1.
int rtmh_t2d(const float* v, const float* d, int nx, int nt, int nzv, int x_st, int x_end,
             float dtt, float dzz, float* out, long long cap) {
  if (!v || !d || nt < 1 || nzv < 1 || x_st < 0 || x_end < x_st || x_end >= nx || !(dzz > 0.f))
    return rtmh::set_error(rtmh::RTMH_E_ARG, "rtmh_t2d: bad arguments");
  const int ntr = x_end - x_st + 1;
  std::vector<float> z0((size_t)ntr * nt, 0.0f);
  float zmax = 0.0f;
  for (int i = x_st; i <= x_end; ++i) {
    float z = 0.0f;
    for (int j = 1; j < nt; ++j) {
      int iz = (int)((float)(int)z / dzz);
      if (iz >= nzv) iz = nzv - 1;
      z += dtt * v[(size_t)i * nzv + iz] / 2;
      z0[(size_t)(i - x_st) * nt + j] = z;
    }
    if (zmax < z) zmax = z;
  }
  const int nzo = (int)(zmax / dzz);
  if (nzo < 1) return rtmh::set_error(rtmh::RTMH_E_ARG, "rtmh_t2d: empty depth axis");
  if (out) {
    if (cap < (long long)ntr * nzo) return rtmh::set_error(rtmh::RTMH_E_ARG, "rtmh_t2d: out too small");
    resample_all(out, z0, d, ntr, nzo, nt, dzz);
  }
  return nzo;
}

// kernel.cu:1163 -> phase_correction_ricker_decon.cpp:153-218.  Each trace is
// zero-padded to n = 2^k >= nt, transformed, bins 1..n/2-1 are rotated by
// angle degrees, bins n/2+1..n-1 are set to their conjugates, and the real
// part of the inverse transform is kept.  Bins 0 and n/2 enter the inverse as
// the amplitude/phase values the forward transform left in place.
void rtmh_phase_rotate(const float* din, float* dout, int ntr, int nt, float angle) {
  const double theta = angle * rtmh::kRefPi / 180;
  int n = 2, k = 1;
  while (n < nt) {
    n *= 2;
    ++k;
  }
  const double cs = std::cos(theta), sn = std::sin(theta);
  const int nh = n / 2;
  std::vector<double> re(n), im(n), fr(n), fi(n);
  for (int tr = 0; tr < ntr; ++tr) {
    for (int i = 0; i < n; ++i) {
      re[i] = i < nt ? (double)din[(size_t)tr * nt + i] : 0.0;
      im[i] = 0.0;
    }
    fft_r2(re.data(), im.data(), n, k, fr.data(), fi.data(), false, true);
    for (int i = 1; i < nh; ++i) {
      re[i] = fr[i] * cs - fi[i] * sn;
      im[i] = fi[i] * cs + fr[i] * sn;
    }
    for (int i = n - 1; i > nh; --i) {
      re[i] = re[n - i];
      im[i] = -im[n - i];
    }
    fft_r2(re.data(), im.data(), n, k, fr.data(), fi.data(), true, false);
    for (int i = 0; i < nt; ++i) dout[(size_t)tr * nt + i] = (float)fr[i];
  }
}

}  // extern "C"
