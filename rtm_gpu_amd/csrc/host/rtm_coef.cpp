// rtm_coef.cpp -- finite-difference coefficient tables for the RTM step.
//
// Adaptive least-squares coefficients (LSMOrCon_rec_2D.cpp:22-524): for each
// velocity bin the operator half-length M is the smallest one whose
// dispersion error stays within eps over the (frequency, angle) grid, found
// by the reference's bisection over bins; the coefficients solve normal
// equations built by 2-D composite Gauss-Legendre quadrature.  Taylor
// coefficients (LSMOrCon_rec_2D.cpp:526-551) for iLSTE = 1.
//
// The floating-point operation sequence is the reference's (the same
// quadrature nodes, accumulation order, elimination with row scaling and
// pivoting), so the float tables agree with the reference bit for bit.  The
// per-bin coefficient solves are independent and run on a thread pool.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "../../../include/rtm_host.h"
#include "rtm_host_internal.h"

namespace {

using rtmh::kRefPi;

// 5-point Gauss-Legendre rule with the reference's 10-digit constants
// (LSMOrCon_rec_2D.cpp:143-146).
const double kNode[5] = {-0.9061798459, -0.5384693101, 0.0, 0.5384693101, 0.9061798459};
const double kWeight[5] = {0.2369268851, 0.4786286705, 0.5688888889, 0.4786286705, 0.2369268851};

class LsDesign {
 public:
  LsDesign(double tao, double h, double df, double eps, int nthita, double hzx, int nfre)
      : tao_(tao), h_(h), df_(df), eps_(eps), nthita_(nthita), hzx_(hzx), nfre_(nfre) {}

  // Solve the normal equations for M coefficients at Courant number r with
  // kh integration limit bb; c[0..M], c0 = -2 sum c_l.
  void solve(double r, double bb, int M, double* c) const {
    std::vector<double> A((size_t)M * M), rhs(M), x(M);
    for (int i = 0; i < M; ++i) {
      for (int j = i; j < M; ++j) A[(size_t)i * M + j] = quad(i + 1, j + 1, bb, r, true);
      rhs[i] = quad(i + 1, 0, bb, r, false);
    }
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < i; ++j) A[(size_t)i * M + j] = A[(size_t)j * M + i];
    eliminate(A, rhs, x, M);
    double s = 0.0;
    for (int l = 1; l <= M; ++l) {
      c[l] = x[l - 1];
      s += c[l];
    }
    c[0] = -2.0 * s;
  }

  // Smallest length >= lmin (up to lmax) meeting the dispersion bound at vel;
  // lmax when none does (LSMOrCon_rec_2D.cpp:317-343).
  int length(double vel, int lmin, int lmax) const {
    std::vector<double> c(lmax + 1, 0.0);
    const double r = vel * tao_ / h_;
    const double bb = 2 * kRefPi * df_ * (nfre_ - 1) * tao_ / r;
    int len = 0;
    for (int M = lmin; M <= lmax; ++M) {
      solve(r, bb, M, c.data());
      len = M;
      if (passes(c.data(), M, vel)) break;
    }
    return len;
  }

 private:
  double integrand(double th, double kh, int ai, int aj, double ar, bool product) const {
    const double hz2 = hzx_ * hzx_;
    const double ct = cos(th), st = sin(th);
    const double denom = pow(ar, -2) * (1 - cos(ar * kh));
    const double fi = ((1 + 1 / hz2) - cos(ai * kh * ct) - cos(hzx_ * ai * kh * st) / hz2) / denom;
    if (!product) return fi;
    const double fj = ((1 + 1 / hz2) - cos(aj * kh * ct) - cos(hzx_ * aj * kh * st) / hz2) / denom;
    return fi * fj;
  }
  // Composite rule over theta in [0, 2 pi] x kh in [0, bb], 4 x 4 panels.
  double quad(int ai, int aj, double bb, double ar, bool product) const {
    const int panels = 4;
    const double half_t = 0.5 * (2 * kRefPi - 0.0) / panels;
    double mid_t = half_t + 0.0, outer = 0.0;
    for (int pt = 0; pt < panels; ++pt) {
      if (pt) mid_t = mid_t + half_t * 2.0;
      for (int nt = 0; nt < 5; ++nt) {
        const double th = half_t * kNode[nt] + mid_t;
        const double half_k = 0.5 * (bb - 0.0) / panels;
        double mid_k = half_k + 0.0, inner = 0.0;
        for (int pk = 0; pk < panels; ++pk) {
          if (pk) mid_k = mid_k + half_k * 2.0;
          for (int nk = 0; nk < 5; ++nk) {
            const double kh = half_k * kNode[nk] + mid_k;
            inner = 1.0 * 1.0 * integrand(th, kh, ai, aj, ar, product) * kWeight[nk] + inner;
          }
        }
        outer = inner * half_k * 1.0 * kWeight[nt] + outer;
      }
    }
    return outer * half_t;
  }
  // Row-scaled Gaussian elimination with partial pivoting and back substitution.
  static void eliminate(std::vector<double>& A, std::vector<double>& b, std::vector<double>& x,
                        int n) {
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    for (int i = 0; i < n; ++i) {
      double big = a(i, 0);
      for (int j = 0; j < n; ++j)
        if (std::fabs(a(i, j)) > std::fabs(big)) big = a(i, j);
      if (std::fabs(big) < 1e-10) return;  // singular row: the reference gives up here
      for (int j = 0; j < n; ++j) a(i, j) = a(i, j) / big;
      b[i] = b[i] / big;
    }
    for (int i = 0; i < n - 1; ++i) {
      int p = i;
      for (int j = i; j < n; ++j)
        if (std::fabs(a(j, i)) > std::fabs(a(p, i))) p = j;
      if (p != i) {
        std::swap(b[i], b[p]);
        for (int j = i; j < n; ++j) std::swap(a(i, j), a(p, j));
      }
      for (int r = i + 1; r < n; ++r) {
        const double m = a(r, i) / a(i, i);
        b[r] = b[r] - m * b[i];
        for (int j = i; j < n; ++j) a(r, j) = a(r, j) - m * a(i, j);
      }
    }
    x[n - 1] = b[n - 1] / a(n - 1, n - 1);
    for (int i = n - 2; i >= 0; --i) {
      double s = 0.0;
      for (int j = i + 1; j < n; ++j) s = s + a(i, j) * x[j];
      x[i] = (b[i] - s) / a(i, i);
    }
  }
  // Dispersion error within eps for every frequency sample k = 1..nfre-1 and
  // every angle n < nthita (the angle n = nthita is not required, as in the
  // reference's break test).
  bool passes(const double* c, int M, double vel) const {
    const double hz2 = hzx_ * hzx_;
    const double dk = 2.0 * kRefPi * df_ * h_ / vel;
    const double r = vel * tao_ / h_, ra = h_ / vel;
    for (int k = 1; k < nfre_; ++k) {
      const double hk = dk * k;
      const double rat = 2 / (r * hk);
      int n = 0;
      for (; n <= nthita_; ++n) {
        const double th = n * kRefPi / (4 * nthita_);
        double mid = 0;
        for (int l = 1; l <= M; ++l) {
          const double sz = sin(l * hzx_ * hk * sin(th) / 2);
          const double sx = sin(l * hk * cos(th) / 2);
          mid = mid + c[l] * (sz * sz / hz2 + sx * sx);
        }
        double e = rat * asin(sqrt(r * r * mid));
        e = std::fabs(ra * (1.0 / e - 1.0));
        if (e > eps_) break;
      }
      if (n < nthita_) return false;
    }
    return true;
  }

  double tao_, h_, df_, eps_;
  int nthita_;
  double hzx_;
  int nfre_;
};

// Operator length of every bin by the reference's interval bisection
// (LSMOrCon_rec_2D.cpp:347-524): lengths of the two end bins first, then
// ranges of equal length are filled between probes, each probe starting its
// search at the length found for the faster neighbour.
void bin_lengths(const LsDesign& D, int nvel, double vmin, double dv, int nfdmin, int nfdmax,
                 int* len) {
  auto probe = [&](int bin1, int lmin) { return D.length(vmin + (bin1 - 1) * dv, lmin, nfdmax); };
  auto fill = [&](int from, int to, int value) {
    for (int i = from; i <= to; ++i) len[i - 1] = value;
  };
  const int first = 1, last = nvel;
  const int l_first = probe(first, nfdmin);
  len[first - 1] = l_first;
  int hi = last, l_hi = probe(last, nfdmin);
  len[last - 1] = l_hi;
  if (l_first == l_hi) {
    fill(first + 1, last, l_first);
    return;
  }
  int step = (last - first) / (2 * (l_first - l_hi));
  if (step < 1) step = 1;
  int lo = 0, l_lo = 0, mark = 0, l_mark = 0;
  bool marked = false;
  enum Next { kStep, kNarrow, kSettle, kDone } next = kStep;
  while (next != kDone) {
    switch (next) {
      case kStep:  // move the lower probe one step below the settled range
        lo = hi - step;
        if (lo < first) {
          lo = first;
          l_lo = l_first;
          if (lo == hi) { next = kDone; break; }
        } else {
          l_lo = probe(lo, l_hi);
        }
        next = kNarrow;
        break;
      case kNarrow:  // bisect (lo, hi) until the length boundary is isolated
        if (l_lo == l_hi) {
          fill(lo, hi - 1, l_lo);
          hi = lo;
          next = kSettle;
        } else if (lo == hi - 1) {
          len[lo - 1] = l_lo;
          hi = lo;
          l_hi = l_lo;
          next = kSettle;
        } else {
          const int mid = (lo + hi) / 2;
          if (!marked) { marked = true; mark = lo; l_mark = l_lo; }
          const int l_mid = probe(mid, l_hi);
          if (l_mid == l_hi) {
            fill(mid, hi - 1, l_hi);
            hi = mid;
          } else {
            lo = mid;
            l_lo = l_mid;
          }
        }
        break;
      case kSettle:
        if (marked) {
          marked = false;
          if (l_mark == l_hi) {
            fill(mark, hi - 1, l_hi);
            hi = mark;
            next = kStep;
          } else {
            lo = mark;
            l_lo = l_mark;
            next = kNarrow;
          }
        } else if (l_hi == l_first) {
          fill(first, hi - 1, l_hi);
          next = kDone;
        } else {
          next = kStep;
        }
        break;
      case kDone:
        break;
    }
  }
}

}  // namespace

extern "C" {

int rtmh_design_ls(int nthita, int nfdmax, int nfdmin, int nvel, double tao, double h, double df,
                   double eps, double fmax, double vmin, double dv, const int* need, double hzx,
                   int* M, int* index, float* c, int c_cap, int nthreads) {
  if (nvel < 1 || nfdmin < 1 || nfdmin > nfdmax || !need || !M || !index || !c)
    return rtmh::set_error(rtmh::RTMH_E_ARG, "design_ls: bad arguments");
  const int nfre = (int)(fmax / df) + 1;
  LsDesign D(tao, h, df, eps, nthita, hzx, nfre);
  bin_lengths(D, nvel, vmin, dv, nfdmin, nfdmax, M);
  for (int i = 0; i < nvel; ++i)
    if (need[i] == 0) M[i] = -1;
  index[0] = 0;
  for (int i = 1; i <= nvel; ++i) index[i] = index[i - 1] + M[i - 1] + 1;
  const int nc = index[nvel];
  if (nc > c_cap) return rtmh::set_error(rtmh::RTMH_E_ARG, "design_ls: %d coefficients > cap %d", nc, c_cap);
  const double rat = tao / h;
  const double ra = 2.0 * kRefPi * fmax * tao;
  std::vector<int> todo;
  for (int i = 0; i < nvel; ++i)
    if (need[i] == 1) todo.push_back(i);
  std::atomic<size_t> next(0);
  auto work = [&]() {
    std::vector<double> cd(nfdmax + 1);
    for (size_t t = next++; t < todo.size(); t = next++) {
      const int i = todo[t];
      const double r = (vmin + i) * rat;  // bin offset i, not i*dv (LSMOrCon_rec_2D.cpp:58)
      D.solve(r, ra / r, M[i], cd.data());
      for (int l = 0; l <= M[i]; ++l) c[index[i] + l] = (float)cd[l];
    }
  };
  int nt = std::max(1, std::min(nthreads, (int)todo.size()));
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
  return nc;
}

void rtmh_taylor(int nfdmax, float* c) {  // LSMOrCon_rec_2D.cpp:526-551, N = 2*nfdmax
  const int half = nfdmax;
  c[0] = 0.0f;
  for (int i = 1; i <= half; ++i) {
    float prod = 1.0f;
    for (int j = 1; j <= half; ++j) {
      if (j == i) continue;
      const double jj = (double)j * j, ii = (double)i * i;
      prod = (float)(prod * std::fabs(jj / (jj - ii)));
    }
    c[i] = (float)(pow(-1, i + 1) / ((double)i * i) * prod);
    c[0] = c[0] - 2 * c[i];
  }
}

// Stability of the time stepping (not in the reference).  A Fourier mode
// exp(i(kx x + kz z)) turns the stencil sum w1 of a cell into
// -(S(kx) + rho S(kz)) times the mode, S(k) = -(c0 + 2 sum_l c_l cos(l k)), so
// the leapfrog update u2 = 2u1 - u0 + P w1 has amplification factors g with
// g + 1/g = 2 - P (S(kx) + rho S(kz)); |g| = 1 for every mode exactly when
// P (S(kx) + rho S(kz)) <= 4, i.e. P (1 + rho) max_k S(k) / 4 <= 1 (frozen
// coefficients, cell by cell).  S is sampled on 4096 + 1 points of [0, pi].
double rtmh_stability(const float* v, long long nxz, const float* c, const int* index, int nvel,
                      float vmin, float dv, int nfdmax, float tao2, float h2, float hzx2_1) {
  constexpr int NK = 4096;
  auto smax_of = [&](const float* row, int taps) {
    double best = 0.0;
    for (int j = 0; j <= NK; ++j) {
      const double k = M_PI * j / NK;
      double sum = row[0];
      for (int l = 1; l <= taps; ++l) sum += 2.0 * row[l] * std::cos(l * k);
      best = std::max(best, -sum);
    }
    return best;
  };
  std::vector<double> smax(index ? (size_t)nvel : 1, -1.0);
  double worst = 0.0;
  for (long long i = 0; i < nxz; ++i) {
    size_t b = 0;
    if (index) {  // the cell's bin, as the kernels select it (kernel.cu:55-56)
      int kb = (int)((double)((v[i] - vmin) / dv) + 0.5);
      b = (size_t)std::min(std::max(kb, 0), nvel - 1);
    }
    if (smax[b] < 0.0)
      smax[b] = index ? smax_of(c + index[b], index[b + 1] - index[b] - 1) : smax_of(c, nfdmax);
    const double p = (double)v[i] * v[i] * tao2 * h2;
    worst = std::max(worst, p * (1.0 + hzx2_1) * smax[b] / 4.0);
  }
  return worst;
}

}  // extern "C"
