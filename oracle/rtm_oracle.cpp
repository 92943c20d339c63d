// rtm_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).
//
// A CPU restatement of the reference RTM path of caixh90/RTM_GPU, written for
// this project.  It follows the reference's kernels thread-index-space by
// thread-index-space so that each loop can be audited against the cited
// file:line.  Not linked by the product; see rtm_oracle.h for who may load it.
//
// Arithmetic conventions (the "FP32 replica", prec = 32):
//   * float storage and float arithmetic, each C operation rounded, evaluated
//     left to right exactly as the reference source is written;
//   * double promotions where the reference writes double literals:
//     (1.0+hzx2_1)*c0*u (kernel.cu:58,91,260,297,360,399); the 2.0*u1 final
//     combine of Add_Con, BKAdd_EFF, BKAdd_EFF_Con (kernel.cu:106,275,312);
//     the bin index (int)((v-vmin)/dv+0.5) (kernel.cu:55-56);
//   * fused multiply-add at exactly three places, as a contracting GPU
//     compiler emits them: s = fma(u[z-l]+u[z+l], hzx2_1, u[x-l]) per tap,
//     w1 = fma(c_l, s, w1), and the FP32 final combine fma(vv, w1, 2u1-u0);
//     plus the imaging accumulations I += a*b (kernel.cu:497-498,511-514).
//   The FP64 oracle (prec = 64) runs the same code with double storage.
//
// Parity pins (tests/): setup integers and LS design tables against the
// compiled reference helpers (oracle/_ref); Taylor closed form; the wave
// engine against closed forms and invariants (linearity, mirror symmetry,
// time-reversal, homogeneous-medium travel time); the image post and stack
// against closed forms (Laplacian of a quadratic, mirror edges, whitening
// constant, mean of constant images); the reference's quirks (mis-indexed ABC
// velocities, the backward wavelet gate, data replacement only for non-zero
// samples, the k+1 data lag, the backward initial state) by known answers on
// tiny grids (tests/test_oracle_quirks.py).
#include "rtm_oracle.h"

#include <pthread.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

const double kPi = 3.1415926535898;  // kernel.cu:16, LSMOrCon_rec_2D.cpp:4

// ---------------------------------------------------------------------------
// Row-parallel helper: a persistent pthread pool with two barriers per pass.
struct Pool {
  int n = 1;
  std::vector<pthread_t> th;
  pthread_barrier_t bar;
  void (*fn)(void*, int, int) = nullptr;
  void* arg = nullptr;
  int total = 0;
  bool quit = false;
  struct Slot { Pool* p; int id; };
  std::vector<Slot> slots;

  static void* worker(void* s) {
    Slot* sl = (Slot*)s;
    Pool* p = sl->p;
    for (;;) {
      pthread_barrier_wait(&p->bar);
      if (p->quit) break;
      p->chunk(sl->id);
      pthread_barrier_wait(&p->bar);
    }
    return nullptr;
  }
  void chunk(int id) {
    int per = (total + n - 1) / n;
    int b = id * per, e = b + per;
    if (e > total) e = total;
    if (b < e) fn(arg, b, e);
  }
  explicit Pool(int nthreads) : n(nthreads < 1 ? 1 : nthreads) {
    if (n > 1) {
      pthread_barrier_init(&bar, nullptr, (unsigned)n);
      slots.resize(n);
      th.resize(n);
      for (int i = 1; i < n; ++i) {
        slots[i] = {this, i};
        pthread_create(&th[i], nullptr, worker, &slots[i]);
      }
    }
  }
  ~Pool() {
    if (n > 1) {
      quit = true;
      pthread_barrier_wait(&bar);
      for (int i = 1; i < n; ++i) pthread_join(th[i], nullptr);
      pthread_barrier_destroy(&bar);
    }
  }
  void run(int count, void (*f)(void*, int, int), void* a) {
    if (n == 1 || count < 2 * n) { f(a, 0, count); return; }
    fn = f; arg = a; total = count;
    pthread_barrier_wait(&bar);
    chunk(0);
    pthread_barrier_wait(&bar);
  }
};

template <typename R>
inline R fmaR(R a, R b, R c) { return std::fma(a, b, c); }

// ---------------------------------------------------------------------------
template <typename R>
struct Engine {
  orc_cfg g;
  int NZ, NX, N2, NFD, NT, NT2, mNZ, mNX;
  size_t NXZ, SS;
  int nslot;
  std::vector<float> v, r1, w, c, wav;
  std::vector<int> index;
  int nvel = 0;
  std::vector<R> F[3];  // forward field, reused as the reconstructed source field
  int f0 = 0, f1 = 1, f2 = 2;
  std::vector<R> Q[3];  // receiver field
  int q0 = 0, q1 = 1, q2 = 2;
  std::vector<R> B;     // one-way scratch (DFWb)
  std::vector<R> SU, SQ, I1, I2;
  // linearised (Born) modelling and its adjoint (not in the reference)
  std::vector<R> D[3];   // scattered field
  int d0 = 0, d1 = 1, d2 = 2;
  std::vector<float> refl;  // reflectivity m, full grid (used on the interior)
  std::vector<R> G;         // sum_k D2u(k+1) * mu(k+2)
  mutable std::vector<R> Gimg;  // G / P (refreshed by field 9 reads)
  bool born = false;
  std::vector<R> strips;
  std::vector<int> snap_slot;  // per step k: snapshot slot or -1
  std::vector<R> snaps;
  std::vector<float> rec, seis;
  std::vector<R> dseis;
  int kf = 2, kb = -1;
  bool fwd_done = false;
  Pool pool;

  explicit Engine(const orc_cfg& cfg) : g(cfg), pool(cfg.nthreads) {
    NZ = g.nz; NX = g.nx; N2 = g.n2; NFD = g.nfdmax; NT = g.nt; NT2 = g.nt2;
    mNZ = NZ - 2 * N2; mNX = NX - 2 * N2;
    NXZ = (size_t)NZ * NX;
    SS = (size_t)2 * NFD * (mNZ + mNX);
    nslot = (g.strip_slots > 0 && g.strip_slots < NT) ? g.strip_slots : NT;
    for (int i = 0; i < 3; ++i) { F[i].assign(NXZ, R(0)); Q[i].assign(NXZ, R(0)); }
    B.assign(NXZ, R(0));
    SU.assign(NXZ, R(0)); SQ.assign(NXZ, R(0)); I1.assign(NXZ, R(0)); I2.assign(NXZ, R(0));
    strips.assign(SS * nslot, R(0));
    snap_slot.assign(NT, -1);
    int ns = 0;
    if (g.snap_every > 0)
      for (int j = g.snap_every; j + 1 <= NT - 3; j += g.snap_every) {
        if (j < 2) continue;  // u(0), u(1) come from forward_begin, not a step
        snap_slot[j] = ns++;
        snap_slot[j + 1] = ns++;
      }
    snaps.assign(NXZ * ns, R(0));
    if (g.record) rec.assign((size_t)g.n * NT, 0.f);
    dseis.assign(g.n > 0 ? g.n : 1, R(0));
  }

  // ---- two-way update: kernel.cu:46-80 (Add), 82-114 (Add_Con),
  //      246-283 (BKAdd_EFF), 286-320 (BKAdd_EFF_Con), 339-418 (BKAdd*)
  inline int bin_of(float vv) const {
    return (int)((double)((vv - g.vmin) / g.dv) + 0.5);  // kernel.cu:55-56
  }
  // Stencil sum w1 at (x, y) of field U1 (x = row z, y = column).
  inline R w1_at(const R* U1, int x, int y) const {
    int top, taps;
    if (g.ilste == 0) {
      int kb_ = bin_of(v[(size_t)x * NX + y]);
      top = index[kb_];
      taps = index[kb_ + 1] - top - 1;  // loop l < M_M (kernel.cu:57-59)
    } else {
      top = 0;
      taps = NFD;  // loop l <= nfdmax (kernel.cu:92)
    }
    const float* cc = &c[top];
    R u = U1[(size_t)x * NX + y];
    R w1;
    if (std::is_same<R, float>::value)
      w1 = (R)((1.0 + (double)g.hzx2_1) * (double)cc[0] * (double)u);
    else
      w1 = (R)((1.0 + (double)g.hzx2_1) * (double)cc[0]) * u;
    const R hz = (R)g.hzx2_1;
    for (int l = 1; l <= taps; ++l) {
      int x1 = x - l, x2 = x + l, y1 = y - l, y2 = y + l;
      if (x1 < 0) x1 = -x1;                 // kernel.cu:65-68
      if (x2 >= NZ) x2 = 2 * NZ - 2 - x2;
      if (y1 < 0) y1 = -y1;
      if (y2 >= NX) y2 = 2 * NX - 2 - y2;
      R a = U1[(size_t)x1 * NX + y] + U1[(size_t)x2 * NX + y];
      R s = fmaR<R>(a, hz, U1[(size_t)x * NX + y1]);
      s = s + U1[(size_t)x * NX + y2];
      w1 = fmaR<R>((R)cc[l], s, w1);
    }
    return w1;
  }
  // Tap l of the coefficient row of cell (x, y) (row-major, x = row z), 0 past
  // the row's operator length; and the longest operator of the table.
  inline float tap_of(int x, int y, int l) const {
    int kb_ = bin_of(v[(size_t)x * NX + y]);
    int M = index[kb_ + 1] - index[kb_] - 1;
    return l <= M ? c[index[kb_] + l] : 0.0f;
  }
  int lmax = 0;  // longest operator of the table (set with the model)
  void set_lmax() {
    lmax = 0;
    for (int k = 0; k < nvel && !index.empty(); ++k) lmax = std::max(lmax, index[k + 1] - index[k] - 1);
  }
  // Transposed stencil sum at interior cell (x, y) for per-cell (LS) rows: the
  // spatial operator (L u)_i = (1+rho) c0(i) u_i + sum_l c_l(i) S_l(u)_i, with
  // S_l(u)_i = rho (u_{z-l} + u_{z+l}) + u_{x-l} + u_{x+l}, has the transpose
  // (L^T u)_i = (1+rho) c0(i) u_i + sum_l [rho (c_l(z-l) u_{z-l} + c_l(z+l) u_{z+l})
  //             + c_l(x-l) u_{x-l} + c_l(x+l) u_{x+l}]:
  // each tap takes the coefficient of the neighbour's own row.  Taps run to the
  // longest operator of the table (zero past a row's length).  Interior cells
  // only (the taps stay in the grid since nfdmax <= n2).
  inline R w1T_at(const R* U1, int x, int y) const {
    int kb_ = bin_of(v[(size_t)x * NX + y]);
    const float c0 = c[index[kb_]];
    R u = U1[(size_t)x * NX + y];
    R w1;
    if (std::is_same<R, float>::value)
      w1 = (R)((1.0 + (double)g.hzx2_1) * (double)c0 * (double)u);
    else
      w1 = (R)((1.0 + (double)g.hzx2_1) * (double)c0) * u;
    const R hz = (R)g.hzx2_1;
    for (int l = 1; l <= lmax; ++l) {
      const R cu = (R)tap_of(x - l, y, l), cd = (R)tap_of(x + l, y, l);
      const R cl = (R)tap_of(x, y - l, l), cr = (R)tap_of(x, y + l, l);
      R vs = fmaR<R>(cu, U1[(size_t)(x - l) * NX + y], cd * U1[(size_t)(x + l) * NX + y]);
      R s = fmaR<R>(vs, hz, cl * U1[(size_t)x * NX + y - l]);
      s = fmaR<R>(cr, U1[(size_t)x * NX + y + l], s);
      w1 = w1 + s;
    }
    return w1;
  }
  // Final combine.  f64comb selects the double form "2.0*u1-u0+P".
  inline R combine(R u1, R u0, float vv, R w1, bool f64comb) const {
    if (std::is_same<R, float>::value) {
      float p = vv * vv * g.tao2 * g.h2;
      if (f64comb)
        return (R)((2.0 * (double)u1 - (double)u0) + (double)(p * (float)w1));
      return fmaR<R>((R)p, w1, (R)2 * u1 - u0);
    } else {
      R p = (R)vv * (R)vv * (R)g.tao2 * (R)g.h2;
      return fmaR<R>(p, w1, (R)2 * u1 - u0);
    }
  }

  // Rows [b, e) of the full-grid forward update (Add / Add_Con).
  struct AddArgs { Engine* e; R* U2; const R* U1; const R* U0; R wav; };
  static void add_rows(void* a_, int b, int e) {
    AddArgs* a = (AddArgs*)a_;
    Engine& E = *a->e;
    bool f64 = (E.g.ilste != 0);  // Add: FP32 combine (kernel.cu:73); Add_Con: 2.0* (kernel.cu:106)
    for (int x = b; x < e; ++x)
      for (int y = 0; y < E.NX; ++y) {
        size_t p = (size_t)x * E.NX + y;
        R w1 = E.w1_at(a->U1, x, y);
        R o = E.combine(a->U1[p], a->U0[p], E.v[p], w1, f64);
        if (x == E.g.r_u && y == E.g.r_x) o = o + a->wav;  // kernel.cu:74-77
        a->U2[p] = o;
      }
  }

  // ---- hybrid ABC: kernel.cu:116-158 (Hybrid1), 160-183 (Hybrid2)
  inline R oneway(float vp, float vq, R U1p, R U0p, R U2i, R U1i, R U0i,
                  R U2ip, R U2im, R U0pp, R U0pm) const {
    R tv = (R)g.taoh * (R)vp;
    R t1 = tv * ((U2i - U0i) + U0p);
    R t2 = (((((R)-2) * U1p + U0p) + U2i) - (R)2 * U1i) + U0i;
    R lap = ((((U2ip - (R)2 * U2i) + U2im) + U0pp) - (R)2 * U0p) + U0pm;
    R t3 = ((R)g.taoh2 * (R)vq) * (R)vq * lap;
    return ((R)1 / (tv + (R)1)) * ((t1 - t2) + t3);
  }
  inline R corner(size_t p, size_t n1, size_t n2, const R* U2, const R* U1) const {
    R rr = (R)r1[p];
    return ((R)1 / ((R)2 * rr + (R)1)) * (U1[p] + rr * (U2[n1] + U2[n2]));
  }
  void hybrid1(R* Bb, const R* U2, const R* U1, const R* U0) const {
    const int W = NX;
    auto P = [W](int z, int x) { return (size_t)z * W + x; };
    for (int x = 1; x <= N2; ++x) {  // top (kernel.cu:122-125)
      int a = N2 - x;
      for (int y = N2 - x + 2; y < NX - N2 - 2 + x; ++y) {
        size_t p = P(a, y), i = P(a + 1, y);
        Bb[p] = oneway(v[p], v[p], U1[p], U0[p], U2[i], U1[i], U0[i], U2[i + 1], U2[i - 1],
                       U0[p + 1], U0[p - 1]);
      }
    }
    for (int y = 1; y <= N2; ++y) {  // left (kernel.cu:126-129), velocity Dv[(N2-y)*NX+x]
      int a = N2 - y;
      for (int x = N2 - y + 2; x < NZ - N2 - 2 + y; ++x) {
        size_t p = P(x, a), i = P(x, a + 1);
        Bb[p] = oneway(v[p], v[(size_t)a * NX + x], U1[p], U0[p], U2[i], U1[i], U0[i],
                       U2[i + W], U2[i - W], U0[p + W], U0[p - W]);
      }
    }
    for (int x = 1; x <= N2; ++x) {  // bottom (kernel.cu:130-133), velocity Dv[(N2-x)*NX+y]
      int bb = NZ - N2 + x - 1;
      for (int y = N2 - x + 2; y < NX - N2 - 2 + x; ++y) {
        size_t p = P(bb, y), i = P(bb - 1, y);
        Bb[p] = oneway(v[p], v[P(N2 - x, y)], U1[p], U0[p], U2[i], U1[i], U0[i], U2[i + 1],
                       U2[i - 1], U0[p + 1], U0[p - 1]);
      }
    }
    for (int y = 1; y <= N2; ++y) {  // right (kernel.cu:134-137), velocity Dv[(N2-y)*NX+x]
      int cc = NX - N2 + y - 1;
      for (int x = N2 - y + 2; x < NZ - N2 - 2 + y; ++x) {
        size_t p = P(x, cc), i = P(x, cc - 1);
        Bb[p] = oneway(v[p], v[(size_t)(N2 - y) * NX + x], U1[p], U0[p], U2[i], U1[i], U0[i],
                       U2[i + W], U2[i - W], U0[p + W], U0[p - W]);
      }
    }
    for (int x = 1; x <= N2; ++x) {  // corners (kernel.cu:138-155)
      int a = N2 - x, b = NZ - 1 - a, cc = NX - 1 - a;
      Bb[P(a + 1, a)] = corner(P(a + 1, a), P(a + 1, a + 1), P(a + 2, a), U2, U1);
      Bb[P(a, a + 1)] = corner(P(a, a + 1), P(a, a + 2), P(a + 1, a + 1), U2, U1);
      Bb[P(a, a)] = corner(P(a, a), P(a, a + 1), P(a + 1, a), U2, U1);
      Bb[P(b - 1, cc)] = corner(P(b - 1, cc), P(b - 1, cc - 1), P(b - 2, cc), U2, U1);
      Bb[P(b, cc - 1)] = corner(P(b, cc - 1), P(b, cc - 2), P(b - 1, cc - 1), U2, U1);
      Bb[P(b, cc)] = corner(P(b, cc), P(b, cc - 1), P(b - 1, cc), U2, U1);
      Bb[P(b, a + 1)] = corner(P(b, a + 1), P(b - 1, a + 1), P(b, a + 2), U2, U1);
      Bb[P(b - 1, a)] = corner(P(b - 1, a), P(b - 2, a), P(b - 1, a + 1), U2, U1);
      Bb[P(b, a)] = corner(P(b, a), P(b - 1, a), P(b, a + 1), U2, U1);
      Bb[P(a + 1, cc)] = corner(P(a + 1, cc), P(a + 1, cc - 1), P(a + 2, cc), U2, U1);
      Bb[P(a, cc - 1)] = corner(P(a, cc - 1), P(a, cc - 2), P(a + 1, cc - 1), U2, U1);
      Bb[P(a, cc)] = corner(P(a, cc), P(a, cc - 1), P(a + 1, cc), U2, U1);
    }
  }
  void hybrid2(R* U2, const R* Bb) const {
    auto blend = [&](size_t p, int i) { U2[p] = ((R)1 - (R)w[i]) * U2[p] + (R)w[i] * Bb[p]; };
    for (int x = 1; x <= N2; ++x)  // top (kernel.cu:166-169)
      for (int y = N2 - x + 1; y < NX - N2 + x; ++y) blend((size_t)(N2 - x) * NX + y, x);
    for (int y = 1; y <= N2; ++y)  // left (kernel.cu:170-173)
      for (int x = N2 - y; x < NZ - N2 + y - 1; ++x) blend((size_t)x * NX + (N2 - y), y);
    for (int x = 1; x <= N2; ++x)  // bottom (kernel.cu:174-177)
      for (int y = N2 - x; y < NX - N2 + x - 1; ++y) blend((size_t)(NZ - N2 + x - 1) * NX + y, x);
    for (int y = 1; y <= N2; ++y)  // right (kernel.cu:178-181)
      for (int x = N2 - y + 1; x < NZ - N2 + y; ++x) blend((size_t)x * NX + NX - N2 + y - 1, y);
  }

  // ---- strips: kernel.cu:18-45 (Equal), 184-208 (Hybrid3), 222-245 (BKEqual)
  // Slot layout: lf[mNZ][NFD], rt[mNZ][NFD], up[NFD][mNX], dw[NFD][mNX].
  void save_strips(int k, const R* U) {
    R* s = &strips[SS * (k % nslot)];
    R* lf = s; R* rt = s + (size_t)mNZ * NFD; R* up = rt + (size_t)mNZ * NFD; R* dw = up + (size_t)NFD * mNX;
    for (int r = 0; r < mNZ; ++r)
      for (int d = 0; d < NFD; ++d) {
        lf[r * NFD + d] = U[(size_t)(N2 + r) * NX + N2 - d - 1];
        rt[r * NFD + d] = U[(size_t)(N2 + r) * NX + d + NX - N2];
      }
    for (int d = 0; d < NFD; ++d)
      for (int cc = 0; cc < mNX; ++cc) {
        up[d * mNX + cc] = U[(size_t)(N2 - d - 1) * NX + cc + N2];
        dw[d * mNX + cc] = U[(size_t)(d + NZ - N2) * NX + cc + N2];
      }
  }
  void restore_strips(int k, R* U) const {
    const R* s = &strips[SS * (k % nslot)];
    const R* lf = s; const R* rt = s + (size_t)mNZ * NFD; const R* up = rt + (size_t)mNZ * NFD;
    const R* dw = up + (size_t)NFD * mNX;
    for (int r = 0; r < mNZ; ++r)
      for (int d = 0; d < NFD; ++d) {
        U[(size_t)(N2 + r) * NX + N2 - d - 1] = lf[r * NFD + d];
        U[(size_t)(N2 + r) * NX + d + NX - N2] = rt[r * NFD + d];
      }
    for (int d = 0; d < NFD; ++d)
      for (int cc = 0; cc < mNX; ++cc) {
        U[(size_t)(N2 - d - 1) * NX + cc + N2] = up[d * mNX + cc];
        U[(size_t)(d + NZ - N2) * NX + cc + N2] = dw[d * mNX + cc];
      }
  }

  void record_step(int k, const R* U) {
    if (!g.record || born) return;
    for (int s = 0; s < g.n; ++s)
      rec[(size_t)s * NT + k] = (float)U[(size_t)g.s_z * NX + g.s_l + s * g.ds];
  }

  // ---- forward: kernel.cu:798-821
  int forward_begin() {
    for (int i = 0; i < 3; ++i) std::fill(F[i].begin(), F[i].end(), R(0));
    f0 = 0; f1 = 1; f2 = 2;
    F[f1][(size_t)g.r_u * NX + g.r_x] = (R)(float)((double)wav[0] / 2.0);  // kernel.cu:803
    save_strips(0, F[f0].data());  // Equal (kernel.cu:808)
    save_strips(1, F[f1].data());
    record_step(0, F[f0].data());
    record_step(1, F[f1].data());
    kf = 2;
    fwd_done = (NT <= 2);
    return kf;
  }
  void forward_step(int k) {
    R wv = (k < NT2) ? (R)wav[k - 1] : R(0);  // kernel.cu:812-813
    AddArgs a{this, F[f2].data(), F[f1].data(), F[f0].data(), wv};
    pool.run(NZ, add_rows, &a);
    hybrid1(B.data(), F[f2].data(), F[f1].data(), F[f0].data());
    hybrid2(F[f2].data(), B.data());
    save_strips(k, F[f2].data());
    record_step(k, F[f2].data());
    if (snap_slot[k] >= 0) std::copy(F[f2].begin(), F[f2].end(), snaps.begin() + NXZ * snap_slot[k]);
    int t = f0; f0 = f1; f1 = f2; f2 = t;  // Deliver (kernel.cu:820) as a rotation
  }
  int forward_advance(int nsteps) {
    for (int i = 0; i < nsteps && kf < NT; ++i, ++kf) forward_step(kf);
    fwd_done = (kf >= NT);
    return kf;
  }

  // ---- backward: kernel.cu:822-931
  int backward_begin(const float* seis_in) {
    seis.assign(seis_in, seis_in + (size_t)g.n * NT);
    // DBW1 = u(NT-2), DBW0 = u(NT-1) (kernel.cu:822-825): keep F as is with
    // the source rotation S0 = F[f1] (u(NT-1)), S1 = F[f0] (u(NT-2)).
    int s0 = f1, s1 = f0, s2 = f2;
    f0 = s0; f1 = s1; f2 = s2;
    for (int i = 0; i < 3; ++i) std::fill(Q[i].begin(), Q[i].end(), R(0));
    q0 = 0; q1 = 1; q2 = 2;
    const R* BW0 = F[f0].data();  // u(NT-1)
    const R* BW1 = F[f1].data();  // u(NT-2)
    size_t ps = (size_t)g.r_u * NX + g.r_x;
    R spike = (R)(float)((double)wav[0] / 2.0);
    for (size_t i = 0; i < NXZ; ++i) {  // kernel.cu:859-876
      R FW1 = (i == ps) ? spike : R(0);
      R FW0 = R(0);
      if (g.icompen == 1) {
        R BW2 = BW0[i] + BW1[i];
        R FW2 = FW0 + FW1;
        SU[i] = BW2;
        SQ[i] = FW2;
        I1[i] = FW2 * BW2 + FW0 * BW0[i];
      } else {
        SU[i] = R(0);
        SQ[i] = R(0);
        I1[i] = FW1 * BW1[i] + FW0 * BW0[i];
      }
      I2[i] = BW1[i] * BW1[i] + BW0[i] * BW0[i];
    }
    kb = NT - 3;
    return kb;
  }

  struct BkArgs { Engine* e; R wav; };
  static void bk_src_rows(void* a_, int b, int e) {  // BKAdd_EFF(_Con), kernel.cu:246-320
    BkArgs* a = (BkArgs*)a_;
    Engine& E = *a->e;
    const R* S1 = E.F[E.f1].data();
    const R* S0 = E.F[E.f0].data();
    R* S2 = E.F[E.f2].data();
    for (int x = b; x < e; ++x) {
      if (x < E.N2 || x >= E.NZ - E.N2) continue;
      for (int y = E.N2; y < E.NX - E.N2; ++y) {
        size_t p = (size_t)x * E.NX + y;
        R w1 = E.w1_at(S1, x, y);
        R o = E.combine(S1[p], S0[p], E.v[p], w1, true);  // 2.0* (kernel.cu:275,312)
        if (x == E.g.r_u && y == E.g.r_x) o = o + a->wav;
        S2[p] = o;
      }
    }
  }
  static void bk_rcv_rows(void* a_, int b, int e) {  // BKAdd(_Con), kernel.cu:339-418
    BkArgs* a = (BkArgs*)a_;
    Engine& E = *a->e;
    const R* Q1 = E.Q[E.q1].data();
    const R* Q0 = E.Q[E.q0].data();
    R* Q2 = E.Q[E.q2].data();
    const orc_cfg& g = E.g;
    for (int x = b; x < e; ++x)
      for (int y = 0; y < E.NX; ++y) {
        size_t p = (size_t)x * E.NX + y;
        int rem = (y - g.s_l) % g.ds;  // receiver spacing test (kernel.cu:349)
        int s = (y - g.s_l) / g.ds;
        if (x == g.s_z && y >= g.s_l && y <= g.s_r && rem == 0 && E.dseis[s] != R(0)) {
          Q2[p] = E.dseis[s];  // replacement (kernel.cu:351-354)
        } else {
          R w1 = E.w1_at(Q1, x, y);
          Q2[p] = E.combine(Q1[p], Q0[p], E.v[p], w1, false);  // FP32 2* (kernel.cu:375,413)
        }
      }
  }
  static void image_rows(void* a_, int b, int e) {  // Rel_Compen / Rel_NonCompen, kernel.cu:489-517
    BkArgs* a = (BkArgs*)a_;
    Engine& E = *a->e;
    const R* u = E.F[E.f1].data();   // DBW2 after Deliver_EFF: u(k)
    const R* q = E.Q[E.q1].data();   // DFW2 after Deliver: new q
    for (int x = b; x < e; ++x) {
      if (x < E.N2 || x >= E.NZ - E.N2) continue;
      for (int y = E.N2; y < E.NX - E.N2; ++y) {
        size_t p = (size_t)x * E.NX + y;
        if (E.g.icompen == 1) {
          E.SU[p] = E.SU[p] + u[p];
          E.SQ[p] = E.SQ[p] + q[p];
          E.I1[p] = fmaR<R>(E.SQ[p], E.SU[p], E.I1[p]);
          E.I2[p] = fmaR<R>(u[p], u[p], E.I2[p]);
        } else {
          E.I1[p] = fmaR<R>(q[p], u[p], E.I1[p]);
          E.I2[p] = fmaR<R>(u[p], u[p], E.I2[p]);
        }
      }
    }
  }
  void backward_step(int k) {
    R wv = (k < NT2) ? (R)wav[k + 1] : R(0);  // kernel.cu:889-890
    restore_strips(k + 1, F[f1].data());       // BKEqual (kernel.cu:892)
    BkArgs a{this, wv};
    pool.run(NZ, bk_src_rows, &a);             // BKAdd_EFF (kernel.cu:896)
    { int t = f0; f0 = f1; f1 = f2; f2 = t; }  // Deliver_EFF (kernel.cu:903)
    if (snap_slot[k] >= 0)  // checkpoint: the reconstructed u(k) is replaced by the saved one
      std::copy(snaps.begin() + NXZ * snap_slot[k], snaps.begin() + NXZ * (snap_slot[k] + 1), F[f1].begin());
    for (int s = 0; s < g.n; ++s) dseis[s] = (R)seis[(size_t)s * NT + k + 1];  // kernel.cu:905-909
    pool.run(NZ, bk_rcv_rows, &a);             // BKAdd (kernel.cu:913)
    hybrid1(B.data(), Q[q2].data(), Q[q1].data(), Q[q0].data());  // BKHybrid1 (kernel.cu:919)
    hybrid2(Q[q2].data(), B.data());                              // BKHybrid2 (kernel.cu:920)
    { int t = q0; q0 = q1; q1 = q2; q2 = t; }  // Deliver (kernel.cu:921)
    pool.run(NZ, image_rows, &a);              // Rel_* (kernel.cu:923-930)
  }
  int backward_advance(int nsteps) {
    for (int i = 0; i < nsteps && kb >= 0; ++i, --kb) backward_step(kb);
    return kb;
  }

  // ---- Born modelling: the scattered field dU follows the background step
  // with the same update and ABC, driven on the interior by m times the
  // background's second time difference:
  //   dU(k) = update(dU(k-1), dU(k-2)) + m * ((u(k) - 2u(k-1)) + u(k-2)).
  // Its receiver record is the Born data.
  int born_begin() {
    for (int i = 0; i < 3; ++i) D[i].assign(NXZ, R(0));
    d0 = 0; d1 = 1; d2 = 2;
    born = false;
    int k = forward_begin();
    born = true;
    if (g.record) std::fill(rec.begin(), rec.end(), 0.f);
    return k;
  }
  struct BornArgs { Engine* e; };
  static void born_rows(void* a_, int b, int e) {
    Engine& E = *((BornArgs*)a_)->e;
    bool f64 = (E.g.ilste != 0);
    const R* U1 = E.D[E.d1].data();
    const R* U0 = E.D[E.d0].data();
    R* U2 = E.D[E.d2].data();
    const R* uk = E.F[E.f1].data();   // after the background rotation: u(k)
    const R* uk1 = E.F[E.f0].data();  // u(k-1)
    const R* uk2 = E.F[E.f2].data();  // u(k-2)
    for (int x = b; x < e; ++x)
      for (int y = 0; y < E.NX; ++y) {
        size_t p = (size_t)x * E.NX + y;
        R w1 = E.w1_at(U1, x, y);
        R o = E.combine(U1[p], U0[p], E.v[p], w1, f64);
        if (x >= E.N2 && x < E.NZ - E.N2 && y >= E.N2 && y < E.NX - E.N2) {
          R d2u = (uk[p] - (R)2 * uk1[p]) + uk2[p];
          o = fmaR<R>((R)E.refl[p], d2u, o);
        }
        U2[p] = o;
      }
  }
  void born_step(int k) {
    forward_step(k);  // background u(k); rotation leaves u(k), u(k-1), u(k-2) in f1, f0, f2
    BornArgs a{this};
    pool.run(NZ, born_rows, &a);
    hybrid1(B.data(), D[d2].data(), D[d1].data(), D[d0].data());
    hybrid2(D[d2].data(), B.data());
    if (g.record)
      for (int s = 0; s < g.n; ++s)
        rec[(size_t)s * NT + k] = (float)D[d2][(size_t)g.s_z * NX + g.s_l + s * g.ds];
    int t = d0; d0 = d1; d1 = d2; d2 = t;
  }
  int born_advance(int nsteps) {
    for (int i = 0; i < nsteps && kf < NT; ++i, ++kf) born_step(kf);
    fwd_done = (kf >= NT);
    return kf;
  }

  // ---- adjoint of Born modelling.  The update is A = 2I + P*L with
  // P = v^2*tao2*h2 per cell, so A^T = 2I + L^T P: the adjoint recursion
  // lambda(k) = A^T lambda(k+1) - lambda(k+2) + R^T r(k) is carried as
  // mu = P*lambda, which takes the update 2I + P L^T and the ABC with P*r(k)
  // added at the receivers.  L^T is the transposed stencil (w1T_at) on the
  // interior for per-cell rows, and L itself for one (symmetric) Taylor row.  The background is reconstructed backward from the strips as in
  // the migration.  G accumulates ((u(k+2) - 2u(k+1)) + u(k)) * mu(k+2) on
  // the interior; the gradient is G / P (adjoint_image).  mu(nt-1) and
  // mu(nt-2) come from two updates before the loop; their image terms pair
  // with mu(nt), mu(nt+1) = 0.
  R pcell(size_t p) const { return (R)(v[p] * v[p] * g.tao2 * g.h2); }
  void adj_inject(int k) {
    for (int s = 0; s < g.n; ++s) {
      size_t p = (size_t)g.s_z * NX + g.s_l + s * g.ds;
      Q[q2][p] = Q[q2][p] + pcell(p) * (R)seis[(size_t)s * NT + k];
    }
  }
  void adj_update(int k) {
    AdjArgs a{this, k};
    pool.run(NZ, adj_rows, &a);
    adj_inject(k);
    hybrid1(B.data(), Q[q2].data(), Q[q1].data(), Q[q0].data());
    hybrid2(Q[q2].data(), B.data());
  }
  int adjoint_begin(const float* resid) {
    born = false;
    backward_begin(resid);  // F roles, strips; Q, images set below
    for (int i = 0; i < 3; ++i) std::fill(Q[i].begin(), Q[i].end(), R(0));
    q0 = 0; q1 = 1; q2 = 2;
    G.assign(NXZ, R(0));
    for (int k = NT - 1; k >= NT - 2; --k) {  // mu(nt-1), mu(nt-2)
      adj_update(k);
      int t = q0; q0 = q1; q1 = q2; q2 = t;
    }
    return kb;
  }
  void adjoint_image(R* out) const {  // G / P on the interior, 0 elsewhere
    for (size_t p = 0; p < NXZ; ++p) out[p] = R(0);
    for (int x = N2; x < NZ - N2; ++x)
      for (int y = N2; y < NX - N2; ++y) {
        size_t p = (size_t)x * NX + y;
        out[p] = G[p] / pcell(p);
      }
  }
  struct AdjArgs { Engine* e; int k; };
  static void adj_rows(void* a_, int b, int e) {
    AdjArgs* a = (AdjArgs*)a_;
    Engine& E = *a->e;
    const R* Q1 = E.Q[E.q1].data();
    const R* Q0 = E.Q[E.q0].data();
    R* Q2 = E.Q[E.q2].data();
    // per-cell LS rows: the interior takes the transposed stencil (exact
    // adjoint of the interior update); one Taylor row is symmetric already.
    // The frame keeps the forward update and ABC.
    const bool tr = E.g.ilste == 0;
    for (int x = b; x < e; ++x)
      for (int y = 0; y < E.NX; ++y) {
        size_t p = (size_t)x * E.NX + y;
        const bool inner = x >= E.N2 && x < E.NZ - E.N2 && y >= E.N2 && y < E.NX - E.N2;
        R w1 = (tr && inner) ? E.w1T_at(Q1, x, y) : E.w1_at(Q1, x, y);
        Q2[p] = E.combine(Q1[p], Q0[p], E.v[p], w1, false);
      }
  }
  static void adj_image_rows(void* a_, int b, int e) {
    Engine& E = *((AdjArgs*)a_)->e;
    const R* u0 = E.F[E.f1].data();  // u(k)
    const R* u1 = E.F[E.f0].data();  // u(k+1)
    const R* u2 = E.F[E.f2].data();  // u(k+2)
    const R* lam = E.Q[E.q0].data(); // lambda(k+2), before the Q rotation
    for (int x = b; x < e; ++x) {
      if (x < E.N2 || x >= E.NZ - E.N2) continue;
      for (int y = E.N2; y < E.NX - E.N2; ++y) {
        size_t p = (size_t)x * E.NX + y;
        R d2u = (u2[p] - (R)2 * u1[p]) + u0[p];
        E.G[p] = fmaR<R>(d2u, lam[p], E.G[p]);
      }
    }
  }
  void adjoint_step(int k) {
    R wv = (k < NT2) ? (R)wav[k + 1] : R(0);  // background reconstruction as in the migration
    restore_strips(k + 1, F[f1].data());
    BkArgs ab{this, wv};
    pool.run(NZ, bk_src_rows, &ab);
    { int t = f0; f0 = f1; f1 = f2; f2 = t; }
    if (snap_slot[k] >= 0)
      std::copy(snaps.begin() + NXZ * snap_slot[k], snaps.begin() + NXZ * (snap_slot[k] + 1), F[f1].begin());
    adj_update(k);
    AdjArgs a{this, k};
    pool.run(NZ, adj_image_rows, &a);
    { int t = q0; q0 = q1; q1 = q2; q2 = t; }
  }
  int adjoint_advance(int nsteps) {
    for (int i = 0; i < nsteps && kb >= 0; ++i, --kb) adjoint_step(kb);
    return kb;
  }

  const R* field_ptr(int which) const {
    if (which == 9) {
      Gimg.resize(NXZ);
      adjoint_image(Gimg.data());
    }
    switch (which) {
      case 7: return D[d1].data();
      case 8: return G.data();
      case 9: return Gimg.data();
      case 0: return F[f1].data();
      case 1: return F[f0].data();
      case 2: return Q[q1].data();
      case 3: return I1.data();
      case 4: return I2.data();
      case 5: return SU.data();
      case 6: return SQ.data();
    }
    return nullptr;
  }
  template <typename T>
  void read_field(int which, T* out) const {
    const R* f = field_ptr(which);
    bool img = which >= 3 && which != 7;  // 7: the scattered wavefield, frame included
    for (int z = 0; z < NZ; ++z)
      for (int x = 0; x < NX; ++x) {
        size_t p = (size_t)z * NX + x;
        bool inside = z >= N2 && z < NZ - N2 && x >= N2 && x < NX - N2;
        out[p] = (img && !inside) ? T(0) : (T)f[p];
      }
  }
};

// ---------------------------------------------------------------------------
// LS coefficient design, restated from LSMOrCon_rec_2D.cpp.

// Gauss-Legendre 5-point nodes and weights with the reference's 10 digits
// (LSMOrCon_rec_2D.cpp:143-146).
const double kGLt[5] = {-0.9061798459, -0.5384693101, 0.0, 0.5384693101, 0.9061798459};
const double kGLw[5] = {0.2369268851, 0.4786286705, 0.5688888889, 0.4786286705, 0.2369268851};

// Integrand of the normal equations, LSMOrCon_rec_2D.cpp:216-233.
inline double ls_integrand(double th, double kh, int ai, int aj, double ar, int al, double hzx) {
  double hzx2 = hzx * hzx;
  double cth = cos(th);
  double sth = sin(th);
  double p = pow(ar, -2);
  double car = cos(ar * kh);
  double f1 = ((1 + 1 / hzx2) - cos(ai * kh * cth) - cos(hzx * ai * kh * sth) / hzx2) / (p * (1 - car));
  if (al == 1) {
    double f2 = ((1 + 1 / hzx2) - cos(aj * kh * cth) - cos(hzx * aj * kh * sth) / hzx2) / (p * (1 - car));
    return f1 * f2;
  }
  return f1;
}

// Composite 2-D Gauss-Legendre over theta in [0, 2pi] and kh in [0, bb],
// 4 x 4 sub-intervals, 5 nodes each, accumulated in the order of
// LSMOrCon_rec_2D.cpp:139-202 (fgaus with n=2, js={4,4}).
double ls_quad(int ai, int aj, double bb, double ar, int al, double hzx) {
  const int nsub = 4;
  double h0 = 0.5 * (2 * kPi - 0.0) / nsub;
  double c0 = h0 + 0.0;
  double acc0 = 0.0;
  for (int s0 = 0; s0 < nsub; ++s0) {
    if (s0 > 0) c0 = c0 + h0 * 2.0;
    for (int k0 = 0; k0 < 5; ++k0) {
      double th = h0 * kGLt[k0] + c0;
      double h1 = 0.5 * (bb - 0.0) / nsub;
      double c1 = h1 + 0.0;
      double acc1 = 0.0;
      for (int s1 = 0; s1 < nsub; ++s1) {
        if (s1 > 0) c1 = c1 + h1 * 2.0;
        for (int k1 = 0; k1 < 5; ++k1) {
          double kh = h1 * kGLt[k1] + c1;
          double p = ls_integrand(th, kh, ai, aj, ar, al, hzx);
          acc1 = 1.0 * 1.0 * p * kGLw[k1] + acc1;
        }
      }
      acc0 = acc1 * h1 * 1.0 * kGLw[k0] + acc0;
    }
  }
  return acc0 * h0;
}

// Row-scaled Gaussian elimination with partial pivoting, LSMOrCon_rec_2D.cpp:74-137.
int gauss_solve(std::vector<std::vector<double>>& A, int n, std::vector<double>& b,
                std::vector<double>& x) {
  for (int i = 0; i < n; ++i) {
    double mx = A[i][0];
    for (int j = 0; j < n; ++j)
      if (fabs(A[i][j]) > fabs(mx)) mx = A[i][j];
    if (fabs(mx) < 1e-10) return 0;
    for (int j = 0; j < n; ++j) A[i][j] = A[i][j] / mx;
    b[i] = b[i] / mx;
  }
  for (int i = 0; i < n - 1; ++i) {
    int k = i;
    for (int j = i; j < n; ++j)
      if (fabs(A[j][i]) > fabs(A[k][i])) k = j;
    if (k != i) {
      std::swap(b[i], b[k]);
      for (int j = i; j < n; ++j) std::swap(A[i][j], A[k][j]);
    }
    for (int p = i + 1; p < n; ++p) {
      double m = A[p][i] / A[i][i];
      b[p] = b[p] - m * b[i];
      for (int j = i; j < n; ++j) A[p][j] = A[p][j] - m * A[i][j];
    }
  }
  x[n - 1] = b[n - 1] / A[n - 1][n - 1];
  for (int i = n - 2; i >= 0; --i) {
    double m = 0.0;
    for (int j = i + 1; j < n; ++j) m = m + A[i][j] * x[j];
    x[i] = (b[i] - m) / A[i][i];
  }
  return 1;
}

// Maximum over the dispersion grid; returns the break state as in
// LSMOrCon_rec_2D.cpp:321-338: true when every frequency passes.
bool disp_ok(const double* c, int M, double vel, double tao, double h, double df, int nfre,
             int nthita, double eps, double hzx) {
  double hzx2 = hzx * hzx;
  double b = 2.0 * kPi * df * h / vel;
  std::vector<double> hk(nfre);
  for (int i = 0; i < nfre; ++i) hk[i] = b * i;
  double r = vel * tao / h;
  double ra = h / vel;
  int k;
  for (k = 1; k < nfre; ++k) {
    double rat = 2 / (r * hk[k]);
    int n;
    for (n = 0; n <= nthita; ++n) {
      double nth = n * kPi / (4 * nthita);
      double mid = 0;
      for (int l = 1; l <= M; ++l)
        mid = mid + c[l] * (pow(sin(l * hzx * hk[k] * sin(nth) / 2), 2) / hzx2 +
                            pow(sin(l * hk[k] * cos(nth) / 2), 2));
      double de = rat * asin(sqrt(r * r * mid));
      de = fabs(ra * (1.0 / de - 1.0));
      if (de > eps) break;
    }
    if (n < nthita) break;
  }
  return k == nfre;
}

}  // namespace

// ===========================================================================
extern "C" {

void orc_derive(int mod_nz, int mod_nx, int n2, int nt1, float tao1, float tao, float h,
                float hz, float f0, int s_l1, int s_z1, int n, int ds, int r_x1,
                int* io, float* fo) {
  int s_l = s_l1 + n2 - 1;  // kernel.cu:607-610
  int r_x = r_x1 + n2 - 1;
  int s_z = s_z1 + n2 - 1;
  int s_r = (n - 1) * ds + s_l;
  int NZ = mod_nz + 2 * n2;  // kernel.cu:611-612
  int NX = mod_nx + 2 * n2;
  int NT = (int)((nt1 - 1) * tao1 / tao + 1.5);  // kernel.cu:613
  float taoh = tao / h;                          // kernel.cu:614-617
  float tao2 = pow(tao, 2);
  float h2 = 1 / pow(h, 2);
  float taoh2 = tao2 * h2 / 2;
  int NT2 = int(2.0 / (f0 * tao)) + 1;  // kernel.cu:624
  float hzx = hz / h;                   // kernel.cu:625-626
  float hzx2_1 = 1 / (hzx * hzx);
  io[0] = NZ; io[1] = NX; io[2] = NT; io[3] = NT2;
  io[4] = s_l; io[5] = s_r; io[6] = s_z; io[7] = r_x;
  fo[0] = taoh; fo[1] = tao2; fo[2] = h2; fo[3] = taoh2; fo[4] = hzx; fo[5] = hzx2_1;
}

int orc_source_row(float depth_m, float hz, int n2) {
  int N = int(depth_m);                              // kernel.cu:794
  return (int)((std::fabs((float)N / hz) + (float)n2) - 1.0f);  // kernel.cu:795
}

float orc_ricker(float t1, float f0) {  // kernel.cu:1261-1266
  float t00 = 1 / f0;
  float y = (1 - 2 * pow(kPi * f0 * (t1 - t00), 2)) * exp(-pow(kPi * f0 * (t1 - t00), 2));
  return y;
}

void orc_pad_model(const float* raw, int mod_nz, int mod_nx, int n2, int ifv, float tao, float h,
                   float* v, float* r1) {
  int NZ = mod_nz + 2 * n2, NX = mod_nx + 2 * n2;
  size_t NXZ = (size_t)NZ * NX;
  size_t idx = 0;
  for (int j = n2; j < NX - n2; ++j)  // column-major read (GPU_velocity_real.cpp:12-18)
    for (int i = n2; i < NZ - n2; ++i) v[(size_t)i * NX + j] = raw[idx++];
  for (int i = 0; i < n2; ++i)  // top, left, bottom, right, corners (:21-83)
    for (int j = n2; j < NX - n2; ++j) v[(size_t)i * NX + j] = v[(size_t)n2 * NX + j];
  for (int j = 0; j < n2; ++j)
    for (int i = n2; i < NZ - n2; ++i) v[(size_t)i * NX + j] = v[(size_t)i * NX + n2];
  for (int i = NZ - n2; i < NZ; ++i)
    for (int j = n2; j < NX - n2; ++j) v[(size_t)i * NX + j] = v[(size_t)(NZ - n2 - 1) * NX + j];
  for (int j = NX - n2; j < NX; ++j)
    for (int i = n2; i < NZ - n2; ++i) v[(size_t)i * NX + j] = v[(size_t)i * NX + (NX - n2 - 1)];
  for (int i = 0; i < n2; ++i)
    for (int j = 0; j < n2; ++j) v[(size_t)i * NX + j] = v[(size_t)n2 * NX + n2];
  for (int i = 0; i < n2; ++i)
    for (int j = NX - n2; j < NX; ++j) v[(size_t)i * NX + j] = v[(size_t)n2 * NX + (NX - n2 - 1)];
  for (int i = NZ - n2; i < NZ; ++i)
    for (int j = 0; j < n2; ++j) v[(size_t)i * NX + j] = v[(size_t)(NZ - n2 - 1) * NX + n2];
  for (int i = NZ - n2; i < NZ; ++i)
    for (int j = NX - n2; j < NX; ++j)
      v[(size_t)i * NX + j] = v[(size_t)(NZ - n2 - 1) * NX + (NX - n2 - 1)];
  if (ifv == 1) {  // :84-100
    std::vector<float> t(NXZ);
    for (int i = 0; i < NZ; ++i)
      for (int j = 0; j < NX; ++j) t[(size_t)i * NX + j] = v[(size_t)i * NX + NX - 1 - j];
    memcpy(v, t.data(), NXZ * sizeof(float));
  }
  for (size_t i = 0; i < NXZ; ++i) {  // :106-117
    float r = v[i] * tao / h;
    float r2 = pow(r, 2) / 2;
    r1[i] = sqrt(r2);
  }
}

void orc_vrange(const float* v, int nxz, float dv, float* vmin_o, float* vmax_o, int* nvel_o) {
  float vmin = v[0], vmax = v[0];  // kernel.cu:704-714
  for (int i = 0; i < nxz; ++i) {
    if (v[i] < vmin) vmin = v[i];
    if (v[i] > vmax) vmax = v[i];
  }
  float vel = (int(vmin / dv)) * dv;  // :715-720
  if (vel > vmin) vmin = vel - dv; else vmin = vel;
  vel = (int(vmax / dv)) * dv;
  if (vel < vmax) vmax = vel + dv; else vmax = vel;
  *vmin_o = vmin; *vmax_o = vmax;
  *nvel_o = int((vmax - vmin) / dv + 1.5);  // :721
}

void orc_bins_needed(const float* v, int nxz, float vmin, float dv, int nvel, int* need) {
  for (int i = 0; i < nvel; ++i) need[i] = 0;  // :727-738
  for (int i = 0; i < nxz; ++i) {
    int k = int((v[i] - vmin) / dv + 0.5);
    if (k >= 0 && k < nvel) need[k] = 1;
  }
}

void orc_blend_weights(int n2, float* w) {
  for (int i = 0; i <= n2; ++i) w[i] = (1.0 * i) / (1.0 * n2);  // kernel.cu:687-691
}

void orc_lscoef(double r, double bb, int M, double hzx, double* c) {  // :236-287
  std::vector<std::vector<double>> A(M, std::vector<double>(M));
  std::vector<double> b(M), x(M);
  for (int i = 0; i < M; ++i) {
    for (int j = i; j < M; ++j) A[i][j] = ls_quad(i + 1, j + 1, bb, r, 1, hzx);
    b[i] = ls_quad(i + 1, M + 1, bb, r, 2, hzx);
  }
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < i; ++j) A[i][j] = A[j][i];
  gauss_solve(A, M, b, x);
  for (int i = 1; i <= M; ++i) c[i] = x[i - 1];
  double s = 0.0;
  for (int i = 1; i <= M; ++i) s += c[i];
  c[0] = -2.0 * s;
}

int orc_lenfd(int nfre, int nfdmin, int nfdmax, double vel, double tao, double h, double df,
              int nthita, double eps, double hzx) {  // :293-346
  std::vector<double> c(nfdmax + 1, 0.0);
  double r = vel * tao / h;
  double b = 2 * kPi * df * (nfre - 1) * tao / r;
  int len = 0;
  for (int j = nfdmin; j <= nfdmax; ++j) {
    orc_lscoef(r, b, j, hzx, c.data());
    bool ok = disp_ok(c.data(), j, vel, tao, h, df, nfre, nthita, eps, hzx);
    len = j;
    if (ok) break;
  }
  return len;
}

double orc_disp_maxerr(const double* c, int M, double vel, double tao, double h, double df,
                       int nfre, int nthita, double hzx) {
  double hzx2 = hzx * hzx;
  double b = 2.0 * kPi * df * h / vel;
  double r = vel * tao / h, ra = h / vel, mx = 0.0;
  for (int k = 1; k < nfre; ++k) {
    double hk = b * k, rat = 2 / (r * hk);
    for (int n = 0; n < nthita; ++n) {
      double nth = n * kPi / (4 * nthita);
      double mid = 0;
      for (int l = 1; l <= M; ++l)
        mid = mid + c[l] * (pow(sin(l * hzx * hk * sin(nth) / 2), 2) / hzx2 +
                            pow(sin(l * hk * cos(nth) / 2), 2));
      double de = fabs(ra * (1.0 / (rat * asin(sqrt(r * r * mid))) - 1.0));
      if (de > mx) mx = de;
    }
  }
  return mx;
}

// Operator-length search over all velocity bins, LSMOrCon_rec_2D.cpp:347-524
// (the bisection of calfdlen_ls restated as an explicit state machine).
static void search_lengths(int nfre, int nfdmax, int nfdmin, double tao, double h, double df,
                           int nthita, double eps, int* len, int nvel, double vmin, double dv,
                           double hzx) {
  auto L = [&](int i, int lmin) {
    double vel = vmin + (i - 1) * dv;
    return orc_lenfd(nfre, lmin, nfdmax, vel, tao, h, df, nthita, eps, hzx);
  };
  const int ibeg = 1, iend = nvel;
  int i1 = ibeg, i2, i3 = 0, i4, l1, l2, l3 = 0, l4, lbeg, lend, iinc, iflag = 0;
  l1 = lbeg = L(i1, nfdmin);
  len[i1 - 1] = l1;
  i2 = iend;
  l2 = lend = L(i2, nfdmin);
  len[i2 - 1] = l2;
  if (l1 == lend) {
    for (int i = ibeg + 1; i <= iend; ++i) len[i - 1] = l1;
    return;
  }
  iinc = (iend - ibeg) / (2 * (l1 - lend));
  if (iinc < 1) iinc = 1;
  enum { S100, S200, SPOST, SDONE } st = S100;
  while (st != SDONE) {
    if (st == S100) {
      i1 = i2 - iinc;
      if (i1 < ibeg) {
        i1 = ibeg;
        l1 = lbeg;
        if (i1 == i2) { st = SDONE; continue; }
      } else {
        l1 = L(i1, l2);
      }
      st = S200;
    } else if (st == S200) {
      if (l1 == l2) {
        for (int i = i1; i <= i2 - 1; ++i) len[i - 1] = l1;
        i2 = i1; l2 = l1;
        st = SPOST;
      } else if (i1 == i2 - 1) {
        len[i1 - 1] = l1;
        i2 = i1; l2 = l1;
        st = SPOST;
      } else {
        i4 = (i1 + i2) / 2;
        if (iflag == 0) { iflag = 1; i3 = i1; l3 = l1; }
        l4 = L(i4, l2);
        if (l4 == l2) {
          for (int i = i4; i <= i2 - 1; ++i) len[i - 1] = l2;
          i2 = i4; l2 = l4;
        } else {
          i1 = i4; l1 = l4;
        }
        st = S200;
      }
    } else {  // SPOST
      if (iflag == 1) {
        iflag = 0;
        if (l3 == l2) {
          for (int i = i3; i <= i2 - 1; ++i) len[i - 1] = l2;
          i2 = i3; l2 = l3;
          st = S100;
        } else {
          i1 = i3; l1 = l3;
          st = S200;
        }
      } else if (l2 == lbeg) {
        for (int i = ibeg; i <= i2 - 1; ++i) len[i - 1] = l2;
        st = SDONE;
      } else {
        st = S100;
      }
    }
  }
}

int orc_design_ls(int nthita, int nfdmax, int nfdmin, int nvel, double tao, double h, double df,
                  double eps, double fmax, double vmin, double dv, const int* need, double hzx,
                  int* M, int* Index, float* c, int c_cap) {  // LSMOrCon_rec_2D.cpp:22-73
  int nfre = (int)(fmax / df) + 1;
  search_lengths(nfre, nfdmax, nfdmin, tao, h, df, nthita, eps, M, nvel, vmin, dv, hzx);
  for (int i = 0; i < nvel; ++i)
    if (need[i] == 0) M[i] = -1;
  Index[0] = 0;
  for (int i = 1; i <= nvel; ++i) Index[i] = Index[i - 1] + M[i - 1] + 1;
  int NC = Index[nvel];
  if (NC > c_cap) return -1;
  double rat = tao / h;
  double ra = 2.0 * kPi * fmax * tao;
  std::vector<double> cd(nfdmax + 1);
  for (int i = 0; i < nvel; ++i) {
    if (need[i] != 1) continue;
    double r = (vmin + i) * rat;  // uses i, not i*dv (LSMOrCon_rec_2D.cpp:58)
    double b = ra / r;
    orc_lscoef(r, b, M[i], hzx, cd.data());
    for (int l = 0; l <= M[i]; ++l) c[l + Index[i]] = (float)cd[l];
  }
  return NC;
}

void orc_taylor(int nfdmax, float* c) {  // LSMOrCon_rec_2D.cpp:526-551 with N = 2*nfdmax
  int N = 2 * nfdmax;
  std::vector<float> x(N / 2 + 1, 1.0f);
  c[0] = 0.0;
  for (int i = 1; i <= N / 2; ++i) {
    for (int j = 1; j <= N / 2; ++j)
      if (j != i) x[i] = x[i] * fabs(pow(j, 2) / (pow(j, 2) - pow(i, 2)));
    c[i] = pow(-1, i + 1) / pow(i, 2) * x[i];
    c[0] = (c[0] - 2 * c[i]);
  }
}

// ---- engine C API -----------------------------------------------------------
struct orc_engine {
  int prec;
  Engine<float>* e32;
  Engine<double>* e64;
};

orc_engine* orc_create(const orc_cfg* cfg, int prec) {
  orc_engine* o = new orc_engine{prec, nullptr, nullptr};
  if (prec == 64) o->e64 = new Engine<double>(*cfg);
  else o->e32 = new Engine<float>(*cfg);
  return o;
}
void orc_destroy(orc_engine* o) {
  if (!o) return;
  delete o->e32;
  delete o->e64;
  delete o;
}

#define ORC_DISPATCH(o, call) ((o)->e64 ? (o)->e64->call : (o)->e32->call)

extern "C++" {
template <typename R>
static void set_model_t(Engine<R>* e, const float* v, const float* r1, const float* w,
                        const float* c, int nc, const int* index, int nvel, const float* wav) {
  size_t nxz = e->NXZ;
  e->v.assign(v, v + nxz);
  e->r1.assign(r1, r1 + nxz);
  e->w.assign(w, w + e->N2 + 1);
  e->c.assign(c, c + nc);
  e->nvel = nvel;
  if (index) e->index.assign(index, index + nvel + 1);
  e->set_lmax();
  e->wav.assign(wav, wav + e->NT + 1);
}
}  // extern "C++"
void orc_set_model(orc_engine* o, const float* v, const float* r1, const float* w,
                   const float* c, int nc, const int* index, int nvel, const float* wav) {
  if (o->e64) set_model_t(o->e64, v, r1, w, c, nc, index, nvel, wav);
  else set_model_t(o->e32, v, r1, w, c, nc, index, nvel, wav);
}
int orc_forward_begin(orc_engine* o) { return ORC_DISPATCH(o, forward_begin()); }
int orc_forward_advance(orc_engine* o, int n) { return ORC_DISPATCH(o, forward_advance(n)); }
int orc_backward_begin(orc_engine* o, const float* seis) { return ORC_DISPATCH(o, backward_begin(seis)); }
int orc_backward_advance(orc_engine* o, int n) { return ORC_DISPATCH(o, backward_advance(n)); }
void orc_set_reflectivity(orc_engine* o, const float* m) {
  if (o->e64) o->e64->refl.assign(m, m + o->e64->NXZ);
  else o->e32->refl.assign(m, m + o->e32->NXZ);
}
int orc_born_begin(orc_engine* o) { return ORC_DISPATCH(o, born_begin()); }
int orc_born_advance(orc_engine* o, int n) { return ORC_DISPATCH(o, born_advance(n)); }
int orc_adjoint_begin(orc_engine* o, const float* resid) { return ORC_DISPATCH(o, adjoint_begin(resid)); }
int orc_adjoint_advance(orc_engine* o, int n) { return ORC_DISPATCH(o, adjoint_advance(n)); }
void orc_read_field(const orc_engine* o, int which, float* out) {
  if (o->e64) o->e64->read_field<float>(which, out);
  else o->e32->read_field<float>(which, out);
}
void orc_read_field64(const orc_engine* o, int which, double* out) {
  if (o->e64) o->e64->read_field<double>(which, out);
  else o->e32->read_field<double>(which, out);
}
void orc_read_record(const orc_engine* o, float* out) {
  const std::vector<float>& r = o->e64 ? o->e64->rec : o->e32->rec;
  memcpy(out, r.data(), r.size() * sizeof(float));
}
void orc_read_strip(const orc_engine* o, int k, float* out) {
  if (o->e64) {
    for (size_t i = 0; i < o->e64->SS; ++i) out[i] = (float)o->e64->strips[o->e64->SS * (k % o->e64->nslot) + i];
  } else {
    memcpy(out, &o->e32->strips[o->e32->SS * (k % o->e32->nslot)], o->e32->SS * sizeof(float));
  }
}

void orc_image_post(const float* I1, const float* I2, const float* v, int NZ, int NX, int N2,
                    float vmax, float whitecoe, float* up, float* down) {
  int mNZ = NZ - 2 * N2, mNX = NX - 2 * N2;
  for (int j = N2; j < NX - N2; ++j)  // kernel.cu:935-949
    for (int i = N2; i < NZ - N2; ++i) {
      int i1 = i - 1, i2 = i + 1, j1 = j - 1, j2 = j + 1;
      if (i1 < N2) i1 = 2 * N2 - i1;
      if (j1 < N2) j1 = 2 * N2 - j1;
      if (i2 >= NZ - N2) i2 = 2 * (NZ - N2 - 1) - i2;
      if (j2 >= NX - N2) j2 = 2 * (NX - N2 - 1) - j2;
      const float* M = I1;
      up[(size_t)(i - N2) * mNX + (j - N2)] =
          -1.0 * (M[(size_t)i * NX + j2] + M[(size_t)i * NX + j1] + M[(size_t)i2 * NX + j] +
                  M[(size_t)i1 * NX + j] - 4 * M[(size_t)i * NX + j]) *
          v[(size_t)i * NX + j] * v[(size_t)i * NX + j] / (vmax * vmax);
    }
  float mx = 0.0f;  // kernel.cu:963-979
  for (int j = N2; j < NX - N2; ++j)
    for (int i = N2; i < NZ - N2; ++i)
      if (std::fabs(I2[(size_t)i * NX + j]) > mx) mx = std::fabs(I2[(size_t)i * NX + j]);
  float stable = mx * whitecoe;
  for (int j = N2; j < NX - N2; ++j)
    for (int i = N2; i < NZ - N2; ++i)
      down[(size_t)(i - N2) * mNX + (j - N2)] = I2[(size_t)i * NX + j] + stable;
  (void)mNZ;
}

void orc_stack(const float* ups, const float* downs, int nshot, int nI, int inorm, float* img,
               float* illum) {  // kernel.cu:1000-1059
  std::vector<float> s1(nI, 0.0f), s2(nI, 0.0f);
  for (int m = 0; m < nshot; ++m)
    for (int i = 0; i < nI; ++i) {
      s1[i] += ups[(size_t)m * nI + i];
      s2[i] += downs[(size_t)m * nI + i];
    }
  for (int i = 0; i < nI; ++i) {
    img[i] = s1[i] / nshot;
    illum[i] = s2[i] / nshot;
    if (inorm == 1) img[i] = img[i] / illum[i];
  }
}

// ---- CPU baseline ------------------------------------------------------------
struct TimeJob {
  orc_cfg cfg;
  const float *v, *r1, *w, *c, *wav, *seis;
  const int* index;
  int nc, nvel, nsteps_f, nsteps_b;
  int row_threads;  // threads of the engine's row pool (1: one shot per thread)
};
static void* time_worker(void* p) {
  TimeJob* j = (TimeJob*)p;
  orc_cfg cfg = j->cfg;
  cfg.nthreads = j->row_threads;
  Engine<float> e(cfg);
  set_model_t(&e, j->v, j->r1, j->w, j->c, j->nc, j->index, j->nvel, j->wav);
  e.forward_begin();
  e.forward_advance(j->nsteps_f);
  if (j->nsteps_b > 0) {
    // The sample runs the backward kernels on the fields reached so far.
    e.kf = e.NT;
    e.backward_begin(j->seis);
    e.backward_advance(j->nsteps_b);
  }
  return nullptr;
}

double orc_time_shots(const orc_cfg* cfg, const float* v, const float* r1, const float* w,
                      const float* c, int nc, const int* index, int nvel, const float* wav,
                      const float* seis, int nshot, int nthreads, int nsteps_f, int nsteps_b) {
  std::vector<TimeJob> jobs(nshot, TimeJob{*cfg, v, r1, w, c, wav, seis, index, nc, nvel,
                                           nsteps_f, nsteps_b, 1});
  auto t0 = std::chrono::steady_clock::now();
  int done = 0;
  while (done < nshot) {
    int batch = std::min(nthreads, nshot - done);
    std::vector<pthread_t> th(batch);
    for (int i = 0; i < batch; ++i) pthread_create(&th[i], nullptr, time_worker, &jobs[done + i]);
    for (int i = 0; i < batch; ++i) pthread_join(th[i], nullptr);
    done += batch;
  }
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

double orc_time_rows(const orc_cfg* cfg, const float* v, const float* r1, const float* w,
                     const float* c, int nc, const int* index, int nvel, const float* wav,
                     const float* seis, int nthreads, int nsteps_f, int nsteps_b) {
  TimeJob job{*cfg, v, r1, w, c, wav, seis, index, nc, nvel, nsteps_f, nsteps_b, nthreads};
  auto t0 = std::chrono::steady_clock::now();
  time_worker(&job);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
