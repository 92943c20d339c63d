// rtm_kernels.hip -- hand-written gfx950 kernels of the 2-D acoustic RTM step.
//
// Reference kernels (caixh90/RTM_GPU kernel.cu) and where their work lives here:
//   Add/Add_Con (46-114)            -> k_interior<EPI_FWD> + k_frame<FR_FWD> (two-way part)
//   Hybrid1/Hybrid2 (116-183)       -> k_frame<FR_FWD> (one-way + blend, from pre-blend values)
//   Equal/Hybrid3 (18-45, 184-208)  -> k_strip_save, and k_frame<FR_FWD> for k >= 2
//   Deliver/Deliver_EFF (210-221, 323-337) -> none: the host rotates buffer pointers
//   BKEqual (222-245)               -> k_strip_restore (into the buffer that becomes u(k+1))
//   BKAdd_EFF(_Con) (246-320)       -> k_interior<EPI_SRC>
//   BKAdd(_Con) (339-418)           -> k_interior<EPI_RCV> + k_frame<FR_RCV>
//   BKHybrid1/2 (421-487)           -> k_frame<FR_RCV>
//   Rel_Compen/Rel_NonCompen (489-517) -> k_interior<EPI_RCV> epilogue (interior only)
//
// Arithmetic follows the reference formulas operation by operation (build
// flags: -ffp-contract=off, IEEE division, denormals kept), with explicit
// fused multiply-adds only at the per-tap stencil sum, the tap accumulation,
// the FP32 final combine and the imaging accumulations.
#include <algorithm>

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "rtm_kernels.h"

namespace rtm {

// ---------------------------------------------------------------------------
// helpers
template <int I>
__device__ __forceinline__ float comp(const float4& v) {
  static_assert(I >= 0 && I < 4, "component");
  return I == 0 ? v.x : (I == 1 ? v.y : (I == 2 ? v.z : v.w));
}
__device__ __forceinline__ float compr(const float4& v, int i) {
  return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Raw buffer access.  Word 3 of a gfx9-family resource: 32-bit data format,
// no swizzle, stride 0 (range check on the byte offset against num_records).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int kRsrcWord3 = 0x00020000;
constexpr unsigned kDrop = 0x80000000u;  // column offset that is out of range of every buffer
__device__ __forceinline__ Rsrc rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, kRsrcWord3);
}
__device__ __forceinline__ float4 bld4(Rsrc r, unsigned vo, unsigned so) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)vo, (int)so, 0);
  return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                     __uint_as_float(v.w));
}
__device__ __forceinline__ uint4 bld4u(Rsrc r, unsigned vo, unsigned so) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)vo, (int)so, 0);
  return make_uint4(v.x, v.y, v.z, v.w);
}
// Stores take the whole offset in the VGPR operand and a zero soffset.  A
// store of more than 64 bits reads its data VGPRs late, so a VALU write to them
// right after the store needs one wait state; the compiler inserts it only for
// stores whose soffset is not a register.  (With the row offset in an SGPR
// soffset the q store of the receiver step lost its first dword in some lanes
// when two blocks shared a CU.)  vo | kDrop stays >= 2^31 after adding so < 2^30.
__device__ __forceinline__ void bst4(Rsrc r, unsigned vo, unsigned so, float a, float b, float c,
                                     float d) {
  const u32x4 v = {__float_as_uint(a), __float_as_uint(b), __float_as_uint(c), __float_as_uint(d)};
  __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)(vo + so), 0, 0);
}
// Streaming variants for data touched once per time step (u(k+1), u0, the
// image and illumination sums): the non-temporal policy keeps them from
// displacing what the shots of a launch share in the caches (the model's P
// and coefficient offsets, the coefficient table, the u1 halo rows).
#ifndef RTM_STREAM_POLICY
#define RTM_STREAM_POLICY 2  // cache-policy bits of the buffer instructions: nt
#endif
__device__ __forceinline__ float4 bld4s(Rsrc r, unsigned vo, unsigned so) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)vo, (int)so, RTM_STREAM_POLICY);
  return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                     __uint_as_float(v.w));
}
__device__ __forceinline__ void bst4s(Rsrc r, unsigned vo, unsigned so, float a, float b, float c,
                                      float d) {
  const u32x4 v = {__float_as_uint(a), __float_as_uint(b), __float_as_uint(c), __float_as_uint(d)};
  __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)(vo + so), 0, RTM_STREAM_POLICY);
}
__device__ __forceinline__ void bst1s_(Rsrc r, unsigned vo, unsigned so, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)(vo + so), 0, RTM_STREAM_POLICY);
}

template <int NFD>
struct CoefRow {
  static constexpr int CS = ((2 + NFD) + 3) & ~3;
  static constexpr int NR = CS / 4;
  float4 r[NR];
  __device__ __forceinline__ void load(const float* ctab, unsigned byte_off) {
    const float4* p = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(ctab) + byte_off);
#pragma unroll
    for (int i = 0; i < NR; ++i) r[i] = p[i];
  }
  __device__ __forceinline__ void load(Rsrc t, unsigned byte_off) {
#pragma unroll
    for (int i = 0; i < NR; ++i) r[i] = bld4(t, byte_off + 16u * i, 0);
  }
  __device__ __forceinline__ double c0d() const {
    return __hiloint2double(__float_as_int(r[0].y), __float_as_int(r[0].x));
  }
  // tap coefficient c_l, l >= 1 (float index 1 + l in the row)
  __device__ __forceinline__ float c(int l) const { return compr(r[(1 + l) >> 2], (1 + l) & 3); }
};

// Final combine of the two-way update.
//  FP32 form (Add, BKAdd, BKAdd_Con: kernel.cu:73,375,413): 2*u1-u0+P.
//  FP64 form (Add_Con, BKAdd_EFF, BKAdd_EFF_Con: kernel.cu:106,275,312): 2.0*u1-u0+P.
// p = v*v*tao2*h2 (evaluated left to right in FP32).
__device__ __forceinline__ float combine_p(float u, float u0, float p, float w1, bool f64) {
  if (f64) return (float)((2.0 * (double)u - (double)u0) + (double)(p * w1));
  return __builtin_fmaf(p, w1, 2.0f * u - u0);
}
__device__ __forceinline__ bool is_receiver(int c, const Rcv& r, int* s) {
  if (c < r.s_l || c > r.s_r) return false;
  int d = c - r.s_l;
  if (d % r.ds != 0) return false;
  *s = d / r.ds;
  return true;
}

// Row-major walk of a tile with `cols` columns by a 256-thread block: thread
// t visits cells t, t+256, ... keeping (row, column) incrementally, so the
// loop carries no integer division.
struct TileWalk {
  int i, j, di, dj, cols;
  __device__ __forceinline__ explicit TileWalk(int c) : cols(c) {
    i = (int)threadIdx.x / c;
    j = (int)threadIdx.x - i * c;
    di = 256 / c;
    dj = 256 - di * c;
  }
  __device__ __forceinline__ void next() {
    i += di;
    j += dj;
    if (j >= cols) { j -= cols; ++i; }
  }
};

__device__ __forceinline__ int mirror_idx(int g, int n) {  // kernel.cu:65-68
  if (g < 0) g = -g;
  if (g >= n) g = 2 * n - 2 - g;
  g = g < 0 ? 0 : g;
  g = g >= n ? n - 1 : g;
  return g;
}

// ---- strip encoding -------------------------------------------------------
// Power-of-two scale of a slot from its max |value|: 2^(e-15) with
// max = f*2^e, so |value/scale| < 2^15 and FP16 cannot overflow.  The scale is
// kept >= 2^-126 so that 1/scale stays finite when the maximum is subnormal
// (slots holding only the tiny precursors of the first steps).
__device__ __forceinline__ float pow2_scale(float amax) {
  float sc = 1.0f;
  if (amax > 0.0f) {
    int e;
    frexpf(amax, &e);
    sc = ldexpf(1.0f, max(e - 15, -126));
  }
  return sc;
}

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

// Strip slot layout: lf[mnz][nfd], rt[mnz][nfd], up[nfd][mnx], dw[nfd][mnx]
// (Dlf/Drt/Dup/Ddw of kernel.cu:199-205); a slot holds ss_of(mnx) elements.
__device__ __forceinline__ long long strip_size(const Grid& g, int mnx) {
  return 2LL * g.nfd * (g.mnz + mnx);
}
// Strip element index -> grid cell of a shot of width nx (interior mnx), in
// 32-bit arithmetic (a slot holds fewer than 2^31 elements).
__device__ __forceinline__ void strip_cell(const Grid& g, int nx, int mnx, unsigned e, int* z,
                                           int* x) {
  const unsigned nfd = (unsigned)g.nfd, a = (unsigned)g.mnz * nfd, b = nfd * (unsigned)mnx;
  if (e < 2 * a) {
    const unsigned side = e >= a ? 1u : 0u, ee = e - side * a;
    const unsigned r = ee / nfd, d = ee - r * nfd;
    *z = g.n2 + (int)r;
    *x = side ? nx - g.n2 + (int)d : g.n2 - 1 - (int)d;
  } else {
    const unsigned bot = e >= 2 * a + b ? 1u : 0u, ee = e - 2 * a - bot * b;
    const unsigned d = ee / (unsigned)mnx, c = ee - d * (unsigned)mnx;
    *z = bot ? g.nz - g.n2 + (int)d : g.n2 - 1 - (int)d;
    *x = g.n2 + (int)c;
  }
}
// Grid cell -> strip element index, or -1.
__device__ __forceinline__ long long strip_index(const Grid& g, int nx, int mnx, int z, int x) {
  const long long a = (long long)g.mnz * g.nfd, b = (long long)g.nfd * mnx;
  const int n2 = g.n2, nfd = g.nfd;
  if (z >= n2 && z < g.nz - n2) {
    if (x >= n2 - nfd && x < n2) return (long long)(z - n2) * nfd + (n2 - 1 - x);
    if (x >= nx - n2 && x < nx - n2 + nfd) return a + (long long)(z - n2) * nfd + (x - (nx - n2));
    return -1;
  }
  if (x >= n2 && x < nx - n2) {
    if (z >= n2 - nfd && z < n2) return 2 * a + (long long)(n2 - 1 - z) * mnx + (x - n2);
    if (z >= g.nz - n2 && z < g.nz - n2 + nfd) return 2 * a + b + (long long)(z - (g.nz - n2)) * mnx + (x - n2);
  }
  return -1;
}
__device__ __forceinline__ void strip_put(const Strips& st, int shot, int k, long long e, float val) {
  if (st.fmt == 0) {
    float* p = (float*)st.base;
    p[((long long)shot * st.nslot + (k % st.nslot)) * st.ss + e] = val;
  } else {
    st.stage[((long long)shot * 2 + (k & 1)) * st.ss + e] = val;
  }
}
__device__ __forceinline__ float strip_get(const Strips& st, int shot, int k, long long e) {
  long long slot = (long long)shot * st.nslot + (k % st.nslot);
  if (st.fmt == 0) return ((const float*)st.base)[slot * st.ss + e];
  const unsigned short* p = (const unsigned short*)st.base;
  unsigned short h = p[slot * st.ss + e];
  float sc = st.scale[slot];
  if (st.fmt == 1) return __half2float(__ushort_as_half(h)) * sc;
  return bf2f(h) * sc;
}
// Strips of slot k into the rings of field f: block b of a launch with bps
// 256-thread blocks per shot restores part of shot b / bps.
__device__ __forceinline__ void strip_restore_body(const Grid& g, float* f, const Strips& st,
                                                   const ShotGeo* geo, int k, unsigned b,
                                                   unsigned bps) {
  const int shot = (int)(b / bps);
  const unsigned blk = b - (unsigned)shot * bps;
  const ShotGeo sg = geo[shot];
  const unsigned ss = (unsigned)strip_size(g, sg.mnx);
  float* F = f + (long long)shot * g.fstride + g.x0;
  for (unsigned e = blk * 256u + threadIdx.x; e < ss; e += bps * 256u) {
    int z, x;
    strip_cell(g, sg.nx, sg.mnx, e, &z, &x);
    F[(long)z * g.pitch + x] = strip_get(st, shot, k, (long long)e);
  }
}

// ---------------------------------------------------------------------------
// Interior update: cells [n2, nz-n2) x [n2, nx-n2).  One wave owns 256
// columns (4 per lane, float4 rows) and marches down a segment of rows.
// Vertical taps come from a register queue of rows; horizontal taps from a
// per-wave LDS row with a halo of H columns each side.  No mirror is needed:
// every tap stays inside the grid because NFD <= n2.
//
// Memory goes through buffer instructions: the per-shot base and size live in
// a scalar resource, the row offset in a scalar register, the column offset
// in a per-lane register that stays fixed along the march.  A lane that must
// not store gets a column offset at or above 2^31, past every buffer's size,
// and the hardware drops the access; the row body therefore has no branch.
//
// Software pipeline per row: (a) the coefficient rows of the row are gathered
// (their byte offsets were fetched PD rows earlier); (b) every stencil operand
// of row z+PD is fetched (u1 row z+NFD+PD, the halo row, u0, the p and offset
// rows), and for the receiver step the imaging operands of row z+1; (c) the
// LDS exchange and the stencil read only data fetched in earlier iterations
// or in (a).  The memory counter is in-order, so (a) goes first: the wait on
// the gathers then leaves the prefetches in flight.  Scheduling barriers keep
// that order.  A segment is a multiple of the queue length QL, so the row
// loop is fully unrolled without a remainder; rows past the interior at the
// end of the last segment are computed from clamped rows and not stored.
template <int EPI>
struct EpiTraits {
  static constexpr bool f64 =
      EPI == EPI_FWD64 || EPI == EPI_SRC || EPI == EPI_SRCI || EPI == EPI_BORN64;  // "2.0*"
  static constexpr bool rcv = EPI == EPI_RCV || EPI == EPI_RCVI;  // receiver step + imaging
  static constexpr bool rcvi = EPI == EPI_RCVI;  // ... without the source-side terms
  static constexpr bool srci = EPI == EPI_SRCI;  // reconstruction + source-side terms
  static constexpr bool born = EPI == EPI_BORN || EPI == EPI_BORN64;
  static constexpr bool adj = EPI == EPI_ADJ || EPI == EPI_ADJT;
  static constexpr bool adjt = EPI == EPI_ADJT;  // transposed stencil (per-cell rows)
  static constexpr bool src = !rcv && !born && !adj;  // injects the source wavelet
  static constexpr int PD = (rcv || born || adj) ? 1 : 2;  // prefetch distance of the stencil operands
};

// Interior body for block bid (k_interior: bid = blockIdx.x; k_step gives
// the interior blocks of a launch shared with the frame their own numbering).
template <int NFD, int EPI>
__device__ __forceinline__ void interior_body(const IntArgs& a, const unsigned bid) {
  using T = EpiTraits<EPI>;
  constexpr int H = (NFD + 3) & ~3;
  constexpr int PD = T::PD;
  constexpr int QL = 2 * NFD + 1 + PD;
  constexpr int W4 = 64 + 2 * (H / 4);
  constexpr int NWIN = 1 + 2 * (H / 4);
  __shared__ float4 lds[4][W4];

  const Grid& g = a.g;
  const Model& m = a.m;
  const int lane = threadIdx.x & 63;
  // wave index: uniform in the wave, so row offsets derived from it stay scalar
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // The four waves of block b march the same rows and columns (tile
  // b / nsg) for four shots (group b % nsg; launches of at least four
  // shots), so their reads of the shared
  // model rows (p, coefficient offsets, coefficient rows) coincide in time and
  // hit the CU's L1.  Workgroups reach the eight XCDs round-robin by b, so
  // the blocks one XCD holds at a time are more shots of the same few tiles,
  // whose model reads hit that XCD's L2.
  // With fewer than four shots the waves of a block take four consecutive
  // row segments of one shot instead (block b: shot b % B, tile b / B).
  int shot, zseg;
  int tile;
  if (a.nshot >= 4) {
    const unsigned nsg = ((unsigned)a.nshot + 3u) >> 2;
    shot = (int)(bid % nsg) * 4 + wv;
    tile = (int)(bid / nsg);
    if (shot >= a.nshot) return;
    zseg = tile / a.nxt;
  } else {
    shot = (int)(bid % (unsigned)a.nshot);
    tile = (int)(bid / (unsigned)a.nshot);
    zseg = (tile / a.nxt) * 4 + wv;
  }
  const int xt = tile % a.nxt;
  const ShotGeo sg = a.geo[shot];
  const int zhi = g.nz - g.n2;
  const int zs = g.n2 + zseg * a.zs;  // a.zs is a multiple of QL
  const int cend = sg.nx - g.n2;                // first column past the shot's interior
  const int cb = g.n2 + xt * 256;
  if (zs >= zhi || cb >= cend) return;          // tiles past a narrower shot's width
  const int ze = zs + a.zs;
  const int c0 = cb + lane * 4;
  const int clast = sg.nx - ((sg.nx - g.n2) & 3);  // largest c == n2 (mod 4) with c <= nx
  const int cl = min(c0, clast);
  const unsigned vcol = (unsigned)(g.x0 + cl) * 4u;
  const unsigned vst = c0 < cend ? (unsigned)(g.x0 + c0) * 4u : kDrop;
  bool colok[4];
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) colok[jj] = c0 + jj < cend;
  // The last four columns of an interior row may reach 1..3 frame columns
  // (interior widths that are not multiples of 4).  Those belong to the frame
  // kernel, which may run in the same launch (k_step), so the lane holding
  // them stores its valid columns one by one; only the wave whose tile holds
  // the interior's right edge (uniform) issues those stores.
  const bool edge_tile = cend < cb + 256;
  const bool part = !colok[3];
  const unsigned vst4 = part ? kDrop : vst;

  // Halo columns: the first 2H/4 lanes fetch them; the others fetch nothing
  // (a dropped load returns 0) and write that to their own slot, which the
  // centre row overwrites right after.
  unsigned vhal = kDrop;
  int hpos = H / 4 + lane;
  if (lane < H / 4) { vhal = (unsigned)(g.x0 + min(cb - H + 4 * lane, clast)) * 4u; hpos = lane; }
  else if (lane < 2 * (H / 4)) { vhal = (unsigned)(g.x0 + min(cb + 256 + 4 * (lane - H / 4), clast)) * 4u; hpos = 64 + lane; }

  const long long so = (long long)shot * g.fstride;
  const unsigned fb = (unsigned)g.fstride * 4u;
  const Rsrc R1 = rsrc(a.u1 + so, fb), R0 = rsrc(a.u0 + so, fb), R2 = rsrc(a.u2 + so, fb);
  const long long mo = (long long)sg.win * m.mstride;
  const Rsrc RP = rsrc(m.pm + mo, fb), RK = rsrc(m.ko + mo, fb);
  const Rsrc RC = rsrc(m.ctab, (unsigned)(m.nrow * m.cs) * 4u);
  const unsigned P4 = (unsigned)g.pitch * 4u;
  auto rowb = [&](int z) -> unsigned { return (unsigned)min(z, g.nz - 1) * P4; };
  Rsrc RTP;
  if (T::adjt) RTP = rsrc(m.tp + (long long)sg.win * m.tpstride, (unsigned)(4 * NFD) * fb);

  const int2 sp = make_int2(sg.r_u, sg.r_x);
  Rcv rr = a.r;  // this shot's receivers; a.r.n stays the stride of the data layout
  rr.s_l = sg.s_l;
  rr.s_r = sg.s_r;
  bool srcc[4];
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) srcc[jj] = c0 + jj == sp.y;

  // receiver step: data replacement of row s_z (kernel.cu:349-354), read once;
  // adjoint: the residual of step k added (times P) at the receivers
  float dval[4] = {0.f, 0.f, 0.f, 0.f};
  bool rflag[4] = {false, false, false, false};
  if ((T::rcv || T::adj) && a.r.s_z >= zs && a.r.s_z < min(ze, zhi)) {
    const int kd = T::rcv ? a.k + 1 : a.k;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      int s;
      if (colok[jj] && is_receiver(c0 + jj, rr, &s)) {
        dval[jj] = a.data[((long long)shot * a.nt + kd) * a.r.n + s];
        rflag[jj] = true;
      }
    }
  }
  Rsrc RU, RI1, RI2, RSU, RSQ;
  unsigned vcs = kDrop, vss = kDrop;
  if (T::rcv || T::srci) {
    RU = rsrc(a.uk + so, fb);
    RI1 = rsrc(a.i1 + so, fb); RI2 = rsrc(a.i2 + so, fb);
    RSU = rsrc(a.su + so, fb); RSQ = rsrc(a.sq + so, fb);
    if (a.icompen) { vcs = vcol; vss = vst; }
  }
  // Born / adjoint operands of the centre row: the background fields and the
  // reflectivity (Born) or the gradient accumulator (adjoint, when imaging)
  Rsrc RA, RB, RCC, RX;
  unsigned vxl = kDrop, vxs = kDrop;
  if (T::born || T::adj) {
    RA = rsrc(a.ua + so, fb); RB = rsrc(a.ub + so, fb); RCC = rsrc(a.uc + so, fb);
    RX = T::born ? rsrc(m.refl + (long long)shot * m.rstride, fb) : rsrc(a.i1 + so, fb);
    if (T::born || a.image) { vxl = vcol; vxs = vst; }
  }

  float4 q[QL];
#pragma unroll
  for (int s = 0; s < 2 * NFD + PD; ++s) q[s] = bld4(R1, vcol, rowb(zs - NFD + s));
  float4 hvb[PD + 1], u0b[PD + 1], pb[PD + 1];
  uint4 kb[PD + 1];
#pragma unroll
  for (int d = 0; d < PD; ++d) {
    kb[d] = bld4u(RK, vcol, rowb(zs + d));
    pb[d] = bld4(RP, vcol, rowb(zs + d));
    u0b[d] = bld4s(R0, vcol, rowb(zs + d));
    hvb[d] = bld4(R1, vhal, rowb(zs + d));
  }
  float4 ukb[1] = {}, i1b[1] = {}, i2b[1] = {}, sub[1] = {}, sqb[1] = {};
  float4 xab[2], xbb[2], xcb[2], xxb[2];
  if (T::born || T::adj) {
    xab[0] = bld4(RA, vxl, rowb(zs));
    xbb[0] = bld4(RB, vxl, rowb(zs));
    xcb[0] = bld4(RCC, vxl, rowb(zs));
    xxb[0] = bld4(RX, vxl, rowb(zs));
  }

  for (int z = zs; z < ze; z += QL) {
#pragma unroll
    for (int j = 0; j < QL; ++j) {
      const int zz = z + j;
      // (a) coefficient rows of this row (transposed stencil: the cells' own
      // c0' and the neighbour-tap planes of the row)
      CoefRow<T::adjt ? 1 : NFD> cr[4];
      float4 tpv[T::adjt ? 4 * NFD : 1];
      if (T::adjt) {
        cr[0].r[0] = bld4(RC, kb[0].x, 0);
        cr[1].r[0] = bld4(RC, kb[0].y, 0);
        cr[2].r[0] = bld4(RC, kb[0].z, 0);
        cr[3].r[0] = bld4(RC, kb[0].w, 0);
#pragma unroll
        for (int t = 0; t < 4 * NFD; ++t) tpv[t] = bld4(RTP, vcol, rowb(zz) + (unsigned)t * fb);
      } else {
        cr[0].load(RC, kb[0].x);
        cr[1].load(RC, kb[0].y);
        cr[2].load(RC, kb[0].z);
        cr[3].load(RC, kb[0].w);
      }
      __builtin_amdgcn_sched_barrier(0);
      // Receiver step: this row's imaging operands, issued ahead of the
      // prefetches so that their wait leaves the prefetches in flight (a
      // row of look-ahead for them would cost 20 VGPRs the kernel needs).
      if (T::rcv) {
        const unsigned rzz = rowb(zz);
        if (!T::rcvi) ukb[0] = bld4(RU, vcol, rzz);  // the split step never needs u(k)
        i1b[0] = bld4s(RI1, vcol, rzz);
        if (!T::rcvi) i2b[0] = bld4s(RI2, vcol, rzz);
        sub[0] = bld4s(RSU, vcs, rzz);
        sqb[0] = bld4s(RSQ, vcs, rzz);
      }
      if (T::srci) {  // this row's sum u and I2, for the source-side image terms
        const unsigned rzz = rowb(zz);
        i2b[0] = bld4s(RI2, vcol, rzz);
        sub[0] = bld4s(RSU, vcs, rzz);
      }
      // (b) operands of later rows
      q[(j + 2 * NFD + PD) % QL] = bld4(R1, vcol, rowb(zz + NFD + PD));
      kb[PD] = bld4u(RK, vcol, rowb(zz + PD));
      pb[PD] = bld4(RP, vcol, rowb(zz + PD));
      u0b[PD] = bld4s(R0, vcol, rowb(zz + PD));
      hvb[PD] = bld4(R1, vhal, rowb(zz + PD));
      if (T::born || T::adj) {
        xab[1] = bld4(RA, vxl, rowb(zz + 1));
        xbb[1] = bld4(RB, vxl, rowb(zz + 1));
        xcb[1] = bld4(RCC, vxl, rowb(zz + 1));
        xxb[1] = bld4(RX, vxl, rowb(zz + 1));
      }
      __builtin_amdgcn_sched_barrier(0);
      // (c) horizontal neighbours through LDS
      const float4 qc = q[(j + NFD) % QL];
      lds[wv][hpos] = hvb[0];
      lds[wv][H / 4 + lane] = qc;
      wave_lds_sync();
      float4 win[NWIN];
#pragma unroll
      for (int t = 0; t < NWIN; ++t) win[t] = lds[wv][lane + t];
      wave_lds_sync();

      const float4 pv = pb[0], u0v = u0b[0];
      float out[4];
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const float u = compr(qc, jj);
        float w1 = (float)(cr[jj].c0d() * (double)u);
        if (T::adjt) {  // each tap with the neighbour's coefficient: (L^T mu)
#pragma unroll
          for (int l = 1; l <= NFD; ++l) {
            const float cu = compr(tpv[4 * (l - 1)], jj), cd = compr(tpv[4 * (l - 1) + 1], jj);
            const float cl = compr(tpv[4 * (l - 1) + 2], jj), cr2 = compr(tpv[4 * (l - 1) + 3], jj);
            const float vs = __builtin_fmaf(cu, compr(q[(j + NFD - l) % QL], jj),
                                            cd * compr(q[(j + NFD + l) % QL], jj));
            float s = __builtin_fmaf(vs, m.hzx2_1,
                                     cl * compr(win[(H + jj - l) >> 2], (H + jj - l) & 3));
            s = __builtin_fmaf(cr2, compr(win[(H + jj + l) >> 2], (H + jj + l) & 3), s);
            w1 = w1 + s;
          }
        } else {
#pragma unroll
          for (int l = 1; l <= NFD; ++l) {
            float s = compr(q[(j + NFD - l) % QL], jj) + compr(q[(j + NFD + l) % QL], jj);
            s = __builtin_fmaf(s, m.hzx2_1, compr(win[(H + jj - l) >> 2], (H + jj - l) & 3));
            s = s + compr(win[(H + jj + l) >> 2], (H + jj + l) & 3);
            w1 = __builtin_fmaf(cr[jj].c(l), s, w1);
          }
        }
        float o = combine_p(u, compr(u0v, jj), compr(pv, jj), w1, T::f64);
        if (T::src)       // source (kernel.cu:74-77, 277-280)
          o = (zz == sp.x && srcc[jj]) ? o + a.wav : o;
        else if (T::rcv)  // data replacement
          o = (zz == a.r.s_z && dval[jj] != 0.0f) ? dval[jj] : o;
        else if (T::adj)  // residual injection, P * r(k)
          o = (zz == a.r.s_z && rflag[jj]) ? o + compr(pv, jj) * dval[jj] : o;
        if (T::born) {    // scattering source m * D2u(k)
          const float d2u = (compr(xab[0], jj) - 2.0f * compr(xbb[0], jj)) + compr(xcb[0], jj);
          o = __builtin_fmaf(compr(xxb[0], jj), d2u, o);
        }
        out[jj] = o;
      }
      const unsigned rowdrop = zz < zhi ? 0u : kDrop;
      const unsigned rz = rowb(zz);
      bst4s(R2, vst4 | rowdrop, rz, out[0], out[1], out[2], out[3]);
      if (edge_tile) {
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
          bst1s_(R2, ((part && colok[jj]) ? vst + 4u * jj : kDrop) | rowdrop, rz, out[jj]);
      }
      if (T::rcv) {  // imaging (kernel.cu:489-517); frame columns keep their values
        const float4 ukv = ukb[0], i1 = i1b[0], i2 = i2b[0], su = sub[0], sq = sqb[0];
        float r1o[4], r2o[4], rsu[4], rsq[4];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const float u = T::rcvi ? 0.0f : compr(ukv, jj);
          // split form (compensated imaging only): sum u already holds this
          // step's u(k), which the reconstruction added
          rsu[jj] = T::rcvi ? compr(su, jj) : (a.icompen ? compr(su, jj) + u : u);
          rsq[jj] = a.icompen ? compr(sq, jj) + out[jj] : out[jj];
          r1o[jj] = colok[jj] ? __builtin_fmaf(rsq[jj], rsu[jj], compr(i1, jj)) : compr(i1, jj);
          if (!T::rcvi) r2o[jj] = colok[jj] ? __builtin_fmaf(u, u, compr(i2, jj)) : compr(i2, jj);
          rsu[jj] = colok[jj] ? rsu[jj] : compr(su, jj);
          rsq[jj] = colok[jj] ? rsq[jj] : compr(sq, jj);
        }
        bst4s(RI1, vst | rowdrop, rz, r1o[0], r1o[1], r1o[2], r1o[3]);
        if (!T::rcvi) {
          bst4s(RI2, vst | rowdrop, rz, r2o[0], r2o[1], r2o[2], r2o[3]);
          bst4s(RSU, vss | rowdrop, rz, rsu[0], rsu[1], rsu[2], rsu[3]);
        }
        bst4s(RSQ, vss | rowdrop, rz, rsq[0], rsq[1], rsq[2], rsq[3]);
      }
      if (T::srci) {  // source-side image terms of step k on the u(k) just computed
        const float4 i2 = i2b[0], su = sub[0];
        float r2o[4], rsu[4];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const float u = out[jj];
          rsu[jj] = colok[jj] ? compr(su, jj) + u : compr(su, jj);
          r2o[jj] = colok[jj] ? __builtin_fmaf(u, u, compr(i2, jj)) : compr(i2, jj);
        }
        bst4s(RI2, vst | rowdrop, rz, r2o[0], r2o[1], r2o[2], r2o[3]);
        bst4s(RSU, vss | rowdrop, rz, rsu[0], rsu[1], rsu[2], rsu[3]);
      }
      if (T::adj) {  // G += D2u(k+1) * mu(k+2); mu(k+2) is this update's u0
        float go[4];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const float d2u = (compr(xab[0], jj) - 2.0f * compr(xbb[0], jj)) + compr(xcb[0], jj);
          const float gi = compr(xxb[0], jj);
          go[jj] = colok[jj] ? __builtin_fmaf(d2u, compr(u0v, jj), gi) : gi;
        }
        bst4(RX, vxs | rowdrop, rz, go[0], go[1], go[2], go[3]);
      }
      if (T::born || T::adj) {
        xab[0] = xab[1]; xbb[0] = xbb[1]; xcb[0] = xcb[1]; xxb[0] = xxb[1];
      }
#pragma unroll
      for (int d = 0; d < PD; ++d) {
        kb[d] = kb[d + 1]; pb[d] = pb[d + 1]; u0b[d] = u0b[d + 1]; hvb[d] = hvb[d + 1];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Frame update (hybrid ABC) as a row march.  The frame is cut into four
// regions: top rows [0, n2) and bottom rows [nz-n2, nz) over all columns,
// left columns [0, n2) and right columns [nx-n2, nx) over the rows between.
// A wave owns 64 consecutive columns of a region (lanes 1..62 write; lanes 0
// and 63 only supply lateral neighbours) and a run of its rows.  Marching
// down, iteration t computes on all 64 columns the pre-blend two-way value T
// of row t -- the value the reference's Hybrid1 reads from the un-blended
// DFW2: Add / BKAdd on the frame with the source, the data replacement, the
// Born source or the adjoint residual -- and then the one-way value and the
// blend of row t-1 (Hybrid1/Hybrid2, kernel.cu:116-183), which read T of rows
// t-2, t-1 and t: the inner neighbour and its laterals sit one row or one
// column away.  Vertical taps come from a register queue of u1 rows,
// horizontal taps from a per-wave LDS row; the reference's outer mirror
// (kernel.cu:65-68) is applied to the row and column indices of the loads.
//
// The row body has no branch (selects; accesses that must not happen get an
// out-of-range buffer offset), so the memory-counter waits stay exact: the
// coefficient row of row t+1 is gathered during row t (its offset arrived a
// row earlier), row operands arrive PD rows ahead and the one-way operands two
// rows ahead, and each row waits only on loads issued a row or more before.
// No block-level barrier: the four waves of a block are independent.
__device__ __forceinline__ float bld1(Rsrc r, unsigned vo, unsigned so) {
  return __uint_as_float((unsigned)__builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ unsigned bld1u(Rsrc r, unsigned vo, unsigned so) {
  return (unsigned)__builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0);
}
__device__ __forceinline__ void bst1(Rsrc r, unsigned vo, unsigned so, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)(vo + so), 0, 0);
}
// streaming variants (u0 rows, u(k+1) and strip writes: touched once per step)
__device__ __forceinline__ float bld1s(Rsrc r, unsigned vo, unsigned so) {
  return __uint_as_float((unsigned)__builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so,
                                                                        RTM_STREAM_POLICY));
}
__device__ __forceinline__ void bst1s(Rsrc r, unsigned vo, unsigned so, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)(vo + so), 0, RTM_STREAM_POLICY);
}

// Strip element of frame cell (z, x) of a shot nx wide, or -1 (strip_index
// without branches, in 32 bits: a slot holds fewer than 2^31 elements).
__device__ __forceinline__ int strip_elem(const Grid& g, int nx, int mnx, int z, int x) {
  const int a = g.mnz * g.nfd, b = g.nfd * mnx, n2 = g.n2, nfd = g.nfd;
  const bool zin = z >= n2 && z < g.nz - n2, xin = x >= n2 && x < nx - n2;
  int e = -1;
  e = (zin && x >= n2 - nfd && x < n2) ? (z - n2) * nfd + (n2 - 1 - x) : e;
  e = (zin && x >= nx - n2 && x < nx - n2 + nfd) ? a + (z - n2) * nfd + (x - (nx - n2)) : e;
  e = (xin && z >= n2 - nfd && z < n2) ? 2 * a + (n2 - 1 - z) * mnx + (x - n2) : e;
  e = (xin && z >= g.nz - n2 && z < g.nz - n2 + nfd) ? 2 * a + b + (z - (g.nz - n2)) * mnx + (x - n2) : e;
  return e;
}

template <int NFD, int MODE>
__device__ __forceinline__ void frame_body(const FrameArgs& a, const unsigned bid) {
  constexpr int PD = 2;                   // rows of prefetch
  constexpr int NB = NFD > 2 ? NFD : 2;   // queue rows kept behind the T row
  // The row loop is unrolled by U, a multiple of 12 that holds the u1 rows
  // t-NB .. t+NFD+PD; every other per-row buffer is circular with a length
  // dividing 12 (2, 3 or 6), so no buffer is ever copied, only renamed.
  constexpr int U = (NB + NFD + 1 + PD + 11) / 12 * 12;
  // LDS row: halo | 64 columns | halo | dump slot; a packed side record
  // (band 4) uses halo | 32 columns | halo | halo | 32 columns | halo | dump
  constexpr int WN = 64 + 4 * NFD + 1;
  __shared__ float win[4][WN];
  const Grid& g = a.g;
  const Model& m = a.m;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ri = (int)bid * 4 + wv;
  if (ri >= a.nrec) return;
  const int4 rec = a.recs[ri];
  const int z0 = rec.x, cx0 = rec.y, band = rec.z & 15, nrows = (rec.z >> 4) & 0xffff, shot = rec.w;
  const int ncol = (rec.z >> 20) & 0xff;  // top/bottom: written lanes 1..ncol (0: 1..62)
  const ShotGeo sg = a.geo[shot];
  const int NX = sg.nx, NZ = g.nz, N2 = g.n2;
  // Band 4 packs both side bands of a row run into one wave (n2 <= 30):
  // lanes 0..31 hold columns -1..30 (the left band and its inner
  // neighbour), lanes 32..63 columns nx-31..nx; lanes 1..30 and 33..62
  // write.  The +-1 lane shuffles of the one-way update stay inside a half.
  const bool packed = band == 4;
  const bool rhalf = packed && lane >= 32;
  const int x = packed ? (rhalf ? NX - 63 + lane : lane - 1) : cx0 + lane;
  const int xm = mirror_idx(x, NX);
  int ccl = 0, ccr = NX;  // output columns of the region
  if (band == 2) ccr = N2;
  else if (band == 3) ccl = NX - N2;
  else if (packed) { ccl = rhalf ? NX - N2 : 0; ccr = rhalf ? NX : N2; }
  const int hl_ = packed ? (lane & 31) : lane;  // lane within its (half) run
  const bool outcol = hl_ >= 1 && hl_ <= (packed ? 30 : (ncol ? ncol : 62)) && x >= ccl && x < ccr;
  // LDS slot of the lane's column (packed: the right half after the left
  // half's right halo and its own left halo)
  const int slot = packed && rhalf ? 3 * NFD + lane : NFD + lane;
  // A top or bottom run whose written columns all lie in [n2+1, nx-n2-2]
  // holds only cells of the top (bottom) edges: no corner cell, no side cell
  // (the ring of (z, x) is z, resp. nz-1-z).  Those waves take a shorter
  // one-way update; the choice is uniform per wave.
  const bool pure_tb = band < 2 && cx0 + 1 >= N2 + 1 && cx0 + 62 <= NX - N2 - 2;

  const long long so = (long long)shot * g.fstride;
  const unsigned fb = (unsigned)g.fstride * 4u;
  const Rsrc R1 = rsrc(a.u1 + so, fb), R0 = rsrc(a.u0 + so, fb), R2 = rsrc(a.u2 + so, fb);
  const long long mo = (long long)sg.win * m.mstride;
  const Rsrc RP = rsrc(m.pm + mo, fb), RK = rsrc(m.ko + mo, fb);
  const Rsrc RV = rsrc(m.v + mo, fb), RR = rsrc(m.r1 + mo, fb);
  const Rsrc RC = rsrc(m.ctab, (unsigned)(m.nrow * m.cs) * 4u);
  const Rsrc RW = rsrc(m.w, (unsigned)(N2 + 1) * 4u);
  const unsigned P4 = (unsigned)g.pitch * 4u;
  auto rowm = [&](int z) -> unsigned { return (unsigned)mirror_idx(z, NZ) * P4; };
  auto colb = [&](int c) -> unsigned { return (unsigned)(g.x0 + c) * 4u; };
  const unsigned vcol = colb(xm);

  // halo columns of the LDS row: lanes 0..NFD-1 the left ones, NFD..2NFD-1
  // the right ones (packed: the left half's two halos, then lanes
  // 2NFD..4NFD-1 the right half's); the other lanes fetch nothing and write
  // the dump slot
  unsigned vhal = kDrop;
  int hpos = WN - 1;
  if (!packed) {
    if (lane < NFD) { vhal = colb(mirror_idx(cx0 - NFD + lane, NX)); hpos = lane; }
    else if (lane < 2 * NFD) { vhal = colb(mirror_idx(cx0 + 64 + lane - NFD, NX)); hpos = 64 + lane; }
  } else {
    if (lane < NFD) { vhal = colb(mirror_idx(-1 - NFD + lane, NX)); hpos = lane; }
    else if (lane < 2 * NFD) { vhal = colb(mirror_idx(31 + lane - NFD, NX)); hpos = 32 + lane; }
    else if (lane < 3 * NFD) { vhal = colb(mirror_idx(NX - 31 - 3 * NFD + lane, NX)); hpos = 32 + lane; }
    else if (lane < 4 * NFD) { vhal = colb(mirror_idx(NX + 1 - 3 * NFD + lane, NX)); hpos = 64 + lane; }
  }

  // receivers on row s_z (kernel.cu:349): this step's data sample of the lane's
  // trace (receiver field: sample k+1; adjoint: sample k), dropped elsewhere
  Rcv rr = a.r;
  rr.s_l = sg.s_l;
  rr.s_r = sg.s_r;
  int rs = 0;
  const bool isrcv = (MODE == FR_RCV || MODE == FR_ADJ) && is_receiver(x, rr, &rs);
  Rsrc RD;
  if (MODE == FR_RCV || MODE == FR_ADJ) {
    const int kd = MODE == FR_RCV ? a.k + 1 : a.k;
    RD = rsrc(a.data + ((long long)shot * a.nt + kd) * a.r.n, (unsigned)a.r.n * 4u);
  }
  const unsigned vrcv = isrcv ? (unsigned)rs * 4u : kDrop;
  // strip slot of this step (FP32) or its FP32 staging (FP16 / BF16)
  Rsrc RS;
  if (MODE == FR_FWD) {
    const long long ss = a.st.ss;
    const float* sb = a.st.fmt == 0
        ? (const float*)a.st.base + ((long long)shot * a.st.nslot + a.k % a.st.nslot) * ss
        : a.st.stage + ((long long)shot * 2 + (a.k & 1)) * ss;
    RS = rsrc(sb, a.save_strips ? (unsigned)ss * 4u : 0u);
  }
  // adjoint with per-cell rows: the interior cells of the march (the rows and
  // columns next to the frame, whose pre-blend values the one-way update
  // reads) take the transposed stencil of the interior kernel
  const bool adjt = MODE == FR_ADJ && a.adjt != 0;
  Rsrc RTP;
  if (MODE == FR_ADJ) RTP = rsrc(m.tp + (long long)sg.win * m.tpstride, adjt ? (unsigned)(4 * NFD) * fb : 0u);
  Rsrc RBX, RBA, RBB, RBC;
  if (MODE == FR_BORN) {
    RBX = rsrc(m.refl + (long long)shot * m.rstride, fb);
    RBA = rsrc(a.ua + so, fb); RBB = rsrc(a.ub + so, fb); RBC = rsrc(a.uc + so, fb);
  }
  constexpr int NE = MODE == FR_BORN ? 4 : 1;  // extra operands of a T row

  // one-way operands of row zr (velocity, r1, the Q1 velocity, blend weight)
  auto abc_ops = [&](int zr, float& vp, float& r1v, float& vq, float& wv_) {
    const int ar = min(min(zr, x), min(NZ - 1 - zr, NX - 1 - x));
    // row zr as a linear index of the model: row zr / NX, column zr % NX (uniform)
    const int zdiv = zr / NX, zmod = zr - zdiv * NX;
    const bool topc = zr == ar, botc = !topc && zr == NZ - 1 - ar;
    // top: the cell itself; bottom: Dv[(N2-x)*NX+y] (kernel.cu:132); sides:
    // Dv[(N2-y)*NX+x] (kernel.cu:128,136), a top-band cell
    int zq = topc ? zr : (botc ? ar : ar + zdiv);
    int xq = (topc || botc) ? x : zmod;
    zq = min(max(zq, 0), NZ - 1);
    xq = min(max(xq, 0), NX - 1);
    vp = bld1(RV, vcol, rowm(zr));
    r1v = bld1(RR, vcol, rowm(zr));
    vq = bld1(RV, colb(xq) + (unsigned)zq * P4, 0u);
    wv_ = bld1(RW, (unsigned)min(max(N2 - ar, 0), N2) * 4u, 0u);
  };
  // extra operands of T row zr: the data sample, or the Born source terms
  auto t_ops = [&](int zr, float* e) {
    if (MODE == FR_RCV || MODE == FR_ADJ) e[0] = bld1(RD, zr == a.r.s_z ? vrcv : kDrop, 0u);
    else if (MODE == FR_BORN) {
      const unsigned rz = rowm(zr);
      e[0] = bld1(RBX, vcol, rz); e[1] = bld1(RBA, vcol, rz);
      e[2] = bld1(RBB, vcol, rz); e[3] = bld1(RBC, vcol, rz);
    } else {
      e[0] = 0.f;
    }
  };

  // Circular buffers, indexed by the iteration number it (T row t0 + it):
  //   q[U]     u1 rows (row t0 - NB + i in slot i % U)
  //   kob[6]   coefficient offsets of T rows it-1 .. it+KD (KD = 4 rows ahead:
  //            the coefficient row of it+1 is gathered from them one row early)
  //   pmb, halb [3]: P and halo of T rows it .. it+PD
  //   u0b[6]   u0 rows it-2 .. it+PD; cr[2] coefficient rows of it, it+1
  //   abc [3]  one-way operands of rows it-1 .. it+1; ext[2] T-row extras
  //   T[3]     T rows it-2 .. it
  const int t0 = z0 - 1;          // T row of the first iteration
  const int niter = nrows + 2;    // T rows z0-1 .. z0+nrows
  float q[U];
#pragma unroll
  for (int s = 0; s < NB + NFD + PD; ++s) q[s] = bld1(R1, vcol, rowm(t0 - NB + s));
  constexpr int KD = 4;
  unsigned kob[6];
#pragma unroll
  for (int d = 0; d < KD; ++d) kob[d] = bld1u(RK, vcol, rowm(t0 + d));
  float pmb[3], halb[3], u0b[6];
#pragma unroll
  for (int d = 0; d < PD; ++d) {
    pmb[d] = bld1(RP, vcol, rowm(t0 + d));
    halb[d] = bld1(R1, vhal, rowm(t0 + d));
    u0b[2 + d] = bld1s(R0, vcol, rowm(t0 + d));   // slot (it + 2) % 6 holds row t0 + it
  }
  u0b[0] = bld1s(R0, vcol, rowm(t0 - 2));
  u0b[1] = bld1s(R0, vcol, rowm(t0 - 1));
  CoefRow<NFD> cr[2];
  cr[0].load(RC, kob[0]);
  float ext[2][NE];
  t_ops(t0, ext[0]);
  float T[3] = {0.f, 0.f, 0.f};
  float vpb[3] = {0.f, 0.f, 0.f}, r1b[3] = {0.f, 0.f, 0.f}, vqb[3] = {0.f, 0.f, 0.f},
        wb[3] = {0.f, 0.f, 0.f};   // slot (it + 1) % 3 holds row t0 + it

  for (int base = 0; base < niter; base += U) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int it = base + j;   // base is a multiple of U, so it % L == j % L for L | 12
      if (it >= niter) break;
      const int tt = t0 + it;
      // (1) coefficient row of T row tt+1 (its offset arrived a row ago)
      cr[(j + 1) % 2].load(RC, kob[(j + 1) % 6]);
      // (2) operands of later rows
      kob[(j + KD) % 6] = bld1u(RK, vcol, rowm(tt + KD));
      q[(j + NB + NFD + PD) % U] = bld1(R1, vcol, rowm(tt + NFD + PD));
      pmb[(j + PD) % 3] = bld1(RP, vcol, rowm(tt + PD));
      halb[(j + PD) % 3] = bld1(R1, vhal, rowm(tt + PD));
      u0b[(j + 2 + PD) % 6] = bld1s(R0, vcol, rowm(tt + PD));
      abc_ops(tt + 1, vpb[(j + 2) % 3], r1b[(j + 2) % 3], vqb[(j + 2) % 3], wb[(j + 2) % 3]);
      t_ops(tt + 1, ext[(j + 1) % 2]);
      __builtin_amdgcn_sched_barrier(0);
      // (3) horizontal neighbours of row tt through LDS
      const float u = q[(j + NB) % U];
      win[wv][slot] = u;
      win[wv][hpos] = halb[j % 3];
      wave_lds_sync();
      float hl[NFD], hr[NFD];
#pragma unroll
      for (int l = 1; l <= NFD; ++l) {
        hl[l - 1] = win[wv][slot - l];
        hr[l - 1] = win[wv][slot + l];
      }
      wave_lds_sync();
      // (4) pre-blend two-way value T of (tt, x) (Add / BKAdd, kernel.cu:46-80, 339-380)
      const CoefRow<NFD>& crc = cr[j % 2];
      float w1 = (float)(crc.c0d() * (double)u);
#pragma unroll
      for (int l = 1; l <= NFD; ++l) {
        float s = q[(j + NB - l + U) % U] + q[(j + NB + l) % U];
        s = __builtin_fmaf(s, m.hzx2_1, hl[l - 1]);
        s = s + hr[l - 1];
        w1 = __builtin_fmaf(crc.c(l), s, w1);
      }
      if (MODE == FR_ADJ) {  // (L^T mu) on interior cells, each tap with the neighbour's row
        const bool inner = tt >= N2 && tt < NZ - N2 && x >= N2 && x < NX - N2;
        float w1t = (float)(crc.c0d() * (double)u);
#pragma unroll
        for (int l = 1; l <= NFD; ++l) {
          const unsigned rz = rowm(tt);
          const float cu = bld1(RTP, vcol, rz + (unsigned)(4 * (l - 1)) * fb);
          const float cd = bld1(RTP, vcol, rz + (unsigned)(4 * (l - 1) + 1) * fb);
          const float cl = bld1(RTP, vcol, rz + (unsigned)(4 * (l - 1) + 2) * fb);
          const float cr2 = bld1(RTP, vcol, rz + (unsigned)(4 * (l - 1) + 3) * fb);
          const float vs = __builtin_fmaf(cu, q[(j + NB - l + U) % U], cd * q[(j + NB + l) % U]);
          float sv = __builtin_fmaf(vs, m.hzx2_1, cl * hl[l - 1]);
          sv = __builtin_fmaf(cr2, hr[l - 1], sv);
          w1t = w1t + sv;
        }
        w1 = (adjt && inner) ? w1t : w1;
      }
      const float u0 = u0b[(j + 2) % 6], pmv = pmb[j % 3];
      const float* ex = ext[j % 2];
      float Tn;
      if (MODE == FR_FWD) {
        Tn = combine_p(u, u0, pmv, w1, a.f64comb != 0);
        Tn = (tt == sg.r_u && x == sg.r_x) ? Tn + a.wav : Tn;
      } else if (MODE == FR_BORN) {  // scattered field: m * D2u on the interior, no source
        Tn = combine_p(u, u0, pmv, w1, a.f64comb != 0);
        const float d2u = (ex[1] - 2.0f * ex[2]) + ex[3];
        const bool inner = tt >= N2 && tt < NZ - N2 && x >= N2 && x < NX - N2;
        Tn = inner ? __builtin_fmaf(ex[0], d2u, Tn) : Tn;
      } else if (MODE == FR_ADJ) {   // adjoint field: P * r(k) added at the receivers
        Tn = combine_p(u, u0, pmv, w1, false);
        Tn = (isrcv && tt == a.r.s_z) ? Tn + pmv * ex[0] : Tn;
      } else {                       // receiver field: data replace where non-zero
        Tn = combine_p(u, u0, pmv, w1, false);
        Tn = (isrcv && tt == a.r.s_z && ex[0] != 0.0f) ? ex[0] : Tn;
      }
      T[j % 3] = Tn;
      // (5) one-way value and blend of row z = tt - 1 (written from the third row on)
      if (pure_tb) {  // top edge (inner row z+1) or bottom edge (inner row z-1) (kernel.cu:122-125,130-133)
        const int z = tt - 1;
        const bool tp = band == 0;
        const float T0 = T[(j + 2) % 3];
        const float Tin = tp ? T[j % 3] : T[(j + 1) % 3];
        const float U1z = q[(j + NB - 1) % U];
        const float U1i = tp ? q[(j + NB) % U] : q[(j + NB - 2) % U];
        const float U0z = u0b[(j + 1) % 6];
        const float U0i = tp ? u0b[(j + 2) % 6] : u0b[j % 6];
        const float Til = __shfl_up(Tin, 1), Tir = __shfl_down(Tin, 1);
        const float U0l = __shfl_up(U0z, 1), U0r = __shfl_down(U0z, 1);
        const int ab = j % 3;
        const float vp = vpb[ab], vq = vqb[ab], wgt = wb[ab];
        const float tv = m.taoh * vp;
        const float t1 = tv * ((Tin - U0i) + U0z);
        const float t2 = ((((-2.0f * U1z) + U0z) + Tin) - 2.0f * U1i) + U0i;
        const float lap = ((((Tir - 2.0f * Tin) + Til) + U0r) - 2.0f * U0z) + U0l;
        const float t3 = ((m.taoh2 * vq) * vq) * lap;
        const float O = (1.0f / (tv + 1.0f)) * ((t1 - t2) + t3);
        const float val = (1.0f - wgt) * T0 + wgt * O;
        const bool out = outcol && it >= 2 && z < z0 + nrows;
        bst1s(R2, out ? vcol : kDrop, rowm(z), val);
        if (MODE == FR_FWD) {  // Hybrid3 (kernel.cu:184-208)
          const int e = strip_elem(g, NX, sg.mnx, z, x);
          bst1s(RS, (out && e >= 0) ? (unsigned)e * 4u : kDrop, 0u, val);
        }
      } else {
        const int z = tt - 1;
        const float Tm = T[(j + 1) % 3], T0 = T[(j + 2) % 3], Tp = T[j % 3];
        const float U1m = q[(j + NB - 2) % U], U1z = q[(j + NB - 1) % U], U1p = q[(j + NB) % U];
        const float U0m = u0b[j % 6], U0z = u0b[(j + 1) % 6], U0p = u0b[(j + 2) % 6];
        const float T0l = __shfl_up(T0, 1), T0r = __shfl_down(T0, 1);
        const float Tml = __shfl_up(Tm, 1), Tmr = __shfl_down(Tm, 1);
        const float Tpl = __shfl_up(Tp, 1), Tpr = __shfl_down(Tp, 1);
        const float U0l = __shfl_up(U0z, 1), U0r = __shfl_down(U0z, 1);
        const float U1l = __shfl_up(U1z, 1), U1r = __shfl_down(U1z, 1);
        const int ar = min(min(z, x), min(NZ - 1 - z, NX - 1 - x));  // ring
        const int br = NZ - 1 - ar, crr = NX - 1 - ar;
        const bool zc = (z - ar <= 1) || (br - z <= 1);
        const bool xc = (x - ar <= 1) || (crr - x <= 1);
        const bool corner = zc && xc;
        const bool top = z == ar, bot = !top && z == br, side = !top && !bot;
        const bool inr = x == ar;  // side cell with its inner neighbour to the right
        const int ab = j % 3;      // one-way operands of row z
        const float vp = vpb[ab], r1v = r1b[ab], vq = vqb[ab], wgt = wb[ab];
        // corner cells (kernel.cu:138-155)
        const float Tx = (x - ar <= 1) ? T0r : T0l;
        const float Tz = (z - ar <= 1) ? Tp : Tm;
        const float num_c = U1z + r1v * (Tx + Tz);
        const float den_c = 2.0f * r1v + 1.0f;
        // one-way update of the edges (kernel.cu:122-137)
        const float Ti = top ? Tp : (bot ? Tm : (inr ? T0r : T0l));
        const float Tip = top ? Tpr : (bot ? Tmr : (inr ? Tpr : Tpl));
        const float Tim = top ? Tpl : (bot ? Tml : (inr ? Tmr : Tml));
        const float U0i = top ? U0p : (bot ? U0m : (inr ? U0r : U0l));
        const float U0pp = side ? U0p : U0r, U0pm = side ? U0m : U0l;
        const float U1i = top ? U1p : (bot ? U1m : (inr ? U1r : U1l));
        const float tv = m.taoh * vp;
        const float t1 = tv * ((Ti - U0i) + U0z);
        const float t2 = ((((-2.0f * U1z) + U0z) + Ti) - 2.0f * U1i) + U0i;
        const float lap = ((((Tip - 2.0f * Ti) + Tim) + U0pp) - 2.0f * U0z) + U0pm;
        const float t3 = ((m.taoh2 * vq) * vq) * lap;
        const float num_s = (t1 - t2) + t3;
        const float den_s = tv + 1.0f;
        const float O = (1.0f / (corner ? den_c : den_s)) * (corner ? num_c : num_s);
        const float val = (1.0f - wgt) * T0 + wgt * O;
        const bool out = outcol && it >= 2 && z < z0 + nrows;
        bst1s(R2, out ? vcol : kDrop, rowm(z), val);
        if (MODE == FR_FWD) {  // Hybrid3 (kernel.cu:184-208)
          const int e = strip_elem(g, NX, sg.mnx, z, x);
          bst1s(RS, (out && e >= 0) ? (unsigned)e * 4u : kDrop, 0u, val);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Top and bottom bands away from the corners, four columns per lane.  A record
// (z0, cs, band | rows << 4, shot) covers the 256 columns cs .. cs+255 (cs ==
// n2 mod 4, so each lane's float4 is aligned like the interior kernel's) and
// writes the columns cs+1 .. cs+252 left of nx-n2-1.  Every written cell is a
// top-edge (bottom-edge) cell of Hybrid1/Hybrid2 (kernel.cu:122-125, 130-133):
// its one-way update reads the inner row and the lateral neighbours x-1, x+1
// of the same row, so lane 0's first column and lane 63's last three only
// supply neighbours.  The row march, the look-ahead of the operands and the
// arithmetic are those of frame_body's top/bottom path, operation for
// operation, so both give the same bits; per cell this one spends a quarter of
// the row addressing, masks and buffer instructions.  FR_FWD and FR_RCV only.
template <int NFD, int MODE, bool BOT>
__device__ __forceinline__ void frame_tb4_body(const FrameArgs& a, const int4 rec) {
  static_assert(MODE == FR_FWD || MODE == FR_RCV, "forward and receiver frames only");
  constexpr int PD = 2;                   // rows of prefetch
  constexpr int NB = NFD > 2 ? NFD : 2;   // queue rows kept behind the T row
  constexpr int U = (NB + NFD + 1 + PD + 11) / 12 * 12;
  constexpr int H = (NFD + 3) & ~3;
  constexpr int W4 = 64 + 2 * (H / 4);
  constexpr int NWIN = 1 + 2 * (H / 4);
  __shared__ float4 lds[4][W4];
  const Grid& g = a.g;
  const Model& m = a.m;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int z0 = rec.x, cs = rec.y, nrows = (rec.z >> 4) & 0xffff, shot = rec.w;
  const ShotGeo sg = a.geo[shot];
  const int NX = sg.nx, NZ = g.nz, N2 = g.n2;
  const int wend = NX - N2 - 1;  // first column past the edge cells
  const int c0 = cs + 4 * lane;
  const int clast = NX - ((NX - N2) & 3);  // largest c == n2 (mod 4) with c <= nx
  const unsigned vcol = (unsigned)(g.x0 + min(c0, clast)) * 4u;
  bool wr[4];
  bool full = true;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    wr[jj] = c0 + jj > cs && c0 + jj <= cs + 252 && c0 + jj < wend;
    full = full && wr[jj];
  }
  const unsigned vst = (unsigned)(g.x0 + c0) * 4u;
  // halo float4s of the LDS row: lanes 0 .. H/4-1 left, H/4 .. 2H/4-1 right
  // (columns past the shot's width only feed unwritten columns)
  unsigned vhal = kDrop;
  int hpos = H / 4 + lane;
  if (lane < H / 4) { vhal = (unsigned)(g.x0 + cs - H + 4 * lane) * 4u; hpos = lane; }
  else if (lane < 2 * (H / 4)) {
    vhal = (unsigned)(g.x0 + min(cs + 256 + 4 * (lane - H / 4), clast)) * 4u;
    hpos = 64 + lane;
  }

  const long long so = (long long)shot * g.fstride;
  const unsigned fb = (unsigned)g.fstride * 4u;
  const Rsrc R1 = rsrc(a.u1 + so, fb), R0 = rsrc(a.u0 + so, fb), R2 = rsrc(a.u2 + so, fb);
  const long long mo = (long long)sg.win * m.mstride;
  const Rsrc RP = rsrc(m.pm + mo, fb), RK = rsrc(m.ko + mo, fb), RV = rsrc(m.v + mo, fb);
  const Rsrc RC = rsrc(m.ctab, (unsigned)(m.nrow * m.cs) * 4u);
  const Rsrc RW = rsrc(m.w, (unsigned)(N2 + 1) * 4u);
  const unsigned P4 = (unsigned)g.pitch * 4u;
  auto rowm = [&](int z) -> unsigned { return (unsigned)mirror_idx(z, NZ) * P4; };
  auto rowc = [&](int z) -> unsigned { return (unsigned)min(max(z, 0), NZ - 1) * P4; };
  // ring of row zr: zr (top edge), nz-1-zr (bottom edge)
  auto ring = [&](int zr) -> int { return BOT ? NZ - 1 - zr : zr; };

  // receiver field: this step's data sample (k+1) of each receiver column,
  // replacing the field on row s_z where non-zero (kernel.cu:349-354)
  // (a column without a receiver reads 0: its dropped load returns 0)
  float ev[4] = {0.f, 0.f, 0.f, 0.f};
  if (MODE == FR_RCV) {
    Rcv rr = a.r;
    rr.s_l = sg.s_l;
    rr.s_r = sg.s_r;
    const Rsrc RD = rsrc(a.data + ((long long)shot * a.nt + a.k + 1) * a.r.n, (unsigned)a.r.n * 4u);
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      int rs = 0;
      const bool isr = is_receiver(c0 + jj, rr, &rs);
      ev[jj] = bld1(RD, isr ? (unsigned)rs * 4u : kDrop, 0u);
    }
  }
  // strip slot of this step (FP32) or its FP32 staging (FP16 / BF16): the
  // rows n2-nfd .. n2-1 (top) and nz-n2 .. nz-n2+nfd-1 (bottom), Hybrid3
  // (kernel.cu:184-208), element 2a + (n2-1-z)*mnx + x-n2, resp.
  // 2a + b + (z-(nz-n2))*mnx + x-n2, as strip_elem
  Rsrc RS;
  if (MODE == FR_FWD) {
    const long long ss = a.st.ss;
    const float* sb = a.st.fmt == 0
        ? (const float*)a.st.base + ((long long)shot * a.st.nslot + a.k % a.st.nslot) * ss
        : a.st.stage + ((long long)shot * 2 + (a.k & 1)) * ss;
    RS = rsrc(sb, a.save_strips ? (unsigned)ss * 4u : 0u);
  }
  const int sa = g.mnz * g.nfd, sbb = g.nfd * sg.mnx;

  const int t0 = z0 - 1;          // T row of the first iteration
  const int niter = nrows + 2;    // T rows z0-1 .. z0+nrows
  float4 q[U];
#pragma unroll
  for (int s = 0; s < NB + NFD + PD; ++s) q[s] = bld4(R1, vcol, rowm(t0 - NB + s));
  uint4 kb[3];
  float4 pmb[3], halb[3], u0b[6];
#pragma unroll
  for (int d = 0; d < PD; ++d) {
    kb[d] = bld4u(RK, vcol, rowm(t0 + d));
    pmb[d] = bld4(RP, vcol, rowm(t0 + d));
    halb[d] = bld4(R1, vhal, rowm(t0 + d));
    u0b[2 + d] = bld4s(R0, vcol, rowm(t0 + d));   // slot (it + 2) % 6 holds row t0 + it
  }
  u0b[0] = bld4s(R0, vcol, rowm(t0 - 2));
  u0b[1] = bld4s(R0, vcol, rowm(t0 - 1));
  CoefRow<NFD> cr[2][4];  // coefficient rows of T rows it, it+1 (slot it % 2)
  cr[0][0].load(RC, kb[0].x);
  cr[0][1].load(RC, kb[0].y);
  cr[0][2].load(RC, kb[0].z);
  cr[0][3].load(RC, kb[0].w);
  const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 T[3] = {z4, z4, z4};
  float4 vpb[3] = {z4, z4, z4}, vqb[3] = {z4, z4, z4};  // slot (it + 1) % 3 holds row t0 + it
  float wb[3] = {0.f, 0.f, 0.f};

  for (int base = 0; base < niter; base += U) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int it = base + j;   // base is a multiple of U, so it % L == j % L for L | 12
      if (it >= niter) break;
      const int tt = t0 + it;
      // (1) coefficient rows of T row tt+1 (their offsets arrived a row ago)
      {
        const uint4 kn = kb[(j + 1) % 3];
        cr[(j + 1) % 2][0].load(RC, kn.x);
        cr[(j + 1) % 2][1].load(RC, kn.y);
        cr[(j + 1) % 2][2].load(RC, kn.z);
        cr[(j + 1) % 2][3].load(RC, kn.w);
      }
      // (2) operands of later rows
      q[(j + NB + NFD + PD) % U] = bld4(R1, vcol, rowm(tt + NFD + PD));
      kb[(j + PD) % 3] = bld4u(RK, vcol, rowm(tt + PD));
      pmb[(j + PD) % 3] = bld4(RP, vcol, rowm(tt + PD));
      halb[(j + PD) % 3] = bld4(R1, vhal, rowm(tt + PD));
      u0b[(j + 2 + PD) % 6] = bld4s(R0, vcol, rowm(tt + PD));
      // one-way operands of row tt+1: velocity, the Q1 velocity (the cell
      // itself on the top edge; Dv[(N2-x)*NX+y], the mirrored top-band cell,
      // on the bottom edge, kernel.cu:132) and the blend weight of its ring
      vpb[(j + 2) % 3] = bld4(RV, vcol, rowm(tt + 1));
      if (BOT) vqb[(j + 2) % 3] = bld4(RV, vcol, rowc(NZ - 1 - (tt + 1)));
      wb[(j + 2) % 3] = bld1(RW, (unsigned)min(max(N2 - ring(tt + 1), 0), N2) * 4u, 0u);
      __builtin_amdgcn_sched_barrier(0);
      // (3) horizontal neighbours of row tt through LDS
      const float4 qc = q[(j + NB) % U];
      lds[wv][hpos] = halb[j % 3];
      lds[wv][H / 4 + lane] = qc;
      wave_lds_sync();
      float4 win[NWIN];
#pragma unroll
      for (int t = 0; t < NWIN; ++t) win[t] = lds[wv][lane + t];
      wave_lds_sync();
      // (4) pre-blend two-way values T of row tt (Add / BKAdd, kernel.cu:46-80, 339-380)
      const float4 u0v = u0b[(j + 2) % 6], pv = pmb[j % 3];
      float tn[4];
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const CoefRow<NFD>& crc = cr[j % 2][jj];
        const float u = compr(qc, jj);
        float w1 = (float)(crc.c0d() * (double)u);
#pragma unroll
        for (int l = 1; l <= NFD; ++l) {
          float s = compr(q[(j + NB - l + U) % U], jj) + compr(q[(j + NB + l) % U], jj);
          s = __builtin_fmaf(s, m.hzx2_1, compr(win[(H + jj - l) >> 2], (H + jj - l) & 3));
          s = s + compr(win[(H + jj + l) >> 2], (H + jj + l) & 3);
          w1 = __builtin_fmaf(crc.c(l), s, w1);
        }
        float Tn;
        if (MODE == FR_FWD) {
          Tn = combine_p(u, compr(u0v, jj), compr(pv, jj), w1, a.f64comb != 0);
          Tn = (tt == sg.r_u && c0 + jj == sg.r_x) ? Tn + a.wav : Tn;
        } else {
          Tn = combine_p(u, compr(u0v, jj), compr(pv, jj), w1, false);
          Tn = (tt == a.r.s_z && ev[jj] != 0.0f) ? ev[jj] : Tn;
        }
        tn[jj] = Tn;
      }
      T[j % 3] = make_float4(tn[0], tn[1], tn[2], tn[3]);
      // (5) one-way value and blend of row z = tt - 1 (written from the third row on)
      const int z = tt - 1;
      const float4 T0v = T[(j + 2) % 3];
      const float4 Tinv = BOT ? T[(j + 1) % 3] : T[j % 3];
      const float4 U1zv = q[(j + NB - 1) % U];
      const float4 U1iv = BOT ? q[(j + NB - 2) % U] : q[(j + NB) % U];
      const float4 U0zv = u0b[(j + 1) % 6];
      const float4 U0iv = BOT ? u0b[j % 6] : u0b[(j + 2) % 6];
      // lateral neighbours across lanes: x-1 of component 0, x+1 of component 3
      const float Tl = __shfl_up(Tinv.w, 1), Tr = __shfl_down(Tinv.x, 1);
      const float Ul = __shfl_up(U0zv.w, 1), Ur = __shfl_down(U0zv.x, 1);
      const int ab = j % 3;
      const float4 vpv = vpb[ab];
      const float4 vqv = BOT ? vqb[ab] : vpb[ab];
      const float wgt = wb[ab];
      float val[4];
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const float T0 = compr(T0v, jj), Tin = compr(Tinv, jj);
        const float U1z = compr(U1zv, jj), U1i = compr(U1iv, jj);
        const float U0z = compr(U0zv, jj), U0i = compr(U0iv, jj);
        const float Til = jj == 0 ? Tl : compr(Tinv, jj - 1);
        const float Tir = jj == 3 ? Tr : compr(Tinv, jj + 1);
        const float U0l = jj == 0 ? Ul : compr(U0zv, jj - 1);
        const float U0r = jj == 3 ? Ur : compr(U0zv, jj + 1);
        const float vp = compr(vpv, jj), vq = compr(vqv, jj);
        const float tv = m.taoh * vp;
        const float t1 = tv * ((Tin - U0i) + U0z);
        const float t2 = ((((-2.0f * U1z) + U0z) + Tin) - 2.0f * U1i) + U0i;
        const float lap = ((((Tir - 2.0f * Tin) + Til) + U0r) - 2.0f * U0z) + U0l;
        const float t3 = ((m.taoh2 * vq) * vq) * lap;
        const float O = (1.0f / (tv + 1.0f)) * ((t1 - t2) + t3);
        val[jj] = (1.0f - wgt) * T0 + wgt * O;
      }
      const bool rowout = it >= 2 && z < z0 + nrows;
      const unsigned rz = rowm(z);
      bst4s(R2, (rowout && full) ? vst : kDrop, rz, val[0], val[1], val[2], val[3]);
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
        bst1s_(R2, (rowout && !full && wr[jj]) ? vst + 4u * jj : kDrop, rz, val[jj]);
      if (MODE == FR_FWD) {  // Hybrid3 (kernel.cu:184-208)
        const bool srow = BOT ? (z >= NZ - N2 && z < NZ - N2 + g.nfd) : (z >= N2 - g.nfd && z < N2);
        const int sbase = BOT ? 2 * sa + sbb + (z - (NZ - N2)) * sg.mnx - N2
                              : 2 * sa + (N2 - 1 - z) * sg.mnx - N2;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
          bst1s(RS, (rowout && srow && wr[jj]) ? (unsigned)(sbase + c0 + jj) * 4u : kDrop, 0u, val[jj]);
      }
    }
  }
}

template <int NFD, int EPI>
__global__ __launch_bounds__(256) void k_interior(const IntArgs a) {
  interior_body<NFD, EPI>(a, blockIdx.x);
}

// One wave per float4 top/bottom record (frame_tb4_body).
template <int NFD, int MODE>
__global__ __launch_bounds__(256) void k_frame4(const FrameArgs a) {
  const int ri = (int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (ri >= a.nrec) return;
  const int4 rec = a.recs[ri];
  if ((rec.z & 15) == 0) frame_tb4_body<NFD, MODE, false>(a, rec);
  else frame_tb4_body<NFD, MODE, true>(a, rec);
}

// Receiver and adjoint frame launches may carry the strip restore of the
// reconstructed field (FrameArgs::rf): its blocks follow the record blocks and
// fill the wave slots the frame's few long waves leave free.
template <int NFD, int MODE>
__global__ __launch_bounds__(256) void k_frame(const FrameArgs a) {
  if ((MODE == FR_RCV || MODE == FR_ADJ) && a.rbps > 0) {
    const unsigned nfb = ((unsigned)a.nrec + 3u) >> 2;
    if (blockIdx.x >= nfb) {
      strip_restore_body(a.g, a.rf, a.st, a.geo, a.k, blockIdx.x - nfb, (unsigned)a.rbps);
      return;
    }
  }
  frame_body<NFD, MODE>(a, blockIdx.x);
}

// One time step's interior and frame work in one launch.  The two write
// disjoint cells of u(k+1) and only read u(k), u(k-1) and the model, so they
// may run concurrently; the interior kernel waits on memory most of the time
// and the frame kernel on instruction issue, so frame blocks placed among the
// interior blocks fill the interior's idle issue slots.  Blocks go in units
// of 8 (one per XCD): every R-th unit is a frame unit while frame units last,
// so an interior block keeps its own index modulo 8 (its XCD, which its L2
// sharing of model rows relies on); indices past a part's block count exit.
template <int NFD, int EPI, int MODE>
__global__ __launch_bounds__(256) void k_step(const IntArgs ia, const FrameArgs fa, unsigned nint,
                                              unsigned nfr) {
  const unsigned nfu = (nfr + 7u) >> 3, niu = (nint + 7u) >> 3;
  const unsigned R = (nfu + niu) / nfu;
  const unsigned u = blockIdx.x >> 3, s8 = blockIdx.x & 7u;
  const unsigned before = min(nfu, (u + R - 1u) / R);  // frame units ahead of unit u
  if (u % R == 0u && u / R < nfu) {
    const unsigned b = (u / R) * 8u + s8;
    if (b < nfr) frame_body<NFD, MODE>(fa, b);
  } else {
    const unsigned b = (u - before) * 8u + s8;
    if (b < nint) interior_body<NFD, EPI>(ia, b);
  }
}

// ---------------------------------------------------------------------------
// Finite-value watch (not in the reference): *flag stays 0xffffffff while every
// checked field is finite; the first check that meets a NaN or an infinity
// stores its step k there.
__global__ void k_flag_reset(unsigned* flag) {
  if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 0xffffffffu;
}
__global__ __launch_bounds__(256) void k_finite_check(const float4* f, long long n4, int k,
                                                      unsigned* flag) {
  bool bad = false;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (long long)gridDim.x * blockDim.x) {
    const float4 v = f[i];
    bad = bad || !(isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w));
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicCAS(flag, 0xffffffffu, (unsigned)k);
}

// ---------------------------------------------------------------------------
// Small kernels.
__global__ void k_fill(float* p, long long n, float val) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long st = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = val;
}

// u(0) = 0, u(1) = f(0)/2 at the source (kernel.cu:798-806).
__global__ void k_fwd_begin(Grid g, float* f0, float* f1, const ShotGeo* geo, float spike) {
  const int shot = blockIdx.y;
  const long long n = (long long)g.nz * g.pitch;
  float* a = f0 + (long long)shot * g.fstride;
  float* b = f1 + (long long)shot * g.fstride;
  const ShotGeo sg = geo[shot];
  const long long ps = (long long)sg.r_u * g.pitch + g.x0 + sg.r_x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    a[i] = 0.0f;
    b[i] = (i == ps) ? spike : 0.0f;
  }
}

__global__ void k_strip_save(Grid g, const float* f, Strips st, const ShotGeo* geo, int k) {
  const int shot = blockIdx.y;
  const ShotGeo sg = geo[shot];
  const unsigned ss = (unsigned)strip_size(g, sg.mnx);
  const float* F = f + (long long)shot * g.fstride + g.x0;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < ss; e += gridDim.x * blockDim.x) {
    int z, x;
    strip_cell(g, sg.nx, sg.mnx, e, &z, &x);
    strip_put(st, shot, k, (long long)e, F[(long)z * g.pitch + x]);
  }
}

// Maximum |value| of the staged fp32 strips of slot k: one atomic per block
// on the bit patterns (non-negative floats order as their bits).
__global__ __launch_bounds__(256) void k_strip_amax(Grid g, Strips st, const ShotGeo* geo, int k) {
  __shared__ unsigned sh[4];
  const int shot = blockIdx.y;
  const long long ss = strip_size(g, geo[shot].mnx);
  const float* s = st.stage + ((long long)shot * 2 + (k & 1)) * st.ss;
  unsigned mx = 0u;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < ss;
       e += (long long)gridDim.x * blockDim.x)
    mx = max(mx, __float_as_uint(fabsf(s[e])));
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, (unsigned)__shfl_xor((int)mx, o));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicMax(&st.amax[(long long)shot * st.nslot + (k % st.nslot)],
              max(max(sh[0], sh[1]), max(sh[2], sh[3])));
}

// Quantise the staged fp32 strips of slot k with the slot's power-of-two scale.
__global__ void k_strip_finish(Grid g, Strips st, const ShotGeo* geo, int k) {
  const int shot = blockIdx.y;
  const long long ss = strip_size(g, geo[shot].mnx);
  const long long slot = (long long)shot * st.nslot + (k % st.nslot);
  const float sc = pow2_scale(__uint_as_float(st.amax[slot]));
  if (blockIdx.x == 0 && threadIdx.x == 0) st.scale[slot] = sc;
  const float inv = 1.0f / sc;     // exact: sc is a power of two
  const float* s = st.stage + ((long long)shot * 2 + (k & 1)) * st.ss;
  unsigned short* d = (unsigned short*)st.base + slot * st.ss;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < ss;
       e += (long long)gridDim.x * blockDim.x) {
    float x = s[e] * inv;
    d[e] = (st.fmt == 1) ? __half_as_ushort(__float2half_rn(x)) : f2bf(x);
  }
}

__global__ __launch_bounds__(256) void k_strip_restore(Grid g, float* f, Strips st,
                                                       const ShotGeo* geo, int k) {
  strip_restore_body(g, f, st, geo, k, blockIdx.y * gridDim.x + blockIdx.x, gridDim.x);
}

// ---- checkpoint snapshots -------------------------------------------------
// A snapshot slot holds nz*nx (widest shot) elements; shot s uses the first
// nz*nx_s of them, row-major over its own width.

__global__ void k_snap_clear(Snaps sn, int slot, int nshot) {  // a kernel, not a memset: capturable
  const int sh = blockIdx.x * blockDim.x + threadIdx.x;
  if (sh < nshot) sn.amax[(long long)sh * sn.nsnap + slot] = 0u;
}

__global__ void k_snap_amax(Grid g, const float* f, Snaps sn, const ShotGeo* geo, int slot) {
  const int shot = blockIdx.y;
  const int nx = geo[shot].nx;
  const float* F = f + (long long)shot * g.fstride + g.x0;
  const long long n = (long long)g.nz * nx;
  float mx = 0.0f;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(e / nx), x = (int)(e % nx);
    mx = fmaxf(mx, fabsf(F[(long)z * g.pitch + x]));
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  if ((threadIdx.x & 63) == 0) atomicMax(&sn.amax[(long long)shot * sn.nsnap + slot], __float_as_uint(mx));
}

__global__ void k_snap_save(Grid g, const float* f, Snaps sn, const ShotGeo* geo, int slot) {
  const int shot = blockIdx.y;
  const int nx = geo[shot].nx;
  const float* F = f + (long long)shot * g.fstride + g.x0;
  const long long n = (long long)g.nz * nx, stride = (long long)g.nz * g.nx;
  const long long so = (long long)shot * sn.nsnap + slot;
  float inv = 1.0f;
  if (sn.fmt != 0) {
    const float sc = pow2_scale(__uint_as_float(sn.amax[so]));
    if (blockIdx.x == 0 && threadIdx.x == 0) sn.scale[so] = sc;
    inv = 1.0f / sc;  // exact: sc is a power of two
  }
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(e / nx), x = (int)(e % nx);
    const float v = F[(long)z * g.pitch + x];
    if (sn.fmt == 0) {
      ((float*)sn.base)[so * stride + e] = v;
    } else {
      const float y = v * inv;
      ((unsigned short*)sn.base)[so * stride + e] =
          (sn.fmt == 1) ? __half_as_ushort(__float2half_rn(y)) : f2bf(y);
    }
  }
}

__global__ void k_snap_restore(Grid g, float* f, Snaps sn, const ShotGeo* geo, int slot) {
  const int shot = blockIdx.y;
  const int nx = geo[shot].nx;
  float* F = f + (long long)shot * g.fstride + g.x0;
  const long long n = (long long)g.nz * nx, stride = (long long)g.nz * g.nx;
  const long long so = (long long)shot * sn.nsnap + slot;
  const float sc = sn.fmt != 0 ? sn.scale[so] : 1.0f;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(e / nx), x = (int)(e % nx);
    float v;
    if (sn.fmt == 0) {
      v = ((const float*)sn.base)[so * stride + e];
    } else {
      const unsigned short h = ((const unsigned short*)sn.base)[so * stride + e];
      v = (sn.fmt == 1 ? __half2float(__ushort_as_half(h)) : bf2f(h)) * sc;
    }
    F[(long)z * g.pitch + x] = v;
  }
}

// Receiver recording of u(k): rec[shot][k][s], s < n of the shot (stride r.n).
__global__ void k_record(Grid g, Rcv r, const ShotGeo* geo, const float* f, float* rec, int k,
                         int nt) {
  const int shot = blockIdx.y;
  const ShotGeo sg = geo[shot];
  const float* F = f + (long long)shot * g.fstride + g.x0;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < sg.n; s += gridDim.x * blockDim.x)
    rec[((long long)shot * nt + k) * r.n + s] = F[(long)r.s_z * g.pitch + sg.s_l + s * r.ds];
}

// Backward initial state (kernel.cu:859-885).  s0 = u(nt-1) (DBW0),
// s1 = u(nt-2) (DBW1); FW0 = 0 and FW1 = spike field are host arrays there.
__global__ void k_bwd_begin(Grid g, const float* s0, const float* s1, float* q0, float* q1,
                            float* su, float* sq, float* i1, float* i2, const ShotGeo* geo,
                            float spike, int icompen) {
  const int shot = blockIdx.y;
  const long long base = (long long)shot * g.fstride;
  const ShotGeo sg = geo[shot];
  const long long ps = (long long)sg.r_u * g.pitch + g.x0 + sg.r_x;
  const long long n = (long long)g.nz * g.pitch;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long p = base + i;
    const float BW0 = s0[p], BW1 = s1[p];
    const float FW1 = (i == ps) ? spike : 0.0f, FW0 = 0.0f;
    q0[p] = 0.0f;
    q1[p] = 0.0f;
    if (icompen == 1) {
      const float BW2 = BW0 + BW1;
      const float FW2 = FW0 + FW1;
      su[p] = BW2;
      sq[p] = FW2;
      i1[p] = FW2 * BW2 + FW0 * BW0;
    } else {
      su[p] = 0.0f;
      sq[p] = 0.0f;
      i1[p] = FW1 * BW1 + FW0 * BW0;
    }
    i2[p] = BW1 * BW1 + BW0 * BW0;
  }
}

// max |I2| over the interior, per shot (kernel.cu:963-970).
__global__ void k_maxabs(Grid g, const ShotGeo* geo, const float* i2, unsigned* out) {
  const int shot = blockIdx.y;
  const int mnx = geo[shot].mnx;
  const float* I = i2 + (long long)shot * g.fstride + g.x0;
  const long long n = (long long)g.mnz * mnx;
  float mx = 0.0f;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    int z = g.n2 + (int)(e / mnx), x = g.n2 + (int)(e % mnx);
    mx = fmaxf(mx, fabsf(I[(long)z * g.pitch + x]));
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  if ((threadIdx.x & 63) == 0) atomicMax(&out[shot], __float_as_uint(mx));
}

// Per-shot image post (kernel.cu:935-979).  up/down: [shot][mnz][g.mnx].
__global__ void k_post(Grid g, const ShotGeo* geo, const float* v, long long mstride,
                       const float* i1, const float* i2, const unsigned* amax, float vmax,
                       float whitecoe, float* up, float* down) {
  const int shot = blockIdx.y;
  const ShotGeo sg = geo[shot];
  const int mnx = sg.mnx, nx = sg.nx;
  const float* M = i1 + (long long)shot * g.fstride + g.x0;
  const float* I = i2 + (long long)shot * g.fstride + g.x0;
  const float* Vv = v + (long long)sg.win * mstride + g.x0;
  const long long n = (long long)g.mnz * mnx;
  const float stable = __uint_as_float(amax[shot]) * whitecoe;
  const int P = g.pitch, n2 = g.n2;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int zi = (int)(e / mnx), xi = (int)(e % mnx);
    const int i = n2 + zi, j = n2 + xi;
    int i1_ = i - 1, i2_ = i + 1, j1 = j - 1, j2 = j + 1;
    if (i1_ < n2) i1_ = 2 * n2 - i1_;
    if (j1 < n2) j1 = 2 * n2 - j1;
    if (i2_ >= g.nz - n2) i2_ = 2 * (g.nz - n2 - 1) - i2_;
    if (j2 >= nx - n2) j2 = 2 * (nx - n2 - 1) - j2;
    const float sum = M[(long)i * P + j2] + M[(long)i * P + j1] + M[(long)i2_ * P + j] +
                      M[(long)i1_ * P + j] - 4.0f * M[(long)i * P + j];
    const float vv = Vv[(long)i * P + j];
    const double t = -1.0 * (double)sum * (double)vv * (double)vv / (double)(vmax * vmax);
    const long long o = ((long long)shot * g.mnz + zi) * g.mnx + xi;
    if (up) up[o] = (float)t;
    if (down) down[o] = I[(long)i * P + j] + stable;
  }
}

// Running stack in shot order (kernel.cu:1026-1039) onto an image of gw
// columns: shot s covers columns x_img .. x_img+mnx-1.
__global__ void k_stack_add(Grid g, const ShotGeo* geo, int gw, const float* up, const float* down,
                            float* s1, float* s2, int nshot) {
  const long long nI = (long long)g.mnz * gw;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nI;
       e += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(e / gw), X = (int)(e % gw);
    float a = s1[e], b = s2[e];
    for (int s = 0; s < nshot; ++s) {
      const int x = X - geo[s].x_img;
      if (x < 0 || x >= geo[s].mnx) continue;
      const long long o = ((long long)s * g.mnz + z) * g.mnx + x;
      a += up[o];
      b += down[o];
    }
    s1[e] = a;
    s2[e] = b;
  }
}

__global__ void k_refl_scatter(Grid g, const ShotGeo* geo, int gw, const float* img, float* dst,
                               long long rstride) {
  const int shot = blockIdx.y;
  const ShotGeo sg = geo[shot];
  float* D = dst + (long long)shot * rstride;
  const long long n = (long long)g.nz * g.pitch;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(i / g.pitch), x = (int)(i % g.pitch) - g.x0;
    float v = 0.0f;
    const int X = x - g.n2 + sg.x_img;
    if (z >= g.n2 && z < g.nz - g.n2 && x >= g.n2 && x < sg.nx - g.n2 && X >= 0 && X < gw)
      v = img[(long long)(z - g.n2) * gw + X];
    D[i] = v;
  }
}

__global__ void k_grad_stack(Grid g, const ShotGeo* geo, int gw, const float* G, const float* pm,
                             long long mstride, int nshot, float* grad) {
  const long long n = (long long)g.mnz * gw;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(e / gw), X = (int)(e % gw);
    float acc = 0.0f;
    for (int s = 0; s < nshot; ++s) {
      const int x = X - geo[s].x_img;
      if (x < 0 || x >= geo[s].mnx) continue;
      const long long pc = (long long)(z + g.n2) * g.pitch + g.x0 + g.n2 + x;
      acc = acc + G[(long long)s * g.fstride + pc] / pm[(long long)geo[s].win * mstride + pc];
    }
    grad[e] = acc;
  }
}

// FP64 dot product of two float arrays in two passes with a fixed reduction
// order (the result does not depend on scheduling): each of the DOT_BLOCKS
// blocks sums a grid-stride slice in double and reduces it in LDS; one block
// then adds the partial sums in a fixed tree.
constexpr int DOT_BLOCKS = 1024;
__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  return sh[0];
}
// Vector v = blockIdx.y of a batch starts at v*stride in a and b; its
// partial sums fill row v of part (gridDim.x = DOT_BLOCKS entries), so every
// vector is reduced in the same order as a single-vector launch.
__global__ __launch_bounds__(256) void k_dot64_part(const float* a, const float* b, long long n,
                                                    long long stride, double* part) {
  __shared__ double sh[256];
  a += (long long)blockIdx.y * stride;
  b += (long long)blockIdx.y * stride;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    acc += (double)a[i] * (double)b[i];
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) part[(long long)blockIdx.y * gridDim.x + blockIdx.x] = t;
}
// One block per vector: out[v] = the fixed-tree sum of row v of part.
__global__ __launch_bounds__(256) void k_dot64_final(const double* part, int np, double* out) {
  __shared__ double sh[256];
  part += (long long)blockIdx.x * np;
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// y = y + a*x (product, then sum)
__global__ void k_axpy(float* y, const float* x, float a, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = y[i] + a * x[i];
}

// out[c][r] = in[r][c] for an rows x cols matrix, through 32 x 32 LDS tiles so
// that both the reads and the writes are coalesced.
__global__ __launch_bounds__(256) void k_transpose(const float* in, float* out, int rows, int cols) {
  __shared__ float tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + j, c = c0 + tx;
    if (r < rows && c < cols) tile[j][tx] = in[(long long)r * cols + c];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + j, r = r0 + tx;
    if (r < rows && c < cols) out[(long long)c * rows + r] = tile[tx][j];
  }
}

__global__ void k_stack_finish(int nI, const float* s1, const float* s2, int nrec, int inorm,
                               float* img, float* ill) {  // kernel.cu:1042-1059
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nI; e += gridDim.x * blockDim.x) {
    float a = s1[e] / nrec;
    float b = s2[e] / nrec;
    // a cell that no shot of a ragged batch covers (both sums 0) stays 0
    // instead of 0/0; every other cell is divided as in the reference
    if (inorm == 1 && !(s1[e] == 0.0f && s2[e] == 0.0f)) a = a / b;
    img[e] = a;
    ill[e] = b;
  }
}

// ---------------------------------------------------------------------------
// launchers
#define RTM_NFD_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

int interior_queue_rows(int epi, int nfd) {
  const bool pd1 = epi == EPI_RCV || epi == EPI_RCVI || epi == EPI_BORN || epi == EPI_BORN64 ||
                   epi == EPI_ADJ || epi == EPI_ADJT;
  return 2 * nfd + 1 + (pd1 ? 1 : 2);  // EpiTraits<EPI>::PD
}

hipError_t launch_interior(int epi, int nfd, IntArgs a, int nshot, hipStream_t s) {
  // rows per wave: the requested segment rounded to a multiple of the queue length
  const int ql = interior_queue_rows(epi, nfd);
  const int mult = a.zs > ql ? (a.zs + ql / 2) / ql : 1;
  a.zs = mult * ql;
  const int nseg = (a.g.mnz + a.zs - 1) / a.zs;
  a.nshot = nshot;
  // four shots per block on shared rows, or four row segments of one shot
  // per block when the launch has fewer than four shots (k_interior)
  dim3 grid(nshot >= 4 ? a.nxt * nseg * ((nshot + 3) / 4) : a.nxt * ((nseg + 3) / 4) * nshot),
      block(256);
#define CASE(N)                                                                             \
  case N:                                                                                   \
    if (epi == EPI_FWD) hipLaunchKernelGGL((k_interior<N, EPI_FWD>), grid, block, 0, s, a);     \
    else if (epi == EPI_FWD64) hipLaunchKernelGGL((k_interior<N, EPI_FWD64>), grid, block, 0, s, a); \
    else if (epi == EPI_SRC) hipLaunchKernelGGL((k_interior<N, EPI_SRC>), grid, block, 0, s, a); \
    else if (epi == EPI_BORN) hipLaunchKernelGGL((k_interior<N, EPI_BORN>), grid, block, 0, s, a); \
    else if (epi == EPI_BORN64) hipLaunchKernelGGL((k_interior<N, EPI_BORN64>), grid, block, 0, s, a); \
    else if (epi == EPI_ADJ) hipLaunchKernelGGL((k_interior<N, EPI_ADJ>), grid, block, 0, s, a); \
    else if (epi == EPI_ADJT) hipLaunchKernelGGL((k_interior<N, EPI_ADJT>), grid, block, 0, s, a); \
    else if (epi == EPI_SRCI) hipLaunchKernelGGL((k_interior<N, EPI_SRCI>), grid, block, 0, s, a); \
    else if (epi == EPI_RCVI) hipLaunchKernelGGL((k_interior<N, EPI_RCVI>), grid, block, 0, s, a); \
    else hipLaunchKernelGGL((k_interior<N, EPI_RCV>), grid, block, 0, s, a);               \
    break;
  switch (nfd) {
    RTM_NFD_CASES(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
  return hipGetLastError();
}

int frame_segment_rows(int nfd) {
  const int nb = nfd > 2 ? nfd : 2;
  return 2 * (nb + nfd + 1 + 2) - 2;  // two register-queue cycles per side run
}
// Float4 windows of the top and bottom bands (k_frame4): first column n2,
// 252 apart, writing [cs+1, min(cs+253, nx-n2-1)).  None when the band's edge
// cells do not reach one window or n2 leaves no room for the column halo.
static int tb4_windows(const Grid& g, int nx) {
  const int wend = nx - g.n2 - 1;
  if (g.n2 < 12 || wend <= g.n2 + 1) return 0;
  return (wend - (g.n2 + 1) + 251) / 252;
}

void frame_records(const Grid& g, int nx, int shot, int seg_tb, int seg, int kind, int4* out,
                   int* count) {
  int n = 0;
  auto put = [&](int z0, int cx0, int band, int rows, int ncol) {
    if (out) out[n] = make_int4(z0, cx0, band | (rows << 4) | (ncol << 20), shot);
    ++n;
  };
  const int nw = kind == FRAME_ALL ? 0 : tb4_windows(g, nx);
  // top band, bottom band, in runs of seg_tb rows (the float4 records: one
  // run of the whole band, each wave writing 252 columns)
  const int run = kind == FRAME_VEC ? g.n2 : seg_tb;
  for (int b = 0; b < 2; ++b) {
    const int zb = b == 0 ? 0 : g.nz - g.n2, ze = b == 0 ? g.n2 : g.nz;
    for (int z0 = zb; z0 < ze; z0 += run) {
      const int rows = std::min(run, ze - z0);
      if (nw == 0) {
        if (kind != FRAME_VEC)
          for (int cx0 = -1; cx0 + 1 < nx; cx0 += 62) put(z0, cx0, b, rows, 0);
      } else if (kind == FRAME_VEC) {
        for (int w = 0; w < nw; ++w) put(z0, g.n2 + 252 * w, b, rows, 0);
      } else {  // the columns left to the scalar kernel: [0, n2+1) and [nx-n2-1, nx)
        for (int c = 0; c < g.n2 + 1; c += 62) put(z0, c - 1, b, rows, std::min(62, g.n2 + 1 - c));
        for (int c = nx - g.n2 - 1; c < nx; c += 62) put(z0, c - 1, b, rows, std::min(62, nx - c));
      }
    }
  }
  if (kind == FRAME_VEC) {
    *count = n;
    return;
  }
  // side bands: both in one wave per row run (band 4) when each fits in half
  // a wave with its inner neighbour and the halves do not overlap
  const bool pack = g.n2 <= 30 && nx >= 64;
  for (int z0 = g.n2; z0 < g.nz - g.n2; z0 += seg) {
    const int rows = std::min(seg, g.nz - g.n2 - z0);
    if (pack) {
      put(z0, 0, 4, rows, 0);
      continue;
    }
    for (int cx0 = -1; cx0 + 1 < g.n2; cx0 += 62) put(z0, cx0, 2, rows, 0);
    for (int cx0 = nx - g.n2 - 1; cx0 + 1 < nx; cx0 += 62) put(z0, cx0, 3, rows, 0);
  }
  *count = n;
}

bool frame4_supported(int nfd) { return nfd >= 1 && nfd <= 6; }

hipError_t launch_frame4(int mode, int nfd, FrameArgs a, int nshot, int nrec, hipStream_t s) {
  a.nshot = nshot;
  a.nrec = nrec;
  if (nrec == 0) return hipSuccess;
  dim3 grid((nrec + 3) / 4), block(256);  // one wave per record
#define CASE(N)                                                                          \
  case N:                                                                                \
    if (mode == FR_FWD) hipLaunchKernelGGL((k_frame4<N, FR_FWD>), grid, block, 0, s, a); \
    else if (mode == FR_RCV) hipLaunchKernelGGL((k_frame4<N, FR_RCV>), grid, block, 0, s, a); \
    else return hipErrorInvalidValue;                                                    \
    break;
  switch (nfd) {
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6)
    default: return hipErrorInvalidValue;
  }
#undef CASE
  return hipGetLastError();
}

hipError_t launch_frame(int mode, int nfd, FrameArgs a, int nshot, int nrec, hipStream_t s) {
  a.nshot = nshot;
  a.nrec = nrec;
  if (mode != FR_RCV && mode != FR_ADJ) a.rbps = 0;
  // one wave per (region run, shot) record, then the restore blocks
  dim3 grid((nrec + 3) / 4 + (a.rbps > 0 ? a.rbps * nshot : 0)), block(256);
#define CASE(N)                                                                          \
  case N:                                                                                \
    if (mode == FR_FWD) hipLaunchKernelGGL((k_frame<N, FR_FWD>), grid, block, 0, s, a);  \
    else if (mode == FR_BORN) hipLaunchKernelGGL((k_frame<N, FR_BORN>), grid, block, 0, s, a); \
    else if (mode == FR_ADJ) hipLaunchKernelGGL((k_frame<N, FR_ADJ>), grid, block, 0, s, a); \
    else hipLaunchKernelGGL((k_frame<N, FR_RCV>), grid, block, 0, s, a);                 \
    break;
  switch (nfd) {
    RTM_NFD_CASES(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
  return hipGetLastError();
}

// The interior grid of launch_interior (rows per wave rounded to the queue
// length) and one wave per frame record, in one k_step launch; only the
// forward and the split receiver step are fused (the other kinds keep two
// launches).
bool step_fusable(int epi, int mode) {
  return (epi == EPI_FWD && mode == FR_FWD) || (epi == EPI_RCVI && mode == FR_RCV);
}

hipError_t launch_step(int epi, int mode, int nfd, IntArgs ia, FrameArgs fa, int nshot, int nrec,
                       hipStream_t s) {
  if (!step_fusable(epi, mode)) return hipErrorInvalidValue;
  const int ql = interior_queue_rows(epi, nfd);
  const int mult = ia.zs > ql ? (ia.zs + ql / 2) / ql : 1;
  ia.zs = mult * ql;
  const int nseg = (ia.g.mnz + ia.zs - 1) / ia.zs;
  ia.nshot = nshot;
  const unsigned nint = (unsigned)(nshot >= 4 ? ia.nxt * nseg * ((nshot + 3) / 4)
                                              : ia.nxt * ((nseg + 3) / 4) * nshot);
  fa.nshot = nshot;
  fa.nrec = nrec;
  const unsigned nfr = (unsigned)((nrec + 3) / 4);
  if (nfr == 0 || nint == 0) return hipErrorInvalidValue;
  const unsigned nfu = (nfr + 7u) / 8u, niu = (nint + 7u) / 8u;
  dim3 grid((nfu + niu) * 8u), block(256);
#define CASE(N)                                                                               \
  case N:                                                                                     \
    if (epi == EPI_FWD) hipLaunchKernelGGL((k_step<N, EPI_FWD, FR_FWD>), grid, block, 0, s, ia, fa, nint, nfr); \
    else hipLaunchKernelGGL((k_step<N, EPI_RCVI, FR_RCV>), grid, block, 0, s, ia, fa, nint, nfr); \
    break;
  switch (nfd) {
    RTM_NFD_CASES(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
  return hipGetLastError();
}

static inline unsigned blocks_for(long long n, int per = 256, unsigned cap = 4096) {
  long long b = (n + per - 1) / per;
  if (b < 1) b = 1;
  return (unsigned)(b > cap ? cap : b);
}

// Neighbour-tap planes: for cell (z, c) of the padded layout (c the storage
// column) and l = 1..nfd, the tap-l coefficient of the rows of (z-l, c),
// (z+l, c), (z, c-l), (z, c+l) -- 0 off the grid.  A row holds [c0' (double) |
// c_1 .. c_NFD | zeros], so tap l sits at float 1 + l of the row.
__global__ void k_tap_planes(Grid g, const unsigned* ko, const float* ctab, int nfd, float* tp) {
  const long long n = (long long)g.nz * g.pitch;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(i / g.pitch), c = (int)(i - (long long)z * g.pitch);
    for (int l = 1; l <= nfd; ++l) {
      const int zz[4] = {z - l, z + l, z, z};
      const int cc[4] = {c, c, c - l, c + l};
      for (int d = 0; d < 4; ++d) {
        float val = 0.0f;
        if (zz[d] >= 0 && zz[d] < g.nz && cc[d] >= 0 && cc[d] < g.pitch)
          val = ctab[ko[(long long)zz[d] * g.pitch + cc[d]] / 4u + 1 + l];
        tp[(long long)(4 * (l - 1) + d) * g.fstride + i] = val;
      }
    }
  }
}
hipError_t launch_tap_planes(const Grid& g, const unsigned* ko, const float* ctab, int nfd,
                             float* tp, hipStream_t s) {
  hipLaunchKernelGGL(k_tap_planes, dim3(blocks_for((long long)g.nz * g.pitch, 256, 2048)),
                     dim3(256), 0, s, g, ko, ctab, nfd, tp);
  return hipGetLastError();
}
hipError_t launch_flag_reset(unsigned* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_flag_reset, dim3(1), dim3(64), 0, s, flag);
  return hipGetLastError();
}
hipError_t launch_finite_check(const float* f, long long n, int k, unsigned* flag, hipStream_t s) {
  const long long n4 = n / 4;
  hipLaunchKernelGGL(k_finite_check, dim3(blocks_for(n4, 256, 2048)), dim3(256), 0, s,
                     reinterpret_cast<const float4*>(f), n4, k, flag);
  return hipGetLastError();
}
hipError_t launch_fill(float* p, long long n, float val, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, s, p, n, val);
  return hipGetLastError();
}
hipError_t launch_fwd_begin(const Grid& g, float* f0, float* f1, const ShotGeo* geo, float spike,
                            int nshot, hipStream_t s) {
  hipLaunchKernelGGL(k_fwd_begin, dim3(blocks_for((long long)g.nz * g.pitch, 256, 1024), nshot),
                     dim3(256), 0, s, g, f0, f1, geo, spike);
  return hipGetLastError();
}
static inline long long strip_size_host(const Grid& g) { return 2LL * g.nfd * (g.mnz + g.mnx); }
int strip_restore_blocks(const Grid& g) { return (int)blocks_for(strip_size_host(g), 1024); }
hipError_t launch_strip_save(const Grid& g, const float* f, Strips st, const ShotGeo* geo, int k,
                             int nshot, hipStream_t s) {
  hipLaunchKernelGGL(k_strip_save, dim3(blocks_for(strip_size_host(g)), nshot), dim3(256), 0, s, g,
                     f, st, geo, k);
  return hipGetLastError();
}
hipError_t launch_strip_finish(const Grid& g, Strips st, const ShotGeo* geo, int k, int nshot,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_strip_amax, dim3(blocks_for(strip_size_host(g), 256, 64), nshot), dim3(256),
                     0, s, g, st, geo, k);
  hipLaunchKernelGGL(k_strip_finish, dim3(blocks_for(strip_size_host(g)), nshot), dim3(256), 0, s,
                     g, st, geo, k);
  return hipGetLastError();
}
hipError_t launch_strip_restore(const Grid& g, float* f, Strips st, const ShotGeo* geo, int k,
                                int nshot, hipStream_t s) {
  hipLaunchKernelGGL(k_strip_restore, dim3(blocks_for(strip_size_host(g)), nshot), dim3(256), 0, s,
                     g, f, st, geo, k);
  return hipGetLastError();
}
hipError_t launch_snap_save(const Grid& g, const float* f, Snaps sn, const ShotGeo* geo, int slot,
                            int nshot, hipStream_t s) {
  const long long n = (long long)g.nz * g.nx;
  if (sn.fmt != 0) {
    hipLaunchKernelGGL(k_snap_clear, dim3(blocks_for(nshot)), dim3(256), 0, s, sn, slot, nshot);
    hipLaunchKernelGGL(k_snap_amax, dim3(blocks_for(n, 256, 512), nshot), dim3(256), 0, s, g, f, sn,
                       geo, slot);
  }
  hipLaunchKernelGGL(k_snap_save, dim3(blocks_for(n, 256, 1024), nshot), dim3(256), 0, s, g, f, sn,
                     geo, slot);
  return hipGetLastError();
}
hipError_t launch_snap_restore(const Grid& g, float* f, Snaps sn, const ShotGeo* geo, int slot,
                               int nshot, hipStream_t s) {
  const long long n = (long long)g.nz * g.nx;
  hipLaunchKernelGGL(k_snap_restore, dim3(blocks_for(n, 256, 1024), nshot), dim3(256), 0, s, g, f,
                     sn, geo, slot);
  return hipGetLastError();
}
hipError_t launch_record(const Grid& g, const Rcv& r, const ShotGeo* geo, const float* f,
                         float* rec, int k, int nt, int nshot, hipStream_t s) {
  hipLaunchKernelGGL(k_record, dim3(blocks_for(r.n), nshot), dim3(256), 0, s, g, r, geo, f, rec, k,
                     nt);
  return hipGetLastError();
}
hipError_t launch_bwd_begin(const Grid& g, const float* s0, const float* s1, float* q0, float* q1,
                            float* su, float* sq, float* i1, float* i2, const ShotGeo* geo,
                            float spike, int icompen, int nshot, hipStream_t s) {
  hipLaunchKernelGGL(k_bwd_begin, dim3(blocks_for((long long)g.nz * g.pitch, 256, 1024), nshot),
                     dim3(256), 0, s, g, s0, s1, q0, q1, su, sq, i1, i2, geo, spike, icompen);
  return hipGetLastError();
}
hipError_t launch_maxabs(const Grid& g, const ShotGeo* geo, const float* i2, unsigned* out,
                         int nshot, hipStream_t s) {
  hipLaunchKernelGGL(k_maxabs, dim3(blocks_for((long long)g.mnz * g.mnx, 256, 256), nshot),
                     dim3(256), 0, s, g, geo, i2, out);
  return hipGetLastError();
}
hipError_t launch_post(const Grid& g, const ShotGeo* geo, const float* v, long long mstride,
                       const float* i1, const float* i2, const unsigned* amax, float vmax,
                       float whitecoe, float* up, float* down, int nshot, hipStream_t s) {
  hipLaunchKernelGGL(k_post, dim3(blocks_for((long long)g.mnz * g.mnx, 256, 1024), nshot),
                     dim3(256), 0, s, g, geo, v, mstride, i1, i2, amax, vmax, whitecoe, up, down);
  return hipGetLastError();
}
hipError_t launch_stack_add(const Grid& g, const ShotGeo* geo, int gw, const float* up,
                            const float* down, float* s1, float* s2, int nshot, hipStream_t s) {
  hipLaunchKernelGGL(k_stack_add, dim3(blocks_for((long long)g.mnz * gw)), dim3(256), 0, s, g, geo,
                     gw, up, down, s1, s2, nshot);
  return hipGetLastError();
}
hipError_t launch_refl_scatter(const Grid& g, const ShotGeo* geo, int gw, const float* img,
                               float* dst, long long rstride, int nshot, hipStream_t s) {
  const int ny = rstride == 0 ? 1 : nshot;
  hipLaunchKernelGGL(k_refl_scatter, dim3(blocks_for((long long)g.nz * g.pitch), ny), dim3(256), 0,
                     s, g, geo, gw, img, dst, rstride);
  return hipGetLastError();
}

hipError_t launch_grad_stack(const Grid& g, const ShotGeo* geo, int gw, const float* G,
                             const float* pm, long long mstride, int nshot, float* grad,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_grad_stack, dim3(blocks_for((long long)g.mnz * gw)), dim3(256), 0, s, g, geo,
                     gw, G, pm, mstride, nshot, grad);
  return hipGetLastError();
}

int dot64_part_size() { return DOT_BLOCKS; }

hipError_t launch_dot64(const float* a, const float* b, long long n, long long stride, int nvec,
                        double* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_dot64_part, dim3(DOT_BLOCKS, nvec), dim3(256), 0, s, a, b, n, stride,
                     part);
  hipLaunchKernelGGL(k_dot64_final, dim3(nvec), dim3(256), 0, s, part, DOT_BLOCKS, out);
  return hipGetLastError();
}

hipError_t launch_axpy(float* y, const float* x, float a, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(blocks_for(n)), dim3(256), 0, s, y, x, a, n);
  return hipGetLastError();
}

hipError_t launch_transpose(const float* in, float* out, int rows, int cols, hipStream_t s) {
  hipLaunchKernelGGL(k_transpose, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(256), 0, s, in, out,
                     rows, cols);
  return hipGetLastError();
}

hipError_t launch_stack_finish(int nI, const float* s1, const float* s2, int nrec, int inorm,
                               float* img, float* illum, hipStream_t s) {
  hipLaunchKernelGGL(k_stack_finish, dim3(blocks_for(nI)), dim3(256), 0, s, nI, s1, s2, nrec, inorm,
                     img, illum);
  return hipGetLastError();
}

}  // namespace rtm
