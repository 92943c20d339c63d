// rtm_segy.cpp -- SEG-Y rev 1 reading and writing for the RTM driver.
//
// Writing follows the reference's WriteSGY (SGYWrite.cpp:3-55): the 3200-byte
// textual header, the 400-byte binary header and the first 240-byte trace
// header are taken from a template file; the binary header gets ns and dt,
// and every trace header is the template's with offset (bytes 37-40),
// sx (73-76), sy (77-80), gx (81-84) and gy (85-88) replaced.  Samples are
// encoded in the template's format code (1 IBM float, 2 int32, 3 int16,
// 5 IEEE float), all big-endian.  IBM conversion truncates toward zero in the
// fraction, as the reference's codec does (segy.cpp:552-601).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../../include/rtm_host.h"
#include "rtm_host_internal.h"

namespace {

constexpr int kTextBytes = 3200, kBinBytes = 400, kTraceHeaderBytes = 240;
constexpr int kBinDt = 16, kBinNs = 20, kBinFormat = 24;  // byte offsets in the binary header
constexpr int kTrOffset = 36, kTrSx = 72, kTrSy = 76, kTrGx = 80, kTrGy = 84;

inline void put_be16(unsigned char* p, int v) {
  const uint16_t u = (uint16_t)(int16_t)v;
  p[0] = (unsigned char)(u >> 8);
  p[1] = (unsigned char)u;
}
inline void put_be32(unsigned char* p, uint32_t u) {
  p[0] = (unsigned char)(u >> 24);
  p[1] = (unsigned char)(u >> 16);
  p[2] = (unsigned char)(u >> 8);
  p[3] = (unsigned char)u;
}
inline int get_be16(const unsigned char* p) { return (int)(int16_t)((p[0] << 8) | p[1]); }
inline uint32_t get_be32(const unsigned char* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

// IEEE single -> IBM System/360 single (sign, excess-64 base-16 exponent,
// 24-bit fraction).  The IEEE significand (with its hidden bit) is aligned
// to a base-16 exponent by shifting; bits shifted out are dropped.
uint32_t ieee_to_ibm(float v) {
  uint32_t x;
  memcpy(&x, &v, 4);
  if ((x & 0x7fffffffu) == 0) return x;  // +-0 kept as is
  uint32_t sign = x & 0x80000000u;
  int e2 = (int)((x >> 23) & 0xff) - 127;
  uint32_t frac = ((x & 0x007fffffu) << 1) | 0x01000000u;  // 1.f as a 25-bit integer
  int e16;
  if (e2 >= 0) {
    frac <<= (e2 & 3);
    e16 = e2 >> 2;
  } else {
    frac >>= ((-e2) & 3);
    e16 = -((-e2) >> 2);
  }
  if (frac & 0x0f000000u) {  // keep 24 fraction bits
    frac >>= 4;
    e16 += 1;
  }
  e16 += 64;
  if (e16 > 127) return sign | 0x7fffffffu;   // overflow: largest magnitude
  if (e16 < 0) return sign;                    // underflow: signed zero
  return sign | ((uint32_t)e16 << 24) | frac;
}

float ibm_to_ieee(uint32_t x) {
  if ((x & 0x7fffffffu) == 0) return 0.0f;
  uint32_t sign = x & 0x80000000u;
  int e = (int)((x >> 24) & 0x7f) - 64;
  uint32_t frac = x & 0x00ffffffu;
  int e2 = (e >= 0) ? (e << 2) : -((-e) << 2);
  e2 -= 1;
  if (frac) {
    while ((frac & 0x00800000u) == 0) {
      frac <<= 1;
      e2 -= 1;
    }
  }
  frac &= 0x007fffffu;
  uint32_t out = sign;
  e2 += 127;
  if (e2 >= 255) out |= 0x7f7fffffu;
  else if (e2 > 0) out |= ((uint32_t)e2 << 23) | frac;
  float f;
  memcpy(&f, &out, 4);
  return f;
}

void encode_sample(unsigned char* p, float v, int format) {
  switch (format) {
    case 1: put_be32(p, ieee_to_ibm(v)); break;
    case 2: put_be32(p, (uint32_t)(int32_t)v); break;
    case 3: put_be16(p, (int)v); break;
    default: {
      uint32_t u;
      memcpy(&u, &v, 4);
      put_be32(p, u);
    }
  }
}
float decode_sample(const unsigned char* p, int format) {
  switch (format) {
    case 1: return ibm_to_ieee(get_be32(p));
    case 2: return (float)(int32_t)get_be32(p);
    case 3: return (float)get_be16(p);
    default: {
      uint32_t u = get_be32(p);
      float f;
      memcpy(&f, &u, 4);
      return f;
    }
  }
}

}  // namespace

extern "C" {

void rtmh_float_to_ibm(const float* in, unsigned char* out, long long n) {
  for (long long i = 0; i < n; ++i) put_be32(out + 4 * i, ieee_to_ibm(in[i]));
}
void rtmh_ibm_to_float(const unsigned char* in, float* out, long long n) {
  for (long long i = 0; i < n; ++i) out[i] = ibm_to_ieee(get_be32(in + 4 * i));
}

int rtmh_write_segy(const char* tpl, const char* path, const float* data, int ntr, int ns,
                    int dt_units, const float* sx, const float* sy, float rx, float ry,
                    const float* dsr) {
  FILE* t = fopen(tpl, "rb");
  if (!t) return rtmh::set_error(rtmh::RTMH_E_IO, "cannot open SEG-Y template %s", tpl);
  unsigned char text[kTextBytes], bin[kBinBytes], trh[kTraceHeaderBytes];
  const bool ok = fread(text, 1, kTextBytes, t) == kTextBytes && fread(bin, 1, kBinBytes, t) == kBinBytes &&
                  fread(trh, 1, kTraceHeaderBytes, t) == kTraceHeaderBytes;
  fclose(t);
  if (!ok) return rtmh::set_error(rtmh::RTMH_E_FORMAT, "%s: short SEG-Y template", tpl);
  const int format = get_be16(bin + kBinFormat);
  put_be16(bin + kBinNs, ns);
  // dt in microseconds when given in seconds (< 1), else the value is taken
  // as milliseconds (set_segydt semantics, segy.cpp:530-538).
  const float dtf = (float)dt_units;
  const double scale = (dtf < 1.0) ? 1000000. : 1000.;
  put_be16(bin + kBinDt, (int)(scale * dtf));
  const int nb = (format == 3) ? 2 : 4;
  FILE* f = fopen(path, "wb");
  if (!f) return rtmh::set_error(rtmh::RTMH_E_IO, "cannot create %s", path);
  fwrite(text, 1, kTextBytes, f);
  fwrite(bin, 1, kBinBytes, f);
  std::vector<unsigned char> buf((size_t)nb * ns);
  for (int i = 0; i < ntr; ++i) {
    put_be32(trh + kTrOffset, (uint32_t)(int32_t)dsr[i]);
    put_be32(trh + kTrSx, (uint32_t)(int32_t)sx[i]);
    put_be32(trh + kTrSy, (uint32_t)(int32_t)sy[i]);
    put_be32(trh + kTrGx, (uint32_t)(int32_t)rx);
    put_be32(trh + kTrGy, (uint32_t)(int32_t)ry);
    fwrite(trh, 1, kTraceHeaderBytes, f);
    const float* tr = data + (size_t)i * ns;
    for (int s = 0; s < ns; ++s) encode_sample(&buf[(size_t)s * nb], tr[s], format);
    fwrite(buf.data(), 1, buf.size(), f);
  }
  const bool bad = ferror(f) != 0;
  fclose(f);
  if (bad) return rtmh::set_error(rtmh::RTMH_E_IO, "write error on %s", path);
  return rtmh::RTMH_OK;
}

int rtmh_read_segy(const char* path, int* ntr, int* ns, int* format, float* data, long long cap) {
  FILE* f = fopen(path, "rb");
  if (!f) return rtmh::set_error(rtmh::RTMH_E_IO, "cannot open %s", path);
  unsigned char text[kTextBytes], bin[kBinBytes];
  if (fread(text, 1, kTextBytes, f) != kTextBytes || fread(bin, 1, kBinBytes, f) != kBinBytes) {
    fclose(f);
    return rtmh::set_error(rtmh::RTMH_E_FORMAT, "%s: short SEG-Y header", path);
  }
  const int fmt = get_be16(bin + kBinFormat);
  const int nb = (fmt == 3) ? 2 : 4;
  fseek(f, 0, SEEK_END);
  const long long size = ftell(f);
  fseek(f, kTextBytes + kBinBytes, SEEK_SET);
  // samples per trace from the binary header (bytes 3221-3222), falling back
  // to the first trace header (bytes 115-116).  WriteSGY patches only the
  // binary header's ns, so the trace headers may carry the template's value.
  unsigned char trh[kTraceHeaderBytes];
  int n = get_be16(bin + kBinNs);
  if (n <= 0 && fread(trh, 1, kTraceHeaderBytes, f) == kTraceHeaderBytes) n = get_be16(trh + 114);
  if (n <= 0) {
    fclose(f);
    return rtmh::set_error(rtmh::RTMH_E_FORMAT, "%s: no sample count", path);
  }
  const long long per = kTraceHeaderBytes + (long long)nb * n;
  const int nt = (int)((size - kTextBytes - kBinBytes) / per);
  *ntr = nt;
  *ns = n;
  *format = fmt;
  if (data) {
    if ((long long)nt * n > cap) {
      fclose(f);
      return rtmh::set_error(rtmh::RTMH_E_ARG, "%s: %d x %d samples exceed capacity", path, nt, n);
    }
    std::vector<unsigned char> buf((size_t)nb * n);
    fseek(f, kTextBytes + kBinBytes, SEEK_SET);
    for (int i = 0; i < nt; ++i) {
      if (fread(trh, 1, kTraceHeaderBytes, f) != kTraceHeaderBytes ||
          fread(buf.data(), 1, buf.size(), f) != buf.size()) {
        fclose(f);
        return rtmh::set_error(rtmh::RTMH_E_FORMAT, "%s: truncated trace %d", path, i);
      }
      for (int s = 0; s < n; ++s) data[(size_t)i * n + s] = decode_sample(&buf[(size_t)s * nb], fmt);
    }
  }
  fclose(f);
  return rtmh::RTMH_OK;
}

}  // extern "C"
