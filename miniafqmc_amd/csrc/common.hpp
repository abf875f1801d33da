// Shared device/host definitions for the gfx950 walker-step kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef short shortx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct zd {
  double re, im;
};
struct zf {
  float re, im;
};

__host__ __device__ inline zd zmul(zd a, zd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline zd zadd(zd a, zd b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ inline zd zsub(zd a, zd b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ inline zd zconj(zd a) { return {a.re, -a.im}; }
__host__ __device__ inline double zcabs1(zd a) { return fabs(a.re) + fabs(a.im); }
// Smith's complex division
__host__ __device__ inline zd zdiv(zd a, zd b)
{
  if (fabs(b.im) <= fabs(b.re)) {
    double r = b.im / b.re, den = b.re + b.im * r;
    return {(a.re + a.im * r) / den, (a.im - a.re * r) / den};
  }
  double r = b.re / b.im, den = b.im + b.re * r;
  return {(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}

// round-to-nearest-even float -> bf16 bit pattern (inputs are finite)
__device__ inline unsigned short f2bf(float x)
{
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short b) { return __uint_as_float(((uint32_t)b) << 16); }

// Split-BF16 complex: x = hi + lo per component, hi = bf16(x),
// lo = bf16(x - hi); each half packed as (re | im << 16).  8 bytes like a
// complex float, so buffers keep their size in either format.
struct zb {
  uint32_t hi, lo;
};
struct bfs {};  // element tag of zb arrays (one complex per zb)

template <class T> struct ztype;
template <> struct ztype<bfs> {
  using type = zb;
};
template <> struct ztype<double> {
  using type = zd;
};
template <> struct ztype<float> {
  using type = zf;
};

__device__ inline zb split_z(float re, float im)
{
  unsigned short hr = f2bf(re), hi = f2bf(im);
  unsigned short lr = f2bf(re - bf2f(hr)), li = f2bf(im - bf2f(hi));
  return zb{(uint32_t)hr | ((uint32_t)hi << 16), (uint32_t)lr | ((uint32_t)li << 16)};
}
// Three-piece split-BF16: x = h + m + l per component, h = bf16(x),
// m = bf16(x - h), l = bf16(x - h - m) (the residuals in FP64), which holds
// 24 significant bits.  (h, m) is stored as a zb, exactly the two-piece split
// of x; l goes to a separate array of one word (bf16(re) | bf16(im) << 16)
// per complex, so two-piece consumers read the same zb array.
__device__ inline unsigned short d2bf(double x) { return f2bf((float)x); }
__device__ __forceinline__ void split3_z(double re, double im, zb& hm, uint32_t& l)
{
  const unsigned short hr = d2bf(re), hi = d2bf(im);
  const double rr = re - (double)bf2f(hr), ri = im - (double)bf2f(hi);
  const unsigned short mr = d2bf(rr), mi = d2bf(ri);
  const unsigned short lr = d2bf(rr - (double)bf2f(mr)), li = d2bf(ri - (double)bf2f(mi));
  hm = zb{(uint32_t)hr | ((uint32_t)hi << 16), (uint32_t)mr | ((uint32_t)mi << 16)};
  l = (uint32_t)lr | ((uint32_t)li << 16);
}
// the same for an FP32 value, in FP32 arithmetic (x - h and x - h - m are
// exact in FP32), with (h, m) equal to split_z's
__device__ __forceinline__ void split3_f(float re, float im, zb& hm, uint32_t& l)
{
  const unsigned short hr = f2bf(re), hi = f2bf(im);
  const float rr = re - bf2f(hr), ri = im - bf2f(hi);
  const unsigned short mr = f2bf(rr), mi = f2bf(ri);
  hm = zb{(uint32_t)hr | ((uint32_t)hi << 16), (uint32_t)mr | ((uint32_t)mi << 16)};
  l = (uint32_t)f2bf(rr - bf2f(mr)) | ((uint32_t)f2bf(ri - bf2f(mi)) << 16);
}
__device__ inline zd to_zd(zd v) { return v; }
__device__ inline zd to_zd(zf v) { return zd{(double)v.re, (double)v.im}; }
__device__ inline zd to_zd(zb v)
{
  return zd{(double)bf2f((unsigned short)(v.hi & 0xffff)) + (double)bf2f((unsigned short)(v.lo & 0xffff)),
            (double)bf2f((unsigned short)(v.hi >> 16)) + (double)bf2f((unsigned short)(v.lo >> 16))};
}
__device__ inline void from_zd(zd v, zd& o) { o = v; }
__device__ inline void from_zd(zd v, zf& o) { o = zf{(float)v.re, (float)v.im}; }
__device__ inline void from_zd(zd v, zb& o) { o = split_z((float)v.re, (float)v.im); }

__device__ inline zd load_z(const zd* p) { return *p; }
__device__ inline zd load_z(const zf* p)
{
  zf v = *p;
  return {(double)v.re, (double)v.im};
}
