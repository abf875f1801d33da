// Counter-based N(0,1) auxiliary-field draws, identical bit for bit on the
// host and on the GPU.
//
// The reference draws the fields inside every substep from one global
// Mersenne-Twister stream through Box-Muller
// (src/miniapps/miniafqmc_base.cpp:282, src/Utilities/RandomGenerator.h:50-62).
// Here draw (n, w) of a substep is a pure function of (seed, substep, walker,
// n), so the GPU can generate the fields of a captured substep on the fly
// while the host generates the same numbers for CPU replays and for walker
// subsets.
//
// Bitwise agreement needs every operation to round the same way on both
// sides.  The functions use only integer operations and IEEE-754 double
// addition, subtraction, multiplication, division and square root, which are
// correctly rounded on x86-64 and on CDNA; contraction into fused
// multiply-adds is switched off inside each function so that neither compiler
// fuses differently.  log and sin/cos are evaluated by fixed polynomials
// (argument reduction by exact bit manipulation, then the series of
// 2 atanh(s) and of sin/cos on [-pi/4, pi/4]) instead of the platform's
// math libraries, whose last bits differ between the host and the GPU.
// Accuracy is a few ulp, far below what the sampling needs.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace afqmc_rng {

// stream tag of the auxiliary fields (other tags belong to the system generator)
constexpr uint64_t kFieldStream = 5;

__host__ __device__ inline uint64_t mix64(uint64_t z)
{
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// uniform in (0, 1]: 53 random bits
__host__ __device__ inline double u01(uint64_t seed, uint64_t stream, uint64_t idx)
{
  const uint64_t h = mix64(mix64(seed * 0x2545f4914f6cdd1dull + stream) ^ (idx * 0xd1342543de82ef95ull));
  return (double)((h >> 11) + 1) * (1.0 / 9007199254740992.0);
}

__host__ __device__ inline double bits_to_double(uint64_t b)
{
  union {
    uint64_t u;
    double d;
  } v;
  v.u = b;
  return v.d;
}
__host__ __device__ inline uint64_t double_to_bits(double d)
{
  union {
    uint64_t u;
    double d;
  } v;
  v.d = d;
  return v.u;
}

// natural log of u in [2^-53, 1] (normal doubles only).  u = m 2^e with m in
// [sqrt(1/2), sqrt(2)) by exact exponent manipulation; log m = 2 atanh(s),
// s = (m - 1) / (m + 1), |s| < 0.1716, summed to s^25; e ln 2 with ln 2 split
// in a high part of 32 significant bits (exact products e * ln2_hi) and a
// low part.
__host__ __device__ inline double det_log(double u)
{
#pragma clang fp contract(off)
  const uint64_t b = double_to_bits(u);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  double m = bits_to_double((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 1.0 / 25.0;
  p = p * z + 1.0 / 23.0;
  p = p * z + 1.0 / 21.0;
  p = p * z + 1.0 / 19.0;
  p = p * z + 1.0 / 17.0;
  p = p * z + 1.0 / 15.0;
  p = p * z + 1.0 / 13.0;
  p = p * z + 1.0 / 11.0;
  p = p * z + 1.0 / 9.0;
  p = p * z + 1.0 / 7.0;
  p = p * z + 1.0 / 5.0;
  p = p * z + 1.0 / 3.0;
  const double lm = 2.0 * s + 2.0 * s * (z * p);
  const double ln2_hi = 6.93147180369123816490e-01;  // 32 significant bits of ln 2
  const double ln2_lo = 1.90821492927058770002e-10;  // ln 2 - ln2_hi
  const double de = (double)e;
  return de * ln2_hi + (de * ln2_lo + lm);
}

// cos and sin of 2 pi t for t in [0, 1]: t = (q + f) / 4 with q the nearest
// integer to 4t (exact), f in [-1/2, 1/2]; phi = (pi/2) f in [-pi/4, pi/4];
// sin phi to phi^19 and cos phi to phi^18, then the quadrant rotation.
__host__ __device__ inline void det_cossin_2pi(double t, double& c, double& s)
{
#pragma clang fp contract(off)
  const double t4 = 4.0 * t;
  const double q = (double)(int64_t)(t4 + 0.5);
  const double f = t4 - q;
  const double phi = f * 1.5707963267948966;
  const double z = phi * phi;
  double ps = -1.0 / 121645100408832000.0;   // -1/19!
  ps = ps * z + 1.0 / 355687428096000.0;      // 1/17!
  ps = ps * z - 1.0 / 1307674368000.0;        // -1/15!
  ps = ps * z + 1.0 / 6227020800.0;           // 1/13!
  ps = ps * z - 1.0 / 39916800.0;             // -1/11!
  ps = ps * z + 1.0 / 362880.0;               // 1/9!
  ps = ps * z - 1.0 / 5040.0;                 // -1/7!
  ps = ps * z + 1.0 / 120.0;                  // 1/5!
  ps = ps * z - 1.0 / 6.0;                    // -1/3!
  const double sp = phi + phi * (z * ps);
  double pc = 1.0 / 6402373705728000.0;       // 1/18!
  pc = pc * z - 1.0 / 20922789888000.0;       // -1/16!
  pc = pc * z + 1.0 / 87178291200.0;          // 1/14!
  pc = pc * z - 1.0 / 479001600.0;            // -1/12!
  pc = pc * z + 1.0 / 3628800.0;              // 1/10!
  pc = pc * z - 1.0 / 40320.0;                // -1/8!
  pc = pc * z + 1.0 / 720.0;                  // 1/6!
  pc = pc * z - 1.0 / 24.0;                   // -1/4!
  pc = pc * z + 0.5;                          // 1/2!
  const double cp = 1.0 - z * pc;
  switch (((int64_t)q) & 3) {
  case 0: c = cp; s = sp; break;
  case 1: c = -sp; s = cp; break;
  case 2: c = -cp; s = -sp; break;
  default: c = sp; s = -cp; break;
  }
}

// Box-Muller pair (a, b) of draw pair idx of a stream
__host__ __device__ inline void normal2(uint64_t seed, uint64_t stream, uint64_t idx, double& a, double& b)
{
#pragma clang fp contract(off)
  const double u1 = u01(seed, stream, 2 * idx), u2 = u01(seed, stream, 2 * idx + 1);
  const double r = __builtin_sqrt(-2.0 * det_log(u1));
  double c, s;
  det_cossin_2pi(u2, c, s);
  a = r * c;
  b = r * s;
}

// the stream of walker w (index in the whole population) in substep `substep`
__host__ __device__ inline uint64_t field_stream(long substep, long w)
{
  return (kFieldStream << 56) ^ ((uint64_t)substep << 24) ^ (uint64_t)w;
}

// draws n and n ^ 1 (n even) of walker w: the pair n / 2 of its stream
__host__ __device__ inline void field_pair(uint64_t seed, long substep, long w, int n, double& a, double& b)
{
  normal2(seed, field_stream(substep, w), (uint64_t)(n >> 1), a, b);
}

} // namespace afqmc_rng
