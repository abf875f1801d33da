// Complex GEMM on CDNA4 matrix cores.
//
//   C(m,n) = alpha * sum_k A(m,k) * B(n,k) + beta * C(m,n)      (batched)
//
// MFMA has no complex type.  The complex product is run as a real GEMM on the
// interleaved storage: A is read as-is (real [m][2k]); B is expanded on the
// fly to B'(2n+s, 2k+t) = {{b_r, -b_i}, {b_i, b_r}}[s][t] while its fragment
// is built from LDS, so the real output C'(m, 2n+s) is the complex C in
// interleaved storage.  No operand is ever stored expanded.
//
// Compute precisions:
//   CP_F64    v_mfma_f64_16x16x4_f64, FP64 accumulation
//   CP_F32    v_mfma_f32_16x16x4_f32, FP32 accumulation
//   CP_BF16X3 v_mfma_f32_16x16x32_bf16 on a hi/lo split of both operands
//             (x = hi + lo, hi = bf16(x), lo = bf16(x - hi)); the three
//             products hi*hi + hi*lo + lo*hi accumulate in FP32.
//
// Output layouts (lane l, accumulator register j, column l%16): the FP32 and
// BF16 16x16 forms hold rows 4*(l/16)+j; the FP64 form holds rows
// 4*j+(l/16).  Within one MFMA, A and B use the same
// lane->k assignment, so the k order inside a fragment is free; it is chosen
// so that each lane reads whole complex numbers (FP64/FP32) or four
// consecutive complex numbers (BF16) from LDS.
#pragma once
#include "common.hpp"
#include <type_traits>

enum { CP_F64 = 0, CP_F32 = 1, CP_BF16X3 = 2, CP_BF16X6 = 3, CP_BF16X1 = 4 };

struct ZGemmArgs {
  const void* A;
  const void* B;
  void* C;
  long sam, sak;  // A(m,k) at A[m*sam + k*sak] (complex elements)
  long sbn, sbk;  // B(n,k)
  long scm, scn;  // C(m,n)
  // batch b -> offset (b / bdiv) * s1 + (b % bdiv) * s2 (complex elements)
  long sa1, sa2, sb1, sb2, sc1, sc2;
  int bdiv;
  int m, n, k;
  double ar, ai, br, bi;
  // optional second output with C's strides: D(m,n) = F(m,n) + alpha*AB
  // (F may equal D; neither may be an operand of the same call).  C may be
  // null when only D is wanted.
  void* D;
  const void* F;
  // tile grid: blockIdx.x runs over MT*NT*batch tiles (filled by the launcher)
  int MT, NT;
  // lower-triangular B: B(n, k) = 0 for k > n, so the tile of columns
  // [n0, n0 + BNC) stops its k loop at n0 + BNC; heavy tiles run first
  int ktri;
  // use conj(B) in place of B
  int conjb;
  // C(m,n) = F(m,n) + alpha*AB (F with C's strides, D unused)
  int cf;
  // F(m,n) at F[m*sfm + n*sfn] (complex elements; batch offsets as C's)
  // when sfm or sfn is nonzero, else with C's strides
  long sfm, sfn;
  // three-piece split-BF16 operands (zgemm_b16dma_kernel with NP = 3): the
  // third pieces of A and B, one 32-bit word per complex element, with A's
  // and B's element strides and batch offsets
  const void* A3;
  const void* B3;
  // and of a split-BF16 output C (when not null): its third pieces, with C's
  // element strides and batch offsets
  void* C3;
};

template <int CP> struct LdsT;
template <> struct LdsT<CP_F64> {
  using T = zd;
  static constexpr int PAD = 1;
};
template <> struct LdsT<CP_F32> {
  using T = zf;
  static constexpr int PAD = 2;
};

template <class Z> __device__ inline void zero_z(Z& z)
{
  z.re = 0;
  z.im = 0;
}
__device__ inline void zero_z(zb& z)
{
  z.hi = 0;
  z.lo = 0;
}

// ADMA (split-BF16 only): A is split-BF16 with unit k stride, 16-byte
// aligned rows and k % 16 == 0 (checked by the launcher), and its k-tiles go
// from global memory straight into LDS (global_load_lds_dwordx4), without
// passing through registers.  BDMA (with ADMA): B likewise.
template <int CP, class AT, class BT, class OT, bool AK, bool BK, int BM, int BNC, int BKC, int WM, int WN,
          bool ADMA = false, bool BDMA = false>
__global__ __launch_bounds__(64 * WM * WN) void zgemm_kernel(ZGemmArgs p)
{
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM;          // real rows per wave
  constexpr int WTN = 2 * BNC / WN;     // real cols per wave
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int EA = (BM * BKC + NT - 1) / NT;
  constexpr int EB = (BNC * BKC + NT - 1) / NT;
  static_assert(WTM % 16 == 0 && WTN % 16 == 0, "wave tile must be a multiple of 16");
  static_assert(CP != CP_BF16X3 || BKC % 16 == 0, "bf16 k-tile must be a multiple of 16 complex");
  static_assert(CP == CP_BF16X3 || BKC % 4 == 0, "k-tile must be a multiple of 4 complex");
  static_assert(CP == CP_BF16X3 || (!std::is_same<AT, bfs>::value && !std::is_same<BT, bfs>::value),
                "split-BF16 operands need the BF16 path");
  static_assert(!ADMA || (CP == CP_BF16X3 && std::is_same<AT, bfs>::value && AK && BKC == 16),
                "direct LDS staging of A needs split-BF16 A with k contiguous and 16-complex k-tiles");
  static_assert(!BDMA || (ADMA && std::is_same<BT, bfs>::value && BK),
                "direct LDS staging of B needs split-BF16 B with k contiguous (and A staged likewise)");
  static_assert(!BDMA || (BNC / WN) % 16 == 0, "each wave's B rows must start at a multiple of 16");
  using AZ = typename ztype<AT>::type;
  using BZ = typename ztype<BT>::type;
  using ACC = typename std::conditional<CP == CP_F64, doublex4, floatx4>::type;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int g = lane >> 4, r = lane & 15, s = r & 1;
  // XCD-aware tile order: consecutive workgroup ids go round-robin over the 8
  // XCDs, so id -> (id % 8) * (T / 8) + id / 8 gives every XCD a contiguous
  // run of tiles; within a run the M tile varies fastest, then the N tile,
  // then the batch item, so tiles sharing a B tile or a batch item share an L2.
  long lin = blockIdx.x;
  {
    const long T = gridDim.x, per = T / 8;
    if (p.ktri) {
      // triangular work falls with the N tile: deal N tiles round-robin to
      // the XCDs (heaviest first) and keep the M tiles of one N tile on one XCD
      if ((p.NT & 7) == 0 && T == (long)p.MT * p.NT) {
        const long x = lin & 7, q = lin >> 3;
        lin = (x + 8 * (q / p.MT)) * p.MT + q % p.MT;
      }
    } else if (per > 0 && lin < per * 8) {
      lin = (lin % 8) * per + lin / 8;
    }
  }
  const int mt = (int)(lin % p.MT);
  const long rest = lin / p.MT;
  const int ntl = (int)(rest % p.NT);
  const int m0 = mt * BM, n0 = (p.ktri ? p.NT - 1 - ntl : ntl) * BNC;
  const int kend = p.ktri ? min(p.k, n0 + BNC) : p.k;
  const int b = (int)(rest / p.NT);
  const long bq = b / p.bdiv, br_ = b % p.bdiv;
  const AZ* __restrict__ A = (const AZ*)p.A + bq * p.sa1 + br_ * p.sa2;
  const BZ* __restrict__ B = (const BZ*)p.B + bq * p.sb1 + br_ * p.sb2;
  constexpr bool OSPLIT = std::is_same<OT, bfs>::value;
  using DT = typename std::conditional<OSPLIT, float, OT>::type;  // D/F element
  using CT = typename std::conditional<OSPLIT, float, OT>::type;  // C element when not split
  const long coff = 2 * (bq * p.sc1 + br_ * p.sc2);
  CT* C = (p.C && !OSPLIT) ? (CT*)p.C + coff : nullptr;
  zb* CS = (p.C && OSPLIT) ? (zb*)p.C + coff / 2 : nullptr;
  DT* D = p.D ? (DT*)p.D + coff : nullptr;
  const DT* F = p.F ? (const DT*)p.F + coff : nullptr;

  ACC acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) acc[i][j][q] = 0;

  if constexpr (CP != CP_BF16X3) {
  using L = LdsT<CP>;
  using LT = typename L::T;
  constexpr int LDK = BKC + L::PAD;
  // two LDS stages (ping-pong): the next k-tile is stored while others compute
  __shared__ __align__(16) LT sA[2][BM * LDK];
  __shared__ __align__(16) LT sB[2][BNC * LDK];
  AZ ra[EA];
  BZ rb[EB];
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < EA; i++) {
      int e = tid + i * NT;
      int row, kc;
      if (AK) { row = e / BKC; kc = e % BKC; } else { kc = e / BM; row = e % BM; }
      int gm = m0 + row, gk = k0 + kc;
      if (e < BM * BKC && gm < p.m && gk < kend) ra[i] = A[(long)gm * p.sam + (long)gk * p.sak];
      else zero_z(ra[i]);
    }
#pragma unroll
    for (int i = 0; i < EB; i++) {
      int e = tid + i * NT;
      int col, kc;
      if (BK) { col = e / BKC; kc = e % BKC; } else { kc = e / BNC; col = e % BNC; }
      int gn = n0 + col, gk = k0 + kc;
      if (e < BNC * BKC && gn < p.n && gk < kend) rb[i] = B[(long)gn * p.sbn + (long)gk * p.sbk];
      else zero_z(rb[i]);
    }
  };
  auto sstore = [&](int st) {
#pragma unroll
    for (int i = 0; i < EA; i++) {
      int e = tid + i * NT;
      if (e >= BM * BKC) break;
      int row, kc;
      if (AK) { row = e / BKC; kc = e % BKC; } else { kc = e / BM; row = e % BM; }
      if constexpr (CP == CP_F64) {
        sA[st][row * LDK + kc] = zd{(double)ra[i].re, (double)ra[i].im};
      } else {
        sA[st][row * LDK + kc] = zf{(float)ra[i].re, (float)ra[i].im};
      }
    }
#pragma unroll
    for (int i = 0; i < EB; i++) {
      int e = tid + i * NT;
      if (e >= BNC * BKC) break;
      int col, kc;
      if (BK) { col = e / BKC; kc = e % BKC; } else { kc = e / BNC; col = e % BNC; }
      if constexpr (CP == CP_F64) {
        const double bi = (double)rb[i].im;
        sB[st][col * LDK + kc] = zd{(double)rb[i].re, p.conjb ? -bi : bi};
      } else {
        const float bi = (float)rb[i].im;
        sB[st][col * LDK + kc] = zf{(float)rb[i].re, p.conjb ? -bi : bi};
      }
    }
  };

  const int nk = (kend + BKC - 1) / BKC;
  gload(0);
  sstore(0);
  __syncthreads();
  const int arow0 = wm * WTM + r;           // + bi*16
  const int bcol0 = (wn * WTN) / 2 + r / 2; // + bj*8  (complex column in tile)
  for (int kt = 0; kt < nk; kt++) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * BKC);
    {
      using RT = typename std::conditional<CP == CP_F64, double, float>::type;
#pragma unroll
      for (int kq = 0; kq < BKC / 4; kq++) {
        LT a[TM], bv[TN];
#pragma unroll
        for (int i = 0; i < TM; i++) a[i] = sA[cur][(arow0 + i * 16) * LDK + kq * 4 + g];
#pragma unroll
        for (int j = 0; j < TN; j++) bv[j] = sB[cur][(bcol0 + j * 8) * LDK + kq * 4 + g];
#pragma unroll
        for (int j = 0; j < TN; j++) {
          RT b0 = s ? bv[j].im : bv[j].re;
          RT b1 = s ? bv[j].re : -bv[j].im;
#pragma unroll
          for (int i = 0; i < TM; i++) {
            if constexpr (CP == CP_F64) {
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].re, b0, acc[i][j], 0, 0, 0);
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].im, b1, acc[i][j], 0, 0, 0);
            } else {
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].re, b0, acc[i][j], 0, 0, 0);
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].im, b1, acc[i][j], 0, 0, 0);
            }
          }
        }
      }
    }
    // stage cur^1 was last read in iteration kt-1, which every thread has
    // finished (barrier below), so it can be refilled now
    if (kt + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }

  } else {
  // Split-BF16: the LDS tiles keep the interleaved (hi, lo) words of each
  // complex number, rows of LDZ = BKC + 2 entries (144 B at BKC = 16), so
  // that the 16-byte stores of k pairs and the 2 x 16-byte fragment reads
  // of 4 consecutive k are both free of bank conflicts.
  constexpr int BKP = BKC / 2;
  constexpr int LDZ = BKC + 2;
  // With ADMA the A tile rows are 16 complex (128 B, eight 16-byte chunks,
  // no padding, as the direct loads fill LDS contiguously) and chunk c of
  // row R is stored at c ^ swz(R), swz(R) = (R & 7) ^ ((R >> 3) & 1): the 16
  // rows that one fragment read touches then fall on 16 distinct 16-byte
  // bank groups, and any 8 consecutive of them on 8.  With BDMA the B tile
  // (rows = columns of the product) is laid out the same way.
  constexpr int PA = ADMA ? 0 : (BM * BKP + NT - 1) / NT;  // A pairs per thread (register path)
  constexpr int PAR = PA > 0 ? PA : 1;
  constexpr int PB = BDMA ? 0 : (BNC * BKP + NT - 1) / NT;  // B pairs per thread (register path)
  constexpr int PBR = PB > 0 ? PB : 1;
  constexpr int LDA = ADMA ? BKC : LDZ;
  constexpr int LDB = BDMA ? BKC : LDZ;
  __shared__ __align__(16) uint2 zA[2][BM * LDA];
  __shared__ __align__(16) uint2 zB[2][BNC * LDB];
  // global -> register prefetch of the next k-tile while the MFMAs run.
  // Each thread's operand pointers are fixed over the k loop; tiles that lie
  // inside the matrix in m / n load without per-element checks except in the
  // k-tile that crosses kend.
  static_assert((BM * BKP) % NT == 0 && (BNC * BKP) % NT == 0, "tile must split evenly over the threads");
  AZ ra[PAR][2];
  BZ rb[PBR][2];
  const AZ* pa[PAR];
  const BZ* pb[PBR];
  int ka[PAR], kb[PBR];
  bool va[PAR], vb[PBR];
#pragma unroll
  for (int i = 0; i < PA; i++) {
    const int e = tid + i * NT;
    int row, kp;
    if (AK) { row = e / BKP; kp = e % BKP; } else { kp = e / BM; row = e % BM; }
    const int gm = m0 + row;
    va[i] = gm < p.m;
    ka[i] = 2 * kp;
    pa[i] = A + (long)(va[i] ? gm : 0) * p.sam + (long)ka[i] * p.sak;
  }
#pragma unroll
  for (int i = 0; i < PB; i++) {
    const int e = tid + i * NT;
    int col, kp;
    if (BK) { col = e / BKP; kp = e % BKP; } else { kp = e / BNC; col = e % BNC; }
    const int gn = n0 + col;
    vb[i] = gn < p.n;
    kb[i] = 2 * kp;
    pb[i] = B + (long)(vb[i] ? gn : 0) * p.sbn + (long)kb[i] * p.sbk;
  }
  // ADMA: 16-byte unit u = i * NT + tid of an A stage is (row u / 8, chunk
  // u % 8), filled from logical chunk (u % 8) ^ swz(row); rows past m read
  // row m - 1 (their outputs are never stored)
  constexpr int AI = ADMA ? BM * BKP / NT : 1;  // direct loads per thread per k-tile
  static_assert(!ADMA || (BM * BKP) % NT == 0, "A tile must split evenly over the threads");
  const AZ* pd[AI];
  if constexpr (ADMA) {
#pragma unroll
    for (int i = 0; i < AI; i++) {
      const int u = i * NT + tid, row = u / BKP, pc = u % BKP;
      const int lc = pc ^ ((row & 7) ^ ((row >> 3) & 1));
      pd[i] = A + (long)min(m0 + row, p.m - 1) * p.sam + 2 * lc;
    }
  }
  // BDMA: the same for B, rows past n reading row n - 1
  constexpr int BI = BDMA ? BNC * BKP / NT : 1;
  static_assert(!BDMA || (BNC * BKP) % NT == 0, "B tile must split evenly over the threads");
  const BZ* pdb[BI];
  if constexpr (BDMA) {
#pragma unroll
    for (int i = 0; i < BI; i++) {
      const int u = i * NT + tid, row = u / BKP, pc = u % BKP;
      const int lc = pc ^ ((row & 7) ^ ((row >> 3) & 1));
      pdb[i] = B + (long)min(n0 + row, p.n - 1) * p.sbn + 2 * lc;
    }
  }
  auto dma = [&](int st, int k0) {
    if constexpr (ADMA) {
      const int w0 = __builtin_amdgcn_readfirstlane(wave);
#pragma unroll
      for (int i = 0; i < AI; i++)
        __builtin_amdgcn_global_load_lds((const void*)(pd[i] + k0),
                                         (__attribute__((address_space(3))) void*)&zA[st][2 * (i * NT + w0 * 64)],
                                         16, 0, 0);
      if constexpr (BDMA) {
#pragma unroll
        for (int i = 0; i < BI; i++)
          __builtin_amdgcn_global_load_lds((const void*)(pdb[i] + k0),
                                           (__attribute__((address_space(3))) void*)&zB[st][2 * (i * NT + w0 * 64)],
                                           16, 0, 0);
      }
    }
  };
  const bool interior = (m0 + BM <= p.m) && (n0 + BNC <= p.n);
  auto gload = [&](int k0) {
    const long oa = (long)k0 * p.sak, ob = (long)k0 * p.sbk;
    if (interior && k0 + BKC <= kend) {
#pragma unroll
      for (int i = 0; i < PA; i++) {
        ra[i][0] = pa[i][oa];
        ra[i][1] = pa[i][oa + p.sak];
      }
#pragma unroll
      for (int i = 0; i < PB; i++) {
        rb[i][0] = pb[i][ob];
        rb[i][1] = pb[i][ob + p.sbk];
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < PA; i++) {
      const int gk = k0 + ka[i];
      if (va[i] && gk < kend) ra[i][0] = pa[i][oa];
      else zero_z(ra[i][0]);
      if (va[i] && gk + 1 < kend) ra[i][1] = pa[i][oa + p.sak];
      else zero_z(ra[i][1]);
    }
#pragma unroll
    for (int i = 0; i < PB; i++) {
      const int gk = k0 + kb[i];
      if (vb[i] && gk < kend) rb[i][0] = pb[i][ob];
      else zero_z(rb[i][0]);
      if (vb[i] && gk + 1 < kend) rb[i][1] = pb[i][ob + p.sbk];
      else zero_z(rb[i][1]);
    }
  };
  auto tozb = [](auto v, bool conj) -> zb {
    if constexpr (std::is_same<decltype(v), zb>::value) {
      const uint32_t flip = conj ? 0x80000000u : 0u;  // sign of the imaginary halves
      return zb{v.hi ^ flip, v.lo ^ flip};
    } else {
      const float im = (float)v.im;
      return split_z((float)v.re, conj ? -im : im);
    }
  };
  auto sstore = [&](int st) {
#pragma unroll
    for (int i = 0; i < PA; i++) {
      const int e = tid + i * NT;
      if (e >= BM * BKP) break;
      int row, kp;
      if (AK) { row = e / BKP; kp = e % BKP; } else { kp = e / BM; row = e % BM; }
      const zb v0 = tozb(ra[i][0], false), v1 = tozb(ra[i][1], false);
      *(uint4*)&zA[st][row * LDA + 2 * kp] = make_uint4(v0.hi, v0.lo, v1.hi, v1.lo);
    }
#pragma unroll
    for (int i = 0; i < PB; i++) {
      const int e = tid + i * NT;
      if (e >= BNC * BKP) break;
      int col, kp;
      if (BK) { col = e / BKP; kp = e % BKP; } else { kp = e / BNC; col = e % BNC; }
      const zb v0 = tozb(rb[i][0], p.conjb != 0), v1 = tozb(rb[i][1], p.conjb != 0);
      *(uint4*)&zB[st][col * LDB + 2 * kp] = make_uint4(v0.hi, v0.lo, v1.hi, v1.lo);
    }
  };

  const int arow0 = wm * WTM + r;           // + bi*16
  const int bcol0 = (wn * WTN) / 2 + r / 2; // + bj*8  (complex column in tile)
  // (conj(B) is taken at the LDS store, so BDMA launches have conjb == 0)
  const uint32_t brot = s ? 16u : 0u, bneg = s ? 0u : 0x80000000u;
  // this lane's rows are 16-aligned blocks + r, so swz(row) = swz(r)
  const int aswz = (r & 7) ^ ((r >> 3) & 1);
  // knext >= 0 (ADMA): A's next k-tile is issued into the other stage once
  // this tile's fragments are read
  auto compute = [&](int cur, int knext) {
#pragma unroll
    for (int kq = 0; kq < BKC / 16; kq++) {
      uint4 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) {
        uint4 u0, u1;
        if constexpr (ADMA) {
          const uint4* pa = (const uint4*)&zA[cur][(arow0 + i * 16) * LDA];
          u0 = pa[(2 * g) ^ aswz];
          u1 = pa[(2 * g + 1) ^ aswz];
        } else {
          const uint4* pa = (const uint4*)&zA[cur][(arow0 + i * 16) * LDA + kq * 16 + 4 * g];
          u0 = pa[0];
          u1 = pa[1];
        }
        ah[i] = make_uint4(u0.x, u0.z, u1.x, u1.z);
        al[i] = make_uint4(u0.y, u0.w, u1.y, u1.w);
      }
#pragma unroll
      for (int j = 0; j < TN; j++) {
        uint4 u0, u1;
        if constexpr (BDMA) {
          // column-block rows bcol0 + 8j: (wn * WTN) / 2 is a multiple of 16,
          // so swz(row) = (r / 2) ^ (j & 1)
          const uint4* pb = (const uint4*)&zB[cur][(bcol0 + j * 8) * LDB];
          const int bsw = (r >> 1) ^ (j & 1);
          u0 = pb[(2 * g) ^ bsw];
          u1 = pb[(2 * g + 1) ^ bsw];
        } else {
          const uint4* pb = (const uint4*)&zB[cur][(bcol0 + j * 8) * LDB + kq * 16 + 4 * g];
          u0 = pb[0];
          u1 = pb[1];
        }
        // lanes s = 0 (real output column) take (b_re, -b_im), lanes s = 1
        // (imaginary column) take (b_im, b_re): one xor and one rotate per
        // word, no branch
        auto ex = [&](uint32_t w) { return __builtin_amdgcn_alignbit(w, w, brot) ^ bneg; };
        bh[j] = make_uint4(ex(u0.x), ex(u0.z), ex(u1.x), ex(u1.z));
        bl[j] = make_uint4(ex(u0.y), ex(u0.w), ex(u1.y), ex(u1.w));
      }
      if (ADMA && kq == 0 && knext >= 0) dma(cur ^ 1, knext);
      // the three products of each accumulator are issued TM*TN MFMAs apart,
      // so consecutive MFMAs never wait on each other's result
#pragma unroll
      for (int t = 0; t < 3; t++)
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++) {
            const bf16x8 xa = __builtin_bit_cast(bf16x8, t == 2 ? al[i] : ah[i]);
            const bf16x8 xb = __builtin_bit_cast(bf16x8, t == 1 ? bl[j] : bh[j]);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa, xb, acc[i][j], 0, 0, 0);
          }
    }
  };
  const int nk = (kend + BKC - 1) / BKC;
  dma(0, 0);
  gload(0);
  sstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * BKC);
    compute(cur, kt + 1 < nk ? (kt + 1) * BKC : -1);
    // stage cur^1 was last read in iteration kt-1, before its barrier
    if (kt + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }
  }

  // epilogue: lanes (2c, 2c+1) hold Re and Im of complex column c.  The
  // values it reads (F, and C when beta != 0) are loaded for the whole tile
  // first: D may alias F, so loads issued between the stores could not
  // overlap them.
  const bool has_beta = (p.br != 0.0 || p.bi != 0.0);
  // FP64 16x16x4 interleaves rows (row = 4*reg + lane/16); the FP32 and BF16
  // 16x16 forms keep 4 consecutive rows per lane group.
  const int rbase = m0 + wm * WTM + ((CP == CP_F64) ? g : 4 * g);
  constexpr int RQ = (CP == CP_F64) ? 4 : 1;  // row step of the accumulator register
  const int cbase = n0 + (wn * WTN + r) / 2;
  // one accumulator row block at a time: its F (and C for beta) values are
  // loaded first, then the block is written
#pragma unroll
  for (int i = 0; i < TM; i++) {
    DT fv[TN][4];
    CT cv[TN][4];
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = rbase + i * 16 + RQ * q, col = cbase + j * 8;
        const bool in = row < p.m && col < p.n;
        const long off = 2 * ((long)row * p.scm + (long)col * p.scn);
        const long foff = (p.sfm || p.sfn) ? 2 * ((long)row * p.sfm + (long)col * p.sfn) : off;
        fv[j][q] = ((D || p.cf) && F && in) ? F[foff + s] : (DT)0;
        if constexpr (!OSPLIT) cv[j][q] = (C && has_beta && in) ? C[off + s] : (CT)0;
      }
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = rbase + i * 16 + RQ * q, col = cbase + j * 8;
        const double v = (double)acc[i][j][q];
        const double pv = __shfl_xor(v, 1);
        const double re = s ? pv : v, im = s ? v : pv;
        const double o = s ? (p.ar * im + p.ai * re) : (p.ar * re - p.ai * im);  // alpha * AB
        double oc = o;                                                              // + beta * C
        if constexpr (!OSPLIT) {
          if (has_beta) {
            const double cs = (double)cv[j][q], cp = __shfl_xor(cs, 1);
            const double cr = s ? cp : cs, ci = s ? cs : cp;
            oc += s ? (p.br * ci + p.bi * cr) : (p.br * cr - p.bi * ci);
          }
        }
        const double ofc = p.cf ? (double)fv[j][q] + o : o;  // C = F + alpha*AB
        const double oi = __shfl_xor(ofc, 1);
        if (row < p.m && col < p.n) {
          const long off = 2 * ((long)row * p.scm + (long)col * p.scn);
          if (D) D[off + s] = (DT)((double)fv[j][q] + o);
          if constexpr (OSPLIT) {
            // lane s=0 writes the whole split complex (no beta), and its
            // third piece when C3 is set
            if (CS && s == 0) {
              if (p.C3) {
                zb h;
                uint32_t t;
                split3_f((float)ofc, (float)oi, h, t);
                CS[off / 2] = h;
                ((uint32_t*)p.C3)[coff / 2 + off / 2] = t;
              } else {
                CS[off / 2] = split_z((float)ofc, (float)oi);
              }
            }
          } else if (C) {
            C[off + s] = (CT)(p.cf ? oc + (double)fv[j][q] : oc);
          }
        }
      }
  }
}
