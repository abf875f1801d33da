// Split-BF16 complex GEMM with both operands staged by direct global-to-LDS
// loads through a ring of LDS stages.
//
//   C(m,n) = alpha * sum_k A(m,k) * B(n,k) (+ epilogue, as zgemm_kernel)
//
// Operands are split-BF16 (zb: hi and lo words, each bf16(re) | bf16(im) << 16)
// with k contiguous, 16-byte aligned rows and k % 16 == 0.  Each k-tile of 16
// complex (128 bytes per row) goes from global memory into LDS by
// global_load_lds_dwordx4 without passing through registers; chunk c (16 B,
// two complex) of tile row R is stored at c ^ swz(R), swz(R) = (R & 7) ^
// ((R >> 3) & 1), so the 16 rows one fragment read touches fall on distinct
// bank groups.  STAGES k-tiles are in flight: the loads of tile t + STAGES - 1
// are issued once every wave has passed the barrier of tile t (so the stage
// they overwrite, read at t - 1, is free), and each tile is waited for with an
// explicit vmcnt count that leaves the later tiles' loads outstanding.  The k
// loop is unrolled by STAGES so each stage is a separate LDS array addressed
// statically.
//
// Complex arithmetic on v_mfma_f32_16x16x32_bf16: the real and imaginary parts
// of a 16 x 16 complex output block have accumulators of their own.  A lane's
// 8 k-slots hold four consecutive complex k as (re, im) pairs; A enters as
// stored, B as (b_re, -b_im) for the real part and (b_im, b_re) for the
// imaginary part (one xor or one rotate per word), so each lane reads the B
// column it owns once.  The three products hi*hi, hi*lo, lo*hi accumulate in
// FP32 in that order per k-tile, with the same k-slot values as zgemm_kernel's
// expanded form, so both kernels give identical results.
//
// NP = 3 (three-piece split-BF16, 24 significant bits): each operand's third
// pieces (one word per complex, p.A3 / p.B3; 64 bytes per row and k-tile)
// get LDS tiles of their own, two 16-row blocks to a 128-byte LDS row: row
// r of the blocks at rows 32q and 32q + 16 goes to LDS row 16q + r, its
// chunk c at (2c) ^ swz(r) and (2c + 1) ^ swz(r) respectively.  A fragment
// read of one block is then the two-piece tiles' read of chunk pair c0 / c1
// (16 rows of 64-byte rows read densely conflict two ways on this LDS, the
// swizzled half-rows do not).  The six products h*h, h*m, m*h, h*l, l*h,
// m*m accumulate in that order.
//
// NP = 1 (one piece, 8 significant bits): the same two-piece tiles are staged
// and only the hi words enter the MFMAs, h*h alone.
#pragma once
#include "gemm.hpp"

namespace b16 {

// s_waitcnt with only vmcnt = n (expcnt and lgkmcnt at their maxima), gfx9
// encoding: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt[5:4] << 14
template <int N> __device__ __forceinline__ void wait_vmcnt()
{
  static_assert(N >= 0 && N < 64, "vmcnt out of range");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

__device__ __forceinline__ int swz(int row) { return (row & 7) ^ ((row >> 3) & 1); }

} // namespace b16

// BM rows x BN complex columns per workgroup, WM x WN waves, STAGES k-tiles
// of 16 complex in LDS.  OT: output element (double, float, or bfs for split
// output).  Launch conditions (checked by the host): A and B split-BF16, unit
// k stride, rows and batch offsets 16-byte aligned, k % 16 == 0, no conj(B);
// NP = 3 also third-piece rows 16-byte aligned.
// WPE: waves per SIMD the register allocation must allow (0: no bound
// beyond the workgroup size).
template <class OT, int BM, int BN, int WM, int WN, int STAGES, int NP = 2, int WPE = 0>
__global__ __launch_bounds__(64 * WM * WN, WPE) void zgemm_b16dma_kernel(ZGemmArgs p)
{
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;  // rows, complex columns per wave
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int KT = 16;                        // complex k per tile
  constexpr int CH = KT / 2;                    // 16-byte chunks per tile row
  constexpr int AL = BM * CH / NT, BL = BN * CH / NT;  // direct loads per thread per tile
  constexpr int C3 = KT / 4;                    // 16-byte chunks per third-piece row
  constexpr int AL3 = NP == 3 ? BM * C3 / NT : 0, BL3 = NP == 3 ? BN * C3 / NT : 0;
  constexpr int LPT = AL + BL + AL3 + BL3;
  // one stage in uint2 units: the two-piece tiles, then the third pieces
  constexpr int STG = (BM + BN) * KT + (NP == 3 ? (BM + BN) * KT / 2 : 0);
  static_assert(NP >= 1 && NP <= 3, "one, two or three pieces");
  static_assert(NP != 3 || ((BM * C3) % NT == 0 && (BN * C3) % NT == 0), "third-piece tiles must split evenly");
  static_assert(NP != 3 || (WTM % 32 == 0 && WTN % 32 == 0), "third pieces pair 16-row blocks within a wave tile");
  static_assert(WTM % 16 == 0 && WTN % 16 == 0, "wave tile must be a multiple of 16");
  static_assert((BM * CH) % NT == 0 && (BN * CH) % NT == 0, "tiles must split evenly over the threads");
  static_assert(STAGES >= 2 && STAGES <= 4, "2 to 4 stages");
  static_assert(LPT * (STAGES - 2) < 64, "outstanding loads exceed vmcnt");

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int g = lane >> 4, r = lane & 15;

  // XCD-aware tile order, as zgemm_kernel
  long lin = blockIdx.x;
  {
    const long T = gridDim.x, per = T / 8;
    if (p.ktri) {
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
  const int m0 = mt * BM, n0 = (p.ktri ? p.NT - 1 - ntl : ntl) * BN;
  const int kend = p.ktri ? min(p.k, n0 + BN) : p.k;
  const int b = (int)(rest / p.NT);
  const long bq = b / p.bdiv, br_ = b % p.bdiv;
  const zb* __restrict__ A = (const zb*)p.A + bq * p.sa1 + br_ * p.sa2;
  const zb* __restrict__ B = (const zb*)p.B + bq * p.sb1 + br_ * p.sb2;
  const uint32_t* __restrict__ A3 = NP == 3 ? (const uint32_t*)p.A3 + bq * p.sa1 + br_ * p.sa2 : nullptr;
  const uint32_t* __restrict__ B3 = NP == 3 ? (const uint32_t*)p.B3 + bq * p.sb1 + br_ * p.sb2 : nullptr;

  __shared__ __align__(16) uint2 s0[STG];
  __shared__ __align__(16) uint2 s1[STG];
  __shared__ __align__(16) uint2 s2[STAGES > 2 ? STG : 1];
  __shared__ __align__(16) uint2 s3[STAGES > 3 ? STG : 1];

  // direct loads: 16-byte unit u = i * NT + tid of a tile is (row u / CH,
  // chunk u % CH), filled from logical chunk (u % CH) ^ swz(row); rows past m
  // (n) read row m - 1 (n - 1), whose outputs are never stored
  const zb* pa[AL];
  const zb* pb[BL];
#pragma unroll
  for (int i = 0; i < AL; i++) {
    const int u = i * NT + tid, row = u / CH, pc = u % CH;
    pa[i] = A + (long)min(m0 + row, p.m - 1) * p.sam + 2 * (pc ^ b16::swz(row));
  }
#pragma unroll
  for (int i = 0; i < BL; i++) {
    const int u = i * NT + tid, row = u / CH, pc = u % CH;
    pb[i] = B + (long)min(n0 + row, p.n - 1) * p.sbn + 2 * (pc ^ b16::swz(row));
  }
  // third pieces: unit u is slot u % 8 of LDS row u / 8 = 16q + r, which
  // holds chunk x / 2 (four complex) of tile row 32q + 16 (x & 1) + r, where
  // x = (u % 8) ^ swz(r)
  auto row3 = [](int u) {
    const int sr = u >> 3, x = (u & 7) ^ b16::swz(sr & 15);
    return 32 * (sr >> 4) + 16 * (x & 1) + (sr & 15);
  };
  auto chunk3 = [](int u) { return ((u & 7) ^ b16::swz((u >> 3) & 15)) >> 1; };
  const uint32_t* pa3[AL3 > 0 ? AL3 : 1];
  const uint32_t* pb3[BL3 > 0 ? BL3 : 1];
#pragma unroll
  for (int i = 0; i < AL3; i++) {
    const int u = i * NT + tid;
    pa3[i] = A3 + (long)min(m0 + row3(u), p.m - 1) * p.sam + 4 * chunk3(u);
  }
#pragma unroll
  for (int i = 0; i < BL3; i++) {
    const int u = i * NT + tid;
    pb3[i] = B3 + (long)min(n0 + row3(u), p.n - 1) * p.sbn + 4 * chunk3(u);
  }
  const int w0 = __builtin_amdgcn_readfirstlane(wave);
  auto issue = [&](uint2* st, int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < AL; i++)
      __builtin_amdgcn_global_load_lds((const void*)(pa[i] + k0),
                                       (__attribute__((address_space(3))) void*)&st[2 * (i * NT + w0 * 64)], 16, 0,
                                       0);
#pragma unroll
    for (int i = 0; i < BL; i++)
      __builtin_amdgcn_global_load_lds((const void*)(pb[i] + k0),
                                       (__attribute__((address_space(3))) void*)&st[BM * KT + 2 * (i * NT + w0 * 64)],
                                       16, 0, 0);
    uint2* st3 = st + (BM + BN) * KT;  // third pieces: A rows, then B rows (8 uint2 per row)
#pragma unroll
    for (int i = 0; i < AL3; i++)
      __builtin_amdgcn_global_load_lds((const void*)(pa3[i] + k0),
                                       (__attribute__((address_space(3))) void*)&st3[2 * (i * NT + w0 * 64)], 16, 0,
                                       0);
#pragma unroll
    for (int i = 0; i < BL3; i++)
      __builtin_amdgcn_global_load_lds(
          (const void*)(pb3[i] + k0),
          (__attribute__((address_space(3))) void*)&st3[BM * KT / 2 + 2 * (i * NT + w0 * 64)], 16, 0, 0);
  };

  floatx4 cr[TM][TN], ci[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        cr[i][j][q] = 0.f;
        ci[i][j][q] = 0.f;
      }

  // fragment rows: this lane's rows are 16-aligned blocks + r, so swz = swz(r)
  const int sw = b16::swz(r);
  const int c0 = (2 * g) ^ sw, c1 = (2 * g + 1) ^ sw;
  const int arow = wm * WTM + r, brow = BM + wn * WTN + r;
  // third pieces: LDS rows of the wave tile's first block pair (A, then B
  // after A's BM / 2 rows); block i of the pair reads chunk pair c0 / c1
  const int arow3 = wm * WTM / 2 + r, brow3 = BM / 2 + wn * WTN / 2 + r;
  // fragments of one k-tile: A rows (hi and lo words of four complex k) and
  // the raw B columns, read from LDS before the next loads are issued
  // (three pieces: the B columns are read per column block inside the MFMA
  // loop instead, which keeps the register count of eight-wave tiles within
  // the budget)
  constexpr int TNF = NP == 3 ? 1 : TN;
  struct Frags {
    uint4 ah[TM], al[TM], b0[TNF], b1[TNF];
    uint4 a3[NP == 3 ? TM : 1];
  };
  auto read = [&](const uint2* st, Frags& f) __attribute__((always_inline)) {
    const uint4* t4 = (const uint4*)st;  // 8 uint4 per tile row
#pragma unroll
    for (int i = 0; i < TM; i++) {
      const uint4 v0 = t4[(arow + 16 * i) * 8 + c0], v1 = t4[(arow + 16 * i) * 8 + c1];
      f.ah[i] = make_uint4(v0.x, v0.z, v1.x, v1.z);
      f.al[i] = make_uint4(v0.y, v0.w, v1.y, v1.w);
      // one piece uses only the hi words: keep the lo words live so that the
      // reads stay 16 bytes wide (narrowed reads of every other word conflict)
      if constexpr (NP == 1) asm volatile("" ::"v"(v0.y), "v"(v0.w), "v"(v1.y), "v"(v1.w));
    }
    if constexpr (NP != 3) {
#pragma unroll
      for (int j = 0; j < TN; j++) {
        f.b0[j] = t4[(brow + 16 * j) * 8 + c0];
        f.b1[j] = t4[(brow + 16 * j) * 8 + c1];
        if constexpr (NP == 1)
          asm volatile("" ::"v"(f.b0[j].y), "v"(f.b0[j].w), "v"(f.b1[j].y), "v"(f.b1[j].w));
      }
    } else {
      const uint4* t3 = (const uint4*)(st + (BM + BN) * KT);  // 8 uint4 per LDS row
#pragma unroll
      for (int i = 0; i < TM; i++) f.a3[i] = t3[(arow3 + 16 * (i >> 1)) * 8 + ((i & 1) ? c1 : c0)];
    }
  };
  auto mfmas = [&](const Frags& f, const uint2* st) __attribute__((always_inline)) {
    // real part: (b_re, -b_im); imaginary part: (b_im, b_re)
    auto ng = [](uint32_t w) { return w ^ 0x80000000u; };
    auto rt = [](uint32_t w) { return __builtin_amdgcn_alignbit(w, w, 16u); };
    auto mm = [&](int i, int j, const uint4& a, const uint4& br, const uint4& bi) __attribute__((always_inline)) {
      const bf16x8 xa = __builtin_bit_cast(bf16x8, a);
      cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa, __builtin_bit_cast(bf16x8, br), cr[i][j], 0, 0, 0);
      ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa, __builtin_bit_cast(bf16x8, bi), ci[i][j], 0, 0, 0);
    };
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const uint4* t4 = (const uint4*)st;
      const uint4 v0 = NP != 3 ? f.b0[j % TNF] : t4[(brow + 16 * j) * 8 + c0];
      const uint4 v1 = NP != 3 ? f.b1[j % TNF] : t4[(brow + 16 * j) * 8 + c1];
      if constexpr (NP == 1) {
        const uint4 rh = make_uint4(ng(v0.x), ng(v0.z), ng(v1.x), ng(v1.z));
        const uint4 ih = make_uint4(rt(v0.x), rt(v0.z), rt(v1.x), rt(v1.z));
#pragma unroll
        for (int i = 0; i < TM; i++) mm(i, j, f.ah[i], rh, ih);
      } else if constexpr (NP == 2) {
        const uint4 rh = make_uint4(ng(v0.x), ng(v0.z), ng(v1.x), ng(v1.z));
        const uint4 rl = make_uint4(ng(v0.y), ng(v0.w), ng(v1.y), ng(v1.w));
        const uint4 ih = make_uint4(rt(v0.x), rt(v0.z), rt(v1.x), rt(v1.z));
        const uint4 il = make_uint4(rt(v0.y), rt(v0.w), rt(v1.y), rt(v1.w));
        // h*h, h*m, m*h (the expanded kernel's order)
#pragma unroll
        for (int t = 0; t < 3; t++)
#pragma unroll
          for (int i = 0; i < TM; i++) mm(i, j, t == 2 ? f.al[i] : f.ah[i], t == 1 ? rl : rh, t == 1 ? il : ih);
      } else {
        // grouped by B's piece so that one piece's two forms are live at a
        // time: h*h, m*h, l*h, then h*m, m*m, then h*l
        {
          const uint4 rh = make_uint4(ng(v0.x), ng(v0.z), ng(v1.x), ng(v1.z));
          const uint4 ih = make_uint4(rt(v0.x), rt(v0.z), rt(v1.x), rt(v1.z));
#pragma unroll
          for (int t = 0; t < 3; t++)
#pragma unroll
            for (int i = 0; i < TM; i++) mm(i, j, t == 0 ? f.ah[i] : t == 1 ? f.al[i] : f.a3[i], rh, ih);
        }
        {
          const uint4 rl = make_uint4(ng(v0.y), ng(v0.w), ng(v1.y), ng(v1.w));
          const uint4 il = make_uint4(rt(v0.y), rt(v0.w), rt(v1.y), rt(v1.w));
#pragma unroll
          for (int t = 0; t < 2; t++)
#pragma unroll
            for (int i = 0; i < TM; i++) mm(i, j, t == 0 ? f.ah[i] : f.al[i], rl, il);
        }
        {
          const uint4 w = ((const uint4*)(st + (BM + BN) * KT))[(brow3 + 16 * (j >> 1)) * 8 + ((j & 1) ? c1 : c0)];
          const uint4 r3 = make_uint4(ng(w.x), ng(w.y), ng(w.z), ng(w.w));
          const uint4 i3 = make_uint4(rt(w.x), rt(w.y), rt(w.z), rt(w.w));
#pragma unroll
          for (int i = 0; i < TM; i++) mm(i, j, f.ah[i], r3, i3);
        }
      }
    }
  };

  const int nk = (kend + KT - 1) / KT;
  // Step j issues tile j into stage j % STAGES and computes tile
  // c = j - (STAGES - 1) from stage c % STAGES.  The first STAGES - 1 steps
  // only issue, so the loads of every tile come from the same instructions
  // (the compiler's tracking of which stage a load writes then holds across
  // the loop head, and it adds no wait of its own).  In a computing step:
  // wait until tile c has landed (later tiles stay in flight), barrier (every
  // wave has also finished reading tile c - 1, whose stage tile j refills),
  // read the fragments, issue tile j, run the MFMAs.
  auto step = [&](auto stc, int j) __attribute__((always_inline)) {
    constexpr int ST = decltype(stc)::value;           // stage of tile j
    constexpr int SC = (ST + 1) % STAGES;              // stage of tile c
    uint2* in_ = ST == 0 ? s0 : ST == 1 ? s1 : ST == 2 ? s2 : s3;
    const uint2* cs = SC == 0 ? s0 : SC == 1 ? s1 : SC == 2 ? s2 : s3;
    const int c = j - (STAGES - 1);
    Frags f;
    if (c >= 0) {
      const int ahead = min(nk - 1 - c, STAGES - 2);  // tiles issued after c
      if (ahead >= 2) b16::wait_vmcnt<(STAGES > 3 ? 2 * LPT : 0)>();
      else if (ahead == 1) b16::wait_vmcnt<(STAGES > 2 ? LPT : 0)>();
      else b16::wait_vmcnt<0>();
      // a bare barrier: a workgroup fence here would wait for the later
      // tiles' loads too
      __builtin_amdgcn_s_barrier();
      read(cs, f);
    }
    if (j < nk) issue(in_, j * KT);
    if (c >= 0) mfmas(f, cs);
  };
  for (int j = 0; j < nk + STAGES - 1; j += STAGES) {
    step(std::integral_constant<int, 0>(), j);
    if (j + 1 < nk + STAGES - 1) step(std::integral_constant<int, 1>(), j + 1);
    if constexpr (STAGES > 2)
      if (j + 2 < nk + STAGES - 1) step(std::integral_constant<int, 2>(), j + 2);
    if constexpr (STAGES > 3)
      if (j + 3 < nk + STAGES - 1) step(std::integral_constant<int, 3>(), j + 3);
  }

  // epilogue: the lane holds (re, im) of complex element (row 4g + q, column
  // r) of each 16 x 16 block; arithmetic in FP64 as zgemm_kernel's epilogue
  constexpr bool OSPLIT = std::is_same<OT, bfs>::value;
  using DT = typename std::conditional<OSPLIT, float, OT>::type;
  const long coff = 2 * (bq * p.sc1 + br_ * p.sc2);
  DT* C = (p.C && !OSPLIT) ? (DT*)p.C + coff : nullptr;
  zb* CS = (p.C && OSPLIT) ? (zb*)p.C + coff / 2 : nullptr;
  // three-piece split output: the third pieces beside C
  uint32_t* CS3 = (OSPLIT && p.C3) ? (uint32_t*)p.C3 + coff / 2 : nullptr;
  DT* D = p.D ? (DT*)p.D + coff : nullptr;
  const DT* F = p.F ? (const DT*)p.F + coff : nullptr;
  const bool has_beta = (p.br != 0.0 || p.bi != 0.0);
  const int rbase = m0 + wm * WTM + 4 * g;
  const int cbase = n0 + wn * WTN + r;
#pragma unroll
  for (int i = 0; i < TM; i++) {
    DT fr[TN][4], fi[TN][4], vr[TN][4], vi[TN][4];
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = rbase + i * 16 + q, col = cbase + j * 16;
        const bool in = row < p.m && col < p.n;
        const long off = 2 * ((long)row * p.scm + (long)col * p.scn);
        const long foff = (p.sfm || p.sfn) ? 2 * ((long)row * p.sfm + (long)col * p.sfn) : off;
        const bool lf = (D || p.cf) && F && in;
        fr[j][q] = lf ? F[foff] : (DT)0;
        fi[j][q] = lf ? F[foff + 1] : (DT)0;
        const bool lc = !OSPLIT && C && has_beta && in;
        vr[j][q] = lc ? C[off] : (DT)0;
        vi[j][q] = lc ? C[off + 1] : (DT)0;
      }
    // column-major split output (the Taylor steps' transposed terms): the
    // lane's four rows are 32 contiguous bytes, stored as two 16-byte words
    bool vec = false;
    if constexpr (OSPLIT) vec = CS && !CS3 && !D && p.scm == 1 && rbase + i * 16 + 3 < p.m;
    if (vec) {
#pragma unroll
        for (int j = 0; j < TN; j++) {
          const int col = cbase + j * 16;
          if (col >= p.n) continue;
          auto piece = [&](int q) __attribute__((always_inline)) -> zb {
            const double re = (double)cr[i][j][q], im = (double)ci[i][j][q];
            const double ore = p.ar * re - p.ai * im, oim = p.ar * im + p.ai * re;
            const double f_r = (double)fr[j][q], f_i = (double)fi[j][q];
            const double sr = p.cf ? f_r + ore : ore, si = p.cf ? f_i + oim : oim;
            return split_z((float)sr, (float)si);
          };
          const zb v0 = piece(0), v1 = piece(1), v2 = piece(2), v3 = piece(3);
          zb* dz = CS + (long)(rbase + i * 16) + (long)col * p.scn;
          if (((uintptr_t)dz & 15) == 0) {
            uint4* dst = (uint4*)dz;
            dst[0] = make_uint4(v0.hi, v0.lo, v1.hi, v1.lo);
            dst[1] = make_uint4(v2.hi, v2.lo, v3.hi, v3.lo);
          } else {
            dz[0] = v0;
            dz[1] = v1;
            dz[2] = v2;
            dz[3] = v3;
          }
        }
    } else {
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = rbase + i * 16 + q, col = cbase + j * 16;
        if (!(row < p.m && col < p.n)) continue;
        const double re = (double)cr[i][j][q], im = (double)ci[i][j][q];
        const double ore = p.ar * re - p.ai * im, oim = p.ar * im + p.ai * re;  // alpha * AB
        double ocr = ore, oci = oim;                                             // + beta * C
        if (!OSPLIT && has_beta) {
          const double cre = (double)vr[j][q], cim = (double)vi[j][q];
          ocr += p.br * cre - p.bi * cim;
          oci += p.br * cim + p.bi * cre;
        }
        const long off = 2 * ((long)row * p.scm + (long)col * p.scn);
        const double f_r = (double)fr[j][q], f_i = (double)fi[j][q];
        if (D) {
          D[off] = (DT)(f_r + ore);
          D[off + 1] = (DT)(f_i + oim);
        }
        if constexpr (OSPLIT) {
          if (CS) {
            const double sr = p.cf ? f_r + ore : ore, si = p.cf ? f_i + oim : oim;
            if (CS3) {
              zb h;
              uint32_t t;
              split3_f((float)sr, (float)si, h, t);
              CS[off / 2] = h;
              CS3[off / 2] = t;
            } else {
              CS[off / 2] = split_z((float)sr, (float)si);
            }
          }
        } else if (C) {
          C[off] = (DT)(p.cf ? ocr + f_r : ocr);
          C[off + 1] = (DT)(p.cf ? oci + f_i : oci);
        }
      }
    }
  }
}
