// Fused split-BF16 propagator: W <- P * E6(V) * P * W for one walker and one
// block of 32 of its 2N columns (both spins side by side) per workgroup
// (afqmc_sys::propagate, src/AFQMC/afqmc_sys.hpp:146-173, with apply_expM,
// src/AFQMC/vHS.hpp:61-91, summed by Horner's rule).
//
// A column block of the product only needs the same column block of its
// right operand, so the eight products of the chain (S = P W, the six Horner
// steps U_k = S + (i/k) V U_{k+1}, W = P U_1) stay inside the workgroup: U
// lives in LDS as split-BF16 [column][row], S in global memory (FP32, read
// back in each Horner epilogue as the chain's GEMMs read it), and the
// left operands (P, the walker's V) go from global memory straight into MFMA
// fragments.  The arithmetic (split of the operands, MFMA order, epilogues in
// FP64) is that of the GEMM chain it replaces, so the walkers it produces
// are the same bit for bit.
#pragma once
#include "common.hpp"

constexpr int PF_NB = 32;     // complex columns per workgroup
constexpr int PF_MMAX = 512;  // largest M (LDS holds U for 32 columns x M rows)
constexpr int PF_LDU = PF_MMAX + 4;

struct PropFusedArgs {
  zd* W;          // [nw][2][M][N], FP64, updated in place
  float* S;       // [nw][2][M][N] complex FP32 scratch for S = P W
  const zb* P;    // [M][M] split-BF16 Propg1
  const zb* V;    // [nw][M][M] split-BF16 vHS (row i, column k) per walker
  int M, N, nw;
  int ncb;        // column blocks per walker, 2N / 32
};

// Each of the NWV waves owns 16*TM rows (TM 16-row tiles) and all 32
// columns (4 tiles of 8 complex columns); lanes follow the GEMM engine's
// split-BF16 fragment layout.  Requires M % 64 == 0, M <= 512,
// 16*TM*NWV >= M, N % 16 == 0.
template <int TM, int NWV>
__global__ __launch_bounds__(64 * NWV) void prop_fused_kernel(PropFusedArgs a)
{
  constexpr int NT = 64 * NWV;
  constexpr int TN = 4;
  __shared__ __align__(16) uint2 Us[PF_NB * PF_LDU];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, r = lane & 15, s = r & 1;
  const int M = a.M, N = a.N;
  // walker-major over the XCDs: the column blocks of one walker run on one
  // XCD, so its V is fetched into one L2
  long lin = blockIdx.x;
  {
    const long T = gridDim.x, per = T / 8;
    if (per > 0 && lin < per * 8) lin = (lin % 8) * per + lin / 8;
  }
  const int w = (int)(lin / a.ncb), cb = (int)(lin % a.ncb);
  zd* Ww = a.W + (long)w * 2 * M * N;
  const zb* Vw = a.V + (long)w * M * M;
  const int c0 = cb * PF_NB;  // first column of the block, column c = (spin c / N, orbital c % N)

  // stage B = W[:, block] as split-BF16, [column][row]
  for (int e = tid; e < PF_NB * M; e += NT) {
    const int cc = e % PF_NB, i = e / PF_NB, c = c0 + cc;
    const zd v = Ww[((long)(c / N) * M + i) * N + (c % N)];
    const zb z = split_z((float)v.re, (float)v.im);
    Us[cc * PF_LDU + i] = make_uint2(z.hi, z.lo);
  }
  __syncthreads();

  const int row0 = wave * 16 * TM;  // first row of this wave
  const uint32_t brot = s ? 16u : 0u, bneg = s ? 0u : 0x80000000u;
  floatx4 acc[TM][TN];
  float* Sw = a.S + 2L * w * 2 * M * N;
  // element (row, column c) of W / S within the walker: (c / N) * M * N +
  // row * N + c % N; the column part of this lane's TN columns, once
  int cbase[TN];
#pragma unroll
  for (int j = 0; j < TN; j++) {
    const int c = c0 + (r >> 1) + j * 8;
    cbase[j] = (c / N) * M * N + (c % N);
  }

  // acc = A * U (A [M][M] split-BF16, row-major) for this wave's rows
  auto product = [&](const zb* A) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const zb* arow[TM];
#pragma unroll
    for (int i = 0; i < TM; i++) {
      const int row = row0 + i * 16 + r;
      arow[i] = A + (long)(row < M ? row : 0) * M + 4 * g;
    }
    // (a register prefetch of the next k-step's A fragments measured slower:
    // it spills at 16 waves)
    for (int k0 = 0; k0 < M; k0 += 16) {
      uint4 ah[TM], al[TM];
#pragma unroll
      for (int i = 0; i < TM; i++) {
        const uint4* q = (const uint4*)(arow[i] + k0);
        uint4 u0 = q[0], u1 = q[1];
        if (row0 + i * 16 + r >= M) u0 = u1 = make_uint4(0, 0, 0, 0);
        ah[i] = make_uint4(u0.x, u0.z, u1.x, u1.z);
        al[i] = make_uint4(u0.y, u0.w, u1.y, u1.w);
      }
      // one column tile of B at a time: its three products over the row
      // tiles, in the order of the GEMM engine (t, then i) per accumulator
#pragma unroll
      for (int j = 0; j < TN; j++) {
        const uint4* q = (const uint4*)&Us[((r >> 1) + j * 8) * PF_LDU + k0 + 4 * g];
        const uint4 u0 = q[0], u1 = q[1];
        auto ex = [&](uint32_t x) { return __builtin_amdgcn_alignbit(x, x, brot) ^ bneg; };
        const uint4 bh = make_uint4(ex(u0.x), ex(u0.z), ex(u1.x), ex(u1.z));
        const uint4 bl = make_uint4(ex(u0.y), ex(u0.w), ex(u1.y), ex(u1.w));
#pragma unroll
        for (int t = 0; t < 3; t++)
#pragma unroll
          for (int i = 0; i < TM; i++) {
            const bf16x8 xa = __builtin_bit_cast(bf16x8, t == 2 ? al[i] : ah[i]);
            const bf16x8 xb = __builtin_bit_cast(bf16x8, t == 1 ? bl : bh);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa, xb, acc[i][j], 0, 0, 0);
          }
      }
    }
  };
  // U = split(S + (0 + i ai) * acc) into LDS (ai = 0: U = S); lane s = 0 of
  // each pair writes the complex value
  auto store_u = [&](double ai, bool first) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int row = row0 + i * 16 + 4 * g + q, cc = (r >> 1) + j * 8;
          const double sv = row < M ? (double)Sw[2 * (cbase[j] + row * N) + s] : 0.0;
          double ofc;
          if (first) {
            ofc = sv;
          } else {
            const double v = (double)acc[i][j][q];
            const double pv = __shfl_xor(v, 1);
            const double re = s ? pv : v, im = s ? v : pv;
            const double o = s ? (0.0 * im + ai * re) : (0.0 * re - ai * im);
            ofc = sv + o;
          }
          const double oi = __shfl_xor(ofc, 1);
          if (s == 0 && row < M) {
            const zb z = split_z((float)ofc, (float)oi);
            Us[cc * PF_LDU + row] = make_uint2(z.hi, z.lo);
          }
          // one element at a time: keeps the epilogue's live values few
          __builtin_amdgcn_sched_barrier(0);
        }
  };

  // S = P W
  product(a.P);
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = row0 + i * 16 + 4 * g + q;
        const double v = (double)acc[i][j][q];
        const double pv = __shfl_xor(v, 1);
        const double re = s ? pv : v, im = s ? v : pv;
        if (row < M) Sw[2 * (cbase[j] + row * N) + s] = (float)(s ? (1.0 * im + 0.0 * re) : (1.0 * re - 0.0 * im));
      }
  __syncthreads();  // S in global memory is read back below by the lanes that wrote it
  store_u(0.0, true);  // U_7 = S
  __syncthreads();
  for (int k = 6; k >= 1; k--) {
    product(Vw);
    __syncthreads();
    store_u(1.0 / (double)k, false);
    __syncthreads();
  }
  // W = P U_1
  product(a.P);
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = row0 + i * 16 + 4 * g + q;
        const double v = (double)acc[i][j][q];
        const double pv = __shfl_xor(v, 1);
        const double re = s ? pv : v, im = s ? v : pv;
        const double o = s ? (1.0 * im + 0.0 * re) : (1.0 * re - 0.0 * im);
        if (row < M) {
          double* dst = (double*)&Ww[cbase[j] + row * N];
          dst[s] = o;
        }
      }
}
