// Non-GEMM kernels of the walker step (FP64 unless stated).
#pragma once
#include "common.hpp"
#include "fieldrng.hpp"

// ---------------------------------------------------------------------------
// Batched in-place Gauss-Jordan inverse + determinant of N x N complex
// matrices, FP64.  Follows LAPACK's pivoting on the column-major view used by
// ma::invert / ma::determinant (src/Numerics/ma_operations.hpp:244-291,
// ma_lapack.hpp:33-79): the matrix factorised is X = T^T, the pivot of step k
// is the first row i >= k maximising |Re|+|Im| of X(i,k) (izamax), and
// det = prod(pivots) * (-1)^(#swaps).  The Gauss-Jordan eliminations below
// the diagonal are those of the LU factorisation, so pivots and determinant
// match getrf; the inverse equals getri's up to rounding.
//
// Thread (i, cg) of the NMAX*CG threads holds row i, columns cg + CG*t of X in
// registers (NMAX/CG complex values).  Rows >= N are padded with the
// identity.  Per elimination step: holders of column k publish |X(i,k)|, every
// wave finds the pivot row p, rows p and k and column k go through LDS, and
// each thread updates its own entries.  The signed pivots also give the
// log-determinant (store_log_det).
// batch b: input T + b*mstride; det written to det[(b/2)*dq + (b%2)*dr] and
// to dcache[b] when dcache != null.
// log det = sum_k log(pv_k), principal logs, imaginary part wrapped to
// [-pi, pi], from the signed pivots pv[0..N) in LDS; summed by one wave in a
// fixed order (every lane returns the same value).  The overlap ratios of
// the weights are taken from these logs, so that a determinant that
// underflows or overflows as a product of N pivots keeps its ratio.
__device__ inline zd log_det_sum(const zd* pv, int N)
{
  const int lane = threadIdx.x & 63;
  double lr = 0.0, li = 0.0;
  for (int q = lane; q < N; q += 64) {
    const zd v = pv[q];
    lr += log(hypot(v.re, v.im));
    li += atan2(v.im, v.re);
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    lr += __shfl_xor(lr, off);
    li += __shfl_xor(li, off);
  }
  return zd{lr, remainder(li, 6.283185307179586476925286766559)};
}

// log det of matrix b (log_det_sum of its signed pivots, {-inf, 0} when
// singular) to ldet[(b/2)*4 + b%2] and lcache[b] where those are set;
// called by every thread after the elimination.
__device__ inline void store_log_det(const zd* pv, int N, bool singular, long b, zd* ldet, zd* lcache)
{
  if ((!ldet && !lcache) || threadIdx.x >= 64) return;
  const zd L = singular ? zd{-INFINITY, 0.0} : log_det_sum(pv, N);
  if (threadIdx.x == 0) {
    if (ldet) ldet[(b / 2) * 4 + (b % 2)] = L;
    if (lcache) lcache[b] = L;
  }
}

// Wave argmax of (v, i) pairs, larger v first, then smaller i (an
// associative, commutative choice, so any pairing order gives the same
// pair): DPP moves within quads, half-rows and rows, then row broadcasts;
// lane 63 holds the wave's pair, and i is returned in every lane.
template <int CTRL, int RM> __device__ __forceinline__ void amx_step(double& bv, int& bi)
{
  const long long u = __double_as_longlong(bv);
  const int lo = (int)u, hi = (int)(u >> 32);
  const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, RM, 0xF, false);
  const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, RM, 0xF, false);
  const int oi = __builtin_amdgcn_update_dpp(bi, bi, CTRL, RM, 0xF, false);
  const double ov = __longlong_as_double((long long)(((unsigned long long)(unsigned)ohi << 32) | (unsigned)olo));
  if (ov > bv || (ov == bv && oi < bi)) {
    bv = ov;
    bi = oi;
  }
}
__device__ __forceinline__ int wave_argmax(double bv, int bi)
{
  amx_step<0xB1, 0xF>(bv, bi);   // quad_perm [1,0,3,2]
  amx_step<0x4E, 0xF>(bv, bi);   // quad_perm [2,3,0,1]
  amx_step<0x141, 0xF>(bv, bi);  // row_half_mirror
  amx_step<0x140, 0xF>(bv, bi);  // row_mirror
  amx_step<0x142, 0xA>(bv, bi);  // row_bcast15 into rows 1 and 3
  amx_step<0x143, 0xC>(bv, bi);  // row_bcast31 into rows 2 and 3
  return __builtin_amdgcn_readlane(bi, 63);
}

template <int NMAX, int CG>
__global__ __launch_bounds__(NMAX* CG) void gj_inverse_kernel(const zd* __restrict__ T, zd* __restrict__ Tinv,
                                                               zd* __restrict__ det, long dq, long dr, int N,
                                                               long mstride, int* __restrict__ info,
                                                               zd* __restrict__ dcache, zd* __restrict__ ldet,
                                                               zd* __restrict__ lcache)
{
  constexpr int NC = NMAX / CG;
  __shared__ double red[NMAX];
  __shared__ zd prow[NMAX], oldk[NMAX], colk[NMAX], pv[NMAX];
  __shared__ int perm[NMAX], dest[NMAX];
  const int tid = threadIdx.x;
  const int i = tid / CG, cg = tid % CG;
  const int lane = tid & 63;
  const long b = blockIdx.x;
  const zd* Tb = T + b * mstride;
  zd x[NC];
#pragma unroll
  for (int t = 0; t < NC; t++) {
    int j = cg + CG * t;
    if (i < N && j < N) x[t] = Tb[(long)j * N + i];  // X = T^T
    else x[t] = zd{(i == j) ? 1.0 : 0.0, 0.0};
  }
  zd dt{1.0, 0.0};
  bool singular = false;
  // Columns are eliminated in blocks of CG: block tk holds columns CG*tk ..
  // CG*tk + CG - 1, one per thread of a row, all in register slot tk.  For the
  // CG steps of a block that slot is moved to xc, a register addressed
  // statically, so only xc needs the column-k special case; the other slots
  // take the plain update (slot tk's copy is refreshed from xc at the end).
  const int nblk = (N + CG - 1) / CG;
  for (int tk = 0; tk < nblk && !singular; tk++) {
    zd xc{0, 0};
#pragma unroll
    for (int t = 0; t < NC; t++)
      if (t == tk) xc = x[t];
    const int jc = cg + CG * tk;  // the column xc holds
#pragma unroll
    for (int ck = 0; ck < CG; ck++) {
      const int k = CG * tk + ck;
      if (k >= N) break;
      const bool own = (cg == ck);  // xc holds column k
      if (own) {
        red[i] = (i >= k) ? zcabs1(xc) : -1.0;
        colk[i] = xc;
      }
      __syncthreads();
      double bv = -2.0;
      int bi = 0;
#pragma unroll
      for (int q = lane; q < NMAX; q += 64) {
        double v = red[q];
        if (v > bv) { bv = v; bi = q; }
      }
      const int p = wave_argmax(bv, bi);
      if (i == p) {
#pragma unroll
        for (int t = 0; t < NC; t++) prow[cg + CG * t] = x[t];
        prow[jc] = xc;  // after slot tk's stale copy: the later store wins
        if (own) pv[k] = (p != k) ? zd{-xc.re, -xc.im} : xc;  // signed pivot
      }
      if (i == k && p != k) {
#pragma unroll
        for (int t = 0; t < NC; t++) oldk[cg + CG * t] = x[t];
        oldk[jc] = xc;
      }
      if (tid == 0) perm[k] = p;
      __syncthreads();
      const zd piv = prow[k];
      if (piv.re == 0.0 && piv.im == 0.0) {
        singular = true;
        dt = zd{0.0, 0.0};
        break;
      }
      dt = zmul(dt, (p != k) ? zd{-piv.re, -piv.im} : piv);
      const zd ipv = zdiv(zd{1.0, 0.0}, piv);
      if (i == k) {
        // new row k = pivot row / piv; column k holds 1/piv
#pragma unroll
        for (int t = 0; t < NC; t++) x[t] = zmul(prow[cg + CG * t], ipv);
        xc = own ? ipv : zmul(prow[jc], ipv);
      } else {
        const bool swapped_in = (i == p);  // p != k here since i != k
        if (swapped_in) {
#pragma unroll
          for (int t = 0; t < NC; t++) x[t] = oldk[cg + CG * t];
          xc = oldk[jc];
        }
        // X(i,k) of the row now at i: the swapped-in row takes old row k's,
        // from oldk (written between this step's barriers), not colk[k],
        // which row k's owner overwrites at the start of the next step with
        // no barrier in between (row k may sit in another wave)
        const zd f = swapped_in ? oldk[k] : colk[i];
        const zd gq = zmul(zd{-f.re, -f.im}, ipv);  // -X(i,k)/piv
#pragma unroll
        for (int t = 0; t < NC; t++) {
          const zd pr = prow[cg + CG * t];
          const double xr = x[t].re, xi = x[t].im;
          x[t].re = xr + gq.re * pr.re - gq.im * pr.im;
          x[t].im = xi + gq.re * pr.im + gq.im * pr.re;
        }
        // column k: X(i,k) = -X(i,k)/piv, i.e. 0 + gq * 1
        const zd pc = prow[jc];
        const double xr = own ? 0.0 : xc.re, xi = own ? 0.0 : xc.im;
        const double pr_r = own ? 1.0 : pc.re, pr_i = own ? 0.0 : pc.im;
        xc.re = xr + gq.re * pr_r - gq.im * pr_i;
        xc.im = xi + gq.re * pr_i + gq.im * pr_r;
      }
    }
#pragma unroll
    for (int t = 0; t < NC; t++)
      if (t == tk) x[t] = xc;
  }
  if (tid == 0) {
    if (singular && info) atomicAdd(info, 1);
    det[(b / 2) * dq + (b % 2) * dr] = dt;
    if (dcache) dcache[b] = dt;
  }
  // the log-determinant last, when the matrix no longer occupies registers
  if (!Tinv) {
    store_log_det(pv, N, singular, b, ldet, lcache);
    return;
  }
  if (singular) {
    // leave a defined value: zeros
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NC; t++) {
      int j = cg + CG * t;
      if (i < N && j < N) Tinv[b * mstride + (long)j * N + i] = zd{0, 0};
    }
    store_log_det(pv, N, singular, b, ldet, lcache);
    return;
  }
  // undo the row interchanges as column interchanges in reverse order
  __syncthreads();
  if (tid == 0) {
    for (int c = 0; c < N; c++) dest[c] = c;  // dest used as colmap first
    for (int k = N - 1; k >= 0; k--) {
      int pk = perm[k];
      int tmp = dest[k];
      dest[k] = dest[pk];
      dest[pk] = tmp;
    }
    for (int c = 0; c < N; c++) perm[dest[c]] = c;  // perm now = inverse map
  }
  __syncthreads();
  // Z(:, c) = Y(:, colmap[c]) -> Y column j goes to column perm[j];
  // T^-1[a][b] = Z[b][a]
#pragma unroll
  for (int t = 0; t < NC; t++) {
    int j = cg + CG * t;
    if (i < N && j < N) Tinv[b * mstride + (long)perm[j] * N + i] = x[t];
  }
  store_log_det(pv, N, singular, b, ldet, lcache);
}

// The same Gauss-Jordan elimination for N too large for the registers of one
// workgroup (N <= GJ_GMAX): the working matrix X = T^T lives in global memory
// (work + b*mstride, X(i, j) at [j*N + i], i.e. column j of X contiguous),
// one workgroup of 1024 threads per matrix; per step the pivot column, the
// pivot row and the old row k go through LDS and every thread updates its
// entries of X with the register kernel's arithmetic, so pivots, determinant
// and inverse follow the same operations.  The result is written to Tinv with
// the row interchanges undone as column interchanges; T may equal Tinv (work
// must not).
constexpr int GJ_GMAX = 256;
__global__ __launch_bounds__(1024) void gj_inverse_global_kernel(const zd* __restrict__ T, zd* __restrict__ Tinv,
                                                                  zd* __restrict__ work, zd* __restrict__ det,
                                                                  long dq, long dr, int N, long mstride,
                                                                  int* __restrict__ info, zd* __restrict__ dcache,
                                                                  zd* __restrict__ ldet, zd* __restrict__ lcache)
{
  __shared__ double red[GJ_GMAX];
  __shared__ zd prow[GJ_GMAX], oldk[GJ_GMAX], colk[GJ_GMAX], pv[GJ_GMAX];
  __shared__ int perm[GJ_GMAX], dest[GJ_GMAX];
  __shared__ int piv_row;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long b = blockIdx.x;
  const zd* Tb = T + b * mstride;
  zd* X = work + b * mstride;
  const long NN = (long)N * N;
  // X(i, j) = T[j][i]: X's column j is T's row j, so the copy is contiguous
  for (long e = tid; e < NN; e += 1024) X[e] = Tb[e];
  __syncthreads();
  zd dt{1.0, 0.0};
  bool singular = false;
  for (int k = 0; k < N; k++) {
    // pivot: first i >= k maximising |Re| + |Im| of X(i, k)
    for (int i = tid; i < GJ_GMAX; i += 1024) {
      const zd xk = i < N ? X[(long)k * N + i] : zd{0, 0};
      if (i < N) colk[i] = xk;
      red[i] = (i >= k && i < N) ? zcabs1(xk) : -1.0;
    }
    __syncthreads();
    if (wave == 0) {
      double bv = -2.0;
      int bi = 0;
      for (int q = lane; q < GJ_GMAX; q += 64) {
        const double v = red[q];
        if (v > bv) { bv = v; bi = q; }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(bv, off);
        const int oi = __shfl_xor(bi, off);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
      }
      if (lane == 0) {
        piv_row = bi;
        perm[k] = bi;
      }
    }
    __syncthreads();
    const int p = piv_row;
    for (int j = tid; j < N; j += 1024) {
      prow[j] = X[(long)j * N + p];
      if (p != k) oldk[j] = X[(long)j * N + k];
    }
    __syncthreads();
    const zd piv = prow[k];
    if (piv.re == 0.0 && piv.im == 0.0) {
      singular = true;
      dt = zd{0.0, 0.0};
      break;
    }
    dt = zmul(dt, (p != k) ? zd{-piv.re, -piv.im} : piv);
    if (tid == 0) pv[k] = (p != k) ? zd{-piv.re, -piv.im} : piv;  // signed pivot
    const zd ipv = zdiv(zd{1.0, 0.0}, piv);
    for (long e = tid; e < NN; e += 1024) {
      const int j = (int)(e / N), i = (int)(e % N);
      zd x;
      if (i == k) {
        // new row k = pivot row / piv; column k holds 1/piv
        x = (j == k) ? ipv : zmul(prow[j], ipv);
      } else {
        const bool swapped_in = (i == p);  // p != k here since i != k
        const zd xo = swapped_in ? oldk[j] : X[e];
        const zd f = swapped_in ? colk[k] : colk[i];
        const zd gq = zmul(zd{-f.re, -f.im}, ipv);  // -X(i,k)/piv
        const zd pr = prow[j];
        const bool kk = (j == k);
        const double xr = kk ? 0.0 : xo.re, xi = kk ? 0.0 : xo.im;
        const double pr_r = kk ? 1.0 : pr.re, pr_i = kk ? 0.0 : pr.im;
        x.re = xr + gq.re * pr_r - gq.im * pr_i;
        x.im = xi + gq.re * pr_i + gq.im * pr_r;
      }
      X[e] = x;
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (singular && info) atomicAdd(info, 1);
    det[(b / 2) * dq + (b % 2) * dr] = dt;
    if (dcache) dcache[b] = dt;
  }
  store_log_det(pv, N, singular, b, ldet, lcache);
  if (!Tinv) return;
  zd* Ob = Tinv + b * mstride;
  if (singular) {
    __syncthreads();
    for (long e = tid; e < NN; e += 1024) Ob[e] = zd{0, 0};
    return;
  }
  __syncthreads();
  if (tid == 0) {
    for (int c = 0; c < N; c++) dest[c] = c;  // column map first
    for (int k = N - 1; k >= 0; k--) {
      const int pk = perm[k];
      const int tmp = dest[k];
      dest[k] = dest[pk];
      dest[pk] = tmp;
    }
    for (int c = 0; c < N; c++) perm[dest[c]] = c;  // inverse map
  }
  __syncthreads();
  // T^-1[perm[j]][i] = X(i, j)
  for (long e = tid; e < NN; e += 1024) {
    const int j = (int)(e / N), i = (int)(e % N);
    Ob[(long)perm[j] * N + i] = X[e];
  }
}

// Three-piece split-BF16 of n complex values: (h, m) to hm (a zb per
// element; may alias the input when IT is zf, each element being read before
// its slot is written by the same thread) and the third pieces to l.
template <class IT>
__global__ void split3_kernel(const IT* in, zb* hm, uint32_t* __restrict__ l, long n)
{
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const IT v = in[i];
    zb h;
    uint32_t t;
    split3_z((double)v.re, (double)v.im, h, t);
    hm[i] = h;
    l[i] = t;
  }
}

// Third pieces of a split-BF16 array from its FP64 source: the stored (h, m)
// pieces are taken as they are, l = bf16(x - h - m) per component.  Element o
// of the split array (and of l) comes from src[(o % R) * C + o / R] when
// trans (the split array is the [C][R] transpose of the [R][C] source), else
// from src[o].
__global__ void third_piece_kernel(const zd* __restrict__ src, const zb* __restrict__ hm, uint32_t* __restrict__ l,
                                   long n, long R, long C, int trans)
{
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < n; o += (long)gridDim.x * blockDim.x) {
    const zd x = src[trans ? (o % R) * C + o / R : o];
    const zb v = hm[o];
    const double rr = x.re - (double)bf2f((unsigned short)(v.hi & 0xffff)) - (double)bf2f((unsigned short)(v.lo & 0xffff));
    const double ri = x.im - (double)bf2f((unsigned short)(v.hi >> 16)) - (double)bf2f((unsigned short)(v.lo >> 16));
    l[o] = (uint32_t)d2bf(rr) | ((uint32_t)d2bf(ri) << 16);
  }
}

// Batched transpose into three-piece split-BF16: in [b][R][C] (row stride C,
// item stride R*C) -> out [b][C][R] (zb) and its third pieces out3, per
// 32 x 32 tile through LDS; grid (C/32, R/32, b).
template <class IT>
__global__ __launch_bounds__(256) void transpose_split3_kernel(const IT* __restrict__ in, zb* __restrict__ out,
                                                               uint32_t* __restrict__ out3, int R, int C)
{
  __shared__ zd tile[32][33];
  const long item = (long)blockIdx.z * R * C;
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int yy = ty; yy < 32; yy += 8) {
    const int rr = by + yy, cc = bx + tx;
    if (rr < R && cc < C) tile[yy][tx] = to_zd(in[item + (long)rr * C + cc]);
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int cc = bx + yy, rr = by + tx;
    if (rr < R && cc < C) {
      const zd v = tile[tx][yy];
      const long o = item + (long)cc * R + rr;
      split3_z(v.re, v.im, out[o], out3[o]);
    }
  }
}

// W_data[w][2+s] = dcache[2w+s] (determinants cached by the last inverse of
// the same walkers)
__global__ void det_restore_kernel(zd* __restrict__ Wd, const zd* __restrict__ dcache, zd* __restrict__ Wl,
                                   const zd* __restrict__ lcache, int nw)
{
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < 2 * nw) {
    Wd[(long)(b / 2) * 8 + 2 + (b % 2)] = dcache[b];
    Wl[(long)(b / 2) * 4 + (b % 2)] = lcache[b];
  }
}

// log-overlaps [nw][4] (current spin 0, 1; saved spin 0, 1) taken from the
// overlaps a caller wrote into W_data (columns 2, 3, 6, 7)
__global__ void wd_logs_kernel(const zd* __restrict__ Wd, zd* __restrict__ Wl, int nw)
{
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 4L * nw) return;
  const long w = e / 4, q = e % 4;
  const zd v = Wd[w * 8 + (q < 2 ? 2 + q : 4 + q)];
  Wl[e] = zd{log(hypot(v.re, v.im)), atan2(v.im, v.re)};
}

// ---------------------------------------------------------------------------
// CSR x dense, mySPBLAS::csrmm 'N' semantics (src/Numerics/sparse.hpp:166-180):
// C(r, w) = alpha * sum_{p in row r, col<K} val[p] * B(col[p], w) + beta*C(r,w)
// B(c, w) at B[c*ldb + w]; C(r, w) at C[r*ldc + w] (row-major, walker last).
// One wave per (row, 64-walker chunk); nonzero indices are wave-uniform.
__global__ __launch_bounds__(256) void csrmm_n_kernel(int nrows, int nw, int K, const int* __restrict__ rowptr,
                                                      const int* __restrict__ col, const zd* __restrict__ val,
                                                      const zd* __restrict__ B, long ldb, zd* __restrict__ C,
                                                      long ldc, zd alpha, zd beta)
{
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int nchunk = (nw + 63) / 64;
  const int r = wave / nchunk, ch = wave % nchunk;
  if (r >= nrows) return;
  const int w = ch * 64 + lane;
  const int p0 = rowptr[0];
  const int pb = __builtin_amdgcn_readfirstlane(rowptr[r] - p0);
  const int pe = __builtin_amdgcn_readfirstlane(rowptr[r + 1] - p0);
  zd acc{0, 0};
  if (w < nw) {
    for (int p = pb; p < pe; p++) {
      int c = col[p];
      if (c >= K) continue;
      zd a = val[p];
      zd bv = B[(long)c * ldb + w];
      acc.re += a.re * bv.re - a.im * bv.im;
      acc.im += a.re * bv.im + a.im * bv.re;
    }
    zd o = zmul(alpha, acc);
    if (beta.re != 0.0 || beta.im != 0.0) o = zadd(o, zmul(beta, C[(long)r * ldc + w]));
    C[(long)r * ldc + w] = o;
  }
}

// csrmm 'N' for long rows: one workgroup of S waves per (row, 64-walker
// chunk).  Wave s sums a contiguous segment of the row's entries into two
// accumulators (even and odd entries, so consecutive FMAs do not wait on
// each other); the S partial sums are reduced through LDS in a fixed order.
// Same semantics as csrmm_n_kernel (entries with col >= K skipped).
template <int S>
__global__ __launch_bounds__(64 * S) void csrmm_n_split_kernel(int nrows, int nw, int K, const int* __restrict__ rowptr,
                                                              const int* __restrict__ col,
                                                              const zd* __restrict__ val, const zd* __restrict__ B,
                                                              long ldb, zd* __restrict__ C, long ldc, zd alpha,
                                                              zd beta)
{
  __shared__ zd part[S][64];
  const int nchunk = (nw + 63) / 64;
  const int r = blockIdx.x / nchunk, ch = blockIdx.x % nchunk;
  const int sw = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int w = ch * 64 + lane;
  const int p0 = rowptr[0];
  const int pb = rowptr[r] - p0, pe = rowptr[r + 1] - p0;
  const int len = pe - pb, per = (len + S - 1) / S;
  const int qb = pb + min(len, sw * per), qe = pb + min(len, (sw + 1) * per);
  zd a0{0, 0}, a1{0, 0};
  if (w < nw) {
    int p = qb;
    for (; p + 1 < qe; p += 2) {
      const int c0 = col[p], c1 = col[p + 1];
      const zd v0 = val[p], v1 = val[p + 1];
      if (c0 < K) {
        const zd b0 = B[(long)c0 * ldb + w];
        a0.re += v0.re * b0.re - v0.im * b0.im;
        a0.im += v0.re * b0.im + v0.im * b0.re;
      }
      if (c1 < K) {
        const zd b1 = B[(long)c1 * ldb + w];
        a1.re += v1.re * b1.re - v1.im * b1.im;
        a1.im += v1.re * b1.im + v1.im * b1.re;
      }
    }
    if (p < qe) {
      const int c0 = col[p];
      const zd v0 = val[p];
      if (c0 < K) {
        const zd b0 = B[(long)c0 * ldb + w];
        a0.re += v0.re * b0.re - v0.im * b0.im;
        a0.im += v0.re * b0.im + v0.im * b0.re;
      }
    }
  }
  part[sw][lane] = zadd(a0, a1);
  __syncthreads();
  if (sw == 0 && w < nw) {
    zd acc = part[0][lane];
#pragma unroll
    for (int q = 1; q < S; q++) acc = zadd(acc, part[q][lane]);
    zd o = zmul(alpha, acc);
    if (beta.re != 0.0 || beta.im != 0.0) o = zadd(o, zmul(beta, C[(long)r * ldc + w]));
    C[(long)r * ldc + w] = o;
  }
}

// value of lane j (j the same in every lane) without going through LDS
__device__ inline int rdlane(int x, int j) { return __builtin_amdgcn_readlane(x, j); }
__device__ inline double rdlane(double x, int j)
{
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffLL), j);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// csrmm 'N' for long rows whose columns are sorted within each row: a
// workgroup of R waves takes R consecutive rows and one chunk of 64 walkers,
// and walks the columns in windows of CB.  Each window of B (CB rows x 64
// walkers) is staged in LDS once and read by all R rows, so B is fetched
// once per R rows instead of once per row.  Entries with col >= K are never
// reached (the reference's csrmm skips them too).
template <int R, int CB>
__global__ __launch_bounds__(64 * R) void csrmm_n_lds_kernel(int nrows, int nw, int K, const int* __restrict__ rowptr,
                                                            const int* __restrict__ col, const zd* __restrict__ val,
                                                            const zd* __restrict__ B, long ldb, zd* __restrict__ C,
                                                            long ldc, zd alpha, zd beta)
{
  __shared__ zd Bs[CB][64];
  const int nchunk = (nw + 63) / 64;
  const int rb = blockIdx.x / nchunk, ch = blockIdx.x % nchunk;
  const int sw = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = rb * R + sw;
  const int w = ch * 64 + lane;
  const int p0 = rowptr[0];
  int p = 0, pe = 0;
  if (r < nrows) {
    p = rowptr[r] - p0;
    pe = rowptr[r + 1] - p0;
  }
  p = __builtin_amdgcn_readfirstlane(p);  // one row per wave
  pe = __builtin_amdgcn_readfirstlane(pe);
  zd a0{0, 0}, a1{0, 0};
  for (int c0 = 0; c0 < K; c0 += CB) {
    for (int e = threadIdx.x; e < CB * 64; e += 64 * R) {
      const int cc = e >> 6, l = e & 63, ww = ch * 64 + l;
      Bs[cc][l] = (c0 + cc < K && ww < nw) ? B[(long)(c0 + cc) * ldb + ww] : zd{0, 0};
    }
    __syncthreads();
    const int cend = c0 + CB;
    // the row's entries in this window, 64 at a time: lane i loads entry
    // p + i, the window's entries form a prefix (columns sorted), and each
    // is read from its lane into scalar registers
    for (;;) {
      const int q = p + lane;
      const bool inw = q < pe && col[q] < cend;
      const int cq = inw ? col[q] - c0 : 0;
      const zd vq = inw ? val[q] : zd{0, 0};
      const int n = __popcll(__ballot(inw));
      for (int j = 0; j + 1 < n; j += 2) {
        const int c_0 = rdlane(cq, j), c_1 = rdlane(cq, j + 1);
        const zd v0{rdlane(vq.re, j), rdlane(vq.im, j)}, v1{rdlane(vq.re, j + 1), rdlane(vq.im, j + 1)};
        const zd b0 = Bs[c_0][lane], b1 = Bs[c_1][lane];
        a0.re += v0.re * b0.re - v0.im * b0.im;
        a0.im += v0.re * b0.im + v0.im * b0.re;
        a1.re += v1.re * b1.re - v1.im * b1.im;
        a1.im += v1.re * b1.im + v1.im * b1.re;
      }
      if (n & 1) {
        const int c_0 = rdlane(cq, n - 1);
        const zd v0{rdlane(vq.re, n - 1), rdlane(vq.im, n - 1)};
        const zd b0 = Bs[c_0][lane];
        a0.re += v0.re * b0.re - v0.im * b0.im;
        a0.im += v0.re * b0.im + v0.im * b0.re;
      }
      p += n;
      if (n < 64) break;
    }
    __syncthreads();
  }
  if (r < nrows && w < nw) {
    zd o = zmul(alpha, zadd(a0, a1));
    if (beta.re != 0.0 || beta.im != 0.0) o = zadd(o, zmul(beta, C[(long)r * ldc + w]));
    C[(long)r * ldc + w] = o;
  }
}

// ---------------------------------------------------------------------------
// Tiled transpose of complex matrices: out[c][r] = in[r][c], in is R x Cc.
// Output precision OT (double or float pairs).
template <class IT, class OT>
__global__ __launch_bounds__(256) void ztranspose_kernel(const IT* __restrict__ in, OT* __restrict__ out, int R,
                                                         int Cc, long ldi, long ldo)
{
  __shared__ zd tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int yy = ty; yy < 32; yy += 8) {
    int rr = by + yy, cc = bx + tx;
    if (rr < R && cc < Cc) tile[yy][tx] = to_zd(in[(long)rr * ldi + cc]);
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    int cc = bx + yy, rr = by + tx;
    if (rr < R && cc < Cc) {
      OT o;
      from_zd(tile[tx][yy], o);
      out[(long)cc * ldo + rr] = o;
    }
  }
}

// elementwise precision conversion / copy of n complex values
template <class IT, class OT>
__global__ void zconvert_kernel(const IT* __restrict__ in, OT* __restrict__ out, long n)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long st = (long)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    OT o;
    from_zd(to_zd(in[i]), o);
    out[i] = o;
  }
}

__global__ void fill_int_kernel(int* __restrict__ p, int n, int v)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void set_long_kernel(long* __restrict__ p, long v)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// out[i] = sum_s in[s*n + i] (FP64 accumulation, fixed order)
template <class PZ>
__global__ void zsum_slices_kernel(const PZ* __restrict__ in, zd* __restrict__ out, long n, int ns)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long st = (long)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    zd acc{0.0, 0.0};
    for (int q = 0; q < ns; q++) acc = zadd(acc, to_zd(in[(long)q * n + i]));
    out[i] = acc;
  }
}

__device__ inline zd ztwice(zd v) { return zd{2.0 * v.re, 2.0 * v.im}; }
__device__ inline zf ztwice(zf v) { return zf{2.0f * v.re, 2.0f * v.im}; }
__device__ inline uint32_t bf2_twice(uint32_t w)
{
  const unsigned short a = f2bf(2.0f * bf2f((unsigned short)(w & 0xffff)));
  const unsigned short b = f2bf(2.0f * bf2f((unsigned short)(w >> 16)));
  return (uint32_t)a | ((uint32_t)b << 16);
}
__device__ inline zb ztwice(zb v) { return zb{bf2_twice(v.hi), bf2_twice(v.lo)}; }

// nmat symmetric n x n matrices -> their lower triangles with the strictly
// lower part doubled, upper part zero, so that
//   sum_pq g_p V_pq g_q = sum_p g_p sum_{q <= p} V'_pq g_q
// (half the work of the full quadratic form).  Scaling by 2 is exact in
// every storage format.
template <class ZT>
__global__ void sym_lower_kernel(ZT* __restrict__ V, long n, long nmat)
{
  const long tot = nmat * n * n;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long st = (long)gridDim.x * blockDim.x;
  for (; i < tot; i += st) {
    const long col = i % n, row = (i / n) % n;
    if (col > row) {
      ZT z;
      zero_z(z);
      V[i] = z;
    } else if (col < row) {
      V[i] = ztwice(V[i]);
    }
  }
}

// y += x (complex, FP64)
__global__ void zaxpy1_kernel(const zd* __restrict__ x, zd* __restrict__ y, long n)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long st = (long)gridDim.x * blockDim.x;
  for (; i < n; i += st) y[i] = zadd(y[i], x[i]);
}

// ---------------------------------------------------------------------------
// X / hybrid weight (src/miniapps/miniafqmc_base.cpp:281-289):
//   X(n, w) = x(n, w) + i vbias(n, w),  hybridW[w] = -sum_n i vb (x + i/2 vb).
// The real draws x come from a staged block (x(n, w) at x[n*nw + w]; with
// xslice != null the block of substep *xslice, an index kept in device memory
// so that a captured graph can be replayed) or, when x is null, are generated
// on the fly from (seed, substep, walker) exactly as afqmc_synth_fields does
// on the host (the reference draws them inside the substep,
// miniafqmc_base.cpp:282).  Non-finite staged draws raise *nonfinite.
struct FieldSrc {
  const double* x;      // staged draws (null: generate)
  const long* xslice;   // device substep index, or null
  long substep;         // substep index when xslice is null
  uint64_t seed;        // generator seed (x == null)
};

__device__ inline void field_term(zd b, double xr, zd& h, zd& Xo)
{
  // t1 = i*vb ; t3 = x + i/2*vb ; hybridW -= t1*t3 ; X = x + i*vb
  const zd t1{-b.im, b.re};
  const zd t3{xr - 0.5 * b.im, 0.5 * b.re};
  const zd t4 = zmul(t1, t3);
  h.re -= t4.re;
  h.im -= t4.im;
  Xo = zd{xr - b.im, b.re};
}

// Walker-first layouts (vbias, X [nw][K]): one workgroup per walker, the
// lanes along n (coalesced), each thread a pair (2j, 2j+1) of draws at a
// time; fixed-order block reduction of hybridW.
template <class XT>
__global__ __launch_bounds__(256) void fields_wfirst_kernel(int K, int nw, const zd* __restrict__ vb, FieldSrc src,
                                                            XT* __restrict__ X, zd* __restrict__ hybridW,
                                                            int* __restrict__ nonfinite,
                                                            uint32_t* __restrict__ X3 = nullptr)
{
  __shared__ zd part[4];
  const int w = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long sub = src.xslice ? *src.xslice : src.substep;
  const double* x = src.x ? src.x + (src.xslice ? sub * (long)K * nw : 0L) : nullptr;
  const zd* vbw = vb + (long)w * K;
  XT* Xw = X + (long)w * K;
  uint32_t* X3w = X3 ? X3 + (long)w * K : nullptr;
  // split X with its third pieces (three-piece split-BF16), else as XT
  auto put = [&](int n, zd o) {
    if constexpr (std::is_same<XT, zb>::value) {
      if (X3w) {
        split3_z(o.re, o.im, Xw[n], X3w[n]);
        return;
      }
    }
    from_zd(o, Xw[n]);
  };
  zd h{0, 0};
  bool bad = false;
  for (int n = 2 * threadIdx.x; n < K; n += 512) {
    double xa, xb = 0.0;
    if (x) {
      xa = x[(long)n * nw + w];
      if (n + 1 < K) xb = x[(long)(n + 1) * nw + w];
      bad |= !isfinite(xa) || !isfinite(xb);
    } else {
      afqmc_rng::field_pair(src.seed, sub, w, n, xa, xb);
    }
    zd o;
    field_term(vbw[n], xa, h, o);
    put(n, o);
    if (n + 1 < K) {
      field_term(vbw[n + 1], xb, h, o);
      put(n + 1, o);
    }
  }
  if (bad) atomicOr(nonfinite, 1);
  for (int off = 32; off >= 1; off >>= 1) {
    h.re += __shfl_xor(h.re, off);
    h.im += __shfl_xor(h.im, off);
  }
  if (lane == 0) part[wv] = h;
  __syncthreads();
  if (threadIdx.x == 0) hybridW[w] = zadd(zadd(part[0], part[1]), zadd(part[2], part[3]));
}

// Walker-last layout (vbias, X [K][nw], the FP64 CSR mode): one workgroup per
// 64 walkers (lanes), its 16 waves taking pairs (2j, 2j+1) of n in turn;
// fixed-order reduction over the waves.
template <class XT>
__global__ __launch_bounds__(1024) void fields_wlast_kernel(int K, int nw, const zd* __restrict__ vb, FieldSrc src,
                                                            XT* __restrict__ X, zd* __restrict__ hybridW,
                                                            int* __restrict__ nonfinite)
{
  __shared__ zd part[16][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int w = blockIdx.x * 64 + lane;
  const long sub = src.xslice ? *src.xslice : src.substep;
  const double* x = src.x ? src.x + (src.xslice ? sub * (long)K * nw : 0L) : nullptr;
  zd h{0, 0};
  bool bad = false;
  if (w < nw) {
    for (int n = 2 * wv; n < K; n += 32) {
      double xa, xb = 0.0;
      if (x) {
        xa = x[(long)n * nw + w];
        if (n + 1 < K) xb = x[(long)(n + 1) * nw + w];
        bad |= !isfinite(xa) || !isfinite(xb);
      } else {
        afqmc_rng::field_pair(src.seed, sub, w, n, xa, xb);
      }
      zd o;
      field_term(vb[(long)n * nw + w], xa, h, o);
      from_zd(o, X[(long)n * nw + w]);
      if (n + 1 < K) {
        field_term(vb[(long)(n + 1) * nw + w], xb, h, o);
        from_zd(o, X[(long)(n + 1) * nw + w]);
      }
    }
  }
  if (bad) atomicOr(nonfinite, 1);
  part[wv][lane] = h;
  __syncthreads();
  if (wv == 0 && w < nw) {
    zd t = part[0][lane];
    for (int q = 1; q < 16; q++) t = zadd(t, part[q][lane]);
    hybridW[w] = t;
  }
}

// W_data[:,5] = W_data[:,4]; W_data[:,6..7] = W_data[:,2..3]; the
// log-overlaps likewise
__global__ void save_overlaps_kernel(zd* __restrict__ Wd, zd* __restrict__ Wl, int nw)
{
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  zd* d = Wd + (long)w * 8;
  d[5] = d[4];
  d[6] = d[2];
  d[7] = d[3];
  zd* l = Wl + (long)w * 4;
  l[2] = l[0];
  l[3] = l[1];
}

// weights / hybrid energy / Eshift (src/miniapps/miniafqmc_base.cpp:317-327).
// The overlap ratio W_data[2]W_data[3] / (W_data[6]W_data[7]) enters as its
// log, from the log-overlaps Wl (log |ratio| and its phase, wrapped to
// [-pi, pi]), so that overlaps beyond the FP64 range keep their ratio.
// One block of up to 1024 threads; fixed-order tree reduction for Eshift.
__global__ __launch_bounds__(1024) void weights_kernel(zd* __restrict__ Wd, const zd* __restrict__ Wl,
                                                       const zd* __restrict__ hybridW, int nw, double dt,
                                                       double* __restrict__ eshift)
{
  __shared__ double red[1024];
  const double Es = *eshift;
  double et = 0.0;
  for (int w = threadIdx.x; w < nw; w += blockDim.x) {
    zd* d = Wd + (long)w * 8;
    const zd* l = Wl + (long)w * 4;
    const double lr = (l[0].re + l[1].re) - (l[2].re + l[3].re);
    const double ang = remainder((l[0].im + l[1].im) - (l[2].im + l[3].im), 6.283185307179586476925286766559);
    double scale = fmax(0.0, cos(ang));
    zd h = hybridW[w];
    d[4] = zd{-(h.re + lr) / dt, -(h.im + ang) / dt};
    double f = scale * exp(-dt * (0.5 * (d[4].re + d[5].re) - Es));
    d[1] = zd{d[1].re * f, d[1].im * f};
    et += d[4].re;
  }
  red[threadIdx.x] = et;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *eshift = red[0] / nw;
}

// Comb population control, selection (one workgroup).  Weights are
// w_i = Re W_data[i][1] (the phaseless weights are >= 0; a negative weight
// counts as 0); cum_i = w_0 + ... + w_i is summed in index order by one
// thread so that every caller gets the same rounding; tooth j sits at
// t_j = (u + j) * (cum_{nw-1} / nw) and takes the first walker with
// cum_i > t_j, or the last walker of positive weight when rounding puts t_j
// at or past the total.  flag[0] is set when the total is not positive and
// finite or a weight is not finite.
__global__ __launch_bounds__(1024) void comb_select_kernel(const zd* __restrict__ Wd, int nw, double u,
                                                            double* __restrict__ cum, int* __restrict__ parent,
                                                            double* __restrict__ wnew, int* __restrict__ flag)
{
  __shared__ int last_pos;
  if (threadIdx.x == 0) {
    double sum = 0.0;
    int lp = -1, bad = 0;
    for (int i = 0; i < nw; i++) {
      const double w = Wd[(long)i * 8 + 1].re;
      if (!isfinite(w)) bad = 1;
      if (w > 0.0) {
        sum += w;
        lp = i;
      }
      cum[i] = sum;
    }
    if (bad || !(sum > 0.0) || !isfinite(sum)) flag[0] = 1;
    *wnew = sum / nw;
    last_pos = lp;
  }
  __syncthreads();
  const double total = cum[nw - 1];
  const double step = total / nw;
  for (int j = threadIdx.x; j < nw; j += blockDim.x) {
    const double t = (u + j) * step;
    int lo = 0, hi = nw;  // first i with cum[i] > t, nw if none
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cum[mid] > t) hi = mid;
      else lo = mid + 1;
    }
    parent[j] = lo < nw ? lo : max(last_pos, 0);
  }
}

// Walkers and W_data rows of the selected parents: Wout[j] = W[parent[j]],
// Wdout[j] = Wd[parent[j]] with the weight replaced by the common weight.
__global__ void comb_gather_kernel(const zd* __restrict__ W, zd* __restrict__ Wout, const zd* __restrict__ Wd,
                                   zd* __restrict__ Wdout, const zd* __restrict__ Wl, zd* __restrict__ Wlout,
                                   const int* __restrict__ parent, const double* __restrict__ wnew, long per, int nw)
{
  const long n = (long)nw * per;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long j = e / per, q = e - j * per;
    Wout[e] = W[(long)parent[j] * per + q];
  }
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (long)nw * 8; e += (long)gridDim.x * blockDim.x) {
    const long j = e / 8, k = e % 8;
    Wdout[e] = (k == 1) ? zd{*wnew, 0.0} : Wd[(long)parent[j] * 8 + k];
  }
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (long)nw * 4; e += (long)gridDim.x * blockDim.x)
    Wlout[e] = Wl[(long)parent[e / 4] * 4 + e % 4];
}

// W = conj(trial) per spin, W_data row = (0, 1, 0, ...), log-overlaps
// log 0 = {-inf, 0}
__global__ void init_walkers_kernel(zd* __restrict__ W, zd* __restrict__ Wd, zd* __restrict__ Wl,
                                    const zd* __restrict__ trial, int nw, long MN)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)nw * 2 * MN;
  if (i < tot) {
    long rem = i % (2 * MN);
    W[i] = zconj(trial[rem]);
  }
  if (i < (long)nw * 8) Wd[i] = zd{(i % 8 == 1) ? 1.0 : 0.0, 0.0};
  if (i < (long)nw * 4) Wl[i] = zd{-INFINITY, 0.0};
}

// ---------------------------------------------------------------------------
// Householder LQ with LAPACK conventions on the column-major view A = W^T
// (N x M) of each W[w][s] ([M][N] row-major): the blocked zgelqf followed by
// the blocked zunglq (ma::gelqf / ma::glq, src/Numerics/ma_lapack.hpp:134-221;
// called by afqmc_sys::orthogonalize, src/AFQMC/afqmc_sys.hpp:175-207).
// A(i, c) is W[c][i]: row i of A is column i of W.  One block of 256 threads
// per matrix.  Panels of NB rows are factorised in LDS (zgelq2 / zungl2);
// the rows below a panel get the block reflector I - V^H T V (zlarft
// 'Forward','Rowwise' + zlarfb 'Right') in two passes through L2.
__device__ inline zd wave_zsum(zd v)
{
  for (int o = 32; o >= 1; o >>= 1) {
    v.re += __shfl_xor(v.re, o);
    v.im += __shfl_xor(v.im, o);
  }
  return v;
}
__device__ inline double block_dsum(double v, double* sh)
{
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  return sh[0] + sh[1] + sh[2] + sh[3];
}

// V(r, c) of a panel held in LDS rows P[r*LP + c]: 1 on the diagonal, the
// stored tail right of it, 0 left of it.
__device__ inline zd vpanel(const zd* P, int LP, int r, int c)
{
  if (c == r) return zd{1.0, 0.0};
  if (c < r) return zd{0.0, 0.0};
  return P[r * LP + c];
}

// zlarfg on panel row r (already conjugated), LAPACK sign rule.
__device__ inline bool householder_row(zd* P, int LP, int r, double* sh, zd& alpha, zd& taui, double& beta)
{
  const int tid = threadIdx.x;
  double ss = 0.0;
  for (int c = r + 1 + tid; c < LP; c += 256) ss += P[r * LP + c].re * P[r * LP + c].re + P[r * LP + c].im * P[r * LP + c].im;
  const double xn2 = block_dsum(ss, sh);
  alpha = P[r * LP + r];
  const bool apply = !(xn2 == 0.0 && alpha.im == 0.0);
  taui = zd{0, 0};
  beta = 0.0;
  if (apply) {
    const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
    beta = -copysign(nrm, alpha.re);
    taui = zd{(beta - alpha.re) / beta, -alpha.im / beta};
    const zd scal = zdiv(zd{1.0, 0.0}, zd{alpha.re - beta, alpha.im});
    for (int c = r + 1 + tid; c < LP; c += 256) P[r * LP + c] = zmul(scal, P[r * LP + c]);
  }
  __syncthreads();
  return apply;
}

// apply I - t v v^H (v = panel row r with unit diagonal) from the right to
// panel rows r+1..ib-1 over columns r..LP-1
template <int NB>
__device__ inline void panel_apply(zd* P, int LP, int r, int ib, zd t, zd* red)
{
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  zd acc[NB];
#pragma unroll
  for (int q = 0; q < NB; q++) acc[q] = zd{0, 0};
  for (int c = r + tid; c < LP; c += 256) {
    const zd vc = (c == r) ? zd{1.0, 0.0} : P[r * LP + c];
#pragma unroll
    for (int q = 0; q < NB; q++)
      if (q > r && q < ib) acc[q] = zadd(acc[q], zmul(P[q * LP + c], vc));
  }
#pragma unroll
  for (int q = 0; q < NB; q++) {
    zd s = wave_zsum(acc[q]);
    if (lane == 0) red[wv * NB + q] = s;
  }
  __syncthreads();
  zd wq[NB];
#pragma unroll
  for (int q = 0; q < NB; q++) wq[q] = zmul(t, zadd(zadd(red[q], red[NB + q]), zadd(red[2 * NB + q], red[3 * NB + q])));
  for (int c = r + tid; c < LP; c += 256) {
    const zd vcc = zconj((c == r) ? zd{1.0, 0.0} : P[r * LP + c]);
#pragma unroll
    for (int q = 0; q < NB; q++)
      if (q > r && q < ib) P[q * LP + c] = zsub(P[q * LP + c], zmul(wq[q], vcc));
  }
  __syncthreads();
}

// T (ib x ib, upper) of zlarft('Forward','Rowwise') from the panel and taus.
// The ib(ib-1)/2 dot products V(j,:) V(i,:)^H (j < i, columns >= i) are
// reduced block-wide, then one thread runs the triangular recursion.
template <int NB>
__device__ inline void form_T(const zd* P, int LP, int ib, const zd* tau, zd* T, zd* red)
{
  constexpr int NP = NB * (NB - 1) / 2;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  zd acc[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) acc[e] = zd{0, 0};
  for (int c = tid; c < LP; c += 256) {
    zd pj[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) pj[j] = (j < ib) ? P[j * LP + c] : zd{0, 0};
#pragma unroll
    for (int i = 1; i < NB; i++) {
      if (i >= ib || c < i) continue;
      const zd vi = zconj((c == i) ? zd{1.0, 0.0} : pj[i]);
#pragma unroll
      for (int j = 0; j < i; j++) acc[i * (i - 1) / 2 + j] = zadd(acc[i * (i - 1) / 2 + j], zmul(pj[j], vi));
    }
  }
#pragma unroll
  for (int e = 0; e < NP; e++) {
    zd s = wave_zsum(acc[e]);
    if (lane == 0) red[wv * NP + e] = s;
  }
  __syncthreads();
  if (tid == 0) {
    for (int i = 0; i < ib; i++) {
      const zd ti = tau[i];
      if (ti.re == 0.0 && ti.im == 0.0) {
        for (int j = 0; j <= i; j++) T[j * NB + i] = zd{0, 0};
        continue;
      }
      // T(0:i-1, i) = -tau_i V(0:i-1, :) V(i, :)^H, then T(0:i-1,0:i-1) times it (in place, ascending rows)
      for (int j = 0; j < i; j++) {
        const int e = i * (i - 1) / 2 + j;
        zd d = zadd(zadd(red[e], red[NP + e]), zadd(red[2 * NP + e], red[3 * NP + e]));
        T[j * NB + i] = zmul(zd{-ti.re, -ti.im}, d);
      }
      for (int j = 0; j < i; j++) {
        zd s{0, 0};
        for (int l = j; l < i; l++) s = zadd(s, zmul(T[j * NB + l], T[l * NB + i]));
        T[j * NB + i] = s;
      }
      T[i * NB + i] = ti;
    }
  }
  __syncthreads();
}

// C <- C (I - V^H op(T) V) for rows rb..N-1 of A, columns c0..M-1 (panel
// columns are local c = global - c0).  conjT: op(T) = T^H (zunglq), else T.
// Threads: 128 rows x 2 column halves; each thread keeps UNR independent
// loads in flight per iteration.
template <int NB>
__device__ inline void trailing_update(zd* W, int N, int rb, int c0, int LP, const zd* P, int ib, const zd* T,
                                       bool conjT, zd* W1)
{
  constexpr int UNR = 8;
  const int tid = threadIdx.x;
  const int nr = N - rb;
  const int half = tid >> 7;
  const int hl = (LP + 1) / 2;           // columns of half 0
  const int cb = half ? hl : 0, ce = half ? LP : hl;
  for (int q0 = 0; q0 < nr; q0 += 128) {
    const int ql = tid & 127;
    const int q = q0 + ql;
    const bool active = q < nr;
    zd acc[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) acc[r] = zd{0, 0};
    for (int c = cb; c < ce; c += UNR) {
      zd cv[UNR];
#pragma unroll
      for (int u = 0; u < UNR; u++)
        cv[u] = (active && c + u < ce) ? W[(long)(c0 + c + u) * N + rb + q] : zd{0, 0};
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        if (c + u >= ce) break;
#pragma unroll
        for (int r = 0; r < NB; r++)
          if (r < ib) acc[r] = zadd(acc[r], zmul(cv[u], zconj(vpanel(P, LP, r, c + u))));
      }
    }
    if (half == 1) {
#pragma unroll
      for (int r = 0; r < NB; r++) W1[ql * NB + r] = acc[r];
    }
    __syncthreads();
    zd w2[NB];
    if (half == 0) {
#pragma unroll
      for (int r = 0; r < NB; r++) acc[r] = zadd(acc[r], W1[ql * NB + r]);
      // W2 = W1 op(T)
#pragma unroll
      for (int r = 0; r < NB; r++) {
        zd s{0, 0};
#pragma unroll
        for (int j = 0; j < NB; j++) {
          if (j >= ib || r >= ib) continue;
          zd tjr = conjT ? ((r <= j) ? zconj(T[r * NB + j]) : zd{0, 0}) : ((j <= r) ? T[j * NB + r] : zd{0, 0});
          s = zadd(s, zmul(acc[j], tjr));
        }
        w2[r] = s;
      }
    }
    __syncthreads();
    if (half == 0) {
#pragma unroll
      for (int r = 0; r < NB; r++) W1[ql * NB + r] = w2[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NB; r++) w2[r] = W1[ql * NB + r];
    if (active) {
      for (int c = cb; c < ce; c += UNR) {
        zd cv[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) cv[u] = (c + u < ce) ? W[(long)(c0 + c + u) * N + rb + q] : zd{0, 0};
#pragma unroll
        for (int u = 0; u < UNR; u++) {
          if (c + u >= ce) break;
          zd s{0, 0};
#pragma unroll
          for (int r = 0; r < NB; r++)
            if (r < ib) s = zadd(s, zmul(w2[r], vpanel(P, LP, r, c + u)));
          W[(long)(c0 + c + u) * N + rb + q] = zsub(cv[u], s);
        }
      }
    }
    __syncthreads();
  }
}

// only != null: matrices b with only[b] == 0 are skipped (the fallback pass
// after the Cholesky route).
template <int NB>
__global__ __launch_bounds__(256) void lq_blocked_kernel(zd* __restrict__ Wall, int M, int N,
                                                         const int* __restrict__ only)
{
  if (only && only[blockIdx.x] == 0) return;
  extern __shared__ zd sm[];
  zd* P = sm;                  // [NB][M]
  zd* W1 = P + NB * M;         // [128][NB] (trailing rows go in chunks of 128)
  zd* T = W1 + 128 * NB;       // [NB][NB]
  zd* tau = T + NB * NB;       // [N]
  zd* red = tau + N;           // [4][NB*NB]
  __shared__ double sh[8];
  zd* W = Wall + (long)blockIdx.x * M * N;
  const int tid = threadIdx.x;
  auto load_panel = [&](int i0, int ib, int LP) {
    for (int e = tid; e < ib * LP; e += 256) {
      const int r = e % ib, c = e / ib;
      P[r * LP + c] = W[(long)(i0 + c) * N + i0 + r];
    }
    __syncthreads();
  };
  auto store_panel = [&](int i0, int ib, int LP) {
    for (int e = tid; e < ib * LP; e += 256) {
      const int r = e % ib, c = e / ib;
      W[(long)(i0 + c) * N + i0 + r] = P[r * LP + c];
    }
    __syncthreads();
  };
  // ---- zgelqf ----
  for (int i0 = 0; i0 < N; i0 += NB) {
    const int ib = min(NB, N - i0), LP = M - i0;
    load_panel(i0, ib, LP);
    for (int r = 0; r < ib; r++) {
      for (int c = r + tid; c < LP; c += 256) P[r * LP + c] = zconj(P[r * LP + c]);
      __syncthreads();
      zd alpha, taui;
      double beta;
      const bool apply = householder_row(P, LP, r, sh, alpha, taui, beta);
      if (tid == 0) tau[i0 + r] = taui;
      if (apply && r < ib - 1) panel_apply<NB>(P, LP, r, ib, taui, red);
      if (tid == 0) P[r * LP + r] = apply ? zd{beta, 0.0} : alpha;
      __syncthreads();
      for (int c = r + tid; c < LP; c += 256) P[r * LP + c] = zconj(P[r * LP + c]);
      __syncthreads();
    }
    store_panel(i0, ib, LP);
    if (i0 + ib < N) {
      form_T<NB>(P, LP, ib, tau + i0, T, red);
      trailing_update<NB>(W, N, i0 + ib, i0, LP, P, ib, T, false, W1);
    }
  }
  // ---- zunglq (k = N reflectors, blocks from last to first) ----
  const int last = ((N - 1) / NB) * NB;
  for (int i0 = last; i0 >= 0; i0 -= NB) {
    const int ib = min(NB, N - i0), LP = M - i0;
    load_panel(i0, ib, LP);
    if (i0 + ib < N) {
      form_T<NB>(P, LP, ib, tau + i0, T, red);
      trailing_update<NB>(W, N, i0 + ib, i0, LP, P, ib, T, true, W1);
    }
    // zungl2 on the panel rows
    for (int r = ib - 1; r >= 0; r--) {
      const zd ti = tau[i0 + r];
      if (r < LP - 1) {
        for (int c = r + 1 + tid; c < LP; c += 256) P[r * LP + c] = zconj(P[r * LP + c]);
        __syncthreads();
        if (r < ib - 1 && !(ti.re == 0.0 && ti.im == 0.0)) panel_apply<NB>(P, LP, r, ib, zconj(ti), red);
        const zd mt{-ti.re, -ti.im};
        for (int c = r + 1 + tid; c < LP; c += 256) P[r * LP + c] = zconj(zmul(mt, P[r * LP + c]));
      }
      __syncthreads();
      if (tid == 0) P[r * LP + r] = zd{1.0 - ti.re, ti.im};
      for (int c = tid; c < r; c += 256) P[r * LP + c] = zd{0, 0};
      __syncthreads();
    }
    store_panel(i0, ib, LP);
    for (int e = tid; e < ib * i0; e += 256) {
      const int r = e % ib, c = e / ib;
      W[(long)c * N + i0 + r] = zd{0, 0};
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Orthogonalisation through Cholesky QR.  For B = conj(W_s) (M x N), the
// reference's zgelqf + zunglq on the column-major view W_s^T equals
// zgeqrf + zungqr on B, so the new walker is conj(Q_h) with B = Q_h R_h and
// R_h's diagonal real with LAPACK's sign rule (beta = -sign(Re alpha) |x|).
// With the positive factorisation B = Q+ R+ (Cholesky of B^H B), Q_h = Q+ D
// for a diagonal D of +-1.  D is recovered without the reflectors: the
// no-pivot LU of the top N x N block X of Q+ in which each pivot is first
// shifted by s_k = -sign(Re X_kk) (the Householder reconstruction of Ballard
// et al., "Reconstructing Householder vectors from tall-skinny QR", 2015)
// yields exactly D_k = s_k, because that shift is the reconstructed
// reflector's choice of beta and the Householder factorisation of B is
// unique once the signs are fixed.

// Cholesky S = R^H R (R upper, positive diagonal) of a batch of Hermitian
// N x N matrices (row-major, upper triangle read) followed by the in-place
// Gauss-Jordan inverse of R: writes U = R^{-1} (upper, row-major).  Thread
// (i, cg) holds row i, columns cg + CG*t in registers.  A pivot below
// tol * max_k S_kk marks the matrix in fail[b] (its U is then undefined).
template <int NMAX, int CG>
__global__ __launch_bounds__(NMAX* CG) void chol_inv_kernel(const zd* __restrict__ S, zd* __restrict__ U, int N,
                                                            double tol, int* __restrict__ fail)
{
  constexpr int NC = NMAX / CG;
  __shared__ zd prow[2][NMAX];
  __shared__ zd colk[2][NMAX];
  __shared__ double dmax;
  const int tid = threadIdx.x;
  const int i = tid / CG, cg = tid % CG;
  const long b = blockIdx.x;
  const zd* Sb = S + b * (long)N * N;
  zd x[NC];
#pragma unroll
  for (int t = 0; t < NC; t++) {
    const int j = cg + CG * t;
    x[t] = (i < N && j < N && j >= i) ? Sb[(long)i * N + j] : zd{(i == j) ? 1.0 : 0.0, 0.0};
  }
  if (tid == 0) {
    double m = 0.0;
    for (int k = 0; k < N; k++) m = fmax(m, Sb[(long)k * N + k].re);
    dmax = m;
  }
  __syncthreads();
  const double thr = tol * dmax;
  bool bad = false;
  // Cholesky, right-looking on the upper triangle
  for (int k = 0; k < N; k++) {
    const int tk = k / CG, ck = k % CG, st = k & 1;
    if (i == k) {
#pragma unroll
      for (int t = 0; t < NC; t++) prow[st][cg + CG * t] = x[t];
    }
    __syncthreads();
    const double d = prow[st][k].re;
    if (!(d > thr)) {
      bad = true;
      break;
    }
    const double r = sqrt(d), rinv = 1.0 / r;
    if (i == k) {
#pragma unroll
      for (int t = 0; t < NC; t++) {
        const int j = cg + CG * t;
        x[t] = (j < k) ? zd{0, 0} : ((j == k) ? zd{r, 0.0} : zd{x[t].re * rinv, x[t].im * rinv});
      }
    } else if (i > k && i < N) {
      // conj(R_ki) R_kj = conj(S_ki) S_kj / d, on every column: the entries
      // left of the diagonal take garbage that is never read and is zeroed
      // before the inverse (padded columns of row k are zero)
      const zd ski = prow[st][i];
      const zd f{ski.re / d, -ski.im / d};
#pragma unroll
      for (int t = 0; t < NC; t++) {
        x[t] = zsub(x[t], zmul(f, prow[st][cg + CG * t]));
        if ((t & 7) == 7) asm volatile("" ::: "memory");  // bound the LDS loads in flight
      }
    }
    (void)tk;
    (void)ck;
  }
  if (bad) {
    if (tid == 0) fail[b] = 1;
    return;
  }
  // the last Cholesky step read prow[(N - 1) & 1], which the inverse's first
  // step (buffer 0) overwrites before its own barrier when N is odd
  __syncthreads();
  // rows i < N now hold R (upper); rows >= N the identity.  In-place
  // Gauss-Jordan inverse, pivots on the diagonal, the columns in blocks of CG
  // as gj_inverse_kernel: slot tk of block tk is held in xc, addressed
  // statically, for the block's CG steps.
#pragma unroll
  for (int t = 0; t < NC; t++) {
    const int j = cg + CG * t;
    if (j < i) x[t] = zd{0, 0};
  }
  const int nblk = (N + CG - 1) / CG;
  for (int tk = 0; tk < nblk; tk++) {
    zd xc{0, 0};
#pragma unroll
    for (int t = 0; t < NC; t++)
      if (t == tk) xc = x[t];
    const int jc = cg + CG * tk;
#pragma unroll
    for (int ck = 0; ck < CG; ck++) {
      const int k = CG * tk + ck, st = k & 1;
      if (k >= N) break;
      const bool own = (cg == ck);
      if (i == k) {
#pragma unroll
        for (int t = 0; t < NC; t++) prow[st][cg + CG * t] = x[t];
        prow[st][jc] = xc;  // after slot tk's stale copy: the later store wins
      }
      if (own) colk[st][i] = xc;
      __syncthreads();
      const double ipv = 1.0 / prow[st][k].re;  // R_kk is real
      if (i == k) {
#pragma unroll
        for (int t = 0; t < NC; t++) x[t] = zd{x[t].re * ipv, x[t].im * ipv};
        xc = own ? zd{ipv, 0.0} : zd{xc.re * ipv, xc.im * ipv};
      } else if (i < k) {
        const zd g = colk[st][i];
        const zd gq{-g.re * ipv, -g.im * ipv};
#pragma unroll
        for (int t = 0; t < NC; t++) {
          x[t] = zadd(x[t], zmul(gq, prow[st][cg + CG * t]));
          if ((t & 7) == 7) asm volatile("" ::: "memory");
        }
        xc = own ? gq : zadd(xc, zmul(gq, prow[st][jc]));
      }
    }
#pragma unroll
    for (int t = 0; t < NC; t++)
      if (t == tk) x[t] = xc;
  }
  zd* Ub = U + b * (long)N * N;
#pragma unroll
  for (int t = 0; t < NC; t++) {
    const int j = cg + CG * t;
    if (i < N && j < N) Ub[(long)i * N + j] = x[t];
  }
}

// Signs D of the LAPACK factor from the shifted no-pivot LU of the top
// N x N block of Q+ (conjugation leaves the real parts of the pivots
// unchanged, so the walker-layout block conj(Q+) is used as is), then
// Wout = Win D for every row.  Matrices with fail[b] set are skipped.
template <int NMAX, int CG>
__global__ __launch_bounds__(NMAX* CG) void ortho_sign_kernel(const zd* __restrict__ Win, zd* __restrict__ Wout,
                                                              int M, int N, const int* __restrict__ fail)
{
  constexpr int NC = NMAX / CG;
  __shared__ zd prow[2][NMAX];
  __shared__ zd colk[2][NMAX];
  __shared__ double sgn[NMAX];
  const long b = blockIdx.x;
  if (fail[b]) return;
  const int tid = threadIdx.x;
  const int i = tid / CG, cg = tid % CG;
  const zd* Wb = Win + b * (long)M * N;
  zd x[NC];
#pragma unroll
  for (int t = 0; t < NC; t++) {
    const int j = cg + CG * t;
    x[t] = (i < N && j < N) ? Wb[(long)i * N + j] : zd{0, 0};
  }
  for (int k = 0; k < N; k++) {
    const int tk = k / CG, ck = k % CG, st = k & 1;
    if (i == k) {
#pragma unroll
      for (int t = 0; t < NC; t++) prow[st][cg + CG * t] = x[t];
    }
    if (cg == ck) {
#pragma unroll
      for (int t = 0; t < NC; t++)
        if (t == tk) colk[st][i] = x[t];
    }
    __syncthreads();
    const zd pk = prow[st][k];
    const double sk = (pk.re >= 0.0) ? -1.0 : 1.0;
    if (tid == 0) sgn[k] = sk;
    if (i > k && i < N) {
      const zd l = zdiv(colk[st][i], zd{pk.re - sk, pk.im});
#pragma unroll
      for (int t = 0; t < NC; t++) {
        const int j = cg + CG * t;
        if (j > k) x[t] = zsub(x[t], zmul(l, prow[st][j]));
      }
    }
  }
  __syncthreads();
  zd* Wo = Wout + b * (long)M * N;
  const long tot = (long)M * N;
  for (long e = tid; e < tot; e += NMAX * CG) {
    const double s = sgn[e % N];
    const zd v = Wb[e];
    Wo[e] = zd{s * v.re, s * v.im};
  }
}

// ---------------------------------------------------------------------------
// Per-walker complex dot products without conjugation, reduction over i:
//   out[w] = scale * sum_i A(i, w) * B(i, w) (+ out[w] if accumulate)
// A(i,w) at A[i*asi + w*asw], same for B.  One block per walker.
template <class AT, class BT>
__global__ __launch_bounds__(256) void coldot_kernel(int n, const AT* __restrict__ A, long asi, long asw,
                                                     const BT* __restrict__ B, long bsi, long bsw, zd* out, long osw,
                                                     double scale, int accumulate)
{
  __shared__ zd sh[256];
  const int w = blockIdx.x;
  zd acc{0, 0};
  for (int i = threadIdx.x; i < n; i += 256) {
    const zd a = to_zd(A[(long)i * asi + (long)w * asw]);
    const zd bb = to_zd(B[(long)i * bsi + (long)w * bsw]);
    acc.re += a.re * bb.re - a.im * bb.im;
    acc.im += a.re * bb.im + a.im * bb.re;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = zadd(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    zd r{sh[0].re * scale, sh[0].im * scale};
    if (accumulate) r = zadd(r, out[(long)w * osw]);
    out[(long)w * osw] = r;
  }
}

// Exchange reduction: for each batch item q (a walker-spin and a chunk of
// Cholesky vectors), sum over n in the chunk of sum_ab Y_n(a,b) Y_n(b,a),
// with Y stored [nq][N][N] per n.  out[q] += result (FP64 accumulation).
template <class YT>
__global__ __launch_bounds__(256) void trace_sq_kernel(const YT* __restrict__ Y, int N, int nq, long ystride,
                                                       zd* __restrict__ out, long ostride, int nbatch)
{
  __shared__ zd sh[256];
  const int q = blockIdx.x;  // batch item
  const YT* Yq = Y + (long)q * ystride;
  zd acc{0, 0};
  const long NN = (long)N * N;
  for (long e = threadIdx.x; e < (long)nq * NN; e += 256) {
    long n = e / NN, rem = e % NN;
    int a = (int)(rem / N), b = (int)(rem % N);
    const zd y1 = to_zd(Yq[n * NN + (long)a * N + b]);
    const zd y2 = to_zd(Yq[n * NN + (long)b * N + a]);
    acc.re += y1.re * y2.re - y1.im * y2.im;
    acc.im += y1.re * y2.im + y1.im * y2.re;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = zadd(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(long)q * ostride] = zadd(out[(long)q * ostride], sh[0]);
}

// out[w] = part[2w] + part[2w+1] (the two spins)
__global__ void pairsum_kernel(const zd* __restrict__ part, zd* __restrict__ out, int nw)
{
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw) out[w] = zadd(part[2 * w], part[2 * w + 1]);
}

// direct term: out[w] += sum_n v(n,w)^2 where v(n,w) at v[n*vsn + w*vsw]
template <class VT>
__global__ __launch_bounds__(256) void sumsq_kernel(int K, const VT* __restrict__ v, long vsn, long vsw,
                                                    zd* __restrict__ out, double scale)
{
  __shared__ zd sh[256];
  const int w = blockIdx.x;
  zd acc{0, 0};
  for (int n = threadIdx.x; n < K; n += 256) {
    const zd x = to_zd(v[(long)n * vsn + (long)w * vsw]);
    acc.re += x.re * x.re - x.im * x.im;
    acc.im += 2.0 * x.re * x.im;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = zadd(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[w].re += scale * sh[0].re;
    out[w].im += scale * sh[0].im;
  }
}

// E_L assembly for the Cholesky form:
// Wd[w][0] = 0.5*(direct[w] - exch[w]) + e1[w]
__global__ void energy_assemble_kernel(zd* __restrict__ Wd, const zd* __restrict__ direct,
                                       const zd* __restrict__ exch, const zd* __restrict__ e1, int nw)
{
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  zd e2{0.5 * (direct[w].re - exch[w].re), 0.5 * (direct[w].im - exch[w].im)};
  Wd[(long)w * 8] = zadd(e2, e1[w]);
}

// ---------------------------------------------------------------------------
// Setup kernels.
// Scatter CSR Spvn [M*M][K] into dense Ld[n][ik] (Cholesky-vector major).
__global__ void csr_densify_kernel(int nrows, const int* __restrict__ rowptr, const int* __restrict__ col,
                                   const zd* __restrict__ val, zd* __restrict__ Ld, long ldn)
{
  int r = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (r >= nrows) return;
  int p0 = rowptr[0];
  for (int p = rowptr[r] - p0 + lane; p < rowptr[r + 1] - p0; p += 64) Ld[(long)col[p] * ldn + r] = val[p];
}

// CSR rows into a dense row-major matrix D[r][col] (zeroed beforehand;
// rowptr starts at 0, no repeated columns within a row)
__global__ void csr_densify_rows_kernel(int nrows, const int* __restrict__ rowptr, const int* __restrict__ col,
                                        const zd* __restrict__ val, zd* __restrict__ D, long ld)
{
  const int r = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= nrows) return;
  for (int p = rowptr[r] + lane; p < rowptr[r + 1]; p += 64) D[(long)r * ld + col[p]] = val[p];
}

// per-row count of |Lt(n, j)| > cutoff, Lt spin-major [2][K][NM]: row n of
// SpvnT is (Lt[0][n][:], Lt[1][n][:]).
__global__ void rot_count_kernel(const zd* __restrict__ Lt, int K, long NM, double cutoff, int* __restrict__ cnt)
{
  __shared__ int sh[256];
  const int n = blockIdx.x;
  int c = 0;
  for (long j = threadIdx.x; j < 2 * NM; j += 256) {
    int sp = (int)(j / NM);
    zd v = Lt[(long)sp * K * NM + (long)n * NM + (j % NM)];
    if (hypot(v.re, v.im) > cutoff) c++;
  }
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) cnt[n] = sh[0];
}

// fill of the CSR rows in ascending column order: one block per row, a
// block-wide prefix over chunks of 256 columns.
__global__ void rot_fill_kernel(const zd* __restrict__ Lt, int K, long NM, double cutoff,
                                const int* __restrict__ rowptr, int* __restrict__ colo, zd* __restrict__ valo)
{
  __shared__ int sh[256];
  const int n = blockIdx.x;
  int base = rowptr[n];
  for (long j0 = 0; j0 < 2 * NM; j0 += 256) {
    long j = j0 + threadIdx.x;
    bool keep = false;
    zd v{0, 0};
    if (j < 2 * NM) {
      int sp = (int)(j / NM);
      v = Lt[(long)sp * K * NM + (long)n * NM + (j % NM)];
      keep = hypot(v.re, v.im) > cutoff;
    }
    sh[threadIdx.x] = keep ? 1 : 0;
    __syncthreads();
    // inclusive scan (Hillis-Steele)
    for (int o = 1; o < 256; o <<= 1) {
      int t = (threadIdx.x >= o) ? sh[threadIdx.x - o] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (keep) {
      int pos = base + sh[threadIdx.x] - 1;
      colo[pos] = (int)j;
      valo[pos] = v;
    }
    int tot = sh[255];
    __syncthreads();
    base += tot;
  }
}
