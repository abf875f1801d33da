/*
 * oracle.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A plain C++ FP64 restatement of the miniAFQMC walker step, written from the
 * reference's algorithm description.  The reference itself cannot be built in
 * this image (it needs Boost), so this file is the oracle.  Each function
 * cites the reference file:line it follows (paths relative to the reference
 * root).  Complex arrays are interleaved doubles; layouts are in oracle.h.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   orc_csrmm                reference known answers (sp_blas.hpp:175-197,
 *                            test_sparse_ops.cpp:226-272,329-385) + dense brute force
 *   orc_product              ma_operations.hpp:324-490 known answers + brute force
 *   orc_getrf/getri/invert/determinant
 *                            ma_operations.hpp:491-505 (A*inv(A)=I at 1e-14),
 *                            numpy/LAPACK det and inverse, permutation/triangular
 *                            closed forms
 *   orc_gelqf/orc_glq        Q*Q^H = I, L*Q = A, real diagonal of L with LAPACK's
 *                            sign rule, numpy QR up to the diagonal phase
 *   mixed density matrix     Gc*conj(Psi) = I, trial special case (Gc = Psi^T,
 *                            ovlp 1), compact vs full agree, numpy inverse
 *   vbias / vHS              brute-force dense products; transposed vs
 *                            non-transposed agree up to the rotation cutoff
 *   apply_expM / propagate   V = 0, diagonal closed form, numpy Taylor sum
 *   orthogonalize            W^T conj(W) = I, G invariant, overlap ratio = det(R)
 *   calculate_energy         Vakbl form == Cholesky form (1e-12), brute force E1
 *   fields / weights         hand-made closed-form rows
 *   halfrotate_cholesky      dense brute force with cutoff
 *   orc_branch               no reference implementation (the driver only marks
 *                            the place, miniafqmc_base.cpp:346): a hand-worked
 *                            comb, and the comb's invariants on random weights
 *                            (copies of walker i = floor or ceil of nw*w_i/W,
 *                            population nw, parents non-decreasing, zero-weight
 *                            walkers dropped, total weight kept)
 *   orc_compress             sorted (row, col) order + rowIndex by brute force
 *   orc_run                  composition of the pinned pieces; no end-to-end
 *                            reference fixture exists (the example afqmc.h5 is
 *                            not in the reference tree), so the trajectory
 *                            itself is "parity unpinned" beyond its pieces.
 */
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <vector>

#include <omp.h>

using cd = std::complex<double>;

namespace {

inline cd cmul(cd a, cd b)
{
  return cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
inline cd cdiv(cd a, cd b)
{
  // Smith's algorithm (what zladiv computes, up to rounding)
  double c = b.real(), d = b.imag();
  if (std::fabs(d) <= std::fabs(c)) {
    double r = d / c, den = c + d * r;
    return cd((a.real() + a.imag() * r) / den, (a.imag() - a.real() * r) / den);
  }
  double r = c / d, den = d + c * r;
  return cd((a.real() * r + a.imag()) / den, (a.imag() * r - a.real()) / den);
}
inline double cabs1(cd a) { return std::fabs(a.real()) + std::fabs(a.imag()); }

using clk = std::chrono::steady_clock;
inline double secs(clk::time_point a, clk::time_point b)
{
  return std::chrono::duration<double>(b - a).count();
}

/* Row-major product C[m][n] = alpha*op(A)*op(B) + beta*C.
 * ma_operations.hpp:94-107 maps product(A,B,C) onto column-major gemm(B,A);
 * the arithmetic is the same sum, written here directly in row-major. */
void product(char opA, char opB, int m, int n, int k, cd alpha, const cd* A, int lda,
             const cd* B, int ldb, cd beta, cd* C, int ldc)
{
  std::vector<cd> acc(n);
  for (int i = 0; i < m; i++) {
    std::fill(acc.begin(), acc.end(), cd(0.0));
    for (int l = 0; l < k; l++) {
      cd a;
      if (opA == 'N') a = A[(size_t)i * lda + l];
      else if (opA == 'T') a = A[(size_t)l * lda + i];
      else a = std::conj(A[(size_t)l * lda + i]);
      if (a == cd(0.0)) continue;
      if (opB == 'N') {
        const cd* Bl = B + (size_t)l * ldb;
        for (int j = 0; j < n; j++) acc[j] += cmul(a, Bl[j]);
      } else if (opB == 'T') {
        for (int j = 0; j < n; j++) acc[j] += cmul(a, B[(size_t)j * ldb + l]);
      } else {
        for (int j = 0; j < n; j++) acc[j] += cmul(a, std::conj(B[(size_t)j * ldb + l]));
      }
    }
    cd* Ci = C + (size_t)i * ldc;
    for (int j = 0; j < n; j++) {
      cd c0 = (beta == cd(0.0)) ? cd(0.0) : cmul(beta, Ci[j]);
      Ci[j] = c0 + cmul(alpha, acc[j]);
    }
  }
}

/* ---- LAPACK restatements on the column-major view (ma_lapack.hpp) ---- */
/* zgetf2 (unblocked right-looking LU with partial pivoting, izamax = first
 * maximum of |re|+|im|).  Column-major m x n, ACM(i,j) = A[j*lda+i]. */
int zgetf2(int m, int n, cd* A, int lda, int* ipiv)
{
#define ACM(i, j) A[(size_t)(j) * lda + (i)]
  int info = 0;
  const double sfmin = std::numeric_limits<double>::min();
  int mn = std::min(m, n);
  for (int j = 0; j < mn; j++) {
    int jp = j;
    double amax = cabs1(ACM(j, j));
    for (int i = j + 1; i < m; i++)
      if (cabs1(ACM(i, j)) > amax) { amax = cabs1(ACM(i, j)); jp = i; }
    ipiv[j] = jp + 1;
    if (ACM(jp, j) != cd(0.0)) {
      if (jp != j)
        for (int c = 0; c < n; c++) std::swap(ACM(j, c), ACM(jp, c));
      if (j < m - 1) {
        if (std::abs(ACM(j, j)) >= sfmin) {
          cd r = cdiv(cd(1.0), ACM(j, j));
          for (int i = j + 1; i < m; i++) ACM(i, j) = cmul(r, ACM(i, j));
        } else {
          for (int i = j + 1; i < m; i++) ACM(i, j) = cdiv(ACM(i, j), ACM(j, j));
        }
      }
    } else if (info == 0) {
      info = j + 1;
    }
    if (j < mn - 1) {
      for (int c = j + 1; c < n; c++) {
        cd u = ACM(j, c);
        if (u == cd(0.0)) continue;
        for (int i = j + 1; i < m; i++) ACM(i, c) -= cmul(ACM(i, j), u);
      }
    }
  }
  return info;
#undef ACM
}

/* ztrti2 (upper, non-unit) + zgetri unblocked path. */
int zgetri(int n, cd* A, int lda, const int* ipiv)
{
#define ACM(i, j) A[(size_t)(j) * lda + (i)]
  for (int j = 0; j < n; j++)
    if (ACM(j, j) == cd(0.0)) return j + 1;
  // ztrti2: inv(U) in place
  for (int j = 0; j < n; j++) {
    ACM(j, j) = cdiv(cd(1.0), ACM(j, j));
    cd ajj = -ACM(j, j);
    // ztrmv('U','N','N', j, A, lda, A(0,j), 1): x := U(0:j,0:j) x
    for (int jj = 0; jj < j; jj++) {
      cd temp = ACM(jj, j);
      if (temp != cd(0.0)) {
        for (int i = 0; i < jj; i++) ACM(i, j) += cmul(temp, ACM(i, jj));
        ACM(jj, j) = cmul(temp, ACM(jj, jj));
      }
    }
    for (int i = 0; i < j; i++) ACM(i, j) = cmul(ajj, ACM(i, j));
  }
  // solve inv(A)*L = inv(U)
  std::vector<cd> work(n);
  for (int j = n - 1; j >= 0; j--) {
    for (int i = j + 1; i < n; i++) {
      work[i] = ACM(i, j);
      ACM(i, j) = cd(0.0);
    }
    if (j < n - 1) {
      for (int l = j + 1; l < n; l++) {
        cd wl = work[l];
        if (wl == cd(0.0)) continue;
        for (int i = 0; i < n; i++) ACM(i, j) -= cmul(ACM(i, l), wl);
      }
    }
  }
  // undo column interchanges
  for (int j = n - 2; j >= 0; j--) {
    int jp = ipiv[j] - 1;
    if (jp != j)
      for (int i = 0; i < n; i++) std::swap(ACM(i, j), ACM(i, jp));
  }
  return 0;
#undef ACM
}

double dnrm2_strided(int n, const cd* x, long inc)
{
  // scaled sum of squares, as dznrm2
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < n; i++) {
    const cd v = x[(size_t)i * inc];
    double parts[2] = {v.real(), v.imag()};
    for (double p : parts) {
      if (p != 0.0) {
        double a = std::fabs(p);
        if (scale < a) {
          ssq = 1.0 + ssq * (scale / a) * (scale / a);
          scale = a;
        } else {
          ssq += (a / scale) * (a / scale);
        }
      }
    }
  }
  return scale * std::sqrt(ssq);
}
double dlapy3(double x, double y, double z)
{
  double xa = std::fabs(x), ya = std::fabs(y), za = std::fabs(z);
  double w = std::max(xa, std::max(ya, za));
  if (w == 0.0) return xa + ya + za;
  return w * std::sqrt((xa / w) * (xa / w) + (ya / w) * (ya / w) + (za / w) * (za / w));
}

/* zlarfg: elementary reflector with real beta, LAPACK's sign rule
 * beta = -sign(|(alpha,x)|, Re alpha). */
void zlarfg(int n, cd& alpha, cd* x, long incx, cd& tau)
{
  if (n <= 0) { tau = 0.0; return; }
  double xnorm = dnrm2_strided(n - 1, x, incx);
  double alphr = alpha.real(), alphi = alpha.imag();
  if (xnorm == 0.0 && alphi == 0.0) { tau = 0.0; return; }
  double beta = -std::copysign(dlapy3(alphr, alphi, xnorm), alphr);
  const double safmin = std::numeric_limits<double>::min() / std::numeric_limits<double>::epsilon();
  const double rsafmn = 1.0 / safmin;
  int knt = 0;
  if (std::fabs(beta) < safmin) {
    do {
      knt++;
      for (int i = 0; i < n - 1; i++) x[(size_t)i * incx] *= rsafmn;
      beta *= rsafmn;
      alphi *= rsafmn;
      alphr *= rsafmn;
    } while (std::fabs(beta) < safmin && knt < 20);
    xnorm = dnrm2_strided(n - 1, x, incx);
    alpha = cd(alphr, alphi);
    beta = -std::copysign(dlapy3(alphr, alphi, xnorm), alphr);
  }
  tau = cd((beta - alphr) / beta, -alphi / beta);
  cd s = cdiv(cd(1.0), cd(alphr - beta, alphi));
  for (int i = 0; i < n - 1; i++) x[(size_t)i * incx] = cmul(s, x[(size_t)i * incx]);
  for (int j = 0; j < knt; j++) beta *= safmin;
  alpha = beta;
}

/* zlarf('Right'): C(mm x nn, col-major ldc) := C (I - tau v v^H), v strided. */
void zlarf_right(int mm, int nn, const cd* v, long incv, cd tau, cd* C, int ldc, std::vector<cd>& w)
{
  if (tau == cd(0.0) || mm <= 0) return;
  w.assign(mm, cd(0.0));
  for (int c = 0; c < nn; c++) {
    cd vc = v[(size_t)c * incv];
    if (vc == cd(0.0)) continue;
    const cd* Cc = C + (size_t)c * ldc;
    for (int r = 0; r < mm; r++) w[r] += cmul(Cc[r], vc);
  }
  for (int c = 0; c < nn; c++) {
    cd f = cmul(-tau, std::conj(v[(size_t)c * incv]));
    cd* Cc = C + (size_t)c * ldc;
    for (int r = 0; r < mm; r++) Cc[r] += cmul(w[r], f);
  }
}

/* zgelq2 on column-major m x n. */
void zgelq2(int m, int n, cd* A, int lda, cd* tau)
{
#define ACM(i, j) A[(size_t)(j) * lda + (i)]
  std::vector<cd> w;
  int k = std::min(m, n);
  for (int i = 0; i < k; i++) {
    for (int c = i; c < n; c++) ACM(i, c) = std::conj(ACM(i, c));
    cd alpha = ACM(i, i);
    zlarfg(n - i, alpha, &ACM(i, std::min(i + 1, n - 1)), lda, tau[i]);
    if (i < m - 1) {
      ACM(i, i) = 1.0;
      zlarf_right(m - i - 1, n - i, &ACM(i, i), lda, tau[i], &ACM(i + 1, i), lda, w);
    }
    ACM(i, i) = alpha;
    for (int c = i; c < n; c++) ACM(i, c) = std::conj(ACM(i, c));
  }
#undef ACM
}

/* zungl2 on column-major m x n with k reflectors. */
void zungl2(int m, int n, int k, cd* A, int lda, const cd* tau)
{
#define ACM(i, j) A[(size_t)(j) * lda + (i)]
  std::vector<cd> w;
  if (k < m) {
    for (int j = 0; j < n; j++) {
      for (int l = k; l < m; l++) ACM(l, j) = 0.0;
      if (j >= k && j < m) ACM(j, j) = 1.0;
    }
  }
  for (int i = k - 1; i >= 0; i--) {
    if (i < n - 1) {
      for (int c = i + 1; c < n; c++) ACM(i, c) = std::conj(ACM(i, c));
      if (i < m - 1) {
        ACM(i, i) = 1.0;
        zlarf_right(m - i - 1, n - i, &ACM(i, i), lda, std::conj(tau[i]), &ACM(i + 1, i), lda, w);
      }
      cd s = -tau[i];
      for (int c = i + 1; c < n; c++) ACM(i, c) = cmul(s, ACM(i, c));
      for (int c = i + 1; c < n; c++) ACM(i, c) = std::conj(ACM(i, c));
    }
    ACM(i, i) = cd(1.0) - std::conj(tau[i]);
    for (int l = 0; l < i; l++) ACM(i, l) = 0.0;
  }
#undef ACM
}

/* ma::invert (ma_operations.hpp:244-260): getrf, det from the pivots, getri */
int invert_sq(int n, cd* A, int* piv, cd* det)
{
  int info = zgetf2(n, n, A, n, piv);
  cd d(1.0);
  for (int i = 0; i < n; i++) d = (piv[i] == i + 1) ? cmul(d, A[(size_t)i * n + i]) : cmul(d, -A[(size_t)i * n + i]);
  if (det) *det = d;
  if (info != 0) return info;
  return zgetri(n, A, n, piv);
}
/* ma::determinant (ma_operations.hpp:276-291) */
int determinant_sq(int n, cd* A, int* piv, cd* det)
{
  int info = zgetf2(n, n, A, n, piv);
  cd d(1.0);
  for (int i = 0; i < n; i++) d = (piv[i] == i + 1) ? cmul(d, A[(size_t)i * n + i]) : cmul(d, -A[(size_t)i * n + i]);
  *det = d;
  return info;
}

/* base::MixedDensityMatrix (mixed_density_matrix.hpp:55-104) */
int mixed_dm(int M, int N, const cd* conjA, const cd* B, cd* C, bool compact, cd* ovlp,
             std::vector<cd>& T1, std::vector<cd>& T2, std::vector<int>& piv)
{
  T1.resize((size_t)N * N);
  piv.resize(N);
  // T1 = T(B) * conjA  (N x N)
  product('T', 'N', N, N, M, 1.0, B, N, conjA, N, 0.0, T1.data(), N);
  int info = invert_sq(N, T1.data(), piv.data(), ovlp);
  if (compact) {
    // C = T1 * T(B)  (N x M)
    product('N', 'T', N, M, N, 1.0, T1.data(), N, B, N, 0.0, C, M);
  } else {
    T2.resize((size_t)N * M);
    product('N', 'T', N, M, N, 1.0, T1.data(), N, B, N, 0.0, T2.data(), M);
    product('N', 'N', M, M, N, 1.0, conjA, N, T2.data(), M, 0.0, C, M);
  }
  return info;
}

/* base::Overlap (mixed_density_matrix.hpp:117-138) */
int overlap(int M, int N, const cd* conjA, const cd* B, cd* ovlp, std::vector<cd>& T1, std::vector<int>& piv)
{
  T1.resize((size_t)N * N);
  piv.resize(N);
  product('T', 'N', N, N, M, 1.0, B, N, conjA, N, 0.0, T1.data(), N);
  return determinant_sq(N, T1.data(), piv.data(), ovlp);
}

/* mySPBLAS::csrmm complex (sparse.hpp:159-216), 0-based ("GxxCxx"). */
void csrmm(char transa, int M, int N, int K, cd alpha, const cd* A, const int* indx,
           const int* pntrb, const int* pntre, const cd* B, int ldb, cd beta, cd* C, int ldc)
{
  int p0 = *pntrb;
  if (transa == 'n' || transa == 'N') {
    // rows are independent: each is summed by one thread in entry order, so
    // the result does not depend on the thread count
#pragma omp parallel for schedule(dynamic, 64)
    for (int nr = 0; nr < M; nr++) {
      cd* Cr = C + (size_t)nr * ldc;
      for (int i = 0; i < N; i++) Cr[i] = cmul(Cr[i], beta);
      for (int i = pntrb[nr] - p0; i < pntre[nr] - p0; i++) {
        if (indx[i] >= K) continue;
        const cd* Bc = B + (size_t)ldb * indx[i];
        cd Arc = cmul(alpha, A[i]);
        for (int k = 0; k < N; k++) Cr[k] += cmul(Arc, Bc[k]);
      }
    }
  } else {
    bool herm = (transa == 'h' || transa == 'H');
    for (int i = 0; i < K; i++)
      for (int j = 0; j < N; j++) C[(size_t)i * ldc + j] = cmul(C[(size_t)i * ldc + j], beta);
    for (int nr = 0; nr < M; nr++, pntrb++, pntre++, B += ldb) {
      for (int i = *pntrb - p0; i < *pntre - p0; i++) {
        if (indx[i] >= K) continue;
        cd* Cc = C + (size_t)ldc * indx[i];
        cd Arc = cmul(alpha, herm ? std::conj(A[i]) : A[i]);
        for (int k = 0; k < N; k++) Cc[k] += cmul(Arc, B[k]);
      }
    }
  }
}

/* base::apply_expM (vHS.hpp:61-91) */
void apply_expM(int M, int N, const cd* V, cd* S, int order, std::vector<cd>& T1, std::vector<cd>& T2)
{
  T1.assign(S, S + (size_t)M * N);
  T2.resize((size_t)M * N);
  cd* r1 = T1.data();
  cd* r2 = T2.data();
  for (int n = 1; n <= order; n++) {
    cd fact = cmul(cd(0.0, 1.0), cd(1.0 / static_cast<double>(n)));
    product('N', 'N', M, N, M, fact, V, M, r1, N, 0.0, r2, N);
    for (size_t i = 0; i < (size_t)M * N; i++) S[i] += r2[i];
    std::swap(r1, r2);
  }
}

struct Sys {
  Sys(int M_, int N_, const cd* a, const cd* b) : M(M_), N(N_), ta(a), tb(b) {}
  int M, N;
  const cd* ta;
  const cd* tb;
  std::vector<cd> T1, T2, TMN, TMM, DM, E1, E2;
  std::vector<int> piv;
};

/* afqmc_sys::calculate_mixed_density_matrix (afqmc_sys.hpp:95-116) */
// Walkers are independent: the walker loops below run in parallel over
// walkers (each walker's arithmetic is done by one thread, in the reference's
// order, so results do not depend on the thread count); `info` is that of the
// first failing walker in index order.
int first_info(const std::vector<int>& r)
{
  for (int x : r)
    if (x) return x;
  return 0;
}

int sys_mdm(Sys& s, int nw, const cd* W, cd* Wd, cd* G, bool compact)
{
  int M = s.M, N = s.N, N_ = compact ? N : M;
  std::vector<int> infos((size_t)2 * nw, 0);
#pragma omp parallel
  {
    Sys t(M, N, s.ta, s.tb);
    t.DM.resize((size_t)N_ * M);
#pragma omp for schedule(dynamic)
    for (int n = 0; n < nw; n++) {
      for (int sp = 0; sp < 2; sp++) {
        const cd* Wn = W + ((size_t)n * 2 + sp) * M * N;
        cd ov;
        infos[(size_t)2 * n + sp] = mixed_dm(M, N, sp == 0 ? t.ta : t.tb, Wn, t.DM.data(), compact, &ov, t.T1, t.T2,
                                             t.piv);
        Wd[(size_t)n * 8 + 2 + sp] = ov;
        for (int a = 0; a < N_; a++)
          for (int k = 0; k < M; k++)
            G[(((size_t)sp * N_ + a) * M + k) * nw + n] = t.DM[(size_t)a * M + k];
      }
    }
  }
  return first_info(infos);
}

/* afqmc_sys::calculate_overlaps (afqmc_sys.hpp:135-144) */
int sys_overlaps(Sys& s, int nw, const cd* W, cd* Wd)
{
  std::vector<int> infos((size_t)2 * nw, 0);
#pragma omp parallel
  {
    Sys t(s.M, s.N, s.ta, s.tb);
#pragma omp for schedule(dynamic)
    for (int n = 0; n < nw; n++)
      for (int sp = 0; sp < 2; sp++) {
        cd ov;
        infos[(size_t)2 * n + sp] =
            overlap(t.M, t.N, sp == 0 ? t.ta : t.tb, W + ((size_t)n * 2 + sp) * t.M * t.N, &ov, t.T1, t.piv);
        Wd[(size_t)n * 8 + 2 + sp] = ov;
      }
  }
  return first_info(infos);
}

/* afqmc_sys::propagate (afqmc_sys.hpp:146-173) */
void sys_propagate(Sys& s, int nw, const cd* P, const cd* vHS, cd* W)
{
  int M = s.M, N = s.N;
  // (walker, spin) pairs are independent
#pragma omp parallel
  {
    Sys t(M, N, s.ta, s.tb);
    t.TMM.resize((size_t)M * M);
    t.TMN.resize((size_t)M * N);
    int wcur = -1;
#pragma omp for schedule(dynamic)
    for (int ws = 0; ws < 2 * nw; ws++) {
      const int w = ws / 2, sp = ws % 2;
      if (w != wcur) {
        for (int i = 0; i < M * M; i++) t.TMM[i] = vHS[(size_t)i * nw + w];
        wcur = w;
      }
      cd* Ws = W + ((size_t)w * 2 + sp) * M * N;
      product('N', 'N', M, N, M, 1.0, P, M, Ws, N, 0.0, t.TMN.data(), N);
      apply_expM(M, N, t.TMM.data(), t.TMN.data(), 6, t.E1, t.E2);
      product('N', 'N', M, N, M, 1.0, P, M, t.TMN.data(), N, 0.0, Ws, N);
    }
  }
}

/* afqmc_sys::orthogonalize (afqmc_sys.hpp:175-207): gelqf + glq per spin on
 * the column-major view (N x M) of W[w][s] ([M][N] row-major). */
void sys_orthogonalize(int M, int N, int nw, cd* W)
{
#pragma omp parallel
  {
    std::vector<cd> tau(std::max(M, N));
#pragma omp for schedule(dynamic)
    for (int ws = 0; ws < 2 * nw; ws++) {
      cd* A = W + (size_t)ws * M * N;
      zgelq2(N, M, A, N, tau.data());
      zungl2(N, M, std::min(M, N), A, N, tau.data());
    }
  }
}

/* base::calculate_energy (energy.hpp:63-102) + afqmc_sys::calculate_energy
 * (afqmc_sys.hpp:118-133) */
double sys_energy_vakbl(int M, int N, int nw, cd* Wd, const cd* Gc, const cd* haj, int vrows,
                        const int* vrp, const int* vcol, const cd* vval)
{
  int NAK = 2 * N * M;
  std::vector<cd> Gcloc((size_t)NAK * nw, cd(0.0));
  for (int n = 0; n < nw; n++) Wd[(size_t)n * 8] = 0.0;
  csrmm('N', vrows, nw, NAK, 1.0, vval, vcol, vrp, vrp + 1, Gc, nw, 0.0, Gcloc.data(), nw);
  for (int i = 0; i < NAK; i++)
    for (int n = 0; n < nw; n++) Wd[(size_t)n * 8] += cmul(Gc[(size_t)i * nw + n], Gcloc[(size_t)i * nw + n]);
  for (int n = 0; n < nw; n++) Wd[(size_t)n * 8] = cmul(Wd[(size_t)n * 8], cd(0.5));
  // gemv: W_data[:,0] += T(Gc) * haj
  for (int n = 0; n < nw; n++) {
    cd acc(0.0);
    for (int i = 0; i < NAK; i++) acc += cmul(Gc[(size_t)i * nw + n], haj[i]);
    Wd[(size_t)n * 8] += acc;
  }
  double eav = 0, wgt = 0;
  for (int n = 0; n < nw; n++) {
    wgt += Wd[(size_t)n * 8 + 1].real();
    eav += Wd[(size_t)n * 8].real() * Wd[(size_t)n * 8 + 1].real();
  }
  return eav / wgt;
}

/* Two-body energy in Cholesky form; algebraically identical to
 * 0.5*Gc^T Vakbl Gc when Vakbl(ak,bl) = sum_n Lt(n,ak)Lt(n,bl)
 * - delta_spin sum_n Lt(n,al)Lt(n,bk).  Lt is unscaled [K][2NM]. */
double sys_energy_chol(int M, int N, int K, int nw, cd* Wd, const cd* Gc, const cd* haj, const cd* Lt)
{
  const int NAK = 2 * N * M, NM = N * M;
  // The sums over n run in kNChunk fixed chunks of consecutive vectors, each
  // chunk summed in order by one thread and the chunks added in order, so the
  // result does not depend on the thread count.  Y^n = Lt^n_s G_s^T is formed
  // as a row-major product with the transposed G (contiguous inner loops).
  constexpr int kNChunk = 64;
  const int per = (K + kNChunk - 1) / kNChunk;
  std::vector<cd> dpart((size_t)nw * kNChunk), xpart((size_t)nw * kNChunk);
  std::vector<cd> gT((size_t)nw * 2 * M * N);  // [w][s][l][b] = G_s(b, l)
  for (int w = 0; w < nw; w++)
    for (int sp = 0; sp < 2; sp++)
      for (int b = 0; b < N; b++)
        for (int l = 0; l < M; l++)
          gT[(((size_t)w * 2 + sp) * M + l) * N + b] = Gc[((size_t)sp * NM + (size_t)b * M + l) * nw + w];
#pragma omp parallel
  {
    std::vector<cd> g(NAK), Y((size_t)N * N);
#pragma omp for schedule(dynamic)
    for (int task = 0; task < nw * kNChunk; task++) {
      const int w = task / kNChunk, ch = task % kNChunk;
      for (int i = 0; i < NAK; i++) g[i] = Gc[(size_t)i * nw + w];
      cd direct(0.0), exch(0.0);
      for (int n = ch * per; n < std::min(K, (ch + 1) * per); n++) {
        const cd* L = Lt + (size_t)n * NAK;
        cd v(0.0);
        for (int i = 0; i < NAK; i++) v += cmul(L[i], g[i]);
        direct += cmul(v, v);
        for (int sp = 0; sp < 2; sp++) {
          // Y[a][b] = sum_l Ls[a][l] G_s[b][l]
          product('N', 'N', N, N, M, 1.0, L + (size_t)sp * NM, M, gT.data() + ((size_t)w * 2 + sp) * M * N, N, 0.0,
                  Y.data(), N);
          for (int a = 0; a < N; a++)
            for (int b = 0; b < N; b++) exch += cmul(Y[(size_t)a * N + b], Y[(size_t)b * N + a]);
        }
      }
      dpart[task] = direct;
      xpart[task] = exch;
    }
  }
  for (int w = 0; w < nw; w++) {
    cd direct(0.0), exch(0.0);
    for (int ch = 0; ch < kNChunk; ch++) {
      direct += dpart[(size_t)w * kNChunk + ch];
      exch += xpart[(size_t)w * kNChunk + ch];
    }
    cd e2 = cmul(cd(0.5), direct - exch);
    cd e1(0.0);
    for (int i = 0; i < NAK; i++) e1 += cmul(Gc[(size_t)i * nw + w], haj[i]);
    Wd[(size_t)w * 8] = e2 + e1;
  }
  double eav = 0, wgt = 0;
  for (int n = 0; n < nw; n++) {
    wgt += Wd[(size_t)n * 8 + 1].real();
    eav += Wd[(size_t)n * 8].real() * Wd[(size_t)n * 8 + 1].real();
  }
  return eav / wgt;
}

/* X / hybrid-weight loop (miniafqmc_base.cpp:281-289); X holds x on entry. */
void fields(int K, int nw, const cd* vbias, cd* X, cd* hybridW)
{
  const cd im(0.0, 1.0), halfim(0.0, 0.5);
  for (int w = 0; w < nw; w++) hybridW[w] = 0.0;
  for (int n = 0; n < K; n++)
    for (int w = 0; w < nw; w++) {
      size_t i = (size_t)n * nw + w;
      hybridW[w] -= cmul(cmul(im, vbias[i]), X[i] + cmul(halfim, vbias[i]));
      X[i] += cmul(im, vbias[i]);
    }
}

/* weight / hybrid energy / Eshift loop (miniafqmc_base.cpp:317-327) */
double weights(int nw, double dt, const cd* hybridW, cd* Wd, double Eshift)
{
  double et = 0.0;
  for (int w = 0; w < nw; w++) {
    cd* d = Wd + (size_t)w * 8;
    cd ratio = cdiv(cmul(d[2], d[3]), cmul(d[6], d[7]));
    double scale = std::max(0.0, std::cos(std::arg(ratio)));
    d[4] = -(hybridW[w] + std::log(ratio)) / dt;
    d[1] = cmul(d[1], cd(scale * std::exp(-dt * (0.5 * (d[4].real() + d[5].real()) - Eshift)), 0.0));
    et += d[4].real();
  }
  return et / nw;
}

/* CSR transpose (stable: entries of each output row keep ascending input row) */
void csr_transpose(int rows, int cols, const int* rp, const int* ci, const cd* v, std::vector<int>& trp,
                   std::vector<int>& tci, std::vector<cd>& tv)
{
  // Rows are split into one contiguous range per thread; each thread counts
  // its entries per column, the counts are prefix-summed column-major over
  // the threads, and each thread scatters its range in row order, so every
  // column keeps its entries in row order (the serial result).
  const long nnz = rp[rows] - rp[0];
  const int nt = std::max(1, std::min(omp_get_max_threads(), rows));
  std::vector<long> cnt((size_t)nt * cols, 0);
  auto lo = [&](int t) { return (int)((long)rows * t / nt); };
#pragma omp parallel for schedule(static, 1) num_threads(nt)
  for (int t = 0; t < nt; t++) {
    long* ct = cnt.data() + (size_t)t * cols;
    for (long p = rp[lo(t)] - rp[0]; p < rp[lo(t + 1)] - rp[0]; p++) ct[ci[p]]++;
  }
  trp.assign(cols + 1, 0);
  long run = 0;
  for (int c = 0; c < cols; c++) {
    trp[c] = (int)run;
    for (int t = 0; t < nt; t++) {
      const long n = cnt[(size_t)t * cols + c];
      cnt[(size_t)t * cols + c] = run;  // thread t's first slot in column c
      run += n;
    }
  }
  trp[cols] = (int)run;
  tci.resize(nnz);
  tv.resize(nnz);
#pragma omp parallel for schedule(static, 1) num_threads(nt)
  for (int t = 0; t < nt; t++) {
    long* pos = cnt.data() + (size_t)t * cols;
    for (int r = lo(t); r < lo(t + 1); r++)
      for (long p = rp[r] - rp[0]; p < rp[r + 1] - rp[0]; p++) {
        const int c = ci[p];
        tci[pos[c]] = r;
        tv[pos[c]] = v[p];
        pos[c]++;
      }
  }
}

} // namespace

#define CZ(p) reinterpret_cast<cd*>(p)
#define CCZ(p) reinterpret_cast<const cd*>(p)

extern "C" {

int orc_csrmm(char transa, int M, int N, int K, const double* alpha, const double* val, const int* indx,
              const int* pntrb, const int* pntre, const double* B, int ldb, const double* beta, double* C, int ldc)
{
  csrmm(transa, M, N, K, cd(alpha[0], alpha[1]), CCZ(val), indx, pntrb, pntre, CCZ(B), ldb, cd(beta[0], beta[1]),
        CZ(C), ldc);
  return 0;
}

void orc_product(char opA, char opB, int m, int n, int k, const double* alpha, const double* A, int lda,
                 const double* B, int ldb, const double* beta, double* C, int ldc)
{
  product(opA, opB, m, n, k, cd(alpha[0], alpha[1]), CCZ(A), lda, CCZ(B), ldb, cd(beta[0], beta[1]), CZ(C), ldc);
}

/* ma::getrf (ma_lapack.hpp:33-47): LAPACK getrf(M=cols, N=rows, lda). */
int orc_getrf(int rows, int cols, double* A, int lda, int* piv) { return zgetf2(cols, rows, CZ(A), lda, piv); }
/* ma::getri (ma_lapack.hpp:63-79) */
int orc_getri(int n, double* A, int lda, const int* piv) { return zgetri(n, CZ(A), lda, piv); }
int orc_invert(int n, double* A, int* piv, double* det_out)
{
  cd d;
  int r = invert_sq(n, CZ(A), piv, &d);
  det_out[0] = d.real();
  det_out[1] = d.imag();
  return r;
}
int orc_determinant(int n, double* A, int* piv, double* det_out)
{
  cd d;
  int r = determinant_sq(n, CZ(A), piv, &d);
  det_out[0] = d.real();
  det_out[1] = d.imag();
  return r;
}
/* ma::gelqf / ma::glq (ma_lapack.hpp:134-221): column-major view (cols x rows) */
int orc_gelqf(int rows, int cols, double* A, int lda, double* tau)
{
  zgelq2(cols, rows, CZ(A), lda, CZ(tau));
  return 0;
}
int orc_glq(int rows, int cols, double* A, int lda, const double* tau)
{
  zungl2(cols, rows, std::min(rows, cols), CZ(A), lda, CCZ(tau));
  return 0;
}

int orc_mixed_density_matrix(int M, int N, const double* conjA, const double* B, double* C, int compact,
                             double* ovlp_out)
{
  std::vector<cd> T1, T2;
  std::vector<int> piv;
  cd ov;
  int r = mixed_dm(M, N, CCZ(conjA), CCZ(B), CZ(C), compact != 0, &ov, T1, T2, piv);
  ovlp_out[0] = ov.real();
  ovlp_out[1] = ov.imag();
  return r;
}

int orc_overlap(int M, int N, const double* conjA, const double* B, double* ovlp_out)
{
  std::vector<cd> T1;
  std::vector<int> piv;
  cd ov;
  int r = overlap(M, N, CCZ(conjA), CCZ(B), &ov, T1, piv);
  ovlp_out[0] = ov.real();
  ovlp_out[1] = ov.imag();
  return r;
}

int orc_calculate_mixed_density_matrix(int M, int N, int nw, const double* ta, const double* tb, const double* W,
                                       double* Wd, double* G, int compact)
{
  Sys s(M, N, CCZ(ta), CCZ(tb));
  return sys_mdm(s, nw, CCZ(W), CZ(Wd), CZ(G), compact != 0);
}

int orc_calculate_overlaps(int M, int N, int nw, const double* ta, const double* tb, const double* W, double* Wd)
{
  Sys s(M, N, CCZ(ta), CCZ(tb));
  return sys_overlaps(s, nw, CCZ(W), CZ(Wd));
}

/* base::get_vbias (vbias.hpp:45-83).  transposed: rows=K, cols=2NM (SpvnT);
 * else rows=M*M, cols=K (Spvn) and G is the full [2*M*M][nw]. */
int orc_get_vbias(int rows, int cols, const int* rp, const int* col, const double* val, int grows, int nw,
                  const double* G, double* v, int transposed)
{
  if (transposed) {
    if (cols != grows) return 1;
    csrmm('N', rows, nw, cols, 1.0, CCZ(val), col, rp, rp + 1, CCZ(G), nw, 0.0, CZ(v), nw);
  } else {
    if (2 * rows != grows) return 1;
    // alpha half: v = T(Spvn) G[0:rows]; beta half: v += T(Spvn) G[rows:2rows]
    csrmm('T', rows, nw, cols, 1.0, CCZ(val), col, rp, rp + 1, CCZ(G), nw, 0.0, CZ(v), nw);
    csrmm('T', rows, nw, cols, 1.0, CCZ(val), col, rp, rp + 1, CCZ(G) + (size_t)rows * nw, nw, 1.0, CZ(v), nw);
  }
  return 0;
}

/* base::get_vHS (vHS.hpp:43-56) */
int orc_get_vhs(int rows, int cols, const int* rp, const int* col, const double* val, int nw, const double* X,
                double* v)
{
  csrmm('N', rows, nw, cols, 1.0, CCZ(val), col, rp, rp + 1, CCZ(X), nw, 0.0, CZ(v), nw);
  return 0;
}

int orc_apply_expM(int M, int N, const double* V, double* S, int order)
{
  std::vector<cd> T1, T2;
  apply_expM(M, N, CCZ(V), CZ(S), order, T1, T2);
  return 0;
}

int orc_propagate(int M, int N, int nw, const double* P, const double* vHS, double* W)
{
  Sys s(M, N, nullptr, nullptr);
  sys_propagate(s, nw, CCZ(P), CCZ(vHS), CZ(W));
  return 0;
}

int orc_orthogonalize(int M, int N, int nw, double* W)
{
  sys_orthogonalize(M, N, nw, CZ(W));
  return 0;
}

int orc_calculate_energy(int M, int N, int nw, double* Wd, const double* Gc, const double* haj, int vrows,
                         const int* vrp, const int* vcol, const double* vval, double* eav_out)
{
  *eav_out = sys_energy_vakbl(M, N, nw, CZ(Wd), CCZ(Gc), CCZ(haj), vrows, vrp, vcol, CCZ(vval));
  return 0;
}

int orc_energy_cholesky(int M, int N, int K, int nw, double* Wd, const double* Gc, const double* haj,
                        const double* Lt, double* eav_out)
{
  *eav_out = sys_energy_chol(M, N, K, nw, CZ(Wd), CCZ(Gc), CCZ(haj), CCZ(Lt));
  return 0;
}

int orc_fields(int K, int nw, const double* vbias, const double* x, double* X, double* hybridW)
{
  cd* Xc = CZ(X);
  for (size_t i = 0; i < (size_t)K * nw; i++) Xc[i] = cd(x[i], 0.0);
  fields(K, nw, CCZ(vbias), Xc, CZ(hybridW));
  return 0;
}

int orc_weights(int nw, double dt, const double* hybridW, double* Wd, double* eshift_inout)
{
  *eshift_inout = weights(nw, dt, CCZ(hybridW), CZ(Wd), *eshift_inout);
  return 0;
}

/* Comb population control at the place the driver marks for branching
 * (miniafqmc_base.cpp:346).  Cumulative weights max(0, Re Wd[i][1]) in index
 * order; tooth j at (u + j) * (W / nw) takes the first walker whose cumulative
 * weight exceeds it (the last walker of positive weight if rounding puts the
 * tooth at or past W).  W and Wd are gathered in place; every new weight is
 * W / nw.  Returns 1 (nothing changed) if W is not positive and finite or a
 * weight is not finite. */
int orc_branch(int M, int N, int nw, double u, double* W_, double* Wd_, int* parent)
{
  cd* W = CZ(W_);
  cd* Wd = CZ(Wd_);
  std::vector<double> cum(nw);
  double sum = 0.0;
  int last = -1;
  bool bad = false;
  for (int i = 0; i < nw; i++) {
    const double w = Wd[(size_t)i * 8 + 1].real();
    if (!std::isfinite(w)) bad = true;
    if (w > 0.0) {
      sum += w;
      last = i;
    }
    cum[i] = sum;
  }
  if (bad || !(sum > 0.0) || !std::isfinite(sum)) return 1;
  const double step = cum[nw - 1] / nw;
  for (int j = 0; j < nw; j++) {
    const double t = (u + j) * step;
    int i = 0;
    while (i < nw && !(cum[i] > t)) i++;
    parent[j] = i < nw ? i : last;
  }
  const size_t per = (size_t)2 * M * N;
  std::vector<cd> Wn((size_t)nw * per), Wdn((size_t)nw * 8);
  for (int j = 0; j < nw; j++) {
    std::copy(W + parent[j] * per, W + (parent[j] + 1) * per, Wn.begin() + j * per);
    std::copy(Wd + (size_t)parent[j] * 8, Wd + (size_t)parent[j] * 8 + 8, Wdn.begin() + (size_t)j * 8);
    Wdn[(size_t)j * 8 + 1] = cd(sum / nw, 0.0);
  }
  std::copy(Wn.begin(), Wn.end(), W);
  std::copy(Wdn.begin(), Wdn.end(), Wd);
  return 0;
}

/* base::halfrotate_cholesky (rotate.hpp:65-169).  The reference counts in a
 * first pass and fills in a second with identical arithmetic; one pass that
 * keeps the kept entries gives the same matrix. */
struct HalfRotation {
  int K;
  std::vector<std::vector<int>> rc;
  std::vector<std::vector<cd>> rv;
};

static HalfRotation* halfrotate(int M, int N, int K, const double* ta_, const double* tb_, const int* rp,
                                const int* col, const double* val, double cutoff)
{
  auto* h = new HalfRotation;
  h->K = K;
  const cd* ta = CCZ(ta_);
  const cd* tb = CCZ(tb_);
  std::vector<int> trp, tci;
  std::vector<cd> tv;
  csr_transpose(M * M, K, rp, col, CCZ(val), trp, tci, tv);
  auto& rc = h->rc;
  auto& rv = h->rv;
  rc.resize(K);
  rv.resize(K);
#pragma omp parallel
  {
    std::vector<cd> C((size_t)N * M), An((size_t)M * M);
    std::vector<double> Cr((size_t)M), Ci((size_t)M);
#pragma omp for schedule(dynamic)
    for (int i = 0; i < K; i++) {
      int nterms = trp[i + 1] - trp[i];
      if (nterms == 0) continue;
      // extract Cholesky vector i as a dense M x M matrix (rotate.hpp:103-111)
      std::fill(An.begin(), An.end(), cd(0.0));
      for (int p = trp[i]; p < trp[i + 1]; p++) An[tci[p]] = tv[p];
      for (int sp = 0; sp < 2; sp++) {
        const cd* al = sp == 0 ? ta : tb;
        // C = T(alpha) * An :  C[a][k] = sum_i alpha[i][a] An[i][k]  (rotate.hpp:116,122)
        for (int a = 0; a < N; a++) {
          std::fill(Cr.begin(), Cr.end(), 0.0);
          std::fill(Ci.begin(), Ci.end(), 0.0);
          for (int ii = 0; ii < M; ii++) {
            const cd c = al[(size_t)ii * N + a];
            if (c == cd(0.0)) continue;
            const double cr = c.real(), ci = c.imag();
            const cd* row = An.data() + (size_t)ii * M;
            for (int k = 0; k < M; k++) {
              const double xr = row[k].real(), xi = row[k].imag();
              Cr[k] += cr * xr - ci * xi;
              Ci[k] += cr * xi + ci * xr;
            }
          }
          for (int k = 0; k < M; k++) C[(size_t)a * M + k] = cd(Cr[k], Ci[k]);
        }
        for (int a = 0; a < N; a++)
          for (int k = 0; k < M; k++)
            if (std::abs(C[(size_t)a * M + k]) > cutoff) {
              rc[i].push_back(sp * N * M + a * M + k);
              rv[i].push_back(C[(size_t)a * M + k]);
            }
      }
    }
  }
  return h;
}

static long halfrot_nnz(const HalfRotation* h)
{
  long nz = 0;
  for (int i = 0; i < h->K; i++) nz += (long)h->rc[i].size();
  return nz;
}

static void halfrot_fill(const HalfRotation* h, int* out_rp, int* out_col, double* out_val)
{
  long p = 0;
  for (int i = 0; i < h->K; i++) {
    out_rp[i] = (int)p;
    for (size_t j = 0; j < h->rc[i].size(); j++, p++) {
      out_col[p] = h->rc[i][j];
      out_val[2 * p] = h->rv[i][j].real();
      out_val[2 * p + 1] = h->rv[i][j].imag();
    }
  }
  out_rp[h->K] = (int)p;
}

long orc_halfrotate_cholesky(int M, int N, int K, const double* ta, const double* tb, const int* rp, const int* col,
                             const double* val, double cutoff, int* out_rp, int* out_col, double* out_val)
{
  HalfRotation* h = halfrotate(M, N, K, ta, tb, rp, col, val, cutoff);
  const long nz = halfrot_nnz(h);
  if (out_rp) halfrot_fill(h, out_rp, out_col, out_val);
  delete h;
  return nz;
}

void* orc_halfrotate_begin(int M, int N, int K, const double* ta, const double* tb, const int* rp, const int* col,
                           const double* val, double cutoff, long* nnz_out)
{
  HalfRotation* h = halfrotate(M, N, K, ta, tb, rp, col, val, cutoff);
  *nnz_out = halfrot_nnz(h);
  return h;
}

void orc_halfrotate_take(void* handle, int* out_rp, int* out_col, double* out_val)
{
  auto* h = static_cast<HalfRotation*>(handle);
  if (out_rp) halfrot_fill(h, out_rp, out_col, out_val);
  delete h;
}

/* SparseMatrix::compress (SparseMatrix.hpp:295-319): sort by (row, col),
 * then rowIndex[i] = first entry of row i (rows past the last entry point at
 * nnz). */
int orc_compress(int nrows, long nnz, int* rows, int* cols, double* vals, int* rowptr)
{
  std::vector<long> idx(nnz);
  for (long i = 0; i < nnz; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](long a, long b) {
    return rows[a] < rows[b] || (!(rows[b] < rows[a]) && cols[a] < cols[b]);
  });
  std::vector<int> r2(nnz), c2(nnz);
  std::vector<double> v2(2 * nnz);
  for (long i = 0; i < nnz; i++) {
    r2[i] = rows[idx[i]];
    c2[i] = cols[idx[i]];
    v2[2 * i] = vals[2 * idx[i]];
    v2[2 * i + 1] = vals[2 * idx[i] + 1];
  }
  std::copy(r2.begin(), r2.end(), rows);
  std::copy(c2.begin(), c2.end(), cols);
  std::copy(v2.begin(), v2.end(), vals);
  int curr = -1;
  for (long n = 0; n < nnz; n++) {
    if (rows[n] != curr) {
      int old = curr;
      curr = rows[n];
      for (int i = old + 1; i <= curr; i++) rowptr[i] = (int)n;
    }
  }
  int last = nnz > 0 ? rows[nnz - 1] : -1;
  for (int i = last + 1; i <= nrows; i++) rowptr[i] = (int)nnz;
  return 0;
}

/* Driver loop, miniafqmc_base.cpp:213-349 (transposed_Spvn = true path).
 * phase_seconds (11 entries, may be NULL) follows the reference's timer list
 * (miniafqmc_base.cpp:50-77): 0 Total, 1 DMc, 2 DM, 3 vbias, 4 vHS, 5 X,
 * 6 Propg, 7 extra, 8 ovlp, 9 ortho, 10 eloc. */
int orc_run(int M, int N, int K, int nw, int nsteps, int nsubsteps, int northo, double dt, const double* ta,
            const double* tb, const double* Propg, const double* haj, const int* sp_rp, const int* sp_col,
            const double* sp_val, const int* spt_rp, const int* spt_col, const double* spt_val, int energy_mode,
            const int* v_rp, const int* v_col, const double* v_val, const double* Lt, const double* x_all,
            double* W_, double* Wd_, double* eav_out, double* eshift_out, double* phase_seconds)
{
  Sys s(M, N, CCZ(ta), CCZ(tb));
  cd* W = CZ(W_);
  cd* Wd = CZ(Wd_);
  int NAK = 2 * N * M;
  std::vector<cd> vbias((size_t)K * nw), vHS((size_t)M * M * nw), Gc((size_t)NAK * nw), X((size_t)K * nw),
      hybridW(nw);
  // no SpvnT: the driver's non-transposed branch (miniafqmc_base.cpp:267-276),
  // full G [2M*M][nw] and vbias = T(Spvn) G_alpha + T(Spvn) G_beta
  const bool transposed = spt_rp != nullptr;
  std::vector<cd> Gfull(transposed ? 0 : (size_t)2 * M * M * nw);
  double tsec[11] = {0};
  bool nonfinite = false;
  auto energy = [&]() {
    if (energy_mode == 2) return 0.0;  // timing samples: no energy evaluation
    if (energy_mode == 0) return sys_energy_vakbl(M, N, nw, Wd, Gc.data(), CCZ(haj), NAK, v_rp, v_col, CCZ(v_val));
    return sys_energy_chol(M, N, K, nw, Wd, Gc.data(), CCZ(haj), CCZ(Lt));
  };
  // walkers start at conj(trialwfn) = Psi, weights 1 (miniafqmc_base.cpp:225-236)
  for (int n = 0; n < nw; n++) {
    for (int i = 0; i < M * N; i++) {
      W[((size_t)n * 2 + 0) * M * N + i] = std::conj(s.ta[i]);
      W[((size_t)n * 2 + 1) * M * N + i] = std::conj(s.tb[i]);
    }
    for (int j = 0; j < 8; j++) Wd[(size_t)n * 8 + j] = 0.0;
    Wd[(size_t)n * 8 + 1] = 1.0;
  }
  int info = sys_mdm(s, nw, W, Wd, Gc.data(), true);
  eav_out[0] = energy();
  double Eshift = 0.0;
  auto t_total0 = clk::now();
  for (int step = 0, step_tot = 0; step < nsteps; step++) {
    for (int sub = 0; sub < nsubsteps; sub++, step_tot++) {
      auto t0 = clk::now();
      int r = 0;
      if (transposed) {
        r = sys_mdm(s, nw, W, Wd, Gc.data(), true);
      } else {
        r = sys_mdm(s, nw, W, Wd, Gfull.data(), false);
      }
      if (r && !info) info = r;
      auto t1 = clk::now();
      if (transposed) {
        csrmm('N', K, nw, NAK, 1.0, CCZ(spt_val), spt_col, spt_rp, spt_rp + 1, Gc.data(), nw, 0.0, vbias.data(), nw);
      } else {
        const int MM = M * M;
        csrmm('T', MM, nw, K, 1.0, CCZ(sp_val), sp_col, sp_rp, sp_rp + 1, Gfull.data(), nw, 0.0, vbias.data(), nw);
        csrmm('T', MM, nw, K, 1.0, CCZ(sp_val), sp_col, sp_rp, sp_rp + 1, Gfull.data() + (size_t)MM * nw, nw, 1.0,
              vbias.data(), nw);
      }
      auto t2 = clk::now();
      const double* x = x_all + (size_t)step_tot * K * nw;
      for (size_t i = 0; i < (size_t)K * nw; i++) {
        if (!std::isfinite(x[i])) nonfinite = true;
        X[i] = cd(x[i], 0.0);
      }
      fields(K, nw, vbias.data(), X.data(), hybridW.data());
      auto t3 = clk::now();
      csrmm('N', M * M, nw, K, 1.0, CCZ(sp_val), sp_col, sp_rp, sp_rp + 1, X.data(), nw, 0.0, vHS.data(), nw);
      auto t4 = clk::now();
      sys_propagate(s, nw, CCZ(Propg), vHS.data(), W);
      auto t5 = clk::now();
      for (int w = 0; w < nw; w++) {
        Wd[(size_t)w * 8 + 5] = Wd[(size_t)w * 8 + 4];
        Wd[(size_t)w * 8 + 6] = Wd[(size_t)w * 8 + 2];
        Wd[(size_t)w * 8 + 7] = Wd[(size_t)w * 8 + 3];
      }
      auto t6 = clk::now();
      r = sys_overlaps(s, nw, W, Wd);
      if (r && !info) info = r;
      auto t7 = clk::now();
      Eshift = weights(nw, dt, hybridW.data(), Wd, Eshift);
      auto t8 = clk::now();
      tsec[1] += secs(t0, t1);
      tsec[3] += secs(t1, t2);
      tsec[5] += secs(t2, t3);
      tsec[4] += secs(t3, t4);
      tsec[6] += secs(t4, t5);
      tsec[7] += secs(t5, t6) + secs(t7, t8);
      tsec[8] += secs(t6, t7);
      if (step_tot > 0 && step_tot % northo == 0) {
        auto o0 = clk::now();
        sys_orthogonalize(M, N, nw, W);
        auto o1 = clk::now();
        r = sys_overlaps(s, nw, W, Wd);
        if (r && !info) info = r;
        auto o2 = clk::now();
        tsec[9] += secs(o0, o1);
        tsec[8] += secs(o1, o2);
      }
    }
    auto e0 = clk::now();
    int r = sys_mdm(s, nw, W, Wd, Gc.data(), true);
    if (r && !info) info = r;
    eav_out[step + 1] = energy();
    tsec[10] += secs(e0, clk::now());
  }
  tsec[0] = secs(t_total0, clk::now());
  if (eshift_out) *eshift_out = Eshift;
  if (phase_seconds)
    for (int i = 0; i < 11; i++) phase_seconds[i] = tsec[i];
  // the reference checks neither (LAPACK info is only asserted, NaN draws
  // pass through); the oracle reports both so that status codes compare
  if (info) return info;
  return nonfinite ? ORC_NONFINITE : 0;
}

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : omp_get_num_procs()); }
int orc_get_threads() { return omp_get_max_threads(); }

} // extern "C"
