// Synthetic AFQMC systems and auxiliary fields (host side).
//
// The reference reads its Hamiltonian from an HDF5 file whose only example
// (example/N32_M64/afqmc.h5) is not shipped, so workloads are generated from a
// seed.  The generator produces exactly the in-memory state that
// afqmc::Initialize returns (src/Matrix/initialize_serial.hpp:35-228):
// trialwfn = conj(Psi) per spin, haj compacted to [2N][M], Propg1 [M][M],
// Spvn CSR [M*M][K] already multiplied by sqrt(dt), and optionally the
// compacted two-body matrix Vakbl CSR [2NM][2NM].
//
// Every random number is a pure function of (seed, stream, index) so any
// piece can be regenerated independently and in parallel.
#include "afqmc_hip.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

using cd = std::complex<double>;

namespace {

inline uint64_t mix64(uint64_t z)
{
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
// uniform in (0,1]: 53 random bits
inline double u01(uint64_t seed, uint64_t stream, uint64_t idx)
{
  uint64_t h = mix64(mix64(seed * 0x2545f4914f6cdd1dull + stream) ^ (idx * 0xd1342543de82ef95ull));
  return ((h >> 11) + 1) * (1.0 / 9007199254740992.0);
}
// standard normal pair via Box-Muller
inline void normal2(uint64_t seed, uint64_t stream, uint64_t idx, double& a, double& b)
{
  double u1 = u01(seed, stream, 2 * idx), u2 = u01(seed, stream, 2 * idx + 1);
  double r = std::sqrt(-2.0 * std::log(u1));
  a = r * std::cos(6.283185307179586 * u2);
  b = r * std::sin(6.283185307179586 * u2);
}

template <class F> void parallel_for(long n, F&& f)
{
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096 || nt == 1) {
    for (long i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> th;
  long chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    long b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=, &f]() {
      for (long i = b; i < e; i++) f(i);
    });
  }
  for (auto& x : th) x.join();
}

// stream tags; 5 is the auxiliary fields' (fieldrng.hpp)
enum Stream : uint64_t { S_H = 1, S_PSI_B = 3, S_L = 4 };

// N orthonormal columns (modified Gram-Schmidt, two passes) of a Gaussian M x N
void random_orbitals(int M, int N, uint64_t seed, uint64_t stream, cd* Psi)
{
  for (long i = 0; i < (long)M * N; i++) {
    double a, b;
    normal2(seed, stream, i, a, b);
    Psi[i] = cd(a, b);
  }
  for (int pass = 0; pass < 2; pass++)
    for (int a = 0; a < N; a++) {
      for (int b = 0; b < a; b++) {
        cd d = 0.0;
        for (int i = 0; i < M; i++) d += std::conj(Psi[(size_t)i * N + b]) * Psi[(size_t)i * N + a];
        for (int i = 0; i < M; i++) Psi[(size_t)i * N + a] -= d * Psi[(size_t)i * N + b];
      }
      double nr = 0.0;
      for (int i = 0; i < M; i++) nr += std::norm(Psi[(size_t)i * N + a]);
      nr = 1.0 / std::sqrt(nr);
      for (int i = 0; i < M; i++) Psi[(size_t)i * N + a] *= nr;
    }
}

} // namespace

struct afqmc_synth {
  afqmc_synth_params p;
  std::vector<double> trial_a, trial_b, haj, propg1, h1;
  std::vector<int> sp_rowptr, sp_col;
  std::vector<double> sp_val;
  std::vector<int> v_rowptr, v_col;
  std::vector<double> v_val;
};

extern "C" {

int afqmc_synth_create(const afqmc_synth_params* prm, afqmc_synth** out)
{
  if (!prm || !out) return AFQMC_ERR_ARG;
  const int M = prm->nmo, N = prm->naea, K = prm->nchol;
  if (M <= 0 || N <= 0 || N > M || K <= 0 || prm->density <= 0.0 || prm->density > 1.0 || prm->dt <= 0.0)
    return AFQMC_ERR_SHAPE;
  if ((long)M * M * (long)K > 2147483647L && prm->density * M * M * (double)K > 2.1e9) return AFQMC_ERR_SHAPE;
  afqmc_synth* s = new (std::nothrow) afqmc_synth;
  if (!s) return AFQMC_ERR_ALLOC;
  s->p = *prm;
  const uint64_t seed = prm->seed;
  const size_t MN = (size_t)M * N, MM = (size_t)M * M;

  // one-body h = U diag(lambda) U^H with U a random unitary (Gram-Schmidt of
  // a complex Gaussian) and lambda ascending at the quantiles of the
  // semicircle law of radius sigma_h*sqrt(2M) (the spectrum of (A + A^H)/2
  // for A complex Gaussian with entries of std sigma_h).  The trial orbitals
  // are the first N columns of U, i.e. the N lowest eigenvectors of h, so the
  // walk projects from a determinant close to the ground state; with a
  // random determinant the phaseless walk loses every walker's weight
  // within ~100 substeps.
  std::vector<cd> U(MM);
  random_orbitals(M, M, seed, S_H, U.data());
  std::vector<double> lam(M);
  {
    const double R = prm->sigma_h * std::sqrt(2.0 * M);
    for (int i = 0; i < M; i++) {
      // F(x) = 1/2 + (x sqrt(1-x^2) + asin x)/pi = (i + 1/2)/M by bisection
      const double target = (i + 0.5) / M;
      double lo = -1.0, hi = 1.0;
      for (int it = 0; it < 100; it++) {
        const double x = 0.5 * (lo + hi);
        const double F = 0.5 + (x * std::sqrt(1.0 - x * x) + std::asin(x)) / M_PI;
        if (F < target) lo = x;
        else hi = x;
      }
      lam[i] = R * 0.5 * (lo + hi);
    }
  }
  std::vector<cd> h(MM);
  parallel_for(M, [&](long i) {
    for (int j = 0; j < M; j++) {
      cd acc = 0.0;
      for (int k = 0; k < M; k++) acc += U[(size_t)i * M + k] * lam[k] * std::conj(U[(size_t)j * M + k]);
      h[(size_t)i * M + j] = acc;
    }
  });
  for (int i = 0; i < M; i++) h[(size_t)i * M + i] = cd(h[(size_t)i * M + i].real(), 0.0);
  s->h1.resize(2 * MM);
  std::memcpy(s->h1.data(), h.data(), MM * sizeof(cd));

  // trial orbitals: Psi_a = U[:, :N]; Psi_b = Psi_a V with V a random N x N
  // unitary (the same determinant, different orbitals); trialwfn = conj(Psi)
  std::vector<cd> pa(MN), pb(MN), V((size_t)N * N);
  random_orbitals(N, N, seed, S_PSI_B, V.data());
  for (int i = 0; i < M; i++)
    for (int a = 0; a < N; a++) {
      pa[(size_t)i * N + a] = U[(size_t)i * M + a];
      cd acc = 0.0;
      for (int b = 0; b < N; b++) acc += U[(size_t)i * M + b] * V[(size_t)b * N + a];
      pb[(size_t)i * N + a] = acc;
    }
  s->trial_a.resize(2 * MN);
  s->trial_b.resize(2 * MN);
  cd* ta = reinterpret_cast<cd*>(s->trial_a.data());
  cd* tb = reinterpret_cast<cd*>(s->trial_b.data());
  for (size_t i = 0; i < MN; i++) {
    ta[i] = std::conj(pa[i]);
    tb[i] = std::conj(pb[i]);
  }

  // haj[a][j] = sum_i trialwfn(i,a) h(i,j)  (= (Psi^H h)(a,j)), alpha rows then beta rows
  s->haj.assign(2 * 2 * MN, 0.0);
  cd* haj = reinterpret_cast<cd*>(s->haj.data());
  for (int sp = 0; sp < 2; sp++) {
    const cd* t = sp == 0 ? ta : tb;
    for (int a = 0; a < N; a++)
      for (int i = 0; i < M; i++) {
        cd ti = t[(size_t)i * N + a];
        for (int j = 0; j < M; j++) haj[((size_t)sp * N + a) * M + j] += ti * h[(size_t)i * M + j];
      }
  }

  // Propg1 = expm(-dt/2 h) = U diag(exp(-dt/2 lambda)) U^H
  {
    s->propg1.resize(2 * MM);
    cd* P = reinterpret_cast<cd*>(s->propg1.data());
    std::vector<double> e(M);
    for (int k = 0; k < M; k++) e[k] = std::exp(-0.5 * prm->dt * lam[k]);
    parallel_for(M, [&](long i) {
      for (int j = 0; j < M; j++) {
        cd acc = 0.0;
        for (int k = 0; k < M; k++) acc += U[(size_t)i * M + k] * e[k] * std::conj(U[(size_t)j * M + k]);
        P[(size_t)i * M + j] = acc;
      }
    });
  }

  // Cholesky vectors L(ik, n): re, im ~ N(0, sigma^2), kept if |L| > tau with
  // tau = sigma*sqrt(-2 ln density) (Rayleigh tail), then * sqrt(dt).
  const double sig = prm->sigma_L;
  const double tau = prm->density >= 1.0 ? -1.0 : sig * std::sqrt(-2.0 * std::log(prm->density));
  const double sdt = std::sqrt(prm->dt);
  const long nrows = (long)MM;
  std::vector<int> rowcount(nrows + 1, 0);
  // Each vector is Hermitian, L(k,i) = conj(L(i,k)) (as the vectors of a
  // factorised real-space Hamiltonian are): entry (i,k) and (k,i) share one
  // draw, the diagonal keeps its real part.
  auto draw = [&](long r, int n, double& a, double& b) {
    long i = r / M, k = r % M;
    long canon = (i <= k) ? r : k * M + i;
    normal2(seed, S_L, (uint64_t)canon * K + n, a, b);
    if (i > k) b = -b;
    if (i == k) b = 0.0;
  };
  parallel_for(nrows, [&](long r) {
    int c = 0;
    for (int n = 0; n < K; n++) {
      double a, b;
      long i = r / M, k = r % M;
      normal2(seed, S_L, (uint64_t)((i <= k) ? r : k * M + i) * K + n, a, b);
      if (std::hypot(a, b) * sig > tau) c++;
    }
    rowcount[r + 1] = c;
  });
  s->sp_rowptr.assign(nrows + 1, 0);
  for (long r = 0; r < nrows; r++) {
    long next = (long)s->sp_rowptr[r] + rowcount[r + 1];
    if (next > 2147483647L) {
      delete s;
      return AFQMC_ERR_SHAPE;
    }
    s->sp_rowptr[r + 1] = (int)next;
  }
  long nnz = s->sp_rowptr[nrows];
  s->sp_col.resize(nnz);
  s->sp_val.resize(2 * (size_t)nnz);
  parallel_for(nrows, [&](long r) {
    long p = s->sp_rowptr[r];
    for (int n = 0; n < K; n++) {
      double a, b;
      long i = r / M, k = r % M;
      normal2(seed, S_L, (uint64_t)((i <= k) ? r : k * M + i) * K + n, a, b);
      if (std::hypot(a, b) * sig > tau) {
        draw(r, n, a, b);
        s->sp_col[p] = n;
        s->sp_val[2 * p] = a * sig * sdt;
        s->sp_val[2 * p + 1] = b * sig * sdt;
        p++;
      }
    }
  });

  // Vakbl(sak, s'bl) = sum_n Lt(n,sak) Lt(n,s'bl) - delta_ss' sum_n Lt(n,sal) Lt(n,sbk)
  // with the unscaled half-rotated Lt(n, s a k) = sum_i trialwfn_s(i,a) L(ik,n).
  if (prm->with_vakbl) {
    const int NAK = 2 * N * M;
    if ((double)NAK * NAK > 3.0e8) {
      delete s;
      return AFQMC_ERR_SHAPE;
    }
    std::vector<cd> Lt((size_t)K * NAK, cd(0.0));
    const cd* val = reinterpret_cast<const cd*>(s->sp_val.data());
    for (int sp = 0; sp < 2; sp++) {
      const cd* t = sp == 0 ? ta : tb;
      parallel_for(K, [&](long n) {
        for (long r = 0; r < nrows; r++) {
          // row r = i*M + k; find column n in row r (sorted)
          const int* b = s->sp_col.data() + s->sp_rowptr[r];
          const int* e = s->sp_col.data() + s->sp_rowptr[r + 1];
          const int* f = std::lower_bound(b, e, (int)n);
          if (f == e || *f != n) continue;
          cd v = val[f - s->sp_col.data()] / sdt;
          int i = (int)(r / M), k = (int)(r % M);
          for (int a = 0; a < N; a++) Lt[(size_t)n * NAK + (size_t)sp * N * M + (size_t)a * M + k] += t[(size_t)i * N + a] * v;
        }
      });
    }
    std::vector<cd> V((size_t)NAK * NAK);
    parallel_for(NAK, [&](long p) {
      const int sp = (int)(p / (N * M)), a = (int)((p % (N * M)) / M), k = (int)(p % M);
      cd* Vp = V.data() + (size_t)p * NAK;
      for (int q = 0; q < NAK; q++) Vp[q] = 0.0;
      for (int n = 0; n < K; n++) {
        const cd* Ln = Lt.data() + (size_t)n * NAK;
        cd lp = Ln[p];
        for (int q = 0; q < NAK; q++) Vp[q] += lp * Ln[q];
        // exchange within the same spin block: q = (sp, b, l): - Lt(n, s a l) Lt(n, s b k)
        const cd* Ls = Ln + (size_t)sp * N * M;
        for (int b = 0; b < N; b++) {
          cd lbk = Ls[(size_t)b * M + k];
          cd* Vq = Vp + (size_t)sp * N * M + (size_t)b * M;
          for (int l = 0; l < M; l++) Vq[l] -= Ls[(size_t)a * M + l] * lbk;
        }
      }
    });
    s->v_rowptr.assign(NAK + 1, 0);
    for (int p = 0; p < NAK; p++) {
      int c = 0;
      for (int q = 0; q < NAK; q++)
        if (V[(size_t)p * NAK + q] != cd(0.0)) c++;
      s->v_rowptr[p + 1] = s->v_rowptr[p] + c;
    }
    s->v_col.resize(s->v_rowptr[NAK]);
    s->v_val.resize(2 * (size_t)s->v_rowptr[NAK]);
    for (int p = 0; p < NAK; p++) {
      int c = s->v_rowptr[p];
      for (int q = 0; q < NAK; q++) {
        cd v = V[(size_t)p * NAK + q];
        if (v == cd(0.0)) continue;
        s->v_col[c] = q;
        s->v_val[2 * (size_t)c] = v.real();
        s->v_val[2 * (size_t)c + 1] = v.imag();
        c++;
      }
    }
  }
  *out = s;
  return AFQMC_OK;
}

int afqmc_synth_destroy(afqmc_synth* s)
{
  delete s;
  return AFQMC_OK;
}

int afqmc_synth_get(const afqmc_synth* s, afqmc_synth_view* v)
{
  if (!s || !v) return AFQMC_ERR_ARG;
  v->nmo = s->p.nmo;
  v->naea = s->p.naea;
  v->nchol = s->p.nchol;
  v->trial_alpha = s->trial_a.data();
  v->trial_beta = s->trial_b.data();
  v->haj = s->haj.data();
  v->propg1 = s->propg1.data();
  v->h1 = s->h1.data();
  v->spvn.nrows = s->p.nmo * s->p.nmo;
  v->spvn.ncols = s->p.nchol;
  v->spvn.nnz = s->sp_rowptr.empty() ? 0 : s->sp_rowptr.back();
  v->spvn.rowptr = s->sp_rowptr.data();
  v->spvn.col = s->sp_col.data();
  v->spvn.val = s->sp_val.data();
  if (s->v_rowptr.empty()) {
    v->vakbl.nrows = v->vakbl.ncols = v->vakbl.nnz = 0;
    v->vakbl.rowptr = nullptr;
    v->vakbl.col = nullptr;
    v->vakbl.val = nullptr;
  } else {
    v->vakbl.nrows = v->vakbl.ncols = 2 * s->p.naea * s->p.nmo;
    v->vakbl.nnz = s->v_rowptr.back();
    v->vakbl.rowptr = s->v_rowptr.data();
    v->vakbl.col = s->v_col.data();
    v->vakbl.val = s->v_val.data();
  }
  return AFQMC_OK;
}

} // extern "C"
