// Context, device memory plan and phase orchestration behind afqmc_hip.h.
//
// Device layouts (HBM):
//   AFQMC_FP64_CSR: the reference's own layouts (walker index last for
//     G [2NM][W], vbias/X [K][W], vHS [M*M][W]) so every csrmm reads rows of
//     walkers contiguously; vHS is transposed once per substep to [W][M*M]
//     for the per-walker propagation.
//   dense modes: walker index first everywhere (G [W][2][N][M],
//     vbias/X [W][K], vHS [W][M][M]) so each contraction is a GEMM whose
//     operands are contiguous along the summed index; the Cholesky vectors
//     are kept as Ld [K][M*M] (vector-major) and the half-rotated
//     Lt [2][K][N*M] (spin-major).
#include "afqmc_hip.h"
#include "common.hpp"
#include "gemm.hpp"
#include "gemm_b16.hpp"
#include "kernels.hpp"
#include "propagate.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace {

struct KStat {
  long long n = 0;
  double ms = 0.0;
};

struct EvPair {
  std::string name;
  hipEvent_t a, b;
};

} // namespace

struct afqmc_ctx {
  int M = 0, N = 0, K = 0, nw = 0, mode = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool have_system = false;
  bool mixed = false;   // FP32 or split-BF16 contractions
  bool bsplit = false;  // split-BF16: G, X, vHS, Taylor terms, Ld, Lt, Vx, Propg1 stored pre-split (zb)
  double dt_sys = 0.0;  // dt the system was scaled with
  std::vector<void*> allocs;

  // system
  zd* trial = nullptr;   // [2][M][N] conj(Psi)
  zd* propg = nullptr;   // [M][M]
  void* propg32 = nullptr; // [M][M] in the contraction format (mixed: zf, or zb when split)
  zd* haj = nullptr;     // [2N][M]
  int* sp_rp = nullptr;
  int* sp_col = nullptr;
  zd* sp_val = nullptr;
  long long sp_nnz = 0;
  bool sp_sorted = false;  // Spvn columns non-decreasing within every row
  int* spt_rp = nullptr;
  int* spt_col = nullptr;
  zd* spt_val = nullptr;
  long long spt_nnz = 0;
  // FP64_CSR: SpvnT as a dense row-major [K][2NM] copy when at least half
  // of it is stored (AFQMC_OPT_CSR_DENSE), for the bias product on MFMA
  zd* sptd = nullptr;
  bool csr_dense = true;
  bool coulomb3 = false;  // AFQMC_OPT_COULOMB_THREE_PIECE
  // non-transposed bias (miniafqmc_base.cpp:267-276): T(Spvn) as CSR [K][M*M]
  // and the full density matrix G [2*M*M][nw]
  bool transposed = true;
  int* sptt_rp = nullptr;
  int* sptt_col = nullptr;
  zd* sptt_val = nullptr;
  zd* Gfull = nullptr;
  int* v_rp = nullptr;
  int* v_col = nullptr;
  zd* v_val = nullptr;
  long long v_nnz = 0;
  void* Ld = nullptr;  // dense modes: [K][M*M]  (zd or zf)
  void* Lt = nullptr;  // [2][K][N*M] (zd or zf); scaled by sqrt(dt)
  zd* Lt64 = nullptr;  // mixed modes with energy_level >= 1: Lt in FP64
  // mixed modes: 0 energy in the contraction precision; 1 FP64 density
  // matrix, one-body and Coulomb terms; 2 and exchange in FP32; 3 and
  // exchange in FP64
  int energy_level = 1;

  // walker state and per-substep buffers
  zd* W = nullptr;
  zd* Wd = nullptr;
  zd* T1 = nullptr;       // [nw][2][N][N] overlap matrices, then their inverses
  zd* dcache = nullptr;   // [nw][2] determinants of the last inversion
  // log-overlaps [nw][4]: log det of the current overlap matrices (spin 0,
  // 1) and of the saved ones (W_data columns 6, 7); the weights take the
  // overlap ratio from them
  zd* Wl = nullptr;
  zd* lcache = nullptr;   // [nw][2] log determinants of the last inversion
  zd* gjwork = nullptr;   // N > 128: [nw][2][N][N] working matrices of the batched inverse
  bool t1inv_valid = false;  // T1 holds the inverses for the current walkers
  bool g_fresh = false;      // G was built from T1 and W as they are now
  zd* Wbr = nullptr;         // population control: gathered walkers [nw][2][M][N]
  zd* Wdbr = nullptr;        // and W_data rows [nw][8]
  zd* Wlbr = nullptr;        // and log-overlaps [nw][4]
  double* brcum = nullptr;   // cumulative weights [nw]
  int* brparent = nullptr;   // parent of each new walker [nw]
  double* brw = nullptr;     // common weight after the comb
  void* G = nullptr;      // CSR: zd [2NM][nw]; FP64_DENSE: zd [nw][2NM]; mixed: zf [nw][2NM]
  zd* vbias = nullptr;    // CSR: [K][nw]; dense: [nw][K]
  void* X = nullptr;      // CSR: zd [K][nw]; dense: [nw][K] (zd / zf)
  void* vHS = nullptr;    // CSR: zd [M*M][nw]; dense: [nw][M*M] (zd / zf)
  zd* vHSw = nullptr;     // CSR: [nw][M*M]
  zd* hybridW = nullptr;
  void* S = nullptr;      // propagation scratch [nw][2][M][N] x 3: S = P W and the Horner terms
  void* Ta = nullptr;
  void* Tb = nullptr;
  double* eshift = nullptr;
  int* flags = nullptr;   // [0] singular count, [1] non-finite fields
  double* xstage = nullptr;
  long nstage = 0;
  // fields drawn on the device from (field_seed, substep, walker) instead of
  // a staged block (afqmc_device_fields)
  bool dev_fields = false;
  uint64_t field_seed = 0;
  // energy scratch
  zd* e_direct = nullptr;
  zd* e_exch = nullptr;
  zd* e_exch_part = nullptr;  // [nw][2]
  zd* e_one = nullptr;
  zd* e_v = nullptr;      // [nw][K] direct-term contractions
  void* Ybuf = nullptr;
  size_t ybuf_bytes = 0;
  // same-spin exchange block of Vakbl, dense [2][NM][NM] (scaled by dt), in
  // the contraction precision; null when it does not fit in device memory
  void* Vx = nullptr;
  // split-BF16 mode, energy level 2: Vx as three-piece split-BF16 (the zb
  // (h, m) pieces in Vx, the third pieces here, [2][NM][NM]) and the
  // walkers' G split likewise for the exchange product ([nw][2NM] each)
  uint32_t* Vx3 = nullptr;
  zb* G3 = nullptr;
  uint32_t* G3l = nullptr;
  // three-piece split-BF16 contractions (AFQMC_OPT_THREE_PIECE, split-BF16
  // mode): bit 1 the bias product, bit 2 the HS potential, bit 4 the
  // propagator; the third pieces of their operands, beside the split arrays
  int three = 0;
  uint32_t* tp_Lt = nullptr;   // [2][K][NM] (bias)
  uint32_t* tp_G = nullptr;    // [nw][2NM] (bias)
  uint32_t* tp_Ld = nullptr;   // [M*M][K] (HS potential)
  uint32_t* tp_X = nullptr;    // [nw][K] (HS potential)
  uint32_t* tp_P = nullptr;    // [M][M] (propagator)
  uint32_t* tp_vHS = nullptr;  // [nw][M*M] (propagator)
  bool vhs3 = false;           // tp_vHS holds the third pieces of the current vHS
  uint32_t* tp_Ta = nullptr;   // [nw][2][N][M] (propagator terms)
  uint32_t* tp_Tb = nullptr;
  // with bit 4 of `three`: which of the propagator's eight products take
  // three pieces (bit 0 P W, bit 7-k the Horner step k = 6..1, bit 7 P U)
  // and which Horner steps take one (the hi words of the same split
  // operands); the others take two pieces
  int three_steps = 0xFF;
  int one_steps = 0;
  bool allow_vx = true;
  zd* Gtmp = nullptr;     // FP64 G [nw][2NM] for the energy (CSR: transposed; mixed: FP64 copy)
  void* vws = nullptr;    // split-k partials of the bias contraction
  size_t vws_bytes = 0;
  // orthogonalisation: Cholesky QR scratch (allocated at the first call)
  bool ortho_householder = false;  // always use the blocked Householder kernel
  bool fused_prop = false;         // split-BF16 propagator in one kernel when the shapes allow
  bool gemm_adma = true;           // split-BF16 GEMMs stage A by direct global -> LDS loads
  bool gemm_ring = true;           // ... and both operands through a ring of LDS stages where they allow
  long graph_gen = 0;              // bumped whenever captured substeps become stale
  long exch_chunk = 2048;          // dense exchange: columns per accumulation chunk (0: one pass)
  zd* oU = nullptr;   // [2nw][N][N] R^{-1}
  zd* oX = nullptr;   // [2][2nw][M][N] the two Cholesky QR passes
  int* ofail = nullptr;  // [2nw] matrices left to the Householder kernel
  zd* Gcloc = nullptr;    // CSR mode Vakbl energy

  // profiling
  bool prof = false;
  std::vector<EvPair> pending;
  std::vector<hipEvent_t> evpool;
  std::map<std::string, KStat> stats;

  // GEMM tile autotuning: call-site key -> candidate index
  std::map<std::string, int> tune;
  bool tuning = false;
  int force_cfg = 0;  // 1-based tile candidate forced on every GEMM (AFQMC_OPT_TILE, afqmc_zgemm)

  // graphs
  bool use_graphs = false;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  long g_substep = -1;
  double g_dt = 0.0;
  double* d_dt = nullptr;
  long* d_sub = nullptr;  // substep index read by the captured fields kernel
  hipGraph_t graph[2] = {nullptr, nullptr};
};

namespace {

#define HIPCHK(ctx, call)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                        \
      return AFQMC_ERR_HIP;                                                                   \
    }                                                                                         \
  } while (0)

#define RET(call)                   \
  do {                              \
    int r_ = (call);                \
    if (r_ != AFQMC_OK) return r_;  \
  } while (0)

int fail(afqmc_ctx* c, int code, const std::string& m)
{
  c->err = m;
  return code;
}

template <class T> int dalloc(afqmc_ctx* c, T** p, size_t n)
{
  void* q = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&q, n * sizeof(T));
  if (e != hipSuccess) return fail(c, AFQMC_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->allocs.push_back(q);
  *p = (T*)q;
  return AFQMC_OK;
}
int dalloc_bytes(afqmc_ctx* c, void** p, size_t bytes)
{
  char* q = nullptr;
  RET(dalloc(c, &q, bytes));
  *p = q;
  return AFQMC_OK;
}
void dfree(afqmc_ctx* c, void* p)
{
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree(p);
}

hipEvent_t get_event(afqmc_ctx* c)
{
  if (!c->evpool.empty()) {
    hipEvent_t e = c->evpool.back();
    c->evpool.pop_back();
    return e;
  }
  hipEvent_t e;
  (void)hipEventCreate(&e);
  return e;
}

// Launch wrapper: brackets the launch with events when profiling.
// Optional ROCTx ranges named after the call site around every launch (with
// AFQMC_ROCTX=1 in the environment), so that `rocprofv3 --kernel-rename`
// reports the walker-step phases rather than template instantiations.
struct Roctx {
  int (*push)(const char*) = nullptr;
  int (*pop)() = nullptr;
};
const Roctx& roctx()
{
  static const Roctx r = []() {
    Roctx x;
    const char* e = std::getenv("AFQMC_ROCTX");
    if (!e || !*e || *e == '0') return x;
    void* h = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!h) return x;
    x.push = (int (*)(const char*))dlsym(h, "roctxRangePushA");
    x.pop = (int (*)())dlsym(h, "roctxRangePop");
    if (!x.push || !x.pop) x.push = nullptr, x.pop = nullptr;
    return x;
  }();
  return r;
}

template <class F> int launch(afqmc_ctx* c, const char* name, F&& f)
{
  const Roctx& rx = roctx();
  if (rx.push) rx.push(c->tuning ? (std::string("tune/") + name).c_str() : name);
  struct Pop {
    const Roctx& r;
    ~Pop()
    {
      if (r.pop) r.pop();
    }
  } pop_{rx};
  hipEvent_t a = nullptr, b = nullptr;
  if (c->prof) {
    a = get_event(c);
    b = get_event(c);
    (void)hipEventRecord(a, c->stream);
  }
  f();
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, AFQMC_ERR_HIP, std::string(name) + ": " + hipGetErrorString(e));
  if (c->prof) {
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({name, a, b});
  }
  return AFQMC_OK;
}

void collect_events(afqmc_ctx* c)
{
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& p : c->pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    auto& s = c->stats[p.name];
    s.n++;
    s.ms += ms;
    c->evpool.push_back(p.a);
    c->evpool.push_back(p.b);
  }
  c->pending.clear();
}

// element pointer advanced by `off` complex numbers (element tags double,
// float or bfs, whose complex types are zd, zf, zb)
template <class T> inline const void* zadv(const void* p, long off)
{
  return (const char*)p + off * (long)sizeof(typename ztype<T>::type);
}

inline int nblk(long n, int t) { return (int)std::min<long>((n + t - 1) / t, 65535L * 8); }

// ---------------------------------------------------------------------------
// GEMM dispatch.  Each call site (name and shape) runs one of a few tile
// shapes; afqmc_tune() times all candidates once per call site and keeps the
// fastest, otherwise a size heuristic picks.
struct TileCfg {
  int BM, BNC, WM, WN;
};
// candidates per precision (BKC is 8 complex for FP64/FP32, 16 for BF16)
constexpr TileCfg kCfgF64[] = {{32, 16, 1, 1}, {64, 32, 2, 2}, {128, 64, 2, 2}};
constexpr TileCfg kCfgF32[] = {{32, 16, 1, 1}, {128, 64, 2, 2}, {128, 128, 2, 2}};
constexpr TileCfg kCfgB16[] = {{32, 16, 1, 1}, {128, 64, 2, 2}, {128, 128, 2, 2}, {256, 64, 4, 2}, {64, 64, 2, 2}};

template <int CP> constexpr int ncfg() { return CP == CP_F64 ? 3 : (CP == CP_F32 ? 3 : 5); }
template <int CP> constexpr TileCfg cfg_at(int i) { return CP == CP_F64 ? kCfgF64[i] : (CP == CP_F32 ? kCfgF32[i] : kCfgB16[i]); }

template <int CP, class AT, class BT, class OT, bool AK, bool BK, int CI>
int zgemm_cfg(afqmc_ctx* c, const char* name, ZGemmArgs a, int batch)
{
  constexpr TileCfg t = cfg_at<CP>(CI);
  constexpr int BKC = (CP == CP_BF16X3) ? 16 : (t.WM * t.WN == 1 ? 16 : 8);
  a.MT = (a.m + t.BM - 1) / t.BM;
  a.NT = (a.n + t.BNC - 1) / t.BNC;
  const long tiles = (long)a.MT * a.NT * batch;
  if (tiles > 2147483647L) return fail(c, AFQMC_ERR_SHAPE, "GEMM grid too large");
  if constexpr (CP == CP_BF16X3 && std::is_same<AT, bfs>::value && AK) {
    // direct global -> LDS staging of A: 16-byte loads of two complex at a
    // time, so every row (and batch item) must start 16-byte aligned, and
    // whole k-tiles (the triangular mode stops at column-block ends, which
    // are multiples of 16 too)
    if (c->gemm_adma && a.sak == 1 && a.k % BKC == 0 && a.sam % 2 == 0 && a.sa1 % 2 == 0 && a.sa2 % 2 == 0 &&
        ((uintptr_t)a.A & 15) == 0 && (!a.ktri || t.BNC % BKC == 0)) {
      if constexpr (std::is_same<BT, bfs>::value && BK) {
        // B too, on the same conditions (conj(B) only through registers)
        if (a.sbk == 1 && a.sbn % 2 == 0 && a.sb1 % 2 == 0 && a.sb2 % 2 == 0 && ((uintptr_t)a.B & 15) == 0 &&
            !a.conjb)
          return launch(c, name, [&]() {
            hipLaunchKernelGGL((zgemm_kernel<CP, AT, BT, OT, AK, BK, t.BM, t.BNC, BKC, t.WM, t.WN, true, true>),
                               dim3((unsigned)tiles), dim3(64 * t.WM * t.WN), 0, c->stream, a);
          });
      }
      return launch(c, name, [&]() {
        hipLaunchKernelGGL((zgemm_kernel<CP, AT, BT, OT, AK, BK, t.BM, t.BNC, BKC, t.WM, t.WN, true>),
                           dim3((unsigned)tiles), dim3(64 * t.WM * t.WN), 0, c->stream, a);
      });
    }
  }
  return launch(c, name, [&]() {
    hipLaunchKernelGGL((zgemm_kernel<CP, AT, BT, OT, AK, BK, t.BM, t.BNC, BKC, t.WM, t.WN>), dim3((unsigned)tiles),
                       dim3(64 * t.WM * t.WN), 0, c->stream, a);
  });
}

template <int CP, class AT, class BT, class OT, bool AK, bool BK>
int zgemm_run(afqmc_ctx* c, const char* name, const ZGemmArgs& a, int batch, int ci)
{
  switch (ci) {
  case 0: return zgemm_cfg<CP, AT, BT, OT, AK, BK, 0>(c, name, a, batch);
  case 1: return zgemm_cfg<CP, AT, BT, OT, AK, BK, 1>(c, name, a, batch);
  case 2: return zgemm_cfg<CP, AT, BT, OT, AK, BK, 2>(c, name, a, batch);
  case 3: if constexpr (ncfg<CP>() > 3) return zgemm_cfg<CP, AT, BT, OT, AK, BK, (ncfg<CP>() > 3 ? 3 : 0)>(c, name, a, batch);
          break;
  case 4: if constexpr (ncfg<CP>() > 4) return zgemm_cfg<CP, AT, BT, OT, AK, BK, (ncfg<CP>() > 4 ? 4 : 0)>(c, name, a, batch);
          break;
  }
  return zgemm_cfg<CP, AT, BT, OT, AK, BK, 0>(c, name, a, batch);
}

template <int CP> int default_cfg(const ZGemmArgs& a)
{
  if (a.m >= 96 && a.n >= 48) return CP == CP_F64 ? 2 : 1;  // 128 x 64 tiles
  return 0;                                                 // one-wave 32 x 16 tiles
}

// Split-BF16 products whose operands are both split-BF16 with k contiguous
// run zgemm_b16dma_kernel (both operands staged by direct loads through a
// ring of LDS stages); candidates (rows, complex columns, waves, stages):
struct RingCfg {
  int BM, BN, WM, WN, S;
};
constexpr RingCfg kRing[] = {{128, 128, 2, 2, 2}, {256, 128, 4, 2, 2}, {128, 256, 2, 4, 2}, {256, 128, 4, 2, 3}};
// three-piece operands take half as much LDS again per stage
constexpr RingCfg kRing3[] = {{128, 128, 2, 2, 2}, {256, 128, 4, 2, 2}, {128, 256, 2, 4, 2}, {128, 128, 2, 2, 3}};
constexpr int kNRing = 4;
template <int NP> constexpr RingCfg ring_at(int i) { return NP == 3 ? kRing3[i] : kRing[i]; }

bool ring_ok(const afqmc_ctx* c, const ZGemmArgs& a)
{
  return c->gemm_adma && c->gemm_ring && a.sak == 1 && a.sbk == 1 && a.k % 16 == 0 && a.sam % 2 == 0 &&
         a.sa1 % 2 == 0 && a.sa2 % 2 == 0 && a.sbn % 2 == 0 && a.sb1 % 2 == 0 && a.sb2 % 2 == 0 &&
         ((uintptr_t)a.A & 15) == 0 && ((uintptr_t)a.B & 15) == 0 && !a.conjb;
}

template <class OT, int RI, int NP> int ring_cfg(afqmc_ctx* c, const char* name, ZGemmArgs a, int batch)
{
  constexpr RingCfg t = ring_at<NP>(RI);
  a.MT = (a.m + t.BM - 1) / t.BM;
  a.NT = (a.n + t.BN - 1) / t.BN;
  const long tiles = (long)a.MT * a.NT * batch;
  if (tiles > 2147483647L) return fail(c, AFQMC_ERR_SHAPE, "GEMM grid too large");
  return launch(c, name, [&]() {
    hipLaunchKernelGGL((zgemm_b16dma_kernel<OT, t.BM, t.BN, t.WM, t.WN, t.S, NP>), dim3((unsigned)tiles),
                       dim3(64 * t.WM * t.WN), 0, c->stream, a);
  });
}

template <class OT, int NP> int ring_run(afqmc_ctx* c, const char* name, const ZGemmArgs& a, int batch, int ri)
{
  switch (ri) {
  case 1: return ring_cfg<OT, 1, NP>(c, name, a, batch);
  case 2: return ring_cfg<OT, 2, NP>(c, name, a, batch);
  case 3: return ring_cfg<OT, 3, NP>(c, name, a, batch);
  default: return ring_cfg<OT, 0, NP>(c, name, a, batch);
  }
}

int ring_default(const ZGemmArgs& a)
{
  if (a.m >= 256 && a.n >= 128) return 1;
  if (a.m >= 128 && a.n >= 256) return 2;
  return 0;
}

template <class OT, int NP = 2> int zgemm_ring(afqmc_ctx* c, const char* name, const ZGemmArgs& a, int batch)
{
  char keyb[160];
  std::snprintf(keyb, sizeof(keyb), "%s/ring%d/%d/%d/%d/%d", name, NP, a.m, a.n, a.k, batch);
  const std::string key(keyb);
  if (c->force_cfg > 0 && c->force_cfg <= kNRing) return ring_run<OT, NP>(c, name, a, batch, c->force_cfg - 1);
  auto it = c->tune.find(key);
  if (it != c->tune.end() && !c->tuning) return ring_run<OT, NP>(c, name, a, batch, it->second);
  if (!c->tuning) return ring_run<OT, NP>(c, name, a, batch, ring_default(a));
  const bool prof = c->prof;
  c->prof = false;
  hipEvent_t e0 = get_event(c), e1 = get_event(c);
  float tms[kNRing];
  for (int ri = 0; ri < kNRing; ri++) tms[ri] = 1e30f;
  for (int round = 0; round < 2; round++)
    for (int ri = 0; ri < kNRing; ri++) {
      RET((ring_run<OT, NP>(c, name, a, batch, ri)));  // warm
      (void)hipEventRecord(e0, c->stream);
      for (int rep = 0; rep < 3; rep++) RET((ring_run<OT, NP>(c, name, a, batch, ri)));
      (void)hipEventRecord(e1, c->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      tms[ri] = std::min(tms[ri], ms);
    }
  int best = ring_default(a);
  for (int ri = 0; ri < kNRing; ri++)
    if (tms[ri] < tms[best]) best = ri;
  c->evpool.push_back(e0);
  c->evpool.push_back(e1);
  c->prof = prof;
  c->tune[key] = best;
  return ring_run<OT, NP>(c, name, a, batch, best);
}

// one-piece products of split operands (CP_BF16X1, the hi words only): the
// staged-ring kernel only, whatever AFQMC_OPT_DIRECT_LDS says
bool ring1_ok(const ZGemmArgs& a)
{
  return a.sak == 1 && a.sbk == 1 && a.k % 16 == 0 && a.sam % 2 == 0 && a.sa1 % 2 == 0 && a.sa2 % 2 == 0 &&
         a.sbn % 2 == 0 && a.sb1 % 2 == 0 && a.sb2 % 2 == 0 && ((uintptr_t)a.A & 15) == 0 &&
         ((uintptr_t)a.B & 15) == 0 && !a.conjb;
}

// three-piece split-BF16 operands (CP_BF16X6): the staged-ring kernel only
bool ring3_ok(const ZGemmArgs& a)
{
  return a.A3 && a.B3 && a.sak == 1 && a.sbk == 1 && a.k % 16 == 0 && a.sam % 4 == 0 && a.sa1 % 4 == 0 &&
         a.sa2 % 4 == 0 && a.sbn % 4 == 0 && a.sb1 % 4 == 0 && a.sb2 % 4 == 0 && ((uintptr_t)a.A & 15) == 0 &&
         ((uintptr_t)a.B & 15) == 0 && ((uintptr_t)a.A3 & 15) == 0 && ((uintptr_t)a.B3 & 15) == 0 && !a.conjb;
}

template <int CP, class AT, class BT, class OT, bool AK, bool BK>
int zgemm(afqmc_ctx* c, const char* name, ZGemmArgs a, int batch)
{
  if (a.m <= 0 || a.n <= 0 || batch <= 0) return AFQMC_OK;
  if (a.bdiv <= 0) a.bdiv = 1;
  if constexpr (CP == CP_BF16X6) {
    static_assert(std::is_same<AT, bfs>::value && std::is_same<BT, bfs>::value && AK && BK,
                  "three-piece products take split operands with k contiguous");
    if (!ring3_ok(a)) return fail(c, AFQMC_ERR_UNSUPPORTED, std::string(name) + ": three-piece operand layout");
    return zgemm_ring<OT, 3>(c, name, a, batch);
  } else if constexpr (CP == CP_BF16X1) {
    static_assert(std::is_same<AT, bfs>::value && std::is_same<BT, bfs>::value && AK && BK,
                  "one-piece products take split operands with k contiguous");
    if (!ring1_ok(a)) return fail(c, AFQMC_ERR_UNSUPPORTED, std::string(name) + ": one-piece operand layout");
    return zgemm_ring<OT, 1>(c, name, a, batch);
  } else {
  if constexpr (CP == CP_BF16X3 && std::is_same<AT, bfs>::value && std::is_same<BT, bfs>::value && AK && BK) {
    if (ring_ok(c, a)) return zgemm_ring<OT>(c, name, a, batch);
  }
  char keyb[160];
  std::snprintf(keyb, sizeof(keyb), "%s/%d/%d/%d/%d/%d/%d%d", name, CP, a.m, a.n, a.k, batch, (int)AK, (int)BK);
  const std::string key(keyb);
  if (c->force_cfg > 0 && c->force_cfg <= ncfg<CP>())
    return zgemm_run<CP, AT, BT, OT, AK, BK>(c, name, a, batch, c->force_cfg - 1);
  auto it = c->tune.find(key);
  if (it != c->tune.end() && !c->tuning) return zgemm_run<CP, AT, BT, OT, AK, BK>(c, name, a, batch, it->second);
  if (!c->tuning) return zgemm_run<CP, AT, BT, OT, AK, BK>(c, name, a, batch, default_cfg<CP>(a));
  // tuning: time each candidate (results of this call are discarded), in
  // two rounds keeping each candidate's faster round
  const bool prof = c->prof;
  c->prof = false;
  hipEvent_t e0 = get_event(c), e1 = get_event(c);
  const int def = default_cfg<CP>(a);
  float tms[8];
  for (int ci = 0; ci < ncfg<CP>(); ci++) tms[ci] = 1e30f;
  for (int round = 0; round < 2; round++)
    for (int ci = 0; ci < ncfg<CP>(); ci++) {
      RET((zgemm_run<CP, AT, BT, OT, AK, BK>(c, name, a, batch, ci)));  // warm
      (void)hipEventRecord(e0, c->stream);
      for (int rep = 0; rep < 3; rep++) RET((zgemm_run<CP, AT, BT, OT, AK, BK>(c, name, a, batch, ci)));
      (void)hipEventRecord(e1, c->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      tms[ci] = std::min(tms[ci], ms);
    }
  int best = def;
  for (int ci = 0; ci < ncfg<CP>(); ci++)
    if (tms[ci] < tms[best]) best = ci;
  c->evpool.push_back(e0);
  c->evpool.push_back(e1);
  c->prof = prof;
  c->tune[key] = best;
  return zgemm_run<CP, AT, BT, OT, AK, BK>(c, name, a, batch, best);
  }
}

ZGemmArgs gargs()
{
  ZGemmArgs a;
  std::memset(&a, 0, sizeof(a));
  a.bdiv = 1;
  a.ar = 1.0;
  return a;
}

// C = alpha * A * B + beta * C for CSR A [nrows][K] and B [K][nw] (walker
// index last).  Rows averaging 512 entries or more: with columns known to be
// sorted within rows, 4 to 16 rows share each LDS-staged window of B;
// otherwise a workgroup of 4 or 16 waves shares each row.  Shorter rows take
// one wave per (row, 64 walkers), which measured faster than staging.
int csrmm_n(afqmc_ctx* c, const char* name, int nrows, int K, long long nnz, const int* rp, const int* col,
            const zd* val, const zd* B, long ldb, zd* C, long ldc, zd alpha, zd beta, bool sorted = false)
{
  const int nw = c->nw;
  const long nchunk = (nw + 63) / 64;
  const double avg = nrows > 0 ? (double)nnz / nrows : 0.0;
  if (sorted && avg >= 512) {
    // rows per workgroup: the most B reuse that still gives 256 workgroups
    if (((nrows + 15) / 16) * nchunk >= 256)
      return launch(c, name, [&]() {
        hipLaunchKernelGGL((csrmm_n_lds_kernel<16, 64>), dim3((unsigned)(((nrows + 15) / 16) * nchunk)), dim3(1024),
                           0, c->stream, nrows, nw, K, rp, col, val, B, ldb, C, ldc, alpha, beta);
      });
    if (((nrows + 7) / 8) * nchunk >= 256)
      return launch(c, name, [&]() {
        hipLaunchKernelGGL((csrmm_n_lds_kernel<8, 64>), dim3((unsigned)(((nrows + 7) / 8) * nchunk)), dim3(512), 0,
                           c->stream, nrows, nw, K, rp, col, val, B, ldb, C, ldc, alpha, beta);
      });
    return launch(c, name, [&]() {
      hipLaunchKernelGGL((csrmm_n_lds_kernel<4, 64>), dim3((unsigned)(((nrows + 3) / 4) * nchunk)), dim3(256), 0,
                         c->stream, nrows, nw, K, rp, col, val, B, ldb, C, ldc, alpha, beta);
    });
  }
  if (avg >= 4096)
    return launch(c, name, [&]() {
      hipLaunchKernelGGL(csrmm_n_split_kernel<16>, dim3((unsigned)(nrows * nchunk)), dim3(1024), 0, c->stream, nrows,
                         nw, K, rp, col, val, B, ldb, C, ldc, alpha, beta);
    });
  if (avg >= 512)
    return launch(c, name, [&]() {
      hipLaunchKernelGGL(csrmm_n_split_kernel<4>, dim3((unsigned)(nrows * nchunk)), dim3(256), 0, c->stream, nrows,
                         nw, K, rp, col, val, B, ldb, C, ldc, alpha, beta);
    });
  const long waves = (long)nrows * nchunk;
  return launch(c, name, [&]() {
    hipLaunchKernelGGL(csrmm_n_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, c->stream, nrows, nw, K, rp,
                       col, val, B, ldb, C, ldc, alpha, beta);
  });
}

// Batched inverse + determinant of the nw*2 overlap matrices T (in place
// when Tinv == T); det of matrix b at det[(b/2)*dq + (b%2)*dr], a count of
// singular matrices in *flags and the determinants in dc when those are set.
int gj_inverse_into(afqmc_ctx* c, zd* T, zd* Tinv, zd* det, long dq, long dr, int* flags, zd* dc, zd* ldet,
                    zd* lc)
{
  const int N = c->N;
  const long batch = 2L * c->nw;
  const long ms = (long)N * N;
  if (N <= 16)
    return launch(c, "gj_inverse", [&]() {
      hipLaunchKernelGGL((gj_inverse_kernel<16, 8>), dim3(batch), dim3(128), 0, c->stream, T, Tinv, det, dq, dr, N, ms,
                         flags, dc, ldet, lc);
    });
  if (N <= 32)
    return launch(c, "gj_inverse", [&]() {
      hipLaunchKernelGGL((gj_inverse_kernel<32, 8>), dim3(batch), dim3(256), 0, c->stream, T, Tinv, det, dq, dr, N, ms,
                         flags, dc, ldet, lc);
    });
  if (N <= 64)
    return launch(c, "gj_inverse", [&]() {
      hipLaunchKernelGGL((gj_inverse_kernel<64, 8>), dim3(batch), dim3(512), 0, c->stream, T, Tinv, det, dq, dr, N, ms,
                         flags, dc, ldet, lc);
    });
  if (N <= 128)
    return launch(c, "gj_inverse", [&]() {
      hipLaunchKernelGGL((gj_inverse_kernel<128, 8>), dim3(batch), dim3(1024), 0, c->stream, T, Tinv, det, dq, dr, N,
                         ms, flags, dc, ldet, lc);
    });
  if (N <= GJ_GMAX) {
    // the working matrices do not fit one workgroup's registers: they stay
    // in global memory (allocated at the first call)
    if (!c->gjwork) RET(dalloc(c, &c->gjwork, (size_t)batch * ms));
    return launch(c, "gj_inverse", [&]() {
      hipLaunchKernelGGL(gj_inverse_global_kernel, dim3(batch), dim3(1024), 0, c->stream, T, Tinv, c->gjwork, det, dq,
                         dr, N, ms, flags, dc, ldet, lc);
    });
  }
  return fail(c, AFQMC_ERR_UNSUPPORTED, "N > 256 electrons per spin is not supported by the batched inverse");
}

// ... of the context's overlap matrices, with its flags and determinant cache
int gj_inverse(afqmc_ctx* c, zd* T, zd* Tinv, zd* det, long dq, long dr)
{
  return gj_inverse_into(c, T, Tinv, det, dq, dr, c->flags, c->dcache, c->Wl, c->lcache);
}

// T1[w][s] = W[w][s]^T * trial_s  (N x N, k = M), all walkers and spins
int overlap_matrices(afqmc_ctx* c)
{
  const int M = c->M, N = c->N;
  ZGemmArgs a = gargs();
  a.A = c->W;
  a.sam = 1;
  a.sak = N;
  a.sa1 = 2L * M * N;
  a.sa2 = (long)M * N;
  a.B = c->trial;
  a.sbn = 1;
  a.sbk = N;
  a.sb1 = 0;
  a.sb2 = (long)M * N;
  a.C = c->T1;
  a.scm = N;
  a.scn = 1;
  a.sc1 = 2L * N * N;
  a.sc2 = (long)N * N;
  a.bdiv = 2;
  a.m = N;
  a.n = N;
  a.k = M;
  return zgemm<CP_F64, double, double, double, false, false>(c, "ovlp_gemm", a, 2 * c->nw);
}

// Gc(a,k) = sum_b T1inv(a,b) W(k,b) for every walker-spin pair; output
// strides are set by the caller.
ZGemmArgs gc_args(const afqmc_ctx* c)
{
  const int M = c->M, N = c->N;
  ZGemmArgs a = gargs();
  a.A = c->T1;
  a.sam = N;
  a.sak = 1;
  a.sa1 = 2L * N * N;
  a.sa2 = (long)N * N;
  a.B = c->W;
  a.sbn = N;
  a.sbk = 1;
  a.sb1 = 2L * M * N;
  a.sb2 = (long)M * N;
  a.bdiv = 2;
  a.m = N;
  a.n = M;
  a.k = N;
  return a;
}

// FP64 copy of G in [nw][2NM] for the energy of the mixed modes: rebuilt
// from the inverses when G is current for the walkers, otherwise widened
// from the stored G (a G uploaded by the caller, or a stale one).
int g_fp64(afqmc_ctx* c)
{
  const long NAK = 2L * c->N * c->M, nw = c->nw;
  if (c->g_fresh && c->t1inv_valid) {
    ZGemmArgs a = gc_args(c);
    a.C = c->Gtmp;
    a.scm = c->M;
    a.scn = 1;
    a.sc1 = NAK;
    a.sc2 = NAK / 2;
    return zgemm<CP_F64, double, double, double, true, true>(c, "gc64_gemm", a, 2 * c->nw);
  }
  return launch(c, "convert", [&]() {
    if (c->bsplit)
      hipLaunchKernelGGL((zconvert_kernel<zb, zd>), dim3(nblk(NAK * nw, 256)), dim3(256), 0, c->stream,
                         (const zb*)c->G, c->Gtmp, NAK * nw);
    else
      hipLaunchKernelGGL((zconvert_kernel<zf, zd>), dim3(nblk(NAK * nw, 256)), dim3(256), 0, c->stream,
                         (const zf*)c->G, c->Gtmp, NAK * nw);
  });
}

// ---------------------------------------------------------------------------
int do_mdm(afqmc_ctx* c)
{
  const int M = c->M, N = c->N, nw = c->nw;
  if (c->t1inv_valid) {
    // the overlap step already inverted W^T conj(Psi) for these walkers
    RET(launch(c, "det_restore", [&]() {
      hipLaunchKernelGGL(det_restore_kernel, dim3((2 * nw + 255) / 256), dim3(256), 0, c->stream, c->Wd, c->dcache,
                         c->Wl, c->lcache, nw);
    }));
  } else {
    RET(overlap_matrices(c));
    RET(gj_inverse(c, c->T1, c->T1, c->Wd + 2, 8, 1));
    c->t1inv_valid = true;
  }
  c->g_fresh = true;
  ZGemmArgs a = gc_args(c);
  a.C = c->G;
  if (c->mode == AFQMC_FP64_CSR) {
    // G[(s*N + a)*M + k][w]
    a.scm = (long)M * nw;
    a.scn = nw;
    a.sc1 = 1;
    a.sc2 = (long)N * M * nw;
    return zgemm<CP_F64, double, double, double, true, true>(c, "gc_gemm", a, 2 * nw);
  }
  a.scm = M;
  a.scn = 1;
  a.sc1 = 2L * N * M;
  a.sc2 = (long)N * M;
  // mixed modes: G only feeds the force bias, which contracts it in the
  // contraction precision, so it is formed in that precision as well (the
  // energy rebuilds an FP64 G from the FP64 inverses)
  if (c->bsplit && (c->three & 1)) {
    // three-piece bias product: G from an FP32 product, split in three
    a.C3 = c->tp_G;
    return zgemm<CP_F32, double, double, bfs, true, true>(c, "gc_gemm", a, 2 * nw);
  }
  if (c->bsplit) return zgemm<CP_BF16X3, double, double, bfs, true, true>(c, "gc_gemm", a, 2 * nw);
  if (c->mixed) return zgemm<CP_F32, double, double, float, true, true>(c, "gc_gemm", a, 2 * nw);
  return zgemm<CP_F64, double, double, double, true, true>(c, "gc_gemm", a, 2 * nw);
}

// afqmc_sys::calculate_overlaps; the inverse is kept for the next density
// matrix of the same walkers.
int do_overlaps(afqmc_ctx* c)
{
  c->g_fresh = false;
  RET(overlap_matrices(c));
  RET(gj_inverse(c, c->T1, c->T1, c->Wd + 2, 8, 1));
  c->t1inv_valid = true;
  return AFQMC_OK;
}

// vbias(w, n) = sum_{s,j} G(w, s, j) Lt(s, n, j) in dense modes.  With few
// output tiles (walkers x vectors) the long j sum is split into S chunks:
// the 2S partial products (spin, chunk) run as one batched GEMM into a
// workspace and are then summed in FP64.
template <int CP, class GT, class LT_>
int dense_contract_vbias(afqmc_ctx* c, const void* Gp, const void* Lt, zd* out, const char* name,
                         const uint32_t* G3 = nullptr, const uint32_t* Lt3 = nullptr,
                         const char* rname = "vbias_reduce")
{
  const long NM = (long)c->N * c->M, nw = c->nw, K = c->K;
  const long tiles = ((nw + 127) / 128) * ((K + 63) / 64);
  int S = 1;
  while (S < 16 && 2 * S * tiles < 2048 && NM % (2 * S) == 0 && (NM / (2 * S)) % 64 == 0) S *= 2;
  if (S == 1) {
    for (int sp = 0; sp < 2; sp++) {
      ZGemmArgs a = gargs();
      a.A = zadv<GT>(Gp, sp * NM);
      a.sam = 2 * NM;
      a.sak = 1;
      a.B = zadv<LT_>(Lt, sp * K * NM);
      a.sbn = NM;
      a.sbk = 1;
      a.C = out;
      a.scm = K;
      a.scn = 1;
      a.m = (int)nw;
      a.n = (int)K;
      a.k = (int)NM;
      a.br = sp == 0 ? 0.0 : 1.0;
      if (G3) {
        a.A3 = G3 + sp * NM;
        a.B3 = Lt3 + sp * K * NM;
      }
      RET((zgemm<CP, GT, LT_, double, true, true>(c, name, a, 1)));
    }
    return AFQMC_OK;
  }
  // partials in FP32 for the FP32/BF16 contractions, FP64 otherwise
  using PT = typename std::conditional<CP == CP_F64, double, float>::type;
  using PZ = typename ztype<PT>::type;
  const size_t need = sizeof(PZ) * 2 * S * nw * K;
  if (c->vws_bytes < need) {
    dfree(c, c->vws);
    c->vws_bytes = need;
    RET(dalloc_bytes(c, &c->vws, need));
  }
  const long chunk = NM / S;
  ZGemmArgs a = gargs();
  a.A = Gp;  // batch b = (spin, chunk): A offset spin*NM + chunk*NM/S
  a.sam = 2 * NM;
  a.sak = 1;
  a.sa1 = NM;
  a.sa2 = chunk;
  a.B = Lt;
  a.sbn = NM;
  a.sbk = 1;
  a.sb1 = K * NM;
  a.sb2 = chunk;
  a.C = c->vws;
  a.scm = K;
  a.scn = 1;
  a.sc1 = S * nw * K;
  a.sc2 = nw * K;
  a.bdiv = S;
  a.m = (int)nw;
  a.n = (int)K;
  a.k = (int)chunk;
  a.A3 = G3;  // third pieces with the same offsets
  a.B3 = Lt3;
  RET((zgemm<CP, GT, LT_, PT, true, true>(c, name, a, 2 * S)));
  return launch(c, rname, [&]() {
    hipLaunchKernelGGL((zsum_slices_kernel<PZ>), dim3(nblk(nw * K, 256)), dim3(256), 0, c->stream, (const PZ*)c->vws, out,
                       nw * K, 2 * S);
  });
}

// afqmc_sys::calculate_mixed_density_matrix(W, W_data, G, false)
// (mixed_density_matrix.hpp:93-101): G_s = conj(A_s) Gc_s per walker and
// spin, [2*M*M][nw] walker-last.  FP64_CSR mode.
int do_mdm_full(afqmc_ctx* c)
{
  const int M = c->M, N = c->N, nw = c->nw;
  const long MM = (long)M * M, MN = (long)M * N;
  if (!c->Gfull) RET(dalloc(c, &c->Gfull, 2 * MM * nw));
  RET(do_mdm(c));
  ZGemmArgs a = gargs();
  a.A = c->trial;  // A(i, a) = trialwfn_s[i][a]
  a.sam = N;
  a.sak = 1;
  a.sa1 = 0;
  a.sa2 = MN;
  a.B = c->G;  // B(k, a) = Gc[(s N + a) M + k][w]
  a.sbn = nw;
  a.sbk = (long)M * nw;
  a.sb1 = 1;
  a.sb2 = MN * nw;
  a.C = c->Gfull;  // C(i, k) = G[(s M + i) M + k][w]
  a.scm = (long)M * nw;
  a.scn = nw;
  a.sc1 = 1;
  a.sc2 = MM * nw;
  a.bdiv = 2;
  a.m = M;
  a.n = M;
  a.k = N;
  return zgemm<CP_F64, double, double, double, true, false>(c, "gfull_gemm", a, 2 * nw);
}

// base::get_vbias(Spvn, G, vbias, false) (vbias.hpp:65-82):
// vbias = T(Spvn) G_alpha + T(Spvn) G_beta, as two csrmm of the transposed
// CSR in the reference's order.
int do_vbias_full(afqmc_ctx* c)
{
  const int nw = c->nw, K = c->K;
  const long MM = (long)c->M * c->M;
  zd one{1, 0}, zero{0, 0};
  for (int sp = 0; sp < 2; sp++)
    RET(csrmm_n(c, "csrmm_vbias_t", K, (int)MM, c->sp_nnz, c->sptt_rp, c->sptt_col, c->sptt_val,
                (const zd*)c->Gfull + sp * MM * nw, (long)nw, c->vbias, (long)nw, one, sp ? one : zero, true));
  return AFQMC_OK;
}

// FP64_CSR bias product with SpvnT kept dense: vbias[K][nw] = SpvnT[K][2NM]
// G[2NM][nw] (walker-last) on FP64 MFMA, the long sum split into S chunks
// (a batched GEMM into a workspace, then summed in FP64 in chunk order).
int dense_csr_vbias(afqmc_ctx* c)
{
  const long NAK = 2L * c->N * c->M, nw = c->nw, K = c->K;
  const long tiles = ((K + 63) / 64) * ((nw + 31) / 32);
  int S = 1;
  while (S < 32 && S * tiles < 1024 && NAK % (2 * S) == 0 && (NAK / (2 * S)) % 64 == 0) S *= 2;
  const long chunk = NAK / S;
  ZGemmArgs a = gargs();
  a.A = c->sptd;
  a.sam = NAK;
  a.sak = 1;
  a.B = c->G;
  a.sbn = 1;
  a.sbk = nw;
  a.m = (int)K;
  a.n = (int)nw;
  a.k = (int)chunk;
  a.scm = nw;
  a.scn = 1;
  if (S == 1) {
    a.C = c->vbias;
    return zgemm<CP_F64, double, double, double, true, false>(c, "vbias_gemm", a, 1);
  }
  const size_t need = sizeof(zd) * S * K * nw;
  if (c->vws_bytes < need) {
    dfree(c, c->vws);
    c->vws_bytes = need;
    RET(dalloc_bytes(c, &c->vws, need));
  }
  a.sa1 = chunk;  // batch b: columns b*chunk .. of SpvnT, rows b*chunk .. of G
  a.sb1 = chunk * nw;
  a.C = c->vws;
  a.sc1 = K * nw;
  RET((zgemm<CP_F64, double, double, double, true, false>(c, "vbias_gemm", a, S)));
  return launch(c, "vbias_reduce", [&]() {
    hipLaunchKernelGGL((zsum_slices_kernel<zd>), dim3(nblk(nw * K, 256)), dim3(256), 0, c->stream, (const zd*)c->vws,
                       c->vbias, nw * K, S);
  });
}

int do_vbias(afqmc_ctx* c)
{
  const int nw = c->nw, K = c->K, NAK = 2 * c->N * c->M;
  if (c->mode == AFQMC_FP64_CSR) {
    if (c->sptd) return dense_csr_vbias(c);
    zd one{1, 0}, zero{0, 0};
    return csrmm_n(c, "csrmm_vbias", K, NAK, c->spt_nnz, c->spt_rp, c->spt_col, c->spt_val, (const zd*)c->G,
                   (long)nw, c->vbias, (long)nw, one, zero, true);
  }
  if (c->mode == AFQMC_FP64_DENSE)
    return dense_contract_vbias<CP_F64, double, double>(c, c->G, c->Lt, c->vbias, "vbias_gemm");
  if (c->mode == AFQMC_MIXED_FP32)
    return dense_contract_vbias<CP_F32, float, float>(c, c->G, c->Lt, c->vbias, "vbias_gemm");
  if (c->three & 1)
    return dense_contract_vbias<CP_BF16X6, bfs, bfs>(c, c->G, c->Lt, c->vbias, "vbias_gemm", c->tp_G, c->tp_Lt);
  return dense_contract_vbias<CP_BF16X3, bfs, bfs>(c, c->G, c->Lt, c->vbias, "vbias_gemm");
}

// Draws of the fields phase: xdev (a device block [K][nw]; with xslice !=
// null the staged block and *xslice its substep index) or, when xdev is null,
// generated on the device from (field_seed, substep or *xslice, walker).
int do_fields(afqmc_ctx* c, const double* xdev, long substep, const long* xslice = nullptr)
{
  const int nw = c->nw, K = c->K;
  FieldSrc src{xdev, xslice, substep, c->field_seed};
  if (c->mode == AFQMC_FP64_CSR)
    return launch(c, "fields", [&]() {
      hipLaunchKernelGGL((fields_wlast_kernel<zd>), dim3((nw + 63) / 64), dim3(1024), 0, c->stream, K, nw, c->vbias,
                         src, (zd*)c->X, c->hybridW, c->flags + 1);
    });
  return launch(c, "fields", [&]() {
    if (!c->mixed)
      hipLaunchKernelGGL((fields_wfirst_kernel<zd>), dim3(nw), dim3(256), 0, c->stream, K, nw, c->vbias, src,
                         (zd*)c->X, c->hybridW, c->flags + 1, (uint32_t*)nullptr);
    else if (c->bsplit)
      hipLaunchKernelGGL((fields_wfirst_kernel<zb>), dim3(nw), dim3(256), 0, c->stream, K, nw, c->vbias, src,
                         (zb*)c->X, c->hybridW, c->flags + 1, c->tp_X);
    else
      hipLaunchKernelGGL((fields_wfirst_kernel<zf>), dim3(nw), dim3(256), 0, c->stream, K, nw, c->vbias, src,
                         (zf*)c->X, c->hybridW, c->flags + 1, (uint32_t*)nullptr);
  });
}

// The fields of substep s are available: drawn on the device, or staged.
bool have_fields(const afqmc_ctx* c, long s)
{
  return s >= 0 && (c->dev_fields || (c->xstage && s < c->nstage));
}
// the staged draws of substep s, or null when they are drawn on the device
const double* stage_at(const afqmc_ctx* c, long s)
{
  return c->dev_fields ? nullptr : c->xstage + s * (long)c->K * c->nw;
}

int do_vhs(afqmc_ctx* c)
{
  const int nw = c->nw, K = c->K;
  const long MM = (long)c->M * c->M;
  if (c->mode == AFQMC_FP64_CSR) {
    zd one{1, 0}, zero{0, 0};
    RET(csrmm_n(c, "csrmm_vhs", (int)MM, K, c->sp_nnz, c->sp_rp, c->sp_col, c->sp_val, (const zd*)c->X, (long)nw,
                (zd*)c->vHS, (long)nw, one, zero, c->sp_sorted));
    dim3 grid((nw + 31) / 32, (unsigned)((MM + 31) / 32));
    return launch(c, "vhs_transpose", [&]() {
      hipLaunchKernelGGL((ztranspose_kernel<zd, zd>), grid, dim3(256), 0, c->stream, (const zd*)c->vHS, c->vHSw,
                         (int)MM, nw, (long)nw, MM);
    });
  }
  // vHS(w, ik) = sum_n X(w, n) Ld(n, ik)
  ZGemmArgs a = gargs();
  a.A = c->X;
  a.sam = K;
  a.sak = 1;
  a.B = c->Ld;
  a.sbn = 1;
  a.sbk = MM;
  a.C = c->vHS;
  a.scm = MM;
  a.scn = 1;
  a.m = nw;
  a.n = (int)MM;
  a.k = K;
  if (c->mode == AFQMC_FP64_DENSE) return zgemm<CP_F64, double, double, double, true, false>(c, "vhs_gemm", a, 1);
  if (c->mode == AFQMC_MIXED_FP32) return zgemm<CP_F32, float, float, float, true, false>(c, "vhs_gemm", a, 1);
  // split-BF16: Ld stored [M*M][K]
  a.sbn = K;
  a.sbk = 1;
  // vHS's third pieces, only when a Horner step of the three-piece
  // propagator reads them (bits 1-6 of the step mask)
  c->vhs3 = c->tp_vHS && (c->three_steps & 0x7E);
  a.C3 = c->vhs3 ? c->tp_vHS : nullptr;
  if (c->three & 2) {
    a.A3 = c->tp_X;
    a.B3 = c->tp_Ld;
    return zgemm<CP_BF16X6, bfs, bfs, bfs, true, true>(c, "vhs_gemm", a, 1);
  }
  return zgemm<CP_BF16X3, bfs, bfs, bfs, true, true>(c, "vhs_gemm", a, 1);
}

// W <- P * T6(V) * P * W per walker and spin (afqmc_sys.hpp:146-173), T6 the
// order-6 Taylor polynomial of exp(i V) applied as in apply_expM
// (vHS.hpp:61-91): T_n = (i/n) V T_{n-1}, S += T_n.
//
// The Horner terms (Ta, Tb) hold each (walker, spin) M x N block
// transposed, column j of the block contiguous ([N][M]): the Taylor steps
// and P U_1 then read them as a right operand with k contiguous, which the
// split-BF16 path stages straight into LDS.  S stays row-major (it is read
// as the first step's right operand, in FP32, and as the epilogues' F).
template <int CP, class ST, class TT, class PT, class VT>
int propagate_chain(afqmc_ctx* c, const void* P, const void* V)
{
  const int M = c->M, N = c->N, nw = c->nw;
  const long MN = (long)M * N, MM = (long)M * M;
  auto base = [&]() {
    ZGemmArgs a = gargs();
    a.sam = M;
    a.sak = 1;
    a.sbn = 1;
    a.sbk = N;
    a.sb1 = 2 * MN;
    a.sb2 = MN;
    a.scm = N;
    a.scn = 1;
    a.sc1 = 2 * MN;
    a.sc2 = MN;
    a.bdiv = 2;
    a.m = M;
    a.n = N;
    a.k = M;
    return a;
  };
  // S = P * W
  {
    ZGemmArgs a = base();
    a.A = P;
    a.B = c->W;
    a.C = c->S;
    RET((zgemm<CP, PT, double, ST, true, false>(c, "prop_pw", a, 2 * nw)));
  }
  // exp(iV) S to sixth order (vHS.hpp:61-91) by Horner's rule:
  // U_7 = S, U_k = S + (i/k) V U_{k+1}, E6(V) S = U_1; each step is one GEMM
  // whose epilogue adds S.
  const void* prev = c->S;
  void* bufs[2] = {c->Ta, c->Tb};
  for (int k = 6; k >= 1; k--) {
    ZGemmArgs a = base();
    a.A = V;
    a.sa1 = MM;
    a.sa2 = 0;
    a.B = prev;
    if (k < 6) {  // U_{k+1} transposed; U_7 = S row-major
      a.sbn = M;
      a.sbk = 1;
    }
    a.C = bufs[k & 1];
    a.scm = 1;
    a.scn = M;
    a.ar = 0.0;
    a.ai = 1.0 / (double)k;
    a.F = c->S;
    a.sfm = N;
    a.sfn = 1;
    a.cf = 1;
    if (k == 6) RET((zgemm<CP, VT, ST, TT, true, false>(c, "prop_taylor", a, 2 * nw)));
    else RET((zgemm<CP, VT, TT, TT, true, true>(c, "prop_taylor", a, 2 * nw)));
    prev = bufs[k & 1];
  }
  // W = P * E6(V) S
  {
    ZGemmArgs a = base();
    a.A = P;
    a.B = prev;
    a.sbn = M;
    a.sbk = 1;
    a.C = c->W;
    RET((zgemm<CP, PT, TT, double, true, true>(c, "prop_ps", a, 2 * nw)));
  }
  return AFQMC_OK;
}

// The propagator in three-piece split-BF16: W is transposed into three-piece
// terms ([N][M] per walker and spin) and P W is written in that layout as
// well (split, plus FP32 for the epilogues' F), so that every product reads
// both operands k-contiguous through the staged-ring kernel.  Each product
// runs in three, two or one pieces as `three_steps` / `one_steps` say (the
// (h, m) words of a three-piece operand are its two-piece split, h alone its
// one-piece form); a term is written with its third pieces only when the
// product that reads it takes three.
int propagate_three(afqmc_ctx* c)
{
  const int M = c->M, N = c->N, nw = c->nw;
  const long MN = (long)M * N, MM = (long)M * M;
  const dim3 tgrid((unsigned)((N + 31) / 32), (unsigned)((M + 31) / 32), (unsigned)(2 * nw));
  // product j in launch order: 0 P W, 7-k the Horner step k, 7 P U
  auto three = [&](int j) { return ((c->three_steps >> j) & 1) != 0; };
  auto one = [&](int j) { return !three(j) && ((c->one_steps >> j) & 1) != 0; };
  zb* Ta = (zb*)c->Ta;
  zb* Tb = (zb*)c->Tb;
  RET(launch(c, "prop_split_w", [&]() {
    hipLaunchKernelGGL(transpose_split3_kernel<zd>, tgrid, dim3(256), 0, c->stream, (const zd*)c->W, Tb, c->tp_Tb, M, N);
  }));
  auto base = [&]() {
    ZGemmArgs a = gargs();
    a.A = c->propg32;
    a.A3 = c->tp_P;
    a.sam = M;
    a.sak = 1;
    a.sbn = M;
    a.sbk = 1;
    a.sb1 = 2 * MN;
    a.sb2 = MN;
    a.bdiv = 2;
    a.m = M;
    a.n = N;
    a.k = M;
    a.scm = 1;  // terms transposed per walker and spin, [N][M]
    a.scn = M;
    a.sc1 = 2 * MN;
    a.sc2 = MN;
    return a;
  };
  {  // S = P W: U_7 = S split (third pieces when U_6's product takes three)
     // into Ta, and S in FP32 in the same layout (D) for the Horner epilogues
    ZGemmArgs a = base();
    a.B = Tb;
    a.B3 = c->tp_Tb;
    a.C = Ta;
    a.C3 = three(1) ? c->tp_Ta : nullptr;
    a.D = c->S;
    if (three(0)) RET((zgemm<CP_BF16X6, bfs, bfs, bfs, true, true>(c, "prop_pw", a, 2 * nw)));
    else RET((zgemm<CP_BF16X3, bfs, bfs, bfs, true, true>(c, "prop_pw", a, 2 * nw)));
  }
  // Horner's rule as propagate_chain: U_7 = S, U_k = S + (i/k) V U_{k+1}
  zb* prev = Ta;
  uint32_t* prev3 = c->tp_Ta;
  zb* bufs[2] = {Tb, Ta};
  uint32_t* bufs3[2] = {c->tp_Tb, c->tp_Ta};
  for (int k = 6; k >= 1; k--) {
    ZGemmArgs a = base();
    a.A = c->vHS;
    a.A3 = c->tp_vHS;
    a.sa1 = MM;
    a.sa2 = 0;
    a.B = prev;
    a.B3 = prev3;
    a.C = bufs[k & 1];
    a.C3 = three(8 - k) ? bufs3[k & 1] : nullptr;
    a.ar = 0.0;
    a.ai = 1.0 / (double)k;
    a.F = c->S;  // same layout as C
    a.cf = 1;
    if (three(7 - k)) RET((zgemm<CP_BF16X6, bfs, bfs, bfs, true, true>(c, "prop_taylor", a, 2 * nw)));
    else if (one(7 - k)) RET((zgemm<CP_BF16X1, bfs, bfs, bfs, true, true>(c, "prop_taylor", a, 2 * nw)));
    else RET((zgemm<CP_BF16X3, bfs, bfs, bfs, true, true>(c, "prop_taylor", a, 2 * nw)));
    prev = bufs[k & 1];
    prev3 = bufs3[k & 1];
  }
  {  // W = P U_1
    ZGemmArgs a = base();
    a.B = prev;
    a.B3 = prev3;
    a.C = c->W;
    a.scm = N;
    a.scn = 1;
    if (three(7)) RET((zgemm<CP_BF16X6, bfs, bfs, double, true, true>(c, "prop_ps", a, 2 * nw)));
    else RET((zgemm<CP_BF16X3, bfs, bfs, double, true, true>(c, "prop_ps", a, 2 * nw)));
  }
  return AFQMC_OK;
}

// W <- P E6(V) P W in one kernel per (walker, 32 columns) (propagate.hpp):
// split-BF16 mode, M a multiple of 64 up to 512, N a multiple of 16.
bool fused_prop_ok(const afqmc_ctx* c)
{
  return c->fused_prop && c->mode == AFQMC_MIXED_BF16 && c->bsplit && c->M % 64 == 0 && c->M <= PF_MMAX &&
         c->N % 16 == 0;
}

int propagate_fused(afqmc_ctx* c)
{
  PropFusedArgs a;
  a.W = c->W;
  a.S = (float*)c->S;
  a.P = (const zb*)c->propg32;
  a.V = (const zb*)c->vHS;
  a.M = c->M;
  a.N = c->N;
  a.nw = c->nw;
  a.ncb = 2 * c->N / PF_NB;
  const unsigned grid = (unsigned)((long)c->nw * a.ncb);
  // 16 waves of 16 or 32 rows each
  return launch(c, "prop_fused", [&]() {
    if (c->M <= 256)
      hipLaunchKernelGGL((prop_fused_kernel<1, 16>), dim3(grid), dim3(1024), 0, c->stream, a);
    else
      hipLaunchKernelGGL((prop_fused_kernel<2, 16>), dim3(grid), dim3(1024), 0, c->stream, a);
  });
}

int do_propagate(afqmc_ctx* c)
{
  c->t1inv_valid = false;
  c->g_fresh = false;
  if (c->three & 4) return propagate_three(c);
  if (fused_prop_ok(c)) return propagate_fused(c);
  if (c->mode == AFQMC_FP64_CSR) return propagate_chain<CP_F64, double, double, double, double>(c, c->propg, c->vHSw);
  if (c->mode == AFQMC_FP64_DENSE)
    return propagate_chain<CP_F64, double, double, double, double>(c, c->propg, c->vHS);
  if (c->mode == AFQMC_MIXED_FP32)
    return propagate_chain<CP_F32, float, float, float, float>(c, c->propg32, c->vHS);
  return propagate_chain<CP_BF16X3, float, bfs, bfs, bfs>(c, c->propg32, c->vHS);
}

int do_save_overlaps(afqmc_ctx* c)
{
  return launch(c, "save_overlaps", [&]() {
    hipLaunchKernelGGL(save_overlaps_kernel, dim3((c->nw + 255) / 256), dim3(256), 0, c->stream, c->Wd, c->Wl,
                       c->nw);
  });
}

// Comb population control (after a step's energy, where the driver marks
// branching: src/miniapps/miniafqmc_base.cpp:346).  Walker i is copied
// floor or ceil of nw * w_i / W times, in index order, and every new weight
// is W / nw.
int do_branch(afqmc_ctx* c, double u, int* parent_out)
{
  const int nw = c->nw;
  const long per = 2L * c->M * c->N;
  if (!(u >= 0.0 && u < 1.0)) return fail(c, AFQMC_ERR_ARG, "comb offset u must lie in [0, 1)");
  if (!c->Wbr) {
    RET(dalloc(c, &c->Wbr, (size_t)nw * per));
    RET(dalloc(c, &c->Wdbr, (size_t)nw * 8));
    RET(dalloc(c, &c->Wlbr, (size_t)nw * 4));
    RET(dalloc(c, &c->brcum, (size_t)nw));
    RET(dalloc(c, &c->brparent, (size_t)nw));
    RET(dalloc(c, &c->brw, 1));
  }
  HIPCHK(c, hipMemsetAsync(c->flags + 2, 0, sizeof(int), c->stream));
  RET(launch(c, "branch_select", [&]() {
    hipLaunchKernelGGL(comb_select_kernel, dim3(1), dim3(256), 0, c->stream, c->Wd, nw, u, c->brcum, c->brparent,
                       c->brw, c->flags + 2);
  }));
  int bad = 0;
  HIPCHK(c, hipMemcpyAsync(&bad, c->flags + 2, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (bad) return fail(c, AFQMC_ERR_NONFINITE, "population control: total weight not positive or a weight not finite");
  const long n = (long)nw * per;
  const unsigned grid = (unsigned)std::min<long>(16384, std::max<long>(1, (n + 255) / 256));
  RET(launch(c, "branch_gather", [&]() {
    hipLaunchKernelGGL(comb_gather_kernel, dim3(grid), dim3(256), 0, c->stream, (const zd*)c->W, c->Wbr,
                       (const zd*)c->Wd, c->Wdbr, (const zd*)c->Wl, c->Wlbr, (const int*)c->brparent,
                       (const double*)c->brw, per, nw);
  }));
  HIPCHK(c, hipMemcpyAsync(c->W, c->Wbr, sizeof(zd) * n, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->Wd, c->Wdbr, sizeof(zd) * nw * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->Wl, c->Wlbr, sizeof(zd) * nw * 4, hipMemcpyDeviceToDevice, c->stream));
  c->t1inv_valid = false;
  c->g_fresh = false;
  if (parent_out) {
    HIPCHK(c, hipMemcpyAsync(parent_out, c->brparent, sizeof(int) * nw, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return AFQMC_OK;
}

int do_weights(afqmc_ctx* c, const double* dtp, double dt)
{
  (void)dtp;
  int th = 64;
  while (th < c->nw && th < 1024) th *= 2;  // power of two for the tree reduction
  return launch(c, "weights", [&]() {
    hipLaunchKernelGGL(weights_kernel, dim3(1), dim3(th), 0, c->stream, c->Wd, c->Wl, c->hybridW, c->nw, dt,
                       c->eshift);
  });
}

constexpr int kLqNB = 8;
size_t lq_lds_bytes(int M, int N)
{
  return sizeof(zd) * ((size_t)kLqNB * M + 128 * kLqNB + kLqNB * kLqNB + std::max(N, 128) + 4 * kLqNB * kLqNB);
}

// Cholesky QR declines a matrix whose pivot falls below this fraction of
// the largest diagonal entry of its Gram matrix (condition number above
// about 1e4), where its forward error would approach 1e-12.
constexpr double kCholTol = 1e-8;

// One Cholesky QR pass on the batch Win [2nw][M][N]: S = Win^T conj(Win)
// (= B^H B for B = conj(Win)), U = R^{-1} from S = R^H R, Wout = Win conj(U)
// (= conj(B R^{-1})).
int cholqr_pass(afqmc_ctx* c, const zd* Win, zd* Wout)
{
  const int M = c->M, N = c->N, nb = 2 * c->nw;
  ZGemmArgs a = gargs();
  a.A = Win;  // A(a, i) = Win[i][a]
  a.sam = 1;
  a.sak = N;
  a.sa1 = (long)M * N;
  a.B = Win;  // B(b, i) = conj(Win[i][b])
  a.sbn = 1;
  a.sbk = N;
  a.sb1 = (long)M * N;
  a.conjb = 1;
  a.C = c->T1;  // S[a][b]
  a.scm = N;
  a.scn = 1;
  a.sc1 = (long)N * N;
  a.m = N;
  a.n = N;
  a.k = M;
  RET((zgemm<CP_F64, double, double, double, false, false>(c, "ortho_gram", a, nb)));
  RET(launch(c, "ortho_chol_inv", [&]() {
    if (N <= 32)
      hipLaunchKernelGGL((chol_inv_kernel<32, 4>), dim3(nb), dim3(128), 0, c->stream, c->T1, c->oU, N, kCholTol,
                         c->ofail);
    else if (N <= 64)
      hipLaunchKernelGGL((chol_inv_kernel<64, 4>), dim3(nb), dim3(256), 0, c->stream, c->T1, c->oU, N, kCholTol,
                         c->ofail);
    else
      hipLaunchKernelGGL((chol_inv_kernel<128, 8>), dim3(nb), dim3(1024), 0, c->stream, c->T1, c->oU, N, kCholTol,
                         c->ofail);
  }));
  ZGemmArgs b = gargs();
  b.A = Win;  // A(i, a) = Win[i][a]
  b.sam = N;
  b.sak = 1;
  b.sa1 = (long)M * N;
  b.B = c->oU;  // B(b, a) = conj(U[a][b]), zero for a > b
  b.sbn = 1;
  b.sbk = N;
  b.sb1 = (long)N * N;
  b.conjb = 1;
  b.ktri = 1;
  b.C = Wout;
  b.scm = N;
  b.scn = 1;
  b.sc1 = (long)M * N;
  b.m = M;
  b.n = N;
  b.k = N;
  return zgemm<CP_F64, double, double, double, true, false>(c, "ortho_apply", b, nb);
}

// afqmc_sys::orthogonalize (afqmc_sys.hpp:175-207): the LAPACK LQ factor of
// every walker-spin matrix, by two Cholesky QR passes and the sign
// reconstruction; matrices too ill-conditioned for Cholesky QR go through
// the blocked Householder kernel.
int do_ortho(afqmc_ctx* c)
{
  c->t1inv_valid = false;
  c->g_fresh = false;
  const size_t lds = lq_lds_bytes(c->M, c->N);
  const int M = c->M, N = c->N, nb = 2 * c->nw;
  if (c->ortho_householder || N > 128) {
    return launch(c, "lq_orthogonalize", [&]() {
      hipLaunchKernelGGL(lq_blocked_kernel<kLqNB>, dim3(nb), dim3(256), lds, c->stream, c->W, M, N,
                         (const int*)nullptr);
    });
  }
  if (!c->oU) {
    RET(dalloc(c, &c->oU, (long)nb * N * N));
    RET(dalloc(c, &c->oX, 2L * nb * M * N));
    RET(dalloc(c, &c->ofail, nb));
  }
  zd* X1 = c->oX;
  zd* X2 = c->oX + (long)nb * M * N;
  RET(launch(c, "ortho_flags", [&]() {
    hipLaunchKernelGGL(fill_int_kernel, dim3((nb + 255) / 256), dim3(256), 0, c->stream, c->ofail, nb, 0);
  }));
  RET(cholqr_pass(c, c->W, X1));
  RET(cholqr_pass(c, X1, X2));
  RET(launch(c, "ortho_sign", [&]() {
    if (N <= 32)
      hipLaunchKernelGGL((ortho_sign_kernel<32, 4>), dim3(nb), dim3(128), 0, c->stream, X2, c->W, M, N, c->ofail);
    else if (N <= 64)
      hipLaunchKernelGGL((ortho_sign_kernel<64, 4>), dim3(nb), dim3(256), 0, c->stream, X2, c->W, M, N, c->ofail);
    else
      hipLaunchKernelGGL((ortho_sign_kernel<128, 8>), dim3(nb), dim3(1024), 0, c->stream, X2, c->W, M, N, c->ofail);
  }));
  return launch(c, "lq_fallback", [&]() {
    hipLaunchKernelGGL(lq_blocked_kernel<kLqNB>, dim3(nb), dim3(256), lds, c->stream, c->W, M, N,
                       (const int*)c->ofail);
  });
}

// Local energy from G in [nw][2][N][M] layout (type GT): one-body haj.G,
// Coulomb term from v = Lv.G (Lv [2][K][NM] of type LVT, scaled by sqrt(dt),
// contracted in precision CPV) and exchange in precision CP, from the dense
// block Vx (type VXT) or from the Cholesky form with Lx (type LT_).
template <int CPV, class LVT, int CP, class GT, class LT_, class YT, class VXT>
int energy_cholesky(afqmc_ctx* c, const void* Gp, const void* Lv, const void* Lx)
{
  const int M = c->M, N = c->N, K = c->K, nw = c->nw;
  const long NM = (long)N * M;
  const double inv_dt = 1.0 / c->dt_sys;
  RET(launch(c, "energy_zero", [&]() {
    (void)hipMemsetAsync(c->e_direct, 0, sizeof(zd) * nw, c->stream);
    (void)hipMemsetAsync(c->e_exch_part, 0, sizeof(zd) * 2 * nw, c->stream);
  }));
  // direct term: v(w,n) = sum Lt G ; direct = sum_n v^2 / dt
  if constexpr (CPV == CP_BF16X6) {
    // three pieces of the split G (built for the exchange) and of Lt
    RET((dense_contract_vbias<CP_BF16X6, bfs, bfs>(c, c->G3, Lv, c->e_v, "energy_v_gemm", c->G3l, c->tp_Lt,
                                                   "energy_v_reduce")));
  } else {
    RET((dense_contract_vbias<CPV, GT, LVT>(c, Gp, Lv, c->e_v, "energy_v_gemm", nullptr, nullptr,
                                            "energy_v_reduce")));
  }
  RET(launch(c, "energy_direct", [&]() {
    hipLaunchKernelGGL((sumsq_kernel<zd>), dim3(nw), dim3(256), 0, c->stream, K, c->e_v, 1L, (long)K, c->e_direct,
                       inv_dt);
  }));
  if (c->Vx) {
    // exchange from the dense same-spin block of Vakbl held in HBM as its
    // lower triangle with the strictly lower part doubled:
    // Z(w, p) = sum_{q <= p} G_s(w, q) Vx'_s(p, q);  e_x(w, s) = sum_p G_s(w, p) Z(w, p)
    // With a chunk size set, the columns q are split into chunks [q0, q0 + kc):
    // each chunk's GEMM accumulates over kc columns only (only p >= q0 meet
    // the triangle) and its dot with G is added in FP64, which bounds the
    // accumulation error of reduced-precision contractions.
    using ZT = typename ztype<YT>::type;
    const long ch = (CP != CP_F64 && c->exch_chunk > 0 && c->exch_chunk < NM) ? c->exch_chunk : NM;
    for (int sp = 0; sp < 2; sp++) {
      for (long q0 = 0; q0 < NM; q0 += ch) {
        const long kc = std::min(ch, NM - q0);
        ZGemmArgs a = gargs();
        a.A = zadv<GT>(Gp, sp * NM + q0);
        a.sam = 2 * NM;
        a.sak = 1;
        a.B = (const typename ztype<VXT>::type*)c->Vx + sp * NM * NM + q0 * NM + q0;
        a.sbn = NM;
        a.sbk = 1;
        a.C = (ZT*)c->Ybuf + q0;
        a.scm = NM;
        a.scn = 1;
        a.m = nw;
        a.n = (int)(NM - q0);
        a.k = (int)kc;
        a.ar = inv_dt;
        a.ktri = 1;
        if constexpr (CP == CP_BF16X6) {
          // the product reads the three-piece G; the dot below keeps the FP64 G
          a.A = c->G3 + sp * NM + q0;
          a.A3 = c->G3l + sp * NM + q0;
          a.B3 = c->Vx3 + sp * NM * NM + q0 * NM + q0;
          RET((zgemm<CP, bfs, VXT, YT, true, true>(c, "energy_vx_gemm", a, 1)));
        } else {
          RET((zgemm<CP, GT, VXT, YT, true, true>(c, "energy_vx_gemm", a, 1)));
        }
        RET(launch(c, "energy_exch_dot", [&]() {
          hipLaunchKernelGGL((coldot_kernel<typename ztype<GT>::type, ZT>), dim3(nw), dim3(256), 0, c->stream,
                             (int)(NM - q0), (const typename ztype<GT>::type*)Gp + sp * NM + q0, 1L, 2 * NM,
                             (const ZT*)c->Ybuf + q0, 1L, NM, c->e_exch_part + sp, 2L, 1.0, q0 > 0 ? 1 : 0);
        }));
      }
    }
  } else {
  // exchange: Y_n,s = Lt_s(n) G_s^T (N x N), sum_ab Y_ab Y_ba, for all
  // walker-spin pairs at once over chunks of nq vectors
  const long per_item = (long)N * N;
  const long ycap = (long)(c->ybuf_bytes / sizeof(typename ztype<YT>::type));
  const int nq = (int)std::max<long>(1, std::min<long>(K, ycap / (2L * nw * per_item)));
  for (int n0 = 0; n0 < K; n0 += nq) {
    const int cnt = std::min(nq, K - n0);
    ZGemmArgs a = gargs();
    a.A = zadv<LT_>(Lx, (long)n0 * NM);
    a.sam = M;
    a.sak = 1;
    a.sa1 = 0;
    a.sa2 = (long)K * NM;  // spin block of Lt
    a.B = Gp;              // G[w][s]
    a.sbn = M;
    a.sbk = 1;
    a.sb1 = 2 * NM;
    a.sb2 = NM;
    a.bdiv = 2;
    a.C = c->Ybuf;
    a.scm = N;
    a.scn = 1;
    a.sc1 = 2L * cnt * per_item;
    a.sc2 = (long)cnt * per_item;
    a.m = cnt * N;
    a.n = N;
    a.k = M;
    a.ar = std::sqrt(inv_dt);  // Y of the unscaled vectors
    if constexpr (CP == CP_BF16X6) {
      return fail(c, AFQMC_ERR_STATE, "three-piece exchange without its dense block");
    } else {
      RET((zgemm<CP, LT_, GT, YT, true, true>(c, "energy_y_gemm", a, 2 * nw)));
    }
    RET(launch(c, "energy_trace", [&]() {
      hipLaunchKernelGGL((trace_sq_kernel<typename ztype<YT>::type>), dim3(2 * nw), dim3(256), 0, c->stream,
                         (const typename ztype<YT>::type*)c->Ybuf, N, cnt, (long)cnt * per_item, c->e_exch_part, 1L,
                         2 * nw);
    }));
  }
  }
  RET(launch(c, "energy_pairsum", [&]() {
    hipLaunchKernelGGL(pairsum_kernel, dim3((nw + 255) / 256), dim3(256), 0, c->stream, c->e_exch_part, c->e_exch, nw);
  }));
  // one-body: e1 = sum_i G(w,i) haj(i)
  RET(launch(c, "energy_one", [&]() {
    hipLaunchKernelGGL((coldot_kernel<typename ztype<GT>::type, zd>), dim3(nw), dim3(256), 0, c->stream, (int)(2 * NM),
                       (const typename ztype<GT>::type*)Gp, 1L, 2 * NM, (const zd*)c->haj, 1L, 0L, c->e_one, 1L, 1.0,
                       0);
  }));
  return launch(c, "energy_assemble", [&]() {
    hipLaunchKernelGGL(energy_assemble_kernel, dim3((nw + 255) / 256), dim3(256), 0, c->stream, c->Wd, c->e_direct,
                       c->e_exch, c->e_one, nw);
  });
}

// Dense same-spin exchange block of Vakbl from the half-rotated vectors
// (the exchange half of the Vakbl construction the reference's input holds,
// initialize_serial.hpp:132-164):  Vx_s[(a,k)][(b,l)] = sum_n Lt_s(n,a,l) Lt_s(n,b,k)
// as N*N batched GEMMs over (a, b).  Lt64 is FP64 [2][K][NM].
template <class OT> int build_vx(afqmc_ctx* c, const zd* Lt64)
{
  const long M = c->M, N = c->N, K = c->K, NM = N * M;
  for (int sp = 0; sp < 2; sp++) {
    ZGemmArgs a = gargs();
    a.A = Lt64 + sp * K * NM;  // A_ab(l, n) = Lt(n, a, l)
    a.sam = 1;
    a.sak = NM;
    a.sa1 = M;
    a.sa2 = 0;
    a.B = Lt64 + sp * K * NM;  // B_ab(k, n) = Lt(n, b, k)
    a.sbn = 1;
    a.sbk = NM;
    a.sb1 = 0;
    a.sb2 = M;
    a.bdiv = (int)N;
    a.C = (typename ztype<OT>::type*)c->Vx + sp * NM * NM;  // C_ab(l, k) -> Vx[(a,k)][(b,l)]
    a.scm = 1;
    a.scn = NM;
    a.sc1 = M * NM;
    a.sc2 = M;
    a.m = (int)M;
    a.n = (int)M;
    a.k = (int)K;
    RET((zgemm<CP_F64, double, double, OT, false, false>(c, "build_vx", a, (int)(N * N))));
  }
  // Vx_s is symmetric (swap (a,k) <-> (b,l)): keep the doubled lower triangle
  using ZT = typename ztype<OT>::type;
  return launch(c, "vx_lower", [&]() {
    hipLaunchKernelGGL((sym_lower_kernel<ZT>), dim3(nblk(2 * NM * NM, 256)), dim3(256), 0, c->stream, (ZT*)c->Vx, NM,
                       2L);
  });
}

// Build Vx when it fits in device memory with room to spare; otherwise the
// energy uses the Cholesky form.
// ZT: storage of Vx; the exchange GEMM writes Z [nw][NM] of zsize bytes per element.
template <class ZT> int maybe_vx(afqmc_ctx* c, const zd* Lt64, size_t zsize)
{
  if (!c->allow_vx) return AFQMC_OK;
  const long NM = (long)c->N * c->M, nw = c->nw;
  const size_t bytes = sizeof(ZT) * 2 * (size_t)NM * NM;
  const size_t ybytes = zsize * (size_t)nw * NM;
  size_t fr = 0, tot = 0;
  HIPCHK(c, hipMemGetInfo(&fr, &tot));
  const size_t need = bytes + ybytes + ((size_t)8 << 30);
  if (need > fr) return AFQMC_OK;
  RET(dalloc_bytes(c, &c->Vx, bytes));
  if (c->ybuf_bytes < ybytes) {
    dfree(c, c->Ybuf);
    c->ybuf_bytes = ybytes;
    RET(dalloc_bytes(c, &c->Ybuf, ybytes));
  }
  if (std::is_same<ZT, zd>::value) return build_vx<double>(c, Lt64);
  if (std::is_same<ZT, zb>::value) return build_vx<bfs>(c, Lt64);
  return build_vx<float>(c, Lt64);
}

// Vx as three-piece split-BF16 (split-BF16 mode, energy level 2): built in
// FP32 like the FP32 block, then split in place into its (h, m) pieces with
// the third pieces alongside; the walkers' G is split likewise per energy.
int maybe_vx3(afqmc_ctx* c, const zd* Lt64)
{
  if (!c->allow_vx) return AFQMC_OK;
  const long NM = (long)c->N * c->M, nw = c->nw;
  const size_t bytes = (sizeof(zb) + sizeof(uint32_t)) * 2 * (size_t)NM * NM;
  const size_t gbytes = (sizeof(zb) + sizeof(uint32_t)) * 2 * (size_t)NM * nw;
  const size_t ybytes = sizeof(zf) * (size_t)nw * NM;
  size_t fr = 0, tot = 0;
  HIPCHK(c, hipMemGetInfo(&fr, &tot));
  if (bytes + gbytes + ybytes + ((size_t)8 << 30) > fr) return AFQMC_OK;
  RET(dalloc_bytes(c, &c->Vx, sizeof(zb) * 2 * (size_t)NM * NM));
  RET(dalloc(c, &c->Vx3, 2 * (size_t)NM * NM));
  RET(dalloc(c, &c->G3, 2 * (size_t)NM * nw));
  RET(dalloc(c, &c->G3l, 2 * (size_t)NM * nw));
  if (c->ybuf_bytes < ybytes) {
    dfree(c, c->Ybuf);
    c->ybuf_bytes = ybytes;
    RET(dalloc_bytes(c, &c->Ybuf, ybytes));
  }
  RET(build_vx<float>(c, Lt64));
  return launch(c, "vx_split3", [&]() {
    hipLaunchKernelGGL((split3_kernel<zf>), dim3(nblk(2 * NM * NM, 256)), dim3(256), 0, c->stream, (const zf*)c->Vx,
                       (zb*)c->Vx, c->Vx3, 2 * NM * NM);
  });
}

int do_energy(afqmc_ctx* c, double* eav_out)
{
  const int nw = c->nw, N = c->N, M = c->M;
  const long NAK = 2L * N * M;
  if (c->mode == AFQMC_FP64_CSR) {
    if (c->v_rp) {
      zd one{1, 0}, zero{0, 0};
      RET(csrmm_n(c, "csrmm_vakbl", (int)NAK, (int)NAK, c->v_nnz, c->v_rp, c->v_col, c->v_val, (const zd*)c->G,
                  (long)nw, c->Gcloc, (long)nw, one, zero));
      // E2 = sum G.(Vakbl G) lands in the direct slot (exchange slot zero),
      // so E_L = E2/2 + E1 as energy.hpp:93-100
      RET(launch(c, "energy_two", [&]() {
        (void)hipMemsetAsync(c->e_exch, 0, sizeof(zd) * nw, c->stream);
        hipLaunchKernelGGL((coldot_kernel<zd, zd>), dim3(nw), dim3(256), 0, c->stream, (int)NAK, (const zd*)c->G,
                           (long)nw, 1L, (const zd*)c->Gcloc, (long)nw, 1L, c->e_direct, 1L, 1.0, 0);
      }));
      RET(launch(c, "energy_one", [&]() {
        hipLaunchKernelGGL((coldot_kernel<zd, zd>), dim3(nw), dim3(256), 0, c->stream, (int)NAK, (const zd*)c->G,
                           (long)nw, 1L, (const zd*)c->haj, 1L, 0L, c->e_one, 1L, 1.0, 0);
      }));
      RET(launch(c, "energy_assemble", [&]() {
        hipLaunchKernelGGL(energy_assemble_kernel, dim3((nw + 255) / 256), dim3(256), 0, c->stream, c->Wd,
                           c->e_direct, c->e_exch, c->e_one, nw);
      }));
    } else {
      dim3 grid((nw + 31) / 32, (unsigned)((NAK + 31) / 32));
      RET(launch(c, "g_transpose", [&]() {
        hipLaunchKernelGGL((ztranspose_kernel<zd, zd>), grid, dim3(256), 0, c->stream, (const zd*)c->G, c->Gtmp,
                           (int)NAK, nw, (long)nw, NAK);
      }));
      RET((energy_cholesky<CP_F64, double, CP_F64, double, double, double, double>(c, c->Gtmp, c->Lt, c->Lt)));
    }
  } else if (c->mode == AFQMC_FP64_DENSE) {
    RET((energy_cholesky<CP_F64, double, CP_F64, double, double, double, double>(c, c->G, c->Lt, c->Lt)));
  } else if (c->Lt64) {
    // FP64 density matrix, one-body and Coulomb terms; the exchange in FP64
    // (level 3), FP32 (level 2) or the contraction precision (level 1)
    RET(g_fp64(c));
    const void *G = c->Gtmp, *L64 = c->Lt64;
    if (c->energy_level >= 3)
      RET((energy_cholesky<CP_F64, double, CP_F64, double, double, double, double>(c, G, L64, L64)));
    else if (c->energy_level == 2 && c->Vx3) {
      // the exchange in three-piece split-BF16 (24-bit operands, FP32
      // accumulation): G split from its FP64 copy
      RET(launch(c, "g_split3", [&]() {
        hipLaunchKernelGGL((split3_kernel<zd>), dim3(nblk(2L * N * M * nw, 256)), dim3(256), 0, c->stream,
                           (const zd*)c->Gtmp, c->G3, c->G3l, 2L * N * M * nw);
      }));
      if (c->coulomb3 && c->tp_Lt)
        RET((energy_cholesky<CP_BF16X6, bfs, CP_BF16X6, double, double, float, bfs>(c, G, c->Lt, L64)));
      else
        RET((energy_cholesky<CP_F64, double, CP_BF16X6, double, double, float, bfs>(c, G, L64, L64)));
    } else if (c->energy_level == 2)
      RET((energy_cholesky<CP_F64, double, CP_F32, double, double, float, float>(c, G, L64, L64)));
    else if (c->bsplit)
      RET((energy_cholesky<CP_F64, double, CP_BF16X3, double, bfs, float, bfs>(c, G, L64, c->Lt)));
    else
      RET((energy_cholesky<CP_F64, double, CP_F32, double, float, float, float>(c, G, L64, c->Lt)));
  } else if (c->mode == AFQMC_MIXED_FP32) {
    RET((energy_cholesky<CP_F32, float, CP_F32, float, float, float, float>(c, c->G, c->Lt, c->Lt)));
  } else {
    RET((energy_cholesky<CP_BF16X3, bfs, CP_BF16X3, bfs, bfs, float, bfs>(c, c->G, c->Lt, c->Lt)));
  }
  if (eav_out) {
    std::vector<zd> wd((size_t)nw * 8);
    HIPCHK(c, hipMemcpyAsync(wd.data(), c->Wd, sizeof(zd) * wd.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double eav = 0, wgt = 0;
    for (int n = 0; n < nw; n++) {
      wgt += wd[(size_t)n * 8 + 1].re;
      eav += wd[(size_t)n * 8].re * wd[(size_t)n * 8 + 1].re;
    }
    *eav_out = eav / wgt;
  }
  return AFQMC_OK;
}

int check_flags(afqmc_ctx* c)
{
  int f[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(f, c->flags, sizeof(f), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (f[1]) {
    HIPCHK(c, hipMemsetAsync(c->flags + 1, 0, sizeof(int), c->stream));
    return fail(c, AFQMC_ERR_NONFINITE, "non-finite auxiliary field draw");
  }
  if (f[0]) {
    HIPCHK(c, hipMemsetAsync(c->flags, 0, sizeof(int), c->stream));
    return fail(c, AFQMC_ERR_SINGULAR, "singular overlap matrix (zero pivot)");
  }
  return AFQMC_OK;
}

int substep_body(afqmc_ctx* c, const double* xdev, long substep, int do_ortho_, double dt,
                 const long* xslice = nullptr)
{
  if (c->transposed) {
    RET(do_mdm(c));
    RET(do_vbias(c));
  } else {
    RET(do_mdm_full(c));
    RET(do_vbias_full(c));
  }
  RET(do_fields(c, xdev, substep, xslice));
  RET(do_vhs(c));
  RET(do_propagate(c));
  RET(do_save_overlaps(c));
  RET(do_overlaps(c));
  RET(do_weights(c, nullptr, dt));
  if (do_ortho_) {
    RET(do_ortho(c));
    RET(do_overlaps(c));
  }
  return AFQMC_OK;
}

bool dense_layout(const afqmc_ctx* c) { return c->mode != AFQMC_FP64_CSR; }

// host split-BF16 packing, same rounding as the device's f2bf
inline uint16_t host_f2bf(float x)
{
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
inline float host_bf2f(uint16_t b)
{
  uint32_t u = ((uint32_t)b) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
inline void host_split(double re, double im, uint32_t w[2])
{
  const float r = (float)re, i = (float)im;
  const uint16_t hr = host_f2bf(r), hi = host_f2bf(i);
  const uint16_t lr = host_f2bf(r - host_bf2f(hr)), li = host_f2bf(i - host_bf2f(hi));
  w[0] = (uint32_t)hr | ((uint32_t)hi << 16);
  w[1] = (uint32_t)lr | ((uint32_t)li << 16);
}
// the third piece of (re, im) given its two-piece split w
inline uint32_t host_third(double re, double im, const uint32_t w[2])
{
  const double rr = re - host_bf2f((uint16_t)(w[0] & 0xffff)) - host_bf2f((uint16_t)(w[1] & 0xffff));
  const double ri = im - host_bf2f((uint16_t)(w[0] >> 16)) - host_bf2f((uint16_t)(w[1] >> 16));
  return (uint32_t)host_f2bf((float)rr) | ((uint32_t)host_f2bf((float)ri) << 16);
}
inline zd host_merge(const uint32_t w[2])
{
  return zd{(double)host_bf2f((uint16_t)(w[0] & 0xffff)) + (double)host_bf2f((uint16_t)(w[1] & 0xffff)),
            (double)host_bf2f((uint16_t)(w[0] >> 16)) + (double)host_bf2f((uint16_t)(w[1] >> 16))};
}

// host transpose helpers: ref [R][C] <-> dev [C][R]
void htranspose(const zd* in, zd* out, long R, long Cc)
{
  for (long r = 0; r < R; r++)
    for (long cc = 0; cc < Cc; cc++) out[cc * R + r] = in[r * Cc + cc];
}

// Drop the captured graphs (the staged block, dt or a buffer changed).
void drop_graphs(afqmc_ctx* c)
{
  c->graph_gen++;  // groups holding this context's substep capture again
  for (int i = 0; i < 2; i++) {
    if (c->gexec[i]) (void)hipGraphExecDestroy(c->gexec[i]);
    if (c->graph[i]) (void)hipGraphDestroy(c->graph[i]);
    c->gexec[i] = nullptr;
    c->graph[i] = nullptr;
  }
}

// Capture one substep (without / with orthogonalisation) on the context's
// stream.  Buffers the substep would allocate on first use are allocated
// here, outside the capture.  The graph starts from the inverses of the
// current walkers' overlap matrices (the state every substep ends in), so
// its density matrix reuses them.
int capture_prep(afqmc_ctx* c, int ortho);

int capture_substep(afqmc_ctx* c, int ortho, double dt)
{
  RET(capture_prep(c, ortho));
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int r = substep_body(c, c->dev_fields ? nullptr : c->xstage, 0, ortho, dt, c->d_sub);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (r != AFQMC_OK) {
    if (g) (void)hipGraphDestroy(g);
    return r;
  }
  if (e != hipSuccess) return fail(c, AFQMC_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  c->graph[ortho] = g;
  HIPCHK(c, hipGraphInstantiate(&c->gexec[ortho], g, nullptr, nullptr, 0));
  return AFQMC_OK;
}

// Everything a captured substep needs that it would otherwise allocate or
// compute on first use, done outside the capture; the context's stream is
// idle afterwards.
int capture_prep(afqmc_ctx* c, int ortho)
{
  const long MN = (long)c->M * c->N, nb = 2L * c->nw;
  if (!c->d_sub) RET(dalloc(c, &c->d_sub, 1));
  if (ortho && c->N <= 128 && !c->ortho_householder && !c->oU) {
    RET(dalloc(c, &c->oU, nb * c->N * c->N));
    RET(dalloc(c, &c->oX, 2L * nb * MN));
    RET(dalloc(c, &c->ofail, nb));
  }
  if (!c->transposed && !c->Gfull) RET(dalloc(c, &c->Gfull, 2L * c->M * c->M * c->nw));
  if (c->mode != AFQMC_FP64_CSR || c->sptd) {
    // the bias contraction's split-k workspace at its largest (at most 32
    // chunks, also for the dense copy of SpvnT)
    const size_t need = sizeof(zd) * 32 * (size_t)c->nw * c->K;
    if (c->vws_bytes < need) {
      dfree(c, c->vws);
      c->vws_bytes = need;
      RET(dalloc_bytes(c, &c->vws, need));
    }
  }
  if (!c->t1inv_valid) RET(do_overlaps(c));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AFQMC_OK;
}

} // namespace

extern "C" {

int afqmc_ctx_create(int nmo, int naea, int nchol, int nwalk, int mode, afqmc_ctx** out)
{
  if (!out) return AFQMC_ERR_ARG;
  *out = nullptr;
  if (nmo <= 0 || naea <= 0 || naea > nmo || nchol <= 0 || nwalk <= 0) return AFQMC_ERR_SHAPE;
  if (mode < AFQMC_FP64_CSR || mode > AFQMC_MIXED_BF16) return AFQMC_ERR_ARG;
  if (naea > GJ_GMAX) return AFQMC_ERR_UNSUPPORTED;
  if (lq_lds_bytes(nmo, naea) > 160 * 1024) return AFQMC_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return AFQMC_ERR_HIP;
  afqmc_ctx* c = new afqmc_ctx;
  c->M = nmo;
  c->N = naea;
  c->K = nchol;
  c->nw = nwalk;
  c->mode = mode;
  c->mixed = (mode == AFQMC_MIXED_FP32 || mode == AFQMC_MIXED_BF16);
  c->bsplit = (mode == AFQMC_MIXED_BF16);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return AFQMC_ERR_HIP;
  }
  if (hipFuncSetAttribute((const void*)lq_blocked_kernel<kLqNB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lq_lds_bytes(nmo, naea)) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return AFQMC_ERR_HIP;
  }
  const long M = nmo, N = naea, K = nchol, nw = nwalk;
  const long MN = M * N, MM = M * M;
  const size_t zs = c->mixed ? sizeof(zf) : sizeof(zd);
  int r = AFQMC_OK;
  auto chk = [&](int x) {
    if (r == AFQMC_OK) r = x;
  };
  chk(dalloc(c, &c->W, nw * 2 * MN));
  chk(dalloc(c, &c->Wd, nw * 8));
  chk(dalloc(c, &c->T1, nw * 2 * N * N));
  chk(dalloc(c, &c->dcache, nw * 2));
  chk(dalloc(c, &c->lcache, nw * 2));
  chk(dalloc(c, &c->Wl, nw * 4));
  chk(dalloc_bytes(c, &c->G, nw * 2 * MN * zs));
  chk(dalloc(c, &c->vbias, nw * K));
  chk(dalloc_bytes(c, &c->X, nw * K * zs));
  chk(dalloc_bytes(c, &c->vHS, nw * MM * zs));
  if (mode == AFQMC_FP64_CSR) chk(dalloc(c, &c->vHSw, nw * MM));
  chk(dalloc(c, &c->hybridW, nw));
  chk(dalloc_bytes(c, &c->S, nw * 2 * MN * zs));
  chk(dalloc_bytes(c, &c->Ta, nw * 2 * MN * zs));
  chk(dalloc_bytes(c, &c->Tb, nw * 2 * MN * zs));
  chk(dalloc(c, &c->eshift, 1));
  chk(dalloc(c, &c->flags, 4));
  chk(dalloc(c, &c->e_direct, nw));
  chk(dalloc(c, &c->e_exch, nw));
  chk(dalloc(c, &c->e_exch_part, 2 * nw));
  chk(dalloc(c, &c->e_one, nw));
  chk(dalloc(c, &c->e_v, nw * K));
  c->ybuf_bytes = std::max<long>(2 * nw * N * N, std::min<long>(2 * nw * K * N * N, 1L << 25)) * sizeof(zd);
  chk(dalloc_bytes(c, &c->Ybuf, c->ybuf_bytes));
  if (r != AFQMC_OK) {
    afqmc_ctx_destroy(c);
    return r;
  }
  (void)hipMemsetAsync(c->eshift, 0, sizeof(double), c->stream);
  (void)hipMemsetAsync(c->flags, 0, 4 * sizeof(int), c->stream);
  (void)hipMemsetAsync(c->Wd, 0, sizeof(zd) * nw * 8, c->stream);
  hipLaunchKernelGGL(wd_logs_kernel, dim3((unsigned)((4L * nw + 255) / 256)), dim3(256), 0, c->stream, c->Wd, c->Wl,
                     nw);
  *out = c;
  return AFQMC_OK;
}

int afqmc_ctx_destroy(afqmc_ctx* c)
{
  if (!c) return AFQMC_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < 2; i++)
    if (c->gexec[i]) (void)hipGraphExecDestroy(c->gexec[i]);
  for (int i = 0; i < 2; i++)
    if (c->graph[i]) (void)hipGraphDestroy(c->graph[i]);
  for (auto& p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->evpool) (void)hipEventDestroy(e);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return AFQMC_OK;
}

const char* afqmc_ctx_error(const afqmc_ctx* c) { return c ? c->err.c_str() : "null context"; }

int afqmc_sync(afqmc_ctx* c)
{
  if (!c) return AFQMC_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AFQMC_OK;
}

int afqmc_set_option(afqmc_ctx* c, int option, double value)
{
  if (!c) return AFQMC_ERR_ARG;
  if (option == AFQMC_OPT_DENSE_EXCHANGE || option == AFQMC_OPT_ENERGY_PRECISION) {
    if (c->have_system) return fail(c, AFQMC_ERR_STATE, "set options before uploading the system");
    if (option == AFQMC_OPT_DENSE_EXCHANGE) {
      c->allow_vx = value != 0.0;
    } else {
      if (value < 0 || value > 3 || value != (int)value) return fail(c, AFQMC_ERR_ARG, "energy precision level is 0..3");
      c->energy_level = (int)value;
    }
    return AFQMC_OK;
  }
  if (option == AFQMC_OPT_CSR_DENSE) {
    if (c->have_system) return fail(c, AFQMC_ERR_STATE, "set options before uploading the system");
    c->csr_dense = value != 0.0;
    return AFQMC_OK;
  }
  if (option == AFQMC_OPT_COULOMB_THREE_PIECE) {
    if (c->have_system) return fail(c, AFQMC_ERR_STATE, "set options before uploading the system");
    if (value != 0.0 && c->mode != AFQMC_MIXED_BF16)
      return fail(c, AFQMC_ERR_UNSUPPORTED, "the three-piece Coulomb term is a split-BF16 option");
    c->coulomb3 = value != 0.0;
    return AFQMC_OK;
  }
  if (option == AFQMC_OPT_TRANSPOSED) {
    if (c->have_system) return fail(c, AFQMC_ERR_STATE, "set options before uploading the system");
    if (value == 0.0 && c->mode != AFQMC_FP64_CSR)
      return fail(c, AFQMC_ERR_UNSUPPORTED, "the non-transposed bias is built for AFQMC_FP64_CSR only");
    c->transposed = value != 0.0;
    return AFQMC_OK;
  }
  if (option == AFQMC_OPT_EXCHANGE_CHUNK) {
    if (value < 0 || value != (long)value || ((long)value % 256) != 0)
      return fail(c, AFQMC_ERR_ARG, "exchange chunk must be 0 or a positive multiple of 256");
    c->exch_chunk = (long)value;
    return AFQMC_OK;
  }
  if (option == AFQMC_OPT_FUSED_PROPAGATOR) {
    drop_graphs(c);
    c->fused_prop = value != 0.0;
    return AFQMC_OK;
  }
  if (option == AFQMC_OPT_THREE_PIECE) {
    if (c->have_system) return fail(c, AFQMC_ERR_STATE, "set options before uploading the system");
    if (value < 0 || value > 7 || value != (int)value) return fail(c, AFQMC_ERR_ARG, "three-piece mask is 0..7");
    if (value != 0 && c->mode != AFQMC_MIXED_BF16)
      return fail(c, AFQMC_ERR_UNSUPPORTED, "three-piece contractions are a split-BF16 option");
    // three-piece products run on the staged-ring kernel only (k-tiles of
    // 16 complex): the propagator contracts over M, the HS potential over K,
    // the bias product over N*M
    const int v = (int)value;
    if (((v & 4) && c->M % 16) || ((v & 2) && c->K % 16) || ((v & 1) && ((long)c->N * c->M) % 16))
      return fail(c, AFQMC_ERR_UNSUPPORTED,
                  "three-piece contractions need their contraction length (M, K, N*M) a multiple of 16");
    c->three = v;
    return AFQMC_OK;
  }
  if (option == AFQMC_OPT_THREE_PIECE_STEPS) {
    if (value < 0 || value > 255 || value != (int)value)
      return fail(c, AFQMC_ERR_ARG, "three-piece propagator step mask is 0..255");
    drop_graphs(c);
    c->three_steps = (int)value;
    return AFQMC_OK;
  }
  if (option == AFQMC_OPT_ONE_PIECE_STEPS) {
    if (value < 0 || value > 255 || value != (int)value || ((int)value & 0x81))
      return fail(c, AFQMC_ERR_ARG, "one-piece step mask takes Horner steps only (bits 1-6)");
    drop_graphs(c);
    c->one_steps = (int)value;
    return AFQMC_OK;
  }
  if (option == AFQMC_OPT_TILE) {
    if (value < 0.0 || value > 8.0) return fail(c, AFQMC_ERR_ARG, "tile candidate out of range");
    c->force_cfg = (int)value;
    drop_graphs(c);
    return AFQMC_OK;
  }
  if (option == AFQMC_OPT_DIRECT_LDS) {
    c->gemm_ring = value >= 2.0;
    drop_graphs(c);
    c->gemm_adma = value != 0.0;
    return AFQMC_OK;
  }
  if (option == AFQMC_OPT_ORTHO_HOUSEHOLDER) {
    drop_graphs(c);
    c->ortho_householder = value != 0.0;
    return AFQMC_OK;
  }
  return fail(c, AFQMC_ERR_ARG, "unknown option");
}

int afqmc_upload_system(afqmc_ctx* c, const double* trial_alpha, const double* trial_beta, const double* propg1,
                        const double* haj, const afqmc_csr_view* spvn, const afqmc_csr_view* vakbl, double cutoff,
                        double dt)
{
  if (!c || !trial_alpha || !trial_beta || !propg1 || !haj || !spvn) return AFQMC_ERR_ARG;
  const long M = c->M, N = c->N, K = c->K;
  const long MN = M * N, MM = M * M, NM = MN;
  if (spvn->nrows != MM || spvn->ncols != K || spvn->nnz < 0 || !spvn->rowptr || (spvn->nnz > 0 && (!spvn->col || !spvn->val)))
    return fail(c, AFQMC_ERR_SHAPE, "Spvn must be CSR [M*M][K]");
  if (spvn->rowptr[MM] - spvn->rowptr[0] != spvn->nnz) return fail(c, AFQMC_ERR_SHAPE, "Spvn rowptr/nnz mismatch");
  if (vakbl && vakbl->nnz > 0 &&
      (vakbl->nrows != 2 * NM || vakbl->ncols != 2 * NM || !vakbl->rowptr || !vakbl->col || !vakbl->val))
    return fail(c, AFQMC_ERR_SHAPE, "Vakbl must be CSR [2NM][2NM]");
  if (vakbl && vakbl->nnz > 0 && vakbl->rowptr[2 * NM] - vakbl->rowptr[0] != vakbl->nnz)
    return fail(c, AFQMC_ERR_SHAPE, "Vakbl rowptr/nnz mismatch");
  // entries of row r sit at rowptr[r] - rowptr[0] from these bases
  const int* scol = spvn->nnz > 0 ? spvn->col + spvn->rowptr[0] : nullptr;
  const double* sval = spvn->nnz > 0 ? spvn->val + 2 * (long)spvn->rowptr[0] : nullptr;
  if (dt <= 0.0) return fail(c, AFQMC_ERR_ARG, "dt must be positive");
  if (c->have_system) return fail(c, AFQMC_ERR_STATE, "system already uploaded");
  c->dt_sys = dt;
  hipStream_t st = c->stream;
  RET(dalloc(c, &c->trial, 2 * MN));
  RET(dalloc(c, &c->propg, MM));
  RET(dalloc(c, &c->haj, 2 * NM));
  HIPCHK(c, hipMemcpyAsync(c->trial, trial_alpha, sizeof(zd) * MN, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->trial + MN, trial_beta, sizeof(zd) * MN, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->propg, propg1, sizeof(zd) * MM, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->haj, haj, sizeof(zd) * 2 * NM, hipMemcpyHostToDevice, st));
  if (c->mixed) {
    RET(dalloc_bytes(c, &c->propg32, sizeof(zf) * MM));
    RET(launch(c, "convert", [&]() {
      if (c->bsplit)
        hipLaunchKernelGGL((zconvert_kernel<zd, zb>), dim3(nblk(MM, 256)), dim3(256), 0, st, c->propg,
                           (zb*)c->propg32, MM);
      else
        hipLaunchKernelGGL((zconvert_kernel<zd, zf>), dim3(nblk(MM, 256)), dim3(256), 0, st, c->propg,
                           (zf*)c->propg32, MM);
    }));
    if (c->three) {
      // three-piece contractions need 16-complex k-tiles of aligned rows
      if (M % 16 || K % 16 || NM % 16) return fail(c, AFQMC_ERR_UNSUPPORTED, "three-piece contractions need M, K and N*M multiples of 16");
      if (c->three & 4) {
        RET(dalloc(c, &c->tp_P, MM));
        RET(launch(c, "third_pieces", [&]() {
          hipLaunchKernelGGL(third_piece_kernel, dim3(nblk(MM, 256)), dim3(256), 0, st, c->propg, (const zb*)c->propg32,
                             c->tp_P, MM, 0L, 0L, 0);
        }));
        RET(dalloc(c, &c->tp_vHS, (size_t)c->nw * MM));
        RET(dalloc(c, &c->tp_Ta, (size_t)c->nw * 2 * MN));
        RET(dalloc(c, &c->tp_Tb, (size_t)c->nw * 2 * MN));
      }
      if (c->three & 2) RET(dalloc(c, &c->tp_X, (size_t)c->nw * K));
      if (c->three & 1) RET(dalloc(c, &c->tp_G, (size_t)c->nw * 2 * NM));
    }
  }
  // Spvn CSR to the device
  const long nnz = spvn->nnz;
  int* rp = nullptr;
  int* col = nullptr;
  zd* val = nullptr;
  RET(dalloc(c, &rp, MM + 1));
  RET(dalloc(c, &col, nnz));
  RET(dalloc(c, &val, nnz));
  {
    std::vector<int> rp0(MM + 1);
    for (long i = 0; i <= MM; i++) rp0[i] = spvn->rowptr[i] - spvn->rowptr[0];
    HIPCHK(c, hipMemcpyAsync(rp, rp0.data(), sizeof(int) * (MM + 1), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(col, scol, sizeof(int) * nnz, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(val, sval, sizeof(zd) * nnz, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipStreamSynchronize(st));
  }
  for (long r = 0; r < MM; r++)
    if (spvn->rowptr[r + 1] < spvn->rowptr[r]) return fail(c, AFQMC_ERR_SHAPE, "Spvn rowptr is not monotone");
  for (long p = 0; p < nnz; p++)
    if (scol[p] < 0 || scol[p] >= K) return fail(c, AFQMC_ERR_SHAPE, "Spvn column index out of range");
  c->sp_sorted = true;
  for (long r = 0; r < MM && c->sp_sorted; r++)
    for (long q = spvn->rowptr[r] - spvn->rowptr[0] + 1; q < spvn->rowptr[r + 1] - spvn->rowptr[0]; q++)
      if (scol[q] < scol[q - 1]) {
        c->sp_sorted = false;
        break;
      }
  // Repeated (row, column) entries: the reference keeps them (SparseMatrix::
  // compress only sorts, SparseMatrix.hpp:295-319) and csrmm sums them, so
  // the CSR products use the matrix as given; the dense forms take each
  // repeated entry's sum, formed here on the host in entry order.
  const int* drp = rp;
  const int* dcol = col;
  const zd* dval = val;
  {
    std::vector<long> last(K, -1), pos(K, 0);
    bool dup = false;
    for (long r = 0; r < MM && !dup; r++)
      for (long q = spvn->rowptr[r] - spvn->rowptr[0]; q < spvn->rowptr[r + 1] - spvn->rowptr[0]; q++) {
        if (last[scol[q]] == r) {
          dup = true;
          break;
        }
        last[scol[q]] = r;
      }
    if (dup) {
      std::vector<int> mrp(MM + 1), mcol;
      std::vector<zd> mval;
      mcol.reserve(nnz);
      mval.reserve(nnz);
      std::fill(last.begin(), last.end(), -1);
      const zd* sv = (const zd*)sval;
      for (long r = 0; r < MM; r++) {
        mrp[r] = (int)mcol.size();
        for (long q = spvn->rowptr[r] - spvn->rowptr[0]; q < spvn->rowptr[r + 1] - spvn->rowptr[0]; q++) {
          const int cq = scol[q];
          if (last[cq] == r) {
            mval[pos[cq]] = zadd(mval[pos[cq]], sv[q]);
          } else {
            last[cq] = r;
            pos[cq] = (long)mcol.size();
            mcol.push_back(cq);
            mval.push_back(sv[q]);
          }
        }
      }
      mrp[MM] = (int)mcol.size();
      int *mrp_d = nullptr, *mcol_d = nullptr;
      zd* mval_d = nullptr;
      RET(dalloc(c, &mrp_d, MM + 1));
      RET(dalloc(c, &mcol_d, (long)mcol.size()));
      RET(dalloc(c, &mval_d, (long)mval.size()));
      HIPCHK(c, hipMemcpyAsync(mrp_d, mrp.data(), sizeof(int) * (MM + 1), hipMemcpyHostToDevice, st));
      HIPCHK(c, hipMemcpyAsync(mcol_d, mcol.data(), sizeof(int) * mcol.size(), hipMemcpyHostToDevice, st));
      HIPCHK(c, hipMemcpyAsync(mval_d, mval.data(), sizeof(zd) * mval.size(), hipMemcpyHostToDevice, st));
      HIPCHK(c, hipStreamSynchronize(st));
      drp = mrp_d;
      dcol = mcol_d;
      dval = mval_d;
    }
  }
  // dense Ld[n][ik] (FP64), then the half rotation Lt[s][n][a][k] = sum_i trial_s(i,a) Ld[n][i][k]
  zd* Ld64 = nullptr;
  RET(dalloc(c, &Ld64, K * MM));
  HIPCHK(c, hipMemsetAsync(Ld64, 0, sizeof(zd) * K * MM, st));
  RET(launch(c, "csr_densify", [&]() {
    hipLaunchKernelGGL(csr_densify_kernel, dim3((unsigned)((MM + 3) / 4)), dim3(256), 0, st, (int)MM, drp, dcol, dval,
                       Ld64, MM);
  }));
  if (drp != rp) {
    HIPCHK(c, hipStreamSynchronize(st));
    dfree(c, (void*)drp);
    dfree(c, (void*)dcol);
    dfree(c, (void*)dval);
  }
  zd* Lt64 = nullptr;
  RET(dalloc(c, &Lt64, 2 * K * NM));
  for (int sp = 0; sp < 2; sp++) {
    ZGemmArgs a = gargs();
    a.A = c->trial + sp * MN;  // A(a, i) = trial[i][a]
    a.sam = 1;
    a.sak = N;
    a.B = Ld64;  // B(k, i) = Ld[n][i][k], batch n
    a.sbn = 1;
    a.sbk = M;
    a.sb1 = MM;
    a.C = Lt64 + sp * K * NM;
    a.scm = M;
    a.scn = 1;
    a.sc1 = NM;
    a.m = (int)N;
    a.n = (int)M;
    a.k = (int)M;
    RET((zgemm<CP_F64, double, double, double, false, false>(c, "halfrotate_gemm", a, (int)K)));
  }
  if (c->mode == AFQMC_FP64_CSR) {
    c->sp_rp = rp;
    c->sp_col = col;
    c->sp_val = val;
    c->sp_nnz = nnz;
    if (!c->transposed) {
      // T(Spvn) as CSR [K][M*M] (counting sort by column; rows stay ordered)
      std::vector<int> trp(K + 1, 0), tcol(nnz);
      std::vector<zd> tval(nnz);
      const int r0 = spvn->rowptr[0];
      for (long q = 0; q < nnz; q++) trp[scol[q] + 1]++;
      for (long n = 0; n < K; n++) trp[n + 1] += trp[n];
      std::vector<int> pos(trp.begin(), trp.end() - 1);
      for (long r = 0; r < MM; r++)
        for (int q = spvn->rowptr[r] - r0; q < spvn->rowptr[r + 1] - r0; q++) {
          const int d = pos[scol[q]]++;
          tcol[d] = (int)r;
          tval[d] = zd{sval[2 * q], sval[2 * q + 1]};
        }
      RET(dalloc(c, &c->sptt_rp, K + 1));
      RET(dalloc(c, &c->sptt_col, nnz));
      RET(dalloc(c, &c->sptt_val, nnz));
      HIPCHK(c, hipMemcpyAsync(c->sptt_rp, trp.data(), sizeof(int) * (K + 1), hipMemcpyHostToDevice, st));
      HIPCHK(c, hipMemcpyAsync(c->sptt_col, tcol.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, st));
      HIPCHK(c, hipMemcpyAsync(c->sptt_val, tval.data(), sizeof(zd) * nnz, hipMemcpyHostToDevice, st));
      HIPCHK(c, hipStreamSynchronize(st));
    }
    // SpvnT CSR [K][2NM] with rotate.hpp's cutoff
    int* cnt = nullptr;
    RET(dalloc(c, &cnt, K));
    RET(launch(c, "rot_count", [&]() {
      hipLaunchKernelGGL(rot_count_kernel, dim3((unsigned)K), dim3(256), 0, st, Lt64, (int)K, NM, cutoff, cnt);
    }));
    std::vector<int> hc(K), hrp(K + 1, 0);
    HIPCHK(c, hipMemcpyAsync(hc.data(), cnt, sizeof(int) * K, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    long tot = 0;
    for (long i = 0; i < K; i++) {
      hrp[i] = (int)tot;
      tot += hc[i];
      if (tot > 2147483647L) return fail(c, AFQMC_ERR_SHAPE, "SpvnT exceeds 2^31 nonzeros");
    }
    hrp[K] = (int)tot;
    c->spt_nnz = tot;
    RET(dalloc(c, &c->spt_rp, K + 1));
    RET(dalloc(c, &c->spt_col, tot));
    RET(dalloc(c, &c->spt_val, tot));
    HIPCHK(c, hipMemcpyAsync(c->spt_rp, hrp.data(), sizeof(int) * (K + 1), hipMemcpyHostToDevice, st));
    RET(launch(c, "rot_fill", [&]() {
      hipLaunchKernelGGL(rot_fill_kernel, dim3((unsigned)K), dim3(256), 0, st, Lt64, (int)K, NM, cutoff, c->spt_rp,
                         c->spt_col, c->spt_val);
    }));
    HIPCHK(c, hipStreamSynchronize(st));
    dfree(c, cnt);
    if (c->csr_dense && c->transposed && 2 * tot >= K * 2 * NM) {
      // a CSR product over rows that are at least half full is a dense
      // product: keep SpvnT dense as well (entries under the cutoff are 0)
      RET(dalloc(c, &c->sptd, K * 2 * NM));
      HIPCHK(c, hipMemsetAsync(c->sptd, 0, sizeof(zd) * K * 2 * NM, st));
      RET(launch(c, "csr_densify", [&]() {
        hipLaunchKernelGGL(csr_densify_rows_kernel, dim3((unsigned)((K + 3) / 4)), dim3(256), 0, st, (int)K,
                           c->spt_rp, c->spt_col, c->spt_val, c->sptd, 2 * NM);
      }));
      HIPCHK(c, hipStreamSynchronize(st));
    }
    dfree(c, Ld64);
    if (vakbl && vakbl->nnz > 0) {
      const long vn = vakbl->nnz, vr = vakbl->nrows;
      const int* vcol = vakbl->col + vakbl->rowptr[0];
      for (long r = 0; r < vr; r++)
        if (vakbl->rowptr[r + 1] < vakbl->rowptr[r]) return fail(c, AFQMC_ERR_SHAPE, "Vakbl rowptr is not monotone");
      for (long q = 0; q < vn; q++)
        if (vcol[q] < 0 || vcol[q] >= vr) return fail(c, AFQMC_ERR_SHAPE, "Vakbl column index out of range");
      RET(dalloc(c, &c->v_rp, vr + 1));
      RET(dalloc(c, &c->v_col, vn));
      RET(dalloc(c, &c->v_val, vn));
      std::vector<int> rp0(vr + 1);
      for (long i = 0; i <= vr; i++) rp0[i] = vakbl->rowptr[i] - vakbl->rowptr[0];
      HIPCHK(c, hipMemcpyAsync(c->v_rp, rp0.data(), sizeof(int) * (vr + 1), hipMemcpyHostToDevice, st));
      HIPCHK(c, hipMemcpyAsync(c->v_col, vakbl->col + vakbl->rowptr[0], sizeof(int) * vn, hipMemcpyHostToDevice, st));
      HIPCHK(c, hipMemcpyAsync(c->v_val, vakbl->val + 2 * (long)vakbl->rowptr[0], sizeof(zd) * vn, hipMemcpyHostToDevice, st));
      c->v_nnz = vn;
      RET(dalloc(c, &c->Gcloc, 2 * NM * c->nw));
      HIPCHK(c, hipStreamSynchronize(st));
      dfree(c, Lt64);
    } else {
      c->Lt = Lt64;  // two-body energy from the vectors
      RET(dalloc(c, &c->Gtmp, 2 * NM * c->nw));
      RET(maybe_vx<zd>(c, Lt64, sizeof(zd)));
    }
  } else {
    dfree(c, rp);
    dfree(c, col);
    dfree(c, val);
    if (!c->mixed || c->energy_level >= 3) RET(maybe_vx<zd>(c, Lt64, sizeof(zd)));
    else if (c->bsplit && c->energy_level <= 1) RET(maybe_vx<zb>(c, Lt64, sizeof(zf)));
    // the three-piece exchange runs on the staged-ring kernel, whose k-tiles
    // are 16 complex (and chunk offsets 16-byte aligned): other sizes keep
    // the FP32 block, as the FP32 mode's level 2
    else if (c->bsplit && NM % 16 == 0) RET(maybe_vx3(c, Lt64));
    else RET(maybe_vx<zf>(c, Lt64, sizeof(zf)));
    if (c->mixed) {
      zf* Ld32 = nullptr;
      zf* Lt32 = nullptr;
      RET(dalloc(c, &Ld32, K * MM));
      RET(dalloc(c, &Lt32, 2 * K * NM));
      RET(launch(c, "convert", [&]() {
        if (c->bsplit) {
          // split-BF16 Ld is stored transposed, [M*M][K]: the vHS GEMM reads
          // it with k contiguous and stages it straight into LDS
          hipLaunchKernelGGL((ztranspose_kernel<zd, zb>), dim3((unsigned)((MM + 31) / 32), (unsigned)((K + 31) / 32)),
                             dim3(256), 0, st, Ld64, (zb*)Ld32, (int)K, (int)MM, MM, K);
          hipLaunchKernelGGL((zconvert_kernel<zd, zb>), dim3(nblk(2 * K * NM, 256)), dim3(256), 0, st, Lt64,
                             (zb*)Lt32, 2 * K * NM);
        } else {
          hipLaunchKernelGGL((zconvert_kernel<zd, zf>), dim3(nblk(K * MM, 256)), dim3(256), 0, st, Ld64, Ld32, K * MM);
          hipLaunchKernelGGL((zconvert_kernel<zd, zf>), dim3(nblk(2 * K * NM, 256)), dim3(256), 0, st, Lt64, Lt32,
                             2 * K * NM);
        }
      }));
      if ((c->three & 1) || c->coulomb3) {
        RET(dalloc(c, &c->tp_Lt, 2 * K * NM));
        RET(launch(c, "third_pieces", [&]() {
          hipLaunchKernelGGL(third_piece_kernel, dim3(nblk(2 * K * NM, 256)), dim3(256), 0, st, Lt64, (const zb*)Lt32,
                             c->tp_Lt, 2 * K * NM, 0L, 0L, 0);
        }));
      }
      if (c->three & 2) {
        // split Ld is [M*M][K], the transpose of the FP64 [K][M*M]
        RET(dalloc(c, &c->tp_Ld, K * MM));
        RET(launch(c, "third_pieces", [&]() {
          hipLaunchKernelGGL(third_piece_kernel, dim3(nblk(K * MM, 256)), dim3(256), 0, st, Ld64, (const zb*)Ld32,
                             c->tp_Ld, K * MM, K, MM, 1);
        }));
      }
      HIPCHK(c, hipStreamSynchronize(st));
      dfree(c, Ld64);
      if (c->energy_level >= 1) {
        c->Lt64 = Lt64;
        RET(dalloc(c, &c->Gtmp, 2 * NM * c->nw));
      } else {
        dfree(c, Lt64);
      }
      c->Ld = Ld32;
      c->Lt = Lt32;
    } else {
      c->Ld = Ld64;
      c->Lt = Lt64;
    }
  }
  HIPCHK(c, hipStreamSynchronize(st));
  c->have_system = true;
  return AFQMC_OK;
}

int afqmc_get_spvnT(afqmc_ctx* c, long long* nnz, int* rowptr, int* col, double* val)
{
  if (!c || !nnz) return AFQMC_ERR_ARG;
  if (c->mode != AFQMC_FP64_CSR || !c->spt_rp) return fail(c, AFQMC_ERR_STATE, "SpvnT exists in FP64_CSR mode only");
  *nnz = c->spt_nnz;
  if (!rowptr) return AFQMC_OK;
  HIPCHK(c, hipMemcpyAsync(rowptr, c->spt_rp, sizeof(int) * (c->K + 1), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(col, c->spt_col, sizeof(int) * c->spt_nnz, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(val, c->spt_val, sizeof(zd) * c->spt_nnz, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AFQMC_OK;
}

int afqmc_init_walkers(afqmc_ctx* c)
{
  if (!c) return AFQMC_ERR_ARG;
  if (!c->have_system) return fail(c, AFQMC_ERR_STATE, "upload a system first");
  const long MN = (long)c->M * c->N;
  c->t1inv_valid = false;
  c->g_fresh = false;
  long tot = std::max<long>((long)c->nw * 2 * MN, (long)c->nw * 8);
  return launch(c, "init_walkers", [&]() {
    hipLaunchKernelGGL(init_walkers_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, c->W, c->Wd,
                       c->Wl, c->trial, c->nw, MN);
  });
}

int afqmc_copy(afqmc_ctx* c, int buffer, int to_device, double* host)
{
  if (!c || !host) return AFQMC_ERR_ARG;
  const long M = c->M, N = c->N, K = c->K, nw = c->nw;
  const long MN = M * N, MM = M * M, NAK = 2 * MN;
  hipStream_t st = c->stream;
  auto raw = [&](void* dev, size_t bytes) -> int {
    if (to_device) HIPCHK(c, hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st));
    else HIPCHK(c, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    return AFQMC_OK;
  };
  // walker-last reference [R][nw] <-> walker-first device [nw][R] of type
  // zd/zf/zb, with the third pieces in dev3 when the buffer has them
  auto wlast = [&](void* dev, long R, bool is32, uint32_t* dev3 = nullptr) -> int {
    std::vector<zd> tmp((size_t)R * nw);
    std::vector<zf> t32;
    std::vector<uint32_t> t3(dev3 && c->bsplit ? tmp.size() : 0);
    if (to_device) {
      htranspose((const zd*)host, tmp.data(), R, nw);
      if (is32) {
        t32.resize(tmp.size());
        for (size_t i = 0; i < tmp.size(); i++) {
          if (c->bsplit) {
            uint32_t w[2];
            host_split(tmp[i].re, tmp[i].im, w);
            std::memcpy(&t32[i], w, sizeof(w));
            if (!t3.empty()) t3[i] = host_third(tmp[i].re, tmp[i].im, w);
          } else {
            t32[i] = zf{(float)tmp[i].re, (float)tmp[i].im};
          }
        }
        HIPCHK(c, hipMemcpyAsync(dev, t32.data(), sizeof(zf) * t32.size(), hipMemcpyHostToDevice, st));
        if (!t3.empty())
          HIPCHK(c, hipMemcpyAsync(dev3, t3.data(), sizeof(uint32_t) * t3.size(), hipMemcpyHostToDevice, st));
      } else {
        HIPCHK(c, hipMemcpyAsync(dev, tmp.data(), sizeof(zd) * tmp.size(), hipMemcpyHostToDevice, st));
      }
      HIPCHK(c, hipStreamSynchronize(st));
    } else {
      if (is32) {
        t32.resize(tmp.size());
        HIPCHK(c, hipMemcpyAsync(t32.data(), dev, sizeof(zf) * t32.size(), hipMemcpyDeviceToHost, st));
        if (!t3.empty())
          HIPCHK(c, hipMemcpyAsync(t3.data(), dev3, sizeof(uint32_t) * t3.size(), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        for (size_t i = 0; i < tmp.size(); i++) {
          if (c->bsplit) {
            uint32_t w[2];
            std::memcpy(w, &t32[i], sizeof(w));
            tmp[i] = host_merge(w);
            if (!t3.empty()) {
              tmp[i].re += (double)host_bf2f((uint16_t)(t3[i] & 0xffff));
              tmp[i].im += (double)host_bf2f((uint16_t)(t3[i] >> 16));
            }
          } else {
            tmp[i] = zd{t32[i].re, t32[i].im};
          }
        }
      } else {
        HIPCHK(c, hipMemcpyAsync(tmp.data(), dev, sizeof(zd) * tmp.size(), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
      }
      htranspose(tmp.data(), (zd*)host, nw, R);
    }
    return AFQMC_OK;
  };
  const bool d = dense_layout(c);
  if (to_device && buffer == AFQMC_BUF_W) c->t1inv_valid = false;
  if (to_device && (buffer == AFQMC_BUF_W || buffer == AFQMC_BUF_G)) c->g_fresh = false;
  switch (buffer) {
  case AFQMC_BUF_W: return raw(c->W, sizeof(zd) * nw * 2 * MN);
  case AFQMC_BUF_WDATA:
    RET(raw(c->Wd, sizeof(zd) * nw * 8));
    if (to_device)  // the log-overlaps follow the overlaps written
      RET(launch(c, "wdata_logs", [&]() {
        hipLaunchKernelGGL(wd_logs_kernel, dim3((unsigned)((4 * nw + 255) / 256)), dim3(256), 0, st, c->Wd, c->Wl,
                           (int)nw);
      }));
    return AFQMC_OK;
  case AFQMC_BUF_HYBRIDW: return raw(c->hybridW, sizeof(zd) * nw);
  case AFQMC_BUF_G: return d ? wlast(c->G, NAK, c->mixed, c->tp_G) : raw(c->G, sizeof(zd) * NAK * nw);
  case AFQMC_BUF_VBIAS: return d ? wlast(c->vbias, K, false) : raw(c->vbias, sizeof(zd) * K * nw);
  case AFQMC_BUF_X: return d ? wlast(c->X, K, c->mixed, c->tp_X) : raw(c->X, sizeof(zd) * K * nw);
  case AFQMC_BUF_VHS:
    if (d) {
      // an upload writes the third pieces; a read adds them only when the
      // last HS potential wrote them
      RET(wlast(c->vHS, MM, c->mixed, (to_device || c->vhs3) ? c->tp_vHS : nullptr));
      if (to_device) c->vhs3 = c->tp_vHS != nullptr;
      return AFQMC_OK;
    }
    RET(raw(c->vHS, sizeof(zd) * MM * nw));
    if (to_device) {
      dim3 grid((nw + 31) / 32, (unsigned)((MM + 31) / 32));
      RET(launch(c, "vhs_transpose", [&]() {
        hipLaunchKernelGGL((ztranspose_kernel<zd, zd>), grid, dim3(256), 0, st, (const zd*)c->vHS, c->vHSw, (int)MM,
                           (int)nw, nw, MM);
      }));
    }
    return AFQMC_OK;
  case AFQMC_BUF_GFULL:
    if (!c->Gfull) return fail(c, AFQMC_ERR_STATE, "no full density matrix (non-transposed bias only)");
    return raw(c->Gfull, sizeof(zd) * 2 * MM * nw);
  case AFQMC_BUF_EPARTS: {
    if (to_device) return fail(c, AFQMC_ERR_ARG, "the energy terms are read-only");
    std::vector<zd> e1(nw), ed(nw), ex(nw);
    HIPCHK(c, hipMemcpyAsync(e1.data(), c->e_one, sizeof(zd) * nw, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(ed.data(), c->e_direct, sizeof(zd) * nw, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(ex.data(), c->e_exch, sizeof(zd) * nw, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    zd* o = (zd*)host;
    for (long w = 0; w < nw; w++) {
      o[3 * w] = e1[w];
      o[3 * w + 1] = zd{0.5 * ed[w].re, 0.5 * ed[w].im};
      o[3 * w + 2] = zd{-0.5 * ex[w].re, -0.5 * ex[w].im};
    }
    return AFQMC_OK;
  }
  default: return fail(c, AFQMC_ERR_ARG, "unknown buffer");
  }
}

#define NEED_SYSTEM(c)                                                                 \
  do {                                                                                 \
    if (!(c)) return AFQMC_ERR_ARG;                                                    \
    if (!(c)->have_system) return fail((c), AFQMC_ERR_STATE, "upload a system first"); \
  } while (0)

int afqmc_mixed_density_matrix(afqmc_ctx* c)
{
  NEED_SYSTEM(c);
  RET(do_mdm(c));
  return check_flags(c);
}
int afqmc_mixed_density_matrix_full(afqmc_ctx* c)
{
  NEED_SYSTEM(c);
  if (c->transposed) return fail(c, AFQMC_ERR_STATE, "set AFQMC_OPT_TRANSPOSED to 0 before the upload");
  return do_mdm_full(c);
}

int afqmc_vbias_full(afqmc_ctx* c)
{
  NEED_SYSTEM(c);
  if (c->transposed) return fail(c, AFQMC_ERR_STATE, "set AFQMC_OPT_TRANSPOSED to 0 before the upload");
  if (!c->Gfull) return fail(c, AFQMC_ERR_STATE, "form the full density matrix first");
  return do_vbias_full(c);
}

int afqmc_vbias(afqmc_ctx* c)
{
  NEED_SYSTEM(c);
  return do_vbias(c);
}
int afqmc_fields(afqmc_ctx* c, const double* x, long substep)
{
  NEED_SYSTEM(c);
  const long KW = (long)c->K * c->nw;
  const double* xdev = nullptr;
  double* tmp = nullptr;
  if (x) {
    RET(dalloc(c, &tmp, KW));
    HIPCHK(c, hipMemcpyAsync(tmp, x, sizeof(double) * KW, hipMemcpyHostToDevice, c->stream));
    xdev = tmp;
  } else {
    if (!have_fields(c, substep)) return fail(c, AFQMC_ERR_STATE, "no staged fields for substep");
    xdev = stage_at(c, substep);
  }
  int r = do_fields(c, xdev, substep);
  if (r == AFQMC_OK) r = check_flags(c);
  if (tmp) {
    (void)hipStreamSynchronize(c->stream);
    dfree(c, tmp);
  }
  return r;
}
int afqmc_vhs(afqmc_ctx* c)
{
  NEED_SYSTEM(c);
  return do_vhs(c);
}
int afqmc_propagate(afqmc_ctx* c)
{
  NEED_SYSTEM(c);
  return do_propagate(c);
}
int afqmc_save_overlaps(afqmc_ctx* c)
{
  NEED_SYSTEM(c);
  return do_save_overlaps(c);
}
int afqmc_calculate_overlaps(afqmc_ctx* c)
{
  NEED_SYSTEM(c);
  RET(do_overlaps(c));
  return check_flags(c);
}
int afqmc_weights(afqmc_ctx* c, double dt, double* eshift_inout)
{
  NEED_SYSTEM(c);
  if (eshift_inout) HIPCHK(c, hipMemcpyAsync(c->eshift, eshift_inout, sizeof(double), hipMemcpyHostToDevice, c->stream));
  RET(do_weights(c, nullptr, dt));
  if (eshift_inout) {
    HIPCHK(c, hipMemcpyAsync(eshift_inout, c->eshift, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return AFQMC_OK;
}
int afqmc_branch(afqmc_ctx* c, double u, int* parent_out)
{
  NEED_SYSTEM(c);
  return do_branch(c, u, parent_out);
}
int afqmc_orthogonalize(afqmc_ctx* c)
{
  NEED_SYSTEM(c);
  return do_ortho(c);
}
int afqmc_energy(afqmc_ctx* c, double* eav_out)
{
  NEED_SYSTEM(c);
  return do_energy(c, eav_out);
}

int afqmc_stage_fields(afqmc_ctx* c, const double* x_all, long nsub)
{
  if (!c || !x_all || nsub <= 0) return AFQMC_ERR_ARG;
  const long KW = (long)c->K * c->nw;
  if (c->xstage) {
    (void)hipStreamSynchronize(c->stream);
    dfree(c, c->xstage);
    c->xstage = nullptr;
    c->nstage = 0;
  }
  drop_graphs(c);  // the captured substeps read the old block
  RET(dalloc(c, &c->xstage, KW * nsub));
  HIPCHK(c, hipMemcpyAsync(c->xstage, x_all, sizeof(double) * KW * nsub, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->nstage = nsub;
  return AFQMC_OK;
}

int afqmc_device_fields(afqmc_ctx* c, uint64_t seed, int enable)
{
  if (!c) return AFQMC_ERR_ARG;
  drop_graphs(c);  // captured substeps read the fields from the old source
  c->dev_fields = enable != 0;
  c->field_seed = seed;
  return AFQMC_OK;
}

int afqmc_substep(afqmc_ctx* c, long substep, int do_ortho_, double dt)
{
  NEED_SYSTEM(c);
  if (!have_fields(c, substep)) return fail(c, AFQMC_ERR_STATE, "no staged fields for substep");
  const int o = do_ortho_ ? 1 : 0;
  if (!c->use_graphs || c->prof || c->tuning) return substep_body(c, stage_at(c, substep), substep, do_ortho_, dt);
  if (c->g_dt != dt) {
    drop_graphs(c);
    c->g_dt = dt;
  }
  if (!c->gexec[o]) RET(capture_substep(c, o, dt));
  // the graph reuses the inverses of the current walkers (see capture_substep)
  if (!c->t1inv_valid) RET(do_overlaps(c));
  RET(launch(c, "set_substep", [&]() {
    hipLaunchKernelGGL(set_long_kernel, dim3(1), dim3(64), 0, c->stream, c->d_sub, substep);
  }));
  HIPCHK(c, hipGraphLaunch(c->gexec[o], c->stream));
  c->t1inv_valid = true;  // the substep ends with the overlaps of the new walkers
  c->g_fresh = false;
  return AFQMC_OK;
}

int afqmc_get_eshift(afqmc_ctx* c, double* e)
{
  if (!c || !e) return AFQMC_ERR_ARG;
  HIPCHK(c, hipMemcpyAsync(e, c->eshift, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AFQMC_OK;
}
int afqmc_set_eshift(afqmc_ctx* c, double e)
{
  if (!c) return AFQMC_ERR_ARG;
  HIPCHK(c, hipMemcpyAsync(c->eshift, &e, sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return AFQMC_OK;
}

int afqmc_run(afqmc_ctx* c, long first, int nsteps, int nsubsteps, int northo, double dt, double* eav_out)
{
  NEED_SYSTEM(c);
  if (nsteps < 0 || nsubsteps < 0 || northo <= 0) return AFQMC_ERR_ARG;
  if (first < 0 || (!c->dev_fields && (!c->xstage || first + (long)nsteps * nsubsteps > c->nstage)))
    return fail(c, AFQMC_ERR_STATE, "not enough staged fields");
  long tot = first;
  for (int step = 0; step < nsteps; step++) {
    for (int sub = 0; sub < nsubsteps; sub++, tot++) {
      int ortho = (tot > 0 && tot % northo == 0) ? 1 : 0;
      RET(afqmc_substep(c, tot, ortho, dt));
    }
    if (eav_out) {
      RET(do_mdm(c));
      RET(do_energy(c, &eav_out[step]));
    }
  }
  return check_flags(c);
}

int afqmc_tune(afqmc_ctx* c, double dt)
{
  NEED_SYSTEM(c);
  if (!have_fields(c, 0)) return fail(c, AFQMC_ERR_STATE, "stage fields before tuning");
  // tune on walkers at the trial (a representative, well-conditioned state)
  drop_graphs(c);
  int r = afqmc_init_walkers(c);
  if (r != AFQMC_OK) return r;
  c->tuning = true;
  r = substep_body(c, stage_at(c, 0), 0, 1, dt);
  if (r == AFQMC_OK) r = do_mdm(c);
  if (r == AFQMC_OK) r = do_energy(c, nullptr);
  c->tuning = false;
  c->t1inv_valid = false;
  c->g_fresh = false;
  if (std::getenv("AFQMC_VERBOSE"))
    for (auto& kv : c->tune) std::fprintf(stderr, "[afqmc] tile %-48s -> candidate %d\n", kv.first.c_str(), kv.second);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemsetAsync(c->flags, 0, 4 * sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(c->eshift, 0, sizeof(double), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return r;
}

int afqmc_use_graphs(afqmc_ctx* c, int enable)
{
  if (!c) return AFQMC_ERR_ARG;
  drop_graphs(c);
  c->use_graphs = enable != 0;
  return AFQMC_OK;
}

int afqmc_profile(afqmc_ctx* c, int enable)
{
  if (!c) return AFQMC_ERR_ARG;
  collect_events(c);
  c->prof = enable != 0;
  if (enable) c->stats.clear();
  return AFQMC_OK;
}

int afqmc_kernel_stats(afqmc_ctx* c, int cap, char (*names)[64], long long* launches, double* total_ms, int* count)
{
  if (!c || !count) return AFQMC_ERR_ARG;
  collect_events(c);
  int i = 0;
  for (auto& kv : c->stats) {
    if (i < cap) {
      if (names) {
        std::strncpy(names[i], kv.first.c_str(), 63);
        names[i][63] = 0;
      }
      if (launches) launches[i] = kv.second.n;
      if (total_ms) total_ms[i] = kv.second.ms;
    }
    i++;
  }
  *count = i;
  return AFQMC_OK;
}

} // extern "C"

namespace {
template <int CP, bool AK, bool BK> int zgemm_host(afqmc_ctx* c, ZGemmArgs a)
{
  return zgemm<CP, double, double, double, AK, BK>(c, "zgemm", a, 1);
}
template <int CP> int zgemm_ops(afqmc_ctx* c, ZGemmArgs a, bool ak, bool bk)
{
  if (ak && bk) return zgemm_host<CP, true, true>(c, a);
  if (ak && !bk) return zgemm_host<CP, true, false>(c, a);
  if (!ak && bk) return zgemm_host<CP, false, true>(c, a);
  return zgemm_host<CP, false, false>(c, a);
}
} // namespace

extern "C" int afqmc_zgemm(int precision, char opA, char opB, int m, int n, int k, const double* alpha,
                           const double* A, int lda, const double* B, int ldb, const double* beta, double* Cm,
                           int ldc)
{
  if (!alpha || !beta || !A || !B || !Cm || m < 0 || n < 0 || k < 0) return AFQMC_ERR_ARG;
  const int tile = (precision >> 8) & 0xff;
  precision &= 0xff;
  if ((opA != 'N' && opA != 'T') || (opB != 'N' && opB != 'T' && opB != 'C' && opB != 'R') || precision < 0 ||
      precision > 2)
    return AFQMC_ERR_ARG;
  // opB 'C': B^H (B stored n x k), 'R': conj(B) (B stored k x n)
  const bool bn = (opB == 'N' || opB == 'R');
  const long ar = (opA == 'N') ? m : k, ac = (opA == 'N') ? k : m;
  const long br = bn ? k : n, bc = bn ? n : k;
  if (lda < ac || ldb < bc || ldc < n) return AFQMC_ERR_SHAPE;
  if (m == 0 || n == 0) return AFQMC_OK;
  afqmc_ctx tmp;
  afqmc_ctx* c = &tmp;
  c->force_cfg = tile;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return AFQMC_ERR_HIP;
  zd *dA = nullptr, *dB = nullptr, *dC = nullptr;
  int r = AFQMC_OK;
  auto run = [&]() -> int {
    RET(dalloc(c, &dA, ar * lda));
    RET(dalloc(c, &dB, br * ldb));
    RET(dalloc(c, &dC, (long)m * ldc));
    HIPCHK(c, hipMemcpyAsync(dA, A, sizeof(zd) * ar * lda, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dB, B, sizeof(zd) * br * ldb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dC, Cm, sizeof(zd) * m * ldc, hipMemcpyHostToDevice, c->stream));
    ZGemmArgs a = gargs();
    a.A = dA;
    a.B = dB;
    a.C = dC;
    const bool ak = (opA == 'N'), bk = !bn;
    a.conjb = (opB == 'C' || opB == 'R') ? 1 : 0;
    a.sam = ak ? lda : 1;
    a.sak = ak ? 1 : lda;
    a.sbn = bk ? ldb : 1;
    a.sbk = bk ? 1 : ldb;
    a.scm = ldc;
    a.scn = 1;
    a.m = m;
    a.n = n;
    a.k = k;
    a.ar = alpha[0];
    a.ai = alpha[1];
    a.br = beta[0];
    a.bi = beta[1];
    if (k == 0) a.ar = a.ai = 0.0;
    if (precision == 0) RET(zgemm_ops<CP_F64>(c, a, ak, bk));
    else if (precision == 1) RET(zgemm_ops<CP_F32>(c, a, ak, bk));
    else RET(zgemm_ops<CP_BF16X3>(c, a, ak, bk));
    HIPCHK(c, hipMemcpyAsync(Cm, dC, sizeof(zd) * m * ldc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AFQMC_OK;
  };
  r = run();
  (void)hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
  return r;
}

// ---------------------------------------------------------------------------
// Allocation-free mixed density matrix.  The overlap matrices (the only
// memory the three kernels need beyond the caller's inputs and outputs) live
// in a file-scope device array, one copy per GPU, held until the process
// ends.  Host code never passes the array's name to a kernel: its device
// address is looked up once per GPU with hipGetSymbolAddress (with that GPU
// current) and kept per device.
namespace {
static __device__ __align__(256) unsigned char scratch[AFQMC_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of this GPU's scratch array (nullptr on failure)
void* scratch_address(int dev)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}
} // namespace

extern "C" int afqmc_mixed_density_matrix_noalloc(int nmo, int naea, int nwalk, const double* trialwfn,
                                                  const double* W, double* Gc, double* ovlp, void* stream)
{
  if (!trialwfn || !W || !Gc || !ovlp || nmo <= 0 || naea <= 0 || nwalk <= 0) return AFQMC_ERR_ARG;
  if (naea > nmo) return AFQMC_ERR_SHAPE;
  if (naea > 128) return AFQMC_ERR_UNSUPPORTED;
  const long M = nmo, N = naea, nw = nwalk;
  const size_t need = sizeof(zd) * 2 * (size_t)nw * N * N;
  if (need > (size_t)AFQMC_NOALLOC_SCRATCH_BYTES) return AFQMC_ERR_SHAPE;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return AFQMC_ERR_HIP;
  zd* T1 = (zd*)scratch_address(dev);
  if (!T1) return AFQMC_ERR_HIP;
  // a context on the stack only carries the stream and sizes into the
  // launchers below (no member of it allocates)
  afqmc_ctx tmp;
  afqmc_ctx* c = &tmp;
  c->stream = (hipStream_t)stream;
  c->M = nmo;
  c->N = naea;
  c->nw = nwalk;
  // T1[w][s] = W[w][s]^T trialwfn_s; every entry is written (beta = 0)
  ZGemmArgs a = gargs();
  a.A = W;
  a.sam = 1;
  a.sak = N;
  a.sa1 = 2 * M * N;
  a.sa2 = M * N;
  a.B = trialwfn;
  a.sbn = 1;
  a.sbk = N;
  a.sb1 = 0;
  a.sb2 = M * N;
  a.C = T1;
  a.scm = N;
  a.scn = 1;
  a.sc1 = 2 * N * N;
  a.sc2 = N * N;
  a.bdiv = 2;
  a.m = naea;
  a.n = naea;
  a.k = nmo;
  int r = zgemm_run<CP_F64, double, double, double, false, false>(c, "ovlp_gemm", a, 2 * nwalk,
                                                                   default_cfg<CP_F64>(a));
  // in-place inverse and determinant (no status flag: nothing is read back)
  if (r == AFQMC_OK) r = gj_inverse_into(c, T1, T1, (zd*)ovlp, 2, 1, nullptr, nullptr, nullptr, nullptr);
  if (r == AFQMC_OK) {
    // Gc[(s N + a) M + k][w] = sum_b T1inv(a, b) W(k, b)
    ZGemmArgs g = gargs();
    g.A = T1;
    g.sam = N;
    g.sak = 1;
    g.sa1 = 2 * N * N;
    g.sa2 = N * N;
    g.B = W;
    g.sbn = N;
    g.sbk = 1;
    g.sb1 = 2 * M * N;
    g.sb2 = M * N;
    g.bdiv = 2;
    g.m = naea;
    g.n = nmo;
    g.k = naea;
    g.C = Gc;
    g.scm = M * nw;
    g.scn = nw;
    g.sc1 = 1;
    g.sc2 = N * M * nw;
    r = zgemm_run<CP_F64, double, double, double, true, true>(c, "gc_gemm", g, 2 * nwalk, default_cfg<CP_F64>(g));
  }
  c->stream = nullptr;
  return r;
}

// ---------------------------------------------------------------------------
// Several contexts advanced together: one HIP graph holds a substep of every
// member, each on its member's own stream (forked from and joined to the
// group's stream inside the capture), so one graph launch advances all the
// systems and their kernels run side by side.
struct afqmc_group {
  std::vector<afqmc_ctx*> ctx;
  std::vector<long> gen[2];  // members' graph generations at capture
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, done = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  double g_dt = 0.0;
  std::string err;
};

namespace {

int gfail(afqmc_group* g, int code, const std::string& m)
{
  g->err = m;
  return code;
}

void group_drop(afqmc_group* g)
{
  for (int i = 0; i < 2; i++) {
    if (g->gexec[i]) (void)hipGraphExecDestroy(g->gexec[i]);
    if (g->graph[i]) (void)hipGraphDestroy(g->graph[i]);
    g->gexec[i] = nullptr;
    g->graph[i] = nullptr;
    g->gen[i].clear();
  }
}

int group_capture(afqmc_group* g, int ortho, double dt)
{
  for (afqmc_ctx* c : g->ctx) {
    const int r = capture_prep(c, ortho);
    if (r != AFQMC_OK) return gfail(g, r, std::string("member: ") + c->err);
  }
  if (hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return gfail(g, AFQMC_ERR_HIP, "group capture did not start");
  int r = hipEventRecord(g->fork, g->stream) == hipSuccess ? AFQMC_OK : AFQMC_ERR_HIP;
  for (size_t i = 0; i < g->ctx.size() && r == AFQMC_OK; i++) {
    afqmc_ctx* c = g->ctx[i];
    if (hipStreamWaitEvent(c->stream, g->fork, 0) != hipSuccess) {
      r = AFQMC_ERR_HIP;
      break;
    }
    r = substep_body(c, c->dev_fields ? nullptr : c->xstage, 0, ortho, dt, c->d_sub);
    if (r != AFQMC_OK) g->err = std::string("member: ") + c->err;
    if (r == AFQMC_OK && (hipEventRecord(g->join[i], c->stream) != hipSuccess ||
                          hipStreamWaitEvent(g->stream, g->join[i], 0) != hipSuccess))
      r = AFQMC_ERR_HIP;
  }
  hipGraph_t gr = nullptr;
  const hipError_t e = hipStreamEndCapture(g->stream, &gr);
  if (r != AFQMC_OK || e != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    return r != AFQMC_OK ? r : gfail(g, AFQMC_ERR_HIP, std::string("group capture: ") + hipGetErrorString(e));
  }
  g->graph[ortho] = gr;
  if (hipGraphInstantiate(&g->gexec[ortho], gr, nullptr, nullptr, 0) != hipSuccess)
    return gfail(g, AFQMC_ERR_HIP, "group graph instantiation failed");
  g->gen[ortho].clear();
  for (afqmc_ctx* c : g->ctx) g->gen[ortho].push_back(c->graph_gen);
  return AFQMC_OK;
}

} // namespace

extern "C" {

int afqmc_group_create(afqmc_ctx* const* ctxs, int n, afqmc_group** out)
{
  if (!out) return AFQMC_ERR_ARG;
  *out = nullptr;
  if (!ctxs || n <= 0) return AFQMC_ERR_ARG;
  for (int i = 0; i < n; i++) {
    if (!ctxs[i] || !ctxs[i]->have_system) return AFQMC_ERR_STATE;
    for (int j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) return AFQMC_ERR_ARG;
  }
  afqmc_group* g = new afqmc_group;
  g->ctx.assign(ctxs, ctxs + n);
  g->join.assign(n, nullptr);
  bool ok = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&g->done, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; i < n && ok; i++) ok = hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    afqmc_group_destroy(g);
    return AFQMC_ERR_HIP;
  }
  *out = g;
  return AFQMC_OK;
}

int afqmc_group_destroy(afqmc_group* g)
{
  if (!g) return AFQMC_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  group_drop(g);
  for (hipEvent_t e : g->join)
    if (e) (void)hipEventDestroy(e);
  if (g->fork) (void)hipEventDestroy(g->fork);
  if (g->done) (void)hipEventDestroy(g->done);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return AFQMC_OK;
}

const char* afqmc_group_error(const afqmc_group* g) { return g ? g->err.c_str() : "null group"; }

int afqmc_group_substep(afqmc_group* g, long substep, int do_ortho, double dt)
{
  if (!g) return AFQMC_ERR_ARG;
  const int o = do_ortho ? 1 : 0;
  for (afqmc_ctx* c : g->ctx)
    if (!have_fields(c, substep))
      return gfail(g, AFQMC_ERR_STATE, "a member has no staged fields for this substep");
  // profiling and tuning run the members' substeps eagerly, one by one
  bool eager = false;
  for (afqmc_ctx* c : g->ctx) eager |= c->prof || c->tuning;
  if (eager) {
    for (afqmc_ctx* c : g->ctx) {
      const int r = afqmc_substep(c, substep, do_ortho, dt);
      if (r != AFQMC_OK) return gfail(g, r, std::string("member: ") + c->err);
    }
    return AFQMC_OK;
  }
  if (g->g_dt != dt) {
    group_drop(g);
    g->g_dt = dt;
  }
  bool stale = !g->gexec[o];
  for (size_t i = 0; i < g->ctx.size() && !stale; i++) stale = g->gen[o][i] != g->ctx[i]->graph_gen;
  if (stale) {
    if (g->gexec[o]) (void)hipGraphExecDestroy(g->gexec[o]);
    if (g->graph[o]) (void)hipGraphDestroy(g->graph[o]);
    g->gexec[o] = nullptr;
    g->graph[o] = nullptr;
    const int r = group_capture(g, o, dt);
    if (r != AFQMC_OK) return r;
  }
  // the graph starts from the inverses of the members' current walkers, and
  // from work already queued on the members' streams
  for (size_t i = 0; i < g->ctx.size(); i++) {
    afqmc_ctx* c = g->ctx[i];
    if (!c->t1inv_valid) {
      const int r = do_overlaps(c);
      if (r != AFQMC_OK) return gfail(g, r, std::string("member: ") + c->err);
    }
    if (hipEventRecord(g->join[i], c->stream) != hipSuccess || hipStreamWaitEvent(g->stream, g->join[i], 0) != hipSuccess)
      return gfail(g, AFQMC_ERR_HIP, "ordering the members' streams");
    hipLaunchKernelGGL(set_long_kernel, dim3(1), dim3(64), 0, g->stream, c->d_sub, substep);
  }
  if (hipGraphLaunch(g->gexec[o], g->stream) != hipSuccess || hipEventRecord(g->done, g->stream) != hipSuccess)
    return gfail(g, AFQMC_ERR_HIP, "group graph launch failed");
  // later work on a member's own stream waits for the group's substep
  for (afqmc_ctx* c : g->ctx) {
    if (hipStreamWaitEvent(c->stream, g->done, 0) != hipSuccess) return gfail(g, AFQMC_ERR_HIP, "ordering the members");
    c->t1inv_valid = true;
    c->g_fresh = false;
  }
  return AFQMC_OK;
}

int afqmc_group_sync(afqmc_group* g)
{
  if (!g) return AFQMC_ERR_ARG;
  if (hipStreamSynchronize(g->stream) != hipSuccess) return gfail(g, AFQMC_ERR_HIP, "group stream");
  for (afqmc_ctx* c : g->ctx)
    if (hipStreamSynchronize(c->stream) != hipSuccess) return gfail(g, AFQMC_ERR_HIP, "member stream");
  return AFQMC_OK;
}

} // extern "C"
