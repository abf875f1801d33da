// miniafqmc -- the miniapp driver (src/miniapps/miniafqmc_base.cpp:93-360)
// restated over the C-ABI of libafqmc_hip.so.
//
// Same flags and loop as the reference (-i -s -w -o -f -t -v -h), with the
// time step, precision mode and synthetic system as extra options since the
// reference's input file is not shipped and dt is hard-coded there.  The
// auxiliary fields come from the counter-based generator (seeded), staged
// on the device one block at a time.
#include "afqmc_h5.h"
#include "afqmc_hip.h"

#include <dlfcn.h>
#include <getopt.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace {

// The reference driver's timers (miniafqmc_base.cpp:50-77) and the library
// call sites whose GPU time each one covers, for the -v report.  Launches at
// upload fall under none.
struct TimerSites {
  const char* timer;
  const char* sites[17];
};
const TimerSites kTimers[] = {
    {"compact Mixed Density Matrix", {"gc_gemm", "det_restore"}},
    {"Mixed Density Matrix", {"gfull_gemm"}},
    {"Bias Potential", {"vbias_gemm", "vbias_reduce", "csrmm_vbias", "csrmm_vbias_t"}},
    {"H-S Potential", {"vhs_gemm", "csrmm_vhs", "vhs_transpose"}},
    {"Sigma", {"fields", "set_substep"}},
    {"Propagation", {"prop_pw", "prop_taylor", "prop_ps", "prop_split_w", "prop_fused"}},
    {"Other", {"save_overlaps", "weights", "branch_select", "branch_gather"}},
    {"Overlap", {"ovlp_gemm", "gj_inverse"}},
    {"Orthgonalization",
     {"ortho_gram", "ortho_chol_inv", "ortho_sign", "ortho_apply", "ortho_flags", "lq_orthogonalize", "lq_fallback"}},
    {"Local Energy",
     {"energy_one", "energy_two", "energy_direct", "energy_exch_dot", "energy_pairsum", "energy_assemble",
      "energy_zero", "energy_trace", "energy_v_gemm", "energy_v_reduce", "energy_vx_gemm", "energy_y_gemm", "csrmm_vakbl", "g_transpose",
      "gc64_gemm", "convert", "g_split3"}},
};

void print_help()
{
  std::printf("miniafqmc - AFQMC walker step on one MI355X\n\n");
  std::printf("Options:\n");
  std::printf("-i                Number of MC steps (default: 10)\n");
  std::printf("-s                Number of substeps (default: 10)\n");
  std::printf("-w                Number of walkers (default: 16)\n");
  std::printf("-o                Number of substeps between orthogonalization (default: 10)\n");
  std::printf("-f                Input file name (afqmc.h5 format; read through libafqmc_h5.so)\n");
  std::printf("-t                yes|no: half-rotated transposed Cholesky matrix for the bias potential\n");
  std::printf("                  ('no': full G and T(Spvn), fp64_csr mode only)\n");
  std::printf("-v                Verbose output (GPU time under the reference's timer names, then per kernel)\n");
  std::printf("--dt D            time step (default 0.005)\n");
  std::printf("--mode M          fp64_csr | fp64_dense | mixed_fp32 | mixed_bf16 (default fp64_csr)\n");
  std::printf("--system S        N16_M64 | N64_M256 | N128_M512 | MxNxK (synthetic, when no -f; default N16_M64)\n");
  std::printf("--density R       kept fraction of Cholesky entries for MxNxK systems (default 0.3)\n");
  std::printf("--seed S          seed of the synthetic system and fields (default 20262202)\n");
  std::printf("--tune            autotune GEMM tiles before the run\n");
  std::printf("--branch          comb population control after each step's energy\n");
}

struct Preset {
  const char* name;
  int M, N, K;
  double density, sigma;
};
const Preset kPresets[] = {
    {"N16_M64", 64, 16, 256, 0.3, 0.02},
    {"N64_M256", 256, 64, 1024, 0.2, 0.0},
    {"N128_M512", 512, 128, 2048, 1.0, 0.0},
};

// sigma of the larger systems: M*K*E|L|^2 matches the 64-orbital case,
// E|L|^2 = 2 sigma^2 rho (1 - ln rho)
double sigma_for(int M, int K, double rho)
{
  const double e0 = 2 * 0.02 * 0.02 * 0.3 * (1 - std::log(0.3));
  const double e = 64.0 * 256.0 * e0 / ((double)M * K);
  return std::sqrt(e / (2 * rho * (1 - std::log(rho))));
}

int mode_of(const std::string& s)
{
  if (s == "fp64_csr") return AFQMC_FP64_CSR;
  if (s == "fp64_dense") return AFQMC_FP64_DENSE;
  if (s == "mixed_fp32") return AFQMC_MIXED_FP32;
  if (s == "mixed_bf16") return AFQMC_MIXED_BF16;
  return -1;
}

#define CHECK(call)                                                                          \
  do {                                                                                       \
    int r_ = (call);                                                                         \
    if (r_ != AFQMC_OK) {                                                                    \
      std::fprintf(stderr, "error %d in %s: %s\n", r_, #call, ctx ? afqmc_ctx_error(ctx) : ""); \
      return 1;                                                                              \
    }                                                                                        \
  } while (0)

// The HDF5 reader lives in libafqmc_h5.so (host code over the HDF5 C
// library), loaded only when an input file is given so that the driver runs
// where HDF5 is absent.
struct H5Api {
  void* so = nullptr;
  decltype(&afqmc_h5_read) read = nullptr;
  decltype(&afqmc_h5_get) get = nullptr;
  decltype(&afqmc_h5_destroy) destroy = nullptr;
  decltype(&afqmc_h5_error) error = nullptr;
};

bool load_h5(H5Api& api)
{
  char exe[4096];
  const ssize_t n = readlink("/proc/self/exe", exe, sizeof(exe) - 1);
  std::string dir = ".";
  if (n > 0) {
    exe[n] = 0;
    dir = std::string(exe);
    dir = dir.substr(0, dir.find_last_of('/'));
  }
  for (const std::string& cand : {dir + "/../libafqmc_h5.so", std::string("libafqmc_h5.so")}) {
    api.so = dlopen(cand.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (api.so) break;
  }
  if (!api.so) {
    std::fprintf(stderr, "Error: cannot load libafqmc_h5.so (%s); it is built where the HDF5 C library is "
                         "installed\n", dlerror());
    return false;
  }
  api.read = (decltype(api.read))dlsym(api.so, "afqmc_h5_read");
  api.get = (decltype(api.get))dlsym(api.so, "afqmc_h5_get");
  api.destroy = (decltype(api.destroy))dlsym(api.so, "afqmc_h5_destroy");
  api.error = (decltype(api.error))dlsym(api.so, "afqmc_h5_error");
  return api.read && api.get && api.destroy && api.error;
}

} // namespace

int main(int argc, char** argv)
{
  int nsteps = 10, nsubsteps = 10, nwalk = 16, northo = 10;
  double dt = 0.005;
  bool verbose = false, tune = false, transposed = true, branch = false;
  std::string mode_s = "fp64_csr", system_s = "N16_M64", init_file;
  double density = 0.3;
  uint64_t seed = 20262202;

  static option longopts[] = {{"dt", required_argument, nullptr, 1},     {"mode", required_argument, nullptr, 2},
                              {"system", required_argument, nullptr, 3}, {"density", required_argument, nullptr, 4},
                              {"seed", required_argument, nullptr, 5},   {"tune", no_argument, nullptr, 6},
                              {"branch", no_argument, nullptr, 7},       {nullptr, 0, nullptr, 0}};
  int opt;
  while ((opt = getopt_long(argc, argv, "t:hvi:s:w:o:f:", longopts, nullptr)) != -1) {
    switch (opt) {
    case 'h': print_help(); return 1;
    case 'i': nsteps = std::atoi(optarg); break;
    case 's': nsubsteps = std::atoi(optarg); break;
    case 'w': nwalk = std::atoi(optarg); break;
    case 'o': northo = std::atoi(optarg); break;
    case 't': transposed = (std::string(optarg) != "no"); break;
    case 'f': init_file = optarg; break;
    case 'v': verbose = true; break;
    case 1: dt = std::atof(optarg); break;
    case 2: mode_s = optarg; break;
    case 3: system_s = optarg; break;
    case 4: density = std::atof(optarg); break;
    case 5: seed = std::strtoull(optarg, nullptr, 10); break;
    case 6: tune = true; break;
    case 7: branch = true; break;
    default: print_help(); return 1;
    }
  }
  afqmc_ctx* ctx = nullptr;
  const int mode = mode_of(mode_s);
  if (mode < 0 || nsteps < 0 || nsubsteps <= 0 || nwalk <= 0 || northo <= 0 || dt <= 0) {
    print_help();
    return 1;
  }

  afqmc_synth_view v{};
  afqmc_synth* syn = nullptr;
  afqmc_h5_system* h5s = nullptr;
  H5Api h5;
  if (!init_file.empty()) {
    // afqmc::Initialize (src/Matrix/initialize_serial.hpp:35-228)
    std::printf("***********************************************************\n");
    std::printf("                 Initializing from HDF5                    \n");
    std::printf("***********************************************************\n");
    if (!load_h5(h5)) return 1;
    if (h5.read(init_file.c_str(), dt, &h5s) != AFQMC_OK) {
      std::fprintf(stderr, "Error reading %s: %s\n", init_file.c_str(), h5.error());
      return 1;
    }
    h5.get(h5s, &v);
  } else {
    afqmc_synth_params sp{};
    bool found = false;
    for (const auto& p : kPresets)
      if (system_s == p.name) {
        sp.nmo = p.M;
        sp.naea = p.N;
        sp.nchol = p.K;
        sp.density = p.density;
        sp.sigma_L = p.sigma > 0 ? p.sigma : sigma_for(p.M, p.K, p.density);
        found = true;
      }
    if (!found) {
      int M = 0, N = 0, K = 0;
      if (std::sscanf(system_s.c_str(), "%dx%dx%d", &M, &N, &K) != 3) {
        std::fprintf(stderr, "unknown system %s\n", system_s.c_str());
        return 1;
      }
      sp.nmo = M;
      sp.naea = N;
      sp.nchol = K;
      sp.density = density;
      sp.sigma_L = sigma_for(M, K, density);
    }
    sp.sigma_h = 1.0;
    sp.dt = dt;
    sp.seed = seed;
    sp.with_vakbl = (mode == AFQMC_FP64_CSR && 2.0 * sp.naea * sp.nmo <= 4096) ? 1 : 0;

    std::printf("***********************************************************\n");
    std::printf("                 Initializing (synthetic %s)              \n", system_s.c_str());
    std::printf("***********************************************************\n");
    if (afqmc_synth_create(&sp, &syn) != AFQMC_OK) {
      std::fprintf(stderr, "Error generating the synthetic system\n");
      return 1;
    }
    afqmc_synth_get(syn, &v);
  }
  const int NMO = v.nmo, NAEA = v.naea, nchol = v.nchol;
  std::printf("\n  NMO: %d  NAEA: %d  NAEB: %d\n", NMO, NAEA, NAEA);
  std::printf("  Execution details: \n    nsteps: %d\n    nsubsteps: %d\n    nwalk: %d\n    northo: %d\n", nsteps,
              nsubsteps, nwalk, northo);
  std::printf("    verbose: %s\n    # Chol Vectors: %d\n    transposed Spvn: %s\n", verbose ? "true" : "false",
              nchol, transposed ? "true" : "false");
  std::printf("    Chol. Matrix Sparsity: %g\n", (double)v.spvn.nnz / ((double)nchol * NMO * NMO));
  std::printf("    Hamiltonian Sparsity: %g\n", v.vakbl.nnz / (NAEA * NAEA * (double)NMO * NMO * 4.0));
  std::printf("    mode: %s  dt: %g\n", mode_s.c_str(), dt);

  CHECK(afqmc_ctx_create(NMO, NAEA, nchol, nwalk, mode, &ctx));
  CHECK(afqmc_set_option(ctx, AFQMC_OPT_TRANSPOSED, transposed ? 1.0 : 0.0));
  CHECK(afqmc_upload_system(ctx, v.trial_alpha, v.trial_beta, v.propg1, v.haj, &v.spvn,
                            v.vakbl.nnz > 0 && mode == AFQMC_FP64_CSR ? &v.vakbl : nullptr, 1e-6, dt));
  if (syn) afqmc_synth_destroy(syn);
  if (h5s) h5.destroy(h5s);

  // fields for one block of substeps at a time
  const long KW = (long)nchol * nwalk;
  std::vector<double> x((size_t)KW * nsubsteps);
  auto stage_block = [&](long first) -> int {
    for (int s = 0; s < nsubsteps; s++)
      afqmc_synth_fields(seed, nchol, nwalk, 0, nwalk, first + s, x.data() + (size_t)s * KW);
    return afqmc_stage_fields(ctx, x.data(), nsubsteps);
  };
  if (tune) {
    CHECK(stage_block(0));
    CHECK(afqmc_tune(ctx, dt));
  }
  CHECK(afqmc_init_walkers(ctx));
  CHECK(afqmc_mixed_density_matrix(ctx));
  double Eav = 0.0;
  CHECK(afqmc_energy(ctx, &Eav));

  std::printf("\n***********************************************************\n");
  std::printf("                     Beginning Steps                       \n");
  std::printf("***********************************************************\n\n");
  std::printf("# Step   Energy   \n");
  if (verbose) CHECK(afqmc_profile(ctx, 1));
  double total = 0.0;
  std::vector<double> eav_steps;
  std::mt19937_64 comb_rng(seed ^ 0x9e3779b97f4a7c15ULL);
  for (int step = 0; step < nsteps; step++) {
    const long first = (long)step * nsubsteps;
    CHECK(stage_block(first));
    CHECK(afqmc_sync(ctx));
    auto t0 = std::chrono::steady_clock::now();
    // substeps: ortho when the running counter is a positive multiple of northo
    for (int sub = 0; sub < nsubsteps; sub++) {
      const long tot = first + sub;
      CHECK(afqmc_substep(ctx, sub, (tot > 0 && tot % northo == 0) ? 1 : 0, dt));
    }
    CHECK(afqmc_mixed_density_matrix(ctx));
    CHECK(afqmc_energy(ctx, &Eav));
    // branching goes here (miniafqmc_base.cpp:346): comb with a uniform offset
    if (branch) CHECK(afqmc_branch(ctx, (double)(comb_rng() >> 11) * 0x1.0p-53, nullptr));
    total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("%d   %.10f\n", step, Eav);
    eav_steps.push_back(Eav);
  }
  if (eav_steps.size() > 1) {
    // mixed estimator over the steps: mean and standard error of the step energies
    double mean = 0, var = 0;
    for (double e : eav_steps) mean += e;
    mean /= eav_steps.size();
    for (double e : eav_steps) var += (e - mean) * (e - mean);
    var /= (eav_steps.size() - 1);
    std::printf("\nMixed estimator over %zu steps: %.10f +/- %.3e\n", eav_steps.size(), mean,
                std::sqrt(var / eav_steps.size()));
  }
  std::printf("\n***********************************************************\n");
  std::printf("                   Finished Calculation                    \n");
  std::printf("***********************************************************\n\n");
  std::printf("Total  %.6f s  (%.2f walker-steps/s)\n", total, total > 0 ? (double)nwalk * nsteps * nsubsteps / total : 0);
  if (verbose) {
    const int cap = 64;
    static char names[cap][64];
    long long n[cap];
    double ms[cap];
    int cnt = 0;
    CHECK(afqmc_kernel_stats(ctx, cap, names, n, ms, &cnt));
    // the reference's timer report: the same GPU time under its phase names
    std::printf("\n%-30s %12s %10s\n", "Timer", "GPU time s", "launches");
    std::printf("%-30s %12.6f %10s\n", "Total (wall)", total, "");
    for (const TimerSites& t : kTimers) {
      double sum = 0.0;
      long long calls = 0;
      for (const char* site : t.sites) {
        if (!site) break;
        for (int i = 0; i < cnt && i < cap; i++)
          if (std::strcmp(names[i], site) == 0) {
            sum += ms[i];
            calls += n[i];
          }
      }
      std::printf("%-30s %12.6f %10lld\n", t.timer, sum * 1e-3, calls);
    }
    std::printf("\n%-24s %10s %14s\n", "kernel", "launches", "total ms");
    for (int i = 0; i < cnt && i < cap; i++) std::printf("%-24s %10lld %14.3f\n", names[i], n[i], ms[i]);
  }
  afqmc_ctx_destroy(ctx);
  return 0;
}
