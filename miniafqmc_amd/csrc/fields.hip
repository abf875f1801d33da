// Host generation of the auxiliary-field draws (afqmc_synth_fields).  The
// same functions (fieldrng.hpp) run inside the fields kernel when the context
// draws its fields on the device, so both give the same numbers bit for bit.
#include "afqmc_hip.h"
#include "fieldrng.hpp"

#include <algorithm>
#include <thread>
#include <vector>

extern "C" int afqmc_synth_fields(uint64_t seed, int nchol, int nwalk, int w0, int wtot, long substep, double* x)
{
  if (!x || nchol <= 0 || nwalk <= 0 || w0 < 0 || w0 + nwalk > wtot) return AFQMC_ERR_ARG;
  auto one = [&](long wl) {
    for (int n = 0; n < nchol; n += 2) {
      double a, b;
      afqmc_rng::field_pair(seed, substep, w0 + wl, n, a, b);
      x[(size_t)n * nwalk + wl] = a;
      if (n + 1 < nchol) x[(size_t)(n + 1) * nwalk + wl] = b;
    }
  };
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if ((long)nwalk * nchol < 65536 || nt == 1) {
    for (long wl = 0; wl < nwalk; wl++) one(wl);
    return AFQMC_OK;
  }
  std::vector<std::thread> th;
  const long chunk = (nwalk + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    const long b = t * chunk, e = std::min<long>(nwalk, b + chunk);
    if (b >= e) break;
    th.emplace_back([&, b, e]() {
      for (long wl = b; wl < e; wl++) one(wl);
    });
  }
  for (auto& t : th) t.join();
  return AFQMC_OK;
}
