// Seeded synthetic spectral fields (host code): the counter-based generator
// shared by every host language of the library.  SplitMix64 keyed by
// (seed, field id, n, m, kappa), Box-Muller in FP64, then made admissible
// (chopped, Hermitian m = 0 plane, Nyquist zero).  The same definition is
// written in numpy in mlegs_amd/inputs.py; the two agree to rounding of the
// elementary functions.
#include <cmath>
#include <cstdint>
#include <vector>

#include "mlx_internal.h"

namespace {

inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline void gaussian_pair(uint64_t seed, int field_id, int64_t n, int64_t m, int64_t kap, double &g1, double &g2) {
  uint64_t h = splitmix64(seed ^ ((uint64_t)field_id << 48));
  h = splitmix64(h ^ (uint64_t)n);
  h = splitmix64(h ^ ((uint64_t)m << 20));
  h = splitmix64(h ^ ((uint64_t)kap << 40));
  const uint64_t h2 = splitmix64(h);
  const double u1 = ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((double)(h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double rad = std::sqrt(-2.0 * std::log(u1));
  const double two_pi = 6.283185307179586;
  g1 = rad * std::cos(two_pi * u2);
  g2 = rad * std::sin(two_pi * u2);
}

}  // namespace

extern "C" int mlx_seed_spectral(double *host, int64_t nrdim, int64_t npdim, int64_t nzdim, const int32_t *chops,
                                 uint64_t seed, int32_t field_id, double amp, double scale, int32_t kind) {
  if (!host || !chops || nrdim <= 0 || npdim <= 0 || nzdim <= 0 || scale <= 0.0 || (kind != 0 && kind != 1))
    return MLX_E_ARG;
  const int nrchop = chops[0], npchop = chops[1], np = chops[3];
  const int64_t nz = nzdim;
  const size_t plane = (size_t)nrdim * npdim;
  auto at = [&](int64_t i, int64_t j, int64_t k) { return 2 * ((size_t)k * plane + (size_t)j * nrdim + i); };
  for (int64_t k = 0; k < nz; ++k) {
    const int64_t kap = (nz > 1) ? (k <= nz / 2 ? k : k - nz) : 0;
    for (int64_t j = 0; j < npdim; ++j)
      for (int64_t i = 0; i < nrdim; ++i) {
        double re = 0.0, im = 0.0;
        const int64_t nrc = nrchop - j > 0 ? nrchop - j : 0;
        if (i < nrc && j < npchop) {
          const int64_t n = i + j;
          double g1, g2;
          gaussian_pair(seed, field_id, n, j, kap, g1, g2);
          const double a = (kind == 0) ? amp * std::exp(-(double)(n + j + (kap < 0 ? -kap : kap)) / scale)
                                       : std::exp(-(double)n / scale);
          re = a * g1 / std::sqrt(2.0);
          im = a * g2 / std::sqrt(2.0);
        }
        host[at(i, j, k)] = re;
        host[at(i, j, k) + 1] = im;
      }
  }
  if (nz > 1) {
    for (int64_t k = 1; k < nz / 2; ++k)
      for (int64_t i = 0; i < nrdim; ++i) {
        host[at(i, 0, nz - k)] = host[at(i, 0, k)];
        host[at(i, 0, nz - k) + 1] = -host[at(i, 0, k) + 1];
      }
    for (int64_t j = 0; j < npdim; ++j)
      for (int64_t i = 0; i < nrdim; ++i) host[at(i, j, nz / 2)] = host[at(i, j, nz / 2) + 1] = 0.0;
  }
  for (int64_t i = 0; i < nrdim; ++i) host[at(i, 0, 0) + 1] = 0.0;
  if (np > 1 && npdim - 1 == np / 2)
    for (int64_t k = 0; k < nz; ++k)
      for (int64_t i = 0; i < nrdim; ++i) host[at(i, npdim - 1, k)] = host[at(i, npdim - 1, k) + 1] = 0.0;
  return MLX_OK;
}
