// gen.cpp -- seeded synthetic configurations (host only; no GPU calls).
//
// The reference starts from equilibrated .gro files (TPL/confs/*.gro).  The
// benchmark workloads are synthetic instead:
//  * a jittered FCC lattice (homogeneous LJ fluid), and
//  * a liquid-vapour slab with uniform random positions inside each region,
//    placed by random sequential addition with a minimum pair distance so that
//    no pair overlaps (forces stay in the fixed-point range without a
//    relaxation step).
// Random numbers: SplitMix64 (Steele, Lea and Flood, OOPSLA 2014).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/sre.h"

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1p-53; }
};

float wrapf(double x, double L) {
  double y = x - std::floor(x / L) * L;
  float f = (float)y;
  if (f >= (float)L) f = 0.f;
  if (f < 0.f) f = 0.f;
  return f;
}

// uniform grid over the box for the minimum-distance test
struct Grid {
  double L[3];
  int n[3];
  double h[3];
  std::vector<std::vector<int>> cells;
  Grid(double lx, double ly, double lz, double cell) {
    L[0] = lx; L[1] = ly; L[2] = lz;
    for (int d = 0; d < 3; ++d) {
      n[d] = (int)std::floor(L[d] / cell);
      if (n[d] < 1) n[d] = 1;
      h[d] = L[d] / n[d];
    }
    cells.resize((size_t)n[0] * n[1] * n[2]);
  }
  int cidx(int x, int y, int z) const { return (x * n[1] + y) * n[2] + z; }
  void cell_of(const float *p, int c[3]) const {
    for (int d = 0; d < 3; ++d) {
      int v = (int)(p[d] / h[d]);
      c[d] = v < 0 ? 0 : (v >= n[d] ? n[d] - 1 : v);
    }
  }
};

}  // namespace

extern "C" sre_status sre_gen_fcc(int ncell, double a, double jitter,
                                  uint64_t seed, float *xyz) {
  if (ncell < 1 || a <= 0 || !xyz) return SRE_ERR_INVALID;
  static const double basis[4][3] = {
      {0.0, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};
  const double L = ncell * a;
  Rng rng(seed);
  size_t k = 0;
  for (int ix = 0; ix < ncell; ++ix)
    for (int iy = 0; iy < ncell; ++iy)
      for (int iz = 0; iz < ncell; ++iz)
        for (int b = 0; b < 4; ++b) {
          const double c[3] = {(ix + basis[b][0]) * a, (iy + basis[b][1]) * a,
                               (iz + basis[b][2]) * a};
          for (int d = 0; d < 3; ++d) {
            const double j = (2.0 * rng.uniform() - 1.0) * jitter;
            xyz[4 * k + d] = wrapf(c[d] + j, L);
          }
          xyz[4 * k + 3] = 0.f;
          ++k;
        }
  return SRE_OK;
}

extern "C" sre_status sre_gen_slab(int n, double lx, double ly, double lz,
                                   double h, double rho_l, double dmin,
                                   uint64_t seed, float *xyz) {
  if (n < 1 || lx <= 0 || ly <= 0 || lz <= 0 || h < 0 || 2 * h > lz ||
      rho_l < 0 || dmin < 0 || !xyz)
    return SRE_ERR_INVALID;
  long long nl = std::llround(rho_l * lx * ly * 2.0 * h);
  if (nl > n) nl = n;
  const double z0 = 0.5 * lz - h, z1 = 0.5 * lz + h;
  const double vap = lz - 2.0 * h;
  Grid g(lx, ly, lz, dmin > 0 ? dmin : 1.0);
  Rng rng(seed);
  const double dmin2 = dmin * dmin;
  const double L[3] = {lx, ly, lz};
  for (int i = 0; i < n; ++i) {
    const bool liquid = i < nl;
    if (!liquid && vap <= 0) return SRE_ERR_INVALID;
    bool placed = false;
    for (int attempt = 0; attempt < 200000 && !placed; ++attempt) {
      float p[3];
      p[0] = wrapf(rng.uniform() * lx, lx);
      p[1] = wrapf(rng.uniform() * ly, ly);
      double z;
      const double u = rng.uniform();
      if (liquid) {
        z = z0 + u * 2.0 * h;
      } else {
        z = u * vap;                       // [0, vap)
        if (z >= z0) z += 2.0 * h;         // skip the liquid region
      }
      p[2] = wrapf(z, lz);
      bool ok = true;
      if (dmin > 0) {
        int c[3];
        g.cell_of(p, c);
        for (int ox = -1; ox <= 1 && ok; ++ox)
          for (int oy = -1; oy <= 1 && ok; ++oy)
            for (int oz = -1; oz <= 1 && ok; ++oz) {
              int cx = (c[0] + ox + g.n[0]) % g.n[0];
              int cy = (c[1] + oy + g.n[1]) % g.n[1];
              int cz = (c[2] + oz + g.n[2]) % g.n[2];
              // small grids: neighbouring offsets may alias; testing a cell
              // twice is harmless
              for (int j : g.cells[g.cidx(cx, cy, cz)]) {
                double r2 = 0;
                for (int d = 0; d < 3; ++d) {
                  double dd = (double)xyz[4 * j + d] - (double)p[d];
                  dd -= std::floor(dd / L[d] + 0.5) * L[d];
                  r2 += dd * dd;
                }
                if (r2 < dmin2) { ok = false; break; }
              }
            }
        if (ok) g.cells[g.cidx(c[0], c[1], c[2])].push_back(i);
      }
      if (ok) {
        xyz[4 * i + 0] = p[0];
        xyz[4 * i + 1] = p[1];
        xyz[4 * i + 2] = p[2];
        xyz[4 * i + 3] = 0.f;
        placed = true;
      }
    }
    if (!placed) return SRE_ERR_INVALID;
  }
  return SRE_OK;
}
