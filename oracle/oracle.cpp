// oracle.cpp -- CPU restatement of the reference's short-range force path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
// file; it is used by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py, as the checker and as the timed CPU path.
//
// What it restates (reference paths relative to the repository root):
//   H  = shortRange.inhomo/template/tools/simulation.hetero.cutoff
//   U  = shortRange.inhomo/template/tools/simulation.uniform.cutoff
//   FC = shortRange.inhomo/template/tools/simulation.fcorr
//   * rc assignment                 H/src/AssignRCut.cu:77-104
//   * Verlet pair set               H/src/NeighborList.cu:327-420 (predicate :391)
//   * LJ force/energy               H/include/NonBondedInteraction.h:866-889
//   * per-atom-rc force kernel      H/src/InteractionEngine_interface.cu:1204-1281
//   * uniform (plain) path          U/include/NonBondedInteraction.h:876-900
//   * minimum image                 H/include/BoxGeometry.h:171-186
//   * virial/pressure convention    H/include/Statistic_interface.h:92-117
//   * mean error force (tail)       FC/src/InteractionEngine_interface.cu:1118,
//                                   FC/src/AssignForceCorr.cu:95-128, PRE:256-262
//
// Precision: the pair SET (neighbor list and cutoff mask) is decided by the
// same IEEE single-precision operation sequence the engine uses (it is a
// definition, restated here operation by operation); forces, energies,
// virials, tail terms and gradients are evaluated in binary64 from the FP32
// input positions.  Compile with -ffp-contract=off.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// ------------------------------------------------------- shared definitions

// Fixed-point lattice coordinate u = floor(x * 2^32/L + 0.5) mod 2^32.
static inline uint32_t or_q(float x, double invq) {
  double t = (double)x * invq;
  double f = std::floor(t + 0.5);
  double w = f - std::floor(f * 0x1p-32) * 0x1p32;
  return (uint32_t)w;
}

static inline float or_disp(uint32_t uj, uint32_t ui, float q) {
  return (float)(int32_t)(uj - ui) * q;
}

static inline float or_d2(float dx, float dy, float dz) {
  return std::fmaf(dz, dz, std::fmaf(dy, dy, dx * dx));
}

void or_quantize(int n, const float *xyz, int stride, const float *box,
                 uint32_t *u) {
  for (int d = 0; d < 3; ++d) {
    const double invq = 4294967296.0 / (double)box[d];
    for (int i = 0; i < n; ++i) u[3 * i + d] = or_q(xyz[(size_t)i * stride + d], invq);
  }
}

// rc bin of one coordinate: H/src/AssignRCut.cu:92-101 (one-image wrap, FP32
// multiply and divide, truncation); the index is clamped to the lattice.
static inline int or_bin(float x, float L, int n) {
  if (x < 0.f) x = x + L;
  else if (x >= L) x = x - L;
  int ix = (int)((x * (float)n) / L);
  if (ix < 0) ix = 0;
  if (ix > n - 1) ix = n - 1;
  return ix;
}

void or_assign_rc(int n, const float *xyz, int stride, const float *box,
                  int nx, int ny, int nz, const float *table, float *rc,
                  int *bin_z) {
  for (int i = 0; i < n; ++i) {
    const float *p = xyz + (size_t)i * stride;
    const int ix = or_bin(p[0], box[0], nx), iy = or_bin(p[1], box[1], ny),
              iz = or_bin(p[2], box[2], nz);
    rc[i] = table[iz + nz * (iy + ny * ix)];
    if (bin_z) bin_z[i] = iz;
  }
}

// FP32 pair predicate shared by list and mask
struct OrSys {
  int n;
  std::vector<uint32_t> u;
  float q[3];
};
static void or_prepare(OrSys &S, int n, const float *xyz, int stride,
                       const float *box) {
  S.n = n;
  S.u.resize((size_t)3 * n);
  or_quantize(n, xyz, stride, box, S.u.data());
  for (int d = 0; d < 3; ++d) S.q[d] = std::ldexp(box[d], -32);
}
static inline float or_pair_d2(const OrSys &S, int i, int j) {
  const uint32_t *a = &S.u[3 * (size_t)i], *b = &S.u[3 * (size_t)j];
  return or_d2(or_disp(b[0], a[0], S.q[0]), or_disp(b[1], a[1], S.q[1]),
               or_disp(b[2], a[2], S.q[2]));
}

// ----------------------------------------------------------- pair lists ---

// O(N^2) build (the reference adaptive engine's own method,
// H/src/NeighborList.cu:373-413): j in L(i) iff j != i and d2 < (rc_i+skin)^2.
// Rows come out ascending in j.  j may be NULL to count only.
int64_t or_build_pairs_brute(int n, const float *xyz, int stride,
                             const float *box, const float *rc, float skin,
                             int64_t *offsets, int32_t *jout) {
  OrSys S;
  or_prepare(S, n, xyz, stride, box);
  std::vector<int64_t> cnt(n, 0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    const float rl = rc[i] + skin;
    const float rl2 = rl * rl;
    int64_t c = 0;
    for (int j = 0; j < n; ++j)
      if (j != i && or_pair_d2(S, i, j) < rl2) ++c;
    cnt[i] = c;
  }
  offsets[0] = 0;
  for (int i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + cnt[i];
  if (jout) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
      const float rl = rc[i] + skin;
      const float rl2 = rl * rl;
      int64_t k = offsets[i];
      for (int j = 0; j < n; ++j)
        if (j != i && or_pair_d2(S, i, j) < rl2) jout[k++] = j;
    }
  }
  return offsets[n];
}

// Cell-list build (the uniform engine's method, U/src/CellList_interface.cu:148-251
// + U/src/NeighborList_interface.cu:361-491) with cells no smaller than the
// largest rc_i + skin and a 27-cell stencil; same predicate, rows sorted.
int64_t or_build_pairs_cells(int n, const float *xyz, int stride,
                             const float *box, const float *rc, float skin,
                             int64_t *offsets, int32_t *jout) {
  OrSys S;
  or_prepare(S, n, xyz, stride, box);
  float rlmax = 0.f;
  for (int i = 0; i < n; ++i) rlmax = std::max(rlmax, rc[i] + skin);
  int nc[3];
  for (int d = 0; d < 3; ++d) {
    nc[d] = (int)std::floor(box[d] / (rlmax * 1.001f));
    if (nc[d] < 3) nc[d] = 1;  // one cell: every column searched once
  }
  const int ncell = nc[0] * nc[1] * nc[2];
  std::vector<int> cell(n), head(ncell + 1, 0), order(n);
  for (int i = 0; i < n; ++i) {
    int c[3];
    for (int d = 0; d < 3; ++d)
      c[d] = (int)(((uint64_t)S.u[3 * (size_t)i + d] * (uint64_t)nc[d]) >> 32);
    cell[i] = (c[0] * nc[1] + c[1]) * nc[2] + c[2];
    head[cell[i] + 1]++;
  }
  for (int c = 0; c < ncell; ++c) head[c + 1] += head[c];
  {
    std::vector<int> fill(head.begin(), head.end() - 1);
    for (int i = 0; i < n; ++i) order[fill[cell[i]]++] = i;
  }
  std::vector<std::vector<int32_t>> rows(n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    const float rl = rc[i] + skin;
    const float rl2 = rl * rl;
    int ci[3];
    ci[2] = cell[i] % nc[2];
    ci[1] = (cell[i] / nc[2]) % nc[1];
    ci[0] = cell[i] / (nc[2] * nc[1]);
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      if (nc[d] == 1) { lo[d] = 0; hi[d] = 0; }
      else { lo[d] = -1; hi[d] = 1; }
    }
    std::vector<int32_t> &row = rows[i];
    for (int ox = lo[0]; ox <= hi[0]; ++ox)
      for (int oy = lo[1]; oy <= hi[1]; ++oy)
        for (int oz = lo[2]; oz <= hi[2]; ++oz) {
          const int cx = (ci[0] + ox + nc[0]) % nc[0];
          const int cy = (ci[1] + oy + nc[1]) % nc[1];
          const int cz = (ci[2] + oz + nc[2]) % nc[2];
          const int c = (cx * nc[1] + cy) * nc[2] + cz;
          for (int k = head[c]; k < head[c + 1]; ++k) {
            const int j = order[k];
            if (j != i && or_pair_d2(S, i, j) < rl2) row.push_back(j);
          }
        }
    std::sort(row.begin(), row.end());
  }
  offsets[0] = 0;
  for (int i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + (int64_t)rows[i].size();
  if (jout)
    for (int i = 0; i < n; ++i)
      std::copy(rows[i].begin(), rows[i].end(), jout + offsets[i]);
  return offsets[n];
}

// ------------------------------------------------------------ tail term ---

struct OrLJ {
  double s6, s12;
};
static inline double oH(double t, const OrLJ &p) {
  const double t2 = t * t, t4 = t2 * t2, t10 = t4 * t4 * t2;
  return p.s6 / t4 - 0.4 * p.s12 / t10;
}
static inline double oK(double t, const OrLJ &p) {
  const double t3 = t * t * t, t9 = t3 * t3 * t3;
  return -p.s6 / (3.0 * t3) + 0.4 * p.s12 / (9.0 * t9);
}
static inline double oU(double r, const OrLJ &p) {
  const double r2 = r * r, r6 = r2 * r2 * r2;
  return 4.0 * (p.s12 / (r6 * r6) - p.s6 / r6);
}
// antiderivative of s * u(max(R,|s|))
static inline double oG(double s, double R, const OrLJ &p) {
  const double a = std::fabs(s);
  if (a <= R) return 0.5 * oU(R, p) * s * s;
  return oH(a, p) - oH(R, p) + 0.5 * oU(R, p) * R * R;
}
// antiderivative of Phi(s) = 2 pi int_{max(R,|s|)}^inf u(r) r dr
static inline double oP(double s, double R, const OrLJ &p) {
  const double a = std::fabs(s);
  const double v = (a <= R) ? oH(R, p) * a : oH(R, p) * R + oK(a, p) - oK(R, p);
  return (s < 0 ? 1.0 : -1.0) * 2.0 * M_PI * v;
}

// Mean error force (z) and half tail energy at height z for an atom of type
// ta with cutoff R, from piecewise-constant rho[tb][k] over nz bins of [0,Lz):
//   F = -2 pi sum_b eps_ab int rho_b(z+s) s u_ab(max(R,|s|)) ds
//   E = 1/2 sum_b eps_ab int rho_b(z+s) Phi_ab(s; R) ds
// Periodic images n = -M..M (M = max(1, ceil(30/Lz))) exactly, the rest with
// the mean density.
void or_tail_point(double z, double R, int ta, int nt, const float *eps,
                   const float *sig, const double *rho, int nz, double Lz,
                   double *F, double *E) {
  const int M = std::max(1, (int)std::ceil(30.0 / Lz));
  const double w = Lz / nz;
  double Ft = 0, Et = 0;
  for (int tb = 0; tb < nt; ++tb) {
    const double sg = sig[ta * nt + tb], ep = eps[ta * nt + tb];
    OrLJ p;
    p.s6 = std::pow(sg, 6);
    p.s12 = p.s6 * p.s6;
    const double *rb = rho + (size_t)tb * nz;
    double fb = 0, eb = 0, rbar = 0;
    for (int k = 0; k < nz; ++k) rbar += rb[k];
    rbar /= nz;
    for (int img = -M; img <= M; ++img)
      for (int k = 0; k < nz; ++k) {
        if (rb[k] == 0.0) continue;
        const double s0 = k * w + img * Lz - z, s1 = s0 + w;
        fb += rb[k] * (oG(s1, R, p) - oG(s0, R, p));
        eb += rb[k] * (oP(s1, R, p) - oP(s0, R, p));
      }
    const double sL = -M * Lz - z, sU = (M + 1) * Lz - z;
    const double Pinf = -2.0 * M_PI * (oH(R, p) * R - oK(R, p));
    fb += rbar * (oG(sL, R, p) - oG(sU, R, p));
    eb += rbar * (2.0 * Pinf - oP(sU, R, p) + oP(sL, R, p));
    Ft += -2.0 * M_PI * ep * fb;
    Et += 0.5 * ep * eb;
  }
  *F = Ft;
  *E = Et;
}

// ------------------------------------------------------------ force path ---

// FP64 forces, per-atom energy and virial diagonal from a CSR pair list.
// Mask: the FP32 predicate d2 < rc_i^2 (H/src/InteractionEngine_interface.cu:1245).
// Pair terms (H/include/NonBondedInteraction.h:866-889):
//   F_i += d (-24 eps (2 s12 - s6) / r^2),  E_i += 4 eps (s12 - s6) + shift
//   W_i,aa += F_a d_a ;  stored E_i/2, W_i/2 (:1276-1279).
// tail (optional): per atom (F_z, E) added, tail[2*i], tail[2*i+1].
// Returns the number of pairs inside rc.
int64_t or_forces(int n, const float *xyz, int stride, const float *box,
                  const int32_t *type, int nt, const float *eps,
                  const float *sig, const float *shift, const float *rc,
                  const int64_t *offsets, const int32_t *jl,
                  const double *tail, double *F, double *atomE,
                  double *atomW) {
  OrSys S;
  or_prepare(S, n, xyz, stride, box);
  const double L[3] = {box[0], box[1], box[2]};
  int64_t nin = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nin)
  for (int i = 0; i < n; ++i) {
    const float rc2 = rc[i] * rc[i];
    const float *pi = xyz + (size_t)i * stride;
    double f[3] = {0, 0, 0}, e = 0, w[3] = {0, 0, 0};
    const int ti = type ? type[i] : 0;
    for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
      const int j = jl[k];
      if (!(or_pair_d2(S, i, j) < rc2)) continue;
      ++nin;
      const float *pj = xyz + (size_t)j * stride;
      double d[3];
      for (int a = 0; a < 3; ++a) {
        double v = (double)pj[a] - (double)pi[a];
        v -= std::floor(v / L[a] + 0.5) * L[a];
        d[a] = v;
      }
      const int tj = type ? type[j] : 0;
      const double ep = eps[ti * nt + tj], sg = sig[ti * nt + tj];
      const double sh = shift ? shift[ti * nt + tj] : 0.0;
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      const double sr2 = sg * sg / r2;
      const double s6 = sr2 * sr2 * sr2, s12 = s6 * s6;
      const double fs = -24.0 * ep * (2.0 * s12 - s6) / r2;
      for (int a = 0; a < 3; ++a) {
        f[a] += fs * d[a];
        w[a] += fs * d[a] * d[a];
      }
      e += 4.0 * ep * (s12 - s6) + sh;
    }
    e *= 0.5;
    if (tail) {
      f[2] += tail[2 * i];
      e += tail[2 * i + 1];
    }
    for (int a = 0; a < 3; ++a) F[3 * (size_t)i + a] = f[a];
    if (atomE) atomE[i] = e;
    if (atomW)
      for (int a = 0; a < 3; ++a) atomW[3 * (size_t)i + a] = 0.5 * w[a];
  }
  return nin;
}

// Per-atom tail (F_z, E) in bin mode: value at the centre of the atom's rc
// bin with that bin's rc (FC/src/AssignForceCorr.cu:95-128 assigns the
// per-bin mean error force by the same bin rule as AssignRCut).
void or_tail_atoms(int n, const float *xyz, int stride, const float *box,
                   const int32_t *type, int nt, const float *eps,
                   const float *sig, int nz_rc, const float *rc_z,
                   const double *rho, int nz_rho, int exact, double *out) {
  std::vector<double> tab((size_t)nz_rc * nt * 2, 0.0);
  if (!exact) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nz_rc; ++b)
      for (int ta = 0; ta < nt; ++ta) {
        const double z = (b + 0.5) * (double)box[2] / nz_rc;
        or_tail_point(z, rc_z[b], ta, nt, eps, sig, rho, nz_rho, box[2],
                      &tab[2 * ((size_t)b * nt + ta)],
                      &tab[2 * ((size_t)b * nt + ta) + 1]);
      }
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    const float *p = xyz + (size_t)i * stride;
    const int b = or_bin(p[2], box[2], nz_rc);
    const int ta = type ? type[i] : 0;
    if (exact) {
      double z = (double)p[2];
      z -= std::floor(z / box[2]) * box[2];
      or_tail_point(z, rc_z[b], ta, nt, eps, sig, rho, nz_rho, box[2],
                    &out[2 * (size_t)i], &out[2 * (size_t)i + 1]);
    } else {
      out[2 * (size_t)i] = tab[2 * ((size_t)b * nt + ta)];
      out[2 * (size_t)i + 1] = tab[2 * ((size_t)b * nt + ta) + 1];
    }
  }
}

// ------------------------------------------------------------- gradients ---

// Analytic FP64 gradient of L = dLdE * E + sum_i dLdF_i . F_i with respect to
// eps[ab], sig[ab] (symmetric: the (a,b) and (b,a) uses are summed into both
// entries) for the pair part with the cutoff mask held fixed.
void or_grad_pairs(int n, const float *xyz, int stride, const float *box,
                   const int32_t *type, int nt, const float *eps,
                   const float *sig, const float *rc, const int64_t *offsets,
                   const int32_t *jl, double dLdE, const double *dLdF,
                   double *g_eps, double *g_sig) {
  OrSys S;
  or_prepare(S, n, xyz, stride, box);
  const double L[3] = {box[0], box[1], box[2]};
  const int npt = nt * nt;
  std::vector<double> ge(npt, 0.0), gs(npt, 0.0);
#pragma omp parallel
  {
    std::vector<double> le(npt, 0.0), ls(npt, 0.0);
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
      const float rc2 = rc[i] * rc[i];
      const float *pi = xyz + (size_t)i * stride;
      const int ti = type ? type[i] : 0;
      const double g[3] = {dLdF ? dLdF[3 * (size_t)i] : 0.0,
                           dLdF ? dLdF[3 * (size_t)i + 1] : 0.0,
                           dLdF ? dLdF[3 * (size_t)i + 2] : 0.0};
      for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
        const int j = jl[k];
        if (!(or_pair_d2(S, i, j) < rc2)) continue;
        const float *pj = xyz + (size_t)j * stride;
        double d[3];
        for (int a = 0; a < 3; ++a) {
          double v = (double)pj[a] - (double)pi[a];
          v -= std::floor(v / L[a] + 0.5) * L[a];
          d[a] = v;
        }
        const int p = ti * nt + (type ? type[j] : 0);
        const double ep = eps[p], sg = sig[p];
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double sr2 = sg * sg / r2;
        const double s6 = sr2 * sr2 * sr2, s12 = s6 * s6;
        const double gd = g[0] * d[0] + g[1] * d[1] + g[2] * d[2];
        // E_i/2 = 2 eps (s12 - s6);  F_i = d * (-24 eps (2 s12 - s6)/r^2)
        le[p] += dLdE * 2.0 * (s12 - s6) - 24.0 * (2.0 * s12 - s6) / r2 * gd;
        ls[p] += dLdE * 12.0 * ep * (2.0 * s12 - s6) / sg -
                 144.0 * ep * (4.0 * s12 - s6) / (sg * r2) * gd;
      }
    }
#pragma omp critical
    for (int p = 0; p < npt; ++p) {
      ge[p] += le[p];
      gs[p] += ls[p];
    }
  }
  for (int a = 0; a < nt; ++a)
    for (int b = 0; b < nt; ++b) {
      const int ab = a * nt + b, ba = b * nt + a;
      g_eps[ab] = ge[ab] + (ab != ba ? ge[ba] : 0.0);
      g_sig[ab] = gs[ab] + (ab != ba ? gs[ba] : 0.0);
    }
}

// Tail part of the same gradient (bin mode), by central differences of
// or_tail_point with step h in each parameter: a method independent of the
// engine's analytic derivative.  g_rc has one entry per rc(z) bin.
void or_grad_tail(int n, const float *xyz, int stride, const float *box,
                  const int32_t *type, int nt, const float *eps,
                  const float *sig, int nz_rc, const float *rc_z,
                  const double *rho, int nz_rho, double dLdE,
                  const double *dLdF, double h, double *g_eps, double *g_sig,
                  double *g_rc) {
  std::vector<double> WF((size_t)nz_rc * nt, 0.0), WE((size_t)nz_rc * nt, 0.0);
  for (int i = 0; i < n; ++i) {
    const int b = or_bin(xyz[(size_t)i * stride + 2], box[2], nz_rc);
    const int ta = type ? type[i] : 0;
    WF[(size_t)b * nt + ta] += dLdF ? dLdF[3 * (size_t)i + 2] : 0.0;
    WE[(size_t)b * nt + ta] += dLdE;
  }
  const int npt = nt * nt;
  std::vector<double> ge(npt, 0.0), gs(npt, 0.0);
  for (int k = 0; k < nz_rc; ++k) g_rc[k] = 0.0;
#pragma omp parallel
  {
    std::vector<double> le(npt, 0.0), ls(npt, 0.0);
    std::vector<float> ep(eps, eps + npt), sg(sig, sig + npt);
#pragma omp for schedule(dynamic, 1)
    for (int b = 0; b < nz_rc; ++b) {
      const double z = (b + 0.5) * (double)box[2] / nz_rc;
      double grc = 0.0;
      for (int ta = 0; ta < nt; ++ta) {
        const double wf = WF[(size_t)b * nt + ta], we = WE[(size_t)b * nt + ta];
        if (wf == 0.0 && we == 0.0) continue;
        auto val = [&](const float *e_, const float *s_, double R) {
          double F, E;
          or_tail_point(z, R, ta, nt, e_, s_, rho, nz_rho, box[2], &F, &E);
          return wf * F + we * E;
        };
        const double R = rc_z[b];
        grc += (val(eps, sig, R + h) - val(eps, sig, R - h)) / (2 * h);
        // FD in binary64 parameters: evaluate the closed form with double
        // parameters by scaling (eps enters linearly, sigma via s6, s12)
        for (int a = 0; a < nt; ++a)
          for (int c = a; c < nt; ++c) {
            if (a != ta && c != ta) continue;
            const int ac = a * nt + c, ca = c * nt + a;
            // d/d eps is exact: T is linear in eps_ab
            std::vector<float> z0(ep.begin(), ep.end());
            for (float &v : z0) v = 0.0f;
            z0[ac] = 1.0f; z0[ca] = 1.0f;
            const double dEps = val(z0.data(), sig, R);
            le[ac] += dEps;
            // d/d sigma by central differences around sigma_ac
            const double s0 = sig[ac];
            const double hs = h * s0;
            std::vector<float> sp(sg.begin(), sg.end()), sm(sg.begin(), sg.end());
            // float parameters cannot carry s0 +- hs exactly; use the
            // float-representable neighbours and the actual step
            const float fp = (float)(s0 + hs), fm = (float)(s0 - hs);
            sp[ac] = fp; sp[ca] = fp;
            sm[ac] = fm; sm[ca] = fm;
            ls[ac] += (val(eps, sp.data(), R) - val(eps, sm.data(), R)) /
                      ((double)fp - (double)fm);
          }
      }
      g_rc[b] = grc;
    }
#pragma omp critical
    for (int p = 0; p < npt; ++p) {
      ge[p] += le[p];
      gs[p] += ls[p];
    }
  }
  for (int a = 0; a < nt; ++a)
    for (int c = 0; c < nt; ++c) {
      const int ac = a * nt + c;
      const int lo = std::min(a, c) * nt + std::max(a, c);
      g_eps[ac] = ge[lo];
      g_sig[ac] = gs[lo];
    }
}

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void or_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

}  // extern "C"
