// ref_driver.cpp -- host driver that runs the reference's own FP64 CPU NEP
// evaluator (third_party/lammps-matpl-overlay/files/src/nep_cpu.cpp,
// NEP3_CPU::compute_for_lammps, nep_cpu.cpp:2385-2479) on a periodic system.
//
// TEST INFRASTRUCTURE ONLY (built into oracle/_ref by oracle/ref.mk from the
// reference sources where they lie; never part of the product).
//
// compute_for_lammps works on raw coordinate differences and Newton-scatters
// forces/virials onto neighbours (nep_cpu.cpp:943-944, 1133-1160), so this
// driver builds explicit periodic ghost images around the box, hands the
// evaluator a full neighbour list of the local atoms, and folds ghost forces
// and virials back onto their owners.
#include "nep_cpu.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Sys {
  std::vector<double> x;   // nall*3
  std::vector<int> owner;  // nall
};

void cross(const double* a, const double* b, double* c)
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

extern "C" {

void* refnep_create(const char* path)
{
  try {
    NEP3_CPU* nep = new NEP3_CPU(std::string(path));
    nep->map_atom_type_idx.resize(nep->paramb.num_types);
    for (int t = 0; t < nep->paramb.num_types; ++t) nep->map_atom_type_idx[t] = t;
    return nep;
  } catch (...) {
    return nullptr;
  }
}

void refnep_destroy(void* h) { delete static_cast<NEP3_CPU*>(h); }

// h: box matrix with the cell vectors as columns (row-major 3x3).
// type: 0-based NEP type per atom.  pos/force SoA (3N), virial SoA (9N,
// xx yy zz xy xz yz yx zx zy), pe (N).  Returns 0 on success.
int refnep_compute(void* handle, const double* h, const int* pbc, int N, const int* type,
                   const double* pos, double* pe, double* force, double* virial)
{
  NEP3_CPU* nep = static_cast<NEP3_CPU*>(handle);
  const double rc = std::max(nep->paramb.rc_radial, nep->paramb.rc_angular);
  const double a[3] = {h[0], h[3], h[6]}, b[3] = {h[1], h[4], h[7]}, c[3] = {h[2], h[5], h[8]};
  double bc[3], ca[3], ab[3];
  cross(b, c, bc);
  cross(c, a, ca);
  cross(a, b, ab);
  const double vol = std::fabs(a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2]);
  const double thick[3] = {vol / std::sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]),
                           vol / std::sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]),
                           vol / std::sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2])};
  // fractional coordinates via the reciprocal rows (cofactors / det)
  double det = h[0] * (h[4] * h[8] - h[5] * h[7]) + h[1] * (h[5] * h[6] - h[3] * h[8]) +
               h[2] * (h[3] * h[7] - h[4] * h[6]);
  double hi[9] = {(h[4] * h[8] - h[5] * h[7]) / det, (h[2] * h[7] - h[1] * h[8]) / det,
                  (h[1] * h[5] - h[2] * h[4]) / det, (h[5] * h[6] - h[3] * h[8]) / det,
                  (h[0] * h[8] - h[2] * h[6]) / det, (h[2] * h[3] - h[0] * h[5]) / det,
                  (h[3] * h[7] - h[4] * h[6]) / det, (h[1] * h[6] - h[0] * h[7]) / det,
                  (h[0] * h[4] - h[1] * h[3]) / det};
  Sys s;
  s.x.resize(3 * N);
  s.owner.resize(N);
  std::vector<double> frac(3 * N);
  for (int i = 0; i < N; ++i) {
    double r[3] = {pos[i], pos[N + i], pos[2 * N + i]};
    for (int d = 0; d < 3; ++d) {
      double f = hi[3 * d] * r[0] + hi[3 * d + 1] * r[1] + hi[3 * d + 2] * r[2];
      if (pbc[d]) f -= std::floor(f);
      frac[3 * i + d] = f;
    }
    for (int d = 0; d < 3; ++d)
      s.x[3 * i + d] = h[3 * d] * frac[3 * i] + h[3 * d + 1] * frac[3 * i + 1] + h[3 * d + 2] * frac[3 * i + 2];
    s.owner[i] = i;
  }
  // ghosts: images whose fractional coordinate lies within rc/thickness of the box
  int nim[3];
  double delta[3];
  for (int d = 0; d < 3; ++d) {
    delta[d] = rc / thick[d] * 1.0000001 + 1e-12;
    nim[d] = pbc[d] ? (int)std::ceil(delta[d]) + 1 : 0;
  }
  for (int ia = -nim[0]; ia <= nim[0]; ++ia)
    for (int ib = -nim[1]; ib <= nim[1]; ++ib)
      for (int ic = -nim[2]; ic <= nim[2]; ++ic) {
        if (ia == 0 && ib == 0 && ic == 0) continue;
        const int sh[3] = {ia, ib, ic};
        for (int i = 0; i < N; ++i) {
          bool keep = true;
          double f[3];
          for (int d = 0; d < 3; ++d) {
            f[d] = frac[3 * i + d] + sh[d];
            if (f[d] < -delta[d] || f[d] > 1.0 + delta[d]) keep = false;
          }
          if (!keep) continue;
          for (int d = 0; d < 3; ++d)
            s.x.push_back(h[3 * d] * f[0] + h[3 * d + 1] * f[1] + h[3 * d + 2] * f[2]);
          s.owner.push_back(i);
        }
      }
  const int nall = (int)s.owner.size();
  // neighbour rows of the local atoms: cell binning of all (local+ghost) atoms
  double lo[3] = {1e300, 1e300, 1e300}, hiv[3] = {-1e300, -1e300, -1e300};
  for (int k = 0; k < nall; ++k)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], s.x[3 * k + d]);
      hiv[d] = std::max(hiv[d], s.x[3 * k + d]);
    }
  int nc[3];
  for (int d = 0; d < 3; ++d) nc[d] = std::max(1, (int)std::floor((hiv[d] - lo[d]) / rc));
  auto cell_of = [&](int k, int* ci) {
    for (int d = 0; d < 3; ++d) {
      int v = (int)std::floor((s.x[3 * k + d] - lo[d]) / (hiv[d] - lo[d] + 1e-9) * nc[d]);
      ci[d] = std::min(std::max(v, 0), nc[d] - 1);
    }
  };
  std::vector<int> head(nc[0] * nc[1] * nc[2], -1), next(nall, -1);
  for (int k = nall - 1; k >= 0; --k) {
    int ci[3];
    cell_of(k, ci);
    int cid = ci[0] + nc[0] * (ci[1] + nc[1] * ci[2]);
    next[k] = head[cid];
    head[cid] = k;
  }
  std::vector<std::vector<int>> nl(N);
  for (int i = 0; i < N; ++i) {
    int ci[3];
    cell_of(i, ci);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int cc[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz};
          if (cc[0] < 0 || cc[1] < 0 || cc[2] < 0 || cc[0] >= nc[0] || cc[1] >= nc[1] || cc[2] >= nc[2])
            continue;
          for (int k = head[cc[0] + nc[0] * (cc[1] + nc[1] * cc[2])]; k >= 0; k = next[k]) {
            if (k == i) continue;
            double dxv = s.x[3 * k] - s.x[3 * i], dyv = s.x[3 * k + 1] - s.x[3 * i + 1],
                   dzv = s.x[3 * k + 2] - s.x[3 * i + 2];
            if (dxv * dxv + dyv * dyv + dzv * dzv < rc * rc) nl[i].push_back(k);
          }
        }
    std::sort(nl[i].begin(), nl[i].end());
  }
  std::vector<int> ilist(N), numneigh(N), lmp_type(nall);
  std::vector<int*> firstneigh(N);
  for (int i = 0; i < N; ++i) {
    ilist[i] = i;
    numneigh[i] = (int)nl[i].size();
    firstneigh[i] = nl[i].data();
  }
  for (int k = 0; k < nall; ++k) lmp_type[k] = type[s.owner[k]] + 1;
  std::vector<double*> xp(nall), fp(nall), vp(nall);
  std::vector<double> fbuf(3 * nall, 0.0), vbuf(9 * nall, 0.0), ebuf(nall, 0.0);
  for (int k = 0; k < nall; ++k) {
    xp[k] = &s.x[3 * k];
    fp[k] = &fbuf[3 * k];
    vp[k] = &vbuf[9 * k];
  }
  double etot = 0.0, vtot[6] = {0, 0, 0, 0, 0, 0};
  nep->compute_for_lammps(N, N, ilist.data(), numneigh.data(), firstneigh.data(), lmp_type.data(),
                          xp.data(), etot, vtot, ebuf.data(), fp.data(), vp.data(), 9, 0,
                          h[0], h[4], h[8], h[1], h[2], h[5], std::string("none"), N,
                          MPI_COMM_WORLD);
  std::memset(force, 0, sizeof(double) * 3 * N);
  std::memset(virial, 0, sizeof(double) * 9 * N);
  for (int i = 0; i < N; ++i) pe[i] = ebuf[i];
  for (int k = 0; k < nall; ++k) {
    int o = s.owner[k];
    for (int d = 0; d < 3; ++d) force[d * N + o] += fbuf[3 * k + d];
    for (int q = 0; q < 9; ++q) virial[q * N + o] += vbuf[9 * k + q];
  }
  return 0;
}

}  // extern "C"
