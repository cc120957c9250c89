// Host check of the element-buffer gather: the plan of egather_plan (exfemus_amd/csrc/egather_host.cpp) built
// from the element -> SELL slot map of a box hierarchy's top level (mesh, DoF map and pattern from the library's
// mesh unit, the SELL layout from its host builder, in a random internal order), and the gather kernel
// k_egather (exfemus_amd/csrc/assemble.hip) replayed on it block by block in a random block order, with every
// array and every LDS access bounds-checked.  The gathered slot values and row rhs are compared bit for bit
// with the direct sums over the elements in element order, the same order the gather keeps.  Prints one JSON
// line.
//   usage: egather_check dim nx ny nz nlev nq2 nq1 W max_lds seed permute
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>
#include "../../exfemus_amd/csrc/egather_host.h"
#include "../../exfemus_amd/csrc/fm_mesh.h"
#include "../../exfemus_amd/csrc/sell_host.h"

static int fail(const std::string &m) {
  std::printf("{\"ok\": false, \"error\": \"%s\"}\n", m.c_str());
  return 1;
}

// LDS of one block with bounds-checked typed access (the dynamic shared memory of k_egather)
struct Lds {
  std::vector<unsigned char> b;
  template <typename T> T &at(size_t byte_off, size_t i) {
    const size_t o = byte_off + i * sizeof(T);
    if (o + sizeof(T) > b.size()) throw std::runtime_error("LDS access out of bounds");
    return *reinterpret_cast<T *>(b.data() + o);
  }
};

int main(int argc, char **argv) {
  if (argc < 12) return fail("usage: dim nx ny nz nlev nq2 nq1 W max_lds seed permute");
  const int dim = atoi(argv[1]), nlev = atoi(argv[5]), nq2 = atoi(argv[6]), nq1 = atoi(argv[7]), W = atoi(argv[8]);
  const size_t max_lds = (size_t)atol(argv[9]);
  const unsigned seed = (unsigned)atoi(argv[10]);
  const bool permute = atoi(argv[11]) != 0;
  const int nx0[3] = {atoi(argv[2]), atoi(argv[3]), dim == 3 ? atoi(argv[4]) : 1};
  const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  fm_mesh m;
  if (fm_build_box(dim, nx0, lo, hi, nlev, &m) != 0) return fail("fm_build_box failed");
  const int L = nlev - 1, nel = m.nel[L], nl = nq2 * m.n2 + nq1 * m.n1;
  std::vector<int> dof;
  int64_t n = 0;
  fm_init_dof(m, L, nq2, nq1, dof, n);
  std::vector<int64_t> rp;
  std::vector<int> ci;
  fm_pattern(m, L, nq2, nq1, dof, rp, ci);
  // internal order
  std::vector<int> ro(n), cr(n);
  std::iota(ro.begin(), ro.end(), 0);
  std::mt19937_64 rng(seed);
  if (permute) std::shuffle(ro.begin(), ro.end(), rng);
  for (int64_t i = 0; i < n; i++) cr[ro[i]] = (int)i;
  SellHost sh;
  std::string err;
  if (sell_build_host(n, rp, ci, ro.data(), cr.data(), nullptr, W, 8, sh, err)) return fail("sell: " + err);
  // SELL position of every CSR entry
  std::vector<int> kof(rp[n], -1);
  for (int64_t s = 0; s < sh.nsl; s++)
    for (int l = 0; l < 64; l++) {
      const int i = sh.perm[64 * s + l];
      if (i < 0) continue;
      for (int k = 0; k < sh.wd[s]; k++) {
        const int64_t slot = sh.off[s] + 64 * (int64_t)k + l;
        if (sh.pos[slot] != 0xFFFF) kof[sh.rbase[i] + sh.pos[slot]] = k;
      }
    }
  // element -> slot map (the library's elslot / elpos)
  std::vector<int> ed(nl);
  std::vector<int> eldof((size_t)nel * nl);
  std::vector<int64_t> slot0((size_t)nel * nl);
  std::vector<unsigned short> pos((size_t)nel * nl * nl);
  for (int e = 0; e < nel; e++) {
    fm_elem_dofs(m, L, e, nq2, nq1, dof, ed.data());
    for (int a = 0; a < nl; a++) {
      eldof[(size_t)e * nl + a] = ed[a];
      const int r = ed[a];
      slot0[(size_t)e * nl + a] = sh.rslot[cr[r]];
      for (int b = 0; b < nl; b++) {
        const auto it = std::lower_bound(ci.begin() + rp[r], ci.begin() + rp[r + 1], ed[b]);
        if (it == ci.begin() + rp[r + 1] || *it != ed[b]) return fail("element entry outside the pattern");
        pos[((size_t)e * nl + a) * nl + b] = (unsigned short)kof[it - ci.begin()];
      }
    }
  }
  EGatherPlan p;
  if (egather_plan(nel, nl, pos, slot0, sh.off, sh.wd, max_lds, p, err)) return fail("plan: " + err);
  size_t oc, os, on;
  const size_t lds = egather_lds(p.maxseg, nl, p.maxctr, p.maxit, &oc, &os, &on);
  if (lds > max_lds) return fail("LDS over budget");
  if ((int64_t)p.bbase.size() != p.nblk + 1 || (int64_t)p.lseg.size() != p.nblk * p.lg + 1) return fail("plan sizes");
  // element buffers: random values
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::vector<double> E((size_t)nel * nl * nl), Fe((size_t)nel * nl);
  for (double &x : E) x = U(rng);
  for (double &x : Fe) x = U(rng);
  // the gather, block by block in a random order
  std::vector<double> v(sh.nslot, std::nan("")), rhs(n, std::nan(""));
  std::vector<float> vf(sh.nslot, std::nanf(""));
  std::vector<int64_t> order(p.nblk);
  std::iota(order.begin(), order.end(), 0);
  std::shuffle(order.begin(), order.end(), rng);
  try {
    for (int64_t g : order) {
      Lds sm;
      sm.b.assign(lds, 0);
      const int lg = p.lg, gpb = 64 / lg;
      const int64_t sl = g / gpb;
      const int l0 = (int)(g % gpb) * lg;
      const int64_t s0 = p.lseg.at(g * lg), s1 = p.lseg.at((g + 1) * lg);
      const int nseg = (int)(s1 - s0);
      const int64_t c0 = p.bbase.at(g), c1 = p.bbase.at(g + 1);
      const int nc = (int)(c1 - c0);
      const int Wd = sh.wd.at(sl), nit = Wd * lg;
      const int64_t off = sh.off.at(sl);
      const size_t og = on + ((size_t)nit + 7) / 8 * 8;  // element row ids
      for (int t = 0; t < nseg; t++) sm.at<int>(og, t) = p.seg.at(s0 + t);
      for (int t = 0; t < nc; t++) sm.at<unsigned short>(oc, t) = p.cidx.at(c0 + t);
      for (int t = 0; t < nit; t++) {
        const int64_t slot = off + 64 * (int64_t)(t / lg) + l0 + t % lg;
        sm.at<unsigned short>(os, t) = p.cst.at(slot);
        sm.at<unsigned char>(on, t) = p.cnt.at(slot);
      }
      for (int t = 0; t < nseg * nl; t++) sm.at<double>(0, t) = E.at((int64_t)sm.at<int>(og, t / nl) * nl + t % nl);
      for (int t = 0; t < lg; t++) {
        const int row = sh.perm.at(64 * sl + l0 + t);
        if (row < 0) continue;
        double f = 0.;
        for (int64_t q = p.lseg.at(g * lg + t); q < p.lseg.at(g * lg + t + 1); q++) f += Fe.at(sm.at<int>(og, q - s0));
        rhs.at(row) = f;
      }
      for (int t = 0; t < nit; t++) {
        const int64_t slot = off + 64 * (int64_t)(t / lg) + l0 + t % lg;
        const int c = sm.at<unsigned char>(on, t), st = sm.at<unsigned short>(os, t);
        double sum = 0.;
        for (int q = 0; q < c; q++) sum += sm.at<double>(0, sm.at<unsigned short>(oc, st + q));
        v.at(slot) = sum;
        vf.at(slot) = (float)sum;
      }
    }
  } catch (const std::exception &ex) {
    return fail(ex.what());
  }
  // direct sums in element order
  std::vector<double> ref(sh.nslot, 0.), rref(n, 0.);
  std::vector<char> hit(n, 0);
  for (int e = 0; e < nel; e++)
    for (int a = 0; a < nl; a++) {
      const size_t r = (size_t)e * nl + a;
      rref[cr[eldof[r]]] += Fe[r];
      hit[cr[eldof[r]]] = 1;
      for (int b = 0; b < nl; b++) ref[slot0[r] + 64 * (int64_t)pos[r * nl + b]] += E[r * nl + b];
    }
  int64_t bad = 0;
  for (int64_t s = 0; s < sh.nslot; s++)
    if (std::memcmp(&v[s], &ref[s], sizeof(double)) != 0 || vf[s] != (float)ref[s]) bad++;
  int64_t badr = 0;
  for (int64_t i = 0; i < n; i++)
    if (hit[i] && std::memcmp(&rhs[i], &rref[i], sizeof(double)) != 0) badr++;
  if (bad || badr) return fail("gathered values differ: " + std::to_string(bad) + " slots, " + std::to_string(badr) + " rows");
  std::printf("{\"ok\": true, \"lg\": %d, \"nblk\": %lld, \"maxseg\": %d, \"maxctr\": %d, \"maxit\": %d, \"lds\": %zu, "
              "\"slots\": %lld, \"contributions\": %lld}\n",
              p.lg, (long long)p.nblk, p.maxseg, p.maxctr, p.maxit, lds, (long long)sh.nslot, (long long)p.nctr);
  return 0;
}
