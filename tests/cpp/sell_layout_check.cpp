// Host check of the SELL-64 layout built by sell_build_host (exfemus_amd/csrc/sell_host.cpp) on structured
// patterns: a box of nx x ny x nz nodes with nv variables per node, every DoF coupled to all DoFs of the nodes
// within Chebyshev distance r (the shape of Q2 rows), rows renumbered by a random permutation (or none), row
// groups at random boundaries.  Checks, for every slot, that the entry it holds is a CSR entry of its row, that
// every CSR entry is held exactly once, that the decoded column (base + lane + 2-bit residual, base + 8/16-bit
// offset, or the 32-bit column) is the entry's internal column, that a row's columns increase with the
// position, that padding is marked and decodes to a column inside the vector, that no slice sits in a class
// below the floor, and that the row slot bases and group slices are consistent.  Prints one JSON line.
//   usage: sell_layout_check nx ny nz nv r seed groups W minbits permute   (minbits 2, 8, 16 or 32)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <string>
#include <vector>
#include "../../exfemus_amd/csrc/sell_host.h"

static int fail(const std::string &m) {
  std::printf("{\"ok\": false, \"error\": \"%s\"}\n", m.c_str());
  return 1;
}

int main(int argc, char **argv) {
  if (argc < 11) return fail("usage: nx ny nz nv r seed groups W minbits permute");
  const int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]), nv = atoi(argv[4]), r = atoi(argv[5]);
  const unsigned seed = (unsigned)atoi(argv[6]);
  const int ngroups = atoi(argv[7]), W = atoi(argv[8]);
  const int minbits = atoi(argv[9]);
  const bool permute = atoi(argv[10]) != 0;
  const int64_t nn = (int64_t)nx * ny * nz, n = nn * nv;
  auto node = [&](int i, int j, int k) { return ((int64_t)k * ny + j) * nx + i; };
  // CSR pattern in the natural numbering (dof = node * nv + v), columns sorted
  std::vector<int64_t> rp(n + 1, 0);
  std::vector<int> ci;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int v = 0; v < nv; v++) {
          const int64_t row = node(i, j, k) * nv + v;
          for (int c = std::max(0, k - r); c <= std::min(nz - 1, k + r); c++)
            for (int b = std::max(0, j - r); b <= std::min(ny - 1, j + r); b++)
              for (int a = std::max(0, i - r); a <= std::min(nx - 1, i + r); a++)
                for (int u = 0; u < nv; u++) ci.push_back((int)(node(a, b, c) * nv + u));
          rp[row + 1] = (int64_t)ci.size();
        }
  // internal order: ro[i] = CSR row of internal row i, cr = its inverse (CSR column -> internal column)
  std::vector<int> ro(n), cr(n);
  std::iota(ro.begin(), ro.end(), 0);
  std::mt19937_64 rng(seed);
  if (permute) std::shuffle(ro.begin(), ro.end(), rng);
  for (int64_t i = 0; i < n; i++) cr[ro[i]] = (int)i;
  std::vector<int64_t> groups{0};
  {
    std::vector<int64_t> cut;
    std::uniform_int_distribution<int64_t> d(1, std::max<int64_t>(1, n - 1));
    for (int g = 1; g < ngroups; g++) cut.push_back(d(rng));
    std::sort(cut.begin(), cut.end());
    cut.erase(std::unique(cut.begin(), cut.end()), cut.end());
    for (int64_t c : cut) if (c > 0 && c < n) groups.push_back(c);
    groups.push_back(n);
  }
  SellHost h;
  std::string err;
  if (sell_build_host(n, rp, ci, ro.data(), cr.data(), &groups, W, minbits, h, err)) return fail("build: " + err);
  const int C = 64;
  if ((int64_t)h.perm.size() != h.nsl * C || (int64_t)h.off.size() != h.nsl + 1 || h.off[h.nsl] != h.nslot)
    return fail("slice arrays");
  std::vector<char> seen(rp[n], 0), row_seen(n, 0);
  int64_t s8 = 0, s16 = 0, s32 = 0, s2 = 0;
  // the column a slot decodes to, as the product kernels decode it
  auto decode = [&](int64_t s, int k, int l) -> int64_t {
    const int b = h.base[h.off[s] / C + k];
    if (h.o2[s] >= 0) return (int64_t)b + l + ((h.c2[h.o2[s] + (int64_t)(k / 16) * C + l] >> (2 * (k % 16))) & 3u);
    if (h.o8[s] >= 0) return b + h.c8[h.o8[s] + (int64_t)(k / 4) * 4 * C + 4 * l + (k % 4)];
    if (h.o16[s] >= 0) return b + h.c16[h.o16[s] + (int64_t)(k / 2) * 2 * C + 2 * l + (k % 2)];
    return h.c32[h.o32[s] + (int64_t)k * C + l];
  };
  for (int64_t s = 0; s < h.nsl; s++) {
    const int cls = (h.o8[s] >= 0) + (h.o16[s] >= 0) + (h.o32[s] >= 0) + (h.o2[s] >= 0);
    if (cls != 1) return fail("slice " + std::to_string(s) + " has no single offset class");
    if (h.off[s + 1] - h.off[s] != (int64_t)h.wd[s] * C) return fail("slice width");
    if (h.o2[s] >= 0 && minbits > 2) return fail("lane-predicted slice below the class floor");
    (h.o2[s] >= 0 ? s2 : h.o8[s] >= 0 ? s8 : h.o16[s] >= 0 ? s16 : s32) += (int64_t)h.wd[s] * C;
    for (int l = 0; l < C; l++) {
      const int i = h.perm[s * C + l];
      if (i >= 0) {
        if (row_seen[i]++) return fail("row placed twice");
        if (h.rslot[i] != h.off[s] + l) return fail("row slot base");
        if (h.rbase[i] != rp[ro[i]]) return fail("row CSR base");
      }
      const int len = i >= 0 ? (int)(rp[ro[i] + 1] - rp[ro[i]]) : 0;
      int64_t prev = -1;
      for (int k = 0; k < h.wd[s]; k++) {
        const int64_t slot = h.off[s] + (int64_t)k * C + l;
        const unsigned t = h.pos[slot];
        if (k >= len) {
          if (t != 0xFFFFu) return fail("padding slot not marked");
          const int64_t pc = decode(s, k, l);   // gathered with value 0: must lie inside the vector
          if (pc < 0 || pc >= n) return fail("padding slot decodes outside the columns");
          continue;
        }
        if (t == 0xFFFFu || (int)t >= len) return fail("slot position out of its row");
        const int64_t q = h.rbase[i] + t;
        if (seen[q]++) return fail("CSR entry held twice");
        const int64_t col = decode(s, k, l);
        if (col != cr[ci[q]]) return fail("decoded column differs from the entry's internal column");
        if (col <= prev) return fail("row columns not increasing");
        prev = col;
      }
    }
  }
  for (int64_t i = 0; i < n; i++) if (!row_seen[i]) return fail("row missing");
  for (int64_t q = 0; q < rp[n]; q++) if (!seen[q]) return fail("CSR entry missing");
  for (size_t g = 0; g + 1 < groups.size(); g++) {
    const int64_t s0 = h.gslice[g];
    if (s0 * C >= (int64_t)h.perm.size() || h.perm[s0 * C] < groups[g] || h.perm[s0 * C] >= groups[g + 1])
      return fail("group does not start its slice");
  }
  std::printf("{\"ok\": true, \"rows\": %lld, \"nnz\": %lld, \"slices\": %lld, \"slots\": %lld, \"c8\": %lld, "
              "\"c16\": %lld, \"c32\": %lld, \"c2\": %lld, \"groups\": %zu}\n",
              (long long)n, (long long)rp[n], (long long)h.nsl, (long long)h.nslot, (long long)s8, (long long)s16,
              (long long)s32, (long long)s2, groups.size() - 1);
  return 0;
}
