// sell_host.cpp -- host construction of the SELL-64 layout (see sell_host.h).
#include "sell_host.h"
#include <algorithm>
#include <climits>
#include <cstdint>
#include <utility>

int sell_build_host(int64_t n, const std::vector<int64_t> &rp, const std::vector<int> &ci, const int *ro,
                    const int *cr, const std::vector<int64_t> *groups, int W, int minbits, SellHost &out,
                    std::string &err) {
  const int C = 64;
  if (minbits != 2 && minbits != 8 && minbits != 16 && minbits != 32) {
    err = "SELL: offset class must be 2, 8, 16 or 32 bits";
    return -1;
  }
  if ((int64_t)rp.size() != n + 1) { err = "SELL build without host pattern"; return -5; }
  auto crow = [&](int64_t i) -> int64_t { return ro ? ro[i] : i; };
  auto rlen = [&](int64_t i) -> int { int64_t r = crow(i); return (int)(rp[r + 1] - rp[r]); };
  std::vector<int64_t> gb = groups ? *groups : std::vector<int64_t>{0, n};
  if (gb.front() != 0 || gb.back() != n) { err = "SELL row groups do not cover the rows"; return -5; }
  // rows of each group, length-sorted inside windows of W that do not cross a group boundary; each group starts a
  // new slice (its last slice is padded with empty lanes)
  std::vector<int> perm;
  perm.reserve(n + (int64_t)C * gb.size());
  std::vector<int64_t> gslice(gb.size(), 0);
  for (size_t g = 0; g + 1 < gb.size(); g++) {
    gslice[g] = (int64_t)perm.size() / C;
    const size_t start = perm.size();
    for (int64_t i = gb[g]; i < gb[g + 1]; i++) perm.push_back((int)i);
    for (size_t w0 = start; w0 < perm.size(); w0 += W) {
      size_t w1 = std::min(perm.size(), w0 + W);
      std::stable_sort(perm.begin() + w0, perm.begin() + w1, [&](int a, int b) { return rlen(a) > rlen(b); });
    }
    while (perm.size() % C) perm.push_back(-1);
  }
  gslice.back() = (int64_t)perm.size() / C;
  const int64_t nsl = (int64_t)perm.size() / C;
  if (nsl >= ((int64_t)1 << 31) / C) { err = "SELL: too many rows"; return -1; }
  for (int64_t i = 0; i < n; i++)
    if (rlen(i) >= 0xFFFF) { err = "SELL: a row has more than 65,534 entries"; return -1; }
  std::vector<int64_t> off(nsl + 1, 0);
  std::vector<int> wd(nsl);
  for (int64_t s = 0; s < nsl; s++) {
    int w = 0;
    for (int l = 0; l < C; l++) {
      const int r = perm[s * C + l];
      if (r >= 0) w = std::max(w, rlen(r));
    }
    wd[s] = w;
    off[s + 1] = off[s] + (int64_t)w * C;
  }
  const int64_t nslot = off[nsl];
  std::vector<int> absc(nslot, -1), base(nslot / C, 0);
  std::vector<unsigned short> pos(nslot, 0xFFFF);
  std::vector<int64_t> rbase(n, 0), rslot(n, -1);
  std::vector<char> wide(nsl, 0), narrow(nsl, 1), pred(nsl, minbits <= 2 ? 1 : 0);
  std::vector<int> pbase(minbits <= 2 ? nslot / C : 0, 0);
#pragma omp parallel
  {
    std::vector<std::pair<int, int>> ent;
#pragma omp for schedule(dynamic, 64)
    for (int64_t s = 0; s < nsl; s++) {
      for (int l = 0; l < C; l++) {
        int i = perm[s * C + l];
        if (i < 0) continue;
        int64_t r = crow(i);
        rbase[i] = rp[r];
        rslot[i] = off[s] + l;
        ent.clear();
        for (int64_t q = rp[r]; q < rp[r + 1]; q++) ent.emplace_back(cr ? cr[ci[q]] : ci[q], (int)(q - rp[r]));
        std::sort(ent.begin(), ent.end());
        for (size_t k = 0; k < ent.size(); k++) {
          int64_t slot = off[s] + (int64_t)k * C + l;
          absc[slot] = ent[k].first;
          pos[slot] = (unsigned short)ent[k].second;
        }
      }
      for (int k = 0; k < wd[s]; k++) {
        int mn = INT32_MAX, mx = -1;
        for (int l = 0; l < C; l++) {
          int cc = absc[off[s] + (int64_t)k * C + l];
          if (cc < 0) continue;
          mn = std::min(mn, cc);
          mx = std::max(mx, cc);
        }
        base[off[s] / C + k] = mn;
        if (mx - mn > 65535) wide[s] = 1;
        if (mx - mn > 255) narrow[s] = 0;
        if (!pred[s]) continue;
        // lane-predicted class: every entry's column is pb + l + d with d in [0, 3] (pb = the smallest
        // column - lane over the position); a padding slot takes the first d whose column lies among the
        // position's real columns, and the slice leaves the class when none does
        int pmn = INT32_MAX, pmx = INT32_MIN;
        for (int l = 0; l < C; l++) {
          int cc = absc[off[s] + (int64_t)k * C + l];
          if (cc < 0) continue;
          pmn = std::min(pmn, cc - l);
          pmx = std::max(pmx, cc - l);
        }
        if (pmx - pmn > 3) { pred[s] = 0; continue; }
        for (int l = 0; l < C && pred[s]; l++) {
          if (absc[off[s] + (int64_t)k * C + l] >= 0) continue;
          bool any = false;
          for (int d = 0; d <= 3 && !any; d++) any = pmn + l + d >= mn && pmn + l + d <= mx;
          if (!any) pred[s] = 0;
        }
        pbase[off[s] / C + k] = pmn;
      }
    }
  }
  // each slice in the narrowest class it allows, but never narrower than minbits (8: no lane-predicted slices,
  // 16: no 8-bit slices either, 32: absolute columns everywhere)
  for (int64_t s = 0; s < nsl; s++) {
    if (minbits == 32) wide[s] = 1;
    if (minbits >= 16) narrow[s] = 0;
    if (wide[s]) pred[s] = 0;
  }
  std::vector<int64_t> o32(nsl, -1), o8(nsl, -1), o16(nsl, -1), o2(nsl, -1);
  int64_t n32 = 0, n8 = 0, n16 = 0, n2 = 0;
  for (int64_t s = 0; s < nsl; s++) {
    if (pred[s]) {
      o2[s] = n2 / 16;  // in 32-bit words
      n2 += (int64_t)((wd[s] + 15) / 16) * 16 * C;
      for (int k = 0; k < wd[s]; k++) base[off[s] / C + k] = pbase[off[s] / C + k];
    } else if (wide[s]) { o32[s] = n32; n32 += (int64_t)wd[s] * C; }
    else if (narrow[s]) { o8[s] = n8; n8 += (int64_t)((wd[s] + 3) / 4) * 4 * C; }
    else { o16[s] = n16; n16 += (int64_t)((wd[s] + 1) / 2) * 2 * C; }
  }
  std::vector<unsigned> c2(std::max<int64_t>(n2 / 16, 1), 0u);
  std::vector<unsigned short> c16(std::max<int64_t>(n16, 1), 0);
  std::vector<int> c32(std::max<int64_t>(n32, 1), 0);
  std::vector<unsigned char> c8(std::max<int64_t>(n8, 1), 0);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < nsl; s++)
    for (int k = 0; k < wd[s]; k++) {
      const int b = base[off[s] / C + k];
      if (o2[s] >= 0) {
        // lane-predicted: 2-bit residual of position k, lane l in bits 2 (k % 16) of word
        // c2[o2[s] + 64 (k / 16) + l]; a padding slot takes the first residual whose column is a real one of
        // the position (its value is 0)
        int mn = INT32_MAX, mx = INT32_MIN;
        for (int l = 0; l < C; l++) {
          const int cc = absc[off[s] + (int64_t)k * C + l];
          if (cc >= 0) { mn = std::min(mn, cc); mx = std::max(mx, cc); }
        }
        for (int l = 0; l < C; l++) {
          const int cc = absc[off[s] + (int64_t)k * C + l];
          int d = cc >= 0 ? cc - b - l : 0;
          if (cc < 0)
            while (d < 3 && !(b + l + d >= mn && b + l + d <= mx)) d++;
          c2[o2[s] + (int64_t)(k / 16) * C + l] |= (unsigned)d << (2 * (k % 16));
        }
        continue;
      }
      for (int l = 0; l < C; l++) {
        int64_t slot = off[s] + (int64_t)k * C + l;
        const int cc = absc[slot] >= 0 ? absc[slot] : b;  // padding: the position's base column, value 0
        // 8-bit (16-bit): positions 4g..4g+3 (2g, 2g+1) of lane l packed in one 32-bit word, the words of a
        // group lane-contiguous
        if (wide[s]) c32[o32[s] + (int64_t)k * C + l] = cc;
        else if (o8[s] >= 0) c8[o8[s] + (int64_t)(k / 4) * 4 * C + 4 * l + (k % 4)] = (unsigned char)(cc - b);
        else c16[o16[s] + (int64_t)(k / 2) * 2 * C + 2 * l + (k % 2)] = (unsigned short)(cc - b);
      }
    }
  out.nsl = nsl; out.nslot = nslot; out.n8 = n8; out.n16 = n16; out.n32 = n32; out.n2 = n2;
  out.o2.swap(o2); out.c2.swap(c2);
  out.perm.swap(perm); out.gslice.swap(gslice); out.off.swap(off); out.wd.swap(wd); out.base.swap(base);
  out.pos.swap(pos); out.rbase.swap(rbase); out.rslot.swap(rslot);
  out.o8.swap(o8); out.o16.swap(o16); out.o32.swap(o32); out.c8.swap(c8); out.c16.swap(c16); out.c32.swap(c32);
  return 0;
}
