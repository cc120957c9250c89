// egather_host.cpp -- host plan of the element-buffer gather (egather_host.h); no device code.
//
// Every element row (e, a) -- a "segment" of NL entries of the element-matrix buffer -- belongs to the SELL row of
// its DoF, hence to one row block (lg lanes of one slice).  Per block the segments are listed by lane, then
// element row; per slot the contributions (entries (e, a, b) with slot(e, a, b) = s) are stored as LDS positions
// in the same order, so a slot sums its contributions in element order whatever the launch does.
#include "egather_host.h"
#include <algorithm>

int egather_plan(int64_t nel, int nl, const std::vector<unsigned short> &pos, const std::vector<int64_t> &slot0,
                 const std::vector<int64_t> &off, const std::vector<int> &wd, size_t max_lds, EGatherPlan &p,
                 std::string &err) {
  const int64_t nsl = (int64_t)wd.size(), nslot = off.back(), nrow_el = nel * nl;
  // slice and lane of every element row (position 0 of its SELL row)
  std::vector<int> rs(nrow_el, -1), rl(nrow_el, -1);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < nrow_el; r++) {
    if (slot0[r] < 0) continue;
    const int64_t s = std::upper_bound(off.begin(), off.end(), slot0[r]) - off.begin() - 1;
    rs[r] = (int)s;
    rl[r] = (int)(slot0[r] - off[s]);
  }
  for (int64_t r = 0; r < nrow_el; r++)
    if (slot0[r] >= 0 && (rl[r] < 0 || rl[r] >= 64)) { err = "element-buffer gather: row slot outside its slice"; return -1; }
  // element rows and contributions per SELL lane ((block, lane) = (slice, lane) in lane order, for any lg)
  std::vector<int64_t> lcount((size_t)nsl * 64 + 1, 0), lctr((size_t)nsl * 64, 0);
  for (int64_t r = 0; r < nrow_el; r++) {
    if (rs[r] < 0) continue;
    const size_t q = (size_t)rs[r] * 64 + rl[r];
    lcount[q + 1]++;
    for (int c = 0; c < nl; c++) lctr[q] += pos[(size_t)r * nl + c] != 0xFFFFu;
  }
  // lane group: the widest whose blocks stage at most max_lds bytes (element rows, contribution stream, and
  // count + start of every slot)
  int lg = 16, maxseg = 0, maxctr = 0, maxit = 0;
  for (;; lg /= 2) {
    maxseg = maxctr = maxit = 0;
    for (int64_t b = 0; b < nsl * (64 / lg); b++) {
      int64_t ns = 0, nc = 0;
      for (int q = 0; q < lg; q++) { ns += lcount[(size_t)b * lg + q + 1]; nc += lctr[(size_t)b * lg + q]; }
      maxseg = (int)std::max<int64_t>(maxseg, ns);
      maxctr = (int)std::max<int64_t>(maxctr, nc);
      maxit = std::max(maxit, wd[b / (64 / lg)] * lg);
    }
    if (egather_lds(maxseg, nl, maxctr, maxit) <= max_lds) break;
    if (lg == 1) { err = "element-buffer gather: the element rows of one SELL row exceed the LDS"; return -1; }
  }
  const int64_t nblk = nsl * (64 / lg);
  // segments by (block, lane) = (slice, lane), then element row
  std::vector<int64_t> lseg(lcount);
  for (size_t i = 1; i < lseg.size(); i++) lseg[i] += lseg[i - 1];
  const int64_t nseg = lseg.back();
  std::vector<int> seg(nseg), segpos(nrow_el, -1);
  {
    std::vector<int64_t> fill(lseg.begin(), lseg.end() - 1);
    for (int64_t r = 0; r < nrow_el; r++) {
      if (rs[r] < 0) continue;
      const int64_t q = fill[(size_t)rs[r] * 64 + rl[r]]++;
      seg[q] = (int)r;
      const int64_t blk = ((int64_t)rs[r] * 64 + rl[r]) / lg;
      segpos[r] = (int)(q - lseg[(size_t)blk * lg]);
    }
  }
  // per block: contributions per slot, their order (slot position in (k, lane) order), the block totals
  std::vector<unsigned char> cnt(nslot, 0);
  std::vector<unsigned short> cst(nslot, 0);
  std::vector<int64_t> btot(nblk + 1, 0);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
  for (int64_t b = 0; b < nblk; b++) {
    const int64_t s = b / (64 / lg);
    const int l0 = (int)(b % (64 / lg)) * lg;
    for (int64_t q = lseg[(size_t)b * lg]; q < lseg[(size_t)(b + 1) * lg]; q++) {
      const int64_t r = seg[q];
      for (int c = 0; c < nl; c++) {
        const unsigned k = pos[(size_t)r * nl + c];
        if (k == 0xFFFFu) continue;
        if ((int)k >= wd[s]) { bad |= 4; continue; }
        const int64_t sl = slot0[r] + 64 * (int64_t)k;
        if (cnt[sl] == 255) bad |= 1;
        else cnt[sl]++;
      }
    }
    int64_t run = 0;
    for (int k = 0; k < wd[s]; k++)
      for (int q = 0; q < lg; q++) {
        const int64_t sl = off[s] + 64 * (int64_t)k + l0 + q;
        if (run > 65535) bad |= 2;
        cst[sl] = (unsigned short)run;
        run += cnt[sl];
      }
    if (run > 65536) bad |= 2;
    btot[b + 1] = run;
  }
  if (bad) {
    err = bad & 4 ? "element-buffer gather: an entry lies beyond its slice" :
                    "element-buffer gather: a slot or block holds too many contributions";
    return -1;
  }
  std::vector<int64_t> bbase(nblk + 1, 0);
  for (int64_t b = 0; b < nblk; b++) bbase[b + 1] = bbase[b] + btot[b + 1];
  const int64_t nctr = bbase[nblk];
  std::vector<unsigned short> cidx(nctr);
#pragma omp parallel
  {
    std::vector<unsigned char> fill;
#pragma omp for schedule(dynamic, 64)
    for (int64_t b = 0; b < nblk; b++) {
      const int64_t s = b / (64 / lg);
      const int l0 = (int)(b % (64 / lg)) * lg;
      // per-slot fill counters of this block's slots, indexed (k, lane)
      fill.assign((size_t)wd[s] * lg, 0);
      for (int64_t q = lseg[(size_t)b * lg]; q < lseg[(size_t)(b + 1) * lg]; q++) {
        const int64_t r = seg[q];
        const int lane = rl[r] - l0;
        for (int c = 0; c < nl; c++) {
          const unsigned k = pos[(size_t)r * nl + c];
          if (k == 0xFFFFu) continue;
          const int64_t sl = slot0[r] + 64 * (int64_t)k;
          const int64_t at = bbase[b] + cst[sl] + fill[(size_t)k * lg + lane]++;
          cidx[at] = (unsigned short)(segpos[r] * nl + c);
        }
      }
    }
  }
  p = EGatherPlan();
  p.lg = lg; p.maxseg = maxseg; p.maxctr = maxctr; p.maxit = maxit; p.nblk = nblk; p.nseg = nseg; p.nctr = nctr;
  p.lseg = std::move(lseg); p.bbase = std::move(bbase); p.seg = std::move(seg);
  p.cnt = std::move(cnt); p.cst = std::move(cst); p.cidx = std::move(cidx);
  return 0;
}
