// egather_host.h -- host plan of the element-buffer gather (no device code; the library uploads the plan, the
// host checker emulates the gather kernel on it).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

// LDS of one gather block: staged element rows (FP64), the contribution stream (16-bit), per slot the start
// (16-bit) and the count (8-bit), the element row ids (32-bit, from the next 8-byte boundary); offsets oc, os, on
// of the stream, the starts and the counts; returns the total
inline size_t egather_lds(int maxseg, int nl, int maxctr, int maxit, size_t *oc = nullptr, size_t *os = nullptr,
                          size_t *on = nullptr) {
  const size_t a = (size_t)maxseg * nl * 8, b = ((size_t)maxctr * 2 + 7) / 8 * 8, c = ((size_t)maxit * 2 + 7) / 8 * 8;
  if (oc) *oc = a;
  if (os) *os = a + b;
  if (on) *on = a + b + c;
  return a + b + c + ((size_t)maxit + 7) / 8 * 8 + (size_t)maxseg * 4;
}

// Gather of one level.  Element rows ("segments" r = e NL + a) belong to the SELL row of their DoF, hence to one
// row block: lg consecutive lanes of one slice (block g: slice g / (64 / lg), lanes lg (g % (64 / lg)) ...).
// seg[lseg[lg g + l] .. lseg[lg g + l + 1]) are the segments of lane l of block g, by element row; block g's
// contribution stream is cidx[bbase[g] .. bbase[g + 1]), ordered by slot ((position, lane) order) and, within a
// slot, by element row; slot s sums the cnt[s] entries from cidx[bbase[g] + cst[s]] on, each an LDS position
// (segment position within the block) NL + column.
struct EGatherPlan {
  int lg = 0, maxseg = 0, maxctr = 0, maxit = 0;
  int64_t nblk = 0, nseg = 0, nctr = 0;
  std::vector<int64_t> lseg, bbase;  // nblk lg + 1, nblk + 1
  std::vector<int> seg;              // nseg
  std::vector<unsigned char> cnt;    // nslot
  std::vector<unsigned short> cst;   // nslot
  std::vector<unsigned short> cidx;  // nctr
};

// pos[r NL + c]: position of entry (r, c) in its SELL row (0xFFFF: none); slot0[r]: slot of position 0 of the row
// of element row r (-1: none); off[nsl + 1] / wd[nsl]: first slot and width of each slice.  The lane group is the
// widest of 16, 8, 4, 2, 1 whose blocks need at most max_lds bytes.  Returns 0, or -1 with err set.
int egather_plan(int64_t nel, int nl, const std::vector<unsigned short> &pos, const std::vector<int64_t> &slot0,
                 const std::vector<int64_t> &off, const std::vector<int> &wd, size_t max_lds, EGatherPlan &p,
                 std::string &err);
