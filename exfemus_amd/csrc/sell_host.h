// sell_host.h -- host construction of the SELL-64 copy of a CSR pattern (no device code; the library
// uploads the result, the host layout checker verifies it).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

// SELL-64 layout of a CSR pattern in internal DoF order: internal row i holds CSR row ro[i] (ro == nullptr:
// identity), CSR column j becomes internal column cr[j] (nullptr: identity).  Rows are length-sorted inside
// windows of W rows that never cross a row group (groups: boundaries in internal rows, nullptr: one group);
// each group starts a new slice.  The entries of a row are sorted by internal column; columns are stored per
// slice in the narrowest class it allows, but never in a class narrower than minbits (2, 8, 16 or 32: 32
// stores absolute columns in every slice):
//   2 bits, lane-predicted: every column of position k is base[k] + lane + d with d in [0, 3] (base[k] = the
//     smallest column - lane over the position; structured-grid slices, whose 64 rows are consecutive nodes of
//     one node class, have this form);
//   8 / 16 bits: base[k] = the smallest column of position k, offsets up to 255 / 65,535;
//   absolute 32-bit columns otherwise.
struct SellHost {
  int64_t nsl = 0, nslot = 0, n8 = 0, n16 = 0, n32 = 0, n2 = 0;
  std::vector<int> perm;          // 64 nsl: internal row of each lane (-1: padding)
  std::vector<int64_t> gslice;    // first slice of each group (+ end)
  std::vector<int64_t> off;       // nsl + 1: first slot of each slice
  std::vector<int> wd;            // nsl: positions per slice
  std::vector<int> base;          // nslot / 64: smallest column of each (slice, position)
  std::vector<unsigned short> pos;  // nslot: position of the slot's entry within its CSR row (0xFFFF: padding)
  std::vector<int64_t> rbase;     // n: CSR row start of internal row i
  std::vector<int64_t> rslot;     // n: slot of position 0 of internal row i
  std::vector<int64_t> o8, o16, o32;  // nsl: offset of the slice in c8 / c16 / c32 (-1: other class)
  std::vector<int64_t> o2;        // nsl: offset of the slice in c2, in 32-bit words (-1: other class)
  std::vector<unsigned> c2;       // 2-bit lane-predicted residuals, positions 16g..16g+15 of a lane in one word
  std::vector<unsigned char> c8;  // 8-bit offsets, positions 4g..4g+3 of a lane packed in one 32-bit word
  std::vector<unsigned short> c16;  // 16-bit offsets, positions 2g, 2g+1 of a lane packed in one word
  std::vector<int> c32;           // absolute columns
};

// returns 0, or -1 (sizes beyond the format) / -5 (inconsistent input) with err set
int sell_build_host(int64_t n, const std::vector<int64_t> &rp, const std::vector<int> &ci, const int *ro,
                    const int *cr, const std::vector<int64_t> *groups, int W, int minbits, SellHost &out,
                    std::string &err);
