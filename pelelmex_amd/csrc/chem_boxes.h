// The chemistry BoxArray of an AMR level (host code): PeleLM::resetCoveredMask builds it at
// Source/PeleLMeX_Utils.cpp:728-788 so that chemistry skips the cells a finer level covers.
//
//  * coarse level: the level's grids minus the finer level's grids coarsened to this level
//    (complementIn(domain, baf) intersected with grids[lev]), chopped to max_grid_size_chem when
//    every component of it is positive, flag 1; then the coarsened finer grids themselves, flag 0
//    (advanceChemistryBAChem converts their state but does not react them, FC = 0,
//    Source/PeleLMeX_Reactions.cpp:260-280);
//  * finest level: its grids chopped to max_grid_size_chem when set, all flag 1.
//
// The box algebra restates AMReX's published semantics (AMReX is a third-party dependency whose
// source is not in the reference tree): cell-centred boxes, coarsening by floor division,
// complement by successive box differences (slabs cut low then high end, x then y then z), and
// BoxList::maxSize, which halves the length and the chunk while both are even, splits the
// coarsened length into ceil(len / chunk) nearly equal pieces (the first len % pieces one cell
// longer, scaled back), cut from the high end, the remainder keeping the low end. The cells a
// box list covers, its flags and the size bound do not depend on how a complement is cut into
// boxes, and those are what the chemistry sees: every cell is integrated on its own.
#ifndef PR_CHEM_BOXES_H
#define PR_CHEM_BOXES_H
#include <algorithm>
#include <vector>

namespace prc {

struct IBox {
  int lo[3], hi[3];
  bool ok() const { return lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2]; }
  long cells() const { return ok() ? (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) : 0; }
};

inline int floor_div(int a, int r) { return (a >= 0) ? a / r : -((-a + r - 1) / r); }

inline IBox coarsen(const IBox& b, int r) {
  IBox c;
  for (int d = 0; d < 3; ++d) {
    c.lo[d] = floor_div(b.lo[d], r);
    c.hi[d] = floor_div(b.hi[d], r);
  }
  return c;
}

inline IBox intersect(const IBox& a, const IBox& b) {
  IBox c;
  for (int d = 0; d < 3; ++d) {
    c.lo[d] = std::max(a.lo[d], b.lo[d]);
    c.hi[d] = std::min(a.hi[d], b.hi[d]);
  }
  return c;
}

// a minus b as disjoint boxes: slabs of a below / above b cut off along x, then y, then z
inline void box_diff(const IBox& a, const IBox& b, std::vector<IBox>& out) {
  if (!intersect(a, b).ok()) {
    out.push_back(a);
    return;
  }
  IBox r = a;
  for (int d = 0; d < 3; ++d) {
    if (r.lo[d] < b.lo[d]) {
      IBox s = r;
      s.hi[d] = b.lo[d] - 1;
      out.push_back(s);
      r.lo[d] = b.lo[d];
    }
    if (r.hi[d] > b.hi[d]) {
      IBox s = r;
      s.lo[d] = b.hi[d] + 1;
      out.push_back(s);
      r.hi[d] = b.hi[d];
    }
  }
}

// the cells of `domain` no box of `cover` holds
inline std::vector<IBox> complement_in(const IBox& domain, const std::vector<IBox>& cover) {
  std::vector<IBox> cur{domain}, nxt;
  for (const IBox& c : cover) {
    nxt.clear();
    for (const IBox& b : cur) box_diff(b, c, nxt);
    cur.swap(nxt);
  }
  return cur;
}

// every non-empty intersection of a box of `a` with a box of `b`, a's boxes outermost
inline std::vector<IBox> intersect_lists(const std::vector<IBox>& a, const std::vector<IBox>& b) {
  std::vector<IBox> out;
  for (const IBox& x : a)
    for (const IBox& y : b) {
      const IBox c = intersect(x, y);
      if (c.ok()) out.push_back(c);
    }
  return out;
}

// BoxList::maxSize: no box longer than chunk[d] cells in direction d
inline void max_size(std::vector<IBox>& boxes, const int chunk[3]) {
  for (int d = 0; d < 3; ++d) {
    std::vector<IBox> added;
    for (IBox& bx : boxes) {
      const int len = bx.hi[d] - bx.lo[d] + 1;
      if (len <= chunk[d]) continue;
      int ratio = 1, bs = chunk[d], nlen = len;
      while (bs % 2 == 0 && nlen % 2 == 0) {
        ratio *= 2;
        bs /= 2;
        nlen /= 2;
      }
      const int numblk = nlen / bs + (nlen % bs ? 1 : 0);
      const int sz = nlen / numblk, extra = nlen % numblk;
      for (int k = 0; k < numblk - 1; ++k) {
        const int ksize = (k < extra ? sz + 1 : sz) * ratio;
        const int pos = bx.hi[d] - ksize + 1;  // the piece [pos, hi] leaves, bx keeps [lo, pos - 1]
        IBox hi_part = bx;
        hi_part.lo[d] = pos;
        bx.hi[d] = pos - 1;
        added.push_back(hi_part);
      }
    }
    boxes.insert(boxes.end(), added.begin(), added.end());
  }
}

// The chemistry boxes of a level and their flags (1: react, 0: covered by the finer level).
// fine: the finer level's grids at its own resolution (empty on the finest level), coarsened
// here by ref_ratio; max_chunk: max_grid_size_chem (chopping applies when all three are > 0).
inline void chem_box_array(const IBox& domain, const std::vector<IBox>& grids, const std::vector<IBox>& fine,
                           int ref_ratio, const int max_chunk[3], bool finest, std::vector<IBox>& boxes,
                           std::vector<int>& flags) {
  const bool chop = max_chunk[0] > 0 && max_chunk[1] > 0 && max_chunk[2] > 0;
  boxes.clear();
  flags.clear();
  if (finest) {
    boxes = grids;
    if (chop) max_size(boxes, max_chunk);
    flags.assign(boxes.size(), 1);
    return;
  }
  std::vector<IBox> baf;
  baf.reserve(fine.size());
  for (const IBox& f : fine) baf.push_back(coarsen(f, ref_ratio));
  std::vector<IBox> unc = intersect_lists(complement_in(domain, baf), grids);
  if (chop) max_size(unc, max_chunk);
  boxes = unc;
  flags.assign(unc.size(), 1);
  boxes.insert(boxes.end(), baf.begin(), baf.end());
  flags.insert(flags.end(), baf.size(), 0);
}

}  // namespace prc
#endif
