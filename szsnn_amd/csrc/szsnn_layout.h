// szsnn_layout.h -- host-side builders of the device formats (plain C++, no HIP).
//
// Replaces the reference's SynapseIndex / GroupData / SynapseData builders
// (szsnn1.1/src/SynapseIndex.cpp:42-63, GroupData.cpp:70-228,
// SynapseData.ipp:41-111):
//   * delay-bucketed CSR: rows keyed by (id, API delay k), row = id*D + k,
//     synapses in input order within a row, each (target, int32 fixed-point
//     weight); zero weights are inert and dropped.  The device push format
//     is the same array with per-id ranges (an id's rows are contiguous) and
//     the delay k packed above the target (target | k << kSynDelayShift);
//   * the delay bitmap per id (bit k set iff row (id, k) is nonempty, the
//     reference's bitmap, SynapseData.ipp:70-83);
//   * the pull layout for dense gathers, a sliced ELLPACK: targets in blocks
//     of 64 * pull_slices (4 to 8 slices, chosen per network by the runtime);
//     per (block, delay k) segment the targets are sorted by their number of
//     delay-k synapses and cut into pull_slices slices of 64 (one target per
//     lane); a slice holds L entries per lane (its longest
//     lane, rounded up to whole 16-byte vectors), stored as vectors
//     [L / V][64 lanes][V entries], so the wave's j-th 16-byte load is
//     contiguous; each lane also records how many vectors hold its own
//     entries (the rest of its column is padding and is not read: loads past
//     a lane's own vectors read one null vector of padding entries stored
//     after the last slice, or, in the count format, an out-of-range offset
//     of a buffer load, which reads zeros).  Entries
//     name the source relative to the block's first staged id, in one of
//     three formats (kPullWide, kPullCount, kPullTable in szsnn_constants.h):
//     8-byte (offset, weight); with at most kPullCountFields distinct weights
//     2-byte running deltas of the offset with a weight tag (skip entries for
//     wider gaps); with at most kPullMaxWeights, 4-byte offset | weight
//     index.  The weight table's index 0 is the inert 0.
// Shared by the runtime and by the CPU layout check (tests/cpp/layout_check.cpp).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/szsnn_hip.h"
#include "szsnn_constants.h"

namespace szsnn {

struct LayoutPair {
  uint32_t x, y;
};

struct HostLayout {
  std::vector<uint32_t> row_off;      // [T*D + 1]
  std::vector<LayoutPair> syn;        // [S] push entries (target | k << kSynDelayShift, weight)
  std::vector<uint32_t> syn_off;      // [T + 1] per-id ranges of syn
  uint64_t active = 0;                // ids with at least one synapse
  std::vector<uint64_t> delay_bits;   // [T]
  uint64_t nonempty = 0;              // nonempty rows
  uint32_t pull_slices = kPullSlices; // slices per (block, delay) segment (set before building)
  std::vector<uint32_t> pull_off;     // [NB*D*pull_slices + 1] slice starts (entries), or empty
  std::vector<uint32_t> pull_lo;      // [NB] first staged raster word of each target block
  std::vector<uint16_t> pull_lane;    // [NB*D][64*pull_slices] slice lane: target (within the block) | vectors << kPullShift
  std::vector<LayoutPair> pull_syn;   // wide entries (8 bytes), or empty
  std::vector<uint32_t> pull_syn32;   // table-format entries (4 bytes), or empty
  std::vector<uint16_t> pull_syn16;   // count-format entries (2 bytes), or empty
  std::vector<int32_t> pull_wtab;     // count, table: weight of each index (index 0: the inert 0)
  uint32_t pull_fmt = kPullWide;      // entry format (kPullWide, kPullCount, kPullTable)
  uint32_t pull_words = 0;
  uint64_t NB = 0;
  std::string error;
};

// Rows, delay bits and the pull layout (targets still unpacked in L.syn).
// Delay groups of the pull gather for NB target blocks: as many (target
// block, group) work items as `slots` resident gather blocks hold in one
// round, the delays split evenly between the groups where that keeps at
// least half of them (equal items end together).
inline uint32_t pull_groups_for(uint64_t NB, uint64_t D, uint64_t slots) {
  const uint64_t want =
      std::max<uint64_t>(1, std::min<uint64_t>({slots / std::max<uint64_t>(NB, 1), D, (uint64_t)kMaxPullGroups}));
  uint64_t even = want;
  while (D % even) --even;
  return (uint32_t)(2 * even >= want ? even : want);
}

// Slices per pull segment for N targets (blocks of 64 * slices targets).  A
// gather launch's time follows its busiest CU, which runs about ceil(items /
// CUs) work items of slices * ceil(D / groups) columns each (a column: one
// slice of one delay); the count taken needs the fewest rounds of the
// resident slots, then the fewest such columns, then the most slices (fewer,
// longer items), among the counts that keep the delay groups of
// kPullSlices.  For 100000 targets, 32 delays and 1792 slots on 256 CUs:
// 8 slices give 1568 items (7 on some CUs, 6 on others) and 7 slices give
// 1792 (7 on every CU), 7 * 7 * 4 columns against 7 * 8 * 4.
inline uint32_t choose_pull_slices(uint64_t N, uint64_t D, uint64_t slots, uint64_t cus) {
  auto blocks = [&](uint32_t s) { return (N + 64 * s - 1) / (64 * s); };
  const uint32_t g_max = pull_groups_for(blocks(kPullSlices), D, slots);
  uint32_t best = kPullSlices;
  uint64_t best_rounds = 0, best_cols = 0;
  for (uint32_t s = kPullSlices; s >= kPullMinSlices; --s) {
    const uint32_t g = pull_groups_for(blocks(s), D, slots);
    if (g < g_max) continue;
    const uint64_t items = blocks(s) * g;
    const uint64_t rounds = (items + slots - 1) / slots;
    const uint64_t cols = (items + cus - 1) / cus * s * ((D + g - 1) / g);
    if (s == kPullSlices || rounds < best_rounds || (rounds == best_rounds && cols < best_cols)) {
      best = s;
      best_rounds = rounds;
      best_cols = cols;
    }
  }
  return best;
}

// choose_slices(fmt, weight-table entries, bound on the staged words), if
// given, sets the slice count once the entry format is known.
using SliceChooser = std::function<uint32_t(uint32_t, uint32_t, uint32_t)>;

inline bool build_rows_and_pull(const szsnn_net_desc& d, HostLayout& L, const SliceChooser& choose_slices = nullptr) {
  const uint64_t N = d.n_neurons, M = d.n_sources, T = N + M, D = d.max_delay;
  const uint64_t nrows = T * D;
  const double scale = std::ldexp(1.0, d.frac_bits);
  L.row_off.assign(nrows + 1, 0);
  std::vector<int32_t> wfix(d.n_syn);
  for (uint64_t s = 0; s < d.n_syn; ++s) {
    const uint32_t src = d.syn_src[s], tgt = d.syn_tgt[s], k = d.syn_delay[s];
    if (src >= T || tgt >= N) {
      std::ostringstream os;
      os << "synapse " << s << " connects nonexisting ids source=" << src << " target=" << tgt;
      L.error = os.str();
      return false;
    }
    if (k >= D) {
      std::ostringstream os;
      os << "synapse " << s << " delay " << k << " >= max_delay " << D;
      L.error = os.str();
      return false;
    }
    const double q = std::nearbyint((double)d.syn_w[s] * scale);
    if (!(std::fabs(q) < 2147483648.0)) {
      std::ostringstream os;
      os << "weight " << d.syn_w[s] << " does not fit int32 at frac_bits " << d.frac_bits;
      L.error = os.str();
      return false;
    }
    wfix[s] = (int32_t)q;
    if (wfix[s] != 0) L.row_off[(uint64_t)src * D + k + 1]++;
  }
  for (uint64_t r = 0; r < nrows; ++r) {
    const uint64_t next = (uint64_t)L.row_off[r + 1] + L.row_off[r];
    if (next > 0xffffffffull) {
      L.error = "more than 2^32-1 synapses";
      return false;
    }
    L.row_off[r + 1] = (uint32_t)next;
  }
  const uint64_t S = L.row_off[nrows];
  L.syn.assign(S, LayoutPair{0, 0});
  {
    std::vector<uint32_t> pos(L.row_off.begin(), L.row_off.end() - 1);
    for (uint64_t s = 0; s < d.n_syn; ++s) {
      if (wfix[s] == 0) continue;
      const uint64_t r = (uint64_t)d.syn_src[s] * D + d.syn_delay[s];
      L.syn[pos[r]++] = LayoutPair{d.syn_tgt[s], (uint32_t)wfix[s]};
    }
  }
  std::vector<int32_t>().swap(wfix);
  L.delay_bits.assign(T, 0);
  L.nonempty = 0;
  for (uint64_t id = 0; id < T; ++id)
    for (uint64_t k = 0; k < D; ++k)
      if (L.row_off[id * D + k + 1] != L.row_off[id * D + k]) {
        L.delay_bits[id] |= 1ull << k;
        ++L.nonempty;
      }

  // pull layout: target blocks of TB = 64 * pull_slices targets; per (target
  // block, delay) segment, pull_slices slices of 64 targets (targets sorted
  // by their synapse count in the segment, so the slices' lengths are
  // close); slice s holds L_s entries per lane, stored as 16-byte vectors
  // [L_s / V][64 lanes][V entries] (V = 8 count, 4 table, 2 wide)
  L.pull_off.clear();
  L.pull_lo.clear();
  L.pull_lane.clear();
  L.pull_syn.clear();
  L.pull_syn32.clear();
  L.pull_syn16.clear();
  L.pull_wtab.clear();
  L.pull_fmt = kPullWide;
  L.pull_words = 0;
  auto slices_ok = [&]() {
    if (L.pull_slices >= kPullMinSlices && L.pull_slices <= kPullSlices) return true;
    L.error = "pull layout: slices per segment out of range";
    return false;
  };
  if (!slices_ok()) return false;
  L.NB = (N + 64 * L.pull_slices - 1) / (64 * L.pull_slices);
  if (d.pull_layout < 0 || S == 0) return true;
  // distinct nonzero weights (up to kPullMaxWeights; more: wide entries),
  // in-degrees and the largest weight, which do not depend on the blocks
  std::vector<uint32_t> indeg(N, 0);
  std::vector<int32_t> wset;
  bool narrow = d.pull_layout == 0;
  int64_t wmax = 0;
  for (uint64_t q = 0; q < S; ++q) {
    indeg[L.syn[q].x]++;
    const int32_t w = (int32_t)L.syn[q].y;
    wmax = std::max<int64_t>(wmax, w < 0 ? -(int64_t)w : (int64_t)w);
    if (narrow && std::find(wset.begin(), wset.end(), w) == wset.end()) {
      wset.push_back(w);
      if (wset.size() > kPullMaxWeights) narrow = false;
    }
  }
  const uint32_t fmt = !narrow ? kPullWide : wset.size() <= kPullCountFields ? kPullCount : kPullTable;
  // the slice count: the caller's choice for this format and size, if it
  // makes one
  if (choose_slices) {
    const uint32_t words_bound = (uint32_t)std::min<uint64_t>((T + 63) / 64, kPullMaxWords);
    L.pull_slices = choose_slices(fmt, (uint32_t)wset.size() + 1, words_bound);
    if (!slices_ok()) return false;
  }
  const uint32_t SL = L.pull_slices, TB = 64 * SL;
  L.NB = (N + TB - 1) / TB;
  const uint64_t NB = L.NB;
  const uint64_t nseg = NB * D;
  std::vector<uint32_t> cnt(nseg * TB, 0);  // synapses per (segment, target)
  std::vector<uint32_t> smin(NB, 0xffffffffu), smax(NB, 0);
  for (uint64_t id = 0; id < T; ++id)
    for (uint64_t k = 0; k < D; ++k)
      for (uint32_t q = L.row_off[id * D + k]; q < L.row_off[id * D + k + 1]; ++q) {
        const uint32_t tg = L.syn[q].x;
        const uint64_t b = tg / TB;
        cnt[(b * D + k) * TB + (tg - b * TB)]++;
        smin[b] = std::min(smin[b], (uint32_t)id);
        smax[b] = std::max(smax[b], (uint32_t)id);
      }
  // the gather sums a target's weights exactly in FP64: |sum| < 2^53
  uint32_t max_indeg = 0;
  for (uint32_t c : indeg) max_indeg = std::max(max_indeg, c);
  if ((double)max_indeg * (double)wmax >= 9007199254740992.0) return true;
  std::vector<uint32_t> lo(NB, 0);
  uint32_t words = 0, max_rel = 0;
  for (uint64_t b = 0; b < NB; ++b) {
    if (smin[b] > smax[b]) continue;  // no synapses into this block
    lo[b] = smin[b] >> 6;
    words = std::max(words, (smax[b] >> 6) - (smin[b] >> 6) + 1);
    max_rel = std::max(max_rel, smax[b] - lo[b] * 64);
  }
  if (words > kPullMaxWords) return true;  // source spans too wide to stage: push only
  if (max_rel >= (uint64_t)words * 64) {
    L.error = "pull layout: source offset beyond the staged words";
    return false;
  }
  // padding entries of the wide and table formats read the zero word staged
  // after the bitmap
  const uint32_t pad_rel = words * 64;
  const uint32_t V = fmt == kPullWide ? 2 : fmt == kPullTable ? 4 : 8;  // entries per 16-byte vector
  std::sort(wset.begin(), wset.end());
  std::vector<int32_t> wtab(1, 0);  // index 0: the inert weight 0
  wtab.insert(wtab.end(), wset.begin(), wset.end());
  auto widx = [&](int32_t w) {
    return (uint32_t)(std::lower_bound(wtab.begin() + 1, wtab.end(), w) - wtab.begin());
  };
  // count format: a word gap wider than kPullMaxDelta takes skip entries
  // first (each advancing kPullMaxDelta words)
  auto skips = [](uint32_t gap) { return gap > kPullMaxDelta ? (gap - 1) / kPullMaxDelta : 0u; };
  std::vector<uint32_t> prev;  // count format: 32-bit word of the last source of each (segment, target)
  if (fmt == kPullCount) {
    // entries per (segment, target) with the skips; sources arrive in
    // ascending order per target (id-major rows)
    prev.assign(nseg * TB, 0u);
    std::fill(cnt.begin(), cnt.end(), 0u);
    for (uint64_t id = 0; id < T; ++id)
      for (uint64_t k = 0; k < D; ++k)
        for (uint32_t q = L.row_off[id * D + k]; q < L.row_off[id * D + k + 1]; ++q) {
          const uint32_t tg = L.syn[q].x;
          const uint64_t b = tg / TB;
          const uint64_t key = (b * D + k) * TB + (tg - b * TB);
          const uint32_t word = ((uint32_t)id - lo[b] * 64) >> 5;
          const uint64_t c = (uint64_t)cnt[key] + skips(word - prev[key]) + 1;
          if (c > 0xffffffffull) return true;
          cnt[key] = (uint32_t)c;
          prev[key] = word;
        }
  }
  // target order per segment and slice offsets; each slice lane records its
  // target and how many of the slice's vectors hold its entries
  std::vector<uint16_t> perm(TB);
  std::vector<uint16_t> lanes(nseg * TB);
  std::vector<uint32_t> off(nseg * SL + 1, 0);
  std::vector<uint32_t> pos_of(nseg * TB);  // (segment, target) -> its lane's first entry
  uint64_t pos = 0;
  for (uint64_t sg = 0; sg < nseg; ++sg) {
    uint16_t* pm = perm.data();
    const uint32_t* c = &cnt[sg * TB];
    for (uint32_t i = 0; i < TB; ++i) pm[i] = (uint16_t)i;
    std::stable_sort(pm, pm + TB, [&](uint16_t a, uint16_t b2) { return c[a] > c[b2]; });
    for (uint32_t sl = 0; sl < SL; ++sl) {
      off[sg * SL + sl] = (uint32_t)pos;
      const uint64_t len = ((uint64_t)c[pm[sl * 64]] + V - 1) / V * V;  // longest lane of the slice
      for (uint32_t lane = 0; lane < 64; ++lane) {
        const uint32_t tl = pm[sl * 64 + lane];
        pos_of[sg * TB + tl] = (uint32_t)(pos + lane * V);
        const uint64_t own = ((uint64_t)c[tl] + V - 1) / V;  // vectors holding the lane's entries
        lanes[sg * TB + sl * 64 + lane] = (uint16_t)(tl | std::min<uint64_t>(own, kPullLaneAll) << kPullShift);
      }
      pos += len * 64;
      if (pos > 0xffffffffull) return true;
    }
  }
  off[nseg * SL] = (uint32_t)pos;
  // count format: the device reads the entries through a buffer resource
  // whose out-of-range offsets (2^31 and up) read zero vectors
  if (fmt == kPullCount && (pos + V) * sizeof(uint16_t) > kPullMaxBufferBytes) return true;
  // entry j of a lane sits at lane_first + (j / V) * 64 * V + j % V
  std::vector<uint32_t> fill(nseg * TB, 0);
  auto place = [&](uint64_t key) {
    const uint32_t j = fill[key]++;
    return (uint64_t)pos_of[key] + (uint64_t)(j / V) * 64 * V + j % V;
  };
  if (fmt == kPullCount) {
    std::vector<uint16_t> ps(pos + V, count_entry(kPullSkipTag, 0, 0));  // padding, null vector: zero skips
    std::fill(prev.begin(), prev.end(), 0u);
    for (uint64_t id = 0; id < T; ++id)
      for (uint64_t k = 0; k < D; ++k)
        for (uint32_t q = L.row_off[id * D + k]; q < L.row_off[id * D + k + 1]; ++q) {
          const uint32_t tg = L.syn[q].x;
          const uint64_t b = tg / TB;
          const uint64_t key = (b * D + k) * TB + (tg - b * TB);
          const uint32_t rel = (uint32_t)id - lo[b] * 64;
          uint32_t gap = (rel >> 5) - prev[key];
          for (uint32_t i = skips(gap); i > 0; --i) {
            ps[place(key)] = count_entry(kPullSkipTag, 0, kPullMaxDelta);
            gap -= kPullMaxDelta;
          }
          ps[place(key)] = count_entry(widx((int32_t)L.syn[q].y), rel & 31u, gap);
          prev[key] = rel >> 5;
        }
    L.pull_syn16.swap(ps);
    L.pull_wtab.swap(wtab);
  } else if (fmt == kPullTable) {
    std::vector<uint32_t> ps(pos + V, pad_rel);  // padding and the null vector: the zero word, index 0
    for (uint64_t id = 0; id < T; ++id)
      for (uint64_t k = 0; k < D; ++k)
        for (uint32_t q = L.row_off[id * D + k]; q < L.row_off[id * D + k + 1]; ++q) {
          const uint32_t tg = L.syn[q].x;
          const uint64_t b = tg / TB;
          ps[place((b * D + k) * TB + (tg - b * TB))] =
              ((uint32_t)id - lo[b] * 64) | widx((int32_t)L.syn[q].y) << kPullRelBits;
        }
    L.pull_syn32.swap(ps);
    L.pull_wtab.swap(wtab);
  } else {
    std::vector<LayoutPair> ps(pos + V, LayoutPair{pad_rel, 0u});  // padding, null vector: the zero word, weight 0
    for (uint64_t id = 0; id < T; ++id)
      for (uint64_t k = 0; k < D; ++k)
        for (uint32_t q = L.row_off[id * D + k]; q < L.row_off[id * D + k + 1]; ++q) {
          const uint32_t tg = L.syn[q].x;
          const uint64_t b = tg / TB;
          ps[place((b * D + k) * TB + (tg - b * TB))] = LayoutPair{(uint32_t)id - lo[b] * 64, L.syn[q].y};
        }
    L.pull_syn.swap(ps);
  }
  L.pull_fmt = fmt;
  L.pull_off.swap(off);
  L.pull_lo.swap(lo);
  L.pull_lane.swap(lanes);
  L.pull_words = words;
  return true;
}

// All device formats.  After the rows and the pull layout, the push format:
// per-id synapse ranges (rows are id-major, so syn_off[id] = row_off[id * D])
// and the API delay packed into each entry's target word.
inline bool build_layout(const szsnn_net_desc& d, HostLayout& L, const SliceChooser& choose_slices = nullptr) {
  if (!build_rows_and_pull(d, L, choose_slices)) return false;
  const uint64_t T = (uint64_t)d.n_neurons + d.n_sources, D = d.max_delay;
  if (d.n_neurons > kMaxNeurons) {
    L.error = "more than 2^26 neurons";
    return false;
  }
  L.syn_off.assign(T + 1, 0);
  L.active = 0;
  for (uint64_t id = 0; id <= T; ++id) L.syn_off[id] = L.row_off[id * D];
  for (uint64_t id = 0; id < T; ++id) {
    L.active += L.syn_off[id + 1] != L.syn_off[id];
    for (uint64_t k = 0; k < D; ++k)
      for (uint32_t q = L.row_off[id * D + k]; q < L.row_off[id * D + k + 1]; ++q)
        L.syn[q].x |= (uint32_t)k << kSynDelayShift;
  }
  return true;
}

}  // namespace szsnn
