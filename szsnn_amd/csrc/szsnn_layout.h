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
//   * the pull layout for dense gathers: targets in blocks of kPullTB, one
//     segment per (block, delay) sorted by target, entries
//     (id << kPullShift | target % kPullTB, weight), segments padded to a
//     multiple of kPullRun entries with inert entries (the block's first
//     source word, weight 0).  Lane-interleaved chunks: a segment is cut into
//     chunks of 64 * kPullRun entries (the last one shorter); lane L of a wave
//     owns the chunk's entries L*kPullRun .. L*kPullRun+kPullRun-1 in sorted
//     order, and entry pair j of lane L is stored at pair j*nl + L of the
//     chunk (nl = lanes in the chunk), so the wave's j-th 16-byte load is
//     contiguous.
// Shared by the runtime and by the CPU layout check (tests/cpp/layout_check.cpp).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
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
  std::vector<uint32_t> pull_off;     // [NB*D + 1] or empty
  std::vector<uint32_t> pull_lo;      // [NB]
  std::vector<LayoutPair> pull_syn;
  uint32_t pull_words = 0;
  uint64_t NB = 0;
  std::string error;
};

// Rows, delay bits and the pull layout (targets still unpacked in L.syn).
inline bool build_rows_and_pull(const szsnn_net_desc& d, HostLayout& L) {
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

  // pull layout
  L.NB = (N + kPullTB - 1) / kPullTB;
  L.pull_off.clear();
  L.pull_lo.clear();
  L.pull_syn.clear();
  L.pull_words = 0;
  if (d.pull_layout < 0 || T >= (1ull << (32 - kPullShift)) || S == 0) return true;
  const uint64_t NB = L.NB;
  const uint64_t nkeys = NB * D * kPullTB;
  std::vector<uint32_t> kcnt(nkeys + 1, 0);
  std::vector<uint32_t> smin(NB, 0xffffffffu), smax(NB, 0);
  for (uint64_t id = 0; id < T; ++id)
    for (uint64_t k = 0; k < D; ++k)
      for (uint32_t q = L.row_off[id * D + k]; q < L.row_off[id * D + k + 1]; ++q) {
        const uint32_t tg = L.syn[q].x;
        const uint64_t b = tg >> kPullShift;
        kcnt[(b * D + k) * kPullTB + (tg & (kPullTB - 1)) + 1]++;
        smin[b] = std::min(smin[b], (uint32_t)id);
        smax[b] = std::max(smax[b], (uint32_t)id);
      }
  std::vector<uint32_t> lo(NB, 0);
  uint32_t words = 0;
  for (uint64_t b = 0; b < NB; ++b) {
    if (smin[b] > smax[b]) continue;  // no synapses into this block
    lo[b] = smin[b] >> 6;
    words = std::max(words, (smax[b] >> 6) - (smin[b] >> 6) + 1);
  }
  if (words > kPullMaxWords) return true;  // source spans too wide to stage: push only
  std::vector<uint32_t> off(NB * D + 1, 0), kpos(nkeys);
  uint64_t pos = 0;
  for (uint64_t sg = 0; sg < NB * D; ++sg) {
    off[sg] = (uint32_t)pos;
    for (uint64_t j = 0; j < kPullTB; ++j) {
      kpos[sg * kPullTB + j] = (uint32_t)pos;
      pos += kcnt[sg * kPullTB + j + 1];
    }
    pos = (pos + kPullRun - 1) / kPullRun * kPullRun;  // pad to whole lane runs
    if (pos > 0xffffffffull) return true;
  }
  off[NB * D] = (uint32_t)pos;
  std::vector<LayoutPair> ps(pos);
  for (uint64_t sg = 0; sg < NB * D; ++sg) {
    const uint64_t b = sg / D;
    for (uint32_t e = off[sg]; e < off[sg + 1]; ++e) ps[e] = LayoutPair{(lo[b] * 64u) << kPullShift, 0u};
  }
  for (uint64_t id = 0; id < T; ++id)
    for (uint64_t k = 0; k < D; ++k)
      for (uint32_t q = L.row_off[id * D + k]; q < L.row_off[id * D + k + 1]; ++q) {
        const uint32_t tg = L.syn[q].x;
        const uint64_t key = ((uint64_t)(tg >> kPullShift) * D + k) * kPullTB + (tg & (kPullTB - 1));
        ps[kpos[key]++] = LayoutPair{((uint32_t)id << kPullShift) | (tg & (kPullTB - 1)), L.syn[q].y};
      }
  // interleave every chunk: logical pair q = lane*(kPullRun/2) + j goes to j*nl + lane
  const uint32_t ppl = kPullRun / 2;  // pairs per lane
  std::vector<LayoutPair> tmp(64 * kPullRun);
  for (uint64_t sg = 0; sg < NB * D; ++sg)
    for (uint64_t cs = off[sg]; cs < off[sg + 1]; cs += 64 * kPullRun) {
      const uint64_t m = std::min<uint64_t>(64 * kPullRun, off[sg + 1] - cs);
      const uint64_t nl = m / kPullRun;
      std::copy(ps.begin() + cs, ps.begin() + cs + m, tmp.begin());
      for (uint64_t q = 0; q < m / 2; ++q) {
        const uint64_t lane = q / ppl, j = q % ppl;
        const uint64_t dst = cs + 2 * (j * nl + lane);
        ps[dst] = tmp[2 * q];
        ps[dst + 1] = tmp[2 * q + 1];
      }
    }
  L.pull_off.swap(off);
  L.pull_lo.swap(lo);
  L.pull_syn.swap(ps);
  L.pull_words = words;
  return true;
}

// All device formats.  After the rows and the pull layout, the push format:
// per-id synapse ranges (rows are id-major, so syn_off[id] = row_off[id * D])
// and the API delay packed into each entry's target word.
inline bool build_layout(const szsnn_net_desc& d, HostLayout& L) {
  if (!build_rows_and_pull(d, L)) return false;
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
