// CPU check of the device formats and of the step kernels' index arithmetic.
//
// Builds layouts with the runtime's builder (szsnn_amd/csrc/szsnn_layout.h)
// for random and edge-case networks, then replays, in plain C++, the index
// computations of the push in update_block and of pull_body from
// szsnn_amd/csrc/szsnn_kernels.hip with an explicit bounds check on every
// array access: spike histories in which some steps push and the others are
// deferred to the pull must deliver, at step t, the naive sum over the
// synapse list.  Exit code 0 and "layout_check OK" on success; any failure
// aborts with a message.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../szsnn_amd/csrc/szsnn_layout.h"

using namespace szsnn;

#define CHECK(cond, ...)                                      \
  do {                                                        \
    if (!(cond)) {                                            \
      std::fprintf(stderr, "FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond); \
      std::fprintf(stderr, __VA_ARGS__);                      \
      std::fprintf(stderr, "\n");                             \
      std::exit(1);                                           \
    }                                                         \
  } while (0)

struct Net {
  uint32_t N, M, D;
  std::vector<uint32_t> src, tgt;
  std::vector<float> w;
  std::vector<uint16_t> dl;
};

// nweights > 0: weights drawn from that many values (a narrow pull layout
// is possible); 0: continuous weights (wide entries).  block_gaps: no
// synapses into every third block of TB targets.
static Net random_net(std::mt19937_64& rng, uint32_t N, uint32_t M, uint32_t D, uint64_t S, bool block_gaps,
                      uint32_t nweights, uint32_t TB) {
  Net n{N, M, D, {}, {}, {}, {}};
  std::uniform_int_distribution<uint32_t> us(0, N + M - 1), ut(0, N - 1), ud(0, D - 1);
  std::uniform_real_distribution<double> uw(-1.0, 1.0);
  std::vector<double> table;
  for (uint32_t i = 0; i < nweights; ++i) table.push_back(uw(rng));
  std::uniform_int_distribution<uint32_t> ui(0, nweights ? nweights - 1 : 0);
  for (uint64_t s = 0; s < S; ++s) {
    uint32_t t = ut(rng);
    if (block_gaps && ((t / TB) % 3 == 1)) continue;  // leave some target blocks empty
    n.src.push_back(us(rng));
    n.tgt.push_back(t);
    double x = nweights ? table[ui(rng)] : uw(rng);
    if (s % 17 == 0) x = 0.0;  // inert
    n.w.push_back((float)std::ldexp(std::nearbyint(std::ldexp(x, 20)), -20));
    n.dl.push_back((uint16_t)ud(rng));
  }
  return n;
}

static szsnn_net_desc desc_of(const Net& n, int32_t pull_layout = 0) {
  szsnn_net_desc d{};
  d.n_neurons = n.N;
  d.n_sources = n.M;
  d.n_syn = n.src.size();
  d.syn_src = n.src.data();
  d.syn_tgt = n.tgt.data();
  d.syn_w = n.w.data();
  d.syn_delay = n.dl.data();
  d.max_delay = n.D;
  d.frac_bits = 20;
  d.pull_layout = pull_layout;
  return d;
}

// push replay of update_block at step s: blocks of BLK ids; the spiking ids'
// synapse ranges are scanned and walked in batches of kBatch per lane; each
// entry adds into ring slot (s + k + 1) % R.  ring = [R][2][N], ev = [R].
static void push_replay(const HostLayout& L, uint32_t N, uint32_t T, uint32_t D, uint64_t s,
                        const std::vector<uint8_t>& fired_s, uint32_t BLK, std::vector<long long>& ring,
                        std::vector<uint64_t>& ev) {
  const uint32_t R = D + 1;
  const uint64_t S = L.syn.size();
  const uint32_t slot = (uint32_t)(s % R);
  const uint32_t first_slot = slot + 1 == R ? 0u : slot + 1;
  constexpr uint32_t kBatch = 4;
  std::vector<uint32_t> start(BLK), pref(BLK + 1);
  for (uint32_t b = 0; b < (T + BLK - 1) / BLK; ++b) {
    uint32_t run = 0;
    for (uint32_t tid = 0; tid < BLK; ++tid) {
      const uint32_t id = b * BLK + tid;
      uint32_t r0 = 0, r1 = 0;
      if (id < T) {
        CHECK((size_t)id + 1 < L.syn_off.size(), "syn_off read id=%u", id);
        r0 = L.syn_off[id];
        r1 = L.syn_off[id + 1];
      }
      start[tid] = r0;
      pref[tid] = run;
      run += (id < T && fired_s[id]) ? r1 - r0 : 0u;
    }
    pref[BLK] = run;
    const uint32_t total = run;
    for (uint32_t e0 = 0; e0 < total; e0 += BLK * kBatch)
      for (uint32_t u = 0; u < kBatch; ++u)
        for (uint32_t tid = 0; tid < BLK; ++tid) {
          const uint32_t e = e0 + u * BLK + tid;
          if (e >= total) continue;
          uint32_t lo = 0, hi = BLK;
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pref[mid] <= e) lo = mid; else hi = mid;
          }
          const uint64_t q = (uint64_t)start[lo] + (e - pref[lo]);
          CHECK(q < S, "syn read q=%llu S=%llu", (unsigned long long)q, (unsigned long long)S);
          const LayoutPair p = L.syn[q];
          const uint32_t k = p.x >> kSynDelayShift, tgt = p.x & kSynTargetMask;
          CHECK(k < D, "packed delay %u", k);
          CHECK(tgt < N, "target %u", tgt);
          uint32_t ds = first_slot + k;
          if (ds >= R) ds -= R;
          CHECK(ds < R && ds != slot, "ring slot %u (own %u)", ds, slot);
          const int32_t w = (int32_t)p.y;
          CHECK(w != 0, "stored zero weight");
          ring[((size_t)ds * 2 + (w < 0 ? 1 : 0)) * N + tgt] += w;
          ++ev[ds];
        }
  }
}

// pull replay for step t with history bitmaps hist[R][W] and due mask `due`
// (bit k: step t-k-1 was deferred); G delay groups, each writing its slot of
// `out` ([G][2][N]); the caller sums the slots.  Follows pull_body's indexing:
// per (block, delay) segment SL = pull_slices slices, lane L of slice s
// serving target pull_lane[seg*TB + s*64 + L] & kPullTMask (TB = 64 * SL
// targets per block), its vector j at 16-byte
// index s0/V + 64*j + L, read for j < pull_lane[...] >> kPullShift, or for
// every j < the slice's vectors when that field is kPullLaneAll (the rest
// of the column must be padding); the spike bit of rel from 32-bit word
// rel / 32 of the staged bitmap, padding from the zero word after it; FP64
// sums as the kernel forms them.
// slice lanes replayed, and those whose word says kPullLaneAll
static uint64_t lanes_seen = 0, lanes_all = 0;

static void pull_replay(const HostLayout& L, uint32_t N, uint32_t T, uint32_t D, uint64_t t, uint64_t due,
                        const std::vector<uint64_t>& hist, uint32_t R, uint32_t W, uint32_t G,
                        std::vector<long long>& out, uint64_t& events) {
  const uint32_t words = L.pull_words;
  std::vector<uint64_t> bits(words);
  const uint32_t SL = L.pull_slices, TB = 64 * SL;
  CHECK(SL >= kPullMinSlices && SL <= kPullSlices, "slices %u", SL);
  std::vector<double> sums(2 * TB);
  const uint32_t fmt = L.pull_fmt;
  const uint32_t V = fmt == kPullWide ? 2 : fmt == kPullTable ? 4 : 8;
  const uint32_t pad_rel = words * 64;
  const uint16_t skip0 = count_entry(kPullSkipTag, 0, 0);
  const size_t nent = fmt == kPullWide ? L.pull_syn.size() : fmt == kPullTable ? L.pull_syn32.size() : L.pull_syn16.size();
  CHECK(L.NB == (N + TB - 1) / TB, "blocks");
  CHECK(L.pull_off.size() == (size_t)L.NB * D * SL + 1, "pull_off size");
  CHECK(L.pull_lane.size() == (size_t)L.NB * D * TB, "pull_lane size");
  // the slices, then one null vector of padding entries (what loads past a
  // lane's own vectors read)
  CHECK(L.pull_off.back() + V == nent, "entries %zu vs offsets %u", nent, L.pull_off.back());
  for (size_t e = L.pull_off.back(); e < nent; ++e) {
    const bool pad = fmt == kPullWide    ? (L.pull_syn[e].x == pad_rel && L.pull_syn[e].y == 0)
                     : fmt == kPullTable ? L.pull_syn32[e] == pad_rel
                                         : L.pull_syn16[e] == skip0;
    CHECK(pad, "null vector entry %zu is not padding", e);
  }
  for (uint32_t bg = 0; bg < L.NB * G; ++bg) {
    const uint32_t b = bg / G, g = bg % G;
    std::fill(sums.begin(), sums.end(), 0.0);
    CHECK(b < L.pull_lo.size(), "pull_lo b=%u", b);
    const uint32_t lo = L.pull_lo[b];
    CHECK(lo <= W, "lo=%u W=%u", lo, W);
    const uint32_t nw = std::min(words, W - lo);
    const uint32_t k0 = (uint32_t)(((uint64_t)g * D) / G), k1 = (uint32_t)(((uint64_t)(g + 1) * D) / G);
    CHECK(k1 <= D, "group end");
    for (uint32_t k = k0; k < k1; ++k) {
      if (!((due >> k) & 1ull)) continue;
      CHECK(t >= (uint64_t)k + 1, "due delay before the first step");
      const size_t seg = (size_t)b * D + k;
      const uint32_t* so = &L.pull_off[seg * SL];
      // the slices' lanes serve every target of the block once
      std::vector<uint8_t> seen(TB, 0);
      for (uint32_t i = 0; i < TB; ++i) {
        const uint32_t tl = L.pull_lane[seg * TB + i] & kPullTMask;
        CHECK(tl < TB && !seen[tl], "perm of segment %zu is not a permutation", seg);
        seen[tl] = 1;
      }
      if (so[0] == so[SL]) continue;
      const uint64_t hbase = ((t - k - 1) % R) * W + lo;
      for (uint32_t w = 0; w < nw; ++w) {
        CHECK(hbase + w < hist.size(), "hist read");
        bits[w] = hist[hbase + w];
      }
      for (uint32_t sl = 0; sl < SL; ++sl) {
        const uint32_t s0 = so[sl], s1 = so[sl + 1];
        CHECK(s0 <= s1 && s1 <= nent, "slice bounds");
        CHECK(s0 % V == 0 && (s1 - s0) % (64 * V) == 0, "slice not made of whole vectors per lane");
        if (s0 == s1) continue;
        const uint32_t nv = (s1 - s0) / (64 * V);
        for (uint32_t lane = 0; lane < 64; ++lane) {
          const uint32_t ln = L.pull_lane[seg * TB + sl * 64 + lane];
          const uint32_t tl = ln & kPullTMask, field = ln >> kPullShift;
          const uint32_t mine = field == kPullLaneAll ? nv : field;
          CHECK(mine <= nv, "lane vectors %u beyond the slice's %u", mine, nv);
          ++lanes_seen;
          lanes_all += field == kPullLaneAll;
          // the column beyond the lane's own vectors is padding
          for (uint32_t j = mine; j < nv; ++j)
            for (uint32_t c = 0; c < V; ++c) {
              const uint64_t e = (uint64_t)(s0 / V + 64 * j + lane) * V + c;
              const bool pad = fmt == kPullWide    ? (L.pull_syn[e].x == pad_rel && L.pull_syn[e].y == 0)
                               : fmt == kPullTable ? L.pull_syn32[e] == pad_rel
                                                   : L.pull_syn16[e] == skip0;
              CHECK(pad, "real entry beyond the lane's vectors");
            }
          double ae = 0.0, ai = 0.0;
          uint32_t run = 0;  // count format: running sum of the deltas
          for (uint32_t j = 0; j < mine; ++j)
            for (uint32_t c = 0; c < V; ++c) {
              const uint64_t e = (uint64_t)(s0 / V + 64 * j + lane) * V + c;
              CHECK(e < s1, "entry beyond the slice");
              uint32_t rel;
              int32_t w;
              if (fmt == kPullCount) {
                const uint16_t x = L.pull_syn16[e];
                run += count_delta(x);  // the running 32-bit word
                const uint32_t tag = count_tag(x);
                CHECK((run >> 1) < nw, "LDS word read word=%u nw=%u", run, nw);
                if (tag == kPullSkipTag) {  // gap or padding: no synapse
                  CHECK(count_bit(x) == 0, "skip entry with a bit");
                  continue;
                }
                CHECK(tag <= kPullCountFields && tag < L.pull_wtab.size(), "count tag %u", tag);
                rel = run * 32 + count_bit(x);
                w = L.pull_wtab[tag];
              } else if (fmt == kPullTable) {
                const uint32_t x = L.pull_syn32[e];
                rel = x & kPullRelMask;
                const uint32_t wi = x == pad_rel ? 0u : x >> kPullRelBits;
                CHECK(wi < L.pull_wtab.size(), "weight index %u of %zu", wi, L.pull_wtab.size());
                w = L.pull_wtab[wi];
              } else {
                rel = L.pull_syn[e].x;
                w = (int32_t)L.pull_syn[e].y;
              }
              if (rel == pad_rel) {
                CHECK(w == 0, "padding entry with weight %d", w);
                continue;  // reads the zero word
              }
              CHECK(w != 0, "real entry with weight 0");
              CHECK((rel >> 6) < nw, "LDS bit read rel=%u nw=%u", rel, nw);
              CHECK((uint64_t)lo * 64 + rel < T, "source id %llu", (unsigned long long)lo * 64 + rel);
              const uint32_t word32 = (uint32_t)(bits[rel >> 6] >> (32 * ((rel >> 5) & 1)));
              if ((word32 >> (rel & 31)) & 1u) {
                if (w > 0) ae += (double)w;
                if (w < 0) ai += (double)w;
                if (w != 0) ++events;
              }
            }
          if (b * TB + tl >= N) CHECK(ae == 0.0 && ai == 0.0, "sum for a target beyond N");
          sums[tl] += ae;
          sums[TB + tl] += ai;
        }
      }
    }
    for (uint32_t i = 0; i < TB; ++i) {
      const uint32_t tgt = b * TB + i;
      if (tgt < N) {
        CHECK((size_t)g * 2 * N + N + tgt < out.size(), "slot write");
        out[(size_t)g * 2 * N + tgt] = (long long)sums[i];
        out[(size_t)g * 2 * N + N + tgt] = (long long)sums[TB + i];
      }
    }
  }
}

// slices: the pull layout's slices per segment, set before the build (odd
// cases) or by the builder's callback once the format is known (even ones)
static void check_case(std::mt19937_64& rng, uint32_t N, uint32_t M, uint32_t D, uint64_t S, bool gaps,
                       uint32_t nweights, int32_t pull_layout, uint32_t slices) {
  Net net = random_net(rng, N, M, D, S, gaps, nweights, 64 * slices);
  const szsnn_net_desc d = desc_of(net, pull_layout);
  HostLayout L;
  uint32_t asked_fmt = ~0u;
  if (slices % 2) {
    L.pull_slices = slices;
    CHECK(build_layout(d, L), "build_layout: %s", L.error.c_str());
  } else {
    CHECK(build_layout(d, L, [&](uint32_t fmt, uint32_t, uint32_t) { asked_fmt = fmt; return slices; }),
          "build_layout: %s", L.error.c_str());
  }
  CHECK(L.pull_slices == slices, "slices %u, asked for %u", L.pull_slices, slices);
  const uint32_t T = N + M, W = (T + 63) / 64, R = D + 1;
  const uint64_t Ss = L.row_off.back();
  CHECK(L.row_off.size() == (size_t)T * D + 1, "row_off size");
  CHECK(L.syn.size() == Ss, "syn size");
  CHECK(L.syn_off.size() == (size_t)T + 1 && L.syn_off[T] == Ss, "syn_off");
  // delay bits agree with rows; packed delays agree with the rows they sit in
  uint64_t active = 0;
  for (uint32_t id = 0; id < T; ++id) {
    active += L.syn_off[id + 1] != L.syn_off[id];
    for (uint32_t k = 0; k < D; ++k) {
      CHECK((((L.delay_bits[id] >> k) & 1) != 0) == (L.row_off[id * D + k + 1] > L.row_off[id * D + k]), "bits");
      for (uint32_t q = L.row_off[id * D + k]; q < L.row_off[id * D + k + 1]; ++q)
        CHECK((L.syn[q].x >> kSynDelayShift) == k, "packed delay of entry %u", q);
    }
  }
  CHECK(active == L.active, "active ids");
  const bool pull = !L.pull_off.empty();
  CHECK(pull, "pull layout expected for N=%u", N);
  // narrow entries exactly when allowed and asked for (these nets' sources fit)
  std::vector<long long> distinct;
  for (float w : net.w) {
    const long long q = (long long)std::nearbyint(std::ldexp((double)w, d.frac_bits));
    if (q != 0) distinct.push_back(q);
  }
  std::sort(distinct.begin(), distinct.end());
  distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
  const bool narrow_expected = pull_layout == 0 && distinct.size() <= kPullMaxWeights;
  const uint32_t fmt_expected = !narrow_expected ? kPullWide
                                : distinct.size() <= kPullCountFields ? kPullCount : kPullTable;
  CHECK(L.pull_fmt == fmt_expected, "format %u expected %u", L.pull_fmt, fmt_expected);
  CHECK(slices % 2 || asked_fmt == fmt_expected, "the callback saw format %u, expected %u", asked_fmt, fmt_expected);
  CHECK(L.pull_fmt == kPullWide || (L.pull_wtab.size() == distinct.size() + 1 && L.pull_wtab[0] == 0),
        "weight table");
  // random spike history up to step t - 1
  const uint64_t t = 3 * D + 2;
  std::vector<uint64_t> hist((size_t)R * W);
  std::bernoulli_distribution sp(0.3);
  std::vector<std::vector<uint8_t>> fired(t, std::vector<uint8_t>(T));
  for (uint64_t s = 0; s < t; ++s) {
    for (uint32_t id = 0; id < T; ++id) fired[s][id] = sp(rng);
    uint64_t* row = &hist[(s % R) * W];
    std::fill(row, row + W, 0);
    for (uint32_t id = 0; id < T; ++id)
      if (fired[s][id]) row[id >> 6] |= 1ull << (id & 63);
  }
  // naive input delivered at step t
  std::vector<long long> naive(2 * (size_t)N, 0);
  uint64_t nev = 0;
  for (size_t s = 0; s < net.src.size(); ++s) {
    const long long wq = (long long)std::nearbyint(std::ldexp((double)net.w[s], 20));
    if (wq == 0) continue;
    const uint64_t k = net.dl[s];
    if (t < k + 1) continue;
    if (!fired[t - k - 1][net.src[s]]) continue;
    naive[(wq < 0 ? (size_t)N : 0) + net.tgt[s]] += wq;
    ++nev;
  }
  // deferral patterns over the steps t-D .. t-1: none, all, random
  for (int pattern = 0; pattern < 3; ++pattern) {
    std::vector<uint8_t> deferred(t, 0);
    std::bernoulli_distribution coin(0.4);
    for (uint64_t s = 0; s < t; ++s) deferred[s] = pattern == 1 ? 1 : pattern == 2 ? coin(rng) : 0;
    uint64_t due = 0;
    for (uint32_t k = 0; k < D; ++k)
      if (t >= (uint64_t)k + 1 && deferred[t - k - 1]) due |= 1ull << k;
    for (uint32_t BLK : {64u, 256u, 1024u}) {
      std::vector<long long> ring((size_t)R * 2 * N, 0);
      std::vector<uint64_t> ev(R, 0);
      for (uint64_t s = (t > D ? t - D : 0); s < t; ++s)
        if (!deferred[s]) push_replay(L, N, T, D, s, fired[s], BLK, ring, ev);
      const size_t slot = t % R;
      for (uint32_t G : {1u, 2u, 3u, 8u}) {
        if (G > D) continue;
        std::vector<long long> slots((size_t)G * 2 * N, -1);
        uint64_t pev = 0;
        if (due) pull_replay(L, N, T, D, t, due, hist, R, W, G, slots, pev);
        std::vector<long long> got(ring.begin() + slot * 2 * N, ring.begin() + (slot + 1) * 2 * N);
        if (due)
          for (uint32_t g = 0; g < G; ++g)
            for (size_t i = 0; i < 2 * (size_t)N; ++i) got[i] += slots[(size_t)g * 2 * N + i];
        CHECK(got == naive, "delivered sum differs (pattern %d, BLK=%u, G=%u)", pattern, BLK, G);
        CHECK(ev[slot] + pev == nev, "events %llu + %llu vs %llu", (unsigned long long)ev[slot],
              (unsigned long long)pev, (unsigned long long)nev);
      }
    }
  }
}

int main() {
  std::mt19937_64 rng(12345);
  struct C { uint32_t N, M, D; uint64_t S; bool gaps; };
  const C cases[] = {{1, 0, 1, 5, false},     {63, 2, 5, 400, false},   {64, 0, 3, 500, false},
                     {255, 7, 64, 3000, false}, {256, 0, 2, 2000, true}, {257, 31, 9, 4000, false},
                     {1000, 100, 32, 30000, true}, {5000, 10, 16, 60000, false}, {3000, 500, 1, 20000, true},
                     // sparse columns over wide source ranges: count-format gaps beyond one delta (skips)
                     {600, 60000, 4, 3000, false}, {300, 390000, 2, 2000, false},
                     // two targets with about 1400 synapses each in one delay: lanes
                     // of more than kPullLaneAll - 1 vectors
                     {2, 3000, 1, 3000, false}};
  // every case at kPullSlices and, in turn, at each other slice count
  uint32_t turn = 0;
  for (const C& c : cases)
    for (uint32_t sl : {kPullSlices, kPullMinSlices + turn++ % (kPullSlices - kPullMinSlices)}) {
      check_case(rng, c.N, c.M, c.D, c.S, c.gaps, 0, 0, sl);    // continuous weights: wide unless few synapses
      check_case(rng, c.N, c.M, c.D, c.S, c.gaps, 3, 0, sl);    // three weights: count format
      check_case(rng, c.N, c.M, c.D, c.S, c.gaps, 4, 0, sl);    // four weights: table format
      check_case(rng, c.N, c.M, c.D, c.S, c.gaps, 200, 0, sl);  // 200 weights: table format
      check_case(rng, c.N, c.M, c.D, c.S, c.gaps, 3, 1, sl);    // wide entries asked for
    }
  CHECK(lanes_all > 0 && lanes_all < lanes_seen, "lanes marked all: %llu of %llu", (unsigned long long)lanes_all,
        (unsigned long long)lanes_seen);
  // the slice chooser against a brute force of its rule: in range, the
  // delay groups of kPullSlices kept, and no allowed count with fewer
  // rounds, or as few rounds and fewer busiest-CU columns, or as many and
  // more slices
  for (uint64_t N : {1ull, 63ull, 640ull, 10000ull, 100000ull, 250000ull, 1000000ull, 5000000ull})
    for (uint64_t D : {1ull, 2ull, 7ull, 16ull, 32ull, 64ull})
      for (uint64_t per_cu : {1ull, 4ull, 7ull, 8ull}) {
        const uint64_t cus = 256, slots = per_cu * cus;
        const uint32_t s = choose_pull_slices(N, D, slots, cus);
        CHECK(s >= kPullMinSlices && s <= kPullSlices, "chose %u", s);
        auto cost = [&](uint32_t x, uint64_t& rounds, uint64_t& cols) {
          const uint64_t nb = (N + 64 * x - 1) / (64 * x);
          const uint32_t g = pull_groups_for(nb, D, slots);
          rounds = (nb * g + slots - 1) / slots;
          cols = (nb * g + cus - 1) / cus * x * ((D + g - 1) / g);
          return g;
        };
        uint64_t r8, c8, rs, cs;
        const uint32_t g8 = cost(kPullSlices, r8, c8);
        CHECK(cost(s, rs, cs) >= g8, "fewer groups than at kPullSlices");
        for (uint32_t x = kPullMinSlices; x <= kPullSlices; ++x) {
          uint64_t r, c;
          if (cost(x, r, c) < g8) continue;
          CHECK(!(r < rs || (r == rs && c < cs) || (r == rs && c == cs && x > s)),
                "N=%llu D=%llu slots=%llu: %u slices beat the chosen %u", (unsigned long long)N,
                (unsigned long long)D, (unsigned long long)slots, x, s);
        }
      }
  // the 100000-neuron network with 32 delays, 7 resident gather blocks on
  // each of 256 CUs: 7 slices, 224 blocks in 8 groups, 7 items on every CU
  CHECK(choose_pull_slices(100000, 32, 7 * 256, 256) == 7, "100k network");
  CHECK(pull_groups_for((100000 + 447) / 448, 32, 7 * 256) == 8, "100k network groups");
  // slice counts outside kPullMinSlices .. kPullSlices are refused
  for (uint32_t sl : {kPullMinSlices - 1, kPullSlices + 1}) {
    Net n = random_net(rng, 100, 0, 2, 500, false, 3, 64 * kPullSlices);
    HostLayout L;
    L.pull_slices = sl;
    CHECK(!build_layout(desc_of(n), L) && !L.error.empty(), "slices %u accepted", sl);
    HostLayout L2;
    CHECK(!build_layout(desc_of(n), L2, [&](uint32_t, uint32_t, uint32_t) { return sl; }) && !L2.error.empty(),
          "slices %u accepted from the callback", sl);
  }
  // count format at the delta boundaries: one target whose sources sit at
  // word gaps 0 (same word, another bit; then a duplicate), 511 (one delta),
  // 512, 1022 (one skip), 1023, 1024 (two skips), with bits 0..31 in their
  // words; the lane's decoded sources and weights must come back
  {
    const uint32_t wgaps[] = {0, 0, 0, 511, 512, 1022, 1023, 1024};
    const uint32_t bits[] = {5, 31, 31, 0, 17, 1, 30, 9};
    Net n{1, 0, 1, {}, {}, {}, {}};
    uint32_t word = 0;
    uint64_t want_entries = 0;
    uint32_t id = 0;
    for (uint32_t i = 0; i < sizeof(wgaps) / sizeof(wgaps[0]); ++i) {
      word += wgaps[i];
      id = word * 32 + bits[i];
      n.src.push_back(id);
      n.tgt.push_back(0);
      n.w.push_back(i % 2 ? 0.25f : -0.5f);
      n.dl.push_back(0);
      want_entries += 1 + (wgaps[i] > kPullMaxDelta ? (wgaps[i] - 1) / kPullMaxDelta : 0);
    }
    n.M = id + 1;
    const szsnn_net_desc d = desc_of(n);
    HostLayout L;
    CHECK(build_layout(d, L), "build: %s", L.error.c_str());
    CHECK(L.pull_fmt == kPullCount && L.pull_lo.size() == 1 && L.pull_lo[0] == 0, "count format expected");
    uint32_t tl = kPullTB, mine = 0, lane = 0, sl = 0;
    for (uint32_t i = 0; i < 64 * L.pull_slices; ++i)
      if ((L.pull_lane[i] & kPullTMask) == 0) {
        tl = 0;
        sl = i / 64;
        lane = i % 64;
        mine = L.pull_lane[i] >> kPullShift;
      }
    CHECK(tl == 0 && mine == (want_entries + 7) / 8, "lane of target 0: %u vectors for %llu entries", mine,
          (unsigned long long)want_entries);
    std::vector<std::pair<uint32_t, int32_t>> got;
    uint32_t run = 0, seen = 0;
    for (uint32_t j = 0; j < mine; ++j)
      for (uint32_t c = 0; c < 8; ++c) {
        const uint16_t x = L.pull_syn16[(size_t)(L.pull_off[sl] / 8 + 64 * j + lane) * 8 + c];
        run += count_delta(x);
        ++seen;
        if (count_tag(x) != kPullSkipTag) got.emplace_back(run * 32 + count_bit(x), L.pull_wtab[count_tag(x)]);
      }
    CHECK(got.size() == n.src.size(), "decoded %zu of %zu sources", got.size(), n.src.size());
    for (size_t i = 0; i < got.size(); ++i) {
      const int32_t wq = (int32_t)std::nearbyint(std::ldexp((double)n.w[i], d.frac_bits));
      CHECK(got[i].first == n.src[i] && got[i].second == wq, "entry %zu: source %u weight %d", i, got[i].first,
            got[i].second);
    }
    CHECK(seen >= want_entries && seen - want_entries < 8, "entries %u for %llu", seen, (unsigned long long)want_entries);
  }
  // a network whose sources span more than the LDS budget gets no pull layout
  {
    Net n{200, 800000, 1, {0, 799999}, {0, 1}, {0.5f, -0.5f}, {0, 0}};
    n.src[1] = 800199;
    const szsnn_net_desc d = desc_of(n);
    HostLayout L;
    CHECK(build_layout(d, L), "build");
    CHECK(L.pull_off.empty(), "wide span must disable the pull layout");
  }
  std::printf("layout_check OK\n");
  return 0;
}
