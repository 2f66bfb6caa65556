// CPU check of the device formats and of the gather kernels' index arithmetic.
//
// Builds layouts with the runtime's builder (szsnn_amd/csrc/szsnn_layout.h)
// for random and edge-case networks, then replays, in plain C++, the index
// computations of gather_body (push) and pull_body (pull) from
// szsnn_amd/csrc/szsnn_kernels.hip with an explicit bounds check on every
// array access, and compares both sums with a naive sum over the synapse list.
// Exit code 0 and "layout_check OK" on success; any failure aborts with a message.
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

static Net random_net(std::mt19937_64& rng, uint32_t N, uint32_t M, uint32_t D, uint64_t S, bool block_gaps) {
  Net n{N, M, D, {}, {}, {}, {}};
  std::uniform_int_distribution<uint32_t> us(0, N + M - 1), ut(0, N - 1), ud(0, D - 1);
  std::uniform_real_distribution<double> uw(-1.0, 1.0);
  for (uint64_t s = 0; s < S; ++s) {
    uint32_t t = ut(rng);
    if (block_gaps && ((t >> kPullShift) % 3 == 1)) continue;  // leave some target blocks empty
    n.src.push_back(us(rng));
    n.tgt.push_back(t);
    double x = uw(rng);
    if (s % 17 == 0) x = 0.0;  // inert
    n.w.push_back((float)std::ldexp(std::nearbyint(std::ldexp(x, 20)), -20));
    n.dl.push_back((uint16_t)ud(rng));
  }
  return n;
}

static szsnn_net_desc desc_of(const Net& n) {
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
  d.pull_layout = 0;
  return d;
}

// push replay: rows due = list of row ids; chunks of `per` <= BLK rows as
// gather_body
static void push_replay(const HostLayout& L, uint32_t N, uint64_t S, const std::vector<uint32_t>& rows,
                        uint64_t q_cap, uint32_t count_in_slot, std::vector<long long>& acc, uint32_t BLK,
                        uint32_t nblocks) {
  const uint32_t nrows = (uint32_t)std::min<uint64_t>(count_in_slot, q_cap);
  const uint32_t per = std::min(BLK, std::max(1u, (nrows + nblocks - 1) / nblocks));
  std::vector<uint32_t> sstart(BLK), spref(BLK + 1);
  for (uint32_t block = 0; block < nblocks; ++block)
    for (uint32_t chunk = block * per; chunk < nrows; chunk += nblocks * per) {
      for (uint32_t tid = 0; tid < BLK; ++tid) {
        const uint32_t r = chunk + tid;
        uint32_t start = 0, len = 0;
        if (tid < per && r < nrows) {
          CHECK(r < rows.size() && r < q_cap, "queue read r=%u", r);
          const uint32_t row = rows[r];
          CHECK((uint64_t)row + 1 < L.row_off.size(), "row_off read row=%u", row);
          start = L.row_off[row];
          len = L.row_off[row + 1] - start;
        }
        sstart[tid] = start;
        spref[tid] = len;
      }
      uint32_t acc_s = 0;
      for (uint32_t tid = 0; tid < BLK; ++tid) {
        const uint32_t v = spref[tid];
        spref[tid] = acc_s;
        acc_s += v;
      }
      spref[BLK] = acc_s;
      const uint32_t total = acc_s;
      for (uint32_t e = 0; e < total; ++e) {
        uint32_t lo = 0, hi = BLK;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (spref[mid] <= e) lo = mid; else hi = mid;
        }
        const uint64_t q = (uint64_t)sstart[lo] + (e - spref[lo]);
        CHECK(q < S, "syn read q=%llu S=%llu", (unsigned long long)q, (unsigned long long)S);
        const LayoutPair p = L.syn[q];
        const int32_t w = (int32_t)p.y;
        CHECK(p.x < N, "target %u", p.x);
        acc[(w < 0 ? (size_t)N : 0) + p.x] += w;
      }
    }
}

// pull replay for step t with history bitmaps hist[R][W]
// pull replay for step t with history bitmaps hist[R][W]; G delay groups, each
// writing its slot of `out` ([G][2][N]); the caller sums the slots
static void pull_replay(const HostLayout& L, uint32_t N, uint32_t T, uint32_t D, uint64_t t,
                        const std::vector<uint64_t>& hist, uint32_t R, uint32_t W, uint32_t G,
                        std::vector<long long>& out, uint64_t& events) {
  const uint32_t words = L.pull_words;
  std::vector<uint64_t> bits(words);
  std::vector<long long> acc(2 * kPullTB);
  for (uint32_t bg = 0; bg < L.NB * G; ++bg) {
    const uint32_t b = bg / G, g = bg % G;
    std::fill(acc.begin(), acc.end(), 0);
    CHECK(b < L.pull_lo.size(), "pull_lo b=%u", b);
    const uint32_t lo = L.pull_lo[b];
    CHECK(lo <= W, "lo=%u W=%u", lo, W);
    const uint32_t nw = std::min(words, W - lo);
    const uint32_t k0 = (uint32_t)(((uint64_t)g * D) / G), k1 = (uint32_t)(((uint64_t)(g + 1) * D) / G);
    CHECK(k1 <= D, "group end");
    for (uint32_t k = k0; k < k1; ++k) {
      if (t < (uint64_t)k + 1) break;
      CHECK((uint64_t)b * D + k + 1 < L.pull_off.size(), "pull_off");
      const uint32_t s0 = L.pull_off[b * D + k], s1 = L.pull_off[b * D + k + 1];
      CHECK(s0 <= s1 && s1 <= L.pull_syn.size(), "segment");
      CHECK(((s1 - s0) % kPullRun) == 0 && (s0 % kPullRun) == 0, "segment not made of whole lane runs");
      if (s0 == s1) continue;
      const uint64_t hbase = ((t - k - 1) % R) * W + lo;
      for (uint32_t w = 0; w < nw; ++w) {
        CHECK(hbase + w < hist.size(), "hist read");
        bits[w] = hist[hbase + w];
      }
      const uint32_t base_id = lo * 64;
      const uint32_t kChunk = 64 * kPullRun;
      for (uint32_t cs = s0; cs < s1; cs += kChunk) {  // every wave's chunk, every lane's run
        const uint32_t nl = std::min(kChunk, s1 - cs) / kPullRun;
        CHECK(nl >= 1 && nl <= 64, "lanes per chunk");
        for (uint32_t lane = 0; lane < nl; ++lane)
        for (uint32_t jj = 0; jj < kPullRun; ++jj) {
          const uint32_t pair = (jj / 2) * nl + lane;
          CHECK(pair < std::min(kChunk, s1 - cs) / 2, "pair index beyond the chunk");
          const uint32_t e = cs + 2 * pair + (jj & 1);
          CHECK(e < s1, "entry beyond the segment");
          const LayoutPair p = L.pull_syn[e];
          const uint32_t tl = p.x & (kPullTB - 1);
          const uint32_t id = p.x >> kPullShift;
          CHECK(id < T, "source id %u", id);
          CHECK(id >= base_id, "source below staged range");
          const uint32_t rel = id - base_id;
          CHECK((rel >> 6) < nw, "LDS bit read rel=%u nw=%u", rel, nw);
          const int32_t w = (int32_t)p.y;
          if (((bits[rel >> 6] >> (rel & 63)) & 1ull) && w != 0) {
            CHECK(b * kPullTB + tl < N, "pull target beyond N");
            acc[(w >= 0 ? 0 : kPullTB) + tl] += w;
            ++events;
          }
        }
      }
    }
    for (uint32_t i = 0; i < kPullTB; ++i) {
      const uint32_t tgt = b * kPullTB + i;
      if (tgt < N) {
        CHECK((size_t)g * 2 * N + N + tgt < out.size(), "slot write");
        out[(size_t)g * 2 * N + tgt] = acc[i];
        out[(size_t)g * 2 * N + N + tgt] = acc[kPullTB + i];
      } else {
        CHECK(acc[i] == 0 && acc[kPullTB + i] == 0, "sum for a target beyond N");
      }
    }
  }
}

static void check_case(std::mt19937_64& rng, uint32_t N, uint32_t M, uint32_t D, uint64_t S, bool gaps) {
  Net net = random_net(rng, N, M, D, S, gaps);
  const szsnn_net_desc d = desc_of(net);
  HostLayout L;
  CHECK(build_layout(d, L), "build_layout: %s", L.error.c_str());
  const uint32_t T = N + M, W = (T + 63) / 64, R = D + 1;
  const uint64_t Ss = L.row_off.back();
  CHECK(L.row_off.size() == (size_t)T * D + 1, "row_off size");
  CHECK(L.syn.size() == Ss, "syn size");
  // delay bits agree with rows
  for (uint32_t id = 0; id < T; ++id)
    for (uint32_t k = 0; k < D; ++k)
      CHECK((((L.delay_bits[id] >> k) & 1) != 0) == (L.row_off[id * D + k + 1] > L.row_off[id * D + k]), "bits");
  const bool pull = !L.pull_off.empty();
  CHECK(pull, "pull layout expected for N=%u", N);
  CHECK(L.pull_off.size() == L.NB * D + 1, "pull_off size");
  // random spike history; step t
  const uint64_t t = 3 * D + 2;
  std::vector<uint64_t> hist((size_t)R * W);
  std::bernoulli_distribution sp(0.3);
  std::vector<std::vector<uint8_t>> fired(t, std::vector<uint8_t>(T));
  for (uint64_t s = 0; s < t; ++s) {
    for (uint32_t id = 0; id < T; ++id) fired[s][id] = sp(rng);
    if (s + R >= t + 1 || true) {
      uint64_t* row = &hist[(s % R) * W];
      std::fill(row, row + W, 0);
      for (uint32_t id = 0; id < T; ++id)
        if (fired[s][id]) row[id >> 6] |= 1ull << (id & 63);
    }
  }
  // naive
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
  // due rows of step t, in queue order of the update kernel: for each spiking id
  // of step t-k-1 with a nonempty row k
  std::vector<uint32_t> rows;
  for (uint32_t k = 0; k < D; ++k) {
    if (t < (uint64_t)k + 1) continue;
    for (uint32_t id = 0; id < T; ++id)
      if (fired[t - k - 1][id] && ((L.delay_bits[id] >> k) & 1)) rows.push_back(id * D + k);
  }
  CHECK(rows.size() <= L.nonempty, "due rows exceed the worst-case capacity");
  for (uint32_t BLK : {128u, 256u, 1024u})
    for (uint32_t nb : {1u, 7u, 300u}) {
      std::vector<long long> push(2 * (size_t)N, 0);
      push_replay(L, N, Ss, rows, std::max<uint64_t>(L.nonempty, 1), (uint32_t)rows.size(), push, BLK, nb);
      CHECK(push == naive, "push sum differs (BLK=%u, blocks=%u)", BLK, nb);
    }
  // overflowed slot: count beyond capacity is clamped, reads stay in bounds
  if (rows.size() > 4) {
    std::vector<uint32_t> part(rows.begin(), rows.begin() + 4);
    std::vector<long long> push(2 * (size_t)N, 0);
    push_replay(L, N, Ss, part, 4, (uint32_t)rows.size() + 1000, push, 256, 3);
  }
  for (uint32_t G : {1u, 2u, 3u, 8u}) {
    if (G > D) continue;
    std::vector<long long> slots((size_t)G * 2 * N, -1);
    uint64_t pev = 0;
    pull_replay(L, N, T, D, t, hist, R, W, G, slots, pev);
    std::vector<long long> pulled(2 * (size_t)N, 0);
    for (uint32_t g = 0; g < G; ++g)
      for (size_t i = 0; i < 2 * (size_t)N; ++i) pulled[i] += slots[(size_t)g * 2 * N + i];
    CHECK(pulled == naive, "pull sum differs (G=%u)", G);
    CHECK(pev == nev, "pull events %llu vs %llu", (unsigned long long)pev, (unsigned long long)nev);
  }
}

int main() {
  std::mt19937_64 rng(12345);
  struct C { uint32_t N, M, D; uint64_t S; bool gaps; };
  const C cases[] = {{1, 0, 1, 5, false},     {63, 2, 5, 400, false},   {64, 0, 3, 500, false},
                     {255, 7, 64, 3000, false}, {256, 0, 2, 2000, true}, {257, 31, 9, 4000, false},
                     {1000, 100, 32, 30000, true}, {5000, 10, 16, 60000, false}, {3000, 500, 1, 20000, true}};
  for (const C& c : cases) check_case(rng, c.N, c.M, c.D, c.S, c.gaps);
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
