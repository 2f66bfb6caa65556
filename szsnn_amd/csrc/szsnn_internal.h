// szsnn_internal.h -- device-side view of an uploaded network and the launch
// interface shared by the runtime (szsnn_runtime.hip) and the kernels
// (szsnn_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "szsnn_constants.h"

namespace szsnn {

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;

// SplitMix64 output mixer (Steele, Lea, Flood 2014); counter-based streams.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
// Output number `idx` (0-based) of a SplitMix64 generator seeded with `seed`.
__host__ __device__ inline uint64_t splitmix_at(uint64_t seed, uint64_t idx) {
  return mix64(seed + kGolden * (idx + 1));
}
// Per-step draw of stream s (0,1 Gaussian; 2 binomial drive; 3 Poisson source).
__host__ __device__ inline uint64_t rng_draw(uint64_t key, uint64_t t, unsigned s) {
  return splitmix_at(key, 4 * t + s);
}

enum : int { kIfCurrExp = 0, kIzhikevich = 1, kIfCondExp = 2 };
enum : int { kSrcNone = 0, kSrcConstRate = 1, kSrcPoisson = 2 };

// Everything a step kernel reads, passed by value as one kernel argument.
//
// Delivery: update(t) pushes the synapses of the ids spiking at t straight
// into a ring of accumulators, slot (t + k + 1) % R for API delay k, and
// update(t') consumes slot t' % R.  A step whose spikes are predicted to be
// dense is deferred instead: its spike bitmap stays in the history ring, the
// step is stamped, and the pull gather of each later step t' sums, per target
// block, the synapses of delay k = t' - s - 1 of every deferred step s.
struct DevNet {
  int32_t model;
  int32_t source_model;
  uint32_t N;        // neurons
  uint32_t M;        // input sources
  uint32_t T;        // N + M ids
  uint32_t D;        // delay slots (API delay k < D)
  uint32_t R;        // ring length = D + 1
  uint32_t W;        // raster words per step = ceil(T / 64)
  uint32_t npad;     // stride of the SoA arrays (N rounded up to 64)
  uint32_t drive_k;
  int32_t frac_bits;
  float acc_scale;   // 2^-frac_bits
  long long drive_w; // fixed-point weight of one drive event
  const float* params;      // [np][npad]
  float* vars;              // [nv][npad]
  long long* acc;           // [R][2][npad] fixed-point E/I input due at the slot's step
  long long* acc_pull;      // [G][2][npad] or null: a pulling step's delay-group sums
  uint32_t pull_groups;     // G: delay groups of the pull gather (1 .. kMaxPullGroups)
  uint32_t* barrier;        // grid barrier words of the persistent loop (kBarWords)
  const uint64_t* keys;     // [T] random-stream keys
  const float* noise_amp;   // [N] or null
  const float* src_rate;    // [M] or null
  float* src_last;          // [M] or null
  const uint64_t* src_thresh;  // [M] or null
  const uint64_t* drive_cdf;   // [drive_k] or null
  const uint32_t* syn_off;  // [T + 1] synapses of id at syn[syn_off[id] .. syn_off[id + 1])
  const uint2* syn;         // [S] (target | k << kSynDelayShift, fixed-point weight as int32 bits)
  const uint64_t* delay_bits;  // [T] bit k set iff id has a synapse of delay k
  uint64_t* dstamp;         // [R] s + 1 in slot s % R when step s was deferred (| kStampLate: deferred late)
  unsigned long long* ring_cnt;  // [R][2][kRingStripes] events and rows due at the slot's step
  unsigned long long* spk_ring;  // [4][kRingStripes] spikes of step s in slot s % 4
  uint32_t defer_spikes;    // defer step t when its predicted spikes >= this (0xffffffff: never)
  uint32_t forced_count;    // ids forced by a host step (for the prediction)
  uint32_t forced_prev;     // ids forced by the host step forced_prev_t (0: none)
  uint64_t forced_prev_t;
  uint64_t volley_t0;       // first step forced with the current volley
  uint64_t* raster;         // [rec_steps][W] or null
  uint64_t rec_first;
  uint64_t rec_steps;
  unsigned long long* counters;  // striped, see counter(): events spikes drive overflow rows
  const uint64_t* t_base;   // device cycle counter
  uint32_t volley;          // ids [0, volley) are forced every step
  // Heavy pushes: in single-step launches a block whose spiking ids carry at
  // least heavy_min synapse entries posts their ranges instead of pushing
  // them (one CU alone pushes only a small share of the device's atomic
  // rate), and the next step's gather launch (or the launch before a
  // persistent loop) pushes them with its whole grid.  Step s posts into
  // list s & 1: heavy_cnt[s & 1] ranges at heavy_list[(s & 1) * T ..].
  uint32_t heavy_min;       // 0: every block pushes its own ids (default kHeavyPushMin)
  uint32_t* heavy_cnt;      // [2]
  uint2* heavy_list;        // [2][T] (first synapse, synapses)
  // Late deferral: when the ids that step s posted for heavy pushes reach
  // half the deferral threshold (defer_late), the gather launch of s + 1
  // pulls step s instead of pushing them (a burst after quiet steps, which
  // the prediction missed).
  // hvy slot s % R holds the spike words of the blocks that posted at s (zero
  // for the others, whose spikes were pushed), and the pull of a step stamped
  // late stages those words instead of the full history.
  uint64_t* hvy;            // [R][W] or null (no heavy pushes)
  const uint64_t* forced_bits;   // [W] forced ids this step, or null
  uint64_t* hist;           // [R][W] spike bitmaps of the last R steps (slot t % R)
  // pull layout (dense gather), a sliced ELLPACK: target blocks of
  // pull_tb = 64 * pull_slices neurons; segment (b, k) holds the delay-k
  // synapses into block b in pull_slices slices of 64 targets, slice s at
  // entries pull_off[(b*D+k)*pull_slices + s] .. (next), lane L serving
  // target pull_lane[(b*D+k)*pull_tb + s*64 + L] & kPullTMask of the block,
  // whose entries fill the first pull_lane[...] >> kPullShift (kPullLaneAll:
  // all the slice's) vectors of its
  // column; 16-byte vectors [L_s/V][64][V].  Entry formats (pull_fmt): see kPullWide, kPullCount,
  // kPullTable; rel counts from source pull_lo[b]*64.
  const uint32_t* pull_off; // [NB*D*pull_slices + 1] or null (no pull layout)
  const uint16_t* pull_lane;  // [NB*D*pull_tb] target | vectors << kPullShift
  const uint2* pull_syn;    // wide entries, or null
  const uint32_t* pull_syn32;  // table-format entries, or null
  const uint16_t* pull_syn16;  // count-format entries, or null
  const int32_t* pull_wtab; // count, table: [pull_wn] weight of each index (index 0: 0)
  uint32_t pull_fmt;        // kPullWide, kPullCount or kPullTable
  uint32_t pull_wn;         // count, table: weight-table entries
  const uint32_t* pull_lo;  // [NB] first raster word of the block's sources
  uint32_t pull_nb;         // NB
  uint32_t pull_slices;     // slices per segment (kPullMinSlices .. kPullSlices)
  uint32_t pull_tb;         // targets per block, 64 * pull_slices
  uint32_t pull_null;       // the null vector (16 bytes of padding entries after the last slice), in vectors
  uint32_t pull_words;      // LDS words staged per block (max over blocks)
  uint32_t pull_stage;      // gather launches: delays whose bitmaps are staged at once (>= 1)
  uint32_t pull_batch;      // gather launches: 16-byte vectors per lane and batch (2 or 4)
};

// 64-bit words of a pull block's LDS bitmap: the staged words, the zero word
// the padding entries read, rounded up to an even count so the FP64 pairs
// after it stay 16-byte aligned.
__host__ __device__ inline uint32_t pull_bitmap_u64(uint32_t words) { return (words + 2) & ~1u; }

// Counter slots.
enum : int {
  kCntEvents = 0, kCntSpikes = 1, kCntDrive = 2, kCntOverflow = 3, kCntRows = 4,
  kCntDeferred = 5,  // steps deferred to the pull (predicted or late)
  kCntLate = 6,      // of them, deferred late
  kCntSlots = 7
};

// Load of data that no kernel writes (the network's layout tables) at a
// wave-uniform address, through the scalar cache (constant address space), so
// the wave's waits for it do not also wait for its vector loads in flight.
template <class T>
__device__ inline T ld_const(const T* p) {
  return *(const __attribute__((address_space(4))) T*)p;
}

// Agent-coherent loads and stores (relaxed agent-scope atomics, `sc1` on
// gfx950) for the state that one workgroup writes and another reads within a
// launch: accumulators, spike history words, deferral stamps and the
// delivery-time counter rings.  They are not served from the per-CU caches
// nor left dirty in a per-XCD L2, so the persistent loop's grid barrier needs
// no L2 write-back or invalidate, only each wave's outstanding accesses
// completed.
template <class T>
__device__ inline T ld_agent(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline void st_agent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier of the persistent loop: kBarLeaves leaf counters, a root
// counter and a generation word (64-bit words), each on its own 128-byte line.
// 32 leaves (8 arrivals each for a 256-block grid) measured fastest for the
// natural step of the 100k network: 5.70 us against 5.81 (16), 5.87 (64),
// 6.32 (128) and 6.43 (one flat counter).
constexpr uint32_t kBarLeaves = 32;
constexpr uint32_t kBarStride = 32;  // u32 words per line
constexpr uint32_t kBarWords = (kBarLeaves + 2) * kBarStride;
// Leaf and root words: arrivals in the low kBarCountBits, the carried sum above.
constexpr uint32_t kBarCountBits = 16;
constexpr unsigned long long kBarCountMask = (1ull << kBarCountBits) - 1;
// Spin limit of a barrier wait (about a second); reaching it (never with a
// co-resident grid) adds kBarTimeout to the overflow counter and ends the
// kernel.
constexpr uint32_t kBarSpinLimit = 1u << 22;
constexpr unsigned long long kBarTimeout = 1ull << 32;

// Counters are striped over kCntStripes cache lines (128 B apart) so that
// blocks adding to one counter do not serialise on a single address; the host
// sums the stripes.  counters = [kCntSlots][kCntStripes][16] u64.  Events and
// rows are counted when delivered (at the step that consumes them).
constexpr uint32_t kCntStripes = 256;
constexpr uint32_t kCntStride = 16;
constexpr size_t kCntWords = (size_t)kCntSlots * kCntStripes * kCntStride;
__host__ __device__ inline unsigned long long* counter(unsigned long long* base, int slot, uint32_t stripe) {
  return base + ((size_t)slot * kCntStripes + (stripe % kCntStripes)) * kCntStride;
}

// Delivery-time counter ring: per slot and stripe one word, the rows due at
// the slot's step in the low 32 bits and the events above (each at most the
// stored synapse count, < 2^32, per step: a synapse delivers at most once per
// step and a row holds at least one synapse), so one atomic adds both; each
// entry on its own 128-byte line.
constexpr uint32_t kRingStripes = 64;
constexpr size_t kRingCntWords(uint32_t R) { return (size_t)R * kRingStripes * kCntStride; }
constexpr size_t kSpkRingWords = (size_t)4 * kRingStripes * kCntStride;
__host__ __device__ inline unsigned long long ring_pack(uint32_t rows, uint32_t events) {
  return (unsigned long long)rows | ((unsigned long long)events << 32);
}
__host__ __device__ inline unsigned long long* ring_counter(unsigned long long* base, uint32_t slot, uint32_t stripe) {
  return base + ((size_t)slot * kRingStripes + (stripe % kRingStripes)) * kCntStride;
}
__host__ __device__ inline unsigned long long* spk_counter(unsigned long long* base, uint64_t step, uint32_t stripe) {
  return base + ((size_t)(step & 3) * kRingStripes + (stripe % kRingStripes)) * kCntStride;
}

// Whether step t is deferred to the pull: its predicted spikes reach
// defer_spikes.  The prediction is the previous step's spikes less the ids
// forced in it (they are forced again or not, as this step says) plus the
// ids forced in this step.  Inside a persistent launch (no host-forced ids,
// the volley forced in every step) it is the previous step's spikes from
// the launch's second step on, which the loop uses directly.
// Deferral stamps: slot s % R holds s + 1 when step s was deferred, with
// kStampLate set when the gather launch of s + 1 deferred it (late deferral).
constexpr uint64_t kStampLate = 1ull << 63;
__host__ __device__ inline uint64_t stamp_step(uint64_t v) { return v & ~kStampLate; }

__device__ inline bool defer_step(const DevNet& n, uint64_t t, unsigned long long spk_prev) {
  if (!n.pull_off) return false;
  const unsigned long long fprev =
      (t > n.volley_t0 ? n.volley : 0u) + (n.forced_prev && t == n.forced_prev_t + 1 ? n.forced_prev : 0u);
  const unsigned long long own = spk_prev > fprev ? spk_prev - fprev : 0ull;
  return own + n.volley + n.forced_count >= n.defer_spikes;
}

// Default heavy-push threshold: synapse entries of one update block's spiking
// ids (one push batch of a 256-thread block).  Lower thresholds move the
// pushes of networks with short rows into gather blocks that each take one
// row per pass; higher ones leave bursts of a few thousand spread ids to the
// update blocks, whose pushes run at a small share of the device's atomic rate.
constexpr uint32_t kHeavyPushMin = 1024;

// Late deferral: step s is pulled instead of pushing its heavy-push list
// when the list holds at least half the deferral threshold's spikes.  The
// threshold prices a step's spikes at the spread update pushes' rate (about
// 2.3e10 synapse events per second); the gather grid pushes a posted list
// at about half that (on 1e8 synapses, 1000 contiguous ids forced every
// step: 97.7 us per step pushed, 60.1 us pulled; 250 ids: 43.1 against 60.4).
__host__ __device__ inline bool defer_late(uint32_t posted, uint32_t defer_spikes) {
  return defer_spikes != 0xffffffffu && 2ull * posted >= defer_spikes;
}

// Launchers (szsnn_kernels.hip).  `j` is the step offset from *t_base.
// gather: the pull of step t (launches nothing without a pull layout).
// With events e0 and e1, the kernel's dispatch records them at its start and
// end (the same for launch_update).
hipError_t launch_gather(const DevNet& n, uint32_t j, uint32_t block, uint32_t grid, hipStream_t s,
                         hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// The heavy pushes posted by step t_base - 1, then both lists emptied (before
// a persistent launch and at threshold changes; the gather launch of a step
// pushes its previous step's list).
hipError_t launch_heavy(const DevNet& n, uint32_t grid, hipStream_t s);
// update: block 64, 128, 256, 512 or 1024; npl (ids per lane) 1, 2 or 4.
hipError_t launch_update(const DevNet& n, uint32_t j, uint32_t block, uint32_t npl, hipStream_t s,
                         hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// One cooperative launch running nsteps steps (block 256 or 512; grid must
// not exceed persistent_max_blocks_per_cu(model, block) * CUs).
hipError_t launch_persistent(const DevNet& n, uint64_t nsteps, uint64_t* t_base, uint32_t block, uint32_t grid,
                             hipStream_t s);
int persistent_max_blocks_per_cu(const DevNet& n, uint32_t block);
// dynamic LDS bytes of a pull with `stage` delays staged at once (the
// persistent loop stages one)
size_t pull_dyn_bytes(const DevNet& n, uint32_t stage);
// The same for a layout of `words` staged bitmap words, entry format `fmt`
// and a weight table of `wn` entries.
size_t pull_dyn_bytes_for(uint32_t words, uint32_t fmt, uint32_t wn, uint32_t stage);
// Resident gather blocks of `block` threads per CU with `dyn` bytes of LDS,
// the given staging depth and batch (vectors per lane: 2 or 4; 0 on error).
int gather_max_blocks_per_cu(size_t dyn, uint32_t block, uint32_t stage, uint32_t batch = 2);
// LDS the delays staged at once by one gather block may take (their bitmaps).
constexpr size_t kPullStageBytes = 51200;
hipError_t launch_advance(uint64_t* t_base, uint64_t k, hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t value, uint64_t n, hipStream_t s);

}  // namespace szsnn
