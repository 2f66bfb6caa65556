// szsnn_kernels.hip -- per-step CDNA4 kernels of the SZSNN core (gfx950, wave64).
//
// Per step t (reference order, szsnn1.1/src/Simulation.ipp:471-483):
//   gather(t)  : only when a deferred (dense) step has synapses due at t: per
//                target block, sum the due delays' synapses of the deferred
//                steps' spiking sources (the pull layout), no atomics.
//   update(t)  : one id per lane, SoA state; consume accumulator slot t % R;
//                neuron model or input-source model; spike words by wavefront
//                ballot into the history ring and raster; then the block
//                pushes its spiking ids' synapses into the accumulator slots
//                of their delivery steps t + k + 1 (fixed-point atomics), or,
//                on a step predicted dense, stamps the step for the pull.
//                Replaces *_update_kernel.cl, scatter_kernel.cl:97-355 and
//                gather_kernel.cl:67-200 (their per-delay queues become the
//                ring of per-delivery-step accumulator slots).
//
// Arithmetic is FP32 with one written evaluation order; the library is built
// with -ffp-contract=off and IEEE division, so results are bit-identical to
// the FP32 CPU restatement of the same equations.
#include "szsnn_internal.h"

#include <hip/hip_ext.h>

namespace szsnn {

namespace {

__device__ inline int32_t gauss_int(uint64_t key, uint64_t t) {
  const uint64_t a = rng_draw(key, t, 0);
  const uint64_t b = rng_draw(key, t, 1);
  uint32_t sum = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    sum += (uint32_t)((a >> (16 * j)) & 0xffffu);
    sum += (uint32_t)((b >> (16 * j)) & 0xffffu);
  }
  return (int32_t)(2u * sum) - 524280;
}

// Unit-variance scale of the centred Irwin-Hall sum: 0.5/sqrt(8*(2^32-1)/12).
constexpr float kGaussScale = 0x1.3988e2p-17f;

// Per-model counts of SoA parameter and state rows.
template <int MODEL>
struct ModelRows {
  static constexpr int kParams = MODEL == kIfCurrExp ? 9 : MODEL == kIzhikevich ? 4 : 11;
  static constexpr int kVars = MODEL == kIzhikevich ? 2 : 4;
};

// One neuron's parameters and state, loaded into registers at the top of the
// update so every load of the step is in flight before the first barrier.
template <int MODEL>
struct NeuronRegs {
  float p[ModelRows<MODEL>::kParams];
  float x[ModelRows<MODEL>::kVars];
};

template <int MODEL>
__device__ inline void load_neuron(const DevNet& n, uint32_t i, NeuronRegs<MODEL>& r) {
#pragma unroll
  for (int k = 0; k < ModelRows<MODEL>::kParams; ++k) r.p[k] = n.params[(size_t)k * n.npad + i];
#pragma unroll
  for (int k = 0; k < ModelRows<MODEL>::kVars; ++k) r.x[k] = n.vars[(size_t)k * n.npad + i];
}

template <int MODEL>
__device__ inline void store_neuron(const DevNet& n, uint32_t i, const NeuronRegs<MODEL>& r) {
#pragma unroll
  for (int k = 0; k < ModelRows<MODEL>::kVars; ++k) n.vars[(size_t)k * n.npad + i] = r.x[k];
}

// IF_curr_exp (IF_curr_exp_update_kernel.cl:230-297) and IF_cond_exp (same
// scheme, conductances in the current slots).  Returns spike.
// The device parameter rows 1, 2, 4, 5 hold 1/cm, 1/tau_m, 1 - 1/tau_syn_E and
// 1 - 1/tau_syn_I, formed once at upload with the same IEEE FP32 operations the
// step would repeat (the reference evaluates them every step).
template <int MODEL>
__device__ inline bool if_update(NeuronRegs<MODEL>& r, float accE, float accI, float iinj, bool forced) {
  const float* p = r.p;
  float* x = r.x;
  const float v_rest = p[0], rcm = p[1], rtau_m = p[2];
  const float tau_refrac = p[3], decay_e = p[4], decay_i = p[5];
  const float i_offset = p[6], v_reset = p[7], v_thresh = p[8];
  float v = x[0];
  const float se = (decay_e * x[1]) + accE;
  const float si = (decay_i * x[2]) + accI;
  float I;
  if constexpr (MODEL == kIfCurrExp) {
    I = ((se + si) + iinj) + i_offset;
  } else {
    const float e_rev_e = p[9], e_rev_i = p[10];
    I = (((se * (e_rev_e - v)) - (si * (e_rev_i - v))) + iinj) + i_offset;
  }
  float fired = x[3];
  const bool refr = fired <= tau_refrac;
  fired = fired + 1.0f;
  if (!refr) {
    // x / y as x * (1/y), the reference's FP32 division (no correctly
    // rounded divide in its build, Simulation.ipp:13-21)
    const float delta = (I * rcm) + ((v_rest - v) * rtau_m);
    v = v + delta;
  }
  const bool spike = (!refr && v >= v_thresh) || forced;
  if (spike) {
    v = v_reset;
    fired = 1.0f;
  }
  x[0] = v;
  x[1] = se;
  x[2] = si;
  x[3] = fired;
  return spike;
}

// Izhikevich (Izhikevich_update_kernel.cl:219,247-251,264-268): one Euler step
// of both equations from the old (v, u).
__device__ inline bool izh_update(NeuronRegs<kIzhikevich>& r, float accE, float accI, float iinj, bool forced) {
  const float a = r.p[0], b = r.p[1];
  float v = r.x[0], u = r.x[1];
  const float I = (accE + accI) + iinj;
  const float dv = ((((0.04f * v) + 5.0f) * v + 140.0f) - u) + I;
  const float du = a * ((b * v) - u);
  v = v + dv;
  u = u + du;
  const bool spike = (v >= 30.0f) || forced;
  if (spike) {
    v = r.p[2];
    u = u + r.p[3];
  }
  r.x[0] = v;
  r.x[1] = u;
  return spike;
}

// s_waitcnt immediate of gfx9-family targets waiting for vmcnt = 0 only
// (expcnt 7 and lgkmcnt 15 mean no wait on those counters).
constexpr int kWaitVmcnt0 = 0x0F70;

// Load of state that other blocks write: inside the persistent loop (AGENT)
// written by the same launch, so agent-coherent; in a single-step launch
// written by earlier launches (a kernel boundary makes them visible), so a
// plain, cacheable load (many blocks reading the same words then hit in L2
// instead of all going to memory).
template <bool AGENT, class T>
__device__ inline T ld_step(const T* p) {
  if constexpr (AGENT) return ld_agent(p);
  else return *p;
}

// x of lane j (j uniform across the wave).
__device__ inline uint64_t readlane64(uint64_t x, uint32_t j) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, (int)j);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), (int)j);
  return ((uint64_t)hi << 32) | lo;
}

// Block-wide exclusive scan of one u32 per thread; returns the total.
template <int BLK>
__device__ inline uint32_t block_exclusive_scan(uint32_t x, uint32_t* s_pref, uint32_t* s_wave) {
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t inc = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(inc, off);
    if (lane >= (uint32_t)off) inc += y;
  }
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int w = 0; w < BLK / 64; ++w) {
      const uint32_t v = s_wave[w];
      s_wave[w] = acc;
      acc += v;
    }
    s_wave[BLK / 64] = acc;
  }
  __syncthreads();
  s_pref[threadIdx.x] = s_wave[wave] + inc - x;
  if (threadIdx.x == 0) s_pref[BLK] = s_wave[BLK / 64];
  __syncthreads();
  return s_pref[BLK];
}

// Bit k set iff step t - k - 1 was deferred, i.e. its synapses of delay k
// are pulled at t.  Every lane of the calling wave gets the mask.
// `late` gets bit k set when that step was deferred late (its pull stages
// the posted spike words, hvy, instead of the history).
template <bool AGENT>
__device__ inline uint64_t due_mask(const DevNet& n, uint64_t t, uint64_t& late) {
  late = 0;
  if (!n.pull_off) return 0;
  const uint32_t k = __lane_id();
  bool due = false, lt = false;
  if (k < n.D && t >= (uint64_t)k + 1) {
    const uint64_t s = t - k - 1;
    const uint64_t v = ld_step<AGENT>(&n.dstamp[s % n.R]);
    due = stamp_step(v) == s + 1;
    lt = due && (v & kStampLate);
  }
  late = __ballot(lt);
  return __ballot(due);
}

// A block pushing at most this many spiking ids walks them from an unordered
// LDS list; more take the block-wide scan.
constexpr uint32_t kSparseIds = 8;

// Push entries loaded per thread and batch.
constexpr uint32_t kPushBatch = 4;

// Adds a thread's batch of loaded push entries into the accumulator ring and
// counts them per delay in the block's LDS ev (flushed once per (block,
// delay)).  Every loaded entry is consumed before the first atomic, so the
// atomics issue back to back.  Consecutive lanes hold consecutive entries,
// sorted by delay within an id, so one LDS add per run of equal delays counts
// the events.
template <uint32_t NB>
__device__ inline void push_entries(const DevNet& n, const uint2 (&sw)[NB], uint32_t first_slot, uint32_t* ev) {
  const uint32_t lane = __lane_id();
  long long* dst[NB];
#pragma unroll
  for (uint32_t u = 0; u < NB; ++u) {
    const uint32_t k = sw[u].x >> kSynDelayShift;
    const bool valid = sw[u].y != 0u;
    const uint32_t kp = __shfl_up(k, 1);
    const bool head = valid && (lane == 0 || kp != k);
    const uint64_t bound = __ballot(head) | ~__ballot(valid);
    uint32_t ds = first_slot + k;
    if (ds >= n.R) ds -= n.R;
    if (head) {
      const uint64_t rest = lane == 63 ? 0ull : bound >> (lane + 1);
      atomicAdd(&ev[k], rest ? (uint32_t)__ffsll((unsigned long long)rest) : 64u - lane);
    }
    const int32_t w = (int32_t)sw[u].y;
    dst[u] = n.acc + (size_t)ds * 2 * n.npad + (w < 0 ? n.npad : 0u) + (sw[u].x & kSynTargetMask);
  }
#pragma unroll
  for (uint32_t u = 0; u < NB; ++u)
    if (sw[u].y != 0u) atomicAdd((unsigned long long*)dst[u], (unsigned long long)(long long)(int32_t)sw[u].y);
}

// Whether a block whose spiking ids carry `total` synapse entries posts them
// as heavy pushes (single-step launches only: inside the persistent loop
// every block pushes its own ids).
__device__ inline bool heavy_block(const DevNet& n, uint32_t total, bool agent) {
  return !agent && n.heavy_min && total >= n.heavy_min;
}

// Posts the synapse ranges (first, count) of the wave's lanes with `has`
// set, for step t: one counter atomic per wave.  Every lane of the wave
// must call it.
__device__ inline void post_heavy(const DevNet& n, uint64_t t, bool has, uint2 range) {
  const uint64_t m = __ballot(has);
  if (!m) return;
  uint32_t base = 0;
  if (__lane_id() == 0) base = atomicAdd(&n.heavy_cnt[t & 1], (uint32_t)__popcll(m));
  base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
  const uint32_t q = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  if (has) n.heavy_list[(size_t)(t & 1) * n.T + q] = range;
}

// Pushes the hc heavy ranges posted by step s: range q by block q (mod the
// grid), all its threads; events counted per (block, delay) into the
// delivery ring like the update's own pushes.  Every thread of the block
// must call it.
template <int BLK>
__device__ inline void heavy_push(const DevNet& n, uint64_t s, uint32_t hc, uint32_t* ev) {
  for (uint32_t i = threadIdx.x; i < 64; i += BLK) ev[i] = 0;
  __syncthreads();
  const uint32_t slot = (uint32_t)(s % n.R);
  const uint32_t first_slot = slot + 1 == n.R ? 0u : slot + 1;
  const uint2* list = n.heavy_list + (size_t)(s & 1) * n.T;
  for (uint32_t q = blockIdx.x; q < hc; q += gridDim.x) {
    const uint2 range = list[q];
    for (uint32_t e0 = 0; e0 < range.y; e0 += BLK * kPushBatch) {
      uint2 sw[kPushBatch];
#pragma unroll
      for (uint32_t u = 0; u < kPushBatch; ++u) {
        const uint32_t e = e0 + u * BLK + threadIdx.x;
        sw[u] = e < range.y ? n.syn[range.x + e] : make_uint2(0u, 0u);
      }
      push_entries<kPushBatch>(n, sw, first_slot, ev);
    }
  }
  __syncthreads();
  if (threadIdx.x < n.D) {
    uint32_t ds = first_slot + threadIdx.x;
    if (ds >= n.R) ds -= n.R;
    if (ev[threadIdx.x]) atomicAdd(ring_counter(n.ring_cnt, ds, blockIdx.x), ring_pack(0u, ev[threadIdx.x]));
  }
}

// LDS of one update block.
template <int BLK>
struct UpdateSmem {
  uint64_t cdf[32];          // the binomial drive's CDF thresholds
  uint32_t cnt[64];          // the block's rows per delay (spiking ids with that delay)
  uint32_t ev[64];           // the block's pushed synapses per delay
  uint32_t start[BLK];       // push: first synapse of each lane's id
  uint32_t pref[BLK + 1];    // push: exclusive scan of the lanes' synapse counts
  uint32_t wave[BLK / 64 + 1];
  uint2 list[kSparseIds];    // push, few spiking ids: (first synapse, count), any order
  uint32_t nlist;            // spiking ids with synapses this step (list holds the first kSparseIds)
  unsigned long long spikes, drive;
  unsigned long long spk_prev;  // spikes of step t - 1
};

// Block-uniform facts of step t known before it starts: the pull's due mask
// (bit k: step t - k - 1 was deferred) and the spikes of step t - 1.  The
// persistent loop carries them across steps (the grid barrier returns each
// step's spike total); in a single-step launch every wave reads the due
// mask itself (update_step) and step_facts reads the spikes.
struct StepFacts {
  uint64_t due;
  uint64_t late;  // the due steps deferred late (due_mask)
  unsigned long long spk_prev;
  bool defer;  // the step's deferral decision (carried: taken by the loop)
};

// Block-uniform facts of step t, gathered by one wave (the block's first)
// with all its loads issued together: the spikes of step t - 1 into
// sm.spk_prev (LDS atomics of the stripes, no wave reduction).  With `fold`
// (block 0), lane s also reads stripe s of the rows and events due at t
// (packed) into fold; fold_finish moves them into the counters at the end of
// the step.
template <int BLK, bool AGENT>
__device__ inline void step_facts(const DevNet& n, uint64_t t, uint32_t slot, bool do_fold, UpdateSmem<BLK>& sm,
                                  unsigned long long& fold) {
  static_assert(kRingStripes == 64, "one stripe per lane");
  const uint32_t lane = __lane_id();
  if (lane == 0) sm.spk_prev = 0;
  // unconditional loads at always-valid addresses (no divergent branch between
  // them, so they are in flight together)
  const unsigned long long spk = ld_step<AGENT>(spk_counter(n.spk_ring, t - 1, lane));
  fold = do_fold ? ld_step<AGENT>(ring_counter(n.ring_cnt, slot, lane)) : 0ull;
  if (t > 0 && spk) atomicAdd(&sm.spk_prev, spk);
}

// Adds a packed delivery-ring word (rows low, events high) to stripe q of the
// rows and events counters.
__device__ inline void fold_counters(const DevNet& n, uint32_t q, unsigned long long v) {
  if (v >> 32) atomicAdd(counter(n.counters, kCntEvents, q), v >> 32);
  if (v & 0xffffffffull) atomicAdd(counter(n.counters, kCntRows, q), v & 0xffffffffull);
}

// Block 0, one wave, end of step t: add the events and rows due at t (read
// by step_facts) to the counters, stripe by stripe, clear their ring
// stripes, and clear the spike-count slot that step t + 1 will fill.  Done
// last so that these writes do not hold up the step's own loads.
__device__ inline void fold_finish(const DevNet& n, uint64_t t, uint32_t slot, unsigned long long fold) {
  const uint32_t lane = __lane_id();
  if (fold) {
    st_agent(ring_counter(n.ring_cnt, slot, lane), 0ull);
    fold_counters(n, lane, fold);
  }
  st_agent(spk_counter(n.spk_ring, t + 1, lane), 0ull);
}

// What a lane keeps about its id across steps: the neuron's parameters and
// state (or the source's), its random-stream key, synapse range and delay
// bits.  Loaded from memory every step, or, in the persistent loop when each
// lane owns one id, once per launch and carried in registers.
template <int MODEL>
struct LaneState {
  NeuronRegs<MODEL> nr;
  uint64_t key, dbits, src_thresh;
  uint32_t r0, r1;
  float amp, src_last, src_rate;
};

template <int MODEL>
__device__ inline void load_lane(const DevNet& n, uint32_t id, LaneState<MODEL>& ls) {
  ls.key = ls.dbits = ls.src_thresh = 0;
  ls.r0 = ls.r1 = 0;
  ls.amp = ls.src_last = ls.src_rate = 0.0f;
  if (id < n.N) {
    load_neuron<MODEL>(n, id, ls.nr);
    if (n.noise_amp) ls.amp = n.noise_amp[id];
  } else if (id < n.T) {
    const uint32_t s = id - n.N;
    if (n.source_model == kSrcConstRate) {
      ls.src_last = n.src_last[s];
      ls.src_rate = n.src_rate[s];
    } else if (n.source_model == kSrcPoisson) {
      ls.src_thresh = n.src_thresh[s];
    }
  }
  if (id < n.T) {
    ls.r0 = n.syn_off[id];
    ls.r1 = n.syn_off[id + 1];
    ls.dbits = n.delay_bits[id];
    ls.key = n.keys[id];
  }
}

// The state a step changed: neuron variables, a Const_rate source's counter.
template <int MODEL>
__device__ inline void store_lane(const DevNet& n, uint32_t id, const LaneState<MODEL>& ls) {
  if (id < n.N)
    store_neuron<MODEL>(n, id, ls.nr);
  else if (id < n.T && n.source_model == kSrcConstRate)
    n.src_last[id - n.N] = ls.src_last;
}

// update(t) for one id per lane; every thread of the block must call it (it
// synchronises the block).  Consumes accumulator slot t % R (plus the pull
// groups' sums when the step pulled), updates the neuron or source, writes
// the spike word of each wave to the history ring and the raster, and then
// either pushes the block's spiking ids' synapses into the accumulator slots
// of their delivery steps or, on a deferred step, stamps the step for the
// pull gather.  Rows and events are added to the delivery-time counter ring
// with one atomic per (block, delay).  CARRY: `ls` already holds the lane's
// state and keeps it (the caller stores it at the end); otherwise it is
// loaded and stored here, and the drive table is staged in LDS here.
// `first`: the block's first id block of the step (a block that serves
// several id blocks folds the delivery-time counters once).
template <int MODEL, int BLK, bool CARRY, bool AGENT>
__device__ inline unsigned long long update_step(const DevNet& n, uint64_t t, uint32_t id, UpdateSmem<BLK>& sm,
                                                 LaneState<MODEL>& ls, const StepFacts& facts,
                                                 unsigned long long& out_drive, bool first = true) {
  const bool folder = first && blockIdx.x == 0;
  const uint32_t lane = __lane_id();
  const uint32_t slot = (uint32_t)(t % n.R);
  // every load the lane needs, issued before the first barrier.  Without
  // the carried facts, every wave reads the deferral stamps first (the reads
  // step_facts makes): when the step pulled, the delay groups' sums are
  // loaded with the lane's other loads.
  uint64_t pulled = 0;
  if (!CARRY && n.pull_off) {
    const uint32_t k = lane < n.D ? lane : 0u;
    const uint64_t v = ld_step<AGENT>(&n.dstamp[(t - k - 1) % n.R]);
    pulled = __ballot(lane < n.D && t >= (uint64_t)lane + 1 && stamp_step(v) == t - lane);
  }
  long long ring_e = 0, ring_i = 0;
  uint64_t fbits = 0;
  long long* ring = n.acc + (size_t)slot * 2 * n.npad;
  if (id < n.N) {
    ring_e = ld_step<AGENT>(&ring[id]);
    ring_i = ld_step<AGENT>(&ring[n.npad + id]);
  }
  long long pe[kMaxPullGroups], pi[kMaxPullGroups];
#pragma unroll
  for (uint32_t g = 0; g < kMaxPullGroups; ++g) pe[g] = pi[g] = 0;
  if (!CARRY && pulled && id < n.N) {
#pragma unroll
    for (uint32_t g = 0; g < kMaxPullGroups; ++g)
      if (g < n.pull_groups) {
        pe[g] = ld_step<AGENT>(&n.acc_pull[(size_t)g * 2 * n.npad + id]);
        pi[g] = ld_step<AGENT>(&n.acc_pull[(size_t)g * 2 * n.npad + n.npad + id]);
      }
  }
  if (!CARRY) {
    load_lane<MODEL>(n, id, ls);
    if (threadIdx.x < n.drive_k) sm.cdf[threadIdx.x] = n.drive_cdf[threadIdx.x];
  }
  if (id < n.T && n.forced_bits) fbits = n.forced_bits[id >> 6];
  unsigned long long fold = 0;
  if (threadIdx.x < 64) {
    sm.cnt[threadIdx.x] = 0;
    sm.ev[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
      sm.spikes = sm.drive = 0;
      sm.nlist = 0;
    }
    if (!CARRY) step_facts<BLK, AGENT>(n, t, slot, folder, sm, fold);
  }
  __syncthreads();
  const unsigned long long spk_prev = CARRY ? facts.spk_prev : sm.spk_prev;
  // step t is deferred when its predicted spikes reach the threshold (the
  // same decision in every block)
  const bool defer = CARRY ? facts.defer : defer_step(n, t, spk_prev);
  const uint32_t r0 = ls.r0, r1 = ls.r1;
  const uint64_t dbits = ls.dbits;
  bool spike = false;
  uint32_t kdrive = 0;
  if (id < n.T) {
    const bool forced = id < n.volley || ((fbits >> (id & 63)) & 1ull);
    if (id < n.N) {
      long long ae = ring_e, ai = ring_i;
      if (CARRY ? facts.due != 0 : pulled != 0) {
        // pulled step: the delay groups' sums (rewritten by every pull); with
        // carried facts they are loaded here, all before the sum
        if (CARRY) {
#pragma unroll
          for (uint32_t g = 0; g < kMaxPullGroups; ++g)
            if (g < n.pull_groups) {
              pe[g] = ld_step<AGENT>(&n.acc_pull[(size_t)g * 2 * n.npad + id]);
              pi[g] = ld_step<AGENT>(&n.acc_pull[(size_t)g * 2 * n.npad + n.npad + id]);
            }
        }
#pragma unroll
        for (uint32_t g = 0; g < kMaxPullGroups; ++g) {
          ae += pe[g];
          ai += pi[g];
        }
      }
      if (n.drive_k) {
        // events = #{j : u >= cdf[j]}, the inverse-CDF index.  The thresholds
        // ascend, so the count stops at the first one no lane of the wave
        // reaches (reads stay uniform LDS broadcasts).
        const uint64_t u = rng_draw(ls.key, t, 2);
        uint32_t k = 0;
        for (uint32_t j = 0; j < n.drive_k; ++j) {
          const bool ge = u >= sm.cdf[j];
          if (!__any(ge)) break;
          k += ge ? 1u : 0u;
        }
        ae += (long long)k * n.drive_w;
        kdrive = k;
      }
      const float accE = (float)ae * n.acc_scale;
      const float accI = (float)ai * n.acc_scale;
      const float iinj = n.noise_amp ? ls.amp * ((float)gauss_int(ls.key, t) * kGaussScale) : 0.0f;
      if constexpr (MODEL == kIzhikevich)
        spike = izh_update(ls.nr, accE, accI, iinj, forced);
      else
        spike = if_update<MODEL>(ls.nr, accE, accI, iinj, forced);
    } else {
      spike = forced;
      if (n.source_model == kSrcConstRate) {
        // Const_rate_update_kernel.cl:213-227
        float last = ls.src_last + 1.0f;
        spike = spike || (last == ls.src_rate);
        if (spike) last = 0.0f;
        ls.src_last = last;
      } else if (n.source_model == kSrcPoisson) {
        spike = spike || (rng_draw(ls.key, t, 3) < ls.src_thresh);
      }
    }
  }
  // Every load of the step so far has been consumed.  Saying so explicitly
  // (vmcnt 0) keeps the compiler from waiting on the stores and atomics issued
  // below at later branch joins, where it cannot tell which conditional loads
  // are still in flight.
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  const uint64_t bits = spike ? dbits : 0;
  // dense spike output and history, one u64 word per wavefront
  // (storeDenseSpikes, IF_curr_exp_update_kernel.cl:58-94)
  const uint64_t mask = __ballot(spike);
  const uint32_t word = id >> 6;
  if (lane == 0 && word < n.W) {
    st_agent(&n.hist[(size_t)slot * n.W + word], mask);
    if (n.raster && t >= n.rec_first && t - n.rec_first < n.rec_steps)
      n.raster[(t - n.rec_first) * n.W + word] = mask;
    if (mask) atomicAdd(&sm.spikes, (unsigned long long)__popcll(mask));
  }
  if (kdrive) atomicAdd(&sm.drive, (unsigned long long)kdrive);
  if (mask) {
    // rows per delay: the wave's spiking ids with a synapse of that delay
    if (__popcll(mask) <= 8) {
      // few spikes: lane k counts bit k of each spiking lane's delay bits
      uint32_t c = 0;
      for (uint64_t m = mask; m; m &= m - 1) {
        const uint64_t b = readlane64(bits, (uint32_t)__ffsll((unsigned long long)m) - 1);
        c += (uint32_t)((b >> lane) & 1ull);
      }
      if (c) atomicAdd(&sm.cnt[lane], c);
    } else {
      // lane k keeps the count of delay k from one ballot per delay (a
      // select, no branch per delay), then one LDS add per lane.  The first
      // 32 delays unrolled (bits of delays past D are zero: counts of 0)
      const uint32_t blo = (uint32_t)bits, bhi = (uint32_t)(bits >> 32);
      uint32_t c = 0;
#pragma unroll
      for (uint32_t k = 0; k < 32; ++k) {
        const uint32_t pk = (uint32_t)__popcll(__ballot((blo >> k) & 1u));
        c = lane == k ? pk : c;
      }
      for (uint32_t k = 32; k < n.D; ++k) {
        const uint32_t pk = (uint32_t)__popcll(__ballot((bhi >> (k - 32)) & 1u));
        c = lane == k ? pk : c;
      }
      if (c) atomicAdd(&sm.cnt[lane], c);
    }
  }
  // spiking ids with synapses, on a pushed step: appended to the block's list
  // (one LDS atomic per wave; the order is free, the sums are exact)
  const bool has = !defer && spike && r1 > r0;
  const uint64_t hm = __ballot(has);
  if (hm) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&sm.nlist, (uint32_t)__popcll(hm));
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
    const uint32_t pos = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(hm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)hm, 0u));
    if (has && pos < kSparseIds) sm.list[pos] = make_uint2(r0, r1 - r0);
  }
  // the consumed slot is cleared now: nothing later in the step reads it,
  // and no push of this step targets it (delays k < D land in the other
  // R - 1 slots)
  if (ring_e) st_agent(&ring[id], 0ll);
  if (ring_i) st_agent(&ring[n.npad + id], 0ll);
  __syncthreads();
  const uint32_t nlist = sm.nlist;
  bool posted = false;  // the block's spiking ids went to the heavy list (block-uniform)
  if (nlist) {
    // push: the synapses of the block's spiking ids, each adding its weight
    // into the slot of step t + k + 1 (scatter_kernel.cl:194-355 and
    // gather_kernel.cl:67-200, folded into the spiking step).  Entry e of
    // the concatenated ranges goes to thread e % BLK, batch e / BLK.
    const uint32_t first_slot = slot + 1 == n.R ? 0u : slot + 1;
    if (nlist <= kSparseIds) {
      // few ids: the ranges come from the list, located by a short linear scan
      uint32_t st[kSparseIds], len[kSparseIds];
      uint32_t total = 0;
#pragma unroll
      for (uint32_t q = 0; q < kSparseIds; ++q) {
        const uint2 v = q < nlist ? sm.list[q] : make_uint2(0u, 0u);
        st[q] = v.x;
        len[q] = v.y;
        total += v.y;
      }
      if (heavy_block(n, total, AGENT)) {
        if (threadIdx.x < 64) post_heavy(n, t, threadIdx.x < nlist, sm.list[threadIdx.x < nlist ? threadIdx.x : 0]);
        total = 0;
        posted = true;
      }
      for (uint32_t e0 = 0; e0 < total; e0 += BLK * kPushBatch) {
        uint2 sw[kPushBatch];
#pragma unroll
        for (uint32_t u = 0; u < kPushBatch; ++u) {
          const uint32_t e = e0 + u * BLK + threadIdx.x;
          sw[u] = make_uint2(0u, 0u);  // weight 0: no synapse (stored weights are nonzero)
          if (e < total) {
            uint32_t src = 0, acc = 0;
#pragma unroll
            for (uint32_t q = 0; q < kSparseIds; ++q) {
              if (e >= acc && e < acc + len[q]) src = st[q] + (e - acc);
              acc += len[q];
            }
            sw[u] = n.syn[src];
          }
        }
        push_entries<kPushBatch>(n, sw, first_slot, sm.ev);
      }
    } else {
      // many ids: block-wide scan of the lanes' counts, binary search per entry
      sm.start[threadIdx.x] = r0;
      uint32_t total = block_exclusive_scan<BLK>(has ? r1 - r0 : 0u, sm.pref, sm.wave);
      if (heavy_block(n, total, AGENT)) {
        post_heavy(n, t, has, make_uint2(r0, r1 - r0));
        total = 0;
        posted = true;
      }
      for (uint32_t e0 = 0; e0 < total; e0 += BLK * kPushBatch) {
        uint2 sw[kPushBatch];
#pragma unroll
        for (uint32_t u = 0; u < kPushBatch; ++u) {
          const uint32_t e = e0 + u * BLK + threadIdx.x;
          sw[u] = make_uint2(0u, 0u);
          if (e < total) {
            uint32_t lo = 0, hi = BLK;
            while (hi - lo > 1) {
              const uint32_t mid = (lo + hi) >> 1;
              if (sm.pref[mid] <= e) lo = mid; else hi = mid;
            }
            sw[u] = n.syn[sm.start[lo] + (e - sm.pref[lo])];
          }
        }
        push_entries<kPushBatch>(n, sw, first_slot, sm.ev);
      }
    }
    __syncthreads();
  }
  // the spike words of a posting block for late deferral (zero for a block
  // that pushed or had no spikes with synapses)
  if (!AGENT && n.hvy && n.heavy_min && lane == 0 && word < n.W) n.hvy[(size_t)slot * n.W + word] = posted ? mask : 0ull;
  // rows and events per delay: one atomic per (block, delay) into the
  // delivery ring, both counts in one word
  if (threadIdx.x < n.D) {
    const uint32_t k = threadIdx.x;
    uint32_t ds = slot + 1 + k;
    if (ds >= n.R) ds -= n.R;
    const uint32_t c = sm.cnt[k], e = sm.ev[k];
    if (c | e) atomicAdd(ring_counter(n.ring_cnt, ds, blockIdx.x), ring_pack(c, e));
  }
  const unsigned long long spikes = sm.spikes, drive = sm.drive;
  if (threadIdx.x == 0) {
    // the deferral stamp, by one block (every block takes the same decision)
    if (defer && folder) {
      st_agent(&n.dstamp[slot], t + 1);
      atomicAdd(counter(n.counters, kCntDeferred, 0), 1ull);
    }
    // the heavy list step t + 1 posts into was drained by this step's gather
    if (!AGENT && n.heavy_min && blockIdx.x == 0) st_agent(&n.heavy_cnt[(t + 1) & 1], 0u);
    // CARRY: the caller sums the block's spikes and drive events over the
    // launch and adds them (and the last step's spike slot) at its end
    if (!CARRY) {
      if (spikes) {
        atomicAdd(counter(n.counters, kCntSpikes, blockIdx.x), spikes);
        atomicAdd(spk_counter(n.spk_ring, t, blockIdx.x), spikes);
      }
      if (drive) atomicAdd(counter(n.counters, kCntDrive, blockIdx.x), drive);
    }
  }
  if (!CARRY && folder && threadIdx.x < 64) fold_finish(n, t, slot, fold);
  // the new state, stored last (nothing in the step waits for these stores)
  if (!CARRY && id < n.T) store_lane<MODEL>(n, id, ls);
  out_drive = drive;
  return spikes;
}

template <int MODEL, int BLK, bool AGENT>
__device__ inline void update_block(const DevNet& n, uint64_t t, uint32_t id, UpdateSmem<BLK>& sm, bool first) {
  LaneState<MODEL> ls;
  unsigned long long drive = 0;
  update_step<MODEL, BLK, false, AGENT>(n, t, id, sm, ls, StepFacts{0, 0, 0, false}, drive, first);
}

// update(t): block b serves the NPL id blocks b * NPL .. b * NPL + NPL - 1 of
// BLK ids each, one after the other (NPL ids per lane: the reference's
// NEURONS_PER_WI tiling knob, locales.h:38-40).
template <int MODEL, int BLK, int NPL>
__global__ void __launch_bounds__(BLK) update_kernel(DevNet n, uint32_t j) {
  __shared__ UpdateSmem<BLK> sm;
  const uint64_t t = *n.t_base + j;
#pragma unroll 1
  for (uint32_t p = 0; p < (uint32_t)NPL; ++p) {
    const uint32_t base = (blockIdx.x * NPL + p) * BLK;
    if (p > 0 && base >= n.T) break;  // block-uniform
    if (p > 0) __syncthreads();       // LDS reuse
    update_block<MODEL, BLK, false>(n, t, base + threadIdx.x, sm, p == 0);
  }
}

// The LDS byte address of a pointer into LDS, and a 32-bit LDS load from one.
typedef __attribute__((address_space(3))) const uint32_t lds_u32;
__device__ inline uint32_t lds_address(const uint32_t* p) {
  return (uint32_t)(uintptr_t)(lds_u32*)p;
}
__device__ inline uint32_t lds_load_u32(uint32_t addr) { return *(lds_u32*)(uintptr_t)addr; }

// Component c (compile-time after unrolling) of a 16-byte load.
__device__ inline uint32_t comp(const uint4& v, uint32_t c) {
  return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
}

// Copies nw words of a step's spike bitmap into LDS for a pull pass, every
// load issued before the first store: one round trip per pass (a loop that
// stored each word before loading the next would wait once per word, and
// each wait also drains the entry loads already in flight).
constexpr uint32_t kStageU = 8;  // words per thread and round: 2048 words per round at 256 threads

template <int BLK, bool AGENT>
__device__ inline void stage_words(uint64_t* dst, const uint64_t* src, uint32_t nw, uint32_t from = 0) {
  constexpr uint32_t kU = kStageU;
  for (uint32_t w0 = from + threadIdx.x; w0 < nw; w0 += BLK * kU) {
    uint64_t v[kU];
#pragma unroll
    for (uint32_t u = 0; u < kU; ++u) v[u] = ld_step<AGENT>(&src[min(w0 + u * BLK, nw - 1u)]);  // no branch per load
#pragma unroll
    for (uint32_t u = 0; u < kU; ++u) {
      const uint32_t w = w0 + u * BLK;
      if (w < nw) dst[w] = v[u];
    }
  }
}

// stage_words split at its first round: the loads into registers, issued
// while the previous pass is still read, and the stores (with any further
// rounds) once the block is done with the previous bitmap.  Entry loads
// issued after load() stay in flight while the wave waits for its words.
// The thread index is taken through an empty asm statement, so that the
// word offsets are formed where they are used instead of being hoisted out
// of the pass loop into registers held across the columns.
template <int BLK, bool AGENT>
struct StagedWords {
  uint64_t v[kStageU];
  __device__ static inline uint32_t tid() {
    uint32_t x = threadIdx.x;
    asm volatile("" : "+v"(x));
    return x;
  }
  // nw >= 1 (a block's source range holds at least one word); no branch, so
  // the compiler's wait counts stay exact across the loads
  __device__ inline void load(const uint64_t* src, uint32_t nw) {
    const uint32_t i = tid();
#pragma unroll
    for (uint32_t u = 0; u < kStageU; ++u) v[u] = ld_step<AGENT>(&src[min(i + u * BLK, nw - 1u)]);
  }
  __device__ inline void store(uint64_t* dst, const uint64_t* src, uint32_t nw) {
    const uint32_t i = tid();
#pragma unroll
    for (uint32_t u = 0; u < kStageU; ++u) {
      const uint32_t w = i + u * BLK;
      if (w < nw) dst[w] = v[u];
    }
    if (nw > BLK * kStageU) stage_words<BLK, AGENT>(dst, src, nw, BLK * kStageU);
  }
};

// The vectors a slice lane reads of its column of nv: its own, or all nv.
__device__ inline uint32_t lane_vectors(uint32_t lane_word, uint32_t nv) {
  const uint32_t m = lane_word >> kPullShift;
  return m == kPullLaneAll ? nv : m;
}

// A wave's column in a pull work item: delay k (from the group's first) and
// slice s, i.e. column q = k * slices + s of the item's stream; a wave walks
// q = wave, wave + waves, ... (stepped without a division).
struct PullCol {
  uint32_t k, s;
};
__device__ inline void col_step(PullCol& c, uint32_t by, uint32_t slices) {
  c.s += by;
  while (c.s >= slices) {
    c.s -= slices;
    ++c.k;
  }
}

// Dense gather (pull) for target block `b` and delay group `g` over the
// sliced layout.  The group's due delays are taken in passes of up to
// `stage` delays: a pass stages the spike bitmaps of steps t-k-1 (the
// block's source range) of all its due delays k in LDS at once, behind one
// block barrier, and the waves then walk their columns of every delay of the
// pass with no barrier in between.  A column is one slice of 64 targets of
// one delay, one target per lane; the item's columns are numbered q = (k -
// k0) * pull_slices + slice and wave w walks q = w, w + waves, ... as one
// stream of 16-byte loads in batches of KB vectors per lane (the wave's j-th
// load is contiguous), held in two buffers that take turns, so one batch is
// in flight while the other is summed; the batch after a column's last is
// the next column's first, loaded before the next pass is staged.  Per
// entry one 32-bit LDS read and one bit-field extract give the source's
// spike bit.  Sums are exact: FP64 sums of integers below 2^53 (wide: E +=
// b * max(w, 0), I += b * min(w, 0); table: the (E, I) pair at 2 * index +
// b, zero for b = 0).  (The count format has its own body, pull_count_body,
// with the same work items, passes and column walk.)  The lane adds its
// column's sums (integers) into the block's LDS sums of its target with an
// LDS atomic: a target sits in different slices for different delays, so
// the columns of one pass that serve it may belong to different waves; the
// block stores its targets' sums for the group into pull slot g: plain
// stores, no global atomics.
// dyn: `stage` bitmaps of pull_bitmap_u64(pull_words) u64 each (the u64 at
// index pull_words of each is zero, for the padding entries) | 2*kPullTB
// int64 sums | narrow formats: 2 * pull_wn (E, I) f64 pairs.
template <int BLK, int FMT, bool AGENT, uint32_t KB, bool MULTI>
__device__ inline void pull_body(const DevNet& n, uint64_t t, uint32_t b, uint32_t g, uint64_t due, uint64_t late,
                                 uint32_t stage, uint64_t* dyn) {
  static_assert(FMT == kPullWide || FMT == kPullTable, "the count format has pull_count_body");
  if (!MULTI) stage = 1;
  constexpr bool kNarrow = FMT != kPullWide;
  constexpr uint32_t kV = FMT == kPullWide ? 2 : 4;  // entries per 16-byte vector
  constexpr uint32_t kB = KB;               // vectors per lane and batch
  constexpr uint32_t kWaves = BLK / 64;
  const uint32_t bstride = pull_bitmap_u64(n.pull_words);  // u64 per staged bitmap
  const uint32_t* bits0 = reinterpret_cast<const uint32_t*>(dyn);
  unsigned long long* sums = reinterpret_cast<unsigned long long*>(dyn + (size_t)stage * bstride);  // [2][kPullTB]
  double2* wt = reinterpret_cast<double2*>(sums + 2 * kPullTB);      // [2 * pull_wn]
  for (uint32_t i = threadIdx.x; i < 2 * kPullTB; i += BLK) sums[i] = 0ull;
  if (threadIdx.x < stage) dyn[(size_t)threadIdx.x * bstride + n.pull_words] = 0ull;  // the padding's zero word
  if (kNarrow)
    for (uint32_t i = threadIdx.x; i < 2 * n.pull_wn; i += BLK) {
      const double w = (i & 1u) ? (double)n.pull_wtab[i >> 1] : 0.0;
      wt[i] = make_double2(w > 0.0 ? w : 0.0, w < 0.0 ? w : 0.0);
    }
  const uint32_t lo = ld_const(&n.pull_lo[b]);
  const uint32_t nw = min(n.pull_words, n.W - lo);
  const uint32_t k0 = (uint32_t)(((uint64_t)g * n.D) / n.pull_groups);
  const uint32_t k1 = (uint32_t)(((uint64_t)(g + 1) * n.D) / n.pull_groups);
  const uint32_t S = n.pull_slices, TB = n.pull_tb, nk = k1 - k0;
  const uint4* vec = FMT == kPullWide ? reinterpret_cast<const uint4*>(n.pull_syn)
                                      : reinterpret_cast<const uint4*>(n.pull_syn32);
  const uint32_t lane = __lane_id();
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // uniform: scalar loads
  uint32_t events = 0;
  // col_from moves column c to the wave's next nonempty column of a due delay
  // at or after it (c.k >= nk: none) and gives its first entry s0 and its
  // vectors per lane nv (the slice's longest lane).
  auto col_from = [&](PullCol& c, uint32_t& s0, uint32_t& nv) {
    for (; c.k < nk; col_step(c, kWaves, S)) {
      const uint32_t k = k0 + c.k;
      if (!((due >> k) & 1ull)) continue;
      const uint32_t* so = n.pull_off + (size_t)(b * n.D + k) * S + c.s;
      s0 = ld_const(&so[0]);
      const uint32_t s1 = ld_const(&so[1]);
      if (s1 != s0) {
        nv = (s1 - s0) / (64 * kV);
        return;
      }
    }
    nv = 0;
  };
  // The lane's pull_lane word of column c, loaded unconditionally: past the
  // end of the stream it reads a valid word and ignores it.
  auto lane_word = [&](PullCol c) {
    if (c.k >= nk) c = PullCol{nk - 1, S - 1};
    return (uint32_t)n.pull_lane[(size_t)(b * n.D + k0 + c.k) * TB + c.s * 64 + lane];
  };
  // Entry loads run a batch ahead of the sums in two buffers, A and B, that
  // take turns (no register copies, so the compiler waits for the batch
  // being summed alone).  Every load is unconditional and every loaded
  // vector is summed, with no branch per vector: a lane's loads past its own
  // vectors read the null vector (padding entries, which add nothing) after
  // the layout's last slice, whose line every lane shares, so the padding
  // of a column is never fetched.  A column is walked in an even number of
  // batches, and the batch after a column's last is the next column's first
  // (its slice's vectors, before its lane words are known; the slice's
  // padding vectors are stored, and beyond the slice the null vector),
  // loaded into A before the next pass is staged; with one delay per pass,
  // when that column is in the next pass, after the loads of its bitmap.
  const uint4* null_vec = vec + n.pull_null;
  PullCol cq{0, 0};
  col_step(cq, wave, S);
  uint32_t cs0 = 0, cnv = 0;
  col_from(cq, cs0, cnv);
  uint32_t cln = lane_word(cq);
  uint4 qa[kB], qb[kB];
  // batch A of the column cq: the lane's own first vectors (these loads
  // follow the lane word's)
  auto load_first = [&]() {
    const uint4* p0 = vec + cs0 / kV + lane;
    const uint32_t m0 = cnv ? lane_vectors(cln, cnv) : 0u;
#pragma unroll
    for (uint32_t u = 0; u < kB; ++u) qa[u] = *(u < m0 ? p0 + 64 * u : null_vec);
  };
  load_first();
  // the spike bitmap of delay k's step and the next due delay from k
  auto ring_row = [&](uint32_t k) {
    const uint64_t* ring = ((late >> k) & 1ull) ? n.hvy : n.hist;  // a late-deferred step: its posted spikes
    return ring + ((t - k - 1) % n.R) * n.W + lo;
  };
  auto next_due = [&](uint32_t k) {
    while (k < k1 && !((due >> k) & 1ull)) ++k;
    return k;
  };
  // One delay per pass (not MULTI): the first pass's bitmap is staged here
  // and each later one at the end of the pass before it, its loads issued
  // ahead of the next column's first batch.
  uint32_t kk = next_due(k0);
  if (!MULTI && kk < k1) {
    __syncthreads();
    stage_words<BLK, AGENT>(dyn, ring_row(kk), nw);
    __syncthreads();
  }
  while (kk < k1) {  // block-uniform: every wave stages every pass
    const uint32_t kend = MULTI ? min(kk + stage, k1) : kk + 1;
    if (MULTI) {
      __syncthreads();  // the previous pass's readers are done with the bitmaps
      for (uint32_t k = kk; k < kend; ++k)
        if ((due >> k) & 1ull) stage_words<BLK, AGENT>(dyn + (size_t)(k - kk) * bstride, ring_row(k), nw);
      __syncthreads();
    }
    const uint32_t kn = next_due(kend);  // the next pass's first delay
    const bool ahead = !MULTI && kn < k1;
    while (cq.k < nk && k0 + cq.k < kend) {  // the wave's columns of this pass
      PullCol nq = cq;
      col_step(nq, kWaves, S);
      uint32_t ns0 = 0, nnv = 0;
      col_from(nq, ns0, nnv);
      const uint32_t nln = lane_word(nq);
      // the next column's first batch follows the column's last, unless the
      // next column is in the next pass, whose bitmap loads go first
      const uint32_t nnv_here = ahead && k0 + nq.k >= kend ? 0u : nnv;
      // one delay per pass: the bitmap at a fixed LDS address (folded into
      // the reads' offsets)
      const uint32_t* bits = MULTI ? bits0 + (size_t)(k0 + cq.k - kk) * bstride * 2 : bits0;
      const uint32_t tl = cln & kPullTMask, mine = lane_vectors(cln, cnv);  // target, the lane's own vectors
      const uint4* p = vec + cs0 / kV + lane;    // the lane's vector j at p[64 * j]
      const uint4* pn = vec + ns0 / kV + lane;   // the next column's
      const uint32_t nv = cnv;
      double ae = 0.0, ai = 0.0;
      // sums of the batch of vectors j0 .. j0 + kB - 1 held in qq
      auto sum_batch = [&](const uint4 (&qq)[kB], uint32_t j0) {
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u) {
          if (j0 + u >= nv) break;  // uniform: past the slice's longest lane
#pragma unroll
          for (uint32_t c = 0; c < kV; ++c) {
            const uint32_t e = comp(qq[u], FMT == kPullWide ? 2 * c : c);
            const uint32_t word = bits[(e >> 5) & (kPullRelMask >> 5)];
            const uint32_t bit = __builtin_amdgcn_ubfe(word, e, 1u);  // bit e % 32
            if (FMT == kPullWide) {
              const int32_t w = (int32_t)comp(qq[u], 2 * c + 1);
              const double bw = (double)(bit ? w : 0);
              ae += bw > 0.0 ? bw : 0.0;
              ai += bw < 0.0 ? bw : 0.0;
              events += bit;
            } else {
              const double2 wv = wt[((e >> kPullRelBits) << 1) | bit];
              ae += wv.x;
              ai += wv.y;
              events += bit;
            }
          }
        }
      };
      // a column has at least one vector: the loop body runs, so every path
      // issues its loads after the lane words (the compiler's waits rely on it)
      uint32_t j0 = 0;
      do {
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u) qb[u] = *(j0 + kB + u < mine ? p + 64 * (j0 + kB + u) : null_vec);
        sum_batch(qa, j0);
        const bool here = j0 + 2 * kB < nv;
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u) {
          const uint32_t j = j0 + 2 * kB + u;
          const uint4* a = here ? (j < mine ? p + 64 * j : null_vec) : (u < nnv_here ? pn + 64 * u : null_vec);
          qa[u] = *a;
        }
        sum_batch(qb, j0 + kB);
        j0 += 2 * kB;
      } while (j0 < nv);
      // exact integers: the order of the additions is free
      if (ae != 0.0) atomicAdd(&sums[tl], (unsigned long long)(long long)ae);
      if (ai != 0.0) atomicAdd(&sums[kPullTB + tl], (unsigned long long)(long long)ai);
      cq = nq;
      cs0 = ns0;
      cnv = nnv;
      cln = nln;
    }
    if (ahead) {
      StagedWords<BLK, AGENT> sw;
      sw.load(ring_row(kn), nw);
      load_first();  // the wave's next column, in a later pass (none: the null vector)
      __syncthreads();  // the block is done with this pass's bitmap
      sw.store(dyn, ring_row(kn), nw);
      __syncthreads();
    }
    kk = kn;
  }
  __syncthreads();
  long long* out = n.acc_pull + (size_t)g * 2 * n.npad;
  for (uint32_t i = threadIdx.x; i < TB; i += BLK) {
    const uint32_t tgt = b * TB + i;
    if (tgt < n.N) {
      st_agent(&out[tgt], (long long)sums[i]);
      st_agent(&out[n.npad + tgt], (long long)sums[kPullTB + i]);  // two's complement: negative sums
    }
  }
  // event count: one atomic per wave
  unsigned long long ev = events;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) ev += __shfl_xor(ev, off);
  if (lane == 0 && ev) atomicAdd(counter(n.counters, kCntEvents, blockIdx.x * 16 + wave), ev);
}

typedef uint32_t vec4u __attribute__((ext_vector_type(4)));

// Buffer-resource word 3 of a raw (untyped, unswizzled) buffer on gfx9 parts.
constexpr int kRawBufferFlags = 0x00020000;

// The field shift of the count entry in the high half of x: the entry's tag
// (bits 16-17 of x) times 8, in the low five bits of the result (one SDWA
// shift of the selected half word; a shift instruction reads only those five
// bits of its amount).
__device__ inline uint32_t high_field_shift(uint32_t x) {
  uint32_t r;
  asm("v_lshlrev_b32_sdwa %0, 3, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1"
      : "=v"(r)
      : "v"(x));
  return r;
}

// The pull of work item (b, g) in the count format (2-byte entries, at most
// kPullCountFields distinct weights): the passes and the column walk of
// pull_body, with the loads, the decode and the sums shaped for the format.
//  * Entries are read by buffer loads with 32-bit byte offsets into one
//    buffer resource; the wave's j-th vector of a column takes j * 1024 as
//    the scalar offset.  A lane past its own vectors loads at
//    kPullOutOfRange and gets a zero vector (skips by 0, no memory access).
//  * Per 16-byte vector the eight running word addresses are formed first
//    and the eight LDS reads issued together.  Per entry: the delta (and,
//    add), the bit's offset (one shift; the low entry of a 32-bit component
//    at bits 11-15, the high one at 27-31), the tag's field shift (one
//    shift), and one bit-field extract added, shifted, into the count
//    register: kPullCountBits-bit fields, tag i's at shift 8 * i, tag 0
//    (skips) in the lowest, which is never read.
//  * The fields are added into per-(tag, target) 32-bit LDS counts at the end
//    of a column, and every kFlushVectors vectors before a field can fill;
//    the weights multiply the counts once per target, in 64-bit integers,
//    when the item stores its sums (exact, so the order is free).
// dyn: `stage` bitmaps of pull_bitmap_u64(pull_words) u64 each |
// kPullCountFields * kPullTB u32 counts.
template <int BLK, bool AGENT, uint32_t KB, bool MULTI>
__device__ inline void pull_count_body(const DevNet& n, uint64_t t, uint32_t b, uint32_t g, uint64_t due,
                                       uint64_t late, uint32_t stage, uint64_t* dyn) {
  if (!MULTI) stage = 1;
  constexpr uint32_t kV = 8;   // entries per 16-byte vector
  constexpr uint32_t kB = KB;  // vectors per lane and batch
  constexpr uint32_t kWaves = BLK / 64;
  constexpr uint32_t kFlushVectors = ((1u << kPullCountBits) - 1) / kV / kB * kB;
  constexpr uint32_t kDeltaMask = kPullMaxDelta << 2;  // the byte delta of the running word
  static_assert(kFlushVectors >= kB, "count fields too narrow for one batch");
  const uint32_t bstride = pull_bitmap_u64(n.pull_words);
  const uint32_t* bits0 = reinterpret_cast<const uint32_t*>(dyn);
  uint32_t* cnts = reinterpret_cast<uint32_t*>(dyn + (size_t)stage * bstride);  // [kPullCountFields][kPullTB]
  for (uint32_t i = threadIdx.x; i < kPullCountFields * kPullTB; i += BLK) cnts[i] = 0u;
  const uint32_t lo = ld_const(&n.pull_lo[b]);
  const uint32_t nw = min(n.pull_words, n.W - lo);
  const uint32_t k0 = (uint32_t)(((uint64_t)g * n.D) / n.pull_groups);
  const uint32_t k1 = (uint32_t)(((uint64_t)(g + 1) * n.D) / n.pull_groups);
  const uint32_t S = n.pull_slices, TB = n.pull_tb, nk = k1 - k0;
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint16_t*>(n.pull_syn16), (short)0, (int)((n.pull_null + 1u) * 16u), kRawBufferFlags);
  const uint32_t lane = __lane_id();
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // uniform: scalar loads
  // col_from moves column c to the wave's next nonempty column of a due delay
  // at or after it (c.k >= nk: none) and gives its first entry s0 and its
  // vectors per lane nv (the slice's longest lane).
  auto col_from = [&](PullCol& c, uint32_t& s0, uint32_t& nv) {
    for (; c.k < nk; col_step(c, kWaves, S)) {
      const uint32_t k = k0 + c.k;
      if (!((due >> k) & 1ull)) continue;
      const uint32_t* so = n.pull_off + (size_t)(b * n.D + k) * S + c.s;
      s0 = ld_const(&so[0]);
      const uint32_t s1 = ld_const(&so[1]);
      if (s1 != s0) {
        nv = (s1 - s0) / (64 * kV);
        return;
      }
    }
    nv = 0;
  };
  // The lane's pull_lane word of column c, loaded unconditionally: past the
  // end of the stream it reads a valid word and ignores it.
  auto lane_word = [&](PullCol c) {
    if (c.k >= nk) c = PullCol{nk - 1, S - 1};
    return (uint32_t)n.pull_lane[(size_t)(b * n.D + k0 + c.k) * TB + c.s * 64 + lane];
  };
  // byte offset of the lane's vector 0 in the slice that starts at entry s0
  // (slices start at whole vectors of all 64 lanes)
  auto voff_of = [&](uint32_t s0) { return s0 * 2u + lane * 16u; };
  auto vload = [&](uint32_t voff, uint32_t j) {
    return __builtin_amdgcn_raw_buffer_load_b128(rsrc, voff, j * 1024u, 0);
  };
  // Entry loads run a batch ahead of the sums in two buffers, A and B, that
  // take turns, as in pull_body; the batch after a column's last is the next
  // column's first (its slice's vectors, before its lane words are known;
  // a slice's stored padding is zero skips).
  PullCol cq{0, 0};
  col_step(cq, wave, S);
  uint32_t cs0 = 0, cnv = 0;
  col_from(cq, cs0, cnv);
  uint32_t cln = lane_word(cq);
  vec4u qa[kB], qb[kB];
  // batch A of the column cq: the lane's own first vectors
  auto load_first = [&]() {
    const uint32_t v0 = voff_of(cs0);
    const uint32_t m0 = cnv ? lane_vectors(cln, cnv) : 0u;
#pragma unroll
    for (uint32_t u = 0; u < kB; ++u) qa[u] = vload(u < m0 ? v0 : kPullOutOfRange, u);
  };
  load_first();
  // the spike bitmap of delay k's step and the next due delay from k
  auto ring_row = [&](uint32_t k) {
    const uint64_t* ring = ((late >> k) & 1ull) ? n.hvy : n.hist;  // a late-deferred step: its posted spikes
    return ring + ((t - k - 1) % n.R) * n.W + lo;
  };
  auto next_due = [&](uint32_t k) {
    while (k < k1 && !((due >> k) & 1ull)) ++k;
    return k;
  };
  // One delay per pass (not MULTI): the first pass's bitmap is staged here
  // and each later one at the end of the pass before it, its loads issued
  // ahead of the next column's first batch.
  uint32_t kk = next_due(k0);
  if (!MULTI && kk < k1) {
    __syncthreads();
    stage_words<BLK, AGENT>(dyn, ring_row(kk), nw);
    __syncthreads();
  }
  while (kk < k1) {  // block-uniform: every wave stages every pass
    const uint32_t kend = MULTI ? min(kk + stage, k1) : kk + 1;
    if (MULTI) {
      __syncthreads();  // the previous pass's readers are done with the bitmaps
      for (uint32_t k = kk; k < kend; ++k)
        if ((due >> k) & 1ull) stage_words<BLK, AGENT>(dyn + (size_t)(k - kk) * bstride, ring_row(k), nw);
      __syncthreads();
    }
    const uint32_t kn = next_due(kend);  // the next pass's first delay
    const bool ahead = !MULTI && kn < k1;
    while (cq.k < nk && k0 + cq.k < kend) {  // the wave's columns of this pass
      const uint32_t* bits = MULTI ? bits0 + (size_t)(k0 + cq.k - kk) * bstride * 2 : bits0;
      const uint32_t tl = cln & kPullTMask, mine = lane_vectors(cln, cnv);  // target, the lane's own vectors
      const uint32_t vcur = voff_of(cs0);
      const uint32_t nv = cnv;
      // the column's second batch goes out before the next column's facts
      // are looked up (their scalar loads and branch stall the wave)
      auto load_b = [&](uint32_t j0) {
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u) {
          const uint32_t j = j0 + kB + u;
          qb[u] = vload(j < mine ? vcur : kPullOutOfRange, j);
        }
      };
      load_b(0);
      PullCol nq = cq;
      col_step(nq, kWaves, S);
      uint32_t ns0 = 0, nnv = 0;
      col_from(nq, ns0, nnv);
      const uint32_t nln = lane_word(nq);
      const uint32_t vnext = voff_of(ns0);
      // the next column's first batch follows the column's last, unless the
      // next column is in the next pass, whose bitmap loads go first
      const uint32_t nnv_here = ahead && k0 + nq.k >= kend ? 0u : nnv;
      uint32_t cnt = 0;                  // kPullCountBits-bit fields per tag
      uint32_t ba = lds_address(bits);   // LDS byte address of the running 32-bit word
      auto sum_vec = [&](const vec4u& q) {
        uint32_t a[2 * 4];
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
          ba += q[c] & kDeltaMask;
          a[2 * c] = ba;
          ba += (q[c] >> 16) & kDeltaMask;
          a[2 * c + 1] = ba;
        }
        uint32_t w[2 * 4];
#pragma unroll
        for (uint32_t i = 0; i < 2 * 4; ++i) w[i] = lds_load_u32(a[i]);
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
          const uint32_t x = q[c];
          cnt += __builtin_amdgcn_ubfe(w[2 * c], x >> 11, 1u) << ((x << 3) & 31u);
          cnt += __builtin_amdgcn_ubfe(w[2 * c + 1], x >> 27, 1u) << (high_field_shift(x) & 31u);
        }
      };
      // sums of the batch of vectors j0 .. j0 + kB - 1 held in qq
      auto sum_batch = [&](const vec4u (&qq)[kB], uint32_t j0) {
        if (j0 >= nv) return;  // uniform: past the slice's longest lane
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u) {
          if (j0 + u >= nv) break;
          sum_vec(qq[u]);
        }
        if ((j0 / kB + 1) % (kFlushVectors / kB) == 0 || j0 + kB >= nv) {
#pragma unroll
          for (uint32_t i = 1; i <= kPullCountFields; ++i)
            if (i < n.pull_wn) atomicAdd(&cnts[(i - 1) * kPullTB + tl], __builtin_amdgcn_ubfe(cnt, 8u * i, 8u));
          cnt = 0;
        }
      };
      // a column has at least one vector; batch A of vectors j0.., batch B
      // of j0 + kB.., then A again from j0 + 2 kB or, after the column's
      // last, the next column's first
      uint32_t j0 = 0;
      for (;;) {
        sum_batch(qa, j0);
        const bool here = j0 + 2 * kB < nv;
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u) {
          // one load either way: its offsets are selected
          const uint32_t j = j0 + 2 * kB + u;
          const uint32_t vo = here ? (j < mine ? vcur : kPullOutOfRange) : (u < nnv_here ? vnext : kPullOutOfRange);
          qa[u] = vload(vo, here ? j : u);
        }
        sum_batch(qb, j0 + kB);
        j0 += 2 * kB;
        if (j0 >= nv) break;
        load_b(j0);
      }
      cq = nq;
      cs0 = ns0;
      cnv = nnv;
      cln = nln;
    }
    if (ahead) {
      StagedWords<BLK, AGENT> sw;
      sw.load(ring_row(kn), nw);
      load_first();  // the wave's next column, in a later pass (none: no vectors, out-of-range loads)
      __syncthreads();  // the block is done with this pass's bitmap
      sw.store(dyn, ring_row(kn), nw);
      __syncthreads();
    }
    kk = kn;
  }
  __syncthreads();
  // sums = counts times weights, per target
  long long* out = n.acc_pull + (size_t)g * 2 * n.npad;
  int32_t wv[kPullCountFields];
#pragma unroll
  for (uint32_t i = 0; i < kPullCountFields; ++i) wv[i] = i + 1 < n.pull_wn ? ld_const(&n.pull_wtab[i + 1]) : 0;
  uint32_t events = 0;
  for (uint32_t i = threadIdx.x; i < TB; i += BLK) {
    const uint32_t tgt = b * TB + i;
    if (tgt < n.N) {
      long long e = 0, in = 0;
#pragma unroll
      for (uint32_t f = 0; f < kPullCountFields; ++f) {
        const uint32_t c = cnts[f * kPullTB + i];
        const long long p = (long long)c * wv[f];
        events += c;
        if (wv[f] > 0) e += p; else in += p;
      }
      st_agent(&out[tgt], e);
      st_agent(&out[n.npad + tgt], in);  // two's complement: negative sums
    }
  }
  // event count: one atomic per wave
  unsigned long long ev = events;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) ev += __shfl_xor(ev, off);
  if (lane == 0 && ev) atomicAdd(counter(n.counters, kCntEvents, blockIdx.x * 16 + wave), ev);
}

// The pull of work item (b, g) in the network's entry format.
template <int BLK, bool AGENT, uint32_t KB, bool MULTI>
__device__ inline void pull_item(const DevNet& n, uint64_t t, uint32_t b, uint32_t g, uint64_t due, uint64_t late,
                                 uint32_t stage, uint64_t* dyn) {
  if (n.pull_fmt == kPullWide)
    pull_body<BLK, kPullWide, AGENT, KB, MULTI>(n, t, b, g, due, late, stage, dyn);
  else if (n.pull_fmt == kPullCount)
    pull_count_body<BLK, AGENT, KB, MULTI>(n, t, b, g, due, late, stage, dyn);
  else
    pull_body<BLK, kPullTable, AGENT, KB, MULTI>(n, t, b, g, due, late, stage, dyn);
}

// gather(t): the pull of the deferred steps whose synapses are due at t, one
// (target block, delay group) work item per block iteration.  Exits at once
// when nothing is due (every block reads the same stamps).  Batches of KB
// vectors per lane (2: at most 72 VGPRs, 7 waves per SIMD; 4: at most 96,
// 5 waves) with one delay staged at a time, or n.pull_stage delays staged
// at once (fewer blocks fit the LDS, so up to 128 VGPRs).
template <int BLK, uint32_t KB, int WPE>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(WPE))) gather_kernel(DevNet n, uint32_t j) {
  extern __shared__ __align__(16) uint64_t dyn[];
  __shared__ uint32_t hev[64];
  const uint64_t t = *n.t_base + j;
  // the heavy pushes of step t - 1 (their delay-0 synapses are due at t), or,
  // when its posted ids reach half the deferral threshold, late deferral:
  // step t - 1 is pulled instead, from its posted spike words (every block
  // takes the same decision from the same count; block 0 stamps the step for
  // the launches after this one)
  const uint32_t hc = n.heavy_min && t > 0 ? n.heavy_cnt[(t - 1) & 1] : 0u;
  const bool late0 = hc && n.hvy && defer_late(hc, n.defer_spikes);
  if (hc && !late0) heavy_push<BLK>(n, t - 1, hc, hev);
  uint64_t late;
  uint64_t due = due_mask<false>(n, t, late);
  if (late0) {
    due |= 1ull;
    late |= 1ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st_agent(&n.dstamp[(t - 1) % n.R], t | kStampLate);
      atomicAdd(counter(n.counters, kCntDeferred, 0), 1ull);
      atomicAdd(counter(n.counters, kCntLate, 0), 1ull);
    }
  }
  if (!due) return;
  for (uint32_t bg = blockIdx.x; bg < n.pull_nb * n.pull_groups; bg += gridDim.x) {
    pull_item<BLK, false, KB, WPE == 4>(n, t, bg / n.pull_groups, bg % n.pull_groups, due, late, n.pull_stage, dyn);
    __syncthreads();  // LDS reuse by the next work item
  }
}

// Grid barrier for a co-resident (cooperative) grid, carrying a sum.
// Arrival: block b adds (1, its spikes) to leaf counter b % kBarLeaves (one
// 64-bit word: arrivals in the low kBarCountBits, the sum above); the leaf's
// last block resets it and adds (1, leaf sum) to the root; the root's last
// block resets it and writes the generation word (generation + 1, grid sum).
// Blocks wait for the generation to change and return the grid sum.
// Spreading arrivals over the leaves avoids funnelling every block through one
// contended address.  State shared across blocks is accessed agent-coherently
// (ld_agent / st_agent / device atomics), so the barrier only has to make
// every wave's accesses complete before its block arrives (s_waitcnt 0); an
// agent-scope fence here would write back and invalidate the XCD's L2 once per
// block and barrier, which costs microseconds.  The wait is bounded: at the
// limit (never reached by a co-resident grid) it records kBarTimeout in the
// overflow counter and returns false, and the caller leaves its step loop.
// sh[0] is the block's copy of the generation, sh[1] its failure flag,
// sh[2..3] the returned sum.
__device__ inline bool grid_barrier(uint32_t* bar, uint32_t* sh, unsigned long long* err, unsigned long long add,
                                    unsigned long long* sum) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t nb = gridDim.x;
    const uint32_t leaf = blockIdx.x % kBarLeaves;
    const uint32_t leaves = min(nb, kBarLeaves);
    const uint32_t leaf_size = (nb - leaf + kBarLeaves - 1) / kBarLeaves;
    unsigned long long* leaf_w = reinterpret_cast<unsigned long long*>(bar + leaf * kBarStride);
    unsigned long long* root_w = reinterpret_cast<unsigned long long*>(bar + kBarLeaves * kBarStride);
    unsigned long long* gen_w = reinterpret_cast<unsigned long long*>(bar + (kBarLeaves + 1) * kBarStride);
    const uint32_t g = sh[0];
    const unsigned long long lo = atomicAdd(leaf_w, 1ull + (add << kBarCountBits));
    if ((lo & kBarCountMask) == leaf_size - 1) {
      atomicExch(leaf_w, 0ull);
      const unsigned long long lsum = (lo >> kBarCountBits) + add;
      const unsigned long long ro = atomicAdd(root_w, 1ull + (lsum << kBarCountBits));
      if ((ro & kBarCountMask) == leaves - 1) {
        atomicExch(root_w, 0ull);
        const unsigned long long tot = (ro >> kBarCountBits) + lsum;
        atomicExch(gen_w, (tot << 32) | (unsigned long long)(g + 1u));
      }
    }
    unsigned long long v = 0;
    uint32_t spins = 0;
    while ((uint32_t)(v = ld_agent(gen_w)) == g) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > kBarSpinLimit) {
        atomicOr(err, kBarTimeout);
        sh[1] = 1;
        break;
      }
    }
    sh[0] = g + 1;
    sh[2] = (uint32_t)(v >> 32);
  }
  __syncthreads();
  *sum = sh[2];
  return sh[1] == 0;
}

// Persistent step loop: one cooperative launch runs `nsteps` steps.  A grid
// barrier ends every update(t); on steps that pull, another one separates
// the pull from update(t).  Every block takes the same barriers.  When the
// grid covers every id with one lane, each lane keeps its id's parameters and
// state in registers for the whole launch and stores the state once at the
// end, and the step facts are carried in registers: the due mask as a shift
// register of deferral bits (step s deferred: the same decision
// in every block, from the barrier's spike total), so a step starts with no
// loads but its accumulators.  Then the delivery-time counters are folded by
// the first 2 * kRingStripes blocks (one stripe each: exchanged with 0 at the
// end of the step, added to the counters after the next barrier) instead of
// by block 0 alone.
template <int MODEL, int BLK>
__global__ void __launch_bounds__(BLK) persistent_kernel(DevNet n, uint64_t nsteps, uint64_t* t_base) {
  __shared__ UpdateSmem<BLK> usm;
  __shared__ uint32_t bar_sh[4];
  __shared__ StepFacts f0;
  unsigned long long* err = counter(n.counters, kCntOverflow, 0);
  const uint64_t t0 = *t_base;
  const bool carry = (uint64_t)gridDim.x * BLK >= n.T;
  if (threadIdx.x < 64) {
    const uint32_t lane = __lane_id();
    if (lane == 0) {
      bar_sh[0] = (uint32_t)ld_agent(reinterpret_cast<unsigned long long*>(n.barrier + (kBarLeaves + 1) * kBarStride));
      bar_sh[1] = 0;
      f0.spk_prev = 0;
      f0.due = 0;
      f0.late = 0;
    }
    if (carry) {
      // the facts of step t0, as step_facts reads them
      const uint32_t k = lane < n.D ? lane : 0u;
      const uint64_t stamp = n.pull_off ? ld_agent(&n.dstamp[(t0 - k - 1) % n.R]) : 0ull;
      const unsigned long long spk = ld_agent(spk_counter(n.spk_ring, t0 - 1, lane));
      const bool ok = n.pull_off && lane < n.D && t0 >= (uint64_t)lane + 1 && stamp_step(stamp) == t0 - lane;
      const uint64_t due = __ballot(ok);
      const uint64_t late = __ballot(ok && (stamp & kStampLate));
      if (lane == 0) {
        f0.due = due;
        f0.late = late;
      }
      if (t0 > 0 && spk) atomicAdd(&f0.spk_prev, spk);
    }
  }
  extern __shared__ __align__(16) uint64_t dyn[];
  const uint32_t id = blockIdx.x * BLK + threadIdx.x;
  LaneState<MODEL> ls;
  if (carry) {
    load_lane<MODEL>(n, id, ls);
    if (threadIdx.x < n.drive_k) usm.cdf[threadIdx.x] = n.drive_cdf[threadIdx.x];
  }
  __syncthreads();
  StepFacts f = f0;
  // the first step's decision; after it a persistent launch forces only its
  // volley, whose ids spike every step, so the prediction is the previous
  // step's spikes
  f.defer = defer_step(n, t0, f.spk_prev);
  const uint64_t dmask = n.D >= 64 ? ~0ull : (1ull << n.D) - 1ull;
  // counter fold of the carry path: lane L of wave 0 owns stripe q
  const uint32_t q = blockIdx.x + __lane_id() * gridDim.x;
  const bool folds = carry && threadIdx.x < 64 && q < kRingStripes;
  unsigned long long fold_v = 0;
  unsigned long long spikes_acc = 0, drive_acc = 0;
  for (uint64_t s = 0; s < nsteps; ++s) {
    const uint64_t t = t0 + s;
    if (fold_v) {
      fold_counters(n, q, fold_v);
      fold_v = 0;
    }
    unsigned long long sum = 0;
    uint64_t late = f.late;
    const uint64_t due = carry ? f.due : due_mask<true>(n, t, late);
    if (due) {
      for (uint32_t bg = blockIdx.x; bg < n.pull_nb * n.pull_groups; bg += gridDim.x) {
        pull_item<BLK, true, 2, false>(n, t, bg / n.pull_groups, bg % n.pull_groups, due, late, 1u, dyn);
        __syncthreads();
      }
      if (!grid_barrier(n.barrier, bar_sh, err, 0, &sum)) return;
    }
    if (carry) {
      unsigned long long drive = 0;
      const unsigned long long spikes = update_step<MODEL, BLK, true, true>(n, t, id, usm, ls, f, drive);
      spikes_acc += spikes;
      drive_acc += drive;
      const uint32_t slot = (uint32_t)(t % n.R);
      if (folds) fold_v = atomicExch(ring_counter(n.ring_cnt, slot, q), 0ull);
      if (!grid_barrier(n.barrier, bar_sh, err, threadIdx.x == 0 ? spikes : 0ull, &sum)) return;
      f.due = ((f.due << 1) | (f.defer ? 1ull : 0ull)) & dmask;
      f.late = (f.late << 1) & dmask;  // the loop's own deferrals stage the history
      f.spk_prev = sum;
      f.defer = n.pull_off && sum >= n.defer_spikes;
    } else {
      for (uint32_t base = blockIdx.x * BLK; base < n.T; base += gridDim.x * BLK) {
        if (base != blockIdx.x * BLK) __syncthreads();  // LDS reuse
        update_block<MODEL, BLK, true>(n, t, base + threadIdx.x, usm, base == blockIdx.x * BLK);
      }
      if (!grid_barrier(n.barrier, bar_sh, err, 0, &sum)) return;
    }
  }
  if (fold_v) fold_counters(n, q, fold_v);
  if (carry) {
    store_lane<MODEL>(n, id, ls);
    if (threadIdx.x == 0) {
      if (spikes_acc) atomicAdd(counter(n.counters, kCntSpikes, blockIdx.x), spikes_acc);
      if (drive_acc) atomicAdd(counter(n.counters, kCntDrive, blockIdx.x), drive_acc);
    }
    // the spike-count slots a following single-step launch reads (step
    // t0 + nsteps - 1: the last total) and fills (the next step: zero)
    if (blockIdx.x == 0 && threadIdx.x < 2 * kRingStripes && nsteps) {
      const uint32_t st = threadIdx.x % kRingStripes;
      const uint64_t tl = t0 + nsteps - 1;
      st_agent(spk_counter(n.spk_ring, threadIdx.x < kRingStripes ? tl : tl + 1, st),
               threadIdx.x == 0 ? f.spk_prev : 0ull);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *t_base = t0 + nsteps;
}

// The heavy pushes posted by step t_base - 1, before a persistent launch
// (whose blocks all push their own ids); heavy_clear_kernel then empties
// both lists.
__global__ void __launch_bounds__(256) heavy_kernel(DevNet n) {
  __shared__ uint32_t hev[64];
  const uint64_t t = *n.t_base;
  const uint32_t hc = t > 0 ? n.heavy_cnt[(t - 1) & 1] : 0u;
  if (hc) heavy_push<256>(n, t - 1, hc, hev);
}

__global__ void heavy_clear_kernel(uint32_t* heavy_cnt) {
  if (threadIdx.x < 2) heavy_cnt[threadIdx.x] = 0;
}

__global__ void advance_kernel(uint64_t* t_base, uint64_t k) { *t_base += k; }

__global__ void fill_u64_kernel(uint64_t* p, uint64_t value, uint64_t count) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

}  // namespace

namespace {
// A launch of kernel(n, j); with events, through hipExtLaunchKernelGGL, which
// has the kernel's dispatch record e0 at its start and e1 at its end (the
// kernel's own duration, without the queue's gaps around it).
template <typename K>
void launch_timed(K kernel, uint32_t grid, uint32_t block, size_t dyn, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                  const DevNet& n, uint32_t j) {
  if (e0)
    hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), (uint32_t)dyn, s, e0, e1, 0u, n, j);
  else
    kernel<<<grid, block, dyn, s>>>(n, j);
}
}  // namespace

hipError_t launch_update(const DevNet& n, uint32_t j, uint32_t block, uint32_t npl, hipStream_t s, hipEvent_t e0,
                         hipEvent_t e1) {
#define SZSNN_UPD_NPL(M, B, P) launch_timed(update_kernel<M, B, P>, (n.T + B * P - 1) / (B * P), B, 0, s, e0, e1, n, j)
#define SZSNN_UPD_B(M, B)                           \
  switch (npl) {                                    \
    case 1: SZSNN_UPD_NPL(M, B, 1); break;          \
    case 2: SZSNN_UPD_NPL(M, B, 2); break;          \
    case 4: SZSNN_UPD_NPL(M, B, 4); break;          \
    default: return hipErrorInvalidValue;           \
  }
#define SZSNN_UPD(M)                                \
  switch (block) {                                  \
    case 64: SZSNN_UPD_B(M, 64) break;              \
    case 128: SZSNN_UPD_B(M, 128) break;            \
    case 256: SZSNN_UPD_B(M, 256) break;            \
    case 512: SZSNN_UPD_B(M, 512) break;            \
    case 1024: SZSNN_UPD_B(M, 1024) break;          \
    default: return hipErrorInvalidValue;           \
  }
  switch (n.model) {
    case kIfCurrExp: SZSNN_UPD(kIfCurrExp) break;
    case kIzhikevich: SZSNN_UPD(kIzhikevich) break;
    case kIfCondExp: SZSNN_UPD(kIfCondExp) break;
    default: return hipErrorInvalidValue;
  }
#undef SZSNN_UPD
#undef SZSNN_UPD_B
#undef SZSNN_UPD_NPL
  return hipGetLastError();
}

size_t pull_dyn_bytes_for(uint32_t words, uint32_t fmt, uint32_t wn, uint32_t stage) {
  const size_t bitmaps = (size_t)stage * pull_bitmap_u64(words) * 8;
  if (fmt == kPullCount) return bitmaps + (size_t)kPullCountFields * kPullTB * 4;  // per-(tag, target) counts
  return bitmaps + (size_t)2 * kPullTB * 8 + (fmt != kPullWide ? (size_t)32 * wn : 0);
}

size_t pull_dyn_bytes(const DevNet& n, uint32_t stage) {
  return n.pull_off ? pull_dyn_bytes_for(n.pull_words, n.pull_fmt, n.pull_wn, stage) : 0;
}

namespace {
// The gather kernel of one block size, staging depth and batch.
template <int BLK>
const void* gather_fn(uint32_t stage, uint32_t batch) {
  if (stage <= 1)
    return batch == 4 ? reinterpret_cast<const void*>(&gather_kernel<BLK, 4, 5>)
                      : reinterpret_cast<const void*>(&gather_kernel<BLK, 2, 7>);
  return batch == 4 ? reinterpret_cast<const void*>(&gather_kernel<BLK, 4, 4>)
                    : reinterpret_cast<const void*>(&gather_kernel<BLK, 2, 4>);
}
const void* gather_fn_for(uint32_t block, uint32_t stage, uint32_t batch) {
  return block == 128    ? gather_fn<128>(stage, batch)
         : block == 256  ? gather_fn<256>(stage, batch)
         : block == 512  ? gather_fn<512>(stage, batch)
         : block == 1024 ? gather_fn<1024>(stage, batch)
                         : nullptr;
}
}  // namespace

int gather_max_blocks_per_cu(size_t dyn, uint32_t block, uint32_t stage, uint32_t batch) {
  int nb = 0;
  const void* f = gather_fn_for(block, stage, batch);
  if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, (int)block, dyn) != hipSuccess) return 0;
  return nb;
}

hipError_t launch_heavy(const DevNet& n, uint32_t grid, hipStream_t s) {
  if (!n.heavy_min) return hipSuccess;
  heavy_kernel<<<grid ? grid : 1, 256, 0, s>>>(n);
  heavy_clear_kernel<<<1, 64, 0, s>>>(n.heavy_cnt);
  return hipGetLastError();
}

hipError_t launch_gather(const DevNet& n, uint32_t j, uint32_t block, uint32_t grid, hipStream_t s, hipEvent_t e0,
                         hipEvent_t e1) {
  if (!n.pull_off) {  // no pull layout: every step pushes (the events bracket nothing)
    hipError_t e = hipSuccess;
    if (e0) e = hipEventRecord(e0, s);
    if (e0 && e == hipSuccess) e = hipEventRecord(e1, s);
    return e;
  }
  if (grid == 0 || n.pull_stage == 0) return hipErrorInvalidValue;
  const size_t dyn = pull_dyn_bytes(n, n.pull_stage);
  const bool wide = n.pull_batch == 4;
#define SZSNN_GATHER(B)                                                              \
  if (n.pull_stage <= 1) {                                                           \
    if (wide) launch_timed(gather_kernel<B, 4, 5>, grid, B, dyn, s, e0, e1, n, j);   \
    else launch_timed(gather_kernel<B, 2, 7>, grid, B, dyn, s, e0, e1, n, j);        \
  } else {                                                                           \
    if (wide) launch_timed(gather_kernel<B, 4, 4>, grid, B, dyn, s, e0, e1, n, j);   \
    else launch_timed(gather_kernel<B, 2, 4>, grid, B, dyn, s, e0, e1, n, j);        \
  }
  switch (block) {
    case 128: SZSNN_GATHER(128) break;
    case 256: SZSNN_GATHER(256) break;
    case 512: SZSNN_GATHER(512) break;
    case 1024: SZSNN_GATHER(1024) break;
    default: return hipErrorInvalidValue;
  }
#undef SZSNN_GATHER
  return hipGetLastError();
}

namespace {
template <int MODEL, int BLK>
hipError_t coop_launch(const DevNet& n, uint64_t nsteps, uint64_t* t_base, uint32_t grid, hipStream_t s) {
  DevNet nn = n;
  uint64_t ns = nsteps;
  uint64_t* tb = t_base;
  void* args[] = {&nn, &ns, &tb};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&persistent_kernel<MODEL, BLK>), dim3(grid),
                                    dim3(BLK), args, pull_dyn_bytes(n, 1), s);
}
template <int MODEL, int BLK>
int coop_max_blocks(size_t dyn) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&persistent_kernel<MODEL, BLK>),
                                                   BLK, dyn) != hipSuccess)
    return 0;
  return nb;
}
}  // namespace

hipError_t launch_persistent(const DevNet& n, uint64_t nsteps, uint64_t* t_base, uint32_t block, uint32_t grid,
                             hipStream_t s) {
#define SZSNN_COOP(M)                                                          \
  switch (block) {                                                             \
    case 256: return coop_launch<M, 256>(n, nsteps, t_base, grid, s);          \
    case 512: return coop_launch<M, 512>(n, nsteps, t_base, grid, s);          \
    default: return hipErrorInvalidValue;                                      \
  }
  switch (n.model) {
    case kIfCurrExp: SZSNN_COOP(kIfCurrExp)
    case kIzhikevich: SZSNN_COOP(kIzhikevich)
    case kIfCondExp: SZSNN_COOP(kIfCondExp)
    default: return hipErrorInvalidValue;
  }
#undef SZSNN_COOP
}

int persistent_max_blocks_per_cu(const DevNet& n, uint32_t block) {
  const size_t dyn = pull_dyn_bytes(n, 1);
  const int model = n.model;
#define SZSNN_OCC(M)                                       \
  switch (block) {                                         \
    case 256: return coop_max_blocks<M, 256>(dyn);            \
    case 512: return coop_max_blocks<M, 512>(dyn);            \
    default: return 0;                                     \
  }
  switch (model) {
    case kIfCurrExp: SZSNN_OCC(kIfCurrExp)
    case kIzhikevich: SZSNN_OCC(kIzhikevich)
    case kIfCondExp: SZSNN_OCC(kIfCondExp)
    default: return 0;
  }
#undef SZSNN_OCC
}

hipError_t launch_advance(uint64_t* t_base, uint64_t k, hipStream_t s) {
  advance_kernel<<<1, 1, 0, s>>>(t_base, k);
  return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t* p, uint64_t value, uint64_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  fill_u64_kernel<<<(uint32_t)blocks, 256, 0, s>>>(p, value, count);
  return hipGetLastError();
}

}  // namespace szsnn
