// szsnn_oracle.hpp -- CPU restatement of the SZSNN per-step simulation core.
//
// TEST INFRASTRUCTURE ONLY.  This header is the checker for the HIP library in
// szsnn_amd/csrc; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may build, load or call it, and never as the thing measured
// or shipped.
//
// One engine, templated on the scalar type R:
//   R = double : the FP64 oracle of the discretised model;
//   R = float  : the FP32 instantiation -- the reference's FLOAT_TYPE -- with one
//                fixed evaluation order and no FMA contraction (build with
//                -ffp-contract=off).  The HIP path must match it bit for bit.
//
// What it restates (reference paths relative to the reference checkout):
//   step order gather -> update(+input model) -> scatter, cycle = completed steps
//       szsnn1.1/src/Simulation.ipp:471-483, :402-460
//   delay convention: API delay k delivered into update(c+k+1)
//       szsnn1.1/src/NetworkGraph.ipp:416,431 (k stored as k+1),
//       szsnn1.1/src/SynapseData.ipp:81 (bit k of the delay bitmap),
//       szsnn1.1/src/cl/scatter_kernel.cl:154-169,204-213,325-335,
//       szsnn1.1/src/cl/gather_kernel.cl:55-58,104
//   sign-split accumulation, zero weights inert, accumulators rebuilt each step
//       szsnn1.1/src/cl/gather_kernel.cl:109-114,172-178,188-199
//   IF_curr_exp update           szsnn1.1/src/cl/IF_curr_exp_update_kernel.cl:224-298
//   Izhikevich update            szsnn1.1/src/cl/Izhikevich_update_kernel.cl:214-289
//   Const_rate input model       szsnn1.1/src/cl/Const_rate_update_kernel.cl:205-230
//   forced (external) spikes     szsnn1.1/src/cl/IF_curr_exp_update_kernel.cl:154-169,281-283
//   dense spike output           szsnn1.1/src/SpikeBuffer.cpp:136-165
//
// Additions that have no reference counterpart (the reference has no IF_cond_exp,
// no Poisson model and no noise source):
//   IF_cond_exp update in the IF_curr_exp discretisation (conductances in the
//   current slots, inhibitory weights stored negative);
//   counter-based random drive: SplitMix64 outputs (Steele, Lea, Flood 2014),
//   Poisson sources (integer threshold), binomial per-neuron drive (inverse CDF
//   over integer thresholds), Gaussian current (Irwin-Hall sum of eight 16-bit
//   uniforms, exact integer sum, one scaling).
//
// Accumulation: the reference sums FP32 weights with LDS compare-and-swap
// atomics in a nondeterministic order (gather_kernel.cl:19-35).  The FP32
// instantiation sums exactly in Q(63-F).F fixed point and rounds once to FP32,
// the order-free value of that sum; the FP64 instantiation sums in double.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include <omp.h>

namespace szsnn_oracle {

// ---------------------------------------------------------------- random drive
constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
// idx-th output (0-based) of a SplitMix64 generator seeded with `seed`.
inline uint64_t splitmix_at(uint64_t seed, uint64_t idx) {
  return mix64(seed + kGolden * (idx + 1));
}
inline uint64_t neuron_key(uint64_t seed, uint64_t local_index) {
  return splitmix_at(seed, local_index);
}
// Stream s (0..3) of step t for one neuron: 0,1 Gaussian, 2 drive, 3 source.
inline uint64_t draw(uint64_t key, uint64_t t, unsigned s) {
  return splitmix_at(key, 4 * t + s);
}
inline int32_t gauss_int(uint64_t key, uint64_t t) {
  const uint64_t a = draw(key, t, 0), b = draw(key, t, 1);
  uint32_t sum = 0;
  for (int j = 0; j < 4; ++j) {
    sum += (uint32_t)((a >> (16 * j)) & 0xffffu);
    sum += (uint32_t)((b >> (16 * j)) & 0xffffu);
  }
  return (int32_t)(2u * sum) - 524280;  // 2*(sum - mean), even integer
}
// 0.5 / sqrt(8 * (65536^2 - 1) / 12): unit variance for the centred sum above.
template <typename R> inline R gauss_scale();
template <> inline float gauss_scale<float>() { return 0x1.3988e2p-17f; }
template <> inline double gauss_scale<double>() { return 0x1.3988e1412ed76p-17; }
template <typename R> inline R gauss(uint64_t key, uint64_t t) {
  return (R)gauss_int(key, t) * gauss_scale<R>();
}

// ---------------------------------------------------------------- models
enum Model { IF_CURR_EXP = 0, IZHIKEVICH = 1, IF_COND_EXP = 2 };
enum SourceModel { SRC_NONE = 0, SRC_CONST_RATE = 1, SRC_POISSON = 2 };

inline int n_params(int model) {
  return model == IF_CURR_EXP ? 9 : model == IZHIKEVICH ? 4 : model == IF_COND_EXP ? 11 : -1;
}
inline int n_vars(int model) {
  return model == IF_CURR_EXP ? 4 : model == IZHIKEVICH ? 2 : model == IF_COND_EXP ? 4 : -1;
}

// One neuron-step.  p: parameters, x: variables (in/out), accE/accI: this
// step's accumulated synaptic input, iinj: injected current.  Returns spike.
// Parameter order IF_curr_exp: v_rest cm tau_m tau_refrac tau_syn_E tau_syn_I
// i_offset v_reset v_thresh (IF_curr_exp_NeuronModel.hpp:12-23); variables
// v ie ii fired.  Izhikevich: a b c d / v u.  IF_cond_exp: the IF_curr_exp
// nine then e_rev_E e_rev_I / v ge gi fired.  All times in steps, cm in
// (capacitance / dt): the reference kernels carry no dt.
template <typename R>
inline bool update_neuron(int model, const R* p, R* x, R accE, R accI, R iinj, bool forced) {
  const R one = R(1);
  if (model == IZHIKEVICH) {
    // Izhikevich_update_kernel.cl:219,247-251,264-268
    R v = x[0], u = x[1];
    const R a = p[0], b = p[1];
    const R I = (accE + accI) + iinj;
    const R dv = ((((R(0.04) * v) + R(5)) * v + R(140)) - u) + I;
    const R du = a * ((b * v) - u);
    v = v + dv;
    u = u + du;
    const bool spike = (v >= R(30)) || forced;
    if (spike) {
      v = p[2];
      u = u + p[3];
    }
    x[0] = v;
    x[1] = u;
    return spike;
  }
  // IF_curr_exp_update_kernel.cl:230-297, IF_cond_exp in the same scheme.
  const R v_rest = p[0], cm = p[1], tau_m = p[2], tau_refrac = p[3];
  const R tau_e = p[4], tau_i = p[5], i_offset = p[6], v_reset = p[7], v_thresh = p[8];
  R v = x[0];
  // Division is evaluated as a product with the reciprocal, x * (1/y): the
  // reference kernels were built without correctly rounded division
  // (Simulation.ipp:13-21), and its FP32 trace (szsnn1.1/bin/myfile.txt~)
  // agrees with this form, not with the IEEE quotient.
  const R se = ((one - one / tau_e) * x[1]) + accE;
  const R si = ((one - one / tau_i) * x[2]) + accI;
  R I;
  if (model == IF_CURR_EXP) {
    I = ((se + si) + iinj) + i_offset;
  } else {
    const R e_rev_e = p[9], e_rev_i = p[10];
    I = (((se * (e_rev_e - v)) - (si * (e_rev_i - v))) + iinj) + i_offset;
  }
  R fired = x[3];
  const bool refr = fired <= tau_refrac;
  fired = fired + one;
  if (!refr) {
    const R delta = (I * (one / cm)) + ((v_rest - v) * (one / tau_m));
    v = v + delta;
  }
  const bool spike = (!refr && v >= v_thresh) || forced;
  if (spike) {
    v = v_reset;
    fired = one;
  }
  x[0] = v;
  x[1] = se;
  x[2] = si;
  x[3] = fired;
  return spike;
}

// ---------------------------------------------------------------- network
struct NetDesc {
  int32_t model;
  uint32_t n_neurons;
  uint32_t n_sources;
  int32_t source_model;
  const float* params;        // [n_params][n_neurons]
  const float* vars;          // [n_vars][n_neurons]
  const float* src_rate;      // [n_sources] Const_rate: rate in steps
  const float* src_last;      // [n_sources] Const_rate: initial `last`
  const uint64_t* src_thresh; // [n_sources] Poisson: spike iff draw < thresh
  uint64_t n_syn;
  const uint32_t* syn_src;    // global source id: neurons 0..N-1, sources N..N+M-1
  const uint32_t* syn_tgt;    // neuron id
  const float* syn_w;
  const uint16_t* syn_delay;  // API delay k: delivered into update(c+k+1)
  uint32_t max_delay;         // D: every k < D
  int32_t frac_bits;          // F of the fixed-point accumulators
  uint64_t seed;
  const float* noise_amp;     // [n_neurons] Gaussian current amplitude, or null
  uint32_t drive_k;           // binomial drive: number of CDF thresholds
  const uint64_t* drive_cdf;  // [drive_k] k = #{j : u >= cdf[j]}
  float drive_w;              // weight of one drive event (into the E channel)
};

template <typename R>
class Engine {
 public:
  using Acc = typename std::conditional<std::is_same<R, float>::value, int64_t, double>::type;

  explicit Engine(const NetDesc& d) : d_(d) {
    if (d.n_neurons == 0) throw std::runtime_error("network has 0 neurons");
    np_ = n_params(d.model);
    nv_ = n_vars(d.model);
    if (np_ < 0) throw std::runtime_error("unknown model");
    if (d.max_delay == 0 || d.max_delay > 64) throw std::runtime_error("max_delay out of range");
    N_ = d.n_neurons;
    M_ = d.n_sources;
    T_ = N_ + M_;
    D_ = d.max_delay;
    RS_ = D_ + 1;
    scale_ = std::ldexp(1.0, d.frac_bits);
    params_.resize((size_t)np_ * N_);
    vars_.resize((size_t)nv_ * N_);
    for (size_t i = 0; i < params_.size(); ++i) params_[i] = (R)d.params[i];
    for (size_t i = 0; i < vars_.size(); ++i) vars_[i] = (R)d.vars[i];
    src_last_.assign(M_, R(0));
    if (d.source_model == SRC_CONST_RATE && d.src_last)
      for (uint32_t j = 0; j < M_; ++j) src_last_[j] = (R)d.src_last[j];
    keys_.resize(T_);
    for (uint32_t i = 0; i < T_; ++i) keys_[i] = neuron_key(d.seed, i);
    // outgoing lists per source, input order kept
    off_.assign((size_t)T_ + 1, 0);
    for (uint64_t s = 0; s < d.n_syn; ++s) {
      if (d.syn_src[s] >= T_ || d.syn_tgt[s] >= N_) throw std::runtime_error("synapse id out of range");
      if (d.syn_delay[s] >= D_) throw std::runtime_error("synapse delay >= max_delay");
      if (d.syn_w[s] != 0.0f) off_[d.syn_src[s] + 1]++;
    }
    for (uint32_t i = 0; i < T_; ++i) off_[i + 1] += off_[i];
    tgt_.resize(off_[T_]);
    w_.resize(off_[T_]);
    wq_.resize(off_[T_]);
    dl_.resize(off_[T_]);
    std::vector<uint64_t> pos(off_.begin(), off_.end() - 1);
    for (uint64_t s = 0; s < d.n_syn; ++s) {
      if (d.syn_w[s] == 0.0f) continue;
      const uint64_t q = pos[d.syn_src[s]]++;
      tgt_[q] = d.syn_tgt[s];
      w_[q] = d.syn_w[s];
      wq_[q] = to_acc(d.syn_w[s]);
      dl_[q] = d.syn_delay[s];
    }
    ring_.assign((size_t)RS_ * 2 * N_, Acc(0));
    ring_events_.assign(RS_, 0);
    drive_acc_ = to_acc(d.drive_w);
  }

  uint64_t words() const { return (T_ + 63) / 64; }
  uint64_t cycle() const { return t_; }
  uint64_t events() const { return events_; }
  uint64_t drive_events() const { return drive_events_; }
  const std::vector<R>& vars() const { return vars_; }
  const std::vector<R>& src_last() const { return src_last_; }

  // One step.  forced: ids (neurons or sources) to fire this step.  raster_row:
  // words() u64 (bit i = id i fired), may be null.  v_row: N values of v after
  // the update, may be null.
  void step(const uint32_t* forced, uint64_t n_forced, uint64_t* raster_row, double* v_row) {
    // --- gather(t): take the slot due now (gather_kernel.cl:101-107,188-199)
    const size_t slot = (size_t)(t_ % RS_);
    Acc* accE = &ring_[(slot * 2 + 0) * N_];
    Acc* accI = &ring_[(slot * 2 + 1) * N_];
    events_ += ring_events_[slot];
    ring_events_[slot] = 0;
    forced_.assign(T_, 0);
    for (uint64_t f = 0; f < n_forced; ++f) {
      if (forced[f] >= T_) throw std::runtime_error("forced spike id not in the network");
      forced_[forced[f]] = 1;
    }
    spikes_.clear();
    // --- update(t) for neurons
    std::vector<R> p(np_), x(nv_);
    for (uint32_t i = 0; i < N_; ++i) {
      for (int k = 0; k < np_; ++k) p[k] = params_[(size_t)k * N_ + i];
      for (int k = 0; k < nv_; ++k) x[k] = vars_[(size_t)k * N_ + i];
      Acc ae = accE[i];
      if (d_.drive_k) {
        const uint64_t u = draw(keys_[i], t_, 2);
        uint32_t kk = 0;  // #{j : u >= cdf[j]}
        for (uint32_t j = 0; j < d_.drive_k; ++j) kk += u >= d_.drive_cdf[j] ? 1u : 0u;
        ae += (Acc)kk * drive_acc_;
        drive_events_ += kk;
      }
      const R e = from_acc(ae), in = from_acc(accI[i]);
      const R iinj = d_.noise_amp ? (R)d_.noise_amp[i] * gauss<R>(keys_[i], t_) : R(0);
      const bool s = update_neuron<R>(d_.model, p.data(), x.data(), e, in, iinj, forced_[i] != 0);
      for (int k = 0; k < nv_; ++k) vars_[(size_t)k * N_ + i] = x[k];
      if (s) spikes_.push_back(i);
      if (v_row) v_row[i] = (double)x[0];
      accE[i] = Acc(0);
      accI[i] = Acc(0);
    }
    // --- input model (Const_rate_update_kernel.cl:205-230; Poisson new)
    for (uint32_t j = 0; j < M_; ++j) {
      const uint32_t id = N_ + j;
      bool s = forced_[id] != 0;
      if (d_.source_model == SRC_CONST_RATE) {
        R last = src_last_[j] + R(1);
        s = s || (last == (R)d_.src_rate[j]);
        if (s) last = R(0);
        src_last_[j] = last;
      } else if (d_.source_model == SRC_POISSON) {
        s = s || (draw(keys_[id], t_, 3) < d_.src_thresh[j]);
      }
      if (s) spikes_.push_back(id);
    }
    // --- scatter(t): API delay k lands in the slot gathered at t+k+1
    for (uint32_t id : spikes_) {
      for (uint64_t q = off_[id]; q < off_[id + 1]; ++q) {
        const size_t sl = (size_t)((t_ + dl_[q] + 1) % RS_);
        const int ch = w_[q] < 0.0f ? 1 : 0;
        ring_[(sl * 2 + ch) * N_ + tgt_[q]] += wq_[q];
        ring_events_[sl] += 1;
      }
    }
    if (raster_row) {
      std::memset(raster_row, 0, words() * 8);
      for (uint32_t id : spikes_) raster_row[id / 64] |= (uint64_t)1 << (id % 64);
    }
    ++t_;
  }

  // The same step on `threads` OpenMP threads: the CPU baseline's step, and
  // the checker's step for large networks.  FP32 engine only: its
  // accumulators are integers, so delivering the spikes with relaxed atomic
  // adds in any order gives the serial step's sums exactly; neurons update
  // independently and the spike list keeps ascending ids (each thread a
  // contiguous range).  Arguments as step().
  void step_parallel(int threads, const uint32_t* forced = nullptr, uint64_t n_forced = 0,
                     uint64_t* raster_row = nullptr, double* v_row = nullptr) {
    if constexpr (!std::is_same<Acc, int64_t>::value) {
      (void)threads, (void)forced, (void)n_forced, (void)raster_row, (void)v_row;
      throw std::runtime_error("step_parallel: FP32 engine only");
    } else {
    const size_t slot = (size_t)(t_ % RS_);
    Acc* accE = &ring_[(slot * 2 + 0) * N_];
    Acc* accI = &ring_[(slot * 2 + 1) * N_];
    events_ += ring_events_[slot];
    ring_events_[slot] = 0;
    if (n_forced) {
      forced_.assign(T_, 0);
      for (uint64_t f = 0; f < n_forced; ++f) {
        if (forced[f] >= T_) throw std::runtime_error("forced spike id not in the network");
        forced_[forced[f]] = 1;
      }
    }
    const bool any_forced = n_forced != 0;
    std::vector<std::vector<uint32_t>> tsp(threads);
    std::vector<uint64_t> tdrive(threads, 0);
#pragma omp parallel num_threads(threads)
    {
      const int th = omp_get_thread_num();
      const uint32_t i0 = (uint32_t)((uint64_t)N_ * th / threads), i1 = (uint32_t)((uint64_t)N_ * (th + 1) / threads);
      std::vector<R> p(np_), x(nv_);
      uint64_t drive = 0;  // the thread's own count (one write at the end: no shared line per neuron)
      for (uint32_t i = i0; i < i1; ++i) {
        for (int k = 0; k < np_; ++k) p[k] = params_[(size_t)k * N_ + i];
        for (int k = 0; k < nv_; ++k) x[k] = vars_[(size_t)k * N_ + i];
        Acc ae = accE[i];
        if (d_.drive_k) {
          const uint64_t u = draw(keys_[i], t_, 2);
          uint32_t kk = 0;
          for (uint32_t j = 0; j < d_.drive_k; ++j) kk += u >= d_.drive_cdf[j] ? 1u : 0u;
          ae += (Acc)kk * drive_acc_;
          drive += kk;
        }
        const R e = from_acc(ae), in = from_acc(accI[i]);
        const R iinj = d_.noise_amp ? (R)d_.noise_amp[i] * gauss<R>(keys_[i], t_) : R(0);
        const bool sp = update_neuron<R>(d_.model, p.data(), x.data(), e, in, iinj, any_forced && forced_[i] != 0);
        for (int k = 0; k < nv_; ++k) vars_[(size_t)k * N_ + i] = x[k];
        if (sp) tsp[th].push_back(i);
        if (v_row) v_row[i] = (double)x[0];
        accE[i] = Acc(0);
        accI[i] = Acc(0);
      }
      tdrive[th] = drive;
    }
    spikes_.clear();
    for (int th = 0; th < threads; ++th) {
      spikes_.insert(spikes_.end(), tsp[th].begin(), tsp[th].end());
      drive_events_ += tdrive[th];
    }
    for (uint32_t j = 0; j < M_; ++j) {
      const uint32_t id = N_ + j;
      bool sp = any_forced && forced_[id] != 0;
      if (d_.source_model == SRC_CONST_RATE) {
        R last = src_last_[j] + R(1);
        sp = sp || last == (R)d_.src_rate[j];
        if (sp) last = R(0);
        src_last_[j] = last;
      } else if (d_.source_model == SRC_POISSON) {
        sp = sp || draw(keys_[id], t_, 3) < d_.src_thresh[j];
      }
      if (sp) spikes_.push_back(id);
    }
    const int64_t ns = (int64_t)spikes_.size();
    std::vector<uint64_t> tev((size_t)threads * RS_, 0);  // events per thread and delivery slot
#pragma omp parallel num_threads(threads)
    {
      uint64_t* ev = &tev[(size_t)omp_get_thread_num() * RS_];
#pragma omp for schedule(dynamic, 4)
      for (int64_t si = 0; si < ns; ++si) {
        const uint32_t id = spikes_[si];
        for (uint64_t q = off_[id]; q < off_[id + 1]; ++q) {
          const size_t sl = (size_t)((t_ + dl_[q] + 1) % RS_);
          const int ch = w_[q] < 0.0f ? 1 : 0;
          __atomic_fetch_add(&ring_[(sl * 2 + ch) * N_ + tgt_[q]], wq_[q], __ATOMIC_RELAXED);
          ev[sl] += 1;
        }
      }
    }
    for (int th = 0; th < threads; ++th)
      for (uint32_t sl = 0; sl < RS_; ++sl) ring_events_[sl] += tev[(size_t)th * RS_ + sl];
    if (raster_row) {
      std::memset(raster_row, 0, words() * 8);
      for (uint32_t id : spikes_) raster_row[id / 64] |= (uint64_t)1 << (id % 64);
    }
    ++t_;
    }
  }

 private:
  Acc to_acc(float w) const {
    if (std::is_same<Acc, int64_t>::value) return (Acc)std::llrint((double)w * scale_);
    return (Acc)w;
  }
  R from_acc(Acc a) const {
    if (std::is_same<Acc, int64_t>::value) return (R)((float)(int64_t)a * (float)(1.0 / scale_));
    return (R)a;
  }

  NetDesc d_;
  int np_ = 0, nv_ = 0;
  uint32_t N_ = 0, M_ = 0, T_ = 0, D_ = 0, RS_ = 0;
  double scale_ = 1.0;
  uint64_t t_ = 0, events_ = 0, drive_events_ = 0;
  Acc drive_acc_ = 0;
  std::vector<R> params_, vars_, src_last_;
  std::vector<uint64_t> keys_, off_;
  std::vector<uint32_t> tgt_;
  std::vector<float> w_;
  std::vector<Acc> wq_;
  std::vector<uint16_t> dl_;
  std::vector<Acc> ring_;
  std::vector<uint64_t> ring_events_;
  std::vector<uint8_t> forced_;
  std::vector<uint32_t> spikes_;
};

}  // namespace szsnn_oracle
