// szsnn.hpp -- C++ host layer over the C ABI (include/szsnn_hip.h).
//
// Keeps the reference's construction and simulation API for C++ drivers,
// without Boost or OpenCL:
//   Context                         szsnn1.1/src/Context.hpp:77-303 (setup)
//   Synapse<floatT>                 szsnn1.1/src/synapse_model.h
//   *_NeuronModel / *_InputModel    szsnn1.1/src/models/*.hpp (param/var order kept)
//   NetworkGraph<NeuronM, InputM>   szsnn1.1/src/NetworkGraph.{hpp,ipp}
//                                   (addNeuron, addSynapse, addInput,
//                                   addInputSynapse, connectInput)
//   Simulation<floatT>              szsnn1.1/src/Simulation.hpp:39-131 (step,
//                                   stepAndFired, getTime)
//   Connection<floatT>              szsnn1.1/src/Connection.hpp:76-174 (.conn files)
//   SpikeRecorder                   szsnn1.1/src/SpikeRecorder.cpp:13-104 (written
//                                   as it records)
//   construction::Population<M>     szsnn1.1/src/Population.hpp:14-397 (population
//                                   and template files)
//   construction::NetworkCreator    szsnn1.1/src/NetworkCreator.hpp:21-188
//   construction::Connection        the Connection above
//   record::SpikeRecorder           szsnn1.1/src/SpikeRecorder.{hpp,cpp} (in
//                                   memory, saved by cycles or by neurons)
//   SzsnnException                  szsnn1.1/src/SzsnnException.hpp:21-41
// Every negative return code of the C ABI becomes an SzsnnException.
//
// Delays are API delays k (a spike of step c reaches the update of step
// c+k+1), as passed to the reference's addSynapse before it adds one.
// Model parameters are in per-step units, as in the reference kernels.
#pragma once

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <map>
#include <set>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/szsnn_hip.h"

namespace szsnn {

class SzsnnException : public std::exception {
 public:
  SzsnnException(const std::string& msg, int code = SZSNN_ERR_INVALID) : msg_(msg), code_(code) {}
  const char* what() const noexcept override { return msg_.c_str(); }
  int code() const { return code_; }

 private:
  std::string msg_;
  int code_;
};

inline void check(int rc, const szsnn_ctx* ctx, const char* what) {
  if (rc != SZSNN_OK) {
    std::ostringstream os;
    os << "SZSNN::" << what << " error: " << szsnn_last_error(ctx);
    throw SzsnnException(os.str(), rc);
  }
}

// ------------------------------------------------------------------ device
class Context {
 public:
  explicit Context(int device = 0) {
    szsnn_ctx* c = nullptr;
    check(szsnn_create(device, &c), nullptr, "Context::setup()");
    h_ = std::shared_ptr<szsnn_ctx>(c, [](szsnn_ctx* p) { szsnn_destroy(p); });
  }
  // Context::setup (Context.cpp:136-224): the device was chosen by the constructor
  void setup() {}
  szsnn_ctx* handle() const { return h_.get(); }
  std::string getDeviceName() const {
    char buf[256];
    check(szsnn_device_name(h_.get(), buf, sizeof(buf)), h_.get(), "Context::getDeviceName()");
    return buf;
  }

 private:
  std::shared_ptr<szsnn_ctx> h_;
};

// ------------------------------------------------------------------ models
template <typename floatT>
struct Synapse {
  floatT weight = 0;
  unsigned delay = 0;
  unsigned synapseIdx = 0;
  unsigned sourceNeuronIdx = 0;
  unsigned targetNeuronIdx = 0;
};

template <typename floatT, int MODEL, unsigned NP, unsigned NV>
class ModelBase {
 public:
  static constexpr int model_id = MODEL;
  static constexpr unsigned param_count = NP;
  static constexpr unsigned var_count = NV;
  floatT getParam(unsigned i) const {
    if (i >= NP) throw SzsnnException("SZSNN::NeuronModel::getParam: paramIdx out of range");
    return params[i];
  }
  floatT getVar(unsigned i) const {
    if (i >= NV) throw SzsnnException("SZSNN::NeuronModel::getVar: varIdx out of range");
    return vars[i];
  }
  void setParam(unsigned i, floatT x) {
    if (i >= NP) throw SzsnnException("SZSNN::NeuronModel::setParam: paramIdx out of range");
    params[i] = x;
  }
  void setVar(unsigned i, floatT x) {
    if (i >= NV) throw SzsnnException("SZSNN::NeuronModel::setVar: varIdx out of range");
    vars[i] = x;
  }
  void setAll(const floatT* p, const floatT* v) {
    std::copy(p, p + NP, params.begin());
    std::copy(v, v + NV, vars.begin());
  }
  std::array<floatT, NP> params{};
  std::array<floatT, NV> vars{};
};

// IF_curr_exp_NeuronModel.hpp:12-54 (defaults kept)
template <typename floatT>
class IF_curr_exp_NeuronModel : public ModelBase<floatT, SZSNN_MODEL_IF_CURR_EXP, 9, 4> {
 public:
  static const char* name() { return "IF_curr_exp_NeuronModel"; }
  enum Params { p_V_REST = 0, p_CM, p_TAU_M, p_TAU_REFRAC, p_TAU_SYN_E, p_TAU_SYN_I, p_I_OFFSET, p_V_RESET, p_V_THRESH };
  enum Vars { v_V = 0, v_IE, v_II, v_FIRED };
  IF_curr_exp_NeuronModel() {
    const floatT p[9] = {-65, 1, 20, 0, 5, 5, 0, -65, -65};
    const floatT v[4] = {0, 0, 0, 10000};
    this->setAll(p, v);
  }
};

// IF_cond_exp (no reference model): the IF_curr_exp parameters plus reversal
// potentials; variables v, ge, gi (stored negative), fired.
template <typename floatT>
class IF_cond_exp_NeuronModel : public ModelBase<floatT, SZSNN_MODEL_IF_COND_EXP, 11, 4> {
 public:
  static const char* name() { return "IF_cond_exp_NeuronModel"; }
  enum Params { p_V_REST = 0, p_CM, p_TAU_M, p_TAU_REFRAC, p_TAU_SYN_E, p_TAU_SYN_I, p_I_OFFSET, p_V_RESET, p_V_THRESH,
                p_E_REV_E, p_E_REV_I };
  enum Vars { v_V = 0, v_GE, v_GI, v_FIRED };
  IF_cond_exp_NeuronModel() {
    const floatT p[11] = {-65, 1, 20, 0, 5, 5, 0, -65, -50, 0, -70};
    const floatT v[4] = {-65, 0, 0, 10000};
    this->setAll(p, v);
  }
};

// Izhikevich_NeuronModel.hpp:12-43 (defaults kept)
template <typename floatT>
class Izhikevich_NeuronModel : public ModelBase<floatT, SZSNN_MODEL_IZHIKEVICH, 4, 2> {
 public:
  static const char* name() { return "Izhikevich_NeuronModel"; }
  enum Params { p_A = 0, p_B, p_C, p_D };
  enum Vars { v_V = 0, v_U };
  Izhikevich_NeuronModel() {
    const floatT p[4] = {0.02, 0.2, -65.0, 2};
    const floatT v[2] = {0, 0};
    this->setAll(p, v);
  }
};

// Const_rate_InputModel.hpp:11-32: fires every `rate` steps
template <typename floatT>
class Const_rate_InputModel : public ModelBase<floatT, SZSNN_SOURCE_CONST_RATE, 1, 1> {
 public:
  static const char* name() { return "Const_rate_InputModel"; }
  enum Params { p_RATE = 0 };
  enum Vars { v_LAST = 0 };
  Const_rate_InputModel() {
    const floatT p[1] = {10};
    const floatT v[1] = {0};
    this->setAll(p, v);
  }
};

// Poisson source (no reference model): param = spike probability per step
template <typename floatT>
class Poisson_InputModel : public ModelBase<floatT, SZSNN_SOURCE_POISSON, 1, 0> {
 public:
  static const char* name() { return "Poisson_InputModel"; }
  enum Params { p_PROB = 0 };
  Poisson_InputModel() { this->params[0] = 0; }
};

// ------------------------------------------------------------------ graph
template <typename NeuronModelT, typename InputModelT>
class NetworkGraph {
 public:
  using float_type = float;

  unsigned addNeuron(const NeuronModelT& n) {
    if (inputConnected_) throw SzsnnException("SZSNN::NetworkGraph::addNeuron(neuron) error: cant execute, after input connection");
    neurons.push_back(n);
    synapses.emplace_back();
    return (unsigned)neurons.size() - 1;
  }
  unsigned addNeuron(size_t idx, const NeuronModelT& n) {
    if (idx < neurons.size()) {
      neurons[idx] = n;
      return (unsigned)neurons.size() - 1;
    }
    if (idx > neurons.size()) {
      std::ostringstream os;
      os << "SZSNN::NetworkGraph::addNeuron(idx, neuron) error: cant add neuron with ID=`" << idx
         << "`, ID's must be consistance";
      throw SzsnnException(os.str());
    }
    return addNeuron(n);
  }
  unsigned addInput(const InputModelT& in) {
    if (inputConnected_) throw SzsnnException("SZSNN::NetworkGraph::addInput(input) error: cant execute, after input connection");
    inputs.push_back(in);
    inputSynapses.emplace_back();
    return (unsigned)inputs.size() - 1;
  }
  unsigned addInput(size_t idx, const InputModelT& in) {
    if (idx < inputs.size()) {
      inputs[idx] = in;
      return (unsigned)inputs.size() - 1;
    }
    if (idx > inputs.size()) {
      std::ostringstream os;
      os << "SZSNN::NetworkGraph::addInput(idx, input) error: cant add input with ID=`" << idx
         << "`, ID's must be consistance";
      throw SzsnnException(os.str());
    }
    return addInput(in);
  }
  void addSynapse(const Synapse<float_type>& s) {
    if (s.sourceNeuronIdx >= neurons.size() || s.targetNeuronIdx >= neurons.size()) {
      std::ostringstream os;
      os << "SZSNN::NetworkGraph::addSynapse(syn) error: cant add synapse connecting nonexisting neurons sourceID=`"
         << s.sourceNeuronIdx << "` targetID=`" << s.targetNeuronIdx << "`";
      throw SzsnnException(os.str());
    }
    synapses[s.sourceNeuronIdx].push_back(s);
    maxDelay_ = std::max(maxDelay_, s.delay);
  }
  void addInputSynapse(const Synapse<float_type>& s) {
    if (s.sourceNeuronIdx >= inputs.size() || s.targetNeuronIdx >= neurons.size()) {
      std::ostringstream os;
      os << "SZSNN::NetworkGraph::addInputSynapse(syn) error: cant add synapse connecting nonexisting neurons sourceID=`"
         << s.sourceNeuronIdx << "` targetID=`" << s.targetNeuronIdx << "`";
      throw SzsnnException(os.str());
    }
    inputSynapses[s.sourceNeuronIdx].push_back(s);
    maxDelay_ = std::max(maxDelay_, s.delay);
  }
  // NetworkGraph.ipp:520-536; input synapses keep their own list here, the
  // id offset is applied when the description is built
  void connectInput() { inputConnected_ = true; }
  size_t getNetworkSize() const { return neurons.size(); }
  size_t getInputCount() const { return inputs.size(); }
  unsigned getInputsIDOffset() const { return (unsigned)neurons.size(); }
  unsigned getMaxDelay() const { return maxDelay_; }
  // debug listings (NetworkGraph.ipp:341-366)
  std::string neuronsToString() const {
    std::ostringstream ss;
    for (size_t i = 0; i < neurons.size(); ++i) ss << "Idx [ " << i << " ], param1 = " << neurons[i].getParam(0) << "\n";
    return ss.str();
  }
  std::string synapsesToString() const {
    std::ostringstream ss;
    for (size_t i = 0; i < synapses.size(); ++i) {
      ss << "Idx [ " << i << " ] ";
      for (size_t j = 0; j < synapses[i].size(); ++j)
        ss << j << "[ " << synapses[i][j].sourceNeuronIdx << " -(" << synapses[i][j].weight << ", "
           << synapses[i][j].delay << ")-> " << synapses[i][j].targetNeuronIdx << "]   ";
      ss << "\n";
    }
    return ss.str();
  }

  std::vector<NeuronModelT> neurons;
  std::vector<InputModelT> inputs;
  std::vector<std::vector<Synapse<float_type>>> synapses;
  std::vector<std::vector<Synapse<float_type>>> inputSynapses;
  uint64_t seed = 0;

 private:
  bool inputConnected_ = false;
  unsigned maxDelay_ = 0;
};

// ------------------------------------------------------------------ simulation
template <typename floatT>
class Simulation {
 public:
  typedef std::vector<unsigned> firing_output;
  typedef std::vector<unsigned> firing_input;

  template <typename NeuronModelT, typename InputModelT>
  Simulation(const Context& context, NetworkGraph<NeuronModelT, InputModelT>& network) : ctx_(context) {
    if (!network.getNetworkSize())
      throw SzsnnException("SZSNN::Simulation::Simulation() error: network has 0 neurons");
    const unsigned N = (unsigned)network.neurons.size(), M = (unsigned)network.inputs.size();
    const unsigned np = NeuronModelT::param_count, nv = NeuronModelT::var_count;
    std::vector<float> params((size_t)np * N), vars((size_t)nv * N);
    for (unsigned i = 0; i < N; ++i) {
      for (unsigned k = 0; k < np; ++k) params[(size_t)k * N + i] = (float)network.neurons[i].params[k];
      for (unsigned k = 0; k < nv; ++k) vars[(size_t)k * N + i] = (float)network.neurons[i].vars[k];
    }
    std::vector<float> rate, last;
    std::vector<uint64_t> thresh;
    const int src_model = M ? InputModelT::model_id : SZSNN_SOURCE_NONE;
    for (unsigned j = 0; j < M; ++j) {
      if (src_model == SZSNN_SOURCE_CONST_RATE) {
        rate.push_back((float)network.inputs[j].params[0]);
        last.push_back(InputModelT::var_count ? (float)network.inputs[j].getVar(0) : 0.0f);
      } else {
        const double p = (double)network.inputs[j].params[0];
        thresh.push_back(p >= 1.0 ? ~0ull : (uint64_t)std::ldexp(std::max(p, 0.0), 64));
      }
    }
    std::vector<uint32_t> src, tgt;
    std::vector<float> w;
    std::vector<uint16_t> dl;
    double maxw = 0;
    auto add = [&](const Synapse<float>& s, uint32_t from) {
      src.push_back(from);
      tgt.push_back(s.targetNeuronIdx);
      w.push_back((float)s.weight);
      dl.push_back((uint16_t)s.delay);
      maxw = std::max(maxw, std::fabs((double)s.weight));
    };
    for (unsigned i = 0; i < N; ++i)
      for (const auto& s : network.synapses[i]) add(s, i);
    for (unsigned j = 0; j < M; ++j)
      for (const auto& s : network.inputSynapses[j]) add(s, N + j);
    szsnn_net_desc d{};
    d.model = NeuronModelT::model_id;
    d.n_neurons = N;
    d.n_sources = M;
    d.source_model = src_model;
    d.params = params.data();
    d.vars = vars.data();
    d.src_rate = rate.empty() ? nullptr : rate.data();
    d.src_last = last.empty() ? nullptr : last.data();
    d.src_thresh = thresh.empty() ? nullptr : thresh.data();
    d.n_syn = src.size();
    d.syn_src = src.data();
    d.syn_tgt = tgt.data();
    d.syn_w = w.data();
    d.syn_delay = dl.data();
    d.max_delay = network.getMaxDelay() + 1;
    // weights as multiples of 2^-F with max|w| * 2^F < 2^30
    d.frac_bits = maxw > 0 ? std::min(40, 30 - (int)std::ceil(std::log2(maxw))) : 32;
    d.seed = network.seed;
    szsnn_net* h = nullptr;
    check(szsnn_net_upload(ctx_.handle(), &d, &h), ctx_.handle(), "Simulation::Simulation()");
    net_ = std::shared_ptr<szsnn_net>(h, [](szsnn_net* p) { szsnn_net_destroy(p); });
    szsnn_net_info info{};
    check(szsnn_net_get_info(h, &info), ctx_.handle(), "Simulation::Simulation()");
    n_neurons_ = info.n_neurons;
    n_vars_ = info.n_vars;
    bitmap_.assign(info.raster_words, 0);
  }

  void step() { do_step(nullptr, nullptr); }
  void step(const firing_input& fInput) { do_step(&fInput, nullptr); }
  const firing_output& stepAndFired() {
    do_step(nullptr, bitmap_.data());
    return decode();
  }
  const firing_output& stepAndFired(const firing_input& fInput) {
    do_step(&fInput, bitmap_.data());
    return decode();
  }
  // n steps in the graph-captured loop
  void run(uint64_t n) {
    const auto t0 = std::chrono::steady_clock::now();
    check(szsnn_run(net_.get(), n), ctx_.handle(), "Simulation::run()");
    wall_ms_ += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  // wall-clock milliseconds spent stepping (SzsnnTimer::getRealTime)
  size_t getTime() const { return (size_t)wall_ms_; }
  szsnn_net* handle() const { return net_.get(); }

  // Single-kernel harnesses (Simulation.ipp:142-210): one kernel of the
  // current step, launched alone and waited for; the cycle does not advance.
  // The scatter is part of the update kernel here, so testScatter launches
  // the update.  Each returns the kernel's time in ms.
  double testUpdate() { return test_kernel(SZSNN_KERNEL_UPDATE, "UPDATE"); }
  double testScatter() { return test_kernel(SZSNN_KERNEL_UPDATE, "SCATTER"); }
  double testGather() { return test_kernel(SZSNN_KERNEL_GATHER, "GATHER"); }

  // State readback (the reference keeps host mirrors of the device state in
  // its public NeuronData / SpikeBuffer members, Simulation.hpp:100-105):
  // neuron variables [var][neuron], the counters since upload, the spike
  // ids of the last stepAndFired.
  unsigned getNeuronCount() const { return n_neurons_; }
  unsigned getVarCount() const { return n_vars_; }
  std::vector<float> getVars() const {
    std::vector<float> v((size_t)n_vars_ * n_neurons_);
    check(szsnn_get_vars(net_.get(), v.data()), ctx_.handle(), "Simulation::getVars()");
    return v;
  }
  void setVars(const std::vector<float>& v) {
    if (v.size() != (size_t)n_vars_ * n_neurons_)
      throw SzsnnException("SZSNN::Simulation::setVars() error: expected var_count x neuron_count values");
    check(szsnn_set_vars(net_.get(), v.data()), ctx_.handle(), "Simulation::setVars()");
  }
  szsnn_counters getCounters() const {
    szsnn_counters c{};
    check(szsnn_get_counters(net_.get(), &c), ctx_.handle(), "Simulation::getCounters()");
    return c;
  }
  uint64_t getCycle() const { return getCounters().cycle; }
  const firing_output& spikes() const { return fired_; }

 private:
  void do_step(const firing_input* in, uint64_t* out) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t* ids = (in && !in->empty()) ? in->data() : nullptr;
    const uint32_t n = ids ? (uint32_t)in->size() : 0;
    check(szsnn_step(net_.get(), ids, n, out), ctx_.handle(), "Simulation::step()");
    wall_ms_ += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  double test_kernel(int32_t which, const char* name) {
    std::cout << "Simulation test -> " << name << " KERNEL" << std::endl;
    double ms = 0;
    check(szsnn_test_kernel(net_.get(), which, &ms), ctx_.handle(), "Simulation::test()");
    return ms;
  }
  // SpikeBuffer::bitvectorToSparse (SpikeBuffer.cpp:146-165): ascending neuron ids
  const firing_output& decode() {
    fired_.clear();
    for (size_t w = 0; w < bitmap_.size(); ++w) {
      uint64_t x = bitmap_[w];
      while (x) {
        const unsigned b = (unsigned)__builtin_ctzll(x);
        const unsigned id = (unsigned)(w * 64 + b);
        if (id < n_neurons_) fired_.push_back(id);
        x &= x - 1;
      }
    }
    return fired_;
  }

  Context ctx_;
  std::shared_ptr<szsnn_net> net_;
  unsigned n_neurons_ = 0, n_vars_ = 0;
  std::vector<uint64_t> bitmap_;
  firing_output fired_;
  double wall_ms_ = 0;
};

// ------------------------------------------------------------------ files
// .conn files (Connection.hpp:76-174): two header lines, then
// "src tgt weight delay" per line; ids relative to the projection's offsets.
template <typename floatT>
class Connection {
 public:
  Connection(unsigned fromStart = 0, unsigned toStart = 0) : from_(fromStart), to_(toStart) {}
  void addConnection(unsigned fromIdx, unsigned toIdx, floatT weight, unsigned delay) {
    connections[{fromIdx, toIdx}] = {weight, delay};
  }
  void loadFromFile(const std::string& filename) {
    std::ifstream in(filename);
    if (!in.is_open())
      throw SzsnnException("SZSNN::CONSTRUCTION::Connection::loadFromFile() error: couldn't open file `" + filename + "`");
    std::string line;
    std::getline(in, line);
    std::getline(in, line);
    size_t lineIdx = 1;
    while (std::getline(in, line)) {
      std::istringstream ss(line);
      std::vector<std::string> v{std::istream_iterator<std::string>(ss), std::istream_iterator<std::string>()};
      if (v.empty()) continue;
      if (v.size() != 4) {
        std::ostringstream os;
        os << "SZSNN::CONSTRUCTION::Connection::loadFromFile() error: wrong format in file `" << filename
           << "` at line #" << lineIdx << " at value #" << v.size();
        throw SzsnnException(os.str());
      }
      addConnection((unsigned)std::stof(v[0]), (unsigned)std::stof(v[1]), (floatT)std::stod(v[2]),
                    (unsigned)std::stof(v[3]));
      ++lineIdx;
    }
  }
  void saveToFile(const std::string& filename) const {
    std::ofstream out(filename);
    out << "#label p1\n#label p2\n";
    for (const auto& kv : connections)  // std::map: sorted by (src, tgt)
      out << kv.first.first << "\t" << kv.first.second << "\t" << kv.second.first << "\t"
          << (float)kv.second.second << "\n";
  }
  typedef typename std::map<std::pair<unsigned, unsigned>, std::pair<floatT, unsigned>>::const_iterator iterator;
  iterator begin() const { return connections.begin(); }
  iterator end() const { return connections.end(); }
  unsigned getFromStartID() const { return from_; }
  unsigned getToStartID() const { return to_; }
  template <typename N, typename I>
  void applyTo(NetworkGraph<N, I>& g) const {
    for (const auto& kv : connections) {
      Synapse<float> s;
      s.sourceNeuronIdx = from_ + kv.first.first;
      s.targetNeuronIdx = to_ + kv.first.second;
      s.weight = (float)kv.second.first;
      s.delay = kv.second.second;
      g.addSynapse(s);
    }
  }
  std::map<std::pair<unsigned, unsigned>, std::pair<floatT, unsigned>> connections;

 private:
  unsigned from_, to_;
};

// SpikeRecorder (SpikeRecorder.cpp:13-104): "cycle<TAB>neuronID" lines after a
// two-line header; the header always ends with a newline.
class SpikeRecorder {
 public:
  SpikeRecorder(const std::string& filename, size_t recordStart = 0, size_t recordStop = 0)
      : out_(filename), start_(recordStart), stop_(recordStop) {
    out_ << "#spikes by cycles\n#record start at: " << start_;
    if (stop_) out_ << "   stop at: " << stop_;
    out_ << "\n";
  }
  void recordCycleSpikes(size_t cycle, const std::vector<unsigned>& spikes) {
    if (cycle < start_ || (stop_ && cycle > stop_)) return;  // the stop cycle is recorded
    if (have_last_ && cycle <= last_)
      throw SzsnnException("SZSNN::SpikeRecorder::recordCycleSpikes error: cycle already recorded");
    for (size_t i = 0; i < spikes.size(); ++i) {
      if (i && spikes[i] <= spikes[i - 1])
        throw SzsnnException("SZSNN::SpikeRecorder::recordCycleSpikes error: spike recorded twice");
      out_ << cycle << "\t" << spikes[i] << "\n";
    }
    last_ = cycle;
    have_last_ = true;
  }

 private:
  std::ofstream out_;
  size_t start_, stop_;
  size_t last_ = 0;
  bool have_last_ = false;
};


// ------------------------------------------------------------------ construction
namespace construction {

using szsnn::Connection;

// Population<NeuronModelT> (Population.hpp:14-397): `size` neurons with ids
// first .. first + size - 1.  Files: `#` comment lines, then the size, the
// first id, the label (one word), then one parameter line and one variable
// line per neuron (the template file: one pair for all neurons).  The
// reference's writer left out the label its reader expects; the files
// written here have it, so they read back.  The files of the pyszsnn Brian
// dumper (pyszsnn/brian/dumper.py:12-126: label or model name first) read
// too; see header().
template <typename NeuronModelT>
class Population {
 public:
  typedef float floatT;
  typedef typename std::vector<NeuronModelT>::iterator iterator;
  typedef typename std::vector<NeuronModelT>::const_iterator const_iterator;
  static constexpr unsigned param_count = NeuronModelT::param_count;
  static constexpr unsigned var_count = NeuronModelT::var_count;

  Population(size_t size, size_t firstNeuronID, std::string label = "")
      : firstNeuronID_(firstNeuronID), label_(std::move(label)), neurons_(size) {
    if (!size) {
      std::ostringstream os;
      os << "SZSNN::CONSTRUCTION::Population() error: Population size=" << size << ", must be > 0";
      throw SzsnnException(os.str());
    }
  }
  Population() = default;

  void clear() {
    firstNeuronID_ = 0;
    label_.clear();
    neurons_.clear();
  }
  // params[param_count], vars[var_count]
  void setNeurons(const floatT* params, const floatT* vars) {
    for (auto& n : neurons_) n.setAll(params, vars);
  }
  void setNeuronParameter(unsigned long idx, unsigned param_idx, floatT val) { at(idx).setParam(param_idx, val); }
  void setNeuronParameters(unsigned param_idx, floatT val) {
    for (auto& n : neurons_) n.setParam(param_idx, val);
  }
  void setNeuronVariables(unsigned var_idx, floatT val) {
    for (auto& n : neurons_) n.setVar(var_idx, val);
  }

  void loadFromFile(const std::string& filename) { load(filename, false, "loadFromFile"); }
  void loadFromTemplateFile(const std::string& filename) { load(filename, true, "loadFromTemplateFile"); }
  void saveToFile(const std::string& filename) const { save(filename, false, "saveToFile"); }
  void saveToTemplateFile(const std::string& filename) const { save(filename, true, "saveToTemplateFile"); }

  iterator begin() { return neurons_.begin(); }
  iterator end() { return neurons_.end(); }
  const_iterator begin() const { return neurons_.begin(); }
  const_iterator end() const { return neurons_.end(); }
  size_t getFirstId() const { return firstNeuronID_; }
  size_t getLastId() const { return firstNeuronID_ + neurons_.size() - 1; }
  std::string getLabel() const { return label_; }
  size_t getSize() const { return neurons_.size(); }

 private:
  NeuronModelT& at(unsigned long idx) {
    if (idx >= neurons_.size()) {
      std::ostringstream os;
      os << "SZSNN::CONSTRUCTION::Population::setNeuronParameter() error: neuron " << idx << " of " << neurons_.size();
      throw SzsnnException(os.str());
    }
    return neurons_[idx];
  }
  [[noreturn]] static void fail(const char* what, const std::string& filename, const std::string& why) {
    throw SzsnnException(std::string("SZSNN::CONSTRUCTION::Population::") + what + "() error: " + why + " in file `" +
                         filename + "`");
  }
  // the next non-empty line that is not a comment
  static bool dataLine(std::istream& in, std::string& line, size_t& lineIdx) {
    while (std::getline(in, line)) {
      ++lineIdx;
      const size_t b = line.find_first_not_of(" \t\r");
      if (b == std::string::npos || line[b] == '#') continue;
      line = line.substr(b);
      return true;
    }
    return false;
  }
  template <unsigned K>
  static void values(const std::string& line, floatT (&out)[K], const char* what, const std::string& filename,
                     size_t lineIdx) {
    std::istringstream ss(line);
    std::vector<std::string> v{std::istream_iterator<std::string>(ss), std::istream_iterator<std::string>()};
    if (v.size() != K) {
      std::ostringstream os;
      os << "wrong format at line #" << lineIdx << ": " << v.size() << " values, expected " << K;
      fail(what, filename, os.str());
    }
    for (unsigned k = 0; k < K; ++k) {
      try {
        out[k] = std::stof(v[k]);
      } catch (const std::exception&) {
        std::ostringstream os;
        os << "wrong format at line #" << lineIdx << " at value #" << k + 1;
        fail(what, filename, os.str());
      }
    }
  }
  static bool isInt(const std::string& t) {
    if (t.empty()) return false;
    size_t i = (t[0] == '-' || t[0] == '+') ? 1 : 0;
    if (i == t.size()) return false;
    for (; i < t.size(); ++i)
      if (t[i] < '0' || t[i] > '9') return false;
    return true;
  }
  static std::vector<std::string> words(const std::string& line) {
    std::istringstream ss(line);
    return {std::istream_iterator<std::string>(ss), std::istream_iterator<std::string>()};
  }
  // The header in any of three orders: size, first id, label (the
  // reference's); or a word (the label of a pyszsnn Brian template, or the
  // model name of a Brian population file) followed by the integers up to
  // the first parameter line, of which the first is the size and the last
  // the first id (templates: size, id; populations: size, size, id).  A
  // reference-order header whose third line has several values has no
  // label (a Brian template written with an empty one).
  void header(const std::vector<std::string>& lines, size_t& k, long long& size, const char* what,
              const std::string& filename) {
    auto num = [&](size_t i) -> long long {
      if (i >= lines.size() || !isInt(lines[i])) fail(what, filename, "wrong header");
      return std::stoll(lines[i]);
    };
    if (lines.empty()) fail(what, filename, "empty file");
    if (isInt(lines[0])) {
      size = num(0);
      firstNeuronID_ = (size_t)num(1);
      if (lines.size() > 2 && words(lines[2]).size() == 1) {
        label_ = words(lines[2])[0];
        k = 3;
      } else {
        k = 2;
      }
      return;
    }
    label_ = words(lines[0])[0];
    std::vector<long long> ints;
    for (k = 1; k < lines.size() && words(lines[k]).size() == 1 && isInt(words(lines[k])[0]); ++k)
      ints.push_back(std::stoll(lines[k]));
    if (!(ints.size() == 2 || (ints.size() == 3 && ints[0] == ints[1]))) fail(what, filename, "wrong header");
    size = ints.front();
    firstNeuronID_ = (size_t)ints.back();
  }
  void load(const std::string& filename, bool tmpl, const char* what) {
    std::ifstream in(filename);
    if (!in.is_open()) fail(what, filename, "couldn't open file");
    clear();
    std::vector<std::string> lines;
    std::vector<size_t> lineNo;
    {
      std::string line;
      size_t lineIdx = 0;
      while (dataLine(in, line, lineIdx)) {
        lines.push_back(line);
        lineNo.push_back(lineIdx);
      }
    }
    size_t k = 0;
    long long size = 0;
    header(lines, k, size, what, filename);
    if (size <= 0) fail(what, filename, "Population size must be > 0");
    neurons_.assign((size_t)size, NeuronModelT());
    floatT p[param_count], v[var_count > 0 ? var_count : 1];
    const size_t per = var_count > 0 ? 2 : 1;
    for (size_t i = 0; i < (tmpl ? 1 : (size_t)size); ++i, k += per) {
      if (k + per > lines.size()) {
        std::ostringstream os;
        os << "not enough data: " << i << " of " << size << " neurons";
        fail(what, filename, os.str());
      }
      values<param_count>(lines[k], p, what, filename, lineNo[k]);
      if constexpr (var_count > 0) values<var_count>(lines[k + 1], v, what, filename, lineNo[k + 1]);
      if (tmpl)
        for (auto& n : neurons_) n.setAll(p, v);
      else
        neurons_[i].setAll(p, v);
    }
  }
  void save(const std::string& filename, bool tmpl, const char* what) const {
    if (neurons_.empty()) fail(what, filename, "trying to save an empty Population");
    std::ofstream out(filename);
    if (!out.is_open()) fail(what, filename, "couldn't open file");
    out << "# Population " << (tmpl ? "Template " : "") << "file for : " << label_ << "\n# Model: "
        << NeuronModelT::name() << "\n# Size: " << neurons_.size() << "\n# First neuron ID: " << firstNeuronID_
        << "\n" << neurons_.size() << "\n" << firstNeuronID_ << "\n" << (label_.empty() ? "population" : label_) << "\n";
    out << std::setprecision(9);
    for (size_t i = 0; i < (tmpl ? 1 : neurons_.size()); ++i) {
      for (unsigned k = 0; k < param_count; ++k) out << (k ? " " : "") << neurons_[i].getParam(k);
      out << "\n";
      if (var_count) {
        for (unsigned k = 0; k < var_count; ++k) out << (k ? " " : "") << neurons_[i].getVar(k);
        out << "\n";
      }
    }
    if (!out) fail(what, filename, "output failure");
  }

  size_t firstNeuronID_ = 0;
  std::string label_;
  std::vector<NeuronModelT> neurons_;
};

// NetworkCreator (NetworkCreator.hpp:21-188): populations and projections
// gathered, then written into a NetworkGraph.  Populations are taken in
// order of their first id and must leave no gap; the input projections
// become input synapses (the reference sorted with a comparator that
// compared a population with itself, left the input size uninitialised and
// added its neuron projections a second time as input synapses: the
// intended behaviour is kept here).
template <typename NeuronModelT, typename InputModelT>
class NetworkCreator {
 public:
  typedef float float_type;
  NetworkCreator() = default;
  NetworkCreator(const NetworkCreator&) = delete;
  NetworkCreator& operator=(const NetworkCreator&) = delete;

  void addPopulation(const Population<NeuronModelT>& population) {
    if (population.getFirstId() > networkSize_) {
      std::ostringstream os;
      os << "SZSNN::CONSTRUCTION::NetworkCreator::addPopulation error: adding population `" << population.getLabel()
         << "` neurons must be consistance: networkSize=" << networkSize_ << " firstId=" << population.getFirstId();
      throw SzsnnException(os.str());
    }
    networkSize_ = std::max(networkSize_, population.getLastId() + 1);
    populations_.push_back(population);
    std::stable_sort(populations_.begin(), populations_.end(),
                     [](const Population<NeuronModelT>& a, const Population<NeuronModelT>& b) {
                       return a.getFirstId() < b.getFirstId();
                     });
  }
  void addInputPopulation(const Population<InputModelT>& population) {
    if (population.getFirstId() > inputSize_) {
      std::ostringstream os;
      os << "SZSNN::CONSTRUCTION::NetworkCreator::addInputPopulation error: adding input population `"
         << population.getLabel() << "` inputs must be consistance: inputSize=" << inputSize_
         << " firstId=" << population.getFirstId();
      throw SzsnnException(os.str());
    }
    inputSize_ = std::max(inputSize_, population.getLastId() + 1);
    inputPopulations_.push_back(population);
    std::stable_sort(inputPopulations_.begin(), inputPopulations_.end(),
                     [](const Population<InputModelT>& a, const Population<InputModelT>& b) {
                       return a.getFirstId() < b.getFirstId();
                     });
  }
  size_t addConnection(const Connection<float_type>& connection) {
    connections_.push_back(connection);
    return connections_.size() - 1;
  }
  size_t addInputConnection(const Connection<float_type>& connection) {
    inputConnections_.push_back(connection);
    return inputConnections_.size() - 1;
  }
  // connection_<i>.conn for every neuron projection
  void saveAllConnections() const {
    for (size_t i = 0; i < connections_.size(); ++i)
      connections_[i].saveToFile("connection_" + std::to_string(i) + ".conn");
  }
  void fillNetworkGraph(NetworkGraph<NeuronModelT, InputModelT>& graph) const {
    for (const auto& pop : populations_) {
      size_t id = pop.getFirstId();
      for (const auto& n : pop) graph.addNeuron(id++, n);
    }
    for (const auto& pop : inputPopulations_) {
      size_t id = pop.getFirstId();
      for (const auto& in : pop) graph.addInput(id++, in);
    }
    for (const auto& con : connections_)
      for (const auto& kv : con) graph.addSynapse(synapse(con, kv));
    for (const auto& con : inputConnections_)
      for (const auto& kv : con) graph.addInputSynapse(synapse(con, kv));
  }

 private:
  template <class KV>
  static Synapse<float_type> synapse(const Connection<float_type>& con, const KV& kv) {
    Synapse<float_type> s;
    s.sourceNeuronIdx = con.getFromStartID() + kv.first.first;
    s.targetNeuronIdx = con.getToStartID() + kv.first.second;
    s.weight = kv.second.first;
    s.delay = kv.second.second;
    return s;
  }
  size_t networkSize_ = 0, inputSize_ = 0;
  std::vector<Population<NeuronModelT>> populations_;
  std::vector<Population<InputModelT>> inputPopulations_;
  std::vector<Connection<float_type>> connections_, inputConnections_;
};

}  // namespace construction

// ------------------------------------------------------------------ record
namespace record {

// SpikeRecorder (SpikeRecorder.{hpp,cpp}): spikes of cycles recordStart ..
// recordStop (0: no end) kept in memory; a cycle may be recorded once and a
// neuron once per cycle.  Files: `#spikes by cycles` or `#spikes by
// neurons`, `#record start at: S[   stop at: E]` (always ended by a newline
// here; the reference's only when a stop was given), then `cycle<TAB>neuron`
// lines ordered by cycle or by neuron.
class SpikeRecorder {
 public:
  typedef unsigned cycleNr_t;
  typedef unsigned neuronID_t;
  explicit SpikeRecorder(cycleNr_t recordStart, cycleNr_t recordStop = 0)
      : start_(recordStart), stop_(recordStop), current_(recordStart) {}

  void recordCycleSpikes(const std::vector<neuronID_t>& spikes) { recordCycleSpikes(current_, spikes); }
  void recordCycleSpikes(cycleNr_t cycle, const std::vector<neuronID_t>& spikes) {
    if (cycle < start_ || (stop_ && cycle > stop_)) return;
    if (byCycle_.count(cycle)) {
      std::ostringstream os;
      os << "SZSNN::RECORD::SpikeRecorder::recordCycleSpikes error: cycle `" << cycle << "` already recorder";
      throw SzsnnException(os.str());
    }
    std::set<neuronID_t> ids(spikes.begin(), spikes.end());
    if (ids.size() != spikes.size()) {
      std::ostringstream os;
      os << "SZSNN::RECORD::SpikeRecorder::recordCycleSpikes error: in cycle `" << cycle
         << "` some spikes were duplicated";
      throw SzsnnException(os.str());
    }
    for (neuronID_t n : ids) byNeuron_[n].insert(cycle);
    byCycle_.emplace(cycle, std::move(ids));
    ++current_;
  }
  void saveToFileByCycles(const std::string& filename) const {
    std::ofstream out = open(filename, "saveToFileByCycles");
    out << "#spikes by cycles\n" << header();
    for (const auto& c : byCycle_)
      for (neuronID_t n : c.second) out << c.first << "\t" << n << "\n";
  }
  void saveToFileByNeurons(const std::string& filename) const {
    std::ofstream out = open(filename, "saveToFileByNeurons");
    out << "#spikes by neurons\n" << header();
    for (const auto& n : byNeuron_)
      for (cycleNr_t c : n.second) out << c << "\t" << n.first << "\n";
  }

 private:
  std::string header() const {
    std::ostringstream os;
    os << "#record start at: " << start_;
    if (stop_) os << "   stop at: " << stop_;
    os << "\n";
    return os.str();
  }
  static std::ofstream open(const std::string& filename, const char* what) {
    std::ofstream out(filename);
    if (!out.is_open())
      throw SzsnnException(std::string("SZSNN::RECORD::SpikeRecorder::") + what + "() error: couldn't open file `" +
                           filename + "`");
    return out;
  }
  cycleNr_t start_, stop_, current_;
  std::map<neuronID_t, std::set<cycleNr_t>> byNeuron_;
  std::map<cycleNr_t, std::set<neuronID_t>> byCycle_;
};

}  // namespace record

}  // namespace szsnn
