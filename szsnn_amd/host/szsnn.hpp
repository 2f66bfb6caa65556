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
//   SpikeRecorder                   szsnn1.1/src/SpikeRecorder.cpp:13-104
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
#include <iterator>
#include <map>
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

 private:
  void do_step(const firing_input* in, uint64_t* out) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t* ids = (in && !in->empty()) ? in->data() : nullptr;
    const uint32_t n = ids ? (uint32_t)in->size() : 0;
    check(szsnn_step(net_.get(), ids, n, out), ctx_.handle(), "Simulation::step()");
    wall_ms_ += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
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
  unsigned n_neurons_ = 0;
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
    if (cycle < start_ || (stop_ && cycle >= stop_)) return;
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

}  // namespace szsnn
