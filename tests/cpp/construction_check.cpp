// CPU check of the C++ host layer's construction and recording classes
// (szsnn_amd/host/szsnn.hpp): Population files, NetworkCreator filling a
// NetworkGraph, and record::SpikeRecorder files.  No device is touched.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>

#include "../../szsnn_amd/host/szsnn.hpp"

#define CHECK(c, ...)                                              \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "%s:%d: ", __FILE__, __LINE__);         \
      std::fprintf(stderr, __VA_ARGS__);                           \
      std::fprintf(stderr, "\n");                                  \
      std::exit(1);                                                \
    }                                                              \
  } while (0)

template <class F>
static bool throws(F&& f) {
  try {
    f();
  } catch (const szsnn::SzsnnException&) {
    return true;
  }
  return false;
}

static std::string slurp(const std::string& path) {
  std::ifstream in(path);
  std::stringstream ss;
  ss << in.rdbuf();
  return ss.str();
}

int main(int argc, char** argv) {
  using namespace szsnn;
  using namespace szsnn::construction;
  typedef Izhikevich_NeuronModel<float> N;
  typedef Const_rate_InputModel<float> I;
  const std::string dir = argc > 1 ? argv[1] : ".";

  // populations: per-neuron values, file round trip (exact), template file
  Population<N> exc(3, 0, "exc"), inh(2, 3, "inh");
  exc.setNeuronParameters(N::p_A, 0.02f);
  exc.setNeuronParameter(1, N::p_D, 6.125f);
  exc.setNeuronVariables(N::v_V, -65.3f);
  inh.setNeuronParameters(N::p_A, 0.1f);
  exc.saveToFile(dir + "/exc.pop");
  Population<N> back;
  back.loadFromFile(dir + "/exc.pop");
  CHECK(back.getSize() == 3 && back.getFirstId() == 0 && back.getLabel() == "exc", "population header");
  for (size_t i = 0; i < 3; ++i) {
    const N& a = *(exc.begin() + i);
    const N& b = *(back.begin() + i);
    for (unsigned k = 0; k < N::param_count; ++k) CHECK(a.getParam(k) == b.getParam(k), "param %u of %zu", k, i);
    for (unsigned k = 0; k < N::var_count; ++k) CHECK(a.getVar(k) == b.getVar(k), "var %u of %zu", k, i);
  }
  CHECK((back.begin() + 1)->getParam(N::p_D) == 6.125f, "per-neuron parameter");
  inh.saveToTemplateFile(dir + "/inh.tmpl");
  Population<N> tmpl;
  tmpl.loadFromTemplateFile(dir + "/inh.tmpl");
  CHECK(tmpl.getSize() == 2 && tmpl.getFirstId() == 3 && (tmpl.begin() + 1)->getParam(N::p_A) == 0.1f, "template");
  {
    std::ofstream(dir + "/short.pop") << "# two announced, one given\n2\n0\nx\n0.02 0.2 -65 8\n-65 -13\n";
  }
  CHECK(throws([&] { Population<N> p; p.loadFromFile(dir + "/short.pop"); }), "missing neuron must throw");
  // the pyszsnn Brian dumper's orders: template (label, size, first id), with
  // and without a label; population (model name, size, size, first id)
  {
    std::ofstream(dir + "/b1.tmpl") << "# population template saved by pyszsnn.brian\n# Izhikevich\nexc\n3\n5\n"
                                       "0.02 0.2 -65 8\n-65 -13\n";
    std::ofstream(dir + "/b2.tmpl") << "# population template saved by pyszsnn.brian\n# Izhikevich\n\n3\n5\n"
                                       "0.02 0.2 -65 8\n-65 -13\n";
    std::ofstream(dir + "/b.pop") << "# population saved by pyszsnn.brian\nIzhikevich\n2\n2\n7\n0.02 0.2 -65 8\n"
                                     "-65 -13\n0.03 0.2 -65 8\n-64 -13\n";
  }
  Population<N> b1, b2, bp;
  b1.loadFromTemplateFile(dir + "/b1.tmpl");
  b2.loadFromTemplateFile(dir + "/b2.tmpl");
  bp.loadFromFile(dir + "/b.pop");
  CHECK(b1.getSize() == 3 && b1.getFirstId() == 5 && b1.getLabel() == "exc" && (b1.begin() + 2)->getParam(3) == 8.0f,
        "Brian template");
  CHECK(b2.getSize() == 3 && b2.getFirstId() == 5 && b2.getLabel().empty() && b2.begin()->getVar(1) == -13.0f,
        "Brian template without a label");
  CHECK(bp.getSize() == 2 && bp.getFirstId() == 7 && (bp.begin() + 1)->getParam(0) == 0.03f &&
            (bp.begin() + 1)->getVar(0) == -64.0f, "Brian population");
  CHECK(throws([&] { Population<N> p; p.loadFromFile(dir + "/none.pop"); }), "missing file must throw");
  CHECK(throws([] { Population<N> p(0, 0); }), "empty population must throw");

  // inputs without variables (Poisson) read back too
  Population<Poisson_InputModel<float>> pois(4, 0, "drive");
  pois.setNeuronParameters(0, 0.001f);
  pois.saveToFile(dir + "/pois.pop");
  Population<Poisson_InputModel<float>> pois2;
  pois2.loadFromFile(dir + "/pois.pop");
  CHECK(pois2.getSize() == 4 && pois2.begin()->getParam(0) == 0.001f, "population without variables");

  // network creator -> graph
  Population<I> in(2, 0, "input");
  in.setNeuronParameters(I::p_RATE, 3.0f);
  Connection<float> ee(0, 0), ie(3, 0), drive(0, 0);
  ee.addConnection(0, 1, 0.5f, 2);
  ee.addConnection(2, 4, 1.0f, 0);
  ie.addConnection(1, 0, -2.0f, 1);  // neuron 4 -> neuron 0
  drive.addConnection(1, 2, 7.0f, 0);
  NetworkCreator<N, I> nc;
  nc.addPopulation(exc);
  CHECK(throws([&] { nc.addPopulation(Population<N>(1, 9, "gap")); }), "a gap in the ids must throw");
  nc.addPopulation(inh);
  nc.addInputPopulation(in);
  nc.addConnection(ee);
  nc.addConnection(ie);
  nc.addInputConnection(drive);
  NetworkGraph<N, I> g;
  nc.fillNetworkGraph(g);
  g.connectInput();
  CHECK(g.getNetworkSize() == 5 && g.getInputCount() == 2, "sizes %zu %zu", g.getNetworkSize(), g.getInputCount());
  CHECK(g.neurons[1].getParam(N::p_D) == 6.125f && g.neurons[4].getParam(N::p_A) == 0.1f, "neurons in id order");
  CHECK(g.synapses[0].size() == 1 && g.synapses[0][0].targetNeuronIdx == 1 && g.synapses[0][0].delay == 2 &&
            g.synapses[0][0].weight == 0.5f, "synapse 0->1");
  CHECK(g.synapses[4].size() == 1 && g.synapses[4][0].targetNeuronIdx == 0 && g.synapses[4][0].weight == -2.0f,
        "offset projection 4->0");
  CHECK(g.inputSynapses[1].size() == 1 && g.inputSynapses[1][0].targetNeuronIdx == 2 &&
            g.inputSynapses[1][0].weight == 7.0f, "input projection");
  size_t nsyn = 0;
  for (const auto& v : g.synapses) nsyn += v.size();
  CHECK(nsyn == 3, "neuron synapses %zu (input projections must not be added again)", nsyn);
  CHECK(g.getMaxDelay() == 2, "max delay");

  // record::SpikeRecorder
  record::SpikeRecorder r(1, 3);
  r.recordCycleSpikes(0, {5});
  r.recordCycleSpikes(1, {2, 7});
  r.recordCycleSpikes(3, {2});
  r.recordCycleSpikes(4, {9});
  CHECK(throws([&] { r.recordCycleSpikes(1, {4}); }), "a cycle recorded twice must throw");
  CHECK(throws([&] { r.recordCycleSpikes(2, {4, 4}); }), "a duplicated spike must throw");
  r.saveToFileByCycles(dir + "/c.txt");
  r.saveToFileByNeurons(dir + "/n.txt");
  CHECK(slurp(dir + "/c.txt") == "#spikes by cycles\n#record start at: 1   stop at: 3\n1\t2\n1\t7\n3\t2\n", "by cycles");
  CHECK(slurp(dir + "/n.txt") == "#spikes by neurons\n#record start at: 1   stop at: 3\n1\t2\n3\t2\n1\t7\n", "by neurons");
  record::SpikeRecorder seq(0);
  seq.recordCycleSpikes({1});
  seq.recordCycleSpikes({});
  seq.recordCycleSpikes({3});
  seq.saveToFileByCycles(dir + "/seq.txt");
  CHECK(slurp(dir + "/seq.txt") == "#spikes by cycles\n#record start at: 0\n0\t1\n2\t3\n", "implicit cycles");
  std::printf("construction_check OK\n");
  return 0;
}
