// Driver in the style of the reference's main.cpp (szsnn1.1/src/main.cpp:312-391):
// build a NetworkGraph from a .conn file, add Const_rate inputs, step with
// stepAndFired (plus forced input at some steps) and record the spikes.
//   host_api_demo <conn file> <steps> <spike file> [vars file]
// With a vars file: after the run, the neuron variables (getVars), then a
// gather and an update of the next step launched alone (testGather,
// testUpdate: the cycle does not advance) and the variables again.
// Exit codes: 0 ok, 2 usage, 3 SzsnnException (message on stderr).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "../../szsnn_amd/host/szsnn.hpp"

using namespace szsnn;

int main(int argc, char** argv) {
  if (argc != 4 && argc != 5) {
    std::cerr << "usage: host_api_demo <conn file> <steps> <spike file> [vars file]\n";
    return 2;
  }
  try {
    NetworkGraph<Izhikevich_NeuronModel<float>, Const_rate_InputModel<float>> g;
    g.seed = 7;
    for (unsigned i = 0; i < 100; ++i) {
      Izhikevich_NeuronModel<float> n;
      n.setParam(Izhikevich_NeuronModel<float>::p_D, i < 80 ? 8.0f : 2.0f);
      n.setVar(Izhikevich_NeuronModel<float>::v_V, -65.0f);
      n.setVar(Izhikevich_NeuronModel<float>::v_U, -13.0f);
      g.addNeuron(n);
    }
    Connection<float> c(0, 0);
    c.loadFromFile(argv[1]);
    // the fixture's weights are Brian volts; scale to the model's current units
    for (auto& kv : c.connections) kv.second.first *= 5e10f;
    c.applyTo(g);
    for (unsigned j = 0; j < 3; ++j) {
      Const_rate_InputModel<float> in;
      in.setParam(0, (float)(5 + 3 * j));
      const unsigned id = g.addInput(in);
      for (unsigned t = 0; t < 20; ++t) {
        Synapse<float> s;
        s.sourceNeuronIdx = id;
        s.targetNeuronIdx = (j * 31 + t * 7) % 100;
        s.weight = 12.0f;
        s.delay = j;
        g.addInputSynapse(s);
      }
    }
    g.connectInput();
    Context ctx(0);
    Simulation<float> sim(ctx, g);
    SpikeRecorder rec(argv[3]);
    const unsigned steps = (unsigned)std::atoi(argv[2]);
    for (unsigned t = 0; t < steps; ++t) {
      const auto& fired = (t % 25 == 3) ? sim.stepAndFired({1u, 50u, 99u}) : sim.stepAndFired();
      rec.recordCycleSpikes(t, fired);
    }
    if (argc == 5) {
      std::ofstream vf(argv[4]);
      vf << std::setprecision(9);
      auto dump = [&](const std::vector<float>& v) {
        for (size_t i = 0; i < v.size(); ++i) vf << (i ? " " : "") << v[i];
        vf << "\n";
      };
      dump(sim.getVars());
      const double g = sim.testGather(), u = sim.testUpdate();
      dump(sim.getVars());
      const szsnn_counters c = sim.getCounters();
      vf << sim.getNeuronCount() << " " << sim.getVarCount() << " " << c.cycle << " " << c.spikes << "\n";
      std::cout << "kernels gather " << g << " ms update " << u << " ms" << std::endl;
    }
    std::cout << "ok " << ctx.getDeviceName() << " " << sim.getTime() << " ms" << std::endl;
  } catch (const SzsnnException& e) {
    std::cerr << e.what() << std::endl;
    return 3;
  }
  return 0;
}
