// A driver written against the reference's construction API (populations,
// projections, NetworkCreator, record::SpikeRecorder; the pattern of the
// reference's main.cpp drivers), built on the C++ host layer.  Compiled on
// the host to check the API; run on a device it steps one Izhikevich neuron
// driven by a Const_rate input and prints the recorded spike lines.
#include <iostream>

#include "../../szsnn_amd/host/szsnn.hpp"

int main(int argc, char** argv) {
  typedef float floatT;
  typedef szsnn::Izhikevich_NeuronModel<floatT> neuronT;
  typedef szsnn::Const_rate_InputModel<floatT> inputT;
  try {
    szsnn::construction::Population<neuronT> cells(1, 0, "izh");
    szsnn::construction::Population<inputT> drive(1, 0, "input");
    cells.setNeuronParameters(neuronT::p_A, 0.02f);
    cells.setNeuronParameters(neuronT::p_B, 0.2f);
    cells.setNeuronParameters(neuronT::p_C, -65.0f);
    cells.setNeuronParameters(neuronT::p_D, 8.0f);
    cells.setNeuronVariables(neuronT::v_V, -80.0f);
    drive.setNeuronParameters(inputT::p_RATE, 2.0f);
    drive.setNeuronVariables(inputT::v_LAST, 0.0f);

    szsnn::construction::Connection<floatT> input(0, 0);
    input.addConnection(0, 0, 20.0f, 1);
    szsnn::construction::NetworkCreator<neuronT, inputT> creator;
    creator.addPopulation(cells);
    creator.addInputPopulation(drive);
    creator.addInputConnection(input);

    szsnn::NetworkGraph<neuronT, inputT> graph;
    creator.fillNetworkGraph(graph);
    graph.connectInput();
    std::cout << graph.neuronsToString() << graph.synapsesToString();

    szsnn::Context ctx;
    ctx.setup();
    szsnn::Simulation<floatT> sim(ctx, graph);
    szsnn::record::SpikeRecorder rec(0, 1000);
    const int steps = argc > 1 ? std::atoi(argv[1]) : 100;
    for (int c = 0; c < steps; ++c) rec.recordCycleSpikes(sim.stepAndFired());
    rec.saveToFileByCycles(argc > 2 ? argv[2] : "creator_spikes.txt");
  } catch (const szsnn::SzsnnException& e) {
    std::cerr << e.what() << std::endl;
    return 3;
  }
  return 0;
}
