// Uses usap::Graph_P exactly as uSAP's main.cpp uses sgp::Graph_P
// (main.cpp:48-63): construct from a file prefix, partition, read labels.
// argv: <file prefix> <mode: load|partition> [seed]
#include <cstdio>
#include <cstdlib>
#include <string>

#include "usap/graph_p.hpp"

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  usap::Graph_P<long> g(argv[1]);
  std::printf("nodes %zu edges %zu truth %zu\n", g.num_nodes(), g.num_edges(),
              g.truePartitions.size());
  if (std::string(argv[2]) == "partition") {
    if (argc > 3) g.seed = std::strtoull(argv[3], nullptr, 10);
    g.partition();
    std::printf("B %d S %.6f\n", g.stats.B, g.stats.S);
    for (size_t x : g.get_partitions()) std::printf("%zu\n", x);
  }
  return 0;
}
