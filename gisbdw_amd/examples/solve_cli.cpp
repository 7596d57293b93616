// solve_cli.cpp -- the reference CLI's dynGPU / bnbGPU paths and its
// highestDegreeHeur / spanningTreeHeur / varianceHeur / bottomUpHeur options
// (Treedepth/main.cpp:178-205) on libtdhip.so through include/tdhip.hpp.
// Reads the reference's graph text ("N;" vertex lines, "a -- b" edges) and
// prints treedepth, root and the decomposition's edges.
//   solve_cli -a dynGPU|bnbGPU|highestDegreeHeur|spanningTreeHeur|varianceHeur|bottomUpHeur
//             <graph file>
#include <fstream>
#include <iostream>
#include <regex>
#include <sstream>
#include <string>

#include "../../include/tdhip.hpp"

int main(int argc, char** argv) {
  if (argc != 4 || std::string(argv[1]) != "-a") {
    std::cerr << "usage: solve_cli -a dynGPU|bnbGPU|highestDegreeHeur|spanningTreeHeur|"
                 "varianceHeur|bottomUpHeur <graph file>\n";
    return 2;
  }
  const std::string alg = argv[2];
  std::ifstream in(argv[3]);
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string text = ss.str();
  // n = max(vertex lines, largest id + 1): trailing isolated vertices kept
  int n = 0;
  std::regex vtx(R"((\d+)\s*;)"), edge(R"((\d+)\s*--\s*(\d+))");
  for (std::sregex_iterator it(text.begin(), text.end(), edge), e; it != e; ++it)
    n = std::max({n, std::stoi((*it)[1]) + 1, std::stoi((*it)[2]) + 1});
  const std::string no_edges = std::regex_replace(text, edge, "");
  for (std::sregex_iterator it(no_edges.begin(), no_edges.end(), vtx), e; it != e; ++it)
    n = std::max(n, std::stoi((*it)[1]) + 1);
  td::Graph g(n);
  for (std::sregex_iterator it(text.begin(), text.end(), edge), e; it != e; ++it) {
    const int a = std::stoi((*it)[1]), b = std::stoi((*it)[2]);
    if (a != b && !(g.adj[a] >> b & 1)) g.AddEdge(a, b);
  }
  try {
    auto ctx = std::make_shared<td::Context>();
    td::Result r;
    bool certified = true;
    if (alg == "dynGPU") {
      td::DynamicGPU dp(ctx);
      dp(g);
      r = dp.GetTDDecomp();
    } else if (alg == "bnbGPU") {
      td::BnBGPU bnb(ctx);
      std::tie(r, certified) = bnb.Solve(g);
    } else if (alg == "highestDegreeHeur" || alg == "spanningTreeHeur" ||
               alg == "varianceHeur") {
      const auto kind = alg == "highestDegreeHeur"  ? td::HeuristicGPU::kHighestDegree
                        : alg == "spanningTreeHeur" ? td::HeuristicGPU::kSpanningTree
                                                    : td::HeuristicGPU::kVariance;
      r = td::HeuristicGPU(kind, ctx).Get(g);
      certified = false;  // an upper bound
    } else if (alg == "bottomUpHeur") {
      r = td::BottomUpHeuristicGPU(ctx).Get(g);
      certified = false;  // an upper bound (UINT_MAX if the levels did not fit)
    } else {
      std::cerr << "Wrong algorithm option specified.\n";
      return 2;
    }
    std::cout << "Vertices: " << g.n << "\nEdges: " << g.NumEdges()
              << "\nTreedepth: " << r.treedepth << "\nRoot: " << r.root
              << "\nCertified: " << (certified ? 1 : 0) << "\n";
    for (auto& e : r.td_decomp) std::cout << e.first << " -- " << e.second << ";\n";
  } catch (const td::Error& e) {
    std::cerr << "error " << e.code << ": " << e.what() << "\n";
    return 1;
  }
  return 0;
}
