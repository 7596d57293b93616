// solve_cli.cpp -- the reference CLI (Treedepth/main.cpp:74-231) on
// libtdhip.so through include/tdhip.hpp:
//
//   solve_cli -a ALGORITHM -o OUTPUT_DIR GRAPH_FILE [GRAPH_FILE ...]
//
// For each graph file (the reference's text: "N;" vertex lines, "a -- b"
// edges) it prints the reference's console lines, then writes
// OUTPUT_DIR/<name>_<epoch><random>.out, the statistics as JSON
// (statistics.hpp:19-30, property-tree layout: 4-space indent, every value a
// string), and <that path>.gviz, the decomposition in graphviz text
// ("graph G {", one "v;" line per vertex, one "a--b ;" line per tree edge).
//
// Algorithms (the reference's -a names, main.cpp:87-98), all on the GPU:
//   dynGPU            subset DP, canonical tree (td_dp_solve)
//   bnbGPU            branch-and-bound seeded with HighestDegree(Variance(
//                     BottomUp)) (main.cpp:180-186; td_bnb_solve), tree too
//   highestDegreeHeur, spanningTreeHeur, varianceHeur, bottomUpHeur
//                     the single heuristics (td_heuristic_batch, td_bottom_up)
//   dynCPU, dynCPUImprov  the reference's CPU subset DPs: same treedepth and
//                     canonical tree, computed here by the GPU subset DP
//   bnbCPU            the reference's CPU branch-and-bound with the seed
//                     HighestDegree(SpanningTree(Variance(BottomUp)))
//                     (main.cpp:166-176): here the GPU search with that seed
// Extension lines on the console: "Root:" and "Certified:" (1 when the
// treedepth is proven, 0 for heuristics and budget-limited searches).
#include <chrono>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <random>
#include <regex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tdhip.hpp"

namespace {
namespace fs = std::filesystem;

void Usage() {
  std::cout << "Usage: solve_cli -a ALGORITHM -o OUTPUT_DIR GRAPH_FILE [GRAPH_FILE ...]\n"
               "ALGORITHM: dynGPU | bnbGPU | highestDegreeHeur | spanningTreeHeur |\n"
               "           varianceHeur | bottomUpHeur | dynCPU | dynCPUImprov | bnbCPU\n"
               "Example usage: ./solve_cli -a bnbGPU -o /path/to/output/dir "
               "/path/to/graph1 /path/to/graph2\n";
  std::exit(1);
}

// "<seconds since epoch><random int>" like GetRandomFilenameSuffix (main.cpp:64-71)
std::string RandomSuffix() {
  const auto epoch = std::chrono::duration_cast<std::chrono::seconds>(
                         std::chrono::system_clock::now().time_since_epoch())
                         .count();
  std::random_device dev;
  std::mt19937 rng(dev());
  std::uniform_int_distribution<int> dist;
  return std::to_string(epoch) + std::to_string(dist(rng));
}

td::Graph ParseGraph(const std::string& text) {
  // n = max(vertex lines, largest id + 1): trailing isolated vertices kept
  int n = 0;
  const std::regex vtx(R"((\d+)\s*;)"), edge(R"((\d+)\s*--\s*(\d+))");
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
  return g;
}

td::Result Solve(const std::string& alg, const td::Graph& g,
                 const std::shared_ptr<td::Context>& ctx, bool* certified) {
  *certified = true;
  if (alg == "dynGPU" || alg == "dynCPU" || alg == "dynCPUImprov") {
    td::DynamicGPU dp(ctx);
    dp(g);
    return dp.GetTDDecomp();
  }
  if (alg == "bnbGPU" || alg == "bnbCPU") {
    const auto seed_kind = alg == "bnbGPU" ? td::HeuristicGPU::kChainBnBGPU
                                           : td::HeuristicGPU::kChainBottomUp;
    const td::Result seed = td::HeuristicGPU(seed_kind, ctx).Get(g);
    td::BnBGPU bnb(ctx);
    auto solved = bnb.Solve(g, seed.treedepth);
    *certified = solved.second;
    return solved.first;
  }
  *certified = false;  // upper bounds
  if (alg == "bottomUpHeur") return td::BottomUpHeuristicGPU(ctx).Get(g);
  if (alg == "highestDegreeHeur") return td::HeuristicGPU(td::HeuristicGPU::kHighestDegree, ctx).Get(g);
  if (alg == "spanningTreeHeur") return td::HeuristicGPU(td::HeuristicGPU::kSpanningTree, ctx).Get(g);
  if (alg == "varianceHeur") return td::HeuristicGPU(td::HeuristicGPU::kVariance, ctx).Get(g);
  std::cerr << "Wrong algorithm option specified.\n";
  Usage();
  return {};
}

}  // namespace

int main(int argc, char** argv) {
  std::string alg, out_dir;
  std::vector<std::string> inputs;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if ((a == "-a" || a == "--algorithm") && i + 1 < argc) {
      alg = argv[++i];
    } else if ((a == "-o" || a == "--output") && i + 1 < argc) {
      out_dir = argv[++i];
    } else if ((a == "-i" || a == "--input") && i + 1 < argc) {
      inputs.push_back(argv[++i]);
    } else if (a == "--help") {
      Usage();
    } else {
      inputs.push_back(a);
    }
  }
  if (alg.empty() || out_dir.empty() || inputs.empty()) Usage();
  if (!fs::is_directory(out_dir)) {
    std::cerr << out_dir << " is not a directory.\n";
    Usage();
  }
  for (const auto& p : inputs)
    if (!fs::is_regular_file(p)) {
      std::cerr << p << " is not a regular file.\n";
      Usage();
    }
  try {
    auto ctx = std::make_shared<td::Context>();
    for (const auto& path : inputs) {
      std::ifstream in(path);
      std::stringstream ss;
      ss << in.rdbuf();
      const td::Graph g = ParseGraph(ss.str());
      fs::path name = fs::path(path);
      while (name != name.stem()) name = name.stem();
      std::cout << "Processing graph " << name.string() << "\nAlgorithm: " << alg
                << "\nVertices: " << g.n << "\nEdges: " << g.NumEdges() << std::endl;
      const auto t1 = std::chrono::high_resolution_clock::now();
      bool certified = true;
      const td::Result r = Solve(alg, g, ctx, &certified);
      const auto t2 = std::chrono::high_resolution_clock::now();
      const double secs =
          std::chrono::duration_cast<std::chrono::milliseconds>(t2 - t1).count() / 1000.0;
      fs::path stats_path = fs::path(out_dir) / name;
      stats_path += "_" + RandomSuffix() + ".out";
      std::cout << "Elapsed time: " << secs << " seconds\nTreedepth: " << r.treedepth
                << "\nRoot: " << r.root << "\nCertified: " << (certified ? 1 : 0)
                << "\nOutput written to: " << stats_path << "\n\n";
      std::ofstream js(stats_path);
      js << "{\n"
         << "    \"algorithmType\": \"" << alg << "\",\n"
         << "    \"graphName\": \"" << name.string() << "\",\n"
         << "    \"timeElapsed\": \"" << secs << "\",\n"
         << "    \"treedepth\": \"" << r.treedepth << "\",\n"
         << "    \"root\": \"" << r.root << "\",\n"
         << "    \"edges\": \"" << g.NumEdges() << "\",\n"
         << "    \"vertices\": \"" << g.n << "\"\n"
         << "}\n";
      fs::path gviz = stats_path;
      gviz += ".gviz";
      std::ofstream gv(gviz);
      gv << "graph G {\n";
      for (std::size_t v = 0; v < r.parent.size(); ++v) gv << v << ";\n";
      for (const auto& e : r.td_decomp) gv << e.first << "--" << e.second << " ;\n";
      gv << "}\n";
    }
  } catch (const td::Error& e) {
    std::cerr << "error " << e.code << ": " << e.what() << "\n";
    return 1;
  }
  return 0;
}
