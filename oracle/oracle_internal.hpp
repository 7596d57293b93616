// oracle_internal.hpp -- shared pieces of the CPU restatement (test
// infrastructure only).
#ifndef GISBDW_ORACLE_INTERNAL_HPP_
#define GISBDW_ORACLE_INTERNAL_HPP_
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

namespace oracle {

uint64_t NChooseK(uint64_t n, uint64_t k);
uint64_t EncodeSorted(const int32_t* s, int32_t k);
uint64_t EncodeExclude(const int32_t* s, int32_t k, int32_t ex);
void Decode(int32_t nverts, int32_t k, uint64_t code, int32_t* out);
uint64_t EncodeMask(uint64_t mask);
uint64_t ComponentOf(const uint64_t* adj, uint64_t set, int v);
std::vector<uint64_t> Components(const uint64_t* adj, uint64_t set);
int EdgesIn(const uint64_t* adj, uint64_t set);
uint64_t FullMask(int n);

// EliminationTree restatement (Treedepth/src/elimination_tree/).
class ElimTree {
 public:
  struct Comp {
    uint64_t mask;
    int depth;
    int creator;
  };
  struct Result {
    std::vector<int> parent;
    int treedepth;
    int root;
  };
  ElimTree(int n, const uint64_t* adj);
  // Eliminate(v): returns the vertex sets of the new components (ordered by
  // lowest vertex), like the list of ComponentIterator it returns.
  std::vector<uint64_t> Eliminate(int v);
  // Merge(): undoes the last elimination; returns the vertex merged back.
  int Merge();
  Result Decompose() const;
  bool HasComponents() const { return !comps_.empty(); }
  const Comp& FirstComponent() const { return comps_.begin()->second; }
  std::vector<Comp> AllComponents() const {
    std::vector<Comp> v;
    for (auto& p : comps_) v.push_back(p.second);
    return v;
  }
  const std::map<int, Comp>& Map() const { return comps_; }
  const uint64_t* Adj() const { return adj_.data(); }
  int N() const { return n_; }

 private:
  static int Low(uint64_t m) { return __builtin_ctzll(m); }
  int n_;
  std::vector<uint64_t> adj_;
  std::map<int, Comp> comps_;  // key: lowest vertex
  std::vector<int> node_depth_, node_parent_, order_;
  std::vector<std::pair<int, Comp>> stack_;
};

}  // namespace oracle
#endif  // GISBDW_ORACLE_INTERNAL_HPP_
