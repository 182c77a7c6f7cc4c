// Host-side structure of the H-matrix: cluster tree (median bisection) and
// block tree (Hackbusch admissibility).  These stay on the CPU and reproduce
// hmat-oss's partition exactly:
//   clustering.cpp:109-114 (stable sort), 136-152 (largestDimension),
//   258-301 (MedianBisectionAlgorithm::partition), 489-565 (build),
//   cluster_tree.cpp:212-289 (AABB), admissibility.cpp:39-117,
//   h_matrix.cpp:144-221 (HMatrix constructor and split).
#pragma once
#include <cstdint>
#include <deque>
#include <memory>
#include <vector>

namespace hmx {

struct Cluster {
  int offset = 0;
  int size = 0;
  int depth = 0;
  int id = -1;            // preorder id
  int leaf_index = -1;    // index among leaves (left to right), -1 for inner nodes
  int first_leaf = 0;     // leaves covered: [first_leaf, first_leaf + nleaves)
  int nleaves = 0;
  std::vector<Cluster*> children;
  Cluster* father = nullptr;
  double bbmin[3] = {0, 0, 0};
  double bbmax[3] = {0, 0, 0};
  bool is_leaf() const { return children.empty(); }
  double diameter(int dim) const;
  double distance_to(const Cluster& o, int dim) const;
};

struct ClusterTree {
  int n = 0;
  int dim = 0;
  int max_leaf = 0;
  std::vector<double> coords;   // user numbering, n * dim
  std::vector<int> indices;     // hmat index -> user index
  std::vector<int> indices_rev; // user index -> hmat index
  std::deque<Cluster> pool;
  Cluster* root = nullptr;
  std::vector<Cluster*> leaves;  // left to right
  int node_count() const { return (int)pool.size(); }
};

// Builds the tree for points given as coords[i * dim + d] (user numbering).
std::unique_ptr<ClusterTree> build_cluster_tree(const double* coords, int dim, int n,
                                                int max_leaf_size);

// A cluster tree from its parts (as serialized, MatrixStructUnmarshaller::
// readClusterTree serialization.cpp:134-160): coordinates in user numbering,
// the hmat -> user permutation, and the nodes in preorder.
struct ClusterNodeDesc {
  int offset, size, nchild;
};
std::unique_ptr<ClusterTree> cluster_tree_from(const double* coords, int dim, int n,
                                               const std::vector<int>& indices,
                                               const std::vector<ClusterNodeDesc>& nodes);

enum class Kind : int8_t { Full = 0, Rk = 1, Node = 2 };

struct Block {
  const Cluster* rows = nullptr;
  const Cluster* cols = nullptr;
  Kind kind = Kind::Node;
  int depth = 0;
  int nrow = 0, ncol = 0;
  bool sym = false;          // isLower: only i >= j children exist
  Block* child[4] = {nullptr, nullptr, nullptr, nullptr};  // index i + j * nrow
  int leaf = -1;             // leaf number (preorder among leaves), -1 for inner nodes
  int id = -1;               // preorder id among all nodes
  Block* get(int i, int j) const { return child[i + j * nrow]; }
  bool is_leaf() const { return kind != Kind::Node; }
  int m() const { return rows->size; }
  int n() const { return cols->size; }
  bool mirrored() const { return rows != cols; }
};

struct AdmissibilityParams {
  double eta = 2.0;
  double ratio = 0.0;
  size_t max_width = (size_t)-1;
  // HODLRAdmissibilityCondition (admissibility.cpp:231-241): every block
  // between two different clusters is low rank, whatever its size
  bool hodlr = false;
};

struct BlockTree {
  std::deque<Block> pool;
  Block* root = nullptr;
  std::vector<Block*> leaves;  // preorder
  int dim = 3;
};

std::unique_ptr<BlockTree> build_block_tree(const ClusterTree& rows, const ClusterTree& cols,
                                            const AdmissibilityParams& p, bool lower_sym);

// A block tree from a serialized H-matrix tree (MatrixStructUnmarshaller::
// readTree serialization.cpp:206-232): nodes in preorder, null children
// included.  rank >= 0: Rk leaf; -1: full leaf; -2: inner node.
struct BlockNodeDesc {
  bool null;
  unsigned char bits;  // isUpper | isLower << 1 | isTriUpper << 2 | isTriLower << 3 |
                       // keepSameRows << 4 | keepSameCols << 5
  int rank;
  int nchild;
};
std::unique_ptr<BlockTree> block_tree_from(const ClusterTree& rows, const ClusterTree& cols,
                                           const std::vector<BlockNodeDesc>& nodes);

// The 2 x 2 real representation of a complex problem: cluster [o, o + s) of
// t becomes [2o, 2o + 2s), and DoF 2u + c (c = 0 real part, 1 imaginary part)
// of the new tree is component c of DoF u.  Coordinates, bounding boxes and
// the node order are those of t, so a complex vector of n entries stored as
// interleaved (re, im) pairs is a real vector in the new tree's numbering.
std::unique_ptr<ClusterTree> embed_cluster_tree(const ClusterTree& t);

// bt (built over rows x cols) over the embedded trees er x ec: the same
// blocks, kinds and leaf order, each block covering twice the rows and columns.
std::unique_ptr<BlockTree> embed_block_tree(const BlockTree& bt, const ClusterTree& rows,
                                            const ClusterTree& er, const ClusterTree& cols,
                                            const ClusterTree& ec);

}  // namespace hmx
