// Cluster tree and block tree construction (host).  Compiled with
// -ffp-contract=off so the bounding-box arithmetic rounds like the reference's
// plain double expressions.
#include "tree.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <stdexcept>
#include <unordered_map>
#include <utility>

namespace hmx {

double Cluster::diameter(int dim) const {
  // AxisAlignedBoundingBox::diameterSqr cluster_tree.cpp:236-246
  double r = 0.0;
  for (int i = 0; i < dim; ++i) {
    double d = bbmin[i] - bbmax[i];
    r += d * d;
  }
  return std::sqrt(r);
}

double Cluster::distance_to(const Cluster& o, int dim) const {
  // AxisAlignedBoundingBox::distanceToSqr cluster_tree.cpp:271-285
  double r = 0.0;
  for (int i = 0; i < dim; ++i) {
    double d = std::max(0.0, bbmin[i] - o.bbmax[i]);
    r += d * d;
    d = std::max(0.0, o.bbmin[i] - bbmax[i]);
    r += d * d;
  }
  return std::sqrt(r);
}

namespace {

void compute_bbox(const ClusterTree& t, Cluster& c) {
  const int* idx = t.indices.data() + c.offset;
  const int dim = t.dim;
  for (int d = 0; d < dim; ++d) {
    double v = t.coords[(size_t)idx[0] * dim + d];
    c.bbmin[d] = v;
    c.bbmax[d] = v;
  }
  for (int i = 1; i < c.size; ++i) {
    const double* p = &t.coords[(size_t)idx[i] * dim];
    for (int d = 0; d < dim; ++d) {
      c.bbmin[d] = std::min(c.bbmin[d], p[d]);
      c.bbmax[d] = std::max(c.bbmax[d], p[d]);
    }
  }
}

// AxisAlignClusteringAlgorithm::largestDimension clustering.cpp:136-152
int largest_dimension(const Cluster& c, int dim, int to_avoid, double avoid_ratio = 1.2) {
  std::vector<std::pair<double, int>> sd(dim);
  for (int i = 0; i < dim; ++i) sd[i] = {c.bbmax[i] - c.bbmin[i], i};
  std::sort(sd.begin(), sd.end());
  if (to_avoid < 0 || dim < 2 || sd[dim - 1].second != to_avoid ||
      sd[dim - 1].first > avoid_ratio * sd[dim - 2].first)
    return sd[dim - 1].second;
  return sd[dim - 2].second;
}

Cluster* new_cluster(ClusterTree& t, int offset, int size, int depth, Cluster* father) {
  t.pool.emplace_back();
  Cluster* c = &t.pool.back();
  c->offset = offset;
  c->size = size;
  c->depth = depth;
  c->father = father;
  return c;
}

// ClusterTreeBuilder::divide_recursive clustering.cpp:551-565 with
// MedianBisectionAlgorithm::partition clustering.cpp:258-301 (divider 2, no groups)
void divide(ClusterTree& t, Cluster* c, int current_axis) {
  if (c->size <= t.max_leaf) return;
  compute_bbox(t, *c);
  const int d = largest_dimension(*c, t.dim, current_axis);
  int* idx = t.indices.data() + c->offset;
  const double* X = t.coords.data();
  const int dim = t.dim;
  std::stable_sort(idx, idx + c->size,
                   [X, dim, d](int a, int b) { return X[(size_t)a * dim + d] < X[(size_t)b * dim + d]; });
  const int mid = c->size * 1 / 2;
  if (mid > 0) c->children.push_back(new_cluster(t, c->offset, mid, c->depth + 1, c));
  c->children.push_back(new_cluster(t, c->offset + mid, c->size - mid, c->depth + 1, c));
  for (Cluster* k : c->children) divide(t, k, d);
}

void finish(ClusterTree& t, Cluster* c, int& id) {
  // admissibility prepare: AABB of every node recomputed from its points
  // (admissibility.cpp:56-66)
  compute_bbox(t, *c);
  c->id = id++;
  c->first_leaf = (int)t.leaves.size();
  if (c->is_leaf()) {
    c->leaf_index = (int)t.leaves.size();
    t.leaves.push_back(c);
  }
  for (Cluster* k : c->children) finish(t, k, id);
  c->nleaves = (int)t.leaves.size() - c->first_leaf;
}

}  // namespace

std::unique_ptr<ClusterTree> build_cluster_tree(const double* coords, int dim, int n,
                                                int max_leaf_size) {
  if (dim < 1 || dim > 3) throw std::runtime_error("cluster tree: dimension must be 1..3");
  if (n < 1) throw std::runtime_error("cluster tree: no points");
  auto t = std::make_unique<ClusterTree>();
  t->n = n;
  t->dim = dim;
  t->max_leaf = max_leaf_size;
  t->coords.assign(coords, coords + (size_t)n * dim);
  t->indices.resize(n);
  for (int i = 0; i < n; ++i) t->indices[i] = i;
  t->root = new_cluster(*t, 0, n, 0, nullptr);
  divide(*t, t->root, -1);
  int id = 0;
  finish(*t, t->root, id);
  t->indices_rev.resize(n);
  for (int i = 0; i < n; ++i) t->indices_rev[t->indices[i]] = i;
  return t;
}

std::unique_ptr<ClusterTree> cluster_tree_from(const double* coords, int dim, int n,
                                               const std::vector<int>& indices,
                                               const std::vector<ClusterNodeDesc>& nodes) {
  if (dim < 1 || dim > 3) throw std::runtime_error("cluster tree: dimension must be 1..3");
  if (n < 1 || (int)indices.size() != n || nodes.empty())
    throw std::runtime_error("cluster tree: inconsistent serialized tree");
  auto t = std::make_unique<ClusterTree>();
  t->n = n;
  t->dim = dim;
  t->coords.assign(coords, coords + (size_t)n * dim);
  t->indices = indices;
  size_t at = 0;
  std::function<Cluster*(Cluster*, int)> rec = [&](Cluster* father, int depth) -> Cluster* {
    if (at >= nodes.size()) throw std::runtime_error("cluster tree: truncated serialized tree");
    const ClusterNodeDesc d = nodes[at++];
    if (d.offset < 0 || d.size < 1 || d.offset + d.size > n || d.nchild < 0)
      throw std::runtime_error("cluster tree: bad node in serialized tree");
    Cluster* c = new_cluster(*t, d.offset, d.size, depth, father);
    for (int k = 0; k < d.nchild; ++k) c->children.push_back(rec(c, depth + 1));
    t->max_leaf = std::max(t->max_leaf, c->is_leaf() ? c->size : 0);
    return c;
  };
  t->root = rec(nullptr, 0);
  int id = 0;
  finish(*t, t->root, id);
  t->indices_rev.assign(n, -1);
  for (int i = 0; i < n; ++i) {
    if (indices[i] < 0 || indices[i] >= n || t->indices_rev[indices[i]] >= 0)
      throw std::runtime_error("cluster tree: the serialized indices are not a permutation");
    t->indices_rev[indices[i]] = i;
  }
  return t;
}

namespace {

struct Builder {
  BlockTree* bt;
  AdmissibilityParams p;
  int dim;

  bool is_low_rank(const Cluster* r, const Cluster* c) const {
    // StandardAdmissibilityCondition::isLowRank admissibility.cpp:110-117
    const double dmin = std::min(r->diameter(dim), c->diameter(dim));
    return dmin > 0.0 && dmin <= p.eta * r->distance_to(*c, dim);
  }

  Block* make(const Cluster* r, const Cluster* c, bool sym, int depth) {
    bt->pool.emplace_back();
    Block* b = &bt->pool.back();
    b->rows = r;
    b->cols = c;
    b->depth = depth;
    const bool same = r->offset == c->offset && r->size == c->size;
    const bool low = p.hodlr ? !same : is_low_rank(r, c);
    // the standard condition stops and keeps full below 2 rows or columns; the
    // HODLR condition keeps the base class's defaults (never)
    const bool stop = !p.hodlr && (r->size < 2 || c->size < 2);
    const bool force_rec = (size_t)r->size > p.max_width || (size_t)c->size > p.max_width;
    if ((r->is_leaf() && c->is_leaf()) || stop || (low && !force_rec)) {
      const bool force_full = !p.hodlr && (r->size < 2 || c->size < 2);
      b->kind = (low && !force_full) ? Kind::Rk : Kind::Full;
      return b;
    }
    // AdmissibilityCondition::splitRowsCols admissibility.cpp:39-49
    bool split_r, split_c;
    if (c->size < p.ratio * r->size) {
      split_r = !r->is_leaf();
      split_c = false;
    } else if (r->size < p.ratio * c->size) {
      split_r = false;
      split_c = !c->is_leaf();
    } else {
      split_r = !r->is_leaf();
      split_c = !c->is_leaf();
    }
    b->kind = Kind::Node;
    b->nrow = split_r ? (int)r->children.size() : 1;
    b->ncol = split_c ? (int)c->children.size() : 1;
    b->sym = sym;
    for (int i = 0; i < b->nrow; ++i)
      for (int j = 0; j < b->ncol; ++j) {
        if (!sym || i >= j) {
          const Cluster* rc = split_r ? r->children[i] : r;
          const Cluster* cc = split_c ? c->children[j] : c;
          b->child[i + j * b->nrow] = make(rc, cc, sym && i == j, depth + 1);
        }
      }
    return b;
  }

  void number(Block* b, int& id) {
    b->id = id++;
    if (b->is_leaf()) {
      b->leaf = (int)bt->leaves.size();
      bt->leaves.push_back(b);
      return;
    }
    for (int k = 0; k < b->nrow * b->ncol; ++k)
      if (b->child[k]) number(b->child[k], id);
  }
};

}  // namespace

std::unique_ptr<BlockTree> block_tree_from(const ClusterTree& rows, const ClusterTree& cols,
                                           const std::vector<BlockNodeDesc>& nodes) {
  auto bt = std::make_unique<BlockTree>();
  bt->dim = rows.dim;
  size_t at = 0;
  std::function<Block*(const Cluster*, const Cluster*, int)> rec =
      [&](const Cluster* r, const Cluster* c, int depth) -> Block* {
    if (at >= nodes.size()) throw std::runtime_error("block tree: truncated serialized tree");
    const BlockNodeDesc d = nodes[at++];
    if (d.null) return nullptr;
    bt->pool.emplace_back();
    Block* b = &bt->pool.back();
    b->rows = r;
    b->cols = c;
    b->depth = depth;
    if (d.nchild == 0) {
      if (d.rank == -1)
        b->kind = Kind::Full;
      else if (d.rank >= 0)
        b->kind = Kind::Rk;
      else
        throw std::runtime_error("block tree: uninitialized leaf in serialized tree");
      return b;
    }
    const bool same_r = d.bits & (1 << 4), same_c = d.bits & (1 << 5);
    b->kind = Kind::Node;
    b->nrow = same_r ? 1 : (int)r->children.size();
    b->ncol = same_c ? 1 : (int)c->children.size();
    if (b->nrow * b->ncol > 4 || b->nrow < 1 || b->ncol < 1 || d.nchild > b->nrow * b->ncol)
      throw std::runtime_error("block tree: node shape does not match the cluster trees");
    for (int k = 0; k < d.nchild; ++k) {
      const int i = k % b->nrow, j = k / b->nrow;
      b->child[k] = rec(same_r ? r : r->children[i], same_c ? c : c->children[j], depth + 1);
    }
    // lower-symmetric storage: a square diagonal node without its upper child
    b->sym = r == c && b->nrow == b->ncol && b->nrow > 1 && !b->get(0, 1);
    return b;
  };
  bt->root = rec(rows.root, cols.root, 0);
  if (!bt->root) throw std::runtime_error("block tree: null root in serialized tree");
  if (at != nodes.size()) throw std::runtime_error("block tree: trailing nodes in serialized tree");
  Builder b{bt.get(), AdmissibilityParams{}, rows.dim};
  int id = 0;
  b.number(bt->root, id);
  return bt;
}

namespace {
// pool node of `from` -> the node at the same pool position of `to`
std::unordered_map<const Cluster*, const Cluster*> pool_map(const ClusterTree& from,
                                                            const ClusterTree& to) {
  std::unordered_map<const Cluster*, const Cluster*> m;
  for (size_t k = 0; k < from.pool.size(); ++k) m[&from.pool[k]] = &to.pool[k];
  return m;
}
}  // namespace

std::unique_ptr<ClusterTree> embed_cluster_tree(const ClusterTree& t) {
  auto e = std::make_unique<ClusterTree>();
  e->n = 2 * t.n;
  e->dim = t.dim;
  e->max_leaf = 2 * t.max_leaf;
  e->coords.resize((size_t)e->n * t.dim);
  e->indices.resize(e->n);
  e->indices_rev.resize(e->n);
  for (int u = 0; u < t.n; ++u)
    for (int c = 0; c < 2; ++c) {
      for (int d = 0; d < t.dim; ++d)
        e->coords[(size_t)(2 * u + c) * t.dim + d] = t.coords[(size_t)u * t.dim + d];
      e->indices[2 * u + c] = 2 * t.indices[u] + c;
      e->indices_rev[2 * u + c] = 2 * t.indices_rev[u] + c;
    }
  for (const Cluster& c : t.pool) {
    e->pool.push_back(c);
    e->pool.back().offset = 2 * c.offset;
    e->pool.back().size = 2 * c.size;
  }
  auto m = pool_map(t, *e);
  auto to = [&](const Cluster* c) { return c ? const_cast<Cluster*>(m.at(c)) : nullptr; };
  for (Cluster& c : e->pool) {
    c.father = to(c.father);
    for (Cluster*& ch : c.children) ch = to(ch);
  }
  e->root = to(t.root);
  for (const Cluster* l : t.leaves) e->leaves.push_back(to(l));
  return e;
}

std::unique_ptr<BlockTree> embed_block_tree(const BlockTree& bt, const ClusterTree& rows,
                                            const ClusterTree& er, const ClusterTree& cols,
                                            const ClusterTree& ec) {
  auto mr = pool_map(rows, er);
  auto mc = pool_map(cols, ec);
  auto e = std::make_unique<BlockTree>();
  e->dim = bt.dim;
  std::unordered_map<const Block*, Block*> mb;
  for (const Block& b : bt.pool) {
    e->pool.push_back(b);
    Block& d = e->pool.back();
    d.rows = mr.at(b.rows);
    d.cols = mc.at(b.cols);
    mb[&b] = &d;
  }
  for (Block& d : e->pool)
    for (Block*& ch : d.child) ch = ch ? mb.at(ch) : nullptr;
  e->root = mb.at(bt.root);
  for (const Block* l : bt.leaves) e->leaves.push_back(mb.at(l));
  return e;
}

std::unique_ptr<BlockTree> build_block_tree(const ClusterTree& rows, const ClusterTree& cols,
                                            const AdmissibilityParams& p, bool lower_sym) {
  auto bt = std::make_unique<BlockTree>();
  bt->dim = rows.dim;
  Builder b{bt.get(), p, rows.dim};
  bt->root = b.make(rows.root, cols.root, lower_sym, 0);
  int id = 0;
  b.number(bt->root, id);
  return bt;
}

}  // namespace hmx
