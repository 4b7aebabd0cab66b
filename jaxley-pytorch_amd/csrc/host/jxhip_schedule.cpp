// Host-side schedule construction of the fused kernels (libjxhost.so): the breadth-first
// DHS solve order with level hops and the segment kernel order. Integer work, O(n).
#include "jxhip_host.h"

#include <algorithm>
#include <cstdint>
#include <vector>

namespace {

// Adjacency in insertion order: per node a short list (tree nodes have few neighbours);
// a duplicate insertion keeps the first position, as a Python dict does.
struct OrderedAdj {
  std::vector<std::vector<int64_t>> nb;
  explicit OrderedAdj(int64_t n) : nb(static_cast<size_t>(n)) {}
  void add(int64_t u, int64_t v) {
    auto& l = nb[static_cast<size_t>(u)];
    if (std::find(l.begin(), l.end(), v) == l.end()) l.push_back(v);
  }
};

}  // namespace

extern "C" int jxhost_bfs_edge_hops(int64_t n, int64_t n_pairs, const int64_t* src,
                                    const int64_t* dst, int64_t root,
                                    int64_t allowed_nodes_per_level, int64_t* edge_parent,
                                    int64_t* edge_child, int64_t* dist, int64_t* edge_level) {
  if (n <= 0 || n_pairs < 0 || root < 0 || root >= n || allowed_nodes_per_level < 1)
    return JXHOST_EINVAL;
  for (int64_t i = 0; i < n_pairs; ++i) {
    if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n) return JXHOST_EINVAL;
  }
  // Successors of every node: first occurrence in row order.
  OrderedAdj succ(n);
  for (int64_t i = 0; i < n_pairs; ++i) succ.add(src[i], dst[i]);
  // Undirected adjacency built node by node (networkx's to_undirected order).
  OrderedAdj adj(n);
  for (int64_t u = 0; u < n; ++u) {
    for (int64_t v : succ.nb[static_cast<size_t>(u)]) {
      adj.add(u, v);
      adj.add(v, u);
    }
  }
  for (int64_t i = 0; i < n; ++i) dist[i] = -1;
  dist[root] = 0;
  std::vector<int64_t> queue;
  queue.reserve(static_cast<size_t>(n));
  queue.push_back(root);
  int64_t ne = 0;
  for (size_t h = 0; h < queue.size(); ++h) {
    const int64_t u = queue[h];
    for (int64_t v : adj.nb[static_cast<size_t>(u)]) {
      if (dist[v] >= 0) continue;
      dist[v] = dist[u] + 1;
      edge_parent[ne] = u;
      edge_child[ne] = v;
      ++ne;
      queue.push_back(v);
    }
  }
  if (ne != n - 1) return JXHOST_ENOTCONNECTED;
  // Levels: a new one when the hop distance grows or the current one is full.
  int64_t depth = 0, since = 0, last = 0;
  for (int64_t i = 0; i < ne; ++i) {
    const int64_t d = dist[edge_child[i]];
    if (d > last) {
      ++depth;
      since = 0;
      last = d;
    } else if (since >= allowed_nodes_per_level) {
      ++depth;
      since = 0;
    }
    ++since;
    edge_level[i] = depth;
  }
  return JXHOST_OK;
}

extern "C" int jxhost_segments(int64_t n, const int64_t* parent, int64_t* korder, int32_t* seg,
                               int32_t* seg_level_ptr, int32_t* seg_child, int64_t* sizes) {
  if (n < 0) return JXHOST_EINVAL;
  for (int64_t i = 0; i < n; ++i) {
    if (parent[i] >= i || parent[i] < -1) return JXHOST_ENOTFOREST;
  }
  // Children of every node in increasing index (CSR).
  std::vector<int64_t> nch(static_cast<size_t>(n), 0), cptr(static_cast<size_t>(n) + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (parent[i] >= 0) ++nch[static_cast<size_t>(parent[i])];
  }
  for (int64_t i = 0; i < n; ++i) cptr[i + 1] = cptr[i] + nch[i];
  std::vector<int64_t> kids(static_cast<size_t>(cptr[n])), fill(cptr.begin(), cptr.end() - 1);
  for (int64_t i = 0; i < n; ++i) {
    if (parent[i] >= 0) kids[static_cast<size_t>(fill[parent[i]]++)] = i;
  }
  // Segments in increasing top index: a top is a root or a child of a branching node.
  std::vector<int64_t> seg_top, seg_len, seg_of(static_cast<size_t>(n), -1);
  std::vector<int64_t> chain_nodes;  // chains back to back, in segment order
  std::vector<int64_t> chain_beg;
  for (int64_t t = 0; t < n; ++t) {
    const bool top = parent[t] < 0 || nch[parent[t]] != 1;
    if (!top) continue;
    const int64_t q = static_cast<int64_t>(seg_top.size());
    chain_beg.push_back(static_cast<int64_t>(chain_nodes.size()));
    int64_t x = t;
    int64_t len = 0;
    while (true) {
      chain_nodes.push_back(x);
      seg_of[x] = q;
      ++len;
      if (nch[x] != 1) break;
      x = kids[cptr[x]];
    }
    seg_top.push_back(t);
    seg_len.push_back(len);
  }
  const int64_t nseg = static_cast<int64_t>(seg_top.size());
  chain_beg.push_back(static_cast<int64_t>(chain_nodes.size()));
  // Levels, children before parents (decreasing top index).
  std::vector<int64_t> level(static_cast<size_t>(nseg), 0);
  for (int64_t q = nseg - 1; q >= 0; --q) {
    const int64_t bot = chain_nodes[chain_beg[q + 1] - 1];
    if (nch[bot] < 2) continue;
    int64_t lv = 0;
    for (int64_t j = cptr[bot]; j < cptr[bot + 1]; ++j) lv = std::max(lv, level[seg_of[kids[j]]]);
    level[q] = 1 + lv;
  }
  // Order: level, then longer first, then top index.
  std::vector<int64_t> order(static_cast<size_t>(nseg));
  for (int64_t q = 0; q < nseg; ++q) order[q] = q;
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    if (level[a] != level[b]) return level[a] < level[b];
    if (seg_len[a] != seg_len[b]) return seg_len[a] > seg_len[b];
    return seg_top[a] < seg_top[b];
  });
  std::vector<int64_t> kidx(static_cast<size_t>(n));
  int64_t pos = 0;
  for (int64_t r = 0; r < nseg; ++r) {
    const int64_t q = order[r];
    for (int64_t j = chain_beg[q]; j < chain_beg[q + 1]; ++j) {
      korder[pos] = chain_nodes[j];
      kidx[chain_nodes[j]] = pos;
      ++pos;
    }
  }
  int64_t nsc = 0, L = 0;
  for (int64_t r = 0; r < nseg; ++r) {
    const int64_t q = order[r];
    const int64_t bot = chain_nodes[chain_beg[q + 1] - 1];
    seg[4 * r + 0] = static_cast<int32_t>(kidx[seg_top[q]]);
    seg[4 * r + 1] = static_cast<int32_t>(kidx[seg_top[q]] + seg_len[q]);
    seg[4 * r + 2] = static_cast<int32_t>(nsc);
    if (nch[bot] >= 2) {
      for (int64_t j = cptr[bot]; j < cptr[bot + 1]; ++j) seg_child[nsc++] = static_cast<int32_t>(kidx[kids[j]]);
    }
    seg[4 * r + 3] = static_cast<int32_t>(nsc);
    L = std::max(L, level[q]);
  }
  // Level pointers: segments of level l are [ptr[l], ptr[l + 1]).
  for (int64_t l = 0; l <= L + 1; ++l) seg_level_ptr[l] = 0;
  for (int64_t q = 0; q < nseg; ++q) ++seg_level_ptr[level[q] + 1];
  for (int64_t l = 0; l <= L; ++l) seg_level_ptr[l + 1] += seg_level_ptr[l];
  sizes[0] = nseg;
  sizes[1] = L;
  sizes[2] = nsc;
  return JXHOST_OK;
}
