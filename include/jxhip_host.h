/* jxhip host library (libjxhost.so): the integer schedule work of the fused kernels,
 * native so that a serving loop can prepare requests quickly. Pure host code (g++, no
 * HIP): preparation worker processes load it without touching the GPU.
 *
 * Plain pointers and sizes; every function returns 0 on success or a JXHOST_E* code.
 */
#ifndef JXHIP_HOST_H
#define JXHIP_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JXHOST_OK 0
#define JXHOST_EINVAL 1        /* an index out of range or a size that is not valid */
#define JXHOST_ENOTCONNECTED 2 /* the graph is not connected */
#define JXHOST_ENOTFOREST 3    /* a parent after its child, or a cycle */

/* Breadth-first solve order of one cell with level hops: replaces networkx's
 * `DiGraph.to_undirected` + `bfs_edges` inside `bfs_edge_hops`
 * (jaxley_pytorch/utils/solver_utils.py:187-223, called from modules/base.py:1143-1200).
 *
 * The undirected graph has the directed pairs (src[i], dst[i]) in row order; a node's
 * neighbours are visited in the order networkx's `to_undirected` builds its adjacency:
 * nodes u = 0 .. n-1 in turn, each adding its successors (first occurrence in row order)
 * to its own and to the successor's neighbour list unless already there. Output: the
 * n - 1 tree edges (parent, child) in discovery order, the hop distance of every node
 * from `root`, and the level of every edge: a new level starts when the distance grows
 * or after `allowed_nodes_per_level` edges of one level.
 */
int jxhost_bfs_edge_hops(int64_t n, int64_t n_pairs, const int64_t* src, const int64_t* dst,
                         int64_t root, int64_t allowed_nodes_per_level, int64_t* edge_parent,
                         int64_t* edge_child, int64_t* dist, int64_t* edge_level);

/* Kernel order of a forest whose parents precede their children (no counterpart in the
 * reference: the fused kernels' own schedule). Segments are maximal chains of
 * single-child nodes (top first); a segment's level is 0 if its bottom node has no
 * children, else 1 + the largest level of its child segments; segments are grouped by
 * level, longer first within a level, then by top index.
 * Outputs (caller-sized for the worst case): korder[n] (kernel index -> input index),
 * seg[4 * n] (first node, end node, child_beg, child_end; kernel indices; child ranges
 * into seg_child), seg_level_ptr[n + 2], seg_child[n] (top nodes of the child segments),
 * sizes[3] = {n_seg, L (largest level), n_seg_child}.
 */
int jxhost_segments(int64_t n, const int64_t* parent, int64_t* korder, int32_t* seg,
                    int32_t* seg_level_ptr, int32_t* seg_child, int64_t* sizes);

#ifdef __cplusplus
}
#endif

#endif /* JXHIP_HOST_H */
