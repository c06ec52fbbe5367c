// tree.hpp -- octree build + interaction-list walk (see tree.hip).
#pragma once
#include "engine.hpp"

namespace gplum {

// Morton keys of d_pos[0..n) sorted; t.perm receives the sorted->input permutation.
int tree_sort_keys(Tree &t, const double *d_pos, int n, hipStream_t st);
// Nodes, moments and i-groups over particles already in sorted order, launched without
// a wait (levels up to one past the previous tree's depth; the arrays stay referenced
// until tree_walk returns).
int tree_build_nodes(Tree &t, const gplum_params &p, const double *d_pos, const double *d_mass,
                     const double *d_rsearch, int n, hipStream_t st);
// EP (near/far) and SP lists of every i-group; reads the level ranges back with its
// results, reports a node overflow and finishes a tree deeper than the levels launched.
int tree_walk(Tree &t, const gplum_params &p, const double *d_pos, const double *d_rsearch,
              hipStream_t st);

}  // namespace gplum
