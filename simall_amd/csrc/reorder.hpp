// reorder.hpp — bandwidth-reducing reordering of a CSR pattern.
#pragma once

#include <cstdint>
#include <vector>

namespace mi
{

// Reverse Cuthill-McKee ordering of the symmetrised pattern.  perm[new] = old.
// Each connected component starts from a pseudo-peripheral node (George-Liu
// level-structure search); neighbours are visited by increasing degree.
void rcm_order(int n, const int* rp, const int* ci, std::vector<int>& perm);

// Multicolour ordering: greedy colouring of the symmetrised pattern visited
// in RCM order, rows grouped by colour (RCM order kept within a colour), so
// that the rows of one colour are mutually independent.  A triangular sweep
// of the reordered matrix then has one dependency level per colour.
// perm[new] = old; *colors receives the colour count.
void color_order(int n, const int* rp, const int* ci, std::vector<int>& perm, int* colors);

// Symmetric permutation B = P A P^T of a CSR pattern: row i of B is row
// perm[i] of A with columns renumbered by the inverse permutation and sorted.
// vmap[k] is the position in A of B's k-th entry.
void permute_pattern(int n, const int* rp, const int* ci, const std::vector<int>& perm,
                     std::vector<int>& brp, std::vector<int>& bci, std::vector<int>& vmap);

// Half-bandwidth max_i max_j |i - j| over the pattern.
int bandwidth(int n, const int* rp, const int* ci);

} // namespace mi
