// Host-side integer topology for the MAS preconditioner: graph partition,
// partition permutation and the multilevel coarse-space hierarchy.
#pragma once
#include <cstdint>
#include <vector>

namespace lss {

// Block adjacency (CSR of nodes; may include self entries).
struct Pattern {
    int32_t n = 0;
    std::vector<int32_t> rowptr;  // n + 1
    std::vector<int32_t> cols;    // ascending per row
};

// Recursive breadth-first bisection into ceil(n / max_part) parts of at most
// max_part nodes, numbered in recursion order (consecutive parts are close at
// every scale).  Replaces KaMinPar (MAS/GraphPartition.cpp:176-289).
void partition_graph(const Pattern &g, int max_part, std::vector<int32_t> &part_of, int32_t &nparts);

// perm[old] = new with nodes sorted by part, stable in node id (MASSolver.cu:90-118).
void partition_permutation(const std::vector<int32_t> &part_of, int32_t nparts, std::vector<int32_t> &perm,
                           std::vector<int32_t> &part_off);

// Relabel a pattern: row perm[i] = sorted perm[cols of row i].
Pattern permute_pattern(const Pattern &g, const std::vector<int32_t> &perm);

struct MasLevel {
    int32_t nodes = 0;
    std::vector<int32_t> bank_off;  // banks of this level (node ranges)
    Pattern op;                     // pattern of this level's operator (level 0: fine operator)
    // Galerkin sources from the level below (level >= 1): slot t of op sums
    // blocks src_blk[src_ptr[t] .. src_ptr[t+1]) of the lower operator.
    std::vector<int64_t> src_ptr;
    std::vector<int32_t> src_blk;
    // restriction from the level below: children of each node, ascending
    std::vector<int64_t> child_ptr;
    std::vector<int32_t> child;
    // map to the next level (empty on the top level)
    std::vector<int32_t> parent;
};

// Multilevel coarse spaces of MAS/MASPreconditioner.cu:128-481: level l+1
// nodes are the connected components of each bank of 32 consecutive level-l
// nodes (level 0 banks = parts), numbered bank by bank and, inside a bank, by
// their smallest member; at most max_coarse coarse levels, stopping when no
// bank merges anything.
void build_mas_levels(const Pattern &fine, const std::vector<int32_t> &part_off, int max_coarse,
                      std::vector<MasLevel> &levels);

}  // namespace lss
