// Host symbolic analysis for the multifrontal Cholesky: fill-reducing
// ordering, elimination tree, postorder, column counts, node partition
// (plain columns, fundamental supernodes, relaxed supernodes) and the front
// structure the numeric phase runs on.  Every result matches the reference's
// symbolic phase bit for bit (CSparse 2.2.3 as vendored in cs.cpp, and
// etree.cpp); see the per-function citations.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mdgpu {

// Compressed-sparse-column pattern (+ optional values), int32 indices.
struct Csc {
    int n = 0;
    std::vector<int> p;   // n+1
    std::vector<int> i;   // p[n]
    int nnz() const { return p.empty() ? 0 : p[n]; }
};

// One node of a partition tree (a set of consecutive pivot columns after the
// final permutation).  Mirrors struct ETree (etree.h:6-19) with indices in
// place of pointers.
struct PartNode {
    int first = 0;              // ETree::nodo: first column of the node
    int ncols = 1;              // ETree::cols
    int rows = 0;               // ETree::rows
    int zeros = 0;              // ETree::ceros (relaxed padding)
    std::vector<int> cols;      // ETree::nodos (relaxed: original column ids)
    std::vector<int> kids;      // child node indices into the node list
};

enum Partition { PART_PLAIN = 0, PART_SUPERNODAL = 1, PART_RELAXED = 2 };

struct AnalyzeOptions {
    bool use_amd = true;
    Partition partition = PART_PLAIN;
    int max_zeros = 0;
    bool single_tree = false;   // reject forests as the reference does
};

// Everything the symbolic phase produces.
struct Symbolic {
    int n = 0;
    int nnz_in = 0;                  // entries of the caller's lower triangle
    AnalyzeOptions opts;
    // reference-visible intermediates
    std::vector<int> amd_perm;       // cs_amd result (identity without AMD)
    std::vector<int> parent, post, count;   // on the AMD-permuted matrix
    std::vector<int> relaxed_perm;   // GetRelaxedETree perm (relaxed only)
    std::vector<int> perm;           // final: new column k is original perm[k]
    std::vector<int> pinv;           // inverse of perm
    std::vector<int> count_final;    // (padded) L column counts, final order
    std::vector<int> lp;             // L column pointers = cumsum(count_final)
    Csc a_low;                       // final permuted lower triangle pattern
    std::vector<int> a_src;          // a_low entry -> caller entry index
    std::vector<PartNode> nodes;     // processing (post)order, root last
    // front structure (one front per node, same indexing)
    std::vector<int> front_parent;   // -1 for the root(s)
    std::vector<int64_t> q_off;      // offsets into q_all (nnodes+1)
    std::vector<int> q_all;          // sorted global row ids of each front
    double flops_counts = 0;         // sum_j c_j^2 over unpadded cs_counts
    double flops_padded = 0;         // reference formula sum (c_j^2 + 1), padded
    int64_t lnz = 0;
};

// cs_transpose (cs.cpp:407-429) on patterns.
Csc transpose(const Csc& a);

// cs_amd(1, A) (cs.cpp:645-991) on the upper triangle of a symmetric matrix.
std::vector<int> amd_order(const Csc& upper);

// cs_etree(A, 0) (cs.cpp:480-507), A = full symmetric pattern.
std::vector<int> etree(const Csc& full);
// cs_post (cs.cpp:432-477).
std::vector<int> postorder(const std::vector<int>& parent);
// cs_counts(A, parent, post, 0) (cs.cpp:338-404).
std::vector<int> column_counts(const Csc& full, const std::vector<int>& parent,
                               const std::vector<int>& post);

// Full analysis (mdgpu.cu:1745-1875 driver glue + front row lists).
// Returns an empty string on success, else an error message.
std::string analyze(int n, const int* Cp, const int* Ci, const AnalyzeOptions& opts,
                    Symbolic& out);

// The relaxed analyses of one matrix for several zero budgets, sharing the
// ordering, elimination tree, postorder and column counts: out[k] equals
// analyze() with partition = PART_RELAXED and max_zeros = max_zeros[k].
std::string analyze_budgets(int n, const int* Cp, const int* Ci, const AnalyzeOptions& opts,
                            const std::vector<int>& max_zeros, std::vector<Symbolic>& out);

// Partition builders on a plain (one column per node) postordered tree:
// fundamental supernodes (GetSuperETree, etree.cpp:50-118) and relaxed
// supernodes (GetRelaxedETree, etree.cpp:194-342; pads `count`, returns the
// new permutation).
void supernodal_partition(const std::vector<PartNode>& plain, std::vector<PartNode>& out);
std::string relaxed_partition(const std::vector<PartNode>& plain, int max_zeros, std::vector<int>& count,
                              std::vector<int>& perm, std::vector<PartNode>& out);
// lower triangle -> full symmetric pattern (make_sym, mdgpu.cu:1630-1638)
Csc full_pattern(const Csc& low);

// Analysis for a caller-supplied ordering and partition (the reference's
// Factor() input: A already permuted, postordered node list): identity
// permutation, front row lists derived, L counts taken from the fronts.
std::string analyze_with_nodes(int n, const int* Cp, const int* Ci, const std::vector<PartNode>& nodes,
                               Symbolic& out);

// Sorted set union of row lists (VectorMerge, mdgpu.cu:535-583): append
// the rows of `list` not yet in `rows` (stamp[r] == tag marks a row already
// there); after the last list, sort_union_tail sorts rows[from..).  The
// front row lists of build_fronts are unions of this kind.
void union_rows(std::vector<int>& rows, std::vector<int>& stamp, int tag, const int* list, int64_t len);
void sort_union_tail(std::vector<int>& rows, size_t from);
// Relative map (VectorMergeIndices, mdgpu.cu:585-595): map[t] = position
// of sub[t] in the sorted list super; false if some sub[t] is missing.
bool relative_map(const int* super, int n, const int* sub, int m, int* map);

// Block-diagonal union of independent analyses (a batch of matrices): the
// columns, rows, caller entries and nodes of part k are shifted by the
// totals of parts 0..k-1, so the result is the analysis of the block
// diagonal matrix diag(A_0, A_1, ...) whose caller values are the
// concatenation of the parts' values.  Each part keeps its own ordering and
// partition exactly.  The front tree is a forest.
void merge_symbolic(const std::vector<Symbolic>& parts, Symbolic& out);

}  // namespace mdgpu
