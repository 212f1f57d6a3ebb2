// Host-side structure and symbolic plan of the batched block LU (built once per network, shared by contexts).
//
// The Jacobian of the stable-row AC system is a 2x2-block matrix over the bus graph: block (i, j) holds
// [dP_i/dphi_j, dP_i/dV_j; dQ_i/dphi_j, dQ_i/dV_j]. Its pattern is structurally symmetric (a branch couples both
// ends both ways), so one fill-reducing order of the bus graph, one elimination tree and one set of block columns
// serve L and U alike. Every scenario shares the plan; only values differ per scenario (outaged branches leave
// zero blocks in the superset pattern, reference docs/gpu-vs-olf.md:14-40).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace olfhip {

// Item record kind of a pair item whose slots are fill-in: the fill writes nothing there, so the factorization
// starts its value from zero instead of loading it.
constexpr int ITEM_FILL_PAIR = -2;

struct Plan {
    int nb = 0;          // buses (= LU nodes)
    int n = 0;           // 2 * nb stable rows
    int nBranch = 0;
    int nShunt = 0;

    // --- order: node k is bus perm[k]; bus b is node pos[b] (minimum degree, renumbered in a postorder of its
    // elimination tree: same fill, subtrees contiguous)
    std::vector<int> perm, pos;

    // --- block LU structure in node numbering
    // S(k): the nodes i > k with a nonzero block L(i,k) (= U(k,i)), ascending.
    std::vector<int> sPtr, sIdx;
    // slots: node k owns [nodeBase[k], nodeBase[k] + 1 + 2|S(k)|): its diagonal block, then for the e-th member
    // i of S(k) the pair L(i,k) at nodeBase[k] + 1 + 2e and U(k,i) right after it.
    std::vector<int> nodeBase;
    int nSlots = 0;

    // --- level schedule (elimination-tree height; leaves are level 0)
    int nLevels = 0;
    std::vector<int> levelNodePtr, levelNodes;     // nodes of each level
    std::vector<int> nodeLevel;                    // level of each node
    std::vector<int> levelPos;                     // position of each node in levelNodes
    // The factors A = L U (L unit block-lower, U with the pivot blocks D(k) on its diagonal) are kept in the form the
    // kernels use: per node k the inverse D(k)^-1 (in the diagonal slot), per pair the unscaled L~(i,k) = L(i,k) D(k)
    // and U(k,i). Every
    // product L(i,j) U(j,k) is then evaluated as L~(i,j) (D(j)^-1 U(j,k)), so a level needs one phase only.
    // A level's items: the diagonal item of node k (D(k) = A(k,k) - sum L~(k,j) D(j)^-1 U(j,k), inverted, and the
    // fused forward solve z(k) = D(k)^-1 (b(k) - sum L~(k,j) z(j))), then the pair items (k, i)
    // (L~(i,k) = A(i,k) - sum L~(i,j) D(j)^-1 U(j,k), U(k,i) = A(k,i) - sum L~(k,j) D(j)^-1 U(j,i)); every
    // contribution list ascends in the contributing node j.
    std::vector<int> levelItemPtr;                 // items of level l in [levelItemPtr[l], levelItemPtr[l+1])
    std::vector<int> itemRec;                      // 4 per item: slot (diag or L), contrib begin, end, node | -1
                                                   // (pair) | ITEM_FILL_PAIR (pair absent from the pattern)
    std::vector<int> itemKI;                       // 2 per item: k and i of L~(i,k) / U(k,i) (diag: k, k)
    std::vector<int> contribRec;                   // 4 per contribution: diag item: L slot of pair (j,k), j,
                                                   // diag slot of j, level of j; pair item: L slot of (j,k), L slot
                                                   // of (j,i), diag slot of j, level of j (U(j,x) sits right after
                                                   // L(x,j))
    std::vector<int> bwdRec;                       // 4 per node in level order: diag slot, S begin, S end, node
    std::vector<int> diagItemOf;                   // diagonal item of each node
    // Threads that share one item's contribution list (splitFor of its length): long lists at the top of the tree
    // are summed by several threads of a wave and reduced across lanes; the kernels cap the split at 64 / T so that
    // an item's threads stay in one wave. Within a level the items are grouped in buckets of one split and one
    // kind, largest split first (diagonal before pair on equal splits): 4 ints per bucket (first item, end, split,
    // 1 for diagonal items), the level's buckets in [levelBucketPtr[l], levelBucketPtr[l+1]). The backward records
    // likewise (split of |S(k)|).
    std::vector<int> bucketRec, levelBucketPtr;
    std::vector<int> bwdBucketRec, levelBwdBucketPtr;

    // --- fill gather tables, per node, half-branches grouped by the far node (ascending)
    std::vector<int> hbPtr;                        // half-branches of node k
    std::vector<int> hbBranch;                     // branch id (for the per-scenario outage lists)
    std::vector<double> hbCoef;                    // 4 per half-branch: A, B, C, psi (see fill)
    std::vector<int> grpPtr;                       // neighbour groups of node k
    std::vector<int> grpEnd;                       // end of the group's half-branches in hbBranch
    std::vector<int> grpFar;                       // far node
    std::vector<int> grpSlot;                      // block slot of (k, far): U(k,far) or L(k,far)
    std::vector<int> fillPtr, fillSlot;            // pairs of node k absent from the pattern (fill-in)
    std::vector<char> fillMask;                    // per slot: 1 in a fill-in pair (never written by the fill)
    std::vector<double> shuntG, shuntB;            // per node, pu
    // per closed branch, for flows at both ends: A1 B C1 psi1 A2 C2 psi2 0 of the own/far form (side 1 at bus1,
    // side 2 at bus2) and the two end nodes
    std::vector<double> branchCoef;
    std::vector<int> branchNode;

    // --- scalar CSR pattern of the stable-row Jacobian (the reference's superset pattern) and its map to slots
    std::vector<int> rowPtr, colIdx;
    std::vector<int> csrSlot;                      // slot * 4 + component of each CSR entry

    // --- statistics
    long long factorMads = 0;                      // 2x2 block multiply-adds of one factorization
    long long contributions = 0;

    // device-independent per-scenario cost, bytes (without the LU value precision factor)
    size_t bytesPerScenario(int luValueBytes) const;
    // The top block for at most maxNodes nodes: the smallest level l0 such that the levels l0.. hold at most
    // maxNodes nodes (an upper set of the elimination tree, so every member's S set stays inside it). Returns l0
    // (nLevels when even the root level does not fit) and the node count in *nodes.
    int topLevel(int maxNodes, int *nodes) const;
    // The items of the top levels l0.. (diagonal items only with diagOnly) ordered by the split of their
    // contributions from below l0, largest first, and their buckets (4 ints: first, end, split, 0 over `items`).
    void topItems(int l0, bool diagOnly, std::vector<int> &items, std::vector<int> &buckets) const;
};

// Threads that share one contribution list of the given length: a power of two, about two contributions each,
// at most 64.
int splitFor(int length);

// Builds the plan from the reference's closed-branch pack (8 doubles, 12 ints per branch) and optional shunts.
// Throws std::invalid_argument on inputs this path does not model.
Plan buildPlan(int n, int nBranch, const double *p8, const int *idx12, int nShunt, const double *shuntP,
               const int *shuntIdx);

// Minimum degree order of a symmetric graph (CSR adjacency), lowest bus index on ties. Also returns, for each
// eliminated node, its neighbour set at elimination (the column structure of the Cholesky factor).
void minimumDegree(int nb, const std::vector<int> &adjPtr, const std::vector<int> &adj, std::vector<int> &perm,
                   std::vector<std::vector<int>> &structureAtElimination);

}  // namespace olfhip
