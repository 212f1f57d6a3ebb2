// Host-side structure and symbolic plan of the batched block LU. See plan.hpp.
// host-only use of the generated term functions
#define OLFHIP_GEN_FN inline
#include "generated/closed_branch_terms.hpp"
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <queue>
#include <stdexcept>
#include <string>
#include <utility>

#include "errors.hpp"

namespace olfhip {

void minimumDegree(int nb, const std::vector<int> &adjPtr, const std::vector<int> &adj, std::vector<int> &perm,
                   std::vector<std::vector<int>> &structureAtElimination) {
    std::vector<std::vector<int>> g(nb);
    for (int i = 0; i < nb; i++) {
        g[i].assign(adj.begin() + adjPtr[i], adj.begin() + adjPtr[i + 1]);
    }
    using Entry = std::pair<int, int>;  // (degree, bus): lowest degree, then lowest bus
    std::priority_queue<Entry, std::vector<Entry>, std::greater<Entry>> heap;
    for (int i = 0; i < nb; i++) heap.push({(int)g[i].size(), i});
    std::vector<char> dead(nb, 0);
    perm.assign(nb, -1);
    structureAtElimination.assign(nb, {});
    std::vector<int> merged;
    for (int step = 0; step < nb; step++) {
        int v = -1;
        while (!heap.empty()) {
            Entry top = heap.top();
            heap.pop();
            if (!dead[top.second] && top.first == (int)g[top.second].size()) {
                v = top.second;
                break;
            }
        }
        if (v < 0) throw std::logic_error("minimum degree: heap exhausted");
        perm[step] = v;
        dead[v] = 1;
        const std::vector<int> nbr = g[v];
        structureAtElimination[v] = nbr;
        for (int u : nbr) {
            merged.clear();
            std::set_union(g[u].begin(), g[u].end(), nbr.begin(), nbr.end(), std::back_inserter(merged));
            g[u].clear();
            for (int c : merged)
                if (c != u && c != v) g[u].push_back(c);
            heap.push({(int)g[u].size(), u});
        }
        std::vector<int>().swap(g[v]);
    }
}

// Renumbers an elimination order into a postorder of its elimination tree (children visited in ascending order, each
// subtree contiguous and ending at its root). Every topological order of the tree eliminates with the same fill and
// the same tree, so only storage positions change: the nodes of a subtree, their factor columns and their vector
// entries sit together, and the contributions an item gathers from its descendants fall on fewer cache lines.
static void postorderElimination(std::vector<int> &perm, const std::vector<std::vector<int>> &structureAtElimination) {
    const int nb = (int)perm.size();
    std::vector<int> pos(nb), parent(nb, -1);
    for (int k = 0; k < nb; k++) pos[perm[k]] = k;
    for (int k = 0; k < nb; k++)
        for (int b : structureAtElimination[perm[k]])
            if (parent[k] < 0 || pos[b] < parent[k]) parent[k] = pos[b];
    std::vector<int> childPtr(nb + 1, 0), child(nb);
    for (int k = 0; k < nb; k++)
        if (parent[k] >= 0) childPtr[parent[k] + 1]++;
    for (int k = 0; k < nb; k++) childPtr[k + 1] += childPtr[k];
    std::vector<int> fillAt(childPtr.begin(), childPtr.end() - 1);
    for (int k = 0; k < nb; k++)
        if (parent[k] >= 0) child[fillAt[parent[k]]++] = k;
    std::vector<int> out;
    out.reserve(nb);
    std::vector<std::pair<int, int>> stack;   // (node, next child position)
    for (int r = 0; r < nb; r++) {
        if (parent[r] >= 0) continue;
        stack.push_back({r, childPtr[r]});
        while (!stack.empty()) {
            std::pair<int, int> &top = stack.back();
            if (top.second < childPtr[top.first + 1]) {
                const int c = child[top.second++];
                stack.push_back({c, childPtr[c]});
            } else {
                out.push_back(top.first);
                stack.pop_back();
            }
        }
    }
    if ((int)out.size() != nb) throw std::logic_error("elimination tree postorder: not a forest");
    std::vector<int> renumbered(nb);
    for (int q = 0; q < nb; q++) renumbered[q] = perm[out[q]];
    perm.swap(renumbered);
}

// Threads that share one contribution list: about two contributions per thread, a power of two, at most 64.
int splitFor(int length) {
    int s = 1;
    while (s < 64 && 2 * s < length) s *= 2;
    return s;
}

static int indexIn(const std::vector<int> &idx, int lo, int hi, int value) {
    auto it = std::lower_bound(idx.begin() + lo, idx.begin() + hi, value);
    if (it == idx.begin() + hi || *it != value) return -1;
    return (int)(it - (idx.begin() + lo));
}

Plan buildPlan(int n, int nBranch, const double *p8, const int *idx12, int nShunt, const double *shuntP,
               const int *shuntIdx) {
    if (n <= 0 || (n & 1)) throw ArgumentError("n must be a positive even row count (2 rows per bus)");
    if (nBranch < 0 || nShunt < 0) throw ArgumentError("negative element count");
    if (nBranch > 0 && (!p8 || !idx12)) throw ArgumentError("closed-branch arrays are null");
    if (nShunt > 0 && (!shuntP || !shuntIdx)) throw ArgumentError("shunt arrays are null");
    Plan P;
    P.n = n;
    P.nb = n / 2;
    P.nBranch = nBranch;
    P.nShunt = nShunt;
    const int nb = P.nb;

    // --- validate the stable-row packing and read the endpoints
    std::vector<int> bus1(nBranch), bus2(nBranch);
    for (int e = 0; e < nBranch; e++) {
        const int *ix = idx12 + 12 * e;
        int a = ix[0] / 2, c = ix[2] / 2;
        bool ok = ix[0] >= 0 && ix[2] >= 0 && ix[0] % 2 == 0 && ix[2] % 2 == 0 && a < nb && c < nb && a != c &&
                  ix[1] == ix[0] + 1 && ix[3] == ix[2] + 1 && ix[4] == ix[0] && ix[5] == ix[1] && ix[6] == ix[2] &&
                  ix[7] == ix[3];
        if (!ok)
            throw UnsupportedError("closed branch " + std::to_string(e) +
                                   ": rows/columns are not stable bus rows (P on phi row 2i, Q/V on row 2i+1)");
        if (ix[8] != -1 || ix[9] != -1 || ix[10] != -1 || ix[11] != -1)
            throw UnsupportedError("closed branch " + std::to_string(e) +
                                   ": phase/ratio control columns and relocated reactive rows are not modelled");
        for (int k = 0; k < 8; k++)
            if (!std::isfinite(p8[8 * e + k])) throw ArgumentError("closed branch " + std::to_string(e) + ": non-finite parameter");
        bus1[e] = a;
        bus2[e] = c;
    }
    P.shuntG.assign(nb, 0.0);
    P.shuntB.assign(nb, 0.0);
    std::vector<double> gBus(nb, 0.0), bBus(nb, 0.0);
    for (int s = 0; s < nShunt; s++) {
        const int *ix = shuntIdx + 3 * s;
        int b = ix[0] / 2;
        if (ix[0] < 0 || ix[0] % 2 || b >= nb || ix[1] != ix[0] + 1 || ix[2] != ix[1])
            throw UnsupportedError("shunt " + std::to_string(s) + ": not on its bus's stable rows");
        gBus[b] += shuntP[2 * s];
        bBus[b] += shuntP[2 * s + 1];
    }

    // --- bus adjacency (unique, sorted)
    std::vector<int> deg(nb + 1, 0);
    for (int e = 0; e < nBranch; e++) {
        deg[bus1[e]]++;
        deg[bus2[e]]++;
    }
    std::vector<std::vector<int>> nbrs(nb);
    for (int i = 0; i < nb; i++) nbrs[i].reserve(deg[i]);
    for (int e = 0; e < nBranch; e++) {
        nbrs[bus1[e]].push_back(bus2[e]);
        nbrs[bus2[e]].push_back(bus1[e]);
    }
    std::vector<int> adjPtr(nb + 1, 0), adj;
    for (int i = 0; i < nb; i++) {
        std::sort(nbrs[i].begin(), nbrs[i].end());
        nbrs[i].erase(std::unique(nbrs[i].begin(), nbrs[i].end()), nbrs[i].end());
        adj.insert(adj.end(), nbrs[i].begin(), nbrs[i].end());
        adjPtr[i + 1] = (int)adj.size();
    }

    // --- fill-reducing order and the symbolic structure
    std::vector<std::vector<int>> sAt;
    minimumDegree(nb, adjPtr, adj, P.perm, sAt);
    postorderElimination(P.perm, sAt);
    P.pos.assign(nb, 0);
    for (int k = 0; k < nb; k++) P.pos[P.perm[k]] = k;
    P.sPtr.assign(nb + 1, 0);
    for (int k = 0; k < nb; k++) {
        std::vector<int> s;
        for (int b : sAt[P.perm[k]]) s.push_back(P.pos[b]);
        std::sort(s.begin(), s.end());
        for (int i : s)
            if (i <= k) throw std::logic_error("symbolic structure: non-ascending member");
        P.sIdx.insert(P.sIdx.end(), s.begin(), s.end());
        P.sPtr[k + 1] = (int)P.sIdx.size();
    }
    P.nodeBase.assign(nb + 1, 0);
    for (int k = 0; k < nb; k++) P.nodeBase[k + 1] = P.nodeBase[k] + 1 + 2 * (P.sPtr[k + 1] - P.sPtr[k]);
    P.nSlots = P.nodeBase[nb];
    auto lslot = [&](int j, int e) { return P.nodeBase[j] + 1 + 2 * e; };

    // --- elimination tree levels
    std::vector<int> level(nb, 0);
    for (int k = 0; k < nb; k++) {
        if (P.sPtr[k + 1] > P.sPtr[k]) {
            int parent = P.sIdx[P.sPtr[k]];
            level[parent] = std::max(level[parent], level[k] + 1);
        }
    }
    P.nLevels = 0;
    for (int k = 0; k < nb; k++) P.nLevels = std::max(P.nLevels, level[k] + 1);

    // --- contributions (ascending contributing node j)
    std::vector<int> pairItemBase(nb + 1, 0);
    for (int k = 0; k < nb; k++) pairItemBase[k + 1] = pairItemBase[k] + (P.sPtr[k + 1] - P.sPtr[k]);
    std::vector<std::vector<std::pair<int, int>>> diagC(nb), pairC(pairItemBase[nb]);
    for (int j = 0; j < nb; j++) {
        const int lo = P.sPtr[j], hi = P.sPtr[j + 1];
        for (int ea = 0; ea < hi - lo; ea++) {
            const int a = P.sIdx[lo + ea];
            diagC[a].push_back({lslot(j, ea), j});
            for (int eb = ea + 1; eb < hi - lo; eb++) {
                const int b = P.sIdx[lo + eb];
                int e = indexIn(P.sIdx, P.sPtr[a], P.sPtr[a + 1], b);
                if (e < 0) throw std::logic_error("symbolic structure is not closed under elimination");
                pairC[pairItemBase[a] + e].push_back({lslot(j, ea), lslot(j, eb)});
            }
        }
    }
    // level-ordered packed records: a level's items (one diagonal item per node, then its pair items) are grouped
    // into buckets of one thread split and one kind, largest split first and diagonal before pair on equal splits
    // (so that every group of threads sharing an item stays aligned inside one wave when buckets are laid end to
    // end, and so is any subsequence of them, such as the diagonal buckets alone)
    std::vector<std::vector<int>> byLevel(P.nLevels);
    for (int k = 0; k < nb; k++) byLevel[level[k]].push_back(k);
    P.levelNodePtr.assign(P.nLevels + 1, 0);
    P.levelItemPtr.assign(P.nLevels + 1, 0);
    P.levelBucketPtr.assign(P.nLevels + 1, 0);
    P.levelBwdBucketPtr.assign(P.nLevels + 1, 0);
    auto pushContrib = [&](int a, int b, int d, int j) {
        P.contribRec.push_back(a);
        P.contribRec.push_back(b);
        P.contribRec.push_back(d);
        P.contribRec.push_back(level[j]);
    };
    P.nodeLevel = level;
    P.levelPos.assign(nb, 0);
    P.diagItemOf.assign(nb, 0);
    struct Item { int slot, c0, c1, node, k, i; };
    auto splitOf = [](const Item &u) { return splitFor(u.c1 - u.c0); };
    auto emit = [&](std::vector<Item> &items) {
        std::stable_sort(items.begin(), items.end(), [&](const Item &u, const Item &v) {
            const int su = splitOf(u), sv = splitOf(v);
            if (su != sv) return su > sv;
            return (u.node >= 0) > (v.node >= 0);
        });
        for (size_t q = 0; q < items.size(); q++) {
            const Item &it = items[q];
            const int idx = (int)(P.itemRec.size() / 4);
            const int diag = it.node >= 0 ? 1 : 0;
            if (q == 0 || splitOf(it) != splitOf(items[q - 1]) || diag != (items[q - 1].node >= 0 ? 1 : 0))
                P.bucketRec.insert(P.bucketRec.end(), {idx, idx, splitOf(it), diag});
            P.bucketRec[P.bucketRec.size() - 3]++;   // the open bucket's end
            P.itemRec.insert(P.itemRec.end(), {it.slot, it.c0, it.c1, it.node});
            P.itemKI.insert(P.itemKI.end(), {it.k, it.i});
            if (diag) P.diagItemOf[it.k] = idx;
        }
    };
    int nContrib = 0;
    std::vector<Item> diagItems, pairItems;
    std::vector<std::pair<int, int>> bwd;   // (split, node)
    for (int l = 0; l < P.nLevels; l++) {
        diagItems.clear();
        pairItems.clear();
        bwd.clear();
        for (int k : byLevel[l]) {
            P.levelPos[k] = (int)P.levelNodes.size();
            P.levelNodes.push_back(k);
            const int c0 = nContrib;
            for (auto &c : diagC[k]) {
                pushContrib(c.first, c.second, P.nodeBase[c.second], c.second);
                nContrib++;
            }
            diagItems.push_back({P.nodeBase[k], c0, nContrib, k, k, k});
            bwd.push_back({splitFor(P.sPtr[k + 1] - P.sPtr[k]), k});
            P.contributions += (long long)diagC[k].size();
            P.factorMads += 2LL * (long long)diagC[k].size() + 1;   // W = D^-1 U, D -= L W; the inverse
        }
        for (int k : byLevel[l]) {
            for (int e = 0; e < P.sPtr[k + 1] - P.sPtr[k]; e++) {
                const auto &lst = pairC[pairItemBase[k] + e];
                if (lst.empty()) continue;   // nothing to subtract: the fill's values are final
                const int c0 = nContrib;
                for (auto &c : lst) {
                    // contributing node j owns both slots: recover it from the slot range
                    const int j = (int)(std::upper_bound(P.nodeBase.begin(), P.nodeBase.end(), c.first) -
                                        P.nodeBase.begin()) - 1;
                    pushContrib(c.first, c.second, P.nodeBase[j], j);
                    nContrib++;
                }
                pairItems.push_back({lslot(k, e), c0, nContrib, -1, k, P.sIdx[P.sPtr[k] + e]});
                P.contributions += (long long)lst.size();
                P.factorMads += 4LL * (long long)lst.size();   // two W = D^-1 U and two updates
            }
        }
        diagItems.insert(diagItems.end(), pairItems.begin(), pairItems.end());
        emit(diagItems);
        // backward records of the level's nodes, grouped by split, largest first
        std::stable_sort(bwd.begin(), bwd.end(), [](const std::pair<int, int> &u, const std::pair<int, int> &v) {
            return u.first > v.first;
        });
        for (size_t q = 0; q < bwd.size(); q++) {
            const int k = bwd[q].second;
            const int idx = (int)(P.bwdRec.size() / 4);
            if (q == 0 || bwd[q].first != bwd[q - 1].first)
                P.bwdBucketRec.insert(P.bwdBucketRec.end(), {idx, idx, bwd[q].first, 0});
            P.bwdBucketRec[P.bwdBucketRec.size() - 3]++;
            P.bwdRec.insert(P.bwdRec.end(), {P.nodeBase[k], P.sPtr[k], P.sPtr[k + 1], k});
        }
        P.levelNodePtr[l + 1] = (int)P.levelNodes.size();
        P.levelItemPtr[l + 1] = (int)(P.itemRec.size() / 4);
        P.levelBucketPtr[l + 1] = (int)(P.bucketRec.size() / 4);
        P.levelBwdBucketPtr[l + 1] = (int)(P.bwdBucketRec.size() / 4);
    }

    // --- per-branch flow coefficients: both ends' own/far forms, read off OLF's closed-branch expressions by the
    // generated closedBranchHalfForms (the same forms the fill gathers)
    P.branchCoef.assign(8 * (size_t)nBranch, 0.0);
    P.branchNode.assign(2 * (size_t)nBranch, 0);
    std::vector<gen::HalfBranchForm> side1(nBranch), side2(nBranch);
    for (int e = 0; e < nBranch; e++) {
        gen::closedBranchHalfForms(p8 + 8 * e, &side1[e], &side2[e]);
        double *c = &P.branchCoef[8 * (size_t)e];
        c[0] = side1[e].A;
        c[1] = side1[e].B;
        c[2] = side1[e].C;
        c[3] = side1[e].psi;
        c[4] = side2[e].A;
        c[5] = side2[e].C;
        c[6] = side2[e].psi;
        P.branchNode[2 * e] = P.pos[bus1[e]];
        P.branchNode[2 * e + 1] = P.pos[bus2[e]];
    }

    // --- fill gather tables
    std::vector<std::vector<int>> inc(nb);
    for (int e = 0; e < nBranch; e++) {
        inc[bus1[e]].push_back(e);
        inc[bus2[e]].push_back(e);
    }
    P.hbPtr.assign(nb + 1, 0);
    P.grpPtr.assign(nb + 1, 0);
    P.fillPtr.assign(nb + 1, 0);
    struct Half { int far, branch; double A, B, C, psi; };
    std::vector<Half> hs;
    for (int k = 0; k < nb; k++) {
        const int b = P.perm[k];
        hs.clear();
        for (int e : inc[b]) {
            // the branch seen from bus b: side 1 (far end bus2) or side 2 (far end bus1)
            const gen::HalfBranchForm &f = bus1[e] == b ? side1[e] : side2[e];
            hs.push_back({P.pos[bus1[e] == b ? bus2[e] : bus1[e]], e, f.A, f.B, f.C, f.psi});
        }
        std::sort(hs.begin(), hs.end(), [](const Half &u, const Half &v) {
            return u.far != v.far ? u.far < v.far : u.branch < v.branch;
        });
        for (size_t h = 0; h < hs.size(); h++) {
            P.hbBranch.push_back(hs[h].branch);
            P.hbCoef.push_back(hs[h].A);
            P.hbCoef.push_back(hs[h].B);
            P.hbCoef.push_back(hs[h].C);
            P.hbCoef.push_back(hs[h].psi);
            bool last = h + 1 == hs.size() || hs[h + 1].far != hs[h].far;
            if (last) {
                const int f = hs[h].far;
                int slot;
                if (f > k) {
                    int e = indexIn(P.sIdx, P.sPtr[k], P.sPtr[k + 1], f);
                    if (e < 0) throw std::logic_error("pattern entry missing from the LU structure");
                    slot = lslot(k, e) + 1;      // U(k, f)
                } else {
                    int e = indexIn(P.sIdx, P.sPtr[f], P.sPtr[f + 1], k);
                    if (e < 0) throw std::logic_error("pattern entry missing from the LU structure");
                    slot = lslot(f, e);          // L(k, f)
                }
                P.grpEnd.push_back((int)P.hbBranch.size());
                P.grpFar.push_back(f);
                P.grpSlot.push_back(slot);
            }
        }
        P.hbPtr[k + 1] = (int)P.hbBranch.size();
        P.grpPtr[k + 1] = (int)P.grpFar.size();
        for (int e = 0; e < P.sPtr[k + 1] - P.sPtr[k]; e++) {
            const int i = P.sIdx[P.sPtr[k] + e];
            if (!std::binary_search(nbrs[b].begin(), nbrs[b].end(), P.perm[i])) P.fillSlot.push_back(lslot(k, e));
        }
        P.fillPtr[k + 1] = (int)P.fillSlot.size();
        P.shuntG[k] = gBus[b];
        P.shuntB[k] = bBus[b];
    }

    // fill-in pairs: flagged in their items and never written by the fill
    P.fillMask.assign(P.nSlots, 0);
    for (int s0 : P.fillSlot) P.fillMask[s0] = P.fillMask[s0 + 1] = 1;
    size_t fillItems = 0;
    for (size_t q = 0; q < P.itemRec.size() / 4; q++)
        if (P.itemRec[4 * q + 3] < 0 && P.fillMask[P.itemRec[4 * q]]) {
            P.itemRec[4 * q + 3] = ITEM_FILL_PAIR;
            fillItems++;
        }
    // every fill-in pair takes contributions, so it has an item that starts it from zero
    if (fillItems != P.fillSlot.size()) throw std::logic_error("a fill-in pair has no factorization item");

    // --- scalar CSR pattern (stable rows) and its slot map
    P.rowPtr.assign(n + 1, 0);
    for (int b = 0; b < nb; b++) {
        std::vector<int> cols = nbrs[b];
        cols.push_back(b);
        std::sort(cols.begin(), cols.end());
        for (int r = 0; r < 2; r++) {
            const int k = P.pos[b];
            for (int c : cols) {
                int slot;
                if (c == b) {
                    slot = P.nodeBase[k];
                } else {
                    const int f = P.pos[c];
                    if (f > k) slot = lslot(k, indexIn(P.sIdx, P.sPtr[k], P.sPtr[k + 1], f)) + 1;
                    else slot = lslot(f, indexIn(P.sIdx, P.sPtr[f], P.sPtr[f + 1], k));
                }
                for (int cc = 0; cc < 2; cc++) {
                    P.colIdx.push_back(2 * c + cc);
                    P.csrSlot.push_back(slot * 4 + 2 * r + cc);
                }
            }
            P.rowPtr[2 * b + r + 1] = (int)P.colIdx.size();
        }
    }
    return P;
}

int Plan::topLevel(int maxNodes, int *nodes) const {
    int l0 = nLevels;
    while (l0 > 0 && nb - levelNodePtr[l0 - 1] <= maxNodes) l0--;
    if (nodes) *nodes = nb - levelNodePtr[l0];
    return l0;
}

void Plan::topItems(int l0, bool diagOnly, std::vector<int> &items, std::vector<int> &buckets) const {
    std::vector<std::pair<int, int>> v;   // (split of the contributions from below l0, item)
    for (int i = levelItemPtr[l0]; i < levelItemPtr[nLevels]; i++) {
        if (diagOnly && itemRec[4 * i + 3] < 0) continue;
        int c = 0;
        for (int q = itemRec[4 * i + 1]; q < itemRec[4 * i + 2]; q++) c += contribRec[4 * q + 3] < l0;
        v.push_back({splitFor(c), i});
    }
    std::stable_sort(v.begin(), v.end(),
                     [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first > b.first; });
    items.clear();
    buckets.clear();
    for (size_t q = 0; q < v.size(); q++) {
        if (q == 0 || v[q].first != v[q - 1].first)
            buckets.insert(buckets.end(), {(int)q, (int)q, v[q].first, 0});
        buckets[buckets.size() - 3]++;
        items.push_back(v[q].second);
    }
}

size_t Plan::bytesPerScenario(int luValueBytes) const {
    size_t s = (size_t)nSlots * 4 * luValueBytes;     // LU values
    s += (size_t)n * 2 * luValueBytes;                // forward-solve vector and first step
    s += (size_t)n * 8 * 6;                           // x, F, step, residual, target, next state
    s += (size_t)nb * 8;                              // reactive injections
    s += (size_t)n;                                   // row modes
    return s + 256;
}

}  // namespace olfhip
