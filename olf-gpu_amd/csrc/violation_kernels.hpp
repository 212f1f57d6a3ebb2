// Device screen of limit violations over the states a batched solve left on the device (reference
// native/cuda/sa/olf_gpu_violations.cu:1-27,170-300, semantics of GpuLimitScanner.java:20-130).
//
// The screen does not decide: it compares against a threshold relaxed by a relative band and keeps emitting a group's
// exceeded limits until one is exceeded beyond the band (certain to be confirmed, and a group reports its first
// confirmed limit only), so it can only hand the host more candidates than the exact decision reports. The host then
// re-tests every candidate with the exact comparison on its own FP64 evaluation of the same state (detectViolations
// in api.hip), which makes the reported set independent of the device arithmetic (FMA contraction, sin/cos).
//
// Records: branch limit l -> l; bus i high voltage -> nLimits + 2i, low voltage -> nLimits + 2i + 1.
#pragma once

#include <hip/hip_runtime.h>

#include "nr_kernels.hpp"

namespace olfhip {

constexpr int QTY_I = 0, QTY_P = 1, QTY_S = 2;

// What a limit group compares with its thresholds, from one side's flow p + jq at voltage v (per unit of the state):
// the current |S| / v, the active power's magnitude |p|, or the apparent power |S|. One definition for the device
// screen and the host's exact decision, so the two cannot drift apart.
__host__ __device__ inline double comparedQuantity(int quantity, double p, double q, double v) {
    if (quantity == QTY_P) return fabs(p);
    const double apparent = sqrt(p * p + q * q);
    return quantity == QTY_I ? apparent / v : apparent;
}

struct DevLimits {
    int nBranches, nGroups, nLimits, nBuses;
    const int *branch;        // table branch -> closed branch
    const int *groupStart;    // [nBranches + 1]
    const int *groupSide;     // 1 or 2
    const int *groupQty;
    const int *limitStart;    // [nGroups + 1]
    const double *threshold;  // [nLimits]
    const double *busHigh, *busLow;   // [nBuses], NaN: none
    const double *coef;       // per closed branch: A1 B C1 psi1 A2 C2 psi2 0
    const int *branchNode;    // per closed branch: end nodes
};

__device__ __forceinline__ void emitRecord(int *records, int *counts, int scenario, int capacity, int record) {
    const int i = atomicAdd(counts + scenario, 1);
    if (i < capacity) records[(size_t)scenario * capacity + i] = record;
}

// Whether node k of lane t is cut off from the slack in this scenario: its angle row is an identity row here but an
// equation in the base case (the island freeze).
template <int T>
__device__ __forceinline__ bool islandedNode(const DevPlan &P, const DevState &S, const int *modeBase, int tile, int k,
                                             int t) {
    return S.mode[vIdx<T>(tile, P.nb, k, t)] == MODE_FIXED && modeBase[2 * P.perm[k]] != MODE_FIXED;
}

// Flows at both ends of closed branch e in lane t of a tile, from the own/far coefficients of the fill: f = p1 q1 v1
// p2 q2 v2.
__device__ __forceinline__ void branchEndFlowsOf(const DevLimits &L, int e, double2 x1, double2 x2, double f[6]) {
    const double *c = L.coef + 8 * (size_t)e;
    double s1, c1, s2, c2;
    sincosReduced(c[3] - x1.x + x2.x, &s1, &c1);
    sincosReduced(c[6] - x2.x + x1.x, &s2, &c2);
    const double bvv = c[1] * x1.y * x2.y;
    f[0] = c[0] * x1.y * x1.y - bvv * s1;
    f[1] = c[2] * x1.y * x1.y - bvv * c1;
    f[2] = x1.y;
    f[3] = c[4] * x2.y * x2.y - bvv * s2;
    f[4] = c[5] * x2.y * x2.y - bvv * c2;
    f[5] = x2.y;
}

template <int T>
__device__ __forceinline__ void branchEndFlows(const DevPlan &P, const DevState &S, const DevLimits &L, int tile, int t,
                                               int e, double f[6]) {
    const int k1 = L.branchNode[2 * e], k2 = L.branchNode[2 * e + 1];
    const double2 x1 = *reinterpret_cast<const double2 *>(S.x + vIdx<T>(tile, P.nb, k1, t));
    const double2 x2 = *reinterpret_cast<const double2 *>(S.x + vIdx<T>(tile, P.nb, k2, t));
    branchEndFlowsOf(L, e, x1, x2, f);
}

// Whether closed branch e is out of service in scenario s: outaged, or an end cut off from the slack.
template <int T>
__device__ __forceinline__ bool branchOutOfService(const DevPlan &P, const DevState &S, const DevLimits &L,
                                                   const int *modeBase, int s, int e) {
    const int tile = s / T, t = s % T;
    bool out = false;
    for (int q = 0; q < S.maxOut; q++) out |= S.outages[(size_t)s * S.maxOut + q] == e;
    return out || islandedNode<T>(P, S, modeBase, tile, L.branchNode[2 * e], t) ||
           islandedNode<T>(P, S, modeBase, tile, L.branchNode[2 * e + 1], t);
}

// One thread per (scenario, table branch). Scenarios whose flag is 0 (not converged) emit nothing.
template <int T>
__global__ void screenBranchesKernel(DevPlan P, DevState S, DevLimits L, const int *modeBase, const int *screenLane,
                                     int nScenarios, double band, int capacity, int *records, int *counts) {
    const size_t total = (size_t)nScenarios * L.nBranches;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(w % L.nBranches);
        const int s = (int)(w / L.nBranches);
        if (!screenLane[s]) continue;
        const int g0 = L.groupStart[b], g1 = L.groupStart[b + 1];
        if (g0 == g1) continue;
        const int e = L.branch[b];
        if (branchOutOfService<T>(P, S, L, modeBase, s, e)) continue;
        double f[6];
        branchEndFlows<T>(P, S, L, s / T, s % T, e, f);
        for (int g = g0; g < g1; g++) {
            const double *end = L.groupSide[g] == 1 ? f : f + 3;
            const double value = comparedQuantity(L.groupQty[g], end[0], end[1], end[2]);
            // thresholds run most severe first and the host reports the first one exceeded: hand it every limit the
            // value may exceed (threshold lowered by the band); past one it exceeds for certain (threshold raised by
            // the band) the rest of the group cannot be reported. A NaN value fails every comparison.
            const int lEnd = L.limitStart[g + 1];
            bool settled = false;
            for (int l = L.limitStart[g]; l < lEnd && !settled; l++) {
                const double margin = fabs(L.threshold[l]) * band;
                if (!(value > L.threshold[l] - margin)) continue;
                emitRecord(records, counts, s, capacity, l);
                settled = value > L.threshold[l] + margin;
            }
        }
    }
}

// One thread per (scenario, bus with limits). A NaN limit compares false (never fires).
template <int T>
__global__ void screenBusesKernel(DevPlan P, DevState S, DevLimits L, const int *modeBase, const int *screenLane,
                                  int nScenarios, double band, int capacity, int *records, int *counts) {
    const size_t total = (size_t)nScenarios * L.nBuses;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(w % L.nBuses);
        const int s = (int)(w / L.nBuses);
        if (!screenLane[s]) continue;
        const int tile = s / T, t = s % T;
        const int k = P.pos[i];
        if (islandedNode<T>(P, S, modeBase, tile, k, t)) continue;
        const double v = S.x[vIdx<T>(tile, P.nb, k, t) + 1];
        const double hi = L.busHigh[i], lo = L.busLow[i];
        if (v > hi - fabs(hi) * band) emitRecord(records, counts, s, capacity, L.nLimits + 2 * i);
        if (v < lo + fabs(lo) * band) emitRecord(records, counts, s, capacity, L.nLimits + 2 * i + 1);
    }
}

// Scenarios of the last solve whose decisions a state difference of relative size `band` could change: a converged
// scenario gets flag 1 when a monitored quantity lies within band x |threshold| of one of its group's thresholds up to
// the first one it exceeds beyond the band (a later threshold cannot be reported), or a bus voltage within band x
// |limit| of its high or low limit, unless the bus holds its voltage (its V row an identity: PV, slack or islanded
// buses, whose voltage is the setpoint whatever the start). only: NULL, or per slot 1 for the scenarios the last solve
// handed back (a bank solve, whose other slots hold scenarios ended earlier or still solving). One thread per
// (scenario, table branch or bus).
template <int T>
__global__ void nearLimitsKernel(DevPlan P, DevState S, DevLimits L, const int *modeBase, int nScenarios, double band,
                                 const int *only, int *flag) {
    const int items = L.nBranches + L.nBuses;
    const size_t total = (size_t)nScenarios * items;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (size_t)gridDim.x * blockDim.x) {
        const int item = (int)(w % items);
        const int s = (int)(w / items);
        if (S.status[s] != ST_CONVERGED || (only && !only[s])) continue;
        const int tile = s / T, t = s % T;
        bool near = false;
        if (item < L.nBranches) {
            const int g0 = L.groupStart[item], g1 = L.groupStart[item + 1];
            const int e = L.branch[item];
            if (g0 == g1 || branchOutOfService<T>(P, S, L, modeBase, s, e)) continue;
            double f[6];
            branchEndFlows<T>(P, S, L, tile, t, e, f);
            for (int g = g0; g < g1 && !near; g++) {
                const double *end = L.groupSide[g] == 1 ? f : f + 3;
                const double value = comparedQuantity(L.groupQty[g], end[0], end[1], end[2]);
                for (int l = L.limitStart[g]; l < L.limitStart[g + 1]; l++) {
                    const double margin = fabs(L.threshold[l]) * band;
                    near |= fabs(value - L.threshold[l]) <= margin;
                    if (near || value > L.threshold[l] + margin) break;
                }
            }
        } else {
            const int i = item - L.nBranches;
            const int k = P.pos[i];
            const size_t vi = vIdx<T>(tile, P.nb, k, t);
            if (S.mode[vi + 1] == MODE_FIXED) continue;   // voltage held at its setpoint (or frozen)
            const double v = S.x[vi + 1];
            near = fabs(v - L.busHigh[i]) <= fabs(L.busHigh[i]) * band ||
                   fabs(v - L.busLow[i]) <= fabs(L.busLow[i]) * band;   // NaN limits compare false
        }
        if (near) flag[s] = 1;
    }
}

// nearLimitsKernel with one workgroup per scenario: the lane's node states (16 B a node) and two flags per node (bit 0:
// cut off from the slack, bit 1: voltage held at its setpoint) staged in LDS (17 B a node: 157 KB at 9,241 buses), so
// that the branches and buses read them there instead of gathering them across the lane's states in global memory.
// The same comparisons on the same values: the flags equal nearLimitsKernel's.
template <int T>
__global__ __launch_bounds__(1024) void nearLimitsStagedKernel(DevPlan P, DevState S, DevLimits L, const int *modeBase,
                                                               double band, const int *only, int *flag) {
    extern __shared__ double2 nearX[];
    __shared__ int sNear;
    const int s = blockIdx.x;
    if (S.status[s] != ST_CONVERGED || (only && !only[s])) return;   // uniform over the workgroup
    const int tile = s / T, t = s % T;
    unsigned char *nodeFlags = reinterpret_cast<unsigned char *>(nearX + P.nb);
    if (threadIdx.x == 0) sNear = 0;
    for (int k = threadIdx.x; k < P.nb; k += blockDim.x) {
        const size_t vi = vIdx<T>(tile, P.nb, k, t);
        nearX[k] = *reinterpret_cast<const double2 *>(S.x + vi);
        nodeFlags[k] = (unsigned char)((S.mode[vi] == MODE_FIXED && modeBase[2 * P.perm[k]] != MODE_FIXED ? 1 : 0) |
                                       (S.mode[vi + 1] == MODE_FIXED ? 2 : 0));
    }
    int outs[MAX_OUTAGES];
    for (int q = 0; q < MAX_OUTAGES; q++) outs[q] = q < S.maxOut ? S.outages[(size_t)s * S.maxOut + q] : -1;
    __syncthreads();
    bool near = false;
#pragma unroll 4
    for (int b = threadIdx.x; b < L.nBranches; b += blockDim.x) {
        const int g0 = L.groupStart[b], g1 = L.groupStart[b + 1];
        if (g0 == g1) continue;
        const int e = L.branch[b];
        const int k1 = L.branchNode[2 * e], k2 = L.branchNode[2 * e + 1];
        bool out = (nodeFlags[k1] & 1) || (nodeFlags[k2] & 1);
        for (int q = 0; q < MAX_OUTAGES; q++) out |= outs[q] == e;
        if (out) continue;
        double f[6];
        branchEndFlowsOf(L, e, nearX[k1], nearX[k2], f);
        for (int g = g0; g < g1 && !near; g++) {
            const double *end = L.groupSide[g] == 1 ? f : f + 3;
            const double value = comparedQuantity(L.groupQty[g], end[0], end[1], end[2]);
            for (int l = L.limitStart[g]; l < L.limitStart[g + 1]; l++) {
                const double margin = fabs(L.threshold[l]) * band;
                near |= fabs(value - L.threshold[l]) <= margin;
                if (near || value > L.threshold[l] + margin) break;
            }
        }
    }
#pragma unroll 4
    for (int i = threadIdx.x; i < L.nBuses; i += blockDim.x) {
        const int k = P.pos[i];
        if (nodeFlags[k] & 2) continue;   // voltage held at its setpoint (or frozen)
        const double v = nearX[k].y;
        near |= fabs(v - L.busHigh[i]) <= fabs(L.busHigh[i]) * band ||
                fabs(v - L.busLow[i]) <= fabs(L.busLow[i]) * band;   // NaN limits compare false
    }
    if (near) sNear = 1;
    __syncthreads();
    if (threadIdx.x == 0 && sNear) flag[s] = 1;
}

// Gathers the (phi, V) pairs of the requested (scenario, node) pairs and whether each node is islanded.
template <int T>
__global__ void gatherNodesKernel(DevPlan P, DevState S, const int *modeBase, int n, const int2 *req, double2 *out,
                                  int *islanded) {
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < n; w += gridDim.x * blockDim.x) {
        const int2 r = req[w];
        const int tile = r.x / T, t = r.x % T;
        out[w] = *reinterpret_cast<const double2 *>(S.x + vIdx<T>(tile, P.nb, r.y, t));
        islanded[w] = islandedNode<T>(P, S, modeBase, tile, r.y, t) ? 1 : 0;
    }
}

// Screen flags of the last solve's scenarios from their device statuses (1: converged, and handed back by that solve
// when `only` is not NULL) and their candidate counts cleared, in one pass.
__global__ void convergedLaneFlagsKernel(const int *status, const int *only, int n, int *flag, int *counts) {
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < n; w += gridDim.x * blockDim.x) {
        flag[w] = status[w] == ST_CONVERGED && (!only || only[w]);
        counts[w] = 0;
    }
}

__global__ void zeroIntKernel(int *p, int n) {
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < n; w += gridDim.x * blockDim.x) p[w] = 0;
}

// Where each scenario's screened records start in one packed list: off[k] = the sum over j < k of min(counts[j],
// capacity), off[K] the total. One workgroup of PACK_THREADS threads: each sums a run of consecutive scenarios, the
// run sums are scanned in LDS (Hillis-Steele), then each thread writes its run's offsets.
constexpr int PACK_THREADS = 1024;
__global__ __launch_bounds__(PACK_THREADS) void packOffsetsKernel(const int *counts, int K, int capacity, int *off) {
    __shared__ int run[PACK_THREADS];
    const int t = threadIdx.x;
    const int per = (K + PACK_THREADS - 1) / PACK_THREADS;
    const int k0 = min(K, t * per), k1 = min(K, k0 + per);
    int sum = 0;
    for (int k = k0; k < k1; k++) sum += min(counts[k], capacity);
    run[t] = sum;
    __syncthreads();
    for (int d = 1; d < PACK_THREADS; d <<= 1) {
        const int add = t >= d ? run[t - d] : 0;
        __syncthreads();
        run[t] += add;
        __syncthreads();
    }
    int at = run[t] - sum;   // exclusive
    for (int k = k0; k < k1; k++) {
        off[k] = at;
        at += min(counts[k], capacity);
    }
    if (t == PACK_THREADS - 1) off[K] = run[t];
}

// Scenario k's records (block k) copied from its row of `capacity` places to its place in the packed list.
__global__ void packRecordsKernel(const int *records, const int *counts, const int *off, int capacity, int *packed) {
    const int k = blockIdx.x;
    const int c = min(counts[k], capacity);
    const int *src = records + (size_t)k * capacity;
    int *dst = packed + off[k];
    for (int q = threadIdx.x; q < c; q += blockDim.x) dst[q] = src[q];
}

}  // namespace olfhip
