// Host driver and C ABI of libolfhip.so (declarations and reference citations in include/olfhip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/olfhip.h"
#include "errors.hpp"
#include "nr_kernels.hpp"
#include "plan.hpp"
#include "violation_kernels.hpp"

namespace olfhip {

static thread_local std::string gLastError;
std::string &lastErrorRef() { return gLastError; }

#define HIP_OK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) {                                                                          \
            if (e_ == hipErrorOutOfMemory) throw MemoryError(std::string(#expr) + ": out of device memory"); \
            throw DeviceError(std::string(#expr) + ": " + hipGetErrorString(e_));                        \
        }                                                                                                \
    } while (0)

static bool traceOn() {
    static int on = -1;
    if (on < 0) {
        const char *e = std::getenv("OLFHIP_TRACE");
        on = (e && *e && std::strcmp(e, "0") != 0) ? 1 : 0;
    }
    return on == 1;
}

// RAII device buffer
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b) {
        if (b <= bytes && p) return;
        release();
        HIP_OK(hipMalloc(&p, b ? b : 16));
        bytes = b;
    }
    template <typename X>
    void upload(const std::vector<X> &v, hipStream_t s = nullptr) {
        alloc(v.size() * sizeof(X));
        if (!v.empty()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice, s));
    }
    template <typename X>
    X *as() const { return reinterpret_cast<X *>(p); }
};

struct StructureImpl {
    Plan plan;
    std::vector<int> bus1, bus2;
    std::vector<int> incPtr, inc;    // branches incident to each bus (island search)
    double jacobianScale = 1.0;      // twice the largest bus admittance sum: the order of ||J||inf near 1 pu
    std::atomic<int> refs{1};   // the caller's handle and every context and limit table made on it
    bool onDevice = false;
    std::mutex deviceLock;      // the plan's first upload, by whichever thread first needs it
    int device = -1;
    std::vector<std::unique_ptr<DBuf>> bufs;
    DevPlan dp{};

    template <typename X>
    const X *put(const std::vector<X> &v) {
        bufs.emplace_back(new DBuf());
        bufs.back()->upload(v);
        return bufs.back()->as<X>();
    }
    void ensureDevice() {
        std::lock_guard<std::mutex> lock(deviceLock);
        if (onDevice) return;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw DeviceError("no HIP device is visible");
        HIP_OK(hipGetDevice(&device));
        const Plan &P = plan;
        dp.nb = P.nb;
        dp.nSlots = P.nSlots;
        dp.nLevels = P.nLevels;
        dp.levelNodePtr = put(P.levelNodePtr);
        dp.levelItemPtr = put(P.levelItemPtr);
        dp.itemRec = reinterpret_cast<const int4 *>(put(P.itemRec));
        dp.contribRec = reinterpret_cast<const int4 *>(put(P.contribRec));
        dp.bwdRec = reinterpret_cast<const int4 *>(put(P.bwdRec));
        dp.bucketRec = reinterpret_cast<const int4 *>(put(P.bucketRec));
        dp.bwdBucketRec = reinterpret_cast<const int4 *>(put(P.bwdBucketRec));
        dp.levelBucketPtr = put(P.levelBucketPtr);
        dp.levelBwdBucketPtr = put(P.levelBwdBucketPtr);
        dp.levelNodes = put(P.levelNodes);
        dp.diagItemOf = put(P.diagItemOf);
        dp.nBuckets = (int)(P.bucketRec.size() / 4);
        dp.nBwdBuckets = (int)(P.bwdBucketRec.size() / 4);
        dp.nodeBase = put(P.nodeBase);
        dp.sPtr = put(P.sPtr);
        dp.sIdx = put(P.sIdx);
        dp.hbPtr = put(P.hbPtr);
        dp.hbBranch = put(P.hbBranch);
        dp.hbCoef = reinterpret_cast<const double4 *>(put(P.hbCoef));
        dp.grpPtr = put(P.grpPtr);
        dp.grpEnd = put(P.grpEnd);
        dp.grpFar = put(P.grpFar);
        dp.grpSlot = put(P.grpSlot);
        dp.shuntG = put(P.shuntG);
        dp.shuntB = put(P.shuntB);
        dp.perm = put(P.perm);
        dp.pos = put(P.pos);
        dp.levelPos = put(P.levelPos);
        dp.nodeLevel = put(P.nodeLevel);
        dp.itemKI = reinterpret_cast<const int2 *>(put(P.itemKI));
        HIP_OK(hipDeviceSynchronize());
        onDevice = true;
    }
};

// The top block a launch of `kernel` (T lanes, NT threads, vtBytes per LU value) can hold in LDS without costing
// occupancy: the LDS one workgroup may use when as many workgroups share a CU as the kernel's registers allow, less
// the kernel's static LDS. OLFHIP_TOP_NODES caps the node count (0: no top block).
struct TopPlan {
    DevTop top{};
    size_t ldsBytes = 0;
    DBuf slots, items, diagItems, buckets, diagBuckets;
    DBuf contribPad, bwdPad, bucketPad, bwdBucketPad;
};

// The sparse levels' records in the unit-major order the kernels read for T lanes per tile (the split of a bucket
// capped at 64 / T as in the kernels): for bucket b with split S, rounds R = the largest ceil(length / S) of its
// items, thread s of item q reads its r-th record at base + (r * nItems + q) * S + s, the same contribution the
// item-major schedule gives it (index s + r S of the item's list), so sums run in the same order. Unused places
// hold x = -1.
static void planPadded(const Plan &P, int T, TopPlan &tp) {
    const int cap = 64 / T;
    std::vector<int> pad, bpad, epad, ebpad;
    for (size_t b = 0; b < P.bucketRec.size() / 4; b++) {
        const int first = P.bucketRec[4 * b], end = P.bucketRec[4 * b + 1];
        const int S = std::min(P.bucketRec[4 * b + 2], cap), nq = end - first;
        int R = 0;
        for (int q = 0; q < nq; q++) {
            const int len = P.itemRec[4 * (first + q) + 2] - P.itemRec[4 * (first + q) + 1];
            R = std::max(R, (len + S - 1) / S);
        }
        const size_t base = pad.size() / 4;
        pad.resize(pad.size() + (size_t)4 * R * nq * S, -1);
        for (int q = 0; q < nq; q++) {
            const int c0 = P.itemRec[4 * (first + q) + 1], len = P.itemRec[4 * (first + q) + 2] - c0;
            for (int c = 0; c < len; c++) {
                const size_t at = base + ((size_t)(c / S) * nq + q) * S + c % S;
                for (int f = 0; f < 4; f++) pad[4 * at + f] = P.contribRec[4 * (size_t)(c0 + c) + f];
            }
        }
        bpad.insert(bpad.end(), {(int)base, R});
    }
    for (size_t b = 0; b < P.bwdBucketRec.size() / 4; b++) {
        const int first = P.bwdBucketRec[4 * b], end = P.bwdBucketRec[4 * b + 1];
        const int S = std::min(P.bwdBucketRec[4 * b + 2], cap), nq = end - first;
        int R = 0;
        for (int q = 0; q < nq; q++) {
            const int len = P.bwdRec[4 * (first + q) + 2] - P.bwdRec[4 * (first + q) + 1];
            R = std::max(R, (len + S - 1) / S);
        }
        const size_t base = epad.size() / 2;
        epad.resize(epad.size() + (size_t)2 * R * nq * S, -1);
        for (int q = 0; q < nq; q++) {
            const int *br = &P.bwdRec[4 * (size_t)(first + q)];
            for (int e = 0; e < br[2] - br[1]; e++) {
                const size_t at = base + ((size_t)(e / S) * nq + q) * S + e % S;
                epad[2 * at] = br[0] + 2 + 2 * e;         // U(k, i): after L~(i, k) in node k's pair e
                epad[2 * at + 1] = P.sIdx[br[1] + e];     // i
            }
        }
        ebpad.insert(ebpad.end(), {(int)base, R});
    }
    if (pad.empty()) pad.assign(4, -1);
    if (epad.empty()) epad.assign(2, -1);
    tp.contribPad.upload(pad);
    tp.bwdPad.upload(epad);
    tp.bucketPad.upload(bpad.empty() ? std::vector<int>{0, 0} : bpad);
    tp.bwdBucketPad.upload(ebpad.empty() ? std::vector<int>{0, 0} : ebpad);
    tp.top.contribPad = tp.contribPad.as<int4>();
    tp.top.bwdPad = tp.bwdPad.as<int2>();
    tp.top.bucketPad = tp.bucketPad.as<int2>();
    tp.top.bwdBucketPad = tp.bwdBucketPad.as<int2>();
}

static void planTop(const StructureImpl *st, const void *kernel, int T, int NT, size_t vtBytes, TopPlan &tp) {
    const Plan &P = st->plan;
    hipFuncAttributes fa{};
    HIP_OK(hipFuncGetAttributes(&fa, kernel));
    int perCU = 0, perBlock = 0;
    HIP_OK(hipDeviceGetAttribute(&perCU, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, st->device));
    HIP_OK(hipDeviceGetAttribute(&perBlock, hipDeviceAttributeMaxSharedMemoryPerBlock, st->device));
    const int vgpr = std::max(8, (fa.numRegs + 7) / 8 * 8);
    const int wavesPerSimd = std::max(1, std::min(8, 512 / vgpr));
    const int blocksPerCU = std::max(1, (4 * wavesPerSimd) / std::max(1, NT / 64));
    const size_t metaBytes = levelMetaBytes(P.nLevels, (int)(P.bucketRec.size() / 4), (int)(P.bwdBucketRec.size() / 4));
    long budget = std::min<long>(perBlock, perCU / blocksPerCU) - (long)fa.sharedSizeBytes - (long)metaBytes - 256;
    int maxNodes = 0;
    while (budget > 0 && topLdsBytes(T, maxNodes + 1, vtBytes) <= (size_t)budget) maxNodes++;
    if (const char *e = std::getenv("OLFHIP_TOP_NODES")) maxNodes = std::min(maxNodes, std::max(0, std::atoi(e)));
    int n = 0;
    const int l0 = maxNodes > 0 ? P.topLevel(maxNodes, &n) : P.nLevels;
    if (l0 >= P.nLevels) n = 0;
    tp.top = DevTop{};
    planPadded(P, T, tp);
    tp.top.l0 = n > 0 ? l0 : P.nLevels;
    tp.top.n = n;
    tp.top.start = n > 0 ? P.levelNodePtr[l0] : P.nb;
    tp.top.metaBytes = (int)metaBytes;
    tp.ldsBytes = metaBytes + (n > 0 ? topLdsBytes(T, n, vtBytes) : 0);
    if (n > 0) {
        std::vector<int> rec;
        for (int a = 0; a < n; a++) {
            const int k = P.levelNodes[tp.top.start + a];
            rec.insert(rec.end(), {P.nodeBase[k], a, a, 0});
            for (int q = P.sPtr[k]; q < P.sPtr[k + 1]; q++) {
                const int b = P.levelPos[P.sIdx[q]] - tp.top.start;
                if (b < 0 || b >= n) throw std::logic_error("top block is not closed under S");
                const int slot = P.nodeBase[k] + 1 + 2 * (q - P.sPtr[k]);
                if (P.fillMask[slot]) continue;               // fill-in: starts from the zeroed LDS block
                rec.insert(rec.end(), {slot, b, a, 0});       // L~(i,k): row i, column k
                rec.insert(rec.end(), {slot + 1, a, b, 0});   // U(k,i): row k, column i
            }
        }
        tp.slots.upload(rec);
        tp.top.slots = tp.slots.as<int4>();
        tp.top.nSlots = (int)(rec.size() / 4);
        std::vector<int> items, buckets;
        P.topItems(l0, false, items, buckets);
        tp.items.upload(items);
        tp.buckets.upload(buckets);
        tp.top.items = tp.items.as<int>();
        tp.top.buckets = tp.buckets.as<int4>();
        tp.top.nBuckets = (int)(buckets.size() / 4);
        P.topItems(l0, true, items, buckets);
        tp.diagItems.upload(items);
        tp.diagBuckets.upload(buckets);
        tp.top.diagItems = tp.diagItems.as<int>();
        tp.top.diagBuckets = tp.diagBuckets.as<int4>();
        tp.top.nDiagBuckets = (int)(buckets.size() / 4);
        if (tp.ldsBytes > 65536)
            HIP_OK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)tp.ldsBytes));
    }
}

static void release(StructureImpl *s) {
    if (s && --s->refs == 0) delete s;
}

// Template dispatch over the lane tile width.
template <typename Fn>
static void withT(int T, Fn &&fn) {
    switch (T) {
        case 1: fn(std::integral_constant<int, 1>{}); break;
        case 2: fn(std::integral_constant<int, 2>{}); break;
        case 4: fn(std::integral_constant<int, 4>{}); break;
        case 8: fn(std::integral_constant<int, 8>{}); break;
        case 16: fn(std::integral_constant<int, 16>{}); break;
        case 32: fn(std::integral_constant<int, 32>{}); break;
        default: throw ArgumentError("lanes_per_tile must be 1, 2, 4, 8, 16 or 32");
    }
}

// Measured on MI355X at 9,241 buses, 2,000 scenarios: T = 2 with 1,024 threads is the fastest shape (the autotuner
// picks per workload).
static int defaultLanesPerTile(int nScenarios) { return nScenarios <= 1 ? 1 : 2; }

static int gridFor(size_t work, int block) {
    size_t g = (work + block - 1) / block;
    return (int)std::max<size_t>(1, std::min<size_t>(g, 65536));
}

struct ContextImpl {
    StructureImpl *st = nullptr;
    int capacity = 0, mode = 0, T = 1, NT = 512, nTiles = 0;
    hipStream_t stream = nullptr;
    int *pinnedCount = nullptr;
    // reactive limits
    int nRL = 0;
    DBuf rlNode, rlMinQ, rlMaxQ, rlLoadQ, rlTargetV;
    // per-lane state
    DBuf x, F, dx, r, target, qinj, mode_, lu, y, dx1;
    DBuf it, status, solves, outer, total, clamps, norm, ratio, outages, tileRunning, count;
    DBuf rlType, rlCount, rlPass;
    // distributed slack (olfhip_set_distributed_slack): participants in node numbering, per-lane loop sequence
    int dsN = 0, slackNode = -1;
    double slackTargetP = 0.0, dsFactorSum = 0.0;
    DBuf dsNode, dsFactor, dsBase, dsMaxD, dsMinD, dsDesc;
    DBuf loopAt, sweepChanged, dsFailed, slackP, dsRetain, dsPrev;
    DBuf xPrev, lsLast, lsHas;   // LINE_SEARCH state (xPrev allocated by the first solve that asks for it)
    DBuf nActive;                // active equations per lane (the 2-norm criterion)
    bool rowCriteria = false;    // per-row stopping thresholds set (olfhip_set_row_criteria), node order [node][2]
    DBuf rowThrFixed, rowThrEq;
    // install inputs
    DBuf x0Base, x0Per, targetBase, modeBase, modeEntries, targetIdx, targetVal, seed, states;
    std::vector<long long> x0SeedIndex;      // olfhip_set_x0_seeds: start-state seeds of the next solves
    std::vector<double> x0SeedValue;
    DBuf x0SeedIdx, x0SeedVal;
    int lastInstalled = 0;
    double lastSolveMs = 0.0;
    int lastSteps = 0;
    double lastStepMs = 0.0;                 // event-timed Newton-step launches of the last solve
    hipEvent_t evStart = nullptr, evStop = nullptr;
    DBuf prof;                               // opt-in section stamps of one tile (OLFHIP_PROFILE=1)
    int profCap = 0;
    int profTile = 0;                        // OLFHIP_PROFILE_TILE, default 0
    DBuf tileClock;                          // opt-in tile timeline (OLFHIP_TILE_TIMELINE=file)
    TopPlan top;                             // the dense top block of the step kernel
    std::vector<int> lastStatus;             // statuses of the last solve's scenarios
    DBuf kryV, kryH;                         // robust mode: the GMRES basis and per-tile dense state
    std::vector<int> lastOutages;            // its outage lists, MAX_OUTAGES per lane (-1 padded)
    // A solve as two graphs, each instantiated once and updated in place by later solves (their kernel arguments
    // change, their sequence does not): the install kernels, and the Newton-step launch with the running-tile count
    // and its copy to the host (replayed when a launch is capped at k steps, OLFHIP_STEPS_PER_LAUNCH).
    hipGraphExec_t installGraph = nullptr, stepGraph = nullptr;
    int graphInstantiations = 0, graphUpdates = 0;
    // The bank (BatchSolveNative.beginBatchInstall / pauseAtEnded / readBatchHandedBackSlots): the slots the next
    // solve installs its scenarios into (bankInstall), which slots hold a scenario not handed back yet (pending), the
    // slots the last solve handed back in the order of its rows, and per slot its outage list and equation count.
    bool bankInstall = false;
    std::vector<int> bankSlots;
    std::vector<char> pending;
    std::vector<long long> installSeq;       // per slot, when its scenario was installed (rows go in install order)
    long long installCounter = 0;
    std::vector<int> handedBack;
    std::vector<int> laneOutages, laneActive;
    DBuf ended, slotOf, handedNow;
    bool bankReadback = false;               // the last solve's readbacks are by slot (a bank or paused solve)

    DevState devState() const {
        DevState S{};
        S.nTiles = nTiles;
        S.nInstalled = lastInstalled;
        S.x = x.as<double>();
        S.F = F.as<double>();
        S.dx = dx.as<double>();
        S.r = r.as<double>();
        S.target = target.as<double>();
        S.qinj = qinj.as<double>();
        S.mode = mode_.as<signed char>();
        S.lu = lu.p;
        S.y = y.p;
        S.dx1 = dx1.p;
        S.it = it.as<int>();
        S.status = status.as<int>();
        S.solves = solves.as<int>();
        S.outer = outer.as<int>();
        S.total = total.as<int>();
        S.clamps = clamps.as<int>();
        S.norm = norm.as<double>();
        S.ratio = ratio.as<double>();
        S.outages = outages.as<int>();
        S.maxOut = MAX_OUTAGES;
        S.prof = prof.as<unsigned long long>();
        S.profCap = profCap;
        S.profTile = profTile;
        S.tileClock = tileClock.as<unsigned long long>();
        S.rlType = rlType.as<signed char>();
        S.rlCount = rlCount.as<signed char>();
        S.rlPass = rlPass.as<signed char>();
        S.tileRunning = tileRunning.as<int>();
        S.ended = ended.as<int>();
        S.kryV = kryV.as<double>();
        S.kryH = kryH.as<double>();
        S.loopAt = loopAt.as<signed char>();
        S.sweepChanged = sweepChanged.as<signed char>();
        S.dsFailed = dsFailed.as<signed char>();
        S.slackP = slackP.as<double>();
        S.dsRetain = dsRetain.as<double>();
        S.dsPrev = dsPrev.as<double>();
        S.xPrev = xPrev.as<double>();
        S.lsLast = lsLast.as<double>();
        S.lsHas = lsHas.as<signed char>();
        S.nActive = nActive.as<int>();
        S.rowThrFixed = rowCriteria ? rowThrFixed.as<double>() : nullptr;
        S.rowThrEq = rowCriteria ? rowThrEq.as<double>() : nullptr;
        return S;
    }
    DevRL devRL() const {
        return DevRL{nRL,
                     rlNode.as<int>(),
                     rlMinQ.as<double>(),
                     rlMaxQ.as<double>(),
                     rlLoadQ.as<double>(),
                     rlTargetV.as<double>(),
                     slackNode >= 0 ? dsDesc.as<DevDS>() : nullptr,
                     slackNode};
    }
    ~ContextImpl() {
        if (installGraph) (void)hipGraphExecDestroy(installGraph);
        if (stepGraph) (void)hipGraphExecDestroy(stepGraph);
        if (pinnedCount) (void)hipHostFree(pinnedCount);
        if (evStart) (void)hipEventDestroy(evStart);
        if (evStop) (void)hipEventDestroy(evStop);
        if (stream) (void)hipStreamDestroy(stream);
        release(st);
    }
};

static size_t luBytes(int mode) { return mode == OLFHIP_LU_FP32_REFINED ? 4 : 8; }
static bool knownMode(int mode) {
    return mode == OLFHIP_LU_FP64 || mode == OLFHIP_LU_FP32_REFINED || mode == OLFHIP_LU_ROBUST;
}
// Static pivoting threshold of the robust mode: sqrt(machine epsilon) times the Jacobian's scale (Li & Demmel's
// choice of sqrt(eps) ||A||).
static double robustPivotDelta(const StructureImpl *st) { return 1.4901161193847656e-08 * st->jacobianScale; }

static ContextImpl *createContext(StructureImpl *st, int nScenarios, int mode, const olfhip_reactive_limits *rl,
                                  const olfhip_tuning *tuning) {
    if (!st) throw ArgumentError("structure is null");
    if (nScenarios <= 0) throw ArgumentError("n_scenarios must be positive");
    if (!knownMode(mode)) throw ArgumentError("unknown linear-solve mode");
    if (mode == OLFHIP_LU_ROBUST && tuning && tuning->lanes_per_tile > 1)
        throw ArgumentError("the robust mode solves one scenario per tile (lanes_per_tile must be 1)");
    st->ensureDevice();
    std::unique_ptr<ContextImpl> c(new ContextImpl());
    st->refs++;
    c->st = st;
    c->capacity = nScenarios;
    c->mode = mode;
    c->T = mode == OLFHIP_LU_ROBUST ? 1
           : (tuning && tuning->lanes_per_tile > 0) ? tuning->lanes_per_tile : defaultLanesPerTile(nScenarios);
    c->NT = (tuning && tuning->threads_per_tile > 0) ? tuning->threads_per_tile : MAX_THREADS;
    withT(c->T, [](auto) {});
    if (c->NT != 256 && c->NT != 512 && c->NT != 1024) throw ArgumentError("threads_per_tile must be 256, 512 or 1024");
    // the step kernel's register budget is fixed for at most MAX_THREADS threads: a larger launch would not fit a CU
    if (c->NT > MAX_THREADS)
        throw ArgumentError("threads_per_tile above this build's maximum of " + std::to_string(MAX_THREADS));
    c->nTiles = (nScenarios + c->T - 1) / c->T;
    HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_OK(hipHostMalloc((void **)&c->pinnedCount, 2 * sizeof(int), hipHostMallocDefault));
    HIP_OK(hipEventCreate(&c->evStart));
    HIP_OK(hipEventCreate(&c->evStop));
    const Plan &P = st->plan;
    const size_t lanes = (size_t)c->nTiles * c->T;
    const size_t nv = lanes * P.nb * 2;
    c->x.alloc(nv * 8);
    c->F.alloc(nv * 8);
    c->dx.alloc(nv * 8);
    c->r.alloc(nv * 8);
    c->target.alloc(nv * 8);
    c->qinj.alloc(lanes * P.nb * 8);
    c->mode_.alloc(nv);
    c->lu.alloc(lanes * (size_t)P.nSlots * 4 * luBytes(mode));
    c->y.alloc(nv * luBytes(mode));
    c->dx1.alloc(nv * luBytes(mode));
    for (DBuf *b : {&c->it, &c->status, &c->solves, &c->outer, &c->total, &c->clamps}) b->alloc(lanes * 4);
    c->norm.alloc(lanes * 8);
    c->ratio.alloc(lanes * 8);
    c->outages.alloc(lanes * MAX_OUTAGES * 4);
    c->tileRunning.alloc((size_t)c->nTiles * 4);
    c->count.alloc(16);
    c->ended.alloc(16);
    c->slotOf.alloc(lanes * 4);
    c->handedNow.alloc(lanes * 4);
    c->pending.assign(lanes, 0);
    c->installSeq.assign(lanes, 0);
    c->laneOutages.assign(lanes * MAX_OUTAGES, -1);
    c->laneActive.assign(lanes, P.n);
    slotsEmptyKernel<<<gridFor(lanes, 256), 256, 0, c->stream>>>(c->status.as<int>(), (int)lanes,
                                                                 c->tileRunning.as<int>(), c->nTiles);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    for (DBuf *b : {&c->loopAt, &c->sweepChanged, &c->dsFailed}) b->alloc(lanes);
    c->slackP.alloc(lanes * 8);
    c->dsRetain.alloc(lanes * 8);
    c->dsPrev.alloc(8);
    c->lsLast.alloc(lanes * 8);
    c->lsHas.alloc(lanes);
    c->nActive.alloc(lanes * 4);
    // reactive limits, entry buses in node numbering
    if (rl && rl->n_entries > 0) {
        if (!rl->voltage_row || !rl->min_q || !rl->max_q || !rl->load_q || !rl->target_v)
            throw ArgumentError("reactive-limit arrays are null");
        c->nRL = rl->n_entries;
        std::vector<int> node(c->nRL);
        for (int e = 0; e < c->nRL; e++) {
            const int vr = rl->voltage_row[e];
            if (vr < 0 || vr >= P.n || (vr & 1) == 0) throw ArgumentError("reactive-limit voltage row is not a V row");
            node[e] = P.pos[vr >> 1];
        }
        c->rlNode.upload(node);
        c->rlMinQ.upload(std::vector<double>(rl->min_q, rl->min_q + c->nRL));
        c->rlMaxQ.upload(std::vector<double>(rl->max_q, rl->max_q + c->nRL));
        c->rlLoadQ.upload(std::vector<double>(rl->load_q, rl->load_q + c->nRL));
        c->rlTargetV.upload(std::vector<double>(rl->target_v, rl->target_v + c->nRL));
    }
    const size_t ne = lanes * (size_t)std::max(c->nRL, 1);
    c->rlType.alloc(ne);
    c->rlCount.alloc(ne);
    c->rlPass.alloc(ne);
    if (mode == OLFHIP_LU_ROBUST) {
        c->kryV.alloc((size_t)(OLFHIP_ROBUST_KRYLOV_DIM + 1) * nv * 8);
        c->kryH.alloc((size_t)c->nTiles * kryHStride(OLFHIP_ROBUST_KRYLOV_DIM) * 8);
        planTop(st, reinterpret_cast<const void *>(&nrStepKernel<1, double, true>), 1, c->NT, 8, c->top);
        c->top.top.pivotDelta = robustPivotDelta(st);
    } else {
        withT(c->T, [&](auto TC) {
            constexpr int T = decltype(TC)::value;
            if (mode == OLFHIP_LU_FP32_REFINED)
                planTop(st, reinterpret_cast<const void *>(&nrStepKernel<T, float>), T, c->NT, 4, c->top);
            else
                planTop(st, reinterpret_cast<const void *>(&nrStepKernel<T, double>), T, c->NT, 8, c->top);
        });
    }
    HIP_OK(hipDeviceSynchronize());
    return c.release();
}

static void solve(ContextImpl *c, const olfhip_solve_inputs *in, olfhip_solve_outputs *out) {
    if (!in) throw ArgumentError("inputs are null");
    const Plan &P = c->st->plan;
    const int n = P.n;
    const int K = in->n_scenarios;
    // the slots olfhip_begin_batch_install chose hold for this solve only, refused or not
    const bool bank = c->bankInstall;
    std::vector<int> bankSlots;
    bankSlots.swap(c->bankSlots);
    c->bankInstall = false;
    if (bank ? (K != (int)bankSlots.size()) : (K <= 0 || K > c->capacity))
        throw ArgumentError(bank ? "n_scenarios must be the count of slots olfhip_begin_batch_install chose"
                                 : "n_scenarios must be in [1, context capacity]");
    if (in->pause_at_ended < 0) throw ArgumentError("pause_at_ended must be >= 0");
    if ((bank || in->pause_at_ended > 0) && out && !out->n_handed_back)
        throw ArgumentError("a bank or paused solve hands back its own count of rows: n_handed_back is needed");
    if (!in->x0_base || !in->target_base || !in->base_row_mode) throw ArgumentError("base state arrays are null");
    if (in->max_iter <= 0) throw ArgumentError("max_iter must be positive");
    if (in->max_outages < 0 || in->max_outages > MAX_OUTAGES)
        throw UnsupportedError("at most 4 outaged branches per scenario are modelled");
    if (in->max_outages > 0 && !in->outaged_branches) throw ArgumentError("outaged_branches is null");
    if (in->n_row_mode_entries > 0 && !in->row_mode_entries) throw ArgumentError("row_mode_entries is null");
    if (in->n_target_overrides > 0 && (!in->target_override_index || !in->target_override_value))
        throw ArgumentError("target override arrays are null");
    if (in->state_vector_scaling_mode != OLFHIP_SCALING_NONE && in->state_vector_scaling_mode != OLFHIP_SCALING_LINE_SEARCH &&
        in->state_vector_scaling_mode != OLFHIP_SCALING_MAX_VOLTAGE_CHANGE)
        throw UnsupportedError("state vector scaling: NONE, LINE_SEARCH and MAX_VOLTAGE_CHANGE are modelled");
    for (int i = 0; i < n; i++)
        if (in->base_row_mode[i] != OLFHIP_ROW_FIXED_VARIABLE && in->base_row_mode[i] != OLFHIP_ROW_EQUATION)
            throw UnsupportedError("base row mode " + std::to_string(in->base_row_mode[i]) + " is not modelled");
    for (int q = 0; q < in->n_row_mode_entries; q++) {
        const int *e = in->row_mode_entries + 3 * q;
        if (e[0] < 0 || e[0] >= K || e[1] < 0 || e[1] >= n || (e[2] != 0 && e[2] != 1))
            throw ArgumentError("row_mode_entries[" + std::to_string(q) + "] out of range");
    }
    for (int q = 0; q < in->n_target_overrides; q++)
        if (in->target_override_index[q] < 0 || in->target_override_index[q] >= (long long)K * n)
            throw ArgumentError("target_override_index[" + std::to_string(q) + "] out of range");
    hipStream_t s = c->stream;
    auto t0 = std::chrono::steady_clock::now();
    const size_t lanes = (size_t)c->nTiles * c->T;
    // the slots this solve installs into: its bank slots, or scenario k into slot k with every other slot emptied
    std::vector<int> slots(K);
    for (int k = 0; k < K; k++) slots[k] = bank ? bankSlots[k] : k;
    for (size_t q = 0; q < (size_t)K * in->max_outages; q++)
        if (in->outaged_branches[q] >= P.nBranch) throw ArgumentError("outaged branch id out of range");
    // a solve refused after this point leaves the bank's host-side slot state as it was
    struct SlotStateRestore {
        ContextImpl *c;
        std::vector<char> pending;
        std::vector<int> outages, active;
        std::vector<long long> seq;
        bool committed = false;
        ~SlotStateRestore() {
            if (committed) return;
            c->pending.swap(pending);
            c->laneOutages.swap(outages);
            c->laneActive.swap(active);
            c->installSeq.swap(seq);
        }
    } restore{c, c->pending, c->laneOutages, c->laneActive, c->installSeq};
    if (!bank) {
        std::fill(c->pending.begin(), c->pending.end(), 0);
        std::fill(c->laneOutages.begin(), c->laneOutages.end(), -1);
        std::fill(c->laneActive.begin(), c->laneActive.end(), n);
    }
    // outage lists, padded to MAX_OUTAGES
    for (int k = 0; k < K; k++) {
        int *o = &c->laneOutages[(size_t)slots[k] * MAX_OUTAGES];
        std::fill(o, o + MAX_OUTAGES, -1);
        for (int q = 0; q < in->max_outages; q++) {
            const int b = in->outaged_branches[(size_t)k * in->max_outages + q];
            o[q] = b < 0 ? -1 : b;
        }
        c->pending[slots[k]] = 1;
        c->installSeq[slots[k]] = c->installCounter++;
    }
    const std::vector<int> &outs = c->laneOutages;
    c->outages.upload(outs, s);
    if (bank && K > 0) c->slotOf.upload(slots, s);
    c->x0Base.upload(std::vector<double>(in->x0_base, in->x0_base + n), s);
    c->targetBase.upload(std::vector<double>(in->target_base, in->target_base + n), s);
    c->modeBase.upload(std::vector<int>(in->base_row_mode, in->base_row_mode + n), s);
    if (in->x0_per_scenario && in->x0_per_scenario_device)
        throw ArgumentError("x0_per_scenario and x0_per_scenario_device are exclusive");
    const double *x0PerDev = in->x0_per_scenario_device;
    if (in->x0_per_scenario) {
        c->x0Per.alloc((size_t)K * n * 8);
        HIP_OK(hipMemcpyAsync(c->x0Per.p, in->x0_per_scenario, (size_t)K * n * 8, hipMemcpyHostToDevice, s));
        x0PerDev = c->x0Per.as<double>();
    }
    const int nSeeds = (int)c->x0SeedIndex.size();
    for (int q = 0; q < nSeeds; q++)
        if (c->x0SeedIndex[q] >= (long long)K * n)
            throw ArgumentError("x0 seed " + std::to_string(q) + " names a scenario beyond n_scenarios");
    if (nSeeds > 0) {
        c->x0SeedIdx.upload(c->x0SeedIndex, s);
        c->x0SeedVal.upload(c->x0SeedValue, s);
    }
    if (in->n_row_mode_entries > 0)
        c->modeEntries.upload(std::vector<int>(in->row_mode_entries, in->row_mode_entries + 3 * in->n_row_mode_entries), s);
    if (in->n_target_overrides > 0) {
        c->targetIdx.upload(std::vector<long long>(in->target_override_index, in->target_override_index + in->n_target_overrides), s);
        c->targetVal.upload(std::vector<double>(in->target_override_value, in->target_override_value + in->n_target_overrides), s);
    }
    const size_t seedRows = in->reactive_limits_seed_shared ? 1 : (size_t)K;
    if (in->reactive_limits_seed_state && c->nRL > 0) {
        c->seed.alloc(seedRows * c->nRL);
        HIP_OK(hipMemcpyAsync(c->seed.p, in->reactive_limits_seed_state, seedRows * c->nRL, hipMemcpyHostToDevice, s));
    }
    NrParams prm{};
    prm.maxIter = in->max_iter;
    prm.tol = in->tol;
    prm.reactiveLimits = !in->skip_reactive_limits;
    prm.rlTol = in->reactive_limit_tol;
    prm.maxOuter = in->max_outer_iterations > 0 ? in->max_outer_iterations : 20;
    prm.maxSwitch = in->max_pq_pv_switch >= 0 ? in->max_pq_pv_switch : 3;
    prm.scaling = in->state_vector_scaling_mode;
    prm.maxDv = in->max_dv;
    prm.maxDphi = in->max_dphi;
    prm.refine = c->mode == OLFHIP_LU_FP32_REFINED;
    // a batch of one: the bank holds one scenario (a bank solve), or the solve installs one (newton_iteration.cu:388,
    // whose nS is the bank size)
    prm.batchOfOne = bank ? c->capacity == 1 : K == 1;
    prm.rowCriteria = c->rowCriteria ? 1 : 0;
    prm.kryDim = OLFHIP_ROBUST_KRYLOV_DIM;
    prm.kryRestarts = OLFHIP_ROBUST_RESTARTS;
    prm.kryTol = OLFHIP_ROBUST_TOL;
    if (std::isnan(in->linear_solve_fail_ratio)) throw ArgumentError("linear_solve_fail_ratio is NaN");
    prm.linFailRatio = in->linear_solve_fail_ratio > 0.0 ? in->linear_solve_fail_ratio : LIN_SOLVE_FAIL_RATIO;
    // tol <= 0 with no other criterion: exactly max_iter steps per scenario (the reference's fixed-iteration mode,
    // native/cuda/solver/outer_loop_driver.cu:165-175)
    prm.fixedIter = !(in->tol > 0.0) && !(in->conv_eps_per_eq > 0.0) && !c->rowCriteria;
    if (in->distributed_slack && c->slackNode < 0)
        throw ArgumentError("distributed_slack is set but the context has no distributed slack (olfhip_set_distributed_slack)");
    if (in->distributed_slack && !(in->slack_dist_tol >= 0.0)) throw ArgumentError("slack_dist_tol must be >= 0");
    if (!(in->conv_eps_per_eq >= 0.0)) throw ArgumentError("conv_eps_per_eq must be >= 0");
    prm.convEps = in->conv_eps_per_eq;
    for (int k = 0; k < K; k++) {
        const int a = in->active_row_counts ? in->active_row_counts[k] : n;
        if (a <= 0 || a > n) throw ArgumentError("active_row_counts[" + std::to_string(k) + "] out of range");
        c->laneActive[slots[k]] = a;
    }
    if (prm.convEps > 0.0) {
        HIP_OK(hipMemcpyAsync(c->nActive.p, c->laneActive.data(), lanes * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    prm.pauseAt = in->pause_at_ended;
    prm.lsMax = in->line_search_max_iteration > 0 ? in->line_search_max_iteration : 10;
    prm.lsFold = in->line_search_step_fold > 1.0 ? in->line_search_step_fold : 4.0 / 3.0;
    if (prm.scaling == OLFHIP_SCALING_LINE_SEARCH) c->xPrev.alloc((size_t)c->nTiles * c->T * P.nb * 2 * 8);
    prm.distributedSlack = in->distributed_slack != 0;
    prm.dsTol = in->slack_dist_tol;
    prm.dsReject = in->reject_slack_distribution_failure != 0;
    const int maxSteps = (prm.maxIter + 1) * (prm.maxOuter + 1) + 1;
    // one launch carries every tile to its end (OLFHIP_STEPS_PER_LAUNCH=k: at most k Newton steps per launch)
    prm.stepsPerLaunch = maxSteps;
    if (const char *e = std::getenv("OLFHIP_STEPS_PER_LAUNCH")) prm.stepsPerLaunch = std::max(1, std::atoi(e));
    const bool profile = std::getenv("OLFHIP_PROFILE") && std::strcmp(std::getenv("OLFHIP_PROFILE"), "0") != 0;
    if (profile) {
        c->profCap = maxSteps;
        const char *pt = std::getenv("OLFHIP_PROFILE_TILE");
        c->profTile = pt ? std::max(0, std::min(c->nTiles - 1, std::atoi(pt))) : 0;
        const size_t bytes = (size_t)PROF_SECTIONS * (maxSteps + 1) * 8;
        c->prof.alloc(bytes);
        HIP_OK(hipMemsetAsync(c->prof.p, 0, bytes, s));   // before any capture: no memset node in the graph
    } else {
        c->profCap = 0;
        c->prof.release();
    }
    const char *timeline = std::getenv("OLFHIP_TILE_TIMELINE");
    if (timeline && *timeline) {
        c->tileClock.alloc((size_t)c->nTiles * 3 * 8);
        HIP_OK(hipMemsetAsync(c->tileClock.p, 0, (size_t)c->nTiles * 3 * 8, s));   // before any capture
    } else {
        c->tileClock.release();
    }
    restore.committed = true;
    const DevPlan dp = c->st->dp;
    DevState S = c->devState();
    S.nInstalled = bank ? 0 : K;   // the default install's scenario count (slots 0 .. K-1)
    const DevRL R = c->devRL();
    const bool useGraph = !std::getenv("OLFHIP_NO_GRAPH");
    int steps = 0;
    withT(c->T, [&](auto TC) {
        constexpr int T = decltype(TC)::value;
        const size_t work = (size_t)c->nTiles * P.nb * T;
        const int *slotMap = bank ? c->slotOf.as<int>() : nullptr;
        const signed char *seedRow =
            (in->reactive_limits_seed_state && c->nRL > 0) ? c->seed.as<signed char>() : nullptr;
        auto enqueueInstall = [&]() {
            zeroIntKernel<<<1, 64, 0, s>>>(S.ended, 1);
            if (bank) {
                if (K > 0)
                    bankInstallKernel<T><<<gridFor((size_t)K * P.nb, 256), 256, 0, s>>>(
                        dp, S, K, slotMap, c->x0Base.as<double>(), x0PerDev, c->targetBase.as<double>(),
                        c->modeBase.as<int>());
            } else {
                installKernel<T><<<gridFor(work, 256), 256, 0, s>>>(dp, S, c->x0Base.as<double>(), x0PerDev,
                                                                    c->targetBase.as<double>(),
                                                                    c->modeBase.as<int>());
            }
            if (in->n_row_mode_entries + in->n_target_overrides + nSeeds > 0)
                overrideKernel<T><<<gridFor(in->n_row_mode_entries + in->n_target_overrides + nSeeds, 256), 256, 0,
                                    s>>>(dp, S, in->n_row_mode_entries, c->modeEntries.as<int>(),
                                         in->n_target_overrides, c->targetIdx.as<long long>(),
                                         c->targetVal.as<double>(), nSeeds, c->x0SeedIdx.as<long long>(),
                                         c->x0SeedVal.as<double>(), slotMap);
            if (bank) {
                if (K > 0)
                    bankLaneInitKernel<T><<<gridFor((size_t)K * (1 + (size_t)c->nRL), 256), 256, 0, s>>>(
                        dp, S, R, K, slotMap, seedRow, seedRows == 1 ? 0 : c->nRL, c->modeBase.as<int>());
                tileRunningKernel<T><<<gridFor(c->nTiles, 256), 256, 0, s>>>(S);
            } else {
                laneInitKernel<T><<<gridFor(lanes * (1 + (size_t)c->nRL), 256), 256, 0, s>>>(
                    dp, S, R, seedRow, seedRows == 1 ? 0 : c->nRL, c->modeBase.as<int>());
            }
        };
        // the per-row criterion runs in the kernels' ROWCRIT variants (the same code plus its test)
        auto launchStep = [&](auto rowCrit) {
            constexpr bool RC = decltype(rowCrit)::value;
            if (c->mode == OLFHIP_LU_FP32_REFINED) {
                nrStepKernel<T, float, false, RC><<<c->nTiles, c->NT, c->top.ldsBytes, s>>>(dp, S, R, prm, c->top.top);
            } else if (c->mode == OLFHIP_LU_ROBUST) {
                if constexpr (T == 1)
                    nrStepKernel<1, double, true, RC><<<c->nTiles, c->NT, c->top.ldsBytes, s>>>(dp, S, R, prm,
                                                                                           c->top.top);
            } else {
                nrStepKernel<T, double, false, RC><<<c->nTiles, c->NT, c->top.ldsBytes, s>>>(dp, S, R, prm, c->top.top);
            }
        };
        auto enqueueStep = [&]() {
            if (prm.rowCriteria) launchStep(std::true_type{});
            else launchStep(std::false_type{});
            countRunningKernel<<<1, 256, 0, s>>>(S.tileRunning, c->nTiles, S.ended, c->count.as<int>());
            HIP_OK(hipMemcpyAsync(c->pinnedCount, c->count.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        };
        // Captures `enqueue` and brings `exec` to it: updated in place when its sequence is the cached graph's
        // (only kernel arguments changed), instantiated again otherwise.
        auto graphOf = [&](hipGraphExec_t &exec, auto &&enqueue) {
            hipGraph_t graph = nullptr;
            HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                enqueue();
            } catch (...) {
                if (hipStreamEndCapture(s, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
                throw;
            }
            HIP_OK(hipStreamEndCapture(s, &graph));
            bool updated = false;
            if (exec) {
                hipGraphExecUpdateResult result;
                hipGraphNode_t failed = nullptr;
                updated = hipGraphExecUpdate(exec, graph, &failed, &result) == hipSuccess &&
                          result == hipGraphExecUpdateSuccess;
                if (!updated) {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(exec);
                    exec = nullptr;
                }
            }
            if (!updated) {
                const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                if (e != hipSuccess) {
                    (void)hipGraphDestroy(graph);
                    HIP_OK(e);
                }
                c->graphInstantiations++;
            } else {
                c->graphUpdates++;
            }
            (void)hipGraphDestroy(graph);
        };
        c->lastStepMs = 0.0;
        const int maxLaunches = (maxSteps + prm.stepsPerLaunch - 1) / prm.stepsPerLaunch;
        if (useGraph) {
            graphOf(c->installGraph, enqueueInstall);
            graphOf(c->stepGraph, enqueueStep);
            HIP_OK(hipGraphLaunch(c->installGraph, s));
        } else {
            enqueueInstall();
        }
        HIP_OK(hipGetLastError());
        for (steps = 0; steps < maxLaunches; steps++) {
            HIP_OK(hipEventRecord(c->evStart, s));
            if (useGraph) HIP_OK(hipGraphLaunch(c->stepGraph, s));
            else enqueueStep();
            HIP_OK(hipEventRecord(c->evStop, s));
            HIP_OK(hipStreamSynchronize(s));
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, c->evStart, c->evStop));
            c->lastStepMs += ms;
            if (traceOn())
                std::fprintf(stderr, "OLFHIP step %d %.3f ms, tiles still running %d\n", steps, ms, *c->pinnedCount);
            if (c->pinnedCount[0] == 0 || (prm.pauseAt > 0 && c->pinnedCount[1] >= prm.pauseAt)) {
                steps++;
                break;
            }
        }
    });
    HIP_OK(hipGetLastError());
    if (profile) {
        std::vector<unsigned long long> h((size_t)PROF_SECTIONS * (maxSteps + 1));
        HIP_OK(hipMemcpy(h.data(), c->prof.p, h.size() * 8, hipMemcpyDeviceToHost));
        int rateKHz = 0;
        HIP_OK(hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, c->st->device));
        const double usPerTick = rateKHz > 0 ? 1e3 / rateKHz : 0.0;
        // stamps in the order a step makes them: the end of each named section
        const int order[13] = {1, 2, 9, 3, 10, 4, 5, 11, 12, 13, 6, 7, 8};
        const char *names[13] = {"fill", "test+outer", "sparse-factor+fwd", "top-factor+fwd", "top-bwd", "sparse-bwd",
                                 "refine-J", "refine-sparse-fwd", "refine-top-fwd", "refine-top-bwd",
                                 "refine-sparse-bwd", "guard", "update"};
        const unsigned long long launches = std::min<unsigned long long>(h[0], (unsigned long long)maxSteps);
        for (unsigned long long k = 0; k < launches; k++) {
            const unsigned long long *r = h.data() + PROF_SECTIONS * (k + 1);
            std::fprintf(stderr, "OLFHIP_PROFILE T=%d NT=%d tile %d launch %llu:", c->T, c->NT, c->profTile, k);
            unsigned long long prev = r[0];
            for (int q = 0; q < 13; q++) {
                const unsigned long long v = r[order[q]];
                if (v == 0) continue;
                std::fprintf(stderr, " %s=%.1fus", names[q], (v - prev) * usPerTick);
                prev = v;
            }
            std::fprintf(stderr, " total=%.1fus\n", (prev - r[0]) * usPerTick);
        }
    }
    if (timeline && *timeline) {
        // the tile timeline, appended to the file as a block per solve: tile, start and end (us from the earliest
        // start), Newton steps
        std::vector<unsigned long long> tc((size_t)c->nTiles * 3);
        HIP_OK(hipMemcpy(tc.data(), c->tileClock.p, tc.size() * 8, hipMemcpyDeviceToHost));
        int rateKHz = 0;
        HIP_OK(hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, c->st->device));
        unsigned long long t0c = ~0ull;
        for (int i = 0; i < c->nTiles; i++)
            if (tc[3 * i]) t0c = std::min(t0c, tc[3 * i]);
        if (FILE *f = std::fopen(timeline, "a")) {
            std::fprintf(f, "solve K=%d T=%d NT=%d tiles=%d t0=%.1f\n", K, c->T, c->NT, c->nTiles,
                         t0c * 1e3 / rateKHz);   // t0: the earliest start, us on the device's wall clock
            for (int i = 0; i < c->nTiles; i++)
                if (tc[3 * i])
                    std::fprintf(f, "%d %.1f %.1f %llu\n", i, (tc[3 * i] - t0c) * 1e3 / rateKHz,
                                 (tc[3 * i + 1] - t0c) * 1e3 / rateKHz, tc[3 * i + 2]);
            std::fclose(f);
        }
    }
    // results: every slot whose scenario has ended and was not handed back yet, in the order the scenarios were
    // installed (a solve that runs to its end in the default install: scenarios 0 .. K-1 in order)
    const int L = (int)lanes;
    std::vector<int> stAll(L), itAll(L), totAll(L);
    std::vector<double> nmAll(L);
    HIP_OK(hipMemcpyAsync(stAll.data(), c->status.p, (size_t)L * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(itAll.data(), c->it.p, (size_t)L * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(totAll.data(), c->total.p, (size_t)L * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(nmAll.data(), c->norm.p, (size_t)L * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const bool paused = prm.pauseAt > 0 && c->pinnedCount[0] > 0 && c->pinnedCount[1] >= prm.pauseAt;
    c->handedBack.clear();
    for (int l = 0; l < L; l++) {
        if (!c->pending[l]) continue;
        if (stAll[l] == LANE_RUNNING) {
            if (paused) continue;                 // goes on at the next call
            stAll[l] = ST_REJECTED;               // batched-iteration cap (newton_iteration.cu:410-429)
        }
        c->handedBack.push_back(l);
        c->pending[l] = 0;
    }
    std::sort(c->handedBack.begin(), c->handedBack.end(),
              [&](int a, int b) { return c->installSeq[a] < c->installSeq[b]; });
    const int H = (int)c->handedBack.size();
    c->lastSteps = steps;
    // the readbacks after the solve (states, outer-loop state, violations): by scenario in the default install,
    // by slot after a bank or paused solve (a slot not handed back by this solve reads as not converged)
    c->bankReadback = bank || prm.pauseAt > 0;
    if (c->bankReadback) {
        c->lastInstalled = c->capacity;
        c->lastStatus.assign(c->capacity, LANE_RUNNING);
        for (int l : c->handedBack) c->lastStatus[l] = stAll[l];
        std::vector<int> now(lanes, 0);
        for (int l : c->handedBack) now[l] = 1;
        c->handedNow.upload(now, s);
    } else {
        c->lastInstalled = K;
        c->lastStatus.assign(stAll.begin(), stAll.begin() + K);
    }
    c->lastOutages = outs;
    const bool wantStates = out && out->results && in->states_to_host && H > 0;
    const int gathered = H > 0 ? *std::max_element(c->handedBack.begin(), c->handedBack.end()) + 1 : 0;
    std::vector<double> xs;
    if (wantStates) {
        c->states.alloc((size_t)gathered * n * 8);
        withT(c->T, [&](auto TC) {
            constexpr int T = decltype(TC)::value;
            gatherStatesKernel<T><<<gridFor((size_t)gathered * P.nb, 256), 256, 0, s>>>(dp, S, gathered,
                                                                                       c->states.as<double>());
        });
        xs.resize((size_t)gathered * n);
        HIP_OK(hipMemcpyAsync(xs.data(), c->states.p, (size_t)gathered * n * 8, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    if (out) {
        if (out->n_handed_back) *out->n_handed_back = H;
        for (int r = 0; r < H; r++) {
            const int l = c->handedBack[r];
            if (out->status) out->status[r] = stAll[l];
            if (out->iterations) out->iterations[r] = itAll[l];
            if (out->total_iterations) out->total_iterations[r] = totAll[l];
            if (out->norm) out->norm[r] = nmAll[l];
            if (out->results) {
                double *row = out->results + (size_t)r * (n + 3);
                if (wantStates) std::memcpy(row, xs.data() + (size_t)l * n, (size_t)n * 8);
                row[n] = itAll[l];
                row[n + 1] = nmAll[l];
                row[n + 2] = stAll[l];
            }
        }
    }
    c->lastSolveMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (traceOn()) std::fprintf(stderr, "OLFHIP solve K=%d T=%d NT=%d steps=%d %.3f ms\n", K, c->T, c->NT, steps, c->lastSolveMs);
}

static void readStates(ContextImpl *c, int K, double *x) {
    if (K <= 0 || K > c->lastInstalled) throw ArgumentError("n_scenarios exceeds the last solve's installed count");
    if (!x) throw ArgumentError("x is null");
    const Plan &P = c->st->plan;
    c->states.alloc((size_t)K * P.n * 8);
    const DevPlan dp = c->st->dp;
    const DevState S = c->devState();
    withT(c->T, [&](auto TC) {
        constexpr int T = decltype(TC)::value;
        gatherStatesKernel<T><<<gridFor((size_t)K * P.nb, 256), 256, 0, c->stream>>>(dp, S, K, c->states.as<double>());
    });
    HIP_OK(hipMemcpyAsync(x, c->states.p, (size_t)K * P.n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
}

// ---- limit violations: the table, the device screen and the exact host confirmation
struct LimitTableImpl {
    StructureImpl *st = nullptr;
    int nBranches = 0, nGroups = 0, nLimits = 0, nBuses = 0;
    std::vector<int> branch, groupStart, groupSide, groupQty, limitStart, groupBranch, limitGroup;
    std::vector<double> threshold, busHigh, busLow;
    int nAngle = 0;
    std::vector<int> angleFrom, angleTo;
    std::vector<double> angleHigh, angleLow;
    DBuf dBranch, dGroupStart, dGroupSide, dGroupQty, dLimitStart, dThreshold, dBusHigh, dBusLow, dCoef, dNode;
    DBuf records, counts, screenLane, req, vals, isl, near, packed, offsets;
    DevLimits dev() const {
        return DevLimits{nBranches, nGroups, nLimits, nBuses, dBranch.as<int>(), dGroupStart.as<int>(),
                         dGroupSide.as<int>(), dGroupQty.as<int>(), dLimitStart.as<int>(), dThreshold.as<double>(),
                         dBusHigh.as<double>(), dBusLow.as<double>(), dCoef.as<double>(), dNode.as<int>()};
    }
    ~LimitTableImpl() { release(st); }
};

static LimitTableImpl *createLimitTable(StructureImpl *st, const olfhip_limit_table_desc *d) {
    if (!st || !d) throw ArgumentError("null argument");
    const Plan &P = st->plan;
    if (d->n_branches < 0 || d->n_buses < 0 || d->n_buses > P.nb) throw ArgumentError("table sizes out of range");
    if (d->n_branches > 0 && (!d->branch || !d->group_start || !d->group_side || !d->group_quantity ||
                              !d->limit_start || !d->limit_threshold))
        throw ArgumentError("branch limit arrays are null");
    if (d->n_buses > 0 && (!d->bus_high || !d->bus_low)) throw ArgumentError("bus limit arrays are null");
    if (d->n_angle_limits < 0) throw ArgumentError("n_angle_limits must not be negative");
    if (d->n_angle_limits > 0 && (!d->angle_from || !d->angle_to || !d->angle_high || !d->angle_low))
        throw ArgumentError("voltage-angle limit arrays are null");
    for (int a = 0; a < d->n_angle_limits; a++)
        if (d->angle_from[a] < 0 || d->angle_from[a] >= P.nb || d->angle_to[a] < 0 || d->angle_to[a] >= P.nb)
            throw ArgumentError("voltage-angle limit bus out of range");
    std::unique_ptr<LimitTableImpl> t(new LimitTableImpl());
    t->nBranches = d->n_branches;
    t->nBuses = d->n_buses;
    if (d->n_branches > 0) {
        if (d->group_start[0] != 0) throw ArgumentError("group_start[0] must be 0");
        t->nGroups = d->group_start[d->n_branches];
        for (int b = 0; b < d->n_branches; b++) {
            if (d->branch[b] < 0 || d->branch[b] >= P.nBranch) throw ArgumentError("table branch out of range");
            if (d->group_start[b + 1] < d->group_start[b]) throw ArgumentError("group_start must not decrease");
        }
        if (t->nGroups > 0 && d->limit_start[0] != 0) throw ArgumentError("limit_start[0] must be 0");
        t->nLimits = t->nGroups > 0 ? d->limit_start[t->nGroups] : 0;
        for (int g = 0; g < t->nGroups; g++) {
            if (d->group_side[g] != 1 && d->group_side[g] != 2) throw ArgumentError("group side must be 1 or 2");
            if (d->group_quantity[g] < OLFHIP_QUANTITY_CURRENT || d->group_quantity[g] > OLFHIP_QUANTITY_APPARENT_POWER)
                throw ArgumentError("unknown group quantity");
            if (d->limit_start[g + 1] < d->limit_start[g]) throw ArgumentError("limit_start must not decrease");
        }
        t->branch.assign(d->branch, d->branch + d->n_branches);
        t->groupStart.assign(d->group_start, d->group_start + d->n_branches + 1);
        t->groupSide.assign(d->group_side, d->group_side + t->nGroups);
        t->groupQty.assign(d->group_quantity, d->group_quantity + t->nGroups);
        t->limitStart.assign(d->limit_start, d->limit_start + t->nGroups + 1);
        t->threshold.assign(d->limit_threshold, d->limit_threshold + t->nLimits);
    } else {
        t->groupStart.assign(1, 0);
        t->limitStart.assign(1, 0);
    }
    t->groupBranch.assign(t->nGroups, 0);
    for (int b = 0; b < t->nBranches; b++)
        for (int g = t->groupStart[b]; g < t->groupStart[b + 1]; g++) t->groupBranch[g] = b;
    t->limitGroup.assign(t->nLimits, 0);
    for (int g = 0; g < t->nGroups; g++)
        for (int l = t->limitStart[g]; l < t->limitStart[g + 1]; l++) t->limitGroup[l] = g;
    if (d->n_buses > 0) {
        t->busHigh.assign(d->bus_high, d->bus_high + d->n_buses);
        t->busLow.assign(d->bus_low, d->bus_low + d->n_buses);
    }
    t->nAngle = d->n_angle_limits;
    if (t->nAngle > 0) {
        t->angleFrom.assign(d->angle_from, d->angle_from + t->nAngle);
        t->angleTo.assign(d->angle_to, d->angle_to + t->nAngle);
        t->angleHigh.assign(d->angle_high, d->angle_high + t->nAngle);
        t->angleLow.assign(d->angle_low, d->angle_low + t->nAngle);
    }
    st->ensureDevice();
    t->dBranch.upload(t->branch);
    t->dGroupStart.upload(t->groupStart);
    t->dGroupSide.upload(t->groupSide);
    t->dGroupQty.upload(t->groupQty);
    t->dLimitStart.upload(t->limitStart);
    t->dThreshold.upload(t->threshold);
    t->dBusHigh.upload(t->busHigh);
    t->dBusLow.upload(t->busLow);
    t->dCoef.upload(P.branchCoef);
    t->dNode.upload(P.branchNode);
    HIP_OK(hipDeviceSynchronize());
    st->refs++;
    t->st = st;
    return t.release();
}

// Screens the last solve's converged scenarios on the device: counts [K], records [K][capacity] on the host.
// Device screen of the last solve. Without `packed`: counts [K] and records [K][capacity] (a row's first counts[k]
// places hold scenario k's records, all of them when counts[k] <= capacity). With `packed`: records holds the rows'
// used places one after another, scenario k's from (*packed)[k] to (*packed)[k + 1].
static void screenViolations(ContextImpl *c, LimitTableImpl *t, double band, int capacity, std::vector<int> &counts,
                             std::vector<int> &records, std::vector<int> *packed = nullptr) {
    if (c->st != t->st) throw ArgumentError("the limit table belongs to another structure");
    const int K = c->lastInstalled;
    if (K <= 0 || (int)c->lastStatus.size() != K) throw ArgumentError("no solve to screen");
    if (capacity <= 0) throw ArgumentError("capacity must be positive");
    if (!(band >= 0.0)) throw ArgumentError("band must be non-negative");
    hipStream_t s = c->stream;
    std::vector<int> lane(K);
    for (int k = 0; k < K; k++) lane[k] = c->lastStatus[k] == ST_CONVERGED;
    t->screenLane.upload(lane, s);
    t->counts.alloc((size_t)K * 4);
    t->records.alloc((size_t)K * capacity * 4);
    zeroIntKernel<<<gridFor(K, 256), 256, 0, s>>>(t->counts.as<int>(), K);
    const DevPlan dp = c->st->dp;
    const DevState S = c->devState();
    const DevLimits L = t->dev();
    withT(c->T, [&](auto TC) {
        constexpr int T = decltype(TC)::value;
        if (t->nBranches > 0)
            screenBranchesKernel<T><<<gridFor((size_t)K * t->nBranches, 256), 256, 0, s>>>(
                dp, S, L, c->modeBase.as<int>(), t->screenLane.as<int>(), K, band, capacity, t->records.as<int>(),
                t->counts.as<int>());
        if (t->nBuses > 0)
            screenBusesKernel<T><<<gridFor((size_t)K * t->nBuses, 256), 256, 0, s>>>(
                dp, S, L, c->modeBase.as<int>(), t->screenLane.as<int>(), K, band, capacity, t->records.as<int>(),
                t->counts.as<int>());
    });
    HIP_OK(hipGetLastError());
    counts.resize(K);
    if (!packed) {
        records.resize((size_t)K * capacity);
        HIP_OK(hipMemcpyAsync(counts.data(), t->counts.p, (size_t)K * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(records.data(), t->records.p, records.size() * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        return;
    }
    // the scenarios' records packed on the device, so that only the records found cross to the host
    packed->assign((size_t)K + 1, 0);
    t->offsets.alloc(((size_t)K + 1) * 4);
    t->packed.alloc((size_t)K * capacity * 4);
    packOffsetsKernel<<<1, PACK_THREADS, 0, s>>>(t->counts.as<int>(), K, capacity, t->offsets.as<int>());
    packRecordsKernel<<<K, 256, 0, s>>>(t->records.as<int>(), t->counts.as<int>(), t->offsets.as<int>(), capacity,
                                        t->packed.as<int>());
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(counts.data(), t->counts.p, (size_t)K * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(packed->data(), t->offsets.p, ((size_t)K + 1) * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const int total = (*packed)[K];
    if (total < 0 || (long long)total > (long long)K * capacity)
        throw DeviceError("screened record count out of range");
    records.resize((size_t)total);
    if (total > 0) {
        HIP_OK(hipMemcpyAsync(records.data(), t->packed.p, (size_t)total * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
}

// Which scenarios of the last solve a state difference of relative size `band` could change the violations of
// (nearLimitsKernel): flags [K] on the host.
static void ambiguousScenarios(ContextImpl *c, LimitTableImpl *t, double band, int nScenarios, int *flags) {
    if (c->st != t->st) throw ArgumentError("the limit table belongs to another structure");
    const int K = c->lastInstalled;
    if (K <= 0 || (int)c->lastStatus.size() != K) throw ArgumentError("no solve to check");
    if (nScenarios != K) throw ArgumentError("n_scenarios must be the scenario count of the context's last solve");
    if (!(band >= 0.0)) throw ArgumentError("band must be non-negative");
    hipStream_t s = c->stream;
    t->near.alloc((size_t)K * 4);
    zeroIntKernel<<<gridFor(K, 256), 256, 0, s>>>(t->near.as<int>(), K);
    const DevPlan dp = c->st->dp;
    const DevState S = c->devState();
    const DevLimits L = t->dev();
    // one workgroup per scenario with its node states in LDS where they fit (OLFHIP_NEAR_GATHER=1: the gathering
    // kernel, one thread per scenario and item)
    const size_t stagedBytes = (size_t)dp.nb * 17;
    int perBlock = 0;
    HIP_OK(hipDeviceGetAttribute(&perBlock, hipDeviceAttributeMaxSharedMemoryPerBlock, c->st->device));
    const char *gather = std::getenv("OLFHIP_NEAR_GATHER");
    const bool staged = stagedBytes <= (size_t)perBlock && !(gather && std::strcmp(gather, "0") != 0);
    const int *only = c->bankReadback ? c->handedNow.as<int>() : nullptr;
    if (t->nBranches + t->nBuses > 0)
        withT(c->T, [&](auto TC) {
            constexpr int T = decltype(TC)::value;
            if (staged) {
                HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void *>(&nearLimitsStagedKernel<T>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)stagedBytes));
                nearLimitsStagedKernel<T><<<K, 1024, stagedBytes, s>>>(dp, S, L, c->modeBase.as<int>(), band, only,
                                                                      t->near.as<int>());
            } else {
                nearLimitsKernel<T><<<gridFor((size_t)K * (t->nBranches + t->nBuses), 256), 256, 0, s>>>(
                    dp, S, L, c->modeBase.as<int>(), K, band, only, t->near.as<int>());
            }
        });
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(flags, t->near.p, (size_t)K * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
}

// ---- the allocation-free screen (olfhip_screen_violations_no_alloc). Its lane flags, candidate counts and candidate
// records live in `scratch`, a file-scope device array the code object carries (each device that loads the library
// holds one until the process ends), so a call allocates no device memory: the context and the limit table bring
// everything else. Every call on one device shares that device's array and leaves its contents behind: the counts are
// cleared by each call, the records are only read up to their counts. Calls that overlap on one device race.
static __device__ __align__(256) unsigned char scratch[OLFHIP_NO_ALLOC_SCRATCH_BYTES];
constexpr int SCRATCH_DEVICES = 64;
static std::mutex gScratchLock;
static unsigned char *gScratchOf[SCRATCH_DEVICES] = {};

// The device address of `scratch` on `device`, which must be current: looked up once per device and kept. (In host
// code the array's name is only a stand-in for the device symbol; kernels and copies take this address.)
static unsigned char *scratchOnDevice(int device) {
    int current = -1;
    HIP_OK(hipGetDevice(&current));
    if (current != device) throw ArgumentError("the context's device is not the current device");
    if (device < 0 || device >= SCRATCH_DEVICES) throw ArgumentError("device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(gScratchLock);
    if (!gScratchOf[device]) {
        void *p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw DeviceError("hipGetSymbolAddress gave no address for the scratch array");
        gScratchOf[device] = static_cast<unsigned char *>(p);
    }
    return gScratchOf[device];
}

static size_t roundUp256(size_t bytes) { return (bytes + 255) / 256 * 256; }

// screenViolations with its buffers in `scratch` and its outputs copied straight into the caller's arrays:
// [lane flags K ints][counts K ints][records K x capacity ints], sections on 256-byte boundaries.
static void screenViolationsNoAlloc(ContextImpl *c, LimitTableImpl *t, double band, int nScenarios, int capacity,
                                    int *counts, int *records) {
    if (c->st != t->st) throw ArgumentError("the limit table belongs to another structure");
    const int K = c->lastInstalled;
    if (K <= 0 || (int)c->lastStatus.size() != K) throw ArgumentError("no solve to screen");
    if (nScenarios != K) throw ArgumentError("n_scenarios must be the scenario count of the context's last solve");
    if (capacity <= 0) throw ArgumentError("capacity must be positive");
    if (!(band >= 0.0)) throw ArgumentError("band must be non-negative");
    const size_t flagBytes = roundUp256((size_t)K * 4);
    const size_t recordBytes = (size_t)K * (size_t)capacity * 4;
    if (recordBytes / 4 / (size_t)capacity != (size_t)K || 2 * flagBytes + recordBytes > OLFHIP_NO_ALLOC_SCRATCH_BYTES)
        throw ArgumentError("n_scenarios x capacity records exceed the allocation-free scratch of " +
                            std::to_string(OLFHIP_NO_ALLOC_SCRATCH_BYTES >> 20) + " MiB (use olfhip_screen_violations)");
    unsigned char *base = scratchOnDevice(c->st->device);
    int *lane = reinterpret_cast<int *>(base);
    int *cnt = reinterpret_cast<int *>(base + flagBytes);
    int *rec = reinterpret_cast<int *>(base + 2 * flagBytes);
    hipStream_t s = c->stream;
    convergedLaneFlagsKernel<<<gridFor(K, 256), 256, 0, s>>>(
        c->status.as<int>(), c->bankReadback ? c->handedNow.as<int>() : nullptr, K, lane, cnt);
    const DevPlan dp = c->st->dp;
    const DevState S = c->devState();
    const DevLimits L = t->dev();
    withT(c->T, [&](auto TC) {
        constexpr int T = decltype(TC)::value;
        if (t->nBranches > 0)
            screenBranchesKernel<T><<<gridFor((size_t)K * t->nBranches, 256), 256, 0, s>>>(
                dp, S, L, c->modeBase.as<int>(), lane, K, band, capacity, rec, cnt);
        if (t->nBuses > 0)
            screenBusesKernel<T><<<gridFor((size_t)K * t->nBuses, 256), 256, 0, s>>>(
                dp, S, L, c->modeBase.as<int>(), lane, K, band, capacity, rec, cnt);
    });
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(counts, cnt, (size_t)K * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(records, rec, recordBytes, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
}

// Exact FP64 flows of closed branch e from its end states: side 1 (p, q, v) and side 2 (p, q, v).
static void branchFlows(const Plan &P, int e, double2 x1, double2 x2, double f[6]) {
    const double *c = &P.branchCoef[8 * (size_t)e];
    const double v1 = x1.y, v2 = x2.y;
    const double t1 = c[3] - x1.x + x2.x, t2 = c[6] - x2.x + x1.x;
    const double bvv = c[1] * v1 * v2;
    f[0] = c[0] * v1 * v1 - bvv * std::sin(t1);
    f[1] = c[2] * v1 * v1 - bvv * std::cos(t1);
    f[2] = v1;
    f[3] = c[4] * v2 * v2 - bvv * std::sin(t2);
    f[4] = c[5] * v2 * v2 - bvv * std::cos(t2);
    f[5] = v2;
}

// Screen, then the exact decisions: every candidate re-tested with the exact comparison on the host's FP64 value,
// the first confirmed limit of a group reported and the rest of the group skipped (GpuLimitScanner.java:178-262,
// 394-410); a scenario whose candidates overflowed the capacity is scanned in full (all its groups and buses in table
// order, GpuLimitScanner.java:95-131). Output ordered by scenario, then branches in table order, then buses, then
// the voltage-angle limits, which the screen does not cover: both buses of every limit are gathered for every
// converged scenario and decided here (GpuLimitScanner.java:239-267).
static void detectViolations(ContextImpl *c, LimitTableImpl *t, double band, int capacity,
                             std::vector<olfhip_violation> &out) {
    std::vector<int> counts, records, offsets;
    screenViolations(c, t, band, capacity, counts, records, &offsets);
    const Plan &P = c->st->plan;
    const int K = c->lastInstalled;
    // per scenario, its sorted candidate records (all records when it overflowed)
    std::vector<std::vector<int>> cand(K);
    std::vector<int> allRecords;
    for (int k = 0; k < K; k++) {
        if (c->lastStatus[k] != ST_CONVERGED) continue;
        if (counts[k] > capacity) {
            if (allRecords.empty()) {
                for (int l = 0; l < t->nLimits; l++) allRecords.push_back(l);
                for (int i = 0; i < t->nBuses; i++) {
                    allRecords.push_back(t->nLimits + 2 * i);
                    allRecords.push_back(t->nLimits + 2 * i + 1);
                }
            }
            cand[k] = allRecords;
        } else {
            cand[k].assign(records.begin() + offsets[k], records.begin() + offsets[k + 1]);
            std::sort(cand[k].begin(), cand[k].end());
        }
    }
    // node states to gather: both ends of each candidate branch, the bus of each candidate voltage
    std::vector<int> req;
    std::vector<std::vector<int>> at(K);   // per scenario, per candidate: index of its first gathered node
    std::vector<int> angleAt(K, -1);       // per converged scenario: index of its first angle-limit node pair
    for (int k = 0; k < K; k++) {
        if (t->nAngle > 0 && c->lastStatus[k] == ST_CONVERGED) {
            angleAt[k] = (int)(req.size() / 2);
            for (int a = 0; a < t->nAngle; a++)
                req.insert(req.end(), {k, P.pos[t->angleFrom[a]], k, P.pos[t->angleTo[a]]});
        }
        at[k].resize(cand[k].size());
        for (size_t q = 0; q < cand[k].size(); q++) {
            const int r = cand[k][q];
            at[k][q] = (int)(req.size() / 2);
            if (r < t->nLimits) {
                const int e = t->branch[t->groupBranch[t->limitGroup[r]]];
                req.insert(req.end(), {k, P.branchNode[2 * e], k, P.branchNode[2 * e + 1]});
            } else {
                req.insert(req.end(), {k, P.pos[(r - t->nLimits) / 2]});
            }
        }
    }
    const int nReq = (int)(req.size() / 2);
    std::vector<double2> xv(nReq);
    std::vector<int> iv(nReq);
    if (nReq > 0) {
        hipStream_t s = c->stream;
        t->req.upload(req, s);
        t->vals.alloc((size_t)nReq * 16);
        t->isl.alloc((size_t)nReq * 4);
        const DevPlan dp = c->st->dp;
        const DevState S = c->devState();
        withT(c->T, [&](auto TC) {
            constexpr int T = decltype(TC)::value;
            gatherNodesKernel<T><<<gridFor(nReq, 256), 256, 0, s>>>(dp, S, c->modeBase.as<int>(), nReq,
                                                                    t->req.as<int2>(), t->vals.as<double2>(),
                                                                    t->isl.as<int>());
        });
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(xv.data(), t->vals.p, (size_t)nReq * 16, hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(iv.data(), t->isl.p, (size_t)nReq * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    out.clear();
    for (int k = 0; k < K; k++) {
        int reportedGroup = -1;
        const int *outs = &c->lastOutages[(size_t)k * MAX_OUTAGES];
        for (size_t q = 0; q < cand[k].size(); q++) {
            const int r = cand[k][q];
            const int a = at[k][q];
            if (r < t->nLimits) {
                const int g = t->limitGroup[r];
                if (g == reportedGroup) continue;   // the group already reported its most serious limit
                const int e = t->branch[t->groupBranch[g]];
                bool isOut = false;
                for (int m = 0; m < MAX_OUTAGES; m++) isOut |= outs[m] == e;
                if (isOut || iv[a] || iv[a + 1]) continue;   // out of service in this scenario
                double f[6];
                branchFlows(P, e, xv[a], xv[a + 1], f);
                const int side = t->groupSide[g], qty = t->groupQty[g];
                const double p = side == 1 ? f[0] : f[3], qq = side == 1 ? f[1] : f[4], v = side == 1 ? f[2] : f[5];
                // the compared value is the screen's (comparedQuantity); active power reports its signed value
                const double compared = comparedQuantity(qty, p, qq, v);
                if (!(compared > t->threshold[r])) continue;   // NaN never fires
                const double value = qty == OLFHIP_QUANTITY_ACTIVE_POWER ? p : compared;
                out.push_back(olfhip_violation{k, OLFHIP_VIOLATION_BRANCH, t->groupBranch[g], side, qty, r, value,
                                               t->threshold[r]});
                reportedGroup = g;
            } else {
                if (iv[a]) continue;
                const int i = (r - t->nLimits) / 2;
                const double v = xv[a].y;
                if (((r - t->nLimits) & 1) == 0) {
                    if (v > t->busHigh[i])
                        out.push_back(olfhip_violation{k, OLFHIP_VIOLATION_HIGH_VOLTAGE, i, 0, 0, -1, v, t->busHigh[i]});
                } else if (v < t->busLow[i]) {
                    out.push_back(olfhip_violation{k, OLFHIP_VIOLATION_LOW_VOLTAGE, i, 0, 0, -1, v, t->busLow[i]});
                }
            }
        }
        if (angleAt[k] < 0) continue;
        for (int a = 0; a < t->nAngle; a++) {
            const int j = angleAt[k] + 2 * a;
            if (iv[j] || iv[j + 1]) continue;      // a bus out of service in this scenario
            const double diff = xv[j + 1].x - xv[j].x;
            // NaN limits make the comparisons false
            if (diff > t->angleHigh[a])
                out.push_back(olfhip_violation{k, OLFHIP_VIOLATION_HIGH_VOLTAGE_ANGLE, a, 0, 0, -1, diff,
                                               t->angleHigh[a]});
            if (diff < t->angleLow[a])
                out.push_back(olfhip_violation{k, OLFHIP_VIOLATION_LOW_VOLTAGE_ANGLE, a, 0, 0, -1, diff,
                                               t->angleLow[a]});
        }
    }
}

static void luFactorizeAndSolve(StructureImpl *st, int mode, int capacity, const double *values, const double *rhs,
                                int nLanes, const int *laneMap, double *solutions, int *clampsOut) {
    if (!st || !values || !rhs || !laneMap || !solutions) throw ArgumentError("null argument");
    if (nLanes <= 0 || capacity <= 0) throw ArgumentError("lane counts must be positive");
    if (!knownMode(mode)) throw ArgumentError("unknown mode");
    if (mode == OLFHIP_LU_ROBUST) throw ArgumentError("the robust mode is a Newton-solve mode (its GMRES needs J)");
    st->ensureDevice();
    const Plan &P = st->plan;
    const int nnz = (int)P.colIdx.size();
    const int T = nLanes >= 8 ? 8 : 1;
    const int nTiles = (nLanes + T - 1) / T;
    const size_t vb = luBytes(mode);
    std::vector<char> lu((size_t)nTiles * P.nSlots * 4 * T * vb, 0);
    std::vector<double> b((size_t)nTiles * P.nb * 2 * T, 0.0);
    for (int i = 0; i < nLanes; i++) {
        const int lane = laneMap[i];
        if (lane < 0 || lane >= capacity) throw ArgumentError("lane_map entry out of range");
        const int tile = i / T, t = i % T;
        for (int q = 0; q < nnz; q++) {
            const int sc = P.csrSlot[q];
            const size_t idx = (((size_t)tile * P.nSlots + sc / 4) * T + t) * 4 + sc % 4;
            const double v = values[(size_t)lane * nnz + q];
            if (vb == 8) reinterpret_cast<double *>(lu.data())[idx] = v;
            else reinterpret_cast<float *>(lu.data())[idx] = (float)v;
        }
        for (int bus = 0; bus < P.nb; bus++)
            for (int c = 0; c < 2; c++)
                b[(((size_t)tile * P.nb + P.pos[bus]) * T + t) * 2 + c] = rhs[(size_t)lane * P.n + 2 * bus + c];
    }
    DBuf dlu, dy, dxo, drhs, dcl;
    dlu.upload(lu);
    drhs.upload(b);
    dy.alloc((size_t)nTiles * P.nb * 2 * T * vb);
    dxo.alloc((size_t)nTiles * P.nb * 2 * T * vb);
    dcl.alloc((size_t)nTiles * T * 4);
    const DevPlan dp = st->dp;
    TopPlan tp;
    withT(T, [&](auto TC) {
        constexpr int TT = decltype(TC)::value;
        if (vb == 8) {
            planTop(st, reinterpret_cast<const void *>(&luSolveKernel<TT, double>), TT, 512, 8, tp);
            luSolveKernel<TT, double><<<nTiles, 512, tp.ldsBytes>>>(dp, dlu.as<double>(), dy.as<double>(),
                                                                    dxo.as<double>(), drhs.as<double>(),
                                                                    dcl.as<int>(), tp.top);
        } else {
            planTop(st, reinterpret_cast<const void *>(&luSolveKernel<TT, float>), TT, 512, 4, tp);
            luSolveKernel<TT, float><<<nTiles, 512, tp.ldsBytes>>>(dp, dlu.as<float>(), dy.as<float>(),
                                                                   dxo.as<float>(), drhs.as<double>(), dcl.as<int>(),
                                                                   tp.top);
        }
    });
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    std::vector<char> xo((size_t)nTiles * P.nb * 2 * T * vb);
    std::vector<int> cl((size_t)nTiles * T);
    HIP_OK(hipMemcpy(xo.data(), dxo.p, xo.size(), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(cl.data(), dcl.p, cl.size() * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < nLanes; i++) {
        const int lane = laneMap[i];
        const int tile = i / T, t = i % T;
        for (int bus = 0; bus < P.nb; bus++)
            for (int c = 0; c < 2; c++) {
                const size_t idx = (((size_t)tile * P.nb + P.pos[bus]) * T + t) * 2 + c;
                solutions[(size_t)lane * P.n + 2 * bus + c] =
                    vb == 8 ? reinterpret_cast<double *>(xo.data())[idx] : reinterpret_cast<float *>(xo.data())[idx];
            }
        if (clampsOut) clampsOut[i] = cl[(size_t)tile * T + t];
    }
}

}  // namespace olfhip

using namespace olfhip;

template <typename Fn>
static int guarded(Fn &&fn) {
    try {
        gLastError.clear();
        fn();
        return OLFHIP_OK;
    } catch (const ArgumentError &e) {
        gLastError = e.what();
        return OLFHIP_E_ARGUMENT;
    } catch (const UnsupportedError &e) {
        gLastError = e.what();
        return OLFHIP_E_UNSUPPORTED;
    } catch (const MemoryError &e) {
        gLastError = e.what();
        return OLFHIP_E_MEMORY;
    } catch (const std::bad_alloc &) {
        gLastError = "host allocation failed";
        return OLFHIP_E_MEMORY;
    } catch (const std::exception &e) {
        gLastError = std::string("GPU batched Newton-Raphson failed: ") + e.what();
        return OLFHIP_E_DEVICE;
    } catch (...) {
        gLastError = "unknown failure";
        return OLFHIP_E_DEVICE;
    }
}

extern "C" {

const char *olfhip_last_error(void) { return gLastError.c_str(); }

int olfhip_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int olfhip_create_structure(int n, int n_closed_branch, const double *closed_branch_parameters,
                            const int *closed_branch_indices, int n_shunt, const double *shunt_parameters,
                            const int *shunt_indices, olfhip_structure **out) {
    return guarded([&] {
        if (!out) throw ArgumentError("out is null");
        *out = nullptr;
        std::unique_ptr<StructureImpl> s(new StructureImpl());
        s->plan = buildPlan(n, n_closed_branch, closed_branch_parameters, closed_branch_indices, n_shunt,
                            shunt_parameters, shunt_indices);
        s->bus1.resize(n_closed_branch);
        s->bus2.resize(n_closed_branch);
        for (int e = 0; e < n_closed_branch; e++) {
            s->bus1[e] = closed_branch_indices[12 * e] / 2;
            s->bus2[e] = closed_branch_indices[12 * e + 2] / 2;
        }
        const int nb = n / 2;
        {
            // bus admittance sums |y| (tap-scaled on side 1) + |g| + |b| of the branch ends and the shunts
            std::vector<double> ysum(nb, 0.0);
            for (int e = 0; e < n_closed_branch; e++) {
                const double *p8 = closed_branch_parameters + 8 * e;
                const double r1 = p8[6] != 0.0 ? p8[6] : 1.0;
                ysum[s->bus1[e]] += (std::fabs(p8[0]) + std::fabs(p8[1]) + std::fabs(p8[3])) * r1 * r1;
                ysum[s->bus2[e]] += std::fabs(p8[0]) + std::fabs(p8[2]) + std::fabs(p8[4]);
            }
            for (int q = 0; q < n_shunt; q++) {
                const int b = shunt_indices[3 * q] / 2;
                if (b >= 0 && b < nb) ysum[b] += std::fabs(shunt_parameters[2 * q]) + std::fabs(shunt_parameters[2 * q + 1]);
            }
            double mx = 0.0;
            for (double v : ysum) mx = std::max(mx, v);
            s->jacobianScale = mx > 0.0 ? 2.0 * mx : 1.0;
        }
        s->incPtr.assign(nb + 1, 0);
        for (int e = 0; e < n_closed_branch; e++) {
            s->incPtr[s->bus1[e] + 1]++;
            s->incPtr[s->bus2[e] + 1]++;
        }
        for (int b = 0; b < nb; b++) s->incPtr[b + 1] += s->incPtr[b];
        s->inc.assign(s->incPtr[nb], 0);
        std::vector<int> fill(s->incPtr.begin(), s->incPtr.end() - 1);
        for (int e = 0; e < n_closed_branch; e++) {
            s->inc[fill[s->bus1[e]]++] = e;
            s->inc[fill[s->bus2[e]]++] = e;
        }
        *out = reinterpret_cast<olfhip_structure *>(s.release());
    });
}

void olfhip_destroy_structure(olfhip_structure *structure) { release(reinterpret_cast<StructureImpl *>(structure)); }

int olfhip_structure_statistics(const olfhip_structure *structure, long long *stats6) {
    return guarded([&] {
        if (!structure || !stats6) throw ArgumentError("null argument");
        const Plan &P = reinterpret_cast<const StructureImpl *>(structure)->plan;
        stats6[0] = P.nb;
        stats6[1] = P.nSlots;
        stats6[2] = P.nLevels;
        stats6[3] = P.factorMads;
        stats6[4] = (long long)P.colIdx.size();
        stats6[5] = 4LL * P.nSlots;
    });
}

int olfhip_structure_order(const olfhip_structure *structure, int *bus_order) {
    return guarded([&] {
        if (!structure || !bus_order) throw ArgumentError("null argument");
        const Plan &P = reinterpret_cast<const StructureImpl *>(structure)->plan;
        std::memcpy(bus_order, P.perm.data(), P.perm.size() * sizeof(int));
    });
}

int olfhip_structure_pattern(const olfhip_structure *structure, int *nnz, int *rowptr, int *colidx) {
    return guarded([&] {
        if (!structure || !nnz) throw ArgumentError("null argument");
        const Plan &P = reinterpret_cast<const StructureImpl *>(structure)->plan;
        *nnz = (int)P.colIdx.size();
        if (rowptr) std::memcpy(rowptr, P.rowPtr.data(), P.rowPtr.size() * sizeof(int));
        if (colidx) std::memcpy(colidx, P.colIdx.data(), P.colIdx.size() * sizeof(int));
    });
}

int olfhip_recommend_batch_size(const olfhip_structure *structure, int max_cap, int mode, int *out) {
    return guarded([&] {
        if (!structure || !out) throw ArgumentError("null argument");
        if (max_cap <= 0) throw ArgumentError("max_cap must be positive");
        const Plan &P = reinterpret_cast<const StructureImpl *>(structure)->plan;
        size_t freeB = 0, totalB = 0;
        HIP_OK(hipMemGetInfo(&freeB, &totalB));
        if (!knownMode(mode)) throw ArgumentError("unknown mode");
        size_t per = P.bytesPerScenario((int)luBytes(mode));
        if (mode == OLFHIP_LU_ROBUST)
            per += (size_t)(OLFHIP_ROBUST_KRYLOV_DIM + 1) * P.nb * 2 * 8 + kryHStride(OLFHIP_ROBUST_KRYLOV_DIM) * 8;
        const size_t fit = (size_t)(0.7 * (double)freeB) / per;
        *out = (int)std::max<size_t>(1, std::min<size_t>(fit, (size_t)max_cap));
    });
}

int olfhip_create_batch_context(olfhip_structure *structure, int n_scenarios, int mode,
                                const olfhip_reactive_limits *reactive_limits, const olfhip_tuning *tuning,
                                olfhip_context **out) {
    return guarded([&] {
        if (!out) throw ArgumentError("out is null");
        *out = nullptr;
        *out = reinterpret_cast<olfhip_context *>(
            createContext(reinterpret_cast<StructureImpl *>(structure), n_scenarios, mode, reactive_limits, tuning));
    });
}

void olfhip_destroy_batch_context(olfhip_context *context) { delete reinterpret_cast<ContextImpl *>(context); }

int olfhip_set_row_criteria(olfhip_context *context, const double *fixed_row_max_mismatch,
                            const double *equation_row_max_mismatch) {
    return guarded([&] {
        ContextImpl *c = reinterpret_cast<ContextImpl *>(context);
        if (!c) throw ArgumentError("context is null");
        if (!fixed_row_max_mismatch && !equation_row_max_mismatch) {
            c->rowCriteria = false;
            return;
        }
        if (!fixed_row_max_mismatch || !equation_row_max_mismatch)
            throw ArgumentError("row criteria need both threshold arrays (or neither)");
        const Plan &P = c->st->plan;
        std::vector<double> fixedNode(P.n), eqNode(P.n);
        for (int r = 0; r < P.n; r++) {
            const double f = fixed_row_max_mismatch[r], e = equation_row_max_mismatch[r];
            if (!(f > 0.0) || !(e > 0.0) || !std::isfinite(f) || !std::isfinite(e))
                throw ArgumentError("row threshold " + std::to_string(r) + " must be finite and > 0");
            const int i = 2 * P.pos[r >> 1] + (r & 1);
            fixedNode[i] = f;
            eqNode[i] = e;
        }
        c->rowThrFixed.upload(fixedNode);
        c->rowThrEq.upload(eqNode);
        HIP_OK(hipDeviceSynchronize());   // the uploads (null stream) complete before any solve on the context's
        c->rowCriteria = true;
    });
}

int olfhip_set_distributed_slack(olfhip_context *context, const olfhip_distributed_slack *ds) {
    return guarded([&] {
        ContextImpl *c = reinterpret_cast<ContextImpl *>(context);
        if (!c) throw ArgumentError("context is null");
        const Plan &P = c->st->plan;
        if (!ds) {
            c->dsN = 0;
            c->slackNode = -1;
            return;
        }
        if (ds->n_buses < 0) throw ArgumentError("n_buses must be >= 0");
        if (ds->slack_bus < 0 || ds->slack_bus >= P.nb) throw ArgumentError("slack_bus out of range");
        if (!std::isfinite(ds->slack_target_p)) throw ArgumentError("slack_target_p must be finite");
        const int m = ds->n_buses;
        if (m > 0 && (!ds->bus || !ds->factor || !ds->base_target_p || !ds->max_delta || !ds->min_delta))
            throw ArgumentError("distributed-slack arrays are null");
        std::vector<int> node(m);
        std::vector<char> seen(P.nb, 0);
        double sum = 0.0;
        for (int i = 0; i < m; i++) {
            const int b = ds->bus[i];
            if (b < 0 || b >= P.nb) throw ArgumentError("participant bus out of range");
            if (b == ds->slack_bus) throw ArgumentError("the slack bus cannot be a participant (its share is 1 - sum of factors)");
            if (seen[b]) throw ArgumentError("participant bus listed twice");
            seen[b] = 1;
            if (!(ds->factor[i] >= 0.0) || !std::isfinite(ds->factor[i])) throw ArgumentError("factors must be finite and >= 0");
            if (!(ds->max_delta[i] >= 0.0) || !(ds->min_delta[i] <= 0.0) || !std::isfinite(ds->base_target_p[i]))
                throw ArgumentError("participant " + std::to_string(i) + ": need max_delta >= 0 >= min_delta and a finite anchor");
            sum += ds->factor[i];
            node[i] = P.pos[b];
        }
        if (sum > 1.0 + 1e-12) throw ArgumentError("participation factors sum above 1");
        c->dsN = m;
        c->slackNode = P.pos[ds->slack_bus];
        c->slackTargetP = ds->slack_target_p;
        c->dsFactorSum = sum;
        c->dsNode.upload(node);
        c->dsFactor.upload(std::vector<double>(ds->factor, ds->factor + m));
        c->dsBase.upload(std::vector<double>(ds->base_target_p, ds->base_target_p + m));
        c->dsMaxD.upload(std::vector<double>(ds->max_delta, ds->max_delta + m));
        c->dsMinD.upload(std::vector<double>(ds->min_delta, ds->min_delta + m));
        c->dsPrev.alloc((size_t)c->nTiles * c->T * std::max(m, 1) * 8);
        DevDS d{m, c->dsNode.as<int>(), c->dsFactor.as<double>(), c->dsBase.as<double>(), c->dsMaxD.as<double>(),
                c->dsMinD.as<double>(), c->slackTargetP, c->dsFactorSum};
        c->dsDesc.alloc(sizeof(DevDS));
        HIP_OK(hipMemcpy(c->dsDesc.p, &d, sizeof(DevDS), hipMemcpyHostToDevice));
        HIP_OK(hipDeviceSynchronize());
    });
}

int olfhip_solve_batch_context(olfhip_context *context, const olfhip_solve_inputs *inputs,
                               olfhip_solve_outputs *outputs) {
    return guarded([&] {
        if (!context) throw ArgumentError("context is null");
        solve(reinterpret_cast<ContextImpl *>(context), inputs, outputs);
    });
}

int olfhip_context_timing(const olfhip_context *context, double *step_ms, int *steps) {
    return guarded([&] {
        if (!context) throw ArgumentError("context is null");
        const ContextImpl *c = reinterpret_cast<const ContextImpl *>(context);
        if (step_ms) *step_ms = c->lastStepMs;
        if (steps) *steps = c->lastSteps;
    });
}

int olfhip_set_x0_seeds(olfhip_context *context, int n_seeds, const long long *flat_index, const double *value) {
    return guarded([&] {
        if (!context) throw ArgumentError("context is null");
        ContextImpl *c = reinterpret_cast<ContextImpl *>(context);
        if (n_seeds < 0 || (n_seeds > 0 && (!flat_index || !value))) throw ArgumentError("bad seed arrays");
        const long long limit = (long long)c->capacity * c->st->plan.n;
        for (int q = 0; q < n_seeds; q++) {
            if (flat_index[q] < 0 || flat_index[q] >= limit) throw ArgumentError("x0 seed index out of range");
            if (!std::isfinite(value[q])) throw ArgumentError("x0 seed value is not finite");
        }
        c->x0SeedIndex.assign(flat_index, flat_index + n_seeds);
        c->x0SeedValue.assign(value, value + n_seeds);
    });
}

int olfhip_begin_batch_install(olfhip_context *context, int n_slots, const int *bank_slots) {
    return guarded([&] {
        if (!context) throw ArgumentError("context is null");
        ContextImpl *c = reinterpret_cast<ContextImpl *>(context);
        if (n_slots < 0) throw ArgumentError("n_slots must be >= 0");
        if (n_slots == 0 || !bank_slots) {
            c->bankInstall = false;   // the whole bank, scenario j into slot j
            c->bankSlots.clear();
            return;
        }
        std::vector<char> taken(c->capacity, 0);
        for (int j = 0; j < n_slots; j++) {
            const int b = bank_slots[j];
            if (b < 0 || b >= c->capacity || taken[b])
                throw ArgumentError("bank slot " + std::to_string(b) + " outside the bank of " +
                                    std::to_string(c->capacity) + " or installed twice");
            if (c->pending[b])
                throw ArgumentError("bank slot " + std::to_string(b) + " holds a scenario not handed back yet");
            taken[b] = 1;
        }
        c->bankSlots.assign(bank_slots, bank_slots + n_slots);
        c->bankInstall = true;
    });
}

int olfhip_read_handed_back_slots(const olfhip_context *context, int *n_slots, int *slots) {
    return guarded([&] {
        if (!context || !n_slots) throw ArgumentError("null argument");
        const ContextImpl *c = reinterpret_cast<const ContextImpl *>(context);
        *n_slots = (int)c->handedBack.size();
        if (slots) std::copy(c->handedBack.begin(), c->handedBack.end(), slots);
    });
}

int olfhip_context_graph_statistics(const olfhip_context *context, int *instantiations, int *updates) {
    return guarded([&] {
        if (!context || !instantiations || !updates) throw ArgumentError("null argument");
        const ContextImpl *c = reinterpret_cast<const ContextImpl *>(context);
        *instantiations = c->graphInstantiations;
        *updates = c->graphUpdates;
    });
}

int olfhip_read_outer_loop_state(olfhip_context *context, int n_scenarios, int *row_mode, double *target,
                                 signed char *reactive_limit_types) {
    return guarded([&] {
        ContextImpl *c = reinterpret_cast<ContextImpl *>(context);
        if (!c || !row_mode || !target) throw ArgumentError("null argument");
        if (n_scenarios <= 0 || n_scenarios > c->lastInstalled)
            throw ArgumentError("n_scenarios exceeds the last solve's installed count");
        const Plan &P = c->st->plan;
        const size_t K = (size_t)n_scenarios;
        DBuf dm, dt, dl;
        dm.alloc(K * P.n * 4);
        dt.alloc(K * P.n * 8);
        dl.alloc(K * std::max(c->nRL, 1));
        const DevPlan dp = c->st->dp;
        const DevState S = c->devState();
        const DevRL R = c->devRL();
        withT(c->T, [&](auto TC) {
            constexpr int T = decltype(TC)::value;
            gatherOuterStateKernel<T><<<gridFor(K * (P.nb + c->nRL), 256), 256, 0, c->stream>>>(
                dp, S, R, n_scenarios, dm.as<int>(), dt.as<double>(), dl.as<signed char>());
        });
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(row_mode, dm.p, K * P.n * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(target, dt.p, K * P.n * 8, hipMemcpyDeviceToHost, c->stream));
        if (reactive_limit_types && c->nRL > 0)
            HIP_OK(hipMemcpyAsync(reactive_limit_types, dl.p, K * c->nRL, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}

int olfhip_read_states(olfhip_context *context, int n_scenarios, double *x) {
    return guarded([&] {
        if (!context) throw ArgumentError("context is null");
        readStates(reinterpret_cast<ContextImpl *>(context), n_scenarios, x);
    });
}

int olfhip_islanded_buses(const olfhip_structure *structure, int slack_bus, int n_outaged, const int *outaged,
                          unsigned char *islanded, int *n_islanded) {
    return guarded([&] {
        if (!structure || !islanded) throw ArgumentError("null argument");
        const StructureImpl *s = reinterpret_cast<const StructureImpl *>(structure);
        const int nb = s->plan.nb;
        if (slack_bus < 0 || slack_bus >= nb) throw ArgumentError("slack bus out of range");
        if (n_outaged > 0 && !outaged) throw ArgumentError("outaged is null");
        std::vector<char> out(s->bus1.size(), 0);
        for (int q = 0; q < n_outaged; q++)
            if (outaged[q] >= 0 && outaged[q] < (int)out.size()) out[outaged[q]] = 1;
        std::vector<int> queue{slack_bus};
        std::vector<char> seen(nb, 0);
        seen[slack_bus] = 1;
        for (size_t h = 0; h < queue.size(); h++) {
            const int u = queue[h];
            for (int q = s->incPtr[u]; q < s->incPtr[u + 1]; q++) {
                const int e = s->inc[q];
                if (out[e]) continue;
                const int w = s->bus1[e] == u ? s->bus2[e] : s->bus1[e];
                if (!seen[w]) {
                    seen[w] = 1;
                    queue.push_back(w);
                }
            }
        }
        for (int b = 0; b < nb; b++) islanded[b] = seen[b] ? 0 : 1;
        if (n_islanded) *n_islanded = nb - (int)queue.size();
    });
}

int olfhip_lu_factorize_and_solve(olfhip_structure *structure, int mode, int capacity, const double *values,
                                  const double *rhs, int n_lanes, const int *lane_map, double *solutions,
                                  int *pivot_clamps) {
    return guarded([&] {
        luFactorizeAndSolve(reinterpret_cast<StructureImpl *>(structure), mode, capacity, values, rhs, n_lanes,
                            lane_map, solutions, pivot_clamps);
    });
}

int olfhip_autotune(olfhip_structure *structure, int mode, const olfhip_reactive_limits *reactive_limits,
                    const olfhip_solve_inputs *inputs, olfhip_tuning *best, double *best_ms) {
    return guarded([&] {
        if (!structure || !inputs || !best) throw ArgumentError("null argument");
        StructureImpl *st = reinterpret_cast<StructureImpl *>(structure);
        const int K = inputs->n_scenarios;
        double bestMs = 1e300;
        olfhip_tuning bestT{0, 0};
        for (int T : {1, 2, 4, 8, 16, 32}) {
            if (T > 1 && (T > K || mode == OLFHIP_LU_ROBUST)) break;   // the robust mode has one lane per tile
            for (int NT : {256, 512, 1024}) {
                if (NT > MAX_THREADS) continue;
                olfhip_tuning tu{T, NT};
                std::unique_ptr<ContextImpl> c(createContext(st, K, mode, reactive_limits, &tu));
                olfhip_solve_inputs in = *inputs;
                in.states_to_host = 0;
                in.pause_at_ended = 0;   // timed to the end
                solve(c.get(), &in, nullptr);   // warm-up (code objects, caches)
                auto t0 = std::chrono::steady_clock::now();
                solve(c.get(), &in, nullptr);
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (traceOn()) std::fprintf(stderr, "OLFHIP autotune T=%d NT=%d %.3f ms\n", T, NT, ms);
                if (ms < bestMs) {
                    bestMs = ms;
                    bestT = tu;
                }
            }
        }
        *best = bestT;
        if (best_ms) *best_ms = bestMs;
    });
}

int olfhip_structure_schedule(const olfhip_structure *structure, int which, int *count, int *values) {
    return guarded([&] {
        if (!structure || !count) throw ArgumentError("null argument");
        const Plan &P = reinterpret_cast<const StructureImpl *>(structure)->plan;
        const std::vector<int> *arrays[] = {&P.itemRec,     &P.bucketRec,    &P.levelBucketPtr, &P.bwdRec,
                                            &P.bwdBucketRec, &P.levelBwdBucketPtr, &P.levelItemPtr,
                                            &P.levelNodePtr, &P.diagItemOf,   &P.contribRec,   &P.levelNodes,
                                            &P.sIdx};
        if (which < 0 || which >= (int)(sizeof(arrays) / sizeof(arrays[0])))
            throw ArgumentError("which selects one of the 12 schedule arrays");
        const std::vector<int> &v = *arrays[which];
        *count = (int)v.size();
        if (values && !v.empty()) std::memcpy(values, v.data(), v.size() * 4);
    });
}

int olfhip_create_limit_table(olfhip_structure *structure, const olfhip_limit_table_desc *desc,
                              olfhip_limit_table **out) {
    return guarded([&] {
        if (!out) throw ArgumentError("out is null");
        *out = nullptr;
        *out = reinterpret_cast<olfhip_limit_table *>(
            createLimitTable(reinterpret_cast<StructureImpl *>(structure), desc));
    });
}

void olfhip_destroy_limit_table(olfhip_limit_table *table) { delete reinterpret_cast<LimitTableImpl *>(table); }

int olfhip_ambiguous_scenarios(olfhip_context *context, olfhip_limit_table *table, double band, int n_scenarios,
                               int *flags) {
    return guarded([&] {
        if (!context || !table || !flags) throw ArgumentError("null argument");
        ambiguousScenarios(reinterpret_cast<ContextImpl *>(context), reinterpret_cast<LimitTableImpl *>(table), band,
                           n_scenarios, flags);
    });
}

int olfhip_screen_violations_no_alloc(olfhip_context *context, olfhip_limit_table *table, double band,
                                      int n_scenarios, int capacity, int *counts, int *records) {
    return guarded([&] {
        if (!context || !table || !counts || !records) throw ArgumentError("null argument");
        screenViolationsNoAlloc(reinterpret_cast<ContextImpl *>(context), reinterpret_cast<LimitTableImpl *>(table),
                                band, n_scenarios, capacity, counts, records);
    });
}

int olfhip_screen_violations(olfhip_context *context, olfhip_limit_table *table, double band, int capacity,
                             int *counts, int *records) {
    return guarded([&] {
        if (!context || !table || !counts || !records) throw ArgumentError("null argument");
        ContextImpl *c = reinterpret_cast<ContextImpl *>(context);
        std::vector<int> cv, rv;
        screenViolations(c, reinterpret_cast<LimitTableImpl *>(table), band, capacity, cv, rv);
        std::memcpy(counts, cv.data(), cv.size() * 4);
        std::memcpy(records, rv.data(), rv.size() * 4);
    });
}

int olfhip_detect_violations(olfhip_context *context, olfhip_limit_table *table, double band, int capacity,
                             int max_violations, olfhip_violation *violations, int *n_violations) {
    return guarded([&] {
        if (!context || !table || !n_violations) throw ArgumentError("null argument");
        if (max_violations < 0 || (max_violations > 0 && !violations)) throw ArgumentError("bad output buffer");
        std::vector<olfhip_violation> v;
        detectViolations(reinterpret_cast<ContextImpl *>(context), reinterpret_cast<LimitTableImpl *>(table), band,
                         capacity, v);
        *n_violations = (int)v.size();
        if ((int)v.size() > max_violations)
            throw ArgumentError("max_violations is smaller than the " + std::to_string(v.size()) +
                                " violations found (their count is in n_violations)");
        if (!v.empty()) std::memcpy(violations, v.data(), v.size() * sizeof(olfhip_violation));
    });
}

}  // extern "C"
