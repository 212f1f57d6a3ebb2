// Device code of the batched Newton-Raphson: one workgroup owns one tile of T scenarios ("lanes") and runs a whole
// Newton step for them — fill, convergence test, reactive-limit check, block LU with the forward solve fused in,
// backward solve, optional FP64-residual refinement, linear-solve guard, step scaling and update — with workgroup
// barriers between elimination levels and no global synchronisation. Tiles are independent, so a launch of
// nTiles workgroups is one batched Newton iteration of the whole chunk.
//
// Layouts (lane fastest, so that the T lanes of one item read one contiguous run):
//   per-node vectors  v[((tile * nb + k) * T + t) * 2 + c]        c = 0: phi / P row, 1: V / Q row
//   LU block slots    lu[((tile * nSlots + s) * T + t) * 4 + c]   c = (P,phi) (P,V) (Q,phi) (Q,V)
//   per-entry state   e[(tile * nEntries + i) * T + t]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "generated/shunt_terms.hpp"   // OLF's shunt terms, rendered by codegen/java2hip.py
#include "plan.hpp"   // ITEM_FILL_PAIR
#include "trig.hpp"

namespace olfhip {

constexpr int LANE_RUNNING = 100;   // still solving
constexpr int LANE_PADDING = 99;    // a slot of the last tile beyond the installed scenarios
constexpr int MAX_OUTAGES = 4;
constexpr int MAX_THREADS = 1024;

// status codes (reference GpuAcSolverStatus, docs/gpu-vs-olf.md:196)
constexpr int ST_CONVERGED = 0, ST_MAX_ITER = 1, ST_DIVERGED = -1, ST_REJECTED = -2, ST_LIN_FAILED = -3;
constexpr signed char MODE_FIXED = 0, MODE_EQUATION = 1;
constexpr signed char QLIM_NONE = 0, QLIM_MIN = 1, QLIM_MAX = 2, QLIM_FROZEN = 5;
constexpr double LIN_SOLVE_FAIL_RATIO = 1e-2;   // reference native/cuda/solver/iterative_refinement.cuh:23

struct DevPlan {
    int nb, nSlots, nLevels;
    const int *levelNodePtr, *levelItemPtr;
    const int4 *itemRec, *contribRec, *bwdRec;
    const int4 *bucketRec, *bwdBucketRec;       // buckets of one thread split (first, end, split, diagonal flag)
    const int *levelBucketPtr, *levelBwdBucketPtr;
    const int *levelNodes, *diagItemOf;
    int nBuckets, nBwdBuckets;
    const int *nodeBase, *sPtr, *sIdx;
    const int *hbPtr, *hbBranch;
    const double4 *hbCoef;
    const int *grpPtr, *grpEnd, *grpFar, *grpSlot;
    const double *shuntG, *shuntB;
    const int *perm, *pos;
    const int *levelPos;        // position of each node in level order
    const int *nodeLevel;       // elimination-tree level of each node
    const int2 *itemKI;         // (k, i) of each item
};

// The top block: the nodes of levels l0.. (an upper set of the elimination tree, `start` = their first position in
// level order, n of them) factored as one dense block matrix per lane in LDS, so that the levels where the tree is
// one or two nodes wide cost one LDS barrier per node instead of one global-memory phase per level. slots lists the
// top nodes' LU slots with their dense coordinates (slot, row, column, 0); S is the thread split of the contribution
// sums from below the block. n = 0: no top block (l0 = nLevels).
struct DevTop {
    int l0, n, start;
    const int4 *slots;
    int nSlots;
    // the top items (all, or the diagonal ones) ordered by the split of their contributions from below, and buckets
    const int *items, *diagItems;
    const int4 *buckets, *diagBuckets;
    int nBuckets, nDiagBuckets;
    int metaBytes;   // LDS before the top block: the level schedule's bucket records
    // The sparse levels' records in unit-major order for this context's lane count: bucket b's contribution records
    // (and the backward pass's (U slot, column node) pairs) for round r, item q and thread s of its split S sit at
    // base_b + (r * nItems_b + q) * S + s, padded with x = -1, so that a thread finds its records from its work unit
    // alone and loads them together with the item record instead of after it. bucketPad / bwdBucketPad hold
    // (base, rounds) per bucket.
    const int4 *contribPad;
    const int2 *bwdPad, *bucketPad, *bwdBucketPad;
    // Static pivoting of the FP64 factors (0: off). A pivot block whose smallest singular value is below about
    // pivotDelta gets pivotDelta added to its diagonal entries (sign-matched), so that the factors are those of
    // J + E with E small and of rank 2 per such block; the robust solve's GMRES then corrects for E.
    double pivotDelta;
};

struct DevState {
    int nTiles, nInstalled;
    double *x, *F, *dx, *r, *target, *qinj;
    signed char *mode;
    void *lu, *y, *dx1;
    int *it, *status, *solves, *outer, *total, *clamps;
    double *norm, *ratio;
    const int *outages;
    int maxOut;
    signed char *rlType, *rlCount, *rlPass;
    int *tileRunning;
    int *ended;   // lanes that ended during the current solve call (the pause test)
    // outer-loop sequence per lane: the loop it is at (0 distributed slack, 1 reactive limits) and whether the current
    // sweep changed something; the slack bus's computed P at the lane's last fill; the slack's applied share (NaN
    // before any distribution), whether the last distribution failed, and each participant's target before the
    // last reset ([tile][participant][lane])
    signed char *loopAt, *sweepChanged, *dsFailed;
    double *slackP, *dsRetain, *dsPrev;
    // LINE_SEARCH: the state before the last step ([tile][node][lane][2]), and per lane the last tested ||F||_2^2 of
    // the current solve and whether there is one
    double *xPrev, *lsLast;
    signed char *lsHas;
    const int *nActive;   // active equations per lane (the 2-norm criterion's count)
    // per-row stopping thresholds in node order ([node][2]: P row, Q row), for identity and for equation rows;
    // null without the per-row criterion
    const double *rowThrFixed, *rowThrEq;
    // robust solve: the Krylov basis (kryDim + 1 vectors, each laid out like x over the whole context) and per tile
    // the Hessenberg matrix, right-hand side, rotations and coefficients of its GMRES cycle
    double *kryV, *kryH;
    // opt-in section timing of one tile: prof[0] counts the Newton steps that recorded, then PROF_SECTIONS stamps per
    // step (wall clock) from prof[PROF_SECTIONS] on, at most profCap steps
    unsigned long long *prof;
    int profCap;
    int profTile;   // the tile whose sections are stamped
    // opt-in tile timeline: per tile its first start and last end (wall clock) and its Newton steps in the solve
    unsigned long long *tileClock;
};
constexpr int PROF_SECTIONS = 16;

// Distributed slack: participating nodes, their factors, anchor targets and room up (>= 0) and down (<= 0) from the
// anchor; the slack's own P target and the sum of the factors (1 - it is the slack's retained share). Kept in device
// memory and read through one pointer, so that the kernel does not hold its fields in registers.
struct DevDS {
    int n;
    const int *node;
    const double *factor, *base, *maxD, *minD;
    double slackTargetP, factorSum;
};

struct DevRL {
    int n;
    const int *node;
    const double *minQ, *maxQ, *loadQ, *targetV;
    const DevDS *ds;   // null without distributed slack
    int slackNode;     // -1 without distributed slack
};

struct NrParams {
    int maxIter;
    double tol;
    int reactiveLimits;
    double rlTol;
    int maxOuter;
    int maxSwitch;
    int scaling;
    double maxDv, maxDphi;
    int refine;
    int batchOfOne;
    int stepsPerLaunch;   // Newton steps one launch may run per tile
    double convEps;       // > 0: converged when ||F||_2 < sqrt(convEps^2 nActive) (OLF's criterion), else ||F||inf < tol
    int lsMax;            // LINE_SEARCH (scaling 1): backtracks per test and the step fold, mu = lsFold^-b
    double lsFold;
    int distributedSlack; // run the distributed-slack outer loop (before the reactive limits)
    double dsTol;         // its gate, slackBusPMaxMismatch in pu
    int dsReject;         // a failed distribution ends the scenario REJECTED
    int rowCriteria;      // converged when every row's |F| is below its threshold (S.rowThrFixed / S.rowThrEq)
    int kryDim;           // robust solve: GMRES basis size, restarts after the first cycle, relative 2-norm target
    int kryRestarts;
    double kryTol;
    double linFailRatio;  // the linear-solve guard's bound on ||J dx + F||inf / ||F||inf
    int fixedIter;        // exactly maxIter steps: no convergence test, no guard, no outer loop
    int pauseAt;          // > 0: a tile stops before a step once this many lanes have ended in the call
};

template <typename VT>
struct Blk {
    VT a, b, c, d;  // [[a b] [c d]]
};
template <typename VT>
struct V2 {
    VT p, q;
};

template <typename VT>
__device__ __forceinline__ Blk<VT> ldBlk(const VT *p);
template <>
__device__ __forceinline__ Blk<float> ldBlk<float>(const float *p) {
    float4 v = *reinterpret_cast<const float4 *>(p);
    return {v.x, v.y, v.z, v.w};
}
template <>
__device__ __forceinline__ Blk<double> ldBlk<double>(const double *p) {
    double2 u = reinterpret_cast<const double2 *>(p)[0];
    double2 v = reinterpret_cast<const double2 *>(p)[1];
    return {u.x, u.y, v.x, v.y};
}
template <typename VT>
__device__ __forceinline__ void stBlk(VT *p, const Blk<VT> &b);
template <>
__device__ __forceinline__ void stBlk<float>(float *p, const Blk<float> &b) {
    *reinterpret_cast<float4 *>(p) = make_float4(b.a, b.b, b.c, b.d);
}
template <>
__device__ __forceinline__ void stBlk<double>(double *p, const Blk<double> &b) {
    reinterpret_cast<double2 *>(p)[0] = make_double2(b.a, b.b);
    reinterpret_cast<double2 *>(p)[1] = make_double2(b.c, b.d);
}
template <typename VT>
__device__ __forceinline__ V2<VT> ldV(const VT *p);
template <>
__device__ __forceinline__ V2<float> ldV<float>(const float *p) {
    float2 v = *reinterpret_cast<const float2 *>(p);
    return {v.x, v.y};
}
template <>
__device__ __forceinline__ V2<double> ldV<double>(const double *p) {
    double2 v = *reinterpret_cast<const double2 *>(p);
    return {v.x, v.y};
}
template <typename VT>
__device__ __forceinline__ void stV(VT *p, V2<VT> v);
template <>
__device__ __forceinline__ void stV<float>(float *p, V2<float> v) {
    *reinterpret_cast<float2 *>(p) = make_float2(v.p, v.q);
}
template <>
__device__ __forceinline__ void stV<double>(double *p, V2<double> v) {
    *reinterpret_cast<double2 *>(p) = make_double2(v.p, v.q);
}

// acc -= X * Y (2x2 blocks)
template <typename VT>
__device__ __forceinline__ void mulSub(Blk<VT> &acc, const Blk<VT> &X, const Blk<VT> &Y) {
    acc.a -= X.a * Y.a + X.b * Y.c;
    acc.b -= X.a * Y.b + X.b * Y.d;
    acc.c -= X.c * Y.a + X.d * Y.c;
    acc.d -= X.c * Y.b + X.d * Y.d;
}
template <typename VT>
__device__ __forceinline__ Blk<VT> mul(const Blk<VT> &X, const Blk<VT> &Y) {
    return {X.a * Y.a + X.b * Y.c, X.a * Y.b + X.b * Y.d, X.c * Y.a + X.d * Y.c, X.c * Y.b + X.d * Y.d};
}
// v -= X * w
template <typename VT>
__device__ __forceinline__ void mulSubV(V2<VT> &v, const Blk<VT> &X, const V2<VT> &w) {
    v.p -= X.a * w.p + X.b * w.q;
    v.q -= X.c * w.p + X.d * w.q;
}

// Inverse of a diagonal block with a static-pivot clamp (the reference clamps near-zero scalar pivots to +-tiny
// rather than failing, native/cuda/lu/lu_pivot.cuh:12-21; here on the block determinant). With delta > 0 (FP64
// factors of the robust solve) a nearly singular block is perturbed first, Li & Demmel's static pivoting (GESP,
// "Making sparse Gaussian elimination scalable by static pivoting", SC'98) on the 2x2 block: |det| <= delta ||D||_1
// bounds the smallest singular value by about delta, and then delta goes onto each diagonal entry with its sign.
template <typename VT>
__device__ __forceinline__ Blk<VT> inv2(Blk<VT> D, bool &clamped, VT delta = (VT)0) {
    VT det = D.a * D.d - D.b * D.c;
    if (delta > (VT)0 && !(fabs(det) > delta * (fabs(D.a) + fabs(D.b) + fabs(D.c) + fabs(D.d)))) {
        D.a += D.a < (VT)0 ? -delta : delta;
        D.d += D.d < (VT)0 ? -delta : delta;
        det = D.a * D.d - D.b * D.c;
    }
    const VT tiny = (VT)1e-26;
    clamped = !(fabs(det) > tiny);
    if (clamped) det = det < (VT)0 ? -tiny : tiny;
    VT id = (VT)1 / det;
    return {D.d * id, -D.b * id, -D.c * id, D.a * id};
}
// The static-pivot perturbation of a context's factors: FP64 only (the FP32 factors keep the plain clamp).
template <typename VT>
__device__ __forceinline__ VT pivotDeltaOf(const DevTop &top) {
    if constexpr (sizeof(VT) == 8) return top.pivotDelta;
    else return (VT)0;
}

template <int T>
__device__ __forceinline__ size_t vIdx(int tile, int nb, int k, int t) {
    return ((size_t)tile * nb + k) * (2 * T) + 2 * t;
}
template <int T>
__device__ __forceinline__ size_t qIdx(int tile, int nb, int k, int t) {
    return ((size_t)tile * nb + k) * T + t;
}
template <int T>
__device__ __forceinline__ size_t bIdx(int tile, int nSlots, int s, int t) {
    return ((size_t)tile * nSlots + s) * (4 * T) + 4 * t;
}

__device__ __forceinline__ bool isOut(int br, const int *o) {
    return br == o[0] || br == o[1] || br == o[2] || br == o[3];
}

// The half-branch terms of node k, lane t: the mismatch sums, the diagonal block and per neighbour group the
// off-diagonal block (WRITE_J), or J * v for node k (APPLY_VEC). Row modes mask the rows of FIXED_VARIABLE (identity)
// rows as the reference's applyRowModeBatch does (native/cuda/equations/fill_launch.cu:82-115). The closed-branch
// formulas are the reference's (native/cuda/equations/generated/closed_branch.cuh:31-229, OLF
// ClosedBranchSide1/2*FlowEquationTerm) in own-bus / far-bus form:
//   p = A vo^2 - B vo vf sin(th),  q = C vo^2 - B vo vf cos(th),  th = psi - phi_o + phi_f.
template <int T, typename VT, bool WRITE_J, bool APPLY_VEC>
__device__ __forceinline__ void nodeTerms(const DevPlan &P, int tile, int k, int t, const double *__restrict__ x,
                                          const int *out, VT *__restrict__ lu, const double *__restrict__ v,
                                          signed char modeP, signed char modeQ, double &sumP, double &sumQ,
                                          double &jvP, double &jvQ) {
    const int nb = P.nb;
    const double2 xo = *reinterpret_cast<const double2 *>(x + vIdx<T>(tile, nb, k, t));
    const double pho = xo.x, vo = xo.y;
    double dPdth = 0, dPdv = 0, dQdth = 0, dQdv = 0;
    double sP = 0, sQ = 0, jP = 0, jQ = 0;
    int h = P.hbPtr[k];
    const int g0 = P.grpPtr[k], g1 = P.grpPtr[k + 1];
    for (int g = g0; g < g1; g++) {
        const int f = P.grpFar[g];
        const int hEnd = P.grpEnd[g];
        const double2 xf = *reinterpret_cast<const double2 *>(x + vIdx<T>(tile, nb, f, t));
        const double phf = xf.x, vf = xf.y;
        double oPth = 0, oPv = 0, oQth = 0, oQv = 0;
        for (; h < hEnd; h++) {
            if (isOut(P.hbBranch[h], out)) continue;
            const double4 cf = P.hbCoef[h];
            const double th = cf.w - pho + phf;
            double s, c;
            sincosReduced(th, &s, &c);
            const double Bvv = cf.y * vo * vf;
            sP += cf.x * vo * vo - Bvv * s;
            sQ += cf.z * vo * vo - Bvv * c;
            dPdv += 2.0 * cf.x * vo - cf.y * vf * s;
            dPdth += Bvv * c;
            dQdv += 2.0 * cf.z * vo - cf.y * vf * c;
            dQdth -= Bvv * s;
            oPv -= cf.y * vo * s;
            oPth -= Bvv * c;
            oQv -= cf.y * vo * c;
            oQth += Bvv * s;
        }
        if (modeP == MODE_FIXED) {
            oPth = 0;
            oPv = 0;
        }
        if (modeQ == MODE_FIXED) {
            oQth = 0;
            oQv = 0;
        }
        if (WRITE_J) {
            Blk<VT> b = {(VT)oPth, (VT)oPv, (VT)oQth, (VT)oQv};
            stBlk<VT>(lu + bIdx<T>(tile, P.nSlots, P.grpSlot[g], t), b);
        }
        if (APPLY_VEC) {
            const double2 vf2 = *reinterpret_cast<const double2 *>(v + vIdx<T>(tile, nb, f, t));
            jP += oPth * vf2.x + oPv * vf2.y;
            jQ += oQth * vf2.x + oQv * vf2.y;
        }
    }
    // bus shunt: OLF's ShuntCompensator terms (generated), p = g v^2, q = -b v^2
    const double gs = P.shuntG[k], bs = P.shuntB[k];
    sP += gen::shP(vo, gs);
    sQ += gen::shQ(vo, bs);
    dPdv += gen::shDpdv(vo, gs);
    dQdv += gen::shDqdv(vo, bs);
    if (modeP == MODE_FIXED) {
        dPdth = 1;
        dPdv = 0;
    }
    if (modeQ == MODE_FIXED) {
        dQdth = 0;
        dQdv = 1;
    }
    if (WRITE_J) {
        Blk<VT> d = {(VT)dPdth, (VT)dPdv, (VT)dQdth, (VT)dQdv};
        stBlk<VT>(lu + bIdx<T>(tile, P.nSlots, P.nodeBase[k], t), d);   // fill-in slots are left alone
    }
    if (APPLY_VEC) {
        const double2 vk = *reinterpret_cast<const double2 *>(v + vIdx<T>(tile, nb, k, t));
        jvP = jP + dPdth * vk.x + dPdv * vk.y;
        jvQ = jQ + dQdth * vk.x + dQdv * vk.y;
    }
    sumP = sP;
    sumQ = sQ;
}

// The lanes a pass works for: every lane of the tile (tm = T lanes per item), or, when exactly one lane is on, that
// lane alone with every thread on it (tm = 1, lane t0), so that a tile whose other lanes have finished runs its
// passes in the work units, and rounds, of one lane. An item's S threads and their sums are the same either way.
struct LaneMap {
    int tm, t0;
};
template <int T>
__device__ __forceinline__ LaneMap laneMapOf(const int *laneOn) {
    if (T == 1 || !laneOn) return LaneMap{T, 0};
    int n = 0, t0 = 0;
    for (int t = 0; t < T; t++)
        if (laneOn[t]) {
            n++;
            t0 = t;
        }
    // the flags are the workgroup's own: every thread reads the same values
    return n == 1 ? LaneMap{1, __builtin_amdgcn_readfirstlane(t0)} : LaneMap{T, 0};
}
// Tree reduction over the item dimension, per lane: threads tid and tid + stride share a lane because strides are
// multiples of T. Lane t's result ends in red[t]. Every thread of the block must call it.
template <int T, bool IS_MAX>
__device__ __forceinline__ void laneReduce(double *red, int tid, int nt) {
    for (int stride = nt >> 1; stride >= T; stride >>= 1) {
        __syncthreads();
        if (tid < stride) red[tid] = IS_MAX ? fmax(red[tid], red[tid + stride]) : red[tid] + red[tid + stride];
    }
    __syncthreads();
}
template <int T>
__device__ __forceinline__ void laneReduceAnd(int *red, int tid, int nt) {
    for (int stride = nt >> 1; stride >= T; stride >>= 1) {
        __syncthreads();
        if (tid < stride) red[tid] = red[tid] & red[tid + stride];
    }
    __syncthreads();
}
// laneReduce / laneReduceAnd after a pass run with lane map lm: for a lane alone (lm.tm = 1) every thread's partial
// was that lane's, so the per-residue results combine into lane t0 and the other lanes get the neutral value their
// threads would have left (0; 1 for the AND).
template <int T, bool IS_MAX>
__device__ __forceinline__ void laneReduceFor(double *red, int tid, int nt, LaneMap lm) {
    laneReduce<T, IS_MAX>(red, tid, nt);
    if (T > 1 && lm.tm == 1) {
        if (tid == 0) {
            double v = red[0];
            for (int t = 1; t < T; t++) v = IS_MAX ? fmax(v, red[t]) : v + red[t];
            for (int t = 0; t < T; t++) red[t] = t == lm.t0 ? v : 0.0;
        }
        __syncthreads();
    }
}
template <int T>
__device__ __forceinline__ void laneReduceAndFor(int *red, int tid, int nt, LaneMap lm) {
    laneReduceAnd<T>(red, tid, nt);
    if (T > 1 && lm.tm == 1) {
        if (tid == 0) {
            int v = red[0];
            for (int t = 1; t < T; t++) v &= red[t];
            for (int t = 0; t < T; t++) red[t] = t == lm.t0 ? v : 1;
        }
        __syncthreads();
    }
}

// Cross-lane sum over the S threads that share an item (lanes t, t + tm, ..., t + (S-1) tm of one wave; tm = T, or 1
// for a pass on one lane alone). The pairs summed are the same for every tm: thread s meets s ^ 1, s ^ 2, ...
template <int T, typename VT>
__device__ __forceinline__ VT groupSum(VT v, int S, int tm = T) {
    for (int off = tm; off < tm * S; off <<= 1) v += __shfl_xor(v, off);
    return v;
}
template <int T, typename VT>
__device__ __forceinline__ void groupSumBlk(Blk<VT> &b, int S, int tm = T) {
    b.a = groupSum<T>(b.a, S, tm);
    b.b = groupSum<T>(b.b, S, tm);
    b.c = groupSum<T>(b.c, S, tm);
    b.d = groupSum<T>(b.d, S, tm);
}

// Work units of a run of buckets laid end to end: unit u is lane t and thread s of the S threads sharing record rec.
// The split is capped at 64 / T so that a record's threads stay in one wave; buckets come largest split first, so
// every record's T * S units start at a multiple of T * S and its threads can reduce with lane shuffles.
struct WorkUnit {
    int rec, t, s, S;
    int b, q;   // bucket and the item's position in it
};
template <int T>
__device__ __forceinline__ int capSplit(int s) {
    return s < 64 / T ? s : 64 / T;
}
template <int T>
__device__ __forceinline__ int bucketUnits(const int4 *bk, int b0, int b1, bool diagOnly = false, int tm = T) {
    int n = 0;
    for (int b = b0; b < b1; b++) {
        const int4 r = bk[b];
        if (diagOnly && r.w != 1) continue;
        n += (r.y - r.x) * tm * capSplit<T>(r.z);
    }
    return n;
}
// unit u of the run (with diagOnly, of its diagonal-item buckets alone), for the lanes of lm
template <int T>
__device__ __forceinline__ WorkUnit unitOf(const int4 *bk, int b0, int b1, int u, bool diagOnly = false,
                                           LaneMap lm = LaneMap{T, 0}) {
    for (int b = b0; b < b1; b++) {
        const int4 r = bk[b];
        if (diagOnly && r.w != 1) continue;
        const int S = capSplit<T>(r.z);
        const int n = (r.y - r.x) * lm.tm * S;
        if (u < n) {
            const int sh = __builtin_ctz(lm.tm * S), st = __builtin_ctz(lm.tm);
            return WorkUnit{r.x + (u >> sh), lm.tm == 1 ? lm.t0 : u % T, (u >> st) & (S - 1), S, b, u >> sh};
        }
        u -= n;
    }
    return WorkUnit{-1, 0, 0, 1, 0, 0};
}

// The level schedule's bucket records, copied to LDS once per launch so that each level's bucket lookups cost LDS
// reads instead of a chain of dependent global loads.
struct LevelMeta {
    const int4 *bucketRec, *bwdBucketRec;
    const int *levelBucketPtr, *levelBwdBucketPtr;
    const int2 *bucketPad, *bwdBucketPad;   // (base, rounds) of each bucket's unit-major records
};
__host__ __device__ constexpr size_t levelMetaBytes(int nLevels, int nBuckets, int nBwdBuckets) {
    return ((size_t)24 * (nBuckets + nBwdBuckets) + (size_t)8 * (nLevels + 1) + 15) / 16 * 16;
}
// Copies the records into the LDS at base (levelMetaBytes of it). Ends with a barrier.
__device__ __forceinline__ LevelMeta loadLevelMeta(const DevPlan &P, const DevTop &top, unsigned char *base, int tid,
                                                   int nt) {
    int4 *b = reinterpret_cast<int4 *>(base);
    int4 *bb = b + P.nBuckets;
    int2 *bp = reinterpret_cast<int2 *>(bb + P.nBwdBuckets);
    int2 *bbp = bp + P.nBuckets;
    int *lp = reinterpret_cast<int *>(bbp + P.nBwdBuckets);
    int *lbp = lp + P.nLevels + 1;
    for (int i = tid; i < P.nBuckets; i += nt) {
        b[i] = P.bucketRec[i];
        bp[i] = top.bucketPad[i];
    }
    for (int i = tid; i < P.nBwdBuckets; i += nt) {
        bb[i] = P.bwdBucketRec[i];
        bbp[i] = top.bwdBucketPad[i];
    }
    for (int i = tid; i <= P.nLevels; i += nt) {
        lp[i] = P.levelBucketPtr[i];
        lbp[i] = P.levelBwdBucketPtr[i];
    }
    __syncthreads();
    return LevelMeta{b, bb, lp, lbp, bp, bbp};
}

// The factorization, left-looking by elimination levels, one phase per level, with the forward solve of
// b = sign * rhs fused into the diagonal items. Stores D(k)^-1 in the diagonal slot, the unscaled L~(i,k) and U(k,i)
// in the pair slots, and z(k) = D(k)^-1 y(k) (y = L^-1 b) in z. An item's contribution list is summed by S threads
// (stride S) and reduced across lanes. Counts clamped pivots per lane into clampLane (shared, may be null). Ends with
// a barrier.
template <int T, typename VT>
__device__ void factorForward(const DevPlan &P, const LevelMeta &M, const DevTop &top, int tile, int tid, int nt,
                              VT *__restrict__ lu, VT *__restrict__ z, const double *__restrict__ rhs, double sign,
                              int *clampLane, int l0, const int *laneOn) {
    const int nb = P.nb;
    const LaneMap lm = laneMapOf<T>(laneOn);
    for (int l = 0; l < l0; l++) {
        const int b0 = M.levelBucketPtr[l], b1 = M.levelBucketPtr[l + 1];
        const int units = bucketUnits<T>(M.bucketRec, b0, b1, false, lm.tm);
        for (int w = tid; w < units; w += nt) {
            const WorkUnit un = unitOf<T>(M.bucketRec, b0, b1, w, false, lm);
            const int t = un.t, s = un.s, S = un.S;
            if (laneOn && !laneOn[t]) continue;
            // the item record and the unit's contribution records are independent loads
            const int4 bk = M.bucketRec[un.b];
            const int2 pad = M.bucketPad[un.b];
            const int4 *cp = top.contribPad + pad.x + (size_t)un.q * S + s;
            const size_t roundStride = (size_t)(bk.y - bk.x) * S;
            int4 next = pad.y > 0 ? cp[0] : make_int4(-1, -1, -1, -1);   // issued before the item record
            const int4 it = P.itemRec[un.rec];
            VT *own = lu + bIdx<T>(tile, P.nSlots, it.x, t);
            if (bk.w) {
                // diagonal item of node it.w: pivot block, its inverse, and the forward-solve entry; the own block
                // and right-hand side are loaded first so that they overlap the contribution loads
                Blk<VT> A = {(VT)0, (VT)0, (VT)0, (VT)0};
                double2 b = make_double2(0.0, 0.0);
                if (s == 0) {
                    A = ldBlk<VT>(own);
                    b = *reinterpret_cast<const double2 *>(rhs + vIdx<T>(tile, nb, it.w, t));
                }
                Blk<VT> D = {(VT)0, (VT)0, (VT)0, (VT)0};
                V2<VT> yk = {(VT)0, (VT)0};
#pragma unroll 2
                for (int r = 0; r < pad.y; r++) {
                    const int4 cr = next;
                    if (cr.x < 0) break;
                    if (r + 1 < pad.y) next = cp[(r + 1) * roundStride];
                    const Blk<VT> Lkj = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, cr.x, t));
                    const Blk<VT> Ujk = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, cr.x + 1, t));
                    const Blk<VT> Dj = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, cr.z, t));
                    const V2<VT> zj = ldV<VT>(z + vIdx<T>(tile, nb, cr.y, t));
                    mulSub(D, Lkj, mul(Dj, Ujk));
                    mulSubV(yk, Lkj, zj);
                }
                if (S > 1) {
                    groupSumBlk<T>(D, S, lm.tm);
                    yk.p = groupSum<T>(yk.p, S, lm.tm);
                    yk.q = groupSum<T>(yk.q, S, lm.tm);
                }
                if (s == 0) {
                    const Blk<VT> Dk = {A.a + D.a, A.b + D.b, A.c + D.c, A.d + D.d};
                    const V2<VT> y = {(VT)(sign * b.x) + yk.p, (VT)(sign * b.y) + yk.q};
                    bool cl;
                    const Blk<VT> Dinv = inv2(Dk, cl, pivotDeltaOf<VT>(top));
                    if (cl && clampLane) atomicAdd(&clampLane[t], 1);
                    stBlk<VT>(own, Dinv);
                    const V2<VT> zk = {Dinv.a * y.p + Dinv.b * y.q, Dinv.c * y.p + Dinv.d * y.q};
                    stV<VT>(z + vIdx<T>(tile, nb, it.w, t), zk);
                }
            } else {
                // pair item: L~(i,k) at it.x, U(k,i) right after (own blocks loaded first; a fill-in pair starts
                // from zero, the fill writes nothing there)
                Blk<VT> A = {(VT)0, (VT)0, (VT)0, (VT)0}, B = {(VT)0, (VT)0, (VT)0, (VT)0};
                if (s == 0 && it.w != ITEM_FILL_PAIR) {
                    A = ldBlk<VT>(own);
                    B = ldBlk<VT>(own + 4 * T);
                }
                Blk<VT> L = {(VT)0, (VT)0, (VT)0, (VT)0};
                Blk<VT> U = {(VT)0, (VT)0, (VT)0, (VT)0};
#pragma unroll 2
                for (int r = 0; r < pad.y; r++) {
                    const int4 c0 = next;
                    if (c0.x < 0) break;
                    if (r + 1 < pad.y) next = cp[(r + 1) * roundStride];
                    const Blk<VT> Lkj = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, c0.x, t));
                    const Blk<VT> Ujk = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, c0.x + 1, t));
                    const Blk<VT> Lij = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, c0.y, t));
                    const Blk<VT> Uji = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, c0.y + 1, t));
                    const Blk<VT> Dj = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, c0.z, t));
                    mulSub(L, Lij, mul(Dj, Ujk));
                    mulSub(U, Lkj, mul(Dj, Uji));
                }
                if (S > 1) {
                    groupSumBlk<T>(L, S, lm.tm);
                    groupSumBlk<T>(U, S, lm.tm);
                }
                if (s == 0) {
                    stBlk<VT>(own, Blk<VT>{A.a + L.a, A.b + L.b, A.c + L.c, A.d + L.d});
                    stBlk<VT>(own + 4 * T, Blk<VT>{B.a + U.a, B.b + U.b, B.c + U.c, B.d + U.d});
                }
            }
        }
        __syncthreads();
    }
}

// Forward solve with existing factors: z(k) = D(k)^-1 (b(k) - sum L~(k,j) z(j)), b = sign * rhs. Ends with a
// barrier.
template <int T, typename VT>
__device__ void forwardOnly(const DevPlan &P, const LevelMeta &M, const DevTop &top, int tile, int tid, int nt,
                            const VT *__restrict__ lu, VT *__restrict__ z, const double *__restrict__ rhs, double sign,
                            int l0, const int *laneOn) {
    const int nb = P.nb;
    const LaneMap lm = laneMapOf<T>(laneOn);
    for (int l = 0; l < l0; l++) {
        const int b0 = M.levelBucketPtr[l], b1 = M.levelBucketPtr[l + 1];
        const int units = bucketUnits<T>(M.bucketRec, b0, b1, true, lm.tm);
        for (int w = tid; w < units; w += nt) {
            const WorkUnit un = unitOf<T>(M.bucketRec, b0, b1, w, true, lm);
            const int t = un.t, s = un.s, S = un.S;
            if (laneOn && !laneOn[t]) continue;
            const int4 bk = M.bucketRec[un.b];
            const int2 pad = M.bucketPad[un.b];
            const int4 *cp = top.contribPad + pad.x + (size_t)un.q * S + s;
            const size_t roundStride = (size_t)(bk.y - bk.x) * S;
            int4 next = pad.y > 0 ? cp[0] : make_int4(-1, -1, -1, -1);   // issued before the item record
            const int4 it = P.itemRec[un.rec];
            double2 b = make_double2(0.0, 0.0);
            Blk<VT> Dinv = {(VT)0, (VT)0, (VT)0, (VT)0};
            if (s == 0) {
                b = *reinterpret_cast<const double2 *>(rhs + vIdx<T>(tile, nb, it.w, t));
                Dinv = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, it.x, t));
            }
            V2<VT> yk = {(VT)0, (VT)0};
#pragma unroll 4
            for (int r = 0; r < pad.y; r++) {
                const int4 cr = next;
                if (cr.x < 0) break;
                if (r + 1 < pad.y) next = cp[(r + 1) * roundStride];
                const Blk<VT> Lkj = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, cr.x, t));
                mulSubV(yk, Lkj, ldV<VT>(z + vIdx<T>(tile, nb, cr.y, t)));
            }
            if (S > 1) {
                yk.p = groupSum<T>(yk.p, S, lm.tm);
                yk.q = groupSum<T>(yk.q, S, lm.tm);
            }
            if (s == 0) {
                const V2<VT> y = {(VT)(sign * b.x) + yk.p, (VT)(sign * b.y) + yk.q};
                const V2<VT> zk = {Dinv.a * y.p + Dinv.b * y.q, Dinv.c * y.p + Dinv.d * y.q};
                stV<VT>(z + vIdx<T>(tile, nb, it.w, t), zk);
            }
        }
        __syncthreads();
    }
}

// The FP64 step of a node as the backward solve finalises its x: dx64 = x (acc false) or dx64 += x (acc true).
template <int T, typename VT>
__device__ __forceinline__ void putStep(double *dx64, bool acc, size_t vi, const V2<VT> &x) {
    if (!dx64) return;
    double2 *p = reinterpret_cast<double2 *>(dx64 + vi);
    const double2 o = acc ? *p : make_double2(0.0, 0.0);
    *p = make_double2(o.x + (double)x.p, o.y + (double)x.q);
}

// Backward solve: x(k) = z(k) - D(k)^-1 sum U(k,i) x(i), levels from the root down; a node's sum is split across
// S threads as in the factorization. With dx64, each x(k) also goes to (or, acc, onto) the FP64 step there (putStep).
// Ends with a barrier.
template <int T, typename VT>
__device__ void backward(const DevPlan &P, const LevelMeta &M, const DevTop &top, int tile, int tid, int nt,
                         const VT *__restrict__ lu, const VT *__restrict__ z, VT *__restrict__ xo, int l0,
                         const int *laneOn, double *dx64 = nullptr, bool accStep = false) {
    const int nb = P.nb;
    const LaneMap lm = laneMapOf<T>(laneOn);
    for (int l = l0 - 1; l >= 0; l--) {
        const int b0 = M.levelBwdBucketPtr[l], b1 = M.levelBwdBucketPtr[l + 1];
        const int units = bucketUnits<T>(M.bwdBucketRec, b0, b1, false, lm.tm);
        for (int w = tid; w < units; w += nt) {
            const WorkUnit un = unitOf<T>(M.bwdBucketRec, b0, b1, w, false, lm);
            const int t = un.t, s = un.s, S = un.S;
            if (laneOn && !laneOn[t]) continue;
            const int4 bk = M.bwdBucketRec[un.b];
            const int2 pad = M.bwdBucketPad[un.b];
            const int2 *ep = top.bwdPad + pad.x + (size_t)un.q * S + s;
            const size_t roundStride = (size_t)(bk.y - bk.x) * S;
            int2 next = pad.y > 0 ? ep[0] : make_int2(-1, -1);   // issued before the node record
            const int4 br = P.bwdRec[un.rec];
            Blk<VT> Dinv = {(VT)0, (VT)0, (VT)0, (VT)0};
            V2<VT> zk = {(VT)0, (VT)0};
            if (s == 0) {
                Dinv = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, br.x, t));
                zk = ldV<VT>(z + vIdx<T>(tile, nb, br.w, t));
            }
            V2<VT> acc = {(VT)0, (VT)0};
#pragma unroll 4
            for (int r = 0; r < pad.y; r++) {
                const int2 e = next;
                if (e.x < 0) break;
                if (r + 1 < pad.y) next = ep[(r + 1) * roundStride];
                const Blk<VT> U = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, e.x, t));
                mulSubV(acc, U, ldV<VT>(xo + vIdx<T>(tile, nb, e.y, t)));
            }
            if (S > 1) {
                acc.p = groupSum<T>(acc.p, S, lm.tm);
                acc.q = groupSum<T>(acc.q, S, lm.tm);
            }
            if (s == 0) {
                const V2<VT> r = {zk.p + Dinv.a * acc.p + Dinv.b * acc.q, zk.q + Dinv.c * acc.p + Dinv.d * acc.q};
                stV<VT>(xo + vIdx<T>(tile, nb, br.w, t), r);
                putStep<T, VT>(dx64, accStep, vIdx<T>(tile, nb, br.w, t), r);
            }
        }
        __syncthreads();
    }
}

// ---- top block in LDS. Per lane: M (n x (n+1) blocks; column n holds the right-hand side in its .a/.c entries),
// the pivot inverses D^-1 (n blocks), z (n pairs) and the backward accumulator (n pairs); lane fastest.
template <int T, typename VT>
struct TopLds {
    VT *base;
    int n;
    __device__ __forceinline__ VT *M(int b, int c, int t) const {
        return base + ((size_t)(b * (n + 1) + c) * T + t) * 4;
    }
    __device__ __forceinline__ VT *Dinv(int a, int t) const {
        return base + ((size_t)n * (n + 1) * T + (size_t)a * T + t) * 4;
    }
    __device__ __forceinline__ VT *Z(int a, int t) const {
        return base + (size_t)n * (n + 2) * T * 4 + ((size_t)a * T + t) * 2;
    }
    __device__ __forceinline__ VT *Acc(int a, int t) const {
        return base + (size_t)n * (n + 2) * T * 4 + (size_t)n * T * 2 + ((size_t)a * T + t) * 2;
    }
};
// LDS bytes of a top block of n nodes for T lanes
__host__ __device__ constexpr size_t topLdsBytes(int T, int n, size_t vtBytes) {
    return (size_t)T * ((size_t)4 * n * (n + 1) + 4 * (size_t)n + 4 * (size_t)n) * vtBytes;
}

// Adds to the top block's right-hand side column (and, with FACTOR, to its matrix blocks) the contributions of the
// nodes below the block: the items of the top levels summed over their contributors j of levels < l0.
template <int T, typename VT, bool FACTOR>
__device__ void topFromBelow(const DevPlan &P, const DevTop &top, const TopLds<T, VT> &L, int tile, int tid, int nt,
                             const VT *__restrict__ lu, const VT *__restrict__ z, const int *laneOn) {
    const int nb = P.nb;
    const LaneMap lm = laneMapOf<T>(laneOn);
    // without FACTOR only the diagonal items (one per top node) matter
    const int4 *bk = FACTOR ? top.buckets : top.diagBuckets;
    const int *list = FACTOR ? top.items : top.diagItems;
    const int nbk = FACTOR ? top.nBuckets : top.nDiagBuckets;
    const int units = bucketUnits<T>(bk, 0, nbk, false, lm.tm);
    for (int w = tid; w < units; w += nt) {
        const WorkUnit un = unitOf<T>(bk, 0, nbk, w, false, lm);
        const int t = un.t, s = un.s, S = un.S;
        if (laneOn && !laneOn[t]) continue;
        const int item = list[un.rec];
        const int4 it = P.itemRec[item];
        const int2 ki = P.itemKI[item];
        const int a = P.levelPos[ki.x] - top.start;
        if (it.w >= 0) {
            Blk<VT> D = {(VT)0, (VT)0, (VT)0, (VT)0};
            V2<VT> yk = {(VT)0, (VT)0};
            for (int c = it.y + s; c < it.z; c += S) {
                const int4 cr = P.contribRec[c];
                if (cr.w >= top.l0) continue;
                const Blk<VT> Lkj = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, cr.x, t));
                if (FACTOR) {
                    const Blk<VT> Ujk = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, cr.x + 1, t));
                    const Blk<VT> Dj = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, cr.z, t));
                    mulSub(D, Lkj, mul(Dj, Ujk));
                }
                mulSubV(yk, Lkj, ldV<VT>(z + vIdx<T>(tile, nb, cr.y, t)));
            }
            if (S > 1) {
                if (FACTOR) groupSumBlk<T>(D, S, lm.tm);
                yk.p = groupSum<T>(yk.p, S, lm.tm);
                yk.q = groupSum<T>(yk.q, S, lm.tm);
            }
            if (s == 0) {
                if (FACTOR) {
                    VT *m = L.M(a, a, t);
                    const Blk<VT> A = ldBlk<VT>(m);
                    stBlk<VT>(m, Blk<VT>{A.a + D.a, A.b + D.b, A.c + D.c, A.d + D.d});
                }
                VT *r = L.M(a, top.n, t);
                const Blk<VT> R = ldBlk<VT>(r);
                stBlk<VT>(r, Blk<VT>{R.a + yk.p, (VT)0, R.c + yk.q, (VT)0});
            }
        } else if (FACTOR) {
            const int b = P.levelPos[ki.y] - top.start;
            Blk<VT> Lx = {(VT)0, (VT)0, (VT)0, (VT)0};
            Blk<VT> Ux = {(VT)0, (VT)0, (VT)0, (VT)0};
            for (int c = it.y + s; c < it.z; c += S) {
                const int4 c0 = P.contribRec[c];
                if (c0.w >= top.l0) continue;
                const Blk<VT> Lkj = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, c0.x, t));
                const Blk<VT> Ujk = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, c0.x + 1, t));
                const Blk<VT> Lij = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, c0.y, t));
                const Blk<VT> Uji = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, c0.y + 1, t));
                const Blk<VT> Dj = ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, c0.z, t));
                mulSub(Lx, Lij, mul(Dj, Ujk));
                mulSub(Ux, Lkj, mul(Dj, Uji));
            }
            if (S > 1) {
                groupSumBlk<T>(Lx, S, lm.tm);
                groupSumBlk<T>(Ux, S, lm.tm);
            }
            if (s == 0) {
                VT *ml = L.M(b, a, t);   // L~(i,k): row i, column k
                VT *mu = L.M(a, b, t);   // U(k,i): row k, column i
                const Blk<VT> A = ldBlk<VT>(ml);
                const Blk<VT> B = ldBlk<VT>(mu);
                stBlk<VT>(ml, Blk<VT>{A.a + Lx.a, A.b + Lx.b, A.c + Lx.c, A.d + Lx.d});
                stBlk<VT>(mu, Blk<VT>{B.a + Ux.a, B.b + Ux.b, B.c + Ux.c, B.d + Ux.d});
            }
        }
    }
    __syncthreads();
}

// Factorization of the top block with its forward solve: loads the block's slots and b = sign * rhs, adds the
// contributions from below, then eliminates right-looking, one barrier per node. Leaves the factors in the stored form
// of the sparse levels (M below the diagonal L~, above it U, D^-1 apart) and z in LDS. Ends with a barrier.
template <int T, typename VT>
__device__ void topFactor(const DevPlan &P, const DevTop &top, const TopLds<T, VT> &L, int tile, int tid, int nt,
                          const VT *__restrict__ lu, const VT *__restrict__ z, const double *__restrict__ rhs,
                          double sign, int *clampLane, const int *laneOn) {
    const int nb = P.nb, n = top.n;
    const Blk<VT> zero = {(VT)0, (VT)0, (VT)0, (VT)0};
    for (int w = tid; w < n * (n + 1) * T; w += nt) stBlk<VT>(L.base + (size_t)w * 4, zero);
    __syncthreads();
    const LaneMap lm = laneMapOf<T>(laneOn);
    for (int w = tid; w < (top.nSlots + n) * lm.tm; w += nt) {
        const int t = lm.tm == 1 ? lm.t0 : w % T, q = lm.tm == 1 ? w : w / T;
        if (laneOn && !laneOn[t]) continue;
        if (q < top.nSlots) {
            const int4 sr = top.slots[q];
            stBlk<VT>(L.M(sr.y, sr.z, t), ldBlk<VT>(lu + bIdx<T>(tile, P.nSlots, sr.x, t)));
        } else {
            const int a = q - top.nSlots;
            const int k = P.levelNodes[top.start + a];
            const double2 b = *reinterpret_cast<const double2 *>(rhs + vIdx<T>(tile, nb, k, t));
            stBlk<VT>(L.M(a, n, t), Blk<VT>{(VT)(sign * b.x), (VT)0, (VT)(sign * b.y), (VT)0});
        }
    }
    __syncthreads();
    topFromBelow<T, VT, true>(P, top, L, tile, tid, nt, lu, z, laneOn);
    if (tid < T && (!laneOn || laneOn[tid])) {
        bool cl;
        stBlk<VT>(L.Dinv(0, tid), inv2(ldBlk<VT>(L.M(0, 0, tid)), cl, pivotDeltaOf<VT>(top)));
        if (cl && clampLane) atomicAdd(&clampLane[tid], 1);
    }
    __syncthreads();
    for (int a = 0; a < n; a++) {
        const int m = n - 1 - a;   // rows below a; columns a+1 .. n (n: right-hand side)
        if (tid < T && (!laneOn || laneOn[tid])) {
            const Blk<VT> Di = ldBlk<VT>(L.Dinv(a, tid));
            const Blk<VT> ya = ldBlk<VT>(L.M(a, n, tid));
            stV<VT>(L.Z(a, tid), V2<VT>{Di.a * ya.a + Di.b * ya.c, Di.c * ya.a + Di.d * ya.c});
        }
        for (int w = tid; w < m * (m + 1) * T; w += nt) {
            const int t = w % T, e = w / T;
            if (laneOn && !laneOn[t]) continue;
            const int b = a + 1 + e / (m + 1), c = a + 1 + e % (m + 1);
            const Blk<VT> Di = ldBlk<VT>(L.Dinv(a, t));
            VT *mb = L.M(b, c, t);
            Blk<VT> acc = ldBlk<VT>(mb);
            mulSub(acc, ldBlk<VT>(L.M(b, a, t)), mul(Di, ldBlk<VT>(L.M(a, c, t))));
            stBlk<VT>(mb, acc);
            if (b == a + 1 && c == a + 1) {
                // the next pivot is final: invert it for the next step
                bool cl;
                stBlk<VT>(L.Dinv(b, t), inv2(acc, cl, pivotDeltaOf<VT>(top)));
                if (cl && clampLane) atomicAdd(&clampLane[t], 1);
            }
        }
        __syncthreads();
    }
}

// Forward solve of the top block with existing factors: b = sign * rhs, contributions from below, then the dense
// forward substitution into z (LDS). Ends with a barrier.
template <int T, typename VT>
__device__ void topForward(const DevPlan &P, const DevTop &top, const TopLds<T, VT> &L, int tile, int tid, int nt,
                           const VT *__restrict__ lu, const VT *__restrict__ z, const double *__restrict__ rhs,
                           double sign, const int *laneOn) {
    const int nb = P.nb, n = top.n;
    for (int w = tid; w < n * T; w += nt) {
        const int t = w % T, a = w / T;
        if (laneOn && !laneOn[t]) continue;
        const int k = P.levelNodes[top.start + a];
        const double2 b = *reinterpret_cast<const double2 *>(rhs + vIdx<T>(tile, nb, k, t));
        stBlk<VT>(L.M(a, n, t), Blk<VT>{(VT)(sign * b.x), (VT)0, (VT)(sign * b.y), (VT)0});
    }
    __syncthreads();
    topFromBelow<T, VT, false>(P, top, L, tile, tid, nt, lu, z, laneOn);
    for (int a = 0; a < n; a++) {
        for (int w = tid; w < (n - a) * T; w += nt) {
            const int t = w % T, b = a + w / T;   // b = a: store z(a); b > a: eliminate
            if (laneOn && !laneOn[t]) continue;
            const Blk<VT> Di = ldBlk<VT>(L.Dinv(a, t));
            const Blk<VT> ya = ldBlk<VT>(L.M(a, n, t));
            const V2<VT> za = {Di.a * ya.a + Di.b * ya.c, Di.c * ya.a + Di.d * ya.c};
            if (b == a) {
                stV<VT>(L.Z(a, t), za);
            } else {
                VT *mb = L.M(b, n, t);
                Blk<VT> r = ldBlk<VT>(mb);
                const Blk<VT> Lba = ldBlk<VT>(L.M(b, a, t));
                r.a -= Lba.a * za.p + Lba.b * za.q;
                r.c -= Lba.c * za.p + Lba.d * za.q;
                stBlk<VT>(mb, r);
            }
        }
        __syncthreads();
    }
}

// Backward solve of the top block: x(a) = z(a) - D(a)^-1 sum_{c > a} U(a,c) x(c), from the last node down, one
// barrier per node; x of the top nodes goes to xo for the sparse levels below. Ends with a barrier.
template <int T, typename VT>
__device__ void topBackward(const DevPlan &P, const DevTop &top, const TopLds<T, VT> &L, int tile, int tid, int nt,
                            VT *__restrict__ xo, const int *laneOn, double *dx64 = nullptr, bool accStep = false) {
    const int nb = P.nb, n = top.n;
    for (int w = tid; w < n * T; w += nt) stV<VT>(L.Acc(w / T, w % T), V2<VT>{(VT)0, (VT)0});
    __syncthreads();
    for (int a = n - 1; a >= 0; a--) {
        for (int w = tid; w < (a + 1) * T; w += nt) {
            const int t = w % T, b = w / T;   // b = a: store x(a); b < a: accumulate U(b,a) x(a)
            if (laneOn && !laneOn[t]) continue;
            const Blk<VT> Di = ldBlk<VT>(L.Dinv(a, t));
            const V2<VT> za = ldV<VT>(L.Z(a, t));
            const V2<VT> ac = ldV<VT>(L.Acc(a, t));
            const V2<VT> xa = {za.p - (Di.a * ac.p + Di.b * ac.q), za.q - (Di.c * ac.p + Di.d * ac.q)};
            if (b == a) {
                const int k = P.levelNodes[top.start + a];
                stV<VT>(xo + vIdx<T>(tile, nb, k, t), xa);
                putStep<T, VT>(dx64, accStep, vIdx<T>(tile, nb, k, t), xa);
            } else {
                const Blk<VT> U = ldBlk<VT>(L.M(b, a, t));
                V2<VT> ab = ldV<VT>(L.Acc(b, t));
                ab.p += U.a * xa.p + U.b * xa.q;
                ab.q += U.c * xa.p + U.d * xa.q;
                stV<VT>(L.Acc(b, t), ab);
            }
        }
        __syncthreads();
    }
}

// ---- the top block by panels of TOP_PANEL nodes: a panel's own sequential work runs inside one wave (or one thread)
// per lane with no workgroup barrier, and what the panel does to the rest of the block is one parallel pass, so a
// panel costs two or three barriers instead of one per node.
constexpr int TOP_PANEL = 8;
// OLFHIP_TOP_BY_NODE=1 at build time: the node-by-node forms throughout (one barrier per node), for comparison
#ifndef OLFHIP_TOP_BY_NODE
#define OLFHIP_TOP_BY_NODE 0
#endif

// Orders the wave's LDS accesses before this point before those after it (the wave's lanes exchange values through
// LDS between the steps of a panel).
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef float topF32x4 __attribute__((ext_vector_type(4)));

// topFactor's elimination by panels (FP32 factors): for each panel [a0, a1)
//   1. its columns, every row below its first node, eliminated node by node inside one wave per lane;
//   2. its rows right of it (the right-hand side column included) by the panel's unit-lower solve, one thread per
//      (column, lane), and z of its nodes;
//   3. the trailing block, rows [a1, n) x columns [a1, n], minus L~(b, a) D(a)^-1 U(a, c) summed over the panel's
//      nodes a: a product of inner size 2 x TOP_PANEL on the matrix cores (v_mfma_f32_16x16x4f32), one wave per
//      (lane, 16-column strip) walking the strip's 16 x 16 tiles, the D^-1 U operand kept in registers.
// The same updates as the node-by-node elimination in another summation order: the stored form of the factors
// (L~ below the diagonal, U above, D^-1 apart) and z are those topFactor leaves, to FP32 rounding.
template <int T>
__device__ void topFactorBlocked(const DevPlan &P, const DevTop &top, const TopLds<T, float> &L, int tile, int tid,
                                 int nt, const float *__restrict__ lu, const float *__restrict__ z,
                                 const double *__restrict__ rhs, double sign, int *clampLane, const int *laneOn) {
    const int nb = P.nb, n = top.n;
    const Blk<float> zero = {0.f, 0.f, 0.f, 0.f};
    for (int w = tid; w < n * (n + 1) * T; w += nt) stBlk<float>(L.base + (size_t)w * 4, zero);
    __syncthreads();
    const LaneMap lm = laneMapOf<T>(laneOn);
    for (int w = tid; w < (top.nSlots + n) * lm.tm; w += nt) {
        const int t = lm.tm == 1 ? lm.t0 : w % T, q = lm.tm == 1 ? w : w / T;
        if (laneOn && !laneOn[t]) continue;
        if (q < top.nSlots) {
            const int4 sr = top.slots[q];
            stBlk<float>(L.M(sr.y, sr.z, t), ldBlk<float>(lu + bIdx<T>(tile, P.nSlots, sr.x, t)));
        } else {
            const int a = q - top.nSlots;
            const int k = P.levelNodes[top.start + a];
            const double2 b = *reinterpret_cast<const double2 *>(rhs + vIdx<T>(tile, nb, k, t));
            stBlk<float>(L.M(a, n, t), Blk<float>{(float)(sign * b.x), 0.f, (float)(sign * b.y), 0.f});
        }
    }
    __syncthreads();
    topFromBelow<T, float, true>(P, top, L, tile, tid, nt, lu, z, laneOn);
    const int wave = tid >> 6, wl = tid & 63, nWaves = nt >> 6;
    for (int a0 = 0; a0 < n; a0 += TOP_PANEL) {
        const int a1 = min(n, a0 + TOP_PANEL), pw = a1 - a0;
        // 1. the panel's columns
        for (int t = wave; t < T; t += nWaves) {
            if (laneOn && !laneOn[t]) continue;
            for (int a = a0; a < a1; a++) {
                bool cl;
                const Blk<float> Di = inv2(ldBlk<float>(L.M(a, a, t)), cl, 0.f);
                if (wl == 0) {
                    stBlk<float>(L.Dinv(a, t), Di);
                    if (cl && clampLane) atomicAdd(&clampLane[t], 1);
                }
                const int rows = n - 1 - a, cols = a1 - 1 - a;
                for (int e = wl; e < rows * cols; e += 64) {
                    const int b = a + 1 + e / cols, c = a + 1 + e % cols;
                    float *mb = L.M(b, c, t);
                    Blk<float> acc = ldBlk<float>(mb);
                    mulSub(acc, ldBlk<float>(L.M(b, a, t)), mul(Di, ldBlk<float>(L.M(a, c, t))));
                    stBlk<float>(mb, acc);
                }
                waveSync();
            }
        }
        __syncthreads();
        // 2. the panel's rows right of it, and z of its nodes (from the right-hand side column n), in place
        for (int w = tid; w < (n + 1 - a1) * T; w += nt) {
            const int t = w % T, c = a1 + w / T;
            if (laneOn && !laneOn[t]) continue;
            for (int q = 0; q < pw; q++) {
                const Blk<float> u = mul(ldBlk<float>(L.Dinv(a0 + q, t)), ldBlk<float>(L.M(a0 + q, c, t)));
                if (c == n) stV<float>(L.Z(a0 + q, t), V2<float>{u.a, u.c});
                for (int r = q + 1; r < pw; r++) {
                    float *mr = L.M(a0 + r, c, t);
                    Blk<float> v = ldBlk<float>(mr);
                    mulSub(v, ldBlk<float>(L.M(a0 + r, a0 + q, t)), u);
                    stBlk<float>(mr, v);
                }
            }
        }
        __syncthreads();
        // 3. the trailing block on the matrix cores
        const int rowsS = 2 * (n - a1), colsS = 2 * (n - a1) + 1, kS = 2 * pw;
        if (rowsS > 0) {
            const int rT = (rowsS + 15) / 16, cT = (colsS + 15) / 16;
            const int i16 = wl & 15, g = wl >> 4;
            for (int job = wave; job < T * cT; job += nWaves) {
                const int t = job % T, ct = job / T;
                if (laneOn && !laneOn[t]) continue;   // uniform per wave
                // D^-1 U operand of the strip: lane (g, i16) holds row k = 4 ks + g, column Cb, of each k step
                const int Cb = 2 * a1 + ct * 16 + i16;
                float bk[TOP_PANEL / 2];
#pragma unroll
                for (int ks = 0; ks < TOP_PANEL / 2; ks++) {
                    const int k = 4 * ks + g;
                    float v = 0.f;
                    if (k < kS && Cb <= 2 * n) {
                        const int a = a0 + (k >> 1), kc = k & 1;
                        const float *d = L.Dinv(a, t);
                        const float *m = L.M(a, Cb >> 1, t);
                        v = d[2 * kc] * m[Cb & 1] + d[2 * kc + 1] * m[2 + (Cb & 1)];
                    }
                    bk[ks] = v;
                }
                for (int rt = 0; rt < rT; rt++) {
                    const int Ra = 2 * a1 + rt * 16 + i16;
                    topF32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int ks = 0; ks < TOP_PANEL / 2; ks++) {
                        if (4 * ks >= kS) break;   // uniform
                        const int k = 4 * ks + g;
                        const float av = (k < kS && Ra < 2 * n)
                                             ? L.M(Ra >> 1, a0 + (k >> 1), t)[2 * (Ra & 1) + (k & 1)]
                                             : 0.f;
                        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bk[ks], acc, 0, 0, 0);
                    }
                    // lane (g, i16) holds rows 4 g + q, column i16 of the tile
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int R = 2 * a1 + rt * 16 + 4 * g + q;
                        if (R < 2 * n && Cb <= 2 * n) {
                            float *p = L.M(R >> 1, Cb >> 1, t) + 2 * (R & 1) + (Cb & 1);
                            *p -= acc[q];
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
}

// topForward by panels: a panel's forward substitution in one thread per lane, then the rows below it minus
// L~(b, a) z(a) over the panel's nodes, one thread per (row, lane). Two barriers per panel.
template <int T, typename VT>
__device__ void topForwardBlocked(const DevPlan &P, const DevTop &top, const TopLds<T, VT> &L, int tile, int tid,
                                  int nt, const VT *__restrict__ lu, const VT *__restrict__ z,
                                  const double *__restrict__ rhs, double sign, const int *laneOn) {
    const int nb = P.nb, n = top.n;
    for (int w = tid; w < n * T; w += nt) {
        const int t = w % T, a = w / T;
        if (laneOn && !laneOn[t]) continue;
        const int k = P.levelNodes[top.start + a];
        const double2 b = *reinterpret_cast<const double2 *>(rhs + vIdx<T>(tile, nb, k, t));
        stBlk<VT>(L.M(a, n, t), Blk<VT>{(VT)(sign * b.x), (VT)0, (VT)(sign * b.y), (VT)0});
    }
    __syncthreads();
    topFromBelow<T, VT, false>(P, top, L, tile, tid, nt, lu, z, laneOn);
    for (int a0 = 0; a0 < n; a0 += TOP_PANEL) {
        const int a1 = min(n, a0 + TOP_PANEL), pw = a1 - a0;
        if (tid < T && (!laneOn || laneOn[tid])) {
            // the panel's forward substitution, in place in the right-hand side column
            const int t = tid;
            for (int q = 0; q < pw; q++) {
                const Blk<VT> Di = ldBlk<VT>(L.Dinv(a0 + q, t));
                const Blk<VT> yq = ldBlk<VT>(L.M(a0 + q, n, t));
                const V2<VT> zq = {Di.a * yq.a + Di.b * yq.c, Di.c * yq.a + Di.d * yq.c};
                stV<VT>(L.Z(a0 + q, t), zq);
                for (int r = q + 1; r < pw; r++) {
                    VT *mr = L.M(a0 + r, n, t);
                    Blk<VT> y = ldBlk<VT>(mr);
                    const Blk<VT> Lrq = ldBlk<VT>(L.M(a0 + r, a0 + q, t));
                    y.a -= Lrq.a * zq.p + Lrq.b * zq.q;
                    y.c -= Lrq.c * zq.p + Lrq.d * zq.q;
                    stBlk<VT>(mr, y);
                }
            }
        }
        __syncthreads();
        if (a1 < n) {
            for (int w = tid; w < (n - a1) * T; w += nt) {
                const int t = w % T, b = a1 + w / T;
                if (laneOn && !laneOn[t]) continue;
                VT *mb = L.M(b, n, t);
                Blk<VT> r = ldBlk<VT>(mb);
                for (int a = a0; a < a1; a++) {
                    const V2<VT> za = ldV<VT>(L.Z(a, t));
                    const Blk<VT> Lba = ldBlk<VT>(L.M(b, a, t));
                    r.a -= Lba.a * za.p + Lba.b * za.q;
                    r.c -= Lba.c * za.p + Lba.d * za.q;
                }
                stBlk<VT>(mb, r);
            }
            __syncthreads();
        }
    }
}

// topBackward by panels, from the last one down: a panel's back substitution in one thread per lane (its rows'
// accumulators already hold the sum over the nodes after the panel), then the accumulators of the rows before it
// plus U(b, a) x(a) over the panel's nodes, one thread per (row, lane). Two barriers per panel.
template <int T, typename VT>
__device__ void topBackwardBlocked(const DevPlan &P, const DevTop &top, const TopLds<T, VT> &L, int tile, int tid,
                                   int nt, VT *__restrict__ xo, const int *laneOn, double *dx64 = nullptr,
                                   bool accStep = false) {
    const int nb = P.nb, n = top.n;
    for (int w = tid; w < n * T; w += nt) stV<VT>(L.Acc(w / T, w % T), V2<VT>{(VT)0, (VT)0});
    __syncthreads();
    const int nPanels = (n + TOP_PANEL - 1) / TOP_PANEL;
    for (int pi = nPanels - 1; pi >= 0; pi--) {
        const int a0 = pi * TOP_PANEL, a1 = min(n, a0 + TOP_PANEL), pw = a1 - a0;
        if (tid < T && (!laneOn || laneOn[tid])) {
            // the panel's back substitution; x(a) replaces row a's accumulator once known
            const int t = tid;
            for (int a = a1 - 1; a >= a0; a--) {
                V2<VT> ac = ldV<VT>(L.Acc(a, t));
                for (int c = a + 1; c < a1; c++) {
                    const Blk<VT> U = ldBlk<VT>(L.M(a, c, t));
                    const V2<VT> xc = ldV<VT>(L.Acc(c, t));
                    ac.p += U.a * xc.p + U.b * xc.q;
                    ac.q += U.c * xc.p + U.d * xc.q;
                }
                const Blk<VT> Di = ldBlk<VT>(L.Dinv(a, t));
                const V2<VT> za = ldV<VT>(L.Z(a, t));
                const V2<VT> xa = {za.p - (Di.a * ac.p + Di.b * ac.q), za.q - (Di.c * ac.p + Di.d * ac.q)};
                stV<VT>(L.Acc(a, t), xa);
                const int k = P.levelNodes[top.start + a];
                stV<VT>(xo + vIdx<T>(tile, nb, k, t), xa);
                putStep<T, VT>(dx64, accStep, vIdx<T>(tile, nb, k, t), xa);
            }
        }
        __syncthreads();
        if (a0 > 0) {
            for (int w = tid; w < a0 * T; w += nt) {
                const int t = w % T, b = w / T;
                if (laneOn && !laneOn[t]) continue;
                V2<VT> ab = ldV<VT>(L.Acc(b, t));
                for (int a = a0; a < a1; a++) {
                    const V2<VT> xa = ldV<VT>(L.Acc(a, t));
                    const Blk<VT> U = ldBlk<VT>(L.M(b, a, t));
                    ab.p += U.a * xa.p + U.b * xa.q;
                    ab.q += U.c * xa.p + U.d * xa.q;
                }
                stV<VT>(L.Acc(b, t), ab);
            }
            __syncthreads();
        }
    }
}

// The whole block LU: sparse levels below the top block, then the top block (n = 0: sparse only).
template <int T, typename VT>
__device__ __forceinline__ void luFactorForward(const DevPlan &P, const LevelMeta &M, const DevTop &top,
                                                const TopLds<T, VT> &L, int tile, int tid, int nt, VT *lu, VT *z,
                                                const double *rhs, double sign, int *clampLane, const int *laneOn,
                                                unsigned long long *prof = nullptr, int mark = 0) {
    factorForward<T, VT>(P, M, top, tile, tid, nt, lu, z, rhs, sign, clampLane, top.l0, laneOn);
    if (prof && tid == 0) prof[mark] = wall_clock64();   // opt-in section stamp between the two parts
    if (top.n > 0) {
        if constexpr (sizeof(VT) == 4) {
            if constexpr (!OLFHIP_TOP_BY_NODE) {
                topFactorBlocked<T>(P, top, L, tile, tid, nt, lu, z, rhs, sign, clampLane, laneOn);
                return;
            }
        }
        topFactor<T, VT>(P, top, L, tile, tid, nt, lu, z, rhs, sign, clampLane, laneOn);
    }
}
template <int T, typename VT>
__device__ __forceinline__ void luForward(const DevPlan &P, const LevelMeta &M, const DevTop &top,
                                          const TopLds<T, VT> &L, int tile, int tid, int nt, const VT *lu, VT *z,
                                          const double *rhs, double sign, const int *laneOn,
                                          unsigned long long *prof = nullptr, int mark = 0) {
    forwardOnly<T, VT>(P, M, top, tile, tid, nt, lu, z, rhs, sign, top.l0, laneOn);
    if (prof && tid == 0) prof[mark] = wall_clock64();
    if (top.n > 0) {
        if constexpr (!OLFHIP_TOP_BY_NODE) topForwardBlocked<T, VT>(P, top, L, tile, tid, nt, lu, z, rhs, sign, laneOn);
        else topForward<T, VT>(P, top, L, tile, tid, nt, lu, z, rhs, sign, laneOn);
    }
}
template <int T, typename VT>
__device__ __forceinline__ void luBackward(const DevPlan &P, const LevelMeta &M, const DevTop &top,
                                           const TopLds<T, VT> &L, int tile, int tid, int nt, const VT *lu,
                                           const VT *z, VT *xo, const int *laneOn, double *dx64 = nullptr,
                                           bool accStep = false, unsigned long long *prof = nullptr,
                                           int mark = 0) {
    if (top.n > 0) {
        if constexpr (!OLFHIP_TOP_BY_NODE) topBackwardBlocked<T, VT>(P, top, L, tile, tid, nt, xo, laneOn, dx64, accStep);
        else topBackward<T, VT>(P, top, L, tile, tid, nt, xo, laneOn, dx64, accStep);
    }
    if (prof && tid == 0) prof[mark] = wall_clock64();
    backward<T, VT>(P, M, top, tile, tid, nt, lu, z, xo, top.l0, laneOn, dx64, accStep);
}

// Mismatch and Jacobian fill of the lanes flagged in fillLane; per-thread partial max|F| and finiteness of the
// thread's lane (tid % T) are returned.
template <int T, typename VT>
__device__ void fillPass(const DevPlan &P, const DevState &S, int tile, int tid, int nt, const int *fillLane,
                         double &pmax, int &pfin, double &psq, int slackNode) {
    const int nb = P.nb;
    VT *lu = reinterpret_cast<VT *>(S.lu);
    // fixed per thread: the stride nt is a multiple of T (a lane alone: every thread on it)
    const LaneMap lm = laneMapOf<T>(fillLane);
    const int t = lm.tm == 1 ? lm.t0 : tid % T;
    if (!fillLane[t]) return;
    int o[MAX_OUTAGES] = {-1, -1, -1, -1};
    for (int q = 0; q < S.maxOut && q < MAX_OUTAGES; q++) o[q] = S.outages[(size_t)(tile * T + t) * S.maxOut + q];
    for (int w = tid; w < nb * lm.tm; w += nt) {
        const int k = lm.tm == 1 ? w : w / T;
        const size_t vi = vIdx<T>(tile, nb, k, t);
        const signed char mP = S.mode[vi], mQ = S.mode[vi + 1];
        double sP, sQ, jP, jQ;
        nodeTerms<T, VT, true, false>(P, tile, k, t, S.x, o, lu, nullptr, mP, mQ, sP, sQ, jP, jQ);
        const double2 tg = *reinterpret_cast<const double2 *>(S.target + vi);
        const double2 xk = *reinterpret_cast<const double2 *>(S.x + vi);
        const double fP = (mP == MODE_FIXED) ? xk.x - tg.x : sP - tg.x;
        const double fQ = (mQ == MODE_FIXED) ? xk.y - tg.y : sQ - tg.y;
        *reinterpret_cast<double2 *>(S.F + vi) = make_double2(fP, fQ);
        S.qinj[qIdx<T>(tile, nb, k, t)] = sQ;
        if (k == slackNode) S.slackP[tile * T + t] = sP;
        if (!isfinite(fP) || !isfinite(fQ)) pfin = 0;
        pmax = fmax(pmax, fmax(fabs(fP), fabs(fQ)));
        psq += fP * fP + fQ * fQ;
    }
}

// The per-row criterion's max |F_i| / threshold of row i's mode for the lanes flagged, per thread (after a fill):
// a pass of its own, so that the fill carries nothing for it when the criterion is not set.
template <int T>
__device__ void rowRatioPass(const DevPlan &P, const DevState &S, int tile, int tid, int nt, const int *laneOn,
                             double &pratio) {
    const int nb = P.nb;
    const LaneMap lm = laneMapOf<T>(laneOn);   // the nodes fillPass gives the thread for the same flags
    const int t = lm.tm == 1 ? lm.t0 : tid % T;
    if (!laneOn[t]) return;
    for (int w = tid; w < nb * lm.tm; w += nt) {
        const int k = lm.tm == 1 ? w : w / T;
        const size_t vi = vIdx<T>(tile, nb, k, t);
        const double2 f = *reinterpret_cast<const double2 *>(S.F + vi);
        const double2 tf = *reinterpret_cast<const double2 *>(S.rowThrFixed + 2 * k);
        const double2 te = *reinterpret_cast<const double2 *>(S.rowThrEq + 2 * k);
        pratio = fmax(pratio, fmax(fabs(f.x) / (S.mode[vi] == MODE_FIXED ? tf.x : te.x),
                                   fabs(f.y) / (S.mode[vi + 1] == MODE_FIXED ? tf.y : te.y)));
    }
}

// out = J * v + add (lanes flagged), per-thread max |out| of the thread's lane.
template <int T, typename VT>
__device__ void applyJPass(const DevPlan &P, const DevState &S, int tile, int tid, int nt, const int *laneOn,
                           const double *v, const double *add, double addSign, double *out, double &pmax) {
    const int nb = P.nb;
    // fixed per thread: the stride nt is a multiple of T (a lane alone: every thread on it)
    const LaneMap lm = laneMapOf<T>(laneOn);
    const int t = lm.tm == 1 ? lm.t0 : tid % T;
    if (!laneOn[t]) return;
    int o[MAX_OUTAGES] = {-1, -1, -1, -1};
    for (int q = 0; q < S.maxOut && q < MAX_OUTAGES; q++) o[q] = S.outages[(size_t)(tile * T + t) * S.maxOut + q];
    for (int w = tid; w < nb * lm.tm; w += nt) {
        const int k = lm.tm == 1 ? w : w / T;
        const size_t vi = vIdx<T>(tile, nb, k, t);
        double sP, sQ, jP, jQ;
        nodeTerms<T, VT, false, true>(P, tile, k, t, S.x, o, nullptr, v, S.mode[vi], S.mode[vi + 1], sP, sQ, jP, jQ);
        const double2 a = *reinterpret_cast<const double2 *>(add + vi);
        const double rP = jP + addSign * a.x, rQ = jQ + addSign * a.y;
        if (out) *reinterpret_cast<double2 *>(out + vi) = make_double2(rP, rQ);
        pmax = fmax(pmax, fmax(fabs(rP), fabs(rQ)));
    }
}

// Per-tile words of the robust solve's small dense GMRES state: the (m + 1) x m Hessenberg matrix (column-major), the
// right-hand side g (m + 1), the rotations' cosines and sines (m each) and the coefficients y (m).
__host__ __device__ inline size_t kryHStride(int m) { return (size_t)(m + 1) * m + (m + 1) + 3 * (size_t)m; }

// The robust solve's correction of dx, the factors' solution of J dx = -F (one lane per tile): restarted GMRES on
// the FP64 Jacobian, applied matrix-free at the lane's state, right-preconditioned by the lane's FP64 factors M
// (Saad & Schultz 1986; the right-preconditioned form of Saad, Iterative Methods for Sparse Linear Systems, 2nd ed.,
// Alg. 9.5), with modified Gram-Schmidt and Givens rotations. It solves J M^-1 u = -(J dx + F) and adds M^-1 u to dx
// until ||J dx + F||_2 <= kryTol ||F||_2, the basis (kryDim vectors) is exhausted kryRestarts + 1 times, or the
// basis breaks down. Where the factors were perturbed by static pivoting, J M^-1 is the identity plus a term of low
// rank, which GMRES removes in a few iterations. Thread tid owns the nodes k = tid (mod nt) in every vector loop, the
// same nodes applyJPass gives it, so a thread reads back what it wrote without a barrier. Ends with a barrier.
template <int T>
__device__ void gmresRefine(const DevPlan &P, const LevelMeta &LM, const DevTop &top, const TopLds<T, double> &TL,
                            const DevState &S, const NrParams &prm, int tile, int tid, int nt, double *lu, double *y,
                            double *z, const int *laneOn, double *red) {
    static_assert(T == 1, "the robust solve runs one lane per tile");
    __shared__ double sRes;
    __shared__ int sK;
    if (!laneOn[0]) return;   // uniform per block
    const int nb = P.nb, m = prm.kryDim;
    const size_t vs = (size_t)S.nTiles * nb * 2;
    double *H = S.kryH + (size_t)tile * kryHStride(m);
    double *g = H + (size_t)(m + 1) * m, *cs = g + (m + 1), *sn = cs + m, *yc = sn + m;
    // vec(j), F, dx, zz and u point at this tile's part and take the node index; the solves and applyJPass take
    // context-wide vectors and index them with vIdx themselves
    auto vec = [&](int j) -> double2 * {
        return reinterpret_cast<double2 *>(S.kryV + (size_t)j * vs + vIdx<1>(tile, nb, 0, 0));
    };
    const double2 *F = reinterpret_cast<const double2 *>(S.F + vIdx<1>(tile, nb, 0, 0));
    double2 *dx = reinterpret_cast<double2 *>(S.dx + vIdx<1>(tile, nb, 0, 0));
    double2 *zz = reinterpret_cast<double2 *>(z + vIdx<1>(tile, nb, 0, 0));
    double2 *u = reinterpret_cast<double2 *>(S.r + vIdx<1>(tile, nb, 0, 0));
    double ss = 0.0;
    for (int k = tid; k < nb; k += nt) ss += F[k].x * F[k].x + F[k].y * F[k].y;
    red[tid] = ss;
    laneReduce<1, false>(red, tid, nt);
    const double target = prm.kryTol * sqrt(red[0]);
    for (int cycle = 0; cycle <= prm.kryRestarts; cycle++) {
        __syncthreads();
        // v0 = -(J dx + F) / beta
        double pm = 0.0;
        applyJPass<1, double>(P, S, tile, tid, nt, laneOn, S.dx, S.F, 1.0, S.kryV, pm);   // indexes from the base
        double2 *v0 = vec(0);
        ss = 0.0;
        for (int k = tid; k < nb; k += nt) ss += v0[k].x * v0[k].x + v0[k].y * v0[k].y;
        red[tid] = ss;
        laneReduce<1, false>(red, tid, nt);
        const double beta = sqrt(red[0]);
        __syncthreads();
        if (!(beta > target)) break;   // converged, or a non-finite residual that the guard reports
        for (int k = tid; k < nb; k += nt) v0[k] = make_double2(-v0[k].x / beta, -v0[k].y / beta);
        if (tid == 0) {
            g[0] = beta;
            sK = 0;
        }
        for (int j = 0; j < m; j++) {
            __syncthreads();   // v_j complete before the solves read it anywhere
            luForward<1, double>(P, LM, top, TL, tile, tid, nt, lu, y, S.kryV + (size_t)j * vs, 1.0, laneOn);
            luBackward<1, double>(P, LM, top, TL, tile, tid, nt, lu, y, z, laneOn);
            __syncthreads();
            // w = J M^-1 v_j into v_{j+1}, orthogonalised against v_0..v_j
            applyJPass<1, double>(P, S, tile, tid, nt, laneOn, z, S.F, 0.0, S.kryV + (size_t)(j + 1) * vs, pm);
            double2 *w = vec(j + 1);
            for (int i = 0; i <= j; i++) {
                const double2 *vi = vec(i);
                double d = 0.0;
                for (int k = tid; k < nb; k += nt) d += w[k].x * vi[k].x + w[k].y * vi[k].y;
                red[tid] = d;
                laneReduce<1, false>(red, tid, nt);
                const double h = red[0];
                __syncthreads();
                for (int k = tid; k < nb; k += nt) w[k] = make_double2(w[k].x - h * vi[k].x, w[k].y - h * vi[k].y);
                if (tid == 0) H[(size_t)j * (m + 1) + i] = h;
            }
            double d = 0.0;
            for (int k = tid; k < nb; k += nt) d += w[k].x * w[k].x + w[k].y * w[k].y;
            red[tid] = d;
            laneReduce<1, false>(red, tid, nt);
            const double hn = sqrt(red[0]);
            if (tid == 0) {
                // the earlier rotations on column j, then the one that zeroes its subdiagonal
                double *h = H + (size_t)j * (m + 1);
                h[j + 1] = hn;
                for (int i = 0; i < j; i++) {
                    const double a = cs[i] * h[i] + sn[i] * h[i + 1];
                    h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                    h[i] = a;
                }
                const double den = hypot(h[j], h[j + 1]);
                const double c = den > 0.0 ? h[j] / den : 1.0, sv = den > 0.0 ? h[j + 1] / den : 0.0;
                cs[j] = c;
                sn[j] = sv;
                h[j] = den;
                h[j + 1] = 0.0;
                g[j + 1] = -sv * g[j];
                g[j] = c * g[j];
                sRes = fabs(g[j + 1]);
                sK = j + 1;
            }
            if (hn > 0.0)
                for (int k = tid; k < nb; k += nt) w[k] = make_double2(w[k].x / hn, w[k].y / hn);
            __syncthreads();
            if (!(sRes > target) || !(hn > 0.0)) break;
        }
        // y = R^-1 g over the sK columns built, then dx += M^-1 (V y)
        const int kk = sK;
        if (tid == 0) {
            for (int i = kk - 1; i >= 0; i--) {
                double acc = g[i];
                for (int l = i + 1; l < kk; l++) acc -= H[(size_t)l * (m + 1) + i] * yc[l];
                const double rii = H[(size_t)i * (m + 1) + i];
                yc[i] = rii != 0.0 ? acc / rii : 0.0;
            }
        }
        __syncthreads();
        for (int k = tid; k < nb; k += nt) {
            double2 a = make_double2(0.0, 0.0);
            for (int i = 0; i < kk; i++) {
                const double2 vv = vec(i)[k];
                a.x += yc[i] * vv.x;
                a.y += yc[i] * vv.y;
            }
            u[k] = a;
        }
        __syncthreads();
        luForward<1, double>(P, LM, top, TL, tile, tid, nt, lu, y, S.r, 1.0, laneOn);
        luBackward<1, double>(P, LM, top, TL, tile, tid, nt, lu, y, z, laneOn);
        __syncthreads();
        for (int k = tid; k < nb; k += nt) dx[k] = make_double2(dx[k].x + zz[k].x, dx[k].y + zz[k].y);
    }
    __syncthreads();
}

constexpr double P_RESIDUE_EPS = 1e-5;   // OLF ActivePowerDistribution.P_RESIDUE_EPS (outer_loop_driver.cu:204)

// One distributed-slack check (DistributedSlackOuterLoop, the reference's per-bus water-fill) of the lanes flagged in
// on[]; every thread of the block calls it. Per lane: the slack's retained share of the total from the anchor
// (native/cuda/outerloop/distributed_slack.cu:131-195), the gate on the participants' share (:215-305), then, when
// it fires, every participant reset to its anchor and the whole total water-filled from there, the factors
// renormalised over the unsaturated participants each pass (:395-420, :530-566), and OLF's outcome: unstable when
// the targets and the slack's applied share moved by more than 0.9 P_RESIDUE_EPS, failed when more than
// P_RESIDUE_EPS stayed unplaced (:568-600). A participant whose P row is not an equation (islanded) takes no part.
// Sums over the participants are per-lane block reductions; participant i of lane t is always handled by the thread
// with tid % T == t, which the strides keep.
template <int T>
__device__ void dsCheck(const DevPlan &P, const DevState &S, const DevRL &R, const NrParams &prm, int tile, int tid,
                        int nt, const int *on, int *unstable, int *failed, double *red) {
    __shared__ double sRem[T], sPrevRetain[T], sSum[T];
    __shared__ int sFire[T], sGo[T];
    const int nb = P.nb;
    const int t = tid % T;
    const DevDS D = *R.ds;
    // (1) the active factors and what the participants hold above their anchors
    double fa = 0.0, ab = 0.0;
    if (on[t]) {
        for (int w = tid; w < D.n * T; w += nt) {
            const int i = w / T;
            const size_t vi = vIdx<T>(tile, nb, D.node[i], t);
            if (S.mode[vi] != MODE_EQUATION) continue;
            fa += D.factor[i];
            ab += S.target[vi] - D.base[i];
        }
    }
    red[tid] = fa;
    laneReduce<T, false>(red, tid, nt);
    if (tid < T) sSum[tid] = red[tid];
    __syncthreads();
    red[tid] = ab;
    laneReduce<T, false>(red, tid, nt);
    // (2) the gate
    if (tid < T) {
        sFire[tid] = 0;
        sGo[tid] = 0;
        if (on[tid]) {
            const int lane = tile * T + tid;
            unstable[tid] = 0;
            failed[tid] = 0;
            const double fsActive = sSum[tid], absorbed = red[tid];
            const double gate = fmax(prm.dsTol, P_RESIDUE_EPS);
            const double m = S.slackP[lane] - D.slackTargetP;
            const double fsAll = D.factorSum, rho = 1.0 - fsAll;
            const double rhoEff = fsActive != fsAll ? (fsActive + rho > 0.0 ? rho / (fsActive + rho) : 0.0) : rho;
            const double retain = rhoEff * (m + absorbed);
            const double rem = m - retain;
            const double fsGlobal =
                fsActive == fsAll ? fsActive : (fsActive + rho > 0.0 ? fsActive / (fsActive + rho) : 0.0);
            const bool noCapacity = !(fsActive > 0.0) && rho < 1e-12;
            const double applied = S.dsRetain[lane];
            if (fsGlobal <= 0.0) {
                if (noCapacity) {
                    if (fabs(m) > prm.dsTol && fabs(m) > P_RESIDUE_EPS) failed[tid] = 1;
                } else {
                    // only the slack participates: it takes the whole mismatch, and the loop re-solves when that
                    // changed
                    const double before = isnan(applied) ? 0.0 : applied;
                    if (fabs(m - before) > gate) {
                        S.dsRetain[lane] = m;
                        unstable[tid] = 1;
                    }
                }
            } else if (fabs(rem) > fsGlobal * gate) {
                sFire[tid] = 1;
                sGo[tid] = 1;
                sPrevRetain[tid] = isnan(applied) ? 0.0 : applied;
                S.dsRetain[lane] = retain;
                sRem[tid] = fsGlobal * (m + absorbed);
            }
        }
    }
    __syncthreads();
    // (3) reset to the anchors
    if (sFire[t]) {
        for (int w = tid; w < D.n * T; w += nt) {
            const int i = w / T;
            const size_t vi = vIdx<T>(tile, nb, D.node[i], t);
            if (S.mode[vi] != MODE_EQUATION) continue;
            S.dsPrev[((size_t)tile * D.n + i) * T + t] = S.target[vi];
            S.target[vi] = D.base[i];
        }
    }
    __syncthreads();
    // (4) the water-fill, at most one round per participant
    for (int iter = 0; iter < D.n; iter++) {
        if (tid < T && sGo[tid] && !(fabs(sRem[tid]) > P_RESIDUE_EPS)) sGo[tid] = 0;
        __syncthreads();
        int any = 0;
        for (int u = 0; u < T; u++) any |= sGo[u];
        if (!any) break;
        double fs = 0.0;
        if (sGo[t]) {
            const double rem = sRem[t];
            for (int w = tid; w < D.n * T; w += nt) {
                const int i = w / T;
                const size_t vi = vIdx<T>(tile, nb, D.node[i], t);
                if (S.mode[vi] != MODE_EQUATION) continue;
                const double cum = S.target[vi] - D.base[i];
                const bool sat = rem > 0 ? cum >= D.maxD[i] : cum <= D.minD[i];
                if (!sat) fs += D.factor[i];
            }
        }
        red[tid] = fs;
        laneReduce<T, false>(red, tid, nt);
        if (tid < T) {
            sSum[tid] = red[tid];
            if (sGo[tid] && !(red[tid] > 0.0)) sGo[tid] = 0;   // every participant saturated: the rest stays
        }
        __syncthreads();
        double done = 0.0;
        if (sGo[t]) {
            const double rem = sRem[t], fsum = sSum[t];
            for (int w = tid; w < D.n * T; w += nt) {
                const int i = w / T;
                const size_t vi = vIdx<T>(tile, nb, D.node[i], t);
                if (S.mode[vi] != MODE_EQUATION) continue;
                const double cum = S.target[vi] - D.base[i];
                const bool sat = rem > 0 ? cum >= D.maxD[i] : cum <= D.minD[i];
                if (sat) continue;
                const double next = fmin(fmax(cum + rem * (D.factor[i] / fsum), D.minD[i]), D.maxD[i]);
                if (next != cum) {
                    S.target[vi] = D.base[i] + next;
                    done += next - cum;
                }
            }
        }
        red[tid] = done;
        laneReduce<T, false>(red, tid, nt);
        if (tid < T && sGo[tid]) sRem[tid] -= red[tid];
        __syncthreads();
    }
    // (5) the outcome
    double mv = 0.0;
    if (sFire[t]) {
        for (int w = tid; w < D.n * T; w += nt) {
            const int i = w / T;
            const size_t vi = vIdx<T>(tile, nb, D.node[i], t);
            if (S.mode[vi] != MODE_EQUATION) continue;
            mv += fabs(S.target[vi] - S.dsPrev[((size_t)tile * D.n + i) * T + t]);
        }
    }
    red[tid] = mv;
    laneReduce<T, false>(red, tid, nt);
    if (tid < T && sFire[tid]) {
        const int lane = tile * T + tid;
        const double moved = red[tid] + fabs(S.dsRetain[lane] - sPrevRetain[tid]);
        unstable[tid] = moved > 0.9 * P_RESIDUE_EPS;
        failed[tid] = fabs(sRem[tid]) > P_RESIDUE_EPS;
    }
    __syncthreads();
}

// Batched Newton steps for the tile of blockIdx.x: up to prm.stepsPerLaunch of them, ending as soon as no lane of the
// tile runs, so that one launch carries every tile to its end and the hardware starts further tiles as earlier ones
// finish (a tile in a long reactive-limit outer loop overlaps the others instead of holding launches of its own).
// Every lane's step count is bounded by its iteration and outer-loop budgets and the loop by prm.stepsPerLaunch, so
// every wave exits. Each step is the reference loop body: native/cuda/solver/newton_iteration.cu:116-203 (test half
// :447-495, testScenario :276-397, solveUniformBatch :580-660), reactive limits
// native/cuda/outerloop/reactive_limits.cu:321-371 plus the forced-PV election, step scaling
// native/cuda/solver/state_vector_scaling.cu:57-97. ROBUST (one lane per tile, FP64): the linear solve runs on
// statically pivoted factors corrected by GMRES (gmresRefine), the role of the reference's ROBUST single-matrix mode
// (native/cuda/common/check.cuh:38-40).
// ROWCRIT: the per-row stopping criterion (prm.rowCriteria), a variant of its own so that the default kernel carries
// no code for it.
// Workgroups of MAX_THREADS threads the step kernel's register budget leaves room for on one CU (a build-time
// tuning knob: 2 caps the kernel at 64 VGPRs, spilling the rest). HIP's second launch bound is the minimum waves per
// SIMD: blocks x MAX_THREADS / 64 waves over the CU's 4 SIMDs.
#ifndef OLFHIP_STEP_MIN_BLOCKS
#define OLFHIP_STEP_MIN_BLOCKS 1
#endif
template <int T, typename VT, bool ROBUST = false, bool ROWCRIT = false>
__global__ __launch_bounds__(MAX_THREADS, OLFHIP_STEP_MIN_BLOCKS *MAX_THREADS / 256) void nrStepKernel(DevPlan P, DevState S, DevRL R, NrParams prm, DevTop top) {
    const int tile = blockIdx.x;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int myLane = tid % T;
    if (S.tileRunning[tile] == 0) return;   // uniform per block
    const unsigned long long tileStart = (S.tileClock && tid == 0) ? wall_clock64() : 0ull;
    int stepsRun = 0;
    extern __shared__ float4 olfhipTopLds[];
    const TopLds<T, VT> TL = {reinterpret_cast<VT *>(reinterpret_cast<unsigned char *>(olfhipTopLds) + top.metaBytes),
                              top.n};

    __shared__ int sStatus[T], sIt[T], sSolves[T], sOuter[T], sTotal[T], sClamp[T];
    __shared__ int sFill[T], sSolve[T], sConv[T], sRl[T], sRemain[T], sSwitched[T], sFin[T];
    __shared__ int sChk[T], sLoop[T], sSweep[T], sResolve[T], sReject[T], sDs[T], sDsUnst[T], sDsFail[T], sBt[T];
    __shared__ double sSsq[T], sRowR[T];
    __shared__ double sNorm[T], sRatio[T], sFactor[T];
    __shared__ int sFlag[4];
    __shared__ double red[MAX_THREADS];
    __shared__ int redI[MAX_THREADS];
    __shared__ unsigned long long *sProf;

    const int nb = P.nb;
    VT *lu = reinterpret_cast<VT *>(S.lu);
    VT *y = reinterpret_cast<VT *>(S.y);
    VT *dx1 = reinterpret_cast<VT *>(S.dx1);
    const LevelMeta LM = loadLevelMeta(P, top, reinterpret_cast<unsigned char *>(olfhipTopLds), tid, nt);
    const bool dsOn = prm.distributedSlack && R.slackNode >= 0;
    const int slackNode = dsOn ? R.slackNode : -1;
    __shared__ int sRunning, sPause;
    __shared__ int sWasRunning[T];
    for (int step = 0; step < prm.stepsPerLaunch; step++) {
    if (prm.pauseAt > 0) {
        // the bank's pause (BatchSolveNative.solveBatchContext's pauseAtEnded): once enough lanes have ended in this
        // call, the tile leaves its running lanes as they are, between two steps, for the next call to go on with
        if (tid == 0) sPause = atomicAdd(S.ended, 0) >= prm.pauseAt;
        __syncthreads();
        if (sPause) break;
    }
    stepsRun++;
    if (tid == 0) {
        sProf = nullptr;
        if (S.prof && tile == S.profTile) {
            const unsigned long long k = atomicAdd(S.prof, 1ull);
            if (k < (unsigned long long)S.profCap) sProf = S.prof + PROF_SECTIONS * (k + 1);
        }
        if (sProf) sProf[0] = wall_clock64();
    }
#define OLFHIP_MARK(i)                                 \
    do {                                               \
        if (tid == 0 && sProf) sProf[i] = wall_clock64(); \
    } while (0)
    if (tid < T) {
        const int lane = tile * T + tid;
        sStatus[tid] = S.status[lane];
        sIt[tid] = S.it[lane];
        sSolves[tid] = S.solves[lane];
        sOuter[tid] = S.outer[lane];
        sTotal[tid] = S.total[lane];
        sRatio[tid] = S.ratio[lane];
        sClamp[tid] = 0;
        sFill[tid] = sStatus[tid] == LANE_RUNNING;
        sWasRunning[tid] = sFill[tid];
        sSolve[tid] = 0;
        sConv[tid] = 0;
        sRl[tid] = 0;
        sRemain[tid] = 0;
        sSwitched[tid] = 0;
        sFactor[tid] = 1.0;
    }
    __syncthreads();

    // ---- 1. mismatch + Jacobian of the running lanes
    {
        double pmax = 0.0, psq = 0.0;
        int pfin = 1;
        const LaneMap lmF = laneMapOf<T>(sFill);
        fillPass<T, VT>(P, S, tile, tid, nt, sFill, pmax, pfin, psq, slackNode);
        red[tid] = pmax;
        redI[tid] = pfin;
        laneReduceFor<T, true>(red, tid, nt, lmF);
        laneReduceAndFor<T>(redI, tid, nt, lmF);
        if (tid < T) {
            sNorm[tid] = red[tid];
            sFin[tid] = redI[tid];
        }
        if constexpr (ROWCRIT) {
            double pratio = 0.0;
            rowRatioPass<T>(P, S, tile, tid, nt, sFill, pratio);   // the thread's own F entries: no barrier needed
            __syncthreads();
            red[tid] = pratio;
            laneReduceFor<T, true>(red, tid, nt, lmF);
            if (tid < T) sRowR[tid] = red[tid];
        }
        if (prm.scaling == 1 || prm.convEps > 0.0) {
            __syncthreads();
            red[tid] = psq;
            laneReduceFor<T, false>(red, tid, nt, lmF);
            if (tid < T) sSsq[tid] = red[tid];
        }
        if (prm.scaling == 1) {
            // LINE_SEARCH (native/cuda/solver/newton_iteration.cu:497-530): after every step but a solve's first, a
            // step that did not lower ||F||_2 (or made F non-finite) is walked back to x_prev + mu dx,
            // mu = lsFold^-b, b = 1..lsMax, refilled at each point until ||F||_2 is below the last tested one
            if (tid < T) {
                const int lane = tile * T + tid;
                sBt[tid] = sFill[tid] && sIt[tid] > 0 && S.lsHas[lane] &&
                           (!sFin[tid] || sSsq[tid] >= S.lsLast[lane]);
            }
            __syncthreads();
            for (int b = 1; b <= prm.lsMax; b++) {
                int any = 0;
                for (int u = 0; u < T; u++) any |= sBt[u];
                if (!any) break;
                const double mu = 1.0 / pow(prm.lsFold, (double)b);
                for (int w = tid; w < nb * T; w += nt) {
                    const int k = w / T, t = w % T;
                    if (!sBt[t]) continue;
                    const size_t vi = vIdx<T>(tile, nb, k, t);
                    const double2 xp = *reinterpret_cast<const double2 *>(S.xPrev + vi);
                    const double2 d = *reinterpret_cast<const double2 *>(S.dx + vi);
                    *reinterpret_cast<double2 *>(S.x + vi) = make_double2(xp.x + mu * d.x, xp.y + mu * d.y);
                }
                __syncthreads();
                double bmax = 0.0, bsq = 0.0;
                int bfin = 1;
                const LaneMap lmB = laneMapOf<T>(sBt);
                fillPass<T, VT>(P, S, tile, tid, nt, sBt, bmax, bfin, bsq, slackNode);
                red[tid] = bmax;
                redI[tid] = bfin;
                laneReduceFor<T, true>(red, tid, nt, lmB);
                laneReduceAndFor<T>(redI, tid, nt, lmB);
                if (tid < T && sBt[tid]) {
                    sNorm[tid] = red[tid];
                    sFin[tid] = redI[tid];
                }
                if constexpr (ROWCRIT) {
                    double bratio = 0.0;
                    rowRatioPass<T>(P, S, tile, tid, nt, sBt, bratio);
                    __syncthreads();
                    red[tid] = bratio;
                    laneReduceFor<T, true>(red, tid, nt, lmB);
                    if (tid < T && sBt[tid]) sRowR[tid] = red[tid];
                }
                __syncthreads();
                red[tid] = bsq;
                laneReduceFor<T, false>(red, tid, nt, lmB);
                if (tid < T && sBt[tid]) {
                    sSsq[tid] = red[tid];
                    if (sFin[tid] && sSsq[tid] < S.lsLast[tile * T + tid]) sBt[tid] = 0;
                }
                __syncthreads();
            }
            if (tid < T && sFill[tid] && sIt[tid] > 0) {
                const int lane = tile * T + tid;
                S.lsLast[lane] = sSsq[tid];
                S.lsHas[lane] = 1;
            }
        }
    }
    OLFHIP_MARK(1);
    // ---- 2. the test of each running lane at its own iteration
    if (tid < T && sFill[tid]) {
        const double norm = sNorm[tid];
        const bool fin = sFin[tid] != 0;
        const int it = sIt[tid];
        sNorm[tid] = norm;
        // the criterion (native/cuda/solver/newton_iteration.cu:241-300): ||F||inf < tol, or with convEps OLF's
        // ||F||_2 < sqrt(convEps^2 nActive), or with row thresholds every |F_i| below its row's (:296-300)
        const bool below = ROWCRIT ? sRowR[tid] < 1.0
                           : prm.convEps > 0.0
                               ? sqrt(sSsq[tid]) < sqrt(prm.convEps * prm.convEps * (double)S.nActive[tile * T + tid])
                               : norm < prm.tol;
        if (prm.fixedIter) {
            // the fixed-iteration mode (native/cuda/solver/newton_iteration.cu:141-151,336-343): no test until the
            // budget, then the status of the final mismatch
            if (it >= prm.maxIter) sStatus[tid] = fin ? ST_MAX_ITER : ST_DIVERGED;
            else sSolve[tid] = 1;
        } else if (!fin) {
            sStatus[tid] = ST_DIVERGED;
        } else if (it > 0 && below && it < prm.maxIter) {
            sConv[tid] = 1;
        } else if (it > 0 && it >= prm.maxIter) {
            const bool bad = !(sRatio[tid] <= prm.linFailRatio);
            sStatus[tid] = bad ? (prm.batchOfOne ? ST_LIN_FAILED : ST_REJECTED) : ST_MAX_ITER;
        } else {
            sSolve[tid] = 1;
        }
    }
    if (tid == 0) {
        int any = 0;
        for (int t = 0; t < T; t++) any |= sConv[t];
        sFlag[0] = any;
    }
    __syncthreads();

    // ---- 3. outer loops of the lanes whose solve converged: distributed slack, then reactive limits, each driven
    // to STABLE with a re-solve after every unstable check, the sweep restarted from the first loop while it changed
    // something (OLF's AcloadFlowEngine.runOuterLoops as the reference's driver runs it,
    // native/cuda/solver/outer_loop_driver.cu:300-363), at most maxOuter re-solves. Without a re-solve a lane's
    // checks are at most: distributed slack, reactive limits, and once more both after a sweep that changed
    // something, so two rounds decide every lane.
    if (sFlag[0]) {
        if (tid < T) {
            const int lane = tile * T + tid;
            sChk[tid] = sConv[tid];
            sLoop[tid] = S.loopAt[lane];
            sSweep[tid] = S.sweepChanged[lane];
            sResolve[tid] = 0;
            sReject[tid] = 0;
            sFill[tid] = 0;
        }
        __syncthreads();
        for (int round = 0; round < 2; round++) {
        if (tid < T) {
            if (sChk[tid] && sOuter[tid] >= prm.maxOuter) sChk[tid] = 0;   // budget spent: the lane ends here
            sDs[tid] = sChk[tid] && sLoop[tid] == 0 && dsOn;
        }
        __syncthreads();
        {
            int anyDs = 0;
            for (int u = 0; u < T; u++) anyDs |= sDs[u];
            if (anyDs) dsCheck<T>(P, S, R, prm, tile, tid, nt, sDs, sDsUnst, sDsFail, red);
        }
        if (tid < T && sChk[tid] && sLoop[tid] == 0) {
            if (sDs[tid]) {
                S.dsFailed[tile * T + tid] = (signed char)sDsFail[tid];
                if (sDsFail[tid] && prm.dsReject) {
                    sReject[tid] = 1;
                    sChk[tid] = 0;
                } else if (sDsUnst[tid]) {
                    sResolve[tid] = 1;
                    sChk[tid] = 0;
                } else {
                    sLoop[tid] = 1;
                }
            } else {
                sLoop[tid] = 1;
            }
        }
        if (tid < T) {
            sRl[tid] = sChk[tid] && sLoop[tid] == 1 && prm.reactiveLimits && R.n > 0;
            sRemain[tid] = 0;
            sSwitched[tid] = 0;
        }
        __syncthreads();
        for (int w = tid; w < R.n * T; w += nt) {
            const int e = w / T, t = w % T;
            if (!sRl[t]) continue;
            const size_t ei = ((size_t)tile * R.n + e) * T + t;
            signed char lt = S.rlType[ei];
            signed char pass = 0;
            if (lt != QLIM_FROZEN) {
                const int k = R.node[e];
                const size_t vi = vIdx<T>(tile, nb, k, t) + 1;
                if (S.mode[vi] == MODE_FIXED && lt == QLIM_NONE) {
                    const double q = S.qinj[qIdx<T>(tile, nb, k, t)] + R.loadQ[e];
                    if (q < R.minQ[e] - prm.rlTol) {
                        S.mode[vi] = MODE_EQUATION;
                        S.target[vi] = R.minQ[e] - R.loadQ[e];
                        S.rlType[ei] = QLIM_MIN;
                        S.rlCount[ei]++;
                        pass = 1;
                    } else if (q > R.maxQ[e] + prm.rlTol) {
                        S.mode[vi] = MODE_EQUATION;
                        S.target[vi] = R.maxQ[e] - R.loadQ[e];
                        S.rlType[ei] = QLIM_MAX;
                        S.rlCount[ei]++;
                        pass = 1;
                    } else {
                        atomicAdd(&sRemain[t], 1);
                    }
                } else if (lt != QLIM_NONE) {
                    const double v = S.x[vi], tv = R.targetV[e];
                    const bool wants = (lt == QLIM_MAX && v > tv) || (lt == QLIM_MIN && v < tv);
                    if (wants && S.rlCount[ei] < prm.maxSwitch) {
                        S.mode[vi] = MODE_FIXED;
                        S.target[vi] = tv;
                        S.rlType[ei] = QLIM_NONE;
                        pass = 2;
                    }
                }
            }
            S.rlPass[ei] = pass;
            if (pass) atomicOr(&sSwitched[t], 1);
        }
        __syncthreads();
        if (tid < T && sRl[tid] && sRemain[tid] == 0) {
            // forced PV: no bus that was PV at this check stays PV -> the strongest bus switched PV -> PQ in this
            // check stays PV (equal nominal voltages: highest P target, then lowest entry)
            int best = -1;
            double bp = -1e300;
            for (int e = 0; e < R.n; e++) {
                const size_t ei = ((size_t)tile * R.n + e) * T + tid;
                if (S.rlPass[ei] != 1) continue;
                const double p = S.target[vIdx<T>(tile, nb, R.node[e], tid)];
                if (p > bp) {
                    bp = p;
                    best = e;
                }
            }
            if (best >= 0) {
                const size_t ei = ((size_t)tile * R.n + best) * T + tid;
                const size_t vi = vIdx<T>(tile, nb, R.node[best], tid) + 1;
                S.mode[vi] = MODE_FIXED;
                S.target[vi] = R.targetV[best];
                S.rlType[ei] = QLIM_NONE;
                S.rlCount[ei]--;
                S.rlPass[ei] = 0;
                int any = 0;
                for (int e = 0; e < R.n; e++) any |= S.rlPass[((size_t)tile * R.n + e) * T + tid] != 0;
                sSwitched[tid] = any;
            }
        }
        __syncthreads();
        if (tid < T && sChk[tid]) {
            if (sLoop[tid] == 1) {
                if (sRl[tid] && sSwitched[tid]) {
                    sResolve[tid] = 1;
                    sChk[tid] = 0;
                } else {
                    sLoop[tid] = 2;
                }
            }
            if (sChk[tid] && sLoop[tid] == 2) {   // end of the sweep
                if (sSweep[tid] && dsOn) {
                    sSweep[tid] = 0;
                    sLoop[tid] = 0;
                } else {
                    sChk[tid] = 0;
                }
            }
        }
        __syncthreads();
        }
        if (tid < T && sConv[tid]) {
            const int lane = tile * T + tid;
            sSolves[tid]++;
            if (sResolve[tid]) {
                sOuter[tid]++;
                sIt[tid] = 0;
                sFill[tid] = 1;     // restarted at its new modes and targets: refill, no test at iteration 0
                sSweep[tid] = 1;
                S.lsHas[lane] = 0;  // a restart retargets the system: no line-search reference yet
            } else if (sReject[tid]) {
                sStatus[tid] = ST_REJECTED;
            } else {
                sStatus[tid] = ST_CONVERGED;
            }
            S.loopAt[lane] = (signed char)sLoop[tid];
            S.sweepChanged[lane] = (signed char)sSweep[tid];
        }
        if (tid == 0) {
            int any = 0;
            for (int t = 0; t < T; t++) any |= sFill[t];
            sFlag[1] = any;
        }
        __syncthreads();
        if (sFlag[1]) {
            double pmax = 0.0, psq = 0.0;
            int pfin = 1;
            const LaneMap lmF = laneMapOf<T>(sFill);
            fillPass<T, VT>(P, S, tile, tid, nt, sFill, pmax, pfin, psq, slackNode);
            red[tid] = pmax;
            redI[tid] = pfin;
            laneReduceFor<T, true>(red, tid, nt, lmF);
            laneReduceAndFor<T>(redI, tid, nt, lmF);
            if (tid < T && sFill[tid]) {
                sNorm[tid] = red[tid];
                if (redI[tid] == 0) sStatus[tid] = ST_DIVERGED;
                else sSolve[tid] = 1;
            }
        }
    }
    if (tid == 0) {
        int any = 0;
        for (int t = 0; t < T; t++) any |= sSolve[t];
        sFlag[2] = any;
    }
    __syncthreads();

    if (sFlag[2]) {
        // ---- 4. factorization with the forward solve of -F, then the backward solve
        OLFHIP_MARK(2);
        luFactorForward<T, VT>(P, LM, top, TL, tile, tid, nt, lu, y, S.F, -1.0, sClamp, sSolve, sProf, 9);
        OLFHIP_MARK(3);
        luBackward<T, VT>(P, LM, top, TL, tile, tid, nt, lu, y, dx1, sSolve, S.dx, false, sProf, 10);   // FP64 step
        OLFHIP_MARK(4);
        // a lane whose factorization met a singular pivot block (clamped) was not factored: the unpivoted batch
        // factorization leaves it, REJECTED in a batch and LINEAR_SOLVE_FAILED alone (the reference routes such
        // lanes out of its fast engine, native/cuda/lu/lu_pivot.cuh:12-21), for a robust solve to take up
        if (tid < T && sSolve[tid] && sClamp[tid] > 0) {
            sStatus[tid] = prm.batchOfOne ? ST_LIN_FAILED : ST_REJECTED;
            sSolve[tid] = 0;
        }
        __syncthreads();
        // ---- 5. the step in FP64, S.dx (+ one FP64-residual refinement step on the FP32 factors, or the robust
        //         solve's GMRES)
        if (prm.refine) {
            double pm = 0.0;
            // FP64 residual r = J dx + F, solved for the correction with sign -1: J dx2 = -r
            applyJPass<T, VT>(P, S, tile, tid, nt, sSolve, S.dx, S.F, 1.0, S.r, pm);
            __syncthreads();
            OLFHIP_MARK(5);
            luForward<T, VT>(P, LM, top, TL, tile, tid, nt, lu, y, S.r, -1.0, sSolve, sProf, 11);
            OLFHIP_MARK(12);
            luBackward<T, VT>(P, LM, top, TL, tile, tid, nt, lu, y, dx1, sSolve, S.dx, true, sProf, 13);   // dx +=
            OLFHIP_MARK(6);
        }
        if constexpr (ROBUST) {
            static_assert(T == 1 && sizeof(VT) == 8, "the robust solve is FP64, one lane per tile");
            gmresRefine<T>(P, LM, top, TL, S, prm, tile, tid, nt, lu, y, dx1, sSolve, red);
        }
        // ---- 6. linear-solve guard at the 2nd iteration of the first solve, and the residual of the solve before
        //         the budget (reference newton_iteration.cu:596-652, :369-389)
        if (tid == 0) {
            int due = 0;
            for (int t = 0; t < T; t++)
                due |= sSolve[t] && ((sSolves[t] == 0 && sIt[t] == 1) || sIt[t] == prm.maxIter - 1);
            sFlag[3] = due && !prm.fixedIter;
        }
        __syncthreads();
        if (sFlag[3]) {
            double pm = 0.0;
            const LaneMap lmS = laneMapOf<T>(sSolve);
            applyJPass<T, VT>(P, S, tile, tid, nt, sSolve, S.dx, S.F, 1.0, nullptr, pm);
            red[tid] = pm;
            laneReduceFor<T, true>(red, tid, nt, lmS);
            if (tid < T && sSolve[tid]) {
                const double ratio = sNorm[tid] > 0.0 ? red[tid] / sNorm[tid] : 0.0;
                if (sSolves[tid] == 0 && sIt[tid] == 1 && !(ratio <= prm.linFailRatio)) {
                    sStatus[tid] = ST_LIN_FAILED;
                    sSolve[tid] = 0;
                }
                sRatio[tid] = ratio;
            }
            __syncthreads();
        }
        OLFHIP_MARK(7);
        // ---- 7. step scaling (MAX_VOLTAGE_CHANGE) and update
        if (prm.scaling == 2) {
            double mdv = 0.0, mdp = 0.0;
            for (int w = tid; w < nb * T; w += nt) {
                const int k = w / T, t = w % T;
                const double2 d = *reinterpret_cast<const double2 *>(S.dx + vIdx<T>(tile, nb, k, t));
                mdp = fmax(mdp, fabs(d.x));
                mdv = fmax(mdv, fabs(d.y));
            }
            red[tid] = mdv;
            laneReduce<T, true>(red, tid, nt);
            if (tid < T) sFactor[tid] = red[tid] > prm.maxDv ? prm.maxDv / red[tid] : 1.0;
            __syncthreads();
            red[tid] = mdp;
            laneReduce<T, true>(red, tid, nt);
            if (tid < T && red[tid] > prm.maxDphi) sFactor[tid] = fmin(sFactor[tid], prm.maxDphi / red[tid]);
            __syncthreads();
        }
        const LaneMap lmU = laneMapOf<T>(sSolve);
        for (int w = tid; w < nb * lmU.tm; w += nt) {
            const int k = lmU.tm == 1 ? w : w / T, t = lmU.tm == 1 ? lmU.t0 : w % T;
            if (!sSolve[t]) continue;
            const double f = sFactor[t];
            const size_t vi = vIdx<T>(tile, nb, k, t);
            const double2 d = *reinterpret_cast<const double2 *>(S.dx + vi);
            double2 *xp = reinterpret_cast<double2 *>(S.x + vi);
            const double2 xv = *xp;
            if (prm.scaling == 1) *reinterpret_cast<double2 *>(S.xPrev + vi) = xv;
            *xp = make_double2(xv.x + f * d.x, xv.y + f * d.y);
        }
        if (tid < T && sSolve[tid]) {
            sIt[tid]++;
            sTotal[tid]++;
        }
    }
    __syncthreads();
    OLFHIP_MARK(8);
#undef OLFHIP_MARK
    (void)myLane;
    if (tid < T) {
        const int lane = tile * T + tid;
        S.status[lane] = sStatus[tid];
        S.it[lane] = sIt[tid];
        S.solves[lane] = sSolves[tid];
        S.outer[lane] = sOuter[tid];
        S.total[lane] = sTotal[tid];
        S.ratio[lane] = sRatio[tid];
        if (sFill[tid] || sConv[tid]) S.norm[lane] = sNorm[tid];
        S.clamps[lane] += sClamp[tid];
    }
    if (tid == 0) {
        int any = 0, newlyEnded = 0;
        for (int t = 0; t < T; t++) {
            any |= sStatus[t] == LANE_RUNNING;
            newlyEnded += sWasRunning[t] && sStatus[t] != LANE_RUNNING;
        }
        S.tileRunning[tile] = any;
        sRunning = any;
        if (newlyEnded) atomicAdd(S.ended, newlyEnded);
    }
    __syncthreads();
    if (!sRunning) break;
    }
    if (S.tileClock && tid == 0) {
        unsigned long long *tc = S.tileClock + 3 * (size_t)tile;
        if (tc[0] == 0) tc[0] = tileStart;
        tc[1] = wall_clock64();
        tc[2] += (unsigned long long)stepsRun;
    }
}

// ---- install: every slot starts from the base (or its own x0), base targets and modes.
template <int T>
__global__ void installKernel(DevPlan P, DevState S, const double *x0Base, const double *x0Per, const double *targetBase,
                              const int *modeBase) {
    const int nb = P.nb;
    const size_t total = (size_t)S.nTiles * nb * T;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(w % T);
        const size_t kt = w / T;
        const int k = (int)(kt % nb);
        const int tile = (int)(kt / nb);
        const int lane = tile * T + t;
        const int b = P.perm[k];
        const size_t vi = vIdx<T>(tile, nb, k, t);
        const size_t n = 2 * (size_t)nb;
        const bool installed = lane < S.nInstalled;
        double2 xv;
        if (installed && x0Per) xv = make_double2(x0Per[lane * n + 2 * b], x0Per[lane * n + 2 * b + 1]);
        else xv = make_double2(x0Base[2 * b], x0Base[2 * b + 1]);
        *reinterpret_cast<double2 *>(S.x + vi) = xv;
        *reinterpret_cast<double2 *>(S.target + vi) = make_double2(targetBase[2 * b], targetBase[2 * b + 1]);
        S.mode[vi] = (signed char)modeBase[2 * b];
        S.mode[vi + 1] = (signed char)modeBase[2 * b + 1];
    }
}

// per-scenario (scenario, row, mode) entries, target overrides (scenario * n + row -> value) and start-state seeds
// (scenario * n + row -> value, over the tiled or per-scenario start)
template <int T>
__global__ void overrideKernel(DevPlan P, DevState S, int nModes, const int *modeEntries, int nTargets,
                               const long long *targetIndex, const double *targetValue, int nSeeds,
                               const long long *seedIndex, const double *seedValue, const int *slotOf) {
    const int nb = P.nb;
    const size_t n = 2 * (size_t)nb;
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < nModes + nTargets + nSeeds;
         w += gridDim.x * blockDim.x) {
        if (w < nModes) {
            const int lane = slotOf ? slotOf[modeEntries[3 * w]] : modeEntries[3 * w], row = modeEntries[3 * w + 1];
            const int tile = lane / T, t = lane % T;
            S.mode[vIdx<T>(tile, nb, P.pos[row >> 1], t) + (row & 1)] = (signed char)modeEntries[3 * w + 2];
        } else {
            const bool target = w < nModes + nTargets;
            const int q = target ? w - nModes : w - nModes - nTargets;
            const long long g = target ? targetIndex[q] : seedIndex[q];
            const int local = (int)(g / (long long)n), row = (int)(g % (long long)n);
            const int lane = slotOf ? slotOf[local] : local;
            const int tile = lane / T, t = lane % T;
            double *dst = target ? S.target : S.x;
            dst[vIdx<T>(tile, nb, P.pos[row >> 1], t) + (row & 1)] = target ? targetValue[q] : seedValue[q];
        }
    }
}

// lane counters and reactive-limit entry state; an entry whose bus's P row a scenario froze (an islanded bus) leaves
// that scenario's outer loop (reference reactive_limits.cu:62-69)
template <int T>
__global__ void laneInitKernel(DevPlan P, DevState S, DevRL R, const signed char *seed, int seedStride,
                               const int *modeBase) {
    const int nb = P.nb;
    const int lanes = S.nTiles * T;
    const size_t totalE = (size_t)lanes * R.n;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < (size_t)lanes + totalE;
         w += (size_t)gridDim.x * blockDim.x) {
        if (w < (size_t)lanes) {
            const int lane = (int)w;
            S.status[lane] = lane < S.nInstalled ? LANE_RUNNING : LANE_PADDING;
            S.it[lane] = 0;
            S.solves[lane] = 0;
            S.outer[lane] = 0;
            S.total[lane] = 0;
            S.clamps[lane] = 0;
            S.norm[lane] = 0.0;
            S.ratio[lane] = 0.0;
            S.loopAt[lane] = 0;
            S.sweepChanged[lane] = 0;
            S.dsFailed[lane] = 0;
            S.dsRetain[lane] = __builtin_nan("");
            S.lsHas[lane] = 0;
            S.slackP[lane] = 0.0;
            if (lane % T == 0) S.tileRunning[lane / T] = lane < S.nInstalled;
        } else {
            const size_t q = w - lanes;
            const int t = (int)(q % T);
            const size_t et = q / T;
            const int e = (int)(et % R.n);
            const int tile = (int)(et / R.n);
            const int lane = tile * T + t;
            const size_t ei = ((size_t)tile * R.n + e) * T + t;
            const int k = R.node[e];
            const int pRow = 2 * P.perm[k];
            signed char lt = (seed && lane < S.nInstalled) ? seed[(size_t)lane * seedStride + e] : QLIM_NONE;
            const signed char pm = S.mode[vIdx<T>(tile, nb, k, t)];
            if (pm == MODE_FIXED && modeBase[pRow] != MODE_FIXED) lt = QLIM_FROZEN;
            S.rlType[ei] = lt;
            S.rlCount[ei] = 0;
            S.rlPass[ei] = 0;
        }
    }
}

// ---- the bank (BatchSolveNative.beginBatchInstall, BatchSolveNative.java:53-59): a solve installs its scenarios
// into the slots slotOf[j] only; every other slot keeps what it holds (a scenario still solving goes on).
template <int T>
__global__ void bankInstallKernel(DevPlan P, DevState S, int nLocal, const int *slotOf, const double *x0Base,
                                  const double *x0Per, const double *targetBase, const int *modeBase) {
    const int nb = P.nb;
    const size_t n = 2 * (size_t)nb;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < (size_t)nLocal * nb;
         w += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(w / nb), k = (int)(w % nb);
        const int lane = slotOf[j];
        const int b = P.perm[k];
        const size_t vi = vIdx<T>(lane / T, nb, k, lane % T);
        const double2 xv = x0Per ? make_double2(x0Per[j * n + 2 * b], x0Per[j * n + 2 * b + 1])
                                 : make_double2(x0Base[2 * b], x0Base[2 * b + 1]);
        *reinterpret_cast<double2 *>(S.x + vi) = xv;
        *reinterpret_cast<double2 *>(S.target + vi) = make_double2(targetBase[2 * b], targetBase[2 * b + 1]);
        S.mode[vi] = (signed char)modeBase[2 * b];
        S.mode[vi + 1] = (signed char)modeBase[2 * b + 1];
    }
}

// the installed slots' counters and reactive-limit entries (laneInitKernel's work for slotOf[j], seed row j)
template <int T>
__global__ void bankLaneInitKernel(DevPlan P, DevState S, DevRL R, int nLocal, const int *slotOf,
                                   const signed char *seed, int seedStride, const int *modeBase) {
    const int nb = P.nb;
    const size_t totalE = (size_t)nLocal * R.n;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < (size_t)nLocal + totalE;
         w += (size_t)gridDim.x * blockDim.x) {
        if (w < (size_t)nLocal) {
            const int lane = slotOf[w];
            S.status[lane] = LANE_RUNNING;
            S.it[lane] = 0;
            S.solves[lane] = 0;
            S.outer[lane] = 0;
            S.total[lane] = 0;
            S.clamps[lane] = 0;
            S.norm[lane] = 0.0;
            S.ratio[lane] = 0.0;
            S.loopAt[lane] = 0;
            S.sweepChanged[lane] = 0;
            S.dsFailed[lane] = 0;
            S.dsRetain[lane] = __builtin_nan("");
            S.lsHas[lane] = 0;
            S.slackP[lane] = 0.0;
        } else {
            const size_t q = w - nLocal;
            const int j = (int)(q / R.n), e = (int)(q % R.n);
            const int lane = slotOf[j];
            const int tile = lane / T, t = lane % T;
            const size_t ei = ((size_t)tile * R.n + e) * T + t;
            const int k = R.node[e];
            const int pRow = 2 * P.perm[k];
            signed char lt = seed ? seed[(size_t)j * seedStride + e] : QLIM_NONE;
            if (S.mode[vIdx<T>(tile, nb, k, t)] == MODE_FIXED && modeBase[pRow] != MODE_FIXED) lt = QLIM_FROZEN;
            S.rlType[ei] = lt;
            S.rlCount[ei] = 0;
            S.rlPass[ei] = 0;
        }
    }
}

// every tile with a running lane runs (after a bank install, or to go on after a pause)
template <int T>
__global__ void tileRunningKernel(DevState S) {
    for (int tile = blockIdx.x * blockDim.x + threadIdx.x; tile < S.nTiles; tile += gridDim.x * blockDim.x) {
        int any = 0;
        for (int t = 0; t < T; t++) any |= S.status[tile * T + t] == LANE_RUNNING;
        S.tileRunning[tile] = any;
    }
}

// a new context's slots: none solving
__global__ void slotsEmptyKernel(int *status, int lanes, int *tileRunning, int nTiles) {
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < lanes; w += gridDim.x * blockDim.x) {
        status[w] = LANE_PADDING;
        if (w < nTiles) tileRunning[w] = 0;
    }
}

// states back in stable-row order, [lane][n]
template <int T>
__global__ void gatherStatesKernel(DevPlan P, DevState S, int nLanes, double *out) {
    const int nb = P.nb;
    const size_t n = 2 * (size_t)nb;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < (size_t)nLanes * nb;
         w += (size_t)gridDim.x * blockDim.x) {
        const int lane = (int)(w / nb), b = (int)(w % nb);
        const int tile = lane / T, t = lane % T;
        const double2 v = *reinterpret_cast<const double2 *>(S.x + vIdx<T>(tile, nb, P.pos[b], t));
        out[lane * n + 2 * b] = v.x;
        out[lane * n + 2 * b + 1] = v.y;
    }
}

// row modes, targets (stable-row order, [lane][n]) and reactive-limit types ([lane][entries]) of the installed lanes
template <int T>
__global__ void gatherOuterStateKernel(DevPlan P, DevState S, DevRL R, int nLanes, int *modeOut, double *targetOut,
                                       signed char *typeOut) {
    const int nb = P.nb;
    const size_t n = 2 * (size_t)nb;
    const size_t rows = (size_t)nLanes * nb;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < rows + (size_t)nLanes * R.n;
         w += (size_t)gridDim.x * blockDim.x) {
        if (w < rows) {
            const int lane = (int)(w / nb), b = (int)(w % nb);
            const int tile = lane / T, t = lane % T;
            const size_t vi = vIdx<T>(tile, nb, P.pos[b], t);
            modeOut[lane * n + 2 * b] = S.mode[vi];
            modeOut[lane * n + 2 * b + 1] = S.mode[vi + 1];
            targetOut[lane * n + 2 * b] = S.target[vi];
            targetOut[lane * n + 2 * b + 1] = S.target[vi + 1];
        } else {
            const size_t q = w - rows;
            const int lane = (int)(q / R.n), e = (int)(q % R.n);
            const int tile = lane / T, t = lane % T;
            const signed char lt = S.rlType[((size_t)tile * R.n + e) * T + t];
            typeOut[q] = lt == QLIM_FROZEN ? QLIM_NONE : lt;
        }
    }
}

// running lanes after a step: sum over tiles into *count (one block)
// count[0]: tiles still running; count[1]: lanes that ended during the call (the pause test)
__global__ void countRunningKernel(const int *tileRunning, int nTiles, const int *ended, int *count) {
    __shared__ int acc[256];
    int s = 0;
    for (int i = threadIdx.x; i < nTiles; i += blockDim.x) s += tileRunning[i] != 0;
    acc[threadIdx.x] = s;
    __syncthreads();
    for (int stride = blockDim.x / 2; stride > 0; stride >>= 1) {
        if ((int)threadIdx.x < stride) acc[threadIdx.x] += acc[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        count[0] = acc[0];
        count[1] = *ended;
    }
}

// standalone batched LU: factor + forward + backward per tile, rhs/solution in the node-ordered vector layout
template <int T, typename VT>
__global__ __launch_bounds__(MAX_THREADS) void luSolveKernel(DevPlan P, VT *lu, VT *y, VT *xo, const double *rhs,
                                                             int *clamps, DevTop top) {
    __shared__ int sClamp[T];
    extern __shared__ float4 olfhipTopLds[];
    const TopLds<T, VT> TL = {reinterpret_cast<VT *>(reinterpret_cast<unsigned char *>(olfhipTopLds) + top.metaBytes),
                              top.n};
    const int tile = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    if (tid < T) sClamp[tid] = 0;
    __syncthreads();
    const LevelMeta LM = loadLevelMeta(P, top, reinterpret_cast<unsigned char *>(olfhipTopLds), tid, nt);
    luFactorForward<T, VT>(P, LM, top, TL, tile, tid, nt, lu, y, rhs, 1.0, sClamp, nullptr);
    luBackward<T, VT>(P, LM, top, TL, tile, tid, nt, lu, y, xo, nullptr);
    if (tid < T) clamps[tile * T + tid] = sClamp[tid];
}

}  // namespace olfhip
