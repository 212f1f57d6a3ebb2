// Learned warm start: a 3-layer MLP trained and served by hand-written gfx950 kernels (BF16 operands on the matrix
// cores, FP32 accumulation, FP32 master weights with Adam). The prediction seeds the Newton solve through the
// per-scenario x0 hook (BatchSolveNative x0PerScenario, BatchSolveNative.java:119); the Newton mismatch alone
// decides the answer.
//
// The network works on the outage's neighbourhood: the K buses electrically nearest the outaged branch (multi-source
// Dijkstra over |x| from both endpoints, endpoints first, -1 padded), computed once per branch at creation. Its N = 2K
// rows are [angle of slot 0..K-1 ; voltage of slot 0..K-1]:
//   input  x_loc = [phi(hood) - phi(endpoint 1) ; V(hood) - 1] of the base state, normalised per row
//   z1 = Win [x_loc ; f] + Emb[branch] + b1,  h1 = gelu(z1)      Win: H x (N + F), Emb: E x H
//   z2 = W2 h1 + b2,                          h2 = gelu(z2)      W2:  H x H
//   y  = W3 h2 + b3                                              W3:  N x H
// with F endpoint features, E branches, H hidden units; loss = mean over batch and rows of (y - target)^2 on
// normalised targets. The warm start keeps the base state outside the neighbourhood and adds
// (y * y_std + y_mean) * scale_b to the neighbourhood's rows, scale_b a per-scenario factor given by the caller.
// A Newton step's remaining error depends on the start's branch-wise differences, so a local prediction with no
// step at its border (angles taken relative to the neighbourhood's outer ring) is what pays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/olfhip.h"
#include "errors.hpp"

namespace olfhip {

std::string &lastErrorRef();

#define MLP_OK(expr)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) {                                                                           \
            if (e_ == hipErrorOutOfMemory) throw MemoryError(std::string(#expr) + ": out of device memory"); \
            throw DeviceError(std::string(#expr) + ": " + hipGetErrorString(e_));                         \
        }                                                                                                 \
    } while (0)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short toBf16(float f) {
    // round to nearest even
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);   // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float fromBf16(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

// ---- GEMM: C[M x N] (+)= A[M x K] B[K x N], BF16 operands with arbitrary strides, FP32 accumulation on
// v_mfma_f32_16x16x16bf16_1k. A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn], C row-major [M][ldc].
// A 256-thread workgroup computes a 64x64 tile (each wave 32x32 = 2x2 MFMA tiles) over K range
// [kz*kChunk, (kz+1)*kChunk), 64 of K per round; with several K chunks the partial tiles are added atomically (C
// zeroed first). Both operands are staged k-contiguous in LDS ([m][k] and [n][k]) so that an MFMA operand (4
// consecutive k of one row or column) is one 8-byte LDS read; a thread brings 8 consecutive elements along the
// operand's contiguous dimension with one 16-byte global load where the run is aligned and inside the matrix, and the
// next round's loads are in flight while the current round's MFMAs run (two LDS buffers, one barrier per round).
constexpr int GT = 64, GK = 64, GKP = GK + 8;   // GKP: padded LDS row (144 bytes, 16-byte aligned)

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// One operand's 64 x 64 slice, loaded as 2 runs of 8 per thread. The slice is X(r, k) for r in [r0, r0 + 64),
// k in [k0, k0 + 64) with X(r, k) = X[r * sr + k * sk]; kc: the run direction is k (sk == 1) or r (sr == 1).
struct Runs {
    u16x8 v[2];
};
__device__ __forceinline__ void loadRuns(Runs &o, const unsigned short *__restrict__ X, long sr, long sk, int r0,
                                         int rEnd, int k0, int kEnd, int tid) {
    const bool kRun = sk == 1;   // else sr == 1, or neither (then element by element)
    for (int h = 0; h < 2; h++) {
        const int e = tid + 256 * h;   // 0..511
        const int a = e >> 3, b = (e & 7) * 8;   // run a (of 64), offset b within it
        const int r = kRun ? r0 + a : r0 + b, k = kRun ? k0 + b : k0 + a;
        const long base = (long)r * sr + (long)k * sk;
        const bool whole = kRun ? (r < rEnd && k + 7 < kEnd) : (sr == 1 && k < kEnd && r + 7 < rEnd);
        if (whole && ((reinterpret_cast<uintptr_t>(X + base) & 15) == 0)) {
            o.v[h] = *reinterpret_cast<const u16x8 *>(X + base);
        } else {
            for (int q = 0; q < 8; q++) {
                const int rr = kRun ? r : r + q, kk = kRun ? k + q : k;
                o.v[h][q] = (rr < rEnd && kk < kEnd) ? X[(long)rr * sr + (long)kk * sk] : (unsigned short)0;
            }
        }
    }
}
// the runs into the LDS slice S[r][k]
__device__ __forceinline__ void storeRuns(unsigned short (*S)[GKP], const Runs &o, bool kRun, int tid) {
    for (int h = 0; h < 2; h++) {
        const int e = tid + 256 * h;
        const int a = e >> 3, b = (e & 7) * 8;
        if (kRun) {
            *reinterpret_cast<u16x8 *>(&S[a][b]) = o.v[h];
        } else {
            for (int q = 0; q < 8; q++) S[b + q][a] = o.v[h][q];
        }
    }
}

__global__ __launch_bounds__(256) void gemmBf16(int M, int N, int K, const unsigned short *__restrict__ A, long sam,
                                                long sak, const unsigned short *__restrict__ B, long sbk, long sbn,
                                                float *__restrict__ C, long ldc, int kChunk, int atomic) {
    __shared__ unsigned short As[2][GT][GKP];   // [m][k]
    __shared__ unsigned short Bs[2][GT][GKP];   // [n][k]
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int kBeg = blockIdx.z * kChunk;
    const int kEnd = min(K, kBeg + kChunk);
    const bool aK = sak == 1, bK = sbk == 1;
    floatx4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
    const int r = lane % 16, g = lane / 16;
    Runs ra, rb;
    // B as the [n][k] operand: X(n, k) = B[n * sbn + k * sbk]
    loadRuns(ra, A, sam, sak, m0, M, kBeg, kEnd, tid);
    loadRuns(rb, B, sbn, sbk, n0, N, kBeg, kEnd, tid);
    storeRuns(As[0], ra, aK, tid);
    storeRuns(Bs[0], rb, bK, tid);
    __syncthreads();
    int buf = 0;
    for (int k0 = kBeg; k0 < kEnd; k0 += GK) {
        const bool more = k0 + GK < kEnd;
        if (more) {
            loadRuns(ra, A, sam, sak, m0, M, k0 + GK, kEnd, tid);
            loadRuns(rb, B, sbn, sbk, n0, N, k0 + GK, kEnd, tid);
        }
#pragma unroll
        for (int kk = 0; kk < GK; kk += 16) {
            bf16x4 a[2], b[2];
            for (int i = 0; i < 2; i++)
                a[i] = *reinterpret_cast<const bf16x4 *>(&As[buf][wm + i * 16 + r][kk + g * 4]);
            for (int j = 0; j < 2; j++)
                b[j] = *reinterpret_cast<const bf16x4 *>(&Bs[buf][wn + j * 16 + r][kk + g * 4]);
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            // the other buffer was last read before the previous round's barrier
            storeRuns(As[buf ^ 1], ra, aK, tid);
            storeRuns(Bs[buf ^ 1], rb, bK, tid);
        }
        __syncthreads();
        buf ^= 1;
    }
    // lane holds D[4g + q][r] of each 16x16 tile
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int q = 0; q < 4; q++) {
                const int gm = m0 + wm + i * 16 + g * 4 + q, gn = n0 + wn + j * 16 + r;
                if (gm < M && gn < N) {
                    if (atomic) atomicAdd(&C[(long)gm * ldc + gn], acc[i][j][q]);
                    else C[(long)gm * ldc + gn] = acc[i][j][q];
                }
            }
}

__global__ void zeroF(float *p, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.f;
}
__global__ void f32ToBf16(const float *in, unsigned short *out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = toBf16(in[i]);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float geluGrad(float x) {
    const float cdf = 0.5f * (1.f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// z = c + bias (+ emb[branch] row); h = gelu(z) as BF16 (the next GEMM's operand); z kept for the backward pass
__global__ void biasGeluFwd(int B, int H, const float *c, const float *bias, const float *emb, const int *branch,
                            float *z, unsigned short *h) {
    const long n = (long)B * H;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int b = (int)(i / H), u = (int)(i % H);
        float v = c[i] + bias[u];
        if (emb) v += emb[(long)branch[b] * H + u];
        z[i] = v;
        h[i] = toBf16(gelu(v));
    }
}

// y = c + bias
__global__ void biasAdd(int B, int N, float *c, const float *bias) {
    const long n = (long)B * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        c[i] += bias[i % N];
}

// dy = 2 (y - t) / (B N) as FP32 and BF16; loss sum into *loss
__global__ void mseGrad(int B, int N, const float *y, const float *t, float *dy, unsigned short *dyh, double *loss) {
    const long n = (long)B * N;
    double part = 0.0;
    const float scale = 2.f / (float)n;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float d = y[i] - t[i];
        part += (double)d * d;
        dy[i] = scale * d;
        dyh[i] = toBf16(scale * d);
    }
    __shared__ double red[256];
    red[threadIdx.x] = part;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// dz = dh * gelu'(z), as FP32 and BF16
__global__ void geluBwd(long n, const float *dh, const float *z, float *dz, unsigned short *dzh) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float v = dh[i] * geluGrad(z[i]);
        dz[i] = v;
        dzh[i] = toBf16(v);
    }
}

// bias gradient: column sums of d [B x H] in two passes with a fixed summation order (no atomics): colSumPart
// writes the sums of chunks of COLSUM_ROWS rows, [chunks][H], each by a 256-thread block of 64 columns x 4 row
// groups; colSumFinal adds the chunks in order.
constexpr int COLSUM_ROWS = 64;
__global__ __launch_bounds__(256) void colSumPart(int B, int H, const float *__restrict__ d, float *__restrict__ part) {
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), rg = threadIdx.x >> 6;
    const int r0 = blockIdx.y * COLSUM_ROWS, r1 = min(B, r0 + COLSUM_ROWS);
    float acc = 0.f;
    if (c < H)
        for (int r = r0 + rg; r < r1; r += 4) acc += d[(long)r * H + c];
    red[rg][threadIdx.x & 63] = acc;
    __syncthreads();
    if (rg == 0 && c < H) {
        const int q = threadIdx.x & 63;
        part[(long)blockIdx.y * H + c] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
    }
}
__global__ void colSumFinal(int chunks, int H, const float *__restrict__ part, float *__restrict__ g) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < H; c += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int q = 0; q < chunks; q++) s += part[(long)q * H + c];
        g[c] = s;
    }
}

// embedding gradient: gEmb[branch[b]] += dz1[b]
__global__ void embScatter(int B, int H, const int *branch, const float *dz, float *gEmb) {
    const long n = (long)B * H;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int b = (int)(i / H), u = (int)(i % H);
        atomicAdd(&gEmb[(long)branch[b] * H + u], dz[i]);
    }
}

// Adam on FP32 masters, refreshed BF16 copy (bf may be null for parameters only used in FP32); four elements per
// thread with 16-byte loads (the buffers come from hipMalloc: 256-byte aligned), the tail element by element
__device__ __forceinline__ float adamOne(float &w, float gi, float &m, float &v, float lr, float b1, float b2, float c1,
                                         float c2, float eps) {
    m = b1 * m + (1.f - b1) * gi;
    v = b2 * v + (1.f - b2) * gi * gi;
    w = w - lr * (m / c1) / (sqrtf(v / c2) + eps);
    return w;
}
__global__ void adam(long n, float *w, const float *g, float *m, float *v, unsigned short *bf, float lr, float b1,
                     float b2, float c1, float c2, float eps) {
    const long n4 = n / 4;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n4; i += (long)gridDim.x * blockDim.x) {
        float4 wi = reinterpret_cast<float4 *>(w)[i];
        const float4 gi = reinterpret_cast<const float4 *>(g)[i];
        float4 mi = reinterpret_cast<float4 *>(m)[i];
        float4 vi = reinterpret_cast<float4 *>(v)[i];
        adamOne(wi.x, gi.x, mi.x, vi.x, lr, b1, b2, c1, c2, eps);
        adamOne(wi.y, gi.y, mi.y, vi.y, lr, b1, b2, c1, c2, eps);
        adamOne(wi.z, gi.z, mi.z, vi.z, lr, b1, b2, c1, c2, eps);
        adamOne(wi.w, gi.w, mi.w, vi.w, lr, b1, b2, c1, c2, eps);
        reinterpret_cast<float4 *>(w)[i] = wi;
        reinterpret_cast<float4 *>(m)[i] = mi;
        reinterpret_cast<float4 *>(v)[i] = vi;
        if (bf) {
            const unsigned long long p = (unsigned long long)toBf16(wi.x) | ((unsigned long long)toBf16(wi.y) << 16) |
                                         ((unsigned long long)toBf16(wi.z) << 32) |
                                         ((unsigned long long)toBf16(wi.w) << 48);
            reinterpret_cast<unsigned long long *>(bf)[i] = p;
        }
    }
    for (long i = 4 * n4 + blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float wi = w[i], mi = m[i], vi = v[i];
        adamOne(wi, g[i], mi, vi, lr, b1, b2, c1, c2, eps);
        w[i] = wi;
        m[i] = mi;
        v[i] = vi;
        if (bf) bf[i] = toBf16(wi);
    }
}

// The neighbourhood's local value of row c (c < K: angle of slot c relative to slot 0; c >= K: voltage of slot c - K
// minus 1) in base state x (stable rows: 2 bus = angle, 2 bus + 1 = voltage); 0 for a padding slot.
__device__ __forceinline__ double localValue(const double *x, const int *hood, int K, int c) {
    if (c < K) {
        const int bus = hood[c];
        return bus < 0 ? 0.0 : x[2 * (long)bus] - x[2 * (long)hood[0]];
    }
    const int bus = hood[c - K];
    return bus < 0 ? 0.0 : x[2 * (long)bus + 1] - 1.0;
}

// layer-1 input rows [x_loc_norm ; f] as BF16: x_loc_norm = (x_loc - mean) / std, x_loc from base state row
// baseRow[b] (row 0 when baseRow is null) of xBase [rows][nFull]
__global__ void buildInput(int B, int K, int F, const double *xBase, long nFull, const int *baseRow, const int *branch,
                           const int *hoods, const float *xMean, const float *xStd, const float *feat,
                           unsigned short *out) {
    const int N = 2 * K;
    const long D = (long)N + F;
    const long n = (long)B * D;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int b = (int)(i / D);
        const int c = (int)(i % D);
        float v;
        if (c < N) {
            const double *x = xBase + (baseRow ? (long)baseRow[b] : 0L) * nFull;
            const double loc = localValue(x, hoods + (long)branch[b] * K, K, c);
            v = (float)((loc - (double)xMean[c]) / (double)xStd[c]);
        } else {
            v = feat[(long)b * F + (c - N)];
        }
        out[i] = toBf16(v);
    }
}

// x0[b] = base state of scenario b (the rows outside its neighbourhood keep their base value)
__global__ void warmStartBase(int B, long nFull, const double *xBase, const int *baseRow, double *x0) {
    const long n = (long)B * nFull;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int b = (int)(i / nFull);
        const long r = i % nFull;
        x0[i] = xBase[(baseRow ? (long)baseRow[b] : 0L) * nFull + r];
    }
}

// x0[b][row of slot c] += (y[b][c] * std_c + mean_c) * scale_b over the neighbourhood (slots of one neighbourhood are
// distinct buses, so no two threads add to one element)
__global__ void warmStartScatter(int B, int K, long nFull, const int *branch, const int *hoods, const float *y,
                                 const float *yMean, const float *yStd, const float *scale, double *x0) {
    const int N = 2 * K;
    const long n = (long)B * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int b = (int)(i / N), c = (int)(i % N);
        const int bus = hoods[(long)branch[b] * K + (c < K ? c : c - K)];
        if (bus < 0) continue;
        const double d = ((double)y[i] * (double)yStd[c] + (double)yMean[c]) * (scale ? (double)scale[b] : 1.0);
        x0[(long)b * nFull + 2 * (long)bus + (c < K ? 0 : 1)] += d;
    }
}

// ---------------------------------------------------------------------------------------------------------------
struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t b) {
        if (p && b <= bytes) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        MLP_OK(hipMalloc(&p, b ? b : 16));
        bytes = b;
    }
    template <typename X>
    X *as() const { return reinterpret_cast<X *>(p); }
};

static int gridOf(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

// Neighbourhoods of every branch: the K buses nearest the branch by the sum of |weight| along paths, from both
// endpoints at distance 0 (endpoint 1, endpoint 2, then in order of settlement; ties by bus index), -1 padded when the
// component is smaller than K.
static std::vector<int> neighbourhoods(int nBus, int E, const int *bus1, const int *bus2, const double *weight, int K) {
    std::vector<int> deg(nBus + 1, 0);
    for (int e = 0; e < E; e++) {
        deg[bus1[e] + 1]++;
        deg[bus2[e] + 1]++;
    }
    for (int i = 0; i < nBus; i++) deg[i + 1] += deg[i];
    std::vector<int> adj(2 * (size_t)E), fill(deg.begin(), deg.end() - 1);
    std::vector<double> w(2 * (size_t)E);
    for (int e = 0; e < E; e++) {
        adj[fill[bus1[e]]] = bus2[e];
        w[fill[bus1[e]]++] = std::fabs(weight[e]);
        adj[fill[bus2[e]]] = bus1[e];
        w[fill[bus2[e]]++] = std::fabs(weight[e]);
    }
    std::vector<int> out((size_t)E * K, -1);
    std::vector<double> dist(nBus, INFINITY);
    std::vector<char> done(nBus, 0);
    std::vector<int> touched;
    typedef std::pair<double, int> QE;
    std::vector<QE> heap;
    auto cmp = [](const QE &a, const QE &b) { return a.first > b.first || (a.first == b.first && a.second > b.second); };
    for (int e = 0; e < E; e++) {
        int *h = &out[(size_t)e * K];
        int k = 0;
        h[k++] = bus1[e];
        if (bus2[e] != bus1[e] && k < K) h[k++] = bus2[e];
        heap.clear();
        for (int s = 0; s < k; s++) {
            dist[h[s]] = 0.0;
            touched.push_back(h[s]);
            heap.push_back(QE(0.0, h[s]));
        }
        std::make_heap(heap.begin(), heap.end(), cmp);
        int settled = 0;
        while (!heap.empty() && settled < K) {
            std::pop_heap(heap.begin(), heap.end(), cmp);
            const QE q = heap.back();
            heap.pop_back();
            const int u = q.second;
            if (done[u] || q.first > dist[u]) continue;
            done[u] = 1;
            settled++;
            if (u != bus1[e] && u != bus2[e]) h[k++] = u;
            for (int p = deg[u]; p < deg[u + 1]; p++) {
                const int v = adj[p];
                const double nd = q.first + w[p];
                if (nd < dist[v]) {
                    if (dist[v] == INFINITY) touched.push_back(v);
                    dist[v] = nd;
                    heap.push_back(QE(nd, v));
                    std::push_heap(heap.begin(), heap.end(), cmp);
                }
            }
        }
        for (int v : touched) {
            dist[v] = INFINITY;
            done[v] = 0;
        }
        touched.clear();
    }
    return out;
}

struct Mlp {
    int K = 0, N = 0, E = 0, F = 0, H = 0;   // N = 2K network rows
    long nFull = 0;                           // state rows of the grid (2 per bus)
    std::vector<int> hoodsHost;
    Buf hoods;                                // [E][K]
    Buf bases;                                // [baseRows][nFull] base states
    int baseRows = 0;
    long D() const { return (long)N + F; }
    // parameters: FP32 masters, BF16 copies of the GEMM weights, Adam moments, gradients
    struct Param {
        long n = 0;
        Buf w, g, m, v, bf;
    };
    Param Win, Emb, b1, W2, b2, W3, b3;
    Buf xMean, xStd, yMean, yStd;
    // activations for a batch of up to cap
    int cap = 0;
    Buf X, c1, z1, h1, c2, z2, h2, y, dy, dyh, dh2, dz2, dz2h, dh1, dz1, dz1h, tgt, branch, baseRow, feat, scale, loss,
        x0, colPart;
    long step = 0;
    hipStream_t s = nullptr;

    void initParam(Param &p, long n, float scale, std::mt19937_64 &rng, bool bf) {
        p.n = n;
        std::vector<float> h(n);
        std::normal_distribution<float> nd(0.f, 1.f);
        for (long i = 0; i < n; i++) h[i] = scale * nd(rng);
        p.w.alloc(n * 4);
        MLP_OK(hipMemcpy(p.w.p, h.data(), n * 4, hipMemcpyHostToDevice));
        p.g.alloc(n * 4);
        p.m.alloc(n * 4);
        p.v.alloc(n * 4);
        zeroF<<<gridOf(n), 256, 0, s>>>(p.m.as<float>(), n);
        zeroF<<<gridOf(n), 256, 0, s>>>(p.v.as<float>(), n);
        if (bf) {
            p.bf.alloc(n * 2);
            f32ToBf16<<<gridOf(n), 256, 0, s>>>(p.w.as<float>(), p.bf.as<unsigned short>(), n);
        }
    }
    void ensureBatch(int B) {
        if (B <= cap) return;
        cap = B;
        X.alloc((size_t)B * D() * 2);
        for (Buf *b : {&c1, &z1, &c2, &z2, &dh2, &dz2, &dh1, &dz1}) b->alloc((size_t)B * H * 4);
        for (Buf *b : {&h1, &h2, &dz2h, &dz1h}) b->alloc((size_t)B * H * 2);
        for (Buf *b : {&y, &dy, &tgt}) b->alloc((size_t)B * N * 4);
        dyh.alloc((size_t)B * N * 2);
        branch.alloc((size_t)B * 4);
        baseRow.alloc((size_t)B * 4);
        scale.alloc((size_t)B * 4);
        feat.alloc((size_t)B * std::max(F, 1) * 4);
        loss.alloc(8);
    }
    // C[M x N] = A B with K chunked so that the grid has at least ~512 workgroups
    void gemm(int M, int Nn, int K, const unsigned short *A, long sam, long sak, const unsigned short *B, long sbk,
              long sbn, float *C, long ldc) {
        const int gx = (Nn + GT - 1) / GT, gy = (M + GT - 1) / GT;
        int splits = 1;
        while (gx * gy * splits < 512 && K / (splits * 2) >= 256) splits *= 2;
        int kChunk = (K + splits - 1) / splits;
        kChunk = (kChunk + GK - 1) / GK * GK;
        splits = (K + kChunk - 1) / kChunk;
        if (splits > 1) zeroF<<<gridOf((long)M * ldc), 256, 0, s>>>(C, (long)M * ldc);
        gemmBf16<<<dim3(gx, gy, splits), 256, 0, s>>>(M, Nn, K, A, sam, sak, B, sbk, sbn, C, ldc, kChunk, splits > 1);
    }
    void forward(int B) {
        const int Hh = H;
        // c1 = X Win^T : [B x D] [D x H]
        gemm(B, Hh, (int)D(), X.as<unsigned short>(), D(), 1, Win.bf.as<unsigned short>(), 1, D(), c1.as<float>(), Hh);
        biasGeluFwd<<<gridOf((long)B * Hh), 256, 0, s>>>(B, Hh, c1.as<float>(), b1.w.as<float>(), Emb.w.as<float>(),
                                                         branch.as<int>(), z1.as<float>(), h1.as<unsigned short>());
        gemm(B, Hh, Hh, h1.as<unsigned short>(), Hh, 1, W2.bf.as<unsigned short>(), 1, Hh, c2.as<float>(), Hh);
        biasGeluFwd<<<gridOf((long)B * Hh), 256, 0, s>>>(B, Hh, c2.as<float>(), b2.w.as<float>(), nullptr, nullptr,
                                                         z2.as<float>(), h2.as<unsigned short>());
        gemm(B, N, Hh, h2.as<unsigned short>(), Hh, 1, W3.bf.as<unsigned short>(), 1, Hh, y.as<float>(), N);
        biasAdd<<<gridOf((long)B * N), 256, 0, s>>>(B, N, y.as<float>(), b3.w.as<float>());
    }
    // column sums of d [B x Hc] into g
    void colSum(int B, int Hc, const float *d, float *g) {
        const int chunks = (B + COLSUM_ROWS - 1) / COLSUM_ROWS;
        colPart.alloc((size_t)chunks * Hc * 4);
        colSumPart<<<dim3((Hc + 63) / 64, chunks), 256, 0, s>>>(B, Hc, d, colPart.as<float>());
        colSumFinal<<<gridOf(Hc), 256, 0, s>>>(chunks, Hc, colPart.as<float>(), g);
    }
    // gradients of the mean squared error into the .g buffers; returns the loss
    double backward(int B) {
        const int Hh = H;
        MLP_OK(hipMemsetAsync(loss.p, 0, 8, s));
        mseGrad<<<gridOf((long)B * N), 256, 0, s>>>(B, N, y.as<float>(), tgt.as<float>(), dy.as<float>(),
                                                   dyh.as<unsigned short>(), loss.as<double>());
        // W3: g = dy^T h2 : [N x B][B x H]
        gemm(N, Hh, B, dyh.as<unsigned short>(), 1, N, h2.as<unsigned short>(), Hh, 1, W3.g.as<float>(), Hh);
        colSum(B, N, dy.as<float>(), b3.g.as<float>());
        // dh2 = dy W3 : [B x N][N x H]
        gemm(B, Hh, N, dyh.as<unsigned short>(), N, 1, W3.bf.as<unsigned short>(), Hh, 1, dh2.as<float>(), Hh);
        geluBwd<<<gridOf((long)B * Hh), 256, 0, s>>>((long)B * Hh, dh2.as<float>(), z2.as<float>(), dz2.as<float>(),
                                                     dz2h.as<unsigned short>());
        gemm(Hh, Hh, B, dz2h.as<unsigned short>(), 1, Hh, h1.as<unsigned short>(), Hh, 1, W2.g.as<float>(), Hh);
        colSum(B, Hh, dz2.as<float>(), b2.g.as<float>());
        gemm(B, Hh, Hh, dz2h.as<unsigned short>(), Hh, 1, W2.bf.as<unsigned short>(), Hh, 1, dh1.as<float>(), Hh);
        geluBwd<<<gridOf((long)B * Hh), 256, 0, s>>>((long)B * Hh, dh1.as<float>(), z1.as<float>(), dz1.as<float>(),
                                                     dz1h.as<unsigned short>());
        gemm(Hh, (int)D(), B, dz1h.as<unsigned short>(), 1, Hh, X.as<unsigned short>(), D(), 1, Win.g.as<float>(), D());
        colSum(B, Hh, dz1.as<float>(), b1.g.as<float>());
        zeroF<<<gridOf(Emb.n), 256, 0, s>>>(Emb.g.as<float>(), Emb.n);
        embScatter<<<gridOf((long)B * Hh), 256, 0, s>>>(B, Hh, branch.as<int>(), dz1.as<float>(), Emb.g.as<float>());
        double l = 0.0;
        MLP_OK(hipMemcpyAsync(&l, loss.p, 8, hipMemcpyDeviceToHost, s));
        MLP_OK(hipStreamSynchronize(s));
        return l / ((double)B * N);
    }
    void adamStep(float lr) {
        step++;
        const float b1c = 0.9f, b2c = 0.999f;
        const float c1 = 1.f - std::pow(b1c, (float)step), c2 = 1.f - std::pow(b2c, (float)step);
        for (Param *p : {&Win, &Emb, &b1, &W2, &b2, &W3, &b3})
            adam<<<gridOf(p->n), 256, 0, s>>>(p->n, p->w.as<float>(), p->g.as<float>(), p->m.as<float>(),
                                              p->v.as<float>(), p->bf.p ? p->bf.as<unsigned short>() : nullptr, lr, b1c,
                                              b2c, c1, c2, 1e-8f);
    }
    // checks a batch's arguments, uploads them and builds the layer-1 input; rows may be null (base row 0)
    void upload(int B, const int *rows, const int *br, const float *ft, const float *tg) {
        if (B <= 0) throw ArgumentError("batch must be positive");
        if (!br || (F > 0 && !ft)) throw ArgumentError("null argument");
        if (baseRows == 0) throw ArgumentError("no base states set (olfhip_mlp_set_base_states)");
        for (int b = 0; b < B; b++) {
            if (br[b] < 0 || br[b] >= E) throw ArgumentError("branch id out of range");
            if (rows && (rows[b] < 0 || rows[b] >= baseRows)) throw ArgumentError("base row out of range");
        }
        ensureBatch(B);
        MLP_OK(hipMemcpyAsync(branch.p, br, (size_t)B * 4, hipMemcpyHostToDevice, s));
        if (rows) MLP_OK(hipMemcpyAsync(baseRow.p, rows, (size_t)B * 4, hipMemcpyHostToDevice, s));
        if (F > 0) MLP_OK(hipMemcpyAsync(feat.p, ft, (size_t)B * F * 4, hipMemcpyHostToDevice, s));
        if (tg) MLP_OK(hipMemcpyAsync(tgt.p, tg, (size_t)B * N * 4, hipMemcpyHostToDevice, s));
        buildInput<<<gridOf((long)B * D()), 256, 0, s>>>(B, K, F, bases.as<double>(), nFull,
                                                        rows ? baseRow.as<int>() : nullptr, branch.as<int>(),
                                                        hoods.as<int>(), xMean.as<float>(), xStd.as<float>(),
                                                        feat.as<float>(), X.as<unsigned short>());
    }
    ~Mlp() {
        if (s) (void)hipStreamDestroy(s);
    }
};

}  // namespace olfhip

using namespace olfhip;

template <typename Fn>
static int mlpGuarded(Fn &&fn) {
    try {
        lastErrorRef().clear();
        fn();
        return OLFHIP_OK;
    } catch (const ArgumentError &e) {
        lastErrorRef() = e.what();
        return OLFHIP_E_ARGUMENT;
    } catch (const MemoryError &e) {
        lastErrorRef() = e.what();
        return OLFHIP_E_MEMORY;
    } catch (const std::exception &e) {
        lastErrorRef() = std::string("warm-start network failed: ") + e.what();
        return OLFHIP_E_DEVICE;
    }
}

extern "C" {

int olfhip_mlp_create(int n_bus, int n_branch, const int *bus1, const int *bus2, const double *weight, int n_local,
                      int n_features, int hidden, unsigned long long seed, olfhip_mlp **out) {
    return mlpGuarded([&] {
        if (!out) throw ArgumentError("out is null");
        *out = nullptr;
        if (!bus1 || !bus2 || !weight) throw ArgumentError("null argument");
        if (n_bus <= 0 || n_branch <= 0 || n_local < 2 || n_local > n_bus || n_features < 0 || hidden <= 0 ||
            hidden % 16)
            throw ArgumentError("sizes must be positive, 2 <= n_local <= n_bus and hidden a multiple of 16");
        for (int e = 0; e < n_branch; e++)
            if (bus1[e] < 0 || bus1[e] >= n_bus || bus2[e] < 0 || bus2[e] >= n_bus)
                throw ArgumentError("branch end out of range");
        std::unique_ptr<Mlp> m(new Mlp());
        m->K = n_local;
        m->N = 2 * m->K;
        m->nFull = 2L * n_bus;
        m->E = n_branch;
        m->F = n_features;
        m->H = hidden;
        m->hoodsHost = neighbourhoods(n_bus, n_branch, bus1, bus2, weight, m->K);
        MLP_OK(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
        m->hoods.alloc(m->hoodsHost.size() * 4);
        MLP_OK(hipMemcpy(m->hoods.p, m->hoodsHost.data(), m->hoodsHost.size() * 4, hipMemcpyHostToDevice));
        std::mt19937_64 rng(seed);
        const int N = m->N;
        const float sIn = 1.f / std::sqrt((float)m->D()), sH = 1.f / std::sqrt((float)hidden);
        m->initParam(m->Win, (long)hidden * m->D(), sIn, rng, true);
        m->initParam(m->Emb, (long)n_branch * hidden, 0.1f, rng, false);
        m->initParam(m->b1, hidden, 0.f, rng, false);
        m->initParam(m->W2, (long)hidden * hidden, sH, rng, true);
        m->initParam(m->b2, hidden, 0.f, rng, false);
        m->initParam(m->W3, (long)N * hidden, sH, rng, true);
        m->initParam(m->b3, N, 0.f, rng, false);
        std::vector<float> zeros(N, 0.f), ones(N, 1.f);
        for (Buf *b : {&m->xMean, &m->yMean}) {
            b->alloc(N * 4);
            MLP_OK(hipMemcpy(b->p, zeros.data(), N * 4, hipMemcpyHostToDevice));
        }
        for (Buf *b : {&m->xStd, &m->yStd}) {
            b->alloc(N * 4);
            MLP_OK(hipMemcpy(b->p, ones.data(), N * 4, hipMemcpyHostToDevice));
        }
        MLP_OK(hipStreamSynchronize(m->s));
        *out = reinterpret_cast<olfhip_mlp *>(m.release());
    });
}

void olfhip_mlp_destroy(olfhip_mlp *mlp) { delete reinterpret_cast<Mlp *>(mlp); }

int olfhip_neighbourhoods(int n_bus, int n_branch, const int *bus1, const int *bus2, const double *weight,
                          int n_local, int *buses) {
    return mlpGuarded([&] {
        if (!bus1 || !bus2 || !weight || !buses) throw ArgumentError("null argument");
        if (n_bus <= 0 || n_branch <= 0 || n_local < 2 || n_local > n_bus)
            throw ArgumentError("sizes must be positive and 2 <= n_local <= n_bus");
        for (int e = 0; e < n_branch; e++)
            if (bus1[e] < 0 || bus1[e] >= n_bus || bus2[e] < 0 || bus2[e] >= n_bus)
                throw ArgumentError("branch end out of range");
        const std::vector<int> h = neighbourhoods(n_bus, n_branch, bus1, bus2, weight, n_local);
        std::memcpy(buses, h.data(), h.size() * 4);
    });
}

int olfhip_mlp_set_normalization(olfhip_mlp *mlp, const float *x_mean, const float *x_std, const float *y_mean,
                                 const float *y_std) {
    return mlpGuarded([&] {
        Mlp *m = reinterpret_cast<Mlp *>(mlp);
        if (!m || !x_mean || !x_std || !y_mean || !y_std) throw ArgumentError("null argument");
        for (int r = 0; r < m->N; r++)
            if (!(x_std[r] > 0.f) || !(y_std[r] > 0.f)) throw ArgumentError("standard deviations must be positive");
        MLP_OK(hipMemcpy(m->xMean.p, x_mean, m->N * 4, hipMemcpyHostToDevice));
        MLP_OK(hipMemcpy(m->xStd.p, x_std, m->N * 4, hipMemcpyHostToDevice));
        MLP_OK(hipMemcpy(m->yMean.p, y_mean, m->N * 4, hipMemcpyHostToDevice));
        MLP_OK(hipMemcpy(m->yStd.p, y_std, m->N * 4, hipMemcpyHostToDevice));
    });
}

int olfhip_mlp_set_base_states(olfhip_mlp *mlp, int n_rows, const double *x_base) {
    return mlpGuarded([&] {
        Mlp *m = reinterpret_cast<Mlp *>(mlp);
        if (!m || !x_base) throw ArgumentError("null argument");
        if (n_rows <= 0) throw ArgumentError("n_rows must be positive");
        MLP_OK(hipStreamSynchronize(m->s));
        m->bases.alloc((size_t)n_rows * m->nFull * 8);
        MLP_OK(hipMemcpy(m->bases.p, x_base, (size_t)n_rows * m->nFull * 8, hipMemcpyHostToDevice));
        m->baseRows = n_rows;
    });
}

int olfhip_mlp_train_step(olfhip_mlp *mlp, int batch, const int *base_row, const int *branch, const float *features,
                          const float *target, float learning_rate, double *loss) {
    return mlpGuarded([&] {
        Mlp *m = reinterpret_cast<Mlp *>(mlp);
        if (!m || !target) throw ArgumentError("null argument");
        m->upload(batch, base_row, branch, features, target);
        m->forward(batch);
        const double l = m->backward(batch);
        m->adamStep(learning_rate);
        MLP_OK(hipStreamSynchronize(m->s));
        if (loss) *loss = l;
    });
}

int olfhip_mlp_forward(olfhip_mlp *mlp, int batch, const int *base_row, const int *branch, const float *features,
                       float *y_out) {
    return mlpGuarded([&] {
        Mlp *m = reinterpret_cast<Mlp *>(mlp);
        if (!m || !y_out) throw ArgumentError("null argument");
        m->upload(batch, base_row, branch, features, nullptr);
        m->forward(batch);
        MLP_OK(hipMemcpyAsync(y_out, m->y.p, (size_t)batch * m->N * 4, hipMemcpyDeviceToHost, m->s));
        MLP_OK(hipStreamSynchronize(m->s));
    });
}

int olfhip_mlp_gradients(olfhip_mlp *mlp, int batch, const int *base_row, const int *branch, const float *features,
                         const float *target, int which, float *grad_out, double *loss) {
    return mlpGuarded([&] {
        Mlp *m = reinterpret_cast<Mlp *>(mlp);
        if (!m || !target || !grad_out) throw ArgumentError("null argument");
        Mlp::Param *ps[7] = {&m->Win, &m->Emb, &m->b1, &m->W2, &m->b2, &m->W3, &m->b3};
        if (which < 0 || which > 6) throw ArgumentError("which selects one of the 7 parameter arrays");
        m->upload(batch, base_row, branch, features, target);
        m->forward(batch);
        const double l = m->backward(batch);
        MLP_OK(hipMemcpy(grad_out, ps[which]->g.p, ps[which]->n * 4, hipMemcpyDeviceToHost));
        if (loss) *loss = l;
    });
}

int olfhip_mlp_parameters(olfhip_mlp *mlp, int which, float *values_out, const float *values_in) {
    return mlpGuarded([&] {
        Mlp *m = reinterpret_cast<Mlp *>(mlp);
        if (!m) throw ArgumentError("null argument");
        Mlp::Param *ps[7] = {&m->Win, &m->Emb, &m->b1, &m->W2, &m->b2, &m->W3, &m->b3};
        if (which < 0 || which > 6) throw ArgumentError("which selects one of the 7 parameter arrays");
        Mlp::Param &p = *ps[which];
        if (values_in) {
            MLP_OK(hipMemcpy(p.w.p, values_in, p.n * 4, hipMemcpyHostToDevice));
            if (p.bf.p) f32ToBf16<<<gridOf(p.n), 256, 0, m->s>>>(p.w.as<float>(), p.bf.as<unsigned short>(), p.n);
            MLP_OK(hipStreamSynchronize(m->s));
        }
        if (values_out) MLP_OK(hipMemcpy(values_out, p.w.p, p.n * 4, hipMemcpyDeviceToHost));
    });
}

int olfhip_mlp_predict_warm_start(olfhip_mlp *mlp, int batch, const int *base_row, const int *branch,
                                  const float *features, const float *scale, double **x0_device) {
    return mlpGuarded([&] {
        Mlp *m = reinterpret_cast<Mlp *>(mlp);
        if (!m || !x0_device) throw ArgumentError("null argument");
        if (scale)
            for (int b = 0; b < batch; b++)
                if (!std::isfinite(scale[b])) throw ArgumentError("scale must be finite");
        m->upload(batch, base_row, branch, features, nullptr);
        if (scale) MLP_OK(hipMemcpyAsync(m->scale.p, scale, (size_t)batch * 4, hipMemcpyHostToDevice, m->s));
        m->forward(batch);
        m->x0.alloc((size_t)batch * m->nFull * 8);
        const int *rows = base_row ? m->baseRow.as<int>() : nullptr;
        warmStartBase<<<gridOf((long)batch * m->nFull), 256, 0, m->s>>>(batch, m->nFull, m->bases.as<double>(), rows,
                                                                        m->x0.as<double>());
        warmStartScatter<<<gridOf((long)batch * m->N), 256, 0, m->s>>>(
            batch, m->K, m->nFull, m->branch.as<int>(), m->hoods.as<int>(), m->y.as<float>(), m->yMean.as<float>(),
            m->yStd.as<float>(), scale ? m->scale.as<float>() : nullptr, m->x0.as<double>());
        MLP_OK(hipStreamSynchronize(m->s));
        *x0_device = m->x0.as<double>();
    });
}

// GEMM on the matrix cores, for the unit test of the lane layout: C = A B, A [M x K] and B [K x N] row-major FP32
// (rounded to BF16 on upload), C FP32.
int olfhip_mlp_gemm_test(int M, int N, int K, const float *A, const float *B, float *C) {
    return mlpGuarded([&] {
        if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) throw ArgumentError("bad arguments");
        Buf a32, b32, a16, b16, c;
        a32.alloc((size_t)M * K * 4);
        b32.alloc((size_t)K * N * 4);
        a16.alloc((size_t)M * K * 2);
        b16.alloc((size_t)K * N * 2);
        c.alloc((size_t)M * N * 4);
        MLP_OK(hipMemcpy(a32.p, A, (size_t)M * K * 4, hipMemcpyHostToDevice));
        MLP_OK(hipMemcpy(b32.p, B, (size_t)K * N * 4, hipMemcpyHostToDevice));
        f32ToBf16<<<gridOf((long)M * K), 256>>>(a32.as<float>(), a16.as<unsigned short>(), (long)M * K);
        f32ToBf16<<<gridOf((long)K * N), 256>>>(b32.as<float>(), b16.as<unsigned short>(), (long)K * N);
        Mlp tmp;
        tmp.s = nullptr;
        tmp.gemm(M, N, K, a16.as<unsigned short>(), K, 1, b16.as<unsigned short>(), N, 1, c.as<float>(), N);
        MLP_OK(hipDeviceSynchronize());
        MLP_OK(hipMemcpy(C, c.p, (size_t)M * N * 4, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
