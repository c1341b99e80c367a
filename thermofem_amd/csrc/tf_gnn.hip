// Learned edge affinities on gfx950: a two-layer message-passing network
// (hidden 32) scores every off-diagonal entry of a matrix with a symmetric
// pattern; the pairwise matching consumes the scores in place of the
// strength weights (src/Multigrid/GpuAggregation.cu:339-364 builds the
// weights that the matching at :402-431 reads; the reference has no learned
// part).  Training runs here too: forward, backward and Adam are kernels of
// this file, so the network is trained on the GPU that uses it.
//
// Network (H = 32, node features FN = 4, edge features FE = 3), edges
// p = (i, j) in CSR order without the diagonal:
//   f_i = [sum_j |a_ij| / |d_i|, max_j |a_ij| / |d_i|, deg_i / 26, 1]
//   s_p = (|a_ij| + |a_ji|) / 2, m_i = max_j |a_ij|,
//   g_p = [s_p / sqrt(|d_i d_j|), min(s_p / m_i, s_p / m_j), max(...)]
//   h0 = relu(W0 f + b0)
//   twice: m_p = relu(Wm [h_j ; g_p] + bm), M_i = mean_{p in row i} m_p,
//          h_i = relu(Wu [h_i ; M_i] + bu)
//   z_p = relu(We [h_i + h_j ; |h_i - h_j| ; g_p] + be), w_p = sigmoid(wo . z_p + bo)
// Weights are kept in FP32 (the Adam master copy) and used rounded to
// bfloat16; every product is accumulated in FP32.  No reduction uses
// atomics: weight gradients are per-block partial sums added in block
// order, and a node's share of its edges' gradients is gathered through
// the twin edge (j, i) of each edge, so training is deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tf_device.hpp"
#include "tf_internal.hpp"

namespace tf {

namespace {

constexpr int H = 32, FN = 4, FE = 3;
constexpr int kLayers = 2;
constexpr int kTileR = 64;        // rows per tile of the dense kernels
constexpr int kGradBlocks = 512;  // blocks (= partials) of a weight-gradient reduction
constexpr int kMaxIn = 2 * H + FE;

// Offsets of the parameter blocks in the flat parameter vector.
struct Layout {
    int W0, b0, Wm[kLayers], bm[kLayers], Wu[kLayers], bu[kLayers], We, be, wo, bo, count;
    Layout() {
        int at = 0;
        auto take = [&](int n) {
            const int r = at;
            at += n;
            return r;
        };
        W0 = take(H * FN);
        b0 = take(H);
        for (int l = 0; l < kLayers; ++l) {
            Wm[l] = take(H * (H + FE));
            bm[l] = take(H);
            Wu[l] = take(H * 2 * H);
            bu[l] = take(H);
        }
        We = take(H * (2 * H + FE));
        be = take(H);
        wo = take(H);
        bo = take(1);
        count = at;
    }
};

// ------------------------------------------------------------ graph build

// The graph kernels give each row a group of kRowLanes lanes that walk its
// entries kRowLanes at a time (coalesced loads and stores); an entry's place
// in the row's edge list is the number of off-diagonal entries before it,
// counted with a ballot over the group.
constexpr int kRowLanes = 16;

struct RowLane {
    int i, lane, gbase;
};
__device__ __forceinline__ RowLane row_lane() {
    const int t = int(blockIdx.x) * kBlock + int(threadIdx.x);
    return RowLane{t / kRowLanes, t % kRowLanes, (int(threadIdx.x) % kWave) / kRowLanes * kRowLanes};
}
// ballot bits of the calling lane's group, and how many of them lie below it
__device__ __forceinline__ unsigned group_bits(bool pred, const RowLane& r) {
    const unsigned long long b = __ballot(pred);
    return unsigned((b >> r.gbase) & ((1ull << kRowLanes) - 1));
}

__global__ void k_offdiag_count(int n, const int* __restrict__ rp, const int* __restrict__ ci, int* cnt) {
    const RowLane r = row_lane();
    if (r.i > n) return;
    if (r.i == n) {
        if (r.lane == 0) cnt[n] = 0;
        return;
    }
    int c = 0;
    for (int j = rp[r.i] + r.lane; j < rp[r.i + 1]; j += kRowLanes) c += ci[j] != r.i;
    c = group_sum<kRowLanes>(c);
    if (r.lane == 0) cnt[r.i] = c;
}

// Edges of row i in CSR order, node features, |d_i| and m_i.
__global__ void k_edges_nodes(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                              const double* __restrict__ v, const int* __restrict__ erp, int* __restrict__ ei,
                              int* __restrict__ ej, double* __restrict__ ea, double* __restrict__ dabs,
                              double* __restrict__ mrow, float* __restrict__ f, float* __restrict__ deg) {
    const RowLane r = row_lane();
    const bool live = r.i < n;  // (every lane of a group takes part in the ballots)
    const int i = live ? r.i : n - 1;
    double d = 0, m = 0, sum = 0;
    int at = erp[i];
    const int s = rp[i], e = rp[i + 1];
    for (int j0 = s; j0 < e; j0 += kRowLanes) {
        const int j = j0 + r.lane;
        const bool valid = j < e;
        const int c = valid ? ci[j] : i;
        const bool off = valid && c != i;
        const unsigned bits = group_bits(off, r);
        if (valid) {
            const double x = v[j];
            if (c == i) {
                d += x;
            }
            else {
                const double ax = fabs(x);
                const int q = at + __popc(bits & ((1u << r.lane) - 1));
                if (live) {
                    ei[q] = i;
                    ej[q] = c;
                    ea[q] = ax;
                }
                m = fmax(m, ax);
                sum += ax;
            }
        }
        at += __popc(bits);
    }
    d = group_sum<kRowLanes>(d);
    sum = group_sum<kRowLanes>(sum);
#pragma unroll
    for (int o = kRowLanes / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, kRowLanes));
    if (!live || r.lane != 0) return;
    const int dg = erp[i + 1] - erp[i];
    d = fabs(d);
    dabs[i] = d;
    mrow[i] = m;
    deg[i] = float(dg);
    f[size_t(i) * FN + 0] = float(sum / d);
    f[size_t(i) * FN + 1] = float(m / d);
    f[size_t(i) * FN + 2] = float(double(dg) / 26.0);
    f[size_t(i) * FN + 3] = 1.0f;
}

// twin edge (j, i) of every edge by binary search in row j (bad when
// absent), and the edge's features from it.
__global__ void k_twins(int64_t E, const int* __restrict__ erp, const int* __restrict__ ei,
                        const int* __restrict__ ej, int* __restrict__ twin, int* bad, const double* __restrict__ ea,
                        const double* __restrict__ dabs, const double* __restrict__ mrow, float* __restrict__ g) {
    const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (p >= E) return;
    const int i = ei[p], j = ej[p];
    int lo = erp[j], hi = erp[j + 1] - 1, found = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int c = ej[mid];
        if (c == i) {
            found = mid;
            break;
        }
        if (c < i)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    if (found < 0) atomicOr(bad, 1);
    const int t = found < 0 ? int(p) : found;
    twin[p] = t;
    const double s = 0.5 * (ea[p] + ea[t]);
    const double ri = s / mrow[i], rj = s / mrow[j];
    g[size_t(p) * FE + 0] = float(s / sqrt(dabs[i] * dabs[j]));
    g[size_t(p) * FE + 1] = float(fmin(ri, rj));
    g[size_t(p) * FE + 2] = float(fmax(ri, rj));
}

// --------------------------------------------------------- dense layers

// Y = relu(X W^T + b) for `rows` rows of IN inputs, H outputs: a tile of 64
// rows and W in LDS; thread t computes outputs 8 (t % 4) .. +7 of row t / 4.
__global__ __launch_bounds__(kBlock) void k_dense(int64_t rows, int in, const float* __restrict__ X,
                                                  const float* __restrict__ W, const float* __restrict__ b,
                                                  float* __restrict__ Y) {
    __shared__ float sw[H * kMaxIn];
    __shared__ float sx[kTileR * kMaxIn];
    for (int q = int(threadIdx.x); q < H * in; q += kBlock) sw[q] = W[q];
    const int64_t r0 = int64_t(blockIdx.x) * kTileR;
    const int64_t left = rows - r0;
    const int tr = int(left < kTileR ? left : kTileR);
    for (int q = int(threadIdx.x); q < tr * in; q += kBlock) sx[q] = X[size_t(r0) * in + q];
    __syncthreads();
    const int r = int(threadIdx.x) / 4, o0 = (int(threadIdx.x) % 4) * 8;
    if (r >= tr) return;
    float acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = b[o0 + u];
    for (int c = 0; c < in; ++c) {
        const float x = sx[r * in + c];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] = fmaf(x, sw[(o0 + u) * in + c], acc[u]);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) Y[size_t(r0 + r) * H + o0 + u] = fmaxf(acc[u], 0.0f);
}

// dX = dY W (rows x H times H x IN).
__global__ __launch_bounds__(kBlock) void k_dense_t(int64_t rows, int in, const float* __restrict__ dY,
                                                    const float* __restrict__ W, float* __restrict__ dX) {
    __shared__ float sw[H * kMaxIn];
    __shared__ float sy[kTileR * H];
    for (int q = int(threadIdx.x); q < H * in; q += kBlock) sw[q] = W[q];
    const int64_t r0 = int64_t(blockIdx.x) * kTileR;
    const int64_t left = rows - r0;
    const int tr = int(left < kTileR ? left : kTileR);
    for (int q = int(threadIdx.x); q < tr * H; q += kBlock) sy[q] = dY[size_t(r0) * H + q];
    __syncthreads();
    const int r = int(threadIdx.x) / 4, part = int(threadIdx.x) % 4;
    if (r >= tr) return;
    for (int c = part; c < in; c += 4) {
        float acc = 0.0f;
        for (int o = 0; o < H; ++o) acc = fmaf(sy[r * H + o], sw[o * in + c], acc);
        dX[size_t(r0 + r) * in + c] = acc;
    }
}

// Weight gradient partials: for the entries e = o * (in + 1) + c of
// [dW | db] (c == in: bias), block b adds dY[r][o] * X[r][c] over the rows
// of its tiles (tiles b, b + grid, ...) in order.
__global__ __launch_bounds__(kBlock) void k_wgrad(int64_t rows, int out, int in, const float* __restrict__ dY,
                                                  const float* __restrict__ X, float* __restrict__ partials) {
    __shared__ float sy[kTileR * H];
    __shared__ float sx[kTileR * kMaxIn];
    constexpr int kPer = (H * (kMaxIn + 1) + kBlock - 1) / kBlock;
    const int entries = out * (in + 1);
    float acc[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) acc[u] = 0.0f;
    const int64_t tiles = (rows + kTileR - 1) / kTileR;
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t r0 = t * kTileR;
        const int64_t left = rows - r0;
        const int tr = int(left < kTileR ? left : kTileR);
        __syncthreads();
        for (int q = int(threadIdx.x); q < tr * out; q += kBlock) sy[q] = dY[size_t(r0) * out + q];
        for (int q = int(threadIdx.x); q < tr * in; q += kBlock) sx[q] = X[size_t(r0) * in + q];
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = int(threadIdx.x) + u * kBlock;
            if (e < entries) {
                const int o = e / (in + 1), c = e % (in + 1);
                float s = acc[u];
                for (int r = 0; r < tr; ++r) s = fmaf(sy[r * out + o], c < in ? sx[r * in + c] : 1.0f, s);
                acc[u] = s;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int e = int(threadIdx.x) + u * kBlock;
        if (e < entries) partials[size_t(blockIdx.x) * (H * (kMaxIn + 1)) + e] = acc[u];
    }
}

// grad[W] and grad[b] from the partials, added in block order.
__global__ void k_wgrad_reduce(int out, int in, int blocks, const float* __restrict__ partials,
                               float* __restrict__ gW, float* __restrict__ gb) {
    const int e = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (e >= out * (in + 1)) return;
    float s = 0.0f;
    for (int b = 0; b < blocks; ++b) s += partials[size_t(b) * (H * (kMaxIn + 1)) + e];
    const int o = e / (in + 1), c = e % (in + 1);
    if (c < in)
        gW[o * in + c] = s;
    else
        gb[o] = s;
}

// ----------------------------------------------------- gathers and heads

// xm_p = [h_j ; g_p]
__global__ void k_gather_xm(int64_t E, const float* __restrict__ h, const int* __restrict__ ej,
                            const float* __restrict__ g, float* __restrict__ xm) {
    const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= E * (H + FE)) return;
    const int64_t p = q / (H + FE);
    const int c = int(q % (H + FE));
    xm[q] = c < H ? h[size_t(ej[p]) * H + c] : g[size_t(p) * FE + (c - H)];
}

// xu_i = [h_i ; mean of m over the edges of row i]
__global__ void k_node_mean(int n, const int* __restrict__ erp, const float* __restrict__ m,
                            const float* __restrict__ h, float* __restrict__ xu) {
    const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= int64_t(n) * H) return;
    const int i = int(q / H), o = int(q % H);
    float s = 0.0f;
    for (int p = erp[i]; p < erp[i + 1]; ++p) s += m[size_t(p) * H + o];
    const int dg = erp[i + 1] - erp[i];
    xu[size_t(i) * 2 * H + o] = h[q];
    xu[size_t(i) * 2 * H + H + o] = s / float(dg > 0 ? dg : 1);
}

// xe_p = [h_i + h_j ; |h_i - h_j| ; g_p]
__global__ void k_gather_xe(int64_t E, const float* __restrict__ h, const int* __restrict__ ei,
                            const int* __restrict__ ej, const float* __restrict__ g, float* __restrict__ xe) {
    const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    constexpr int IN = 2 * H + FE;
    if (q >= E * IN) return;
    const int64_t p = q / IN;
    const int c = int(q % IN);
    float val;
    if (c < 2 * H) {
        const int o = c % H;
        const float a = h[size_t(ei[p]) * H + o], b = h[size_t(ej[p]) * H + o];
        val = c < H ? a + b : fabsf(a - b);
    }
    else {
        val = g[size_t(p) * FE + (c - 2 * H)];
    }
    xe[q] = val;
}

// w_p = sigmoid(wo . z_p + bo); with a target, the loss partial per block
// and ds_p = dL/ds_p for L = mean (w - t)^2.
__global__ __launch_bounds__(kBlock) void k_score(int64_t E, const float* __restrict__ z, const float* __restrict__ wo,
                                                  const float* __restrict__ bo, float* __restrict__ w,
                                                  const float* __restrict__ target, float* __restrict__ ds,
                                                  double* __restrict__ lossPartials) {
    const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    double l = 0.0;
    if (p < E) {
        float s = bo[0];
        for (int o = 0; o < H; ++o) s = fmaf(wo[o], z[size_t(p) * H + o], s);
        const float wp = 1.0f / (1.0f + expf(-s));
        w[p] = wp;
        if (target) {
            const float d = wp - target[p];
            l = double(d) * double(d);
            ds[p] = (2.0f * d / float(E)) * wp * (1.0f - wp);
        }
    }
    if (target) {
        l = block_sum(l);
        if (threadIdx.x == 0) lossPartials[blockIdx.x] = l;
    }
}

// ------------------------------------------------ fused inference (MFMA)
//
// Inference never materialises the edge activations: one kernel per
// message-passing layer forms each edge's message, its row mean and the node
// update, and one kernel forms the scorer's features, its hidden layer and
// the score.  Every product runs on the matrix cores as
// v_mfma_f32_16x16x32_bf16 (16 rows x 32 inputs x 16 outputs; lane l holds
// row l % 16, inputs 8 (l / 16) .. +7 of A and of B's column l % 16; the
// accumulator's lane l register q is row 4 (l / 16) + q, column l % 16 --
// the layout was checked on gfx950 with distinct values).  The weights are
// bfloat16 values already (the rounded copy the training path uses), so they
// go in exactly; an FP32 activation x goes in as two bfloat16 terms,
// hi = bf16(x) and lo = bf16(x - hi), whose products are accumulated in FP32:
// x is represented to 2^-17 of itself, so the scores agree with the FP32
// reference to ~1e-6 while the products run at the bfloat16 matrix rate.
typedef float mf4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ mf4 mfma32(bf8 a, bf8 b, mf4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
}

// acc += x W^T for one 32-input chunk: x split into hi + lo.
__device__ __forceinline__ mf4 mfma_split(const float (&x)[8], bf8 w, mf4 acc) {
    bf8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = static_cast<__bf16>(x[e]);
        hi[e] = h;
        lo[e] = static_cast<__bf16>(x[e] - static_cast<float>(h));
    }
    acc = mfma32(hi, w, acc);
    return mfma32(lo, w, acc);
}

// B operand of a 32-input chunk for output half hf: inputs k0 + 8a .. +7 of
// output 16 hf + c (0 past the layer's `in` inputs).
__device__ __forceinline__ bf8 wchunk(const float* __restrict__ W, int in, int hf, int c, int k0) {
    bf8 b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = k0 + e;
        b[e] = static_cast<__bf16>(k < in ? W[(16 * hf + c) * in + k] : 0.0f);
    }
    return b;
}

__device__ __forceinline__ void load8(const float* __restrict__ row, int a, float (&v)[8]) {
    const float4 p = *reinterpret_cast<const float4*>(row + 8 * a);
    const float4 q = *reinterpret_cast<const float4*>(row + 8 * a + 4);
    v[0] = p.x; v[1] = p.y; v[2] = p.z; v[3] = p.w;
    v[4] = q.x; v[5] = q.y; v[6] = q.z; v[7] = q.w;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

constexpr int kFuseWaves = kBlock / kWave;

// Work items [0, count) for a persistent grid (a multiple of kXcds blocks):
// XCD x walks the x-th contiguous eighth, its waves interleaved, so the rows
// of h that neighbouring items gather stay in that XCD's L2.
struct XcdRange {
    int64_t first, end, step;
};
__device__ __forceinline__ XcdRange xcd_range(int64_t count, int wave) {
    const int per = int(gridDim.x) / kXcds;
    const int x = int(blockIdx.x % kXcds), idx = int(blockIdx.x / kXcds);
    const int64_t span = (count + kXcds - 1) / kXcds;
    const int64_t begin = x * span;
    return XcdRange{begin + int64_t(idx) * kFuseWaves + wave, min(count, begin + span), int64_t(per) * kFuseWaves};
}

// One message-passing layer for chunks of 16 rows, a wave each:
//   m_p = relu(Wm [h_j ; g_p] + bm) for the chunk's edges, 16 a step;
//   the chunk's row sums M += S m with S[r][p] = 1 when edge p is in row r
//   (a matrix-core product too, so every row's messages are reduced in one
//   fixed order), then M_i /= deg_i;
//   hout_i = relu(Wu [h_i ; M_i] + bu).
__global__ __launch_bounds__(kBlock) void k_gnn_layer(int n, const int* __restrict__ erp, const int* __restrict__ ei,
                                                      const int* __restrict__ ej, const float* __restrict__ g,
                                                      const float* __restrict__ h, const float* __restrict__ Wm,
                                                      const float* __restrict__ bm, const float* __restrict__ Wu,
                                                      const float* __restrict__ bu, float* __restrict__ hout) {
    constexpr int INM = H + FE, INU = 2 * H;
    __shared__ float sm[kFuseWaves][16][H + 1];
    __shared__ int srow[kFuseWaves][16];
    const int wv = int(threadIdx.x) / kWave, l = int(threadIdx.x) % kWave;
    const int a = l / 16, c = l % 16;
    bf8 wM[2][2], wU[2][2];
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
        wM[hf][0] = wchunk(Wm, INM, hf, c, 8 * a);
        wM[hf][1] = wchunk(Wm, INM, hf, c, H + 8 * a);
        wU[hf][0] = wchunk(Wu, INU, hf, c, 8 * a);
        wU[hf][1] = wchunk(Wu, INU, hf, c, H + 8 * a);
    }
    const float bm0 = bm[c], bm1 = bm[16 + c], bu0 = bu[c], bu1 = bu[16 + c];
    const XcdRange rg = xcd_range((n + 15) / 16, wv);
    for (int ch = int(rg.first); ch < rg.end; ch += int(rg.step)) {
        const int r0 = ch * 16, rows = min(16, n - r0);
        const int e0 = erp[r0], e1 = erp[r0 + rows];
        mf4 M0 = {0, 0, 0, 0}, M1 = {0, 0, 0, 0};
        for (int t0 = e0; t0 < e1; t0 += 16) {
            const int cnt = min(16, e1 - t0);
            const int e = t0 + min(c, cnt - 1);
            float x[8];
            load8(h + size_t(ej[e]) * H, a, x);
            float gx[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) gx[q] = (a == 0 && q < FE) ? g[size_t(e) * FE + q] : 0.0f;
            if (l < 16) srow[wv][l] = l < cnt ? ei[e] - r0 : -1;
            mf4 acc0 = {bm0, bm0, bm0, bm0}, acc1 = {bm1, bm1, bm1, bm1};
            acc0 = mfma_split(x, wM[0][0], acc0);
            acc1 = mfma_split(x, wM[1][0], acc1);
            acc0 = mfma_split(gx, wM[0][1], acc0);
            acc1 = mfma_split(gx, wM[1][1], acc1);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                sm[wv][4 * a + q][c] = fmaxf(acc0[q], 0.0f);
                sm[wv][4 * a + q][16 + c] = fmaxf(acc1[q], 0.0f);
            }
            wave_lds_sync();
            // M += S m: A = S (row c, edges 8a..8a+7), B = m (edges 8a..+7, channel c / 16 + c);
            // edges 16..31 of the product are zero
            bf8 S;
            float m0[8], m1[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int p = 8 * a + q;
                const bool in = p < 16;
                S[q] = static_cast<__bf16>(in && srow[wv][in ? p : 0] == c ? 1.0f : 0.0f);
                m0[q] = in ? sm[wv][in ? p : 0][c] : 0.0f;
                m1[q] = in ? sm[wv][in ? p : 0][16 + c] : 0.0f;
            }
            bf8 h0, l0, h1, l1;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                h0[q] = static_cast<__bf16>(m0[q]);
                l0[q] = static_cast<__bf16>(m0[q] - static_cast<float>(h0[q]));
                h1[q] = static_cast<__bf16>(m1[q]);
                l1[q] = static_cast<__bf16>(m1[q] - static_cast<float>(h1[q]));
            }
            M0 = mfma32(S, h0, M0);
            M0 = mfma32(S, l0, M0);
            M1 = mfma32(S, h1, M1);
            M1 = mfma32(S, l1, M1);
            wave_lds_sync();
        }
        // M (rows 4a + q, channels c / 16 + c) / deg -> LDS -> the update's A operand
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = min(r0 + 4 * a + q, n - 1);
            const int dg = erp[row + 1] - erp[row];
            const float d = float(dg > 0 ? dg : 1);
            sm[wv][4 * a + q][c] = M0[q] / d;
            sm[wv][4 * a + q][16 + c] = M1[q] / d;
        }
        wave_lds_sync();
        const int r = min(r0 + c, n - 1);
        float x[8], mx[8];
        load8(h + size_t(r) * H, a, x);
#pragma unroll
        for (int q = 0; q < 8; ++q) mx[q] = sm[wv][c][8 * a + q];
        mf4 acc0 = {bu0, bu0, bu0, bu0}, acc1 = {bu1, bu1, bu1, bu1};
        acc0 = mfma_split(x, wU[0][0], acc0);
        acc1 = mfma_split(x, wU[1][0], acc1);
        acc0 = mfma_split(mx, wU[0][1], acc0);
        acc1 = mfma_split(mx, wU[1][1], acc1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = r0 + 4 * a + q;
            if (row < n) {
                hout[size_t(row) * H + c] = fmaxf(acc0[q], 0.0f);
                hout[size_t(row) * H + 16 + c] = fmaxf(acc1[q], 0.0f);
            }
        }
        wave_lds_sync();
    }
}

// The scorer for 16 edges a step per wave: z_p = relu(We [h_i + h_j ;
// |h_i - h_j| ; g_p] + be), w_p = sigmoid(wo . z_p + bo).
__global__ __launch_bounds__(kBlock) void k_gnn_score(int64_t E, const int* __restrict__ ei,
                                                      const int* __restrict__ ej, const float* __restrict__ g,
                                                      const float* __restrict__ h, const float* __restrict__ We,
                                                      const float* __restrict__ be, const float* __restrict__ wo,
                                                      const float* __restrict__ bo, float* __restrict__ w) {
    constexpr int IN = 2 * H + FE;
    const int wv = int(threadIdx.x) / kWave, l = int(threadIdx.x) % kWave;
    const int a = l / 16, c = l % 16;
    bf8 wE[2][3];
#pragma unroll
    for (int hf = 0; hf < 2; ++hf)
#pragma unroll
        for (int k = 0; k < 3; ++k) wE[hf][k] = wchunk(We, IN, hf, c, 32 * k + 8 * a);
    const float be0 = be[c], be1 = be[16 + c], wo0 = wo[c], wo1 = wo[16 + c], b0 = bo[0];
    const int64_t tiles = (E + 15) / 16;
    const XcdRange rg = xcd_range(tiles, wv);
    for (int64_t t = rg.first; t < rg.end; t += rg.step) {
        const int64_t t0 = t * 16;
        const int64_t e = min(t0 + c, E - 1);
        float hi[8], hj[8], xs[8], xd[8], gx[8];
        load8(h + size_t(ei[e]) * H, a, hi);
        load8(h + size_t(ej[e]) * H, a, hj);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            xs[q] = hi[q] + hj[q];
            xd[q] = fabsf(hi[q] - hj[q]);
            gx[q] = (a == 0 && q < FE) ? g[size_t(e) * FE + q] : 0.0f;
        }
        mf4 acc0 = {be0, be0, be0, be0}, acc1 = {be1, be1, be1, be1};
        acc0 = mfma_split(xs, wE[0][0], acc0);
        acc1 = mfma_split(xs, wE[1][0], acc1);
        acc0 = mfma_split(xd, wE[0][1], acc0);
        acc1 = mfma_split(xd, wE[1][1], acc1);
        acc0 = mfma_split(gx, wE[0][2], acc0);
        acc1 = mfma_split(gx, wE[1][2], acc1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float sc = fmaf(wo0, fmaxf(acc0[q], 0.0f), wo1 * fmaxf(acc1[q], 0.0f));
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) sc += __shfl_xor(sc, o, 16);
            const int64_t p = t0 + 4 * a + q;
            if (c == 0 && p < E) w[p] = 1.0f / (1.0f + expf(-(b0 + sc)));
        }
    }
}

// dz_p = ds_p wo masked by z_p > 0
__global__ void k_dz(int64_t E, const float* __restrict__ ds, const float* __restrict__ wo,
                     const float* __restrict__ z, float* __restrict__ dz) {
    const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= E * H) return;
    dz[q] = z[q] > 0.0f ? ds[q / H] * wo[q % H] : 0.0f;
}

// Node gradient from the scorer input gradient dxe: node i collects, over
// the edges q of its row, dxe_q[o] + sgn_q dxe_q[H + o] (as endpoint i of q)
// and dxe_t[o] - sgn_t dxe_t[H + o] of the twin t = (j, i) (as endpoint j of t).
__global__ void k_collect_xe(int n, const int* __restrict__ erp, const int* __restrict__ twin,
                             const int* __restrict__ ei, const int* __restrict__ ej, const float* __restrict__ h,
                             const float* __restrict__ dxe, float* __restrict__ dh) {
    const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= int64_t(n) * H) return;
    constexpr int IN = 2 * H + FE;
    const int i = int(q / H), o = int(q % H);
    float s = 0.0f;
    for (int p = erp[i]; p < erp[i + 1]; ++p) {
        const float hi = h[size_t(i) * H + o], hj = h[size_t(ej[p]) * H + o];
        const float sg = hi > hj ? 1.0f : (hi < hj ? -1.0f : 0.0f);
        s += dxe[size_t(p) * IN + o] + sg * dxe[size_t(p) * IN + H + o];
        const int t = twin[p];  // edge (j, i): i is its second endpoint
        const float ti = h[size_t(ei[t]) * H + o];
        const float sgt = ti > hi ? 1.0f : (ti < hi ? -1.0f : 0.0f);
        s += dxe[size_t(t) * IN + o] - sgt * dxe[size_t(t) * IN + H + o];
    }
    dh[q] = s;
}

// dY masked by Y > 0 (relu), in place.
__global__ void k_relu_mask(int64_t count, const float* __restrict__ Y, float* __restrict__ dY) {
    const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (q < count && !(Y[q] > 0.0f)) dY[q] = 0.0f;
}

// dm_p = dxu_i[H + o] / deg_i masked by m_p > 0; dh_i = dxu_i[o].
__global__ void k_split_xu(int n, const int* __restrict__ erp, const float* __restrict__ dxu,
                           const float* __restrict__ m, float* __restrict__ dm, float* __restrict__ dh) {
    const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= int64_t(n) * H) return;
    const int i = int(q / H), o = int(q % H);
    const int dg = erp[i + 1] - erp[i];
    const float gM = dxu[size_t(i) * 2 * H + H + o] / float(dg > 0 ? dg : 1);
    for (int p = erp[i]; p < erp[i + 1]; ++p) dm[size_t(p) * H + o] = m[size_t(p) * H + o] > 0.0f ? gM : 0.0f;
    dh[q] = dxu[size_t(i) * 2 * H + o];
}

// dh_j += sum over the edges t = (i, j) into j of dxm_t[o]: gathered through
// the twins of row j's edges.
__global__ void k_collect_xm(int n, const int* __restrict__ erp, const int* __restrict__ twin,
                             const float* __restrict__ dxm, float* __restrict__ dh) {
    const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= int64_t(n) * H) return;
    const int j = int(q / H), o = int(q % H);
    float s = dh[q];
    for (int p = erp[j]; p < erp[j + 1]; ++p) s += dxm[size_t(twin[p]) * (H + FE) + o];
    dh[q] = s;
}

// bfloat16 rounding (nearest even) of the master weights.
__global__ void k_quantize(int count, const float* __restrict__ master, float* __restrict__ wq) {
    const int q = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (q >= count) return;
    const float x = master[q];
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7F800000u) != 0x7F800000u) u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    wq[q] = __uint_as_float(u);
}

// Adam (Kingma & Ba 2015) on the master weights.
__global__ void k_adam(int count, float* __restrict__ master, const float* __restrict__ grad, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int q = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (q >= count) return;
    const float g = grad[q];
    const float a = b1 * m1[q] + (1.0f - b1) * g;
    const float b = b2 * m2[q] + (1.0f - b2) * g * g;
    m1[q] = a;
    m2[q] = b;
    master[q] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// Affinities: c_ij = (x_i . x_j)^2 / ((x_i . x_i)(x_j . x_j)) over the k
// columns of a row-major block, per stored entry (0 on the diagonal).
__global__ void k_affinity(int n, int k, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ X, float* __restrict__ out) {
    const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (i >= n) return;
    double ni = 0;
    for (int q = 0; q < k; ++q) ni += X[size_t(i) * k + q] * X[size_t(i) * k + q];
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        const int j = ci[p];
        if (j == i) {
            out[p] = 0.0f;
            continue;
        }
        double nj = 0, d = 0;
        for (int q = 0; q < k; ++q) {
            nj += X[size_t(j) * k + q] * X[size_t(j) * k + q];
            d += X[size_t(i) * k + q] * X[size_t(j) * k + q];
        }
        out[p] = float(d * d / (ni * nj));
    }
}

// Scatter of the edge weights to the matrix positions (0 on the diagonal).
__global__ void k_to_positions(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                               const int* __restrict__ erp, const float* __restrict__ w, double* __restrict__ out) {
    const RowLane r = row_lane();
    const bool live = r.i < n;
    const int i = live ? r.i : n - 1;
    int at = erp[i];
    const int e = rp[i + 1];
    for (int j0 = rp[i]; j0 < e; j0 += kRowLanes) {
        const int j = j0 + r.lane;
        const bool valid = j < e;
        const bool off = valid && ci[j] != i;
        const unsigned bits = group_bits(off, r);
        if (live && valid) out[j] = off ? double(w[at + __popc(bits & ((1u << r.lane) - 1))]) : 0.0;
        at += __popc(bits);
    }
}

// Edge targets from matrix-position targets.
__global__ void k_from_positions(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                 const int* __restrict__ erp, const float* __restrict__ t, float* __restrict__ out) {
    const RowLane r = row_lane();
    const bool live = r.i < n;
    const int i = live ? r.i : n - 1;
    int at = erp[i];
    const int e = rp[i + 1];
    for (int j0 = rp[i]; j0 < e; j0 += kRowLanes) {
        const int j = j0 + r.lane;
        const bool valid = j < e;
        const bool off = valid && ci[j] != i;
        const unsigned bits = group_bits(off, r);
        if (live && off) out[at + __popc(bits & ((1u << r.lane) - 1))] = t[j];
        at += __popc(bits);
    }
}

unsigned blocks_of(int64_t count) { return unsigned((count + kBlock - 1) / kBlock); }

}  // namespace

// ------------------------------------------------------------ the model

struct GnnModel::Impl {
    Layout L;
    tf_ctx* c;
    float *master = nullptr, *wq = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    float* wpart = nullptr;     // weight-gradient partials
    double* lpart = nullptr;    // loss partials
    int steps = 0;
    // graph workspace
    int n = 0;
    int64_t E = 0, capN = 0, capE = 0;
    int *erp = nullptr, *ei = nullptr, *ej = nullptr, *twin = nullptr, *cnt = nullptr;
    double *ea = nullptr, *dabs = nullptr, *mrow = nullptr;
    float *f = nullptr, *deg = nullptr, *g = nullptr;
    // activations and gradients
    float* h[kLayers + 1] = {};
    float* xm[kLayers] = {};
    float* m[kLayers] = {};
    float* xu[kLayers] = {};
    float *xe = nullptr, *z = nullptr, *w = nullptr, *t = nullptr, *ds = nullptr;
    float *dE = nullptr, *dE2 = nullptr, *dN = nullptr, *dN2 = nullptr, *dNu = nullptr;
    std::vector<void*> owned;
    bool capTrain = false;  // training (gradient) buffers allocated

    float* falloc(int64_t count) {
        float* p = dev_alloc<float>(*c, size_t(std::max<int64_t>(count, 1)));
        owned.push_back(p);
        return p;
    }
    void free_graph() {
        (void)hipStreamSynchronize(c->stream);
        for (void* p : owned) dev_free(*c, p);
        owned.clear();
        capN = capE = 0;
        capTrain = false;
    }
    void reserve(int nn, int64_t ee, bool training) {
        if (nn <= capN && ee <= capE && (capTrain || !training)) return;
        free_graph();
        capTrain = training;
        capN = std::max<int64_t>(nn, 1);
        capE = std::max<int64_t>(ee, 1);
        auto ialloc = [&](int64_t count) {
            int* p = dev_alloc<int>(*c, size_t(std::max<int64_t>(count, 1)));
            owned.push_back(p);
            return p;
        };
        auto dalloc = [&](int64_t count) {
            double* p = dev_alloc<double>(*c, size_t(std::max<int64_t>(count, 1)));
            owned.push_back(p);
            return p;
        };
        erp = ialloc(capN + 1);
        cnt = ialloc(capN + 1);
        ei = ialloc(capE);
        ej = ialloc(capE);
        twin = ialloc(capE);
        ea = dalloc(capE);
        dabs = dalloc(capN);
        mrow = dalloc(capN);
        f = falloc(capN * FN);
        deg = falloc(capN);
        g = falloc(capE * FE);
        for (int l = 0; l <= kLayers; ++l) h[l] = falloc(capN * H);
        w = falloc(capE);
        if (!training) return;  // inference keeps no edge activations
        for (int l = 0; l < kLayers; ++l) {
            xm[l] = falloc(capE * (H + FE));
            m[l] = falloc(capE * H);
            xu[l] = falloc(capN * 2 * H);
        }
        xe = falloc(capE * (2 * H + FE));
        z = falloc(capE * H);
        t = falloc(capE);
        ds = falloc(capE);
        dE = falloc(capE * (2 * H + FE));
        dE2 = falloc(capE * H);
        dN = falloc(capN * 2 * H);
        dN2 = falloc(capN * H);
        dNu = falloc(capN * H);
    }

    void graph(int nn, int64_t nnz, const int* rp, const int* ci, const double* v, bool training) {
        require(nn >= 1, "the network needs a nonempty matrix");
        // edge count
        int* tmp = dev_alloc<int>(*c, size_t(nn) + 1);
        hipLaunchKernelGGL(k_offdiag_count, dim3(blocks_of((int64_t(nn) + 1) * kRowLanes)), dim3(kBlock), 0,
                           c->stream, nn, rp, ci, tmp);
        int* trp = dev_alloc<int>(*c, size_t(nn) + 1);
        scan_excl(*c, nn + 1, tmp, trp);
        const int64_t ee = c->read_int(trp + nn);
        (void)nnz;
        reserve(nn, ee, training);
        check(hipMemcpyAsync(erp, trp, (size_t(nn) + 1) * sizeof(int), hipMemcpyDeviceToDevice, c->stream),
              "edge offsets copy");
        c->sync();
        dev_free(*c, tmp);
        dev_free(*c, trp);
        n = nn;
        E = ee;
        hipLaunchKernelGGL(k_edges_nodes, dim3(blocks_of(int64_t(n) * kRowLanes)), dim3(kBlock), 0, c->stream, n, rp,
                           ci, v, erp, ei, ej,
                           ea, dabs, mrow, f, deg);
        c->clear_flags(1);
        if (E > 0)
            hipLaunchKernelGGL(k_twins, dim3(blocks_of(E)), dim3(kBlock), 0, c->stream, E, erp, ei, ej, twin,
                               c->dflags, ea, dabs, mrow, g);
        check(hipGetLastError(), "graph build launch");
        c->read_flags(1);
        require(c->hflags[0] == 0, "learned affinities need a matrix with a symmetric pattern");
    }

    void dense(int64_t rows, int in, const float* X, int Woff, int boff, float* Y) {
        if (rows <= 0) return;
        hipLaunchKernelGGL(k_dense, dim3(unsigned((rows + kTileR - 1) / kTileR)), dim3(kBlock), 0, c->stream, rows, in,
                           X, wq + Woff, wq + boff, Y);
    }
    void dense_t(int64_t rows, int in, const float* dY, int Woff, float* dX) {
        if (rows <= 0) return;
        hipLaunchKernelGGL(k_dense_t, dim3(unsigned((rows + kTileR - 1) / kTileR)), dim3(kBlock), 0, c->stream, rows,
                           in, dY, wq + Woff, dX);
    }
    void wgrad(int64_t rows, int out, int in, const float* dY, const float* X, int Woff, int boff) {
        const int64_t tiles = (rows + kTileR - 1) / kTileR;
        const int blocks = int(std::max<int64_t>(1, std::min<int64_t>(tiles, kGradBlocks)));
        hipLaunchKernelGGL(k_wgrad, dim3(unsigned(blocks)), dim3(kBlock), 0, c->stream, rows, out, in, dY, X, wpart);
        hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_of(int64_t(out) * (in + 1))), dim3(kBlock), 0, c->stream, out,
                           in, blocks, wpart, grad + Woff, grad + boff);
    }

    // forward pass; with target, also the loss and ds; returns nothing
    void forward(const float* target) {
        dense(n, FN, f, L.W0, L.b0, h[0]);
        for (int l = 0; l < kLayers; ++l) {
            if (E > 0) {
                hipLaunchKernelGGL(k_gather_xm, dim3(blocks_of(E * (H + FE))), dim3(kBlock), 0, c->stream, E, h[l], ej,
                                   g, xm[l]);
                dense(E, H + FE, xm[l], L.Wm[l], L.bm[l], m[l]);
            }
            hipLaunchKernelGGL(k_node_mean, dim3(blocks_of(int64_t(n) * H)), dim3(kBlock), 0, c->stream, n, erp, m[l],
                               h[l], xu[l]);
            dense(n, 2 * H, xu[l], L.Wu[l], L.bu[l], h[l + 1]);
        }
        if (E > 0) {
            hipLaunchKernelGGL(k_gather_xe, dim3(blocks_of(E * (2 * H + FE))), dim3(kBlock), 0, c->stream, E,
                               h[kLayers], ei, ej, g, xe);
            dense(E, 2 * H + FE, xe, L.We, L.be, z);
            hipLaunchKernelGGL(k_score, dim3(blocks_of(E)), dim3(kBlock), 0, c->stream, E, z, wq + L.wo, wq + L.bo, w,
                               target, ds, lpart);
        }
        check(hipGetLastError(), "network forward launch");
    }

    // inference: the fused layer and scorer kernels
    void forward_infer() {
        dense(n, FN, f, L.W0, L.b0, h[0]);
        const int chunks = (n + 15) / 16;
        const unsigned lb = unsigned(round_xcd(std::min(2048, (chunks + kFuseWaves - 1) / kFuseWaves)));
        for (int l = 0; l < kLayers; ++l)
            hipLaunchKernelGGL(k_gnn_layer, dim3(lb), dim3(kBlock), 0, c->stream, n, erp, ei, ej, g, h[l],
                               wq + L.Wm[l], wq + L.bm[l], wq + L.Wu[l], wq + L.bu[l], h[l + 1]);
        if (E > 0) {
            const int64_t tiles = (E + 15) / 16;
            const unsigned sb = unsigned(round_xcd(int(std::min<int64_t>(8192, (tiles + kFuseWaves - 1) / kFuseWaves))));
            hipLaunchKernelGGL(k_gnn_score, dim3(sb), dim3(kBlock), 0, c->stream, E, ei, ej, g, h[kLayers],
                               wq + L.We, wq + L.be, wq + L.wo, wq + L.bo, w);
        }
        check(hipGetLastError(), "network inference launch");
    }

    double loss_value() {
        const unsigned nb = blocks_of(E);
        std::vector<double> host(nb);
        check(hipMemcpyAsync(host.data(), lpart, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream),
              "loss download");
        c->sync();
        double s = 0;
        for (double x : host) s += x;
        return s / double(E);
    }

    void backward() {
        fill_zero(*c, grad, size_t(L.count) * sizeof(float));
        // scorer: wo, bo from ds and z; dz; We, be; dxe -> dh
        wgrad(E, 1, H, ds, z, L.wo, L.bo);
        hipLaunchKernelGGL(k_dz, dim3(blocks_of(E * H)), dim3(kBlock), 0, c->stream, E, ds, wq + L.wo, z, dE2);
        wgrad(E, H, 2 * H + FE, dE2, xe, L.We, L.be);
        dense_t(E, 2 * H + FE, dE2, L.We, dE);
        hipLaunchKernelGGL(k_collect_xe, dim3(blocks_of(int64_t(n) * H)), dim3(kBlock), 0, c->stream, n, erp, twin, ei,
                           ej, h[kLayers], dE, dN2);
        for (int l = kLayers - 1; l >= 0; --l) {
            // dN2 = dL/dh[l+1]
            hipLaunchKernelGGL(k_relu_mask, dim3(blocks_of(int64_t(n) * H)), dim3(kBlock), 0, c->stream,
                               int64_t(n) * H, h[l + 1], dN2);
            wgrad(n, H, 2 * H, dN2, xu[l], L.Wu[l], L.bu[l]);
            dense_t(n, 2 * H, dN2, L.Wu[l], dN);
            hipLaunchKernelGGL(k_split_xu, dim3(blocks_of(int64_t(n) * H)), dim3(kBlock), 0, c->stream, n, erp, dN,
                               m[l], dE2, dNu);
            if (E > 0) {
                wgrad(E, H, H + FE, dE2, xm[l], L.Wm[l], L.bm[l]);
                dense_t(E, H + FE, dE2, L.Wm[l], dE);
                hipLaunchKernelGGL(k_collect_xm, dim3(blocks_of(int64_t(n) * H)), dim3(kBlock), 0, c->stream, n, erp,
                                   twin, dE, dNu);
            }
            std::swap(dN2, dNu);  // dN2 = dL/dh[l]
        }
        hipLaunchKernelGGL(k_relu_mask, dim3(blocks_of(int64_t(n) * H)), dim3(kBlock), 0, c->stream, int64_t(n) * H,
                           h[0], dN2);
        wgrad(n, H, FN, dN2, f, L.W0, L.b0);
        check(hipGetLastError(), "network backward launch");
    }

    void quantize() {
        hipLaunchKernelGGL(k_quantize, dim3(blocks_of(L.count)), dim3(kBlock), 0, c->stream, L.count, master, wq);
    }
};

GnnModel::GnnModel(tf_ctx* ctx) : p_(new Impl) {
    p_->c = ctx;
    const int cnt = p_->L.count;
    p_->master = dev_alloc<float>(*ctx, size_t(cnt));
    p_->wq = dev_alloc<float>(*ctx, size_t(cnt));
    p_->grad = dev_alloc<float>(*ctx, size_t(cnt));
    p_->m1 = dev_alloc<float>(*ctx, size_t(cnt));
    p_->m2 = dev_alloc<float>(*ctx, size_t(cnt));
    p_->wpart = dev_alloc<float>(*ctx, size_t(kGradBlocks) * H * (kMaxIn + 1));
    p_->lpart = dev_alloc<double>(*ctx, size_t(1) << 22);
    for (float* q : {p_->master, p_->wq, p_->grad, p_->m1, p_->m2}) fill_zero(*ctx, q, size_t(cnt) * sizeof(float));
    ctx->sync();
}

GnnModel::~GnnModel() {
    tf_ctx& c = *p_->c;
    p_->free_graph();
    for (void* q : {static_cast<void*>(p_->master), static_cast<void*>(p_->wq), static_cast<void*>(p_->grad),
                    static_cast<void*>(p_->m1), static_cast<void*>(p_->m2), static_cast<void*>(p_->wpart),
                    static_cast<void*>(p_->lpart)})
        dev_free(c, q);
    delete p_;
}

int GnnModel::param_count() { return Layout().count; }

tf_ctx* GnnModel::context() const { return p_->c; }

void GnnModel::set_params(const float* host) {
    check(hipMemcpyAsync(p_->master, host, size_t(p_->L.count) * sizeof(float), hipMemcpyHostToDevice, p_->c->stream),
          "parameter upload");
    fill_zero(*p_->c, p_->m1, size_t(p_->L.count) * sizeof(float));
    fill_zero(*p_->c, p_->m2, size_t(p_->L.count) * sizeof(float));
    p_->steps = 0;
    p_->quantize();
    p_->c->sync();
}

void GnnModel::get_params(float* host) {
    check(hipMemcpyAsync(host, p_->master, size_t(p_->L.count) * sizeof(float), hipMemcpyDeviceToHost, p_->c->stream),
          "parameter download");
    p_->c->sync();
}

void GnnModel::infer(int n, int64_t nnz, const int* rp, const int* ci, const double* v, double* out) {
    p_->graph(n, nnz, rp, ci, v, false);
    p_->forward_infer();
    hipLaunchKernelGGL(k_to_positions, dim3(blocks_of(int64_t(n) * kRowLanes)), dim3(kBlock), 0, p_->c->stream, n, rp,
                       ci, p_->erp, p_->w, out);
    check(hipGetLastError(), "edge weight scatter launch");
}

double GnnModel::loss_grad(int n, int64_t nnz, const int* rp, const int* ci, const double* v, const float* target,
                           float* hostGrad) {
    require(nnz >= 0, "negative nonzero count");
    p_->graph(n, nnz, rp, ci, v, true);
    hipLaunchKernelGGL(k_from_positions, dim3(blocks_of(int64_t(n) * kRowLanes)), dim3(kBlock), 0, p_->c->stream, n,
                       rp, ci, p_->erp, target, p_->t);
    require(p_->E > 0 && p_->E <= (int64_t(1) << 22) * kBlock, "training needs 1 .. 2^30 edges");
    p_->forward(p_->t);
    const double l = p_->loss_value();
    p_->backward();
    if (hostGrad) {
        check(hipMemcpyAsync(hostGrad, p_->grad, size_t(p_->L.count) * sizeof(float), hipMemcpyDeviceToHost,
                             p_->c->stream),
              "gradient download");
        p_->c->sync();
    }
    return l;
}

double GnnModel::train_step(int n, int64_t nnz, const int* rp, const int* ci, const double* v, const float* target,
                            float lr) {
    const double l = loss_grad(n, nnz, rp, ci, v, target, nullptr);
    const float b1 = 0.9f, b2 = 0.999f;
    ++p_->steps;
    const float c1 = 1.0f - std::pow(b1, float(p_->steps)), c2 = 1.0f - std::pow(b2, float(p_->steps));
    hipLaunchKernelGGL(k_adam, dim3(blocks_of(p_->L.count)), dim3(kBlock), 0, p_->c->stream, p_->L.count, p_->master,
                       p_->grad, p_->m1, p_->m2, lr, b1, b2, 1e-8f, c1, c2);
    p_->quantize();
    check(hipGetLastError(), "adam launch");
    p_->c->sync();
    return l;
}

// ------------------------------------------------------------ affinities

void affinities(tf_ctx& c, int n, int64_t nnz, const int* rp, const int* ci, const double* v, int k, int sweeps,
                double omega, const double* X0, float* out) {
    require(n >= 1 && k >= 1 && sweeps >= 0, "affinities need n >= 1, k >= 1, sweeps >= 0");
    double* inv = dev_alloc<double>(c, size_t(n));
    double* X = dev_alloc<double>(c, size_t(n) * size_t(k));
    double* Y = dev_alloc<double>(c, size_t(n) * size_t(k));
    double* Z = dev_alloc<double>(c, size_t(n) * size_t(k));  // the zero right-hand side
    fill_zero(c, Z, size_t(n) * size_t(k) * sizeof(double));
    c.clear_flags(1);
    inverse_diagonal<double>(c, n, rp, ci, v, inv, c.dflags);
    c.read_flags(1);
    const bool bad = c.hflags[0] != 0;
    if (!bad) {
        convert<double, double>(c, size_t(n) * size_t(k), X0, X);
        for (int s = 0; s < sweeps; ++s) {
            block_jacobi<double>(c, n, rp, ci, v, inv, omega, k, Z, X, Y);
            std::swap(X, Y);
        }
        hipLaunchKernelGGL(k_affinity, dim3(blocks_of(n)), dim3(kBlock), 0, c.stream, n, k, rp, ci, X, out);
        check(hipGetLastError(), "affinity launch");
    }
    c.sync();
    dev_free(c, inv);
    dev_free(c, X);
    dev_free(c, Y);
    dev_free(c, Z);
    (void)nnz;
    require(!bad, "affinities need a nonzero finite diagonal");
}

}  // namespace tf
