// Coarse-level dense solver setup: partial-pivoting LU in FP64 with the
// pivot rule of GpuAggregation.cu:827-893 (first maximal |a_ik| over
// i >= k, rejected when non-finite or <= 64 eps max|A|; whole rows swapped;
// unit-lower multipliers stored below the diagonal), blocked in 16-column
// panels with the trailing updates on the FP64 matrix cores, and the
// explicit inverse A^-1 = U^-1 L^-1 P by blocked substitution with one
// workgroup per 16 columns of P I, whose off-diagonal block updates also
// run on the matrix cores.
//
// Factorisation (one 512-thread workgroup, matrix row-major in global
// memory, resident in L2): per panel of 16 columns
//   1. the panel (rows kb..n-1) goes to LDS column-major;
//   2. one wave factors it there column by column (pivot search as a DPP
//      wave argmax, row swap, scaling, rank-1 update inside the panel), its
//      first lane composing the panel's row permutation;
//   3. every other column gets the panel's permutation (one gather of the
//      touched rows) and, right of the panel, U12 = L11^-1 A12 by
//      substitution in row order;
//   4. A22 -= L21 U12 over 16x16 tiles with v_mfma_f64_16x16x4f64 (K = 16 in
//      four steps), one wave per strip of eight tiles.
// Each element's updates are applied in pivot order as in the unblocked
// algorithm; inside a trailing update the 16 products are accumulated by the
// matrix core before the subtraction, so factors agree with the sequential
// elimination to rounding.
//
// Orders up to kLuSmall (the coarsest levels of typical hierarchies) take a
// latency-lean path instead: one 768-thread workgroup, a group of kColLanes
// lanes holding each column in registers, and the unblocked elimination with
// one workgroup barrier per pivot (the pivot column's group searches, swaps
// and divides, publishing the multipliers in LDS; every other group then
// swaps and updates its own column).  The inverse columns are solved by the
// same groups in registers, reading the factors from LDS, with no barrier at
// all (a column's lanes share a wave).  Its
// arithmetic is the sequential algorithm's operation for operation, so the
// factors and inverse columns equal the unblocked elimination's bit for bit.
#include "tf_device.hpp"
#include "tf_internal.hpp"

namespace tf {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kNb = 16;           // panel width = matrix-core tile edge
constexpr int kFactorThreads = 512;
constexpr int kInvThreads = 256;  // one workgroup per 16 inverse columns
constexpr int kTouched = 2 * kNb; // rows a panel's permutation can move
constexpr int kStrip = 8;         // trailing-update tiles per wave and round
constexpr int kLuSmall = 96;      // orders factored and inverted by one workgroup (factors in LDS, 72 KB)
constexpr int kSmallThreads = 768;  // 12 waves: one lane group per column of order kLuSmall
constexpr int kColLanes = 8;      // lanes of a column group (kSmallThreads / kColLanes >= kLuSmall)
constexpr int kSlots = kLuSmall / kColLanes;  // rows of a column per lane
static_assert(kSmallThreads / kColLanes >= kLuSmall, "one lane group per column");

// acc(16x16) += A(16x4) B(4x16): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds acc[4 q + l / 16][l % 16] in acc[q] (lane
// layout checked on gfx950 with distinct operand values).
__device__ __forceinline__ d4 mfma_16x16x4(double a, double b, d4 acc) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}

// One step of a wave argmax over (value, row) pairs through DPP: larger
// value wins, ties go to the smaller row, row < 0 means no candidate (the
// operation is associative and commutative, so the reduction order does not
// change the result).  Lanes outside RowMask keep their pair.
template <int Ctrl, int RowMask>
__device__ __forceinline__ void argmax_dpp(double& v, int& i) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(int(b), int(b), Ctrl, RowMask, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(int(b >> 32), int(b >> 32), Ctrl, RowMask, 0xf, false);
    const int oi = __builtin_amdgcn_update_dpp(i, i, Ctrl, RowMask, 0xf, false);
    const double ov = __longlong_as_double(
        static_cast<long long>((static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                               static_cast<unsigned>(lo)));
    const bool take = (oi >= 0) & ((i < 0) | (ov > v) | ((ov == v) & (oi < i)));
    v = take ? ov : v;
    i = take ? oi : i;
}

// Wave argmax: quad exchanges (xor 1, xor 2), half-row and row mirrors,
// then row broadcasts 15 and 31 leave the result in lane 63.
__device__ __forceinline__ int wave_argmax(double v, int i) {
    argmax_dpp<0xB1, 0xF>(v, i);   // quad_perm [1,0,3,2]
    argmax_dpp<0x4E, 0xF>(v, i);   // quad_perm [2,3,0,1]
    argmax_dpp<0x141, 0xF>(v, i);  // row_half_mirror
    argmax_dpp<0x140, 0xF>(v, i);  // row_mirror
    argmax_dpp<0x142, 0xA>(v, i);  // row_bcast:15 into rows 1 and 3
    argmax_dpp<0x143, 0xC>(v, i);  // row_bcast:31 into rows 2 and 3
    return __builtin_amdgcn_readlane(i, kWave - 1);
}

// Orders this wave's LDS accesses before and after (lanes exchange data
// through LDS inside the single-wave panel factorisation).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A zero the compiler cannot prove uniform: LDS indices offset by it load
// into vector registers, which keeps the many wave-uniform coefficients of
// the unrolled permutation and substitution loops out of the scalar
// register file (where they would be spilled inside the pivot loop).
__device__ __forceinline__ int lane_zero() {
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

__global__ void k_dense_fill(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                             const double* __restrict__ v, double* __restrict__ m) {
    const int r = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (r >= n) return;
    for (int j = 0; j < n; ++j) m[size_t(r) * n + j] = 0.0;
    for (int j = rp[r]; j < rp[r + 1]; ++j) m[size_t(r) * n + ci[j]] = v[j];
}

__global__ __launch_bounds__(kFactorThreads) void k_lu_factor(int n, double* __restrict__ m, int* __restrict__ piv,
                                                              int* bad) {
    __shared__ double pan[kCoarseMax * kNb];  // panel, column-major, leading dimension n
    __shared__ int src[kCoarseMax];           // panel-relative row now holding original row src[i]
    __shared__ int tl[kTouched], ts[kTouched];
    __shared__ int ppiv[kNb];
    __shared__ double sink[kWave];           // discarded stores of the panel update
    __shared__ double sv[kFactorThreads / kWave];
    __shared__ double scale;
    __shared__ int failed, touched;
    const int tid = int(threadIdx.x);
    const int lane = tid & (kWave - 1), w = tid / kWave;
    const int z = lane_zero();
    double mx = 0;
    for (int i = tid; i < n * n; i += kFactorThreads) mx = fmax(mx, fabs(m[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, kWave));
    if (lane == 0) sv[w] = mx;
    __syncthreads();
    if (tid == 0) {
        double s = 0;
        for (int i = 0; i < kFactorThreads / kWave; ++i) s = fmax(s, sv[i]);
        scale = s;
        failed = 0;
    }
    __syncthreads();
    const double thr = 64 * 2.2204460492503131e-16 * scale;
    for (int kb = 0; kb < n; kb += kNb) {
        const int nb = min(kNb, n - kb), rows = n - kb;
        for (int idx = tid; idx < rows * nb; idx += kFactorThreads) {
            const int i = idx / nb, c = idx % nb;
            pan[c * n + i] = m[size_t(kb + i) * n + kb + c];
        }
        for (int i = tid; i < rows; i += kFactorThreads) src[i] = i;
        __syncthreads();
        if (w == 0) {
            for (int c = 0; c < nb; ++c) {
                // first maximal |a_ic| over i >= c (strict comparisons skip
                // NaN); a NaN diagonal keeps p = c
                double bv = -1.0;
                int bi = -1;
                for (int i = c + lane; i < rows; i += kWave) {
                    const double a = fabs(pan[c * n + i]);
                    const bool take = a > bv;
                    bv = take ? a : bv;
                    bi = take ? i : bi;
                }
                int p = wave_argmax(bv, bi);
                if (p < 0 || isnan(pan[c * n + c])) p = c;
                const double pv = pan[c * n + p];
                if (!isfinite(pv) || fabs(pv) <= thr) {
                    if (lane == 0) failed = 1;
                    break;
                }
                if (lane == 0) {
                    ppiv[c] = kb + p;
                    const int t = src[c];
                    src[c] = src[p];
                    src[p] = t;
                }
                if (p != c && lane < nb) {
                    const double t = pan[lane * n + c];
                    pan[lane * n + c] = pan[lane * n + p];
                    pan[lane * n + p] = t;
                }
                wave_sync();
                // scale and rank-1 update without branches: row c's entries
                // in registers, each row's 16 loads issued together (column
                // indices clamped into the panel), the stores of columns <= c
                // redirected to this lane's sink slot
                const double d = pan[c * n + c];
                double u[kNb];
#pragma unroll
                for (int j = 0; j < kNb; ++j) u[j] = pan[min(j, nb - 1) * n + c + z];
                for (int i = c + 1 + lane; i < rows; i += kWave) {
                    const double l = pan[c * n + i] / d;
                    double a[kNb];
#pragma unroll
                    for (int j = 0; j < kNb; ++j) a[j] = pan[min(j, nb - 1) * n + i];
                    pan[c * n + i] = l;
#pragma unroll
                    for (int j = 0; j < kNb; ++j) {
                        const bool on = (j > c) & (j < nb);
                        double* dst = on ? pan + j * n + i : sink + lane;
                        *dst = a[j] - l * u[j];
                    }
                }
                wave_sync();
            }
            wave_sync();
            // rows the permutation moved: the panel's top rows, then the
            // distinct pivot rows below them
            if (lane < kNb) {
                const int c = lane;
                int below = 0;
                if (c < nb && !failed) {
                    const int p = ppiv[c] - kb;
                    below = p >= nb ? 1 : 0;
                    for (int q = 0; q < c; ++q) below = ppiv[q] - kb == p ? 0 : below;
                }
                const unsigned long long mask = __ballot(below);
                if (c < nb) {
                    tl[c] = c;
                    ts[c] = src[c];
                    if (!failed) piv[kb + c] = ppiv[c];
                }
                if (below) {
                    const int slot = nb + __popcll(mask & ((1ull << c) - 1));
                    const int p = ppiv[c] - kb;
                    tl[slot] = p;
                    ts[slot] = src[p];
                }
                if (c == 0) touched = nb + __popcll(mask);
            }
        }
        __syncthreads();
        if (failed) {
            if (tid == 0) *bad = 1;
            return;
        }
        for (int idx = tid; idx < rows * nb; idx += kFactorThreads) {
            const int i = idx / nb, c = idx % nb;
            m[size_t(kb + i) * n + kb + c] = pan[c * n + i];
        }
        // the other columns: permuted rows (all loads before the stores),
        // then right of the panel U12 = L11^-1 A12 in row order
        const int cnt = touched;
        for (int j = tid; j < n - nb; j += kFactorThreads) {
            const int col = j < kb ? j : j + nb;
            double t[kTouched];
#pragma unroll
            for (int q = 0; q < kTouched; ++q)
                t[q] = q < cnt && ts[q + z] != tl[q + z] ? m[size_t(kb + ts[q + z]) * n + col] : 0.0;
#pragma unroll
            for (int q = 0; q < kTouched; ++q)
                if (q < cnt && ts[q + z] != tl[q + z]) m[size_t(kb + tl[q + z]) * n + col] = t[q];
            if (col < kb) continue;
            double u[kNb];
#pragma unroll
            for (int r = 0; r < kNb; ++r) u[r] = r < nb ? m[size_t(kb + r) * n + col] : 0.0;
#pragma unroll
            for (int r = 1; r < kNb; ++r)
#pragma unroll
                for (int c = 0; c < r; ++c)
                    if (r < nb) u[r] -= pan[c * n + r + z] * u[c];
#pragma unroll
            for (int r = 1; r < kNb; ++r)
                if (r < nb) m[size_t(kb + r) * n + col] = u[r];
        }
        __syncthreads();
        // A22 -= L21 U12 on the matrix cores: a wave takes a strip of kStrip
        // 16x16 tiles side by side (one A operand), all loads before the
        // products
        const int m2 = n - kb - nb;
        const int tiles = (m2 + 15) / 16, strips = (tiles + kStrip - 1) / kStrip;
        const int li = lane & 15, lk = lane >> 4;
        for (int t = w; t < tiles * strips; t += kFactorThreads / kWave) {
            const int r0 = kb + nb + (t / strips) * 16, cb = kb + nb + (t % strips) * 16 * kStrip;
            double a[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int kk = 4 * s + lk, ar = r0 + li;
                a[s] = kk < nb && ar < n ? -pan[kk * n + (ar - kb)] : 0.0;
            }
            d4 acc[kStrip];
            double bm[kStrip][4];
#pragma unroll
            for (int v = 0; v < kStrip; ++v) {
                const int col = cb + 16 * v + li;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = r0 + 4 * q + lk;
                    acc[v][q] = r < n && col < n ? m[size_t(r) * n + col] : 0.0;
                }
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int kk = 4 * s + lk;
                    bm[v][s] = kk < nb && col < n ? m[size_t(kb + kk) * n + col] : 0.0;
                }
            }
#pragma unroll
            for (int v = 0; v < kStrip; ++v)
#pragma unroll
                for (int s = 0; s < 4; ++s) acc[v] = mfma_16x16x4(a[s], bm[v][s], acc[v]);
#pragma unroll
            for (int v = 0; v < kStrip; ++v) {
                const int col = cb + 16 * v + li;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = r0 + 4 * q + lk;
                    if (r < n && col < n) m[size_t(r) * n + col] = acc[v][q];
                }
            }
        }
        __syncthreads();
    }
    int nonfinite = 0;
    for (int i = tid; i < n * n; i += kFactorThreads)
        if (!isfinite(m[i])) nonfinite = 1;
    if (nonfinite) atomicOr(bad, 1);
}

// Columns j0..j0+15 of A^-1: X = P I (the factorisation's row swaps in
// order), L Y = X and U Z = Y by 16-row blocks.  The diagonal blocks are
// solved by one thread per column in the unblocked algorithm's order (rows
// ascending below / descending above, each update then the division); the
// blocks off the diagonal are applied as 16x16x16 matrix-core products.
__global__ __launch_bounds__(kInvThreads) void k_lu_inverse(int n, const double* __restrict__ lu,
                                                            const int* __restrict__ piv, double* __restrict__ inv,
                                                            const int* failed) {
    if (failed && *failed) return;  // (queued without a host check: no valid factors)
    __shared__ double X[kCoarseMax * kNb];  // row-major, 16 columns
    __shared__ int pv[kCoarseMax];
    const int tid = int(threadIdx.x);
    const int lane = tid & (kWave - 1), w = tid / kWave;
    const int li = lane & 15, lk = lane >> 4;
    const int j0 = int(blockIdx.x) * kNb;
    const int live = min(kNb, n - j0);
    for (int i = tid; i < n; i += kInvThreads) pv[i] = piv[i];
    for (int i = tid; i < n * kNb; i += kInvThreads) X[i] = 0.0;
    __syncthreads();
    if (tid < live) {
        // where the unit entry of column j0 + tid ends up after the swaps
        int q = j0 + tid;
        for (int k = 0; k < n; ++k) {
            const int p = pv[k];
            q = q == k ? p : (q == p ? k : q);
        }
        X[q * kNb + tid] = 1.0;
    }
    __syncthreads();
    const int blocks = (n + kNb - 1) / kNb;
    for (int b = 0; b < blocks; ++b) {
        const int rb = b * kNb, bs = min(kNb, n - rb);
        if (tid < kNb) {
            double x[kNb];
#pragma unroll
            for (int r = 0; r < kNb; ++r) x[r] = r < bs ? X[(rb + r) * kNb + tid] : 0.0;
#pragma unroll
            for (int r = 1; r < kNb; ++r)
#pragma unroll
                for (int c = 0; c < r; ++c)
                    if (r < bs) x[r] -= lu[size_t(rb + r) * n + rb + c] * x[c];
#pragma unroll
            for (int r = 1; r < kNb; ++r)
                if (r < bs) X[(rb + r) * kNb + tid] = x[r];
        }
        __syncthreads();
        // rows below, four 16-row tiles per wave and round (all loads first)
        const int tiles = (n - rb - bs + 15) / 16;
        double bx[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int kk = 4 * s + lk;
            bx[s] = kk < bs ? X[(rb + kk) * kNb + li] : 0.0;
        }
        for (int t0 = 4 * w; t0 < tiles; t0 += kInvThreads / kWave * 4) {
            d4 acc[4];
            double a[4][4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r0 = rb + bs + 16 * (t0 + v);
                const bool on = t0 + v < tiles;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = r0 + 4 * q + lk;
                    acc[v][q] = on && r < n ? X[r * kNb + li] : 0.0;
                }
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int kk = 4 * s + lk, ar = r0 + li;
                    a[v][s] = on && kk < bs && ar < n ? -lu[size_t(ar) * n + rb + kk] : 0.0;
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int s = 0; s < 4; ++s) acc[v] = mfma_16x16x4(a[v][s], bx[s], acc[v]);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (t0 + v >= tiles) continue;
                const int r0 = rb + bs + 16 * (t0 + v);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = r0 + 4 * q + lk;
                    if (r < n) X[r * kNb + li] = acc[v][q];
                }
            }
        }
        __syncthreads();
    }
    for (int b = blocks - 1; b >= 0; --b) {
        const int rb = b * kNb, bs = min(kNb, n - rb);
        if (tid < kNb) {
            double x[kNb];
#pragma unroll
            for (int r = 0; r < kNb; ++r) x[r] = r < bs ? X[(rb + r) * kNb + tid] : 0.0;
#pragma unroll
            for (int r = kNb - 1; r >= 0; --r) {
                if (r >= bs) continue;
#pragma unroll
                for (int c = kNb - 1; c > r; --c)
                    if (c < bs) x[r] -= lu[size_t(rb + r) * n + rb + c] * x[c];
                x[r] /= lu[size_t(rb + r) * n + rb + r];
            }
#pragma unroll
            for (int r = 0; r < kNb; ++r)
                if (r < bs) X[(rb + r) * kNb + tid] = x[r];
        }
        __syncthreads();
        // rows above (all within 16-row tiles: rb is a multiple of 16)
        const int tiles = rb / 16;
        double bx[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int kk = 4 * s + lk;
            bx[s] = kk < bs ? X[(rb + kk) * kNb + li] : 0.0;
        }
        for (int t0 = 4 * w; t0 < tiles; t0 += kInvThreads / kWave * 4) {
            d4 acc[4];
            double a[4][4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r0 = 16 * (t0 + v);
                const bool on = t0 + v < tiles;
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[v][q] = on ? X[(r0 + 4 * q + lk) * kNb + li] : 0.0;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int kk = 4 * s + lk;
                    a[v][s] = on && kk < bs ? -lu[size_t(r0 + li) * n + rb + kk] : 0.0;
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int s = 0; s < 4; ++s) acc[v] = mfma_16x16x4(a[v][s], bx[s], acc[v]);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (t0 + v >= tiles) continue;
                const int r0 = 16 * (t0 + v);
#pragma unroll
                for (int q = 0; q < 4; ++q) X[(r0 + 4 * q + lk) * kNb + li] = acc[v][q];
            }
        }
        __syncthreads();
    }
    // column-major output: column j0 + t is contiguous
    for (int i = tid; i < n * kNb; i += kInvThreads) {
        const int t = i / n, r = i - t * n;
        if (t < live) inv[size_t(j0 + t) * n + r] = X[r * kNb + t];
    }
}

// Value of row r of a column held by a lane group (row r in register r / 8
// of lane r % 8), returned to every lane of the group.
__device__ __forceinline__ double group_row(const double (&x)[kSlots], int r, int base) {
    double mine = x[0];
#pragma unroll
    for (int q = 1; q < kSlots; ++q) mine = q == r / kColLanes ? x[q] : mine;
    return __shfl(mine, base + r % kColLanes, kWave);
}

// Sets row r of the group's column to val (in the lane holding it).
__device__ __forceinline__ void group_set(double (&x)[kSlots], int r, int sub, double val) {
#pragma unroll
    for (int q = 0; q < kSlots; ++q) x[q] = (q == r / kColLanes && sub == r % kColLanes) ? val : x[q];
}

// Factor and (when inv is given) invert an order-n matrix, n <= kLuSmall,
// from CSR in one workgroup.  Lane group g = tid / kColLanes owns column g
// in registers: lane sub holds rows sub + kColLanes q (q < kSlots).  Step k:
// group k finds the first maximal |a_ik| (i >= k) with the pivot rule above,
// swaps rows k and p of its column, divides the rows below by the pivot and
// publishes those multipliers and p through LDS; after the barrier every
// other group swaps rows k and p of its column and, right of the pivot,
// subtracts l_i * a_kj from a_ij (rounded product, then the difference).  A
// group's step-k update precedes its own pivot work at step k + 1 in program
// order, and the multipliers alternate between two buffers, so one barrier
// per step orders everything.  The inverse columns are solved in registers
// by the same groups, reading L and U from LDS, with no barrier.
__global__ __launch_bounds__(kSmallThreads) void k_lu_small(int n, const int* __restrict__ rp,
                                                            const int* __restrict__ ci,
                                                            const double* __restrict__ v, double* __restrict__ lu,
                                                            int* __restrict__ piv, double* __restrict__ inv,
                                                            int* bad) {
#pragma clang fp contract(off)
    __shared__ double M[kLuSmall * kLuSmall];  // row-major, leading dimension n
    __shared__ double lcol[2][kLuSmall + 1];   // multipliers of the step's pivot column (+ a discard slot)
    __shared__ int pk[kLuSmall];
    __shared__ int rowsrc[kLuSmall];
    __shared__ double sv[kSmallThreads / kWave];
    __shared__ int failed;
    const int tid = int(threadIdx.x);
    const int lane = tid & (kWave - 1), w = tid / kWave;
    const int g = tid / kColLanes, sub = tid % kColLanes, base = lane - sub;
    for (int i = tid; i < n * n; i += kSmallThreads) M[i] = 0.0;
    if (tid == 0) failed = 0;
    __syncthreads();
    double mx = 0;
    for (int r = tid; r < n; r += kSmallThreads)
        for (int j = rp[r]; j < rp[r + 1]; ++j) {
            M[r * n + ci[j]] = v[j];
            mx = fmax(mx, fabs(v[j]));
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, kWave));
    if (lane == 0) sv[w] = mx;
    __syncthreads();
    double scale = 0;
#pragma unroll
    for (int i = 0; i < kSmallThreads / kWave; ++i) scale = fmax(scale, sv[i]);
    const double thr = 64 * 2.2204460492503131e-16 * scale;
    const int gc = min(g, n - 1);  // (groups past n keep a copy of the last column; never stored)
    double x[kSlots];
#pragma unroll
    for (int q = 0; q < kSlots; ++q) x[q] = M[min(sub + kColLanes * q, n - 1) * n + gc];

    for (int k = 0; k < n; ++k) {
        double* lk = lcol[k & 1];
        if (g == k) {
            double bv = -1.0;
            int bi = -1;
#pragma unroll
            for (int q = 0; q < kSlots; ++q) {
                const int i = sub + kColLanes * q;
                const double a = fabs(x[q]);
                const bool take = i >= k && i < n && a > bv;  // strict: NaN never wins, the first maximum stays
                bv = take ? a : bv;
                bi = take ? i : bi;
            }
            // the group's 8 lanes: quad exchanges, then the half-row mirror
            argmax_dpp<0xB1, 0xF>(bv, bi);
            argmax_dpp<0x4E, 0xF>(bv, bi);
            argmax_dpp<0x141, 0xF>(bv, bi);
            const double akk = group_row(x, k, base);
            const int p = (bi < 0 || isnan(akk)) ? k : bi;
            const double pv = group_row(x, p, base);
            if (!isfinite(pv) || fabs(pv) <= thr) {
                if (sub == 0) failed = 1;
            }
            else {
                group_set(x, p, sub, akk);
                group_set(x, k, sub, pv);
                // every slot divides (the quotients go out together); rows
                // outside (k, n) keep their value and store to the discard slot
                double l[kSlots];
#pragma unroll
                for (int q = 0; q < kSlots; ++q) l[q] = x[q] / pv;
#pragma unroll
                for (int q = 0; q < kSlots; ++q) {
                    const int i = sub + kColLanes * q;
                    const bool below = unsigned(i - k - 1) < unsigned(n - k - 1);
                    x[q] = below ? l[q] : x[q];
                    lk[below ? i : kLuSmall] = l[q];
                }
                if (sub == 0) pk[k] = p;
            }
        }
        __syncthreads();
        if (failed) {
            if (tid == 0) *bad = 1;
            return;
        }
        if (g < n && g != k) {
            const int p = pk[k];
            const double vk = group_row(x, k, base), vp = group_row(x, p, base);
            group_set(x, p, sub, vk);
            group_set(x, k, sub, vp);
            if (g > k) {
                double l[kSlots];
#pragma unroll
                for (int q = 0; q < kSlots; ++q) l[q] = lk[min(sub + kColLanes * q, n - 1)];
#pragma unroll
                for (int q = 0; q < kSlots; ++q) {
                    const int i = sub + kColLanes * q;
                    const double prod = l[q] * vp;
                    const double y = x[q] - prod;
                    x[q] = unsigned(i - k - 1) < unsigned(n - k - 1) ? y : x[q];
                }
            }
        }
    }
    int nonfinite = 0;
    if (g < n) {
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            const int i = sub + kColLanes * q;
            if (i < n) {
                nonfinite |= isfinite(x[q]) ? 0 : 1;
                M[i * n + g] = x[q];
                lu[size_t(i) * n + g] = x[q];
            }
        }
    }
    for (int i = tid; i < n; i += kSmallThreads) piv[i] = pk[i];
    if (__syncthreads_or(nonfinite)) {
        if (tid == 0) *bad = 1;
        return;
    }
    if (!inv) return;

    // X = P I (the row swaps in pivot order), then L Y = X and U Z = Y by
    // columns, each as the sequential substitution: forward updates in row
    // order, backward the division then the updates above.
    if (tid == 0) {
        for (int i = 0; i < n; ++i) rowsrc[i] = i;
        for (int k = 0; k < n; ++k) {
            const int p = pk[k], t = rowsrc[k];
            rowsrc[k] = rowsrc[p];
            rowsrc[p] = t;
        }
    }
    __syncthreads();
    if (g < n) {
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            const int i = sub + kColLanes * q;
            x[q] = i < n && rowsrc[i] == g ? 1.0 : 0.0;
        }
        for (int k = 0; k < n; ++k) {
            double l[kSlots];
#pragma unroll
            for (int q = 0; q < kSlots; ++q) l[q] = M[min(sub + kColLanes * q, n - 1) * n + k];
            const double xk = group_row(x, k, base);
#pragma unroll
            for (int q = 0; q < kSlots; ++q) {
                const int i = sub + kColLanes * q;
                const double prod = l[q] * xk;
                const double y = x[q] - prod;
                x[q] = unsigned(i - k - 1) < unsigned(n - k - 1) ? y : x[q];
            }
        }
        for (int k = n - 1; k >= 0; --k) {
            double u[kSlots];
#pragma unroll
            for (int q = 0; q < kSlots; ++q) u[q] = M[min(sub + kColLanes * q, n - 1) * n + k];
            const double xk = group_row(x, k, base) / M[k * n + k];
            group_set(x, k, sub, xk);
#pragma unroll
            for (int q = 0; q < kSlots; ++q) {
                const int i = sub + kColLanes * q;
                const double prod = u[q] * xk;
                const double y = x[q] - prod;
                x[q] = i < k ? y : x[q];
            }
        }
        // column-major output: column g is contiguous
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            const int i = sub + kColLanes * q;
            if (i < n) inv[size_t(g) * n + i] = x[q];
        }
    }
}

}  // namespace

bool coarse_lu_factor(tf_ctx& c, int n, const int* rp, const int* ci, const double* v, double* lu, int* piv,
                      double* inverse) {
    require(n >= 1 && n <= kCoarseMax, "GPU LU requires order 1..1024");
    c.clear_flags(1);
    if (n <= kLuSmall) {
        hipLaunchKernelGGL(k_lu_small, dim3(1), dim3(kSmallThreads), 0, c.stream, n, rp, ci, v, lu, piv, inverse,
                           c.dflags);
        check(hipGetLastError(), "small lu launch");
        c.read_flags(1);
        return c.hflags[0] == 0;
    }
    hipLaunchKernelGGL(k_dense_fill, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c.stream, n, rp, ci, v, lu);
    check(hipGetLastError(), "dense fill launch");
    hipLaunchKernelGGL(k_lu_factor, dim3(1), dim3(kFactorThreads), 0, c.stream, n, lu, piv, c.dflags);
    check(hipGetLastError(), "lu factor launch");
    c.read_flags(1);
    const bool ok = c.hflags[0] == 0;
    if (ok && inverse) {
        hipLaunchKernelGGL(k_lu_inverse, dim3((n + kNb - 1) / kNb), dim3(kInvThreads), 0, c.stream, n, lu, piv,
                           inverse, static_cast<const int*>(nullptr));
        check(hipGetLastError(), "lu inverse launch");
    }
    return ok;
}

// The factorisation's verdict slot in the context's flags (apart from the
// low slots the main stream's setup kernels use meanwhile).
constexpr int kLuFlagSlot = 48;

void coarse_lu_launch(tf_ctx& c, int n, const int* rp, const int* ci, const double* v, double* lu, int* piv,
                      double* inverse) {
    require(n >= 1 && n <= kCoarseMax, "GPU LU requires order 1..1024");
    hipStream_t s = c.side_stream();
    int* flag = c.dflags + kLuFlagSlot;
    check(hipMemsetAsync(flag, 0, sizeof(int), c.stream), "lu flag clear");
    check(hipEventRecord(c.ev_fork, c.stream), "fork record");
    check(hipStreamWaitEvent(s, c.ev_fork, 0), "fork wait");
    if (n <= kLuSmall) {
        hipLaunchKernelGGL(k_lu_small, dim3(1), dim3(kSmallThreads), 0, s, n, rp, ci, v, lu, piv, inverse, flag);
    }
    else {
        hipLaunchKernelGGL(k_dense_fill, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, rp, ci, v, lu);
        hipLaunchKernelGGL(k_lu_factor, dim3(1), dim3(kFactorThreads), 0, s, n, lu, piv, flag);
        if (inverse)
            hipLaunchKernelGGL(k_lu_inverse, dim3((n + kNb - 1) / kNb), dim3(kInvThreads), 0, s, n, lu, piv, inverse,
                               static_cast<const int*>(flag));
    }
    check(hipGetLastError(), "lu launch");
    check(hipEventRecord(c.ev_join, s), "join record");
}

bool coarse_lu_join(tf_ctx& c) {
    check(hipStreamWaitEvent(c.stream, c.ev_join, 0), "join wait");
    check(hipMemcpyAsync(c.hflags + kLuFlagSlot, c.dflags + kLuFlagSlot, sizeof(int), hipMemcpyDeviceToHost, c.stream),
          "lu flag download");
    c.sync();
    return c.hflags[kLuFlagSlot] == 0;
}

}  // namespace tf
