// common.hip.h -- device-side building blocks shared by the gfx950 kernels.
//
//  * Mfma<T>: one wave computes a 16x16 block D += A(16x4) * B(4x16) with
//    v_mfma_f64_16x16x4_f64 (T = double) or v_mfma_f32_16x16x4_f32
//    (T = float).  Operand lane layout (both instructions):
//      A: lane l holds A[l % 16][l / 16]
//      B: lane l holds B[l / 16][l % 16]
//    Result: lane l, element q holds D[row][l % 16] with
//      row = l / 16 + 4 * q       for the f64 instruction,
//      row = 4 * (l / 16) + q     for the f32 instruction
//    (both measured on gfx950 with an in-bounds probe).
//  * View<T>: a strided 2-D window, element (r, c) at p[r * rs + c * cs].
//    lower_view() returns tile (m, n) of the lower-equivalent factor
//    L = (stored triangle) for PlasmaLower, or U^T for PlasmaUpper, so that
//    every algorithm below is written once, for L.
//  * GemmNT: acc += X * Y^T for a BM x BN block, K streamed in chunks of
//    KC through double-buffered LDS; 256 threads = 4 waves.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "plasma_descriptor.h"

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;

template <> struct Mfma<double> {
    typedef double4_t acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int q)
    {
        return lane / 16 + 4 * q;
    }
};

template <> struct Mfma<float> {
    typedef float4_t acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int q)
    {
        return 4 * (lane / 16) + q;
    }
};

template <typename T> struct View {
    T *p;
    int rs, cs;
    __device__ __forceinline__ T &at(int r, int c) const
    {
        return p[(ptrdiff_t)r * rs + (ptrdiff_t)c * cs];
    }
};

// True when the kernels must stop: a previous operation of the sequence
// failed (reference: every task body tests sequence->status first).
__device__ __forceinline__ bool seq_failed(const int *status)
{
    return __atomic_load_n(status, __ATOMIC_RELAXED) != 0;
}

// Block-uniform form of seq_failed: thread 0 reads the word once and every
// thread of the workgroup gets the same answer, so no wave can be left
// waiting at a barrier another wave skipped (the word may be set by a kernel
// running concurrently on the lookahead stream).
__device__ __forceinline__ bool block_should_stop(const int *status)
{
    __shared__ int s_stop;
    if (threadIdx.x == 0) s_stop = seq_failed(status) ? 1 : 0;
    __syncthreads();
    return s_stop != 0;
}

// Tile (m, n) of the lower-equivalent factor, m >= n.
template <typename T>
__device__ __forceinline__ View<T> lower_view(const plasma_desc_t &A,
                                              bool upper, int m, int n)
{
    View<T> v;
    if (!upper) {
        v.p = (T *)plasma_tile_addr(A, m, n);
        v.rs = 1;
        v.cs = plasma_tile_mmain(A, m);
    }
    else {
        v.p = (T *)plasma_tile_addr(A, n, m);
        v.rs = plasma_tile_mmain(A, n);
        v.cs = 1;
    }
    return v;
}

// Plain column-major tile (m, n) of a general descriptor.
template <typename T>
__device__ __forceinline__ View<T> tile_view(const plasma_desc_t &A, int m,
                                             int n)
{
    View<T> v;
    v.p = (T *)plasma_tile_addr(A, m, n);
    v.rs = 1;
    v.cs = plasma_tile_mmain(A, m);
    return v;
}

// Decode a linear index over the trapezoid of tiles {(m, n): n in
// [n0, n1), m in [n, mt)} enumerated column by column.
__device__ __forceinline__ void trapezoid_decode(int t, int mt, int n0,
                                                 int *m, int *n)
{
    long T0 = mt - n0;  // tiles in the first column
    // P(j) = j*T0 - j*(j-1)/2 tiles precede column j
    double b = 2.0 * (double)T0 + 1.0;
    long j = (long)((b - sqrt(b * b - 8.0 * (double)t)) * 0.5);
    if (j < 0) j = 0;
    while (j > 0 && j * T0 - j * (j - 1) / 2 > t) j--;
    while ((j + 1) * T0 - (j + 1) * j / 2 <= t) j++;
    long first = j * T0 - j * (j - 1) / 2;
    *n = n0 + (int)j;
    *m = *n + (int)(t - first);
}

// ---------------------------------------------------------------------------
// Block GEMM core: acc += X(r0:r0+BM, 0:K) * Y(c0:c0+BN, 0:K)^T
// X, Y are Views positioned at their first row; rows beyond xm / ym and k
// beyond K read as zero.  256 threads, 4 waves in a (BM/WM) x (BN/WN) grid.
// ---------------------------------------------------------------------------
template <typename T, int BM, int BN, int KC, int WM, int WN>
struct GemmNT {
    typedef typename Mfma<T>::acc_t acc_t;
    static constexpr int NT = 256;
    static constexpr int FM = WM / 16;  // MFMA blocks per wave, rows
    static constexpr int FN = WN / 16;  // MFMA blocks per wave, cols
    static constexpr int LX = BM * KC / NT;  // elements per thread per chunk
    static constexpr int LY = BN * KC / NT;
    static_assert((BM / WM) * (BN / WN) == 4, "four waves");
    static_assert(LX >= 1 && LY >= 1, "chunk too small");

    // LDS rows padded by 16 bytes: row starts stay 16-byte aligned for the
    // vector stores, and the transposed stores (one lane per k value, W
    // rows apart) no longer land in one bank
    static constexpr int PAD = 16 / (int)sizeof(T);
    struct Smem {
        T x[2][KC][BM + PAD];
        T y[2][KC][BN + PAD];
    };

    // Chunk loads into registers, W elements per 16-byte vector.  With
    // `rowmap` (unit row stride, or no unit stride at all) consecutive
    // threads walk rows: vector v covers rows r..r+W-1 of chunk column kk;
    // otherwise (unit k stride) they walk k: vector v covers k values
    // kk..kk+W-1 of row r.  Full, 16-byte aligned chunks with a unit stride
    // use vector loads; the rest is read element by element with bounds.
    static constexpr int W = 16 / (int)sizeof(T);
    static __device__ __forceinline__ bool rowmap(const View<T> &V)
    {
        return V.rs == 1 || V.cs != 1;
    }

    template <int ROWS, int L>
    static __device__ __forceinline__ void load(T (&reg)[L], const View<T> &V,
                                                int nrows, int k0, int K)
    {
        static_assert(L % W == 0, "whole vectors per thread");
        typedef T vec_t __attribute__((ext_vector_type(W)));
        const int t = threadIdx.x;
        const bool rm = rowmap(V);
        const int ld = rm ? V.cs : V.rs;
        const bool fast = nrows >= ROWS && k0 + KC <= K &&
                          (rm ? V.rs == 1 : true) &&
                          ((uintptr_t)V.p % 16 == 0) && (ld % W == 0);
        // one uniform branch around straight-line load sequences
        if (rm && fast) {
#pragma unroll
            for (int i = 0; i < L / W; i++) {
                const int v = i * NT + t;
                const int kk = v / (ROWS / W), r = (v % (ROWS / W)) * W;
                vec_t x = *reinterpret_cast<const vec_t *>(
                    V.p + r + (ptrdiff_t)(k0 + kk) * V.cs);
#pragma unroll
                for (int u = 0; u < W; u++) reg[i * W + u] = x[u];
            }
        }
        else if (rm) {
#pragma unroll
            for (int i = 0; i < L / W; i++) {
                const int v = i * NT + t;
                const int kk = v / (ROWS / W), r = (v % (ROWS / W)) * W;
                const int kg = k0 + kk;
#pragma unroll
                for (int u = 0; u < W; u++)
                    reg[i * W + u] = (r + u < nrows && kg < K)
                        ? V.p[(ptrdiff_t)(r + u) * V.rs + (ptrdiff_t)kg * V.cs]
                        : T(0);
            }
        }
        else if (fast) {
#pragma unroll
            for (int i = 0; i < L / W; i++) {
                const int v = i * NT + t;
                const int r = v / (KC / W), kk = (v % (KC / W)) * W;
                vec_t x = *reinterpret_cast<const vec_t *>(
                    V.p + (ptrdiff_t)r * V.rs + (k0 + kk));
#pragma unroll
                for (int u = 0; u < W; u++) reg[i * W + u] = x[u];
            }
        }
        else {
#pragma unroll
            for (int i = 0; i < L / W; i++) {
                const int v = i * NT + t;
                const int r = v / (KC / W), kk = (v % (KC / W)) * W;
                const int kg = k0 + kk;
#pragma unroll
                for (int u = 0; u < W; u++)
                    reg[i * W + u] = (r < nrows && kg + u < K)
                        ? V.p[(ptrdiff_t)r * V.rs + (kg + u)]
                        : T(0);
            }
        }
    }

    // Registers -> LDS chunk s[kk][row], with the mapping load() used.
    template <int ROWS, int L, int LD>
    static __device__ __forceinline__ void store(const T (&reg)[L],
                                                 T (*s)[LD], bool rm)
    {
        const int t = threadIdx.x;
        if (rm) {
            typedef T vec_t __attribute__((ext_vector_type(W)));
#pragma unroll
            for (int i = 0; i < L / W; i++) {
                const int v = i * NT + t;
                const int kk = v / (ROWS / W), r = (v % (ROWS / W)) * W;
                vec_t x;
#pragma unroll
                for (int u = 0; u < W; u++) x[u] = reg[i * W + u];
                *reinterpret_cast<vec_t *>(&s[kk][r]) = x;
            }
        }
        else {
#pragma unroll
            for (int i = 0; i < L / W; i++) {
                const int v = i * NT + t;
                const int r = v / (KC / W), kk = (v % (KC / W)) * W;
#pragma unroll
                for (int u = 0; u < W; u++) s[kk + u][r] = reg[i * W + u];
            }
        }
    }

    static __device__ __forceinline__ void zero(acc_t (&acc)[FM][FN])
    {
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++)
                acc[i][j] = acc_t{0, 0, 0, 0};
    }

    static __device__ __forceinline__ void run(acc_t (&acc)[FM][FN],
                                               const View<T> &X, int xm,
                                               const View<T> &Y, int ym,
                                               int K, Smem &sm)
    {
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
        const int wr = wave / (BN / WN), wc = wave % (BN / WN);
        const int nk = (K + KC - 1) / KC;
        const bool xu = rowmap(X), yu = rowmap(Y);
        T rx[LX], ry[LY];

        load<BM>(rx, X, xm, 0, K);
        load<BN>(ry, Y, ym, 0, K);
        store<BM>(rx, sm.x[0], xu);
        store<BN>(ry, sm.y[0], yu);
        __syncthreads();

        for (int kc = 0; kc < nk; kc++) {
            const int cur = kc & 1;
            if (kc + 1 < nk) {
                load<BM>(rx, X, xm, (kc + 1) * KC, K);
                load<BN>(ry, Y, ym, (kc + 1) * KC, K);
            }
#pragma unroll
            for (int kk = 0; kk < KC; kk += 4) {
                T a[FM], b[FN];
#pragma unroll
                for (int i = 0; i < FM; i++)
                    a[i] = sm.x[cur][kk + lane / 16][wr * WM + i * 16 + lane % 16];
#pragma unroll
                for (int j = 0; j < FN; j++)
                    b[j] = sm.y[cur][kk + lane / 16][wc * WN + j * 16 + lane % 16];
#pragma unroll
                for (int i = 0; i < FM; i++)
#pragma unroll
                    for (int j = 0; j < FN; j++)
                        acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
            }
            if (kc + 1 < nk) {
                store<BM>(rx, sm.x[cur ^ 1], xu);
                store<BN>(ry, sm.y[cur ^ 1], yu);
            }
            __syncthreads();
        }
    }

    // Row / column (within the BM x BN block) of accumulator element q of
    // MFMA block (i, j) held by this lane.
    static __device__ __forceinline__ int acc_row(int i, int q)
    {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        return (wave / (BN / WN)) * WM + i * 16 + Mfma<T>::row(lane, q);
    }
    static __device__ __forceinline__ int acc_col(int j)
    {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        return (wave % (BN / WN)) * WN + j * 16 + lane % 16;
    }
};
