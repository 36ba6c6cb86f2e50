// tilegemm.hip.h -- the MFMA block GEMM shared by the trailing update and
// the panel solve: one workgroup (4 waves, 2 x 2 of 64 x 64) computes
//     C_out(128 x 128 block) = [C_in] -/+ X Y^T
// where X and Y are "rows x K" operands (any strides) and K streams through
// double-buffered LDS in chunks of KC.
//
//   * With ACC the C block is loaded into the accumulators before the K loop
//     (X negated while it is staged when NEG), so the epilogue is a store.
//   * CROW (C has unit row stride, i.e. column-major tiles) forms the
//     product transposed (Y X^T), which puts 16 consecutive rows of one C
//     column in 16 consecutive lanes: C moves in 128-byte runs.  Without
//     CROW (C has unit column stride) X Y^T gives the same property.
//   * 16-byte global loads when a block is full and aligned.
//   * A Decode functor maps blockIdx.x to one job (views, bounds, mask).
#pragma once

#include "common.hip.h"

namespace tg {

constexpr int NT = 256;
constexpr int BM = 128;

template <typename T> struct Vec;
template <> struct Vec<double> { typedef double2 t; static constexpr int W = 2; };
template <> struct Vec<float>  { typedef float4  t; static constexpr int W = 4; };

template <typename T> struct Cfg;
template <> struct Cfg<double> { static constexpr int KC = 16; };
template <> struct Cfg<float>  { static constexpr int KC = 32; };

template <typename T>
struct Job {
    View<T> X, Y;      // operand rows, positioned at the block's first row
    View<T> C;         // positioned at the block's (0, 0)
    int xrows, yrows;  // valid operand rows (>= BM: full block)
    int K;
    bool lower;        // store only r >= c (block-relative + diag offset)
    int diag_off;      // c - r offset of the block origin for the mask
};

template <typename T>
struct Tile {
    static constexpr int KC = Cfg<T>::KC;
    static constexpr int LD = BM;
    static constexpr int NE = BM * KC / NT;
    static constexpr int W = Vec<T>::W;
    static constexpr int NV = NE / W;
    typedef typename Vec<T>::t vec_t;

    struct Stage { T v[NE]; };

    // UNIT: the operand's rows are contiguous (V.rs == 1); otherwise its k
    // index is (V.cs == 1).
    template <bool UNIT>
    static __device__ __forceinline__ void load(Stage &g, const View<T> &V,
                                                int nrows, int k0, int K,
                                                bool fast)
    {
        const int t = threadIdx.x;
        if (fast && UNIT) {
#pragma unroll
            for (int i = 0; i < NV; i++) {
                const int v = i * NT + t;
                const int kk = v / (BM / W), row = (v % (BM / W)) * W;
                vec_t x = *reinterpret_cast<const vec_t *>(
                    V.p + row + (ptrdiff_t)(k0 + kk) * V.cs);
#pragma unroll
                for (int u = 0; u < W; u++) g.v[i * W + u] = x[u];
            }
        }
        else if (fast) {
#pragma unroll
            for (int i = 0; i < NV; i++) {
                const int v = i * NT + t;
                const int row = v / (KC / W), kk = (v % (KC / W)) * W;
                vec_t x = *reinterpret_cast<const vec_t *>(
                    V.p + (ptrdiff_t)row * V.rs + k0 + kk);
#pragma unroll
                for (int u = 0; u < W; u++) g.v[i * W + u] = x[u];
            }
        }
        else {
#pragma unroll
            for (int i = 0; i < NV; i++) {
                const int v = i * NT + t;
#pragma unroll
                for (int u = 0; u < W; u++) {
                    int row, kk;
                    if (UNIT) { kk = v / (BM / W); row = (v % (BM / W)) * W + u; }
                    else      { row = v / (KC / W); kk = (v % (KC / W)) * W + u; }
                    const int kg = k0 + kk;
                    g.v[i * W + u] = (row < nrows && kg < K)
                        ? (UNIT ? V.p[row + (ptrdiff_t)kg * V.cs]
                                : V.p[(ptrdiff_t)row * V.rs + kg])
                        : T(0);
                }
            }
        }
    }

    template <bool UNIT, bool NEGATE>
    static __device__ __forceinline__ void store(const Stage &g, T *s)
    {
        const int t = threadIdx.x;
        const T sg = NEGATE ? T(-1) : T(1);
#pragma unroll
        for (int i = 0; i < NV; i++) {
            const int v = i * NT + t;
            if (UNIT) {
                const int kk = v / (BM / W), row = (v % (BM / W)) * W;
                vec_t x;
#pragma unroll
                for (int u = 0; u < W; u++) x[u] = sg * g.v[i * W + u];
                *reinterpret_cast<vec_t *>(&s[kk * LD + row]) = x;
            }
            else {
                const int row = v / (KC / W), kk = (v % (KC / W)) * W;
#pragma unroll
                for (int u = 0; u < W; u++)
                    s[(kk + u) * LD + row] = sg * g.v[i * W + u];
            }
        }
    }
};

template <typename T, bool UNIT>
__device__ __forceinline__ bool vec_ok(const View<T> &V)
{
    const int ld = UNIT ? V.cs : V.rs;
    return ((uintptr_t)V.p % 16 == 0) && (ld % Vec<T>::W == 0);
}

// The block GEMM body; returns after the epilogue.
template <typename T, bool CROW, bool ACC, bool NEG, bool XU, bool YU>
__device__ __forceinline__ void run(const Job<T> &job, T (*sx)[Tile<T>::KC * BM],
                                    T (*sy)[Tile<T>::KC * BM])
{
    typedef Tile<T> U;
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int KC = U::KC, LD = U::LD;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int K = job.K;

    // C element of accumulator (i, j, q): row from X, column from Y
    auto crow = [&](int i, int q) {
        return CROW ? wr * 64 + i * 16 + lane % 16
                    : wr * 64 + i * 16 + Mfma<T>::row(lane, q);
    };
    auto ccol = [&](int j, int q) {
        return CROW ? wc * 64 + j * 16 + Mfma<T>::row(lane, q)
                    : wc * 64 + j * 16 + lane % 16;
    };
    const int mrows = job.xrows, ncols = job.yrows;

    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                T v = T(0);
                if (ACC) {
                    const int r = crow(i, q), c = ccol(j, q);
                    if (r < mrows && c < ncols) v = job.C.at(r, c);
                }
                acc[i][j][q] = v;
            }

    const bool xvec = vec_ok<T, XU>(job.X) && job.xrows >= BM;
    const bool yvec = vec_ok<T, YU>(job.Y) && job.yrows >= BM;
    const int nk = (K + KC - 1) / KC;
    typename U::Stage gx, gy;
    U::template load<XU>(gx, job.X, job.xrows, 0, K, xvec && KC <= K);
    U::template load<YU>(gy, job.Y, job.yrows, 0, K, yvec && KC <= K);
    U::template store<XU, NEG>(gx, sx[0]);
    U::template store<YU, false>(gy, sy[0]);
    __syncthreads();

    for (int kc = 0; kc < nk; kc++) {
        const int cur = kc & 1;
        if (kc + 1 < nk) {
            const int k0 = (kc + 1) * KC;
            U::template load<XU>(gx, job.X, job.xrows, k0, K, xvec && k0 + KC <= K);
            U::template load<YU>(gy, job.Y, job.yrows, k0, K, yvec && k0 + KC <= K);
        }
        const T *bx = sx[cur], *by = sy[cur];
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            T a[4], b[4];
            const int kr = (kk + lane / 16) * LD + lane % 16;
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = bx[kr + wr * 64 + i * 16];
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = by[kr + wc * 64 + j * 16];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = CROW ? Mfma<T>::mma(b[j], a[i], acc[i][j])
                                     : Mfma<T>::mma(a[i], b[j], acc[i][j]);
        }
        if (kc + 1 < nk) {
            U::template store<XU, NEG>(gx, sx[cur ^ 1]);
            U::template store<YU, false>(gy, sy[cur ^ 1]);
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = crow(i, q), c = ccol(j, q);
                if (r < mrows && c < ncols &&
                    (!job.lower || c + job.diag_off <= r))
                    job.C.at(r, c) = acc[i][j][q];
            }
}

// Workgroups are dispatched to the eight XCDs of the chip round robin
// (workgroup b to XCD b % 8), and each XCD has its own L2.  Decoders number
// their jobs so that neighbours share operands (the blocks of one tile,
// the tiles of one column), so workgroup b runs job
//     xcd_job(b) = (b % 8) * chunk + b / 8,   chunk = ceil(njobs / 8),
// which gives every XCD one contiguous range of jobs.  Grids are padded to
// a multiple of 8 (8 * chunk); indices past njobs do nothing.
constexpr int NXCD = 8;

__host__ __device__ __forceinline__ int xcd_grid(int njobs)
{
    return (njobs + NXCD - 1) / NXCD * NXCD;
}

// XU / YU: the decoder's X / Y views have unit row stride (else unit k
// stride); fixed at compile time so each loader has one code path.  The
// workgroups loop over b = blockIdx.x, blockIdx.x + gridDim.x, ... below
// the padded job count: a grid smaller than that keeps its workgroups
// resident and leaves the rest of the chip's workgroup slots to other
// streams (the jobs are independent, so any grid size is correct).
template <typename T, bool CROW, bool ACC, bool NEG, bool XU, bool YU,
          typename Decode>
__global__ __launch_bounds__(NT, 2) void kernel(Decode dec, int njobs,
                                                const int *status)
{
    __shared__ __attribute__((aligned(16))) T sx[2][Tile<T>::KC * BM];
    __shared__ __attribute__((aligned(16))) T sy[2][Tile<T>::KC * BM];
    if (block_should_stop(status)) return;
    const int padded = xcd_grid(njobs), chunk = padded / NXCD;
    for (int b = blockIdx.x; b < padded; b += gridDim.x) {
        const int jb = (b % NXCD) * chunk + b / NXCD;
        Job<T> job;
        if (jb >= njobs || !dec(jb, job)) continue;
        // run() ends with a barrier after its last LDS read
        run<T, CROW, ACC, NEG, XU, YU>(job, sx, sy);
    }
}

}  // namespace tg
