// skinny.hip.h -- the block product of the super-block sweeps on gfx950:
// one workgroup computes a 64-row x 32-column block
//     P(r, c) = sum_{k in [klo, khi)} X(r, k) Y(k, c)
// with a long K (a super-block, 1024) and few columns, in FP32.
//
// The K range is split across the 8 waves of a 512-thread workgroup; each
// wave walks its slice in 16-deep steps with its operands loaded straight
// from memory into the v_mfma_f32_16x16x4_f32 operand layout, one step
// ahead, with no LDS and no barrier in the loop.  The partial blocks of the
// 8 waves are added in LDS at the end, in a fixed order.
//
// Operand layout.  The instruction takes A[m][k] from lane l at m = l % 16,
// k = l / 16, and B[k][n] at k = l / 16, n = l % 16; D[m][n] comes back in
// lane l, element q, at m = 4 (l / 16) + q, n = l % 16.  The 64 rows are
// four 16-row MFMA blocks i whose slot m holds row 4 m + i, so the four
// blocks' A values of one lane are four consecutive rows: one 16-byte load
// when rows are contiguous.  The K slots are permuted too: in substep t of
// a step at k0, slot s is k = k0 + 4 s + t, so a lane's four substeps read
// four consecutive k: one 16-byte load of Y (K contiguous), and of X when
// X is stored K-contiguous.  Any assignment that agrees between A and B
// gives the same sum; the result rows of block i, element q are
// 4 (4 (l / 16) + q) + i = 16 (l / 16) + 4 q + i.
#pragma once

#include "common.hip.h"

namespace skinny {

constexpr int NW = 8;                 // waves per workgroup
constexpr int NT = 64 * NW;
constexpr int BR = 64, BC = 32;       // rows x columns of the block
constexpr int STEP = 16;              // k per step
constexpr int RLD = BR + 4;           // padded LDS column (floats)

struct Smem {
    float red[NW][BC][RLD];
};

// Operand window of one step: element (r, k) / (k, c) at p[r rs + k ks] /
// p[k + c cs].  A step whose loads are all in bounds and 16-byte aligned
// reads vectors; any other reads element by element with bounds.
struct XStep {
    const float *p;
    int rs, ks;            // row stride, k stride (one of them is 1)
};
struct YStep {
    const float *p;        // k contiguous
    int cs;                // column stride
};

typedef float4_t acc_t;

// Operands of one 16-deep step in MFMA order: a[t][i] (substep t, row
// block i), b[j][t] (column block j).
struct Frag {
    float a[4][4], b[2][4];
};

template <bool XROWS>
__device__ __forceinline__ void load(Frag &f, const XStep &x, const YStep &y,
                                     int xrows, int kn, int ncols, bool fast)
{
    const int lane = threadIdx.x & 63;
    const int rl = 4 * (lane % 16);          // first row of this lane
    const int kl = 4 * (lane / 16);          // first k of this lane
    if (fast) {
        if (XROWS) {                         // rows contiguous: one vector per t
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const float4_t v = *reinterpret_cast<const float4_t *>(
                    x.p + rl + (ptrdiff_t)(kl + t) * x.ks);
#pragma unroll
                for (int i = 0; i < 4; i++) f.a[t][i] = v[i];
            }
        }
        else {                               // k contiguous: one vector per i
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float4_t v = *reinterpret_cast<const float4_t *>(
                    x.p + (ptrdiff_t)(rl + i) * x.rs + kl);
#pragma unroll
                for (int t = 0; t < 4; t++) f.a[t][i] = v[t];
            }
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const float4_t v = *reinterpret_cast<const float4_t *>(
                y.p + kl + (ptrdiff_t)(16 * j + lane % 16) * y.cs);
#pragma unroll
            for (int t = 0; t < 4; t++) f.b[j][t] = v[t];
        }
    }
    else {
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int k = kl + t;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = rl + i;
                f.a[t][i] = (r < xrows && k < kn)
                    ? x.p[(ptrdiff_t)r * x.rs + (ptrdiff_t)k * x.ks] : 0.0f;
            }
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int c = 16 * j + lane % 16;
                f.b[j][t] = (k < kn && c < ncols)
                    ? y.p[k + (ptrdiff_t)c * y.cs] : 0.0f;
            }
        }
    }
}

__device__ __forceinline__ void compute(acc_t (&acc)[4][2], const Frag &f)
{
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                    f.a[t][i], f.b[j][t], acc[i][j], 0, 0, 0);
}

__device__ __forceinline__ bool aligned16(const void *p)
{
    return ((uintptr_t)p & 15) == 0;
}

// The whole block: waves take K slices of [klo, khi) (multiples of 16
// from klo).  XF(k) / YF(k) give the step windows at absolute k (a step
// never crosses a tile of the operands: tiles are multiples of 16 deep and
// klo is a multiple of 16).  On return the sum of the waves' partial
// blocks is in sm.red[0][c][r] (r < 64, c < 32), ready for store().
template <bool XROWS, typename XF, typename YF>
__device__ __forceinline__ void block(Smem &sm, int klo, int khi, int xrows,
                                      int ncols, XF xf, YF yf)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    acc_t acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = acc_t{0, 0, 0, 0};
    const int nsteps = (khi - klo + STEP - 1) / STEP;
    const int per = (nsteps + NW - 1) / NW;
    const int s0 = wave * per, s1 = min(nsteps, s0 + per);
    // step s's operands are loaded while step s - 1 computes
    auto fetch = [&](Frag &f, int s) {
        const int k0 = klo + s * STEP;
        const XStep x = xf(k0);
        const YStep y = yf(k0);
        const int kn = min(STEP, khi - k0);
        const bool fast =
            kn == STEP && xrows >= BR && ncols >= BC &&
            aligned16(x.p) && aligned16(y.p) && (y.cs & 3) == 0 &&
            (XROWS ? (x.ks & 3) == 0 : (x.rs & 3) == 0);
        load<XROWS>(f, x, y, xrows, kn, ncols, fast);
    };
    Frag cur;
    if (s0 < s1) fetch(cur, s0);
    for (int s = s0; s < s1; s++) {
        Frag nxt;
        if (s + 1 < s1) fetch(nxt, s + 1);
        compute(acc, cur);
        cur = nxt;
    }
    // partial blocks into LDS: lane l, block (i, j), element q is row
    // 16 (l / 16) + 4 q + i, column 16 j + l % 16
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float4_t v;
#pragma unroll
            for (int i = 0; i < 4; i++) v[i] = acc[i][j][q];
            *reinterpret_cast<float4_t *>(
                &sm.red[wave][16 * j + lane % 16][16 * (lane / 16) + 4 * q]) =
                v;
        }
    __syncthreads();
    // fixed-order sum over the waves into red[0]: thread t owns rows
    // 4 (t % 16) .. + 3 of column t / 16
    {
        const int c = threadIdx.x / 16, r = 4 * (threadIdx.x % 16);
        float4_t s = *reinterpret_cast<const float4_t *>(&sm.red[0][c][r]);
#pragma unroll
        for (int w = 1; w < NW; w++)
            s += *reinterpret_cast<const float4_t *>(&sm.red[w][c][r]);
        __syncthreads();
        *reinterpret_cast<float4_t *>(&sm.red[0][c][r]) = s;
    }
    __syncthreads();
}

// C(r0 + r, c0 + c) (-)= red[0][c][r] for r < mv - r0, c < nv - c0; C
// column major with unit row stride.
__device__ __forceinline__ void store(const Smem &sm, float *C, int ld,
                                      int rows, int cols, bool subtract)
{
    const int c = threadIdx.x / 16, r = 4 * (threadIdx.x % 16);
    if (c >= cols || r >= rows) return;
    const float4_t v = *reinterpret_cast<const float4_t *>(&sm.red[0][c][r]);
    float *p = C + r + (ptrdiff_t)c * ld;
    if (r + 4 <= rows && aligned16(p)) {
        float4_t o = v;
        if (subtract) o = *reinterpret_cast<const float4_t *>(p) - v;
        *reinterpret_cast<float4_t *>(p) = o;
    }
    else {
        for (int u = 0; u < 4 && r + u < rows; u++)
            p[u] = subtract ? p[u] - v[u] : v[u];
    }
}

}  // namespace skinny
