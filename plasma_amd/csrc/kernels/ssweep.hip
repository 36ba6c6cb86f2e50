// ssweep.hip -- triangular sweeps of potrs by super-blocks of S tiles, for
// solves that reuse one factor many times (the refinement loop of
// plasma_dsposv).
//
// The reference sweeps tile by tile (compute/pztrsm.c:113-177): per block
// row k a trsm with the diagonal tile, then a gemm update of every block row
// still to be solved, mt steps per sweep.  On the GPU each step is a launch
// whose work shrinks to a few microseconds, so the sweep is bound by the
// chain of launches.  Here the diagonal blocks are super-blocks of S tiles
// (S nb = 2048 rows by default):
//   * once per factor, V_K = L_KK^{-1} for every super-block K (a dense
//     S nb x S nb lower triangle), built from the tile inverses W_kk by
//     recursive doubling: [A 0; B C]^{-1} = [A^-1 0; -C^-1 (B A^-1) C^-1],
//     log2(S) levels of two batched products (sbinv_*);
//   * a sweep is then two launches per super-block: Y = op(V_K) B_K into a
//     scratch slab, and one launch that updates every block row still to be
//     solved with K = S nb (B(m) -= L(m,K) Y forward, B(m) -= L(K,m)^T Y
//     backward) and copies Y back into B_K.
// mt / S + mt / S launches per sweep instead of mt, each with S times the
// work.  Same lower-equivalent convention as the other kernels (uplo ==
// PlasmaUpper reads L = U^T).  Requires nb a multiple of 64.
#include <stdlib.h>

#include "common.hip.h"
#include "skinny.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NT = 256;
constexpr int BR = 64, BC = 32, KC = 32;      // rows x right-hand sides

template <typename T, int KCV = KC>
using Gemm = GemmNT<T, BC, BR, KCV, 16, 32>;

// Rows of super-block K (S tiles from tile K S, the last one clipped).
__device__ __forceinline__ int sb_rows(const plasma_desc_t &A, int S, int K)
{
    const int t0 = K * S, t1 = min(A.mt, t0 + S);
    return (t1 - t0 - 1) * A.mb + plasma_tile_mview(A, t1 - 1);
}

// acc (transposed product: rows of the result along acc_col) stored into C.
template <typename T, typename G = Gemm<T>>
__device__ __forceinline__ void store_block(
    const typename G::acc_t (&acc)[G::FM][G::FN], const View<T> &C, int r0,
    int c0, int mv, int nv, bool subtract)
{
#pragma unroll
    for (int i = 0; i < G::FM; i++)
#pragma unroll
        for (int j = 0; j < G::FN; j++) {
            const int rr = r0 + G::acc_col(j);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int cc = c0 + G::acc_row(i, q);
                if (rr < mv && cc < nv) {
                    if (subtract) C.at(rr, cc) -= acc[i][j][q];
                    else          C.at(rr, cc) = acc[i][j][q];
                }
            }
        }
}

// ---------------------------------------------------------------------------
// V_K: slab K of V (ld = S nb, column major).  Init: every entry zero except
// the diagonal tiles, which get the lower triangles of W_kk.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NT) void sbinv_init_kernel(plasma_desc_t A,
                                                        const T *inv, T *V,
                                                        int S,
                                                        const int *status)
{
    if (block_should_stop(status)) return;
    const int i = blockIdx.x, j = blockIdx.y, K = blockIdx.z;
    const int nb = A.mb, ld = S * nb;
    T *slab = V + (size_t)K * ld * ld + (size_t)i * nb + (size_t)j * nb * ld;
    const int k = K * S + i;
    const bool diag = i == j && k < A.mt;
    const int mv = diag ? plasma_tile_mview(A, k) : 0;
    const T *w = inv + (size_t)k * A.mb * A.mb;
    for (int e = threadIdx.x; e < nb * nb; e += NT) {
        const int r = e % nb, c = e / nb;
        T v = T(0);
        if (diag && r >= c && r < mv) v = w[r + (size_t)c * A.mb];
        slab[r + (size_t)c * ld] = v;
    }
}

// One doubling level, blocks of h tiles: in every super-block K, the pairs
// p of aligned blocks A = tiles [2ph, 2ph + h), C = [2ph + h, 2ph + 2h)
// (clipped to the super-block): phase 0 forms T = L(C, A) V(A, A) into the
// scratch, phase 1 V(C, A) = -V(C, C) T (V(C, A) is zero before).  Jobs are
// 64-row x 32-column blocks of T / V(C, A).
template <typename T>
__global__ __launch_bounds__(NT) void sbinv_level_kernel(
    plasma_desc_t A, int upper, T *V, T *Tscr, int S, int h, int phase,
    int npairs, int nrb, int ncb, const int *status)
{
    typedef Gemm<T> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int nb = A.mb, ld = S * nb, hb = h * nb;
    int b = blockIdx.x;
    const int cb = b % ncb;  b /= ncb;
    const int rb = b % nrb;  b /= nrb;
    const int p = b % npairs;
    const int K = b / npairs;
    const int St = min(S, A.mt - K * S);
    const int a0 = 2 * p * h, ct = a0 + h;
    if (ct >= St) return;
    const int ch = min(h, St - ct);
    const int crows = (ch - 1) * nb + plasma_tile_mview(A, K * S + ct + ch - 1);
    const int r0 = rb * BR, c0 = cb * BC;
    if (r0 >= crows) return;
    T *slab = V + (size_t)K * ld * ld;
    T *Tp = Tscr + ((size_t)K * npairs + p) * hb * hb;
    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    if (phase == 0) {
        // T(r, c) = sum_kt L(C tile of r, A tile kt)(r, :) V(A rows kt, c)
        const int gi = K * S + ct + r0 / nb;
        for (int kt = 0; kt < h; kt++) {
            View<T> X = lower_view<T>(A, upper != 0, gi, K * S + a0 + kt);
            X.p += (ptrdiff_t)(r0 % nb) * X.rs;
            View<T> Y;                       // Y(c, kk) = V(a0 nb + kt nb + kk, a0 nb + c0 + c)
            Y.p = slab + (size_t)(a0 + kt) * nb + (size_t)(a0 * nb + c0) * ld;
            Y.rs = ld;
            Y.cs = 1;
            G::run(acc, Y, hb - c0, X, min(BR, crows - r0), nb, sm);
        }
        View<T> C;
        C.p = Tp;
        C.rs = 1;
        C.cs = hb;
        store_block<T>(acc, C, r0, c0, crows, hb, false);
    }
    else {
        // V(C, A)(r, c) = - sum_kk V(C, C)(r, kk) T(kk, c), kk <= r
        const int khi = min(r0 + BR, crows);
        View<T> X;                           // X(r, kk) = V(ct nb + r0 + r, ct nb + kk)
        X.p = slab + (size_t)ct * nb + r0 + (size_t)ct * nb * ld;
        X.rs = 1;
        X.cs = ld;
        View<T> Y;                           // Y(c, kk) = T(kk, c0 + c)
        Y.p = Tp + (size_t)c0 * hb;
        Y.rs = hb;
        Y.cs = 1;
        G::run(acc, Y, hb - c0, X, min(BR, crows - r0), khi, sm);
        View<T> C;                           // V(ct nb + r, a0 nb + c)
        C.p = slab + (size_t)ct * nb + (size_t)a0 * nb * ld;
        C.rs = 1;
        C.cs = ld;
        store_block<T>(acc, C, r0, c0, crows, hb, true);
    }
}

// ---------------------------------------------------------------------------
// Sweep step K, diagonal part: Y(r, c) = op(V_K)(r, :) B_K(:, c) into the
// scratch slab (ld = S nb, global right-hand-side column c).  Forward
// op(V) = V (lower: k < r0 + 64), backward V^T (k >= r0).  B_K spans S
// tiles of B, read tile by tile.
// ---------------------------------------------------------------------------
template <typename T, int KCV>
__global__ __launch_bounds__(NT) void ssweep_diag_kernel(
    plasma_desc_t B, const T *V, T *Y, int S, int K, int backward,
    int cchunks, const int *status)
{
    typedef Gemm<T, KCV> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int nb = B.mb, ld = S * nb;
    const int rows = sb_rows(B, S, K);
    const int r0 = blockIdx.x * BR;
    const int tn = blockIdx.y / cchunks;
    const int c0 = (blockIdx.y % cchunks) * BC;
    const int nv = plasma_tile_nview(B, tn);
    if (r0 >= rows || c0 >= nv) return;
    const T *slab = V + (size_t)K * ld * ld;
    const int klo = backward ? r0 : 0;
    const int khi = backward ? rows : min(r0 + BR, rows);
    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    for (int kt = klo / nb; kt * nb < khi; kt++) {
        const int kb = max(klo, kt * nb), ke = min(khi, (kt + 1) * nb);
        View<T> X;                           // X(r, kk) = op(V)(r0 + r, kb + kk)
        if (!backward) {
            X.p = const_cast<T *>(slab) + r0 + (size_t)kb * ld;
            X.rs = 1;
            X.cs = ld;
        }
        else {
            X.p = const_cast<T *>(slab) + kb + (size_t)r0 * ld;
            X.rs = ld;
            X.cs = 1;
        }
        const View<T> Bt = tile_view<T>(B, K * S + kt, tn);
        View<T> Yv;                          // Yv(c, kk) = B_K(kb + kk, c0 + c)
        Yv.p = Bt.p + (kb - kt * nb) + (ptrdiff_t)c0 * Bt.cs;
        Yv.rs = Bt.cs;
        Yv.cs = 1;
        G::run(acc, Yv, nv - c0, X, min(BR, rows - r0), ke - kb, sm);
    }
    View<T> C;
    C.p = Y + (size_t)tn * B.nb * ld;
    C.rs = 1;
    C.cs = ld;
    store_block<T, G>(acc, C, r0, c0, rows, nv, false);
}

// Sweep step K, update part.  Forward: B(m) -= sum_kt L(m, K S + kt) Y_kt
// for tiles m >= (K + 1) S; backward: B(m) -= sum_kt L(K S + kt, m)^T Y_kt
// for m < K S.  The workgroups past the update copy Y into B_K.
template <typename T, int KCV>
__global__ __launch_bounds__(NT) void ssweep_update_kernel(
    plasma_desc_t A, plasma_desc_t B, int upper, const T *Y, int S, int K,
    int backward, int ntiles, int rchunks, int cchunks, const int *status)
{
    typedef Gemm<T, KCV> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int nb = B.mb, ld = S * nb;
    const int tn = blockIdx.y / cchunks;
    const int c0 = (blockIdx.y % cchunks) * BC;
    const int nv = plasma_tile_nview(B, tn);
    const T *Ys = Y + (size_t)tn * B.nb * ld;
    const int St = min(S, B.mt - K * S);

    if ((int)blockIdx.x >= ntiles * rchunks) {
        // copy Y -> B_K
        const int rows = sb_rows(B, S, K);
        const int r0 = ((int)blockIdx.x - ntiles * rchunks) * BR;
        if (r0 >= rows || c0 >= nv) return;
        for (int e = threadIdx.x; e < BR * BC; e += NT) {
            const int r = r0 + e % BR, c = c0 + e / BR;
            if (r < rows && c < nv) {
                const View<T> Bt = tile_view<T>(B, K * S + r / nb, tn);
                Bt.at(r % nb, c) = Ys[r + (size_t)c * ld];
            }
        }
        return;
    }
    const int m = backward ? (int)(blockIdx.x / rchunks)
                           : (K + 1) * S + (int)(blockIdx.x / rchunks);
    const int r0 = (blockIdx.x % rchunks) * BR;
    const int mv = plasma_tile_mview(B, m);
    if (r0 >= mv || c0 >= nv) return;
    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    for (int kt = 0; kt < St; kt++) {
        const int gk = K * S + kt;
        View<T> X;
        if (!backward) {                     // X(r, kk) = L(m, gk)(r0 + r, kk)
            X = lower_view<T>(A, upper != 0, m, gk);
            X.p += (ptrdiff_t)r0 * X.rs;
        }
        else {                               // X(r, kk) = L(gk, m)(kk, r0 + r)
            const View<T> L = lower_view<T>(A, upper != 0, gk, m);
            X.p = L.p + (ptrdiff_t)r0 * L.cs;
            X.rs = L.cs;
            X.cs = L.rs;
        }
        View<T> Yv;                          // Yv(c, kk) = Y(kt nb + kk, c0 + c)
        Yv.p = const_cast<T *>(Ys) + (size_t)kt * nb + (size_t)c0 * ld;
        Yv.rs = ld;
        Yv.cs = 1;
        G::run(acc, Yv, nv - c0, X, mv - r0, plasma_tile_mview(B, gk), sm);
    }
    store_block<T, G>(acc, tile_view<T>(B, m, tn), r0, c0, mv, nv, true);
}

// FP32 forms of the two sweep kernels on the skinny block product
// (skinny.hip.h): the K range of a block split across 8 waves, operands
// loaded straight into the MFMA layout.  XROWS: the factor operand has
// unit row stride (forward sweep of lower storage, backward of upper).
template <bool XROWS>
__global__ __launch_bounds__(skinny::NT) void ssweep2_diag_kernel(
    plasma_desc_t B, const float *V, float *Y, int S, int K, int backward,
    int cchunks, const int *status)
{
    __shared__ skinny::Smem sm;
    if (block_should_stop(status)) return;
    const int nb = B.mb, ld = S * nb;
    const int rows = sb_rows(B, S, K);
    const int r0 = blockIdx.x * BR;
    const int tn = blockIdx.y / cchunks;
    const int c0 = (blockIdx.y % cchunks) * BC;
    const int nv = plasma_tile_nview(B, tn);
    if (r0 >= rows || c0 >= nv) return;
    const float *slab = V + (size_t)K * ld * ld;
    const int klo = backward ? r0 : 0;
    const int khi = backward ? rows : min(r0 + BR, rows);
    auto xf = [&](int k) {                   // op(V)(r0 + r, k)
        skinny::XStep x;
        if (!backward) {
            x.p = slab + r0 + (size_t)k * ld;
            x.rs = 1;
            x.ks = ld;
        }
        else {
            x.p = slab + k + (size_t)r0 * ld;
            x.rs = ld;
            x.ks = 1;
        }
        return x;
    };
    auto yf = [&](int k) {                   // B_K(k, c0 + c)
        const int kt = k / nb;
        const View<float> Bt = tile_view<float>(B, K * S + kt, tn);
        skinny::YStep y;
        y.p = Bt.p + (k - kt * nb) + (ptrdiff_t)c0 * Bt.cs;
        y.cs = Bt.cs;
        return y;
    };
    skinny::block<XROWS>(sm, klo, khi, rows - r0, nv - c0, xf, yf);
    skinny::store(sm, Y + (size_t)tn * B.nb * ld + r0 + (size_t)c0 * ld, ld,
                  rows - r0, nv - c0, false);
}

template <bool XROWS>
__global__ __launch_bounds__(skinny::NT) void ssweep2_update_kernel(
    plasma_desc_t A, plasma_desc_t B, int upper, const float *Y, int S,
    int K, int backward, int ntiles, int rchunks, int cchunks,
    const int *status)
{
    __shared__ skinny::Smem sm;
    if (block_should_stop(status)) return;
    const int nb = B.mb, ld = S * nb;
    const int tn = blockIdx.y / cchunks;
    const int c0 = (blockIdx.y % cchunks) * BC;
    const int nv = plasma_tile_nview(B, tn);
    const float *Ys = Y + (size_t)tn * B.nb * ld;
    const int rows = sb_rows(B, S, K);

    if ((int)blockIdx.x >= ntiles * rchunks) {
        // copy Y -> B_K
        const int r0 = ((int)blockIdx.x - ntiles * rchunks) * BR;
        if (r0 >= rows || c0 >= nv) return;
        for (int e = threadIdx.x; e < BR * BC; e += skinny::NT) {
            const int r = r0 + e % BR, c = c0 + e / BR;
            if (r < rows && c < nv) {
                const View<float> Bt = tile_view<float>(B, K * S + r / nb, tn);
                Bt.at(r % nb, c) = Ys[r + (size_t)c * ld];
            }
        }
        return;
    }
    const int m = backward ? (int)(blockIdx.x / rchunks)
                           : (K + 1) * S + (int)(blockIdx.x / rchunks);
    const int r0 = (blockIdx.x % rchunks) * BR;
    const int mv = plasma_tile_mview(B, m);
    if (r0 >= mv || c0 >= nv) return;
    auto xf = [&](int k) {
        const int kt = k / nb, kk = k - kt * nb, gk = K * S + kt;
        skinny::XStep x;
        if (!backward) {                     // L(m, gk)(r0 + r, kk)
            const View<float> L = lower_view<float>(A, upper != 0, m, gk);
            x.p = L.p + (ptrdiff_t)r0 * L.rs + (ptrdiff_t)kk * L.cs;
            x.rs = L.rs;
            x.ks = L.cs;
        }
        else {                               // L(gk, m)(kk, r0 + r)
            const View<float> L = lower_view<float>(A, upper != 0, gk, m);
            x.p = L.p + (ptrdiff_t)kk * L.rs + (ptrdiff_t)r0 * L.cs;
            x.rs = L.cs;
            x.ks = L.rs;
        }
        return x;
    };
    auto yf = [&](int k) {                   // Y(k, c0 + c)
        skinny::YStep y;
        y.p = Ys + k + (size_t)c0 * ld;
        y.cs = ld;
        return y;
    };
    skinny::block<XROWS>(sm, 0, rows, mv - r0, nv - c0, xf, yf);
    const View<float> Bt = tile_view<float>(B, m, tn);
    skinny::store(sm, Bt.p + r0 + (ptrdiff_t)c0 * Bt.cs, Bt.cs, mv - r0,
                  nv - c0, true);
}

template <typename T>
int launch_sbinv(plasma_enum_t uplo, plasma_desc_t A, const void *inv,
                 void *V, void *Tscr, int S, const int *status, hipStream_t s)
{
    const int Q = (A.mt + S - 1) / S;
    hipLaunchKernelGGL(sbinv_init_kernel<T>, dim3(S, S, Q), dim3(NT), 0, s, A,
                       (const T *)inv, (T *)V, S, status);
    for (int h = 1; h < S; h *= 2) {
        const int npairs = (S + 2 * h - 1) / (2 * h);
        const int nrb = h * A.mb / BR, ncb = h * A.mb / BC;
        const long jobs = (long)Q * npairs * nrb * ncb;
        if (jobs > 0x7fffffffL) return PlasmaErrorNotSupported;
        for (int phase = 0; phase < 2; phase++)
            hipLaunchKernelGGL(sbinv_level_kernel<T>, dim3((unsigned)jobs),
                               dim3(NT), 0, s, A, uplo == PlasmaUpper ? 1 : 0,
                               (T *)V, (T *)Tscr, S, h, phase, npairs, nrb,
                               ncb, status);
    }
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T, int KCV>
void launch_ssweep_kc(plasma_enum_t uplo, int backward, plasma_desc_t A,
                      plasma_desc_t B, const void *V, void *Y, int S, int K,
                      const int *status, hipStream_t s)
{
    const int cchunks = (min(B.nb, B.n) + BC - 1) / BC;
    const int rows_sb = S * B.mb;
    hipLaunchKernelGGL((ssweep_diag_kernel<T, KCV>),
                       dim3((rows_sb + BR - 1) / BR, B.nt * cchunks),
                       dim3(NT), 0, s, B, (const T *)V, (T *)Y, S, K,
                       backward, cchunks, status);
    const int ntiles = backward ? K * S : max(B.mt - (K + 1) * S, 0);
    const int rchunks = (B.mb + BR - 1) / BR;
    const int ncopy = (rows_sb + BR - 1) / BR;
    hipLaunchKernelGGL((ssweep_update_kernel<T, KCV>),
                       dim3(ntiles * rchunks + ncopy, B.nt * cchunks),
                       dim3(NT), 0, s, A, B, uplo == PlasmaUpper ? 1 : 0,
                       (const T *)Y, S, K, backward, ntiles, rchunks, cchunks,
                       status);
}

// K chunk depth of the sweep kernels: a workgroup walks K = S nb in
// chunks with one chunk prefetched, so deeper chunks shorten its chain of
// memory round trips (PLASMA_SWEEP_KC: 32, 64 or 128).
int sweep_kc()
{
    static int kc = -1;
    if (kc < 0) {
        const char *e = getenv("PLASMA_SWEEP_KC");
        kc = e != nullptr ? atoi(e) : 64;
        if (kc != 32 && kc != 64 && kc != 128) kc = 64;
    }
    return kc;
}

template <bool XROWS>
void launch_ssweep2(plasma_enum_t uplo, int backward, plasma_desc_t A,
                    plasma_desc_t B, const void *V, void *Y, int S, int K,
                    const int *status, hipStream_t s)
{
    const int cchunks = (min(B.nb, B.n) + BC - 1) / BC;
    const int rows_sb = S * B.mb;
    // the diagonal product reads V (forward: unit row stride) or V^T
    if (!backward)
        hipLaunchKernelGGL(ssweep2_diag_kernel<true>,
                           dim3((rows_sb + BR - 1) / BR, B.nt * cchunks),
                           dim3(skinny::NT), 0, s, B, (const float *)V,
                           (float *)Y, S, K, backward, cchunks, status);
    else
        hipLaunchKernelGGL(ssweep2_diag_kernel<false>,
                           dim3((rows_sb + BR - 1) / BR, B.nt * cchunks),
                           dim3(skinny::NT), 0, s, B, (const float *)V,
                           (float *)Y, S, K, backward, cchunks, status);
    const int ntiles = backward ? K * S : max(B.mt - (K + 1) * S, 0);
    const int rchunks = (B.mb + BR - 1) / BR;
    const int ncopy = (rows_sb + BR - 1) / BR;
    hipLaunchKernelGGL(ssweep2_update_kernel<XROWS>,
                       dim3(ntiles * rchunks + ncopy, B.nt * cchunks),
                       dim3(skinny::NT), 0, s, A, B,
                       uplo == PlasmaUpper ? 1 : 0, (const float *)Y, S, K,
                       backward, ntiles, rchunks, cchunks, status);
}

// FP32 sweeps: the skinny block product unless PLASMA_SWEEP_KERNEL=gemm
// selects the LDS-staged GemmNT kernels (kept for comparison)
bool sweep_skinny()
{
    static int on = -1;
    if (on < 0) {
        const char *e = getenv("PLASMA_SWEEP_KERNEL");
        on = !(e != nullptr && e[0] == 'g');
    }
    return on != 0;
}

template <typename T>
int launch_ssweep(plasma_enum_t uplo, int backward, plasma_desc_t A,
                  plasma_desc_t B, const void *V, void *Y, int S, int K,
                  const int *status, hipStream_t s)
{
    if (sizeof(T) == 4 && sweep_skinny()) {
        // the factor operand has unit row stride for the forward sweep of
        // lower storage and the backward sweep of upper storage
        if ((backward != 0) == (uplo == PlasmaUpper))
            launch_ssweep2<true>(uplo, backward, A, B, V, Y, S, K, status, s);
        else
            launch_ssweep2<false>(uplo, backward, A, B, V, Y, S, K, status,
                                  s);
        return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
    }
    switch (sweep_kc()) {
    case 32:
        launch_ssweep_kc<T, 32>(uplo, backward, A, B, V, Y, S, K, status, s);
        break;
    case 128:   // FP32 only: 128-deep FP64 chunks exceed the LDS
        if constexpr (sizeof(T) == 4)
            launch_ssweep_kc<T, 128>(uplo, backward, A, B, V, Y, S, K, status,
                                     s);
        else
            launch_ssweep_kc<T, 64>(uplo, backward, A, B, V, Y, S, K, status,
                                    s);
        break;
    default:
        launch_ssweep_kc<T, 64>(uplo, backward, A, B, V, Y, S, K, status, s);
    }
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

}  // namespace

extern "C" size_t plasma_sbinv_scratch_elems(plasma_desc_t A, int S)
{
    const int Q = (A.mt + S - 1) / S;
    size_t most = 0;
    for (int h = 1; h < S; h *= 2) {
        const size_t npairs = (S + 2 * h - 1) / (2 * h);
        const size_t hb = (size_t)h * A.mb;
        most = max(most, npairs * hb * hb);
    }
    return (size_t)Q * most;
}

extern "C" int plasma_kern_sbinv(plasma_enum_t uplo, plasma_desc_t A,
                                 const void *inv, void *V, void *Tscr, int S,
                                 const int *status, hipStream_t s)
{
    if (A.mb != A.nb || A.mb % BR != 0 || S < 1)
        return PlasmaErrorNotSupported;
    if (A.precision == PlasmaRealDouble)
        return launch_sbinv<double>(uplo, A, inv, V, Tscr, S, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_sbinv<float>(uplo, A, inv, V, Tscr, S, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_ssweep(plasma_enum_t uplo, int backward,
                                  plasma_desc_t A, plasma_desc_t B,
                                  const void *V, void *Y, int S, int K,
                                  const int *status, hipStream_t s)
{
    if (A.precision != B.precision || B.mb != A.mb || A.mb % BR != 0)
        return PlasmaErrorNotSupported;
    if (A.precision == PlasmaRealDouble)
        return launch_ssweep<double>(uplo, backward, A, B, V, Y, S, K, status,
                                     s);
    if (A.precision == PlasmaRealFloat)
        return launch_ssweep<float>(uplo, backward, A, B, V, Y, S, K, status,
                                    s);
    return PlasmaErrorNotSupported;
}
