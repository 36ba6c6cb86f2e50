// potrf.hip -- tile Cholesky kernels for gfx950: the diagonal tile
// factorization (with the inverse of the factor), the panel triangular
// solve and the MFMA trailing update.
//
// They replace the tile tasks of the reference's right-looking tile
// Cholesky (compute/pzpotrf.c:37-80 lower, :84-126 upper): one launch per
// step and task class processes every tile of that step, instead of one
// OpenMP task per tile.  All kernels use the lower-equivalent convention of
// common.hip.h (upper storage is read as L = U^T).
#include "common.hip.h"
#include "tilegemm.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NTHREADS = 256;
// Threads of the diagonal-tile factorization and its minimum waves per
// SIMD (the second __launch_bounds__ argument).  It runs beside the
// trailing update, whose workgroups take 256 registers per lane on each
// SIMD (one wave each): a workgroup that needs the whole register file of
// its CU can only start on a CU the update has left entirely, i.e. near the
// update's end.  Both precisions run 4 waves held to 256 registers (at
// least 2 waves per SIMD), so they fit beside one update workgroup.  FP32
// fits without spills (n = 32768 eager, lookahead 1: factor 134.7 ->
// 125.2 ms); FP64 spills 10 registers and still gains (depth 3: 244.9 ->
// 236.2 ms eager, 240.1 -> 234.8 ms graph).
template <typename T> struct DiagNT {
    static constexpr int v = 256, min_waves = 2;
};
constexpr int JB = 32;          // column block of the diagonal tile

// ---------------------------------------------------------------------------
// Wave-level 32 x 32 kernels.  Lane l holds row r = l % 32 of a block in
// registers d[0..31] (lanes l and l + 32 compute the same row and store
// the same values, so the code has no divergent branches, which keeps the
// unrolled register code free of spills); rows and columns beyond jb are
// padded with the identity.  Values of other rows are broadcast through
// LDS: one store per row, then same-address reads, which the LDS serves to
// all lanes at once.  Blocks move between global memory and registers
// through an LDS copy (column major, stride LDB = JB + 1), so the register
// code addresses memory with immediate offsets only.  Only the calling wave
// touches its LDS buffers; the wavefront-scope fences order its LDS
// accesses across lanes (a wave's LDS operations complete in program
// order).
// ---------------------------------------------------------------------------
constexpr int LDB = JB + 1;

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Pivot arithmetic.  FP32 uses the hardware square root and reciprocal
// (1 ulp, as LAPACK's xPOTF2 it scales the column by the reciprocal of the
// pivot); FP64 keeps the correctly rounded operations.
__device__ __forceinline__ float pivot_sqrt(float x)
{
    return __builtin_amdgcn_sqrtf(x);
}
__device__ __forceinline__ double pivot_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float pivot_rcp(float x)
{
    return __builtin_amdgcn_rcpf(x);
}
__device__ __forceinline__ double pivot_rcp(double x) { return 1.0 / x; }

// Value of x in lane `src` (a compile-time lane after unrolling), read into
// a scalar register.
__device__ __forceinline__ float lane_value(float x, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}
__device__ __forceinline__ double lane_value(double x, int src)
{
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)v, src);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Right-looking Cholesky of the block held in d; returns 0 or the 1-based
// column of the first non-positive (or NaN) pivot (LAPACK dpotf2 rule).
// The column loop is straight-line code: the pivot comes from lane j by
// readlane, every lane updates all of its trailing columns (entries above
// the diagonal receive values that are never read: the caller keeps the
// lower triangle only), and a failed pivot is recorded, not returned from
// inside the loop.  Column j of L reaches the other lanes by readlane; sb
// (2 * JB elements of LDS) is not used.
template <typename T>
__device__ __forceinline__ int wave_potrf32(T (&d)[JB], int jb, T *sb)
{
    const int r = threadIdx.x & 31;
    int info = 0;
#pragma unroll
    for (int j = 0; j < JB; j++) {
        const T djj = lane_value(d[j], j);
        if (info == 0 && j < jb && !(djj > T(0))) info = j + 1;
        const T ljj = pivot_sqrt(djj);
        const T rl = pivot_rcp(ljj);
        const T lrj = (r == j) ? ljj : d[j] * rl;
        d[j] = lrj;
        // column j of L from its lanes by readlane (lane c holds L(c, j)):
        // no LDS round trip on the pivot chain
#pragma unroll
        for (int c = j + 1; c < JB; c++) d[c] -= lrj * lane_value(lrj, c);
    }
    (void)sb;
    return info;
}

// Inverse of the lower-triangular block held in d (rows on lanes): lane l
// returns column l % 32 of L^{-1} in inv[0..31].  Row r of L comes from
// lane r by readlane, so nothing goes through LDS (sl is not used).
template <typename T>
__device__ __forceinline__ void wave_trtri32(const T (&d)[JB], T (&inv)[JB],
                                             T *sl)
{
    const int cl = threadIdx.x & 31;
    (void)sl;
#pragma unroll
    for (int r = 0; r < JB; r++) {
        // two partial sums halve the dependent chain of the row
        T s0 = (cl == r) ? T(1) : T(0), s1 = T(0);
#pragma unroll
        for (int l = 0; l + 1 < r; l += 2) {
            s0 -= lane_value(d[l], r) * inv[l];
            s1 -= lane_value(d[l + 1], r) * inv[l + 1];
        }
        if (r & 1) s0 -= lane_value(d[r - 1], r) * inv[r - 1];
        inv[r] = (s0 + s1) * pivot_rcp(lane_value(d[r], r));
    }
}

// sD <- the jb x jb diagonal block at (j0, j0) of the tile view L: lower
// triangle, identity padding; two columns per pass, coalesced, in-bounds
// addresses with the bounds applied to the values.
template <typename T>
__device__ __forceinline__ void wave_load_block(const View<T> &L, T *sD,
                                                int j0, int jb)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 31, h = lane >> 5;
    // all 16 loads of the lane in flight before the first LDS store
    T x[JB / 2];
#pragma unroll
    for (int u = 0; u < JB / 2; u++) {
        const int c = 2 * u + h;
        x[u] = L.at(j0 + min(r, jb - 1), j0 + min(c, jb - 1));
    }
#pragma unroll
    for (int u = 0; u < JB / 2; u++) {
        const int c = 2 * u + h;
        T v;
        if (r < jb && c < jb) v = (c <= r) ? x[u] : T(0);
        else                  v = (r == c) ? T(1) : T(0);
        sD[c * LDB + r] = v;
    }
    wave_sync_lds();
}

// Registers -> LDS block: row lane % 32 of d into sD (column major).
template <typename T>
__device__ __forceinline__ void wave_rows_to_lds(const T (&d)[JB], T *sD)
{
    const int r = threadIdx.x & 31;
#pragma unroll
    for (int c = 0; c < JB; c++) sD[c * LDB + r] = d[c];
}

// LDS block -> registers: row `lane` of sD into d.
template <typename T>
__device__ __forceinline__ void wave_rows_from_lds(T (&d)[JB], const T *sD)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < JB; c++) d[c] = sD[c * LDB + (lane & 31)];
}

// One 16 x 16 MFMA block: acc += A(16 x K) B(16 x K)^T with the operands
// given as Views positioned at their first row; rows beyond the bounds read
// as zero.  Every lane loads from a clamped, in-bounds address and the
// bounds select the value afterwards, so the loads issue back to back
// instead of each sitting in its own branch.  Result element q of the lane is at (Mfma<T>::row(lane, q),
// lane % 16).  Operands are fetched 32 k-values at a time before the MFMAs
// so the loads of a chunk overlap.
template <typename T>
__device__ __forceinline__ void wave_mma_nt(typename Mfma<T>::acc_t &acc,
                                            const View<T> &Av, int arows,
                                            const View<T> &Bv, int brows,
                                            int K)
{
    const int lane = threadIdx.x & 63;
    const int r = lane % 16, kq = lane / 16;
    for (int k = 0; k < K; k += 32) {
        T a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int kk = k + 4 * u + kq;
            const T x = Av.at(min(r, arows - 1), min(kk, K - 1));
            const T y = Bv.at(min(r, brows - 1), min(kk, K - 1));
            a[u] = (r < arows && kk < K) ? x : T(0);
            b[u] = (r < brows && kk < K) ? y : T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) acc = Mfma<T>::mma(a[u], b[u], acc);
    }
}

// 2 x 2 blocks at once: acc[u][v] += A_u B_v^T with A_u = Av rows 16u.., B_v
// = Bv rows 16v..; the loads of all four blocks of a 32-k chunk are issued
// before their MFMAs.
template <typename T>
__device__ __forceinline__ void wave_mma_nt22(typename Mfma<T>::acc_t (&acc)[2][2],
                                              const View<T> &Av, int arows,
                                              const View<T> &Bv, int brows,
                                              int K)
{
    const int lane = threadIdx.x & 63;
    const int r = lane % 16, kq = lane / 16;
    for (int k = 0; k < K; k += 32) {
        T a[2][8], b[2][8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int kk = k + 4 * u + kq;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int ra = r + 16 * h;
                const T x = Av.at(min(ra, arows - 1), min(kk, K - 1));
                const T y = Bv.at(min(ra, brows - 1), min(kk, K - 1));
                a[h][u] = (ra < arows && kk < K) ? x : T(0);
                b[h][u] = (ra < brows && kk < K) ? y : T(0);
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int g = 0; g < 2; g++)
                    acc[h][g] = Mfma<T>::mma(a[h][u], b[g][u], acc[h][g]);
    }
}

template <typename T>
__device__ __forceinline__ View<T> sub(View<T> v, int r, int c)
{
    v.p += (ptrdiff_t)r * v.rs + (ptrdiff_t)c * v.cs;
    return v;
}

// Factor and invert the jb x jb diagonal block at (j0, j0) of the tile view
// L (one wave): the block's L goes to sD (column major, stride LDB) and its
// inverse to sW (W^T column major, i.e. W row major: W(r, c) =
// sW[r * LDB + c]); store_diag_blocks writes both to global memory.  sb: 2 *
// JB elements of LDS.  Returns 0 or the block-local 1-based pivot column.
template <typename T>
__device__ __forceinline__ int diag_block32(const View<T> &L, T *sD, T *sW,
                                            T *sb, int j0, int jb)
{
    wave_load_block(L, sD, j0, jb);
    T d[JB], iv[JB];
    wave_rows_from_lds(d, sD);
    const int info = wave_potrf32(d, jb, sb);
    if (info) return info;
    wave_trtri32(d, iv, sD);                      // sD: rows of L, row major
    wave_rows_to_lds(iv, sW);                     // column lane % 32 of W ...
    wave_rows_to_lds(d, sD);                      // ... and L, column major
    wave_sync_lds();
    return 0;
}

// The factored diagonal block (lower triangle of sD) and its inverse (sW)
// to the tile and the workspace; all threads of the workgroup, after the
// barrier that follows diag_block32.
template <typename T, int NT>
__device__ __forceinline__ void store_diag_blocks(const View<T> &L,
                                                  const View<T> &Wv,
                                                  const T *sD, const T *sW,
                                                  int j0, int jb)
{
    for (int e = threadIdx.x; e < JB * JB; e += NT) {
        const int r = e % JB, c = e / JB;
        if (r < jb && c < jb) {
            if (r >= c) L.at(j0 + r, j0 + c) = sD[c * LDB + r];
            Wv.at(j0 + r, j0 + c) = sW[r * LDB + c];
        }
    }
}

// Invert the factored jb x jb diagonal block at (j0, j0) of L into the
// workspace tile Wv (one wave; sD, sW: LDS block scratch).
template <typename T>
__device__ __forceinline__ void inv_block32(const View<T> &L, const View<T> &Wv,
                                            T *sD, T *sW, int j0, int jb)
{
    const int lane = threadIdx.x & 63;
    wave_load_block(L, sD, j0, jb);
    T d[JB], iv[JB];
    wave_rows_from_lds(d, sD);
    wave_trtri32(d, iv, sD);
    wave_rows_to_lds(iv, sW);
    wave_sync_lds();
    const int r = lane & 31, h = lane >> 5;
    for (int c = h; c < jb; c += 2)
        if (r < jb) Wv.at(j0 + r, j0 + c) = sW[r * LDB + c];
}

// ---------------------------------------------------------------------------
// One workgroup factors a tile of order n <= 512: L L^T = A by 32-column
// blocks (a wave factors and inverts the 32 x 32 block in registers, the
// rows below are multiplied by the block's inverse on MFMA, and the rest of
// the tile is updated on MFMA); then W = L^{-1} (lower, zeros above) by
// block columns: W_JJ = L_JJ^{-1}, W_IJ = -W_II sum_{K=J}^{I-1} L_IK W_KJ.
// ---------------------------------------------------------------------------
// mode 0: invert an already factored tile; 1: factor and invert; 2:
// factor, with only the 32 x 32 diagonal blocks of W inverted.  L: the tile
// (lower-equivalent view), n its order; Wv: the n x n workspace of W (unit
// row stride); sP: dynamic LDS of max(n, 8 * JB) x (JB + 1) elements.
// Returns 0 or the 1-based column of the first failed pivot (the same value
// in every thread).  All threads of the workgroup must call it.
template <typename T, int NT = NTHREADS>
__device__ int factor_tile(const View<T> &L, const int n, const View<T> &Wv,
                           const int mode, T *sP)
{
    constexpr int NW = NT / 64;          // waves of the workgroup
    const int nbb = (n + JB - 1) / JB;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    typedef typename Mfma<T>::acc_t acc_t;

    constexpr int LDP = JB + 1;          // padded row stride in LDS
    __shared__ T sDinv[JB * LDB];       // inverse of the current block
    __shared__ T sDblk[JB * LDB];       // the current block, for wave 0
    __shared__ T sB[2 * JB];            // broadcast buffer of wave 0
    __shared__ int s_info;
    if (t == 0) s_info = 0;

    View<T> Dv;                          // Dv(c, l) = Dinv(c, l)
    Dv.p = sDinv;                        // row major, stride LDB
    Dv.rs = LDB;
    Dv.cs = 1;
    View<T> Pv;                          // Pv(r, l) = panel row r, column l
    Pv.p = sP;
    Pv.rs = LDP;
    Pv.cs = 1;

    if (mode != 0) {
        for (int J = 0; J < nbb; J++) {
            const int j0 = J * JB, jb = min(JB, n - j0);
            const int i0 = j0 + jb, nrows = n - i0;
            // (a) one wave: factor and invert the diagonal block
            if (wave == 0) {
                const int info = diag_block32<T>(L, sDblk, sDinv, sB, j0,
                                                 jb);
                if (info && lane == 0) s_info = j0 + info;
            }
            else {
                // the other waves stage the panel rows below the block,
                // four loads in flight per thread
                constexpr int STEP = NT - 64;
                const int total = nrows * JB;
                for (int e0 = t - 64; e0 < total; e0 += 4 * STEP) {
                    T x[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int e = min(e0 + u * STEP, total - 1);
                        const int r = e % nrows, c = e / nrows;
                        x[u] = L.at(i0 + r, j0 + min(c, jb - 1));
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int e = e0 + u * STEP;
                        const int r = e % nrows, c = e / nrows;
                        if (e < total) sP[r * LDP + c] = (c < jb) ? x[u] : T(0);
                    }
                }
            }
            __syncthreads();
            if (s_info) return s_info;
            store_diag_blocks<T, NT>(L, Wv, sDblk, sDinv, j0, jb);
            // (b) P = A_panel Dinv^T, one 16-row strip per wave (both
            //     16-column halves); P goes back to LDS and to the tile
            for (int st = wave; st * 16 < nrows; st += NW) {
                acc_t a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
                const View<T> P = sub(Pv, st * 16, 0);
                const int pr = min(16, nrows - st * 16);
                wave_mma_nt<T>(a0, P, pr, Dv, 16, JB);
                wave_mma_nt<T>(a1, P, pr, sub(Dv, 16, 0), 16, JB);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int r = Mfma<T>::row(lane, q), c = lane % 16;
                    if (r < pr) {
                        sP[(st * 16 + r) * LDP + c] = a0[q];
                        sP[(st * 16 + r) * LDP + c + 16] = a1[q];
                        if (c < jb) L.at(i0 + st * 16 + r, j0 + c) = a0[q];
                        if (c + 16 < jb)
                            L.at(i0 + st * 16 + r, j0 + 16 + c) = a1[q];
                    }
                }
            }
            __syncthreads();
            // (c) trailing part of the tile: A22 -= P P^T, lower 16x16 blocks;
            //     the C values of a block are fetched before its MFMAs
            //     in groups of UB blocks per wave: the C loads of a group
            //     are issued together, so a wave waits for one memory
            //     round trip per group instead of one per block
            constexpr int UB = 4;
            const int nb16 = (nrows + 15) / 16;
            const int nblk = nb16 * (nb16 + 1) / 2;
            for (int b0 = wave * UB; b0 < nblk; b0 += NW * UB) {
                int bi[UB], bj[UB];
                T cv[UB][4];
#pragma unroll
                for (int u = 0; u < UB; u++) {
                    const int b = min(b0 + u, nblk - 1);
                    int i = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
                    while (i * (i + 1) / 2 > b) i--;
                    while ((i + 1) * (i + 2) / 2 <= b) i++;
                    bi[u] = i;
                    bj[u] = b - i * (i + 1) / 2;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int r = bi[u] * 16 + Mfma<T>::row(lane, q);
                        const int c = bj[u] * 16 + lane % 16;
                        const T x = L.at(i0 + min(r, nrows - 1),
                                         i0 + min(c, nrows - 1));
                        cv[u][q] = (r < nrows && c <= r) ? x : T(0);
                    }
                }
#pragma unroll
                for (int u = 0; u < UB; u++) {
                    if (b0 + u >= nblk) break;
                    acc_t acc = {0, 0, 0, 0};
                    wave_mma_nt<T>(acc, sub(Pv, bi[u] * 16, 0),
                                   nrows - bi[u] * 16, sub(Pv, bj[u] * 16, 0),
                                   nrows - bj[u] * 16, JB);
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int r = bi[u] * 16 + Mfma<T>::row(lane, q);
                        const int c = bj[u] * 16 + lane % 16;
                        if (r < nrows && c <= r)
                            L.at(i0 + r, i0 + c) = cv[u][q] - acc[q];
                    }
                }
            }
            __syncthreads();
        }
    }
    else {
        // invert-only: the diagonal blocks' inverses first
        for (int J = wave; J < nbb; J += NW) {
            const int j0 = J * JB, jb = min(JB, n - j0);
            // per-wave block scratch in the (still unused) panel area
            inv_block32<T>(L, Wv, sP + 2 * wave * JB * LDB,
                           sP + (2 * wave + 1) * JB * LDB, j0, jb);
        }
    }

    if (mode == 2) return 0;

    // zero the strict upper triangle of W
    for (int c = wave; c < n; c += NW)
        for (int r = lane; r < c; r += 64) Wv.at(r, c) = T(0);
    __syncthreads();
    T (*sT)[JB][LDP] = reinterpret_cast<T (*)[JB][LDP]>(sP);  // per wave

    // off-diagonal blocks of W; block columns are paired (J, nbb-1-J) per
    // wave (slots 2w and 2w+1, then 2w+8 and 2w+9) so the longest and
    // shortest dependency chains share a wave
    for (int slot0 = 2 * wave; slot0 < nbb; slot0 += 2 * NW)
    for (int slot = slot0; slot < min(slot0 + 2, nbb); slot++) {
        // slot s -> column s/2 (even s) or nbb - 1 - s/2 (odd s)
        const int J = (slot & 1) ? nbb - 1 - slot / 2 : slot / 2;
        const int j0 = J * JB, jb = min(JB, n - j0);
        for (int I = J + 1; I < nbb; I++) {
            const int r0 = I * JB, ib = min(JB, n - r0);
            // T = sum_{K=J}^{I-1} L_IK W_KJ  (rows of I, columns of J)
            acc_t acc[2][2];
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int v = 0; v < 2; v++) acc[u][v] = acc_t{0, 0, 0, 0};
            View<T> Wt;                 // (c, kk) = W(j0 + kk, j0 + c)
            Wt.p = Wv.p + j0 + (ptrdiff_t)j0 * Wv.cs;
            Wt.rs = Wv.cs;
            Wt.cs = 1;
            wave_mma_nt22<T>(acc, sub(L, r0, j0), ib, Wt, jb, r0 - j0);
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int v = 0; v < 2; v++)
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        sT[wave][u * 16 + Mfma<T>::row(lane, q)][v * 16 + lane % 16] =
                            acc[u][v][q];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // W_IJ = -W_II T
            View<T> Tt;                 // (c, kk) = T(kk, c)
            Tt.p = &sT[wave][0][0];
            Tt.rs = 1;
            Tt.cs = LDP;
            acc_t o[2][2];
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int v = 0; v < 2; v++) o[u][v] = acc_t{0, 0, 0, 0};
            wave_mma_nt22<T>(o, sub(Wv, r0, r0), ib, Tt, jb, ib);
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int v = 0; v < 2; v++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int r = u * 16 + Mfma<T>::row(lane, q);
                        const int c = v * 16 + lane % 16;
                        if (r < ib && c < jb) Wv.at(r0 + r, j0 + c) = -o[u][v][q];
                    }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Diagonal tile k: with `factor`, L L^T = A(k,k) and W = L^{-1}; without,
// W only (one workgroup per tile of a batch).  A failed pivot at local
// column j sets the status word to k*nb + j + 1 (reference
// core_zpotrf.c:72-79, iinfo = A.nb * k).
// ---------------------------------------------------------------------------
template <typename T, int NT, int MINB = 1>
__global__ __launch_bounds__(NT, MINB) void potrf_diag_kernel(
    plasma_desc_t A, int upper, int k0, int factor, T *inv, int *status)
{
    if (block_should_stop(status)) return;
    const int k = k0 + blockIdx.x;
    View<T> Wv;
    Wv.p = inv + (size_t)k * A.mb * A.mb;
    Wv.rs = 1;
    Wv.cs = A.mb;
    extern __shared__ unsigned char smem_raw[];
    T *sP = reinterpret_cast<T *>(smem_raw);
    const int info = factor_tile<T, NT>(lower_view<T>(A, upper != 0, k, k),
                                        plasma_tile_mview(A, k), Wv, factor,
                                        sP);
    if (info && threadIdx.x == 0) atomicCAS(status, 0, A.nb * k + info);
}

// ---------------------------------------------------------------------------
// Job decoders for the block GEMM (tilegemm.hip.h).
// ---------------------------------------------------------------------------

// Trailing update by the panels of a column group: tiles (m, n) of columns
// [n_first, ...) enumerated column by column, bpt x bpt blocks of 128 per
// tile, C(m, n) -= P(m) P(n)^T with P(m) the first `kdepth` columns of the
// group's panel buffer tile m (at panel + (m - pbase - 1) * mb * pcols,
// column major, ld = mb; its columns [j * nb, (j + 1) * nb) hold the solved
// panel of tile column pbase + j).  Both operands have unit row stride for
// either storage triangle.  The first `ncopy` workgroups copy panel columns
// back into A(m, pbase + j), m > pbase + j, j < ncols, instead (transposed
// for upper storage).
template <typename T>
struct UpdateDecode {
    plasma_desc_t A;
    int pbase, pcols, kdepth, ncols, n_first, bpt, upper, ncopy;
    const T *panel;
    __device__ View<T> panel_tile(int m) const
    {
        View<T> v;
        v.p = const_cast<T *>(panel) + (size_t)(m - pbase - 1) * A.mb * pcols;
        v.rs = 1;
        v.cs = A.mb;
        return v;
    }
    __device__ bool operator()(int b, tg::Job<T> &j) const
    {
        if (b < ncopy) {
            // copy job b: tile column pbase + jc, tile m > pbase + jc
            int jc = 0, bb = b;
            while (jc + 1 < ncols && bb >= A.mt - (pbase + jc) - 1) {
                bb -= A.mt - (pbase + jc) - 1;
                jc++;
            }
            const int col = pbase + jc, m = col + 1 + bb;
            const int mv = plasma_tile_mview(A, m), nb = plasma_tile_mview(A, col);
            const View<T> D = lower_view<T>(A, upper != 0, m, col);
            View<T> S = panel_tile(m);
            S.p += (size_t)jc * A.nb * A.mb;
            if (!upper) {
                const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
                for (int c = wave; c < nb; c += tg::NT / 64)
                    for (int r = lane; r < mv; r += 64) D.at(r, c) = S.at(r, c);
            }
            else {
                // D.at(r, c) = p[r * ld + c]: walk c fastest
                for (int e = threadIdx.x; e < mv * nb; e += tg::NT) {
                    const int r = e / nb, c = e % nb;
                    D.at(r, c) = S.at(r, c);
                }
            }
            return false;
        }
        b -= ncopy;
        const int per_tile = bpt * bpt;
        const int sub_ = b % per_tile;
        const int bi = sub_ / bpt, bj = sub_ % bpt;
        int m, n;
        trapezoid_decode(b / per_tile, A.mt, n_first, &m, &n);
        if (m == n && bj > bi) return false;
        const int mv = plasma_tile_mview(A, m), nv = plasma_tile_mview(A, n);
        const int r0 = bi * tg::BM, c0 = bj * tg::BM;
        if (r0 >= mv || c0 >= nv) return false;
        j.X = sub(panel_tile(m), r0, 0);
        j.Y = sub(panel_tile(n), c0, 0);
        j.C = sub(lower_view<T>(A, upper, m, n), r0, c0);
        j.xrows = mv - r0;
        j.yrows = nv - c0;
        j.K = kdepth;
        j.lower = (m == n);
        j.diag_off = c0 - r0;
        return true;
    }
};

// Panel solve of step k: P(m) = A(m,k) W_kk^T for m > k into column block
// `pcol` of the panel buffer tile m (at panel + (m - pbase - 1) * mb * pcols
// + pcol * mb, column major, ld = mb).
template <typename T>
struct PanelDecode {
    plasma_desc_t A;
    int k, bpt, upper, pbase, pcols, pcol;
    const T *inv;
    T *panel;
    __device__ bool operator()(int b, tg::Job<T> &j) const
    {
        const int per_tile = bpt * bpt;
        const int m = k + 1 + b / per_tile;
        const int sub_ = b % per_tile;
        const int bi = sub_ / bpt, bj = sub_ % bpt;
        const int mv = plasma_tile_mview(A, m), nb = plasma_tile_mview(A, k);
        const int r0 = bi * tg::BM, c0 = bj * tg::BM;
        if (r0 >= mv || c0 >= nb) return false;
        View<T> W;
        W.p = const_cast<T *>(inv) + (size_t)k * A.mb * A.mb;
        W.rs = 1;
        W.cs = A.mb;
        View<T> S;
        S.p = panel + (size_t)(m - pbase - 1) * A.mb * pcols +
              (size_t)pcol * A.mb;
        S.rs = 1;
        S.cs = A.mb;
        j.X = sub(lower_view<T>(A, upper, m, k), r0, 0);
        j.Y = sub(W, c0, 0);
        j.C = sub(S, r0, c0);
        j.xrows = mv - r0;
        j.yrows = nb - c0;
        j.K = nb;
        j.lower = false;
        j.diag_off = 0;
        return true;
    }
};

template <typename T>
int launch_potrf_diag(plasma_enum_t uplo, plasma_desc_t A, int k0, int count,
                      int factor, void *inv, int *status, hipStream_t s)
{
    if (count <= 0) return 0;
    // panel rows (mb x 33) in the factor phase; 8 blocks of 32 x 33 for the
    // invert-only form, one per wave in the inverse (8 with the 512-thread
    // factorization); the static part is sDinv, sDblk, sB and the flags
    const size_t lds = (size_t)max(A.mb, 8 * JB) * (JB + 1) * sizeof(T);
    const size_t lds_static = (2 * JB * (JB + 1) + 2 * JB) * sizeof(T) + 16;
    if (lds + lds_static > 160 * 1024) return PlasmaErrorNotSupported;
    // the factorization (one tile, on the critical path) runs 8 waves; the
    // invert-only form (a batch of tiles) 4 per tile
    const void *fn =
        factor ? (const void *)potrf_diag_kernel<T, DiagNT<T>::v,
                                                 DiagNT<T>::min_waves>
               : (const void *)potrf_diag_kernel<T, NTHREADS>;
    if (lds > 65536) {
        static int attr_set[2][2] = {{0, 0}, {0, 0}};
        int &done = attr_set[sizeof(T) == 8][factor != 0];
        if (!done) {
            hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
            done = 1;
        }
    }
    if (factor)
        hipLaunchKernelGGL((potrf_diag_kernel<T, DiagNT<T>::v,
                                              DiagNT<T>::min_waves>),
                           dim3(count),
                           dim3(DiagNT<T>::v), lds, s, A, uplo == PlasmaUpper, k0,
                           factor, (T *)inv, status);
    else
        hipLaunchKernelGGL((potrf_diag_kernel<T, NTHREADS>), dim3(count),
                           dim3(NTHREADS), lds, s, A, uplo == PlasmaUpper, k0,
                           factor, (T *)inv, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

// ---------------------------------------------------------------------------
// Blocked panel solve: P(m) = A(m,k) L_kk^{-T} from the 32 x 32 diagonal-block
// inverses D_J = L_JJ^{-1} alone (factor_tile mode 2), so the diagonal
// kernel need not form the whole W_kk on the critical path.  One workgroup
// per 32-row block of a panel tile; with X the solved rows,
//     T   = A_J - X(:, 0:j0) L_kk(J, 0:j0)^T        (block column J)
//     X_J = T D_J^T
// for J = 0, 1, ...  The row block of A (becoming X) and the block row of
// L_kk are staged in LDS (row major, k contiguous); each of the 4 waves
// computes one 16 x 16 quarter of T and of X_J on MFMA.
// ---------------------------------------------------------------------------
constexpr int PB_RB = 32;           // rows per workgroup

__host__ __device__ constexpr int pb_ld(int nb) { return nb + 4; }

template <typename T>
size_t pb_lds_bytes(int nb)
{
    return (2 * (size_t)PB_RB * pb_ld(nb) + 2 * JB * (JB + 1)) * sizeof(T);
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void panel_blocked_kernel(
    plasma_desc_t A, int upper, int k, const T *inv, T *panel, int pbase,
    int pcols, int pcol, int nrb, const int *status)
{
    if (block_should_stop(status)) return;
    const int m = k + 1 + blockIdx.x / nrb;
    const int r0 = (blockIdx.x % nrb) * PB_RB;
    const int mv = plasma_tile_mview(A, m), kv = plasma_tile_mview(A, k);
    if (r0 >= mv) return;
    const int rows = min(PB_RB, mv - r0);
    const int LDX = pb_ld(A.nb);
    extern __shared__ unsigned char smem_raw[];
    T *sX = reinterpret_cast<T *>(smem_raw);           // PB_RB x LDX
    T *sL = sX + PB_RB * LDX;                           // JB x LDX
    T *sT = sL + JB * LDX;                              // JB x (JB + 1)
    T *sD = sT + JB * (JB + 1);                         // JB x (JB + 1)
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const View<T> Amk = lower_view<T>(A, upper != 0, m, k);
    const View<T> Lkk = lower_view<T>(A, upper != 0, k, k);
    const T *W = inv + (size_t)k * A.mb * A.mb;         // column major, ld mb
    T *P = panel + (size_t)(m - pbase - 1) * A.mb * pcols +
           (size_t)pcol * A.mb;                          // column major, ld mb

    // the row block of A(m,k); rows past the tile read as zero
    for (int e = t; e < PB_RB * kv; e += NTHREADS) {
        const int r = e % PB_RB, c = e / PB_RB;
        sX[r * LDX + c] = (r < rows) ? Amk.at(r0 + r, c) : T(0);
    }
    const int wr = wave >> 1, wc = wave & 1;             // 16 x 16 quarter
    typedef typename Mfma<T>::acc_t acc_t;
    for (int j0 = 0; j0 < kv; j0 += JB) {
        const int jb = min(JB, kv - j0);
        // block row J of L_kk left of the diagonal, and D_J (zero padded)
        for (int e = t; e < JB * j0; e += NTHREADS) {
            const int c = e % JB, kk = e / JB;
            sL[c * LDX + kk] = (c < jb) ? Lkk.at(j0 + c, kk) : T(0);
        }
        for (int e = t; e < JB * JB; e += NTHREADS) {
            const int c = e % JB, l = e / JB;
            sD[c * (JB + 1) + l] = (c < jb && l < jb)
                ? W[(size_t)(j0 + l) * A.mb + j0 + c] : T(0);
        }
        __syncthreads();
        // T = A_J - X(:, 0:j0) L(J, 0:j0)^T
        acc_t acc = {0, 0, 0, 0};
        for (int kk = 0; kk < j0; kk += 4) {
            const T a = sX[(wr * 16 + lane % 16) * LDX + kk + lane / 16];
            const T b = sL[(wc * 16 + lane % 16) * LDX + kk + lane / 16];
            acc = Mfma<T>::mma(a, b, acc);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = wr * 16 + Mfma<T>::row(lane, q);
            const int c = wc * 16 + lane % 16;
            const T aj = (c < jb) ? sX[r * LDX + j0 + c] : T(0);
            sT[r * (JB + 1) + c] = aj - acc[q];
        }
        __syncthreads();
        // X_J = T D_J^T
        acc_t x = {0, 0, 0, 0};
#pragma unroll
        for (int l = 0; l < JB; l += 4) {
            const T a = sT[(wr * 16 + lane % 16) * (JB + 1) + l + lane / 16];
            const T b = sD[(wc * 16 + lane % 16) * (JB + 1) + l + lane / 16];
            x = Mfma<T>::mma(a, b, x);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = wr * 16 + Mfma<T>::row(lane, q);
            const int c = wc * 16 + lane % 16;
            if (c < jb) {
                sX[r * LDX + j0 + c] = x[q];
                if (r < rows) P[(size_t)(j0 + c) * A.mb + r0 + r] = x[q];
            }
        }
        __syncthreads();
    }
}

template <typename T>
int launch_panel_blocked(plasma_enum_t uplo, plasma_desc_t A, int k,
                         const void *inv, void *panel, int pbase, int pcols,
                         int pcol, const int *status, hipStream_t s)
{
    const int ntiles = A.mt - k - 1;
    if (ntiles <= 0) return 0;
    const size_t lds = pb_lds_bytes<T>(A.nb);
    if (lds > 160 * 1024) return PlasmaErrorNotSupported;
    static int attr_done = 0;
    if (!attr_done && lds > 65536) {
        hipFuncSetAttribute((const void *)panel_blocked_kernel<T>,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
        attr_done = 1;
    }
    const int nrb = (A.mb + PB_RB - 1) / PB_RB;
    hipLaunchKernelGGL(panel_blocked_kernel<T>, dim3(ntiles * nrb),
                       dim3(NTHREADS), lds, s, A, uplo == PlasmaUpper, k,
                       (const T *)inv, (T *)panel, pbase, pcols, pcol, nrb,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T>
int launch_trsm_panel(plasma_enum_t uplo, plasma_desc_t A, int k,
                      const void *inv, void *panel, int pbase, int pcols,
                      int pcol, const int *status, hipStream_t s)
{
    const int ntiles = A.mt - k - 1;
    if (ntiles <= 0) return 0;
    const int bpt = (A.mb + tg::BM - 1) / tg::BM;
    PanelDecode<T> d{A, k, bpt, uplo == PlasmaUpper, pbase, pcols, pcol,
                     (const T *)inv, (T *)panel};
    if (uplo == PlasmaUpper)
        hipLaunchKernelGGL((tg::kernel<T, true, false, false, false, true,
                                       PanelDecode<T>>),
                           dim3(tg::xcd_grid(ntiles * bpt * bpt)),
                           dim3(tg::NT), 0, s, d, ntiles * bpt * bpt, status);
    else
        hipLaunchKernelGGL((tg::kernel<T, true, false, false, true, true,
                                       PanelDecode<T>>),
                           dim3(tg::xcd_grid(ntiles * bpt * bpt)),
                           dim3(tg::NT), 0, s, d, ntiles * bpt * bpt, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T>
int launch_trailing_update(plasma_enum_t uplo, plasma_desc_t A, int pbase,
                           int kdepth, int n_first, int n_end,
                           const void *panel, int pcols, int copy_cols,
                           int max_wgs, const int *status, hipStream_t s)
{
    n_end = min(n_end, A.mt);
    long nblocks = 0;
    const int bpt = (A.mb + tg::BM - 1) / tg::BM;
    if (n_first < n_end) {
        const long T0 = A.mt - n_first;
        const long ncols = n_end - n_first;
        const long ntiles = ncols * T0 - ncols * (ncols - 1) / 2;
        nblocks = ntiles * bpt * bpt;
    }
    int ncopy = 0;
    for (int j = 0; j < copy_cols; j++)
        ncopy += max(A.mt - (pbase + j) - 1, 0);
    nblocks += ncopy;
    if (nblocks == 0) return 0;
    if (nblocks > 0x7fffffffL) return PlasmaErrorNotSupported;
    const long grid = max_wgs > 0 ? min((long)tg::xcd_grid((int)nblocks),
                                        (long)max_wgs)
                                  : (long)tg::xcd_grid((int)nblocks);
    UpdateDecode<T> d{A, pbase, pcols, kdepth, max(copy_cols, 1), n_first,
                      bpt, uplo == PlasmaUpper, ncopy, (const T *)panel};
    if (uplo == PlasmaUpper)
        hipLaunchKernelGGL((tg::kernel<T, false, true, true, true, true,
                                       UpdateDecode<T>>),
                           dim3((unsigned)grid), dim3(tg::NT), 0, s, d,
                           (int)nblocks, status);
    else
        hipLaunchKernelGGL((tg::kernel<T, true, true, true, true, true,
                                       UpdateDecode<T>>),
                           dim3((unsigned)grid), dim3(tg::NT), 0, s, d,
                           (int)nblocks, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

// ---------------------------------------------------------------------------
// Batched SPD solves of uneven order n_i <= 512 (no reference counterpart:
// the reference factors one matrix at a time).  Each system is one
// workgroup's work: factor_tile in mode 2 (factor, inverses of the 32 x 32
// diagonal blocks), then the two triangular solves by 32-row blocks,
// y_J = W_JJ (b_J - L_J,<J y_<J) and x_J = W_JJ^T (y_J - L_>J,J^T x_>J),
// with the solution vector in LDS.  A persistent grid takes systems from a
// ticket counter in the order `order` (largest first; NULL: index order);
// every workgroup exits when the tickets run out.
//
// Workspace: only the 32 x 32 diagonal blocks of W are formed (mode 2), so
// each workgroup keeps them in its own slab of W, reused system after
// system: the view (r, c) -> r + 32 c puts block J, (r, c) in
// [32 J, 32 J + 32)^2, at 1056 J + (r - 32 J) + 32 (c - 32 J), inside
// [1056 J, 1056 J + 1024), so the blocks of an order-512 system fit in
// VB_WSLAB = 33 * 512 elements.  A system whose order is outside [0, 512]
// is skipped with info -3 (the argument position of n).
// ---------------------------------------------------------------------------
constexpr int VB_NMAX = 512;
constexpr int VB_WSLAB = 33 * VB_NMAX;
template <typename T>
__device__ void solve_tile(const View<T> &L, const int n, const View<T> &Wv,
                           T *Bp, const int ldb, const int nrhs, T *sw)
{
    const int t = threadIdx.x;
    const int nbb = (n + JB - 1) / JB;
    T *sy = sw;                    // the vector being solved, n <= 512
    T *spart = sw + 512;           // 8 x 32 partial sums
    T *st = spart + 8 * JB;        // 32
    for (int c = 0; c < nrhs; c++) {
        T *b = Bp + (size_t)c * ldb;
        for (int i = t; i < n; i += NTHREADS) sy[i] = b[i];
        __syncthreads();
        for (int J = 0; J < nbb; J++) {            // L y = b
            const int j0 = J * JB, jb = min(JB, n - j0);
            const int r = t % JB, p = t / JB;
            T acc = T(0);
            if (r < jb)
                for (int k = p; k < j0; k += NTHREADS / JB)
                    acc += L.at(j0 + r, k) * sy[k];
            spart[p * JB + r] = acc;
            __syncthreads();
            if (t < jb) {
                T v = sy[j0 + t];
                for (int q = 0; q < NTHREADS / JB; q++) v -= spart[q * JB + t];
                st[t] = v;
            }
            __syncthreads();
            if (t < jb) {
                T y = T(0);
                for (int cc = 0; cc <= t; cc++)
                    y += Wv.at(j0 + t, j0 + cc) * st[cc];
                sy[j0 + t] = y;
            }
            __syncthreads();
        }
        for (int J = nbb - 1; J >= 0; J--) {       // L^T x = y
            const int j0 = J * JB, jb = min(JB, n - j0);
            const int r = t / (NTHREADS / JB), p = t % (NTHREADS / JB);
            T acc = T(0);
            if (r < jb)
                for (int k = j0 + jb + p; k < n; k += NTHREADS / JB)
                    acc += L.at(k, j0 + r) * sy[k];
            spart[p * JB + r] = acc;
            __syncthreads();
            if (t < jb) {
                T v = sy[j0 + t];
                for (int q = 0; q < NTHREADS / JB; q++) v -= spart[q * JB + t];
                st[t] = v;
            }
            __syncthreads();
            if (t < jb) {
                T x = T(0);
                for (int cc = t; cc < jb; cc++)
                    x += Wv.at(j0 + cc, j0 + t) * st[cc];
                sy[j0 + t] = x;
            }
            __syncthreads();
        }
        for (int i = t; i < n; i += NTHREADS) b[i] = sy[i];
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void vbatched_kernel(
    int upper, int batch, const int *nvec, T *A, const size_t *aoff, T *W,
    int nrhs, T *B, const size_t *boff, int *info, const int *order,
    int *ticket)
{
    extern __shared__ unsigned char smem_raw[];
    T *sP = reinterpret_cast<T *>(smem_raw);
    __shared__ int s_i;
    for (;;) {
        if (threadIdx.x == 0) s_i = atomicAdd(ticket, 1);
        __syncthreads();
        const int i = s_i;
        __syncthreads();
        if (i >= batch) break;
        const int sys = order != nullptr ? order[i] : i;
        const int n = nvec[sys];
        int inf = (n < 0 || n > VB_NMAX) ? -3 : 0;
        if (n > 0 && inf == 0) {
            View<T> L;
            L.p = A + aoff[sys];
            L.rs = upper ? n : 1;
            L.cs = upper ? 1 : n;
            View<T> Wv;
            Wv.p = W + (size_t)blockIdx.x * VB_WSLAB;
            Wv.rs = 1;
            Wv.cs = JB;
            inf = factor_tile<T>(L, n, Wv, 2, sP);
            if (inf == 0 && nrhs > 0)
                solve_tile<T>(L, n, Wv, B + boff[sys], n, nrhs, sP);
        }
        if (threadIdx.x == 0) info[sys] = inf;
        __syncthreads();
    }
}

template <typename T>
int launch_vbatched(int upper, int batch, const int *n, void *A,
                    const size_t *aoff, void *W, int nrhs, void *B,
                    const size_t *boff, int *info, const int *order,
                    int *ticket, int grid, hipStream_t s)
{
    // factor_tile's panel area for n <= 512, and the solve's vectors
    const size_t lds = (size_t)8 * JB * 2 * (JB + 1) * sizeof(T);
    static int attr_set[2] = {0, 0};
    int &done = attr_set[sizeof(T) == 8];
    if (!done) {
        hipFuncSetAttribute((const void *)vbatched_kernel<T>,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
        done = 1;
    }
    hipLaunchKernelGGL(vbatched_kernel<T>, dim3(grid), dim3(NTHREADS), lds,
                       s, upper, batch, n, (T *)A, aoff, (T *)W, nrhs,
                       (T *)B, boff, info, order, ticket);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}


// ---------------------------------------------------------------------------
// Inverses of the lower triangular tiles (0, k) of a one-tile-row descriptor
// D (the diagonal tiles of a triangular matrix, as plasma_kern_trsm_diag_prep
// leaves them), one workgroup per tile: W_k = D(0,k)^{-1}, lower, at
// inv + k mb^2 with leading dimension mb.  factor_tile's invert-only form.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NTHREADS) void trtri_tiles_kernel(
    plasma_desc_t D, T *inv, const int *status)
{
    if (block_should_stop(status)) return;
    const int k = blockIdx.x;
    View<T> L;
    L.p = (T *)plasma_tile_addr(D, 0, k);
    L.rs = 1;
    L.cs = plasma_tile_mmain(D, 0);
    View<T> Wv;
    Wv.p = inv + (size_t)k * D.mb * D.mb;
    Wv.rs = 1;
    Wv.cs = D.mb;
    extern __shared__ unsigned char smem_raw[];
    factor_tile<T, NTHREADS>(L, plasma_tile_nview(D, k), Wv, 0,
                             reinterpret_cast<T *>(smem_raw));
}

template <typename T>
int launch_trtri_tiles(plasma_desc_t D, void *inv, const int *status,
                       hipStream_t s)
{
    if (D.nt <= 0) return 0;
    if (D.mb != D.nb || D.mb > 512) return PlasmaErrorNotSupported;
    const size_t lds = (size_t)max(D.mb, 8 * JB) * (JB + 1) * sizeof(T);
    if (lds > 65536) {
        static int attr_set[2] = {0, 0};
        int &done = attr_set[sizeof(T) == 8];
        if (!done) {
            hipFuncSetAttribute((const void *)trtri_tiles_kernel<T>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
            done = 1;
        }
    }
    hipLaunchKernelGGL(trtri_tiles_kernel<T>, dim3(D.nt), dim3(NTHREADS), lds,
                       s, D, (T *)inv, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

}  // namespace

extern "C" int plasma_kern_potrf_diag(plasma_enum_t uplo, plasma_desc_t A,
                                      int k, void *inv, int *status,
                                      hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_potrf_diag<double>(uplo, A, k, 1, 1, inv, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_potrf_diag<float>(uplo, A, k, 1, 1, inv, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_potrf_diag_blocks(plasma_enum_t uplo,
                                             plasma_desc_t A, int k,
                                             void *inv, int *status,
                                             hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_potrf_diag<double>(uplo, A, k, 1, 2, inv, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_potrf_diag<float>(uplo, A, k, 1, 2, inv, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_panel_blocked_ok(plasma_desc_t A)
{
    if (A.mb != A.nb) return 0;
    if (A.precision == PlasmaRealDouble)
        return pb_lds_bytes<double>(A.nb) <= 160 * 1024;
    if (A.precision == PlasmaRealFloat)
        return pb_lds_bytes<float>(A.nb) <= 160 * 1024;
    return 0;
}

extern "C" int plasma_kern_trsm_panel_blocked(plasma_enum_t uplo,
                                              plasma_desc_t A, int k,
                                              const void *inv, void *panel,
                                              int pbase, int pcols, int pcol,
                                              const int *status,
                                              hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_panel_blocked<double>(uplo, A, k, inv, panel, pbase,
                                            pcols, pcol, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_panel_blocked<float>(uplo, A, k, inv, panel, pbase,
                                           pcols, pcol, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_trtri_diag(plasma_enum_t uplo, plasma_desc_t A,
                                      void *inv, const int *status,
                                      hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_potrf_diag<double>(uplo, A, 0, A.mt, 0, inv,
                                         (int *)status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_potrf_diag<float>(uplo, A, 0, A.mt, 0, inv,
                                        (int *)status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_trsm_panel(plasma_enum_t uplo, plasma_desc_t A,
                                      int k, const void *inv, void *panel,
                                      int pbase, int pcols, int pcol,
                                      const int *status, hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_trsm_panel<double>(uplo, A, k, inv, panel, pbase,
                                         pcols, pcol, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_trsm_panel<float>(uplo, A, k, inv, panel, pbase, pcols,
                                        pcol, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_trailing_update(plasma_enum_t uplo,
                                           plasma_desc_t A, int pbase,
                                           int kdepth, int n_first,
                                           int n_end, const void *panel,
                                           int pcols, int copy_cols,
                                           int max_wgs, const int *status,
                                           hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_trailing_update<double>(uplo, A, pbase, kdepth,
                                              n_first, n_end, panel, pcols,
                                              copy_cols, max_wgs, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_trailing_update<float>(uplo, A, pbase, kdepth, n_first,
                                             n_end, panel, pcols, copy_cols,
                                             max_wgs, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" size_t plasma_kern_vbatched_lds(plasma_enum_t precision)
{
    const size_t es = precision == PlasmaRealDouble ? 8 : 4;
    return (size_t)8 * JB * 2 * (JB + 1) * es +
           (2 * JB * (JB + 1) + 2 * JB) * es + 32;
}

extern "C" size_t plasma_kern_vbatched_wslab(plasma_enum_t precision)
{
    return (size_t)VB_WSLAB * (precision == PlasmaRealDouble ? 8 : 4);
}

extern "C" int plasma_kern_vbatched_posv(plasma_enum_t precision, int upper,
                                         int batch, const int *n, void *A,
                                         const size_t *aoff, void *W,
                                         int nrhs, void *B,
                                         const size_t *boff, int *info,
                                         const int *order, int *ticket,
                                         int grid, hipStream_t s)
{
    if (precision == PlasmaRealDouble)
        return launch_vbatched<double>(upper, batch, n, A, aoff, W, nrhs, B,
                                       boff, info, order, ticket, grid, s);
    if (precision == PlasmaRealFloat)
        return launch_vbatched<float>(upper, batch, n, A, aoff, W, nrhs, B,
                                      boff, info, order, ticket, grid, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_trtri_tiles(plasma_desc_t D, void *inv,
                                       const int *status, hipStream_t s)
{
    if (D.precision == PlasmaRealDouble)
        return launch_trtri_tiles<double>(D, inv, status, s);
    if (D.precision == PlasmaRealFloat)
        return launch_trtri_tiles<float>(D, inv, status, s);
    return PlasmaErrorNotSupported;
}
