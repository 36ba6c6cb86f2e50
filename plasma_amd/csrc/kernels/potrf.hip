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
constexpr int JB = 32;          // column block of the diagonal tile

// ---------------------------------------------------------------------------
// Wave-level 32 x 32 kernels.  Lane r (r < 32) holds row r of a block in
// d[0..31]; rows and columns beyond jb are padded with the identity.
// Values of other lanes are fetched with v_readlane (no LDS, no barriers).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T readlane(T v, int l);

template <>
__device__ __forceinline__ double readlane<double>(double v, int l)
{
    int2 u = __builtin_bit_cast(int2, v);
    u.x = __builtin_amdgcn_readlane(u.x, l);
    u.y = __builtin_amdgcn_readlane(u.y, l);
    return __builtin_bit_cast(double, u);
}

template <>
__device__ __forceinline__ float readlane<float>(float v, int l)
{
    return __builtin_bit_cast(float,
        __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// Right-looking Cholesky of the block held in d; returns 0 or the 1-based
// column of the first non-positive (or NaN) pivot (LAPACK dpotf2 rule).
template <typename T>
__device__ __forceinline__ int wave_potrf32(T (&d)[JB], int jb)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < JB; j++) {
        const T djj = readlane(d[j], j);
        if (j < jb && !(djj > T(0))) return j + 1;
        const T ljj = sqrt(djj);
        if (lane == j) d[j] = ljj;
        else if (lane > j) d[j] = d[j] / ljj;
#pragma unroll
        for (int c = j + 1; c < JB; c++) {
            const T lcj = readlane(d[j], c);
            if (lane >= c) d[c] -= d[j] * lcj;
        }
        // keep this column's scalar reads next to their use
        __builtin_amdgcn_sched_barrier(0);
    }
    return 0;
}

// Inverse of the lower-triangular block held in d (rows on lanes):
// lane c returns column c of L^{-1} in inv[0..31].
template <typename T>
__device__ __forceinline__ void wave_trtri32(const T (&d)[JB], T (&inv)[JB])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < JB; r++) {
        T s = (lane == r) ? T(1) : T(0);
#pragma unroll
        for (int l = 0; l < r; l++) s -= readlane(d[l], r) * inv[l];
        inv[r] = s / readlane(d[r], r);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// One 16 x 16 MFMA block: acc += A(16 x K) B(16 x K)^T with the operands
// given as Views positioned at their first row; rows beyond the bounds read
// as zero.  Result element q of the lane is at (Mfma<T>::row(lane, q),
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
            a[u] = (r < arows && kk < K) ? Av.at(r, kk) : T(0);
            b[u] = (r < brows && kk < K) ? Bv.at(r, kk) : T(0);
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
                a[h][u] = (ra < arows && kk < K) ? Av.at(ra, kk) : T(0);
                b[h][u] = (ra < brows && kk < K) ? Bv.at(ra, kk) : T(0);
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

// ---------------------------------------------------------------------------
// Diagonal tile k: with `factor`, L L^T = A(k,k) by 32-column blocks (a wave
// factors the 32 x 32 block with readlane, the rows below are multiplied by
// the block's inverse on MFMA, and the rest of the tile is updated on MFMA);
// then, always, W = L^{-1} (lower, zeros above) into the workspace tile
// inv + k * mb * mb (ld = mb), by block columns: W_JJ = L_JJ^{-1},
// W_IJ = -W_II sum_{K=J}^{I-1} L_IK W_KJ.  The invert-only form runs one
// workgroup per tile of a batch.
// A non-positive or NaN pivot at local column j sets the status word to
// k*nb + j + 1 (reference core_zpotrf.c:72-79, iinfo = A.nb * k).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NTHREADS) void potrf_diag_kernel(
    plasma_desc_t A, int upper, int k0, int factor, T *inv, int *status)
{
    if (block_should_stop(status)) return;
    const int k = k0 + blockIdx.x;
    const int n = plasma_tile_mview(A, k);
    const int nbb = (n + JB - 1) / JB;
    const View<T> L = lower_view<T>(A, upper != 0, k, k);
    View<T> Wv;
    Wv.p = inv + (size_t)k * A.mb * A.mb;
    Wv.rs = 1;
    Wv.cs = A.mb;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    typedef typename Mfma<T>::acc_t acc_t;

    constexpr int LDP = JB + 1;          // padded row stride in LDS
    __shared__ T sDinv[JB][LDP];        // inverse of the current block
    __shared__ int s_info;
    extern __shared__ unsigned char smem_raw[];
    T *sP = reinterpret_cast<T *>(smem_raw);   // panel rows, n x LDP
    if (t == 0) s_info = 0;

    View<T> Dv;                          // Dv(c, l) = Dinv(c, l)
    Dv.p = &sDinv[0][0];
    Dv.rs = LDP;
    Dv.cs = 1;
    View<T> Pv;                          // Pv(r, l) = panel row r, column l
    Pv.p = sP;
    Pv.rs = LDP;
    Pv.cs = 1;

    if (factor) {
        for (int J = 0; J < nbb; J++) {
            const int j0 = J * JB, jb = min(JB, n - j0);
            const int i0 = j0 + jb, nrows = n - i0;
            // (a) one wave: factor and invert the diagonal block
            if (wave == 0) {
                T d[JB];
#pragma unroll
                for (int c = 0; c < JB; c++) {
                    T v;
                    if (lane < jb && c < jb) v = (c <= lane) ? L.at(j0 + lane, j0 + c) : T(0);
                    else                     v = (lane == c) ? T(1) : T(0);
                    d[c] = v;
                }
                const int info = wave_potrf32(d, jb);
                if (info) {
                    if (lane == 0) s_info = j0 + info;
                }
                else {
                    T iv[JB];
                    wave_trtri32(d, iv);
                    if (lane < JB) {
#pragma unroll
                        for (int c = 0; c < JB; c++) {
                            if (lane < jb && c <= lane && c < jb)
                                L.at(j0 + lane, j0 + c) = d[c];
                            sDinv[c][lane] = iv[c];   // row c of the inverse
                            if (lane < jb && c < jb)
                                Wv.at(j0 + c, j0 + lane) = iv[c];
                        }
                    }
                }
            }
            else {
                // the other waves stage the panel rows below the block
                for (int e = t - 64; e < nrows * JB; e += NTHREADS - 64) {
                    const int r = e % nrows, c = e / nrows;
                    sP[r * LDP + c] = (c < jb) ? L.at(i0 + r, j0 + c) : T(0);
                }
            }
            __syncthreads();
            if (s_info) {
                if (t == 0) atomicCAS(status, 0, A.nb * k + s_info);
                return;
            }
            // (b) P = A_panel Dinv^T, one 16-row strip per wave (both
            //     16-column halves); P goes back to LDS and to the tile
            for (int st = wave; st * 16 < nrows; st += 4) {
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
            const int nb16 = (nrows + 15) / 16;
            const int nblk = nb16 * (nb16 + 1) / 2;
            for (int b = wave; b < nblk; b += 4) {
                int bi = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
                while (bi * (bi + 1) / 2 > b) bi--;
                while ((bi + 1) * (bi + 2) / 2 <= b) bi++;
                const int bj = b - bi * (bi + 1) / 2;
                T cv[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int r = bi * 16 + Mfma<T>::row(lane, q);
                    const int c = bj * 16 + lane % 16;
                    cv[q] = (r < nrows && c <= r) ? L.at(i0 + r, i0 + c) : T(0);
                }
                acc_t acc = {0, 0, 0, 0};
                wave_mma_nt<T>(acc, sub(Pv, bi * 16, 0), nrows - bi * 16,
                               sub(Pv, bj * 16, 0), nrows - bj * 16, JB);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int r = bi * 16 + Mfma<T>::row(lane, q);
                    const int c = bj * 16 + lane % 16;
                    if (r < nrows && c <= r) L.at(i0 + r, i0 + c) = cv[q] - acc[q];
                }
            }
            __syncthreads();
        }
    }
    else {
        // invert-only: the diagonal blocks' inverses first
        for (int J = wave; J < nbb; J += 4) {
            const int j0 = J * JB, jb = min(JB, n - j0);
            T d[JB], iv[JB];
#pragma unroll
            for (int c = 0; c < JB; c++) {
                T v;
                if (lane < jb && c < jb) v = (c <= lane) ? L.at(j0 + lane, j0 + c) : T(0);
                else                     v = (lane == c) ? T(1) : T(0);
                d[c] = v;
            }
            wave_trtri32(d, iv);
            if (lane < jb) {
#pragma unroll
                for (int c = 0; c < JB; c++)
                    if (c < jb) Wv.at(j0 + c, j0 + lane) = iv[c];
            }
        }
    }

    // zero the strict upper triangle of W
    for (int c = wave; c < n; c += 4)
        for (int r = lane; r < c; r += 64) Wv.at(r, c) = T(0);
    __syncthreads();
    T (*sT)[JB][LDP] = reinterpret_cast<T (*)[JB][LDP]>(sP);  // per wave

    // off-diagonal blocks of W; block columns are paired (J, nbb-1-J) per
    // wave so the longest and shortest dependency chains share a wave
    for (int slot = wave; slot < nbb; slot += 4) {
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
}

// ---------------------------------------------------------------------------
// Job decoders for the block GEMM (tilegemm.hip.h).
// ---------------------------------------------------------------------------

// Trailing update of step k, tile columns [n_first, ...): tiles (m, n)
// enumerated column by column, bpt x bpt blocks of 128 per tile.
template <typename T>
struct UpdateDecode {
    plasma_desc_t A;
    int k, n_first, bpt, upper;
    __device__ bool operator()(int b, tg::Job<T> &j) const
    {
        const int per_tile = bpt * bpt;
        const int sub_ = b % per_tile;
        const int bi = sub_ / bpt, bj = sub_ % bpt;
        int m, n;
        trapezoid_decode(b / per_tile, A.mt, n_first, &m, &n);
        if (m == n && bj > bi) return false;
        const int mv = plasma_tile_mview(A, m), nv = plasma_tile_mview(A, n);
        const int r0 = bi * tg::BM, c0 = bj * tg::BM;
        if (r0 >= mv || c0 >= nv) return false;
        j.X = sub(lower_view<T>(A, upper, m, k), r0, 0);
        j.Y = sub(lower_view<T>(A, upper, n, k), c0, 0);
        j.C = sub(lower_view<T>(A, upper, m, n), r0, c0);
        j.xrows = mv - r0;
        j.yrows = nv - c0;
        j.K = A.nb;
        j.lower = (m == n);
        j.diag_off = c0 - r0;
        return true;
    }
};

// Panel solve of step k: S(m) = A(m,k) W_kk^T for m > k into the scratch
// panel (tile m at scratch + (m-k-1) * mb * nb, column major, ld = mb).
template <typename T>
struct PanelDecode {
    plasma_desc_t A;
    int k, bpt, upper;
    const T *inv;
    T *scratch;
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
        S.p = scratch + (size_t)(m - k - 1) * A.mb * A.nb;
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

// Copy the solved panel back: A(m,k) = S(m), m > k.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void panel_copy_kernel(
    plasma_desc_t A, int upper, int k, const T *scratch, const int *status)
{
    if (block_should_stop(status)) return;
    const int m = k + 1 + blockIdx.x;
    const int mv = plasma_tile_mview(A, m), nb = plasma_tile_mview(A, k);
    const View<T> D = lower_view<T>(A, upper != 0, m, k);
    const T *s = scratch + (size_t)(m - k - 1) * A.mb * A.nb;
    if (!upper) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int c = wave; c < nb; c += NTHREADS / 64)
            for (int r = lane; r < mv; r += 64)
                D.at(r, c) = s[r + (size_t)c * A.mb];
    }
    else {
        // D.at(r, c) = p[r * ld + c]: walk c fastest
        for (int e = threadIdx.x; e < mv * nb; e += NTHREADS) {
            const int r = e / nb, c = e % nb;
            D.at(r, c) = s[r + (size_t)c * A.mb];
        }
    }
}

template <typename T>
int launch_potrf_diag(plasma_enum_t uplo, plasma_desc_t A, int k0, int count,
                      int factor, void *inv, int *status, hipStream_t s)
{
    if (count <= 0) return 0;
    // panel rows (mb x 33) in the factor phase; 4 x 32 x 33 in the inverse
    const size_t lds = (size_t)max(A.mb, 4 * JB) * (JB + 1) * sizeof(T);
    if (lds > 65536) {
        static int attr_set[2] = {0, 0};
        int &done = attr_set[sizeof(T) == 8];
        if (!done) {
            hipFuncSetAttribute((const void *)potrf_diag_kernel<T>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
            done = 1;
        }
        if (lds > 150 * 1024) return PlasmaErrorNotSupported;
    }
    hipLaunchKernelGGL(potrf_diag_kernel<T>, dim3(count), dim3(NTHREADS), lds,
                       s, A, uplo == PlasmaUpper, k0, factor, (T *)inv,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T>
int launch_trsm_panel(plasma_enum_t uplo, plasma_desc_t A, int k,
                      const void *inv, void *scratch, const int *status,
                      hipStream_t s)
{
    const int ntiles = A.mt - k - 1;
    if (ntiles <= 0) return 0;
    const int bpt = (A.mb + tg::BM - 1) / tg::BM;
    PanelDecode<T> d{A, k, bpt, uplo == PlasmaUpper, (const T *)inv,
                     (T *)scratch};
    if (uplo == PlasmaUpper)
        hipLaunchKernelGGL((tg::kernel<T, true, false, false, false, true,
                                       PanelDecode<T>>),
                           dim3(ntiles * bpt * bpt), dim3(tg::NT), 0, s, d,
                           status);
    else
        hipLaunchKernelGGL((tg::kernel<T, true, false, false, true, true,
                                       PanelDecode<T>>),
                           dim3(ntiles * bpt * bpt), dim3(tg::NT), 0, s, d,
                           status);
    hipLaunchKernelGGL(panel_copy_kernel<T>, dim3(ntiles), dim3(NTHREADS), 0,
                       s, A, uplo == PlasmaUpper, k, (const T *)scratch,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T>
int launch_trailing_update(plasma_enum_t uplo, plasma_desc_t A, int k,
                           int n_first, int n_end, const int *status,
                           hipStream_t s)
{
    if (n_first >= n_end || n_first >= A.mt) return 0;
    n_end = min(n_end, A.mt);
    const long T0 = A.mt - n_first;
    const long ncols = n_end - n_first;
    const long ntiles = ncols * T0 - ncols * (ncols - 1) / 2;
    const int bpt = (A.mb + tg::BM - 1) / tg::BM;
    const long nblocks = ntiles * bpt * bpt;
    if (nblocks > 0x7fffffffL) return PlasmaErrorNotSupported;
    UpdateDecode<T> d{A, k, n_first, bpt, uplo == PlasmaUpper};
    if (uplo == PlasmaUpper)
        hipLaunchKernelGGL((tg::kernel<T, false, true, true, false, false,
                                       UpdateDecode<T>>),
                           dim3((unsigned)nblocks), dim3(tg::NT), 0, s, d,
                           status);
    else
        hipLaunchKernelGGL((tg::kernel<T, true, true, true, true, true,
                                       UpdateDecode<T>>),
                           dim3((unsigned)nblocks), dim3(tg::NT), 0, s, d,
                           status);
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
                                      int k, const void *inv, void *scratch,
                                      const int *status, hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_trsm_panel<double>(uplo, A, k, inv, scratch, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_trsm_panel<float>(uplo, A, k, inv, scratch, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_trailing_update(plasma_enum_t uplo,
                                           plasma_desc_t A, int k,
                                           int n_first, int n_end,
                                           const int *status, hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_trailing_update<double>(uplo, A, k, n_first, n_end,
                                              status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_trailing_update<float>(uplo, A, k, n_first, n_end,
                                             status, s);
    return PlasmaErrorNotSupported;
}
