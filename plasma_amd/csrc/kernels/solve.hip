// solve.hip -- triangular sweeps of potrs / posv for gfx950.
//
// The reference solves L Y = B and L^T X = Y with pztrsm
// (compute/pztrsm.c:113-144 Left-Lower-NoTrans, :148-177 Left-Lower-Trans,
// :43-107 the upper variants): per block row k a tile trsm with the
// diagonal tile, then a gemm update of the block rows still to be solved.
// Here the diagonal solve multiplies by the inverse of the diagonal tile
// (computed with the factorization, or by plasma_kern_trtri_diag) on MFMA
// into a scratch block row S = W_kk B(k,:) (or W_kk^T B(k,:)), and one
// launch then updates every block row still to be solved from S and copies
// S back into B(k,:).  Rows of a block row are independent, so both
// launches split the work into blocks of 64 rows x 32 right-hand sides.
#include "common.hip.h"
#include "tilegemm.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NTHREADS = 256;

template <typename T>
__device__ __forceinline__ View<T> sub(View<T> v, int r, int c)
{
    v.p += (ptrdiff_t)r * v.rs + (ptrdiff_t)c * v.cs;
    return v;
}

// Both sweep kernels use GemmNT on blocks of 64 rows x 32 right-hand sides
// with 32-deep K chunks: the per-step kernels are latency-bound, so small
// blocks (more workgroups, less serial work each) measured fastest.  The
// product is formed transposed (Y X^T), so the accumulator's 16 consecutive
// lanes hold 16 consecutive rows of the column-major result.
constexpr int SW_BM = 64, SW_BN = 32, SW_KC = 32;

template <typename G, typename T>
__device__ __forceinline__ void store_rows(
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

// S(r0.., c0.. of tile column tn) = op(W_kk) B(k, tn), S column major with
// ld = mb and global column tn * nb + c.  op(W) = W (forward) is lower
// triangular, so rows r0.. need k < r0 + 64; W^T (backward) needs k >= r0.
// Tile k = k0 + blockIdx.z; its S slab starts at scratch + blockIdx.z *
// kstride.  With from_s the right-hand side is read from the slab instead
// of B(k, :), with to_b the product is stored into B(k, :) instead of the
// slab (the two are never both set).
template <typename T>
__global__ __launch_bounds__(NTHREADS) void sweep_diag_kernel(
    plasma_desc_t A, plasma_desc_t B, int backward, int k0, int cchunks,
    const T *inv, T *scratch, size_t kstride, int from_s, int to_b,
    const int *status)
{
    typedef GemmNT<T, SW_BN, SW_BM, SW_KC, 16, 32> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int k = k0 + blockIdx.z;
    const int r0 = blockIdx.x * SW_BM;
    const int tn = blockIdx.y / cchunks;
    const int c0 = (blockIdx.y % cchunks) * SW_BN;
    const int mv = plasma_tile_mview(B, k), nv = plasma_tile_nview(B, tn);
    if (r0 >= mv || c0 >= nv) return;
    const int klo = backward ? r0 : 0;
    const int khi = backward ? mv : min(r0 + SW_BM, mv);
    View<T> X;                                  // X(r, kk) = op(W)(r0 + r, klo + kk)
    const T *w = inv + (size_t)k * A.mb * A.mb;
    if (!backward) { X.rs = 1; X.cs = A.mb; }
    else           { X.rs = A.mb; X.cs = 1; }
    X.p = const_cast<T *>(w) + (ptrdiff_t)r0 * X.rs + (ptrdiff_t)klo * X.cs;
    View<T> S;
    S.p = scratch + blockIdx.z * kstride + (size_t)tn * B.nb * A.mb;
    S.rs = 1;
    S.cs = A.mb;
    const View<T> Bk = tile_view<T>(B, k, tn);
    const View<T> &Src = from_s ? S : Bk;
    View<T> Y;                                  // Y(c, kk) = Src(klo + kk, c0 + c)
    Y.p = Src.p + klo + (ptrdiff_t)c0 * Src.cs;
    Y.rs = Src.cs;
    Y.cs = 1;
    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    G::run(acc, Y, nv - c0, X, mv - r0, khi - klo, sm);
    store_rows<G, T>(acc, to_b ? Bk : S, r0, c0, mv, nv, false);
}

// ---------------------------------------------------------------------------
// Sweep update, reading the solved block row from S.
// Forward: B(m,:) -= L(m,k) S for m > k.  Backward: B(m,:) -= L(k,m)^T S
// for m < k.  The last row of workgroups copies S into B(k,:).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NTHREADS) void sweep_update_kernel(
    plasma_desc_t A, plasma_desc_t B, int upper, int backward, int k,
    int mlo, int ntiles, int rchunks, int cchunks, const T *scratch,
    const int *status)
{
    typedef GemmNT<T, SW_BN, SW_BM, SW_KC, 16, 32> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;

    const int tn = blockIdx.y / cchunks;
    const int c0 = (blockIdx.y % cchunks) * SW_BN;
    const int nv = plasma_tile_nview(B, tn);
    const int K = plasma_tile_mview(B, k);
    View<T> S;                                   // S(kk, c) of tile column tn
    S.p = const_cast<T *>(scratch) + (size_t)tn * B.nb * A.mb;
    S.rs = 1;
    S.cs = A.mb;

    if ((int)blockIdx.x >= ntiles * rchunks) {
        // copy S -> B(k, tn)
        const int r0 = ((int)blockIdx.x - ntiles * rchunks) * SW_BM;
        if (r0 >= K || c0 >= nv) return;
        const View<T> Bk = tile_view<T>(B, k, tn);
        for (int e = threadIdx.x; e < SW_BM * SW_BN; e += NTHREADS) {
            const int r = r0 + e % SW_BM, c = c0 + e / SW_BM;
            if (r < K && c < nv) Bk.at(r, c) = S.at(r, c);
        }
        return;
    }

    const int m = backward ? mlo + (int)(blockIdx.x / rchunks)
                           : k + 1 + (int)(blockIdx.x / rchunks);
    const int r0 = (blockIdx.x % rchunks) * SW_BM;
    const int mv = plasma_tile_mview(B, m);
    if (r0 >= mv || c0 >= nv) return;

    View<T> X;
    if (!backward) {
        X = lower_view<T>(A, upper != 0, m, k);        // X(r, kk) = L(m,k)
    }
    else {
        View<T> v = lower_view<T>(A, upper != 0, k, m);  // L(k,m)
        X.p = v.p; X.rs = v.cs; X.cs = v.rs;             // transpose
    }
    X.p += (ptrdiff_t)r0 * X.rs;
    View<T> Y;                                   // Y(c, kk) = S(kk, c0 + c)
    Y.p = S.p + (ptrdiff_t)c0 * S.cs;
    Y.rs = S.cs;
    Y.cs = 1;
    const View<T> C = tile_view<T>(B, m, tn);

    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    G::run(acc, Y, nv - c0, X, mv - r0, K, sm);
    store_rows<G, T>(acc, C, r0, c0, mv, nv, true);
}

// ---------------------------------------------------------------------------
// Sweeps with many right-hand sides (B.n >= WIDE_RHS, e.g. the n columns of
// the Cholesky adjoint): the diagonal product S = op(W_kk) B(k, :) and the
// update B(m, :) -= op(L) S of one sweep step run as 128 x 128 MFMA block
// GEMMs (tilegemm.hip.h), one job per (tile row, 128-row block, 128-column
// block), instead of the 64 x 32 blocks sized for a few dozen columns.
// ---------------------------------------------------------------------------
constexpr int WIDE_RHS = 1024;

template <typename T>
struct SweepDecode {
    plasma_desc_t A, B;
    const T *inv;
    T *S;
    int k, backward, upper, diag, bpt, cbpt, mlo;
    __device__ bool operator()(int b, tg::Job<T> &j) const
    {
        const int ncb = B.nt * cbpt;
        const int rb = b % bpt, cb = (b / bpt) % ncb, t = b / bpt / ncb;
        const int tn = cb / cbpt, c0 = (cb % cbpt) * tg::BM;
        const int r0 = rb * tg::BM;
        const int nv = plasma_tile_nview(B, tn);
        const int kv = plasma_tile_mview(B, k);
        if (c0 >= nv) return false;
        View<T> Sv;                                  // S(kk, c), tile column tn
        Sv.p = S + (size_t)tn * B.nb * A.mb;
        Sv.rs = 1;
        Sv.cs = A.mb;
        if (diag) {
            // S(r0.., c0..) = op(W_kk)(r0.., :) B(k, tn)(:, c0..)
            if (r0 >= kv) return false;
            View<T> X;
            if (!backward) { X.rs = 1; X.cs = A.mb; }
            else           { X.rs = A.mb; X.cs = 1; }
            X.p = const_cast<T *>(inv) + (size_t)k * A.mb * A.mb +
                  (ptrdiff_t)r0 * X.rs;
            const View<T> Bk = tile_view<T>(B, k, tn);
            j.X = X;
            j.Y.p = Bk.p + (ptrdiff_t)c0 * Bk.cs;  // Y(c, kk) = B(k)(kk, c0 + c)
            j.Y.rs = Bk.cs;
            j.Y.cs = 1;
            j.C.p = Sv.p + r0 + (ptrdiff_t)c0 * Sv.cs;
            j.C.rs = 1;
            j.C.cs = Sv.cs;
            j.xrows = kv - r0;
        }
        else {
            // forward: B(m) -= L(m,k) S, m > k; backward: B(m) -= L(k,m)^T S
            const int m = backward ? mlo + t : k + 1 + t;
            const int mv = plasma_tile_mview(B, m);
            if (r0 >= mv) return false;
            View<T> X;
            if (!backward) {
                X = lower_view<T>(A, upper != 0, m, k);
            }
            else {
                const View<T> v = lower_view<T>(A, upper != 0, k, m);
                X.p = v.p;
                X.rs = v.cs;
                X.cs = v.rs;
            }
            X.p += (ptrdiff_t)r0 * X.rs;
            const View<T> Bm = tile_view<T>(B, m, tn);
            j.X = X;
            j.Y.p = Sv.p + (ptrdiff_t)c0 * Sv.cs;   // Y(c, kk) = S(kk, c0 + c)
            j.Y.rs = Sv.cs;
            j.Y.cs = 1;
            j.C.p = Bm.p + r0 + (ptrdiff_t)c0 * Bm.cs;
            j.C.rs = 1;
            j.C.cs = Bm.cs;
            j.xrows = mv - r0;
        }
        j.yrows = nv - c0;
        j.K = kv;
        j.lower = false;
        j.diag_off = 0;
        return true;
    }
};

// One launch of a wide sweep step; ACC/NEG and the X layout (XU: unit row
// stride) are compile-time in the block GEMM.
template <typename T, bool ACC, bool XU>
int launch_wide(const SweepDecode<T> &d, int njobs, const int *status,
                hipStream_t s)
{
    if (njobs <= 0) return 0;
    hipLaunchKernelGGL((tg::kernel<T, true, ACC, ACC, XU, false,
                                   SweepDecode<T>>),
                       dim3(tg::xcd_grid(njobs)), dim3(tg::NT), 0, s, d, njobs,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T>
int launch_sweep_diag(int backward, plasma_desc_t A, plasma_desc_t B, int k,
                      const void *inv, void *scratch, const int *status,
                      hipStream_t s)
{
    if (B.n >= WIDE_RHS) {
        const int bpt = (A.mb + tg::BM - 1) / tg::BM;
        const int cbpt = (B.nb + tg::BM - 1) / tg::BM;
        SweepDecode<T> d{A, B, (const T *)inv, (T *)scratch, k, backward, 0,
                         1, bpt, cbpt, 0};
        const int njobs = bpt * B.nt * cbpt;
        return backward ? launch_wide<T, false, false>(d, njobs, status, s)
                        : launch_wide<T, false, true>(d, njobs, status, s);
    }
    const int rchunks = (B.mb + SW_BM - 1) / SW_BM;
    const int cchunks = (min(B.nb, B.n) + SW_BN - 1) / SW_BN;
    hipLaunchKernelGGL(sweep_diag_kernel<T>, dim3(rchunks, B.nt * cchunks),
                       dim3(NTHREADS), 0, s, A, B, backward, k, cchunks,
                       (const T *)inv, (T *)scratch, (size_t)0, 0, 0,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

// Rows updated: forward m in [k + 1, mhi), backward m in [mlo, k).
template <typename T>
int launch_sweep_update(plasma_enum_t uplo, int backward, plasma_desc_t A,
                        plasma_desc_t B, int k, int mlo, int mhi,
                        const void *scratch, const int *status, hipStream_t s)
{
    int ntiles = backward ? k - mlo : mhi - k - 1;
    if (ntiles < 0) ntiles = 0;
    const int rchunks = (B.mb + SW_BM - 1) / SW_BM;
    const int cchunks = (min(B.nb, B.n) + SW_BN - 1) / SW_BN;
    if (B.n >= WIDE_RHS) {
        // the update as block GEMMs; sweep_update_kernel below then only
        // copies S into B(k, :)
        const int upper = (uplo == PlasmaUpper);
        const int bpt = (A.mb + tg::BM - 1) / tg::BM;
        const int cbpt = (B.nb + tg::BM - 1) / tg::BM;
        SweepDecode<T> d{A, B, nullptr, (T *)scratch, k, backward, upper, 0,
                         bpt, cbpt, mlo};
        const int njobs = bpt * B.nt * cbpt * ntiles;
        // X has unit row stride for forward-lower and backward-upper
        const int err = (backward == upper)
            ? launch_wide<T, true, true>(d, njobs, status, s)
            : launch_wide<T, true, false>(d, njobs, status, s);
        if (err) return err;
        ntiles = 0;
    }
    hipLaunchKernelGGL(sweep_update_kernel<T>,
                       dim3(ntiles * rchunks + rchunks, B.nt * cchunks),
                       dim3(NTHREADS), 0, s, A, B, uplo == PlasmaUpper,
                       backward, k, mlo, ntiles, rchunks, cchunks,
                       (const T *)scratch, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

// ---------------------------------------------------------------------------
// Sweeps with the scaled panels M(m,k) = L(m,k) W_kk, m > k (all strictly
// lower tiles, packed column by column: tile (m, k) at
// M + (k (mt-1) - k (k-1) / 2 + m - k - 1) * mb * nb, column major, ld mb).
// With L = M D, D = diag(L_kk), the forward sweep is Z = M^{-1} B (unit
// block diagonal: one update launch per step), Y = D^{-1} Z, and the
// backward sweep solves M^T X = D^{-T} Y: per step one launch again.
// ---------------------------------------------------------------------------
__device__ __forceinline__ size_t mpanel_index(int mt, int m, int k)
{
    return (size_t)k * (mt - 1) - (size_t)k * (k - 1) / 2 + (m - k - 1);
}

template <typename T>
struct ScaleDecode {
    plasma_desc_t A;
    int bpt, upper;
    const T *inv;
    T *M;
    __device__ bool operator()(int b, tg::Job<T> &j) const
    {
        const int per_tile = bpt * bpt;
        const int t = b / per_tile, sub_ = b % per_tile;
        const int bi = sub_ / bpt, bj = sub_ % bpt;
        int mm, k;
        trapezoid_decode(t, A.mt - 1, 0, &mm, &k);
        const int m = mm + 1;
        const int mv = plasma_tile_mview(A, m), nb = plasma_tile_mview(A, k);
        const int r0 = bi * tg::BM, c0 = bj * tg::BM;
        if (r0 >= mv || c0 >= nb) return false;
        View<T> Y;                              // Y(c, kk) = W_kk(kk, c)
        Y.p = const_cast<T *>(inv) + (size_t)k * A.mb * A.mb +
              (size_t)c0 * A.mb;
        Y.rs = A.mb;
        Y.cs = 1;
        View<T> C;
        C.p = M + (size_t)t * A.mb * A.nb + r0 + (size_t)c0 * A.mb;
        C.rs = 1;
        C.cs = A.mb;
        j.X = sub(lower_view<T>(A, upper, m, k), r0, 0);
        j.Y = Y;
        j.C = C;
        j.xrows = mv - r0;
        j.yrows = nb - c0;
        j.K = nb;
        j.lower = false;
        j.diag_off = 0;
        return true;
    }
};

template <typename T>
int launch_scale_panels(plasma_enum_t uplo, plasma_desc_t A, const void *inv,
                        void *M, const int *status, hipStream_t s)
{
    if (A.mt < 2) return 0;
    const int bpt = (A.mb + tg::BM - 1) / tg::BM;
    const long nblocks = (long)A.mt * (A.mt - 1) / 2 * bpt * bpt;
    if (nblocks > 0x7fffffffL) return PlasmaErrorNotSupported;
    ScaleDecode<T> d{A, bpt, uplo == PlasmaUpper, (const T *)inv, (T *)M};
    if (uplo == PlasmaUpper)
        hipLaunchKernelGGL((tg::kernel<T, true, false, false, false, false,
                                       ScaleDecode<T>>),
                           dim3(tg::xcd_grid((int)nblocks)), dim3(tg::NT), 0,
                           s, d, (int)nblocks, status);
    else
        hipLaunchKernelGGL((tg::kernel<T, true, false, false, true, false,
                                       ScaleDecode<T>>),
                           dim3(tg::xcd_grid((int)nblocks)), dim3(tg::NT), 0,
                           s, d, (int)nblocks, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

// Forward: B(m,:) -= M(m,k) B(k,:) for m > k.  Backward: B(j,:) -=
// M(k,j)^T B(k,:) for j < k.  B(k,:) is final when step k runs.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void msweep_update_kernel(
    plasma_desc_t A, plasma_desc_t B, int backward, int k, int rchunks,
    int cchunks, const T *M, const int *status)
{
    typedef GemmNT<T, SW_BN, SW_BM, SW_KC, 16, 32> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int tn = blockIdx.y / cchunks;
    const int c0 = (blockIdx.y % cchunks) * SW_BN;
    const int nv = plasma_tile_nview(B, tn);
    const int K = plasma_tile_mview(B, k);
    const int m = backward ? (int)(blockIdx.x / rchunks)
                           : k + 1 + (int)(blockIdx.x / rchunks);
    const int r0 = (blockIdx.x % rchunks) * SW_BM;
    const int mv = plasma_tile_mview(B, m);
    if (r0 >= mv || c0 >= nv) return;
    View<T> X;
    if (!backward) {                            // X(r, kk) = M(m,k)(r, kk)
        X.p = const_cast<T *>(M) + mpanel_index(A.mt, m, k) * A.mb * A.nb +
              r0;
        X.rs = 1;
        X.cs = A.mb;
    }
    else {                                      // X(r, kk) = M(k,m)(kk, r)
        X.p = const_cast<T *>(M) + mpanel_index(A.mt, k, m) * A.mb * A.nb +
              (size_t)r0 * A.mb;
        X.rs = A.mb;
        X.cs = 1;
    }
    const View<T> Bk = tile_view<T>(B, k, tn);
    View<T> Y;                                   // Y(c, kk) = B(k)(kk, c0 + c)
    Y.p = Bk.p + (ptrdiff_t)c0 * Bk.cs;
    Y.rs = Bk.cs;
    Y.cs = 1;
    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    G::run(acc, Y, nv - c0, X, mv - r0, K, sm);
    store_rows<G, T>(acc, tile_view<T>(B, m, tn), r0, c0, mv, nv, true);
}

template <typename T>
int launch_msweep_update(int backward, plasma_desc_t A, plasma_desc_t B,
                         int k, const void *M, const int *status,
                         hipStream_t s)
{
    const int ntiles = backward ? k : B.mt - k - 1;
    if (ntiles <= 0) return 0;
    const int rchunks = (B.mb + SW_BM - 1) / SW_BM;
    const int cchunks = (min(B.nb, B.n) + SW_BN - 1) / SW_BN;
    hipLaunchKernelGGL(msweep_update_kernel<T>,
                       dim3(ntiles * rchunks, B.nt * cchunks), dim3(NTHREADS),
                       0, s, A, B, backward, k, rchunks, cchunks,
                       (const T *)M, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

// B(k,:) = W_kk^T W_kk B(k,:) for every k: two launches over all tiles
// through the scratch (mt slabs of mb x (nt * nb)).
template <typename T>
int launch_diag_both(plasma_desc_t A, plasma_desc_t B, const void *inv,
                     void *scratch, const int *status, hipStream_t s)
{
    const int rchunks = (B.mb + SW_BM - 1) / SW_BM;
    const int cchunks = (min(B.nb, B.n) + SW_BN - 1) / SW_BN;
    const size_t kstride = (size_t)A.mb * B.nt * B.nb;
    dim3 grid(rchunks, B.nt * cchunks, B.mt);
    hipLaunchKernelGGL(sweep_diag_kernel<T>, grid, dim3(NTHREADS), 0, s, A, B,
                       0, 0, cchunks, (const T *)inv, (T *)scratch, kstride,
                       0, 0, status);
    hipLaunchKernelGGL(sweep_diag_kernel<T>, grid, dim3(NTHREADS), 0, s, A, B,
                       1, 0, cchunks, (const T *)inv, (T *)scratch, kstride,
                       1, 1, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

}  // namespace

extern "C" int plasma_kern_trsm_diag_left(plasma_enum_t uplo, int backward,
                                          plasma_desc_t A, plasma_desc_t B,
                                          int k, const void *inv,
                                          void *scratch, const int *status,
                                          hipStream_t s)
{
    (void)uplo;
    if (A.precision != B.precision) return PlasmaErrorIllegalValue;
    if (A.precision == PlasmaRealDouble)
        return launch_sweep_diag<double>(backward, A, B, k, inv, scratch,
                                         status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_sweep_diag<float>(backward, A, B, k, inv, scratch,
                                        status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_sweep_update(plasma_enum_t uplo, int backward,
                                        plasma_desc_t A, plasma_desc_t B,
                                        int k, const void *scratch,
                                        const int *status, hipStream_t s)
{
    if (A.precision != B.precision) return PlasmaErrorIllegalValue;
    if (A.precision == PlasmaRealDouble)
        return launch_sweep_update<double>(uplo, backward, A, B, k, 0, B.mt,
                                           scratch, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_sweep_update<float>(uplo, backward, A, B, k, 0, B.mt,
                                          scratch, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_sweep_update_range(plasma_enum_t uplo,
                                              int backward, plasma_desc_t A,
                                              plasma_desc_t B, int k, int mlo,
                                              int mhi, const void *scratch,
                                              const int *status,
                                              hipStream_t s)
{
    if (A.precision != B.precision) return PlasmaErrorIllegalValue;
    mlo = max(mlo, 0);
    mhi = min(mhi, B.mt);
    if (A.precision == PlasmaRealDouble)
        return launch_sweep_update<double>(uplo, backward, A, B, k, mlo, mhi,
                                           scratch, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_sweep_update<float>(uplo, backward, A, B, k, mlo, mhi,
                                          scratch, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_scale_panels(plasma_enum_t uplo, plasma_desc_t A,
                                        const void *inv, void *M,
                                        const int *status, hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_scale_panels<double>(uplo, A, inv, M, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_scale_panels<float>(uplo, A, inv, M, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_msweep_update(int backward, plasma_desc_t A,
                                         plasma_desc_t B, int k,
                                         const void *M, const int *status,
                                         hipStream_t s)
{
    if (A.precision != B.precision) return PlasmaErrorIllegalValue;
    if (A.precision == PlasmaRealDouble)
        return launch_msweep_update<double>(backward, A, B, k, M, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_msweep_update<float>(backward, A, B, k, M, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_diag_normal(plasma_desc_t A, plasma_desc_t B,
                                       const void *inv, void *scratch,
                                       const int *status, hipStream_t s)
{
    if (A.precision != B.precision) return PlasmaErrorIllegalValue;
    if (A.precision == PlasmaRealDouble)
        return launch_diag_both<double>(A, B, inv, scratch, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_diag_both<float>(A, B, inv, scratch, status, s);
    return PlasmaErrorNotSupported;
}
