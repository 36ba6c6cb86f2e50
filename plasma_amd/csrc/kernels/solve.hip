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
template <typename T>
__global__ __launch_bounds__(NTHREADS) void sweep_diag_kernel(
    plasma_desc_t A, plasma_desc_t B, int backward, int k, int cchunks,
    const T *inv, T *scratch, const int *status)
{
    typedef GemmNT<T, SW_BN, SW_BM, SW_KC, 16, 32> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
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
    const View<T> Bk = tile_view<T>(B, k, tn);
    View<T> Y;                                  // Y(c, kk) = B(k)(klo + kk, c0 + c)
    Y.p = Bk.p + klo + (ptrdiff_t)c0 * Bk.cs;
    Y.rs = Bk.cs;
    Y.cs = 1;
    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    G::run(acc, Y, nv - c0, X, mv - r0, khi - klo, sm);
    View<T> S;
    S.p = scratch + (size_t)tn * B.nb * A.mb;
    S.rs = 1;
    S.cs = A.mb;
    store_rows<G, T>(acc, S, r0, c0, mv, nv, false);
}

// ---------------------------------------------------------------------------
// Sweep update, reading the solved block row from S.
// Forward: B(m,:) -= L(m,k) S for m > k.  Backward: B(m,:) -= L(k,m)^T S
// for m < k.  The last row of workgroups copies S into B(k,:).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NTHREADS) void sweep_update_kernel(
    plasma_desc_t A, plasma_desc_t B, int upper, int backward, int k,
    int ntiles, int rchunks, int cchunks, const T *scratch,
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

    const int m = backward ? (int)(blockIdx.x / rchunks)
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

template <typename T>
int launch_sweep_diag(int backward, plasma_desc_t A, plasma_desc_t B, int k,
                      const void *inv, void *scratch, const int *status,
                      hipStream_t s)
{
    const int rchunks = (B.mb + SW_BM - 1) / SW_BM;
    const int cchunks = (B.nb + SW_BN - 1) / SW_BN;
    hipLaunchKernelGGL(sweep_diag_kernel<T>, dim3(rchunks, B.nt * cchunks),
                       dim3(NTHREADS), 0, s, A, B, backward, k, cchunks,
                       (const T *)inv, (T *)scratch, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T>
int launch_sweep_update(plasma_enum_t uplo, int backward, plasma_desc_t A,
                        plasma_desc_t B, int k, const void *scratch,
                        const int *status, hipStream_t s)
{
    const int ntiles = backward ? k : B.mt - k - 1;
    const int rchunks = (B.mb + SW_BM - 1) / SW_BM;
    const int cchunks = (B.nb + SW_BN - 1) / SW_BN;
    hipLaunchKernelGGL(sweep_update_kernel<T>,
                       dim3(ntiles * rchunks + rchunks, B.nt * cchunks),
                       dim3(NTHREADS), 0, s, A, B, uplo == PlasmaUpper,
                       backward, k, ntiles, rchunks, cchunks,
                       (const T *)scratch, status);
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
        return launch_sweep_update<double>(uplo, backward, A, B, k, scratch,
                                           status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_sweep_update<float>(uplo, backward, A, B, k, scratch,
                                          status, s);
    return PlasmaErrorNotSupported;
}
