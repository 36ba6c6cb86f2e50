// solve.hip -- triangular sweeps of potrs / posv for gfx950.
//
// The reference solves L Y = B and L^T X = Y with pztrsm
// (compute/pztrsm.c:113-144 Left-Lower-NoTrans, :148-177 Left-Lower-Trans,
// :43-107 the upper variants): per block row k a tile trsm with the
// diagonal tile, then a gemm update of the block rows still to be solved.
// Here each of those two task classes is one launch per step: a trsm over
// all tile columns of B(k,:), and an MFMA update over every block row
// m > k (forward) or m < k (backward) and every tile column of B.
#include "common.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NTHREADS = 256;
constexpr int SOLVE_RB = 32;   // rows per step of the diagonal solve
constexpr int SOLVE_NC = 8;    // right-hand sides per workgroup

// ---------------------------------------------------------------------------
// Diagonal solve for block row k: Y = L(k,k)^{-1} B(k,:) (forward) or
// L(k,k)^{-T} B(k,:) (backward).  One workgroup per 8 right-hand sides; the
// 8 columns live in LDS; rows are processed in blocks of 32 (a gemv-like
// accumulation with the rows already solved, then a 32-row substitution).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NTHREADS) void trsm_diag_left_kernel(
    plasma_desc_t A, plasma_desc_t B, int upper, int backward, int k,
    int chunks_per_tilecol, const int *status)
{
    if (block_should_stop(status)) return;
    extern __shared__ unsigned char smem_raw[];
    T *Y = reinterpret_cast<T *>(smem_raw);  // n x SOLVE_NC, row major

    const int tn = blockIdx.x / chunks_per_tilecol;
    const int c0 = (blockIdx.x % chunks_per_tilecol) * SOLVE_NC;
    const int nv = plasma_tile_nview(B, tn);
    if (c0 >= nv) return;
    const int ncols = min(SOLVE_NC, nv - c0);
    const int n = plasma_tile_mview(A, k);
    const View<T> L = lower_view<T>(A, upper != 0, k, k);
    View<T> Bk = tile_view<T>(B, k, tn);
    Bk.p += (ptrdiff_t)c0 * Bk.cs;
    const int t = threadIdx.x;
    const int r = t % SOLVE_RB, c = t / SOLVE_RB;  // 32 rows x 8 columns

    for (int e = t; e < n * SOLVE_NC; e += NTHREADS) {
        int rr = e % n, cc = e / n;
        Y[rr * SOLVE_NC + cc] = (cc < ncols) ? Bk.at(rr, cc) : T(0);
    }
    __syncthreads();

    if (!backward) {
        for (int i0 = 0; i0 < n; i0 += SOLVE_RB) {
            const int ib = min(SOLVE_RB, n - i0);
            if (r < ib && i0 > 0) {
                T acc = T(0);
                for (int l = 0; l < i0; l++)
                    acc += L.at(i0 + r, l) * Y[l * SOLVE_NC + c];
                Y[(i0 + r) * SOLVE_NC + c] -= acc;
            }
            __syncthreads();
            if (r == 0) {
                for (int i = 0; i < ib; i++) {
                    T v = Y[(i0 + i) * SOLVE_NC + c];
                    for (int l = 0; l < i; l++)
                        v -= L.at(i0 + i, i0 + l) * Y[(i0 + l) * SOLVE_NC + c];
                    Y[(i0 + i) * SOLVE_NC + c] = v / L.at(i0 + i, i0 + i);
                }
            }
            __syncthreads();
        }
    }
    else {
        const int nblk = (n + SOLVE_RB - 1) / SOLVE_RB;
        for (int bI = nblk - 1; bI >= 0; bI--) {
            const int i0 = bI * SOLVE_RB;
            const int ib = min(SOLVE_RB, n - i0);
            const int i1 = i0 + ib;
            if (r < ib && i1 < n) {
                T acc = T(0);
                for (int l = i1; l < n; l++)
                    acc += L.at(l, i0 + r) * Y[l * SOLVE_NC + c];
                Y[(i0 + r) * SOLVE_NC + c] -= acc;
            }
            __syncthreads();
            if (r == 0) {
                for (int i = ib - 1; i >= 0; i--) {
                    T v = Y[(i0 + i) * SOLVE_NC + c];
                    for (int l = i + 1; l < ib; l++)
                        v -= L.at(i0 + l, i0 + i) * Y[(i0 + l) * SOLVE_NC + c];
                    Y[(i0 + i) * SOLVE_NC + c] = v / L.at(i0 + i, i0 + i);
                }
            }
            __syncthreads();
        }
    }

    for (int e = t; e < n * SOLVE_NC; e += NTHREADS) {
        int rr = e % n, cc = e / n;
        if (cc < ncols) Bk.at(rr, cc) = Y[rr * SOLVE_NC + cc];
    }
}

// ---------------------------------------------------------------------------
// Sweep update on MFMA.  Forward: B(m,:) -= L(m,k) B(k,:) for m > k.
// Backward: B(m,:) -= L(k,m)^T B(k,:) for m < k.  64x64 output blocks.
// ---------------------------------------------------------------------------
constexpr int SW_BM = 64, SW_BN = 64;

template <typename T>
__global__ __launch_bounds__(NTHREADS) void sweep_update_kernel(
    plasma_desc_t A, plasma_desc_t B, int upper, int backward, int k,
    int rchunks, int cchunks, const int *status)
{
    typedef GemmNT<T, SW_BM, SW_BN, 16, 32, 32> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;

    const int m = backward ? (int)(blockIdx.x / rchunks)
                           : k + 1 + (int)(blockIdx.x / rchunks);
    const int r0 = (blockIdx.x % rchunks) * SW_BM;
    const int tn = blockIdx.y / cchunks;
    const int c0 = (blockIdx.y % cchunks) * SW_BN;
    const int mv = plasma_tile_mview(B, m);
    const int nv = plasma_tile_nview(B, tn);
    if (r0 >= mv || c0 >= nv) return;
    const int K = plasma_tile_mview(B, k);

    View<T> X;
    if (!backward) {
        X = lower_view<T>(A, upper != 0, m, k);        // X(r, kk) = L(m,k)
    }
    else {
        View<T> v = lower_view<T>(A, upper != 0, k, m);  // L(k,m)
        X.p = v.p; X.rs = v.cs; X.cs = v.rs;             // transpose
    }
    X.p += (ptrdiff_t)r0 * X.rs;
    const View<T> Bk = tile_view<T>(B, k, tn);
    View<T> Y;                         // Y(c, kk) = B(k, tn)(kk, c)
    Y.p = Bk.p + (ptrdiff_t)c0 * Bk.cs;
    Y.rs = Bk.cs;
    Y.cs = 1;
    View<T> C = tile_view<T>(B, m, tn);

    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    G::run(acc, X, mv - r0, Y, nv - c0, K, sm);
#pragma unroll
    for (int i = 0; i < G::FM; i++)
#pragma unroll
        for (int j = 0; j < G::FN; j++) {
            const int cc = c0 + G::acc_col(j);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int rr = r0 + G::acc_row(i, q);
                if (rr < mv && cc < nv) C.at(rr, cc) -= acc[i][j][q];
            }
        }
}

template <typename T>
int launch_trsm_diag(plasma_enum_t uplo, int backward, plasma_desc_t A,
                     plasma_desc_t B, int k, const int *status, hipStream_t s)
{
    const int chunks = (B.nb + SOLVE_NC - 1) / SOLVE_NC;
    const size_t lds = (size_t)A.mb * SOLVE_NC * sizeof(T);
    if (lds > 65536) return PlasmaErrorNotSupported;
    hipLaunchKernelGGL(trsm_diag_left_kernel<T>, dim3(B.nt * chunks),
                       dim3(NTHREADS), lds, s, A, B, uplo == PlasmaUpper,
                       backward, k, chunks, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T>
int launch_sweep_update(plasma_enum_t uplo, int backward, plasma_desc_t A,
                        plasma_desc_t B, int k, const int *status,
                        hipStream_t s)
{
    const int ntiles = backward ? k : B.mt - k - 1;
    if (ntiles <= 0) return 0;
    const int rchunks = (B.mb + SW_BM - 1) / SW_BM;
    const int cchunks = (B.nb + SW_BN - 1) / SW_BN;
    hipLaunchKernelGGL(sweep_update_kernel<T>, dim3(ntiles * rchunks,
                       B.nt * cchunks), dim3(NTHREADS), 0, s, A, B,
                       uplo == PlasmaUpper, backward, k, rchunks, cchunks,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

}  // namespace

extern "C" int plasma_kern_trsm_diag_left(plasma_enum_t uplo, int backward,
                                          plasma_desc_t A, plasma_desc_t B,
                                          int k, const int *status,
                                          hipStream_t s)
{
    if (A.precision != B.precision) return PlasmaErrorIllegalValue;
    if (A.precision == PlasmaRealDouble)
        return launch_trsm_diag<double>(uplo, backward, A, B, k, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_trsm_diag<float>(uplo, backward, A, B, k, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_sweep_update(plasma_enum_t uplo, int backward,
                                        plasma_desc_t A, plasma_desc_t B,
                                        int k, const int *status,
                                        hipStream_t s)
{
    if (A.precision != B.precision) return PlasmaErrorIllegalValue;
    if (A.precision == PlasmaRealDouble)
        return launch_sweep_update<double>(uplo, backward, A, B, k, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_sweep_update<float>(uplo, backward, A, B, k, status, s);
    return PlasmaErrorNotSupported;
}
