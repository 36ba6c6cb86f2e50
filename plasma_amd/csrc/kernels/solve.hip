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
// S back into B(k,:).
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

// S(:, cols of tile column tn) = op(W_kk) B(k, tn); S column major with
// ld = mb, global column index tn * nb + c.
template <typename T>
struct SweepDiagDecode {
    plasma_desc_t A, B;
    int k, backward, rblocks, cblocks;
    const T *inv;
    T *scratch;
    __device__ bool operator()(int b, tg::Job<T> &j) const
    {
        const int per_tn = rblocks * cblocks;
        const int tn = b / per_tn, rem = b % per_tn;
        const int r0 = (rem / cblocks) * tg::BM, c0 = (rem % cblocks) * tg::BM;
        const int mv = plasma_tile_mview(B, k), nv = plasma_tile_nview(B, tn);
        if (r0 >= mv || c0 >= nv) return false;
        View<T> W;
        W.p = const_cast<T *>(inv) + (size_t)k * A.mb * A.mb;
        if (!backward) { W.rs = 1; W.cs = A.mb; }     // X(r, kk) = W(r, kk)
        else           { W.rs = A.mb; W.cs = 1; }     // X(r, kk) = W(kk, r)
        const View<T> Bk = tile_view<T>(B, k, tn);
        View<T> Y;                                    // Y(c, kk) = Bk(kk, c)
        Y.p = Bk.p;
        Y.rs = Bk.cs;
        Y.cs = 1;
        View<T> S;
        S.p = scratch + (size_t)tn * B.nb * A.mb;
        S.rs = 1;
        S.cs = A.mb;
        j.X = sub(W, r0, 0);
        j.Y = sub(Y, c0, 0);
        j.C = sub(S, r0, c0);
        j.xrows = mv - r0;
        j.yrows = nv - c0;
        j.K = mv;
        j.lower = false;
        j.diag_off = 0;
        return true;
    }
};

// ---------------------------------------------------------------------------
// Sweep update on MFMA, reading the solved block row from S.
// Forward: B(m,:) -= L(m,k) S for m > k.  Backward: B(m,:) -= L(k,m)^T S
// for m < k.  The last row of workgroups copies S into B(k,:).
// 64 x 64 output blocks.
// ---------------------------------------------------------------------------
constexpr int SW_BM = 64, SW_BN = 64;

template <typename T>
__global__ __launch_bounds__(NTHREADS) void sweep_update_kernel(
    plasma_desc_t A, plasma_desc_t B, int upper, int backward, int k,
    int ntiles, int rchunks, int cchunks, const T *scratch,
    const int *status)
{
    typedef GemmNT<T, SW_BM, SW_BN, 16, 32, 32> G;
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
int launch_sweep_diag(int backward, plasma_desc_t A, plasma_desc_t B, int k,
                      const void *inv, void *scratch, const int *status,
                      hipStream_t s)
{
    const int rblocks = (B.mb + tg::BM - 1) / tg::BM;
    const int cblocks = (B.nb + tg::BM - 1) / tg::BM;
    SweepDiagDecode<T> d{A, B, k, backward, rblocks, cblocks,
                         (const T *)inv, (T *)scratch};
    if (backward)
        hipLaunchKernelGGL((tg::kernel<T, true, false, false, false, false,
                                       SweepDiagDecode<T>>),
                           dim3(B.nt * rblocks * cblocks), dim3(tg::NT), 0, s,
                           d, status);
    else
        hipLaunchKernelGGL((tg::kernel<T, true, false, false, true, false,
                                       SweepDiagDecode<T>>),
                           dim3(B.nt * rblocks * cblocks), dim3(tg::NT), 0, s,
                           d, status);
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
