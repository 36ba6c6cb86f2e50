// potrf.hip -- tile Cholesky kernels for gfx950: the diagonal tile
// factorization, the panel triangular solve and the MFMA trailing update.
//
// They replace the tile tasks of the reference's right-looking tile
// Cholesky (compute/pzpotrf.c:37-80 lower, :84-126 upper): one launch per
// step and task class processes every tile of that step, instead of one
// OpenMP task per tile.  All kernels use the lower-equivalent convention of
// common.hip.h (upper storage is read as L = U^T).
#include "common.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int DIAG_JB = 32;   // column block of the diagonal factorization
constexpr int NTHREADS = 256;

// ---------------------------------------------------------------------------
// Diagonal tile: L(k,k) L(k,k)^T = A(k,k), one workgroup, blocked
// right-looking with jb-column blocks (jb <= 32, chosen so the panel rows of
// one block fit in 64 KiB of LDS).  The jb x jb diagonal block is factored
// in LDS, the rows below it are solved against it in LDS, written back, and
// the remaining lower triangle of the tile is updated from LDS with 4x4
// register blocks.
// On a non-positive or NaN pivot at local column j the status word becomes
// k*nb + j + 1 (reference core_zpotrf.c:72-79 with iinfo = A.nb*k), the
// LAPACK dpotrf convention for the order of the failing leading minor.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NTHREADS) void potrf_diag_kernel(
    plasma_desc_t A, int upper, int k, int JB, int *status)
{
    if (block_should_stop(status)) return;
    extern __shared__ unsigned char smem_raw[];
    const int n = plasma_tile_mview(A, k);
    T *P = reinterpret_cast<T *>(smem_raw);  // (n - jb) x JB, ld = n
    const View<T> L = lower_view<T>(A, upper != 0, k, k);
    const int t = threadIdx.x;

    __shared__ T D[DIAG_JB][DIAG_JB + 1];
    __shared__ int s_info;
    if (t == 0) s_info = 0;

    for (int j0 = 0; j0 < n; j0 += JB) {
        const int jb = min(JB, n - j0);
        const int i0 = j0 + jb;
        const int nrows = n - i0;

        // (a) diagonal block and the panel rows below it into LDS
        for (int e = t; e < jb * jb; e += NTHREADS) {
            int r = e % jb, c = e / jb;
            D[r][c] = (r >= c) ? L.at(j0 + r, j0 + c) : T(0);
        }
        for (int e = t; e < nrows * jb; e += NTHREADS) {
            int r = e % nrows, c = e / nrows;
            P[r + c * n] = L.at(i0 + r, j0 + c);
        }
        __syncthreads();

        // (b) unblocked right-looking Cholesky of the block, rows on lanes
        for (int j = 0; j < jb; j++) {
            T d = D[j][j];
            if (!(d > T(0))) {  // also catches NaN
                if (t == 0) s_info = j0 + j + 1;
                break;  // uniform: every thread read the same d
            }
            d = sqrt(d);
            __syncthreads();
            if (t == j) D[j][j] = d;
            if (t > j && t < jb) D[t][j] /= d;
            __syncthreads();
            if (t > j && t < jb) {
                T lij = D[t][j];
                for (int l = j + 1; l <= t; l++) D[t][l] -= lij * D[l][j];
            }
            __syncthreads();
        }
        __syncthreads();
        if (s_info != 0) {
            if (t == 0)
                atomicCAS(status, 0, A.nb * k + s_info);
            return;
        }

        // (c) panel rows: X D^T = P, column by column (right-looking)
        for (int c = 0; c < jb; c++) {
            const T dcc = D[c][c];
            for (int r = t; r < nrows; r += NTHREADS) P[r + c * n] /= dcc;
            __syncthreads();
            const int rest = jb - c - 1;
            for (int e = t; e < nrows * rest; e += NTHREADS) {
                int r = e % nrows, l = c + 1 + e / nrows;
                P[r + l * n] -= P[r + c * n] * D[l][c];
            }
            __syncthreads();
        }

        // write the factored block and the solved panel back
        for (int e = t; e < jb * jb; e += NTHREADS) {
            int r = e % jb, c = e / jb;
            if (r >= c) L.at(j0 + r, j0 + c) = D[r][c];
        }
        for (int e = t; e < nrows * jb; e += NTHREADS) {
            int r = e % nrows, c = e / nrows;
            L.at(i0 + r, j0 + c) = P[r + c * n];
        }

        // (d) trailing lower triangle of the tile: A22 -= P P^T, 4x4 blocks
        const int nb4 = (nrows + 3) / 4;
        const int nblk = nb4 * (nb4 + 1) / 2;
        for (int b = t; b < nblk; b += NTHREADS) {
            // decode (bi >= bj) from the row-major lower enumeration
            int bi = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
            while (bi * (bi + 1) / 2 > b) bi--;
            while ((bi + 1) * (bi + 2) / 2 <= b) bi++;
            int bj = b - bi * (bi + 1) / 2;
            T acc[4][4];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] = T(0);
            for (int l = 0; l < jb; l++) {
                T pr[4], pc[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    int r = min(4 * bi + u, nrows - 1);
                    int c = min(4 * bj + u, nrows - 1);
                    pr[u] = P[r + l * n];
                    pc[u] = P[c + l * n];
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[u][v] += pr[u] * pc[v];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    int r = 4 * bi + u, c = 4 * bj + v;
                    if (r < nrows && c <= r)
                        L.at(i0 + r, i0 + c) -= acc[u][v];
                }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Panel: X L(k,k)^T = A(m,k) for every m > k.  One workgroup per 32 rows
// (fewer for nb > 256) of one tile; the row block lives in LDS while the
// columns are solved left to right in 32-wide blocks.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NTHREADS) void trsm_panel_kernel(
    plasma_desc_t A, int upper, int k, int rb, int chunks_per_tile,
    const int *status)
{
    if (block_should_stop(status)) return;
    extern __shared__ unsigned char smem_raw[];
    T *X = reinterpret_cast<T *>(smem_raw);  // rb x nb, column major

    const int m = k + 1 + blockIdx.x / chunks_per_tile;
    const int r0 = (blockIdx.x % chunks_per_tile) * rb;
    const int mv = plasma_tile_mview(A, m);
    if (r0 >= mv) return;
    const int rows = min(rb, mv - r0);
    const int nb = plasma_tile_mview(A, k);  // order of L(k,k) (== nb)
    const View<T> Lkk = lower_view<T>(A, upper != 0, k, k);
    View<T> B = lower_view<T>(A, upper != 0, m, k);
    B.p += (ptrdiff_t)r0 * B.rs;
    const int t = threadIdx.x;

    for (int e = t; e < rb * nb; e += NTHREADS) {
        int r = e % rb, c = e / rb;
        X[r + c * rb] = (r < rows) ? B.at(r, c) : T(0);
    }
    __syncthreads();

    const int ngroups = NTHREADS / rb;
    const int r = t % rb, g = t / rb;
    for (int j0 = 0; j0 < nb; j0 += DIAG_JB) {
        const int jb = min(DIAG_JB, nb - j0);
        // solve the jb columns of this block, one row per thread
        if (t < rb) {
            for (int c = j0; c < j0 + jb; c++) {
                T v = X[r + c * rb];
                for (int l = j0; l < c; l++) v -= X[r + l * rb] * Lkk.at(c, l);
                X[r + c * rb] = v / Lkk.at(c, c);
            }
        }
        __syncthreads();
        // update the columns to the right
        for (int c = j0 + jb + g; c < nb; c += ngroups) {
            T acc = T(0);
            for (int l = j0; l < j0 + jb; l++) acc += X[r + l * rb] * Lkk.at(c, l);
            X[r + c * rb] -= acc;
        }
        __syncthreads();
    }

    for (int e = t; e < rb * nb; e += NTHREADS) {
        int rr = e % rb, c = e / rb;
        if (rr < rows) B.at(rr, c) = X[rr + c * rb];
    }
}

// ---------------------------------------------------------------------------
// Trailing update on MFMA: for tiles (m, n), n in [n_first, n_end),
// m in [n, mt): A(m,n) -= A(m,k) A(n,k)^T; on diagonal tiles only the lower
// triangle is written (the reference's syrk with uplo = Lower leaves the
// strictly upper part of diagonal tiles untouched).
// 128x128 output blocks, 4 waves of 64x64, K = nb streamed in chunks.
// ---------------------------------------------------------------------------
template <typename T> struct UpdCfg;
template <> struct UpdCfg<double> { static constexpr int KC = 16; };
template <> struct UpdCfg<float>  { static constexpr int KC = 32; };

constexpr int UPD_BM = 128;

template <typename T>
__global__ __launch_bounds__(NTHREADS) void trailing_update_kernel(
    plasma_desc_t A, int upper, int k, int n_first, int bpt,
    const int *status)
{
    typedef GemmNT<T, UPD_BM, UPD_BM, UpdCfg<T>::KC, 64, 64> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;

    const int per_tile = bpt * bpt;
    const int tile = blockIdx.x / per_tile;
    const int sub = blockIdx.x % per_tile;
    const int bi = sub / bpt, bj = sub % bpt;
    int m, n;
    trapezoid_decode(tile, A.mt, n_first, &m, &n);
    if (m == n && bj > bi) return;
    const int mv = plasma_tile_mview(A, m);
    const int nv = plasma_tile_mview(A, n);
    const int r0 = bi * UPD_BM, c0 = bj * UPD_BM;
    if (r0 >= mv || c0 >= nv) return;

    View<T> X = lower_view<T>(A, upper != 0, m, k);
    View<T> Y = lower_view<T>(A, upper != 0, n, k);
    View<T> C = lower_view<T>(A, upper != 0, m, n);
    X.p += (ptrdiff_t)r0 * X.rs;
    Y.p += (ptrdiff_t)c0 * Y.rs;
    const int K = A.nb;

    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    G::run(acc, X, mv - r0, Y, nv - c0, K, sm);

    const bool diag = (m == n);
#pragma unroll
    for (int i = 0; i < G::FM; i++)
#pragma unroll
        for (int j = 0; j < G::FN; j++) {
            const int c = c0 + G::acc_col(j);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = r0 + G::acc_row(i, q);
                if (r < mv && c < nv && (!diag || r >= c))
                    C.at(r, c) -= acc[i][j][q];
            }
        }
}

template <typename T>
int launch_potrf_diag(plasma_enum_t uplo, plasma_desc_t A, int k, int *status,
                      hipStream_t s)
{
    // column block: panel rows (n - jb) x jb must fit in 64 KiB of LDS
    const int n = A.mb;
    int jb = DIAG_JB;
    while (jb > 1 && (size_t)n * jb * sizeof(T) > 65536) jb /= 2;
    const size_t lds = (size_t)n * jb * sizeof(T);
    if (lds > 65536) return PlasmaErrorNotSupported;
    hipLaunchKernelGGL(potrf_diag_kernel<T>, dim3(1), dim3(NTHREADS), lds, s,
                       A, uplo == PlasmaUpper, k, jb, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T>
int launch_trsm_panel(plasma_enum_t uplo, plasma_desc_t A, int k,
                      const int *status, hipStream_t s)
{
    const int ntiles = A.mt - k - 1;
    if (ntiles <= 0) return 0;
    const int nb = A.nb;
    // rows per workgroup: keep the LDS row block at 64 KiB or less
    int rb = 64;
    while (rb > 8 && (size_t)rb * nb * sizeof(T) > 65536) rb /= 2;
    if ((size_t)rb * nb * sizeof(T) > 65536 || NTHREADS % rb) return PlasmaErrorNotSupported;
    const int chunks = (A.mb + rb - 1) / rb;
    const size_t lds = (size_t)rb * nb * sizeof(T);
    hipLaunchKernelGGL(trsm_panel_kernel<T>, dim3(ntiles * chunks),
                       dim3(NTHREADS), lds, s, A, uplo == PlasmaUpper, k, rb,
                       chunks, status);
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
    const int bpt = (A.mb + UPD_BM - 1) / UPD_BM;
    const long nblocks = ntiles * bpt * bpt;
    if (nblocks > 0x7fffffffL) return PlasmaErrorNotSupported;
    hipLaunchKernelGGL(trailing_update_kernel<T>, dim3((unsigned)nblocks),
                       dim3(NTHREADS), 0, s, A, uplo == PlasmaUpper, k,
                       n_first, bpt, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

}  // namespace

extern "C" int plasma_kern_potrf_diag(plasma_enum_t uplo, plasma_desc_t A,
                                      int k, int *status, hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_potrf_diag<double>(uplo, A, k, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_potrf_diag<float>(uplo, A, k, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_trsm_panel(plasma_enum_t uplo, plasma_desc_t A,
                                      int k, const int *status, hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_trsm_panel<double>(uplo, A, k, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_trsm_panel<float>(uplo, A, k, status, s);
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
