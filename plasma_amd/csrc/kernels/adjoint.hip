// adjoint.hip -- kernels of the Cholesky and potrs adjoints for gfx950 (no
// reference counterpart; the reference has no derivatives).
//
//  * ltl_lower_kernel: X(I,J) = sum_{K >= I} L(K,I)^T Lbar(K,J) for the
//    tiles I >= J, with L lower triangular (the factor, lower-equivalent
//    view) and Lbar lower triangular (upper storage: Ubar read transposed),
//    and writes Q = the symmetric matrix whose lower triangle is tril(X):
//    Q(I,J) = X(I,J) and Q(J,I) = X(I,J)^T.  64 x 64 output blocks, K in
//    16-deep chunks through double-buffered LDS, MFMA accumulation; the
//    diagonal tiles of L and Lbar are read with their triangles masked.
//  * tile_axpby_t_kernel: Z(I,J) = a X(I,J) + b Y(J,I)^T over all tiles of
//    square general descriptors (transpose, symmetric part), or the
//    symmetric matrix with lower triangle a X, through LDS so both sides
//    move in contiguous runs.
//  * syr2k_lapack_kernel: C -= s (U V^T + V U^T) on LAPACK-layout arrays
//    (the potrs adjoint's symmetric rank-2k term, nrhs small).
//  * Steps of the blocked reverse sweep (lower storage): panel_w_kernel
//    P(t) = Lbar(k+1+t, k) W_kk; gram_partial_kernel + gram_reduce_
//    kernel Lbar(k,k) -= sum_t P(t)^T L(k+1+t, k), split over the tiles t
//    into partial sums added in a fixed order; copy_panel_kernel P ->
//    Lbar(k+1.., k); tile_xw_kernel (twice) Lbar(k,k) = W_kk^T Q W_kk for
//    the diagonal tile's own adjoint.
#include "common.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NTHREADS = 256;
constexpr int AB = 64, AKC = 16;

template <typename T>
__global__ __launch_bounds__(NTHREADS) void ltl_lower_kernel(
    plasma_desc_t L, int upper, plasma_desc_t Lbar, plasma_desc_t Q,
    int bpt, const int *status)
{
    typedef GemmNT<T, AB, AB, AKC, 32, 32> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int per_tile = bpt * bpt;
    const int sub_ = blockIdx.x % per_tile;
    const int rb = sub_ / bpt, cb = sub_ % bpt;
    int I, J;
    trapezoid_decode(blockIdx.x / per_tile, L.mt, 0, &I, &J);  // I >= J
    if (I == J && cb > rb) return;
    const int iv = plasma_tile_mview(L, I), jv = plasma_tile_mview(L, J);
    const int r0 = rb * AB, c0 = cb * AB;
    if (r0 >= iv || c0 >= jv) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave / 2, wc = wave % 2;
    const int t = threadIdx.x;
    const int cpt = (L.mb + AKC - 1) / AKC;
    const int nchunks = (L.mt - I) * cpt;
    constexpr int W = G::W;

    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    T rp[G::LX], rq[G::LY];
    bool pu = false, qu = false;

    // chunk c: P(r, k) = L(K,I)(k0 + k, r0 + r),
    //          Qo(c, k) = Lbar(K,J)(k0 + k, c0 + c)
    auto load = [&](int ch) {
        const int K = I + ch / cpt, k0 = (ch % cpt) * AKC;
        const int kv = plasma_tile_mview(L, K);
        const View<T> Lv = lower_view<T>(L, upper != 0, K, I);
        View<T> P;                                  // transposed view
        P.p = Lv.p + (ptrdiff_t)r0 * Lv.cs;
        P.rs = Lv.cs;
        P.cs = Lv.rs;
        View<T> Qo;
        if (!upper) {
            const View<T> B = tile_view<T>(Lbar, K, J);
            Qo.p = B.p + (ptrdiff_t)c0 * B.cs;
            Qo.rs = B.cs;
            Qo.cs = 1;
        }
        else {                                      // Lbar(K,J) = Ubar(J,K)^T
            const View<T> B = tile_view<T>(Lbar, J, K);
            Qo.p = B.p + c0;
            Qo.rs = 1;
            Qo.cs = B.cs;
        }
        if (K != I && K != J) {
            G::template load<AB>(rp, P, iv - r0, k0, kv);
            G::template load<AB>(rq, Qo, jv - c0, k0, kv);
            pu = G::rowmap(P);
            qu = G::rowmap(Qo);
            return;
        }
        // diagonal tiles: L(I,I)(k, r) = 0 for k < r, Lbar(J,J)(k, c) = 0
        // for k < c; element loads in the row mapping of GemmNT::load
#pragma unroll
        for (int i = 0; i < G::LX / W; i++) {
            const int v = i * NTHREADS + t;
            const int kk = v / (AB / W), r = (v % (AB / W)) * W;
            const int kg = k0 + kk;
#pragma unroll
            for (int u = 0; u < W; u++) {
                const int rr = r0 + r + u;
                T x = T(0);
                if (r + u < iv - r0 && kg < kv && (K != I || kg >= rr))
                    x = P.at(r + u, kg);
                rp[i * W + u] = x;
                const int cc = c0 + r + u;
                T y = T(0);
                if (r + u < jv - c0 && kg < kv && (K != J || kg >= cc))
                    y = Qo.at(r + u, kg);
                rq[i * W + u] = y;
            }
        }
        pu = qu = true;
    };
    auto store = [&](int buf) {
        G::template store<AB>(rp, sm.x[buf], pu);
        G::template store<AB>(rq, sm.y[buf], qu);
    };

    load(0);
    store(0);
    __syncthreads();
    for (int ch = 0; ch < nchunks; ch++) {
        const int cur = ch & 1;
        if (ch + 1 < nchunks) load(ch + 1);
#pragma unroll
        for (int kk = 0; kk < AKC; kk += 4) {
            T a[G::FM], b[G::FN];
#pragma unroll
            for (int i = 0; i < G::FM; i++)
                a[i] = sm.x[cur][kk + lane / 16][wr * 32 + i * 16 + lane % 16];
#pragma unroll
            for (int j = 0; j < G::FN; j++)
                b[j] = sm.y[cur][kk + lane / 16][wc * 32 + j * 16 + lane % 16];
#pragma unroll
            for (int i = 0; i < G::FM; i++)
#pragma unroll
                for (int j = 0; j < G::FN; j++)
                    acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
        }
        if (ch + 1 < nchunks) store(cur ^ 1);
        __syncthreads();
    }

    const View<T> Qij = tile_view<T>(Q, I, J);
    const View<T> Qji = tile_view<T>(Q, J, I);
#pragma unroll
    for (int i = 0; i < G::FM; i++)
#pragma unroll
        for (int j = 0; j < G::FN; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = r0 + G::acc_row(i, q), c = c0 + G::acc_col(j);
                if (r >= iv || c >= jv) continue;
                if (I == J && c > r) continue;
                const T v = acc[i][j][q];
                Qij.at(r, c) = v;
                if (I != J || c != r) Qji.at(c, r) = v;
            }
}

// Z(I,J) = a X(I,J) + b Y(J,I)^T, one 32 x 32 block per workgroup step.
// With mirror, element (i, j) (global indices) is a X(i, j) for i >= j and
// b Y(j, i) for i < j: the symmetric matrix whose lower triangle is a X
// (a = b, X = Y), reading only the lower triangle.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void tile_axpby_t_kernel(
    T a, plasma_desc_t X, T b, plasma_desc_t Y, plasma_desc_t Z, int mirror,
    const int *status)
{
    __shared__ T sb[32][33];
    if (block_should_stop(status)) return;
    const int bpt = (Z.mb + 31) / 32;
    const int I = blockIdx.x, J = blockIdx.y;
    const int mv = plasma_tile_mview(Z, I), nv = plasma_tile_nview(Z, J);
    const View<T> Zv = tile_view<T>(Z, I, J);
    const View<T> Yv = tile_view<T>(Y, J, I);      // (c, r) element order
    const View<T> Xv = tile_view<T>(X, I, J);
    const bool need_y = !mirror || I <= J;         // some element from Y^T
    const bool need_x = a != T(0) && (!mirror || I >= J);
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;   // 32 x 8
    for (int blk = 0; blk < bpt * bpt; blk++) {
        const int r0 = (blk / bpt) * 32, c0 = (blk % bpt) * 32;
        if (r0 >= mv || c0 >= nv) continue;
        if (need_y) {
            // Y(J,I) rows c0.., columns r0..: read along its columns
            for (int cc = ty; cc < 32; cc += 8) {
                const int yr = c0 + tx, yc = r0 + cc;
                sb[cc][tx] = (yr < nv && yc < mv) ? Yv.at(yr, yc) : T(0);
            }
        }
        __syncthreads();
        for (int cc = ty; cc < 32; cc += 8) {
            const int r = r0 + tx, c = c0 + cc;
            if (r < mv && c < nv) {
                T v;
                if (mirror) {
                    const bool low = (I > J) || (I == J && r >= c);
                    v = low ? a * Xv.at(r, c) : b * sb[tx][cc];
                }
                else {
                    v = b * sb[tx][cc];
                    if (need_x) v += a * Xv.at(r, c);
                }
                Zv.at(r, c) = v;
            }
        }
        __syncthreads();
    }
}

// C(i, j) -= s * sum_r (U(i, r) V(j, r) + V(i, r) U(j, r))
template <typename T>
__global__ __launch_bounds__(NTHREADS) void syr2k_lapack_kernel(
    int n, int k, T s, const T *U, int ldu, const T *V, int ldv, T *C,
    int ldc)
{
    const int i = blockIdx.x * 64 + threadIdx.x % 64;
    const int j = blockIdx.y * 4 + threadIdx.x / 64;
    if (i >= n || j >= n) return;
    T acc = T(0);
    for (int r = 0; r < k; r++)
        acc += U[i + (size_t)r * ldu] * V[j + (size_t)r * ldv] +
               V[i + (size_t)r * ldv] * U[j + (size_t)r * ldu];
    C[i + (size_t)j * ldc] -= s * acc;
}

// part[z](i, j) = sum over tiles t of split z of P(t)^T L(k+1+t, k), 64 x 64
// blocks of the nb x nb result (block (bi, bj) = blockIdx.x).
template <typename T>
__global__ __launch_bounds__(NTHREADS) void gram_partial_kernel(
    plasma_desc_t L, int k, int ntiles, int nsplit, const T *P, T *part,
    const int *status)
{
    typedef GemmNT<T, AB, AB, AKC, 32, 32> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int kv = plasma_tile_nview(L, k);
    const int bpt = (L.nb + AB - 1) / AB;
    const int i0 = (blockIdx.x / bpt) * AB, j0 = (blockIdx.x % bpt) * AB;
    const int z = blockIdx.y;
    const int t0 = (int)((long)ntiles * z / nsplit);
    const int t1 = (int)((long)ntiles * (z + 1) / nsplit);
    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    if (i0 < kv && j0 < kv) {
        for (int t = t0; t < t1; t++) {
            const int mv = plasma_tile_mview(L, k + 1 + t);
            View<T> X;                  // X(i, kk) = P(t)(kk, i0 + i)
            X.p = const_cast<T *>(P) + (size_t)t * L.mb * L.nb +
                  (size_t)i0 * L.mb;
            X.rs = L.mb;
            X.cs = 1;
            const View<T> Lt = lower_view<T>(L, false, k + 1 + t, k);
            View<T> Y;                  // Y(j, kk) = L(k+1+t, k)(kk, j0 + j)
            Y.p = Lt.p + (ptrdiff_t)j0 * Lt.cs;
            Y.rs = Lt.cs;
            Y.cs = Lt.rs;
            G::run(acc, X, kv - i0, Y, kv - j0, mv, sm);
        }
    }
    T *pz = part + (size_t)z * L.nb * L.nb;     // column major, ld nb
#pragma unroll
    for (int i = 0; i < G::FM; i++)
#pragma unroll
        for (int j = 0; j < G::FN; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = i0 + G::acc_row(i, q), c = j0 + G::acc_col(j);
                if (r < kv && c < kv) pz[(size_t)c * L.nb + r] = acc[i][j][q];
            }
}

// Lbar(k,k) -= sum_z part[z]
template <typename T>
__global__ __launch_bounds__(NTHREADS) void gram_reduce_kernel(
    plasma_desc_t Lb, int k, int nsplit, const T *part, const int *status)
{
    if (block_should_stop(status)) return;
    const int kv = plasma_tile_nview(Lb, k);
    const int e = blockIdx.x * NTHREADS + threadIdx.x;
    if (e >= kv * kv) return;
    const int r = e % kv, c = e / kv;
    T sum = T(0);
    for (int z = 0; z < nsplit; z++)
        sum += part[(size_t)z * Lb.nb * Lb.nb + (size_t)c * Lb.nb + r];
    tile_view<T>(Lb, k, k).at(r, c) -= sum;
}

// Lbar(k+1+t, k) = P(t)
template <typename T>
__global__ __launch_bounds__(NTHREADS) void copy_panel_kernel(
    plasma_desc_t Lb, int k, const T *P, const int *status)
{
    if (block_should_stop(status)) return;
    const int t = blockIdx.y;
    const int mv = plasma_tile_mview(Lb, k + 1 + t);
    const int kv = plasma_tile_nview(Lb, k);
    const View<T> D = tile_view<T>(Lb, k + 1 + t, k);
    const T *S = P + (size_t)t * Lb.mb * Lb.nb;
    for (int e = blockIdx.x * NTHREADS + threadIdx.x; e < mv * kv;
         e += gridDim.x * NTHREADS) {
        const int r = e % mv, c = e / mv;
        D.at(r, c) = S[(size_t)c * Lb.mb + r];
    }
}

// One 64 x 64 block (i0, j0) of C = X W: C(i, j) = sum_kk X(i, kk) W(kk, j)
// over kk < kv, for i < m rows and j < kv columns.
template <typename T, typename G>
__device__ __forceinline__ void xw_block(View<T> X, const View<T> &C,
                                         const T *W, int ldw, int m, int kv,
                                         int i0, int j0,
                                         typename G::Smem &sm)
{
    X.p += (ptrdiff_t)i0 * X.rs;
    View<T> Y;                          // Y(j, kk) = W(kk, j0 + j)
    Y.p = const_cast<T *>(W) + (size_t)j0 * ldw;
    Y.rs = ldw;
    Y.cs = 1;
    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    G::run(acc, X, m - i0, Y, kv - j0, kv, sm);
#pragma unroll
    for (int i = 0; i < G::FM; i++)
#pragma unroll
        for (int j = 0; j < G::FN; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = i0 + G::acc_row(i, q), c = j0 + G::acc_col(j);
                if (r < m && c < kv) C.at(r, c) = acc[i][j][q];
            }
}

// C = X W^T-view on one kv x kv tile, 64 x 64 blocks (blockIdx.x): C(i, j)
// = sum_kk X(i, kk) W(kk, j), X given by its strides (Q, or T1 read
// transposed), W the inverse of a diagonal tile (column major, ld ldw).
template <typename T>
__global__ __launch_bounds__(NTHREADS) void tile_xw_kernel(
    View<T> X, View<T> C, const T *W, int ldw, int kv, const int *status)
{
    typedef GemmNT<T, AB, AB, AKC, 32, 32> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int bpt = (kv + AB - 1) / AB;
    const int i0 = (blockIdx.x / bpt) * AB, j0 = (blockIdx.x % bpt) * AB;
    xw_block<T, G>(X, C, W, ldw, kv, kv, i0, j0, sm);
}

// P(t) = Lbar(k+1+t, k) W_kk for the panel below tile k, 64 x 64 blocks:
// blockIdx.x the block of the tile, blockIdx.y the tile t.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void panel_w_kernel(
    plasma_desc_t Lb, int k, const T *W, T *P, const int *status)
{
    typedef GemmNT<T, AB, AB, AKC, 32, 32> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int t = blockIdx.y;
    const int mv = plasma_tile_mview(Lb, k + 1 + t);
    const int kv = plasma_tile_nview(Lb, k);
    const int bpc = (Lb.nb + AB - 1) / AB;      // column blocks per tile
    const int i0 = (blockIdx.x / bpc) * AB, j0 = (blockIdx.x % bpc) * AB;
    if (i0 >= mv || j0 >= kv) return;
    const View<T> X = tile_view<T>(Lb, k + 1 + t, k);
    View<T> C;
    C.p = P + (size_t)t * Lb.mb * Lb.nb;
    C.rs = 1;
    C.cs = Lb.mb;
    xw_block<T, G>(X, C, W, Lb.mb, mv, kv, i0, j0, sm);
}


// Lb(k,k) = W^T Q W = (Q W)^T W (Q symmetric): T1 = Q W, then T1^T W.
template <typename T>
int launch_adjoint_diag_sandwich(plasma_desc_t Lb, int k, const T *W,
                                 const T *Q, T *T1, const int *status,
                                 hipStream_t s)
{
    const int kv = plasma_tile_nview(Lb, k);
    const int bpt = (kv + AB - 1) / AB;
    const int ld = Lb.mb;
    View<T> X, C;
    X.p = const_cast<T *>(Q);           // X(i, kk) = Q(i, kk)
    X.rs = 1;
    X.cs = ld;
    C.p = T1;
    C.rs = 1;
    C.cs = ld;
    hipLaunchKernelGGL(tile_xw_kernel<T>, dim3(bpt * bpt), dim3(NTHREADS), 0,
                       s, X, C, W, ld, kv, status);
    X.p = T1;                           // X(i, kk) = T1(kk, i)
    X.rs = ld;
    X.cs = 1;
    C.p = (T *)plasma_tile_addr(Lb, k, k);
    C.rs = 1;
    C.cs = plasma_tile_mmain(Lb, k);
    hipLaunchKernelGGL(tile_xw_kernel<T>, dim3(bpt * bpt), dim3(NTHREADS), 0,
                       s, X, C, W, ld, kv, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T>
int launch_adjoint_panel_step(plasma_desc_t L, plasma_desc_t Lb, int k,
                              const T *W, T *P, T *part, int maxsplit,
                              const int *status, hipStream_t s)
{
    const int ntiles = L.mt - k - 1;
    if (ntiles <= 0) return 0;
    // P = Lbar(k+1.., k) W_kk
    const int bpc = (L.nb + AB - 1) / AB, bpr = (L.mb + AB - 1) / AB;
    hipLaunchKernelGGL(panel_w_kernel<T>, dim3(bpr * bpc, ntiles),
                       dim3(NTHREADS), 0, s, Lb, k, W, P, status);
    // Lbar(k,k) -= sum_t P(t)^T L(k+1+t, k)
    const int gb = (L.nb + AB - 1) / AB;
    const int nsplit = max(1, min(ntiles, maxsplit));
    hipLaunchKernelGGL(gram_partial_kernel<T>, dim3(gb * gb, nsplit),
                       dim3(NTHREADS), 0, s, L, k, ntiles, nsplit,
                       (const T *)P, part, status);
    const int kv = plasma_tile_nview(L, k);
    hipLaunchKernelGGL(gram_reduce_kernel<T>,
                       dim3((kv * kv + NTHREADS - 1) / NTHREADS),
                       dim3(NTHREADS), 0, s, Lb, k, nsplit,
                       (const T *)part, status);
    // Lbar(k+1.., k) = P
    hipLaunchKernelGGL(copy_panel_kernel<T>, dim3(8, ntiles), dim3(NTHREADS),
                       0, s, Lb, k, (const T *)P, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

}  // namespace

extern "C" int plasma_kern_adjoint_panel_step(plasma_desc_t L,
                                              plasma_desc_t Lb, int k,
                                              const void *W, void *P,
                                              void *part, int maxsplit,
                                              const int *status,
                                              hipStream_t s)
{
    if (L.precision != Lb.precision) return PlasmaErrorIllegalValue;
    if (L.precision == PlasmaRealDouble)
        return launch_adjoint_panel_step<double>(
            L, Lb, k, (const double *)W, (double *)P, (double *)part,
            maxsplit, status, s);
    if (L.precision == PlasmaRealFloat)
        return launch_adjoint_panel_step<float>(
            L, Lb, k, (const float *)W, (float *)P, (float *)part, maxsplit,
            status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_adjoint_diag_sandwich(plasma_desc_t Lb, int k,
                                                 const void *W, const void *Q,
                                                 void *T1, const int *status,
                                                 hipStream_t s)
{
    if (Lb.precision == PlasmaRealDouble)
        return launch_adjoint_diag_sandwich<double>(
            Lb, k, (const double *)W, (const double *)Q, (double *)T1,
            status, s);
    if (Lb.precision == PlasmaRealFloat)
        return launch_adjoint_diag_sandwich<float>(
            Lb, k, (const float *)W, (const float *)Q, (float *)T1, status,
            s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_ltl_lower(plasma_enum_t uplo, plasma_desc_t L,
                                     plasma_desc_t Lbar, plasma_desc_t Q,
                                     const int *status, hipStream_t s)
{
    const int bpt = (L.mb + AB - 1) / AB;
    const long ntiles = (long)L.mt * (L.mt + 1) / 2;
    const long nblocks = ntiles * bpt * bpt;
    if (nblocks > 0x7fffffffL) return PlasmaErrorNotSupported;
    if (L.precision == PlasmaRealDouble)
        hipLaunchKernelGGL(ltl_lower_kernel<double>, dim3((unsigned)nblocks),
                           dim3(NTHREADS), 0, s, L, uplo == PlasmaUpper, Lbar,
                           Q, bpt, status);
    else if (L.precision == PlasmaRealFloat)
        hipLaunchKernelGGL(ltl_lower_kernel<float>, dim3((unsigned)nblocks),
                           dim3(NTHREADS), 0, s, L, uplo == PlasmaUpper, Lbar,
                           Q, bpt, status);
    else
        return PlasmaErrorNotSupported;
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_tile_axpby_t(double a, plasma_desc_t X, double b,
                                        plasma_desc_t Y, plasma_desc_t Z,
                                        int mirror, const int *status,
                                        hipStream_t s)
{
    dim3 grid(Z.mt, Z.nt);
    if (Z.precision == PlasmaRealDouble)
        hipLaunchKernelGGL(tile_axpby_t_kernel<double>, grid, dim3(NTHREADS),
                           0, s, a, X, b, Y, Z, mirror, status);
    else if (Z.precision == PlasmaRealFloat)
        hipLaunchKernelGGL(tile_axpby_t_kernel<float>, grid, dim3(NTHREADS), 0,
                           s, (float)a, X, (float)b, Y, Z, mirror, status);
    else
        return PlasmaErrorNotSupported;
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_syr2k_lapack(plasma_enum_t precision, int n, int k,
                                        double scale, const void *U, int ldu,
                                        const void *V, int ldv, void *C,
                                        int ldc, hipStream_t s)
{
    if (n == 0) return 0;
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    if (precision == PlasmaRealDouble)
        hipLaunchKernelGGL(syr2k_lapack_kernel<double>, grid, dim3(NTHREADS),
                           0, s, n, k, scale, (const double *)U, ldu,
                           (const double *)V, ldv, (double *)C, ldc);
    else if (precision == PlasmaRealFloat)
        hipLaunchKernelGGL(syr2k_lapack_kernel<float>, grid, dim3(NTHREADS), 0,
                           s, n, k, (float)scale, (const float *)U, ldu,
                           (const float *)V, ldv, (float *)C, ldc);
    else
        return PlasmaErrorNotSupported;
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}
