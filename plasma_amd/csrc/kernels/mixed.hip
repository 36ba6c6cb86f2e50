// mixed.hip -- FP64 pieces of the mixed-precision solver for gfx950:
// ||A||_inf of a symmetric matrix held in one triangle, column max-norms,
// the residual R <- R - A X on MFMA, and the convergence test.
//
// Reference: compute/pzlanhe.c:83-131 (Inf norm through per-tile row / column
// sums), compute/pdzamax.c:36-52 with core_dzamax.c:20-61 (column max |x|),
// compute/pzhemm.c:44-80 (R -= A X with alpha = -1, beta = 1, A symmetric,
// one triangle), compute/zcposv.c:275-287 and :476 (the test).
#include "common.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NTHREADS = 256;

__global__ void fill_double_kernel(double *p, double v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// One workgroup per stored tile (m, n), m >= n of the lower-equivalent
// factor storage: |A| row sums of tile rows go to work[rows of m]; for an
// off-diagonal tile its column sums are the row sums of the mirrored tile
// (rows of n).  Diagonal tiles are read symmetrically from their lower part.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void lansy_rowsum_kernel(
    plasma_desc_t A, int upper, double *work, const int *status)
{
    if (block_should_stop(status)) return;
    int m, n;
    trapezoid_decode(blockIdx.x, A.mt, 0, &m, &n);
    const int mv = plasma_tile_mview(A, m), nv = plasma_tile_mview(A, n);
    const View<T> L = lower_view<T>(A, upper != 0, m, n);
    const int t = threadIdx.x;
    // row sums (one thread per row)
    for (int r = t; r < mv; r += NTHREADS) {
        double s = 0;
        for (int c = 0; c < nv; c++) {
            if (m == n && c > r) s += fabs((double)L.at(c, r));
            else                 s += fabs((double)L.at(r, c));
        }
        atomicAdd(&work[m * A.mb + r], s);
    }
    if (m != n) {
        for (int c = t; c < nv; c += NTHREADS) {
            double s = 0;
            for (int r = 0; r < mv; r++) s += fabs((double)L.at(r, c));
            atomicAdd(&work[n * A.mb + c], s);
        }
    }
}

// value = max_i work[i]; a NaN row sum propagates (LAPACK's disnan rule)
__global__ __launch_bounds__(NTHREADS) void max_kernel(const double *work,
                                                       int n, double *value,
                                                       const int *status)
{
    if (block_should_stop(status)) return;
    __shared__ double red[NTHREADS];
    double v = 0;
    for (int i = threadIdx.x; i < n; i += NTHREADS) {
        double w = work[i];
        if (w > v || w != w) v = w;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = NTHREADS / 2; s > 0; s /= 2) {
        if (threadIdx.x < s) {
            double o = red[threadIdx.x + s];
            if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *value = red[0];
}

// values[j] = max_i |A(i, j)|, one workgroup per column (reference
// core_dzamax.c: a NaN entry does not replace a larger value).
template <typename T>
__global__ __launch_bounds__(NTHREADS) void amax_cols_kernel(
    plasma_desc_t A, double *values, const int *status)
{
    if (block_should_stop(status)) return;
    const int j = blockIdx.x;
    const int tn = j / A.nb, c = j % A.nb;
    __shared__ double red[NTHREADS];
    double v = 0;
    for (int tm = 0; tm < A.mt; tm++) {
        const View<T> V = tile_view<T>(A, tm, tn);
        const int mv = plasma_tile_mview(A, tm);
        for (int r = threadIdx.x; r < mv; r += NTHREADS) {
            double a = fabs((double)V.at(r, c));
            if (a > v) v = a;
        }
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = NTHREADS / 2; s > 0; s /= 2) {
        if (threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x])
            red[threadIdx.x] = red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) values[j] = red[0];
}

// ---------------------------------------------------------------------------
// R(m, tn) -= sum_k Asym(m, k) X(k, tn) with Asym(m,k) = L(m,k) (k < m),
// L(k,m)^T (k > m), and the diagonal tile read symmetrically from its lower
// part.  64 x 64 output blocks, K streamed over every tile column through
// double-buffered LDS, MFMA accumulation in the precision of A.
// ---------------------------------------------------------------------------
constexpr int SY_BM = 64, SY_BN = 64, SY_KC = 16;

template <typename T>
__global__ __launch_bounds__(NTHREADS) void symm_residual_kernel(
    plasma_desc_t A, plasma_desc_t X, plasma_desc_t R, int upper,
    int rchunks, int cchunks, const int *status)
{
    typedef GemmNT<T, SY_BM, SY_BN, SY_KC, 32, 32> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int m = blockIdx.x / rchunks;
    const int r0 = (blockIdx.x % rchunks) * SY_BM;
    const int tn = blockIdx.y / cchunks;
    const int c0 = (blockIdx.y % cchunks) * SY_BN;
    const int mv = plasma_tile_mview(R, m), nv = plasma_tile_nview(R, tn);
    if (r0 >= mv || c0 >= nv) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave / 2, wc = wave % 2;
    const int t = threadIdx.x;
    const int cpt = (A.mb + SY_KC - 1) / SY_KC;   // chunks per tile column
    const int nchunks = A.mt * cpt;

    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    T ra[G::LX], rb[G::LY];
    bool xu = true, yu = false;

    // Operand views of chunk c: X(row, kk) = Asym(m, kt)(r0 + row, k0 + kk),
    // Y(col, kk) = X(kt, tn)(k0 + kk, c0 + col).
    auto load = [&](int c) {
        const int kt = c / cpt, k0 = (c % cpt) * SY_KC;
        const int K = plasma_tile_mview(A, kt);
        View<T> V;
        if (kt < m) {
            V = lower_view<T>(A, upper != 0, m, kt);
        }
        else {
            const View<T> W = lower_view<T>(A, upper != 0, kt, m);
            V.p = W.p; V.rs = W.cs; V.cs = W.rs;  // transposed
        }
        if (kt != m) {
            View<T> Vr = V;
            Vr.p += (ptrdiff_t)r0 * V.rs;
            G::template load<SY_BM>(ra, Vr, mv - r0, k0, K);
            xu = (Vr.rs == 1);
        }
        else {
            // diagonal tile: symmetric read of the stored lower triangle
#pragma unroll
            for (int i = 0; i < G::LX; i++) {
                int e = i * NTHREADS + t;
                int kk = e % SY_KC, r = e / SY_KC;
                int kg = k0 + kk, row = r0 + r;
                T v = T(0);
                if (row < mv && kg < K)
                    v = row >= kg ? V.at(row, kg) : V.at(kg, row);
                ra[i] = v;
            }
            xu = false;
        }
        const View<T> Xk = tile_view<T>(X, kt, tn);
        View<T> Y;
        Y.p = Xk.p + (ptrdiff_t)c0 * Xk.cs;
        Y.rs = Xk.cs;
        Y.cs = 1;
        G::template load<SY_BN>(rb, Y, nv - c0, k0, K);
        yu = (Y.rs == 1);
    };
    auto store = [&](int buf) {
        G::template store<SY_BM>(ra, sm.x[buf], xu);
        G::template store<SY_BN>(rb, sm.y[buf], yu);
    };

    load(0);
    store(0);
    __syncthreads();
    for (int c = 0; c < nchunks; c++) {
        const int cur = c & 1;
        if (c + 1 < nchunks) load(c + 1);
#pragma unroll
        for (int kk = 0; kk < SY_KC; kk += 4) {
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
        if (c + 1 < nchunks) store(cur ^ 1);
        __syncthreads();
    }

    const View<T> C = tile_view<T>(R, m, tn);
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

__global__ void conv_test_kernel(const double *rnorm, const double *xnorm,
                                 int nrhs, const double *anorm,
                                 double eps_sqrtn, int *flag,
                                 const int *status)
{
    if (threadIdx.x != 0) return;
    if (seq_failed(status)) return;
    const double cte = (*anorm) * eps_sqrtn;
    int ok = 1;
    for (int j = 0; j < nrhs; j++)
        if (rnorm[j] > xnorm[j] * cte) { ok = 0; break; }
    *flag = ok;
}

template <typename T>
int launch_symm(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t X,
                plasma_desc_t R, const int *status, hipStream_t s)
{
    const int rchunks = (R.mb + SY_BM - 1) / SY_BM;
    const int cchunks = (R.nb + SY_BN - 1) / SY_BN;
    hipLaunchKernelGGL(symm_residual_kernel<T>,
                       dim3(R.mt * rchunks, R.nt * cchunks), dim3(NTHREADS),
                       0, s, A, X, R, uplo == PlasmaUpper, rchunks, cchunks,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

}  // namespace

extern "C" int plasma_kern_lansy_inf(plasma_enum_t uplo, plasma_desc_t A,
                                     double *work, double *value,
                                     const int *status, hipStream_t s)
{
    const int n = A.mt * A.mb;
    hipLaunchKernelGGL(fill_double_kernel, dim3((n + 255) / 256), dim3(256),
                       0, s, work, 0.0, n);
    const long ntiles = (long)A.mt * (A.mt + 1) / 2;
    if (A.precision == PlasmaRealDouble)
        hipLaunchKernelGGL(lansy_rowsum_kernel<double>, dim3((unsigned)ntiles),
                           dim3(NTHREADS), 0, s, A, uplo == PlasmaUpper, work,
                           status);
    else if (A.precision == PlasmaRealFloat)
        hipLaunchKernelGGL(lansy_rowsum_kernel<float>, dim3((unsigned)ntiles),
                           dim3(NTHREADS), 0, s, A, uplo == PlasmaUpper, work,
                           status);
    else
        return PlasmaErrorNotSupported;
    hipLaunchKernelGGL(max_kernel, dim3(1), dim3(NTHREADS), 0, s, work, n,
                       value, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_amax_cols(plasma_desc_t A, double *values,
                                     const int *status, hipStream_t s)
{
    if (A.n == 0) return 0;
    if (A.precision == PlasmaRealDouble)
        hipLaunchKernelGGL(amax_cols_kernel<double>, dim3(A.n), dim3(NTHREADS),
                           0, s, A, values, status);
    else if (A.precision == PlasmaRealFloat)
        hipLaunchKernelGGL(amax_cols_kernel<float>, dim3(A.n), dim3(NTHREADS),
                           0, s, A, values, status);
    else
        return PlasmaErrorNotSupported;
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_symm_residual(plasma_enum_t uplo, plasma_desc_t A,
                                         plasma_desc_t X, plasma_desc_t R,
                                         const int *status, hipStream_t s)
{
    if (A.precision != X.precision || A.precision != R.precision)
        return PlasmaErrorIllegalValue;
    if (A.precision == PlasmaRealDouble)
        return launch_symm<double>(uplo, A, X, R, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_symm<float>(uplo, A, X, R, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_conv_test(const double *rnorm, const double *xnorm,
                                     int nrhs, const double *anorm,
                                     double eps_sqrtn, int *flag,
                                     const int *status, hipStream_t s)
{
    hipLaunchKernelGGL(conv_test_kernel, dim3(1), dim3(64), 0, s, rnorm, xnorm,
                       nrhs, anorm, eps_sqrtn, flag, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}
