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

// Sum of |V(r, c)| over c < cols for each r < rows into out[r].  With a unit
// row stride one thread walks each row (consecutive threads read consecutive
// rows); otherwise one wave walks each row with its lanes along the row and
// reduces across lanes.  Coalesced either way.
template <typename T>
__device__ __forceinline__ void abs_row_sums(const View<T> &V, int rows,
                                             int cols, double *out)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (V.rs == 1) {
        for (int r = t; r < rows; r += NTHREADS) {
            double s = 0;
            for (int c = 0; c < cols; c++) s += fabs((double)V.at(r, c));
            out[r] = s;
        }
    }
    else {
        for (int r = wave; r < rows; r += NTHREADS / 64) {
            double s = 0;
            for (int c = lane; c < cols; c += 64) s += fabs((double)V.at(r, c));
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) out[r] = s;
        }
    }
}

// One workgroup per stored tile (m, n), m >= n, of the lower-equivalent
// storage.  work is an mt x (mt * mb) array of partial row sums of |A|:
// slot n of the rows of m holds the row sums of L(m,n), and for m > n slot m
// of the rows of n holds its column sums (the row sums of the mirrored tile
// L(m,n)^T).  Every (slot, row) is written by exactly one tile, so no
// atomics are needed and the result does not depend on scheduling.
// Diagonal tiles are read symmetrically from their stored triangle.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void lansy_tile_kernel(
    plasma_desc_t A, int upper, double *work, const int *status)
{
    if (block_should_stop(status)) return;
    int m, n;
    trapezoid_decode(blockIdx.x, A.mt, 0, &m, &n);
    const int mv = plasma_tile_mview(A, m), nv = plasma_tile_mview(A, n);
    const View<T> L = lower_view<T>(A, upper != 0, m, n);
    const size_t ld = (size_t)A.mt * A.mb;
    if (m == n) {
        for (int r = threadIdx.x; r < mv; r += NTHREADS) {
            double s = 0;
            for (int c = 0; c < nv; c++)
                s += fabs((double)(c <= r ? L.at(r, c) : L.at(c, r)));
            work[n * ld + (size_t)m * A.mb + r] = s;
        }
        return;
    }
    if (L.rs == 1 && mv <= 256) {
        // One pass over a column-major tile: wave w takes the columns
        // c = w, w + 4, ...; lane l the rows l, l + 64, ... (consecutive
        // lanes read consecutive rows).  A column's sum is a wave
        // reduction; a row's sum adds the four waves' partial sums in a
        // fixed order.
        __shared__ double part[NTHREADS / 64][256];
        const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
        double rsum[4] = {0, 0, 0, 0};
        double *csum = work + m * ld + (size_t)n * A.mb;
#pragma unroll 2
        for (int c = wave; c < nv; c += NTHREADS / 64) {
            const T *col = L.p + (ptrdiff_t)c * L.cs;
            double s = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int r = lane + 64 * u;
                const double x = r < mv ? fabs((double)col[r]) : 0.0;
                rsum[u] += x;
                s += x;
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) csum[c] = s;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) part[wave][lane + 64 * u] = rsum[u];
        __syncthreads();
        double *rout = work + n * ld + (size_t)m * A.mb;
        for (int r = t; r < mv; r += NTHREADS) {
            double s = part[0][r];
#pragma unroll
            for (int w = 1; w < NTHREADS / 64; w++) s += part[w][r];
            rout[r] = s;
        }
        return;
    }
    abs_row_sums<T>(L, mv, nv, work + n * ld + (size_t)m * A.mb);
    View<T> Lt;
    Lt.p = L.p;
    Lt.rs = L.cs;
    Lt.cs = L.rs;
    abs_row_sums<T>(Lt, nv, mv, work + m * ld + (size_t)n * A.mb);
}

// partial[b] = max over the rows i of block b of sum_slot work[slot][i];
// a NaN row sum propagates (LAPACK's disnan rule).
__global__ __launch_bounds__(NTHREADS) void lansy_reduce_kernel(
    const double *work, int mt, size_t ld, int n, double *partial,
    const int *status)
{
    if (block_should_stop(status)) return;
    __shared__ double red[NTHREADS];
    const int i = blockIdx.x * NTHREADS + threadIdx.x;
    double v = 0;
    if (i < n) {
        for (int slot = 0; slot < mt; slot++) v += work[slot * ld + i];
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
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// value = max_i work[i]; a NaN propagates (LAPACK's disnan rule)
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
    // rows of the column as one index i = tm mb + r: every thread takes
    // rows i = t, t + NTHREADS, ..., four passes at a time so that their
    // loads are in flight together
    const int rows = A.mt * A.mb;
    for (int i0 = threadIdx.x; i0 < rows; i0 += 4 * NTHREADS) {
        double a[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * NTHREADS;
            const int tm = i / A.mb, r = i - tm * A.mb;
            a[u] = 0;
            if (i < rows && r < plasma_tile_mview(A, tm))
                a[u] = fabs((double)tile_view<T>(A, tm, tn).at(r, c));
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (a[u] > v) v = a[u];
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
// L(k,m)^T (k > m), and the diagonal tile read symmetrically from its
// stored triangle.  Output blocks of 64 right-hand sides x 64 (FP64) or
// 128 (FP32) rows; K runs over every tile column in
// 16-deep (FP64) or 32-deep (FP32) chunks through double-buffered LDS with
// 16-byte loads (with the padded rows 33 KB of LDS in FP64, 51 KB in FP32:
// four or three workgroups share a CU and hide each other's loads); MFMA
// accumulation in the precision of A.  The product is formed transposed
// (X^T Asym^T) so the accumulator's consecutive lanes are consecutive rows
// of R.
// ---------------------------------------------------------------------------
constexpr int SY_BN = 64;
constexpr int SY_WM = 32;               // per wave: right-hand sides

template <typename T, int SY_KC, int SY_BM>
__global__ __launch_bounds__(NTHREADS) void symm_residual_kernel(
    plasma_desc_t A, plasma_desc_t X, plasma_desc_t R, int upper,
    int rchunks, int cchunks, int ksplit, T *part, const int *status)
{
    // first operand: 64 right-hand sides; second: SY_BM rows of Asym
    constexpr int SY_WN = SY_BM / 2;    // per wave: rows
    typedef GemmNT<T, SY_BN, SY_BM, SY_KC, SY_WM, SY_WN> G;
    __shared__ typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int m = blockIdx.x / rchunks;
    const int r0 = (blockIdx.x % rchunks) * SY_BM;
    const int tn = blockIdx.y / cchunks;
    const int c0 = (blockIdx.y % cchunks) * SY_BN;
    const int mv = plasma_tile_mview(R, m), nv = plasma_tile_nview(R, tn);
    if (r0 >= mv || c0 >= nv) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave / (SY_BM / SY_WN), wc = wave % (SY_BM / SY_WN);
    const int t = threadIdx.x;
    const int cpt = (A.mb + SY_KC - 1) / SY_KC;   // chunks per tile column
    const int nchunks = A.mt * cpt;
    // split-K: workgroup z sums chunks [cbeg, cend) into its own partial
    const int cbeg = (int)((long)nchunks * blockIdx.z / ksplit);
    const int cend = (int)((long)nchunks * (blockIdx.z + 1) / ksplit);
    constexpr int W = G::W;

    typename G::acc_t acc[G::FM][G::FN];
    G::zero(acc);
    T rp[G::LX], rq[G::LY];
    bool pu = false, qu = true;

    // Chunk c: P(col, kk) = X(kt, tn)(k0 + kk, c0 + col),
    //          Q(row, kk) = Asym(m, kt)(r0 + row, k0 + kk).
    auto load = [&](int c) {
        const int kt = c / cpt, k0 = (c % cpt) * SY_KC;
        const int K = plasma_tile_mview(A, kt);
        const View<T> Xk = tile_view<T>(X, kt, tn);
        View<T> P;
        P.p = Xk.p + (ptrdiff_t)c0 * Xk.cs;
        P.rs = Xk.cs;
        P.cs = 1;
        G::template load<SY_BN>(rp, P, nv - c0, k0, K);
        pu = G::rowmap(P);
        if (kt != m) {
            View<T> Q;
            if (kt < m) {
                Q = lower_view<T>(A, upper != 0, m, kt);
            }
            else {
                const View<T> V = lower_view<T>(A, upper != 0, kt, m);
                Q.p = V.p; Q.rs = V.cs; Q.cs = V.rs;  // transposed
            }
            Q.p += (ptrdiff_t)r0 * Q.rs;
            G::template load<SY_BM>(rq, Q, mv - r0, k0, K);
            qu = G::rowmap(Q);
        }
        else {
            // diagonal tile: symmetric read of the stored triangle, with
            // the row mapping of GemmNT::load
            const View<T> V = lower_view<T>(A, upper != 0, m, m);
#pragma unroll
            for (int i = 0; i < G::LY / W; i++) {
                const int v = i * NTHREADS + t;
                const int kk = v / (SY_BM / W), r = (v % (SY_BM / W)) * W;
                const int kg = k0 + kk;
#pragma unroll
                for (int u = 0; u < W; u++) {
                    const int row = r0 + r + u;
                    T x = T(0);
                    if (row < mv && kg < K)
                        x = row >= kg ? V.at(row, kg) : V.at(kg, row);
                    rq[i * W + u] = x;
                }
            }
            qu = true;
        }
    };
    auto store = [&](int buf) {
        G::template store<SY_BN>(rp, sm.x[buf], pu);
        G::template store<SY_BM>(rq, sm.y[buf], qu);
    };

    if (cbeg < cend) {
        load(cbeg);
        store(0);
    }
    __syncthreads();
    for (int c = cbeg; c < cend; c++) {
        const int cur = (c - cbeg) & 1;
        if (c + 1 < cend) load(c + 1);
#pragma unroll
        for (int kk = 0; kk < SY_KC; kk += 4) {
            T a[G::FM], b[G::FN];
#pragma unroll
            for (int i = 0; i < G::FM; i++)
                a[i] = sm.x[cur][kk + lane / 16][wr * SY_WM + i * 16 + lane % 16];
#pragma unroll
            for (int j = 0; j < G::FN; j++)
                b[j] = sm.y[cur][kk + lane / 16][wc * SY_WN + j * 16 + lane % 16];
#pragma unroll
            for (int i = 0; i < G::FM; i++)
#pragma unroll
                for (int j = 0; j < G::FN; j++)
                    acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
        }
        if (c + 1 < cend) store(cur ^ 1);
        __syncthreads();
    }

    if (ksplit > 1) {
        // partial z of element (global row, global column) at
        // part[(z * R.n + col) * R.m + row]
        T *pz = part + (size_t)blockIdx.z * R.m * R.n;
#pragma unroll
        for (int i = 0; i < G::FM; i++)
#pragma unroll
            for (int j = 0; j < G::FN; j++) {
                const int rr = r0 + G::acc_col(j);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int cc = c0 + G::acc_row(i, q);
                    if (rr < mv && cc < nv)
                        pz[(size_t)(tn * R.nb + cc) * R.m + m * R.mb + rr] =
                            acc[i][j][q];
                }
            }
        return;
    }
    const View<T> C = tile_view<T>(R, m, tn);
#pragma unroll
    for (int i = 0; i < G::FM; i++)
#pragma unroll
        for (int j = 0; j < G::FN; j++) {
            const int rr = r0 + G::acc_col(j);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int cc = c0 + G::acc_row(i, q);
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

// R -= sum_z part_z, the partials added in a fixed order (deterministic).
template <typename T>
__global__ __launch_bounds__(NTHREADS) void symm_reduce_kernel(
    plasma_desc_t R, int ksplit, const T *part, const int *status)
{
    if (block_should_stop(status)) return;
    const size_t e = (size_t)blockIdx.x * NTHREADS + threadIdx.x;
    const size_t total = (size_t)R.m * R.n;
    if (e >= total) return;
    const int row = (int)(e % R.m), col = (int)(e / R.m);
    T sum = T(0);
    for (int z = 0; z < ksplit; z++) sum += part[(size_t)z * total + e];
    const View<T> C = tile_view<T>(R, row / R.mb, col / R.nb);
    C.at(row % R.mb, col % R.nb) -= sum;
}

// The grid has one workgroup per block of R (column blocks only
// over the right-hand sides present: min(nb, n) columns per tile column);
// when that is small (few right-hand sides) and work has room, K is split
// so that about SY_TARGET workgroups share the chip.
constexpr int SY_TARGET = 4096, SY_MAXSPLIT = 8;

template <typename T, int SY_BM>
int launch_symm_rb(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t X,
                   plasma_desc_t R, void *work, size_t work_bytes,
                   const int *status, hipStream_t s)
{
    const int rchunks = (R.mb + SY_BM - 1) / SY_BM;
    const int cchunks = (min(R.nb, R.n) + SY_BN - 1) / SY_BN;
    const long blocks = (long)R.mt * rchunks * R.nt * cchunks;
    const size_t part_bytes = (size_t)R.m * R.n * sizeof(T);
    int ksplit = (int)((SY_TARGET + blocks - 1) / blocks);
    ksplit = min(ksplit, SY_MAXSPLIT);
    if (work == nullptr || R.i % R.mb != 0 || R.j % R.nb != 0) ksplit = 1;
    else ksplit = (int)min((size_t)ksplit, work_bytes / part_bytes);
    if (ksplit < 1) ksplit = 1;
    // chunk depth: 32 in FP32 (half the MFMA work per chunk of FP64, so
    // a deeper chunk hides the load and barrier), 16 in FP64.  Measured on
    // the Cholesky adjoint's 16384 x 16384 symmetric product: FP32 63 ms
    // (16) -> 56 ms (32) -> 67 ms (64); FP64 99 ms (16), 102 ms (32),
    // 121 ms (64).
    constexpr int KC = sizeof(T) == 4 ? 32 : 16;
    hipLaunchKernelGGL((symm_residual_kernel<T, KC, SY_BM>),
                       dim3(R.mt * rchunks, R.nt * cchunks, ksplit),
                       dim3(NTHREADS), 0, s, A, X, R, uplo == PlasmaUpper,
                       rchunks, cchunks, ksplit, (T *)work, status);
    if (ksplit > 1) {
        const size_t total = (size_t)R.m * R.n;
        hipLaunchKernelGGL(symm_reduce_kernel<T>,
                           dim3((unsigned)((total + NTHREADS - 1) / NTHREADS)),
                           dim3(NTHREADS), 0, s, R, ksplit,
                           (const T *)work, status);
    }
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

// Rows of R per workgroup: 128 in FP32 (every chunk of X then feeds twice
// the rows), 64 in FP64.  Measured on the Cholesky adjoint's 16384 x 16384
// symmetric product: FP32 56.1 -> 52.6 ms with 128 rows; FP64 98.6 ->
// 101.2 ms; the dsposv residual (n = 32768, 64 right-hand sides) is the
// same with either.
template <typename T>
int launch_symm(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t X,
                plasma_desc_t R, void *work, size_t work_bytes,
                const int *status, hipStream_t s)
{
    constexpr int RB = sizeof(T) == 4 ? 128 : 64;
    return launch_symm_rb<T, RB>(uplo, A, X, R, work, work_bytes, status, s);
}

}  // namespace

extern "C" size_t plasma_lansy_work_doubles(plasma_desc_t A)
{
    const size_t ld = (size_t)A.mt * A.mb;
    return (size_t)A.mt * ld + (ld + NTHREADS - 1) / NTHREADS;
}

extern "C" int plasma_kern_lansy_inf(plasma_enum_t uplo, plasma_desc_t A,
                                     double *work, double *value,
                                     const int *status, hipStream_t s)
{
    const size_t ld = (size_t)A.mt * A.mb;
    const int nblk = (int)((ld + NTHREADS - 1) / NTHREADS);
    double *partial = work + (size_t)A.mt * ld;
    const long ntiles = (long)A.mt * (A.mt + 1) / 2;
    if (plasma_is_complex(A.precision)) {
        const int err = plasma_kern_zlanhe_tiles(uplo, A, work, status, s);
        if (err) return err;
    }
    else if (A.precision == PlasmaRealDouble)
        hipLaunchKernelGGL(lansy_tile_kernel<double>, dim3((unsigned)ntiles),
                           dim3(NTHREADS), 0, s, A, uplo == PlasmaUpper, work,
                           status);
    else if (A.precision == PlasmaRealFloat)
        hipLaunchKernelGGL(lansy_tile_kernel<float>, dim3((unsigned)ntiles),
                           dim3(NTHREADS), 0, s, A, uplo == PlasmaUpper, work,
                           status);
    else
        return PlasmaErrorNotSupported;
    hipLaunchKernelGGL(lansy_reduce_kernel, dim3(nblk), dim3(NTHREADS), 0, s,
                       work, A.mt, ld, A.m, partial, status);
    hipLaunchKernelGGL(max_kernel, dim3(1), dim3(NTHREADS), 0, s, partial,
                       nblk, value, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_amax_cols(plasma_desc_t A, double *values,
                                     const int *status, hipStream_t s)
{
    if (A.n == 0) return 0;
    if (plasma_is_complex(A.precision))
        return plasma_kern_zamax_cols(A, values, status, s);
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
                                         void *work, size_t work_bytes,
                                         const int *status, hipStream_t s)
{
    if (A.precision != X.precision || A.precision != R.precision)
        return PlasmaErrorIllegalValue;
    if (plasma_is_complex(A.precision))
        return plasma_kern_zhemm_residual(uplo, A, X, R, status, s);
    if (A.precision == PlasmaRealDouble)
        return launch_symm<double>(uplo, A, X, R, work, work_bytes, status,
                                   s);
    if (A.precision == PlasmaRealFloat)
        return launch_symm<float>(uplo, A, X, R, work, work_bytes, status,
                                  s);
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
