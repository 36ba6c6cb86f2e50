// norms.hip -- matrix norms and row / column max-norms of tile descriptors
// for gfx950: the device side of plasma_{d,s}lansy, plasma_{d,s}lange and
// plasma_{d,s}amax.
//
// Reference: compute/pzlansy.c:27-172 and compute/pzlange.c (per-tile
// partial results combined in a final step), compute/pdzamax.c:24-53 with
// core_blas/core_dzamax.c.  Every kernel is one pass over the stored tiles
// (HBM-bound); partial results go to slots that exactly one workgroup
// writes, and the final reductions run in a fixed order, so the values do
// not depend on scheduling.
//
//  * Max norm: each tile's max |a| into partial[job]; max over the jobs.
//  * One / Inf norm of a general matrix: each tile's column sums into slot
//    m of the column sums, its row sums into slot n of the row sums; the
//    slots are added per column (row) and the largest sum is the norm.  The
//    symmetric One / Inf norm is the lansy kernel of the mixed solver.
//  * Frobenius norm: each tile's (scale, ssq) pair, sum_ij a_ij^2 =
//    scale^2 ssq with scale = max |a| (LAPACK's dlassq representation: no
//    overflow for large entries); off-diagonal entries of a symmetric
//    matrix count twice; the pairs are combined in a fixed order.
// NaN propagates through the norms (LAPACK's disnan rule); in the
// max-norms of rows and columns a NaN entry does not replace a larger value
// (core_dzamax.c compares with >).
#include "common.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NT = 256;

// v = max(v, a) with NaN propagating
__device__ __forceinline__ double nanmax(double v, double a)
{
    return (a > v || a != a) ? a : v;
}

__device__ __forceinline__ double block_nanmax(double v, double *red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = nanmax(red[threadIdx.x],
                                                       red[threadIdx.x + s]);
        __syncthreads();
    }
    return red[0];
}

// Tile of job j: symmetric storage walks the lower-equivalent trapezoid
// m >= n; a general descriptor walks all tiles, column by column.
__device__ __forceinline__ void job_tile(const plasma_desc_t &A, bool sym,
                                         int j, int *m, int *n)
{
    if (sym) {
        trapezoid_decode(j, A.mt, 0, m, n);
    }
    else {
        *m = j % A.mt;
        *n = j / A.mt;
    }
}

// Tile (m, n) as a view: lower-equivalent for symmetric storage (stored
// entries r >= c in diagonal tiles), plain column-major otherwise.
template <typename T>
__device__ __forceinline__ View<T> job_view(const plasma_desc_t &A, bool sym,
                                            bool upper, int m, int n)
{
    return sym ? lower_view<T>(A, upper, m, n) : tile_view<T>(A, m, n);
}

template <typename T>
__global__ __launch_bounds__(NT) void tile_max_kernel(plasma_desc_t A,
                                                      int sym, int upper,
                                                      double *partial,
                                                      const int *status)
{
    __shared__ double red[NT];
    if (block_should_stop(status)) return;
    int m, n;
    job_tile(A, sym, blockIdx.x, &m, &n);
    const int mv = plasma_tile_mview(A, m);
    const int nv = sym ? plasma_tile_mview(A, n) : plasma_tile_nview(A, n);
    const View<T> V = job_view<T>(A, sym, upper, m, n);
    const bool diag = sym && m == n;
    double v = 0.0;
    for (int c = 0; c < nv; c++)
        for (int r = threadIdx.x; r < mv; r += NT)
            if (!diag || r >= c) v = nanmax(v, fabs((double)V.at(r, c)));
    v = block_nanmax(v, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// (scale, ssq) of a tile; weight 2 for the entries that stand for two
// entries of a symmetric matrix
template <typename T>
__global__ __launch_bounds__(NT) void tile_ssq_kernel(plasma_desc_t A,
                                                      int sym, int upper,
                                                      double *pairs,
                                                      const int *status)
{
    __shared__ double red[NT];
    if (block_should_stop(status)) return;
    int m, n;
    job_tile(A, sym, blockIdx.x, &m, &n);
    const int mv = plasma_tile_mview(A, m);
    const int nv = sym ? plasma_tile_mview(A, n) : plasma_tile_nview(A, n);
    const View<T> V = job_view<T>(A, sym, upper, m, n);
    const bool diag = sym && m == n;
    double v = 0.0;
    for (int c = 0; c < nv; c++)
        for (int r = threadIdx.x; r < mv; r += NT)
            if (!diag || r >= c) v = nanmax(v, fabs((double)V.at(r, c)));
    const double scale = block_nanmax(v, red);
    __syncthreads();
    double q = 0.0;
    if (scale > 0.0 && scale == scale && scale <= 1.7976931348623157e308) {
        for (int c = 0; c < nv; c++)
            for (int r = threadIdx.x; r < mv; r += NT) {
                if (diag && r < c) continue;
                const double a = (double)V.at(r, c) / scale;
                const double w = (sym && (m != n || r != c)) ? 2.0 : 1.0;
                q += w * a * a;
            }
    }
    red[threadIdx.x] = q;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pairs[2 * blockIdx.x] = scale;
        pairs[2 * blockIdx.x + 1] =
            (scale == scale && scale > 1.7976931348623157e308) ? 1.0 : red[0];
    }
}

// (s1, q1) + (s2, q2) in the (scale, ssq) representation
__device__ __forceinline__ void ssq_combine(double &s1, double &q1, double s2,
                                            double q2)
{
    if (s1 != s1 || s2 != s2) {           // NaN
        s1 = s1 != s1 ? s1 : s2;
        q1 = 1.0;
        return;
    }
    if (s2 == 0.0) return;
    if (s1 >= s2) {
        const double r = s2 / s1;
        q1 += q2 * r * r;
    }
    else {
        const double r = s1 / s2;
        q1 = q2 + q1 * r * r;
        s1 = s2;
    }
}

__global__ __launch_bounds__(NT) void ssq_final_kernel(const double *pairs,
                                                       int njobs,
                                                       double *value,
                                                       const int *status)
{
    __shared__ double rs[NT], rq[NT];
    if (block_should_stop(status)) return;
    double s = 0.0, q = 0.0;
    for (int j = threadIdx.x; j < njobs; j += NT)
        ssq_combine(s, q, pairs[2 * j], pairs[2 * j + 1]);
    rs[threadIdx.x] = s;
    rq[threadIdx.x] = q;
    __syncthreads();
    for (int h = NT / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            double a = rs[threadIdx.x], b = rq[threadIdx.x];
            ssq_combine(a, b, rs[threadIdx.x + h], rq[threadIdx.x + h]);
            rs[threadIdx.x] = a;
            rq[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *value = rs[0] * sqrt(rq[0]);
}

// value = NaN-propagating max of work[0 .. n)
__global__ __launch_bounds__(NT) void final_max_kernel(const double *work,
                                                       int n, double *value,
                                                       const int *status)
{
    __shared__ double red[NT];
    if (block_should_stop(status)) return;
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) v = nanmax(v, work[i]);
    v = block_nanmax(v, red);
    if (threadIdx.x == 0) *value = v;
}

// General tile (m, n): column sums of |a| into colw[m * ldc + n * nb + c],
// row sums into roww[n * ldr + m * mb + r].
template <typename T>
__global__ __launch_bounds__(NT) void ge_sums_kernel(plasma_desc_t A,
                                                     double *roww, size_t ldr,
                                                     double *colw, size_t ldc,
                                                     const int *status)
{
    if (block_should_stop(status)) return;
    const int m = blockIdx.x % A.mt, n = blockIdx.x / A.mt;
    const int mv = plasma_tile_mview(A, m), nv = plasma_tile_nview(A, n);
    const View<T> V = tile_view<T>(A, m, n);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // rows: consecutive threads read consecutive rows of a column
    for (int r = threadIdx.x; r < mv; r += NT) {
        double s = 0.0;
        for (int c = 0; c < nv; c++) s += fabs((double)V.at(r, c));
        roww[n * ldr + (size_t)m * A.mb + r] = s;
    }
    // columns: one wave per column, lanes along it
    for (int c = wave; c < nv; c += NT / 64) {
        double s = 0.0;
        for (int r = lane; r < mv; r += 64) s += fabs((double)V.at(r, c));
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) colw[m * ldc + (size_t)n * A.nb + c] = s;
    }
}

// part[b] = max over the entries i of block b of sum_slot work[slot][i]
__global__ __launch_bounds__(NT) void slots_reduce_kernel(const double *work,
                                                          int nslots,
                                                          size_t ld, int count,
                                                          double *part,
                                                          const int *status)
{
    __shared__ double red[NT];
    if (block_should_stop(status)) return;
    const int i = blockIdx.x * NT + threadIdx.x;
    double v = 0.0;
    if (i < count)
        for (int s = 0; s < nslots; s++) v += work[s * ld + i];
    v = block_nanmax(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// values[i] = max_j |a_ij| over the rows of tile row blockIdx.x
template <typename T>
__global__ __launch_bounds__(NT) void amax_rows_kernel(plasma_desc_t A,
                                                       double *values,
                                                       const int *status)
{
    if (block_should_stop(status)) return;
    const int m = blockIdx.x;
    const int mv = plasma_tile_mview(A, m);
    for (int r = threadIdx.x; r < mv; r += NT) {
        double v = 0.0;
        for (int n = 0; n < A.nt; n++) {
            const View<T> V = tile_view<T>(A, m, n);
            const int nv = plasma_tile_nview(A, n);
            for (int c = 0; c < nv; c++) {
                const double a = fabs((double)V.at(r, c));
                if (a > v) v = a;
            }
        }
        values[(size_t)m * A.mb + r] = v;
    }
}

template <typename T>
int launch_norm(plasma_enum_t norm, plasma_enum_t uplo, plasma_desc_t A,
                double *work, double *value, const int *status, hipStream_t s)
{
    const bool sym = uplo != PlasmaGeneral;
    const int upper = uplo == PlasmaUpper;
    const long njobs = sym ? (long)A.mt * (A.mt + 1) / 2 : (long)A.mt * A.nt;
    if (njobs <= 0 || njobs > 0x7fffffffL) return PlasmaErrorNotSupported;
    const dim3 grid((unsigned)njobs), blk(NT);
    switch (norm) {
    case PlasmaMaxNorm:
        hipLaunchKernelGGL(tile_max_kernel<T>, grid, blk, 0, s, A, (int)sym,
                           upper, work, status);
        hipLaunchKernelGGL(final_max_kernel, dim3(1), blk, 0, s, work,
                           (int)njobs, value, status);
        break;
    case PlasmaFrobeniusNorm:
        hipLaunchKernelGGL(tile_ssq_kernel<T>, grid, blk, 0, s, A, (int)sym,
                           upper, work, status);
        hipLaunchKernelGGL(ssq_final_kernel, dim3(1), blk, 0, s, work,
                           (int)njobs, value, status);
        break;
    case PlasmaOneNorm:
    case PlasmaInfNorm:
        if (sym)
            return plasma_kern_lansy_inf(uplo, A, work, value, status, s);
        {
            const size_t ldr = (size_t)A.mt * A.mb, ldc = (size_t)A.nt * A.nb;
            double *roww = work, *colw = work + (size_t)A.nt * ldr;
            double *part = colw + (size_t)A.mt * ldc;
            hipLaunchKernelGGL(ge_sums_kernel<T>, grid, blk, 0, s, A, roww,
                               ldr, colw, ldc, status);
            const bool one = norm == PlasmaOneNorm;
            const int count = one ? A.n : A.m;
            const int nblk = (count + NT - 1) / NT;
            hipLaunchKernelGGL(slots_reduce_kernel, dim3(nblk), blk, 0, s,
                               one ? colw : roww, one ? A.mt : A.nt,
                               one ? ldc : ldr, count, part, status);
            hipLaunchKernelGGL(final_max_kernel, dim3(1), blk, 0, s, part,
                               nblk, value, status);
        }
        break;
    default:
        return PlasmaErrorIllegalValue;
    }
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T>
int launch_amax(plasma_enum_t colrow, plasma_desc_t A, double *values,
                const int *status, hipStream_t s)
{
    // per column: the mixed solver's kernel (kernels/mixed.hip), so the
    // public routine and the refinement test share one implementation
    if (colrow == PlasmaColumnwise)
        return plasma_kern_amax_cols(A, values, status, s);
    else {
        if (A.mt > 0)
            hipLaunchKernelGGL(amax_rows_kernel<T>, dim3(A.mt), dim3(NT), 0,
                               s, A, values, status);
    }
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

}  // namespace

extern "C" size_t plasma_norm_work_doubles(plasma_desc_t A)
{
    const size_t jobs = (size_t)A.mt * A.nt;
    const size_t ldr = (size_t)A.mt * A.mb, ldc = (size_t)A.nt * A.nb;
    size_t w = 2 * jobs;
    const size_t sums = A.nt * ldr + A.mt * ldc +
                        (ldr > ldc ? ldr : ldc) / NT + 1;
    if (sums > w) w = sums;
    const size_t lansy = plasma_lansy_work_doubles(A);
    if (lansy > w) w = lansy;
    return w + 16;
}

extern "C" int plasma_kern_norm(plasma_enum_t norm, plasma_enum_t uplo,
                                plasma_desc_t A, double *work, double *value,
                                const int *status, hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_norm<double>(norm, uplo, A, work, value, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_norm<float>(norm, uplo, A, work, value, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_amax(plasma_enum_t colrow, plasma_desc_t A,
                                double *values, const int *status,
                                hipStream_t s)
{
    if (A.precision == PlasmaRealDouble)
        return launch_amax<double>(colrow, A, values, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_amax<float>(colrow, A, values, status, s);
    return PlasmaErrorNotSupported;
}
