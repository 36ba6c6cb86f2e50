// gmres.hip -- FP64 kernels of plasma_dhposv (no reference counterpart):
// symmetric diagonal equilibration of a tile matrix and the column-wise
// vector operations of GMRES with many right-hand sides in lockstep.
//
//  * equil_diag_kernel: s_i = 1 / sqrt(a_ii) (1 where a_ii <= 0, so a
//    matrix that is not positive definite still fails in the factor).
//  * scale_sym_kernel: a_ij *= s_i s_j over the stored tiles of a
//    triangular descriptor; scale_rows_kernel: b_ij *= s_i (general).
//  * coldots_kernel: out[i][c] = sum_r V_i(r, c) Z(r, c), one workgroup per
//    (c, i), the rows summed in a fixed order (deterministic).
//  * colcombine_kernel: Y = beta Y + alpha sum_i coef[i][c] V_i (beta = 0
//    does not read Y).
// Vectors are n x nrhs general descriptors of one geometry: V_i is the
// descriptor Z with its storage pointer replaced (VecSet).
#include "common.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NTHREADS = 256;

__device__ __forceinline__ double &elem(const plasma_desc_t &D, int r, int c)
{
    return tile_view<double>(D, r / D.mb, c / D.nb).at(r % D.mb, c % D.nb);
}

__device__ __forceinline__ plasma_desc_t with_storage(plasma_desc_t D,
                                                      const double *p)
{
    D.matrix = const_cast<double *>(p);
    return D;
}

__global__ __launch_bounds__(NTHREADS) void equil_diag_kernel(
    plasma_desc_t A, int upper, double *s, const int *status)
{
    if (block_should_stop(status)) return;
    const int i = blockIdx.x * NTHREADS + threadIdx.x;
    if (i >= A.m) return;
    const double a = lower_view<double>(A, upper != 0, i / A.mb, i / A.mb)
                         .at(i % A.mb, i % A.mb);
    s[i] = a > 0.0 ? 1.0 / sqrt(a) : 1.0;
}

// one workgroup per stored tile: (I, J), I >= J, of the lower-equivalent
// numbering; the storage tile is (I, J) (lower) or (J, I) (upper)
__global__ __launch_bounds__(NTHREADS) void scale_sym_kernel(
    plasma_desc_t A, int upper, const double *s, const int *status)
{
    if (block_should_stop(status)) return;
    int I, J;
    trapezoid_decode(blockIdx.x, A.mt, 0, &I, &J);
    const int m = upper ? J : I, n = upper ? I : J;
    const int mv = plasma_tile_mview(A, m), nv = plasma_tile_nview(A, n);
    const View<double> T = tile_view<double>(A, m, n);
    for (int e = threadIdx.x; e < mv * nv; e += NTHREADS) {
        const int r = e % mv, c = e / mv;
        T.at(r, c) *= s[m * A.mb + r] * s[n * A.nb + c];
    }
}

__global__ __launch_bounds__(NTHREADS) void scale_rows_kernel(
    plasma_desc_t B, const double *s, const int *status)
{
    if (block_should_stop(status)) return;
    const long total = (long)B.m * B.n;
    for (long e = (long)blockIdx.x * NTHREADS + threadIdx.x; e < total;
         e += (long)gridDim.x * NTHREADS) {
        const int r = (int)(e % B.m), c = (int)(e / B.m);
        elem(B, r, c) *= s[r];
    }
}

__global__ __launch_bounds__(NTHREADS) void coldots_kernel(
    plasma_desc_t Z, plasma_vecset_t V, double *out, const int *status)
{
    __shared__ double red[NTHREADS];
    if (block_should_stop(status)) return;
    const int c = blockIdx.x, i = blockIdx.y;
    const plasma_desc_t Vi = with_storage(Z, V.p[i]);
    double sum = 0.0;
    for (int r = threadIdx.x; r < Z.m; r += NTHREADS)
        sum += elem(Vi, r, c) * elem(Z, r, c);
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int h = NTHREADS / 2; h > 0; h /= 2) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)i * Z.n + c] = red[0];
}

__global__ __launch_bounds__(NTHREADS) void colcombine_kernel(
    plasma_desc_t Y, double beta, double alpha, plasma_vecset_t V,
    const double *coef, const int *status)
{
    if (block_should_stop(status)) return;
    const long total = (long)Y.m * Y.n;
    for (long e = (long)blockIdx.x * NTHREADS + threadIdx.x; e < total;
         e += (long)gridDim.x * NTHREADS) {
        const int r = (int)(e % Y.m), c = (int)(e / Y.m);
        double acc = 0.0;
        for (int i = 0; i < V.nv; i++)
            acc += coef[(size_t)i * Y.n + c] *
                   elem(with_storage(Y, V.p[i]), r, c);
        double &y = elem(Y, r, c);
        y = (beta == 0.0 ? 0.0 : beta * y) + alpha * acc;
    }
}

int elementwise_grid(const plasma_desc_t &D)
{
    const long blocks = ((long)D.m * D.n + NTHREADS - 1) / NTHREADS;
    return (int)(blocks < 8192 ? blocks : 8192);
}

}  // namespace

extern "C" int plasma_kern_equil_diag(plasma_enum_t uplo, plasma_desc_t A,
                                      double *s, const int *status,
                                      hipStream_t st)
{
    if (A.precision != PlasmaRealDouble || A.m == 0) return 0;
    hipLaunchKernelGGL(equil_diag_kernel,
                       dim3((A.m + NTHREADS - 1) / NTHREADS), dim3(NTHREADS),
                       0, st, A, uplo == PlasmaUpper, s, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_scale_sym(plasma_enum_t uplo, plasma_desc_t A,
                                     const double *s, const int *status,
                                     hipStream_t st)
{
    if (A.precision != PlasmaRealDouble || A.mt == 0) return 0;
    const long ntiles = (long)A.mt * (A.mt + 1) / 2;
    hipLaunchKernelGGL(scale_sym_kernel, dim3((unsigned)ntiles),
                       dim3(NTHREADS), 0, st, A, uplo == PlasmaUpper, s,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_scale_rows(plasma_desc_t B, const double *s,
                                      const int *status, hipStream_t st)
{
    if (B.precision != PlasmaRealDouble || B.m == 0 || B.n == 0) return 0;
    hipLaunchKernelGGL(scale_rows_kernel, dim3(elementwise_grid(B)),
                       dim3(NTHREADS), 0, st, B, s, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_coldots(plasma_desc_t Z, plasma_vecset_t V,
                                   double *out, const int *status,
                                   hipStream_t st)
{
    if (Z.precision != PlasmaRealDouble || V.nv < 1 ||
        V.nv > PLASMA_VECSET_MAX)
        return PlasmaErrorIllegalValue;
    if (Z.m == 0 || Z.n == 0) return 0;
    hipLaunchKernelGGL(coldots_kernel, dim3(Z.n, V.nv), dim3(NTHREADS), 0,
                       st, Z, V, out, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_colcombine(plasma_desc_t Y, double beta,
                                      double alpha, plasma_vecset_t V,
                                      const double *coef, const int *status,
                                      hipStream_t st)
{
    if (Y.precision != PlasmaRealDouble || V.nv < 0 ||
        V.nv > PLASMA_VECSET_MAX)
        return PlasmaErrorIllegalValue;
    if (Y.m == 0 || Y.n == 0) return 0;
    hipLaunchKernelGGL(colcombine_kernel, dim3(elementwise_grid(Y)),
                       dim3(NTHREADS), 0, st, Y, beta, alpha, V, coef,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}
