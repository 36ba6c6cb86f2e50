// layout.hip -- HBM-bound tile kernels for gfx950: LAPACK <-> tile layout
// translation, FP64 <-> FP32 conversion, copy and add.
//
// Reference: compute/pzge2desc.c:22-55, pzdesc2ge.c, pztr2desc.c:22-50,
// pzdesc2tr.c:22-51 (one lacpy task per tile), pzlag2c.c:47-85 with
// core_zlag2c.c:48-79, pclag2z.c:28-67, pzlacpy.c:118-140, pzgeadd.c:28-73.
// Here one workgroup moves one tile; a wave copies whole tile columns, so
// both the column-major LAPACK array and the tile are read and written in
// contiguous 512-byte runs.
#include <float.h>
#include "common.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NTHREADS = 256;

__global__ void fill_int_kernel(int *p, int v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ bool tile_in_region(int region, int m, int n)
{
    if (region == PlasmaLower) return n <= m;
    if (region == PlasmaUpper) return m <= n;
    return true;
}

// Copy between the LAPACK array of the view and its tiles.
template <typename T, bool TO_TILES>
__global__ __launch_bounds__(NTHREADS) void layout_kernel(
    plasma_desc_t A, T *pA, int lda, int region, const int *status)
{
    if (block_should_stop(status)) return;
    const int m = blockIdx.x % A.mt, n = blockIdx.x / A.mt;
    if (!tile_in_region(region, m, n)) return;
    const int mv = plasma_tile_mview(A, m), nv = plasma_tile_nview(A, n);
    T *tile = (T *)plasma_tile_addr(A, m, n);
    const int ldt = plasma_tile_mmain(A, m);
    T *f77 = pA + (size_t)A.nb * lda * n + (size_t)A.mb * m;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = wave; c < nv; c += NTHREADS / 64) {
        T *a = f77 + (size_t)c * lda;
        T *b = tile + (size_t)c * ldt;
        for (int r = lane; r < mv; r += 64) {
            if (TO_TILES) b[r] = a[r];
            else          a[r] = b[r];
        }
    }
}

// LAPACK band array AB <-> the tiles of a band descriptor (reference
// core_blas/core_zlacpy_band.c, lapack2tile_band / tile2lapack_band): A(i,j)
// of the band (lower: 0 <= i - j <= kd, upper: 0 <= j - i <= kd) is
// AB[off + i - j + j ldab] with off = kd for upper storage, 0 for lower.
// Tile (m, n) of the band, one workgroup each; towards the tiles every
// entry outside the band (or past the matrix) is written as zero, so the
// factorization finds zeros there.
template <typename T, bool TO_TILES>
__global__ __launch_bounds__(NTHREADS) void band_layout_kernel(
    plasma_desc_t A, T *pAB, int ldab, int kd, int upper, const int *status)
{
    if (block_should_stop(status)) return;
    const int bt = upper ? A.kut : A.klt;
    const int n = blockIdx.x / bt, d = blockIdx.x % bt;
    const int m = upper ? n - d : n + d;
    if (m < 0 || m >= A.mt || n >= A.nt) return;
    T *tile = (T *)plasma_tile_addr(A, m, n);
    const int ldt = plasma_tile_mmain(A, m);
    const int off = upper ? kd : 0;
    // a tile of the last tile column is only n % nb wide in the band
    // storage (the next tile follows it): no column past the view
    const int ncols = plasma_tile_nview(A, n);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = wave; c < ncols; c += NTHREADS / 64) {
        const int j = n * A.nb + c;
        for (int r = lane; r < A.mb; r += 64) {
            const int i = m * A.mb + r;
            const int dist = upper ? j - i : i - j;
            const bool in = i < A.m && j < A.n && dist >= 0 && dist <= kd;
            T *ab = pAB + (ptrdiff_t)(off + i - j) + (ptrdiff_t)j * ldab;
            if (TO_TILES) tile[r + (size_t)c * ldt] = in ? *ab : T(0);
            else if (in)  *ab = tile[r + (size_t)c * ldt];
        }
    }
}

// Tile-wise precision conversion between two general descriptors of the
// same shape and tiling.  FP64 -> FP32 rounds to nearest; an entry with
// x > FLT_MAX or x < -FLT_MAX (so also +-Inf, but not NaN) sets the status
// word to 1, the LAPACK dlag2s convention (reference core_zlag2c.c:68-75
// fails the sequence with info = 1).
template <typename S, typename D>
__global__ __launch_bounds__(NTHREADS) void lag2_kernel(
    plasma_desc_t A, plasma_desc_t B, int *status, int check)
{
    if (block_should_stop(status)) return;
    const int m = blockIdx.x % A.mt, n = blockIdx.x / A.mt;
    // triangular descriptors hold only the tiles of their triangle
    if (!tile_in_region(A.type, m, n)) return;
    const int mv = plasma_tile_mview(A, m), nv = plasma_tile_nview(A, n);
    const S *a = (const S *)plasma_tile_addr(A, m, n);
    D *b = (D *)plasma_tile_addr(B, m, n);
    const int lda = plasma_tile_mmain(A, m), ldb = plasma_tile_mmain(B, m);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bool overflow = false;
    for (int c = wave; c < nv; c += NTHREADS / 64) {
        for (int r = lane; r < mv; r += 64) {
            S v = a[r + (size_t)c * lda];
            if (check && (v > (S)FLT_MAX || v < -(S)FLT_MAX)) overflow = true;
            b[r + (size_t)c * ldb] = (D)v;
        }
    }
    if (overflow) atomicCAS(status, 0, 1);
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void geadd_kernel(
    T alpha, plasma_desc_t A, T beta, plasma_desc_t B, int copy,
    const int *status)
{
    if (block_should_stop(status)) return;
    const int m = blockIdx.x % B.mt, n = blockIdx.x / B.mt;
    const int mv = plasma_tile_mview(B, m), nv = plasma_tile_nview(B, n);
    const T *a = (const T *)plasma_tile_addr(A, m, n);
    T *b = (T *)plasma_tile_addr(B, m, n);
    const int lda = plasma_tile_mmain(A, m), ldb = plasma_tile_mmain(B, m);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = wave; c < nv; c += NTHREADS / 64)
        for (int r = lane; r < mv; r += 64) {
            T av = a[r + (size_t)c * lda];
            T *bp = &b[r + (size_t)c * ldb];
            *bp = copy ? av : alpha * av + beta * (*bp);
        }
}

template <typename T, bool TO_TILES>
int launch_layout(plasma_enum_t region, void *pA, int lda, plasma_desc_t A,
                  const int *status, hipStream_t s)
{
    const long ntiles = (long)A.mt * A.nt;
    if (ntiles == 0) return 0;
    hipLaunchKernelGGL((layout_kernel<T, TO_TILES>), dim3((unsigned)ntiles),
                       dim3(NTHREADS), 0, s, A, (T *)pA, lda, (int)region,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

}  // namespace

__global__ void copy_int_kernel(int *dst, const int *src)
{
    if (threadIdx.x == 0) *dst = *src;
}

extern "C" int plasma_kern_copy_int(int *dst, const int *src, hipStream_t s)
{
    hipLaunchKernelGGL(copy_int_kernel, dim3(1), dim3(64), 0, s, dst, src);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_fill_int(int *p, int v, int n, hipStream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       s, p, v, n);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_lapack2tile(plasma_enum_t region, const void *pA,
                                       int lda, plasma_desc_t A,
                                       const int *status, hipStream_t s)
{
    if (plasma_is_complex(A.precision)) {
        A = plasma_desc_real_view(A);
        lda *= 2;
    }
    if (A.precision == PlasmaRealDouble)
        return launch_layout<double, true>(region, (void *)pA, lda, A, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_layout<float, true>(region, (void *)pA, lda, A, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_tile2lapack(plasma_enum_t region, plasma_desc_t A,
                                       void *pA, int lda, const int *status,
                                       hipStream_t s)
{
    if (plasma_is_complex(A.precision)) {
        A = plasma_desc_real_view(A);
        lda *= 2;
    }
    if (A.precision == PlasmaRealDouble)
        return launch_layout<double, false>(region, pA, lda, A, status, s);
    if (A.precision == PlasmaRealFloat)
        return launch_layout<float, false>(region, pA, lda, A, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_lag2(plasma_desc_t src, plasma_desc_t dst,
                                int *status, hipStream_t s)
{
    // complex: both parts converted, an overflow of either sets the status
    // (LAPACK zlag2c, reference core_zlag2c.c:48-79)
    if (plasma_is_complex(src.precision) != plasma_is_complex(dst.precision))
        return PlasmaErrorIllegalValue;
    src = plasma_desc_real_view(src);
    dst = plasma_desc_real_view(dst);
    const long ntiles = (long)src.mt * src.nt;
    if (ntiles == 0) return 0;
    if (src.precision == PlasmaRealDouble && dst.precision == PlasmaRealFloat)
        hipLaunchKernelGGL((lag2_kernel<double, float>), dim3((unsigned)ntiles),
                           dim3(NTHREADS), 0, s, src, dst, status, 1);
    else if (src.precision == PlasmaRealFloat &&
             dst.precision == PlasmaRealDouble)
        hipLaunchKernelGGL((lag2_kernel<float, double>), dim3((unsigned)ntiles),
                           dim3(NTHREADS), 0, s, src, dst, status, 0);
    else
        return PlasmaErrorIllegalValue;
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_lacpy(plasma_desc_t A, plasma_desc_t B,
                                 const int *status, hipStream_t s)
{
    A = plasma_desc_real_view(A);
    B = plasma_desc_real_view(B);
    const long ntiles = (long)B.mt * B.nt;
    if (ntiles == 0) return 0;
    if (A.precision == PlasmaRealDouble)
        hipLaunchKernelGGL(geadd_kernel<double>, dim3((unsigned)ntiles),
                           dim3(NTHREADS), 0, s, 1.0, A, 0.0, B, 1, status);
    else if (A.precision == PlasmaRealFloat)
        hipLaunchKernelGGL(geadd_kernel<float>, dim3((unsigned)ntiles),
                           dim3(NTHREADS), 0, s, 1.0f, A, 0.0f, B, 1, status);
    else
        return PlasmaErrorNotSupported;
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_geadd(double alpha, plasma_desc_t A, double beta,
                                 plasma_desc_t B, const int *status,
                                 hipStream_t s)
{
    // real alpha and beta: a complex descriptor is added part by part
    A = plasma_desc_real_view(A);
    B = plasma_desc_real_view(B);
    const long ntiles = (long)B.mt * B.nt;
    if (ntiles == 0) return 0;
    if (A.precision == PlasmaRealDouble)
        hipLaunchKernelGGL(geadd_kernel<double>, dim3((unsigned)ntiles),
                           dim3(NTHREADS), 0, s, alpha, A, beta, B, 0, status);
    else if (A.precision == PlasmaRealFloat)
        hipLaunchKernelGGL(geadd_kernel<float>, dim3((unsigned)ntiles),
                           dim3(NTHREADS), 0, s, (float)alpha, A, (float)beta,
                           B, 0, status);
    else
        return PlasmaErrorNotSupported;
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_band_layout(plasma_desc_t A, void *pAB, int ldab,
                                       int kd, int upper, int to_tiles,
                                       const int *status, hipStream_t s)
{
    const int bt = upper ? A.kut : A.klt;
    const long nblocks = (long)A.nt * bt;
    if (nblocks <= 0) return 0;
    if (nblocks > 0x7fffffffL) return PlasmaErrorNotSupported;
    const dim3 grid((unsigned)nblocks), block(NTHREADS);
    if (A.precision == PlasmaRealDouble) {
        if (to_tiles)
            hipLaunchKernelGGL((band_layout_kernel<double, true>), grid, block,
                               0, s, A, (double *)pAB, ldab, kd, upper,
                               status);
        else
            hipLaunchKernelGGL((band_layout_kernel<double, false>), grid,
                               block, 0, s, A, (double *)pAB, ldab, kd, upper,
                               status);
    }
    else if (A.precision == PlasmaRealFloat) {
        if (to_tiles)
            hipLaunchKernelGGL((band_layout_kernel<float, true>), grid, block,
                               0, s, A, (float *)pAB, ldab, kd, upper, status);
        else
            hipLaunchKernelGGL((band_layout_kernel<float, false>), grid, block,
                               0, s, A, (float *)pAB, ldab, kd, upper, status);
    }
    else {
        return PlasmaErrorNotSupported;
    }
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}
