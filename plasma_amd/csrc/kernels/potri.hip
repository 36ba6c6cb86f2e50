// potri.hip -- pieces of the inverse from the Cholesky factor for gfx950
// (plasma_{d,s}potri, plasma_{d,s}poinv; reference compute/zpotri.c,
// compute/zpoinv.c).  The products themselves run on the library's MFMA
// kernels: W = L^{-1} by the wide-right-hand-side forward sweep from the
// identity, and A^{-1} = W^T W by the triangular Gram product of the
// adjoint (the tiles I >= J of sum_{K >= I} W(K,I)^T W(K,J)).  Here:
//  * identity_kernel: a general tile matrix set to I;
//  * diag_zero_kernel: the first zero on the factor's diagonal (LAPACK
//    dtrtri's singularity check) into the sequence's status word;
//  * copy_tri_kernel: the uplo triangle of the symmetric result Q into A's
//    stored tiles (the other triangle of A's diagonal tiles untouched).
#include "common.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NT = 256;

template <typename T>
__global__ __launch_bounds__(NT) void identity_kernel(plasma_desc_t B,
                                                      const int *status)
{
    if (block_should_stop(status)) return;
    const int m = blockIdx.x, n = blockIdx.y;
    const int mv = plasma_tile_mview(B, m), nv = plasma_tile_nview(B, n);
    const View<T> V = tile_view<T>(B, m, n);
    for (int c = 0; c < nv; c++)
        for (int r = threadIdx.x; r < mv; r += NT)
            V.at(r, c) = (m == n && r == c) ? T(1) : T(0);
}

// One workgroup: the smallest i with L(i,i) == 0, reported as i + 1 (the
// first failure of the sequence wins, as for the factorization's pivots).
template <typename T>
__global__ __launch_bounds__(NT) void diag_zero_kernel(plasma_desc_t A,
                                                       int upper, int *status)
{
    __shared__ int red[NT];
    if (block_should_stop(status)) return;
    int first = 0x7fffffff;
    for (int i = threadIdx.x; i < A.m; i += NT) {
        const int k = i / A.mb, r = i % A.mb;
        const View<T> D = lower_view<T>(A, upper != 0, k, k);
        if (D.at(r, r) == T(0) && i < first) first = i;
    }
    red[threadIdx.x] = first;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s && red[threadIdx.x + s] < red[threadIdx.x])
            red[threadIdx.x] = red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] != 0x7fffffff)
        atomicCAS(status, 0, red[0] + 1);
}

// A's stored tile (m, n), m >= n in the lower-equivalent view, from Q's
// tile (m, n) (Q symmetric, general storage); diagonal tiles: r >= c only.
template <typename T>
__global__ __launch_bounds__(NT) void copy_tri_kernel(plasma_desc_t Q,
                                                      plasma_desc_t A,
                                                      int upper,
                                                      const int *status)
{
    if (block_should_stop(status)) return;
    int m, n;
    trapezoid_decode(blockIdx.x, A.mt, 0, &m, &n);
    const int mv = plasma_tile_mview(A, m), nv = plasma_tile_mview(A, n);
    const View<T> S = tile_view<T>(Q, m, n);
    const View<T> D = lower_view<T>(A, upper != 0, m, n);
    if (!upper) {
        for (int c = 0; c < nv; c++)
            for (int r = threadIdx.x; r < mv; r += NT)
                if (m != n || r >= c) D.at(r, c) = S.at(r, c);
    }
    else {
        // D has unit column stride: consecutive threads along c
        for (int r = 0; r < mv; r++)
            for (int c = threadIdx.x; c < nv; c += NT)
                if (m != n || r >= c) D.at(r, c) = S.at(r, c);
    }
}

template <typename T>
int launch_potri_pieces(int which, plasma_enum_t uplo, plasma_desc_t A,
                        plasma_desc_t B, int *status, hipStream_t s)
{
    const int upper = uplo == PlasmaUpper;
    switch (which) {
    case 0:
        if (B.mt > 0 && B.nt > 0)
            hipLaunchKernelGGL(identity_kernel<T>, dim3(B.mt, B.nt), dim3(NT),
                               0, s, B, status);
        break;
    case 1:
        hipLaunchKernelGGL(diag_zero_kernel<T>, dim3(1), dim3(NT), 0, s, A,
                           upper, status);
        break;
    default: {
        const long ntiles = (long)A.mt * (A.mt + 1) / 2;
        if (ntiles > 0)
            hipLaunchKernelGGL(copy_tri_kernel<T>, dim3((unsigned)ntiles),
                               dim3(NT), 0, s, B, A, upper, status);
    }
    }
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

int dispatch(int which, plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
             int *status, hipStream_t s)
{
    const int prec = which == 0 ? B.precision : A.precision;
    if (prec == PlasmaRealDouble)
        return launch_potri_pieces<double>(which, uplo, A, B, status, s);
    if (prec == PlasmaRealFloat)
        return launch_potri_pieces<float>(which, uplo, A, B, status, s);
    return PlasmaErrorNotSupported;
}

}  // namespace

extern "C" int plasma_kern_identity(plasma_desc_t B, const int *status,
                                    hipStream_t s)
{
    return dispatch(0, PlasmaGeneral, B, B, (int *)status, s);
}

extern "C" int plasma_kern_diag_zero(plasma_enum_t uplo, plasma_desc_t A,
                                     int *status, hipStream_t s)
{
    return dispatch(1, uplo, A, A, status, s);
}

extern "C" int plasma_kern_copy_tri(plasma_enum_t uplo, plasma_desc_t Q,
                                    plasma_desc_t A, const int *status,
                                    hipStream_t s)
{
    return dispatch(2, uplo, A, Q, (int *)status, s);
}
