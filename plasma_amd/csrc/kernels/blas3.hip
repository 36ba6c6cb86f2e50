// blas3.hip -- the tile GEMM behind plasma_{d,s}gemm / syrk / trsm for
// gfx950, and the diagonal-tile preparation of the tile TRSM.
//
// gemm_tiles_kernel: C = alpha op(A) op(B) + beta C on general tile
// descriptors (reference compute/pzgemm.c and pzsyrk.c: one tile gemm /
// syrk task per (m, n, k); here one workgroup per 128 x 128 block of a C
// tile with the whole K sum held in its accumulators, so C is read and
// written once).  Operands stream through double-buffered LDS by GemmNT
// (common.hip.h) one K tile at a time; the product is formed as
// C^T = op(B)^T op(A)^T so that 16 consecutive rows of one C column sit in
// 16 consecutive lanes (C tiles are column major).  With a triangle (syrk)
// only the tiles of that triangle are touched and, inside diagonal tiles,
// only its elements (core_zsyrk updates one triangle of a diagonal tile).
// alpha == 0 reads no operand; beta == 0 reads no C (BLAS semantics).
//
// trsm_diag_prep_kernel: tile (k, k) of a triangular A as its lower-
// equivalent triangle L_kk (L = A for PlasmaLower, A^T for PlasmaUpper),
// unit diagonal for PlasmaUnit, zeros above, into tile (0, k) of a one-tile-
// row scratch descriptor; the tile inverses are then formed by
// plasma_kern_trtri_tiles (kernels/potrf.hip).
#include "common.hip.h"
#include "tilegemm.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NT = 256;
constexpr int BM = 128;

template <typename T> struct KcOf;
template <> struct KcOf<double> { static constexpr int v = 16; };
template <> struct KcOf<float>  { static constexpr int v = 32; };

template <typename T>
struct GemmArgs {
    plasma_desc_t A, B, C;
    int ta, tb;          // op(A) = A^T / op(B) = B^T
    T alpha, beta;
    int kt;              // K tiles
    int tri;             // 0: all of C; 1: lower triangle; 2: upper
    int bpm, bpn;        // 128-row / -column blocks per C tile
};

// Rows r0.. of tile row m of op(A) for K tile l: element (r, k) of the view
template <typename T>
__device__ __forceinline__ void p_view(const GemmArgs<T> &g, int m, int l,
                                       int r0, View<T> &v, int &rows, int &kv)
{
    if (!g.ta) {
        v.p = (T *)plasma_tile_addr(g.A, m, l) + r0;
        v.rs = 1;
        v.cs = plasma_tile_mmain(g.A, m);
        rows = plasma_tile_mview(g.A, m) - r0;
        kv = plasma_tile_nview(g.A, l);
    }
    else {
        const int ld = plasma_tile_mmain(g.A, l);
        v.p = (T *)plasma_tile_addr(g.A, l, m) + (ptrdiff_t)r0 * ld;
        v.rs = ld;
        v.cs = 1;
        rows = plasma_tile_nview(g.A, m) - r0;
        kv = plasma_tile_mview(g.A, l);
    }
}

// Columns c0.. of tile column n of op(B) for K tile l, as rows: element
// (r, k) of the view is op(B)(k, c0 + r)
template <typename T>
__device__ __forceinline__ void q_view(const GemmArgs<T> &g, int n, int l,
                                       int c0, View<T> &v, int &rows, int &kv)
{
    if (!g.tb) {
        const int ld = plasma_tile_mmain(g.B, l);
        v.p = (T *)plasma_tile_addr(g.B, l, n) + (ptrdiff_t)c0 * ld;
        v.rs = ld;
        v.cs = 1;
        rows = plasma_tile_nview(g.B, n) - c0;
        kv = plasma_tile_mview(g.B, l);
    }
    else {
        v.p = (T *)plasma_tile_addr(g.B, n, l) + c0;
        v.rs = 1;
        v.cs = plasma_tile_mmain(g.B, n);
        rows = plasma_tile_mview(g.B, n) - c0;
        kv = plasma_tile_nview(g.B, l);
    }
}

template <typename T>
__global__ __launch_bounds__(NT, 2) void gemm_tiles_kernel(GemmArgs<T> g,
                                                           int njobs,
                                                           const int *status)
{
    typedef GemmNT<T, BM, BM, KcOf<T>::v, 64, 64> G;
    __shared__ __attribute__((aligned(16))) typename G::Smem sm;
    if (block_should_stop(status)) return;
    const int padded = tg::xcd_grid(njobs), chunk = padded / tg::NXCD;
    const int per = g.bpm * g.bpn;
    for (int b = blockIdx.x; b < padded; b += gridDim.x) {
        const int jb = (b % tg::NXCD) * chunk + b / tg::NXCD;
        if (jb >= njobs) continue;
        // tiles column by column, the blocks of a tile together (they share
        // op(A) rows or op(B) columns)
        const int tile = jb / per, sub = jb % per;
        const int m = tile % g.C.mt, n = tile / g.C.mt;
        const int bi = sub % g.bpm, bj = sub / g.bpm;
        const int mv = plasma_tile_mview(g.C, m), nv = plasma_tile_nview(g.C, n);
        const int r0 = bi * BM, c0 = bj * BM;
        if (r0 >= mv || c0 >= nv) continue;
        const int rows = min(BM, mv - r0), cols = min(BM, nv - c0);
        // block origin in the view's coordinates (views are tile aligned)
        const long gr = (long)m * g.C.mb + r0, gc = (long)n * g.C.nb + c0;
        if (g.tri == 1 && gc > gr + rows - 1) continue;    // all above
        if (g.tri == 2 && gr > gc + cols - 1) continue;    // all below

        typename G::acc_t acc[G::FM][G::FN];
        G::zero(acc);
        if (g.alpha != T(0)) {
            for (int l = 0; l < g.kt; l++) {
                View<T> P, Q;
                int prow, pk, qrow, qk;
                p_view(g, m, l, r0, P, prow, pk);
                q_view(g, n, l, c0, Q, qrow, qk);
                // run() ends with a barrier after its last LDS read
                G::run(acc, Q, qrow, P, prow, min(pk, qk), sm);
            }
        }
        View<T> Cv = tile_view<T>(g.C, m, n);
        Cv.p += r0 + (ptrdiff_t)c0 * Cv.cs;
#pragma unroll
        for (int i = 0; i < G::FM; i++)
#pragma unroll
            for (int j = 0; j < G::FN; j++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int c = G::acc_row(i, q), r = G::acc_col(j);
                    if (r >= rows || c >= cols) continue;
                    if (g.tri == 1 && gc + c > gr + r) continue;
                    if (g.tri == 2 && gr + r > gc + c) continue;
                    T v = g.alpha * acc[i][j][q];
                    if (g.beta != T(0)) v += g.beta * Cv.at(r, c);
                    Cv.at(r, c) = v;
                }
    }
}

template <typename T>
int launch_gemm(int ta, int tb, double alpha, plasma_desc_t A,
                plasma_desc_t B, double beta, plasma_desc_t C, int tri,
                const int *status, hipStream_t s)
{
    if (C.mt <= 0 || C.nt <= 0) return 0;
    GemmArgs<T> g;
    g.A = A;
    g.B = B;
    g.C = C;
    g.ta = ta;
    g.tb = tb;
    g.alpha = (T)alpha;
    g.beta = (T)beta;
    g.kt = ta ? A.mt : A.nt;
    g.tri = tri;
    g.bpm = (C.mb + BM - 1) / BM;
    g.bpn = (C.nb + BM - 1) / BM;
    const long njobs = (long)C.mt * C.nt * g.bpm * g.bpn;
    if (njobs > 0x7ffffff0L) return PlasmaErrorNotSupported;
    const long grid = tg::xcd_grid((int)njobs);
    hipLaunchKernelGGL(gemm_tiles_kernel<T>, dim3((unsigned)grid), dim3(NT),
                       0, s, g, (int)njobs, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename T>
__global__ __launch_bounds__(NT) void trsm_diag_prep_kernel(plasma_desc_t A,
                                                            int upper,
                                                            int unit,
                                                            plasma_desc_t D,
                                                            const int *status)
{
    if (block_should_stop(status)) return;
    const int k = blockIdx.x;
    const int nk = plasma_tile_mview(A, k);
    const View<T> S = lower_view<T>(A, upper != 0, k, k);
    View<T> Dv;
    Dv.p = (T *)plasma_tile_addr(D, 0, k);
    Dv.rs = 1;
    Dv.cs = plasma_tile_mmain(D, 0);
    for (int c = 0; c < nk; c++)
        for (int r = threadIdx.x; r < nk; r += NT) {
            T v = T(0);
            if (r > c) v = S.at(r, c);
            else if (r == c) v = unit ? T(1) : S.at(r, c);
            Dv.at(r, c) = v;
        }
}

}  // namespace

extern "C" int plasma_kern_gemm(int transa, int transb, double alpha,
                                plasma_desc_t A, plasma_desc_t B, double beta,
                                plasma_desc_t C, plasma_enum_t tri,
                                const int *status, hipStream_t s)
{
    const int ta = transa != PlasmaNoTrans, tb = transb != PlasmaNoTrans;
    const int t = tri == PlasmaLower ? 1 : tri == PlasmaUpper ? 2 : 0;
    if (A.precision != C.precision || B.precision != C.precision)
        return PlasmaErrorNotSupported;
    if (C.precision == PlasmaRealDouble)
        return launch_gemm<double>(ta, tb, alpha, A, B, beta, C, t, status, s);
    if (C.precision == PlasmaRealFloat)
        return launch_gemm<float>(ta, tb, alpha, A, B, beta, C, t, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_trsm_diag_prep(plasma_enum_t uplo,
                                          plasma_enum_t diag,
                                          plasma_desc_t A, plasma_desc_t D,
                                          const int *status, hipStream_t s)
{
    if (A.mt <= 0) return 0;
    const int upper = uplo == PlasmaUpper, unit = diag == PlasmaUnit;
    if (A.precision == PlasmaRealDouble)
        hipLaunchKernelGGL(trsm_diag_prep_kernel<double>, dim3(A.mt),
                           dim3(NT), 0, s, A, upper, unit, D, status);
    else if (A.precision == PlasmaRealFloat)
        hipLaunchKernelGGL(trsm_diag_prep_kernel<float>, dim3(A.mt),
                           dim3(NT), 0, s, A, upper, unit, D, status);
    else
        return PlasmaErrorNotSupported;
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}
