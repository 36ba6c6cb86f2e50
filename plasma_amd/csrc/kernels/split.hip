// split.hip -- the FP32 trailing update on bf16 MFMA for gfx950.
//
// Every element x of a solved FP32 panel is written once as three bf16
// terms, hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid) (each
// difference is exact in FP32, and x - hi - mid - lo is below 2^-26 |x|).
// A product x y is then formed from the six terms whose weight is at least
// 2^-16 of hi*hi:
//     hi*hi + hi*mid + mid*hi + mid*mid + hi*lo + lo*hi
// each exact in FP32 (8-bit significands) and summed in FP32 by
// v_mfma_f32_16x16x32_bf16; the dropped terms (mid*lo, lo*mid, lo*lo) are
// below 2^-24 of |x y|.  The update C -= P(m) P(n)^T therefore keeps FP32
// accuracy while running on the bf16 matrix cores, whose dense rate is 16
// times the FP32 one (six products: 2.7 times the FP32 rate).
//
//  * split_panel_kernel: the panel columns of one factorization step into
//    three bf16 planes per panel tile, row major (row stride pstride, a
//    multiple of 8), so a lane reads its 8 consecutive k values with one
//    16-byte load.  32 x 32 blocks transposed through LDS.
//  * split_update_kernel: the trailing update of one column range, jobs as
//    in the FP32 update (copy-back of the group's panels first, then
//    128 x 128 blocks of the tiles (m, n), n >= n_first, column by column),
//    workgroup b running job (b % 8) * chunk + b / 8.  K streams in chunks
//    of 32 through LDS (3 planes x 128 rows x 32 k per operand, rows padded
//    to 40 elements), the next chunk prefetched into registers.  C is
//    loaded into the accumulators negated and stored negated, so the
//    products are only added.
//
// MFMA operand layout (measured on gfx950): lane l supplies row l % 16 of
// A and column l % 16 of B with 8 k values each; the k assignment only has
// to agree between A and B.  Result element q of lane l is D[4 (l / 16) +
// q][l % 16].
#include "common.hip.h"
#include "tilegemm.hip.h"
#include "plasma_kernels.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;
constexpr int BM = 128;
constexpr int KC = 32;              // k per chunk = one MFMA k step
constexpr int LDP = KC + 8;         // LDS row stride (bf16 elements)

// Plane q of panel tile t (tile m = pbase + 1 + t): row r, column c at
// planes[((t * 3 + q) * mb + r) * pstride + c].
__device__ __forceinline__ size_t plane_off(int t, int q, int mb, int r,
                                            int pstride)
{
    return ((size_t)(t * 3 + q) * mb + r) * pstride;
}

// Three-term split of x (see the file comment).
__device__ __forceinline__ void split3(float x, __bf16 &h, __bf16 &m,
                                       __bf16 &l)
{
    h = (__bf16)x;
    const float r1 = x - (float)h;
    m = (__bf16)r1;
    const float r2 = r1 - (float)m;
    l = (__bf16)r2;
}

// ---------------------------------------------------------------------------
// Panel columns [col0, col0 + ncol) of the panel tiles t in [t0, t1) into
// the planes.  Block (t, rb, cb): rows rb * 32.., columns cb * 32...
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void split_panel_kernel(
    plasma_desc_t A, int pbase, int t0, int pcols, int col0, int ncol,
    const float *panel, __bf16 *planes, int pstride, const int *status)
{
    __shared__ float sb[32][33];
    if (block_should_stop(status)) return;
    const int t = t0 + blockIdx.x;
    const int r0 = blockIdx.y * 32, c0 = blockIdx.z * 32;
    const int mv = plasma_tile_mview(A, pbase + 1 + t);
    if (r0 >= mv || c0 >= ncol) return;
    const float *src = panel + (size_t)t * A.mb * pcols;   // ld = mb
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;   // 32 x 8
    for (int cc = ty; cc < 32; cc += 8) {
        const int r = r0 + tx, c = c0 + cc;
        sb[cc][tx] = (r < mv && c < ncol)
            ? src[(size_t)(col0 + c) * A.mb + r] : 0.0f;
    }
    __syncthreads();
    for (int rr = ty; rr < 32; rr += 8) {
        const int r = r0 + rr, c = c0 + tx;
        if (r < mv && c < ncol) {
            __bf16 h, m, l;
            split3(sb[tx][rr], h, m, l);
            planes[plane_off(t, 0, A.mb, r, pstride) + col0 + c] = h;
            planes[plane_off(t, 1, A.mb, r, pstride) + col0 + c] = m;
            planes[plane_off(t, 2, A.mb, r, pstride) + col0 + c] = l;
        }
    }
}

// ---------------------------------------------------------------------------
// Trailing update on the planes.
// ---------------------------------------------------------------------------
// Tile t of the trapezoid {(m, n): n in [n_first, n_first + ncols),
// m >= n} in pair order: columns are taken two at a time, and within a
// pair (c, c + 1) the tiles go (c, c), (c+1, c), (c+1, c+1), (c+2, c),
// (c+2, c+1), ...  so a panel row block X(m) serves both columns back to
// back while the two Y blocks stay in the XCD's L2.  With T = mt - n_first,
// pair p holds 2 T - 4 p - 1 tiles (a last, unpaired column T - 2 p), and
// P(p) = p (2 T + 1 - 2 p) tiles precede it.
__device__ __forceinline__ void pair_decode(int t, int mt, int n_first,
                                            int ncols, int *m, int *n)
{
    const long T = mt - n_first;
    const double b = 2.0 * T + 1.0;
    long p = (long)((b - sqrt(b * b - 8.0 * (double)t)) * 0.25);
    if (p < 0) p = 0;
    auto pre = [&](long q) { return q * (2 * T + 1 - 2 * q); };
    while (p > 0 && pre(p) > t) p--;
    while (2 * (p + 1) < ncols && pre(p + 1) <= t) p++;
    const int c = n_first + 2 * (int)p;
    const int idx = (int)(t - pre(p));
    if (2 * p + 1 >= ncols) {                    // unpaired last column
        *n = c;
        *m = c + idx;
    }
    else if (idx == 0) {
        *n = c;
        *m = c;
    }
    else {
        *m = c + 1 + (idx - 1) / 2;
        *n = c + (idx - 1) % 2;
    }
}

struct SplitUpdate {
    plasma_desc_t A;
    int pbase, pcols, kdepth, ncols, n_first, bpt, upper, ncopy, pstride;
    int ucols;                       // updated tile columns from n_first
    const float *panel;
    const __bf16 *planes;
    int single;                      // one product (bf16 inputs): hi x hi
};

// Operand chunk of one workgroup: 3 planes x 128 rows x 32 k; thread t
// holds row t / 2, k half (t % 2) * 16, two 16-byte vectors per plane.
struct Stage {
    bf16x8 v[3][2];
};

__device__ __forceinline__ void load_stage(Stage &g, const __bf16 *base,
                                           int mb, int pstride, int t,
                                           int nrows, int k0, int nq)
{
    const int row = threadIdx.x >> 1, kh = (threadIdx.x & 1) * 16;
    const int r = min(row, nrows - 1);
    const bool ok = row < nrows;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        if (q >= nq) break;
        const __bf16 *p = base + plane_off(t, q, mb, r, pstride) + k0 + kh;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            bf16x8 x = *reinterpret_cast<const bf16x8 *>(p + 8 * h);
            if (!ok) x = bf16x8{};
            g.v[q][h] = x;
        }
    }
}

__device__ __forceinline__ void store_stage(const Stage &g, __bf16 *s,
                                            int nq)
{
    const int row = threadIdx.x >> 1, kh = (threadIdx.x & 1) * 16;
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int h = 0; h < 2; h++)
            if (q < nq)
            *reinterpret_cast<bf16x8 *>(s + (q * BM + row) * LDP + kh +
                                        8 * h) = g.v[q][h];
}

__device__ __forceinline__ f32x4 mma(const bf16x8 &a, const bf16x8 &b,
                                     f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// C block (128 x 128 at (r0, c0) of tile (m, n)) -= X Y^T, X = panel rows of
// tile m, Y = panel rows of tile n, K = kdepth.  CROW: C has unit row
// stride (product formed as Y X^T so 16 consecutive rows of a C column sit
// in 16 consecutive lanes).
template <bool CROW>
__device__ __forceinline__ void run_split(const SplitUpdate &u, int tx_,
                                          int ty_, int xrows, int yrows,
                                          const View<float> &C, bool lower,
                                          int diag_off, __bf16 *sx,
                                          __bf16 *sy, int xr0, int yr0)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    auto crow = [&](int i, int q) {
        return CROW ? wr * 64 + i * 16 + lane % 16
                    : wr * 64 + i * 16 + 4 * (lane / 16) + q;
    };
    auto ccol = [&](int j, int q) {
        return CROW ? wc * 64 + j * 16 + 4 * (lane / 16) + q
                    : wc * 64 + j * 16 + lane % 16;
    };
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = crow(i, q), c = ccol(j, q);
                acc[i][j][q] = (r < xrows && c < yrows) ? -C.at(r, c) : 0.0f;
            }

    // the panel rows of the block: tile tx_ rows xr0.., tile ty_ rows yr0..
    const __bf16 *xb = u.planes + (size_t)xr0 * u.pstride;
    const __bf16 *yb = u.planes + (size_t)yr0 * u.pstride;
    const int nk = u.kdepth / KC;
    const int nq = u.single ? 1 : 3;     // planes read
    Stage gx, gy;
    load_stage(gx, xb, u.A.mb, u.pstride, tx_, xrows, 0, nq);
    load_stage(gy, yb, u.A.mb, u.pstride, ty_, yrows, 0, nq);
    for (int kc = 0; kc < nk; kc++) {
        __syncthreads();                 // the previous chunk's reads done
        store_stage(gx, sx, nq);
        store_stage(gy, sy, nq);
        __syncthreads();
        if (kc + 1 < nk) {
            load_stage(gx, xb, u.A.mb, u.pstride, tx_, xrows, (kc + 1) * KC,
                       nq);
            load_stage(gy, yb, u.A.mb, u.pstride, ty_, yrows, (kc + 1) * KC,
                       nq);
        }
        const int ko = 8 * (lane / 16);
        bf16x8 a[4][3], b[4][3];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int q = 0; q < 3; q++)
                a[i][q] = q < nq ? *reinterpret_cast<const bf16x8 *>(
                    sx + (q * BM + wr * 64 + i * 16 + lane % 16) * LDP + ko)
                                 : bf16x8{};
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 3; q++)
                b[j][q] = q < nq ? *reinterpret_cast<const bf16x8 *>(
                    sy + (q * BM + wc * 64 + j * 16 + lane % 16) * LDP + ko)
                                 : bf16x8{};
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // first operand: rows of D (C columns for CROW)
                const bf16x8 *f = CROW ? b[j] : a[i];
                const bf16x8 *s = CROW ? a[i] : b[j];
                f32x4 v = acc[i][j];
                if (!u.single) {
                    v = mma(f[2], s[0], v);   // lo * hi
                    v = mma(f[0], s[2], v);   // hi * lo
                    v = mma(f[1], s[1], v);   // mid * mid
                    v = mma(f[1], s[0], v);   // mid * hi
                    v = mma(f[0], s[1], v);   // hi * mid
                }
                v = mma(f[0], s[0], v);       // hi * hi
                acc[i][j] = v;
            }
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = crow(i, q), c = ccol(j, q);
                if (r < xrows && c < yrows && (!lower || c + diag_off <= r))
                    C.at(r, c) = -acc[i][j][q];
            }
    __syncthreads();                     // LDS free for the next job
}

template <bool CROW>
__global__ __launch_bounds__(NT, 2) void split_update_kernel(
    SplitUpdate u, int njobs, const int *status)
{
    __shared__ __attribute__((aligned(16))) __bf16 sx[3 * BM * LDP];
    __shared__ __attribute__((aligned(16))) __bf16 sy[3 * BM * LDP];
    if (block_should_stop(status)) return;
    const int padded = tg::xcd_grid(njobs), chunk = padded / tg::NXCD;
    for (int b = blockIdx.x; b < padded; b += gridDim.x) {
        int jb = (b % tg::NXCD) * chunk + b / tg::NXCD;
        if (jb >= njobs) continue;
        if (jb < u.ncopy) {
            // copy job: tile column pbase + jc, tile m > pbase + jc, from
            // the FP32 panel back into A
            int jc = 0, bb = jb;
            while (jc + 1 < u.ncols && bb >= u.A.mt - (u.pbase + jc) - 1) {
                bb -= u.A.mt - (u.pbase + jc) - 1;
                jc++;
            }
            const int col = u.pbase + jc, m = col + 1 + bb;
            const int mv = plasma_tile_mview(u.A, m);
            const int nb = plasma_tile_mview(u.A, col);
            const View<float> D = lower_view<float>(u.A, u.upper != 0, m, col);
            const float *S = u.panel + (size_t)(m - u.pbase - 1) * u.A.mb *
                                           u.pcols + (size_t)jc * u.A.nb * u.A.mb;
            if (!u.upper) {
                const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
                for (int c = wave; c < nb; c += NT / 64)
                    for (int r = lane; r < mv; r += 64)
                        D.at(r, c) = S[(size_t)c * u.A.mb + r];
            }
            else {
                for (int e = threadIdx.x; e < mv * nb; e += NT) {
                    const int r = e / nb, c = e % nb;
                    D.at(r, c) = S[(size_t)c * u.A.mb + r];
                }
            }
            continue;
        }
        jb -= u.ncopy;
        const int per_tile = u.bpt * u.bpt;
        const int sub_ = jb % per_tile;
        const int bi = sub_ / u.bpt, bj = sub_ % u.bpt;
        int m, n;
        pair_decode(jb / per_tile, u.A.mt, u.n_first, u.ucols, &m, &n);
        if (m == n && bj > bi) continue;
        const int mv = plasma_tile_mview(u.A, m), nv = plasma_tile_mview(u.A, n);
        const int r0 = bi * BM, c0 = bj * BM;
        if (r0 >= mv || c0 >= nv) continue;
        View<float> C = lower_view<float>(u.A, u.upper != 0, m, n);
        C.p += (ptrdiff_t)r0 * C.rs + (ptrdiff_t)c0 * C.cs;
        run_split<CROW>(u, m - u.pbase - 1, n - u.pbase - 1, mv - r0, nv - c0,
                        C, m == n, c0 - r0, sx, sy, r0, c0);
    }
}

}  // namespace

extern "C" size_t plasma_split_pstride(int pcols)
{
    return (size_t)((pcols + 7) / 8 * 8);
}

extern "C" size_t plasma_split_bytes(plasma_desc_t A, int pcols)
{
    return (size_t)(A.mt > 1 ? A.mt - 1 : 1) * 3 * A.mb *
           plasma_split_pstride(pcols) * sizeof(__bf16);
}

extern "C" int plasma_kern_split_panel(plasma_desc_t A, int k, int pbase,
                                       int pcols, const void *panel,
                                       void *planes, const int *status,
                                       hipStream_t s)
{
    if (A.precision != PlasmaRealFloat) return PlasmaErrorNotSupported;
    const int t0 = k - pbase, ntiles = A.mt - k - 1;
    if (ntiles <= 0) return 0;
    const int ncol = plasma_tile_mview(A, k);
    dim3 grid(ntiles, (A.mb + 31) / 32, (ncol + 31) / 32);
    hipLaunchKernelGGL(split_panel_kernel, grid, dim3(NT), 0, s, A, pbase,
                       t0, pcols, (k - pbase) * A.nb, ncol,
                       (const float *)panel, (__bf16 *)planes,
                       (int)plasma_split_pstride(pcols), status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_trailing_update_split(
    plasma_enum_t uplo, plasma_desc_t A, int pbase, int kdepth, int n_first,
    int n_end, const void *panel, const void *planes, int pcols,
    int copy_cols, int max_wgs, int single, const int *status, hipStream_t s)
{
    if (A.precision != PlasmaRealFloat || kdepth % KC != 0)
        return PlasmaErrorNotSupported;
    n_end = min(n_end, A.mt);
    const int bpt = (A.mb + BM - 1) / BM;
    long ntiles = 0;
    for (int n = n_first; n < n_end; n++) ntiles += A.mt - n;
    int ncopy = 0;
    for (int jc = 0; jc < copy_cols; jc++) ncopy += A.mt - (pbase + jc) - 1;
    const long njobs = ntiles * bpt * bpt + ncopy;
    if (njobs <= 0) return 0;
    if (njobs > 0x7fffffffL) return PlasmaErrorNotSupported;
    SplitUpdate u{A, pbase, pcols, kdepth, copy_cols > 0 ? copy_cols : 1,
                  n_first, bpt, uplo == PlasmaUpper, ncopy,
                  (int)plasma_split_pstride(pcols), max(n_end - n_first, 0),
                  (const float *)panel, (const __bf16 *)planes, single};
    const long padded = tg::xcd_grid((int)njobs);
    const long grid = max_wgs > 0 ? min(padded, (long)max_wgs) : padded;
    if (uplo == PlasmaUpper)
        hipLaunchKernelGGL(split_update_kernel<false>, dim3((unsigned)grid),
                           dim3(NT), 0, s, u, (int)njobs, status);
    else
        hipLaunchKernelGGL(split_update_kernel<true>, dim3((unsigned)grid),
                           dim3(NT), 0, s, u, (int)njobs, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}
