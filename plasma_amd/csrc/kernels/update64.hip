// update64.hip -- the FP64 trailing update of the tile Cholesky for gfx950
// with operands loaded straight into LDS.
//
// C(m,n) -= P(m) P(n)^T for the tiles of a column range (reference
// compute/pzpotrf.c:58-78, the syrk and gemm tasks of one step, here for
// the g panels of a column group at once, K = g nb).  Same jobs as the
// staged FP64 update (kernels/potrf.hip, UpdateDecode): the group's panels
// copied back into A first, then 128 x 128 blocks of the tiles (m, n),
// column by column, workgroup b running job (b % 8) * chunk + b / 8.
//
// Operands.  The panel buffer holds P(m) column major (ld mb).  A fragment
// is 16 rows x 8 k values (1 KB); one global_load_lds_dwordx4 moves it:
// lane l reads rows 2 (l % 8), 2 (l % 8) + 1 of column l / 8 (16 contiguous
// bytes; 8 lanes cover a column's 128 bytes) and the hardware writes it to
// LDS at 16 l, i.e. the fragment lands column major, element (r, k) at
// double 16 k + r.  v_mfma_f64_16x16x4_f64 wants, for k values k0..k0+3 of
// a fragment, row l % 16 and k value k0 + l / 16 in lane l: double
// 16 k0 + l, a lane-linear ds_read_b64 (no bank conflicts).  Per 16-deep K
// chunk a workgroup loads 8 row blocks x 2 k groups of X and of Y: 32
// fragments, eight per wave.  Two LDS stages (64 KB): each chunk's
// fragments are read into registers, then the next chunk's loads start
// into the other stage and run under the chunk's 64 MFMAs per wave.
//
// C is loaded into the accumulators negated and stored negated, so the
// products are only added.  mb must be a multiple of 128 (whole row blocks
// inside each panel tile); rows of a short last tile beyond its view feed
// only C rows that are not stored.
#include "common.hip.h"
#include "tilegemm.hip.h"
#include "plasma_kernels.h"

namespace {

#define LDS_AS __attribute__((address_space(3)))

constexpr int NT = 256;
constexpr int BM = 128;
constexpr int KC = 16;                  // k per chunk
constexpr int FRAG = 128;               // doubles per fragment (16 x 8)
constexpr int NF = 32;                  // fragments per chunk (X and Y)
constexpr int DW = NF / 4;              // loads per wave per chunk
constexpr int STAGE = NF * FRAG;        // doubles per LDS stage

__device__ __forceinline__ void wait_vm0()
{
    __builtin_amdgcn_s_waitcnt(0x0F70);
}
__device__ __forceinline__ void wait_lgkm0()
{
    __builtin_amdgcn_s_waitcnt(0xC07F);
}
// workgroup barrier without a memory fence (explicit waits around it); the
// empty asm statements keep the compiler from moving memory accesses across
__device__ __forceinline__ void bar()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

struct Upd64 {
    plasma_desc_t A;
    int pbase, pcols, kdepth, ncols, n_first, bpt, upper, ncopy;
    const double *panel;
};

// C block (128 x 128 at (r0, c0) of tile (m, n)) -= X Y^T with X rows r0..
// of panel tile tx and Y rows c0.. of panel tile ty.  CROW: C has unit row
// stride; the product is formed as Y X^T so 16 consecutive rows of one C
// column sit in 16 consecutive lanes.
template <bool CROW>
__device__ __forceinline__ void run64(const Upd64 &u, const double *xp,
                                      const double *yp, int xrows, int yrows,
                                      const View<double> &C, bool lower,
                                      int diag_off, double *lds)
{
    typedef Mfma<double>::acc_t acc_t;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int mb = u.A.mb;
    auto crow = [&](int i, int q) {
        return CROW ? wr * 64 + i * 16 + lane % 16
                    : wr * 64 + i * 16 + Mfma<double>::row(lane, q);
    };
    auto ccol = [&](int j, int q) {
        return CROW ? wc * 64 + j * 16 + Mfma<double>::row(lane, q)
                    : wc * 64 + j * 16 + lane % 16;
    };

    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = crow(i, q), c = ccol(j, q);
                acc[i][j][q] = (r < xrows && c < yrows) ? C.at(r, c) : 0.0;
            }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) acc[i][j][q] = -acc[i][j][q];

    // fragment f of a chunk: operand f / 16 (0: X, 1: Y), k group
    // (f / 8) % 2, row block f % 8.  Wave w loads f = 8 w .. 8 w + 7: one
    // operand and k group, the eight row blocks 128 bytes apart, so one
    // address per lane and constant offsets
    const int o = wave / 2, kg = wave % 2;
    const double *src = (o ? yp : xp) + (size_t)(kg * 8 + lane / 8) * mb +
                        (lane % 8) * 2;
    auto issue = [&](int kc, int stage) {
        const double *p = src + (size_t)kc * KC * mb;
#pragma unroll
        for (int i = 0; i < DW; i++)
            __builtin_amdgcn_global_load_lds(
                (const void *)(p + i * 16),
                (LDS_AS void *)(lds + stage * STAGE + (wave * DW + i) * FRAG),
                16, 0, 0);
    };

    const int nk = u.kdepth / KC;
    issue(0, 0);
    for (int kc = 0; kc < nk; kc++) {
        const int cur = kc & 1;
        wait_vm0();
        bar();                           // chunk kc is in LDS
        const double *s = lds + cur * STAGE + lane;
        double a[4][4], b[4][4];         // [k step][block]
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int kg = ks / 2, k0 = (ks % 2) * 4;
#pragma unroll
            for (int i = 0; i < 4; i++)
                a[ks][i] = s[(kg * 8 + wr * 4 + i) * FRAG + k0 * 16];
#pragma unroll
            for (int j = 0; j < 4; j++)
                b[ks][j] = s[(16 + kg * 8 + wc * 4 + j) * FRAG + k0 * 16];
        }
        asm volatile("" ::: "memory");
        if (kc + 1 < nk) issue(kc + 1, cur ^ 1);
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = CROW
                        ? Mfma<double>::mma(b[ks][j], a[ks][i], acc[i][j])
                        : Mfma<double>::mma(a[ks][i], b[ks][j], acc[i][j]);
    }
    wait_lgkm0();
    bar();                               // LDS free for the next job

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
}

template <bool CROW>
__global__ __launch_bounds__(NT, 2) void upd64_kernel(Upd64 u, int njobs,
                                                      const int *status)
{
    __shared__ __attribute__((aligned(1024))) double lds[2 * STAGE];
    if (block_should_stop(status)) return;
    const int padded = tg::xcd_grid(njobs), chunk = padded / tg::NXCD;
    for (int b = blockIdx.x; b < padded; b += gridDim.x) {
        int jb = (b % tg::NXCD) * chunk + b / tg::NXCD;
        if (jb >= njobs) continue;
        if (jb < u.ncopy) {
            // copy job: tile column pbase + jc, tile m > pbase + jc, from
            // the panel back into A
            int jc = 0, bb = jb;
            while (jc + 1 < u.ncols && bb >= u.A.mt - (u.pbase + jc) - 1) {
                bb -= u.A.mt - (u.pbase + jc) - 1;
                jc++;
            }
            const int col = u.pbase + jc, m = col + 1 + bb;
            const int mv = plasma_tile_mview(u.A, m);
            const int nb = plasma_tile_mview(u.A, col);
            const View<double> D = lower_view<double>(u.A, u.upper != 0, m,
                                                      col);
            const double *S = u.panel +
                              (size_t)(m - u.pbase - 1) * u.A.mb * u.pcols +
                              (size_t)jc * u.A.nb * u.A.mb;
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
        trapezoid_decode(jb / per_tile, u.A.mt, u.n_first, &m, &n);
        if (m == n && bj > bi) continue;
        const int mv = plasma_tile_mview(u.A, m), nv = plasma_tile_mview(u.A, n);
        const int r0 = bi * BM, c0 = bj * BM;
        if (r0 >= mv || c0 >= nv) continue;
        View<double> C = lower_view<double>(u.A, u.upper != 0, m, n);
        C.p += (ptrdiff_t)r0 * C.rs + (ptrdiff_t)c0 * C.cs;
        const double *xp = u.panel +
                           (size_t)(m - u.pbase - 1) * u.A.mb * u.pcols + r0;
        const double *yp = u.panel +
                           (size_t)(n - u.pbase - 1) * u.A.mb * u.pcols + c0;
        run64<CROW>(u, xp, yp, mv - r0, nv - c0, C, m == n, c0 - r0, lds);
    }
}

}  // namespace

extern "C" int plasma_update64_ok(plasma_desc_t A, int kdepth)
{
    return A.precision == PlasmaRealDouble && A.mb == A.nb &&
           A.mb % BM == 0 && kdepth % KC == 0;
}

extern "C" int plasma_kern_trailing_update64(
    plasma_enum_t uplo, plasma_desc_t A, int pbase, int kdepth, int n_first,
    int n_end, const void *panel, int pcols, int copy_cols, int max_wgs,
    const int *status, hipStream_t s)
{
    if (!plasma_update64_ok(A, kdepth) || kdepth > pcols)
        return PlasmaErrorNotSupported;
    n_end = min(n_end, A.mt);
    const int bpt = A.mb / BM;
    long ntiles = 0;
    for (int n = n_first; n < n_end; n++) ntiles += A.mt - n;
    int ncopy = 0;
    for (int jc = 0; jc < copy_cols; jc++) ncopy += max(A.mt - (pbase + jc) - 1, 0);
    const long njobs = ntiles * bpt * bpt + ncopy;
    if (njobs <= 0) return 0;
    if (njobs > 0x7fffffffL) return PlasmaErrorNotSupported;
    Upd64 u{A, pbase, pcols, kdepth, copy_cols > 0 ? copy_cols : 1, n_first,
            bpt, uplo == PlasmaUpper, ncopy, (const double *)panel};
    const long padded = tg::xcd_grid((int)njobs);
    const long grid = max_wgs > 0 ? min(padded, (long)max_wgs) : padded;
    if (uplo == PlasmaUpper)
        hipLaunchKernelGGL(upd64_kernel<false>, dim3((unsigned)grid),
                           dim3(NT), 0, s, u, (int)njobs, status);
    else
        hipLaunchKernelGGL(upd64_kernel<true>, dim3((unsigned)grid),
                           dim3(NT), 0, s, u, (int)njobs, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}
