// split16.hip -- the FP32 trailing update on fp16 MFMA for gfx950, from
// row-scaled fp16 panel terms loaded straight into LDS.
//
// Scaling.  For an SPD matrix every entry of row r of its Cholesky factor
// satisfies |L(r,j)| <= sqrt(a_rr), because sum_j L(r,j)^2 = a_rr.  With
// the power of two sigma_r = 2^-ceil(log2 sqrt(a_rr)) (from the diagonal of
// A before the factorization) every scaled entry x = sigma_r L(r,j) has
// |x| <= 1, whatever the scaling of A, so fp16's range is enough.
//
// Terms.  x is stored as hi = fp16(x) and, in the two-term form,
// lo = fp16(x - hi) (the difference is exact in FP32): 22 significant bits,
// |x - hi - lo| <= max(2^-22 |x|, 2^-25).  The update
//     C(r,c) -= sum_j L(r,j) L(c,j)
//            = (sigma_r sigma_c)^-1 sum_j x(r,j) x(c,j)
// takes the products hi*hi + hi*lo + lo*hi (each exact in FP32, summed in
// FP32 by v_mfma_f32_16x16x32_f16; the dropped lo*lo is below 2^-22 of
// |x y|): three products instead of the six of the bf16 three-term split,
// and 4 bytes of operand per element instead of 6.  C is scaled by
// sigma_r sigma_c on entry and back on exit (powers of two: exact).  The
// one-term form (hi only) is the fp16-input factor of the GMRES-IR solver.
//
// Layout of the terms ("planes"): MFMA fragments.  A fragment is 16 rows x
// 32 k values of one plane, 1 KB, stored in the order the instruction's
// lanes consume it: lane l holds row l % 16, k values 8 (l / 16) .. +7, so
// element (r, k) of the fragment is at ((r % 16) + 16 (k / 8)) * 8 + k % 8.
// Fragment (t, kb, q, rb): panel tile t, k block kb, plane q, row block rb,
// at index ((t * kbs + kb) * np + q) * rbp + rb, with rbp the row blocks of
// a tile padded to a multiple of 8 (rows past the tile are zero).
//
// Update kernel.  128 x 128 C blocks, 4 waves of 64 x 64.  Per 32-deep K
// chunk each operand needs 8 row blocks x np planes of fragments; each is
// one global_load_lds_dwordx4 (64 lanes x 16 contiguous bytes, written to
// LDS in lane order), so operands go from L2 to LDS without passing
// through registers, and every LDS read is a lane-linear ds_read_b128 (no
// bank conflicts).  Two LDS stages: each chunk's fragments are read into
// registers, then the loads of the next chunk start into the other stage
// and run under the chunk's products; one barrier per chunk.  The waits
// are explicit around bare s_barrier instructions (a __syncthreads()
// fence would also wait for stores in flight).  C is scaled by exact
// reciprocals of the row scales (no division); an interior block of the
// lower triangle is read and written through LDS with 16-byte loads and
// non-temporal stores (a quarter of the memory instructions of the
// accumulator layout, whose lanes hold four columns each).
#include "common.hip.h"
#include "tilegemm.hip.h"
#include "plasma_kernels.h"

namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
#define LDS_AS __attribute__((address_space(3)))

constexpr int NT = 256;
constexpr int BM = 128;
constexpr int KC = 32;
constexpr int FRAG = 512;           // fp16 elements per fragment
constexpr int SST = 68;             // column stride of the C staging area
constexpr int STAGE_BYTES = 4 * 64 * SST * 4;   // four waves' 64 x 64 blocks

struct Geom {
    int rbp;                         // row blocks per tile (multiple of 8)
    int kbs;                         // k blocks per panel buffer row
    int np;                          // planes
};

__device__ __forceinline__ size_t frag_index(const Geom &g, int t, int kb,
                                             int q, int rb)
{
    return (((size_t)t * g.kbs + kb) * g.np + q) * g.rbp + rb;
}

// s_waitcnt immediates (gfx9 encoding: vmcnt in [3:0] and [15:14],
// expcnt in [6:4], lgkmcnt in [11:8]; an all-ones field does not wait)
template <int N>
__device__ __forceinline__ void wait_vm()
{
    static_assert(N >= 0 && N < 16, "vmcnt");
    __builtin_amdgcn_s_waitcnt(0x0F70 | N);
}
__device__ __forceinline__ void wait_lgkm0()
{
    __builtin_amdgcn_s_waitcnt(0xC07F);
}
// workgroup barrier without a memory fence; the empty asm statements keep
// the compiler from moving memory accesses across it
__device__ __forceinline__ void bar()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// ---------------------------------------------------------------------------
// Row scales sigma_r from the diagonal of the lower-equivalent tiles; rows
// past A.m (tile padding) get 1.  A diagonal that is not positive (the
// factorization will fail there) also gets 1.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void split16_scale_kernel(plasma_desc_t A,
                                                           int upper,
                                                           float *rscale,
                                                           int nrows,
                                                           const int *status)
{
    if (block_should_stop(status)) return;
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= nrows) return;
    float s = 1.0f;
    if (i < A.m) {
        const int k = i / A.mb, r = i % A.mb;
        const View<float> D = lower_view<float>(A, upper != 0, k, k);
        const float d = D.at(r, r);
        if (d > 0.0f && d <= 3.4e38f) {
            int e;
            const float f = frexpf(sqrtf(d), &e);   // sqrt(d) = f 2^e
            if (f == 0.5f) e -= 1;                   // sqrt(d) <= 2^e
            s = ldexpf(1.0f, -e);
        }
    }
    rscale[i] = s;
}

// ---------------------------------------------------------------------------
// The panel columns [col0, col0 + ncol) of panel tiles t0 + blockIdx.x into
// np fp16 planes.  Workgroup (t, y, z): row blocks 8y .. 8y+7, k block
// col0 / 32 + z; each wave writes whole fragments (one 16-byte store per
// lane and plane).
// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(NT) void split16_panel_kernel(
    plasma_desc_t A, int pbase, int t0, int pcols, int col0, int ncol,
    const float *panel, _Float16 *planes, Geom g, const float *rscale,
    const int *status)
{
    if (block_should_stop(status)) return;
    const int t = t0 + blockIdx.x, m = pbase + 1 + t;
    const int mv = plasma_tile_mview(A, m);
    const int kb = col0 / KC + blockIdx.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float *src = panel + (size_t)t * A.mb * pcols;
    const int c0 = kb * KC + 8 * (lane >> 4);       // panel-buffer column
    for (int rb = blockIdx.y * 8 + wave; rb < blockIdx.y * 8 + 8; rb += 4) {
        const int r = rb * 16 + (lane & 15);
        const bool ok = r < mv;
        const float sc = ok ? rscale[(size_t)m * A.mb + r] : 0.0f;
        f16x8 h, l;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int c = c0 + e;
            const float x = (ok && c - col0 < ncol)
                ? src[(size_t)c * A.mb + r] * sc : 0.0f;
            h[e] = (_Float16)x;
            l[e] = (_Float16)(x - (float)h[e]);
        }
        _Float16 *d = planes + frag_index(g, t, kb, 0, rb) * FRAG + lane * 8;
        *reinterpret_cast<f16x8 *>(d) = h;
        if (NP > 1)
            *reinterpret_cast<f16x8 *>(d + (size_t)g.rbp * FRAG) = l;
    }
}

// ---------------------------------------------------------------------------
// Trailing update.
// ---------------------------------------------------------------------------
// Tile t of the trapezoid {(m, n): n in [n_first, n_first + ncols), m >= n}
// in column-pair order: within a pair (c, c+1) the tiles go (c, c),
// (c+1, c), (c+1, c+1), (c+2, c), (c+2, c+1), ... so a panel row block
// serves both columns back to back.  With T = mt - n_first, pair p holds
// 2T - 4p - 1 tiles (a last, unpaired column T - 2p), and
// P(p) = p (2T + 1 - 2p) tiles precede it.
__device__ __forceinline__ void pair_decode16(int t, int mt, int n_first,
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
    if (2 * p + 1 >= ncols) {
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

struct Update16 {
    plasma_desc_t A;
    int pbase, pcols, kdepth, ncols, n_first, bpt, upper, ncopy, ucols;
    const float *panel;
    const _Float16 *planes;
    Geom g;
    const float *rscale;
};

__device__ __forceinline__ f32x4 mma(const f16x8 &a, const f16x8 &b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// C block (128 x 128 at (r0, c0) of tile (m, n)) -= X Y^T.  X: row blocks
// rbx0.. of panel tile tx, Y: row blocks rby0.. of panel tile ty.  CROW: C
// has unit row stride (product formed as Y X^T, so 16 consecutive rows of
// a C column sit in 16 consecutive lanes).
template <bool CROW, int NP>
__device__ __forceinline__ void run16(const Update16 &u, int tx, int ty,
                                      int rbx0, int rby0, int xrows, int yrows,
                                      const View<float> &C, bool lower,
                                      int diag_off, const float *sx_,
                                      const float *sy_, _Float16 *lds)
{
    constexpr int NF = 2 * NP * 8;      // fragments per chunk
    constexpr int DW = NF / 4;          // loads per wave per chunk
    constexpr int STAGE = NF * FRAG;    // fp16 elements per LDS stage
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

    // C block and the row / column scales of this lane's elements.  The
    // accumulator layout gives each lane four C columns.  An interior block
    // of the lower triangle goes through LDS: each wave reads its 64 x 64
    // part as 16-byte row quads (four columns of 256 contiguous bytes per
    // load) into its own region (column stride SST), then its elements in
    // the accumulator layout; the operand stages reuse that LDS after a
    // barrier.  Other blocks load element by element (masked, no
    // arithmetic between the loads, so they issue back to back).  The
    // scaling follows once they are all in.
    const bool staged = CROW && !lower && xrows >= BM && yrows >= BM &&
                        C.rs == 1 && (C.cs & 3) == 0 &&
                        (((uintptr_t)C.p) & 15) == 0;
    f32x4 acc[4][4];
    if (staged) {
        float *w = reinterpret_cast<float *>(lds) + wave * 64 * SST;
        const float *cb = C.p + wr * 64 + (ptrdiff_t)(wc * 64) * C.cs;
        f32x4 quad[16];
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int cc = 4 * t + lane / 16, rr = 4 * (lane % 16);
            quad[t] = *reinterpret_cast<const f32x4 *>(cb +
                                                       (ptrdiff_t)cc * C.cs +
                                                       rr);
        }
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int cc = 4 * t + lane / 16, rr = 4 * (lane % 16);
            *reinterpret_cast<f32x4 *>(w + cc * SST + rr) = quad[t];
        }
        wait_lgkm0();   // the region is the wave's own: no barrier
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int q = 0; q < 4; q++)
                    acc[i][j][q] = w[(j * 16 + 4 * (lane / 16) + q) * SST +
                                     i * 16 + lane % 16];
        wait_lgkm0();
        bar();          // every wave has read its region: stages reusable
    }
    else {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int r = crow(i, q), c = ccol(j, q);
                    acc[i][j][q] = (r < xrows && c < yrows) ? C.at(r, c)
                                                            : 0.0f;
                }
    }
    // scale of C element (i, j, q): sx_[crow] * sy_[ccol], indices clamped
    // to the valid rows (the clamped values are never used)
    float fx[4][4], fy[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int q = 0; q < 4; q++)
            fx[i][q] = sx_[min(crow(i, q), xrows - 1)];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int q = 0; q < 4; q++)
            fy[j][q] = sy_[min(ccol(j, q), yrows - 1)];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                acc[i][j][q] *= -(fx[i][q] * fy[j][q]);

    // fragment f of a chunk: operand f / (8 NP) (0: X, 1: Y), plane
    // (f / 8) % NP, row block f % 8; wave w loads f = w DW .. w DW + DW-1
    const _Float16 *src[DW];
#pragma unroll
    for (int i = 0; i < DW; i++) {
        const int f = wave * DW + i;
        const int o = f / (8 * NP), q = (f / 8) % NP, rb = f % 8;
        src[i] = u.planes +
                 frag_index(u.g, o ? ty : tx, 0, q, (o ? rby0 : rbx0) + rb) *
                     FRAG + lane * 8;
    }
    const size_t kstep = (size_t)u.g.np * u.g.rbp * FRAG;   // one k block
    auto issue = [&](int kc, int stage) {
#pragma unroll
        for (int i = 0; i < DW; i++) {
            const int f = wave * DW + i;
            __builtin_amdgcn_global_load_lds(
                (const void *)(src[i] + (size_t)kc * kstep),
                (LDS_AS void *)(lds + stage * STAGE + f * FRAG), 16, 0, 0);
        }
    };

    // Chunk kc: wait for its loads, barrier (every wave's loads of kc are
    // in LDS, and every wave has read stage (kc+1) & 1 in step kc-1), read
    // its fragments into registers, start the loads of kc+1 into the other
    // stage, multiply.  The loads of kc+1 run under the products of kc.
    const int nk = u.kdepth / KC;
    issue(0, 0);
    for (int kc = 0; kc < nk; kc++) {
        const int cur = kc & 1;
        wait_vm<0>();
        bar();
        const _Float16 *s = lds + cur * STAGE + lane * 8;
        f16x8 a[4][NP], b[4][NP];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int q = 0; q < NP; q++)
                a[i][q] = *reinterpret_cast<const f16x8 *>(
                    s + (q * 8 + wr * 4 + i) * FRAG);
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < NP; q++)
                b[j][q] = *reinterpret_cast<const f16x8 *>(
                    s + (8 * NP + q * 8 + wc * 4 + j) * FRAG);
        asm volatile("" ::: "memory");
        if (kc + 1 < nk) issue(kc + 1, cur ^ 1);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // first operand: rows of D (C columns for CROW)
                const f16x8 *f = CROW ? b[j] : a[i];
                const f16x8 *g = CROW ? a[i] : b[j];
                f32x4 v = acc[i][j];
                if (NP > 1) {
                    v = mma(f[1], g[0], v);       // lo * hi
                    v = mma(f[0], g[1], v);       // hi * lo
                }
                v = mma(f[0], g[0], v);           // hi * hi
                acc[i][j] = v;
            }
    }
    // every wave's last reads of the stages are done before LDS is reused
    wait_lgkm0();
    bar();
    // exact reciprocals of the power-of-two scales: 2^k = 0.5 * 2^(k+1)
    // has frexp exponent k + 1
    float ix[4][4], iy[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            ix[i][q] = __builtin_ldexpf(
                1.0f, 1 - __builtin_amdgcn_frexp_expf(fx[i][q]));
            iy[i][q] = __builtin_ldexpf(
                1.0f, 1 - __builtin_amdgcn_frexp_expf(fy[i][q]));
        }
    if (staged) {
        // Interior block: each wave stages its 64 x 64 result in its own
        // LDS region (column stride SST floats) and writes 16-byte row
        // quads, four columns of 256 contiguous bytes per store, marked
        // non-temporal (C is not read again in this launch).
        float *w = reinterpret_cast<float *>(lds) + wave * 64 * SST;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int q = 0; q < 4; q++)
                    w[(j * 16 + 4 * (lane / 16) + q) * SST + i * 16 +
                      lane % 16] = -acc[i][j][q] * ix[i][q] * iy[j][q];
        wait_lgkm0();   // the region is the wave's own: no barrier
        float *cb = C.p + wr * 64 + (ptrdiff_t)(wc * 64) * C.cs;
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int cc = 4 * t + lane / 16, rr = 4 * (lane % 16);
            const f32x4 v = *reinterpret_cast<const f32x4 *>(w + cc * SST +
                                                             rr);
            __builtin_nontemporal_store(
                v, reinterpret_cast<f32x4 *>(cb + (ptrdiff_t)cc * C.cs + rr));
        }
        // the next job's operand loads write LDS
        bar();
        return;
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = crow(i, q), c = ccol(j, q);
                if (r < xrows && c < yrows && (!lower || c + diag_off <= r))
                    C.at(r, c) = -acc[i][j][q] * ix[i][q] * iy[j][q];
            }
}

template <bool CROW, int NP>
__global__ __launch_bounds__(NT, 2) void split16_update_kernel(
    Update16 u, int njobs, const int *status)
{
    // two operand stages, reused as the C staging area after the products
    constexpr int OPS = 2 * 2 * NP * 8 * FRAG, STG = STAGE_BYTES / 2;
    __shared__ __attribute__((aligned(1024))) _Float16 lds[OPS > STG ? OPS
                                                                     : STG];
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
        const int sub = jb % per_tile;
        const int bi = sub / u.bpt, bj = sub % u.bpt;
        int m, n;
        pair_decode16(jb / per_tile, u.A.mt, u.n_first, u.ucols, &m, &n);
        if (m == n && bj > bi) continue;
        const int mv = plasma_tile_mview(u.A, m), nv = plasma_tile_mview(u.A, n);
        const int r0 = bi * BM, c0 = bj * BM;
        if (r0 >= mv || c0 >= nv) continue;
        View<float> C = lower_view<float>(u.A, u.upper != 0, m, n);
        C.p += (ptrdiff_t)r0 * C.rs + (ptrdiff_t)c0 * C.cs;
        run16<CROW, NP>(u, m - u.pbase - 1, n - u.pbase - 1, r0 / 16,
                        c0 / 16, mv - r0, nv - c0, C, m == n, c0 - r0,
                        u.rscale + (size_t)m * u.A.mb + r0,
                        u.rscale + (size_t)n * u.A.mb + c0, lds);
    }
}

Geom geom_of(plasma_desc_t A, int pcols, int np)
{
    Geom g;
    g.rbp = (A.mb + BM - 1) / BM * (BM / 16);
    g.kbs = pcols / KC;
    g.np = np;
    return g;
}

}  // namespace

extern "C" size_t plasma_split16_bytes(plasma_desc_t A, int pcols, int np)
{
    const Geom g = geom_of(A, pcols, np);
    return (size_t)(A.mt > 1 ? A.mt - 1 : 1) * g.kbs * g.np * g.rbp * FRAG *
           sizeof(_Float16);
}

extern "C" size_t plasma_split16_scale_elems(plasma_desc_t A)
{
    return (size_t)A.mt * A.mb;
}

extern "C" int plasma_kern_split16_scale(plasma_enum_t uplo, plasma_desc_t A,
                                         float *rscale, const int *status,
                                         hipStream_t s)
{
    if (A.precision != PlasmaRealFloat) return PlasmaErrorNotSupported;
    const int nrows = (int)plasma_split16_scale_elems(A);
    if (nrows <= 0) return 0;
    hipLaunchKernelGGL(split16_scale_kernel, dim3((nrows + NT - 1) / NT),
                       dim3(NT), 0, s, A, uplo == PlasmaUpper, rscale, nrows,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_split16_panel(plasma_desc_t A, int k, int pbase,
                                         int pcols, const void *panel,
                                         void *planes, int np,
                                         const float *rscale,
                                         const int *status, hipStream_t s)
{
    if (A.precision != PlasmaRealFloat || A.nb % KC != 0 || pcols % KC != 0 ||
        (np != 1 && np != 2))
        return PlasmaErrorNotSupported;
    const int t0 = k - pbase, ntiles = A.mt - k - 1;
    if (ntiles <= 0) return 0;
    const int ncol = plasma_tile_mview(A, k);
    const Geom g = geom_of(A, pcols, np);
    dim3 grid(ntiles, g.rbp / 8, (ncol + KC - 1) / KC);
    const int col0 = (k - pbase) * A.nb;
    if (np == 2)
        hipLaunchKernelGGL(split16_panel_kernel<2>, grid, dim3(NT), 0, s, A,
                           pbase, t0, pcols, col0, ncol, (const float *)panel,
                           (_Float16 *)planes, g, rscale, status);
    else
        hipLaunchKernelGGL(split16_panel_kernel<1>, grid, dim3(NT), 0, s, A,
                           pbase, t0, pcols, col0, ncol, (const float *)panel,
                           (_Float16 *)planes, g, rscale, status);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

extern "C" int plasma_kern_trailing_update_split16(
    plasma_enum_t uplo, plasma_desc_t A, int pbase, int kdepth, int n_first,
    int n_end, const void *panel, const void *planes, int pcols,
    int copy_cols, int max_wgs, int np, const float *rscale,
    const int *status, hipStream_t s)
{
    if (A.precision != PlasmaRealFloat || kdepth % KC != 0 ||
        pcols % KC != 0 || kdepth > pcols || (np != 1 && np != 2))
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
    Update16 u{A, pbase, pcols, kdepth, copy_cols > 0 ? copy_cols : 1,
               n_first, bpt, uplo == PlasmaUpper, ncopy,
               max(n_end - n_first, 0), (const float *)panel,
               (const _Float16 *)planes, geom_of(A, pcols, np), rscale};
    const long padded = tg::xcd_grid((int)njobs);
    const long grid = max_wgs > 0 ? min(padded, (long)max_wgs) : padded;
    const dim3 gd((unsigned)grid), bd(NT);
    if (np == 2) {
        if (uplo == PlasmaUpper)
            hipLaunchKernelGGL((split16_update_kernel<false, 2>), gd, bd, 0, s,
                               u, (int)njobs, status);
        else
            hipLaunchKernelGGL((split16_update_kernel<true, 2>), gd, bd, 0, s,
                               u, (int)njobs, status);
    }
    else {
        if (uplo == PlasmaUpper)
            hipLaunchKernelGGL((split16_update_kernel<false, 1>), gd, bd, 0, s,
                               u, (int)njobs, status);
        else
            hipLaunchKernelGGL((split16_update_kernel<true, 1>), gd, bd, 0, s,
                               u, (int)njobs, status);
    }
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}
