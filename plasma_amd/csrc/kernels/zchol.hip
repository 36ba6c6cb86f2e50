// zchol.hip -- the complex (Hermitian) Cholesky path on gfx950: the
// factorization of plasma_{z,c}potrf, the sweeps of plasma_{z,c}potrs, and
// the Hermitian residual and norms of the mixed solver plasma_zcposv.
//
// Reference: compute/pzpotrf.c:27-127 with core_zpotrf / ztrsm / zherk /
// zgemm, compute/pztrsm.c:43-177 (Left, ConjTrans / NoTrans), pzhemm.c,
// pzlanhe.c:83-131 and core_dzamax.c.
//
// Elements are interleaved (re, im) pairs, the reference's
// plasma_complex64_t / plasma_complex32_t.  gfx950 has no complex matrix
// instruction, so a complex 16x16x4 block product is four real
// v_mfma_f64_16x16x4_f64 (or _f32) products on one lane layout,
//     Re += Xr Yr - Xi Yi,   Im += Xr Yi + Xi Yr,
// with every lane loading one whole complex element per operand slot.  A
// complex product carries four times the real one's arithmetic per operand
// byte, and an FP64 MFMA takes 64 cycles, so the operands come straight from
// memory (L2) into registers one 4-deep step ahead, without LDS.
//
// Lower-equivalent convention: for uplo == PlasmaUpper (A = U^H U) the
// kernels work on L = U^H, reading and writing the stored upper tiles
// conjugate-transposed, so each algorithm is written once, for L.
#include "common.hip.h"
#include "plasma_kernels.h"

namespace {

constexpr int NT = 256;
constexpr int BS = 32;       // diagonal block of the tile factorization
constexpr int BLD = BS + 1;  // padded LDS row

template <typename R> struct alignas(2 * sizeof(R)) cx {
    R re, im;
};

template <typename R> __device__ __forceinline__ cx<R> czero()
{
    return cx<R>{R(0), R(0)};
}

// a * b and a * conj(b)
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b)
{
    return cx<R>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__device__ __forceinline__ cx<R> cmulc(cx<R> a, cx<R> b)
{
    return cx<R>{a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}

// A strided complex window: element (r, c) at p[r rs + c cs], conjugated
// on read and write when `conj`.  With `herm` the window is the Hermitian
// matrix whose lower triangle (r >= c) is stored: r < c reads the conjugate
// of (c, r) and the diagonal reads as real (LAPACK ignores the imaginary
// part of a Hermitian diagonal).
template <typename R> struct CView {
    cx<R> *p;
    int rs, cs;
    int conj;
    int herm;

    __device__ __forceinline__ cx<R> raw(int r, int c) const
    {
        cx<R> v = p[(ptrdiff_t)r * rs + (ptrdiff_t)c * cs];
        if (conj) v.im = -v.im;
        return v;
    }
    __device__ __forceinline__ cx<R> get(int r, int c) const
    {
        if (herm) {
            if (r == c) {
                cx<R> v = raw(r, r);
                v.im = R(0);
                return v;
            }
            if (r < c) {
                cx<R> v = raw(c, r);
                v.im = -v.im;
                return v;
            }
        }
        return raw(r, c);
    }
    __device__ __forceinline__ void put(int r, int c, cx<R> v) const
    {
        if (conj) v.im = -v.im;
        p[(ptrdiff_t)r * rs + (ptrdiff_t)c * cs] = v;
    }
    // the window moved to start at (r, c) (plain windows only)
    __device__ __forceinline__ CView at(int r, int c) const
    {
        CView v = *this;
        v.p = p + (ptrdiff_t)r * rs + (ptrdiff_t)c * cs;
        return v;
    }
};

// Conjugate transpose of a window (a Hermitian window is its own).
template <typename R>
__device__ __forceinline__ CView<R> herm_of(CView<R> v)
{
    if (v.herm) return v;
    CView<R> h = v;
    h.rs = v.cs;
    h.cs = v.rs;
    h.conj = !v.conj;
    return h;
}

// Tile (m, n), m >= n, of the lower-equivalent factor L.
template <typename R>
__device__ __forceinline__ CView<R> lower_cview(const plasma_desc_t &A,
                                                bool upper, int m, int n)
{
    CView<R> v;
    v.herm = 0;
    if (!upper) {
        v.p = (cx<R> *)plasma_tile_addr(A, m, n);
        v.rs = 1;
        v.cs = plasma_tile_mmain(A, m);
        v.conj = 0;
    }
    else {
        v.p = (cx<R> *)plasma_tile_addr(A, n, m);
        v.rs = plasma_tile_mmain(A, n);
        v.cs = 1;
        v.conj = 1;
    }
    return v;
}

template <typename R>
__device__ __forceinline__ CView<R> tile_cview(const plasma_desc_t &A, int m,
                                               int n)
{
    CView<R> v;
    v.p = (cx<R> *)plasma_tile_addr(A, m, n);
    v.rs = 1;
    v.cs = plasma_tile_mmain(A, m);
    v.conj = 0;
    v.herm = 0;
    return v;
}

// A column-major buffer with leading dimension ld.
template <typename R>
__device__ __forceinline__ CView<R> buf_cview(cx<R> *p, int ld)
{
    CView<R> v;
    v.p = p;
    v.rs = 1;
    v.cs = ld;
    v.conj = 0;
    v.herm = 0;
    return v;
}

// ---------------------------------------------------------------------------
// Wave-level complex block product: MI x NJ blocks of 16 x 16.
// ---------------------------------------------------------------------------
template <typename R, int MI, int NJ> struct CAcc {
    typedef typename Mfma<R>::acc_t acc_t;
    acc_t re[MI][NJ], im[MI][NJ];

    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int i = 0; i < MI; i++)
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                re[i][j] = acc_t{0, 0, 0, 0};
                im[i][j] = acc_t{0, 0, 0, 0};
            }
    }
};

// acc += X(xr0 + r, k) Y(k, yc0 + c) over k in [0, K), r < 16 MI, c < 16 NJ;
// rows r >= xrows and columns c >= ycols read as zero.  Lane l loads
// X(16 i + l % 16, k0 + l / 16) and Y(k0 + l / 16, 16 j + l % 16): the MFMA
// operand layout of A[l % 16][l / 16] and B[l / 16][l % 16].  The operands
// of PF steps are in flight at once (a ring of PF register slots, the loop
// unrolled by PF so every slot index is a constant): one 4-deep step is
// only 4 MI NJ products, shorter than an L2 round trip.  FP64 (64 cycles
// per product) gains from three steps in flight: zposv n = 16384 239 ->
// 207 ms; the FP32 update did not (cposv 135 -> 140 ms) and keeps one.
template <typename R, int MI, int NJ, int PF = (sizeof(R) == 8 ? 3 : 1)>
__device__ __forceinline__ void cprod(CAcc<R, MI, NJ> &acc,
                                      const CView<R> &X, int xr0, int xrows,
                                      const CView<R> &Y, int yc0, int ycols,
                                      int K)
{
    if (K <= 0) return;
    const int lane = threadIdx.x & 63;
    const int l16 = lane & 15, lk = lane >> 4;
    cx<R> a[PF][MI], b[PF][NJ];
    // operands of the step at k0 (zeros past K: nothing is loaded there)
    auto fetch = [&](cx<R>(&fa)[MI], cx<R>(&fb)[NJ], int k0) {
        const int k = k0 + lk;
        const bool kin = k < K;
#pragma unroll
        for (int i = 0; i < MI; i++) {
            const int r = 16 * i + l16;
            fa[i] = (kin && r < xrows) ? X.get(xr0 + r, k) : czero<R>();
        }
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int c = 16 * j + l16;
            fb[j] = (kin && c < ycols) ? Y.get(k, yc0 + c) : czero<R>();
        }
    };
#pragma unroll
    for (int s = 0; s < PF; s++) fetch(a[s], b[s], 4 * s);
    for (int k0 = 0; k0 < K; k0 += 4 * PF) {
#pragma unroll
        for (int s = 0; s < PF; s++) {
            if (k0 + 4 * s >= K) break;
            // two passes over the blocks, so consecutive products go to
            // different accumulators (a product waits for the previous one
            // on its accumulator)
#pragma unroll
            for (int i = 0; i < MI; i++)
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    acc.re[i][j] =
                        Mfma<R>::mma(a[s][i].re, b[s][j].re, acc.re[i][j]);
                    acc.im[i][j] =
                        Mfma<R>::mma(a[s][i].re, b[s][j].im, acc.im[i][j]);
                }
#pragma unroll
            for (int i = 0; i < MI; i++)
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    acc.re[i][j] =
                        Mfma<R>::mma(-a[s][i].im, b[s][j].im, acc.re[i][j]);
                    acc.im[i][j] =
                        Mfma<R>::mma(a[s][i].im, b[s][j].re, acc.im[i][j]);
                }
            // the slot takes the step PF ahead
            fetch(a[s], b[s], k0 + 4 * (s + PF));
        }
    }
}

// C(cr0 + r, cc0 + c) = alpha acc (+ C when beta) for r < rows, c < cols;
// lower_only keeps to cr0 + r >= cc0 + c.
template <typename R, int MI, int NJ>
__device__ __forceinline__ void cstore(const CAcc<R, MI, NJ> &acc,
                                       const CView<R> &C, int cr0, int rows,
                                       int cc0, int cols, R alpha, int beta,
                                       int lower_only)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < MI; i++)
#pragma unroll
        for (int j = 0; j < NJ; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = 16 * i + Mfma<R>::row(lane, q);
                const int c = 16 * j + (lane & 15);
                if (r >= rows || c >= cols) continue;
                if (lower_only && cr0 + r < cc0 + c) continue;
                cx<R> v{alpha * acc.re[i][j][q], alpha * acc.im[i][j][q]};
                if (beta) {
                    const cx<R> o = C.get(cr0 + r, cc0 + c);
                    v.re += o.re;
                    v.im += o.im;
                }
                C.put(cr0 + r, cc0 + c, v);
            }
}


__host__ __device__ __forceinline__ int cdiv(int a, int b)
{
    return (a + b - 1) / b;
}

// ---------------------------------------------------------------------------
// Block jobs.  A launch is a list of 64 x 64 blocks of output windows; block
// b is C = alpha sum_kt X_kt Y_kt (+ C when beta), or a copy of operand 0's
// X.  The four waves of a workgroup take its 32 x 32 quarters.  ti / tj are
// the tile indices a decoder needs to find the operands.
// ---------------------------------------------------------------------------
template <typename R> struct Job {
    CView<R> C;
    int r0, c0, rows, cols;   // block origin in C and its extent (<= 64)
    int lower_only;           // only C(r, c) with r >= c
    int copy;                 // C = X of operand 0, no product
    int nk;                   // operand pairs
    int beta;
    R alpha;
    int ti, tj;
};

template <typename R> struct Operand {
    CView<R> X, Y;            // X(r, k) and Y(k, c) in C's coordinates
    int K;
};

template <typename R>
__device__ __forceinline__ void job_block(Job<R> &j, int r0, int c0,
                                          int mrows, int ncols)
{
    j.r0 = r0;
    j.c0 = c0;
    j.rows = min(64, mrows - r0);
    j.cols = min(64, ncols - c0);
}

template <typename R, typename Dec>
__global__ __launch_bounds__(NT) void cjobs_kernel(Dec d, const int *status)
{
    if (block_should_stop(status)) return;
    Job<R> j;
    if (!d.job(blockIdx.x, j)) return;
    if (j.rows <= 0 || j.cols <= 0) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int rows = min(32, j.rows - wr), cols = min(32, j.cols - wc);
    if (rows <= 0 || cols <= 0) return;
    const int r0 = j.r0 + wr, c0 = j.c0 + wc;
    if (j.lower_only && r0 + rows - 1 < c0) return;
    if (j.copy) {
        const Operand<R> o = d.operand(j, 0);
        for (int e = lane; e < 32 * 32; e += 64) {
            const int r = e & 31, c = e >> 5;
            if (r < rows && c < cols)
                j.C.put(r0 + r, c0 + c, o.X.get(r0 + r, c0 + c));
        }
        return;
    }
    CAcc<R, 2, 2> acc;
    acc.zero();
    for (int kt = 0; kt < j.nk; kt++) {
        const Operand<R> o = d.operand(j, kt);
        cprod<R, 2, 2, Dec::PF>(acc, o.X, r0, rows, o.Y, c0, cols, o.K);
    }
    cstore<R, 2, 2>(acc, j.C, r0, rows, c0, cols, j.alpha, j.beta,
                    j.lower_only);
}

// Panel of step k: P(m) = L(m,k) W_kk^H for m in (k, mt) into slot m-k-1 of
// the panel buffer (mb x nb per slot, ld mb): the trsm tasks of
// pzpotrf.c:47-57 (upper: :90-99), as products with the tile inverse.
template <typename R> struct PanelDec {
    // operand steps in flight (cprod): FP64 three, FP32 one
    static constexpr int PF = (sizeof(R) == 8 ? 3 : 1);
    plasma_desc_t A;
    int upper, k, nbr;
    cx<R> *P;
    const cx<R> *W;

    __device__ bool job(int b, Job<R> &j) const
    {
        const int per = nbr * nbr;
        const int m = k + 1 + b / per, blk = b % per;
        j.ti = m;
        j.tj = k;
        j.C = buf_cview(P + (size_t)(m - k - 1) * A.mb * A.nb, A.mb);
        job_block(j, 64 * (blk % nbr), 64 * (blk / nbr),
                  plasma_tile_mview(A, m), plasma_tile_mview(A, k));
        j.lower_only = 0;
        j.copy = 0;
        j.nk = 1;
        j.alpha = R(1);
        j.beta = 0;
        return true;
    }
    __device__ Operand<R> operand(const Job<R> &j, int) const
    {
        Operand<R> o;
        o.X = lower_cview<R>(A, upper != 0, j.ti, k);
        o.Y = herm_of(buf_cview((cx<R> *)W, A.mb));
        o.K = plasma_tile_mview(A, k);
        return o;
    }
};

// Update of step k: L(m,n) -= P(m) P(n)^H for n in [n_first, n_end),
// n <= m < mt (n_first > k), lower triangle only on diagonal tiles (the
// herk and gemm tasks of pzpotrf.c:58-78; upper: :100-125).  The first
// ncopy jobs copy the panels back into L(m,k); the update reads the
// buffer, not those tiles.
template <typename R> struct UpdateDec {
    // operand steps in flight (cprod): FP64 three, FP32 one
    static constexpr int PF = (sizeof(R) == 8 ? 3 : 1);
    plasma_desc_t A;
    int upper, k, nbr, ncopy, n_first;
    const cx<R> *P;

    __device__ bool job(int b, Job<R> &j) const
    {
        const int per = nbr * nbr;
        int m, n;
        j.lower_only = 0;
        j.copy = 0;
        j.nk = 1;
        j.alpha = R(-1);
        j.beta = 1;
        if (b < ncopy) {
            m = k + 1 + b / per;
            n = k;
            j.copy = 1;
        }
        else {
            b -= ncopy;
            trapezoid_decode(b / per, A.mt, n_first, &m, &n);
        }
        const int blk = b % per, br = blk % nbr, bc = blk / nbr;
        if (!j.copy && m == n) {
            if (br < bc) return false;
            j.lower_only = 1;
        }
        j.ti = m;
        j.tj = n;
        j.C = lower_cview<R>(A, upper != 0, m, n);
        job_block(j, 64 * br, 64 * bc, plasma_tile_mview(A, m),
                  plasma_tile_mview(A, n));
        return true;
    }
    __device__ Operand<R> operand(const Job<R> &j, int) const
    {
        Operand<R> o;
        const size_t slot = (size_t)A.mb * A.nb;
        o.X = buf_cview((cx<R> *)P + (j.ti - k - 1) * slot, A.mb);
        if (j.copy) {
            o.Y = o.X;
            o.K = 0;
            return o;
        }
        o.Y = herm_of(buf_cview((cx<R> *)P + (j.tj - k - 1) * slot, A.mb));
        o.K = plasma_tile_mview(A, k);
        return o;
    }
};

// potrs sweep, diagonal step k: S = W_kk B(k,:) (forward, L Y = B) or
// W_kk^H B(k,:) (backward, L^H X = Y) into S (mb x B.n, ld mb): the
// diagonal trsm tasks of pztrsm.c:113-177, as products with the inverse.
template <typename R> struct SweepDiagDec {
    // operand steps in flight (cprod): three (few blocks, K chains
    // of a whole tile)
    static constexpr int PF = 3;
    plasma_desc_t B;
    int k, backward, nbr, nbc;
    const cx<R> *W;
    cx<R> *S;

    __device__ bool job(int b, Job<R> &j) const
    {
        const int per = nbr * nbc;
        const int jt = b / per, blk = b % per;
        j.ti = k;
        j.tj = jt;
        j.C = buf_cview(S + (size_t)jt * B.nb * B.mb, B.mb);
        job_block(j, 64 * (blk % nbr), 64 * (blk / nbr),
                  plasma_tile_mview(B, k), plasma_tile_nview(B, jt));
        j.lower_only = 0;
        j.copy = 0;
        j.nk = 1;
        j.alpha = R(1);
        j.beta = 0;
        return true;
    }
    __device__ Operand<R> operand(const Job<R> &j, int) const
    {
        Operand<R> o;
        const CView<R> Wv = buf_cview((cx<R> *)W, B.mb);
        o.X = backward ? herm_of(Wv) : Wv;
        o.Y = tile_cview<R>(B, k, j.tj);
        o.K = plasma_tile_mview(B, k);
        return o;
    }
};

// potrs sweep, update of step k: B(m,:) -= L(m,k) S for m > k (forward) or
// L(k,m)^H S for m < k (backward), and the first ncopy jobs write S into
// B(k,:) (pztrsm.c:113-177, the gemm tasks).
template <typename R> struct SweepUpdDec {
    // operand steps in flight (cprod): three (few blocks per launch)
    static constexpr int PF = 3;
    plasma_desc_t A, B;
    int upper, k, backward, nbr, nbc, ncopy;
    const cx<R> *S;

    __device__ bool job(int b, Job<R> &j) const
    {
        const int per = nbr * nbc * B.nt;
        int m;
        j.copy = 0;
        if (b < ncopy) {
            m = k;
            j.copy = 1;
        }
        else {
            b -= ncopy;
            m = backward ? b / per : k + 1 + b / per;
        }
        const int rest = b % per;
        const int jt = rest / (nbr * nbc), blk = rest % (nbr * nbc);
        j.ti = m;
        j.tj = jt;
        j.C = tile_cview<R>(B, m, jt);
        job_block(j, 64 * (blk % nbr), 64 * (blk / nbr),
                  plasma_tile_mview(B, m), plasma_tile_nview(B, jt));
        j.lower_only = 0;
        j.nk = 1;
        j.alpha = R(-1);
        j.beta = 1;
        return true;
    }
    __device__ Operand<R> operand(const Job<R> &j, int) const
    {
        Operand<R> o;
        const CView<R> Sv =
            buf_cview((cx<R> *)S + (size_t)j.tj * B.nb * B.mb, B.mb);
        if (j.copy) {
            o.X = Sv;
            o.Y = Sv;
            o.K = 0;
            return o;
        }
        o.X = backward ? herm_of(lower_cview<R>(A, upper != 0, k, j.ti))
                       : lower_cview<R>(A, upper != 0, j.ti, k);
        o.Y = Sv;
        o.K = plasma_tile_mview(A, k);
        return o;
    }
};

// R(m,:) -= sum_kt A(m,kt) X(kt,:) with A Hermitian in its lower-equivalent
// storage: L(m,kt) for kt < m, L(kt,m)^H for kt > m, the diagonal tile read
// as Hermitian from its stored triangle (the hemm of zcposv.c:466-472,
// pzhemm.c Left).
template <typename R> struct ResidualDec {
    // operand steps in flight (cprod): three
    static constexpr int PF = 3;
    plasma_desc_t A, X, Rd;
    int upper, nbr, nbc;

    __device__ bool job(int b, Job<R> &j) const
    {
        const int per = nbr * nbc;
        const int t = b / per, blk = b % per;
        const int m = t % Rd.mt, jt = t / Rd.mt;
        j.ti = m;
        j.tj = jt;
        j.C = tile_cview<R>(Rd, m, jt);
        job_block(j, 64 * (blk % nbr), 64 * (blk / nbr),
                  plasma_tile_mview(Rd, m), plasma_tile_nview(Rd, jt));
        j.lower_only = 0;
        j.copy = 0;
        j.nk = A.mt;
        j.alpha = R(-1);
        j.beta = 1;
        return true;
    }
    __device__ Operand<R> operand(const Job<R> &j, int kt) const
    {
        Operand<R> o;
        const bool up = upper != 0;
        if (kt < j.ti) {
            o.X = lower_cview<R>(A, up, j.ti, kt);
        }
        else if (kt > j.ti) {
            o.X = herm_of(lower_cview<R>(A, up, kt, j.ti));
        }
        else {
            o.X = lower_cview<R>(A, up, kt, kt);
            o.X.herm = 1;
        }
        o.Y = tile_cview<R>(X, kt, j.tj);
        o.K = plasma_tile_mview(A, kt);
        return o;
    }
};

// ---------------------------------------------------------------------------
// Diagonal tile: LAPACK zpotrf semantics (the diagonal's real part is used,
// the factor's diagonal is real, the strictly upper part of the tile is not
// referenced) and W = L_kk^{-1} (lower triangular, zeros above, ld mb).
//
// Blocked by 32 columns, one workgroup: each 32 x 32 diagonal block is
// factored and inverted in LDS, the rows below it within the tile are
// solved as X <- X D^{-H} (a product with the block inverse, one 16-row
// strip per wave, in place: a wave reads only the rows it writes), and the
// tile's trailing lower part is updated by 32 x 32 block products.  W's
// off-diagonal blocks follow by block forward substitution,
//     W(r,c) = -D_r^{-1} sum_{l=c}^{r-1} L(r,l) W(l,c),
// one wave per block column.  factor == 0: the tile is already factored
// and only W is formed (one workgroup per tile, block k = blockIdx.x).
// A non positive (or NaN) pivot at tile column j sets the status word to
// nb k + j + 1 (reference core_zpotrf.c:72-81 with pzpotrf.c's offset).
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ cx<R> cdivide(cx<R> s, cx<R> d)
{
    if (d.im == R(0)) return cx<R>{s.re / d.re, s.im / d.re};
    const R q = d.re * d.re + d.im * d.im;
    const cx<R> t = cmulc(s, d);
    return cx<R>{t.re / q, t.im / q};
}

template <typename R>
__global__ __launch_bounds__(NT) void zdiag_kernel(plasma_desc_t A, int upper,
                                                   int kfirst, cx<R> *inv,
                                                   int factor, int *status)
{
    __shared__ cx<R> D[BS][BLD], Di[BS][BLD];
    __shared__ int s_fail;
    if (block_should_stop(status)) return;
    const int k = kfirst + blockIdx.x;
    const int n = plasma_tile_mview(A, k);
    const int ld = A.mb;
    const CView<R> T = lower_cview<R>(A, upper != 0, k, k);
    cx<R> *W = inv + (size_t)k * A.mb * A.mb;
    const CView<R> Wv = buf_cview(W, ld);
    const int tid = threadIdx.x, wave = tid >> 6;
    for (int e = tid; e < n * n; e += NT)
        W[e % n + (size_t)(e / n) * ld] = czero<R>();
    if (tid == 0) s_fail = 0;
    __syncthreads();

    for (int jb = 0; jb < n; jb += BS) {
        const int w = min(BS, n - jb);
        for (int e = tid; e < BS * BS; e += NT) {
            const int r = e % BS, c = e / BS;
            D[r][c] = (r < w && c < w && r >= c) ? T.get(jb + r, jb + c)
                                                 : czero<R>();
        }
        __syncthreads();
        if (factor) {
            for (int j = 0; j < w; j++) {
                if (tid == 0) {
                    const R d = D[j][j].re;
                    if (!(d > R(0)))
                        s_fail = jb + j + 1;
                    else
                        D[j][j] = cx<R>{(R)sqrt(d), R(0)};
                }
                __syncthreads();
                if (s_fail) break;
                if (tid > j && tid < w) {
                    const R d = D[j][j].re;
                    D[tid][j].re /= d;
                    D[tid][j].im /= d;
                }
                __syncthreads();
                for (int e = tid; e < BS * BS; e += NT) {
                    const int r = e % BS, c = e / BS;
                    if (c > j && c <= r && r < w) {
                        const cx<R> t = cmulc(D[r][j], D[c][j]);
                        D[r][c].re -= t.re;
                        D[r][c].im -= t.im;
                    }
                }
                __syncthreads();
            }
            if (s_fail) {
                if (tid == 0) atomicCAS(status, 0, A.nb * k + s_fail);
                return;
            }
        }
        // Di = D^{-1} by forward elimination on the identity, every entry
        // of a step in parallel: row j of Di is final after step j,
        //     Di(j,:) /= D(j,j);  Di(i,:) -= D(i,j) Di(j,:), i > j
        for (int e = tid; e < BS * BS; e += NT) {
            const int r = e % BS, c = e / BS;
            Di[r][c] = (r == c) ? cx<R>{R(1), R(0)} : czero<R>();
        }
        __syncthreads();
        for (int j = 0; j < w; j++) {
            if (tid <= j) Di[j][tid] = cdivide(Di[j][tid], D[j][j]);
            __syncthreads();
            for (int e = tid; e < BS * BS; e += NT) {
                const int r = e % BS, c = e / BS;
                if (r > j && r < w && c <= j) {
                    const cx<R> t = cmul(D[r][j], Di[j][c]);
                    Di[r][c].re -= t.re;
                    Di[r][c].im -= t.im;
                }
            }
            __syncthreads();
        }
        for (int e = tid; e < BS * BS; e += NT) {
            const int r = e % BS, c = e / BS;
            if (r < w && c < w) {
                if (factor && r >= c) T.put(jb + r, jb + c, D[r][c]);
                W[(jb + r) + (size_t)(jb + c) * ld] = Di[r][c];
            }
        }
        __syncthreads();
        if (factor && jb + w < n) {
            // rows below the block: X <- X Di^H, one 16-row strip per wave
            CView<R> Dv;
            Dv.p = &Di[0][0];
            Dv.rs = BLD;
            Dv.cs = 1;
            Dv.conj = 0;
            Dv.herm = 0;
            const CView<R> Yd = herm_of(Dv);
            const CView<R> X = T.at(0, jb);
            for (int s0 = jb + w + 16 * wave; s0 < n; s0 += 64) {
                CAcc<R, 1, 2> acc;
                acc.zero();
                const int rows = min(16, n - s0);
                cprod<R, 1, 2>(acc, X, s0, rows, Yd, 0, w, w);
                cstore<R, 1, 2>(acc, X, s0, rows, 0, w, R(1), 0, 0);
            }
            __syncthreads();
            // trailing lower part of the tile: T(r,c) -= X(r,:) X(c,:)^H
            const int t0 = jb + w, tb = cdiv(n - t0, 32);
            const int nblk = tb * (tb + 1) / 2;
            for (int b = wave; b < nblk; b += NT / 64) {
                int bi = 0;
                while ((bi + 1) * (bi + 2) / 2 <= b) bi++;
                const int bj = b - bi * (bi + 1) / 2;
                const int r0 = t0 + 32 * bi, c0 = t0 + 32 * bj;
                const int rows = min(32, n - r0), cols = min(32, n - c0);
                CAcc<R, 2, 2> acc;
                acc.zero();
                cprod<R, 2, 2>(acc, X, r0, rows, herm_of(X), c0, cols, w);
                cstore<R, 2, 2>(acc, T, r0, rows, c0, cols, R(-1), 1,
                                bi == bj);
            }
        }
        __syncthreads();
    }

    // off-diagonal blocks of W, one wave per block column
    const int nbk = cdiv(n, BS);
    for (int cb = wave; cb < nbk; cb += NT / 64) {
        for (int rb = cb + 1; rb < nbk; rb++) {
            const int r0 = BS * rb, c0 = BS * cb;
            const int rows = min(BS, n - r0), cols = min(BS, n - c0);
            CAcc<R, 2, 2> acc;
            acc.zero();
            // T1 = sum_{l = c0}^{r0 - 1} L(r, l) W(l, c)
            cprod<R, 2, 2>(acc, T.at(0, c0), r0, rows, Wv.at(c0, 0), c0, cols,
                           r0 - c0);
            cstore<R, 2, 2>(acc, Wv, r0, rows, c0, cols, R(1), 0, 0);
            __threadfence_block();
            // W(r, c) = -Di_r T1 (Di_r: W's diagonal block rb)
            CAcc<R, 2, 2> acc2;
            acc2.zero();
            cprod<R, 2, 2>(acc2, Wv.at(0, r0), r0, rows, Wv.at(r0, 0), c0,
                           cols, rows);
            cstore<R, 2, 2>(acc2, Wv, r0, rows, c0, cols, R(-1), 0, 0);
            __threadfence_block();
        }
    }
}

// ---------------------------------------------------------------------------
// Norms of the mixed solver.
// ---------------------------------------------------------------------------
template <typename R> __device__ __forceinline__ double cabs_of(cx<R> v)
{
    return hypot((double)v.re, (double)v.im);
}

// Row sums of |A| for the Hermitian Inf norm (= One norm), the layout of
// the real lansy work array: slot n of the rows of m holds the row sums of
// L(m,n); for m > n slot m of the rows of n holds its column sums.  The
// diagonal counts its real part (LAPACK zlanhe, pzlanhe.c:83-131).
template <typename R>
__global__ __launch_bounds__(NT) void zlanhe_tile_kernel(plasma_desc_t A,
                                                         int upper,
                                                         double *work,
                                                         const int *status)
{
    if (block_should_stop(status)) return;
    int m, n;
    trapezoid_decode(blockIdx.x, A.mt, 0, &m, &n);
    const int mv = plasma_tile_mview(A, m), nv = plasma_tile_mview(A, n);
    CView<R> L = lower_cview<R>(A, upper != 0, m, n);
    const size_t ld = (size_t)A.mt * A.mb;
    if (m == n) L.herm = 1;
    if (m != n && mv <= 256) {
        // One pass over an off-diagonal tile: wave w takes the columns
        // c = w, w + 4, ...; lane l the rows l, l + 64, ... (consecutive
        // rows of a column-major tile to consecutive lanes).  A column's
        // sum is a wave reduction; a row's sum adds the four waves' partial
        // sums in a fixed order.
        __shared__ double part[NT / 64][256];
        const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
        double rsum[4] = {0, 0, 0, 0};
        for (int c = wave; c < nv; c += NT / 64) {
            double s = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int r = lane + 64 * u;
                const double x = r < mv ? cabs_of(L.get(r, c)) : 0.0;
                rsum[u] += x;
                s += x;
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) work[m * ld + (size_t)n * A.mb + c] = s;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) part[wave][lane + 64 * u] = rsum[u];
        __syncthreads();
        for (int r = t; r < mv; r += NT) {
            double s = part[0][r];
#pragma unroll
            for (int w = 1; w < NT / 64; w++) s += part[w][r];
            work[n * ld + (size_t)m * A.mb + r] = s;
        }
        return;
    }
    for (int r = threadIdx.x; r < mv; r += NT) {
        double s = 0;
        for (int c = 0; c < nv; c++) s += cabs_of(L.get(r, c));
        work[n * ld + (size_t)m * A.mb + r] = s;
    }
    if (m == n) return;
    for (int c = threadIdx.x; c < nv; c += NT) {
        double s = 0;
        for (int r = 0; r < mv; r++) s += cabs_of(L.get(r, c));
        work[m * ld + (size_t)n * A.mb + c] = s;
    }
}

// values[j] = max_i |Re A(i,j)| + |Im A(i,j)| (core_dzamax.c with dcabs1;
// a NaN entry does not replace a larger value), one workgroup per column.
template <typename R>
__global__ __launch_bounds__(NT) void zamax_cols_kernel(plasma_desc_t A,
                                                        double *values,
                                                        const int *status)
{
    __shared__ double red[NT];
    if (block_should_stop(status)) return;
    const int j = blockIdx.x;
    const int tn = j / A.nb, c = j % A.nb;
    double v = 0;
    // rows of the column as one index i = tm mb + r: every thread takes
    // rows i = t, t + NT, ..., four passes at a time so that their loads
    // are in flight together
    const int rows = A.mt * A.mb;
    for (int i0 = threadIdx.x; i0 < rows; i0 += 4 * NT) {
        double a[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * NT;
            const int tm = i / A.mb, r = i - tm * A.mb;
            a[u] = 0;
            if (i < rows && r < plasma_tile_mview(A, tm)) {
                const cx<R> z = tile_cview<R>(A, tm, tn).get(r, c);
                a[u] = fabs((double)z.re) + fabs((double)z.im);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (a[u] > v) v = a[u];
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x])
            red[threadIdx.x] = red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) values[j] = red[0];
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
int done()
{
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}

template <typename R, typename Dec>
int launch_jobs(const Dec &d, long njobs, const int *status, hipStream_t s)
{
    if (njobs <= 0) return 0;
    if (njobs > 0x7fffffffL) return PlasmaErrorNotSupported;
    hipLaunchKernelGGL((cjobs_kernel<R, Dec>), dim3((unsigned)njobs), dim3(NT),
                       0, s, d, status);
    return done();
}

template <typename R>
int panel(plasma_enum_t uplo, plasma_desc_t A, int k, const void *inv,
          void *P, const int *status, hipStream_t s)
{
    PanelDec<R> d;
    d.A = A;
    d.upper = uplo == PlasmaUpper;
    d.k = k;
    d.nbr = cdiv(A.mb, 64);
    d.P = (cx<R> *)P;
    d.W = (const cx<R> *)inv + (size_t)k * A.mb * A.mb;
    return launch_jobs<R>(d, (long)(A.mt - k - 1) * d.nbr * d.nbr, status, s);
}

template <typename R>
int update(plasma_enum_t uplo, plasma_desc_t A, int k, int n_first,
           int n_end, int copy, const void *P, const int *status,
           hipStream_t s)
{
    UpdateDec<R> d;
    d.A = A;
    d.upper = uplo == PlasmaUpper;
    d.k = k;
    d.n_first = n_first;
    d.nbr = cdiv(A.mb, 64);
    d.P = (const cx<R> *)P;
    const long per = (long)d.nbr * d.nbr;
    const long rest = (long)A.mt - k - 1;
    d.ncopy = copy ? (int)(rest * per) : 0;
    // tiles of columns [n_first, n_end): column n holds mt - n of them
    long tiles = 0;
    for (int n = n_first; n < n_end; n++) tiles += A.mt - n;
    return launch_jobs<R>(d, d.ncopy + tiles * per, status, s);
}

template <typename R>
int sweep_diag(int backward, plasma_desc_t B, int k, const void *inv,
               void *S, const int *status, hipStream_t s)
{
    SweepDiagDec<R> d;
    d.B = B;
    d.k = k;
    d.backward = backward;
    d.nbr = cdiv(B.mb, 64);
    d.nbc = cdiv(B.nb, 64);
    d.W = (const cx<R> *)inv + (size_t)k * B.mb * B.mb;
    d.S = (cx<R> *)S;
    return launch_jobs<R>(d, (long)B.nt * d.nbr * d.nbc, status, s);
}

template <typename R>
int sweep_update(plasma_enum_t uplo, int backward, plasma_desc_t A,
                 plasma_desc_t B, int k, const void *S, const int *status,
                 hipStream_t s)
{
    SweepUpdDec<R> d;
    d.A = A;
    d.B = B;
    d.upper = uplo == PlasmaUpper;
    d.k = k;
    d.backward = backward;
    d.nbr = cdiv(B.mb, 64);
    d.nbc = cdiv(B.nb, 64);
    d.S = (const cx<R> *)S;
    const long per = (long)d.nbr * d.nbc * B.nt;
    d.ncopy = (int)per;
    const long targets = backward ? k : (long)A.mt - k - 1;
    return launch_jobs<R>(d, per + targets * per, status, s);
}

template <typename R>
int residual(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t X,
             plasma_desc_t Rd, const int *status, hipStream_t s)
{
    ResidualDec<R> d;
    d.A = A;
    d.X = X;
    d.Rd = Rd;
    d.upper = uplo == PlasmaUpper;
    d.nbr = cdiv(Rd.mb, 64);
    d.nbc = cdiv(Rd.nb, 64);
    return launch_jobs<R>(d, (long)Rd.mt * Rd.nt * d.nbr * d.nbc, status, s);
}

bool is_z(plasma_desc_t A) { return A.precision == PlasmaComplexDouble; }
bool is_c(plasma_desc_t A) { return A.precision == PlasmaComplexFloat; }

}  // namespace

extern "C" int plasma_kern_zpotrf_diag(plasma_enum_t uplo, plasma_desc_t A,
                                       int k, void *inv, int *status,
                                       hipStream_t s)
{
    const int upper = uplo == PlasmaUpper;
    if (is_z(A))
        hipLaunchKernelGGL(zdiag_kernel<double>, dim3(1), dim3(NT), 0, s, A,
                           upper, k, (cx<double> *)inv, 1, status);
    else if (is_c(A))
        hipLaunchKernelGGL(zdiag_kernel<float>, dim3(1), dim3(NT), 0, s, A,
                           upper, k, (cx<float> *)inv, 1, status);
    else
        return PlasmaErrorNotSupported;
    return done();
}

extern "C" int plasma_kern_ztrtri_diag(plasma_enum_t uplo, plasma_desc_t A,
                                       void *inv, int *status, hipStream_t s)
{
    const int upper = uplo == PlasmaUpper;
    if (A.mt <= 0) return 0;
    if (is_z(A))
        hipLaunchKernelGGL(zdiag_kernel<double>, dim3(A.mt), dim3(NT), 0, s,
                           A, upper, 0, (cx<double> *)inv, 0, status);
    else if (is_c(A))
        hipLaunchKernelGGL(zdiag_kernel<float>, dim3(A.mt), dim3(NT), 0, s, A,
                           upper, 0, (cx<float> *)inv, 0, status);
    else
        return PlasmaErrorNotSupported;
    return done();
}

extern "C" int plasma_kern_zpanel(plasma_enum_t uplo, plasma_desc_t A, int k,
                                  const void *inv, void *panel_buf,
                                  const int *status, hipStream_t s)
{
    if (is_z(A)) return panel<double>(uplo, A, k, inv, panel_buf, status, s);
    if (is_c(A)) return panel<float>(uplo, A, k, inv, panel_buf, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_zupdate(plasma_enum_t uplo, plasma_desc_t A, int k,
                                   int n_first, int n_end, int copy,
                                   const void *panel_buf, const int *status,
                                   hipStream_t s)
{
    if (n_end > A.mt) n_end = A.mt;
    if (n_first <= k || (n_first >= n_end && !copy)) return 0;
    if (is_z(A))
        return update<double>(uplo, A, k, n_first, n_end, copy, panel_buf,
                              status, s);
    if (is_c(A))
        return update<float>(uplo, A, k, n_first, n_end, copy, panel_buf,
                             status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_zsweep_diag(int backward, plasma_desc_t B, int k,
                                       const void *inv, void *S,
                                       const int *status, hipStream_t s)
{
    if (is_z(B)) return sweep_diag<double>(backward, B, k, inv, S, status, s);
    if (is_c(B)) return sweep_diag<float>(backward, B, k, inv, S, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_zsweep_update(plasma_enum_t uplo, int backward,
                                         plasma_desc_t A, plasma_desc_t B,
                                         int k, const void *S,
                                         const int *status, hipStream_t s)
{
    if (is_z(A))
        return sweep_update<double>(uplo, backward, A, B, k, S, status, s);
    if (is_c(A))
        return sweep_update<float>(uplo, backward, A, B, k, S, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_zhemm_residual(plasma_enum_t uplo, plasma_desc_t A,
                                          plasma_desc_t X, plasma_desc_t R,
                                          const int *status, hipStream_t s)
{
    if (is_z(A)) return residual<double>(uplo, A, X, R, status, s);
    if (is_c(A)) return residual<float>(uplo, A, X, R, status, s);
    return PlasmaErrorNotSupported;
}

extern "C" int plasma_kern_zlanhe_tiles(plasma_enum_t uplo, plasma_desc_t A,
                                        double *work, const int *status,
                                        hipStream_t s)
{
    const long ntiles = (long)A.mt * (A.mt + 1) / 2;
    if (ntiles == 0) return 0;
    const int upper = uplo == PlasmaUpper;
    if (is_z(A))
        hipLaunchKernelGGL(zlanhe_tile_kernel<double>, dim3((unsigned)ntiles),
                           dim3(NT), 0, s, A, upper, work, status);
    else if (is_c(A))
        hipLaunchKernelGGL(zlanhe_tile_kernel<float>, dim3((unsigned)ntiles),
                           dim3(NT), 0, s, A, upper, work, status);
    else
        return PlasmaErrorNotSupported;
    return done();
}

extern "C" int plasma_kern_zamax_cols(plasma_desc_t A, double *values,
                                      const int *status, hipStream_t s)
{
    if (A.n == 0) return 0;
    if (is_z(A))
        hipLaunchKernelGGL(zamax_cols_kernel<double>, dim3(A.n), dim3(NT), 0,
                           s, A, values, status);
    else if (is_c(A))
        hipLaunchKernelGGL(zamax_cols_kernel<float>, dim3(A.n), dim3(NT), 0, s,
                           A, values, status);
    else
        return PlasmaErrorNotSupported;
    return done();
}
