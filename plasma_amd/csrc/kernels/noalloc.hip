// noalloc.hip -- the device scratch of plasma_dposv_vbatched_noalloc (no
// counterpart in the reference; compute/batched.c holds the entry point).
//
// Callers that cannot allocate device memory (code under hipGraph capture,
// serving loops, harnesses that forbid hipMalloc) get a batched solve whose
// scratch lives in one file-scope device array, sized for at most
// PLASMA_VBATCH_NOALLOC_MAX systems and PLASMA_VBATCH_NOALLOC_GRID resident
// workgroups.  Every device the library runs on holds its own copy of the
// array until the process ends.  In host code the array's name is only a
// stand-in for the symbol, so the device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and that address is
// what every kernel argument and copy uses.
//
// Layout (byte offsets, each region 256-byte aligned):
//   W      grid slabs of the diagonal-block inverses (FP64)
//   ctrl   ticket, invalid-order flag
//   n      orders (when the caller's array is on the host)
//   order  the systems, largest order first
//   info   per-system info (when the caller's array is on the host)
//   aoff, boff  offsets (when the caller's arrays are on the host)
#include <hip/hip_runtime.h>

#include "plasma_kernels.h"

namespace {

constexpr int VBMAX_N = 512;                   // largest system order
constexpr size_t WSLAB_BYTES = (size_t)33 * VBMAX_N * sizeof(double);
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

constexpr size_t OFF_W = 0;
constexpr size_t OFF_CTRL =
    OFF_W + align256(WSLAB_BYTES * PLASMA_VBATCH_NOALLOC_GRID);
constexpr size_t OFF_N = OFF_CTRL + 256;
constexpr size_t OFF_ORDER =
    OFF_N + align256(sizeof(int) * PLASMA_VBATCH_NOALLOC_MAX);
constexpr size_t OFF_INFO =
    OFF_ORDER + align256(sizeof(int) * PLASMA_VBATCH_NOALLOC_MAX);
constexpr size_t OFF_AOFF =
    OFF_INFO + align256(sizeof(int) * PLASMA_VBATCH_NOALLOC_MAX);
constexpr size_t OFF_BOFF =
    OFF_AOFF + align256(sizeof(size_t) * PLASMA_VBATCH_NOALLOC_MAX);
constexpr size_t SCRATCH_BYTES =
    OFF_BOFF + align256(sizeof(size_t) * PLASMA_VBATCH_NOALLOC_MAX);

constexpr int MAX_DEVICES = 64;
void *scratch_ptr[MAX_DEVICES];

// Systems by order, largest first (the order only balances the persistent
// grid: every system's result is the same whichever workgroup solves it).
// One workgroup: count the orders, turn the counts into descending start
// positions, place each system; zero the ticket; flag an order outside
// [0, 512].
constexpr int ORDER_NT = 1024;
__global__ __launch_bounds__(ORDER_NT) void order_kernel(const int *n,
                                                         int batch,
                                                         int *order,
                                                         int *ctrl)
{
    __shared__ int start[VBMAX_N + 1];
    __shared__ int bad;
    const int t = threadIdx.x;
    for (int v = t; v <= VBMAX_N; v += ORDER_NT) start[v] = 0;
    if (t == 0) bad = 0;
    __syncthreads();
    for (int i = t; i < batch; i += ORDER_NT) {
        int v = n[i];
        if (v < 0 || v > VBMAX_N) {
            atomicOr(&bad, 1);
            v = 0;
        }
        atomicAdd(&start[v], 1);
    }
    __syncthreads();
    if (t == 0) {
        int s = 0;
        for (int v = VBMAX_N; v >= 0; v--) {
            const int c = start[v];
            start[v] = s;
            s += c;
        }
    }
    __syncthreads();
    for (int i = t; i < batch; i += ORDER_NT) {
        int v = n[i];
        if (v < 0 || v > VBMAX_N) v = 0;
        order[atomicAdd(&start[v], 1)] = i;
    }
    if (t == 0) {
        ctrl[0] = 0;       // ticket
        ctrl[1] = bad;     // invalid-order flag
    }
}

}  // namespace

// The scratch array itself.  Host code never passes this name to a kernel
// or a copy: it uses the address plasma_noalloc_setup looked up.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

extern "C" int plasma_noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return PlasmaErrorNotSupported;
    if (scratch_ptr[device] != nullptr) return 0;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) {
        (void)hipGetLastError();
        return PlasmaErrorInternal;
    }
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess ||
        p == nullptr) {
        (void)hipGetLastError();
        return PlasmaErrorInternal;
    }
    scratch_ptr[device] = p;
    return 0;
}

extern "C" int plasma_noalloc_regions(int device, plasma_noalloc_t *r)
{
    if (device < 0 || device >= MAX_DEVICES || scratch_ptr[device] == nullptr)
        return PlasmaErrorInternal;
    char *base = (char *)scratch_ptr[device];
    r->W = base + OFF_W;
    r->wslab_bytes = WSLAB_BYTES;
    r->grid_max = PLASMA_VBATCH_NOALLOC_GRID;
    r->ctrl = (int *)(base + OFF_CTRL);
    r->n = (int *)(base + OFF_N);
    r->order = (int *)(base + OFF_ORDER);
    r->info = (int *)(base + OFF_INFO);
    r->aoff = (size_t *)(base + OFF_AOFF);
    r->boff = (size_t *)(base + OFF_BOFF);
    r->bytes = SCRATCH_BYTES;
    return 0;
}

extern "C" int plasma_kern_vbatched_order(const int *n, int batch, int *order,
                                          int *ctrl, hipStream_t s)
{
    hipLaunchKernelGGL(order_kernel, dim3(1), dim3(ORDER_NT), 0, s, n, batch,
                       order, ctrl);
    return hipGetLastError() == hipSuccess ? 0 : PlasmaErrorInternal;
}
