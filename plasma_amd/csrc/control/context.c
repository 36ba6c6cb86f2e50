/*
 * context.c -- library initialisation, parameters and the device binding.
 *
 * Mirrors the reference's single global context (control/context.c:22-202):
 * plasma_init / plasma_finalize / plasma_set / plasma_get with the same
 * return codes and defaults (tuning enabled, nb = 256, ib = 64).  On top of
 * that the context owns the MI355X resources: the HIP device, two streams
 * of default priority (updates / lookahead panel), the device status words
 * of the sequences, a small device scratch area and the workspace pool.
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static int plasma_initialized_g = 0;
static struct plasma_context_s plasma_context_g;

int plasma_hip_check(hipError_t err, const char *what)
{
    if (err == hipSuccess)
        return PlasmaSuccess;
    fprintf(stderr, "PLASMA ERROR: %s failed: %s\n", what,
            hipGetErrorString(err));
    return PlasmaErrorInternal;
}

plasma_context_t *plasma_context_self(void)
{
    return plasma_initialized_g ? &plasma_context_g : NULL;
}

void plasma_version(int *major, int *minor, int *micro)
{
    if (major) *major = PLASMA_VERSION_MAJOR;
    if (minor) *minor = PLASMA_VERSION_MINOR;
    if (micro) *micro = PLASMA_VERSION_PATCH;
}

plasma_enum_t plasma_uplo_const(char lapack_char)
{
    switch (lapack_char) {
    case 'U': case 'u': return PlasmaUpper;
    case 'L': case 'l': return PlasmaLower;
    default:            return PlasmaGeneral;
    }
}

static void context_defaults(plasma_context_t *ctx)
{
    memset(ctx, 0, sizeof(*ctx));
    ctx->tuning = PlasmaEnabled;
    ctx->nb = 256;
    ctx->ib = 64;
    ctx->inplace_outplace = PlasmaOutplace;
    ctx->max_threads = 1;
    ctx->max_panel_threads = 1;
    ctx->householder_mode = PlasmaFlatHouseholder;
    ctx->lookahead = 1;
    ctx->group = 2;
    ctx->split_update = PLASMA_DEFAULT_SPLIT;
    ctx->half_input = PlasmaSplitBf16;
    ctx->low_factor = 0;
    ctx->graph_capture = PlasmaDisabled;
}

int plasma_init(void)
{
    if (plasma_initialized_g)
        return PlasmaErrorNotInitialized;

    plasma_context_t *ctx = &plasma_context_g;
    context_defaults(ctx);

    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        plasma_error("no HIP device available: this library runs on "
                     "an MI355X (gfx950) GPU");
        return PlasmaErrorEnvironment;
    }
    int dev = 0;
    const char *env = getenv("PLASMA_DEVICE");
    if (env != NULL)
        dev = atoi(env);
    else if (hipGetDevice(&dev) != hipSuccess)
        dev = 0;
    if (dev < 0 || dev >= count) {
        plasma_error("PLASMA_DEVICE out of range");
        return PlasmaErrorEnvironment;
    }
    if (plasma_hip_check(hipSetDevice(dev), "hipSetDevice"))
        return PlasmaErrorEnvironment;
    ctx->device = dev;

    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess)
        ctx->num_cus = prop.multiProcessorCount;

    /* Both streams keep the default priority. */
    if (plasma_hip_check(hipStreamCreateWithFlags(&ctx->stream,
                                                  hipStreamNonBlocking),
                         "hipStreamCreate") ||
        plasma_hip_check(hipStreamCreateWithFlags(&ctx->panel_stream,
                                                  hipStreamNonBlocking),
                         "hipStreamCreate"))
        return PlasmaErrorEnvironment;
    ctx->owns_streams = 1;

    /* one word per sequence slot, plus the word of the cached graphs */
    if (plasma_hip_check(hipMalloc((void **)&ctx->d_status,
                                   (PLASMA_STATUS_SLOTS + 1) * sizeof(int)),
                         "hipMalloc(status)") ||
        plasma_hip_check(hipMalloc((void **)&ctx->d_scalars,
                                   PLASMA_SCALARS * sizeof(double)),
                         "hipMalloc(scalars)"))
        return PlasmaErrorOutOfMemory;
    if (plasma_kern_fill_int(ctx->d_status, 0, PLASMA_STATUS_SLOTS + 1,
                             ctx->stream) != 0 ||
        plasma_hip_check(hipStreamSynchronize(ctx->stream), "init sync"))
        return PlasmaErrorEnvironment;

    /* the address of this device's allocation-free scratch
     * (kernels/noalloc.hip), looked up once with the device current */
    if (plasma_noalloc_setup(dev) != 0)
        plasma_warning("allocation-free batched scratch unavailable");

    plasma_initialized_g = 1;
    plasma_tuning_init(ctx);
    return PlasmaSuccess;
}

int plasma_finalize(void)
{
    if (!plasma_initialized_g)
        return PlasmaErrorNotInitialized;
    plasma_context_t *ctx = &plasma_context_g;
    hipStreamSynchronize(ctx->stream);
    hipStreamSynchronize(ctx->panel_stream);
    plasma_graph_release_all(ctx);
    plasma_pool_release_all(ctx);
    hipFree(ctx->d_status);
    hipFree(ctx->d_scalars);
    if (ctx->owns_streams) {
        hipStreamDestroy(ctx->stream);
        hipStreamDestroy(ctx->panel_stream);
    }
    plasma_initialized_g = 0;
    return PlasmaSuccess;
}

int plasma_set(plasma_enum_t param, int value)
{
    if (!plasma_initialized_g)
        return PlasmaErrorNotInitialized;
    plasma_context_t *ctx = &plasma_context_g;
    switch (param) {
    case PlasmaTuning:
        if (value != PlasmaEnabled && value != PlasmaDisabled) {
            plasma_error("invalid tuning flag");
            return PlasmaErrorIllegalValue;
        }
        ctx->tuning = value;
        break;
    case PlasmaNb:
        if (value <= 0) {
            plasma_error("invalid tile size");
            return PlasmaErrorIllegalValue;
        }
        ctx->nb = value;
        break;
    case PlasmaIb:
        if (value <= 0) {
            plasma_error("invalid inner block size");
            return PlasmaErrorIllegalValue;
        }
        ctx->ib = value;
        break;
    case PlasmaNumPanelThreads:
        if (value <= 0) {
            plasma_error("invalid number of panel threads");
            return PlasmaErrorIllegalValue;
        }
        ctx->max_panel_threads = value;
        break;
    case PlasmaHouseholderMode:
        if (value != PlasmaFlatHouseholder && value != PlasmaTreeHouseholder) {
            plasma_error("invalid Householder mode");
            return PlasmaErrorIllegalValue;
        }
        ctx->householder_mode = value;
        break;
    case PlasmaLookahead:
        if (value < 0 || value > 4) {
            plasma_error("invalid lookahead depth");
            return PlasmaErrorIllegalValue;
        }
        ctx->lookahead = value;
        break;
    case PlasmaUpdateDepth:
        if (value < 1 || value > 4) {
            plasma_error("invalid update depth");
            return PlasmaErrorIllegalValue;
        }
        ctx->group = value;
        break;
    case PlasmaGraphCapture:
        if (value != PlasmaEnabled && value != PlasmaDisabled) {
            plasma_error("invalid graph capture flag");
            return PlasmaErrorIllegalValue;
        }
        ctx->graph_capture = value;
        break;
    case PlasmaSplitUpdate:
        if (value != PlasmaEnabled && value != PlasmaDisabled &&
            value != PlasmaSplitBf16 && value != PlasmaSplitFp16) {
            plasma_error("invalid split update mode");
            return PlasmaErrorIllegalValue;
        }
        ctx->split_update = value == PlasmaEnabled ? PLASMA_DEFAULT_SPLIT
                                                   : value;
        break;
    case PlasmaHalfInput:
        if (value != PlasmaSplitBf16 && value != PlasmaSplitFp16) {
            plasma_error("invalid half-precision input format");
            return PlasmaErrorIllegalValue;
        }
        ctx->half_input = value;
        break;
    default:
        plasma_error("unknown parameter");
        return PlasmaErrorIllegalValue;
    }
    return PlasmaSuccess;
}

int plasma_get(plasma_enum_t param, int *value)
{
    if (!plasma_initialized_g)
        return PlasmaErrorNotInitialized;
    plasma_context_t *ctx = &plasma_context_g;
    switch (param) {
    case PlasmaTuning:          *value = ctx->tuning;            break;
    case PlasmaNb:              *value = ctx->nb;                break;
    case PlasmaIb:              *value = ctx->ib;                break;
    case PlasmaNumPanelThreads: *value = ctx->max_panel_threads; break;
    case PlasmaHouseholderMode: *value = ctx->householder_mode;  break;
    case PlasmaLookahead:       *value = ctx->lookahead;         break;
    case PlasmaGraphCapture:    *value = ctx->graph_capture;     break;
    case PlasmaDevice:          *value = ctx->device;            break;
    case PlasmaUpdateDepth:     *value = ctx->group;             break;
    case PlasmaSplitUpdate:     *value = ctx->split_update;      break;
    case PlasmaHalfInput:       *value = ctx->half_input;        break;
    default:
        plasma_error("Unknown parameter");
        return PlasmaErrorIllegalValue;
    }
    return PlasmaSuccess;
}
