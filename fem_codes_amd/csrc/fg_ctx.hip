// fg_ctx.hip -- context, device memory, error reporting, precision casts.
#include <stdarg.h>
#include <string.h>

#include <atomic>

#include "fg_internal.h"

static std::atomic<unsigned long long> g_cfg_gen_value{0};

static thread_local char g_err[1024] = "";

void fg_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

extern "C" const char *fg_last_error(void) { return g_err; }
extern "C" const char *fg_version(void) { return "femgpu 0.1 gfx950"; }

extern "C" int fg_ctx_create(int device, fg_ctx **out)
{
    FG_ARG(out != nullptr, "out");
    int ndev = 0;
    FG_HIP(hipGetDeviceCount(&ndev));
    FG_ARG(device >= 0 && device < ndev, "device index");
    FG_HIP(hipSetDevice(device));
    fg_ctx *c = new fg_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    FG_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cu = prop.multiProcessorCount;
    c->red_blocks = c->num_cu * 8;
    if (c->red_blocks > FG_RED_MAX) c->red_blocks = FG_RED_MAX;
    FG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    FG_HIP(hipMalloc(&c->red_part, sizeof(double) * FG_RED_MAX));
    FG_HIP(hipMalloc(&c->red_ticket, sizeof(unsigned) * 4));
    FG_HIP(hipMemsetAsync(c->red_ticket, 0, sizeof(unsigned) * 4, c->stream));
    FG_HIP(hipEventCreate(&c->ev0));
    FG_HIP(hipEventCreate(&c->ev1));
    FG_HIP(hipStreamSynchronize(c->stream));
    // once per device: the device address of the allocation-free solver's
    // scratch array, looked up with this device current
    {
        const int rc = fg_noalloc_init(device);
        if (rc != FG_OK) {
            hipFree(c->red_part);
            hipFree(c->red_ticket);
            hipEventDestroy(c->ev0);
            hipEventDestroy(c->ev1);
            hipStreamDestroy(c->stream);
            delete c;
            return rc;
        }
    }
    *out = c;
    return FG_OK;
}

unsigned long long fg_cfg_bump() { return ++g_cfg_gen_value; }

unsigned long long fg_cfg_gen() { return g_cfg_gen_value.load(); }

void fg_cycle_graph_clear(fg_cycle_graph &G)
{
    if (G.ge) hipGraphExecDestroy(G.ge);
    if (G.g) hipGraphDestroy(G.g);
    G = fg_cycle_graph();
}

extern "C" int fg_ctx_destroy(fg_ctx *c)
{
    if (!c) return FG_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    fg_cycle_graph_clear(c->cg);
    hipFree(c->red_part);
    hipFree(c->red_ticket);
    hipFree(c->ws);
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    hipStreamDestroy(c->stream);
    delete c;
    return FG_OK;
}

int fg_ctx_workspace(fg_ctx *c, size_t bytes, void **p)
{
    if (bytes == 0) bytes = 256;
    if (bytes > c->ws_bytes) {
        FG_HIP(hipStreamSynchronize(c->stream));
        if (c->ws) FG_HIP(hipFree(c->ws));
        c->ws = nullptr;
        c->ws_bytes = 0;
        c->ws_gen++;
        FG_HIP(hipMalloc(&c->ws, bytes));
        c->ws_bytes = bytes;
    }
    *p = c->ws;
    return FG_OK;
}

extern "C" int fg_ctx_trim(fg_ctx *c)
{
    FG_ARG(c, "ctx");
    FG_HIP(hipSetDevice(c->device));
    FG_HIP(hipStreamSynchronize(c->stream));
    fg_cycle_graph_clear(c->cg);
    if (c->ws) FG_HIP(hipFree(c->ws));
    c->ws = nullptr;
    c->ws_bytes = 0;
    c->ws_gen++;
    return FG_OK;
}

extern "C" int fg_ctx_sync(fg_ctx *c)
{
    FG_ARG(c, "ctx");
    FG_HIP(hipStreamSynchronize(c->stream));
    return FG_OK;
}

extern "C" int fg_malloc(fg_ctx *c, size_t bytes, void **dptr)
{
    FG_ARG(c && dptr, "ctx/dptr");
    FG_HIP(hipSetDevice(c->device));
    FG_HIP(hipMalloc(dptr, bytes ? bytes : 1));
    return FG_OK;
}

extern "C" int fg_free(fg_ctx *c, void *dptr)
{
    FG_ARG(c, "ctx");
    if (!dptr) return FG_OK;
    FG_HIP(hipStreamSynchronize(c->stream));
    FG_HIP(hipFree(dptr));
    return FG_OK;
}

extern "C" int fg_memcpy_h2d(fg_ctx *c, void *dst, const void *src, size_t bytes)
{
    FG_ARG(c, "ctx");
    FG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    return FG_OK;
}

extern "C" int fg_memcpy_d2h(fg_ctx *c, void *dst, const void *src, size_t bytes)
{
    FG_ARG(c, "ctx");
    FG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    return FG_OK;
}

extern "C" int fg_memcpy_d2d(fg_ctx *c, void *dst, const void *src, size_t bytes)
{
    FG_ARG(c, "ctx");
    FG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return FG_OK;
}

extern "C" int fg_memset(fg_ctx *c, void *dst, int value, size_t bytes)
{
    FG_ARG(c, "ctx");
    FG_HIP(hipMemsetAsync(dst, value, bytes, c->stream));
    return FG_OK;
}

int fg_vec_grid(fg_ctx *c, long long n)
{
    long long g = (n + FG_TPB * 4 - 1) / (FG_TPB * 4);
    if (g < 1) g = 1;
    if (g > c->red_blocks) g = c->red_blocks;
    return (int)g;
}

__global__ void k_cast_64_32(int n, const double *__restrict__ x, float *__restrict__ y)
{
    for (int i = blockIdx.x * FG_TPB + threadIdx.x; i < n; i += gridDim.x * FG_TPB)
        y[i] = (float)x[i];
}

__global__ void k_cast_32_64(int n, const float *__restrict__ x, double *__restrict__ y)
{
    for (int i = blockIdx.x * FG_TPB + threadIdx.x; i < n; i += gridDim.x * FG_TPB)
        y[i] = (double)x[i];
}

extern "C" int fg_cast_f64_to_f32(fg_ctx *c, int n, const double *x, float *y)
{
    FG_ARG(c && n >= 0, "cast");
    if (n == 0) return FG_OK;
    FG_LAUNCH(k_cast_64_32, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, x, y);
    return FG_OK;
}

extern "C" int fg_cast_f32_to_f64(fg_ctx *c, int n, const float *x, double *y)
{
    FG_ARG(c && n >= 0, "cast");
    if (n == 0) return FG_OK;
    FG_LAUNCH(k_cast_32_64, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, x, y);
    return FG_OK;
}
