// Library runtime: device selection, the library stream, error reporting and
// device-memory helpers (utils.c:555 phgInit / utils.c:1996 phgError roles).
#include "phg_hip_internal.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

namespace phg {

static Context g_ctx;
static bool g_ready = false;
static thread_local char g_msg[512] = "";

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    return set_error(e == hipErrorOutOfMemory ? PHG_HIP_E_NOMEM : PHG_HIP_E_HIP,
                     "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file,
                     line);
}

Context *ctx() { return g_ready ? &g_ctx : nullptr; }

int require_ctx()
{
    if (!g_ready)
        return set_error(PHG_HIP_E_STATE, "phg_hip_init has not been called");
    return PHG_HIP_OK;
}

int dmalloc(void **p, size_t bytes)
{
    *p = nullptr;
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(PHG_HIP_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes,
                         hipGetErrorString(e));
    }
    return PHG_HIP_OK;
}

void dfree(void *p)
{
    if (p)
        (void)hipFree(p);
}

// Host arrays of several MB (load vectors, Dirichlet values, coefficients
// as PHG hands them over) go through a pinned staging buffer of the
// library: host threads copy a piece into it, then one DMA moves the piece;
// a pageable copy of the same size runs at a fraction of the link's rate.
constexpr size_t STAGE_BYTES = size_t(64) << 20;
constexpr size_t STAGE_MIN = size_t(4) << 20;
static void *g_stage = nullptr;

static void parallel_memcpy(void *dst, const void *src, size_t bytes)
{
    const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, bytes / (size_t(1) << 20)));
    std::vector<std::thread> th;
    auto piece = [&](int t) {
        const size_t lo = bytes * t / nt, hi = bytes * (t + 1) / nt;
        memcpy((char *)dst + lo, (const char *)src + lo, hi - lo);
    };
    for (int t = 1; t < nt; t++) {
        try {
            th.emplace_back(piece, t);
        }
        catch (...) {
            piece(t);
        }
    }
    piece(0);
    for (auto &x : th)
        x.join();
}

int upload(void *d, const void *h, size_t bytes)
{
    if (bytes == 0)
        return PHG_HIP_OK;
    if (bytes >= STAGE_MIN && g_stage == nullptr &&
        hipHostMalloc(&g_stage, STAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        g_stage = nullptr;
    }
    if (bytes < STAGE_MIN || g_stage == nullptr) {
        PHG_HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice,
                                     g_ctx.stream));
        PHG_HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
        return PHG_HIP_OK;
    }
    for (size_t off = 0; off < bytes; off += STAGE_BYTES) {
        const size_t len = std::min(STAGE_BYTES, bytes - off);
        parallel_memcpy(g_stage, (const char *)h + off, len);
        PHG_HIP_CHECK(hipMemcpyAsync((char *)d + off, g_stage, len,
                                     hipMemcpyHostToDevice, g_ctx.stream));
        PHG_HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
    }
    return PHG_HIP_OK;
}

static double now_ms()
{
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

PhaseTrace::PhaseTrace(const char *sc)
    : scope(sc), on(getenv("PHG_HIP_SETUP_TRACE") != nullptr), t_last(0.0)
{
    if (on) {
        (void)hipStreamSynchronize(g_ctx.stream);
        t_last = now_ms();
    }
}

void PhaseTrace::mark(const char *what)
{
    if (!on)
        return;
    (void)hipStreamSynchronize(g_ctx.stream);
    const double t = now_ms();
    fprintf(stderr, "%s: %-28s %8.2f ms\n", scope, what, t - t_last);
    t_last = t;
}

int download(void *h, const void *d, size_t bytes)
{
    if (bytes == 0)
        return PHG_HIP_OK;
    PHG_HIP_CHECK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost,
                                 g_ctx.stream));
    PHG_HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
    return PHG_HIP_OK;
}

}  // namespace phg

extern "C" int phg_hip_init(int32_t device)
{
    using namespace phg;
    if (g_ready) {
        if (device == g_ctx.device)
            return PHG_HIP_OK;
        return set_error(PHG_HIP_E_STATE,
                         "already initialised on device %d", g_ctx.device);
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return set_error(PHG_HIP_E_HIP, "no HIP device available (%s)",
                         hipGetErrorString(e));
    if (device < 0 || device >= n)
        return set_error(PHG_HIP_E_ARG, "device %d out of range [0,%d)",
                         device, n);
    PHG_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    PHG_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    g_ctx.device = device;
    g_ctx.num_cus = prop.multiProcessorCount;
    // default priority stream
    PHG_HIP_CHECK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    int rc = noalloc_init(device);
    if (rc) {
        (void)hipStreamDestroy(g_ctx.stream);
        g_ctx = Context();
        return rc;
    }
    PHG_HIP_CHECK(hipHostMalloc(&g_ctx.host_mailbox, 4096, hipHostMallocDefault));
    g_ready = true;
    return PHG_HIP_OK;
}

namespace phg {
int aux_streams(hipStream_t **pool, int *n)
{
    int rc;
    if ((rc = require_ctx()))
        return rc;
    while (g_ctx.naux < Context::NAUX) {
        hipStream_t st;
        PHG_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        g_ctx.aux[g_ctx.naux++] = st;
    }
    *pool = g_ctx.aux;
    *n = g_ctx.naux;
    return PHG_HIP_OK;
}
}  // namespace phg

extern "C" int phg_hip_finalize(void)
{
    using namespace phg;
    if (!g_ready)
        return PHG_HIP_OK;
    (void)hipStreamSynchronize(g_ctx.stream);
    for (int i = 0; i < g_ctx.naux; i++) {
        (void)hipStreamSynchronize(g_ctx.aux[i]);
        (void)hipStreamDestroy(g_ctx.aux[i]);
    }
    (void)hipHostFree(g_ctx.host_mailbox);
    if (g_stage != nullptr)
        (void)hipHostFree(g_stage);
    g_stage = nullptr;
    (void)hipStreamDestroy(g_ctx.stream);
    g_ctx = Context();
    g_ready = false;
    return PHG_HIP_OK;
}

extern "C" const char *phg_hip_last_error(void) { return phg::g_msg; }

extern "C" int phg_hip_synchronize(void)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    PHG_HIP_CHECK(hipStreamSynchronize(phg::g_ctx.stream));
    return PHG_HIP_OK;
}

//------------------------------------------------------------------ vectors

extern "C" int phg_hip_vec_create(int64_t n, phg_hip_vec **v)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (v == nullptr || n <= 0)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_vec_create: bad args");
    phg_hip_vec *x = new phg_hip_vec();
    x->n = n;
    if ((rc = phg::dmalloc((void **)&x->d, (size_t)n * sizeof(double)))) {
        delete x;
        return rc;
    }
    PHG_HIP_CHECK(hipMemsetAsync(x->d, 0, (size_t)n * sizeof(double),
                                 phg::g_ctx.stream));
    *v = x;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_vec_destroy(phg_hip_vec *v)
{
    if (v == nullptr)
        return PHG_HIP_OK;
    if (v->batch_refs > 0)
        return phg::set_error(PHG_HIP_E_STATE, "phg_hip_vec_destroy: the vector is in a "
                                               "batch; destroy the batch first");
    phg::dfree(v->d);
    delete v;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_vec_upload(phg_hip_vec *v, const double *host)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (v == nullptr || host == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null vector or buffer");
    return phg::upload(v->d, host, (size_t)v->n * sizeof(double));
}

extern "C" int phg_hip_vec_download(const phg_hip_vec *v, double *host)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (v == nullptr || host == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null vector or buffer");
    return phg::download(host, v->d, (size_t)v->n * sizeof(double));
}

extern "C" int phg_hip_mat_rhs(const phg_hip_mat *A, phg_hip_vec *b)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (A == nullptr || b == nullptr || b->n != A->n)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_mat_rhs: size mismatch");
    PHG_HIP_CHECK(hipMemcpyAsync(b->d, A->d_rhs, (size_t)A->n * sizeof(double),
                                 hipMemcpyDeviceToDevice, phg::g_ctx.stream));
    return PHG_HIP_OK;
}
