// Context lifetime, error latching, scratch memory, timing and grid fills.
#include <string.h>

#include "pic_common.hpp"
#include "pic_internal.hpp"

namespace pic {

void* scratch(pic_context* ctx, size_t bytes)
{
    if (bytes <= ctx->scratch_bytes) return ctx->scratch;
    // The stream may still use the old buffer.
    if (ctx->scratch) {
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return nullptr;
        (void)hipFree(ctx->scratch);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    size_t want = bytes + bytes / 4 + (1u << 20);
    if (hipMalloc(&ctx->scratch, want) != hipSuccess) {
        ctx->scratch = nullptr;
        return nullptr;
    }
    ctx->scratch_bytes = want;
    return ctx->scratch;
}

static hipEvent_t take_event(pic_context* ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TimedLaunch::TimedLaunch(pic_context* c, const char* name) : ctx(c), idx(SIZE_MAX)
{
    if (!ctx->timing) return;
    pic_context::Rec r{name, take_event(ctx), take_event(ctx)};
    if (!r.start || !r.stop) return;
    (void)hipEventRecord(r.start, ctx->stream);
    idx = ctx->recs.size();
    ctx->recs.push_back(r);
}

TimedLaunch::~TimedLaunch()
{
    if (idx != SIZE_MAX) (void)hipEventRecord(ctx->recs[idx].stop, ctx->stream);
}

// Grid-window fill with zeros.  A kernel rather than hipMemsetAsync: windows
// are strided, and memset nodes misbehave in captured graphs.
template <typename T>
__global__ void k_zero_grid(GView<T> g, int ncomp)
{
    const int64_t n12 = int64_t(g.n[1]) * g.n[2];
    const int64_t total = int64_t(g.n[0]) * n12 * ncomp;
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * blockDim.x) {
        // k fastest so consecutive lanes write consecutive k for unit stride.
        const int64_t k = t % g.n[2];
        const int64_t r = t / g.n[2];
        const int64_t j = r % g.n[1];
        const int64_t r2 = r / g.n[1];
        const int64_t i = r2 % g.n[0];
        const int64_t c = r2 / g.n[0];
        g.data[i * g.s[0] + j * g.s[1] + k * g.s[2] + c * g.cs] = T(0);
    }
}

int zero_grid(pic_context* ctx, int dtype, const pic_grid& g, int ncomp)
{
    const int64_t total = int64_t(g.n[0]) * g.n[1] * g.n[2] * ncomp;
    unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
    TimedLaunch tl(ctx, "zero_grid");
    if (dtype == PIC_F64)
        k_zero_grid<double><<<blocks, 256, 0, ctx->stream>>>(make_gview<double>(g), ncomp);
    else
        k_zero_grid<float><<<blocks, 256, 0, ctx->stream>>>(make_gview<float>(g), ncomp);
    return launch_status();
}

}  // namespace pic

using namespace pic;

extern "C" {

int pic_api_version(void) { return PIC_API_VERSION; }

const char* pic_status_string(int s)
{
    switch (s) {
    case PIC_OK: return "ok";
    case PIC_ERR_ARG: return "invalid argument";
    case PIC_ERR_BOUNDS: return "stencil or cell outside its window";
    case PIC_ERR_HIP: return "HIP runtime error";
    case PIC_ERR_ALLOC: return "device allocation failed";
    case PIC_ERR_UNSUPPORTED: return "unsupported combination";
    case PIC_ERR_CAPACITY: return "size exceeds a documented maximum";
    default: return "unknown status";
    }
}

int pic_context_create(pic_context** out, int device)
{
    if (!out) return PIC_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return PIC_ERR_ARG;
    DeviceGuard g(device);
    if (!g.ok) return PIC_ERR_HIP;
    pic_context* ctx = new pic_context();
    ctx->device = device;
    if (hipMalloc(&ctx->d_err, sizeof(int)) != hipSuccess) {
        delete ctx;
        return PIC_ERR_ALLOC;
    }
    if (hipMemset(ctx->d_err, 0, sizeof(int)) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(ctx->d_err);
        delete ctx;
        return PIC_ERR_HIP;
    }
    *out = ctx;
    return PIC_OK;
}

int pic_context_destroy(pic_context* ctx)
{
    if (!ctx) return PIC_ERR_ARG;
    DeviceGuard g(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& r : ctx->recs) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->d_err) (void)hipFree(ctx->d_err);
    delete ctx;
    return PIC_OK;
}

int pic_set_stream(pic_context* ctx, void* stream)
{
    if (!ctx) return PIC_ERR_ARG;
    ctx->stream = static_cast<hipStream_t>(stream);
    return PIC_OK;
}

int pic_synchronize(pic_context* ctx)
{
    if (!ctx) return PIC_ERR_ARG;
    DeviceGuard g(ctx->device);
    PIC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    int err = 0;
    PIC_HIP_CHECK(hipMemcpy(&err, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        PIC_HIP_CHECK(hipMemset(ctx->d_err, 0, sizeof(int)));
        PIC_HIP_CHECK(hipDeviceSynchronize());
        return PIC_ERR_BOUNDS;
    }
    return PIC_OK;
}

int pic_set_timing(pic_context* ctx, int enable)
{
    if (!ctx) return PIC_ERR_ARG;
    ctx->timing = enable != 0;
    return PIC_OK;
}

int pic_reset_timing(pic_context* ctx)
{
    if (!ctx) return PIC_ERR_ARG;
    DeviceGuard g(ctx->device);
    PIC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (auto& r : ctx->recs) {
        ctx->event_pool.push_back(r.start);
        ctx->event_pool.push_back(r.stop);
    }
    ctx->recs.clear();
    return PIC_OK;
}

int pic_kernel_time(pic_context* ctx, const char* name, double* ms, int64_t* launches)
{
    if (!ctx || !name || !ms || !launches) return PIC_ERR_ARG;
    DeviceGuard g(ctx->device);
    PIC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double total = 0.0;
    int64_t count = 0;
    for (auto& r : ctx->recs) {
        if (r.name != name) continue;
        float t = 0.f;
        PIC_HIP_CHECK(hipEventElapsedTime(&t, r.start, r.stop));
        total += t;
        ++count;
    }
    *ms = total;
    *launches = count;
    return PIC_OK;
}

}  // extern "C"
