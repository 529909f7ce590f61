// Context lifetime, error latching, scratch memory, timing and grid fills.
#include <string.h>

#include "pic_common.hpp"
#include "pic_internal.hpp"

namespace pic {

void* scratch(pic_context* ctx, size_t bytes)
{
    if (bytes <= ctx->scratch_bytes) return ctx->scratch;
    // No allocation inside a graph capture, and no reallocation while a
    // captured graph still points into the buffer.
    if (ctx->capturing || ctx->live_graphs > 0) return nullptr;
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
    if (!ctx->timing || ctx->capturing) return;
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

// Grid-window fills with zeros.  Kernels rather than hipMemsetAsync: windows
// are strided, and memset nodes misbehave in captured graphs.

// Strided window: x runs along k (one wave per 64 nodes), y along j, z over
// (i, component).
template <typename T>
__global__ __launch_bounds__(256) void k_zero_grid3(GView<T> g)
{
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int i = blockIdx.z % g.n[0];
    const int c = blockIdx.z / g.n[0];
    if (k < g.n[2] && j < g.n[1])
        g.data[int64_t(i) * g.s[0] + int64_t(j) * g.s[1] + int64_t(k) * g.s[2] +
               int64_t(c) * g.cs] = T(0);
}

// Window whose elements form one dense range: 16-byte vector stores.
__global__ __launch_bounds__(256) void k_zero_flat(uint4* p, int64_t n16, char* tail, int ntail)
{
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n16;
         q += int64_t(gridDim.x) * blockDim.x)
        p[q] = z;
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

bool dense_window(const pic_grid& g, int ncomp, int64_t* count)
{
    int64_t ext[4] = {g.n[0], g.n[1], g.n[2], ncomp};
    int64_t str[4] = {g.stride[0], g.stride[1], g.stride[2], ncomp > 1 ? g.cstride : 0};
    int order[4] = {0, 1, 2, 3};
    for (int a = 0; a < 4; ++a)
        for (int b = a + 1; b < 4; ++b)
            if (str[order[b]] < str[order[a]]) std::swap(order[a], order[b]);
    int64_t expect = 1, total = 1;
    for (int q = 0; q < 4; ++q) {
        const int d = order[q];
        if (ext[d] == 1) continue;
        if (str[d] != expect) return false;
        expect *= ext[d];
        total *= ext[d];
    }
    *count = total;
    return true;
}

int zero_grid(pic_context* ctx, int dtype, const pic_grid& g, int ncomp)
{
    TimedLaunch tl(ctx, "zero_grid");
    const size_t esz = dtype == PIC_F64 ? 8 : 4;
    int64_t count = 0;
    if (dense_window(g, ncomp, &count) && (reinterpret_cast<uintptr_t>(g.data) & 15) == 0) {
        const int64_t bytes = count * int64_t(esz);
        const int64_t n16 = bytes / 16;
        const unsigned blocks = (unsigned)std::max<int64_t>(
            1, std::min<int64_t>((n16 + 255) / 256, 8192));
        k_zero_flat<<<blocks, 256, 0, ctx->stream>>>(static_cast<uint4*>(g.data), n16,
                                                      static_cast<char*>(g.data) + n16 * 16,
                                                      int(bytes - n16 * 16));
        return launch_status();
    }
    const dim3 block(64, 4);
    const dim3 grid((g.n[2] + 63) / 64, (g.n[1] + 3) / 4, unsigned(g.n[0]) * unsigned(ncomp));
    if (dtype == PIC_F64)
        k_zero_grid3<double><<<grid, block, 0, ctx->stream>>>(make_gview<double>(g));
    else
        k_zero_grid3<float><<<grid, block, 0, ctx->stream>>>(make_gview<float>(g));
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
    if (hipDeviceGetAttribute(&ctx->num_cu, hipDeviceAttributeMultiprocessorCount, device) !=
        hipSuccess) {
        delete ctx;
        return PIC_ERR_HIP;
    }
    if (hipMalloc(&ctx->d_err, sizeof(int)) != hipSuccess) {
        delete ctx;
        return PIC_ERR_ALLOC;
    }
    if (hipMemset(ctx->d_err, 0, sizeof(int)) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess || noalloc_setup(device) != PIC_OK) {
        (void)hipFree(ctx->d_err);
        delete ctx;
        return PIC_ERR_HIP;
    }
    *out = ctx;
    return PIC_OK;
}

int pic_context_destroy(pic_context* ctx)
{
    if (!ctx || ctx->live_graphs > 0) return PIC_ERR_ARG;
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
    if (ctx->capturing) return static_cast<hipStream_t>(stream) == ctx->stream ? PIC_OK : PIC_ERR_ARG;
    ctx->stream = static_cast<hipStream_t>(stream);
    return PIC_OK;
}

int pic_synchronize(pic_context* ctx)
{
    if (!ctx || ctx->capturing) return PIC_ERR_ARG;
    DeviceGuard g(ctx->device);
    PIC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    int err = 0;
    PIC_HIP_CHECK(hipMemcpy(&err, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        PIC_HIP_CHECK(hipMemset(ctx->d_err, 0, sizeof(int)));
        PIC_HIP_CHECK(hipDeviceSynchronize());
        return (err & ~ERRBIT_CAPACITY) ? PIC_ERR_BOUNDS : PIC_ERR_CAPACITY;
    }
    return PIC_OK;
}

int pic_set_option(pic_context* ctx, int option, int value)
{
    if (!ctx) return PIC_ERR_ARG;
    if (option == PIC_OPT_G2P_GROUP) {
        if (value != 0 && value != 128 && value != 256) return PIC_ERR_ARG;
        ctx->g2p_group = value;
        return PIC_OK;
    }
    return PIC_ERR_ARG;
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

// ---------------------------------------------------------------------------
// Step graphs: capture of the context's calls into a HIP graph.
struct pic_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
    pic_context* ctx = nullptr;  // whose scratch the captured kernels use
};

int pic_graph_begin(pic_context* ctx)
{
    if (!ctx || ctx->capturing || !ctx->stream) return PIC_ERR_ARG;
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    PIC_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    return PIC_OK;
}

int pic_graph_end(pic_context* ctx, pic_graph** out)
{
    if (!ctx || !ctx->capturing || !out) return PIC_ERR_ARG;
    *out = nullptr;
    DeviceGuard g(ctx->device);
    ctx->capturing = false;
    if (!g.ok) return PIC_ERR_HIP;
    hipGraph_t graph = nullptr;
    PIC_HIP_CHECK(hipStreamEndCapture(ctx->stream, &graph));
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return PIC_ERR_HIP;
    }
    pic_graph* pg = new pic_graph();
    pg->graph = graph;
    pg->exec = exec;
    pg->device = ctx->device;
    pg->ctx = ctx;
    ++ctx->live_graphs;
    *out = pg;
    return PIC_OK;
}

int pic_graph_launch(pic_context* ctx, const pic_graph* graph)
{
    if (!ctx || !graph || ctx->capturing || graph->ctx != ctx) return PIC_ERR_ARG;
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    PIC_HIP_CHECK(hipGraphLaunch(graph->exec, ctx->stream));
    return PIC_OK;
}

int pic_graph_destroy(pic_graph* graph)
{
    if (!graph) return PIC_ERR_ARG;
    DeviceGuard g(graph->device);
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    if (graph->ctx) --graph->ctx->live_graphs;
    delete graph;
    return PIC_OK;
}
