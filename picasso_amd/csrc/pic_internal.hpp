// Host-side context, argument validation and launch bookkeeping shared by the
// translation units of the library.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <stdint.h>

#include <algorithm>
#include <string>
#include <map>
#include <utility>
#include <vector>

#include "picasso_amd.h"

struct pic_context {
    int device = 0;
    int num_cu = 0;  // compute units of the device
    hipStream_t stream = nullptr;
    int* d_err = nullptr;  // device-side error word (ERRBIT_*)

    // Growable device scratch, reused across calls on this context.
    void* scratch = nullptr;
    size_t scratch_bytes = 0;

    // Optional per-launch event timing.
    bool timing = false;
    struct Rec {
        std::string name;
        hipEvent_t start, stop;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> event_pool;

    // Between pic_graph_begin and pic_graph_end: calls are being captured
    // (no scratch growth, no timing events, no synchronisation).
    bool capturing = false;
    // Step graphs made from this context and not yet destroyed.  They hold
    // pointers into the scratch buffer, so it may not be reallocated while
    // any is alive.
    int live_graphs = 0;

    // Work-items per workgroup of the collocated g2p brick kernels (0: the
    // measured default of each precision; pic_set_option).
    int g2p_group = 0;

    // Resident workgroups per compute unit, per (kernel, dynamic LDS bytes).
    std::map<std::pair<const void*, size_t>, int> occupancy;
};

namespace pic {

#define PIC_HIP_CHECK(expr)                                \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return PIC_ERR_HIP;          \
    } while (0)

// A roctx range around one library call (rocprofv3 --marker-trace shows
// it): named after the function it replaces, the GridOperator apply with the
// reference's own region name (Picasso_GridOperator.hpp:464).
struct Region {
    explicit Region(const char* name) { roctxRangePush(name); }
    ~Region() { roctxRangePop(); }
    Region(const Region&) = delete;
    Region& operator=(const Region&) = delete;
};

// Make the context's device current for the duration of a call.
struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) return;
        ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Ensure at least `bytes` of scratch; returns nullptr on failure, and when
// the buffer would have to grow during a capture or while a step graph of
// the context is alive (its replays use the current buffer).
void* scratch(pic_context* ctx, size_t bytes);

// Launch bracketing for optional timing.
struct TimedLaunch {
    pic_context* ctx;
    size_t idx;
    TimedLaunch(pic_context* c, const char* name);
    ~TimedLaunch();
};

// Launch error check after a kernel launch.
inline int launch_status()
{
    return hipGetLastError() == hipSuccess ? PIC_OK : PIC_ERR_HIP;
}

inline bool valid_pfield(const pic_pfield* f)
{
    return f && f->data && f->vlen_log2 >= 0 && f->vlen_log2 < 31;
}

inline bool valid_grid(const pic_grid* g)
{
    return g && g->data && g->n[0] > 0 && g->n[1] > 0 && g->n[2] > 0;
}

inline bool valid_entity(int e) { return e >= PIC_NODE && e <= PIC_EDGE_K; }

inline unsigned blocks_for(int64_t n, int threads)
{
    return (unsigned)((n + threads - 1) / threads);
}

// Workgroup count for a persistent launch that walks `items` work items with
// a grid-stride loop: as many workgroups as the device holds at once (from
// the occupancy query, cached per kernel), and at most `items`.
template <typename K>
unsigned persistent_blocks(pic_context* ctx, K kernel, int threads, size_t dyn_lds, int64_t items)
{
    const auto key = std::make_pair(reinterpret_cast<const void*>(kernel), dyn_lds);
    auto it = ctx->occupancy.find(key);
    int per_cu;
    if (it != ctx->occupancy.end()) {
        per_cu = it->second;
    } else {
        per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, dyn_lds) !=
                hipSuccess || per_cu < 1)
            per_cu = 1;
        ctx->occupancy[key] = per_cu;
    }
    const int64_t g = int64_t(per_cu) * std::max(ctx->num_cu, 1);
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, items));
}

// Resolve, once per device and with that device current, the device address
// of the allocation-free sort's scratch array (pic_bin_sort_noalloc).
int noalloc_setup(int device);

// Zero a grid window on the stream (used by PIC_P2G_ZERO_FIRST).
int zero_grid(pic_context* ctx, int dtype, const pic_grid& g, int ncomp);

// Do the window's elements (ncomp components) cover one dense range?  Its
// element count goes to *count.
bool dense_window(const pic_grid& g, int ncomp, int64_t* count);

}  // namespace pic
