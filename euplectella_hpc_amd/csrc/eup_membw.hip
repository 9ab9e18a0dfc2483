// HBM bandwidth probe: copy, read-only and write-only streams of a given size (measurement
// utility; no reference counterpart).  Used to state the achievable fraction of the MI355X HBM
// peak next to the hot kernel's roofline, and as the known-size kernel that calibrates the
// FETCH_SIZE / WRITE_SIZE counters of rocprofv3.
#include "eup_internal.h"

namespace eup {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

// one 16-byte vector per lane, one launch over the whole buffer (the fastest streaming form on
// MI355X; grid-stride loops and several vectors per lane are slower)
__global__ __launch_bounds__(256) void k_copy4(const v4f *__restrict__ a, v4f *__restrict__ b, int64_t n4)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) __builtin_nontemporal_store(a[i], &b[i]);
}

__global__ __launch_bounds__(256) void k_read4(const v4f *__restrict__ a, int64_t n4, float *__restrict__ sink)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n4) return;
    const v4f v = a[i];
    const float acc = v.x + v.y + v.z + v.w;
    if (acc == 1234.5f) sink[blockIdx.x & 4095] = acc; // keeps the loads; never true for the fill value
}

__global__ __launch_bounds__(256) void k_write4(v4f *__restrict__ b, int64_t n4)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const v4f v = {0.f, 0.f, 0.f, 0.f};
    if (i < n4) __builtin_nontemporal_store(v, &b[i]);
}

} // namespace
} // namespace eup

using namespace eup;

extern "C" int eup_membw(int32_t device, int64_t bytes, int32_t reps, double out[3])
{
    if (!out || bytes < (1 << 20) || bytes > ((int64_t)1 << 40) || reps < 1)
        return fail(nullptr, EUP_EINVAL, "eup_membw: bad argument");
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, EUP_EHIP, "eup_membw: hipSetDevice");
    const int64_t n4 = bytes / 16;
    v4f *a = nullptr, *b = nullptr;
    float *sink = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&a, n4 * 16) != hipSuccess || hipMalloc((void **)&b, n4 * 16) != hipSuccess ||
        hipMalloc((void **)&sink, 1 << 16) != hipSuccess || hipStreamCreate(&st) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail(nullptr, EUP_EHIP, "eup_membw: allocation failed");
    const unsigned grid = (unsigned)((n4 + 255) / 256);
    for (int kind = 0; kind < 3 && !rc; ++kind) {
        auto launch = [&]() {
            if (kind == 0) hipLaunchKernelGGL(k_copy4, dim3(grid), dim3(256), 0, st, a, b, n4);
            if (kind == 1) hipLaunchKernelGGL(k_read4, dim3(grid), dim3(256), 0, st, a, n4, sink);
            if (kind == 2) hipLaunchKernelGGL(k_write4, dim3(grid), dim3(256), 0, st, b, n4);
        };
        if (kind == 0) hipLaunchKernelGGL(k_write4, dim3(grid), dim3(256), 0, st, a, n4); // defined input
        launch();
        hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) launch();
        hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(nullptr, EUP_EHIP, "eup_membw: kernel failed");
            break;
        }
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        const double moved = (kind == 0 ? 2.0 : 1.0) * (double)n4 * 16.0 * reps;
        out[kind] = moved / (ms / 1e3) / 1e9;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    hipFree(a);
    hipFree(b);
    hipFree(sink);
    return rc;
}
