// Copy probe: one plain device-to-device copy kernel of a known size, used
// to calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters (the counters of
// this kernel should read the copied bytes once each way).
//   usage: copy_probe [MiB]   (default 1024)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const double4* __restrict__ src, double4* __restrict__ dst,
                       size_t n)
{
    size_t const stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += stride)
        dst[i] = src[i];
}

static void ok(hipError_t e, const char* what)
{
    if (e != hipSuccess)
    {
        std::fprintf(stderr, "copy_probe: %s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

int main(int argc, char** argv)
{
    size_t const mib = argc > 1 ? (size_t)std::atol(argv[1]) : 1024;
    size_t const bytes = mib << 20;
    size_t const n = bytes / sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    ok(hipMalloc(&a, bytes), "hipMalloc");
    ok(hipMalloc(&b, bytes), "hipMalloc");
    ok(hipMemset(a, 1, bytes), "hipMemset");
    ok(hipMemset(b, 0, bytes), "hipMemset");
    ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hipEvent_t e0, e1;
    ok(hipEventCreate(&e0), "hipEventCreate");
    ok(hipEventCreate(&e1), "hipEventCreate");
    ok(hipEventRecord(e0), "hipEventRecord");
    k_copy<<<2048, 256>>>(a, b, n);
    ok(hipGetLastError(), "k_copy launch");
    ok(hipEventRecord(e1), "hipEventRecord");
    ok(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.0f;
    ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    std::printf("copy_probe: %zu bytes read, %zu bytes written, %.3f ms, "
                "%.1f GB/s (read+write)\n", bytes, bytes, ms,
                2.0 * bytes / (ms * 1e-3) / 1e9);
    ok(hipFree(a), "hipFree");
    ok(hipFree(b), "hipFree");
    return 0;
}
