// copy_probe.hip — a plain device-to-device copy of a known size, for calibrating the
// FETCH_SIZE / WRITE_SIZE counters that bench.py's roofline traffic figure is read from:
// run it under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE) and divide the counter by the
// bytes it moved.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void kCopy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char** argv) {
    const size_t bytes = (argc > 1 ? (size_t)atoll(argv[1]) : 256ull) << 20;  // MiB
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int reps = 5;
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) kCopy<<<4096, 256>>>(a, b, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    printf("copy_probe bytes_read=%zu bytes_written=%zu per_launch_ms=%.4f GB_s=%.1f\n", bytes, bytes, ms / reps,
           2.0 * bytes / (ms / reps * 1e-3) / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
