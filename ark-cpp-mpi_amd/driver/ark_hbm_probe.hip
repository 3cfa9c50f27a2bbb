// HBM bandwidth probe: what plain streaming kernels reach on this GPU, the
// practical ceiling for the stencil kernels (whose roofline is quoted against
// the published 8 TB/s).  Copy (plain and non-temporal), read-only and
// write-only over two 4 GiB buffers, 16-byte accesses, several launch shapes;
// prints one line per case and, last, one JSON line with the best of each.
#include <hip/hip_runtime.h>

#include <cstdio>

typedef float v4 __attribute__((ext_vector_type(4)));

__global__ void k_copy(const v4* __restrict__ a, v4* __restrict__ b, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}
__global__ void k_copy_nt(const v4* __restrict__ a, v4* __restrict__ b, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}
__global__ void k_read(const v4* __restrict__ a, float* out, long long n) {
    float s = 0.f;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const v4 v = a[i];
        s += v.x + v.y + v.z + v.w;
    }
    if (s == 12345.f) out[0] = s;   // keeps the loads; the buffer holds zeros
}
__global__ void k_write(v4* __restrict__ b, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = v4{1.f, 2.f, 3.f, 4.f};
}

int main() {
    const long long bytes = 4LL << 30, n = bytes / 16;
    v4 *a = nullptr, *b = nullptr;
    float* o = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess || hipMalloc(&o, 64) != hipSuccess) {
        std::printf("hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(a, 0, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    const char* name[4] = {"copy", "copy_nt", "read", "write"};
    double best[4] = {0, 0, 0, 0};
    for (int grid : {1024, 4096, 16384})
        for (int blk : {256, 512})
            for (int kind = 0; kind < 4; ++kind) {
                auto run = [&]() {
                    if (kind == 0) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(blk), 0, 0, a, b, n);
                    else if (kind == 1) hipLaunchKernelGGL(k_copy_nt, dim3(grid), dim3(blk), 0, 0, a, b, n);
                    else if (kind == 2) hipLaunchKernelGGL(k_read, dim3(grid), dim3(blk), 0, 0, a, o, n);
                    else hipLaunchKernelGGL(k_write, dim3(grid), dim3(blk), 0, 0, b, n);
                };
                run();
                run();
                if (hipDeviceSynchronize() != hipSuccess) return 1;
                (void)hipEventRecord(e0);
                for (int r = 0; r < 5; ++r) run();
                (void)hipEventRecord(e1);
                if (hipEventSynchronize(e1) != hipSuccess) return 1;
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                ms /= 5.f;
                const double gbs = (kind < 2 ? 2.0 : 1.0) * bytes / ms / 1e6;
                if (gbs > best[kind]) best[kind] = gbs;
                std::printf("%-8s grid %5d block %3d: %.3f ms  %.0f GB/s\n", name[kind], grid, blk, ms, gbs);
                std::fflush(stdout);
            }
    std::printf("{\"copy_GBs\": %.0f, \"copy_nt_GBs\": %.0f, \"read_GBs\": %.0f, \"write_GBs\": %.0f, \"bytes\": %lld}\n",
                best[0], best[1], best[2], best[3], bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(o);
    return 0;
}
