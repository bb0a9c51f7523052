// copy_probe.hip -- a plain device-to-device copy of known size, the
// reference point for reading HBM traffic counters (rocprofv3 --pmc
// FETCH_SIZE / WRITE_SIZE): each launch reads and writes exactly `bytes`.
// With a third argument `widths`, two more kernels of the same size follow:
// k_copy2 (the same copy at 2 bytes per lane) and k_rows2 (every 128-byte row
// read once in a scattered order at 2 bytes per lane, as the relaxation's
// label gathers read, and written in order), to calibrate the counters at
// the access width of the gathers.
//
//   emv-copyprobe [bytes = 1 GiB] [launches = 3] [widths]
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

__global__ void k_copy16(const uint4 *__restrict__ a, uint4 *__restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

__global__ void k_copy2(const uint16_t *__restrict__ a, uint16_t *__restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

// rows is a power of two; i -> (i * odd) mod rows is a permutation, so each
// row is read exactly once.  One wave per row, lane = 16-bit column.
__global__ void k_rows2(const uint16_t *__restrict__ a, uint16_t *__restrict__ b, size_t rows) {
    const int lane = threadIdx.x & 63;
    size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const size_t stride = (size_t)gridDim.x * blockDim.x / 64;
    for (; i < rows; i += stride) {
        const size_t p = (size_t)((uint32_t)i * 2654435761u) & (rows - 1);
        b[i * 64 + lane] = a[p * 64 + lane];
    }
}

__global__ void k_init16(uint4 *__restrict__ a, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) a[i] = make_uint4((uint32_t)i, 1u, 2u, 3u);
}

#define CK(x)                                                                 \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) {                                               \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
            return 1;                                                         \
        }                                                                     \
    } while (0)

int main(int argc, char **argv) {
    const size_t bytes = (argc > 1) ? strtoull(argv[1], nullptr, 10) : ((size_t)1 << 30);
    const int launches = (argc > 2) ? atoi(argv[2]) : 3;
    const size_t n = bytes / sizeof(uint4);
    if (n == 0 || launches < 1) {
        fprintf(stderr, "usage: emv-copyprobe [bytes] [launches]\n");
        return 2;
    }
    uint4 *a = nullptr, *b = nullptr;
    CK(hipMalloc((void **)&a, n * sizeof(uint4)));
    CK(hipMalloc((void **)&b, n * sizeof(uint4)));
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, 0));
    const int grid = prop.multiProcessorCount * 8;
    hipLaunchKernelGGL(k_init16, dim3(grid), dim3(256), 0, 0, a, n);
    hipLaunchKernelGGL(k_init16, dim3(grid), dim3(256), 0, 0, b, n);
    CK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    CK(hipEventRecord(e0, 0));
    for (int r = 0; r < launches; r++) hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, 0, a, b, n);
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / launches;
    printf("copy_probe bytes=%zu launches=%d avg_ms=%.4f read+write GB/s=%.1f\n", n * sizeof(uint4),
           launches, per, 2.0 * (double)(n * sizeof(uint4)) / (per * 1e6));
    if (argc > 3) {
        const size_t n2 = n * sizeof(uint4) / 2, rows = n * sizeof(uint4) / 128;
        for (int r = 0; r < launches; r++)
            hipLaunchKernelGGL(k_copy2, dim3(grid), dim3(256), 0, 0, (const uint16_t *)a,
                               (uint16_t *)b, n2);
        if (rows > 0 && (rows & (rows - 1)) == 0 && rows <= ((size_t)1 << 32)) {
            for (int r = 0; r < launches; r++)
                hipLaunchKernelGGL(k_rows2, dim3(grid), dim3(256), 0, 0, (const uint16_t *)a,
                                   (uint16_t *)b, rows);
        } else {
            fprintf(stderr, "k_rows2 skipped: bytes / 128 must be a power of two\n");
        }
        CK(hipDeviceSynchronize());
        printf("copy_probe widths: k_copy2 and k_rows2 ran %d launches each\n", launches);
    }
    CK(hipFree(a));
    CK(hipFree(b));
    return 0;
}
