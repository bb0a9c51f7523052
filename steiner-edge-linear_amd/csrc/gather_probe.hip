// gather_probe.hip -- the ceiling of a pull relaxation round: random
// 128-byte row gathers (64 lanes x 16-bit labels), eight per item plus the
// item's own row, one row written per item, over label arrays of several
// sizes.  Neighbour ids come from a seeded xorshift, each in [0, rows).
// Prints items/s and the cache-line rate as TB/s; the vertex-wise relaxation
// kernel is compared against it.
//
//   emv-gatherprobe [launches = 10]
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#define CK(x)                                                        \
    do {                                                             \
        hipError_t e_ = (x);                                         \
        if (e_ != hipSuccess) {                                      \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));  \
            return 1;                                                \
        }                                                            \
    } while (0)

// Two items per wave; idx[item * 8 + q] is neighbour q of the item.
__global__ __launch_bounds__(256) void k_gather(const uint16_t *__restrict__ L,
                                                uint16_t *__restrict__ out,
                                                const int32_t *__restrict__ idx, int64_t nitems) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const int64_t nw = (int64_t)gridDim.x * 4;
    for (int64_t it0 = wave * 2; it0 < nitems; it0 += nw * 2) {
        int32_t my = 0;
        if (lane < 16 && it0 * 8 + lane < nitems * 8) my = idx[it0 * 8 + lane];
        uint16_t d[2][8];
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int u = __builtin_amdgcn_readlane(my, t * 8 + q);
                d[t][q] = L[(int64_t)u * 64 + lane];
            }
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int64_t it = it0 + t;
            if (it < nitems) {
                uint16_t m = L[it * 64 + lane];
#pragma unroll
                for (int q = 0; q < 8; q++) m = d[t][q] + 1 < m ? d[t][q] + 1 : m;
                out[it * 64 + lane] = m;
            }
        }
    }
}

int main(int argc, char **argv) {
    const int reps = argc > 1 ? atoi(argv[1]) : 10;
    const int64_t sizes[] = {1 << 18, 1 << 21, 1 << 24};  // 32 MiB, 256 MiB, 2 GiB of labels
    for (int64_t rows : sizes) {
        uint16_t *L = nullptr, *out = nullptr;
        int32_t *idx = nullptr;
        CK(hipMalloc(&L, rows * 128));
        CK(hipMalloc(&out, rows * 128));
        CK(hipMalloc(&idx, rows * 8 * sizeof(int32_t)));
        std::vector<int32_t> h((size_t)rows * 8);
        uint64_t x = 12345;
        for (auto &v : h) {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            v = (int32_t)(x % (uint64_t)rows);
        }
        CK(hipMemcpy(idx, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        CK(hipMemset(L, 1, rows * 128));
        hipEvent_t a, b;
        CK(hipEventCreate(&a));
        CK(hipEventCreate(&b));
        const int grid = 256 * 64;
        for (int w = 0; w < 2; w++)
            hipLaunchKernelGGL(k_gather, dim3(grid), dim3(256), 0, 0, L, out, idx, rows);
        CK(hipEventRecord(a));
        for (int w = 0; w < reps; w++)
            hipLaunchKernelGGL(k_gather, dim3(grid), dim3(256), 0, 0, L, out, idx, rows);
        CK(hipEventRecord(b));
        CK(hipEventSynchronize(b));
        CK(hipGetLastError());
        float ms = 0;
        CK(hipEventElapsedTime(&ms, a, b));
        ms /= reps;
        const double lines = (double)rows * 10.0;  // 9 rows read + 1 written
        printf("labels %5.0f MiB: %.3f ms per launch, %.3g items/s, %.2f TB/s of 128-byte lines\n",
               rows * 128 / 1048576.0, ms, rows / (ms * 1e-3), lines * 128 / (ms * 1e-3) / 1e12);
        fflush(stdout);
        CK(hipEventDestroy(a));
        CK(hipEventDestroy(b));
        CK(hipFree(L));
        CK(hipFree(out));
        CK(hipFree(idx));
    }
    return 0;
}
