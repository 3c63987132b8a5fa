// copy_calibration.hip -- a plain copy kernel of known size (256 MiB read,
// 256 MiB written per launch) used to calibrate rocprofv3's FETCH_SIZE and
// WRITE_SIZE counters: measured / expected = the scale factor applied to the
// solver kernels' counter readings.
#include <hip/hip_runtime.h>
#include <cstdio>

__global__ void copy_f4(const float4* __restrict__ a, float4* __restrict__ b, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

int main()
{
    const size_t bytes = 256ull << 20, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    (void)hipMemset(a, 1, bytes);
    (void)hipMemset(b, 0, bytes);
    for (int r = 0; r < 3; ++r) copy_f4<<<(unsigned)((n + 255) / 256), 256>>>(a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) return 2;
    std::printf("copied %zu bytes x 3\n", bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
