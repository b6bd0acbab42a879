// Evaluates skt::powNear1 and skt::powPos (skewtcloudsim_amd/csrc/skt_device.h)
// on the device for host arrays of bases and exponents, so that the tests
// can measure them against the exact power. Test infrastructure, built by
// __graft_entry__.build() next to this file.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../skewtcloudsim_amd/csrc/skt_device.h"

__global__ __launch_bounds__(256) void k_pow_eval(const float* x, const float* y, float* near1, float* pos, long n) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    near1[i] = skt::powNear1(x[i], y[i]);
    pos[i] = skt::powPos(x[i], y[i]);
}

// n pairs (x[i], y[i]) -> near1[i], pos[i]; returns 0 on success.
extern "C" int pow_eval(const float* x, const float* y, float* near1, float* pos, long n) {
    if (n <= 0) return 0;
    float* d = nullptr;
    const size_t b = (size_t)n * sizeof(float);
    if (hipMalloc((void**)&d, 4 * b) != hipSuccess) return 1;
    int rc = 0;
    if (hipMemcpy(d, x, b, hipMemcpyHostToDevice) != hipSuccess || hipMemcpy(d + n, y, b, hipMemcpyHostToDevice) != hipSuccess) {
        rc = 1;
    } else {
        const unsigned nb = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_pow_eval, dim3(nb), dim3(256), 0, 0, d, d + n, d + 2 * n, d + 3 * n, n);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = 2;
        else if (hipMemcpy(near1, d + 2 * n, b, hipMemcpyDeviceToHost) != hipSuccess ||
                 hipMemcpy(pos, d + 3 * n, b, hipMemcpyDeviceToHost) != hipSuccess)
            rc = 1;
    }
    (void)hipFree(d);
    return rc;
}
