// Exhaustive check of skt::faceOfSum (the pressure operator's face
// coefficient with the short division) against the IEEE expression of the
// reference, -max(1, 1.225 / (2 / s)), over float bit patterns; the same for
// skt::recipRNfast against 1.0f / s. Test infrastructure, built by
// __graft_entry__.build() next to this file.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../skewtcloudsim_amd/csrc/skt_device.h"

__global__ __launch_bounds__(256) void k_face_check(uint32_t lo, uint64_t count, unsigned long long* bad,
                                                     uint32_t* first) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256ull) {
        const uint32_t bits = lo + (uint32_t)i;
        const float s = __uint_as_float(bits);
        const float a = skt::faceOfSum(s);
        const float b = -fmaxf(1.0f, 1.225f / (2.0f / s));
        const bool same = __float_as_uint(a) == __float_as_uint(b) || (a != a && b != b);
        if (!same) {
            atomicAdd(bad, 1ull);
            atomicMin(first, bits);
        }
    }
}

// Checks bit patterns lo .. lo+count-1; returns 0 on success.
extern "C" int face_check(uint32_t lo, uint64_t count, unsigned long long* bad_out, uint32_t* first_out) {
    unsigned long long* bad = nullptr;
    uint32_t* first = nullptr;
    if (hipMalloc((void**)&bad, sizeof(*bad)) != hipSuccess) return 1;
    if (hipMalloc((void**)&first, sizeof(*first)) != hipSuccess) return 1;
    const unsigned long long zero = 0;
    const uint32_t big = 0xFFFFFFFFu;
    if (hipMemcpy(bad, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) return 1;
    if (hipMemcpy(first, &big, sizeof(big), hipMemcpyHostToDevice) != hipSuccess) return 1;
    hipLaunchKernelGGL(k_face_check, dim3(16384), dim3(256), 0, 0, lo, count, bad, first);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 2;
    if (hipMemcpy(bad_out, bad, sizeof(*bad), hipMemcpyDeviceToHost) != hipSuccess) return 1;
    if (hipMemcpy(first_out, first, sizeof(*first), hipMemcpyDeviceToHost) != hipSuccess) return 1;
    (void)hipFree(bad);
    (void)hipFree(first);
    return 0;
}

// recipRNfast(s) against the IEEE 1.0f / s for the bit patterns lo ..
// lo+count-1 that lie in its range (others are counted as skipped).
__global__ __launch_bounds__(256) void k_recip_check(uint32_t lo, uint64_t count, unsigned long long* bad,
                                                      uint32_t* first, unsigned long long* skipped) {
    unsigned long long nb = 0, ns = 0;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256ull) {
        const uint32_t bits = lo + (uint32_t)i;
        const float s = __uint_as_float(bits);
        if (!skt::recipOk(s)) {
            ns++;
            continue;
        }
        const float a = skt::recipRNfast(s);
        const float b = 1.0f / s;
        if (__float_as_uint(a) != __float_as_uint(b)) {
            nb++;
            atomicMin(first, bits);
        }
    }
    if (nb) atomicAdd(bad, nb);
    if (ns) atomicAdd(skipped, ns);
}

// recipRNfast over the bit patterns lo..lo+count-1; returns 0 on success.
extern "C" int recip_check(uint64_t lo, uint64_t count, unsigned long long* bad_out, uint32_t* first_out,
                           unsigned long long* skipped_out) {
    unsigned long long* cnt = nullptr;
    uint32_t* first = nullptr;
    if (hipMalloc((void**)&cnt, 2 * sizeof(*cnt)) != hipSuccess) return 1;
    if (hipMalloc((void**)&first, sizeof(*first)) != hipSuccess) return 1;
    const unsigned long long zero[2] = {0, 0};
    const uint32_t big = 0xFFFFFFFFu;
    if (hipMemcpy(cnt, zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) return 1;
    if (hipMemcpy(first, &big, sizeof(big), hipMemcpyHostToDevice) != hipSuccess) return 1;
    hipLaunchKernelGGL(k_recip_check, dim3(16384), dim3(256), 0, 0, (uint32_t)lo, count, cnt, first, cnt + 1);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 2;
    unsigned long long h[2];
    if (hipMemcpy(h, cnt, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 1;
    if (hipMemcpy(first_out, first, sizeof(*first), hipMemcpyDeviceToHost) != hipSuccess) return 1;
    *bad_out = h[0];
    *skipped_out = h[1];
    (void)hipFree(cnt);
    (void)hipFree(first);
    return 0;
}
