// Calibration kernel of the measurement: the VALU issue rate that bounds the splat kernels.
// Every lane runs eight independent FMA chains (FP32 or FP64) and writes one value, so the
// launch is issue-bound; bench.py --workload valu times it.
#include "dsplat_internal.h"

namespace dsplat {

template <typename T>
__global__ __launch_bounds__(256) void valu_chains(T* __restrict__ out, int iters) {
    T a[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = (T)threadIdx.x * (T)1e-3 + (T)i;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = fma(a[i], (T)0.999, (T)0.25);
    }
    T s = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) s += a[i];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace dsplat

using namespace dsplat;

extern "C" int dsplat_calibrate_valu(int32_t fp64, int32_t blocks, int32_t iters, void* out, void* stream) {
    if (blocks <= 0 || iters <= 0 || !out || (fp64 != 0 && fp64 != 1)) return hipErrorInvalidValue;
    if (fp64)
        hipLaunchKernelGGL(valu_chains<double>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           static_cast<double*>(out), iters);
    else
        hipLaunchKernelGGL(valu_chains<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           static_cast<float*>(out), iters);
    return hipGetLastError();
}
