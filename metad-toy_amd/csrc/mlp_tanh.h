// tanh of the learned-bias network, shared by its training and inference kernels.
// tanh|z| = (1 - e) / (1 + e) with e = exp(-2|z|) from the hardware base-2 exponential,
// and the odd Taylor polynomial z - z^3/3 + 2 z^5/15 below |z| = 0.05, where 1 - e would
// cancel (the polynomial's truncation there is below 1e-10 relative).  Relative error
// a few FP32 ulp, against the network's BF16 hi + lo products.
#pragma once
#include <hip/hip_runtime.h>

namespace mdt {

__device__ __forceinline__ float mlp_tanh(float z) {
    const float a = fabsf(z);
    const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * a);   // 2 / ln 2
    const float t = (1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e);
    const float z2 = z * z;
    const float p = z * (1.0f + z2 * (-1.0f / 3.0f + z2 * (2.0f / 15.0f)));
    return a < 0.05f ? p : copysignf(t, z);
}

}  // namespace mdt
