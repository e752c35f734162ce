// lanes.h — cross-lane helpers shared by the sweep kernels (device code only).
#pragma once

#include <hip/hip_runtime.h>

namespace e2e {

// lane l takes lane l-n of its row of 16 (row_shr:n); lanes l < n of the row keep `id`
__device__ __forceinline__ int dpp_row_shr(int v, int n, int id) {
    switch (n) {
        case 1: return __builtin_amdgcn_update_dpp(id, v, 0x111, 0xf, 0xf, false);
        case 2: return __builtin_amdgcn_update_dpp(id, v, 0x112, 0xf, 0xf, false);
        case 4: return __builtin_amdgcn_update_dpp(id, v, 0x114, 0xf, 0xf, false);
        default: return __builtin_amdgcn_update_dpp(id, v, 0x118, 0xf, 0xf, false);
    }
}
__device__ __forceinline__ float dpp_row_shr(float v, int n, float id) {
    return __int_as_float(dpp_row_shr(__float_as_int(v), n, __float_as_int(id)));
}
__device__ __forceinline__ double dpp_row_shr(double v, int n, double id) {
    const long long b = __double_as_longlong(v), ib = __double_as_longlong(id);
    const int lo = dpp_row_shr((int)(b & 0xffffffffll), n, (int)(ib & 0xffffffffll));
    const int hi = dpp_row_shr((int)(b >> 32), n, (int)(ib >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

}  // namespace e2e
