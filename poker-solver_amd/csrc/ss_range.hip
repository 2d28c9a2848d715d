// ss_range.hip -- Bayesian range narrowing (python/range_narrowing.py:76-122).
//
// One workgroup: each lane takes hands h = lane, lane + 1024: w * p (IEEE
// float64 product, as the reference's Python float product), the floor, a
// workgroup max (order-free: max is exact), then the division by the max.
#include <hip/hip_runtime.h>

#include "ss_internal.h"
#include "../../include/street_solve_range.h"

namespace ss {

namespace {
constexpr int RANGE_BLOCK = 1024;
}

__global__ __launch_bounds__(RANGE_BLOCK) void k_range_update(double *__restrict__ w,
                                                              const double *__restrict__ p,
                                                              const unsigned char *__restrict__ has,
                                                              double p_missing, double floor_w, int n) {
    __shared__ double part[RANGE_BLOCK / 64];
    __shared__ double wmax;
    double nw[2] = {0.0, 0.0};
    double m = -1.0;
    for (int k = 0; k < 2; ++k) {
        int h = threadIdx.x + k * RANGE_BLOCK;
        if (h >= n) break;
        double old = w[h];
        double pa = (!has || has[h]) ? p[h] : p_missing;
        double v = old * pa;
        if (v < floor_w && old > 0.0) v = floor_w;
        nw[k] = v;
        m = v > m ? v : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = part[0];
        for (int i = 1; i < RANGE_BLOCK / 64; ++i) x = part[i] > x ? part[i] : x;
        wmax = x;
    }
    __syncthreads();
    const double mx = wmax;
    for (int k = 0; k < 2; ++k) {
        int h = threadIdx.x + k * RANGE_BLOCK;
        if (h >= n) break;
        w[h] = mx > 0.0 ? nw[k] / mx : nw[k];
    }
}

}  // namespace ss

extern "C" int ss_range_update(int num_hands, double *weights, const double *p_action, const unsigned char *has_p,
                               double p_missing, double weight_floor) {
    using namespace ss;
    clear_error();
    if (num_hands < 0 || num_hands > SS_RANGE_MAX_HANDS)
        return set_error(SS_ERR_ARG, "num_hands %d (0..%d)", num_hands, SS_RANGE_MAX_HANDS);
    if (num_hands == 0) return SS_OK;
    if (!weights || !p_action) return set_error(SS_ERR_ARG, "NULL weights or p_action");
    double *dw = nullptr, *dp = nullptr;
    unsigned char *dh = nullptr;
    size_t nb = (size_t)num_hands * sizeof(double);
    hipError_t e = hipMalloc((void **)&dw, nb);
    if (e == hipSuccess) e = hipMalloc((void **)&dp, nb);
    if (e == hipSuccess && has_p) e = hipMalloc((void **)&dh, (size_t)num_hands);
    if (e == hipSuccess) e = hipMemcpy(dw, weights, nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, p_action, nb, hipMemcpyHostToDevice);
    if (e == hipSuccess && has_p) e = hipMemcpy(dh, has_p, (size_t)num_hands, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_range_update<<<1, RANGE_BLOCK>>>(dw, dp, dh, p_missing, weight_floor, num_hands);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(weights, dw, nb, hipMemcpyDeviceToHost);
    if (dw) (void)hipFree(dw);
    if (dp) (void)hipFree(dp);
    if (dh) (void)hipFree(dh);
    if (e != hipSuccess) return set_error(SS_ERR_HIP, "range update: %s", hipGetErrorString(e));
    return SS_OK;
}
