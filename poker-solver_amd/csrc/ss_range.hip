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

// The same update with P(action | hand) taken on the device from a resident
// solve's strategy sums: for each of the acting player's nsolver hands the
// average strategy exactly as the host output normalises it (street_solve.cu:
// 1597-1610: positive parts summed in action order in float32, times the
// float32 reciprocal, uniform when the sum is 0), kept when above `threshold`
// and 0 otherwise (the 0.001 cut of python/street_solver_gpu.py:484 and the
// .get(label, 0.0) of :550-553; action < 0 = a label the node does not
// have: 0 for every hand). Hands of the narrowed range that the solve does
// not hold (solver_hand < 0) take 1 / max(number of distinct probabilities,
// 2), as range_narrowing.py:83-84 computes it over the probability dict.
__global__ __launch_bounds__(RANGE_BLOCK) void k_range_from_strategy(const float *__restrict__ SS, int row0, int na,
                                                                     int action, int Mp, int nsolver,
                                                                     const int *__restrict__ solver_hand,
                                                                     double *__restrict__ w, double floor_w,
                                                                     double threshold, int n) {
    __shared__ double prob[SS_RANGE_MAX_HANDS];
    __shared__ int cpart[RANGE_BLOCK / 64];
    __shared__ double part[RANGE_BLOCK / 64];
    __shared__ double wmax, pmiss;
    for (int j = threadIdx.x; j < nsolver; j += RANGE_BLOCK) {
        float sum = 0.f;
        for (int a = 0; a < na; ++a) {
            float v = SS[(size_t)(row0 + a) * Mp + j];
            sum += v > 0.f ? v : 0.f;
        }
        float avg = 0.f;
        if (action >= 0) {
            float inv = sum > 0.f ? 1.0f / sum : 1.0f / (float)na;
            float v = SS[(size_t)(row0 + action) * Mp + j];
            if (v < 0.f) v = 0.f;
            avg = sum > 0.f ? v * inv : inv;
        }
        const double d = (double)avg;
        prob[j] = d > threshold ? d : 0.0;
    }
    __syncthreads();
    // distinct values: a hand counts when no earlier hand has its value
    int cnt = 0;
    for (int j = threadIdx.x; j < nsolver; j += RANGE_BLOCK) {
        const double x = prob[j];
        int first = 1;
        for (int k = 0; k < j; ++k)
            if (prob[k] == x) {
                first = 0;
                break;
            }
        cnt += first;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0) cpart[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int i = 0; i < RANGE_BLOCK / 64; ++i) t += cpart[i];
        pmiss = 1.0 / (double)(t > 2 ? t : 2);
    }
    __syncthreads();
    double nw[2] = {0.0, 0.0};
    double m = -1.0;
    for (int k = 0; k < 2; ++k) {
        int h = threadIdx.x + k * RANGE_BLOCK;
        if (h >= n) break;
        const int sh = solver_hand[h];
        const double old = w[h];
        const double pa = sh >= 0 ? prob[sh] : pmiss;
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

int range_update_from_strategy(const float *d_SS, int row0, int na, int action, int Mp, int nsolver,
                               const int *solver_hand, double *weights, int n, double floor_w, double threshold,
                               hipStream_t st) {
    double *dw = nullptr;
    int *dsh = nullptr;
    hipError_t e = hipMalloc((void **)&dw, (size_t)n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&dsh, (size_t)n * sizeof(int));
    if (e == hipSuccess) e = hipMemcpyAsync(dw, weights, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dsh, solver_hand, (size_t)n * sizeof(int), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        k_range_from_strategy<<<1, RANGE_BLOCK, 0, st>>>(d_SS, row0, na, action, Mp, nsolver, dsh, dw, floor_w,
                                                          threshold, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(weights, dw, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (dw) (void)hipFree(dw);
    if (dsh) (void)hipFree(dsh);
    if (e != hipSuccess) return set_error(SS_ERR_HIP, "range update from strategy: %s", hipGetErrorString(e));
    return SS_OK;
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
