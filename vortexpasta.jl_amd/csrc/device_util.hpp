// Device helpers shared by the short-range and learned-correction kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace vpbs {

// Minimum-image separation along one period (Filaments/min_distance.jl:54-66): two
// selects, ≥ L/2 → −L, < −L/2 → +L.
__device__ inline double fold_sep(double r, double L, double Lh) {
    r = (r >= Lh) ? r - L : r;
    r = (r < -Lh) ? r + L : r;
    return r;
}

// r² rounded exactly as (x² + y²) + z² (no contraction), as in the reference.
__device__ inline double dist2_nocontract(double a, double b, double c) {
#pragma clang fp contract(off)
    double s = a * a + b * b;
    return s + c * c;
}

__device__ inline int cell_index(double x, double L, int N) { return (int)((x / L) * (double)N); }

// Inclusive prefix sum over the 64 lanes of a wavefront with DPP row operations: shifts by
// 1, 2, 4 and 8 lanes within each row of 16, then the last lane of row 0 (row 2) added to
// row 1 (row 3), then lane 31 added to rows 2 and 3. Lanes a shift does not reach add 0.
__device__ inline int wave_incl_scan(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return v;
}

// Sum over the 64 lanes in a fixed order (the DPP steps of wave_incl_scan on the two halves of
// each double); the total is in lane 63.
__device__ inline double wave_sum_f64_last(double v) {
    auto step = [&v](auto ctrl, auto rows) {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)b, decltype(ctrl)::value, decltype(rows)::value, 0xf, false);
        const int hi =
            __builtin_amdgcn_update_dpp(0, (int)(b >> 32), decltype(ctrl)::value, decltype(rows)::value, 0xf, false);
        v += __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
    };
    step(std::integral_constant<int, 0x111>{}, std::integral_constant<int, 0xf>{});  // row_shr:1
    step(std::integral_constant<int, 0x112>{}, std::integral_constant<int, 0xf>{});  // row_shr:2
    step(std::integral_constant<int, 0x114>{}, std::integral_constant<int, 0xf>{});  // row_shr:4
    step(std::integral_constant<int, 0x118>{}, std::integral_constant<int, 0xf>{});  // row_shr:8
    step(std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xa>{});  // row_bcast:15 into rows 1, 3
    step(std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xc>{});  // row_bcast:31 into rows 2, 3
    return v;
}

// The same fixed-order sum for FP32 values (one DPP move per step).
__device__ inline float wave_sum_f32_last(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x112, 0xf, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x114, 0xf, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x118, 0xf, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0xa, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x143, 0xc, 0xf, false));
    return v;
}
__device__ inline double wave_sum_last(double v) { return wave_sum_f64_last(v); }
__device__ inline float wave_sum_last(float v) { return wave_sum_f32_last(v); }

// Orders this wavefront's LDS accesses for the compiler (no wait: the LDS executes one
// wavefront's instructions in issue order). For one-wavefront blocks exchanging values
// between lanes through LDS.
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// Inclusive max-scan over the 64 lanes (identity −1), same DPP steps as the sum.
__device__ inline int wave_incl_max(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x118, 0xf, 0xf, false));  // row_shr:8
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x142, 0xa, 0xf, false));  // row_bcast:15 into rows 1, 3
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x143, 0xc, 0xf, false));  // row_bcast:31 into rows 2, 3
    return v;
}

__device__ inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

}  // namespace vpbs
