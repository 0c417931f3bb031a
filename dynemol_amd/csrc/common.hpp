// Shared definitions for the dynemol_hip gfx950 library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace dh {

void set_error(const char* fmt, ...);
const char* last_error();

struct cplx {
    double re, im;
};

__host__ __device__ inline cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// Wavefront-wide sum (64 lanes) with DPP-free shuffles; result valid in every lane.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline float wave_sum_f(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace dh

#define DH_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            dh::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_));      \
            return DH_ERR_HIP;                                                                    \
        }                                                                                         \
    } while (0)
