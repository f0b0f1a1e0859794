// Minimal host stand-in for the HIP device header, so the device-side DFT
// headers (jaxmaterials_amd/csrc/fft_lds.hpp, fft_reg.hpp) compile with g++ as
// plain host code for tests/test_host_dft.py.  Test infrastructure only.
#pragma once
#include <cmath>
#include <utility>
#define __device__
#define __host__
#define __forceinline__ inline
#define __global__
#define __shared__ static
#define __launch_bounds__(x)
#define __restrict__
struct float2 { float x, y; };
struct double2 { double x, y; };
struct float4 { float x, y, z, w; };
inline float2 make_float2(float a, float b) { return {a, b}; }
inline double2 make_double2(double a, double b) { return {a, b}; }
inline unsigned __umulhi(unsigned a, unsigned b) { return (unsigned)(((unsigned long long)a * b) >> 32); }
struct Dim3 { unsigned x = 0, y = 0, z = 0; };
static Dim3 threadIdx, blockIdx, blockDim, gridDim;
inline void __syncthreads() {}
inline double __shfl_xor(double v, int, int) { return v; }
inline float __shfl_xor(float v, int, int) { return v; }
