// Build-time compile check of odemi_device.h with a sample traced problem
// (Lorenz, n = 3): instantiates every lane kernel the runtime can request.
#include <hip/hip_runtime.h>
#define ODEMI_N 3
#define ODEMI_NP 3
template <typename T>
__device__ __forceinline__ void odemi_f(T* __restrict__ du, const T* __restrict__ u, const T* __restrict__ p, T t) {
    du[0] = p[0] * (u[1] - u[0]);
    du[1] = u[0] * (p[1] - u[2]) - u[1];
    du[2] = u[0] * u[1] - p[2] * u[2];
}
template <typename T>
__device__ __forceinline__ void odemi_jac(T* __restrict__ J, const T* __restrict__ u, const T* __restrict__ p, T t) {
    J[0] = -p[0]; J[1] = p[0]; J[2] = T(0);
    J[3] = p[1] - u[2]; J[4] = T(-1); J[5] = -u[0];
    J[6] = u[1]; J[7] = u[0]; J[8] = -p[2];
}
template <typename T>
__device__ __forceinline__ void odemi_tgrad(T* __restrict__ dT, const T* __restrict__ u, const T* __restrict__ p, T t) {
    dT[0] = T(0); dT[1] = T(0); dT[2] = T(0);
}
#include "odemi_device.h"
extern "C" __global__ void __launch_bounds__(256) k1(odemi::Params P) { odemi::tsit5_persistent<false, double, double, 3>(P); }
extern "C" __global__ void __launch_bounds__(256) k2(odemi::Params P) { odemi::tsit5_persistent<false, double, float, 3>(P); }
extern "C" __global__ void __launch_bounds__(256) k3(odemi::Params P) { odemi::tsit5_plain_init<3>(P); }
extern "C" __global__ void __launch_bounds__(256) k4(odemi::Params P) { odemi::tsit5_plain_step<3>(P); }
extern "C" __global__ void __launch_bounds__(256) k5(odemi::Params P) { odemi::rodas_persistent<double, 3>(P); }
extern "C" __global__ void __launch_bounds__(256) k6(odemi::Params P) { odemi::rodas_persistent<float, 3>(P); }
extern "C" __global__ void __launch_bounds__(256) k7(odemi::Params P) { odemi::rodas_plain_init<3>(P); }
extern "C" __global__ void __launch_bounds__(256) k8(odemi::Params P) { odemi::rodas_plain_step<3>(P); }
