// Build-time compile check of odemi_gj.h: gj_load, gj_invert and gj_solve at
// every tile size (the form is chosen by ODEMI_GJ_MFMA on the command line).
#include <hip/hip_runtime.h>
#define ODEMI_N 128
#include "odemi_gj.h"

template <int R>
__device__ void gj_check_r(const float* J, int n, double* x) {
    extern __shared__ float sh[];
    int* pk = (int*)sh;
    float* gjbuf = sh + 128;
    odemi::GJTile<R> G;
    odemi::gj_load<R>(G, J, n, 1.0);
    odemi::gj_invert<R>(G, n, pk, gjbuf);
    odemi::gj_solve<R>(G, n, pk, x, gjbuf + odemi::gj_zbuf_off<R>());
}

extern "C" __global__ void __launch_bounds__(256) gj_check(const float* J, int n, double* x) {
    switch ((n + 15) / 16) {
        case 1: gj_check_r<1>(J, n, x); break;
        case 2: gj_check_r<2>(J, n, x); break;
        case 3: gj_check_r<3>(J, n, x); break;
        case 4: gj_check_r<4>(J, n, x); break;
        case 5: gj_check_r<5>(J, n, x); break;
        case 6: gj_check_r<6>(J, n, x); break;
        case 7: gj_check_r<7>(J, n, x); break;
        default: gj_check_r<8>(J, n, x); break;
    }
}
