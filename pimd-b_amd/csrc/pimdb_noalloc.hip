// pimdb_noalloc.hip -- the exchange forward pass and its adjoint on caller
// device arrays with no device allocation: the kernels of
// pimdb_xplan_forward / pimdb_xplan_vb_grad, with their workspace (V, Vb, the
// ring's link prefix sums and the overflow flag) in a fixed device array of
// this file instead of buffers a plan allocates.  For callers that cannot
// allocate (code under graph capture, serving loops).
//
// Every GPU holds its own copy of the array until the process ends, at its own
// address, looked up once per device with that device current.  Calls on one
// GPU share it: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pimdb_gpu.h"
#include "ctx_common.hpp"

using namespace pimdb;

namespace {

// Largest ring: V and Vb (N+1 each), Shi and Slo (N each), then the flag.
constexpr int NOALLOC_MAX_N = 2048;
constexpr size_t NOALLOC_FLAG = 4 * (size_t)(NOALLOC_MAX_N + 1);
constexpr size_t NOALLOC_DOUBLES = NOALLOC_FLAG + 1;

__device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The flag starts every call at zero (a fill kernel, so that the call can be
// recorded into a graph).
__global__ void k_clear_flag(int* flag) {
    if (threadIdx.x == 0) *flag = 0;
}

// Device address of this device's scratch, looked up on its first use.
int scratch_base(double** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return PIMDB_E_HIP;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_create_error("pimdb_exchange_vb_grad_noalloc: device index beyond " +
                         std::to_string(NOALLOC_MAX_DEVICES));
        return PIMDB_E_HIP;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_create_error(std::string("pimdb_exchange_vb_grad_noalloc: hipGetSymbolAddress: ") +
                             hipGetErrorString(e));
            return PIMDB_E_HIP;
        }
        scratch_ptr[dev] = (double*)p;
    }
    *out = scratch_ptr[dev];
    return PIMDB_OK;
}

template <int D>
void launch(int N, double bt, double k, int pbc, double L, const double* x1, const double* xP, double* V,
            double* Vb, double* Shi, double* Slo, int* flag, double* grad, hipStream_t s) {
    hipLaunchKernelGGL(k_exchange_recursions<D>, dim3(2), dim3(1024), rec_lds_bytes(N, D), s, N, bt, k, pbc, L, x1,
                       xP, V, Vb, Shi, Slo, flag);
    if (grad)
        hipLaunchKernelGGL(k_exchange_vb_grad<D>, dim3((2 * N + 3) / 4), dim3(256), 0, s, N, bt, k, pbc, L, Shi,
                           Slo, V, Vb, x1, xP, grad);
}

}  // namespace

extern "C" int pimdb_exchange_vb_grad_noalloc(int natoms, int ndim, double beta_t, double spring_k, int pbc,
                                              double size, const double* x1, const double* xP, double* VB,
                                              double* grad, void* stream) {
    if (natoms < 1 || ndim < 1 || ndim > 3 || !(beta_t > 0.0) || (pbc && !(size > 0.0)) || !x1 || !xP) {
        set_create_error("pimdb_exchange_vb_grad_noalloc: invalid argument");
        return PIMDB_E_ARG;
    }
    if (natoms > NOALLOC_MAX_N) {
        set_create_error("pimdb_exchange_vb_grad_noalloc: natoms " + std::to_string(natoms) +
                         " exceeds the no-allocation maximum of " + std::to_string(NOALLOC_MAX_N));
        return PIMDB_E_ARG;
    }
    if (rec_lds_bytes(natoms, ndim) > device_lds_limit()) {
        set_create_error("pimdb_exchange_vb_grad_noalloc: natoms " + std::to_string(natoms) +
                         " exceeds the recursion's LDS workspace");
        return PIMDB_E_ARG;
    }
    double* base = nullptr;
    int rc = scratch_base(&base);
    if (rc) return rc;
    const int N = natoms;
    double* V = base;
    double* Vb = V + (N + 1);
    double* Shi = Vb + (N + 1);
    double* Slo = Shi + N;
    int* flag = (int*)(base + NOALLOC_FLAG);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, flag);
    switch (ndim) {
    case 1: launch<1>(N, beta_t, spring_k, pbc, size, x1, xP, V, Vb, Shi, Slo, flag, grad, s); break;
    case 2: launch<2>(N, beta_t, spring_k, pbc, size, x1, xP, V, Vb, Shi, Slo, flag, grad, s); break;
    default: launch<3>(N, beta_t, spring_k, pbc, size, x1, xP, V, Vb, Shi, Slo, flag, grad, s); break;
    }
    if (hipGetLastError() != hipSuccess) return PIMDB_E_HIP;
    if (VB && hipMemcpyAsync(VB, V + N, sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess) return PIMDB_E_HIP;
    int st = 0;
    if (hipMemcpyAsync(&st, flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return PIMDB_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return PIMDB_E_HIP;
    if (st) {
        set_create_error("pimdb_exchange_vb_grad_noalloc: non-finite bosonic exchange potential");
        return PIMDB_E_NONFINITE;
    }
    return PIMDB_OK;
}
