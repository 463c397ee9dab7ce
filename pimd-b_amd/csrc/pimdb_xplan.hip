// pimdb_xplan.hip -- the exchange on caller-provided device arrays: the
// forward V_B kernel chain and its hand-written adjoint, for the autograd
// harness (pimd-b_amd/pimdb_amd/autograd.py).
#include <hip/hip_runtime.h>

#include <string>

#include "pimdb_gpu.h"
#include "ctx_common.hpp"

using namespace pimdb;

struct pimdb_xplan {
    int N = 0, D = 0, pbc = 0;
    double bt = 0.0, k = 0.0, L = 0.0;
    double *V = nullptr, *Vb = nullptr, *Shi = nullptr, *Slo = nullptr;
    double* ws = nullptr;   // the recursions' workspace in device memory, when it exceeds LDS
    int* status = nullptr;
};

static int xplan_forward_internal(pimdb_xplan* pl, const double* x1, const double* xP, hipStream_t s) {
    const size_t lds = rec_lds_bytes(pl->N, pl->D);
    if (pl->ws) {
        switch (pl->D) {
        case 1: hipLaunchKernelGGL(k_exchange_recursions_gws<1>, dim3(2), dim3(1024), 0, s, pl->N, pl->bt, pl->k, pl->pbc, pl->L, x1, xP, pl->V, pl->Vb, pl->Shi, pl->Slo, pl->status, pl->ws); break;
        case 2: hipLaunchKernelGGL(k_exchange_recursions_gws<2>, dim3(2), dim3(1024), 0, s, pl->N, pl->bt, pl->k, pl->pbc, pl->L, x1, xP, pl->V, pl->Vb, pl->Shi, pl->Slo, pl->status, pl->ws); break;
        default: hipLaunchKernelGGL(k_exchange_recursions_gws<3>, dim3(2), dim3(1024), 0, s, pl->N, pl->bt, pl->k, pl->pbc, pl->L, x1, xP, pl->V, pl->Vb, pl->Shi, pl->Slo, pl->status, pl->ws); break;
        }
    } else switch (pl->D) {
    case 1: hipLaunchKernelGGL(k_exchange_recursions<1>, dim3(2), dim3(1024), lds, s, pl->N, pl->bt, pl->k, pl->pbc, pl->L, x1, xP, pl->V, pl->Vb, pl->Shi, pl->Slo, pl->status); break;
    case 2: hipLaunchKernelGGL(k_exchange_recursions<2>, dim3(2), dim3(1024), lds, s, pl->N, pl->bt, pl->k, pl->pbc, pl->L, x1, xP, pl->V, pl->Vb, pl->Shi, pl->Slo, pl->status); break;
    default: hipLaunchKernelGGL(k_exchange_recursions<3>, dim3(2), dim3(1024), lds, s, pl->N, pl->bt, pl->k, pl->pbc, pl->L, x1, xP, pl->V, pl->Vb, pl->Shi, pl->Slo, pl->status); break;
    }
    if (hipGetLastError() != hipSuccess) return PIMDB_E_HIP;
    int st = 0;
    if (hipMemcpyAsync(&st, pl->status, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return PIMDB_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return PIMDB_E_HIP;
    if (st) {
        (void)hipMemsetAsync(pl->status, 0, sizeof(int), s);
        (void)hipStreamSynchronize(s);
        return PIMDB_E_NONFINITE;
    }
    return PIMDB_OK;
}

extern "C" {

int pimdb_xplan_create(int natoms, int ndim, double beta_t, double spring_k, int pbc, double size,
                       pimdb_xplan** out) {
    if (!out || natoms < 1 || ndim < 1 || ndim > 3 || !(beta_t > 0.0) || (pbc && !(size > 0.0))) return PIMDB_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PIMDB_E_HIP;
    pimdb_xplan* pl = new pimdb_xplan();
    pl->N = natoms; pl->D = ndim; pl->pbc = pbc; pl->bt = beta_t; pl->k = spring_k; pl->L = size;
    if (hipMalloc((void**)&pl->Shi, natoms * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&pl->Slo, natoms * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&pl->V, (natoms + 1) * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&pl->Vb, (natoms + 1) * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&pl->status, sizeof(int)) != hipSuccess || hipMemset(pl->status, 0, sizeof(int)) != hipSuccess ||
        // a ring whose recursion workspace exceeds a workgroup's LDS runs on device memory
        (rec_lds_bytes(natoms, ndim) > device_lds_limit() &&
         hipMalloc((void**)&pl->ws, 2 * rec_lds_bytes(natoms, ndim)) != hipSuccess)) {
        pimdb_xplan_destroy(pl);
        return PIMDB_E_HIP;
    }
    *out = pl;
    return PIMDB_OK;
}

void pimdb_xplan_destroy(pimdb_xplan* pl) {
    if (!pl) return;
    (void)hipDeviceSynchronize();
    if (pl->Shi) (void)hipFree(pl->Shi);
    if (pl->Slo) (void)hipFree(pl->Slo);
    if (pl->V) (void)hipFree(pl->V);
    if (pl->Vb) (void)hipFree(pl->Vb);
    if (pl->status) (void)hipFree(pl->status);
    if (pl->ws) (void)hipFree(pl->ws);
    delete pl;
}

int pimdb_xplan_forward(pimdb_xplan* pl, const double* x1, const double* xP, double* VB, double* V, double* Vb,
                        void* stream) {
    if (!pl || !x1 || !xP) return PIMDB_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    int rc = xplan_forward_internal(pl, x1, xP, s);
    if (rc) return rc;
    const int N = pl->N;
    if (VB && hipMemcpyAsync(VB, pl->V + N, sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess) return PIMDB_E_HIP;
    if (V && hipMemcpyAsync(V, pl->V, (N + 1) * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess) return PIMDB_E_HIP;
    if (Vb && hipMemcpyAsync(Vb, pl->Vb, (N + 1) * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess) return PIMDB_E_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? PIMDB_OK : PIMDB_E_HIP;
}

int pimdb_xplan_vb_grad(pimdb_xplan* pl, const double* x1, const double* xP, double* grad, void* stream) {
    if (!pl || !x1 || !xP || !grad) return PIMDB_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    int rc = xplan_forward_internal(pl, x1, xP, s);
    if (rc) return rc;
    const int N = pl->N;
    dim3 grid((2 * N + 3) / 4), block(256);
    switch (pl->D) {
    case 1: hipLaunchKernelGGL(k_exchange_vb_grad<1>, grid, block, 0, s, N, pl->bt, pl->k, pl->pbc, pl->L, pl->Shi, pl->Slo, pl->V, pl->Vb, x1, xP, grad); break;
    case 2: hipLaunchKernelGGL(k_exchange_vb_grad<2>, grid, block, 0, s, N, pl->bt, pl->k, pl->pbc, pl->L, pl->Shi, pl->Slo, pl->V, pl->Vb, x1, xP, grad); break;
    default: hipLaunchKernelGGL(k_exchange_vb_grad<3>, grid, block, 0, s, N, pl->bt, pl->k, pl->pbc, pl->L, pl->Shi, pl->Slo, pl->V, pl->Vb, x1, xP, grad); break;
    }
    if (hipGetLastError() != hipSuccess) return PIMDB_E_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? PIMDB_OK : PIMDB_E_HIP;
}

}  // extern "C"
