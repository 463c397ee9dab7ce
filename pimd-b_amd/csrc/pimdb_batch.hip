// pimdb_batch.hip -- C ABI of the batched persistent form (include/pimdb_gpu.h):
// replicas of uneven N, P and T in one persistent kernel per MD step,
// replayed from a hipGraph, with the workgroup size and residency chosen by a
// small autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

#include "pimdb_gpu.h"
#include "batch.hpp"
#include "ctx_common.hpp"

using namespace pimdb;

struct pimdb_batch {
    std::vector<pimdb_params> prm;
    std::vector<RepDesc> desc;
    int R = 0, D = 3, kind = 0;
    int threads = 1024, resident = 0;   // resident workgroups (0: all CUs)
    int cus = 0;
    int per_cu = 1;                     // workgroups that share one CU's LDS
    size_t nx = 0, nV = 0, nC = 0, nR = 0, nH = 0;
    double* d_fx = nullptr;             // physical forces of normal-mode-propagator replicas
    double* d_rot = nullptr;            // their rotation tables
    double* d_nh = nullptr;             // Nose-Hoover chain states: eta, eta_dot, eta_dot_dot
    double* d_modes = nullptr;          // normal modes of Nose-Hoover replicas with normal-mode coupling
    double* d_xws = nullptr;            // exchange workspaces in device memory (RepDesc::xws)
    size_t xws_cap = 0;                 // doubles held by d_xws
    size_t lds = 0;
    BatchPtrs bp{};
    RepDesc* d_desc = nullptr;
    int *d_sched = nullptr, *d_sched_off = nullptr;
    int grid = 0;
    hipStream_t s = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool fused_any = false;       // some replica takes the fused normal-mode path
    bool nmprop_any = false;      // some replica takes the normal-mode propagator
    bool nh_any = false;          // some replica takes Nose-Hoover chains
    bool xdev_any = false;        // some replica's exchange workspace is in device memory
    bool ext_any = false;         // any of the three: the EXT kernels
    bool maybe_pending = false;   // a step ran since the last finishing launch
    std::string err;
    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
};

#define BCK(expr)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return b->fail(PIMDB_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Cost model for the schedule: ordered pairs dominate, the exchange adds N^2.
static double rep_cost(const pimdb_params& p) {
    const double N = p.natoms, P = p.nbeads;
    return P * N * N + (p.bosonic && p.nbeads > 1 ? 8.0 * N * N : 0.0) + 4.0 * P * P * N;
}

// Longest-processing-time-first assignment of replicas to `grid` workgroups.
static void build_schedule(pimdb_batch* b, std::vector<int>& sched, std::vector<int>& off) {
    std::vector<int> order(b->R);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return rep_cost(b->prm[a]) > rep_cost(b->prm[c]); });
    using Item = std::pair<double, int>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
    for (int g = 0; g < b->grid; ++g) heap.push({0.0, g});
    std::vector<std::vector<int>> lists(b->grid);
    for (int r : order) {
        Item it = heap.top();
        heap.pop();
        lists[it.second].push_back(r);
        heap.push({it.first + rep_cost(b->prm[r]), it.second});
    }
    sched.clear();
    off.assign(1, 0);
    for (auto& l : lists) {
        sched.insert(sched.end(), l.begin(), l.end());
        off.push_back((int)sched.size());
    }
}

// LDS plan of every replica for `threads` threads per workgroup and
// `per_cu` workgroups per compute unit (each then gets 1/per_cu of its LDS):
// whether the pair forces of every wave's slice stay in LDS, how many chunk
// slots the exchange's backward pass gets, and the largest need.  A replica
// whose exchange layout does not fit even with one chunk slot gets its
// workspace in device memory instead (one slot per wave; its size in doubles
// into (*xdev)[r], 0 for the others).  Returns the dynamic LDS bytes, or 0
// (with *why) when some replica does not fit.
static size_t plan_lds(pimdb_batch* b, int threads, int per_cu, std::vector<RepDesc>& desc, std::string* why,
                       std::vector<size_t>* xdev = nullptr) {
    const size_t budget = device_lds_limit() / (size_t)per_cu / sizeof(double);   // doubles
    const size_t avail = budget > 16 ? budget - 16 : 0;                             // after the reduction slots
    const int nw = threads / WAVE;
    size_t need = 0;
    for (int r = 0; r < b->R; ++r) {
        const pimdb_params& p = b->prm[r];
        const int N = p.natoms, D = p.ndim;
        RepDesc& d = desc[r];
        d.xslots = 1;
        if (xdev) (*xdev)[r] = 0;
        if (d.bosonic) {
            // the exchange's layout with as many chunk slots (up to one per
            // wave) as fit
            if (N > XB_MAXPER * threads) {
                *why = "replica " + std::to_string(r) + ": natoms above " + std::to_string(XB_MAXPER) +
                       " per thread of a batched workgroup";
                return 0;
            }
            int slots = nw;
            size_t x = (size_t)xb_layout(N, D, nw, slots).total;
            while (slots > 1 && x > avail) x = (size_t)xb_layout(N, D, nw, --slots).total;
            if (x > avail) {   // the workspace in device memory
                slots = nw;
                if (xdev) (*xdev)[r] = (size_t)xb_layout(N, D, nw, slots).total;
                x = 0;
            }
            d.xslots = slots;
            need = std::max(need, x);
        }
        size_t nm = batch_nm_lds_doubles(p.nbeads, nw);
        if (d.nmprop) nm = std::max(nm, 4 * (size_t)p.nbeads * NMD_TC);   // wg_nm_drift's tiles
        if (p.thermostat >= PIMDB_THERMO_NOSE_HOOVER && p.nmthermostat)
            nm = std::max(nm, (size_t)p.nbeads * NMD_TC);                 // wg_nh_ostep's tiles
        if (nm > avail) {
            *why = "replica " + std::to_string(r) + ": nbeads too large for the batched workgroup";
            return 0;
        }
        need = std::max(need, nm);
        // pair forces of every wave's slice in LDS when they fit (+ the tile
        // order and its inverse, N ints each)
        const size_t pl = (size_t)nw * pair_fstride(D) * N + N;
        d.plds = pl <= avail ? 1 : 0;
        if (d.plds) need = std::max(need, pl);
    }
    return sizeof(double) * (16 + need);
}

static int upload_schedule(pimdb_batch* b) {
    const int resident = b->resident > 0 ? b->resident : b->cus;
    b->grid = std::max(1, std::min(b->R, resident));
    // the LDS is planned for the residency asked for (each CU's share)
    b->per_cu = (resident + b->cus - 1) / b->cus;
    {
        std::vector<RepDesc> desc = b->desc;
        std::string why;
        std::vector<size_t> xdev(b->R, 0);
        const size_t lds = plan_lds(b, b->threads, b->per_cu, desc, &why, &xdev);
        if (!lds) return b->fail(PIMDB_E_ARG, why);
        // the device-memory exchange workspaces of this plan
        size_t total = 0;
        for (size_t n : xdev) total += n;
        if (total > b->xws_cap) {
            BCK(hipStreamSynchronize(b->s));
            if (b->d_xws) (void)hipFree(b->d_xws);
            b->d_xws = nullptr;
            b->xws_cap = 0;
            BCK(hipMalloc((void**)&b->d_xws, total * sizeof(double)));
            b->xws_cap = total;
        }
        size_t o = 0;
        for (int r = 0; r < b->R; ++r) {
            if (desc[r].bosonic) desc[r].xws = xdev[r] ? b->d_xws + o : nullptr;   // (fx is the nmprop one's)
            o += xdev[r];
        }
        b->xdev_any = total > 0;
        b->ext_any = b->nmprop_any || b->nh_any || b->xdev_any;
        b->desc = desc;
        b->lds = lds;
        if (b->d_desc)
            BCK(hipMemcpy(b->d_desc, b->desc.data(), sizeof(RepDesc) * b->R, hipMemcpyHostToDevice));
    }
    std::vector<int> sched, off;
    build_schedule(b, sched, off);
    if (b->d_sched) (void)hipFree(b->d_sched);
    if (b->d_sched_off) (void)hipFree(b->d_sched_off);
    b->d_sched = nullptr;
    b->d_sched_off = nullptr;
    BCK(hipMalloc((void**)&b->d_sched, sizeof(int) * sched.size()));
    BCK(hipMalloc((void**)&b->d_sched_off, sizeof(int) * off.size()));
    BCK(hipMemcpy(b->d_sched, sched.data(), sizeof(int) * sched.size(), hipMemcpyHostToDevice));
    BCK(hipMemcpy(b->d_sched_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
    b->bp.sched = b->d_sched;
    b->bp.sched_off = b->d_sched_off;
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    return PIMDB_OK;
}

template <int D, int KIND>
static void launch_batch_dk(pimdb_batch* b, int mode) {
    dim3 grid(b->grid), block(b->threads);
    if (b->ext_any && (mode == BM_STEP || mode == BM_FORCES || mode == BM_INTEGRATOR || mode == BM_EXCHANGE)) {
        switch (mode) {
        case BM_STEP: hipLaunchKernelGGL((k_batch_step<D, KIND, BM_STEP, true>), grid, block, b->lds, b->s, b->bp); break;
        case BM_FORCES: hipLaunchKernelGGL((k_batch_step<D, KIND, BM_FORCES, true>), grid, block, b->lds, b->s, b->bp); break;
        case BM_EXCHANGE: hipLaunchKernelGGL((k_batch_step<D, KIND, BM_EXCHANGE, true>), grid, block, b->lds, b->s, b->bp); break;
        default: hipLaunchKernelGGL((k_batch_step<D, KIND, BM_INTEGRATOR, true>), grid, block, b->lds, b->s, b->bp); break;
        }
        return;
    }
    switch (mode) {
    case BM_STEP: hipLaunchKernelGGL((k_batch_step<D, KIND, BM_STEP>), grid, block, b->lds, b->s, b->bp); break;
    case BM_FORCES: hipLaunchKernelGGL((k_batch_step<D, KIND, BM_FORCES>), grid, block, b->lds, b->s, b->bp); break;
    case BM_PAIRS: hipLaunchKernelGGL((k_batch_step<D, KIND, BM_PAIRS>), grid, block, b->lds, b->s, b->bp); break;
    case BM_EXCHANGE: hipLaunchKernelGGL((k_batch_step<D, KIND, BM_EXCHANGE>), grid, block, b->lds, b->s, b->bp); break;
    case BM_FINISH: hipLaunchKernelGGL((k_batch_step<D, KIND, BM_FINISH>), grid, block, b->lds, b->s, b->bp); break;
    default: hipLaunchKernelGGL((k_batch_step<D, KIND, BM_INTEGRATOR>), grid, block, b->lds, b->s, b->bp); break;
    }
}

template <int D>
static void launch_batch_d(pimdb_batch* b, int mode) {
    switch (b->kind) {
    case POT_HARMONIC: launch_batch_dk<D, POT_HARMONIC>(b, mode); break;
    case POT_DIPOLE: launch_batch_dk<D, POT_DIPOLE>(b, mode); break;
    case POT_AZIZ: launch_batch_dk<D, POT_AZIZ>(b, mode); break;
    default: launch_batch_dk<D, POT_FREE>(b, mode); break;
    }
}

static int launch_batch(pimdb_batch* b, int mode) {
    switch (b->D) {
    case 1: launch_batch_d<1>(b, mode); break;
    case 2: launch_batch_d<2>(b, mode); break;
    default: launch_batch_d<3>(b, mode); break;
    }
    BCK(hipGetLastError());
    return PIMDB_OK;
}

static int capture(pimdb_batch* b) {
    if (b->gexec) return PIMDB_OK;
    BCK(hipStreamSynchronize(b->s));
    BCK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
    int rc = launch_batch(b, BM_STEP);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return b->fail(PIMDB_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    BCK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    return PIMDB_OK;
}

static int check_batch_status(pimdb_batch* b) {
    std::vector<int> st(b->R);
    BCK(hipStreamSynchronize(b->s));
    BCK(hipMemcpy(st.data(), b->bp.status, sizeof(int) * b->R, hipMemcpyDeviceToHost));
    for (int r = 0; r < b->R; ++r) {
        if (st[r]) {
            BCK(hipMemset(b->bp.status, 0, sizeof(int) * b->R));
            return b->fail(PIMDB_E_NONFINITE, "non-finite bosonic exchange potential in replica " + std::to_string(r));
        }
    }
    return PIMDB_OK;
}

extern "C" {

int pimdb_batch_create(const pimdb_params* reps, int nrep, pimdb_batch** out) {
    if (!reps || !out || nrep < 1) return PIMDB_E_ARG;
    *out = nullptr;
    std::string why;
    for (int r = 0; r < nrep; ++r) {
        if (!validate_params(reps[r], &why)) {
            set_create_error("replica " + std::to_string(r) + ": " + why);
            return PIMDB_E_ARG;
        }
        if (reps[r].bosonic && reps[r].exchange_algo != PIMDB_EXCHANGE_QUADRATIC) {
            set_create_error("batched replicas use the Feldman-Hirshberg exchange");
            return PIMDB_E_ARG;
        }
        if (reps[r].ndim != reps[0].ndim || reps[r].int_kind != reps[0].int_kind) {
            set_create_error("all replicas of a batch share ndim and the interaction kind");
            return PIMDB_E_ARG;
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_create_error("no HIP device available");
        return PIMDB_E_HIP;
    }
    pimdb_batch* b = new pimdb_batch();
    b->R = nrep;
    b->D = reps[0].ndim;
    b->kind = reps[0].int_kind;
    b->prm.assign(reps, reps + nrep);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        delete b;
        set_create_error("hipGetDeviceProperties failed");
        return PIMDB_E_HIP;
    }
    b->cus = prop.multiProcessorCount;
    int Nmax = 1;
    std::vector<size_t> off_R, off_H;   // per replica: its block of the rotation tables and of the chains
    for (int r = 0; r < nrep; ++r) {
        const pimdb_params& p = reps[r];
        const Derived dv = derive(p);
        RepDesc d{};
        d.N = p.natoms; d.P = p.nbeads; d.bosonic = dv.bosonic; d.pbc = p.pbc;
        d.thermostat = p.thermostat; d.nm = p.nmthermostat; d.fixcom = p.fixcom;
        d.nmprop = p.propagator == PIMDB_PROP_NORMAL_MODES ? 1 : 0;   // (distinguishable only: validate_params)
        d.fused = (p.thermostat == PIMDB_THERMO_LANGEVIN && p.nmthermostat && p.nbeads <= NM_MFMA_MAXP &&
                   !d.nmprop) ? 1 : 0;
        if (d.fused) b->fused_any = true;
        d.L = p.size; d.invL = 1.0 / p.size; d.mass = p.mass; d.dt = p.dt;
        d.k = dv.k; d.bt = dv.bt; d.c1 = dv.c1; d.c2 = dv.c2;
        d.ext = dv.ext; d.int_k = dv.int_k; d.strength = p.int_strength; d.cutoff = dv.cutoff;
        d.cut2 = cutoff_threshold(dv.cutoff);
        d.seed = p.seed; d.rid = 0;   // noise stream keyed by the seed, as in the single-system form
        d.off_x = b->nx; b->nx += (size_t)p.nbeads * p.ndim * p.natoms;
        d.off_V = b->nV; b->nV += (size_t)p.natoms + 1;
        d.off_C = b->nC; b->nC += (size_t)p.nbeads * p.nbeads;
        off_R.push_back(b->nR); b->nR += 4 * (size_t)p.nbeads;
        off_H.push_back(b->nH);
        if (p.thermostat >= PIMDB_THERMO_NOSE_HOOVER) {
            // chains per bead and their degrees of freedom (nose_hoover.cpp:10-32, 152-163, 198-208)
            const int groups = p.thermostat == PIMDB_THERMO_NOSE_HOOVER ? 1
                               : p.thermostat == PIMDB_THERMO_NOSE_HOOVER_NP ? p.natoms : p.natoms * p.ndim;
            const int ndof = p.thermostat == PIMDB_THERMO_NOSE_HOOVER ? p.ndim * p.natoms
                             : p.thermostat == PIMDB_THERMO_NOSE_HOOVER_NP ? p.ndim : 1;
            d.nh_nc = p.nchains;
            d.nh_tbeta = dv.thermo_beta;            // beta / P
            d.nh_Qi = 1.0 * 1.0 * dv.beta / p.nbeads;   // hbar^2 beta / P
            d.nh_Q1 = ndof * d.nh_Qi;
            d.nh_required = ndof / d.nh_tbeta;
            b->nH += (size_t)p.nbeads * groups * p.nchains;
            b->nh_any = true;
        }
        b->desc.push_back(d);
        Nmax = std::max(Nmax, p.natoms);
    }
    int rc = PIMDB_OK;
    {
        // the default launch shape (1024 threads, one workgroup per CU) must fit
        std::vector<RepDesc> desc = b->desc;
        std::string why;
        b->lds = plan_lds(b, b->threads, 1, desc, &why);
        if (!b->lds) rc = b->fail(PIMDB_E_ARG, why);
        else b->desc = desc;
    }
    auto alloc = [&](double** ptr, size_t n) {
        if (rc) return;
        if (hipMalloc((void**)ptr, std::max<size_t>(1, n) * sizeof(double)) != hipSuccess ||
            hipMemset(*ptr, 0, std::max<size_t>(1, n) * sizeof(double)) != hipSuccess)
            rc = b->fail(PIMDB_E_HIP, "device allocation failed");
    };
    double *Ct = nullptr, *CTt = nullptr, *rot = nullptr;
    alloc(&b->bp.x, b->nx); alloc(&b->bp.p, b->nx); alloc(&b->bp.f, b->nx);
    alloc(&b->bp.V, b->nV);
    alloc(&b->bp.cmp, 8 * (size_t)nrep);
    alloc(&Ct, b->nC + NM_BLOCK); alloc(&CTt, b->nC + NM_BLOCK);   // padded for block reads
    alloc(&rot, b->nR);
    b->d_rot = rot;
    for (const RepDesc& d : b->desc) b->nmprop_any = b->nmprop_any || d.nmprop;
    if (b->nmprop_any) alloc(&b->d_fx, b->nx);
    if (b->nh_any) alloc(&b->d_nh, 3 * b->nH);   // the chains start at rest (alloc zeroes them)
    bool nhnm_any = false;   // Nose-Hoover with normal-mode coupling: the modes' buffer
    for (int r = 0; r < nrep; ++r) nhnm_any = nhnm_any || (reps[r].thermostat >= PIMDB_THERMO_NOSE_HOOVER && reps[r].nmthermostat);
    if (nhnm_any) alloc(&b->d_modes, b->nx);
    for (int r = 0; r < nrep && !rc; ++r) {   // the EXT paths' per-replica arrays
        RepDesc& d = b->desc[r];
        d.fx = d.nmprop ? b->d_fx + d.off_x : nullptr;   // (shares its slot with xws, bosonic replicas')
        d.rot = rot + off_R[r];
        d.nh = b->nh_any ? b->d_nh + off_H[r] : nullptr;
        d.nh_stride = b->nH;
        d.modes = nhnm_any ? b->d_modes + d.off_x : nullptr;
    }
    b->ext_any = b->nmprop_any || b->nh_any;
    b->bp.Ct = Ct;
    b->bp.CTt = CTt;
    if (!rc && (hipMalloc((void**)&b->bp.status, sizeof(int) * nrep) != hipSuccess ||
                hipMemset(b->bp.status, 0, sizeof(int) * nrep) != hipSuccess ||
                hipMalloc((void**)&b->bp.ctl, 2 * sizeof(unsigned long long)) != hipSuccess ||
                hipMemset(b->bp.ctl, 0, 2 * sizeof(unsigned long long)) != hipSuccess ||
                hipMalloc((void**)&b->d_desc, sizeof(RepDesc) * nrep) != hipSuccess ||
                hipMemcpy(b->d_desc, b->desc.data(), sizeof(RepDesc) * nrep, hipMemcpyHostToDevice) != hipSuccess))
        rc = b->fail(PIMDB_E_HIP, "device allocation failed");
    if (!rc) {
        b->bp.reps = b->d_desc;
        std::vector<double> hc(b->nC), hct(b->nC);
        for (int r = 0; r < nrep && !rc; ++r) {
            const int P = reps[r].nbeads;
            std::vector<double> C0((size_t)P * P), CT0((size_t)P * P);
            nm_matrices(P, C0.data(), CT0.data());
            transpose_square(P, C0.data(), hc.data() + b->desc[r].off_C);
            transpose_square(P, CT0.data(), hct.data() + b->desc[r].off_C);
        }
        // the free ring polymer's rotation per mode (nm_rotation, as the
        // single-system form builds it)
        std::vector<double> hr(std::max<size_t>(1, b->nR));
        for (int r = 0; r < nrep; ++r) nm_rotation(reps[r], derive(reps[r]), hr.data() + off_R[r]);
        if (hipMemcpy(Ct, hc.data(), sizeof(double) * b->nC, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(CTt, hct.data(), sizeof(double) * b->nC, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(rot, hr.data(), sizeof(double) * b->nR, hipMemcpyHostToDevice) != hipSuccess)
            rc = b->fail(PIMDB_E_HIP, "upload of normal-mode tables failed");
    }
    if (!rc && (hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess))
        rc = b->fail(PIMDB_E_HIP, "stream/event creation failed");
    if (!rc) rc = upload_schedule(b);
    if (rc) {
        set_create_error(b->err);
        pimdb_batch_destroy(b);
        return rc;
    }
    *out = b;
    return PIMDB_OK;
}

void pimdb_batch_destroy(pimdb_batch* b) {
    if (!b) return;
    if (b->s) (void)hipStreamSynchronize(b->s);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    double* bufs[] = {b->bp.x, b->bp.p, b->bp.f, b->bp.V, b->bp.cmp,
                      const_cast<double*>(b->bp.Ct), const_cast<double*>(b->bp.CTt), b->d_rot, b->d_fx, b->d_nh,
                      b->d_modes, b->d_xws};
    for (double* p : bufs)
        if (p) (void)hipFree(p);
    if (b->bp.status) (void)hipFree(b->bp.status);
    if (b->bp.ctl) (void)hipFree(b->bp.ctl);
    if (b->d_desc) (void)hipFree(b->d_desc);
    if (b->d_sched) (void)hipFree(b->d_sched);
    if (b->d_sched_off) (void)hipFree(b->d_sched_off);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
}

const char* pimdb_batch_last_error(const pimdb_batch* b) { return b ? b->err.c_str() : create_error(); }

}  // extern "C"

// The second B and O-step that fused replicas leave owed after a step, done
// before the host reads or replaces a state.
static int finish_pending(pimdb_batch* b) {
    if (!b->fused_any || !b->maybe_pending) return PIMDB_OK;
    int rc = launch_batch(b, BM_FINISH);
    if (rc) return rc;
    BCK(hipStreamSynchronize(b->s));
    b->maybe_pending = false;
    return PIMDB_OK;
}

extern "C" {

int pimdb_batch_set_state(pimdb_batch* b, int rep, const double* x, const double* p) {
    if (!b || rep < 0 || rep >= b->R) return PIMDB_E_ARG;
    const int frc = finish_pending(b);
    if (frc) return frc;
    const pimdb_params& q = b->prm[rep];
    const size_t n = (size_t)q.nbeads * q.ndim * q.natoms;
    std::vector<double> st(n);
    const double* src[2] = {x, p};
    double* dst[2] = {b->bp.x, b->bp.p};
    for (int k = 0; k < 2; ++k) {
        if (!src[k]) continue;
        to_soa(src[k], st.data(), q.natoms, q.nbeads, q.ndim);
        BCK(hipMemcpyAsync(dst[k] + b->desc[rep].off_x, st.data(), sizeof(double) * n, hipMemcpyHostToDevice, b->s));
        BCK(hipStreamSynchronize(b->s));
    }
    if (p) {
        // no COM pending; the sum of p per axis for the next first O-step
        double cmp[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (size_t e = 0; e < n; ++e) cmp[4 + e % q.ndim] += p[e];
        BCK(hipMemcpy(b->bp.cmp + 8 * (size_t)rep, cmp, sizeof(cmp), hipMemcpyHostToDevice));
    }
    return PIMDB_OK;
}

int pimdb_batch_get_state(pimdb_batch* b, int rep, double* x, double* p, double* f) {
    if (!b || rep < 0 || rep >= b->R) return PIMDB_E_ARG;
    const int frc = finish_pending(b);
    if (frc) return frc;
    const pimdb_params& q = b->prm[rep];
    const size_t n = (size_t)q.nbeads * q.ndim * q.natoms;
    std::vector<double> st(n);
    double* dst[3] = {x, p, f};
    const double* src[3] = {b->bp.x, b->bp.p, b->bp.f};
    for (int k = 0; k < 3; ++k) {
        if (!dst[k]) continue;
        BCK(hipMemcpyAsync(st.data(), src[k] + b->desc[rep].off_x, sizeof(double) * n, hipMemcpyDeviceToHost, b->s));
        BCK(hipStreamSynchronize(b->s));
        from_soa(st.data(), dst[k], q.natoms, q.nbeads, q.ndim);
    }
    if (p) {
        // the COM momentum still pending from the last O-step
        double cmp[8];
        BCK(hipMemcpy(cmp, b->bp.cmp + 8 * (size_t)rep, sizeof(cmp), hipMemcpyDeviceToHost));
        for (size_t e = 0; e < n; ++e) p[e] -= cmp[e % q.ndim];
    }
    return PIMDB_OK;
}

int pimdb_batch_forces(pimdb_batch* b) {
    if (!b) return PIMDB_E_ARG;
    int rc = launch_batch(b, BM_FORCES);
    if (rc) return rc;
    return check_batch_status(b);
}

int pimdb_batch_step(pimdb_batch* b, int nsteps) {
    if (!b || nsteps < 0) return PIMDB_E_ARG;
    int rc = capture(b);
    if (rc) return rc;
    for (int s = 0; s < nsteps; ++s) BCK(hipGraphLaunch(b->gexec, b->s));
    if (nsteps > 0) b->maybe_pending = true;
    return check_batch_status(b);
}

int pimdb_batch_step_timed(pimdb_batch* b, int nsteps, double* device_ms) {
    if (!b || nsteps < 1 || !device_ms) return PIMDB_E_ARG;
    int rc = capture(b);
    if (rc) return rc;
    BCK(hipEventRecord(b->t0, b->s));
    for (int s = 0; s < nsteps; ++s) BCK(hipGraphLaunch(b->gexec, b->s));
    BCK(hipEventRecord(b->t1, b->s));
    b->maybe_pending = true;
    BCK(hipEventSynchronize(b->t1));
    float ms = 0.0f;
    BCK(hipEventElapsedTime(&ms, b->t0, b->t1));
    *device_ms = (double)ms;
    return check_batch_status(b);
}

int pimdb_batch_configure(pimdb_batch* b, int threads, int resident) {
    if (!b) return PIMDB_E_ARG;
    if (threads && threads != 256 && threads != 512 && threads != 1024)
        return b->fail(PIMDB_E_ARG, "threads must be 256, 512 or 1024");
    if (resident < 0) return b->fail(PIMDB_E_ARG, "resident must be >= 0");
    BCK(hipStreamSynchronize(b->s));
    const int old_threads = b->threads, old_resident = b->resident;
    if (threads) b->threads = threads;
    if (resident) b->resident = resident;
    const int rc = upload_schedule(b);
    if (rc == PIMDB_E_ARG) {   // this shape's LDS plan does not fit: keep the previous one
        const std::string why = b->err;
        b->threads = old_threads;
        b->resident = old_resident;
        const int rc2 = upload_schedule(b);
        if (rc2) return rc2;
        return b->fail(PIMDB_E_ARG, why);
    }
    return rc;
}

int pimdb_batch_time_step(pimdb_batch* b, int reps, double* ms_step) {
    if (!b || reps < 1 || !ms_step) return PIMDB_E_ARG;
    int rc = capture(b);
    if (rc) return rc;
    BCK(hipGraphLaunch(b->gexec, b->s));
    b->maybe_pending = true;
    BCK(hipEventRecord(b->t0, b->s));
    for (int r = 0; r < reps; ++r) BCK(hipGraphLaunch(b->gexec, b->s));
    BCK(hipEventRecord(b->t1, b->s));
    BCK(hipEventSynchronize(b->t1));
    float ms = 0.0f;
    BCK(hipEventElapsedTime(&ms, b->t0, b->t1));
    *ms_step = (double)ms / reps;
    return check_batch_status(b);
}

// Device ms of one launch of a phase (0 step, 1 pair forces, 2 exchange,
// 3 integrator), averaged over reps launches, from a saved state that is
// restored afterwards.
int pimdb_batch_time_phase(pimdb_batch* b, int phase, int reps, double* ms_out) {
    if (!b || reps < 1 || !ms_out || phase < 0 || phase > 3) return PIMDB_E_ARG;
    static const int modes[4] = {BM_STEP, BM_PAIRS, BM_EXCHANGE, BM_INTEGRATOR};
    double *sx = nullptr, *sp = nullptr, *sf = nullptr, *sfx = nullptr;
    unsigned long long sctl[2];
    BCK(hipStreamSynchronize(b->s));
    BCK(hipMalloc((void**)&sx, sizeof(double) * b->nx));
    BCK(hipMalloc((void**)&sp, sizeof(double) * b->nx));
    BCK(hipMalloc((void**)&sf, sizeof(double) * b->nx));
    BCK(hipMemcpy(sx, b->bp.x, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    BCK(hipMemcpy(sp, b->bp.p, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    BCK(hipMemcpy(sf, b->bp.f, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    if (b->d_fx) {   // the normal-mode replicas' physical forces
        BCK(hipMalloc((void**)&sfx, sizeof(double) * b->nx));
        BCK(hipMemcpy(sfx, b->d_fx, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    }
    std::vector<double> seta;   // the Nose-Hoover chains
    if (b->nh_any) {
        seta.resize(3 * b->nH);
        BCK(hipMemcpy(seta.data(), b->d_nh, sizeof(double) * 3 * b->nH, hipMemcpyDeviceToHost));
    }
    BCK(hipMemcpy(sctl, b->bp.ctl, sizeof(sctl), hipMemcpyDeviceToHost));
    std::vector<double> scmp(8 * (size_t)b->R);
    BCK(hipMemcpy(scmp.data(), b->bp.cmp, sizeof(double) * scmp.size(), hipMemcpyDeviceToHost));
    int rc = launch_batch(b, modes[phase]);   // warm
    if (!rc) {
        BCK(hipEventRecord(b->t0, b->s));
        for (int r = 0; r < reps && !rc; ++r) rc = launch_batch(b, modes[phase]);
        BCK(hipEventRecord(b->t1, b->s));
        BCK(hipEventSynchronize(b->t1));
        float ms = 0.0f;
        BCK(hipEventElapsedTime(&ms, b->t0, b->t1));
        *ms_out = (double)ms / reps;
    }
    BCK(hipStreamSynchronize(b->s));
    BCK(hipMemcpy(b->bp.x, sx, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    BCK(hipMemcpy(b->bp.p, sp, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    BCK(hipMemcpy(b->bp.f, sf, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    if (sfx) BCK(hipMemcpy(b->d_fx, sfx, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    if (b->nh_any) BCK(hipMemcpy(b->d_nh, seta.data(), sizeof(double) * 3 * b->nH, hipMemcpyHostToDevice));
    BCK(hipMemcpy(b->bp.ctl, sctl, sizeof(sctl), hipMemcpyHostToDevice));
    BCK(hipMemcpy(b->bp.cmp, scmp.data(), sizeof(double) * scmp.size(), hipMemcpyHostToDevice));
    (void)hipFree(sx);
    (void)hipFree(sp);
    (void)hipFree(sf);
    if (sfx) (void)hipFree(sfx);
    if (rc) return rc;
    return check_batch_status(b);
}

// Tries every workgroup size and residency (1 or 2 workgroups per CU) for
// trial_steps steps from a saved state, keeps the fastest and restores the
// state, so the batch's trajectory is unaffected.
int pimdb_batch_autotune(pimdb_batch* b, int trial_steps, int* threads_out, int* resident_out) {
    if (!b || trial_steps < 1) return PIMDB_E_ARG;
    double *sx = nullptr, *sp = nullptr, *sf = nullptr, *sfx = nullptr;
    unsigned long long sctl[2];
    BCK(hipStreamSynchronize(b->s));
    BCK(hipMalloc((void**)&sx, sizeof(double) * b->nx));
    BCK(hipMalloc((void**)&sp, sizeof(double) * b->nx));
    BCK(hipMalloc((void**)&sf, sizeof(double) * b->nx));
    BCK(hipMemcpy(sx, b->bp.x, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    BCK(hipMemcpy(sp, b->bp.p, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    BCK(hipMemcpy(sf, b->bp.f, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    if (b->d_fx) {   // the normal-mode replicas' physical forces
        BCK(hipMalloc((void**)&sfx, sizeof(double) * b->nx));
        BCK(hipMemcpy(sfx, b->d_fx, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    }
    std::vector<double> seta;   // the Nose-Hoover chains
    if (b->nh_any) {
        seta.resize(3 * b->nH);
        BCK(hipMemcpy(seta.data(), b->d_nh, sizeof(double) * 3 * b->nH, hipMemcpyDeviceToHost));
    }
    BCK(hipMemcpy(sctl, b->bp.ctl, sizeof(sctl), hipMemcpyDeviceToHost));
    std::vector<double> scmp(8 * (size_t)b->R);
    BCK(hipMemcpy(scmp.data(), b->bp.cmp, sizeof(double) * scmp.size(), hipMemcpyDeviceToHost));
    // launch shapes: threads per workgroup x workgroups per CU, at most 16
    // waves per CU (the kernel's 128 VGPRs allow 4 waves per SIMD); shapes
    // whose LDS plan does not fit, or that give the same grid as one already
    // timed, are skipped
    const int cand[6][2] = {{1024, 1}, {512, 1}, {512, 2}, {256, 1}, {256, 2}, {256, 4}};
    double best = 1e300;
    int bt = b->threads, br = b->resident;
    int rc = PIMDB_OK;
    std::vector<std::pair<int, int>> tried;
    for (int ci = 0; ci < 6 && !rc; ++ci) {
        const int threads = cand[ci][0], per_cu = cand[ci][1], resident = b->cus * per_cu;
        const int grid = std::max(1, std::min(b->R, resident));
        // more workgroups per CU than the batch fills only tightens the plan
        if (per_cu > 1 && grid <= b->cus * (per_cu - 1)) continue;
        if (std::find(tried.begin(), tried.end(), std::make_pair(threads, grid)) != tried.end()) continue;
        tried.push_back({threads, grid});
        std::vector<RepDesc> desc = b->desc;
        std::string why;
        if (!plan_lds(b, threads, per_cu, desc, &why)) continue;
        b->threads = threads;
        b->resident = resident;
        rc = upload_schedule(b);
        double ms = 0.0;
        if (!rc) rc = pimdb_batch_time_step(b, trial_steps, &ms);
        if (!rc && ms < best) {
            best = ms;
            bt = b->threads;
            br = b->resident;
        }
    }
    b->threads = bt;
    b->resident = br;
    if (!rc) rc = upload_schedule(b);
    BCK(hipMemcpy(b->bp.x, sx, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    BCK(hipMemcpy(b->bp.p, sp, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    BCK(hipMemcpy(b->bp.f, sf, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    if (sfx) BCK(hipMemcpy(b->d_fx, sfx, sizeof(double) * b->nx, hipMemcpyDeviceToDevice));
    if (b->nh_any) BCK(hipMemcpy(b->d_nh, seta.data(), sizeof(double) * 3 * b->nH, hipMemcpyHostToDevice));
    BCK(hipMemcpy(b->bp.ctl, sctl, sizeof(sctl), hipMemcpyHostToDevice));
    BCK(hipMemcpy(b->bp.cmp, scmp.data(), sizeof(double) * scmp.size(), hipMemcpyHostToDevice));
    (void)hipFree(sx);
    (void)hipFree(sp);
    (void)hipFree(sf);
    if (sfx) (void)hipFree(sfx);
    if (threads_out) *threads_out = b->threads;
    if (resident_out) *resident_out = b->resident;
    return rc;
}

}  // extern "C"
