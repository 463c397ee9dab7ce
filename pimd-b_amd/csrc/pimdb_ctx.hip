// pimdb_ctx.hip -- the single-system ("plain" per-kernel) form of the C ABI in
// include/pimdb_gpu.h.
//
// One MD step (Simulation::run body, src/simulation.cpp:246-259) is
//   K1 k_momenta      O-step (+ deferred zeroMomentum of the previous step)
//   K2 k_kick_drift   zeroMomentum, B, A          (or k_nm_propagate)
//   K3 forces         s1: k_exchange_recursions (cycle energies on the fly)
//                     s0: k_bead_forces (all beads, all pairs)
//                     s0 after join: k_exchange_forces (beads 1 and P)
//   K4 k_momenta      B, O-step, partial sums for zeroMomentum
// The exchange chain (one long-latency workgroup) overlaps the pair forces
// on a second stream of default priority.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pimdb_gpu.h"
#include "ctx_common.hpp"
#include "exchange.hpp"
#include "factorial.hpp"
#include "forces.hpp"
#include "nose_hoover.hpp"
#include "integrator.hpp"
#include "observables.hpp"

using namespace pimdb;

struct pimdb_ctx {
    pimdb_params prm{};
    Derived dv{};
    int N = 0, P = 0, D = 0;
    size_t nbd = 0;
    hipStream_t s0 = nullptr, s1 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, t0 = nullptr, t1 = nullptr;
    double *x = nullptr, *p = nullptr, *f = nullptr, *fext = nullptr;
    double *E = nullptr, *ET = nullptr, *V = nullptr, *Vb = nullptr, *prob = nullptr, *fex = nullptr;
    double *Shi = nullptr, *Slo = nullptr;   // link prefix sums of the ring (exterior forces)
    double* xws = nullptr;   // the recursions' workspace in device memory, when it exceeds LDS
    double *C = nullptr, *CT = nullptr, *rot = nullptr, *com0 = nullptr, *com1 = nullptr;
    double *noise = nullptr, *obs = nullptr;
    size_t noise_cap = 0;
    int* status = nullptr;
    unsigned long long* ctr = nullptr;
    int tc = 1, ntiles = 1;
    int ncom = 1;                  // rows of the partial momentum sums (com0 / com1)
    NhArgs nh{};                   // Nose-Hoover chains (thermostat >= 2)
    double* modes = nullptr;       // normal modes of p, [k][a][i] (Nose-Hoover with nmthermostat)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<double> stage;
    double* pinned = nullptr;
    size_t pinned_cap = 0;
    std::string err;
    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
};

#define CK(expr)                                                                                   \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return ctx->fail(PIMDB_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------ launches

template <int D>
static void launch_bead_forces_d(pimdb_ctx* ctx, hipStream_t s, double* f, double* fext) {
    ForceArgs fa = make_force_args(ctx->prm, ctx->dv);
    dim3 grid((ctx->N + FORCE_TILE - 1) / FORCE_TILE, ctx->P);
    const size_t red = sizeof(double) * (size_t)(FORCE_WAVES - 1) * D * FORCE_TILE;
    const size_t lds = sizeof(double) * (size_t)D * ctx->N + red;
    const dim3 blk(FORCE_TILE * FORCE_WAVES);
    if (lds > device_lds_limit()) {   // the slice read in place
        switch (ctx->prm.int_kind) {
        case POT_HARMONIC: hipLaunchKernelGGL((k_bead_forces<D, POT_HARMONIC, true>), grid, blk, red, s, fa, ctx->x, f, fext); break;
        case POT_DIPOLE: hipLaunchKernelGGL((k_bead_forces<D, POT_DIPOLE, true>), grid, blk, red, s, fa, ctx->x, f, fext); break;
        case POT_AZIZ: hipLaunchKernelGGL((k_bead_forces<D, POT_AZIZ, true>), grid, blk, red, s, fa, ctx->x, f, fext); break;
        default: hipLaunchKernelGGL((k_bead_forces<D, POT_FREE, true>), grid, blk, red, s, fa, ctx->x, f, fext); break;
        }
        return;
    }
    switch (ctx->prm.int_kind) {
    case POT_HARMONIC: hipLaunchKernelGGL((k_bead_forces<D, POT_HARMONIC>), grid, blk, lds, s, fa, ctx->x, f, fext); break;
    case POT_DIPOLE: hipLaunchKernelGGL((k_bead_forces<D, POT_DIPOLE>), grid, blk, lds, s, fa, ctx->x, f, fext); break;
    case POT_AZIZ: hipLaunchKernelGGL((k_bead_forces<D, POT_AZIZ>), grid, blk, lds, s, fa, ctx->x, f, fext); break;
    default: hipLaunchKernelGGL((k_bead_forces<D, POT_FREE>), grid, blk, lds, s, fa, ctx->x, f, fext); break;
    }
}

static void launch_bead_forces(pimdb_ctx* ctx, hipStream_t s) {
    double* fext = ctx->prm.propagator == PIMDB_PROP_NORMAL_MODES ? ctx->fext : nullptr;
    switch (ctx->D) {
    case 1: launch_bead_forces_d<1>(ctx, s, ctx->f, fext); break;
    case 2: launch_bead_forces_d<2>(ctx, s, ctx->f, fext); break;
    default: launch_bead_forces_d<3>(ctx, s, ctx->f, fext); break;
    }
}

// Both recursions (cycle energies recomputed from the ring's prefix sums,
// which block 0 also leaves in Shi/Slo for the exterior forces).
static void launch_exchange_forward(pimdb_ctx* ctx, hipStream_t s) {
    const int N = ctx->N, P = ctx->P, D = ctx->D;
    const double* x1 = ctx->x;
    const double* xP = ctx->x + (size_t)(P - 1) * D * N;
    if (ctx->xws) {   // a ring too large for the LDS of one workgroup
        switch (D) {
        case 1: hipLaunchKernelGGL(k_exchange_recursions_gws<1>, dim3(2), dim3(1024), 0, s, N, ctx->dv.bt, ctx->dv.k, ctx->prm.pbc, ctx->prm.size, x1, xP, ctx->V, ctx->Vb, ctx->Shi, ctx->Slo, ctx->status, ctx->xws); break;
        case 2: hipLaunchKernelGGL(k_exchange_recursions_gws<2>, dim3(2), dim3(1024), 0, s, N, ctx->dv.bt, ctx->dv.k, ctx->prm.pbc, ctx->prm.size, x1, xP, ctx->V, ctx->Vb, ctx->Shi, ctx->Slo, ctx->status, ctx->xws); break;
        default: hipLaunchKernelGGL(k_exchange_recursions_gws<3>, dim3(2), dim3(1024), 0, s, N, ctx->dv.bt, ctx->dv.k, ctx->prm.pbc, ctx->prm.size, x1, xP, ctx->V, ctx->Vb, ctx->Shi, ctx->Slo, ctx->status, ctx->xws); break;
        }
        return;
    }
    const size_t lds = rec_lds_bytes(N, D);
    switch (D) {
    case 1: hipLaunchKernelGGL(k_exchange_recursions<1>, dim3(2), dim3(1024), lds, s, N, ctx->dv.bt, ctx->dv.k, ctx->prm.pbc, ctx->prm.size, x1, xP, ctx->V, ctx->Vb, ctx->Shi, ctx->Slo, ctx->status); break;
    case 2: hipLaunchKernelGGL(k_exchange_recursions<2>, dim3(2), dim3(1024), lds, s, N, ctx->dv.bt, ctx->dv.k, ctx->prm.pbc, ctx->prm.size, x1, xP, ctx->V, ctx->Vb, ctx->Shi, ctx->Slo, ctx->status); break;
    default: hipLaunchKernelGGL(k_exchange_recursions<3>, dim3(2), dim3(1024), lds, s, N, ctx->dv.bt, ctx->dv.k, ctx->prm.pbc, ctx->prm.size, x1, xP, ctx->V, ctx->Vb, ctx->Shi, ctx->Slo, ctx->status); break;
    }
}

// The stored triangle E (and its row-major copy), for the outputs and
// estimators that take E itself.
static void launch_cycle_energies_ctx(pimdb_ctx* ctx, hipStream_t s) {
    const int N = ctx->N, P = ctx->P;
    const double* x1 = ctx->x;
    const double* xP = ctx->x + (size_t)(P - 1) * ctx->D * N;
    launch_cycle_energies(ctx->D, N, x1, xP, ctx->dv.k, ctx->prm.pbc, ctx->prm.size, ctx->E, ctx->ET, s);
}

template <int D>
static void launch_exchange_forces_d(pimdb_ctx* ctx, hipStream_t s, double* f, double* fex) {
    const int N = ctx->N;
    hipLaunchKernelGGL((k_exchange_forces<D>), dim3((2 * N + 3) / 4), dim3(256), 0, s, N, ctx->P, ctx->dv.bt,
                       ctx->dv.k, ctx->prm.pbc, ctx->prm.size, ctx->Shi, ctx->Slo, ctx->V, ctx->Vb, ctx->x, f, fex);
}

static void launch_exchange_forces(pimdb_ctx* ctx, hipStream_t s, double* f, double* fex) {
    switch (ctx->D) {
    case 1: launch_exchange_forces_d<1>(ctx, s, f, fex); break;
    case 2: launch_exchange_forces_d<2>(ctx, s, f, fex); break;
    default: launch_exchange_forces_d<3>(ctx, s, f, fex); break;
    }
}

static bool factorial(const pimdb_ctx* ctx) {
    return ctx->dv.bosonic && ctx->prm.exchange_algo == PIMDB_EXCHANGE_FACTORIAL;
}

// The factorial exchange (one wave): forces added into f1 / fP, or written to
// fex; V_eff to VN; the estimator triple to est.
static void launch_factorial(pimdb_ctx* ctx, hipStream_t s, double* f, double* fex, double* VN, double* est) {
    const int N = ctx->N, P = ctx->P, D = ctx->D;
    const size_t bs = (size_t)D * N;
    const double* x1 = ctx->x;
    const double* x2 = ctx->x + bs * (1 % P);
    const double* xPm1 = ctx->x + bs * ((P - 2 + P) % P);
    const double* xP = ctx->x + bs * (P - 1);
    double* f1 = f ? f : nullptr;
    double* fP = f ? f + bs * (P - 1) : nullptr;
    const size_t lds = sizeof(double) * factorial_lds_doubles(N, D);
    const double bt = ctx->dv.bt, k = ctx->dv.k, L = ctx->prm.size;
    const int pbc = ctx->prm.pbc;
    switch (D) {
    case 1: hipLaunchKernelGGL(k_factorial_exchange<1>, dim3(1), dim3(FACT_LANES), lds, s, N, P, bt, k, pbc, L, x1, x2, xPm1, xP, f1, fP, fex, VN, est); break;
    case 2: hipLaunchKernelGGL(k_factorial_exchange<2>, dim3(1), dim3(FACT_LANES), lds, s, N, P, bt, k, pbc, L, x1, x2, xPm1, xP, f1, fP, fex, VN, est); break;
    default: hipLaunchKernelGGL(k_factorial_exchange<3>, dim3(1), dim3(FACT_LANES), lds, s, N, P, bt, k, pbc, L, x1, x2, xPm1, xP, f1, fP, fex, VN, est); break;
    }
}

// Forces of every bead (updateForces on all ranks); f and (NM propagator) fext.
static int enqueue_forces(pimdb_ctx* ctx) {
    if (factorial(ctx)) {
        launch_bead_forces(ctx, ctx->s0);
        launch_factorial(ctx, ctx->s0, ctx->f, nullptr, nullptr, nullptr);
    } else if (ctx->dv.bosonic) {
        CK(hipEventRecord(ctx->ev_fork, ctx->s0));
        CK(hipStreamWaitEvent(ctx->s1, ctx->ev_fork, 0));
        launch_exchange_forward(ctx, ctx->s1);
        CK(hipEventRecord(ctx->ev_join, ctx->s1));
        launch_bead_forces(ctx, ctx->s0);
        CK(hipStreamWaitEvent(ctx->s0, ctx->ev_join, 0));
        launch_exchange_forces(ctx, ctx->s0, ctx->f, nullptr);
    } else {
        launch_bead_forces(ctx, ctx->s0);
    }
    CK(hipGetLastError());
    return PIMDB_OK;
}

static MomentaArgs momenta_args(pimdb_ctx* ctx, int ostep, const double* noise) {
    MomentaArgs m{};
    m.N = ctx->N; m.P = ctx->P; m.D = ctx->D; m.tc = ctx->tc;
    m.half_dt = 0.5 * ctx->prm.dt;
    m.c1 = ctx->dv.c1; m.c2 = ctx->dv.c2;
    m.thermostat = ctx->prm.thermostat;
    m.nm = ctx->prm.nmthermostat;
    m.Ct = ctx->C; m.CTt = ctx->CT;
    m.noise = noise;
    m.step_ctr = ctx->ctr;
    m.ostep = (uint32_t)ostep;
    m.seed = ctx->prm.seed;
    m.replica = 0;
    return m;
}

static size_t momenta_lds(pimdb_ctx* ctx) {
    return sizeof(double) * (2 * (size_t)ctx->P * ctx->tc + 256 / WAVE);
}

// One OBABO step; noise0/noise1: device noise of the two O-steps or null.
// One Nose-Hoover O-step: the pending zeroMomentum (com_in), the half kick
// (kick_f), the chains, the partial sums of the result (com_out).  Under
// normal-mode coupling the chains act on the modes: transform (k_momenta
// stage 1), chains, back-transform with the partial sums (stage 2).
static void enqueue_nh_ostep(pimdb_ctx* ctx, const double* com_in, const double* kick_f, double* com_out) {
    NhArgs h = ctx->nh;
    h.half_dt = 0.5 * ctx->prm.dt;
    if (!ctx->prm.nmthermostat) {
        h.com_in = com_in;
        h.ncom_in = ctx->ncom;
        h.kick_f = kick_f;
        h.com_out = com_out;
        hipLaunchKernelGGL(k_nh_ostep, dim3(ctx->P), dim3(256), 0, ctx->s0, h, ctx->p);
        return;
    }
    dim3 tgrid(ctx->ntiles, ctx->D);
    MomentaArgs m = momenta_args(ctx, 0, nullptr);
    m.thermostat = 0;
    m.stage = 1;
    m.modes = ctx->modes;
    m.com_in = com_in;
    m.ncom_in = ctx->ncom;
    m.kick_f = kick_f;
    hipLaunchKernelGGL(k_momenta, tgrid, dim3(256), momenta_lds(ctx), ctx->s0, m, ctx->p);
    h.com_in = nullptr;
    h.kick_f = nullptr;
    h.com_out = nullptr;
    hipLaunchKernelGGL(k_nh_ostep, dim3(ctx->P), dim3(256), 0, ctx->s0, h, ctx->modes);
    m.stage = 2;
    m.com_in = nullptr;
    m.kick_f = nullptr;
    m.com_out = com_out;
    hipLaunchKernelGGL(k_momenta, tgrid, dim3(256), momenta_lds(ctx), ctx->s0, m, ctx->p);
}

static int enqueue_step(pimdb_ctx* ctx, const double* noise0, const double* noise1) {
    const bool fixcom = ctx->prm.fixcom != 0;
    const int N = ctx->N, P = ctx->P, D = ctx->D;
    dim3 tgrid(ctx->ntiles, D);
    // K1: deferred zeroMomentum + O-step
    if (ctx->prm.thermostat >= PIMDB_THERMO_NOSE_HOOVER) {
        enqueue_nh_ostep(ctx, fixcom ? ctx->com1 : nullptr, nullptr, fixcom ? ctx->com0 : nullptr);
    } else {
        MomentaArgs m1 = momenta_args(ctx, 0, noise0);
        m1.com_in = fixcom ? ctx->com1 : nullptr;
        m1.ncom_in = ctx->ncom;
        m1.com_out = fixcom ? ctx->com0 : nullptr;
        hipLaunchKernelGGL(k_momenta, tgrid, dim3(256), momenta_lds(ctx), ctx->s0, m1, ctx->p);
    }
    // K2: zeroMomentum + B + A   (or the normal-mode propagator)
    if (ctx->prm.propagator == PIMDB_PROP_NORMAL_MODES) {
        NmPropArgs a{};
        a.N = N; a.P = P; a.D = D; a.tc = ctx->tc;
        a.half_dt = 0.5 * ctx->prm.dt; a.dt = ctx->prm.dt; a.mass = ctx->prm.mass;
        a.Ct = ctx->C; a.CTt = ctx->CT; a.rot = ctx->rot;
        a.com_in = fixcom ? ctx->com0 : nullptr; a.ncom_in = ctx->ncom;
        a.fext = ctx->fext; a.step_ctr = ctx->ctr;
        hipLaunchKernelGGL(k_nm_propagate, tgrid, dim3(256), sizeof(double) * 4 * (size_t)P * ctx->tc, ctx->s0, a,
                           ctx->x, ctx->p);
    } else {
        DriftArgs d{};
        d.N = N; d.P = P; d.D = D;
        d.half_dt = 0.5 * ctx->prm.dt; d.dt = ctx->prm.dt; d.mass = ctx->prm.mass;
        d.com_in = fixcom ? ctx->com0 : nullptr; d.ncom_in = ctx->ncom;
        d.step_ctr = ctx->ctr;
        hipLaunchKernelGGL(k_kick_drift, dim3((unsigned)((ctx->nbd + 255) / 256)), dim3(256), 0, ctx->s0, d, ctx->x,
                           ctx->p, ctx->f);
    }
    // K3: forces
    int rc = enqueue_forces(ctx);
    if (rc) return rc;
    // K4: B + O-step + partial COM sums
    const double* kick = ctx->prm.propagator == PIMDB_PROP_NORMAL_MODES ? ctx->fext : ctx->f;
    if (ctx->prm.thermostat >= PIMDB_THERMO_NOSE_HOOVER) {
        enqueue_nh_ostep(ctx, nullptr, kick, fixcom ? ctx->com1 : nullptr);
    } else {
        MomentaArgs m2 = momenta_args(ctx, 1, noise1);
        m2.kick_f = kick;
        m2.com_out = fixcom ? ctx->com1 : nullptr;
        hipLaunchKernelGGL(k_momenta, tgrid, dim3(256), momenta_lds(ctx), ctx->s0, m2, ctx->p);
    }
    CK(hipGetLastError());
    return PIMDB_OK;
}

// Apply the pending zeroMomentum of the last O-step and clear the partials.
static int finish_steps(pimdb_ctx* ctx) {
    if (ctx->prm.fixcom) {
        hipLaunchKernelGGL(k_com_subtract, dim3((unsigned)((ctx->nbd + 255) / 256)), dim3(256), 0, ctx->s0, ctx->N,
                           ctx->P, ctx->D, ctx->com1, ctx->ncom, ctx->p);
        CK(hipMemsetAsync(ctx->com1, 0, sizeof(double) * (size_t)ctx->ncom * ctx->D, ctx->s0));
    }
    CK(hipGetLastError());
    return PIMDB_OK;
}

static int check_status(pimdb_ctx* ctx) {
    int st = 0;
    CK(hipMemcpyAsync(&st, ctx->status, sizeof(int), hipMemcpyDeviceToHost, ctx->s0));
    CK(hipStreamSynchronize(ctx->s0));
    if (st) {
        CK(hipMemsetAsync(ctx->status, 0, sizeof(int), ctx->s0));
        CK(hipStreamSynchronize(ctx->s0));
        return ctx->fail(PIMDB_E_NONFINITE, "Invalid sig_denom in bosonic exchange potential (non-finite V)");
    }
    return PIMDB_OK;
}

static double* pinned_buffer(pimdb_ctx* ctx, size_t n) {
    if (ctx->pinned_cap < n) {
        if (ctx->pinned) (void)hipHostFree(ctx->pinned);
        ctx->pinned = nullptr;
        ctx->pinned_cap = 0;
        if (hipHostMalloc((void**)&ctx->pinned, n * sizeof(double), hipHostMallocDefault) != hipSuccess) return nullptr;
        ctx->pinned_cap = n;
    }
    return ctx->pinned;
}

// ------------------------------------------------------------------ C ABI

extern "C" {

int pimdb_ctx_create(const pimdb_params* prm, pimdb_ctx** out) {
    if (!prm || !out) return PIMDB_E_ARG;
    *out = nullptr;
    std::string why;
    if (!validate_params(*prm, &why)) {
        set_create_error(why);
        return PIMDB_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_create_error("no HIP device available");
        return PIMDB_E_HIP;
    }
    pimdb_ctx* ctx = new pimdb_ctx();
    ctx->prm = *prm;
    ctx->dv = derive(*prm);
    ctx->N = prm->natoms; ctx->P = prm->nbeads; ctx->D = prm->ndim;
    ctx->nbd = (size_t)ctx->N * ctx->P * ctx->D;
    ctx->tc = column_tile(ctx->P);
    ctx->ntiles = (ctx->N + ctx->tc - 1) / ctx->tc;
    const int N = ctx->N, P = ctx->P, D = ctx->D;
    int rc = PIMDB_OK;
    auto alloc = [&](auto** ptr, size_t bytes) {
        if (rc) return;
        if (hipMalloc((void**)ptr, bytes) != hipSuccess || hipMemset(*ptr, 0, bytes) != hipSuccess)
            rc = ctx->fail(PIMDB_E_HIP, "device allocation failed");
    };
    alloc(&ctx->x, sizeof(double) * ctx->nbd);
    alloc(&ctx->p, sizeof(double) * ctx->nbd);
    alloc(&ctx->f, sizeof(double) * ctx->nbd);
    alloc(&ctx->fext, sizeof(double) * ctx->nbd);
    alloc(&ctx->E, sizeof(double) * ((size_t)N * (N + 1) / 2));
    alloc(&ctx->ET, sizeof(double) * ((size_t)N * (N + 1) / 2));
    alloc(&ctx->V, sizeof(double) * (N + 1));
    alloc(&ctx->Vb, sizeof(double) * (N + 1));
    alloc(&ctx->Shi, sizeof(double) * N);
    alloc(&ctx->Slo, sizeof(double) * N);
    // both recursions' workspaces in device memory when one exceeds the LDS
    // of a workgroup (the same kernel body; slower, no size limit of its own)
    if (ctx->dv.bosonic && prm->exchange_algo != PIMDB_EXCHANGE_FACTORIAL && rec_lds_bytes(N, D) > device_lds_limit())
        alloc(&ctx->xws, 2 * rec_lds_bytes(N, D));
    alloc(&ctx->prob, sizeof(double) * (size_t)N * N);
    alloc(&ctx->fex, sizeof(double) * 2 * (size_t)D * N);
    alloc(&ctx->C, sizeof(double) * (size_t)P * P);
    alloc(&ctx->CT, sizeof(double) * (size_t)P * P);
    alloc(&ctx->rot, sizeof(double) * 4 * (size_t)P);
    // partial momentum sums: one row per k_momenta tile, or per bead when the
    // Cartesian Nose-Hoover O-step (one workgroup per bead) produces them
    const bool nh = prm->thermostat >= PIMDB_THERMO_NOSE_HOOVER;
    ctx->ncom = (nh && !prm->nmthermostat) ? P : ctx->ntiles;
    alloc(&ctx->com0, sizeof(double) * (size_t)ctx->ncom * D);
    alloc(&ctx->com1, sizeof(double) * (size_t)ctx->ncom * D);
    if (nh) {
        NhArgs& h = ctx->nh;
        h.N = N; h.P = P; h.D = D; h.nc = prm->nchains;
        h.kind = prm->thermostat;
        h.groups = prm->thermostat == PIMDB_THERMO_NOSE_HOOVER ? 1 : prm->thermostat == PIMDB_THERMO_NOSE_HOOVER_NP ? N : N * D;
        const int ndof = prm->thermostat == PIMDB_THERMO_NOSE_HOOVER ? D * N : prm->thermostat == PIMDB_THERMO_NOSE_HOOVER_NP ? D : 1;
        h.mass = prm->mass; h.dt = prm->dt;
        h.thermo_beta = ctx->dv.thermo_beta;   // beta / P
        h.Qi = 1.0 * 1.0 * ctx->dv.beta / P;    // hbar^2 beta / P, nose_hoover.cpp:17
        h.Q1 = ndof * h.Qi;
        h.required = ndof / h.thermo_beta;
        const size_t nch = (size_t)P * h.groups * h.nc;
        alloc(&h.eta, sizeof(double) * nch);
        alloc(&h.eta_dot, sizeof(double) * nch);
        alloc(&h.eta_dot_dot, sizeof(double) * nch);
        if (prm->nmthermostat) alloc(&ctx->modes, sizeof(double) * ctx->nbd);
    }
    alloc(&ctx->obs, sizeof(double) * (2 * (size_t)P + 3 * (size_t)P * ((N + FORCE_TILE - 1) / FORCE_TILE) + 5));
    alloc(&ctx->status, sizeof(int));
    alloc(&ctx->ctr, sizeof(unsigned long long));
    if (!rc && (hipStreamCreateWithFlags(&ctx->s0, hipStreamNonBlocking) != hipSuccess ||
                hipStreamCreateWithFlags(&ctx->s1, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
                hipEventCreate(&ctx->t0) != hipSuccess || hipEventCreate(&ctx->t1) != hipSuccess))
        rc = ctx->fail(PIMDB_E_HIP, "stream/event creation failed");
    if (!rc) {
        std::vector<double> C0((size_t)P * P), CT0((size_t)P * P), Cm((size_t)P * P), CTm((size_t)P * P),
            rot(4 * (size_t)P);
        nm_matrices(P, C0.data(), CT0.data());
        transpose_square(P, C0.data(), Cm.data());    // device holds C^T: [j][k]
        transpose_square(P, CT0.data(), CTm.data());  // and CT^T: [k][b]
        nm_rotation(*prm, ctx->dv, rot.data());
        if (hipMemcpy(ctx->C, Cm.data(), sizeof(double) * Cm.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(ctx->CT, CTm.data(), sizeof(double) * CTm.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(ctx->rot, rot.data(), sizeof(double) * rot.size(), hipMemcpyHostToDevice) != hipSuccess)
            rc = ctx->fail(PIMDB_E_HIP, "upload of normal-mode tables failed");
    }
    if (!rc && ctx->dv.bosonic && prm->exchange_algo == PIMDB_EXCHANGE_FACTORIAL &&
        sizeof(double) * factorial_lds_doubles(N, D) > device_lds_limit())
        rc = ctx->fail(PIMDB_E_ARG, "natoms too large for the factorial exchange workgroup");
    if (rc) {
        set_create_error(ctx->err);
        pimdb_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return PIMDB_OK;
}

void pimdb_ctx_destroy(pimdb_ctx* ctx) {
    if (!ctx) return;
    if (ctx->s0) (void)hipStreamSynchronize(ctx->s0);
    if (ctx->s1) (void)hipStreamSynchronize(ctx->s1);
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    double* bufs[] = {ctx->x, ctx->p, ctx->f, ctx->fext, ctx->E, ctx->ET, ctx->V, ctx->Vb, ctx->prob, ctx->fex,
                      ctx->C, ctx->CT, ctx->rot, ctx->com0, ctx->com1, ctx->noise, ctx->obs, ctx->Shi, ctx->Slo,
                      ctx->nh.eta, ctx->nh.eta_dot, ctx->nh.eta_dot_dot, ctx->modes, ctx->xws};
    for (double* b : bufs)
        if (b) (void)hipFree(b);
    if (ctx->status) (void)hipFree(ctx->status);
    if (ctx->ctr) (void)hipFree(ctx->ctr);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->t0) (void)hipEventDestroy(ctx->t0);
    if (ctx->t1) (void)hipEventDestroy(ctx->t1);
    if (ctx->s0) (void)hipStreamDestroy(ctx->s0);
    if (ctx->s1) (void)hipStreamDestroy(ctx->s1);
    delete ctx;
}

const char* pimdb_last_error(const pimdb_ctx* ctx) { return ctx ? ctx->err.c_str() : create_error(); }

int pimdb_set_state(pimdb_ctx* ctx, const double* x, const double* p) {
    if (!ctx) return PIMDB_E_ARG;
    const double* src[2] = {x, p};
    double* dst[2] = {ctx->x, ctx->p};
    for (int k = 0; k < 2; ++k) {
        if (!src[k]) continue;
        double* st = pinned_buffer(ctx, ctx->nbd);
        if (!st) return ctx->fail(PIMDB_E_HIP, "pinned allocation failed");
        to_soa(src[k], st, ctx->N, ctx->P, ctx->D);
        CK(hipMemcpyAsync(dst[k], st, sizeof(double) * ctx->nbd, hipMemcpyHostToDevice, ctx->s0));
        CK(hipStreamSynchronize(ctx->s0));
    }
    return PIMDB_OK;
}

int pimdb_get_state(pimdb_ctx* ctx, double* x, double* p, double* f) {
    if (!ctx) return PIMDB_E_ARG;
    double* dst[3] = {x, p, f};
    const double* src[3] = {ctx->x, ctx->p, ctx->f};
    for (int k = 0; k < 3; ++k) {
        if (!dst[k]) continue;
        double* st = pinned_buffer(ctx, ctx->nbd);
        if (!st) return ctx->fail(PIMDB_E_HIP, "pinned allocation failed");
        CK(hipMemcpyAsync(st, src[k], sizeof(double) * ctx->nbd, hipMemcpyDeviceToHost, ctx->s0));
        CK(hipStreamSynchronize(ctx->s0));
        from_soa(st, dst[k], ctx->N, ctx->P, ctx->D);
    }
    return PIMDB_OK;
}

int pimdb_forces(pimdb_ctx* ctx) {
    if (!ctx) return PIMDB_E_ARG;
    int rc = enqueue_forces(ctx);
    if (rc) return rc;
    return check_status(ctx);
}

int pimdb_exchange_forward(pimdb_ctx* ctx, double* VB, double* V, double* Vb, double* prob, double* E) {
    if (!ctx) return PIMDB_E_ARG;
    if (!ctx->dv.bosonic) return ctx->fail(PIMDB_E_STATE, "exchange requires bosonic=true and nbeads > 1");
    const int N = ctx->N;
    if (factorial(ctx)) {
        if (V || Vb || prob || E)
            return ctx->fail(PIMDB_E_ARG, "the factorial exchange has no V, Vb, E or connection probabilities");
        launch_factorial(ctx, ctx->s0, nullptr, nullptr, ctx->V + N, nullptr);
        CK(hipGetLastError());
        const int rc = check_status(ctx);   // waits for the stream
        if (rc) return rc;
        if (VB) CK(hipMemcpy(VB, ctx->V + N, sizeof(double), hipMemcpyDeviceToHost));
        return PIMDB_OK;
    }
    launch_exchange_forward(ctx, ctx->s0);
    if (prob || E) launch_cycle_energies_ctx(ctx, ctx->s0);
    if (prob) {
        const size_t nn = (size_t)N * N;
        hipLaunchKernelGGL(k_exchange_probabilities, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, ctx->s0, N,
                           ctx->dv.bt, ctx->E, ctx->V, ctx->Vb, ctx->prob);
    }
    CK(hipGetLastError());
    int rc = check_status(ctx);
    if (rc) return rc;
    if (VB) CK(hipMemcpy(VB, ctx->V + N, sizeof(double), hipMemcpyDeviceToHost));
    if (V) CK(hipMemcpy(V, ctx->V, sizeof(double) * (N + 1), hipMemcpyDeviceToHost));
    if (Vb) CK(hipMemcpy(Vb, ctx->Vb, sizeof(double) * (N + 1), hipMemcpyDeviceToHost));
    if (prob) CK(hipMemcpy(prob, ctx->prob, sizeof(double) * (size_t)N * N, hipMemcpyDeviceToHost));
    if (E) CK(hipMemcpy(E, ctx->E, sizeof(double) * (size_t)N * (N + 1) / 2, hipMemcpyDeviceToHost));
    return PIMDB_OK;
}

int pimdb_exchange_adjoint(pimdb_ctx* ctx, double* f_first, double* f_last) {
    if (!ctx) return PIMDB_E_ARG;
    if (!ctx->dv.bosonic) return ctx->fail(PIMDB_E_STATE, "exchange requires bosonic=true and nbeads > 1");
    if (factorial(ctx)) {
        launch_factorial(ctx, ctx->s0, nullptr, ctx->fex, nullptr, nullptr);
    } else {
        launch_exchange_forward(ctx, ctx->s0);
        launch_exchange_forces(ctx, ctx->s0, nullptr, ctx->fex);
    }
    CK(hipGetLastError());
    int rc = check_status(ctx);
    if (rc) return rc;
    const int N = ctx->N, D = ctx->D;
    std::vector<double> h(2 * (size_t)D * N);
    CK(hipMemcpy(h.data(), ctx->fex, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int w = 0; w < 2; ++w) {
        double* dst = w == 0 ? f_first : f_last;
        if (!dst) continue;
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) dst[(size_t)i * D + a] = h[((size_t)w * D + a) * N + i];
    }
    return PIMDB_OK;
}

int pimdb_step(pimdb_ctx* ctx, int nsteps, const double* host_noise) {
    if (!ctx || nsteps < 0) return PIMDB_E_ARG;
    const size_t per = 2 * ctx->nbd;   // noise of one step
    if (host_noise) {
        const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)nsteps, (64u << 20) / (per * 8) + 1));
        const size_t need = per * (size_t)chunk;
        if (ctx->noise_cap < need) {
            if (ctx->noise) (void)hipFree(ctx->noise);
            ctx->noise = nullptr;
            ctx->noise_cap = 0;
            CK(hipMalloc((void**)&ctx->noise, need * sizeof(double)));
            ctx->noise_cap = need;
        }
        for (int s0 = 0; s0 < nsteps; s0 += chunk) {
            const int n = std::min(chunk, nsteps - s0);
            CK(hipStreamSynchronize(ctx->s0));   // previous chunk's copy and kernels are done
            double* st = pinned_buffer(ctx, per * n);
            if (!st) return ctx->fail(PIMDB_E_HIP, "pinned allocation failed");
            for (size_t blk = 0; blk < 2 * (size_t)n; ++blk)
                to_soa(host_noise + ((size_t)s0 * 2 + blk) * ctx->nbd, st + blk * ctx->nbd, ctx->N, ctx->P, ctx->D);
            CK(hipMemcpyAsync(ctx->noise, st, sizeof(double) * per * n, hipMemcpyHostToDevice, ctx->s0));
            for (int s = 0; s < n; ++s) {
                const double* nz = ctx->noise + per * s;
                int rc = enqueue_step(ctx, nz, nz + ctx->nbd);
                if (rc) return rc;
            }
        }
    } else {
        for (int s = 0; s < nsteps; ++s) {
            if (ctx->gexec) {
                CK(hipGraphLaunch(ctx->gexec, ctx->s0));
            } else {
                int rc = enqueue_step(ctx, nullptr, nullptr);
                if (rc) return rc;
            }
        }
    }
    int rc = finish_steps(ctx);
    if (rc) return rc;
    return check_status(ctx);
}

int pimdb_graph_capture(pimdb_ctx* ctx) {
    if (!ctx) return PIMDB_E_ARG;
    if (ctx->gexec) return PIMDB_OK;
    CK(hipStreamSynchronize(ctx->s0));
    CK(hipStreamBeginCapture(ctx->s0, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(ctx, nullptr, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->s0, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return ctx->fail(PIMDB_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    ctx->graph = g;
    CK(hipGraphInstantiate(&ctx->gexec, ctx->graph, nullptr, nullptr, 0));
    return PIMDB_OK;
}

int pimdb_observables(pimdb_ctx* ctx, double* out) {
    if (!ctx || !out) return PIMDB_E_ARG;
    const int N = ctx->N, P = ctx->P, D = ctx->D;
    const int nt = (N + FORCE_TILE - 1) / FORCE_TILE;
    double* sums = ctx->obs;                 // [P][2]
    double* epart = ctx->obs + 2 * (size_t)P;    // [P][nt][3]
    double* prim = epart + 3 * (size_t)P * nt;   // [3]
    launch_bead_sums(D, N, P, ctx->dv.k, ctx->prm.pbc, ctx->prm.size, ctx->x, ctx->p, sums, ctx->s0);
    launch_bead_energies(D, ctx->prm, ctx->dv, ctx->x, epart, ctx->s0);
    if (factorial(ctx)) {
        launch_factorial(ctx, ctx->s0, nullptr, nullptr, ctx->V + N, prim);
    } else if (ctx->dv.bosonic) {
        launch_exchange_forward(ctx, ctx->s0);
        launch_cycle_energies_ctx(ctx, ctx->s0);
        if (sizeof(double) * (N + 1 + 16) > device_lds_limit())
            return ctx->fail(PIMDB_E_ARG, "natoms too large for the estimator workgroup");
        hipLaunchKernelGGL(k_prim_estimator, dim3(1), dim3(1024), sizeof(double) * (N + 1 + 16), ctx->s0, N, P,
                           ctx->dv.bt, ctx->E, ctx->V, prim);
    }
    double* nh_out = prim + 3;   // [1]
    if (ctx->prm.thermostat >= PIMDB_THERMO_NOSE_HOOVER) {
        const int ndof = ctx->nh.kind == PIMDB_THERMO_NOSE_HOOVER ? D * N : ctx->nh.kind == PIMDB_THERMO_NOSE_HOOVER_NP ? D : 1;
        hipLaunchKernelGGL(k_nh_energy, dim3(1), dim3(256), 0, ctx->s0, ctx->nh, ndof, nh_out);
    }
    CK(hipGetLastError());
    int rc = check_status(ctx);
    if (rc) return rc;
    std::vector<double> h(2 * (size_t)P + 3 * (size_t)P * nt + 4);
    CK(hipMemcpy(h.data(), ctx->obs, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    double VN = 0.0;
    if (ctx->dv.bosonic) CK(hipMemcpy(&VN, ctx->V + N, sizeof(double), hipMemcpyDeviceToHost));
    combine_observables(ctx->prm, ctx->dv, h.data(), h.data() + 2 * (size_t)P, nt, h.data() + 2 * (size_t)P + 3 * (size_t)P * nt,
                        VN, out);
    out[10] = ctx->prm.thermostat >= PIMDB_THERMO_NOSE_HOOVER ? h[2 * (size_t)P + 3 * (size_t)P * nt + 3] : 0.0;
    return PIMDB_OK;
}

int pimdb_exchange_estimators(pimdb_ctx* ctx, double* prim, double* prob_distinct, double* prob_longest) {
    if (!ctx) return PIMDB_E_ARG;
    if (!ctx->dv.bosonic) return ctx->fail(PIMDB_E_STATE, "exchange estimators need a bosonic system");
    const int N = ctx->N, P = ctx->P;
    const int nt = (N + FORCE_TILE - 1) / FORCE_TILE;
    double* out = ctx->obs + 2 * (size_t)P + 3 * (size_t)P * nt;   // [3], as pimdb_observables
    if (factorial(ctx)) {
        launch_factorial(ctx, ctx->s0, nullptr, nullptr, nullptr, out);
    } else {
        launch_exchange_forward(ctx, ctx->s0);
        launch_cycle_energies_ctx(ctx, ctx->s0);
        if (sizeof(double) * (N + 1 + 16) > device_lds_limit())
            return ctx->fail(PIMDB_E_ARG, "natoms too large for the estimator workgroup");
        hipLaunchKernelGGL(k_prim_estimator, dim3(1), dim3(1024), sizeof(double) * (N + 1 + 16), ctx->s0, N, P,
                           ctx->dv.bt, ctx->E, ctx->V, out);
    }
    CK(hipGetLastError());
    int rc = check_status(ctx);
    if (rc) return rc;
    double h[3];
    CK(hipMemcpy(h, out, sizeof(h), hipMemcpyDeviceToHost));
    if (prim) *prim = h[0];
    if (prob_distinct) *prob_distinct = h[1];
    if (prob_longest) *prob_longest = h[2];
    return PIMDB_OK;
}

}  // extern "C"

// Device time per repetition of `body`, from `reps` repetitions captured in
// one hipGraph and replayed on the context stream between two events (so
// host launch latency is not counted).
template <class F>
static int time_captured(pimdb_ctx* ctx, int reps, F body, double* ms_out) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    CK(hipStreamSynchronize(ctx->s0));
    CK(hipStreamBeginCapture(ctx->s0, hipStreamCaptureModeThreadLocal));
    int rc = PIMDB_OK;
    for (int r = 0; r < reps && !rc; ++r) rc = body();
    hipError_t e = hipStreamEndCapture(ctx->s0, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return ctx->fail(PIMDB_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return ctx->fail(PIMDB_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    float ms = 0.0f;
    (void)hipGraphLaunch(ge, ctx->s0);   // warm
    (void)hipEventRecord(ctx->t0, ctx->s0);
    (void)hipGraphLaunch(ge, ctx->s0);
    (void)hipEventRecord(ctx->t1, ctx->s0);
    e = hipEventSynchronize(ctx->t1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->t0, ctx->t1);
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return ctx->fail(PIMDB_E_HIP, std::string("graph timing: ") + hipGetErrorString(e));
    *ms_out = (double)ms / reps;
    return PIMDB_OK;
}

extern "C" {

int pimdb_time_kernels(pimdb_ctx* ctx, int reps, double* ms_forces, double* ms_exchange, double* ms_step) {
    if (!ctx || reps < 1) return PIMDB_E_ARG;
    int rc = PIMDB_OK;
    if (ms_forces) {
        rc = time_captured(ctx, reps, [&]() { launch_bead_forces(ctx, ctx->s0); return PIMDB_OK; }, ms_forces);
        if (rc) return rc;
    }
    if (ms_exchange) {
        *ms_exchange = 0.0;
        if (ctx->dv.bosonic) {
            rc = time_captured(ctx, reps, [&]() {
                if (factorial(ctx)) {
                    launch_factorial(ctx, ctx->s0, nullptr, ctx->fex, nullptr, nullptr);
                } else {
                    launch_exchange_forward(ctx, ctx->s0);
                    launch_exchange_forces(ctx, ctx->s0, nullptr, ctx->fex);
                }
                return PIMDB_OK;
            }, ms_exchange);
            if (rc) return rc;
        }
    }
    if (ms_forces || ms_exchange) {
        // the force timing overwrote f with the physical and spring part
        // alone: the full forces of the current positions again
        rc = enqueue_forces(ctx);
        if (rc) return rc;
    }
    if (ms_step) {
        // whole steps (device noise), as the benchmark replays them; this
        // advances the state by 2 * reps steps
        rc = time_captured(ctx, reps, [&]() { return enqueue_step(ctx, nullptr, nullptr); }, ms_step);
        if (rc) return rc;
        rc = finish_steps(ctx);
        if (rc) return rc;
    }
    CK(hipGetLastError());
    return check_status(ctx);
}

}  // extern "C"
