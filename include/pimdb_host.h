/*
 * pimdb_host.h -- C entry points of the host library (libpimdb_host.so),
 * the C++ side that drives the GPU path the way the reference's Simulation
 * drives its ranks: configuration, units, RanMars noise streams, initial
 * conditions and output files.  The GPU work itself goes through
 * pimdb_gpu.h.  These functions exist so that callers without C++ (the
 * Python tests, a ctypes user) reach the same code; C++ callers use the
 * classes in pimd-b_amd/host directly.
 *
 * Return values: 0 on success, PIMDB_E_ARG for what the reference reports
 * as std::invalid_argument, PIMDB_E_NONFINITE for std::overflow_error,
 * PIMDB_E_HIP for anything else; the message goes to err (may be NULL).
 */
#ifndef PIMDB_HOST_H
#define PIMDB_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

#include "pimdb_gpu.h"

/* Units::convertToInternal (src/units.cpp:42-59): number in `unit` of
 * `family` ("energy", "length", ... or "number") to atomic units. */
int pimdb_host_to_internal(const char* family, const char* unit, double number, double* out, char* err, int errlen);

/* RanMars (libs/random_mars.cpp:11-103): the noise stream of one bead. */
typedef struct pimdb_host_ranmars pimdb_host_ranmars;
pimdb_host_ranmars* pimdb_host_ranmars_new(int seed);   /* NULL for an invalid seed */
double pimdb_host_ranmars_uniform(pimdb_host_ranmars* r);
double pimdb_host_ranmars_gaussian(pimdb_host_ranmars* r);
void pimdb_host_ranmars_free(pimdb_host_ranmars* r);

/* Params::Params (src/params.cpp:8-258) followed by the Simulation
 * constructor's derivations (src/simulation.cpp:16-101): the GPU parameters
 * of a config.ini plus its step count and save frequency. */
int pimdb_host_load_config(const char* path, pimdb_params* out, long* steps, long* sfreq, char* err, int errlen);

/* Initial positions and momenta of every bead as the reference's ranks draw
 * them (mt19937(seed + bead), random/xyz/grid positions, Maxwell-Boltzmann or
 * file momenta, zeroMomentum; simulation.cpp:118-217, 708-753), host layout
 * [bead][particle][axis].  Needs the GPU library loaded but not a GPU. */
int pimdb_host_initial_state(const char* path, double* x, double* p, char* err, int errlen);

/* Gaussian deviates of the beads' RanMars streams for nsteps MD steps,
 * [step][O-step][bead][particle][axis] (langevin.cpp:18-26). */
int pimdb_host_draw_noise(const char* path, long nsteps, double* out, char* err, int errlen);

/* Simulation::run (src/simulation.cpp:222-290) on the GPU: writes
 * simulation.out, report.txt and the enabled position/velocity/force files
 * under output_dir.  steps >= 0 replaces the configured step count.  flags:
 * PIMDB_HOST_DEVICE_NOISE selects on-device Philox noise instead of the
 * RanMars streams; PIMDB_HOST_FACTORIAL the factorial bosonic exchange (the
 * reference's FACTORIAL_BOSONIC_ALGORITHM build, simulation.cpp:694-698). */
enum { PIMDB_HOST_DEVICE_NOISE = 1, PIMDB_HOST_FACTORIAL = 2 };
int pimdb_host_run(const char* path, long steps, int flags, const char* output_dir, char* err, int errlen);

#ifdef __cplusplus
}
#endif
#endif
