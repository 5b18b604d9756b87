/*
 * fluidsim_host.h -- C entry points of the C++ host layer (libfluidsim_host.so):
 * the Config/Parameter readers and the Sim driver that sit above the device
 * C ABI of fluidsim_amd.h.  The `fluidsim` executable is main() over
 * fsh_run(); the other functions let bindings (ctypes, tests) reach the
 * same readers the solver uses.
 */
#ifndef FLUIDSIM_HOST_H
#define FLUIDSIM_HOST_H

#include "fluidsim_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call on this thread. */
const char *fsh_last_error(void);

/* Parameter::initiate (include/parameter.h:28-33) on a case directory
 * <root>/config/input.ini + <root>/data/{fluid,solid,virt}_particle.dat:
 * particle counts from the .dat headers, keys from the ini, derived
 * constants.  Returns FSA_ERR_INVALID with the reference's message
 * ("FAILED TO READ ...") on a bad case. */
int fsh_parameter_initiate(const char *root, fsa_params *out);

/* The same derivation with counts given directly and the ini file at an
 * arbitrary path (for generated scenes that are not written as .dat). */
int fsh_parameter_from_ini(const char *ini_path, int n_fluid, int n_solid, int n_virt, fsa_params *out);

/* Config::Read<T>(key, default) on an ini file (src/config.cpp:59-129,
 * include/config.h:140-159).  *found is set to 1 if the key exists. */
int fsh_config_read_string(const char *ini_path, const char *key, char *buf, int buflen, int *found);
int fsh_config_read_float(const char *ini_path, const char *key, float fallback, float *out);
int fsh_config_read_int(const char *ini_path, const char *key, int fallback, int *out);
int fsh_config_read_bool(const char *ini_path, const char *key, int fallback, int *out);

/* Kernel::calculateCoeff (src/kernel.cpp:25-41) for ratio K. */
float fsh_kernel_coefficient(float impact_length_hsml_ratio);

/* main() of the reference (src/main.cu:17-30) on a case directory: read,
 * run istart..iend with .vtk output every result_interval steps into
 * <root>/out.  verbose=0 silences the progress text.  Optional ini keys
 * with no counterpart in the reference: output_fluid_pressure = true
 * (write p instead of the always-zero color into fluid_pressure-*.vtk),
 * vtk_crlf = true (end .vtk lines with CR LF, as the reference's text-mode
 * streams do on Windows), checkpoint_interval = N (every N steps write the
 * state as a resumable case <root>/checkpoint/step-<i>: input.ini with
 * istart = i and the three .dat files). */
int fsh_run(const char *root, int verbose);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDSIM_HOST_H */
