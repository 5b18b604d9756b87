// host_capi.cpp -- C entry points of the host layer (fluidsim_host.h).
#include "fluidsim_host.h"

#include <cmath>
#include <cstring>
#include <exception>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "config.h"
#include "parameter.h"
#include "sim.h"

namespace {
thread_local std::string g_err;

int fail(const std::string &msg, int code = FSA_ERR_INVALID) {
    g_err = msg;
    return code;
}

// runs fn and maps the reference's exception types to status codes
template <class F> int guarded(F &&fn) {
    try {
        fn();
        g_err.clear();
        return FSA_OK;
    } catch (const fsa_host::Config::File_not_found &e) {
        return fail("file not found: " + e.filename);
    } catch (const fsa_host::Config::Key_not_found &e) {
        return fail("key not found: " + e.key);
    } catch (const std::invalid_argument &e) {
        return fail(e.what());
    } catch (const std::exception &e) {
        return fail(e.what(), FSA_ERR_STATE);
    }
}

// keeps the reference's progress text off stdout when not verbose
struct Quiet {
    std::streambuf *saved = nullptr;
    std::ostringstream sink;
    explicit Quiet(bool on) {
        if (on) saved = std::cout.rdbuf(sink.rdbuf());
    }
    ~Quiet() {
        if (saved) std::cout.rdbuf(saved);
    }
};
}  // namespace

extern "C" {

const char *fsh_last_error(void) { return g_err.c_str(); }

int fsh_parameter_initiate(const char *root, fsa_params *out) {
    if (!root || !out) return fail("fsh_parameter_initiate: null argument");
    Quiet q(true);
    return guarded([&] {
        fsa_host::Parameter p;
        p.initiate(root);
        *out = p.toDevice();
    });
}

int fsh_parameter_from_ini(const char *ini_path, int nf, int ns, int nv, fsa_params *out) {
    if (!ini_path || !out || nf < 0 || ns < 0 || nv < 0) return fail("fsh_parameter_from_ini: bad argument");
    Quiet q(true);
    return guarded([&] {
        fsa_host::Parameter p;
        p.fluid.number = nf;
        p.solid.number = ns;
        p.virt.number = nv;
        p.readFieldParameter(std::string(ini_path));
        p.initiateParameter();
        *out = p.toDevice();
    });
}

int fsh_config_read_string(const char *ini_path, const char *key, char *buf, int buflen, int *found) {
    if (!ini_path || !key || !buf || buflen <= 0) return fail("fsh_config_read_string: bad argument");
    return guarded([&] {
        fsa_host::Config c{std::string(ini_path)};
        const bool f = c.KeyExists(key);
        if (found) *found = f ? 1 : 0;
        const std::string v = f ? c.Read<std::string>(key) : std::string();
        std::strncpy(buf, v.c_str(), (size_t)buflen - 1);
        buf[buflen - 1] = '\0';
    });
}

int fsh_config_read_float(const char *ini_path, const char *key, float fallback, float *out) {
    if (!ini_path || !key || !out) return fail("fsh_config_read_float: bad argument");
    return guarded([&] { *out = fsa_host::Config{std::string(ini_path)}.Read<float>(key, fallback); });
}

int fsh_config_read_int(const char *ini_path, const char *key, int fallback, int *out) {
    if (!ini_path || !key || !out) return fail("fsh_config_read_int: bad argument");
    return guarded([&] { *out = fsa_host::Config{std::string(ini_path)}.Read<int>(key, fallback); });
}

int fsh_config_read_bool(const char *ini_path, const char *key, int fallback, int *out) {
    if (!ini_path || !key || !out) return fail("fsh_config_read_bool: bad argument");
    return guarded([&] { *out = fsa_host::Config{std::string(ini_path)}.Read<bool>(key, fallback != 0) ? 1 : 0; });
}

float fsh_kernel_coefficient(float ratio) {
    fsa_host::Kernel k;
    k.impact_length_hsml_ratio = ratio;
    k.kernel_coefficient_1 = std::exp(-std::pow(ratio, 2.0f));
    return k.calculateCoeff();
}

int fsh_run(const char *root, int verbose) {
    if (!root) return fail("fsh_run: null root");
    Quiet q(!verbose);
    return guarded([&] {
        fsa_host::Sim sim(root);
        sim.verbose = verbose != 0;
        {
            fsa_host::Config c(std::string(root) + "/config/input.ini");
            sim.output_true_pressure = c.Read("output_fluid_pressure", false);
            sim.vtk_crlf = c.Read("vtk_crlf", false);
            sim.checkpoint_interval = c.Read("checkpoint_interval", 0);
        }
        sim.initiate();
        sim.calculation();
    });
}

}  // extern "C"
