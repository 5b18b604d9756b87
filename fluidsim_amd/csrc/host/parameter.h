// parameter.h -- the reference's parameter classes, read from
// config/input.ini and the .dat headers, and flattened into the fsa_params
// block that the device library takes.
//
// Classes, keys, defaults and derivations follow include/parameter.h:17-49,
// include/{physics,properties,domain,sim_time,kernel}.h and the matching
// src/*.cpp files (cited per member function in parameter.cpp).
#pragma once

#include <string>

#include "config.h"
#include "fluidsim_amd.h"

namespace fsa_host {

// include/global.h:4-10
constexpr float kPi = 3.141593f;
constexpr int kVirtLayer = 8;
// skin of the fluid neighbour lists when the ini has no neighbour_skin key
// (a fraction of the cut-off)
constexpr float kDefaultNeighbourSkin = 0.05f;

struct Physics {
    float gravity[3] = {0.0f, 0.0f, 0.0f};
    bool readParameter(const Config &config);
    void initiate() {}
};

struct FluidProperties {
    int number = 0;
    float sound_speed = 0, reference_density = 0, reference_density_inverse = 0;
    float gamma = 0, viscosity = 0, gamma_inv = 0;
    float coefficient_p2rho = 0, coefficient_rho2p = 0;
    float min_pressure = 0, min_density = 0, max_density = 0;
    bool is_fluid_viscid = false;
    bool readNumber(const std::string &data_dir);
    bool readParameter(const Config &config);
    void initiate();
};

struct SolidProperties {
    int number = 0;
    float reference_density = 0, youngs_modulus = 0, poisson_ratio = 0;
    float lambda = 0, nu = 0, bulk_modulus = 0;
    float artificial_viscocity[2] = {0, 0};
    float sound_speed = 0;
    bool readNumber(const std::string &data_dir);
    bool readParameter(const Config &config);
    void initiate();
};

struct VirtualProperties {
    int number = 0;
    bool readNumber(const std::string &data_dir);
    bool readParameter(const Config &) { return false; }
    void initiate() {}
};

struct Kernel {
    float smoothing_length = 0, particle_diameter = 0, particle_volume = 0, diameter_square = 0;
    int thread_num = 32;
    int precision_form = 0;  // FSA_FORM_* bits (optional key, not in the reference)
    // fluid neighbour lists with this skin (fraction of the cut-off; optional
    // key, not in the reference; 0 = walk the cells every step)
    float neighbour_skin = kDefaultNeighbourSkin;
    float smoothing_length_square = 0, smoothing_length_inverse = 0, smoothing_length_inverse_square = 0;
    float impact_length_hsml_ratio = 0, impact_distance_square = 0;
    float kernel_coefficient_1 = 0, kernel_coefficient_2 = 0, kernel_differential = 0;
    bool readParameter(const Config &config);
    void initiate();
    float calculateCoeff() const;
};

struct Domain {
    float cell_min[3] = {0, 0, 0}, cell_max[3] = {0, 0, 0};
    float domain_min[3] = {0, 0, 0}, domain_max[3] = {0, 0, 0};
    int cell_number[3] = {0, 0, 0};
    float interval[3] = {0, 0, 0}, interval_inv[3] = {0, 0, 0};
    int cell_number_total = 0;
    bool readParameter(const Config &config);
    void initiate(const Kernel &kernel);
};

struct SimTime {
    float dt = 0, current_time = 0;
    int istart = 0, iend = 0, i = 0;
    int file_step = 0;
    int solid_sub_step = 1;
    int result_interval = 1, sort_every_step = 1;
    bool readParameter(const Config &config);
    void initiate();
    bool addStep();
    bool isFirstStep() const { return i == istart + 1; }
    bool isOutputData() const { return !((i - 1) % result_interval); }
};

class Parameter {
public:
    Physics physics;
    FluidProperties fluid;
    SolidProperties solid;
    VirtualProperties virt;
    Domain domain;
    SimTime time;
    Kernel kernel;

    // Parameter::initiate (include/parameter.h:28-33) on a case directory
    // laid out as <root>/config/input.ini and <root>/data/*.dat
    void initiate(const std::string &root) {
        readParticleNumber(root + "/data");
        readFieldParameter(root + "/config/input.ini");
        initiateParameter();
    }
    void readParticleNumber(const std::string &data_dir);
    void readFieldParameter(const std::string &ini_path);
    void readFieldParameter(const Config &config);
    void initiateParameter();
    void showInfo() const;

    bool hasFluid() const { return fluid.number != 0; }
    bool hasSolid() const { return solid.number != 0; }
    bool hasVirtual() const { return virt.number != 0; }

    fsa_params toDevice() const;
};

}  // namespace fsa_host
