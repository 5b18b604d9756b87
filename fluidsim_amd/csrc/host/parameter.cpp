// parameter.cpp -- see parameter.h.
#include "parameter.h"

#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace fsa_host {

namespace {
// *Properties::readNumber (src/properties.cpp:7-32, 68-93, 107-132): the
// first line of a .dat file is the particle count and must equal the number
// of lines that follow.
bool read_count(const std::string &path, const char *what, int &number) {
    std::ifstream in(path);
    if (!in.is_open()) {
        std::cout << "failed to read " << what << " particles parameter" << std::endl;
        return true;
    }
    std::string line;
    std::getline(in, line);
    std::istringstream head(line);
    number = 0;
    head >> number;
    if (number == 0) return false;
    int lines = 0;
    while (std::getline(in, line)) ++lines;
    if (lines != number) {
        std::cout << "mismatch " << what << " particle numbers" << std::endl;
        return true;
    }
    return false;
}
}  // namespace

// src/physics.cpp:9-19: the key is a magnitude, gravity points along -y
bool Physics::readParameter(const Config &config) {
    const float g = config.Read("gravity", 9.81f);
    gravity[0] = 0.0f;
    gravity[1] = -g;
    gravity[2] = 0.0f;
    return false;
}

bool FluidProperties::readNumber(const std::string &dir) { return read_count(dir + "/fluid_particle.dat", "fluid", number); }
bool SolidProperties::readNumber(const std::string &dir) { return read_count(dir + "/solid_particle.dat", "solid", number); }
bool VirtualProperties::readNumber(const std::string &dir) { return read_count(dir + "/virt_particle.dat", "virtual", number); }

// src/properties.cpp:34-45
bool FluidProperties::readParameter(const Config &config) {
    is_fluid_viscid = config.Read("is_viscid", false);
    min_pressure = config.Read("min_pressure", -1.013e5f);
    sound_speed = config.Read("fluid_sound_speed", 1500.0f);
    reference_density = config.Read("fluid_density", 1000.0f);
    gamma = config.Read("gamma", 7.0f);
    viscosity = config.Read("viscosity", 1.01e-3f);
    return false;
}

// src/properties.cpp:47-57: Tait constants and the +-10 % density clamp
void FluidProperties::initiate() {
    const float variation_percent = 0.1f;
    coefficient_p2rho = gamma * std::pow(sound_speed, -2.0f) / reference_density;
    coefficient_rho2p = 1.0f / coefficient_p2rho;
    gamma_inv = 1.0f / gamma;
    reference_density_inverse = 1.0f / reference_density;
    min_density = (1.0f - variation_percent) * reference_density;
    max_density = (1.0f + variation_percent) * reference_density;
}

// src/properties.cpp:95-104
bool SolidProperties::readParameter(const Config &config) {
    reference_density = config.Read("solid_density", 2700.0f);
    youngs_modulus = config.Read("youngs_modulus", 67.5e9f);
    poisson_ratio = config.Read("poisson_ratio", 0.34f);
    artificial_viscocity[0] = config.Read("arti_vis_alpha", 2.5f);
    artificial_viscocity[1] = config.Read("arti_vis_beta", 2.5f);
    return false;
}

// src/properties.cpp:59-66: Lame constants and the solid sound speed
void SolidProperties::initiate() {
    lambda = youngs_modulus * poisson_ratio / (1.0f - 2.0f * poisson_ratio) / (1.0f + poisson_ratio);
    nu = youngs_modulus / (1.0f + poisson_ratio) / 2.0f;
    bulk_modulus = youngs_modulus / (3.0f * (1.0f - 2.0f * poisson_ratio));
    sound_speed = std::sqrt(bulk_modulus / reference_density);
}

// src/kernel.cpp:9-23
bool Kernel::readParameter(const Config &config) {
    smoothing_length = config.Read("smoothing_length", 0.0f);
    particle_diameter = config.Read("particle_diameter", 0.0f);
    impact_length_hsml_ratio = config.Read("impact_length_hsml_ratio", 3.0f);
    thread_num = config.Read("thread_num", 32);
    precision_form = config.Read("precision_form", 0);
    neighbour_skin = config.Read("neighbour_skin", kDefaultNeighbourSkin);
    if (smoothing_length == 0.0f || particle_diameter == 0.0f) {
        std::cout << "failed to read kernel information" << std::endl;
        return true;
    }
    if (precision_form < 0 || precision_form > (FSA_FORM_PACKED | FSA_FORM_COMPENSATED)) {
        std::cout << "precision_form must be 0..3" << std::endl;
        return true;
    }
    if (!(neighbour_skin >= 0.0f && neighbour_skin <= 1.0f)) {
        std::cout << "neighbour_skin must be in [0, 1]" << std::endl;
        return true;
    }
    return false;
}

// src/kernel.cpp:25-41: normalisation of the truncated, shifted Gaussian
//   W(q) = coef * (exp(-q^2) - exp(-K^2)) / h^3,   q < K,
// with 4*pi*int_0^K x^2 exp(-x^2) dx taken as a left Riemann sum of 10^4
// FP32 terms (this sum, not the erf closed form, is what the solver uses).
float Kernel::calculateCoeff() const {
    const int n = 10000;
    const float K = impact_length_hsml_ratio;
    const float dx = K / (float)n;
    float sum = 0.0f;
    for (int i = 0; i < n; ++i) {
        const float x = float(i) * dx;
        sum += x * x * std::exp(-x * x) * dx;
    }
    sum *= 4.0f * kPi;
    sum -= kernel_coefficient_1 * 4.0f / 3.0f * kPi * std::pow(K, 3.0f);
    return 1.0f / sum;
}

// src/kernel.cpp:43-60
void Kernel::initiate() {
    particle_volume = std::pow(particle_diameter, 3.0f);
    diameter_square = std::pow(particle_diameter, 2.0f);
    smoothing_length_square = std::pow(smoothing_length, 2.0f);
    smoothing_length_inverse = 1.0f / smoothing_length;
    smoothing_length_inverse_square = std::pow(smoothing_length_inverse, 2.0f);
    impact_distance_square = std::pow(impact_length_hsml_ratio, 2.0f) * std::pow(smoothing_length, 2.0f);
    const float hsml_diameter_ratio = smoothing_length / particle_diameter;
    kernel_coefficient_1 = std::exp(-std::pow(impact_length_hsml_ratio, 2.0f));
    const float coef = calculateCoeff();
    kernel_coefficient_2 = coef * std::pow(hsml_diameter_ratio, -3.0f);
    kernel_differential =
        2.0f * coef * std::pow(hsml_diameter_ratio, -3.0f) * std::pow(smoothing_length, -2.0f);
}

// src/domain.cpp:9-26
bool Domain::readParameter(const Config &config) {
    domain_min[0] = config.Read("xmin", 0.0f);
    domain_min[1] = config.Read("ymin", 0.0f);
    domain_min[2] = config.Read("zmin", 0.0f);
    domain_max[0] = config.Read("xmax", 0.0f);
    domain_max[1] = config.Read("ymax", 0.0f);
    domain_max[2] = config.Read("zmax", 0.0f);
    if (domain_max[0] - domain_min[0] == 0.0f || domain_max[1] - domain_min[1] == 0.0f ||
        domain_max[2] - domain_min[2] == 0.0f) {
        std::cout << "failed to read domain information" << std::endl;
        return true;
    }
    return false;
}

// src/domain.cpp:28-47: the grid is padded by VIRT_LAYER particle spacings;
// cells are at least K*h wide
void Domain::initiate(const Kernel &kernel) {
    const float thick = (float)kVirtLayer * kernel.particle_diameter;
    for (int a = 0; a < 3; ++a) {
        cell_min[a] = domain_min[a] - thick;
        cell_max[a] = domain_max[a] + thick;
    }
    const float cell_size = kernel.impact_length_hsml_ratio * kernel.smoothing_length;
    for (int a = 0; a < 3; ++a) {
        cell_number[a] = (int)((cell_max[a] - cell_min[a]) / cell_size);
        interval[a] = (cell_max[a] - cell_min[a]) / (float)cell_number[a];
        interval_inv[a] = 1.0f / interval[a];
    }
    cell_number_total = cell_number[0] * cell_number[1] * cell_number[2];
}

// src/sim_time.cpp:8-24
bool SimTime::readParameter(const Config &config) {
    dt = config.Read("dt", 0.0f);
    istart = config.Read("istart", 0);
    iend = config.Read("iend", 0);
    solid_sub_step = config.Read("solid_sub_step", 1);
    result_interval = config.Read("result_interval", 1);
    sort_every_step = config.Read("sort_every_step", 1);
    if (dt == 0.0f || istart - iend == 0) {
        std::cout << "failed to read time information" << std::endl;
        return true;
    }
    return false;
}

// src/sim_time.cpp:26-39
void SimTime::initiate() {
    i = istart;
    current_time = (float)istart * dt;
    file_step = i / result_interval;
}

bool SimTime::addStep() {
    i++;
    current_time = i * dt;
    file_step = i / result_interval;
    return i <= iend;
}

// src/parameter.cpp:54-63
void Parameter::readParticleNumber(const std::string &dir) {
    bool error = fluid.readNumber(dir);
    error = solid.readNumber(dir) || error;
    error = virt.readNumber(dir) || error;
    if (error) throw std::invalid_argument("FAILED TO READ PARTICLES PARAMETER");
}

// src/parameter.cpp:36-52
void Parameter::readFieldParameter(const std::string &ini_path) {
    Config config(ini_path);
    readFieldParameter(config);
}

void Parameter::readFieldParameter(const Config &config) {
    bool error = physics.readParameter(config);
    error = fluid.readParameter(config) || error;
    error = solid.readParameter(config) || error;
    error = virt.readParameter(config) || error;
    error = domain.readParameter(config) || error;
    error = time.readParameter(config) || error;
    error = kernel.readParameter(config) || error;
    if (error) throw std::invalid_argument("FAILED TO READ INPUT.INI");
}

// src/parameter.cpp:25-34 (kernel before domain: the grid needs K*h)
void Parameter::initiateParameter() {
    physics.initiate();
    fluid.initiate();
    solid.initiate();
    virt.initiate();
    time.initiate();
    kernel.initiate();
    domain.initiate(kernel);
}

// src/parameter.cpp:3-23
void Parameter::showInfo() const {
    std::cout << "Parameters reading finished" << std::endl;
    std::cout << "\n[physics]" << std::endl;
    std::cout << "gravity=" << physics.gravity[1] << std::endl;
    std::cout << "\n[particles]" << std::endl;
    std::cout << "number total:" << fluid.number + solid.number + virt.number << std::endl;
    std::cout << "fluid number:" << fluid.number << std::endl;
    std::cout << "solid number:" << solid.number << std::endl;
    std::cout << "virtual number:" << virt.number << std::endl;
    std::cout << "\n[domain]" << std::endl;
    std::cout << "x(" << domain.domain_min[0] << "," << domain.domain_max[0];
    std::cout << ") y(" << domain.domain_min[1] << "," << domain.domain_max[1];
    std::cout << ") z(" << domain.domain_min[2] << "," << domain.domain_max[2] << ")" << std::endl;
    std::cout << "block number:" << domain.cell_number_total << "  ";
    std::cout << domain.cell_number[0] << "*" << domain.cell_number[1] << "*" << domain.cell_number[2] << std::endl;
    std::cout << "\n[time]" << std::endl;
    std::cout << "dt=" << time.dt << "  current step=" << time.i << std::endl;
    std::cout << "start time=" << (float)time.istart * time.dt;
    std::cout << "  end time=" << (float)time.iend * time.dt << std::endl;
}

fsa_params Parameter::toDevice() const {
    fsa_params p{};
    p.n_fluid = fluid.number;
    p.n_solid = solid.number;
    p.n_virt = virt.number;
    for (int a = 0; a < 3; ++a) p.gravity[a] = physics.gravity[a];
    p.fluid_sound_speed = fluid.sound_speed;
    p.fluid_reference_density = fluid.reference_density;
    p.fluid_reference_density_inverse = fluid.reference_density_inverse;
    p.fluid_gamma = fluid.gamma;
    p.fluid_gamma_inv = fluid.gamma_inv;
    p.fluid_viscosity = fluid.viscosity;
    p.fluid_coefficient_p2rho = fluid.coefficient_p2rho;
    p.fluid_coefficient_rho2p = fluid.coefficient_rho2p;
    p.fluid_min_pressure = fluid.min_pressure;
    p.fluid_min_density = fluid.min_density;
    p.fluid_max_density = fluid.max_density;
    p.fluid_is_viscid = fluid.is_fluid_viscid ? 1 : 0;
    p.solid_reference_density = solid.reference_density;
    p.solid_youngs_modulus = solid.youngs_modulus;
    p.solid_poisson_ratio = solid.poisson_ratio;
    p.solid_lambda = solid.lambda;
    p.solid_nu = solid.nu;
    p.solid_bulk_modulus = solid.bulk_modulus;
    p.solid_sound_speed = solid.sound_speed;
    p.solid_artificial_viscosity[0] = solid.artificial_viscocity[0];
    p.solid_artificial_viscosity[1] = solid.artificial_viscocity[1];
    for (int a = 0; a < 3; ++a) {
        p.domain_min[a] = domain.domain_min[a];
        p.domain_max[a] = domain.domain_max[a];
        p.cell_min[a] = domain.cell_min[a];
        p.cell_max[a] = domain.cell_max[a];
        p.interval[a] = domain.interval[a];
        p.interval_inv[a] = domain.interval_inv[a];
        p.cell_number[a] = domain.cell_number[a];
    }
    p.cell_number_total = domain.cell_number_total;
    p.dt = time.dt;
    p.istart = time.istart;
    p.iend = time.iend;
    p.solid_sub_step = time.solid_sub_step;
    p.result_interval = time.result_interval;
    p.sort_every_step = time.sort_every_step;
    p.smoothing_length = kernel.smoothing_length;
    p.particle_diameter = kernel.particle_diameter;
    p.impact_length_hsml_ratio = kernel.impact_length_hsml_ratio;
    p.smoothing_length_square = kernel.smoothing_length_square;
    p.smoothing_length_inverse = kernel.smoothing_length_inverse;
    p.smoothing_length_inverse_square = kernel.smoothing_length_inverse_square;
    p.impact_distance_square = kernel.impact_distance_square;
    p.kernel_coefficient_1 = kernel.kernel_coefficient_1;
    p.kernel_coefficient_2 = kernel.kernel_coefficient_2;
    p.kernel_differential = kernel.kernel_differential;
    p.thread_num = kernel.thread_num;
    p.precision_form = kernel.precision_form;
    p.neighbour_skin = kernel.neighbour_skin;
    return p;
}

}  // namespace fsa_host
