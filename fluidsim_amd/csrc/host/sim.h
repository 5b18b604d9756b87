// sim.h -- the Sim driver: reads a case, owns one device context, runs the
// step loop with its output cadence, and writes the reference's .vtk files.
//
// Mirrors Sim (include/sim.cuh:18-58, src/sim.cu), DeviceCalculation
// (include/calculation.cuh:9-21, src/calculation.cu:14-35) and Output
// (include/output.cuh:9-25, src/output.cu:6-335).  The device work goes
// through the C ABI of fluidsim_amd.h only.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "fluidsim_amd.h"
#include "parameter.h"

namespace fsa_host {

// host-side particle lists as read from the .dat files (include/host_data.cuh:10-163)
struct FluidHostList {
    std::vector<float> position, velocity, density, pressure;
    void addList(const std::string &line);
};
struct SolidHostList {
    std::vector<float> position, velocity, reference_position;
    std::vector<int32_t> solid_type;
    void addList(const std::string &line);
};
struct VirtHostList {
    std::vector<float> position;
    std::vector<int32_t> virt_type;
    void addList(const std::string &line);
};

class Sim {
public:
    explicit Sim(const std::string &root) : root_(root) {}
    ~Sim();
    Sim(const Sim &) = delete;
    Sim &operator=(const Sim &) = delete;

    Parameter parameter;

    // Parameter::initiate + Sim::initiate (include/sim.cuh:33-43)
    void initiate();
    // DeviceCalculation::calculation (src/calculation.cu:20-35)
    void calculation();

    fsa_sim *context() const { return ctx_; }
    const std::string &root() const { return root_; }
    bool verbose = true;
    // write the real fluid pressure into fluid_pressure-*.vtk instead of the
    // always-zero `color` field of the reference (src/output.cu:75)
    bool output_true_pressure = false;
    // end .vtk lines with "\r\n", the bytes the reference's text-mode
    // fstream writes on its Windows build (src/output.cu:32-35)
    bool vtk_crlf = false;
    // every checkpoint_interval steps (0: never), write the state after the
    // step as a case directory <root>/checkpoint/step-<i>: the input.ini with
    // istart = i and the three .dat files (include/host_data.cuh:25, 84, 152),
    // which `fluidsim` resumes from
    int checkpoint_interval = 0;
    void writeCheckpoint(const std::string &dir);

private:
    void readParticleData();
    void copyData();
    void outputData();
    void check(int rc, const char *what) const;

    std::string root_;
    fsa_sim *ctx_ = nullptr;
    FluidHostList fluid_host_;
    SolidHostList solid_host_;
    VirtHostList virt_host_;
    // the columns of the .dat files that never change on the device, kept
    // for checkpoints: solid reference positions and types, walls
    std::vector<float> solid_ref_, virt_pos_;
    std::vector<int32_t> solid_type_, virt_type_;
};

// ASCII POLYDATA writers of src/output.cu:6-297 (std::to_string formatting)
namespace vtk {
void write_vectors(const std::string &path, const char *title, const char *name, const float *points,
                   const float *vectors, int n, bool crlf = false);
void write_scalars(const std::string &path, const char *name, const float *points, const float *scalars, int n,
                   bool crlf = false);
}  // namespace vtk

}  // namespace fsa_host
