// sim.cpp -- see sim.h.
#include "sim.h"

#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace fsa_host {

// Fluid_Host_List::addList (include/host_data.cuh:20-38): x y z ux uy uz rho p
void FluidHostList::addList(const std::string &line) {
    std::istringstream in(line);
    float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (float &f : v) in >> f;
    position.insert(position.end(), v, v + 3);
    velocity.insert(velocity.end(), v + 3, v + 6);
    density.push_back(v[6]);
    pressure.push_back(v[7]);
}

// Solid_Host_List::addList (include/host_data.cuh:79-101): x y z ux uy uz X Y Z type
void SolidHostList::addList(const std::string &line) {
    std::istringstream in(line);
    float v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int type = 0;
    for (float &f : v) in >> f;
    in >> type;
    position.insert(position.end(), v, v + 3);
    velocity.insert(velocity.end(), v + 3, v + 6);
    reference_position.insert(reference_position.end(), v + 6, v + 9);
    solid_type.push_back(type);
}

// Virt_Host_List::addList (include/host_data.cuh:147-163): x y z type
void VirtHostList::addList(const std::string &line) {
    std::istringstream in(line);
    float v[3] = {0, 0, 0};
    int type = 0;
    for (float &f : v) in >> f;
    in >> type;
    position.insert(position.end(), v, v + 3);
    virt_type.push_back(type);
}

Sim::~Sim() {
    if (ctx_) fsa_destroy(ctx_);
}

void Sim::check(int rc, const char *what) const {
    if (rc != FSA_OK) throw std::runtime_error(std::string(what) + ": " + fsa_last_error());
}

namespace {
template <class List> bool read_list(const std::string &path, int number, const char *what, List &list) {
    if (number == 0) {
        std::cout << "\nno " << what << " particle" << std::endl;
        return false;
    }
    if (number < 0) {
        std::cout << what << " particle number should >= 0" << std::endl;
        return true;
    }
    std::ifstream in(path);
    if (!in.is_open()) {
        std::cout << "failed to read " << what << " particles data" << std::endl;
        return true;
    }
    std::string line;
    std::getline(in, line);
    while (std::getline(in, line)) list.addList(line);
    return false;
}
}  // namespace

// Sim::readParticleData, src/sim.cu:40-151
void Sim::readParticleData() {
    if (verbose) std::cout << "\nReading Particles..." << std::endl;
    const std::string dir = root_ + "/data";
    bool error = read_list(dir + "/fluid_particle.dat", parameter.fluid.number, "fluid", fluid_host_);
    error = read_list(dir + "/solid_particle.dat", parameter.solid.number, "solid", solid_host_) || error;
    error = read_list(dir + "/virt_particle.dat", parameter.virt.number, "virtual", virt_host_) || error;
    if (error) throw std::invalid_argument("FAILED TO READ PARTICLES DATA");
}

// Sim::copyData + setPointer + MemoryCopy::copyConstant (src/sim.cu:304-321,
// src/memory_copy.cu:8-12): one context holds parameters and buffers
void Sim::copyData() {
    const fsa_params p = parameter.toDevice();
    check(fsa_create(&p, nullptr, &ctx_), "fsa_create");
    if (p.n_fluid > 0)
        check(fsa_upload_fluid(ctx_, fluid_host_.position.data(), fluid_host_.velocity.data(),
                               fluid_host_.density.data(), fluid_host_.pressure.data()),
              "fsa_upload_fluid");
    if (p.n_solid > 0)
        check(fsa_upload_solid(ctx_, solid_host_.position.data(), solid_host_.velocity.data(),
                               solid_host_.reference_position.data(), solid_host_.solid_type.data()),
              "fsa_upload_solid");
    if (p.n_virt > 0)
        check(fsa_upload_virt(ctx_, virt_host_.position.data(), virt_host_.virt_type.data()), "fsa_upload_virt");
    if (checkpoint_interval > 0) {
        solid_ref_ = solid_host_.reference_position;
        solid_type_ = solid_host_.solid_type;
        virt_pos_ = virt_host_.position;
        virt_type_ = virt_host_.virt_type;
    }
    // Sim::clearHost, src/sim.cu:323-330
    fluid_host_ = FluidHostList();
    solid_host_ = SolidHostList();
    virt_host_ = VirtHostList();
}

void Sim::initiate() {
    parameter.initiate(root_);
    if (verbose) parameter.showInfo();
    readParticleData();
    copyData();
    if (verbose) std::cout << "\nInitialization completed" << std::endl;
}

// DeviceCalculation::calculation, src/calculation.cu:20-35.  Output happens
// at the start of a step whose (i-1) is a multiple of result_interval, and
// once more after the loop.
void Sim::calculation() {
    SimTime &time = parameter.time;
    while (time.addStep()) {
        outputData();
        check(fsa_particle_cell_update(ctx_), "particleCellUpdate");
        check(fsa_fill_cell(ctx_), "fillCell");
        check(fsa_couple_calculation(ctx_), "coupleCalculation");
        check(fsa_virt_update(ctx_), "virtUpdate");
        check(fsa_fluid_calculation(ctx_), "fluidCalculation");
        check(fsa_solid_calculation(ctx_), "solidCalculation");
        if (checkpoint_interval > 0 && time.i % checkpoint_interval == 0)
            writeCheckpoint(root_ + "/checkpoint/step-" + std::to_string(time.i));
    }
    outputData();
    check(fsa_synchronize(ctx_), "synchronize");
}

namespace {
void make_dir(const std::string &path) { ::mkdir(path.c_str(), 0755); }

inline void append_float(std::string &s, float v) {
    // std::to_string(float) formats the value with "%f"
    char buf[64];
    int n = std::snprintf(buf, sizeof buf, "%f", (double)v);
    s.append(buf, (size_t)n);
}

std::string header(const char *title, int n) {
    return std::string("# vtk DataFile Version 2.0\n") + title + "\nASCII\nDATASET POLYDATA\nPOINTS " +
           std::to_string(n) + " float\n";
}

void append_points(std::string &s, const float *p, int n) {
    for (int i = 0; i < n; ++i) {
        append_float(s, p[3 * i]);
        s += ' ';
        append_float(s, p[3 * i + 1]);
        s += ' ';
        append_float(s, p[3 * i + 2]);
        s += '\n';
    }
}

void flush(const std::string &path, const std::string &s, bool crlf) {
    std::ofstream out(path, std::ios::out | std::ios::binary);
    if (!crlf) {
        out << s;
        return;
    }
    std::string t;
    t.reserve(s.size() + s.size() / 16);
    for (char c : s) {
        if (c == '\n') t += '\r';
        t += c;
    }
    out << t;
}

// one row of a .dat file: "%.9g" round-trips every float through the
// reader's istream >> float
void append_row(std::string &s, const float *v, int n, const int32_t *tail) {
    char buf[32];
    for (int k = 0; k < n; ++k) {
        const int m = std::snprintf(buf, sizeof buf, k ? " %.9g" : "%.9g", (double)v[k]);
        s.append(buf, (size_t)m);
    }
    if (tail) s += " " + std::to_string(*tail);
    s += '\n';
}

void write_text(const std::string &path, const std::string &s) {
    std::ofstream out(path, std::ios::out | std::ios::binary);
    out << s;
    if (!out) throw std::runtime_error("cannot write " + path);
}
}  // namespace

// The state after step i as a case directory: config/input.ini with
// istart = i (SimTime::initiate, src/sim_time.cpp:26-31, then continues the
// step and file numbering) and data/{fluid,solid,virt}_particle.dat in the
// formats Sim::readParticleData reads (include/host_data.cuh:25, 84, 152).
// F, P and L are recomputed from x and X in every sub-step, the coupling
// fields and wall states in every step, and the sort order is a function of
// the positions and .dat indices, so a resumed plain-form run repeats the
// uninterrupted one bit for bit.  (The compensated form's carried rounding
// errors, below an ulp, are not part of the format: it resumes with none.)
void Sim::writeCheckpoint(const std::string &dir) {
    make_dir(root_ + "/checkpoint");
    make_dir(dir);
    make_dir(dir + "/config");
    make_dir(dir + "/data");
    {
        std::ifstream in(root_ + "/config/input.ini");
        std::string line, text;
        bool replaced = false;
        while (std::getline(in, line)) {
            const std::string body = line.substr(0, line.find('#'));
            const size_t eq = body.find('=');
            std::string key = eq == std::string::npos ? std::string() : body.substr(0, eq);
            key.erase(0, key.find_first_not_of(" \t"));
            key.erase(key.find_last_not_of(" \t\r") + 1);
            if (key == "istart") {
                line = "istart = " + std::to_string(parameter.time.i);
                replaced = true;
            }
            text += line + "\n";
        }
        if (!replaced) text += "[time]\nistart = " + std::to_string(parameter.time.i) + "\n";
        write_text(dir + "/config/input.ini", text);
    }
    const int nf = parameter.fluid.number, ns = parameter.solid.number, nv = parameter.virt.number;
    {
        std::vector<float> x(3 * (size_t)nf), u(3 * (size_t)nf), rho(nf), p(nf);
        if (nf > 0) check(fsa_download_fluid(ctx_, x.data(), u.data(), rho.data(), p.data()), "download fluid");
        std::string s = std::to_string(nf) + "\n";
        s.reserve(s.size() + (size_t)nf * 96);
        for (int i = 0; i < nf; ++i) {
            const float v[8] = {x[3 * i], x[3 * i + 1], x[3 * i + 2], u[3 * i], u[3 * i + 1], u[3 * i + 2], rho[i], p[i]};
            append_row(s, v, 8, nullptr);
        }
        write_text(dir + "/data/fluid_particle.dat", s);
    }
    {
        std::vector<float> x(3 * (size_t)ns), u(3 * (size_t)ns);
        if (ns > 0)
            check(fsa_download_solid(ctx_, x.data(), u.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                     nullptr),
                  "download solid");
        std::string s = std::to_string(ns) + "\n";
        for (int i = 0; i < ns; ++i) {
            const float v[9] = {x[3 * i], x[3 * i + 1], x[3 * i + 2], u[3 * i], u[3 * i + 1], u[3 * i + 2],
                                solid_ref_[3 * i], solid_ref_[3 * i + 1], solid_ref_[3 * i + 2]};
            append_row(s, v, 9, &solid_type_[i]);
        }
        write_text(dir + "/data/solid_particle.dat", s);
    }
    {
        std::string s = std::to_string(nv) + "\n";
        s.reserve(s.size() + (size_t)nv * 40);
        for (int i = 0; i < nv; ++i) append_row(s, &virt_pos_[3 * (size_t)i], 3, &virt_type_[i]);
        write_text(dir + "/data/virt_particle.dat", s);
    }
    if (verbose) std::cout << "checkpoint: " << dir << std::endl;
}

namespace vtk {
void write_vectors(const std::string &path, const char *title, const char *name, const float *points,
                   const float *vectors, int n, bool crlf) {
    std::string s = header(title, n);
    s.reserve(s.size() + (size_t)n * 80);
    append_points(s, points, n);
    s += "POINT_DATA " + std::to_string(n) + "\n";
    s += std::string("VECTORS ") + name + " float\n";
    append_points(s, vectors, n);
    flush(path, s, crlf);
}

void write_scalars(const std::string &path, const char *name, const float *points, const float *scalars, int n,
                   bool crlf) {
    std::string s = header("scalar vector file", n);
    s.reserve(s.size() + (size_t)n * 60);
    append_points(s, points, n);
    s += "POINT_DATA " + std::to_string(n) + "\n";
    s += std::string("SCALARS ") + name + " float 1\n";
    s += "LOOKUP_TABLE default\n";
    for (int i = 0; i < n; ++i) {
        append_float(s, scalars[i]);
        s += '\n';
    }
    flush(path, s, crlf);
}
}  // namespace vtk

// Output::outputData, src/output.cu:299-335: nine files per output step
void Sim::outputData() {
    const SimTime &time = parameter.time;
    if (!time.isOutputData()) return;
    const auto t0 = std::chrono::steady_clock::now();
    if (verbose) {
        std::cout << "Outputing Data..." << std::endl;
        std::cout << "--time=" << time.current_time - time.dt << "s step=" << time.i - 1 << "--" << std::endl;
    }
    const std::string out = root_ + "/out";
    make_dir(out);
    const std::string idx = std::to_string(time.file_step);
    if (parameter.hasFluid()) {
        const int n = parameter.fluid.number;
        std::vector<float> x(3 * (size_t)n), u(3 * (size_t)n), rho(n), p(n);
        check(fsa_download_fluid(ctx_, x.data(), u.data(), rho.data(), p.data()), "download fluid");
        const std::string dir = out + "/fluid";
        make_dir(dir);
        vtk::write_vectors(dir + "/fluid_velocity-" + idx + ".vtk", "Velocity vector file", "velocity", x.data(),
                           u.data(), n, vtk_crlf);
        vtk::write_scalars(dir + "/fluid_density-" + idx + ".vtk", "density", x.data(), rho.data(), n, vtk_crlf);
        // the reference writes the `color` field here, which its kernels set to 0
        if (!output_true_pressure) std::fill(p.begin(), p.end(), 0.0f);
        vtk::write_scalars(dir + "/fluid_pressure-" + idx + ".vtk", "pressure", x.data(), p.data(), n, vtk_crlf);
    }
    if (parameter.hasSolid()) {
        const int n = parameter.solid.number;
        std::vector<float> x(3 * (size_t)n), u(3 * (size_t)n), p(n), nrm(3 * (size_t)n);
        check(fsa_download_solid(ctx_, x.data(), u.data(), p.data(), nrm.data(), nullptr, nullptr, nullptr, nullptr,
                                 nullptr),
              "download solid");
        const std::string dir = out + "/solid";
        make_dir(dir);
        vtk::write_vectors(dir + "/solid_velocity-" + idx + ".vtk", "Velocity vector file", "velocity", x.data(),
                           u.data(), n, vtk_crlf);
        vtk::write_scalars(dir + "/solid_pressure-" + idx + ".vtk", "pressure", x.data(), p.data(), n, vtk_crlf);
        std::vector<float> rho(n, parameter.solid.reference_density);
        vtk::write_scalars(dir + "/solid_density-" + idx + ".vtk", "density", x.data(), rho.data(), n, vtk_crlf);
        vtk::write_vectors(dir + "/solid_normal-" + idx + ".vtk", "Velocity vector file", "normal", x.data(),
                           nrm.data(), n, vtk_crlf);
    }
    if (parameter.hasVirtual()) {
        const int n = parameter.virt.number;
        std::vector<float> x(3 * (size_t)n), nrm(3 * (size_t)n), p(n);
        check(fsa_download_virt(ctx_, x.data(), nrm.data(), nullptr, nullptr, p.data()), "download virt");
        const std::string dir = out + "/virtual";
        make_dir(dir);
        vtk::write_vectors(dir + "/virt_normal-" + idx + ".vtk", "Velocity vector file", "normal", x.data(),
                           nrm.data(), n, vtk_crlf);
        vtk::write_scalars(dir + "/virt_pressure-" + idx + ".vtk", "pressure", x.data(), p.data(), n, vtk_crlf);
    }
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (verbose) std::cout << "write time=" << (long long)ms << "ms" << std::endl;
}

}  // namespace fsa_host
