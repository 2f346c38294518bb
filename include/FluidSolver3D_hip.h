/*
 * FluidSolver3D_hip.h -- header-only C++ class with the public surface of the
 * reference's FluidSolver3D (src/FluidSim/FluidSolver3D.cuh:157-186), over the
 * fs_* C ABI (fluidsim.h).  A program written against the reference class
 * (main.cpp:18-57: construct, set the public fields, init(labels), run(frames))
 * includes this header in place of FluidSolver3D.cuh / utility.cuh and links
 * libfluidsim_hip.so (or, for checking, the CPU oracle, which exports the same
 * ABI).
 *
 * Compiles as plain C++17 (g++ or hipcc); float3 comes from the HIP vector
 * types (-I/opt/rocm/include; with a host compiler also -D__HIP_PLATFORM_AMD__).
 *
 * Errors: init() and run() throw std::runtime_error with the ABI's message,
 * as the reference's init/loadSDF throw (FluidSolver3D.cu:84, utility.cuh:242,
 * 301).  A pressure solve that misses its tolerance, which the reference
 * neither detects nor reports, completes the run and is then reported the
 * same way.
 */
#ifndef FLUIDSOLVER3D_HIP_H
#define FLUIDSOLVER3D_HIP_H

#include <hip/hip_vector_types.h>

#include <cstdio>
#include <stdexcept>
#include <string>

#include "fluidsim.h"

/* utility.cuh:58-61 */
enum FileOutputFormat { PLY = FS_FMT_PLY, OFF = FS_FMT_OFF };

class FluidSolver3D {
public:
    /* FluidSolver3D.cuh:160-166, same names, types and initial values */
    float PIC_WEIGHT = 0.5f;
    int iterPerFrame = 10;
    FileOutputFormat outputFormat = PLY;
    std::string outputTemplate = "InputData/particles_";
    bool activeBody = false;
    float3 initialBodyPos = make_float3(0.5f, 0.5f, 0.5f);
    float bodyMass = 50.0f;
    /* the reference hard-codes this path (FluidSolver3D.cu:160) */
    std::string bodySdf = "InputData/ball.sdf";

    /* FluidSolver3D.cu:4-43: the grid size is replaced by the label file's */
    FluidSolver3D(int, int, int, float dx, float dt) {
        fs_config_default(&cfg_);
        cfg_.dx = dx;
        cfg_.dt = dt;
    }
    ~FluidSolver3D() {
        if (h_) fs_destroy(h_);
    }
    FluidSolver3D(const FluidSolver3D&) = delete;
    FluidSolver3D& operator=(const FluidSolver3D&) = delete;

    /* FluidSolver3D::init FluidSolver3D.cu:81-204: label file, rigid body,
     * seeding, InputData/particles_-1.ply */
    void init(const std::string& labels) {
        cfg_.pic_weight = PIC_WEIGHT;
        cfg_.iter_per_frame = iterPerFrame;
        cfg_.output_format = outputFormat == OFF ? FS_FMT_OFF : FS_FMT_PLY;
        std::snprintf(cfg_.output_template, sizeof cfg_.output_template, "%s", outputTemplate.c_str());
        cfg_.active_body = activeBody ? 1 : 0;
        if (h_) fs_destroy(h_);
        h_ = nullptr;
        check(fs_create(&cfg_, &h_));
        check(fs_load_labels_file(h_, labels.c_str()));
        if (activeBody)
            check(fs_load_body_sdf(h_, bodySdf.c_str(), initialBodyPos.x, initialBodyPos.y, initialBodyPos.z,
                                   bodyMass));
        check(fs_seed_particles(h_, cfg_.particles_per_cell, cfg_.seed));
        std::printf("Number of particles is%lld\n", (long long)fs_num_particles(h_));
        check(fs_write_particles(h_, "InputData/particles_-1.ply", FS_FMT_PLY));
    }

    /* FluidSolver3D::run FluidSolver3D.cu:3100-3163 */
    void run(int max_steps) {
        if (!h_) throw std::runtime_error("FluidSolver3D::run before init");
        check(fs_run(h_, max_steps));
    }

    /* the handle, for the rest of the fs_* ABI (read-back, stage timers) */
    void* handle() const { return h_; }

private:
    static void check(int rc) {
        if (rc != FS_OK) throw std::runtime_error(fs_last_error());
    }
    fs_config cfg_{};
    void* h_ = nullptr;
};

#endif /* FLUIDSOLVER3D_HIP_H */
