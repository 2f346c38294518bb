// fluidsim -- command-line driver, the counterpart of the reference's main.cpp:9-60.
//
//   fluidsim [config.json]      (default InputData/config.json, as main.cpp:11)
//
// Reads the config keys of main.cpp:18-52, loads the label scene and the body
// SDF, seeds particles, writes the initial particle file
// (FluidSolver3D.cu:199-201) and runs frameAmount frames (FluidSolver3D::run).
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../../include/fluidsim.h"

static int die(const char* what) {
    std::cerr << what << ": " << fs_last_error() << std::endl;
    return 1;
}

int main(int argc, char** argv) {
    const char* cfg_path = argc > 1 ? argv[1] : "InputData/config.json";
    fs_config cfg;
    fs_config_default(&cfg);
    if (fs_config_load_json(&cfg, cfg_path) != FS_OK) return die("config");
    std::cout << "dx=" << cfg.dx << "\n";
    std::cout << "alpha = " << cfg.pic_weight << std::endl;
    void* s = nullptr;
    if (fs_create(&cfg, &s) != FS_OK) return die("create");
    if (fs_load_labels_file(s, cfg.input_label_data) != FS_OK) return die("labels");
    if (cfg.active_body &&
        fs_load_body_sdf(s, cfg.body_sdf_path, cfg.body_x, cfg.body_y, cfg.body_z, cfg.body_mass) != FS_OK)
        return die("body");
    if (fs_seed_particles(s, cfg.particles_per_cell, cfg.seed) != FS_OK) return die("seed");
    std::cout << "Number of particles is" << fs_num_particles(s) << std::endl;
    if (fs_write_particles(s, "InputData/particles_-1.ply", FS_FMT_PLY) != FS_OK) return die("output");
    if (fs_run(s, cfg.frame_amount) != FS_OK) return die("run");
    double ms[FS_NUM_STAGES];
    fs_get_stage_times(s, ms, FS_NUM_STAGES);
    const char* names[FS_NUM_STAGES] = {"labelTime", "p2gTime", "prsslvTime", "applprsTime",
                                        "updbdTime", "g2pAdvctTime", "sortTime"};
    double total = 0;
    for (int i = 0; i < FS_NUM_STAGES; ++i) total += ms[i];
    std::cout << "total  frames = " << cfg.frame_amount * cfg.iter_per_frame << std::endl;
    for (int i = 0; i < FS_NUM_STAGES; ++i) std::cout << names[i] << " = " << ms[i] / 1000.0 << std::endl;
    std::cout << "steps/s (solver only) = " << (cfg.frame_amount * cfg.iter_per_frame + 1) / (total / 1000.0)
              << std::endl;
    fs_destroy(s);
    return 0;
}
