// A program written against the reference's FluidSolver3D class, following
// main.cpp:9-60 step by step (config keys and defaults of main.cpp:18-52,
// the public fields, init(labels), run(frames)), compiled against
// include/FluidSolver3D_hip.h.  The config is read with fs_config_load_json
// (same keys and defaults) in place of the reference's jsoncpp reader.
#include <iostream>
#include <string>

#include "FluidSolver3D_hip.h"

int main() {
    std::string configPath = "InputData/config.json";
    fs_config json_root;
    fs_config_default(&json_root);
    if (fs_config_load_json(&json_root, configPath.c_str()) != FS_OK) {
        std::cerr << fs_last_error() << std::endl;
        return 1;
    }
    float dt = json_root.dt;
    float dx = json_root.dx;

    FluidSolver3D solver(5, 5, 5, dx, dt);
    solver.PIC_WEIGHT = json_root.pic_weight;
    solver.iterPerFrame = json_root.iter_per_frame;
    if (json_root.output_format == FS_FMT_PLY) {
        solver.outputFormat = PLY;
    } else {
        solver.outputFormat = OFF;
    }
    solver.outputTemplate = json_root.output_template;
    std::string inputLabelData = json_root.input_label_data;
    int frameAmount = json_root.frame_amount;
    std::cout << "alpha = " << solver.PIC_WEIGHT << std::endl;

    bool activeBody = json_root.active_body != 0;
    solver.activeBody = activeBody;
    if (activeBody) {
        float3 body_pos = make_float3(json_root.body_x, json_root.body_y, json_root.body_z);
        solver.initialBodyPos = body_pos;
        solver.bodyMass = json_root.body_mass;
    }
    try {
        solver.init(inputLabelData);
        solver.run(frameAmount);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
