// Sanitizer check of the CPU oracle (test infrastructure): built with
// -fsanitize=address,undefined together with the oracle's sources
// (make -C oracle sanitized), it drives every part of the fs_* ABI the tests
// use on small scenes and exits 0 when no sanitizer reported an error.
//   oracle_sanitized <inputdata dir> <scratch dir>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../include/fluidsim.h"

#define CHECK(x)                                                                            \
    do {                                                                                    \
        int rc_ = (x);                                                                      \
        if (rc_ != FS_OK) {                                                                 \
            std::fprintf(stderr, "%s:%d %s -> %d (%s)\n", __FILE__, __LINE__, #x, rc_,      \
                         fs_last_error());                                                  \
            return 1;                                                                       \
        }                                                                                   \
    } while (0)

static std::vector<float> sphere_sdf(int n, float cs, float r) {
    std::vector<float> sdf((size_t)n * n * n);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                float x = (i - 0.5f * n) * cs, y = (j - 0.5f * n) * cs, z = (k - 0.5f * n) * cs;
                sdf[(size_t)i + (size_t)j * n + (size_t)k * n * n] = std::sqrt(x * x + y * y + z * z) - r;
            }
    return sdf;
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string in = argv[1], out = argv[2];
    fs_config cfg;
    fs_config_default(&cfg);
    CHECK(fs_config_load_json(&cfg, (in + "/config.json").c_str()));
    // every label file, then the enclosed (singular) system
    const char* files[] = {"labels_simple.txt", "labels_simple3d.txt", "labels_3d_sphere.txt",
                           "labels_bathymetry.txt", "labels_colimn.txt"};
    for (const char* f : files) {
        void* s = nullptr;
        fs_config c;
        fs_config_default(&c);
        CHECK(fs_create(&c, &s));
        CHECK(fs_load_labels_file(s, (in + "/" + f).c_str()));
        int W, H, D;
        CHECK(fs_get_dims(s, &W, &H, &D));
        std::vector<uint8_t> lab((size_t)W * H * D);
        CHECK(fs_get_grid(s, FS_GRID_LABELS, lab.data(), (int64_t)lab.size()));
        std::vector<float> b(lab.size());
        for (size_t i = 0; i < b.size(); ++i) b[i] = lab[i] == FS_FLUID ? (float)((i * 7919) % 13) - 6.0f : 0.0f;
        std::vector<double> p(lab.size());
        fs_solve_stats st;
        int rc = fs_solve_pressure_system(s, lab.data(), b.data(), p.data(), W, H, D, 1e-10, &st);
        if (rc != FS_OK && rc != FS_ERR_NO_FLUID && rc != FS_ERR_NOT_CONVERGED) CHECK(rc);
        fs_destroy(s);
    }
    // a dam break with a falling body, stepped, written and run
    const int N = 20;
    cfg.dx = 1.0f / N;
    cfg.dt = 1.0f / 120;
    cfg.pic_weight = 0.05f;
    cfg.cfl = 1.0f;
    cfg.iter_per_frame = 1;
    std::snprintf(cfg.output_template, sizeof cfg.output_template, "%s/particles_", out.c_str());
    void* s = nullptr;
    CHECK(fs_create(&cfg, &s));
    std::vector<uint8_t> lab((size_t)N * N * N, FS_AIR);
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                size_t c = (size_t)i + (size_t)j * N + (size_t)k * N * N;
                if (i == 0 || j == 0 || k == 0 || i == N - 1 || j == N - 1 || k == N - 1) lab[c] = FS_SOLID;
                else if (i < N / 2 && j < N / 2) lab[c] = FS_FLUID;
            }
    lab[(size_t)15 + (size_t)15 * N + (size_t)15 * N * N] = FS_FLUID;   // an isolated-ish cell in the air
    CHECK(fs_set_labels(s, lab.data(), N, N, N));
    const int n = 12;
    std::vector<float> sdf = sphere_sdf(n, cfg.dx, 0.12f);
    CHECK(fs_set_body_sdf(s, sdf.data(), n, n, n, -0.5f * n * cfg.dx, -0.5f * n * cfg.dx, -0.5f * n * cfg.dx, cfg.dx,
                          0.6f, 0.45f, 0.5f, 1.5f));
    CHECK(fs_seed_particles(s, 4, 9));
    for (int step = 0; step < 6; ++step) {
        int rc = fs_step(s, 1);
        if (rc != FS_OK && rc != FS_ERR_NOT_CONVERGED) CHECK(rc);
    }
    CHECK(fs_stage_label(s));
    CHECK(fs_stage_p2g(s));
    CHECK(fs_stage_rhs(s));
    int rc = fs_stage_pressure(s);
    if (rc != FS_OK && rc != FS_ERR_NOT_CONVERGED) CHECK(rc);
    CHECK(fs_stage_project(s));
    CHECK(fs_stage_body(s));
    CHECK(fs_stage_g2p_advect(s));
    const int64_t np = fs_num_particles(s);
    std::vector<float> pos((size_t)np * 3), vel((size_t)np * 3);
    int64_t got = 0;
    CHECK(fs_get_particles(s, pos.data(), vel.data(), np, &got));
    std::vector<int> keys((size_t)np);
    CHECK(fs_get_grid(s, FS_GRID_CELLKEY, keys.data(), np));
    CHECK(fs_write_particles(s, (out + "/p.ply").c_str(), FS_FMT_PLY));
    CHECK(fs_write_particles(s, (out + "/p.off").c_str(), FS_FMT_OFF));
    CHECK(fs_write_particles(s, (out + "/p.bin.ply").c_str(), FS_FMT_PLY_BINARY));
    fs_body_state bs;
    CHECK(fs_get_body_state(s, &bs));
    CHECK(fs_set_body_state(s, &bs));
    rc = fs_run(s, 2);
    if (rc != FS_OK && rc != FS_ERR_NOT_CONVERGED) CHECK(rc);
    double ms[FS_NUM_STAGES];
    CHECK(fs_get_stage_times(s, ms, FS_NUM_STAGES));
    fs_destroy(s);
    // error paths
    void* e = nullptr;
    CHECK(fs_create(&cfg, &e));
    if (fs_load_labels_file(e, (in + "/no_such_file.txt").c_str()) != FS_ERR_IO) return 3;
    if (fs_load_body_sdf(e, (in + "/config.json").c_str(), 0.5f, 0.5f, 0.5f, 1.0f) == FS_OK) return 4;
    fs_destroy(e);
    std::printf("oracle sanitizer run: %lld particles, OK\n", (long long)np);
    return 0;
}
