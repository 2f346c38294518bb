/*
 * fluidsim.h -- C ABI of the PIC/FLIP step (one handle = one solver on one device).
 *
 * Two libraries export exactly these symbols:
 *   fluidprjct_amd/libfluidsim_hip.so  -- the product: hand-written gfx950 HIP kernels
 *   oracle/_build/libfs_oracle.so      -- the CPU restatement used only as a checker
 *
 * The reference exposes this path as the C++ class FluidSolver3D
 * (src/FluidSim/FluidSolver3D.cuh:13-186) driven by main.cpp:9-60.  Each entry
 * point below names the reference member or routine it replaces.
 *
 * Conventions
 *   - Every function returns FS_OK (0) or a negative FS_ERR_* code; the message of
 *     the last failure on the calling thread is returned by fs_last_error().
 *     Nothing throws across the ABI (the reference throws std::runtime_error from
 *     init/loadSDF/ctor: FluidSolver3D.cu:84, utility.cuh:242,301, FluidSolver3D.cu:25).
 *   - Arrays passed in are copied; arrays passed out are caller-owned buffers.
 *   - Grid storage is i-fastest: cell (i,j,k) -> i + j*W + k*W*H; u is (W+1)xHxD,
 *     v is Wx(H+1)xD, w is WxHx(D+1) (utility.cuh:142-149, FluidSolver3D.cu:512,550,588).
 *   - A handle is not re-entrant; one stream per handle.
 */
#ifndef FLUIDSIM_H
#define FLUIDSIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FS_OK                  0
#define FS_ERR_ARG            -1
#define FS_ERR_IO             -2
#define FS_ERR_NO_FLUID       -3   /* pressureSolve's "No fluid cells found!" -1 (FluidSolver3D.cu:1875-1878) */
#define FS_ERR_NOT_CONVERGED  -4
#define FS_ERR_DEVICE         -5
#define FS_ERR_STATE          -6

/* cell labels (utility.cuh:65-68) */
#define FS_SOLID 0
#define FS_FLUID 1
#define FS_AIR   2
#define FS_BODY  3

/* particle output formats (utility.cuh:58-61) */
#define FS_FMT_PLY 0
#define FS_FMT_OFF 1
#define FS_FMT_PLY_BINARY 2   /* binary_little_endian PLY, same vertex order (no reference counterpart) */

/* pressure preconditioners (replace the cuDSS direct solve, FluidSolver3D.cu:1930-2008) */
#define FS_PRECOND_MG       0   /* multigrid V-cycle inside PCG (default)            */
#define FS_PRECOND_LEARNED  1   /* learned conv net inside flexible PCG, MG fallback  */
#define FS_PRECOND_JACOBI   2   /* diagonal; plain PCG                                */
#define FS_PRECOND_HYBRID   3   /* V-cycle + learned correction, flexible PCG, MG fallback */

/* grids readable / writable through fs_get_grid / fs_set_grid */
#define FS_GRID_U        0   /* float (W+1)*H*D  */
#define FS_GRID_V        1   /* float W*(H+1)*D  */
#define FS_GRID_W        2   /* float W*H*(D+1)  */
#define FS_GRID_USAVED   3   /* float, as U      */
#define FS_GRID_VSAVED   4
#define FS_GRID_WSAVED   5
#define FS_GRID_P        6   /* float W*H*D, pressure as applied (0 off fluid)       */
#define FS_GRID_LABELS   7   /* uint8 W*H*D, FS_SOLID/FLUID/AIR/BODY                 */
#define FS_GRID_RHS      8   /* float W*H*D, b of A p = b (0 off fluid)              */
#define FS_GRID_P64      9   /* double W*H*D, pressure solution before rounding      */
#define FS_GRID_CELLKEY 10   /* int32 Np, cell index of each particle's current position (id order) */
#define FS_GRID_SOLID   11   /* uint8 W*H*D, static SOLID mask (1 = SOLID)            */
#define FS_GRID_OPEN_U  12   /* float, as U: open fraction of each face in the pressure operator (read only;
                                0 where a SOLID/BODY cell or the domain boundary closes it, the body's
                                SDF fraction with FS_STEP_FACE_FRACTIONS, else 1; no reference counterpart) */
#define FS_GRID_OPEN_V  13   /* float, as V */
#define FS_GRID_OPEN_W  14   /* float, as W */

/* per-stage timer slots, in the reference's order (FluidSolver3D.cu:3030-3098) */
#define FS_STAGE_LABEL    0
#define FS_STAGE_P2G      1   /* P2G + saveVelocities + applyForces, fused */
#define FS_STAGE_PRESSURE 2   /* RHS + solve */
#define FS_STAGE_PROJECT  3
#define FS_STAGE_BODY     4
#define FS_STAGE_G2P      5   /* G2P + advection, fused */
#define FS_STAGE_SORT     6   /* cell sort (no reference counterpart) */
#define FS_NUM_STAGES     7

/* Scene and run parameters.  The first block mirrors the keys and defaults of
 * main.cpp:18-52; the second block holds the constants the reference hard-codes
 * (FluidSolver3D.cuh:72-80, FluidSolver3D.cu:3092) plus the solver settings that
 * replace cuDSS. */
typedef struct fs_config {
    float dt;                    /* "dt", default 0.02                               */
    float dx;                    /* "dx", default 0.1                                */
    float pic_weight;            /* "PIC_WEIGHT", default 0.01 (alpha = PIC share)   */
    int   iter_per_frame;        /* "iterPerFrame", default 1                        */
    int   output_format;         /* "outputFormat": FS_FMT_PLY unless "OFF"          */
    char  output_template[256];  /* "outputTemplate", default "InputData/particles_" */
    char  input_label_data[256]; /* "inputLabelData", default "InputData/labels.txt" */
    int   frame_amount;          /* "frameAmount", default 100                       */
    int   active_body;           /* "activeBody", default false                      */
    float body_x, body_y, body_z;/* "body_x/y/z", default 0.5                        */
    float body_mass;             /* "body_mass", default 50                          */
    char  body_sdf_path[256];    /* reference hard-codes "InputData/ball.sdf"        */

    int      particles_per_cell; /* 4 (FluidSolver3D.cuh:72)                         */
    uint64_t seed;               /* counter-based seeding key                        */
    float    gravity_y;          /* -9.81 (FluidSolver3D.cuh:78)                     */
    float    density;            /* 1000 (FluidSolver3D.cuh:80)                      */
    float    advect_c;           /* 0.1 (FluidSolver3D.cu:3092)                      */
    double   solver_tol;         /* relative residual ||b-Ap||/||b||, default 1e-6   */
    int      solver_max_iter;    /* default 500                                      */
    int      precond;            /* FS_PRECOND_*                                     */
    int      stall_window;       /* learned-FPCG fallback after this many non-decreasing iterations (5) */
    float    cfl;                /* > 0: dt = min(dt, cfl*dx/max|v_particle|) each step; 0: fixed dt
                                    (no reference counterpart: the reference's dt is fixed) */
    int      solver_flags;       /* FS_SOLVER_* bits */
    int      step_flags;         /* FS_STEP_* bits (0: the reference's behaviour) */
} fs_config;

/* solver_flags */
#define FS_SOLVER_NO_GRAPH 1     /* launch the inner PCG iterations eagerly instead of as a HIP graph */
#define FS_SOLVER_NO_TAIL  2     /* run every multigrid level as separate launches (no single-workgroup tail) */
#define FS_SOLVER_NO_FUSE  4     /* one launch per sweep on every level (no temporally blocked smoothing) */
#define FS_SOLVER_COLD_START 8   /* start every step's solve from p = 0 instead of the previous step's pressure */

/* step_flags (options with no counterpart in the reference's default path) */
#define FS_STEP_DETERMINISTIC 1  /* particles of a cell kept in original-id order: runs repeat bit for bit
                                    (one more pass over the particles per step) */
#define FS_STEP_BODY_WALLS    2  /* handleBodyWallCollisions (FluidSolver3D.cu:2848-2895, which the
                                    reference has switched off) after each body update */
#define FS_STEP_BODY_INERTIA  4  /* full inertia tensor of the body's SDF voxels and a quaternion
                                    orientation in place of the scalar sphere inertia and the Euler-angle
                                    rates of RigidBody::integrate (utility.cuh:408-467) */
#define FS_STEP_FACE_FRACTIONS 8 /* the body couples to the pressure solve through the open fractions of
                                    the faces between FLUID/AIR cells (from its SDF at the face corners):
                                    A's coefficient on such a face is its fraction th, the divergence
                                    takes th u + (1 - th) u_body, fully closed faces take u_body.  BODY
                                    cells (centre inside, MarkBodyCellsFunctor FluidSolver3D.cu:45-78)
                                    stay walls.  Off: the reference's cell-centre voxels only. */

typedef struct fs_body_state {
    float pos[3], vel[3], omega[3];
    float theta, psi, phi;
    float force[3], torque[3];   /* last step's totals before integration */
    float submerged_volume;      /* last step's buoyancy volume           */
    float rotation[9];
} fs_body_state;

typedef struct fs_solve_stats {
    int    iterations;           /* inner (preconditioned) iterations, last solve       */
    int    outer;                /* FP64 refinement passes, last solve                  */
    int    fallback;             /* 1 if the learned preconditioner fell back to MG    */
    int    converged;
    double rel_residual;         /* FP64 ||b - A p|| / ||b||                           */
    double solve_ms;             /* wall time of the last solve (device events)        */
    int    fluid_cells;
    int    singular;             /* 1 if no fluid cell touches AIR (pure Neumann)       */
} fs_solve_stats;

/* ---- configuration ------------------------------------------------------ */
/* main.cpp:18-52 defaults */
void        fs_config_default(fs_config* cfg);
/* main.cpp:11-16: parse a flat JSON object; unknown keys are ignored */
int         fs_config_load_json(fs_config* cfg, const char* path);
const char* fs_last_error(void);
/* "hip" or "oracle" */
const char* fs_backend(void);

/* ---- lifecycle (FluidSolver3D ctor/dtor, FluidSolver3D.cu:4-43) ------------ */
int  fs_create(const fs_config* cfg, void** out_solver);
void fs_destroy(void* solver);

/* ---- scene ---------------------------------------------------------------- */
/* FluidSolver3D::init label part (FluidSolver3D.cu:82-153): "W H D" then W*H*D
 * characters S/F/A, the first one being cell (W-1,H-1,D-1); unknown characters
 * leave SOLID.  A 2D file ("W H") is read as D = 1. */
int fs_load_labels_file(void* s, const char* path);
/* same scene from memory, i-fastest, values FS_SOLID/FS_FLUID/FS_AIR */
int fs_set_labels(void* s, const uint8_t* labels, int W, int H, int D);
/* level sets -> labels (DistrGen/fluidLabels.py:28-33): phi_solid<0 -> SOLID,
 * else phi_fluid<0 -> FLUID, else AIR.  i-fastest, W*H*D each. */
int fs_set_level_sets(void* s, const float* phi_solid, const float* phi_fluid, int W, int H, int D);
int fs_get_dims(void* s, int* W, int* H, int* D);

/* ---- rigid body (RigidBody::loadSDF utility.cuh:231-338 + init FluidSolver3D.cu:157-196) */
int fs_load_body_sdf(void* s, const char* path, float x, float y, float z, float mass);
/* values k->j->i, i-fastest, w*h*d */
int fs_set_body_sdf(void* s, const float* sdf, int w, int h, int d,
                    float origin_x, float origin_y, float origin_z, float cell_size,
                    float x, float y, float z, float mass);
int fs_get_body_state(void* s, fs_body_state* out);
int fs_set_body_state(void* s, const fs_body_state* in);

/* ---- particles -------------------------------------------------------------- */
/* FluidSolver3D::seedParticles (FluidSolver3D.cu:213-297) with a counter-based
 * generator keyed by (seed, cell, slot) in place of std::random_device. */
int     fs_seed_particles(void* s, int ppc, uint64_t seed);
/* pos, vel: n*3 floats (x,y,z interleaved, as Particle3D utility.cuh:76-80) */
int     fs_set_particles(void* s, const float* pos, const float* vel, int64_t n);
int     fs_get_particles(void* s, float* pos, float* vel, int64_t cap, int64_t* n);
int64_t fs_num_particles(void* s);

/* ---- stepping (FluidSolver3D::frameStep FluidSolver3D.cu:3030-3098) --------- */
int fs_step(void* s, int nsteps);
/* FluidSolver3D::run (FluidSolver3D.cu:3100-3163): frames*iterPerFrame+1 steps,
 * particle file (+ body PLY, CSV) every iterPerFrame steps. */
int fs_run(void* s, int frames);

/* per-stage entry points, in frameStep order */
int fs_stage_label(void* s);        /* labelGrid :340-397 (+ cell sort)                 */
int fs_stage_p2g(void* s);          /* particlesToGrid :607-685, saveVelocities :399-415, applyForces :700-761 */
int fs_stage_rhs(void* s);          /* constructRHS :1518-1596 (b on the full grid)     */
int fs_stage_pressure(void* s);     /* pressureSolve :1815-2062 minus RHS (solve only)  */
int fs_stage_project(void* s);      /* applyPressure :2340-2464                         */
int fs_stage_body(void* s);         /* updateBody :2793-2846                            */
int fs_stage_g2p_advect(void* s);   /* gridToParticles :983-1039 + advectParticles :1334-1357 */

/* ---- read-back / state injection --------------------------------------------- */
int fs_get_grid(void* s, int which, void* out, int64_t count);
int fs_set_grid(void* s, int which, const void* in, int64_t count);
int fs_get_stage_times(void* s, double* ms, int n);   /* accumulated ms per FS_STAGE_* */
/* simulated time, the dt of the last step and the number of steps taken */
int fs_get_clock(void* s, double* sim_time, float* last_dt, int64_t* steps);
int fs_get_solve_stats(void* s, fs_solve_stats* out);

/* ---- output (save3dParticlesToPLY / OFF, utility.cu:91-127) ------------------ */
int fs_write_particles(void* s, const char* path, int format);

/* ---- standalone pressure system (A p = b of constructA :1718-1777) ------------ */
/* Solve the reference-assembled system for given labels and b (full grid, b=0 off
 * fluid).  p: double W*H*D out. */
int fs_solve_pressure_system(void* s, const uint8_t* labels, const float* b, double* p,
                             int W, int H, int D, double tol, fs_solve_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* FLUIDSIM_H */
