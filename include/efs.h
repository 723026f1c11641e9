/*
 * efs.h — C ABI of the MI355X (gfx950) HIP backend for the Eulerian liquid step.
 *
 * The library replaces, for one GPU, the backend slot that the reference fills with
 * CPUSimulator3D / CudaSimulator3D behind `ISimulator` (reference:
 * include/simulation/simulator/ISimulator.h:8-53).  Every entry point takes plain pointers and
 * sizes; fields cross the boundary in the reference layout: `float` per cell for scalar fields
 * and `float[3]` (AoS, glm::vec3) per cell for velocity and forces, cell (x,y,z) at index
 * z*W*H + y*W + x (include/simulation/data/ScalField3D.h:19-25).  Device memory holds the fields
 * as structure-of-arrays; conversion happens inside the library.
 *
 * Return codes: 0 ok; 1 pressure solve did not converge (the caller sets PAUSED, as
 * CPUSimulator3D.cpp:541-547 does); 2 non-finite residual; negative = error (-1 invalid
 * argument, -2 out of memory, -3 library/extension unavailable, -(1000 + hipError_t) HIP error).
 * efs_last_error() returns a message for the last failing call on the calling thread.
 */
#ifndef EFS_H
#define EFS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct efs_sim* efs_handle;

/* Field ids: the eight accessors of IGrid3D (include/simulation/data/IGrid.h:52-59). */
enum efs_field {
    EFS_FIELD_VEL = 0,        /* float[3] per cell: face velocities u (x-face), v, w      */
    EFS_FIELD_RHO = 1,
    EFS_FIELD_GAMMA = 2,
    EFS_FIELD_BETA = 3,
    EFS_FIELD_DIVERGENCE = 4,
    EFS_FIELD_PRESSURE = 5,
    EFS_FIELD_FORCES = 6,     /* float[3] per cell                                        */
    EFS_FIELD_BOUNDARY = 7
};

/* Stages in ISimulator::update order (ISimulator.h:15-35); bit s of disable_mask skips stage s
 * like the DISABLE_* properties (include/simulation/Properties.h:186-192). */
enum efs_stage {
    EFS_STAGE_VEL_DECAY = 0,
    EFS_STAGE_VEL_EXTRAPOLATE = 1,
    EFS_STAGE_RHO_SHARPEN = 2,
    EFS_STAGE_RHO_ADVECT = 3,
    EFS_STAGE_VEL_ADVECT = 4,
    EFS_STAGE_ADD_FORCE = 5,
    EFS_STAGE_PROJECT = 6
};

/* Solver ids keep the reference enum values (Properties.h:475-481).  MULTIGRID, which the reference
 * lists (Properties.cpp:144) but no reference backend implements, runs CG preconditioned by one
 * geometric multigrid V-cycle per iteration (3D only; the same solve as SELF_CG with
 * precond = EFS_PRECOND_MG). */
enum efs_solver { EFS_SOLVER_SELF_CG = 2, EFS_SOLVER_JACOBI = 3, EFS_SOLVER_MULTIGRID = 4 };

/* CG stop criterion. */
enum efs_criterion {
    EFS_CRIT_REFERENCE = 0, /* r.r < solver_threshold (CGSolver.cpp:434-456)           */
    EFS_CRIT_RELATIVE = 1   /* ||r||_2 / ||b||_2 <= solver_threshold                   */
};

/* CG preconditioner (3D).  MG: one geometric multigrid V-cycle per iteration (red-black
 * Gauss-Seidel, Dirichlet-first coarse labels); LEARNED: the U-Net below. */
enum efs_precond { EFS_PRECOND_NONE = 0, EFS_PRECOND_LEARNED = 1, EFS_PRECOND_MG = 2 };

/* Initial CG guess on the fluid rows. */
enum efs_initial_guess {
    EFS_GUESS_DIVERGENCE = 0, /* x0 = b: the CPU backend (project clears p, CG sets p = b, CGSolver.cpp:429) */
    EFS_GUESS_PREVIOUS = 1    /* x0 = last step's pressure: the CUDA backend (CudaSolver.cu:427)           */
};

/* The step's parameter set (Properties.h:136-218).  gravity_force is GRAVITY_FORCE in
 * cells/s^2 = GRAVITY_FORCE_PHYSICAL / CELL_SIZE (src/Application.cpp:89). */
typedef struct {
    float dt;
    float gravity_force;
    float lambda_minus;
    float lambda_plus;
    float eta;
    float liquid_threshold;
    float rho_d;
    float vel_decay;
    int32_t gravity_enabled;
    int32_t solver_type;      /* efs_solver                                                */
    int32_t max_steps;        /* MAX_STEPS: CG iteration cap; Jacobi runs max_steps/2 passes */
    float solver_threshold;   /* SOLVER_THRESHOLD                                          */
    uint32_t disable_mask;    /* bit per efs_stage                                         */
    int32_t criterion;        /* efs_criterion                                             */
    int32_t precond;          /* efs_precond                                               */
    int32_t true_residual;    /* 1 = FP64 true residual ||b - A p|| / ||b|| after each solve */
    int32_t initial_guess;    /* efs_initial_guess                                         */
    int32_t re_calc_div;      /* RE_CALC_DIV: 2D only, divergence recomputed after the projection
                                 (CPUSimulator2D.cpp:503-521)                              */
} efs_params;

typedef struct {
    int32_t status;           /* return code of the step / solve                           */
    int32_t iterations;       /* CG iterations run (Jacobi: passes)                        */
    double rr;                /* final r.r of the recurrence, FP64 (the reference's returned err) */
    double rel_residual;      /* sqrt(rr) / ||b||                                          */
    double true_rel_residual; /* ||b - A p||_2 / ||b||_2 in FP64, or -1 if not computed     */
    double bnorm;             /* ||b||_2                                                   */
    int32_t fallback;         /* the preconditioned solve (learned or multigrid) stalled and
                                 fell back to plain CG                                       */
    int32_t warm_started;     /* learned warm start used                                   */
    float stage_ms[7];        /* device time per stage (0 when timing is off)              */
} efs_stats;

/* Defaults of Properties.h:136-218 with the step's own parameters. */
void efs_default_params(efs_params* p);

/* Reference: CPUSimulator3D(w,h,d) / CudaSimulator3D ctor + Grid3D ctor (Grid3D.cpp:10-48):
 * allocates the grid, gamma = 1, border cells solid with rho 0. */
int efs_create_3d(int32_t width, int32_t height, int32_t depth, const efs_params* params, efs_handle* out);
/* Reference: CPUSimulator2D(w,h) + Grid2D ctor (Grid2D.cpp:11-60), the 2D backend with
 * CPUSimulator2D's semantics (face-wise forces and pressure gradient, two-sided wall
 * enforcement, 7 diffusion iterations).  Fields cross the boundary as `float` per cell and
 * `float[2]` (glm::vec2) per cell for velocity and forces, cell (x,y) at y*W + x.  efs_dims
 * reports depth 1.  The learned preconditioner is 3D only. */
int efs_create_2d(int32_t width, int32_t height, const efs_params* params, efs_handle* out);
/* Reference: CudaSimulator3D::~CudaSimulator3D (CudaSimulator3D.cu:51-80). */
int efs_destroy(efs_handle h);
/* Reference: the PROPERTY_* reads each stage performs (Properties.h:16-32). */
int efs_set_params(efs_handle h, const efs_params* params);
/* Reference: ISimulator::availableSolvers (src/simulation/simulator/ISimulator.cpp:5-30, the list
 * each backend's ctor passes, e.g. CPUSimulator3D.cpp:18-21): writes up to `capacity` solver ids
 * for a 2D or 3D backend, first the default, and returns how many there are.  A solver_type outside
 * the list is rejected by efs_create_* and efs_set_params with -1; the reference's fallback to the
 * first entry is the caller's (the C++ adapter does it). */
int efs_available_solvers(int32_t dim, int32_t* solvers, int32_t capacity);
int efs_get_params(efs_handle h, efs_params* params);
int efs_dims(efs_handle h, int32_t dims[3]);

/* Reference: CudaGrid3D::upload / download (CudaGrid3D.cu:247-269), one field at a time, in the
 * reference layout; bytes must equal the field size (N*4, or N*12 for VEL/FORCES; N*8 in 2D). */
int efs_upload(efs_handle h, int32_t field, const void* host, size_t bytes);
int efs_download(efs_handle h, int32_t field, void* host, size_t bytes);
/* Device-to-device variants for callers that keep fields in HBM (same layout). */
int efs_upload_device(efs_handle h, int32_t field, const void* device_ptr, size_t bytes);
int efs_download_device(efs_handle h, int32_t field, void* device_ptr, size_t bytes);

/* Render readback.  Reference: the buffers MarchingCubes::updateMesh (MarchingCubes.cpp:54-105)
 * feeds its shader for the FluidNode: the density (N floats, rho) and, for the gradient colour
 * mode, the velocity repacked per cell as vec4 (x, y, z, 0) by a host loop over N cells
 * (:75-81).  Here the packing runs on the device, and either pointer may be NULL (vel4: N x 16
 * bytes; 2D: z = 0).  _device writes device buffers (e.g. a shared graphics buffer) without a
 * host copy. */
int efs_render_feed(efs_handle h, float* rho, float* vel4);
int efs_render_feed_device(efs_handle h, float* rho_device, float* vel4_device);

/* Reference: ISimulator::update (ISimulator.h:12-38): the seven stages with disable_mask. */
int efs_step(efs_handle h, efs_stats* stats);
/* One stage in isolation (the protected stage methods of ISimulator, ISimulator.h:43-50). */
int efs_stage(efs_handle h, int32_t stage, efs_stats* stats);
/* Reference: ISolver3D::solve (include/simulation/solver/ISolver.h:40), on the current
 * divergence; pressure is cleared first as project() does (CPUSimulator3D.cpp:513). */
int efs_solve_pressure(efs_handle h, efs_stats* stats);
/* The first half of project() (CPUSimulator3D.cpp:513-539): wall enforcement and the
 * divergence with density correction, without the solve. */
int efs_project_divergence(efs_handle h);
/* Reference: CGSolverCPU3D::impl_mat_mul (CGSolver.cpp:467-497) on device vectors of N floats. */
int efs_apply_operator(efs_handle h, const float* x_device, float* y_device);

/* Streams: the library launches on its own non-blocking stream unless the caller sets one
 * (a hipStream_t passed as void*). */
int efs_set_stream(efs_handle h, void* hip_stream);
void* efs_get_stream(efs_handle h);
int efs_synchronize(efs_handle h);
/* 1 = record device time per stage into efs_stats.stage_ms (adds event pairs per stage). */
int efs_set_timing(efs_handle h, int32_t on);
/* Per-kernel event timing of the dominant kernels: on = 1 records start and end events on each
 * launch of the tracked kernels; on = n > 1 records them in every n-th efs_step only (the first, then
 * every n-th after it; a timed launch is followed by a gap of several microseconds, so sampling
 * keeps the step rate of the untimed steps); 0 stops.  Every call clears what was recorded.
 * efs_kernel_time reads the accumulated count and milliseconds of kernel
 * 0 (CG direction + SpMV), 1 (CG update), 2 (Jacobi), 3 (multigrid fine up sweep), 4 (multigrid
 * fine down sweep with restriction), 5 (rho_advect backward trace), 6 (rho_advect forward trace),
 * 7 (project's wall enforcement + divergence) or 8 (project's gradient + wall enforcement). */
int efs_set_kernel_timing(efs_handle h, int32_t on);
int efs_kernel_time(efs_handle h, int32_t kernel, int64_t* launches, double* total_ms);

/* Seeded scene inputs (scenes are setup, not the step; they replace the editor rasterisation of
 * box spawns, Editor.cpp:176-182, 246-257).  Primitives are evaluated at interior cell centres
 * ((i - 0.5) h, h = 1/(n - 2)) in order; a solid primitive sets boundary 1 and rho 0, a fluid
 * one sets rho 1 on non-solid cells.  The border stays solid. */
typedef struct {
    int32_t kind;  /* 0 box [a, b), 1 sphere centre a radius b[0] */
    int32_t solid;
    double a[3];
    double b[3];
} efs_prim;
int efs_scene_rasterise(efs_handle h, const efs_prim* prims, int32_t nprims);
/* N(0, sigma) noise on faces that touch no solid, counter-based from seed (stream 1). */
int efs_scene_face_noise(efs_handle h, uint64_t seed, double sigma);
/* Reference: the editor's mesh voxelisation (assets/shaders/sdf.comp, dispatched by
 * Editor::initFluid / initBoundaries, src/core/Editor.cpp:90-258).  triangles: ntris x 9 floats
 * (v0, v1, v2) in world space; the domain box (SimulationDomainNode position = centre, scale =
 * size) maps to the W x H x D cells.  field (EFS_FIELD_RHO or EFS_FIELD_BOUNDARY) is overwritten:
 * value_inside where the cell centre is inside the mesh (|winding number| > pi, or on a
 * triangle), value_outside elsewhere and on the last layer of each axis. */
int efs_scene_voxelise(efs_handle h, const float* triangles, int32_t ntris, const float domain_center[3],
                       const float domain_size[3], int32_t field, float value_inside, float value_outside);
/* Reference: Editor::updateForceFields for one active ForceFieldNode (src/core/Editor.cpp:260-330,
 * assets/shaders/forcefield_cells.comp): forces = force on the cells inside the field's mesh
 * (the same inside test as efs_scene_voxelise; the node's force is direction * strength).  A
 * field covering more than max_cells cells is skipped, as MAX_FORCEFIELD_CELLS does
 * (max_cells <= 0: no limit); *ncells receives the count.  3D only.  Clear the forces first
 * (efs_upload of zeros) when the reference would (a field moved). */
int efs_scene_force_region(efs_handle h, const float* triangles, int32_t ntris, const float domain_center[3],
                           const float domain_size[3], const float force[3], int32_t max_cells, int32_t* ncells);

/* Reference: QueryNode::update's per-frame statistics (src/simulation/scene/nodes/QueryNode.cpp:
 * 45-84) over the cells [max(0, lo), min(n - 1, hi)) of each axis (the node's position and
 * position + scale, truncated), skipping rho < 0.5 when only_fluid: cell count, max / min / mean
 * cell coordinates, mean face velocity (u, v, w stored at the cell) and mean pressure (sums in
 * FP64).  With no cell, every value is 0. */
typedef struct {
    int32_t count;
    float max_pos[3];
    float min_pos[3];
    float avg_pos[3];
    float vel[3];
    float pressure;
} efs_query_result;
int efs_query_box(efs_handle h, const int32_t lo[3], const int32_t hi[3], int32_t only_fluid, efs_query_result* out);
/* No counterpart in the reference.  efs_query_box for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * statistics, bit-identical, with the block partials in the library's static device scratch (512
 * partials and their fold, 45 KB per GPU, held until the process ends) instead of a per-call
 * allocation.  Any box fits the scratch; a box of more than 2^31 - 1 cells is refused (-1).  Calls
 * that overlap on one GPU share the scratch and race: serialise them. */
int efs_query_box_noalloc(efs_handle h, const int32_t lo[3], const int32_t hi[3], int32_t only_fluid, efs_query_result* out);

/* ---- Learned preconditioner (no counterpart in the reference: it accelerates the solve that
 * CGSolverCPU3D::solve performs, CGSolver.cpp:416-465) ------------------------------------
 * A 4-level 3D U-Net (channel widths w0..w3, 3x3x3 convolutions, BF16 matrix cores with FP32
 * accumulation) on the half-resolution residual, used inside flexible CG with a warm start and
 * a fallback to plain CG.  Network tensors are channels-last FP32 on the device. */
typedef struct efs_unet_s* efs_unet;
int efs_unet_create(const int32_t widths[4], uint64_t seed, efs_unet* out);  /* widths NULL = 16,32,48,64 */
int efs_unet_destroy(efs_unet u);
int64_t efs_unet_num_params(efs_unet u);
int efs_unet_get_params(efs_unet u, float* host);
int efs_unet_set_params(efs_unet u, const float* host);
int efs_unet_get_grads(efs_unet u, float* host);
/* in: (D, H, W, 16), out: (D, H, W); D, H, W multiples of 8 */
int efs_unet_forward_device(efs_unet u, const float* in, int32_t D, int32_t H, int32_t W, float* out);
/* loss = mean over voxels of mask * (out - target)^2 (mask NULL = 1); gradients replace the
 * gradient buffer (efs_unet_get_grads) */
int efs_unet_loss_grad_device(efs_unet u, const float* in, const float* target, const float* mask, int32_t D, int32_t H,
                              int32_t W, double* loss);
int efs_unet_adam_step(efs_unet u, float lr);
int efs_unet_synchronize(efs_unet u);
void* efs_unet_stream(efs_unet u);

typedef struct {
    float out_scale;      /* network output unit: pressure = out_scale * sigma * prediction   */
    float jacobi_weight;  /* weight of the diagonal term - r / nCount                         */
    int32_t warm_start;   /* x0 = M(b) instead of b                                            */
    int32_t max_iters;    /* flexible iterations before falling back to plain CG (0 = max_steps) */
    int32_t every;        /* network applied every `every` flexible iterations (the others use the
                             diagonal term only); 0 = warm start only, then plain CG     */
    int32_t stall_window; /* fall back when r.r fails to drop below stall_ratio times its value */
    float stall_ratio;    /*   stall_window iterations earlier (0 window = never)              */
} efs_learned_opts;
void efs_default_learned_opts(efs_learned_opts* o);
/* Attaches a network (owned by the caller) as the preconditioner used when params.precond is
 * EFS_PRECOND_LEARNED.  When the grid's interior is a multiple of 16 per axis it also allocates
 * the flexible CG's vectors and pinned check slots and sizes the network's buffers for the coarse
 * grid, so the solves allocate nothing; a failed allocation returns its error code here.  The
 * network's buffers follow the last grid it was sized for: a network shared by simulators of
 * different sizes is resized (allocated again) whenever a solve switches it to another grid. */
int efs_set_learned_preconditioner(efs_handle h, efs_unet net, const efs_learned_opts* opts);
/* Half-resolution dims (interior / 2; the interior must be a multiple of 16 per axis). */
int efs_learned_coarse_dims(efs_handle h, int32_t dims[3]);
/* Training triple from the current divergence and the last CG solution: network input
 * (nz, ny, nx, 16), target (nz, ny, nx) in units of out_scale * sigma_b, mask = fluid fraction;
 * device FP32 buffers.  *sigma = RMS of the divergence. */
int efs_learned_coarse_pair(efs_handle h, float out_scale, float* in16, float* target, float* mask, double* sigma);

const char* efs_last_error(void);
const char* efs_version(void);

#ifdef __cplusplus
}
#endif

#endif /* EFS_H */
