/*
 * clip_amd.h — C-ABI of the MI355X-native coupled two-phase LBM step
 * (Allen–Cahn phase field + weighted-MRT Navier–Stokes, D2Q9 / D3Q19).
 *
 * Every entry point returns CLIP_OK (0) or a negative status; the message of
 * the last failure on the calling thread is returned by clip_last_error().
 * No C++ exception crosses this boundary.  A handle owns its device memory;
 * caller buffers stay caller-owned.  A handle is not re-entrant: calls on one
 * handle must be serialised by the caller.
 *
 * The reference exposes a C++ class surface instead of a C-ABI; each entry
 * point below names the reference member it replaces (paths relative to the
 * CLIP source tree).
 */
#ifndef CLIP_AMD_H
#define CLIP_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLIP_MAX_DIM 3
#define CLIP_NUM_SIDES 6
#define CLIP_MAX_GEOMETRIES 16
#define CLIP_MAX_INLET_MODES 8

/* ---- status codes ------------------------------------------------------ */
#define CLIP_OK 0
#define CLIP_ERR_ARG (-1)       /* invalid argument / config value          */
#define CLIP_ERR_PARSE (-2)     /* config.txt syntax or missing key/side    */
#define CLIP_ERR_HIP (-3)       /* HIP runtime failure                      */
#define CLIP_ERR_IO (-4)        /* file missing, truncated or mismatched    */
#define CLIP_ERR_STATE (-5)     /* call out of order (e.g. step before init)*/
#define CLIP_ERR_DIVERGED (-6)  /* NaN/Inf found by the health check        */
#define CLIP_ERR_NOMEM (-7)     /* device allocation failed                 */

/* ---- enums (values equal the reference's enums) ------------------------- */
/* InputData::CaseType, include/input/InputData.cuh:56-63 */
enum clip_case_type {
    CLIP_CASE_BUBBLE = 0, CLIP_CASE_DROP = 1, CLIP_CASE_RTI = 2,
    CLIP_CASE_JET = 3, CLIP_CASE_UNKNOWN = 4
};
/* Boundary::Objects, include/boundary/Boundary.cuh:66-76 */
enum clip_side {
    CLIP_X_MINUS = 0, CLIP_X_PLUS = 1, CLIP_Y_MINUS = 2,
    CLIP_Y_PLUS = 3, CLIP_Z_MINUS = 4, CLIP_Z_PLUS = 5
};
/* Boundary::Type, include/boundary/Boundary.cuh:81-92 */
enum clip_bc_type {
    CLIP_BC_WALL = 0, CLIP_BC_SLIP_WALL = 1, CLIP_BC_FREE_CONVECT = 2,
    CLIP_BC_PERIODIC = 3, CLIP_BC_NEUMANN = 4, CLIP_BC_VELOCITY = 5,
    CLIP_BC_DO_NOTHING = 6, CLIP_BC_UNKNOWN = 7
};
/* Geometry::Type, include/geometry/Geometry.cuh:54-62 */
enum clip_geometry_type {
    CLIP_GEOM_CIRCLE = 0, CLIP_GEOM_SPHERE = 1, CLIP_GEOM_SQUARE = 2,
    CLIP_GEOM_CUBE = 3, CLIP_GEOM_PERTURBATION = 4, CLIP_GEOM_UNKNOWN = 5
};

/* Storage / arithmetic policy of a solver handle. */
enum clip_precision {
    CLIP_PREC_F64 = 0,        /* FP64 populations, FP64 arithmetic (drop-in)     */
    CLIP_PREC_F32 = 1,        /* FP32 populations, FP32 arithmetic               */
    CLIP_PREC_MIXED_F16 = 2,  /* FP16 population deviations, FP32 arithmetic     */
    CLIP_PREC_MIXED_BF16 = 3  /* BF16 population deviations, FP32 arithmetic     */
};

/* Fields reachable through clip_get_field / clip_set_field.  Names follow the
 * reference's DataArray::deviceDataArray (include/data/DataArray.cuh:234-250). */
enum clip_field {
    CLIP_FIELD_PHI = 0,    /* dev_c      (1 dof)   */
    CLIP_FIELD_RHO = 1,    /* dev_rho    (1 dof)   */
    CLIP_FIELD_P = 2,      /* dev_p      (1 dof)   */
    CLIP_FIELD_U = 3,      /* dev_vel    (DIM dof) */
    CLIP_FIELD_MU = 4,     /* dev_mu     (1 dof)   */
    CLIP_FIELD_DPHI = 5,   /* dev_dc     (DIM dof) */
    CLIP_FIELD_NORMAL = 6, /* dev_normal (DIM dof) */
    CLIP_FIELD_F = 7,      /* dev_f      (Q dof)   */
    CLIP_FIELD_G = 8,      /* dev_g      (Q dof)   */
    CLIP_FIELD_F_POST = 9, /* dev_f_post (Q dof)   */
    CLIP_FIELD_G_POST = 10 /* dev_g_post (Q dof)   */
};

/* Host-buffer layouts. */
enum clip_layout {
    CLIP_LAYOUT_AOS = 0,  /* CLIP's ((i*ext_y+j)*ext_z+k)*ndof+dof, ghosts included
                             (include/domain/Domain.cuh:63-68)                       */
    CLIP_LAYOUT_SOA = 1   /* ((dof*ext_z+k)*ext_y+j)*ext_x+i, ghosts included       */
};

/* ---- configuration ------------------------------------------------------
 * Mirrors InputData::SimParams (include/input/InputData.cuh:69-116), the
 * Boundary::BCTypeMap (Boundary.cuh:109-113) and Geometry::GeometryDevice
 * (Geometry.cuh:81-90) in one POD struct.  Derived members are filled by
 * clip_config_derive() (InputData.cu:76-114). */
typedef struct clip_config {
    int dim;                       /* 2 or 3 (the reference's ENABLE_2D/3D)    */
    int case_type;                 /* enum clip_case_type                      */
    unsigned final_step, output_interval, report_interval;
    unsigned checkpoint_interval, checkpoint_copy;
    double t_final;
    double N[CLIP_MAX_DIM];        /* 2D: interior size; 3D: size incl. ghosts */
    double reference_length, reference_velocity;
    double Re, We, Pe, Ca, Bo, Mo;
    double rho_ratio, mu_ratio;
    double gravity[CLIP_MAX_DIM];  /* body-force acceleration                  */
    double interface_width, mobility, amplitude;
    /* optional overrides (key present => flag set) */
    int has_sigma, has_mu_h, has_gravity_scale;
    double sigma_override, mu_h_override, gravity_scale;
    /* derived */
    double rho_h, rho_l, sigma, mu_h, mu_l, tau_h, tau_l, beta, kappa;
    /* boundary blocks */
    int bc_type[CLIP_NUM_SIDES];
    int bc_present[CLIP_NUM_SIDES];
    double bc_value[CLIP_NUM_SIDES][CLIP_MAX_DIM];
    /* geometry blocks */
    int num_geometries;
    int geom_type[CLIP_MAX_GEOMETRIES];
    int geom_id[CLIP_MAX_GEOMETRIES];
    double geom_center[CLIP_MAX_GEOMETRIES][CLIP_MAX_DIM];
    double geom_length[CLIP_MAX_GEOMETRIES][CLIP_MAX_DIM];
    double geom_radius[CLIP_MAX_GEOMETRIES];
    double geom_amplitude[CLIP_MAX_GEOMETRIES];
    /* inlet modulation (extension, no reference counterpart):
       u_in = value * (1 + sum_m amp[m] cos((m+1)*theta + phase[m])),
       theta measured around geometry 0's centre in the inlet plane. */
    int inlet_modes;
    double inlet_mode_amp[CLIP_MAX_INLET_MODES];
    double inlet_mode_phase[CLIP_MAX_INLET_MODES];
    /* optional mobility override (no reference counterpart): the RTI case
       derives the mobility from Pe (InputData.cu:76-114); with the flag set
       clip_config_derive keeps mobility_override instead, for every case */
    int has_mobility_override;
    double mobility_override;
    /* slip sides (no reference counterpart in effect): 0 = the reference's
       behaviour, a slip side takes the wall bounce-back when any side is a
       wall (kernelHalfBounceBack with wallBCMap, Solver.cu:248-333) and is
       otherwise left as streamed, since Solver::slipWallBoundary
       (Solver.cu:1012-1016) is never launched; 1 = slip sides reflect
       specularly through WMRT::slipWallBCMap (WMRT.cuh:243-264) in the wall
       pass.  Config key `slipWall = specular`. */
    int slip_specular;
} clip_config;

/* Sizes of a created solver. */
typedef struct clip_info {
    int dim, q;
    int precision;
    unsigned ext[CLIP_MAX_DIM];       /* extents incl. ghosts (Domain.cu:22-82) */
    unsigned long long nodes;         /* ext_x*ext_y*ext_z                      */
    unsigned long long interior;      /* interior (updated) nodes               */
    unsigned long long device_bytes;  /* device memory owned by the handle      */
    unsigned long long step;          /* TimeInfo::simInfo::currentStep         */
} clip_info;

typedef struct clip_solver clip_solver;

/* ---- configuration (host only) ----------------------------------------- */
/* Parse CLIP's config.txt: replaces InputData::read_config (InputData.cu:57-115),
 * Boundary::readBoundaries (Boundary.cu:146-279) and Geometry::readGeometries
 * (Geometry.cu:72-182).  dim_override = 2 or 3 forces the dimension (the
 * reference's compile-time switch); 0 infers 2 when N[2] <= 1. */
int clip_config_from_file(const char* path, int dim_override, clip_config* out);
/* Same, from a string holding the file's text. */
int clip_config_from_string(const char* text, int dim_override, clip_config* out);
/* Recompute derived parameters (InputData.cu:76-114) and validate. */
int clip_config_derive(clip_config* cfg);

/* Geometry::sdf (include/geometry/Geometry.cuh:102-161): signed distance
 * (positive outside) from (x, y, z) to the first geometry object of `cfg`
 * carrying `id`; 1e10 when none does.  Host only. */
int clip_geometry_sdf(const clip_config* cfg, int id, double x, double y, double z, double* out);

/* Measurement helper (no counterpart in the reference): GB/s (read + write)
 * of a 16-byte-per-lane streaming copy of `bytes` on the current device,
 * timed over `reps` launches: the practical HBM ceiling that bench.py
 * compares the step kernels with.  Allocates and frees 2 x bytes. */
int clip_copy_rate(unsigned long long bytes, int reps, double* gbs);

/* ---- Solver launchers (include/solver/Solver.cuh:61-107) ----------------
 * The boundary launchers of the reference's Solver class, on the caller's own
 * arrays: device pointers to FP64 arrays in CLIP's AoS layout with the ghost
 * ring (((i*ext_y+j)*ext_z+k)*ndof+dof, Domain.cuh:63-68) for the handle's
 * configuration, which sets the extents, side types, inlet and geometry.
 * The solver's own step does not use them (it resolves every boundary in its
 * pull); they serve code written against CLIP's Solver that applies a
 * boundary to arrays it owns.  Each runs only when the reference's launcher
 * would (its Boundary flag, Solver.cu:985-1062) and is a no-op otherwise.
 * dev_b / dev_b_post (or dev_b_prev) may be NULL.  Work is enqueued on the
 * handle's stream and finished on return.  Host pointers are CLIP_ERR_ARG.
 * Where the reference's kernel reads what the same launch writes (periodic
 * edges and corners; inner neighbours of free-convect / Neumann nodes that
 * are themselves on such a side) these read the values from before the
 * launch. */
/* Solver::periodicBoundary<Q> (Solver.cu:75-141, 985-990): ghost layers of
 * periodic axes take the opposite interior layer, X then Y then Z; ndof =
 * values per node (Q for populations, 1 for phi). */
int clip_periodic_boundary(clip_solver* s, int ndof, double* dev_a, double* dev_b);
/* Solver::wallBoundary<Q, dof> (Solver.cu:248-333, 998-1003): half-way
 * bounce-back with WMRT::wallBCMap at nodes next to wall and slip-wall sides,
 * a_post[in] = a[opposite], when some side is a wall. */
int clip_wall_boundary(clip_solver* s, const double* dev_a, double* dev_a_post, const double* dev_b,
                       double* dev_b_post);
/* Solver::slipWallBoundary<Q, dof> (Solver.cu:1012-1016): the same kernel with
 * WMRT::slipWallBCMap (WMRT.cuh:243-264), when some side is a slip wall. */
int clip_slip_wall_boundary(clip_solver* s, const double* dev_a, double* dev_a_post, const double* dev_b,
                            double* dev_b_post);
/* Solver::freeConvectBoundary<Q, dof> (Solver.cu:342-453, 1025-1029):
 * a[x, in] = (a_prev[x, in] + |u_n| a[x + n, in]) / (1 + |u_n|) at nodes next
 * to free-convect sides; dev_vel holds DIM components per node. */
int clip_free_convect_boundary(clip_solver* s, const double* dev_vel, double* dev_a, const double* dev_a_prev,
                               double* dev_b, const double* dev_b_prev);
/* Solver::NeumannBoundary<Q, dof> (Solver.cu:462-542, 1038-1042):
 * a[x, in] = a[x + n, in] at nodes next to Neumann sides. */
int clip_neumann_boundary(clip_solver* s, double* dev_a, double* dev_b);
/* Solver::velocityBoundary (JetBoundary, Solver.cu:550-921, 1058-1062): inlet
 * populations of f and g on the ghost planes of velocity sides from phi
 * (dev_c) and the ghost populations, with each side's own inlet velocity. */
int clip_velocity_boundary(clip_solver* s, double* dev_c, double* dev_f, double* dev_g);
/* Solver::mirrorBoundary (Solver.cu:927-978, 1048-1052): ghost cells of wall,
 * free-convect, Neumann, slip and do-nothing sides copy the boundary cell,
 * when some side is a wall, free-convect, slip wall or Neumann side. */
int clip_mirror_boundary(clip_solver* s, double* dev_a);

/* ---- solver lifecycle --------------------------------------------------- */
/* Replaces the object graph built in main.cu:61-65 (Domain, DataArray,
 * NSAllen): allocates every device buffer for the configuration. */
int clip_create(const clip_config* cfg, int precision, clip_solver** out);
/* Bind the handle to a HIP stream (hipStream_t passed as void*; NULL = the
 * handle's own stream). */
int clip_set_stream(clip_solver* s, void* hip_stream);
/* NSAllen::initialCondition + DataArray::updateDevice + NSAllen::deviceInitializer
 * (NsAllen.cu:768-824, DataArray.cu:124-128, NsAllen.cu:897-908). */
int clip_init_fresh(clip_solver* s);
/* NSAllen::deviceInitializer (NsAllen.cu:897-908, KernelInitializeDistributions
 * 137-202) from a phase field the caller gives on every node, ghost ring
 * included, in CLIP's AoS layout: what the reference computes from dev_c after
 * initialCondition() and updateDevice() filled it.  clip_init_fresh equals
 * this call with clip_initial_phi's field. */
int clip_init_from_phi(clip_solver* s, const double* phi_host);
/* NSAllen::initialCondition (NsAllen.cu:768-824): the initial phase field of a
 * configuration on every node, ghosts included (CLIP AoS layout, host). */
int clip_initial_phi(const clip_config* cfg, double* phi_host);
/* nsteps x NSAllen::solve (NsAllen.cu:881-892), TimeInfo::increment each. */
int clip_step(clip_solver* s, int nsteps);
/* Block until the handle's stream has drained. */
int clip_synchronize(clip_solver* s);
/* Copy a field to host memory (double, layout as requested, ghosts included;
 * host_out holds ndof*nodes doubles).  Replaces DataArray::updateHost
 * (DataArray.cu:131-138). */
int clip_get_field(clip_solver* s, int field, int layout, double* host_out);
/* Overwrite one of the reference's device arrays from host memory (same shape
 * as clip_get_field): PHI (dev_c, ghost ring included), U (dev_vel), P
 * (dev_p), F / G (dev_f / dev_g), F_POST / G_POST (dev_f_post / dev_g_post).
 * The arrays keep their roles in the reference's solve(): the next collision
 * reads f_post, g_post, p, u and phi (mu, grad phi and the normal are
 * recomputed from the phi given) and overwrites the interior of dev_f /
 * dev_g, whose ghost cells later pulls read.  Nothing is recomputed from the
 * populations: call clip_macroscopic for that.  This is how
 * CheckPointer::load (src/io/CheckPointer.cu:84-102) fills dev_f, dev_g,
 * dev_f_post and dev_g_post before NSAllen::macroscopic (main.cu:67-71).
 * RHO, MU, DPHI and NORMAL are derived and rejected (CLIP_ERR_ARG) before
 * host_in is read. */
int clip_set_field(clip_solver* s, int field, int layout, const double* host_in);
/* NSAllen::macroscopic (NsAllen.cu:866-877, include/equation/NsAllen.cuh:89):
 * phi = sum g_post at interior nodes (ghost ring left as it is), mu and
 * grad phi from it, p = sum f_post, u from f_post with u_old = the current u
 * (kernelMacroscopicf, NsAllen.cu:369-511); then the pending collision.
 * clip_checkpoint_load already leaves the handle as CheckPointer::load +
 * macroscopic() do (main.cu:67-71). */
int clip_macroscopic(clip_solver* s);
/* FP64 sum of a scalar field over interior nodes: Reporter::sum
 * (Reporter.cu:95-114).  field = PHI or RHO. */
int clip_sum(clip_solver* s, int field, double* out);
/* Allocation-free variant of clip_sum.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  FP64 sum
 * (same kernels and order as clip_sum) of a device-resident scalar field over
 * the interior nodes of a grid with CLIP's one-node ghost ring:
 * `field` holds ext[0]*ext[1]*ext[2] values, x fastest (double for
 * CLIP_PREC_F64, float otherwise; a handle's PHI in CLIP_LAYOUT_SOA order),
 * dim 2 (ext[2] = 1) or 3, each used extent in [3, 2^20] (CLIP_ERR_ARG
 * otherwise).  Writes the sum to `out` and, if non-NULL, the number of
 * non-finite values to `nonfinite`, both device pointers, asynchronously on
 * `hip_stream` (NULL = the null stream).  It allocates nothing: the reduction
 * partials live in a fixed 8 KiB device array that each GPU keeps until the
 * process ends and that every call on that GPU shares, so calls that overlap
 * on one GPU race. */
int clip_sum_noalloc(const void* field, int precision, int dim, const unsigned ext[3], double* out,
                     unsigned* nonfinite, void* hip_stream);
/* Health check of Reporter::print (Reporter.cu:124-144): sums PHI and returns
 * CLIP_ERR_DIVERGED (instead of throwing) when phi holds a NaN or Inf. */
int clip_check_health(clip_solver* s, double* phi_sum);
/* CheckPointer::save / load (CheckPointer.cu:53-102): CLIP's file set in
 * `dir` (dev_f.bin, dev_g.bin, dev_f_post.bin, dev_g_post.bin,
 * [dev_f_prev.bin, dev_g_prev.bin], timeInfo.bin, domainSize.bin, info.txt)
 * plus clip_amd_state.bin (phi of the last two steps with their ghost rings
 * and u_old; for 16-bit storage also the stored bytes of the pulled f and g;
 * the outflow planes go to dev_f_prev / dev_g_prev) so that a resume is
 * bit-exact.  Loading without the extension file reproduces the reference's
 * macroscopic() restart with u_old = 0. */
int clip_checkpoint_save(clip_solver* s, const char* dir);
int clip_checkpoint_load(clip_solver* s, const char* dir);
/* CheckPointer::rotateFolders (CheckPointer.cu:108-138). */
int clip_checkpoint_rotate(const char* base, unsigned copies);
int clip_get_info(const clip_solver* s, clip_info* out);
/* Time of the device work of the last clip_step call, measured with HIP events
 * around each kernel class on the handle's stream: total, and the three step
 * kernels (ghost refresh, phase sum, fused collide-stream). */
int clip_last_step_timing(const clip_solver* s, double* total_ms, double* ghost_ms,
                          double* phase_ms, double* fused_ms);
int clip_enable_timing(clip_solver* s, int on);
/* Step through a HIP graph (no counterpart in the reference): clip_step then
 * replays a captured graph of two steps (the ping-pong period of the buffers)
 * per pair, cutting the host launch cost of small grids where a step is a
 * few microseconds per kernel.  Results are identical to plain stepping.
 * Ignored while per-kernel timing is on. */
int clip_enable_graph(clip_solver* s, int on);
/* VTSwriter::writeVTSBinaryFile (src/io/VTSwriter.cu:84-205): write the
 * current Velocity, C (phi), P and Rho of the interior nodes to one ASCII VTK
 * StructuredGrid file at `path`.  Fields leave the device only on this call. */
int clip_write_vts(clip_solver* s, const char* path);

/* ---- reverse mode (adjoint) ----------------------------------------------
 * No counterpart in the reference: the differentiable-solver extension for
 * gradient-based callers (parameter fitting, inverse problems, inflow
 * control).  CLIP_PREC_F64 or CLIP_PREC_F32 handles; periodic, wall, velocity
 * inlet, free-convect, Neumann and do-nothing sides (a slip side only
 * together with a wall side, where it acts as one).
 * clip_adjoint_forward builds the initial state of the case
 * (KernelInitializeDistributions, NsAllen.cu:137-202, with the first
 * collision) from phi0 on the interior nodes (a ghost cell copies phi0 at its
 * wrapped position when every axis it lies outside of is periodic, else it
 * keeps the case's initial phase field), runs nsteps steps keeping the state
 * (populations, phi with its ghost ring, u, outflow planes) every
 * checkpoint_every steps, and writes phi after the last step (interior) to
 * phi_out.
 * clip_adjoint_backward takes dL/dphi_out and returns dL/dphi0 and, if the
 * pointers are non-NULL, dL/dsigma (through beta, kappa and the initial p)
 * and dL/dmobility; it recomputes each segment between kept states once
 * (twice with an inner interval, clip_adjoint_set_nesting).
 * The interior arrays are host FP64 arrays of the interior nodes in the
 * order `layout` names: CLIP_INTERIOR_X_FASTEST (element i + nx (j + ny k),
 * i, j, k counted from the first interior node) or CLIP_INTERIOR_X_SLOWEST
 * (element (i ny + j) nz + k, a C array [nx][ny][nz], or [nx][ny] in 2D).
 * Each crosses the bus once; the scatter to and gather from the device's
 * node layout run on the device. */
#define CLIP_INTERIOR_X_FASTEST 0
#define CLIP_INTERIOR_X_SLOWEST 1
int clip_adjoint_forward(clip_solver* s, const double* phi0, int nsteps, int checkpoint_every, double* phi_out,
                         int layout);
int clip_adjoint_backward(clip_solver* s, const double* grad_phi_out, double* grad_phi0, double* grad_sigma,
                          double* grad_mobility, int layout);
/* Regularisation of the interface normal's derivative in the reverse sweep
 * (no counterpart in the reference).  0 (default): the exact derivative of
 * the step.  c > 0: where |grad phi| < c 4 phi (1 - phi) / W the derivative
 * through n = grad phi / |grad phi| is scaled by |grad phi| / (c 4 phi (1 - phi) / W),
 * which bounds it on flat plateaus (drop centres, far field) where the exact
 * derivative ~ 1 / |grad phi| amplifies rounding without bound over long
 * horizons; elsewhere the gradient is unchanged.  c = 0.1 is a good choice. */
int clip_adjoint_set_normal_floor(clip_solver* s, double normal_floor);
/* Second checkpoint level of the reverse sweep: 0 < inner_every <
 * checkpoint_every keeps, while a segment is recomputed, a state every
 * inner_every steps and recomputes each of those sub-segments once more.
 * Memory: nsteps / checkpoint_every + checkpoint_every / inner_every +
 * inner_every states instead of nsteps / checkpoint_every + checkpoint_every.
 * 0 (default): one level.  Read by the next clip_adjoint_forward. */
int clip_adjoint_set_nesting(clip_solver* s, int inner_every);
/* dL/dA_m of the inlet modulation u_in = value (1 + sum_m A_m cos((m+1) theta
 * + phase_m)) (clip_config inlet_mode_amp) from the last clip_adjoint_backward,
 * m < n (entries past inlet_modes are 0).  The amplitudes enter through the
 * inlet ghost populations (JetBoundary, Solver.cu:578-619).  CLIP_ERR_STATE
 * before a backward sweep. */
int clip_adjoint_inlet_gradient(clip_solver* s, double* grad, int n);
void clip_destroy(clip_solver* s);
const char* clip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CLIP_AMD_H */
