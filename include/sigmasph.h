/*
 * sigmasph.h -- C ABI of libsigmasph.so, the MI355X (gfx950) particle time step.
 *
 * The reference keeps this path behind Python objects that hold Taichi fields and has
 * no FFI of its own.  Each entry point below replaces one group of its Taichi kernels;
 * the citation names the reference function whose behaviour it reproduces.  The
 * Python facade (sigmasph_amd/) binds these with ctypes and keeps the reference's
 * ParticleSystem / DSPHSolver names on top.
 *
 * Conventions
 *   - Every pointer in a state or workspace struct is a DEVICE pointer owned by the
 *     caller (torch tensors); the context owns only scratch it allocates at create time.
 *   - Per-particle vectors are [n][4] (the 4th lane is padding), scalars are [n].
 *   - "Real" fields are double in SPH_PREC_FP64 and float in SPH_PREC_MIXED; position,
 *     velocity, displacement and reference position are double in both modes.
 *   - Functions return 0 on success and a negative SPH_E* code otherwise;
 *     sph_last_error() gives the message.  Device faults surface at the next call
 *     that synchronises.
 *   - `stream` is a hipStream_t (NULL = default stream).  Nothing allocates or
 *     synchronises the host inside a step.
 */
#ifndef SIGMASPH_H
#define SIGMASPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---------------------------------------------------------------- */
#define SPH_OK          0
#define SPH_EINVAL     -1   /* bad argument (null pointer, size, flag combination) */
#define SPH_EHIP       -2   /* a HIP runtime call failed                            */
#define SPH_ENOMEM     -3   /* scratch allocation failed at create time             */
#define SPH_ECAPACITY  -4   /* an output buffer was too small (count still written) */

/* ---- precision ------------------------------------------------------------------- */
#define SPH_PREC_FP64   0   /* every field double                                   */
#define SPH_PREC_MIXED  1   /* pair sums float (compensated), x/v/u/x0 double       */

/* ---- geometry flags (sph_dims.flags) --------------------------------------------- */
#define SPH_TWO_D          (1 << 0)  /* single z = 0 layer: plane strain or axisymmetry */
#define SPH_AXISYMMETRIC   (1 << 1)  /* (x, r) meridional half-plane about the x axis   */
#define SPH_MIRROR_Y       (1 << 2)  /* image particles across y = 0                    */
#define SPH_MIRROR_Z       (1 << 3)  /* image particles across z = 0                    */
#define SPH_PERIODIC_X     (1 << 4)
#define SPH_PERIODIC_Y     (1 << 5)
#define SPH_PERIODIC_Z     (1 << 6)
#define SPH_PLANE_STRAIN   (1 << 7)

/* ---- device scalar slots (sph_state.scal, double[SPH_NSCAL]) ---------------------- */
#define SPH_S_DT            0   /* dt[None]: read before and rewritten inside a step    */
#define SPH_S_DIFF_WORK     1   /* f64 running totals of the energy budget              */
#define SPH_S_LEFTOVER      2
#define SPH_S_VISC_WORK     3
#define SPH_S_WALL_WORK     4
#define SPH_S_MAX_C2        5   /* last step's max pair sound speed squared             */
#define SPH_S_MAX_ACC       6   /* last step's max |a| over on-grid particles           */
#define SPH_S_MAX_V         7   /* max |v| after advection (global Mach number)         */
#define SPH_S_OFF_GRID_NUM  8   /* particles quarantined (non-finite position)         */
#define SPH_S_OFF_GRID_MASS 9
#define SPH_S_OUT_OF_BOX   10   /* particles binned in the hashed overflow table        */
#define SPH_NSCAL          32

/* ---- reductions (sph_reduce `which`) ---------------------------------------------- */
#define SPH_R_KINETIC        0  /* Sum 1/2 m |v|^2                          -> out[0]   */
#define SPH_R_POTENTIAL      1  /* -Sum m g.x                               -> out[0]   */
#define SPH_R_MASS           2  /* Sum m                                    -> out[0]   */
#define SPH_R_MOMENTUM       3  /* Sum m v                                  -> out[0..2]*/
#define SPH_R_GROSS_MOMENTUM 4  /* Sum m |v|                                -> out[0]   */
#define SPH_R_TAIT_ENERGY    5  /* Sum m e_tait(rho), clamped below rho0    -> out[0]   */
#define SPH_R_KINETIC_OFFGRID 6 /* kinetic energy of non-finite particles   -> out[0]   */

/* Cell lattice.  A particle at x sits in cell c = floor((x - origin) / cell) per axis.
 * Cells inside the dense box [box_lo, box_lo + box_n) are addressed directly; every
 * other finite cell goes to a hashed overflow table of 2^hash_bits buckets (the hash of
 * particle_system.py:1283-1293), so the lattice is unbounded like the reference's. */
typedef struct {
    int64_t n;              /* particle count                                        */
    int32_t precision;      /* SPH_PREC_*                                            */
    int32_t flags;          /* SPH_TWO_D | ...                                       */
    double  h;              /* support radius: W = 0 for r >= h (sph_base.py:57-64)  */
    double  cell[3];        /* cell edge per axis, >= h                              */
    double  origin[3];      /* lattice origin (0 for non-periodic axes)              */
    double  period[3];      /* box length on periodic axes (= cell * box_n)          */
    int32_t box_lo[3];      /* first dense cell per axis                             */
    int32_t box_n[3];       /* dense cells per axis                                  */
    int32_t hash_bits;      /* log2 of the overflow table size                       */
    int32_t pad_;
    double  dx0;            /* lattice spacing: sizes the cell tiles (0 = h / 2.6)    */
} sph_dims;

/* One of the two buffers every sorted (persistent) field is double-buffered in.  The
 * counting sort gathers buf[cur] into buf[1 - cur] and flips `cur` -- the registry of
 * _alloc_sorted (particle_system.py:913-927) with a gather instead of scatter+copy. */
typedef struct {
    int32_t* pid;           /* stable particle id (creation index)                   */
    int32_t* obj;           /* object id                                             */
    double*  x;             /* [n][4] position                                       */
    double*  x0;            /* [n][4] reference position                             */
    double*  u;             /* [n][4] displacement, x = x0 + u                       */
    double*  v;             /* [n][4] velocity                                       */
    void*    shift;         /* [n][4] Real: PST shift of the previous step           */
    void*    V;             /* [n] Real volume                                       */
    void*    V_lo;          /* [n] Real rounding residual of V (mixed mode)          */
    void*    m;             /* [n] Real mass                                         */
    void*    p;             /* [n] Real pressure                                     */
} sph_sorted;

typedef struct {
    sph_sorted buf[2];
    int32_t    cur;         /* index of the live buffer                              */
    int32_t    pad_;
    /* derived per step, never carried across a sort */
    void*   acc;            /* [n][4] Real acceleration                              */
    void*   lam;            /* [n]    Real lambda = min eig E, clamped to [0, 1]     */
    void*   L6;             /* [n][8] Real regularised inverse of E (00,11,22,01,02,12) */
    void*   lam_grad;       /* [n][4] Real <grad lambda>                             */
    void*   div;            /* [n]    Real -div v (the volumetric strain rate)       */
    void*   de_visc;        /* [n]    Real viscous specific power v.a_visc           */
    void*   V_new;          /* [n]    Real volume after this step's update           */
    void*   rho;            /* [n]    Real m / V at the sort (written by sph_sort)   */
    void*   rho_new;        /* [n]    Real m / V after the volume update             */
    double* scal;           /* [SPH_NSCAL] device scalars                            */
} sph_state;

/* Parameters of the delta-plus-SPH step (DSPH.py:50-540 constructor; params.py). */
typedef struct {
    double rho0, c0, gamma, stiffness;   /* Tait EOS p = S((rho/rho0)^gamma - 1)      */
    double viscosity;                    /* Morris kinematic viscosity                 */
    double delta_sph;                    /* delta-SPH density diffusion coefficient    */
    double cfl;
    double c_signal;                     /* CFL floor speed                            */
    double g[3];                         /* gravitation                                */
    double dx0;                          /* lattice spacing 2 * particleRadius         */
    double axi_r_min;                    /* floor of r in axisymmetric 1/r terms        */
    double pst_R, pst_max_shift, pst_ma_fixed, pst_lambda_lo, pst_lambda_hi;
    double pst_m_delta, l_eig_tol_lam;
    double restitution, padding;
    double domain_size[3];
    int32_t pst_n;
    int32_t allow_negative_pressure;
    int32_t track_visc_work;
    int32_t dt_pair;                     /* pair sound-speed estimate (1) or floor only  */
    int32_t pst_enabled;
    int32_t pst_colle;                   /* 0 = Sun, 1 = Colle                           */
    int32_t pst_ma_mode;                 /* 0 = local, 1 = global, 2 = fixed             */
    int32_t pst_surface_none;
    int32_t pst_clamp_disp, pst_clamp_vel;
    int32_t pst_ale_momentum;
    int32_t walls;                       /* Domain.boundary "reflect"                    */
    int32_t displacement_positions;
    int32_t compensated_volume;
} sph_fluid_params;

/* ---- context ---------------------------------------------------------------------- */
typedef struct sph_ctx sph_ctx;

/* Allocate every piece of scratch the step needs (keys, bucket counts and starts,
 * permutation, reduction partials) for `dims`.  Replaces the allocation half of
 * ParticleSystem.__init__ (particle_system.py:337-350, 360-466). */
int  sph_ctx_create(sph_ctx** out, const sph_dims* dims);
int  sph_ctx_destroy(sph_ctx* ctx);
const char* sph_last_error(const sph_ctx* ctx);
/* Library build string, usable without a GPU. */
const char* sph_version(void);

/* ---- neighbour search (A2-A6) ------------------------------------------------------ */
/* Bin every particle into its cell, count, scan, rank by (cell key, pid) and gather
 * every sorted field of buf[cur] into buf[1-cur]; flips st->cur.  Replaces
 * ParticleSystem.initialize_particle_system (particle_system.py:1342-1345:
 * update_grid_id 1300-1318, PrefixSumExecutor 1344, counting_sort 1320-1340).  Unlike
 * the reference the order within a cell is deterministic (ascending pid). */
int  sph_sort(sph_ctx* ctx, sph_state* st, void* stream);

/* Directed neighbour pairs of the live buffer as (pid_i, pid_j, mirror code) triples,
 * enumerated exactly as every pair pass visits them (for_all_neighbors,
 * particle_system.py:1348-1490).  `pairs` is device int32[cap][3]; `count` is a device
 * int64 that receives the total (which may exceed cap: SPH_ECAPACITY).  With pairs NULL
 * and cap 0 only the count is produced. */
int  sph_neighbor_pairs(sph_ctx* ctx, const sph_state* st, int32_t* pairs, int64_t cap,
                        int64_t* count, void* stream);

/* ---- delta-plus-SPH step (A9-A18) ---------------------------------------------------- */
/* The whole step: sort, then DSPHSolver.substep (DSPH.py:1854-1883), then the boundary
 * kernels of SPHBase.step (sph_base.py:622-655).  Reads dt from scal[SPH_S_DT] and
 * leaves the adaptive dt there, as solver.dt[None] behaves. */
int  sph_fluid_step(sph_ctx* ctx, sph_state* st, const sph_fluid_params* prm, void* stream);

/* The phases of sph_fluid_step, callable one at a time on an already sorted state. */
/* compute_lam first pass (DSPH.py:611-689) fused with compute_volume_evolution
 * (DSPH.py:1432-1495) and the EOS loop of compute_pressure_accelerations (1514-1522). */
int  sph_fluid_pass_volume(sph_ctx* ctx, sph_state* st, const sph_fluid_params* prm, void* stream);
/* compute_lam second pass (DSPH.py:690-702) fused with the pair loop of
 * compute_pressure_accelerations (1532-1640) and compute_adaptive_dt's reductions
 * (1365-1430), then the dt itself. */
int  sph_fluid_pass_force(sph_ctx* ctx, sph_state* st, const sph_fluid_params* prm, void* stream);
/* accumulate_step_work (DSPH.py:1714-1763) + advect (1652-1665) + compute_max_velocity (712-723). */
int  sph_fluid_advect(sph_ctx* ctx, sph_state* st, const sph_fluid_params* prm, void* stream);
/* apply_pst first pass (DSPH.py:943-1043). */
int  sph_fluid_pass_shift(sph_ctx* ctx, sph_state* st, const sph_fluid_params* prm, void* stream);
/* apply_pst second pass (DSPH.py:1045-1049) + the boundary kernels of SPHBase.step
 * (sph_base.py:622-655) + banking of the step's f64 work partials. */
int  sph_fluid_finish(sph_ctx* ctx, sph_state* st, const sph_fluid_params* prm, void* stream);

/* ---- diagnostics (A29) -------------------------------------------------------------- */
/* Deterministic f64 reduction over the live buffer into device double out[0..2]
 * (sph_base.py:425-620, DSPH.py:1670-1712).  `g` is the gravitation for POTENTIAL;
 * `prm` supplies the EOS for TAIT_ENERGY (may be NULL otherwise). */
int  sph_reduce(sph_ctx* ctx, const sph_state* st, int32_t which, const double* g,
                const sph_fluid_params* prm, double* out, void* stream);

/* ---- timing helper for the benchmark ------------------------------------------------ */
/* Average device duration in milliseconds of the pair kernels launched since the last
 * reset, measured with HIP events on the stream they run on (which = 0: volume pass,
 * 1: force pass, 2: shift pass, 3: sort).  Enabled by sph_timing(ctx, 1). */
int  sph_timing(sph_ctx* ctx, int32_t enable);
int  sph_timing_read(sph_ctx* ctx, int32_t which, double* avg_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif /* SIGMASPH_H */
