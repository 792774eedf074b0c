/*
 * sph07_amd.h -- C ABI of the MI355X-native per-timestep particle loop of the
 * sph07 SPH solver.
 *
 * Every entry point replaces one member function of the reference's
 * MSolution / MSimulationInit surface (MSolution.h:25-65,
 * MSimulationInit.h) and keeps its meaning: the same phase order, the same
 * error convention (0 = OK, 1 = ERROR as in global.h:10-11, plus the specific
 * codes below), and state that is mutated in place -- here device-resident
 * per context instead of the MSimulationData singleton.
 *
 * Plain C types only: no HIP or torch types cross this boundary.  One context
 * is one scene bound to one HIP stream; calls on a context must be serialised
 * by the caller; several contexts may live side by side.
 */
#ifndef SPH07_AMD_H_
#define SPH07_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (0/1 as the reference; the rest are specific) ---- */
#define SPH_OK              0
#define SPH_ERROR           1   /* generic, message in sph_last_error */
#define SPH_ERR_NBR_OVERFLOW 2  /* "Neighbour list exceeded maximum." MSolution.cpp:201-206 */
#define SPH_ERR_OUT_OF_GRID 3   /* particle outside the grid (unchecked UB in the reference) */
#define SPH_ERR_UNSUPPORTED 4   /* deck option outside the hot path */
#define SPH_ERR_HIP         5   /* HIP runtime failure */
#define SPH_ERR_CAPACITY    6   /* input larger than an entry point's fixed capacity */

/* One material (material cards, MSimulationData.cpp:454-536, after
   MMaterial::CheckDefaults MMaterial.cpp:71-81). */
typedef struct sph_material_desc {
    int32_t model;            /* 9 = fluid */
    int32_t eos;              /* 28 Tait (Monaghan 1994) or 29 linear (Morris 1997) */
    int32_t visc_type;        /* 1 Morris viscosity; 2 adds Monaghan alpha (extension) */
    int32_t pad_;
    double  rho;              /* rho0 */
    double  av_l;             /* kinematic viscosity nu */
    double  av_q;
    double  surface_tension;  /* sigma (CSF) */
    double  mass;             /* total material mass (massopt 0) */
    double  h;                /* smoothing length */
    double  eosinput[4];
    double  strinput0;        /* pressure cut-off applied at initialisation */
    double  av_alpha;         /* Monaghan alpha (visc_type 2) */
} sph_material_desc;

/* One scene: the content of a .mcm deck (MSimulationData::GetInput,
   MSimulationData.cpp:186-219).  Arrays are 0-based, particle-major. */
typedef struct sph_scene_desc {
    int32_t ndim, disctype, tcrit_opt, maxnbr;          /* cards 1, 5, 6 */
    int32_t surface_tension, rigid_body, boundary, baseaccel;
    int32_t massopt, init_h_opt, init_rhoe, nummat;     /* card 4 */
    int32_t np, rb_id, veloc_opt, h_opt;
    int32_t boundary_code[2][3];                        /* boundary card: 0 free, 2 periodic */
    double  endtime, tssf, itss;                        /* card 2 */
    double  boundary_x[2][3];
    double  base_a[3];                                  /* base acceleration card */
    double  rb_total_mass, rb_inertia;                  /* rigid body card */
    const sph_material_desc *mat;  /* nummat entries; material n is mat[n-1] */
    const double  *x;              /* np*3 node coordinates */
    const double  *v;              /* np*3 initial velocities, or NULL */
    const int32_t *mat_id;         /* np, 1-based material number */
    const int32_t *color;          /* np, CSF colour (NULL: all 1) */
    const int32_t *dispbc;         /* np: 0 free, 1-7 as ConstrainQuantity, 8 moving wall */
    const double  *node_mass;      /* np, massopt 1 */
    const double  *node_h;         /* reserved (init_h_opt must be 0) */
    const double  *node_rho;       /* np, init_rhoe 1 */
} sph_scene_desc;

/* Scalar state (MGlobalVars.h:18-31, MRigidBody.h:18-30). */
typedef struct sph_globals {
    double  dt, dtold, ptime, critts;
    double  coord_minmax[2][3];
    double  hmax;
    double  rb_R[3], rb_V[3], rb_OMEGA[3], rb_A[3], rb_ALPHA[3];
    int32_t timestep, ngp, done, error;
    int32_t np, grid_lim[3];
} sph_globals;

/* Arithmetic arm of the pair sweeps. */
#define SPH_ARM_FP64        0   /* reference arithmetic, reference summation order */
#define SPH_ARM_FP32C       1   /* FP32 pair sums with TwoSum compensation; FP64 state,
                                   integration and rigid body.  Applies to sph_solution;
                                   the phase-by-phase entry points run the FP64 sweeps. */

/* Fields for sph_download (np entries; vector fields np*3). */
enum {
    SPH_F_X = 0, SPH_F_V, SPH_F_A, SPH_F_RHO, SPH_F_P, SPH_F_C, SPH_F_MASS, SPH_F_H,
    SPH_F_TRD, SPH_F_MINDIST, SPH_F_NORMAL, SPH_F_NNORMAL, SPH_F_N, SPH_F_KAPPA,
    SPH_F_CRITTS, SPH_F_RHOOLD, SPH_F_ACTIVE, SPH_F_VABS, SPH_F_NNBR,
    /* Classic discretisation (disctype 0): the artificial viscosity q (np),
       and the symmetric stress and rate of deformation as vector fields
       (np*3): diagonal xx, yy, zz or off-diagonal xy, xz, yz.  With disctype
       2 the stress is -p I, q and the rate of deformation are zero. */
    SPH_F_Q, SPH_F_SIGMA_DIAG, SPH_F_SIGMA_OFF, SPH_F_ROD_DIAG, SPH_F_ROD_OFF
};

/* --- lifetime --- */

/* MSimulationInit::Initialize (MSimulationInit.cpp:31-99): uploads the deck,
   computes mass, rho/p/c, h, the first neighbour pass, the critical step and
   the first dt on the device.  *out is NULL on error. */
int  sph_create(const sph_scene_desc *desc, int arm, void **out, char *err, int errlen);
/* MSolution::DestroyInstance (MSolution.h:18) */
void sph_destroy(void *ctx);
/* last error message of a context (PrintScreenLog text, MSolution.cpp:217-221) */
const char *sph_last_error(void *ctx);
int  sph_desc_size(void);
int  sph_material_desc_size(void);

/* --- MSolution phase surface (MSolution.h:25-65) --- */

/* MSolution::Neighbours -> SetupLinkedList (MSolution.cpp:96-148):
   grid (MNeighbourVars::InitGrid, MNeighbourVars.cpp:18-66) and binning. */
int  sph_neighbours(void *ctx);
/* MSolution::GhostSetup (MSolution.cpp:811-918): periodic images. */
int  sph_ghost_setup(void *ctx);
/* MSolution::LinkedListNeighbours (MSolution.cpp:150-215). */
int  sph_linked_list_neighbours(void *ctx);
/* MSolution::Strain + ClassicStrain (MSolution.cpp:327-413); with disctype 0
   the full rate of deformation. */
int  sph_strain(void *ctx);
/* MSolution::CalculateSurfaceNormals (MSolution.cpp:1429-1493). */
int  sph_calculate_surface_normals(void *ctx);
/* MSolution::Constitutive + RhoUpdate (MSolution.cpp:436-477): the EOS;
   with disctype 0 also F3DM9, AViscosity and StressUp of model 9
   (MConstitutive.cpp:10-68,123-133,210-222). */
int  sph_constitutive(void *ctx);
/* MSolution::CalcCriticalTimestep (MSolution.cpp:34-94); critts may be NULL. */
int  sph_calc_critical_timestep(void *ctx, double *critts);
/* The dt rule inside MSolution::Solution (MSolution.cpp:275-277). */
int  sph_time_advance(void *ctx);
/* MSolution::Momentum (MSolution.cpp:479-504): MorrisAcceleration incl.
   RigidBodyAcceleration (disctype 2) or ClassicAcceleration (disctype 0,
   :506-548; the FP64 arm, no rigid body). */
int  sph_momentum(void *ctx);
/* MSolution::BoundCond (MSolution.cpp:700-726). */
int  sph_boundcond(void *ctx);
/* MSolution::UpdateVelocity (MSolution.cpp:602-698) incl. RigidBodyVelocity. */
int  sph_update_velocity(void *ctx);
/* MSolution::MoveParticles (MSolution.cpp:769-809). */
int  sph_move_particles(void *ctx);
/* timestep++ and the stop test of MSolution::Solution (MSolution.cpp:317-319). */
int  sph_end_step(void *ctx);
/* MSolution::Solution (MSolution.cpp:238-324) without output: up to
   max_steps fused device-resident iterations; stops early when the stop test
   fires.  *steps_done receives the iterations executed. */
int  sph_solution(void *ctx, int max_steps, int *steps_done);

/* --- state access --- */
int  sph_get_globals(void *ctx, sph_globals *out);
/* Copies one field, in the deck's particle order, to host memory. */
int  sph_download(void *ctx, int field, double *host_out);
/* Neighbour lists of the last LinkedListNeighbours, in summation order.
   Canonical ids as the oracle: real j -> j; periodic image of j shifted by
   (sx,sy,sz) -> -(1 + j*27 + (sx+1)*9 + (sy+1)*3 + (sz+1)). */
int  sph_get_neighbours(void *ctx, int32_t *nnbr, int64_t *canon, int maxnbr);
/* Blocks until all work queued on the context's stream has finished. */
int  sph_synchronize(void *ctx);

/* --- MNeighbourVars surface (MNeighbourVars.h:13-28): the grid and the
   linked lists of the last MSolution::Neighbours / GhostSetup, as the
   reference holds them.  Ids are the reference's: real particles by 1-based
   deck number, periodic image g as -g, SPH_LINKEDLISTEND ends a list. --- */
#define SPH_LINKEDLISTEND (-1000000)   /* global.h:13 */
/* sph_gridmin, sph_gridsize, sph_gridlim after InitGrid (MNeighbourVars.cpp:18-56);
   any pointer may be NULL. */
int  sph_grid_info(void *ctx, double *gridmin, double *gridsize, int32_t *gridlim);
/* MNeighbourVars::GetGridValue (MNeighbourVars.cpp:70-73): the head of cell
   (i,j,k)'s list: its newest image, else its highest-numbered real particle,
   else SPH_LINKEDLISTEND.  SPH_ERR_OUT_OF_GRID outside the grid. */
int  sph_get_grid_value(void *ctx, int i, int j, int k, int32_t *value);
/* The whole sph_llgrid (ncell >= gridlim[0]*gridlim[1]*gridlim[2] entries,
   index gridlim[2]*(gridlim[1]*i+j)+k), the llpointer of every real particle
   (real_next[np], indexed by 0-based deck number; MSolution.cpp:135-144) and
   of every image (ghost_next[ngp]) with the image's source particle
   (ghost_par[ngp], MGhostParticleData.cpp:16,30-31).  Any output may be NULL
   (ghost_next and ghost_par together); ngp is the capacity of the image arrays. */
int  sph_get_linked_list(void *ctx, int32_t *llgrid, long long ncell, int32_t *real_next,
                         int32_t *ghost_next, int32_t *ghost_par, int ngp);
/* MNeighbourVars::GetBoxCoords (MNeighbourVars.cpp:83-89): cell coordinates
   of x[3] on the current grid, by truncation, unclamped. */
int  sph_get_box_coords(void *ctx, const double *x, int32_t *boxcrd);
/* MNeighbourVars::GetLoopLimits (MNeighbourVars.cpp:91-103): looplim[0][m]
   and looplim[1][m] (row-major 2x3) clamped to the grid, 0 beyond ndim. */
int  sph_get_loop_limits(void *ctx, int ndim, const int32_t *boxcrd, int32_t *looplim);

/* --- MKernel surface: MKernelBSpline::Kernel and ::GradW, the static forms
   (MKernelBSpline.cpp:98-187; MKernel.h:12-13), evaluated on the device for
   n pairs.  xi, xj: n*3 coordinates (host memory); outputs w[n], dwdx[n*3]
   (0 beyond ndim), dwdr[n]; any output may be NULL.  ndim 1-3, havg > 0. --- */
int  sph_kernel_eval(int ndim, double havg, int n, const double *xi, const double *xj,
                     double *w, double *dwdx, double *dwdr);

/* sph_kernel_eval without any device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The staging of coordinates and results lives in one static
   device array per GPU, sized for SPH_KERNEL_EVAL_NOALLOC_MAX pairs
   (11 doubles per pair, 5.8 MB) and held until the process ends; n above
   that returns SPH_ERR_CAPACITY.  Runs on the calling thread's default
   stream and returns when the results are in host memory.  Every call on one
   GPU shares the array: calls that overlap on one GPU (several host threads)
   race and must be serialised by the caller. */
#define SPH_KERNEL_EVAL_NOALLOC_MAX 65536
int  sph_kernel_eval_noalloc(int ndim, double havg, int n, const double *xi, const double *xj,
                             double *w, double *dwdx, double *dwdr);

/* --- batched scenes (no reference counterpart: the reference runs one scene
   per process, MSolution.h:16-22 singletons).  Many independent scenes of one
   ndim and arm are advanced together: every step is one sequence of kernel
   launches over all live scenes (per-scene parameter tables and block maps),
   captured in a HIP graph.  Each scene keeps its own dt, step counter and
   state, so its results do not depend on which other scenes share the batch. --- */
int  sph_batch_create(int ndim, int arm, void **batch);
void sph_batch_destroy(void *batch);
const char *sph_batch_last_error(void *batch);
/* Initialises a scene (as sph_create) and makes it live; step_limit > 0 stops
   it after that many steps.  *slot identifies it until sph_batch_remove. */
int  sph_batch_add(void *batch, const sph_scene_desc *desc, int step_limit, int *slot,
                   char *err, int errlen);
int  sph_batch_remove(void *batch, int slot);
/* Advances every live scene by up to nsteps steps; *live receives the number
   of scenes still running (scenes admitted since the last step included). */
int  sph_batch_step(void *batch, int nsteps, int *live);
/* sph_batch_step in two halves, so that the host admits and prepares scenes
   while the device steps the batch: sph_batch_launch queues nsteps steps of
   every live scene (admitted scenes join first) and returns without waiting;
   sph_batch_sync waits for them and updates the live set.  While steps are
   queued, sph_batch_add initialises the new scene on its own stream (it joins
   at the next launch) and sph_batch_remove defers the free to the next sync.
   A launch while steps are queued syncs first. */
int  sph_batch_launch(void *batch, int nsteps);
int  sph_batch_sync(void *batch, int *live);
int  sph_batch_status(void *batch, int slot, sph_globals *out);
/* The scene context of a slot, for sph_download / sph_get_neighbours. */
void *sph_batch_scene(void *batch, int slot);
int  sph_batch_live(void *batch);
/* Writes up to max slots that are resident but stopped (finished or failed);
   returns the count. */
int  sph_batch_finished(void *batch, int *slots, int max);
/* Device memory a scene of this deck will hold, in bytes (admission control). */
long long sph_scene_footprint(const sph_scene_desc *desc);
/* Launch shapes of a batch (timing only, never results): shape[0] = slots
   per wave of the batched cell kernel (0: the per-particle neighbour walk
   for every scene), shape[1] = cell storage order (0
   Morton, 1 row-major), shape[2] = 1 once fixed.  A batch of 65536 live
   particles or more times each candidate on one 8-step graph of its own run
   and keeps the cheapest per particle-step.  reps == 0 fixes the current
   shapes (no tuning); reps < 0 only reports.  shape holds 3 entries. */
int  sph_batch_autotune(void *batch, int reps, int32_t *shape);

/* --- .mcm deck reader (MSimulationData::GetInput, MSimulationData.cpp:186-1068) --- */
/* Reads a deck into an owned descriptor; *deck is freed with sph_deck_free. */
int  sph_deck_read(const char *path, void **deck, char *err, int errlen);
/* The descriptor to pass to sph_create (valid until sph_deck_free). */
const sph_scene_desc *sph_deck_desc(void *deck);
/* Title card (:194-196) and status interval of control card 3 (:274-275). */
const char *sph_deck_title(void *deck);
int  sph_deck_status_interval(void *deck);
/* Time between state plots and the state-output format of control card 3
   (MSimulationData.cpp:274-275; format 2 = EnSight CASE, MOutput.cpp:45-67). */
double sph_deck_state_time(void *deck);
int  sph_deck_state_opt(void *deck);
void sph_deck_free(void *deck);

/* --- state output (MOutput, MOutput.cpp:39-720; host side, cold path) --- */
/* Opens the EnSight CASE output of a scene under the base name `base` (the
   reference uses the deck path): files base.case, base.geo, base_time.txt and
   base_{sig,dis,vel,acc,eps,rho,egy,snd,mas,bnd,nbr,tmp}.ens.  The current
   positions become xzero of the displacements.  mat_id: np 1-based materials. */
int  sph_output_open(void *ctx, const char *base, int32_t ndim, int32_t nummat,
                     const int32_t *mat_id, void **out);
/* MOutput::StateOutput (:73-114): the geometry with the first state, the case
   file with every state's time, one time step of each variable file.  Fields
   the path does not carry are written as zero (efps, energy, temperature, the
   boundary flag) or as -p on the diagonal (stress: no deviatoric part). */
int  sph_output_state(void *out);
/* MSolution::PlotDroppingCylinder (MSolution.cpp:1638-1668): x(0) and rho
   of particles 1..np-1, "%13.5e%13.5e" per line. */
int  sph_output_cylinder_plot(void *out, const char *path);
/* Appends time, R, V, OMEGA of the rigid body to path (%13.5e each). */
int  sph_output_rigid(void *out, const char *path);
const char *sph_output_last_error(void *out);
void sph_output_close(void *out);
/* Globals and x, v, a, rho, p of every particle (deck order) as hexadecimal
   floats: a full-precision state dump. */
int  sph_write_state_dump(void *ctx, const char *path);

/* --- autotuner (no reference counterpart) ---
   Launch shapes that change timing but never results: shape[0] = slots per
   wave of the cell-centric neighbour search (0: the per-particle walk, the
   same lists in the same order), shape[1] / shape[2] = neighbour
   records in flight per lane in the continuity / momentum sweeps of the
   scene's arm (FP64 or FP32c), shape[3] = cell storage order (0 Morton,
   1 row-major).  shape holds 4 entries.
   reps > 0 times every alternative reps times on the scene's current state
   and keeps the fastest (scenes of 65536 particles or more are tuned this way
   on their first sph_solution call); reps == 0 keeps the current shapes and
   turns the automatic tuning off; reps < 0 only reports.  shape may be NULL. */
int  sph_autotune(void *ctx, int reps, int32_t *shape);
/* Sets the four launch shapes of sph_autotune explicitly (slots per wave
   4-4096 or 0 for the walk, batches 2/4/8, order 0/1) and turns the automatic
   tuning off. */
int  sph_set_launch_shape(void *ctx, const int32_t *shape);

/* --- measurement (no reference counterpart: the reference only prints total
   wall seconds, MSimulationData.cpp:110-122) --- */
/* Turns HIP-event timing of the kernel groups of sph_solution on or off and
   clears the totals. */
int  sph_set_timing(void *ctx, int enable);
/* slot < 0: returns the number of timed groups.  Otherwise fills the group's
   name, total milliseconds and launch count. */
int  sph_get_timing(void *ctx, int slot, char *name, int namelen, double *total_ms,
                    long long *count);

#ifdef __cplusplus
}
#endif
#endif
