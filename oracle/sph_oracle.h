/*
 * sph_oracle.h -- CPU FP64 restatement of sph07's per-timestep particle loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product never links, imports or calls it.
 *
 * The deck layout below mirrors the .mcm cards the reference reads
 * (MSimulationData.cpp:244-323, 441-541, 655-1068).  All indices in the
 * arrays passed in are 0-based; internally the oracle keeps the reference's
 * 1-based particle numbering so that the linked-list order matches.
 */
#ifndef SPH_ORACLE_H_
#define SPH_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_material {
    int32_t model;            /* 9 = fluid (only model on the path) */
    int32_t eos;              /* 28 (Monaghan/Tait) or 29 (Morris linear) */
    int32_t visc_type;        /* 1: Morris viscosity; 2: + Monaghan alpha term */
    int32_t pad_;
    double  rho;              /* rho0 */
    double  av_l;             /* kinematic viscosity nu (after CheckDefaults) */
    double  av_q;
    double  surface_tension;  /* sigma */
    double  mass;             /* total material mass (massopt 0) */
    double  h;                /* smoothing length */
    double  eosinput[4];
    double  strinput0;        /* pressure cut-off used at init */
    double  av_alpha;         /* Monaghan alpha (visc_type 2) */
} orc_material;

typedef struct orc_deck {
    int32_t ndim, disctype, tcrit_opt, maxnbr;
    int32_t surface_tension, rigid_body, boundary, baseaccel;
    int32_t massopt, init_h_opt, init_rhoe, nummat;
    int32_t np, rb_id, veloc_opt, h_opt;
    int32_t boundary_code[2][3];
    double  endtime, tssf, itss;
    double  boundary_x[2][3];
    double  base_a[3];
    double  rb_total_mass, rb_inertia;
    const orc_material *mat;  /* nummat entries; material n is mat[n-1] */
    const double  *x;         /* np*3, particle-major */
    const double  *v;         /* np*3 or NULL */
    const int32_t *mat_id;    /* np, 1-based material number */
    const int32_t *color;     /* np */
    const int32_t *dispbc;    /* np */
    const double  *node_mass; /* np or NULL (massopt 1) */
    const double  *node_h;    /* unused (init_h_opt must be 0) */
    const double  *node_rho;  /* np or NULL (init_rhoe 1) */
} orc_deck;

typedef struct orc_scene orc_scene;

/* Phases of one MSolution::Solution iteration (MSolution.cpp:250-321). */
enum {
    ORC_PH_NEIGHBOURS = 0,   /* SetupLinkedList + GhostSetup + LinkedListNeighbours */
    ORC_PH_STRAIN,           /* Strain (half-step shift, ClassicStrain, shift back) */
    ORC_PH_NORMALS,          /* CalculateSurfaceNormals (if surface tension) */
    ORC_PH_CONSTITUTIVE,     /* RhoUpdate + EOS */
    ORC_PH_TIMESTEP,         /* CalcCriticalTimestep + dt rule + ptime */
    ORC_PH_MOMENTUM,         /* MorrisAcceleration [+ RigidBodyAcceleration] */
    ORC_PH_BOUNDCOND,        /* base acceleration + dispbc */
    ORC_PH_VELOCITY,         /* UpdateVelocity [+ RigidBodyVelocity] */
    ORC_PH_MOVE,             /* MoveParticles + timestep++ + stop test */
    ORC_NPHASES
};

int         orc_deck_size(void);
int         orc_material_size(void);
orc_scene  *orc_create(const orc_deck *deck, char *err, int errlen);
void        orc_destroy(orc_scene *s);
const char *orc_last_error(orc_scene *s);
int         orc_phase(orc_scene *s, int phase);
/* Runs up to nsteps full iterations; returns 0 OK, 1 error. *done set when
   the reference's stop test (t - dt >= endtime) fired. */
int         orc_run(orc_scene *s, int nsteps, int *steps_done, int *done);
int         orc_np(orc_scene *s);
/* OpenMP threads the per-particle loops use (1 without OpenMP) */
int orc_threads(void);
int         orc_ngp(orc_scene *s);
/* Globals: [dt, dtold, ptime, critts, timestep, hmax,
             cmm_min0..2, cmm_max0..2, ngp, rbR0..2, rbV0..2, rbW0..2, rbA0..2, rbAL0..2] */
int         orc_globals(orc_scene *s, double *out, int n);
/* Per-particle field download (0-based, np entries, vectors np*3). */
enum { ORC_F_X = 0, ORC_F_V, ORC_F_A, ORC_F_RHO, ORC_F_P, ORC_F_C, ORC_F_MASS,
       ORC_F_H, ORC_F_TRD, ORC_F_MINDIST, ORC_F_NORMAL, ORC_F_NNORMAL,
       ORC_F_N, ORC_F_KAPPA, ORC_F_CRITTS, ORC_F_RHOOLD, ORC_F_ACTIVE, ORC_F_VABS,
       ORC_F_Q, ORC_F_SIGMA_DIAG, ORC_F_SIGMA_OFF, ORC_F_ROD_DIAG, ORC_F_ROD_OFF };
/* ORC_F_SIGMA_* / ORC_F_ROD_* are vector fields (np*3): the diagonal xx, yy,
   zz or the off-diagonal xy, xz, yz of the symmetric tensors (disctype 0;
   with disctype 2 sigma is -p I and rod is not kept). */
int         orc_get_field(orc_scene *s, int field, double *out);
/* Test hook: overwrites the velocities (ORC_F_V only) between phases. */
int         orc_set_field(orc_scene *s, int field, const double *in);
/* Neighbour lists in the reference's order.  canon ids: a real neighbour j is
   j (0-based); a periodic image of real j shifted by (sx,sy,sz) in {-1,0,1}^3
   is encoded as -(1 + j*27 + (sx+1)*9 + (sy+1)*3 + (sz+1)). */
int         orc_get_neighbours(orc_scene *s, int32_t *nnbr, int64_t *canon, int maxnbr);
/* MKernelBSpline statics (MKernelBSpline.cpp:98-187): W, grad W (3 entries,
   0 beyond ndim) and dW/dr of one pair. */
double      orc_kernel_w(const double *xi, const double *xj, int ndim, double havg);
void        orc_kernel_gradw(double *dwdx, const double *xi, const double *xj, int ndim, double havg);
double      orc_kernel_dwdr(const double *xi, const double *xj, int ndim, double havg);
/* Linked-list state of the last SetupLinkedList + GhostSetup: grid geometry,
   sph_llgrid (ncell >= glim product), llpointer of the real particles (np, by
   0-based deck number; reference ids: 1-based reals, -g images, -1000000
   end) and of the images (ngp), each image's 1-based source particle.
   Any output may be NULL; returns 1 if an array is too small. */
int         orc_get_linked_list(orc_scene *s, double *gmin, double *gsize, int32_t *glim,
                                int32_t *llgrid, long long ncell, int32_t *real_next,
                                int32_t *ghost_next, int32_t *ghost_par, int ngp);

#ifdef __cplusplus
}
#endif
#endif
