/*
 * sph07_amd.hpp -- the reference's C++ call surface (MSimulationInit,
 * MSolution, MNeighbourVars, MKernelBSpline) as thin header-only classes over
 * the C ABI of sph07_amd.h.  A driver written against the reference's class
 * names calls these; every method forwards to one extern "C" entry point and
 * returns the reference's status codes (0 = OK, 1 = ERROR, plus the specific
 * SPH_ERR_* codes).  Vectors are plain double[3] / int[3] in place of the
 * uBLAS bounded_vector<.,3> of global.h:33-36.
 */
#ifndef SPH07_AMD_HPP_
#define SPH07_AMD_HPP_

#include "sph07_amd.h"

namespace sph07_amd {

/* MKernelBSpline (MKernelBSpline.h:7-18; base MKernel.h:6-18). */
class MKernelBSpline {
public:
    MKernelBSpline(int ndim, double havg) : m_ndim(ndim), m_havg(havg) {}
    virtual ~MKernelBSpline() {}

    /* MKernelBSpline::Kernel (MKernelBSpline.cpp:12-47) */
    double Kernel(const double xi[3], const double xj[3]) const { return Kernel(xi, xj, m_ndim, m_havg); }
    /* MKernelBSpline::GradW (MKernelBSpline.cpp:49-96) */
    void GradW(double dWdx[3], double &dWdr, const double xi[3], const double xj[3]) const
    {
        GradW(dWdx, dWdr, xi, xj, m_ndim, m_havg);
    }
    /* the static forms (MKernelBSpline.cpp:98-187) */
    static double Kernel(const double xi[3], const double xj[3], int ndim, double havg)
    {
        double w = 0.0;
        sph_kernel_eval(ndim, havg, 1, xi, xj, &w, nullptr, nullptr);
        return w;
    }
    static void GradW(double dWdx[3], double &dWdr, const double xi[3], const double xj[3], int ndim,
                      double havg)
    {
        sph_kernel_eval(ndim, havg, 1, xi, xj, nullptr, dWdx, &dWdr);
    }
    /* n pairs per device call: the form to use off the single-pair paths */
    static int Evaluate(int ndim, double havg, int n, const double *xi, const double *xj, double *w,
                        double *dWdx, double *dWdr)
    {
        return sph_kernel_eval(ndim, havg, n, xi, xj, w, dWdx, dWdr);
    }

    int m_ndim;
    double m_havg;
};

/* MNeighbourVars (MNeighbourVars.h:7-29): the grid of the last
   MSolution::Neighbours of a scene.  Refresh() re-reads the members after a
   Neighbours() call, as the reference's members change inside InitGrid. */
class MNeighbourVars {
public:
    explicit MNeighbourVars(void *ctx) : m_ctx(ctx)
    {
        for (int m = 0; m < 3; m++) {
            sph_gridsize[m] = sph_gridmin[m] = 0.0;
            sph_gridlim[m] = 1;
        }
    }
    int Refresh() { return sph_grid_info(m_ctx, sph_gridmin, sph_gridsize, sph_gridlim); }
    /* MNeighbourVars::GetGridValue (MNeighbourVars.cpp:70-73), by value */
    int GetGridValue(int i, int j, int k) const
    {
        int32_t v = SPH_LINKEDLISTEND;
        sph_get_grid_value(m_ctx, i, j, k, &v);
        return v;
    }
    /* MNeighbourVars::GetBoxCoords (MNeighbourVars.cpp:83-89) */
    int GetBoxCoords(const double x[3], int32_t boxcrd[3]) const { return sph_get_box_coords(m_ctx, x, boxcrd); }
    /* MNeighbourVars::GetLoopLimits (MNeighbourVars.cpp:91-103) */
    int GetLoopLimits(int ndim, const int32_t boxcrd[3], int32_t looplim[2][3]) const
    {
        return sph_get_loop_limits(m_ctx, ndim, boxcrd, &looplim[0][0]);
    }

    double sph_gridsize[3];
    double sph_gridmin[3];
    int32_t sph_gridlim[3];

private:
    void *m_ctx;
};

/* MSolution (MSolution.h:12-84): the phase methods of one scene. */
class MSolution {
public:
    MSolution(void *ctx, int ndim) : m_ctx(ctx), m_nv(ctx), m_ndim(ndim) {}

    int Neighbours()
    {
        const int rc = sph_neighbours(m_ctx);
        return rc ? rc : m_nv.Refresh();
    }
    int GhostSetup() { return sph_ghost_setup(m_ctx); }
    int LinkedListNeighbours() { return sph_linked_list_neighbours(m_ctx); }
    int Strain() { return sph_strain(m_ctx); }
    int CalculateSurfaceNormals() { return sph_calculate_surface_normals(m_ctx); }
    int Constitutive() { return sph_constitutive(m_ctx); }
    int CalcCriticalTimestep(double *critts = nullptr) { return sph_calc_critical_timestep(m_ctx, critts); }
    int Momentum() { return sph_momentum(m_ctx); }
    int BoundCond() { return sph_boundcond(m_ctx); }
    int UpdateVelocity() { return sph_update_velocity(m_ctx); }
    int MoveParticles() { return sph_move_particles(m_ctx); }
    /* MSolution::GenerateKernel (MSolution.cpp:415-434): the caller deletes it */
    MKernelBSpline *GenerateKernel(double havg) const { return new MKernelBSpline(m_ndim, havg); }

    /* MSolution::Solution (MSolution.cpp:238-324) without output: up to
       max_steps device-resident iterations, fused and captured in graphs. */
    int Solution(int max_steps, int *steps_done) { return sph_solution(m_ctx, max_steps, steps_done); }

    /* The same loop body, one phase entry point at a time, in the order of
       MSolution.cpp:250-319 (the FP64 sweeps). */
    int SolutionByPhases(int max_steps, int *steps_done)
    {
        int n = 0, rc = SPH_OK;
        sph_globals g;
        for (; n < max_steps; n++) {
            if ((rc = Neighbours())) break;
            if ((rc = GhostSetup())) break;           /* no-op without periodic planes */
            if ((rc = LinkedListNeighbours())) break;
            if ((rc = Strain())) break;
            if ((rc = CalculateSurfaceNormals())) break;  /* no-op without surface tension */
            if ((rc = Constitutive())) break;
            if ((rc = CalcCriticalTimestep())) break;
            if ((rc = sph_time_advance(m_ctx))) break;   /* dt rule, :275-277 */
            if ((rc = Momentum())) break;
            if ((rc = BoundCond())) break;
            if ((rc = UpdateVelocity())) break;
            if ((rc = MoveParticles())) break;
            if ((rc = sph_end_step(m_ctx))) break;       /* timestep++ and the stop test */
            if ((rc = sph_get_globals(m_ctx, &g))) break;
            if (g.done) {
                n++;
                break;
            }
        }
        if (steps_done) *steps_done = n;
        return rc;
    }

    MNeighbourVars &NeighbourVars() { return m_nv; }

private:
    void *m_ctx;
    MNeighbourVars m_nv;
    int m_ndim;
};

/* MSimulationInit::Initialize (MSimulationInit.cpp:31-99): creates the
   device scene of a deck; Context() is the handle the classes above take. */
class MSimulationInit {
public:
    MSimulationInit() {}
    ~MSimulationInit()
    {
        if (m_ctx) sph_destroy(m_ctx);
    }
    MSimulationInit(const MSimulationInit &) = delete;
    MSimulationInit &operator=(const MSimulationInit &) = delete;

    int Initialize(const sph_scene_desc *desc, int arm, char *err, int errlen)
    {
        if (m_ctx) sph_destroy(m_ctx);
        m_ctx = nullptr;
        return sph_create(desc, arm, &m_ctx, err, errlen);
    }
    void *Context() const { return m_ctx; }

private:
    void *m_ctx = nullptr;
};

}  // namespace sph07_amd

#endif
