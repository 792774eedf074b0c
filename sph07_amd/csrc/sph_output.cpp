// sph_output.cpp -- state output of a device scene (host side, cold path):
//   * the EnSight CASE set the reference writes for state_opt 2
//     (MOutput::WriteCaseGeo / WriteCase / WriteCase1D-3D, MOutput.cpp:116-710;
//     file names of MFileHandlingVars::OpenCaseFiles, MFileHandlingVars.cpp:87-103),
//   * the cylinder plot of MSolution::PlotDroppingCylinder (MSolution.cpp:1638-1668),
//   * a rigid-body trajectory log (the reference has the line commented out,
//     MSolution.cpp:1660-1661; here one line per call),
//   * a full-precision dump of the state (hexadecimal floats).
// Fields come from sph_download in deck order.  Quantities the device path
// does not carry are written as the reference's model-9 fluid would hold
// them or as zero: efps, energy, temperature and the boundary flag (never
// assigned in the reference).  The stress is -p I for the Morris
// discretisation (it needs only tr D) and the full stress for the Classic one.
#include "../../include/sph07_amd.h"

#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

namespace {

struct Output {
    void *ctx = nullptr;
    std::string base;
    int ndim = 0, np = 0, nummat = 0;
    int istate = 0;                     // MOutputVars::sph_istate
    std::vector<double> timenum;        // MOutputVars::sph_timenum
    std::vector<double> x0;             // xzero: positions when the output was opened
    std::vector<int> mat;               // material of each particle (1-based)
    std::string err;
};

FILE *open_file(Output *o, const char *ext, bool append)
{
    const std::string name = o->base + ext;
    FILE *f = fopen(name.c_str(), append ? "at" : "wt");
    if (!f) o->err = "cannot open " + name;
    return f;
}

struct Fields {
    std::vector<double> x, v, a, rho, p, c, m, nnbr, sd, so;   // sd, so: stress diagonal, off-diagonal
};

int fetch(Output *o, Fields &F)
{
    const size_t n = (size_t)o->np;
    F.x.resize(3 * n);
    F.v.resize(3 * n);
    F.a.resize(3 * n);
    F.rho.resize(n);
    F.p.resize(n);
    F.c.resize(n);
    F.m.resize(n);
    F.nnbr.resize(n);
    F.sd.resize(3 * n);
    F.so.resize(3 * n);
    int rc = sph_download(o->ctx, SPH_F_X, F.x.data());
    if (!rc) rc = sph_download(o->ctx, SPH_F_V, F.v.data());
    if (!rc) rc = sph_download(o->ctx, SPH_F_A, F.a.data());
    if (!rc) rc = sph_download(o->ctx, SPH_F_RHO, F.rho.data());
    if (!rc) rc = sph_download(o->ctx, SPH_F_P, F.p.data());
    if (!rc) rc = sph_download(o->ctx, SPH_F_C, F.c.data());
    if (!rc) rc = sph_download(o->ctx, SPH_F_MASS, F.m.data());
    if (!rc) rc = sph_download(o->ctx, SPH_F_NNBR, F.nnbr.data());
    if (!rc) rc = sph_download(o->ctx, SPH_F_SIGMA_DIAG, F.sd.data());
    if (!rc) rc = sph_download(o->ctx, SPH_F_SIGMA_OFF, F.so.data());
    if (rc) o->err = sph_last_error(o->ctx);
    return rc;
}

// MOutput::WriteCaseGeo (MOutput.cpp:653-710): node coordinates, then one
// part of point elements per material
int write_geo(Output *o, const Fields &F, bool append)
{
    FILE *f = open_file(o, ".geo", append);
    if (!f) return SPH_ERROR;
    std::vector<int> per_mat(o->nummat + 1, 0);
    fprintf(f, "BEGIN TIME STEP\n");
    fprintf(f, "Title1\n");
    fprintf(f, "Title2\n");
    fprintf(f, "node id given\n");
    fprintf(f, "element id given\n");
    fprintf(f, "coordinates\n");
    fprintf(f, "%d\n", o->np);
    for (int j = 0; j < o->np; j++) {
        fprintf(f, "%8d %12.5e %12.5e %12.5e\n", j + 1, F.x[3 * (size_t)j], F.x[3 * (size_t)j + 1],
                F.x[3 * (size_t)j + 2]);
        per_mat[o->mat[j]]++;
    }
    for (int m = 1; m <= o->nummat; m++) {
        fprintf(f, "part %8d\n", m);
        fprintf(f, "SPH particles in part %8d\n", m);
        fprintf(f, "point\n");
        fprintf(f, "%8d\n", per_mat[m]);
        for (int i = 0; i < o->np; i++)
            if (o->mat[i] == m) fprintf(f, "%8d%8d\n", i + 1, i + 1);
    }
    fprintf(f, "END TIME STEP\n");
    fclose(f);
    return SPH_OK;
}

// MOutput::WriteCase (MOutput.cpp:116-167): the case file, rewritten with
// the time values of every state so far
int write_case(Output *o, double ptime)
{
    FILE *f = open_file(o, ".case", false);
    if (!f) return SPH_ERROR;
    const std::string &b = o->base;
    fprintf(f, "FORMAT\n");
    fprintf(f, "type:   ensight\n\n");
    fprintf(f, "GEOMETRY\n");
    fprintf(f, "model:                                    %26s\n\n", (b + ".geo").c_str());
    fprintf(f, "VARIABLES\n");
    const char *rows[][2] = {{"scalar per node:         1 1 mass         ", "_mas.ens"},
                             {"scalar per node:         1 1 boundary     ", "_bnd.ens"},
                             {"scalar per node:         1 1 nnbr         ", "_nbr.ens"},
                             {"vector per node:         1 1 displacement ", "_dis.ens"},
                             {"vector per node:         1 1 velocity     ", "_vel.ens"},
                             {"vector per node:         1 1 acceleration ", "_acc.ens"},
                             {"scalar per node:         1 1 efps         ", "_eps.ens"},
                             {"scalar per node:         1 1 density      ", "_rho.ens"},
                             {"scalar per node:         1 1 energy       ", "_egy.ens"},
                             {"scalar per node:         1 1 soundspeed   ", "_snd.ens"},
                             {"scalar per node:         1 1 temperature  ", "_tmp.ens"}};
    for (auto &r : rows) fprintf(f, "%s%26s\n", r[0], (b + r[1]).c_str());
    fprintf(f, "tensor symm per node:    1 1 stress       %26s\n\n", (b + "_sig.ens").c_str());
    fprintf(f, "TIME\n");
    fprintf(f, "time set:                1\n");
    fprintf(f, "number of steps:     %5d\n", o->istate + 1);
    fprintf(f, "filename start number:   1\n");
    fprintf(f, "filename increment:      1\n");
    if (o->istate == 0) {
        fprintf(f, "time values:  %12.5e\n", ptime);
    } else {
        fprintf(f, "time values:  %12.5e\n", 0.0);
        for (size_t i = 1; i < o->timenum.size(); i++) fprintf(f, "              %12.5e\n", o->timenum[i]);
        fprintf(f, "              %12.5e\n\n", ptime);
    }
    fprintf(f, "FILE\n");
    fprintf(f, "file set:                1\n");
    fprintf(f, "filename index:          1\n");
    fprintf(f, "number of steps:     %5d\n", o->istate + 1);
    fclose(f);
    return SPH_OK;
}

// MOutput::WriteCase1D / 2D / 3D (MOutput.cpp:169-651): one time step of
// every variable file, new files for the first state, appended after
int write_vars(Output *o, const Fields &F, double ptime)
{
    const bool app = o->istate != 0;
    FILE *ft = open_file(o, "_time.txt", app);
    if (!ft) return SPH_ERROR;
    fprintf(ft, "%13.6e\n", ptime);
    fclose(ft);
    const char *ext[12] = {"_sig.ens", "_dis.ens", "_vel.ens", "_acc.ens", "_eps.ens", "_rho.ens",
                           "_egy.ens", "_snd.ens", "_mas.ens", "_bnd.ens", "_nbr.ens", "_tmp.ens"};
    const char *title[12] = {"stress", "displacement", "velocity", "acceleration", "efps", "density",
                             "energy", "speed of sound", "mass", "boundary particles", "number of neighbours",
                             "temperaure"};
    FILE *fv[12];
    for (int k = 0; k < 12; k++) {
        fv[k] = open_file(o, ext[k], app);
        if (!fv[k]) {
            for (int q = 0; q < k; q++) fclose(fv[q]);
            return SPH_ERROR;
        }
    }
    FILE *fsig = fv[0], *fdis = fv[1], *fvel = fv[2], *facc = fv[3], *feps = fv[4], *frho = fv[5], *fegy = fv[6],
         *fsnd = fv[7], *fmas = fv[8], *fbnd = fv[9], *fnbr = fv[10], *ftmp = fv[11];
    for (int k = 0; k < 12; k++) fprintf(fv[k], "BEGIN TIME STEP\n");
    for (int k = 0; k < 12; k++) fprintf(fv[k], "%s\n", title[k]);
    const int D = o->ndim, n = o->np;
    for (int i = 1; i <= n; i++) {
        const size_t b = 3 * (size_t)(i - 1);
        double dis[3], vel[3], acc[3];
        for (int d = 0; d < 3; d++) {
            const bool on = d < D;
            dis[d] = on ? F.x[b + d] - o->x0[b + d] : 0.0;
            vel[d] = on ? F.v[b + d] : 0.0;
            acc[d] = on ? F.a[b + d] : 0.0;
        }
        fprintf(fdis, "%12.3e%12.3e%12.3e", dis[0], dis[1], dis[2]);
        fprintf(fvel, "%12.3e%12.3e%12.3e", vel[0], vel[1], vel[2]);
        fprintf(facc, "%12.3e%12.3e%12.3e", acc[0], acc[1], acc[2]);
        fprintf(fmas, "%12.3e", F.m[i - 1]);
        fprintf(fbnd, "%12.3e", 0.0);
        fprintf(fnbr, "%12.3e", F.nnbr[i - 1]);
        if (i % 2 == 0) {
            fprintf(fdis, "\n");
            fprintf(fvel, "\n");
            fprintf(facc, "\n");
        }
        if (i % 6 == 0) {
            fprintf(fmas, "\n");
            fprintf(fbnd, "\n");
            fprintf(fnbr, "\n");
        }
    }
    if (n % 2 != 0) {
        fprintf(fdis, "\n");
        fprintf(fvel, "\n");
        fprintf(facc, "\n");
    }
    if (n % 6 != 0) {
        fprintf(fmas, "\n");
        fprintf(fbnd, "\n");
        fprintf(fnbr, "\n");
    }
    FILE *per[5] = {frho, feps, fegy, fsnd, ftmp};
    for (int m = 1; m <= o->nummat; m++) {
        for (FILE *f : {frho, fsig, feps, fegy, fsnd, ftmp}) fprintf(f, "part %8d\n", m);
        for (FILE *f : {frho, fsig, feps, fegy, fsnd, ftmp}) fprintf(f, "point\n");
        int count = 0, i = 1;
        for (; i <= n; i++) {
            if (o->mat[i - 1] != m) continue;
            count++;
            // sigma: -p I for the Morris discretisation, with the deviatoric
            // part for the Classic one
            const double *sd = &F.sd[3 * (size_t)(i - 1)], *so = &F.so[3 * (size_t)(i - 1)];
            if (D == 1)
                fprintf(fsig, "%12.4e%12.4e%12.4e%12.4e%12.4e%12.4e\n", sd[0], 0.0, 0.0, 0.0, 0.0, 0.0);
            else    // the 2D and 3D writers both write (00, 11, 0, 01, 0, 0)
                fprintf(fsig, "%12.4e%12.4e%12.4e%12.4e%12.4e%12.4e\n", sd[0], sd[1], 0.0, so[0], 0.0, 0.0);
            fprintf(frho, "%12.4e", F.rho[i - 1]);
            fprintf(feps, "%12.4e", 0.0);
            fprintf(fegy, "%12.4e", 0.0);
            fprintf(fsnd, "%12.4e", F.c[i - 1]);
            fprintf(ftmp, "%12.4e", 0.0);
            // the 3D writer breaks lines on the particle number, 1D/2D on the count
            const int brk = D == 3 ? i : count;
            if (brk % 6 == 0)
                for (FILE *f : per) fprintf(f, "\n");
        }
        const int tail = D == 3 ? i : count;    // i = np + 1 after the loop in 3D
        if (tail % 6 != 0)
            for (FILE *f : per) fprintf(f, "\n");
    }
    for (int k = 0; k < 12; k++) {
        fprintf(fv[k], "END TIME STEP\n");
        fclose(fv[k]);
    }
    return SPH_OK;
}

Output *out_of(void *p) { return (Output *)p; }

}  // namespace

extern "C" {

int sph_output_open(void *ctx, const char *base, int32_t ndim, int32_t nummat, const int32_t *mat_id, void **out)
{
    if (out) *out = nullptr;
    if (!ctx || !base || !out || ndim < 1 || ndim > 3 || nummat < 1 || !mat_id) return SPH_ERROR;
    sph_globals g;
    int rc = sph_get_globals(ctx, &g);
    if (rc) return rc;
    Output *o = new Output();
    o->ctx = ctx;
    o->base = base;
    o->np = g.np;
    o->nummat = nummat;
    o->mat.assign(mat_id, mat_id + g.np);
    for (int v : o->mat)
        if (v < 1 || v > nummat) {
            delete o;
            return SPH_ERROR;
        }
    o->ndim = ndim;
    o->x0.resize(3 * (size_t)g.np);           // xzero (MSimulationInit.cpp)
    rc = sph_download(ctx, SPH_F_X, o->x0.data());
    if (rc) {
        delete o;
        return rc;
    }
    *out = o;
    return SPH_OK;
}

// MOutput::InitStateOutput + StateOutput (MOutput.cpp:39-114): the geometry
// with the first state, the case file and one time step of the variables.
int sph_output_state(void *out)
{
    Output *o = out_of(out);
    if (!o) return SPH_ERROR;
    sph_globals g;
    int rc = sph_get_globals(o->ctx, &g);
    if (rc) return rc;
    Fields F;
    if (fetch(o, F)) return SPH_ERROR;
    if (o->istate == 0 && write_geo(o, F, false)) return SPH_ERROR;
    if (write_case(o, g.ptime)) return SPH_ERROR;
    if (write_vars(o, F, g.ptime)) return SPH_ERROR;
    o->istate++;                       // MOutput::IncreaseStateCounter (:712-720)
    o->timenum.push_back(g.ptime);
    return SPH_OK;
}

// MSolution::PlotDroppingCylinder (MSolution.cpp:1638-1668): x(0) and rho
// of particles 1 .. np-1 (the reference's loop stops before the last one)
int sph_output_cylinder_plot(void *out, const char *path)
{
    Output *o = out_of(out);
    if (!o || !path) return SPH_ERROR;
    Fields F;
    if (fetch(o, F)) return SPH_ERROR;
    FILE *f = fopen(path, "wt");
    if (!f) {
        o->err = std::string("cannot open ") + path;
        return SPH_ERROR;
    }
    for (int j = 1; j < o->np; j++) fprintf(f, "%13.5e%13.5e\n", F.x[3 * (size_t)(j - 1)], F.rho[j - 1]);
    fclose(f);
    return SPH_OK;
}

// One line of the rigid-body trajectory: time, R, V, OMEGA (%13.5e, the
// widths of the reference's commented-out line, MSolution.cpp:1660-1661)
int sph_output_rigid(void *out, const char *path)
{
    Output *o = out_of(out);
    if (!o || !path) return SPH_ERROR;
    sph_globals g;
    int rc = sph_get_globals(o->ctx, &g);
    if (rc) return rc;
    FILE *f = fopen(path, "at");
    if (!f) return SPH_ERROR;
    fprintf(f, "%13.5e", g.ptime);
    for (int m = 0; m < 3; m++) fprintf(f, "%13.5e", g.rb_R[m]);
    for (int m = 0; m < 3; m++) fprintf(f, "%13.5e", g.rb_V[m]);
    for (int m = 0; m < 3; m++) fprintf(f, "%13.5e", g.rb_OMEGA[m]);
    fprintf(f, "\n");
    fclose(f);
    return SPH_OK;
}

const char *sph_output_last_error(void *out)
{
    Output *o = out_of(out);
    return o ? o->err.c_str() : "null output";
}

void sph_output_close(void *out) { delete out_of(out); }

// Full-precision dump: the globals, then x, v, a (3 per particle), rho, p
// per particle in deck order, every value as a C99 hexadecimal float.
int sph_write_state_dump(void *ctx, const char *path)
{
    if (!ctx || !path) return SPH_ERROR;
    sph_globals g;
    int rc = sph_get_globals(ctx, &g);
    if (rc) return rc;
    const size_t n = (size_t)g.np;
    std::vector<double> x(3 * n), v(3 * n), a(3 * n), rho(n), p(n);
    rc = sph_download(ctx, SPH_F_X, x.data());
    if (!rc) rc = sph_download(ctx, SPH_F_V, v.data());
    if (!rc) rc = sph_download(ctx, SPH_F_A, a.data());
    if (!rc) rc = sph_download(ctx, SPH_F_RHO, rho.data());
    if (!rc) rc = sph_download(ctx, SPH_F_P, p.data());
    if (rc) return rc;
    FILE *f = fopen(path, "wt");
    if (!f) return SPH_ERROR;
    fprintf(f, "sph07_amd state dump\n");
    fprintf(f, "np %d timestep %d\n", g.np, g.timestep);
    fprintf(f, "ptime %a dt %a dtold %a critts %a\n", g.ptime, g.dt, g.dtold, g.critts);
    fprintf(f, "rbR %a %a %a rbV %a %a %a rbOMEGA %a %a %a\n", g.rb_R[0], g.rb_R[1], g.rb_R[2], g.rb_V[0],
            g.rb_V[1], g.rb_V[2], g.rb_OMEGA[0], g.rb_OMEGA[1], g.rb_OMEGA[2]);
    for (size_t i = 0; i < n; i++)
        fprintf(f, "%a %a %a %a %a %a %a %a %a %a %a\n", x[3 * i], x[3 * i + 1], x[3 * i + 2], v[3 * i],
                v[3 * i + 1], v[3 * i + 2], a[3 * i], a[3 * i + 1], a[3 * i + 2], rho[i], p[i]);
    fclose(f);
    return SPH_OK;
}

}  // extern "C"
