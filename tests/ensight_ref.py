"""The reference's EnSight CASE writer restated in Python, as a checker of
the library's state output (sph_output.cpp).  Written from MOutput.cpp: the
geometry (WriteCaseGeo, :653-710), the case file (WriteCase, :116-167) and
one time step of the variable files (WriteCase1D/2D/3D, :169-651), with the
file set of MFileHandlingVars::OpenCaseFiles.  Python's %-formatting of
%e / %d / %s conversions is C's printf.

The model keeps every file as a string, opened new or appended as the
reference opens them: the geometry with the first state only, the case file
rewritten at every state, the time list and the variable files new at the
first state and appended after it.  Values per particle come from the
caller, so a test can feed it the fields the library downloaded at the same
moment and compare the files byte for byte.
"""
import numpy as np

VAR_FILES = [("_sig.ens", "stress"), ("_dis.ens", "displacement"), ("_vel.ens", "velocity"),
             ("_acc.ens", "acceleration"), ("_eps.ens", "efps"), ("_rho.ens", "density"),
             ("_egy.ens", "energy"), ("_snd.ens", "speed of sound"), ("_mas.ens", "mass"),
             ("_bnd.ens", "boundary particles"), ("_nbr.ens", "number of neighbours"),
             ("_tmp.ens", "temperaure")]

# WriteCase's variable table (name column, file suffix)
CASE_VARS = [("scalar per node:         1 1 mass         ", "_mas.ens"),
             ("scalar per node:         1 1 boundary     ", "_bnd.ens"),
             ("scalar per node:         1 1 nnbr         ", "_nbr.ens"),
             ("vector per node:         1 1 displacement ", "_dis.ens"),
             ("vector per node:         1 1 velocity     ", "_vel.ens"),
             ("vector per node:         1 1 acceleration ", "_acc.ens"),
             ("scalar per node:         1 1 efps         ", "_eps.ens"),
             ("scalar per node:         1 1 density      ", "_rho.ens"),
             ("scalar per node:         1 1 energy       ", "_egy.ens"),
             ("scalar per node:         1 1 soundspeed   ", "_snd.ens"),
             ("scalar per node:         1 1 temperature  ", "_tmp.ens"),
             ("tensor symm per node:    1 1 stress       ", "_sig.ens")]


class EnsightModel:
    def __init__(self, base, ndim, mat, nummat, xzero):
        self.base, self.ndim, self.nummat = base, ndim, nummat
        self.mat = [int(m) for m in mat]            # 1-based material of particle i + 1
        self.xzero = np.asarray(xzero, float).reshape(-1, 3)
        self.np = len(self.mat)
        self.istate = 0
        self.timenum = []
        self.files = {}

    # ---- MOutput::WriteCaseGeo ------------------------------------------
    def _geo(self, x):
        out = ["BEGIN TIME STEP\n", "Title1\n", "Title2\n", "node id given\n", "element id given\n",
               "coordinates\n", "%d\n" % self.np]
        per_mat = [0] * (self.nummat + 1)
        for j in range(1, self.np + 1):
            out.append("%8d %12.5e %12.5e %12.5e\n" % (j, x[j - 1, 0], x[j - 1, 1], x[j - 1, 2]))
            per_mat[self.mat[j - 1]] += 1
        for m in range(1, self.nummat + 1):
            out += ["part %8d\n" % m, "SPH particles in part %8d\n" % m, "point\n", "%8d\n" % per_mat[m]]
            out += ["%8d%8d\n" % (i, i) for i in range(1, self.np + 1) if self.mat[i - 1] == m]
        out.append("END TIME STEP\n")
        return "".join(out)

    # ---- MOutput::WriteCase ---------------------------------------------
    def _case(self, ptime):
        b = self.base
        out = ["FORMAT\n", "type:   ensight\n\n", "GEOMETRY\n",
               "model:                                    %26s\n\n" % (b + ".geo"), "VARIABLES\n"]
        for k, (name, ext) in enumerate(CASE_VARS):
            out.append("%s%26s\n" % (name, b + ext) + ("\n" if k == len(CASE_VARS) - 1 else ""))
        out += ["TIME\n", "time set:                1\n", "number of steps:     %5d\n" % (self.istate + 1),
                "filename start number:   1\n", "filename increment:      1\n"]
        if self.istate == 0:
            out.append("time values:  %12.5e\n" % ptime)
        else:
            out.append("time values:  %12.5e\n" % 0.0)
            out += ["              %12.5e\n" % t for t in self.timenum[1:]]
            out.append("              %12.5e\n\n" % ptime)
        out += ["FILE\n", "file set:                1\n", "filename index:          1\n",
                "number of steps:     %5d\n" % (self.istate + 1)]
        return "".join(out)

    # ---- MOutput::WriteCase1D / 2D / 3D ---------------------------------
    def _vars(self, f):
        D, n = self.ndim, self.np
        w = {ext: ["BEGIN TIME STEP\n", title + "\n"] for ext, title in VAR_FILES}
        for i in range(1, n + 1):
            q = i - 1
            dis = [f["x"][q, d] - self.xzero[q, d] if d < D else 0.0 for d in range(3)]
            vel = [f["v"][q, d] if d < D else 0.0 for d in range(3)]
            acc = [f["a"][q, d] if d < D else 0.0 for d in range(3)]
            w["_dis.ens"].append("%12.3e%12.3e%12.3e" % tuple(dis))
            w["_vel.ens"].append("%12.3e%12.3e%12.3e" % tuple(vel))
            w["_acc.ens"].append("%12.3e%12.3e%12.3e" % tuple(acc))
            w["_mas.ens"].append("%12.3e" % f["mass"][q])
            w["_bnd.ens"].append("%12.3e" % f["boundary"][q])
            w["_nbr.ens"].append("%12.3e" % f["nnbr"][q])
            if i % 2 == 0:
                for e in ("_dis.ens", "_vel.ens", "_acc.ens"):
                    w[e].append("\n")
            if i % 6 == 0:
                for e in ("_mas.ens", "_bnd.ens", "_nbr.ens"):
                    w[e].append("\n")
        if n % 2 != 0:
            for e in ("_dis.ens", "_vel.ens", "_acc.ens"):
                w[e].append("\n")
        if n % 6 != 0:
            for e in ("_mas.ens", "_bnd.ens", "_nbr.ens"):
                w[e].append("\n")
        per = ("_rho.ens", "_eps.ens", "_egy.ens", "_snd.ens", "_tmp.ens")
        for m in range(1, self.nummat + 1):
            for e in ("_rho.ens", "_sig.ens", "_eps.ens", "_egy.ens", "_snd.ens", "_tmp.ens"):
                w[e].append("part %8d\n" % m)
            for e in ("_rho.ens", "_sig.ens", "_eps.ens", "_egy.ens", "_snd.ens", "_tmp.ens"):
                w[e].append("point\n")
            count = 0
            for i in range(1, n + 1):
                q = i - 1
                if self.mat[q] != m:
                    continue
                count += 1
                s = f["sigma"][q]          # (00, 11, 01)
                if D == 1:
                    row = (s[0], 0.0, 0.0, 0.0, 0.0, 0.0)
                else:                      # 2D and 3D: (00, 11, 0, 01, 0, 0)
                    row = (s[0], s[1], 0.0, s[2], 0.0, 0.0)
                w["_sig.ens"].append("%12.4e%12.4e%12.4e%12.4e%12.4e%12.4e\n" % row)
                w["_rho.ens"].append("%12.4e" % f["rho"][q])
                w["_eps.ens"].append("%12.4e" % f["efps"][q])
                w["_egy.ens"].append("%12.4e" % f["energy"][q])
                w["_snd.ens"].append("%12.4e" % f["c"][q])
                w["_tmp.ens"].append("%12.4e" % f["temper"][q])
                # WriteCase3D tests the particle number, 1D and 2D the count
                if (i if D == 3 else count) % 6 == 0:
                    for e in per:
                        w[e].append("\n")
            # after WriteCase3D's loop the particle number is np + 1
            if ((n + 1) if D == 3 else count) % 6 != 0:
                for e in per:
                    w[e].append("\n")
        for ext, _ in VAR_FILES:
            w[ext].append("END TIME STEP\n")
        return {ext: "".join(v) for ext, v in w.items()}

    def state(self, fields, ptime):
        """One state (MOutput::StateOutput): fields holds per particle x, v,
        a (n x 3), mass, boundary, nnbr, rho, efps, energy, c, temper (n) and
        sigma (n x 3: 00, 11, 01)."""
        if self.istate == 0:
            self.files[".geo"] = self._geo(np.asarray(fields["x"], float))
        self.files[".case"] = self._case(ptime)
        t = "%13.6e\n" % ptime
        self.files["_time.txt"] = t if self.istate == 0 else self.files["_time.txt"] + t
        for ext, txt in self._vars(fields).items():
            self.files[ext] = txt if self.istate == 0 else self.files[ext] + txt
        self.istate += 1
        self.timenum.append(ptime)
