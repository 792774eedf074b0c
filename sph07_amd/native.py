"""ctypes binding of the C ABI in ``include/sph07_amd.h``.

``Scene`` mirrors the reference's MSolution surface (MSolution.h:25-65):
one method per phase, the same order, the same OK/ERROR convention (raised
as :class:`SphError` here).  The library is the HIP build in this package;
there is no CPU fallback -- if it is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

from .deck import CDeck, CMaterial, Deck

HERE = os.path.dirname(os.path.abspath(__file__))
# SPH07_LIB overrides the library path (alternative builds of the same sources)
LIB_PATH = os.environ.get("SPH07_LIB") or os.path.join(HERE, "libsph07_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "sph07_amd.h")

ARM_FP64 = 0
ARM_FP32C = 1

F_X, F_V, F_A, F_RHO, F_P, F_C, F_MASS, F_H, F_TRD, F_MINDIST, F_NORMAL, F_NNORMAL, \
    F_N, F_KAPPA, F_CRITTS, F_RHOOLD, F_ACTIVE, F_VABS, F_NNBR, F_Q, F_SIGMA_DIAG, F_SIGMA_OFF, \
    F_ROD_DIAG, F_ROD_OFF = range(24)
_VEC = {F_X, F_V, F_A, F_NORMAL, F_NNORMAL, F_SIGMA_DIAG, F_SIGMA_OFF, F_ROD_DIAG, F_ROD_OFF}


class SphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class CGlobals(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("dtold", ctypes.c_double), ("ptime", ctypes.c_double),
        ("critts", ctypes.c_double), ("coord_minmax", (ctypes.c_double * 3) * 2),
        ("hmax", ctypes.c_double),
        ("rb_R", ctypes.c_double * 3), ("rb_V", ctypes.c_double * 3),
        ("rb_OMEGA", ctypes.c_double * 3), ("rb_A", ctypes.c_double * 3),
        ("rb_ALPHA", ctypes.c_double * 3),
        ("timestep", ctypes.c_int32), ("ngp", ctypes.c_int32), ("done", ctypes.c_int32),
        ("error", ctypes.c_int32), ("np", ctypes.c_int32), ("grid_lim", ctypes.c_int32 * 3),
    ]


_lib = None


def header_symbols():
    """Names of the functions declared in include/sph07_amd.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(
        r"^\s*(?:int|void|long long|double|const char|const sph_scene_desc)\s*\**\s*(sph_\w+)\s*\(",
        txt, re.M)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                          "(python -m sph07_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    L.sph_create.argtypes = [ctypes.c_void_p, ip, ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ip]
    L.sph_destroy.argtypes = [vp]
    L.sph_destroy.restype = None
    L.sph_last_error.argtypes = [vp]
    L.sph_last_error.restype = ctypes.c_char_p
    for name in ("sph_neighbours", "sph_ghost_setup", "sph_linked_list_neighbours", "sph_strain",
                 "sph_calculate_surface_normals", "sph_constitutive", "sph_time_advance",
                 "sph_momentum", "sph_boundcond", "sph_update_velocity", "sph_move_particles",
                 "sph_end_step", "sph_synchronize"):
        getattr(L, name).argtypes = [vp]
    L.sph_calc_critical_timestep.argtypes = [vp, dp]
    L.sph_solution.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_int)]
    L.sph_get_globals.argtypes = [vp, ctypes.POINTER(CGlobals)]
    L.sph_download.argtypes = [vp, ip, ctypes.c_void_p]
    L.sph_get_neighbours.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p, ip]
    L.sph_set_timing.argtypes = [vp, ip]
    L.sph_get_timing.argtypes = [vp, ip, ctypes.c_char_p, ip, dp, ctypes.POINTER(ctypes.c_longlong)]
    assert L.sph_desc_size() == ctypes.sizeof(CDeck), "sph_scene_desc layout mismatch"
    assert L.sph_material_desc_size() == ctypes.sizeof(CMaterial), "material layout mismatch"
    _lib = L
    return L


class Scene:
    """One device-resident scene (MSimulationInit::Initialize on creation)."""

    def __init__(self, deck: Deck, arm: int = ARM_FP64):
        L = lib()
        c, self._keep = deck.to_c()
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = L.sph_create(ctypes.byref(c), arm, ctypes.byref(h), err, 512)
        if rc != 0:
            raise SphError(rc, err.value.decode())
        self._h = h
        self.np = deck.np
        self.ndim = deck.ndim
        self.maxnbr = deck.maxnbr
        self.arm = arm

    def close(self):
        if getattr(self, "_h", None):
            lib().sph_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise SphError(rc, lib().sph_last_error(self._h).decode())

    # --- MSolution phase surface ---
    def Neighbours(self):
        self._chk(lib().sph_neighbours(self._h))

    def GhostSetup(self):
        self._chk(lib().sph_ghost_setup(self._h))

    def LinkedListNeighbours(self):
        self._chk(lib().sph_linked_list_neighbours(self._h))

    def Strain(self):
        self._chk(lib().sph_strain(self._h))

    def CalculateSurfaceNormals(self):
        self._chk(lib().sph_calculate_surface_normals(self._h))

    def Constitutive(self):
        self._chk(lib().sph_constitutive(self._h))

    def CalcCriticalTimestep(self) -> float:
        out = ctypes.c_double(0.0)
        self._chk(lib().sph_calc_critical_timestep(self._h, ctypes.byref(out)))
        return out.value

    def TimeAdvance(self):
        self._chk(lib().sph_time_advance(self._h))

    def Momentum(self):
        self._chk(lib().sph_momentum(self._h))

    def BoundCond(self):
        self._chk(lib().sph_boundcond(self._h))

    def UpdateVelocity(self):
        self._chk(lib().sph_update_velocity(self._h))

    def MoveParticles(self):
        self._chk(lib().sph_move_particles(self._h))

    def EndStep(self):
        self._chk(lib().sph_end_step(self._h))

    def phase(self, k: int):
        """Phase k of one iteration, numbered as the oracle's phases."""
        [lambda: (self.Neighbours(), self.GhostSetup(), self.LinkedListNeighbours()),
         self.Strain, self.CalculateSurfaceNormals, self.Constitutive,
         lambda: (self.CalcCriticalTimestep(), self.TimeAdvance()),
         self.Momentum, self.BoundCond, self.UpdateVelocity,
         lambda: (self.MoveParticles(), self.EndStep())][k]()

    def Solution(self, nsteps: int) -> int:
        done = ctypes.c_int(0)
        self._chk(lib().sph_solution(self._h, nsteps, ctypes.byref(done)))
        return done.value

    def synchronize(self):
        self._chk(lib().sph_synchronize(self._h))

    # --- state access ---
    def globals(self) -> dict:
        g = CGlobals()
        self._chk(lib().sph_get_globals(self._h, ctypes.byref(g)))
        out = {k: getattr(g, k) for k in ("dt", "dtold", "ptime", "critts", "hmax", "timestep",
                                          "ngp", "done", "error", "np")}
        for s, nm in ((0, "cmm_min"), (1, "cmm_max")):
            for m in range(3):
                out[f"{nm}{m}"] = g.coord_minmax[s][m]
        for nm, arr in (("rbR", g.rb_R), ("rbV", g.rb_V), ("rbW", g.rb_OMEGA),
                        ("rbA", g.rb_A), ("rbAL", g.rb_ALPHA)):
            for m in range(3):
                out[f"{nm}{m}"] = arr[m]
        out["grid_lim"] = tuple(g.grid_lim)
        return out

    def field(self, f: int) -> np.ndarray:
        out = np.zeros((self.np, 3) if f in _VEC else self.np)
        self._chk(lib().sph_download(self._h, f, out.ctypes.data))
        return out

    def set_timing(self, on: bool):
        self._chk(lib().sph_set_timing(self._h, 1 if on else 0))

    def autotune(self, reps: int = 3) -> dict:
        """sph_autotune: reps > 0 tunes now, 0 freezes the current launch
        shapes (no automatic tuning), < 0 only reports them."""
        L = lib()
        L.sph_autotune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
        shape = (ctypes.c_int32 * 4)()
        self._chk(L.sph_autotune(self._h, reps, shape))
        return {"cell_slots": shape[0], "strain_batch": shape[1], "momentum_batch": shape[2],
                "cell_order": "row-major" if shape[3] == 1 else "morton"}

    def set_launch_shape(self, cell_slots: int = 32, strain_batch: int = None, momentum_batch: int = None,
                         cell_order: str = "morton"):
        """sph_set_launch_shape: fixes the launch shapes (no automatic tuning).
        cell_slots 0 selects the per-particle neighbour walk (the same lists).
        The batches are the neighbour records in flight per lane of the
        scene's arm's continuity and momentum sweeps; None keeps the arm's
        untuned default (FP64 4 / 2, FP32c 8 / 8)."""
        L = lib()
        L.sph_set_launch_shape.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
        fp32c = self.arm == ARM_FP32C
        if strain_batch is None:
            strain_batch = 8 if fp32c else 4
        if momentum_batch is None:
            momentum_batch = 8 if fp32c else 2
        shape = (ctypes.c_int32 * 4)(cell_slots, strain_batch, momentum_batch,
                                     1 if cell_order == "row-major" else 0)
        self._chk(L.sph_set_launch_shape(self._h, shape))

    def timing(self) -> dict:
        """{group name: (total ms, launches)} from HIP events on the scene's stream."""
        L = lib()
        out = {}
        for k in range(L.sph_get_timing(self._h, -1, None, 0, None, None)):
            name = ctypes.create_string_buffer(64)
            ms = ctypes.c_double(0.0)
            cnt = ctypes.c_longlong(0)
            self._chk(L.sph_get_timing(self._h, k, name, 64, ctypes.byref(ms), ctypes.byref(cnt)))
            out[name.value.decode()] = (ms.value, cnt.value)
        return out

    def neighbours(self):
        nn = np.zeros(self.np, np.int32)
        canon = np.zeros((self.np, self.maxnbr), np.int64)
        self._chk(lib().sph_get_neighbours(self._h, nn.ctypes.data, canon.ctypes.data, self.maxnbr))
        return nn, canon

    # --- MNeighbourVars surface (MNeighbourVars.h:13-28) ---
    def grid_info(self) -> dict:
        """sph_gridmin, sph_gridsize, sph_gridlim of the last InitGrid."""
        gmin, gsize = np.zeros(3), np.zeros(3)
        glim = np.zeros(3, np.int32)
        self._chk(_surface().sph_grid_info(self._h, gmin.ctypes.data, gsize.ctypes.data,
                                           glim.ctypes.data))
        return {"gridmin": gmin, "gridsize": gsize, "gridlim": glim}

    def linked_list(self) -> dict:
        """The reference's linked lists: llgrid (gridlim[0] x gridlim[1] x
        gridlim[2], cell heads), real_next (llpointer per deck particle),
        ghost_next / ghost_par (per image).  Reference ids: 1-based reals, -g
        for image g, LINKEDLISTEND at a list's end."""
        glim = self.grid_info()["gridlim"]
        ngp = self.globals()["ngp"]
        ncell = int(np.prod(glim.astype(np.int64)))
        grid = np.zeros(ncell, np.int32)
        real_next = np.zeros(self.np, np.int32)
        gnext = np.zeros(max(ngp, 1), np.int32)
        gpar = np.zeros(max(ngp, 1), np.int32)
        self._chk(_surface().sph_get_linked_list(self._h, grid.ctypes.data, ncell,
                                                 real_next.ctypes.data, gnext.ctypes.data,
                                                 gpar.ctypes.data, max(ngp, 1)))
        return {"llgrid": grid.reshape(tuple(int(g) for g in glim)), "real_next": real_next,
                "ghost_next": gnext[:ngp], "ghost_par": gpar[:ngp]}

    def GetGridValue(self, i: int, j: int, k: int) -> int:
        out = ctypes.c_int32(0)
        self._chk(_surface().sph_get_grid_value(self._h, i, j, k, ctypes.byref(out)))
        return out.value

    def GetBoxCoords(self, x) -> np.ndarray:
        xv = np.zeros(3)
        xv[:len(x)] = x
        out = np.zeros(3, np.int32)
        self._chk(_surface().sph_get_box_coords(self._h, xv.ctypes.data, out.ctypes.data))
        return out

    def GetLoopLimits(self, ndim: int, boxcrd) -> np.ndarray:
        b = np.ascontiguousarray(boxcrd, np.int32)
        out = np.zeros((2, 3), np.int32)
        self._chk(_surface().sph_get_loop_limits(self._h, ndim, b.ctypes.data, out.ctypes.data))
        return out


class StateOutput:
    """The reference's state output of a scene (MOutput, EnSight CASE format;
    sph_output_*): open it after creating the scene (the current positions are
    the displacements' origin), then call state() at each output time."""

    def __init__(self, scene: "Scene", base: str, deck: Deck):
        L = lib()
        vp = ctypes.c_void_p
        L.sph_output_open.argtypes = [vp, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, vp,
                                      ctypes.POINTER(vp)]
        for name in ("sph_output_state",):
            getattr(L, name).argtypes = [vp]
        L.sph_output_cylinder_plot.argtypes = [vp, ctypes.c_char_p]
        L.sph_output_rigid.argtypes = [vp, ctypes.c_char_p]
        L.sph_output_last_error.argtypes = [vp]
        L.sph_output_last_error.restype = ctypes.c_char_p
        L.sph_output_close.argtypes = [vp]
        L.sph_output_close.restype = None
        self._mat = np.ascontiguousarray(deck.mat_id, np.int32)
        h = vp()
        rc = L.sph_output_open(scene._h, base.encode(), deck.ndim, len(deck.materials),
                               self._mat.ctypes.data, ctypes.byref(h))
        if rc != 0:
            raise SphError(rc, "sph_output_open failed")
        self._h = h

    def _chk(self, rc):
        if rc != 0:
            raise SphError(rc, lib().sph_output_last_error(self._h).decode())

    def state(self):
        self._chk(lib().sph_output_state(self._h))

    def cylinder_plot(self, path: str):
        self._chk(lib().sph_output_cylinder_plot(self._h, path.encode()))

    def rigid(self, path: str):
        self._chk(lib().sph_output_rigid(self._h, path.encode()))

    def close(self):
        if getattr(self, "_h", None):
            lib().sph_output_close(self._h)
            self._h = None

    def __del__(self):
        self.close()


def write_state_dump(scene: "Scene", path: str):
    """sph_write_state_dump: globals and x, v, a, rho, p as hexadecimal floats."""
    L = lib()
    L.sph_write_state_dump.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    rc = L.sph_write_state_dump(scene._h, path.encode())
    if rc != 0:
        raise SphError(rc, "sph_write_state_dump failed")


def read_state_dump(path: str) -> dict:
    """Parses a state dump written by write_state_dump."""
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[0] == "sph07_amd state dump"
    head = lines[1].split()
    out = {"np": int(head[1]), "timestep": int(head[3])}
    g = lines[2].split()
    for k in range(0, len(g), 2):
        out[g[k]] = float.fromhex(g[k + 1])
    vals = np.array([[float.fromhex(t) for t in ln.split()] for ln in lines[4:4 + out["np"]]])
    out["x"], out["v"], out["a"] = vals[:, 0:3], vals[:, 3:6], vals[:, 6:9]
    out["rho"], out["p"] = vals[:, 9], vals[:, 10]
    return out


LINKEDLISTEND = -1000000   # global.h:13


def _surface():
    """Argument types of the MNeighbourVars / MKernel entry points."""
    L = lib()
    if not getattr(L, "_surface_typed", False):
        vp, ip = ctypes.c_void_p, ctypes.c_int
        L.sph_grid_info.argtypes = [vp, vp, vp, vp]
        L.sph_get_linked_list.argtypes = [vp, vp, ctypes.c_longlong, vp, vp, vp, ip]
        L.sph_get_grid_value.argtypes = [vp, ip, ip, ip, ctypes.POINTER(ctypes.c_int32)]
        L.sph_get_box_coords.argtypes = [vp, vp, vp]
        L.sph_get_loop_limits.argtypes = [vp, ip, vp, vp]
        L.sph_kernel_eval.argtypes = [ip, ctypes.c_double, ip, vp, vp, vp, vp, vp]
        L._surface_typed = True
    return L


def kernel_eval(ndim: int, havg: float, xi, xj):
    """MKernelBSpline::Kernel and ::GradW (statics, MKernelBSpline.cpp:98-187)
    evaluated on the GPU for every row pair of xi, xj (n x 3): returns
    (W[n], dWdx[n, 3], dWdr[n])."""
    xi = np.ascontiguousarray(xi, np.float64).reshape(-1, 3)
    xj = np.ascontiguousarray(xj, np.float64).reshape(-1, 3)
    assert xi.shape == xj.shape
    n = xi.shape[0]
    w, dwdx, dwdr = np.zeros(n), np.zeros((n, 3)), np.zeros(n)
    rc = _surface().sph_kernel_eval(ndim, havg, n, xi.ctypes.data, xj.ctypes.data, w.ctypes.data,
                                    dwdx.ctypes.data, dwdr.ctypes.data)
    if rc != 0:
        raise SphError(rc, "sph_kernel_eval failed")
    return w, dwdx, dwdr


def read_mcm(path: str) -> Deck:
    """Reads a .mcm deck with the library's reader (sph_deck_read) into a Deck."""
    from .deck import Material
    L = lib()
    L.sph_deck_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p,
                                ctypes.c_int]
    L.sph_deck_desc.argtypes = [ctypes.c_void_p]
    L.sph_deck_desc.restype = ctypes.POINTER(CDeck)
    L.sph_deck_title.argtypes = [ctypes.c_void_p]
    L.sph_deck_title.restype = ctypes.c_char_p
    L.sph_deck_free.argtypes = [ctypes.c_void_p]
    L.sph_deck_free.restype = None
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(512)
    rc = L.sph_deck_read(path.encode(), ctypes.byref(h), err, 512)
    if rc != 0:
        raise SphError(rc, err.value.decode())
    try:
        d = L.sph_deck_desc(h).contents
        n = d.np

        def arr(ptr, count, dtype):
            if not ptr:
                return None
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype).copy()

        mats = []
        for k in range(d.nummat):
            m = d.mat[k]
            mats.append(Material(model=m.model, eos=m.eos, visc_type=m.visc_type, rho=m.rho,
                                 av_l=m.av_l, av_q=m.av_q, surface_tension=m.surface_tension,
                                 mass=m.mass, h=m.h, eosinput=tuple(m.eosinput),
                                 strinput0=m.strinput0, av_alpha=m.av_alpha))
        deck = Deck(ndim=d.ndim, x=arr(d.x, 3 * n, np.float64).reshape(n, 3),
                    v=arr(d.v, 3 * n, np.float64).reshape(n, 3),
                    mat_id=arr(d.mat_id, n, np.int32), materials=mats,
                    color=arr(d.color, n, np.int32), dispbc=arr(d.dispbc, n, np.int32),
                    disctype=d.disctype, tcrit_opt=d.tcrit_opt, maxnbr=d.maxnbr,
                    surface_tension=d.surface_tension, rigid_body=d.rigid_body,
                    boundary=d.boundary, baseaccel=d.baseaccel, massopt=d.massopt,
                    init_h_opt=d.init_h_opt, init_rhoe=d.init_rhoe, veloc_opt=d.veloc_opt,
                    h_opt=d.h_opt, endtime=d.endtime, tssf=d.tssf, itss=d.itss, rb_id=d.rb_id,
                    rb_total_mass=d.rb_total_mass, rb_inertia=d.rb_inertia,
                    node_mass=arr(d.node_mass, n, np.float64),
                    node_rho=arr(d.node_rho, n, np.float64),
                    name=os.path.splitext(os.path.basename(path))[0])
        deck.boundary_code = np.array([[d.boundary_code[s][m] for m in range(3)] for s in range(2)],
                                      np.int32)
        deck.boundary_x = np.array([[d.boundary_x[s][m] for m in range(3)] for s in range(2)])
        deck.base_a = np.array([d.base_a[m] for m in range(3)])
        deck.meta["title"] = L.sph_deck_title(h).decode().rstrip()
        deck.n_fluid = int((deck.dispbc == 0).sum())
        return deck
    finally:
        L.sph_deck_free(h)


class Batch:
    """Many independent scenes advanced together (sph_batch_* entry points)."""

    def __init__(self, ndim: int, arm: int = ARM_FP64):
        L = lib()
        L.sph_batch_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.sph_batch_destroy.argtypes = [ctypes.c_void_p]
        L.sph_batch_destroy.restype = None
        L.sph_batch_last_error.argtypes = [ctypes.c_void_p]
        L.sph_batch_last_error.restype = ctypes.c_char_p
        L.sph_batch_add.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
        L.sph_batch_remove.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.sph_batch_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.sph_batch_status.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(CGlobals)]
        L.sph_batch_scene.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.sph_batch_scene.restype = ctypes.c_void_p
        L.sph_batch_live.argtypes = [ctypes.c_void_p]
        L.sph_scene_footprint.argtypes = [ctypes.c_void_p]
        L.sph_scene_footprint.restype = ctypes.c_longlong
        h = ctypes.c_void_p()
        rc = L.sph_batch_create(ndim, arm, ctypes.byref(h))
        if rc != 0:
            raise SphError(rc, "sph_batch_create failed")
        self._h = h
        self.ndim = ndim
        self.arm = arm
        self.decks = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().sph_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise SphError(rc, lib().sph_batch_last_error(self._h).decode())

    def add(self, deck: Deck, step_limit: int = 0) -> int:
        c, keep = deck.to_c()
        slot = ctypes.c_int(-1)
        err = ctypes.create_string_buffer(512)
        rc = lib().sph_batch_add(self._h, ctypes.byref(c), step_limit, ctypes.byref(slot), err, 512)
        if rc != 0:
            raise SphError(rc, err.value.decode())
        self.decks[slot.value] = deck
        return slot.value

    def remove(self, slot: int):
        self._chk(lib().sph_batch_remove(self._h, slot))
        self.decks.pop(slot, None)

    def step(self, nsteps: int) -> int:
        live = ctypes.c_int(0)
        self._chk(lib().sph_batch_step(self._h, nsteps, ctypes.byref(live)))
        return live.value

    def launch(self, nsteps: int):
        """sph_batch_launch: queues nsteps steps and returns at once."""
        L = lib()
        L.sph_batch_launch.argtypes = [ctypes.c_void_p, ctypes.c_int]
        self._chk(L.sph_batch_launch(self._h, nsteps))

    def sync(self) -> int:
        """sph_batch_sync: waits for the queued steps; returns the live count."""
        L = lib()
        L.sph_batch_sync.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        live = ctypes.c_int(0)
        self._chk(L.sph_batch_sync(self._h, ctypes.byref(live)))
        return live.value

    def status(self, slot: int) -> dict:
        g = CGlobals()
        self._chk(lib().sph_batch_status(self._h, slot, ctypes.byref(g)))
        return {k: getattr(g, k) for k in ("dt", "ptime", "timestep", "done", "error", "np")}

    def field(self, slot: int, f: int) -> np.ndarray:
        h = lib().sph_batch_scene(self._h, slot)
        n = self.decks[slot].np
        out = np.zeros((n, 3) if f in _VEC else n)
        rc = lib().sph_download(h, f, out.ctypes.data)
        if rc != 0:
            raise SphError(rc, lib().sph_last_error(h).decode())
        return out

    @property
    def live(self) -> int:
        return lib().sph_batch_live(self._h)

    def autotune(self, reps: int = -1) -> dict:
        """sph_batch_autotune: 0 fixes the current shapes (no online tuning),
        < 0 only reports."""
        L = lib()
        L.sph_batch_autotune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
        shape = (ctypes.c_int32 * 3)()
        self._chk(L.sph_batch_autotune(self._h, reps, shape))
        return {"cell_slots": shape[0], "cell_order": "row-major" if shape[1] == 1 else "morton",
                "tuned": bool(shape[2])}

    def finished(self):
        L = lib()
        L.sph_batch_finished.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        buf = np.zeros(max(len(self.decks), 1), np.int32)
        n = L.sph_batch_finished(self._h, buf.ctypes.data, len(buf))
        return buf[:n].tolist()


def footprint(deck: Deck) -> int:
    """Device bytes a scene of this deck will hold (sph_scene_footprint)."""
    L = lib()
    L.sph_scene_footprint.argtypes = [ctypes.c_void_p]
    L.sph_scene_footprint.restype = ctypes.c_longlong
    c, keep = deck.to_c()
    return int(L.sph_scene_footprint(ctypes.byref(c)))
