"""Scene deck: the content of one sph07 ``.mcm`` input deck, as arrays.

The fields are the cards ``MSimulationData::GetInput`` reads
(MSimulationData.cpp:186-219): control cards 1-6 (:244-323), material cards
(:441-541, with ``MMaterial::CheckDefaults`` MMaterial.cpp:71-81 already
applied), node cards (:655-763), velocities (:765-881), boundary planes
(:883-1003), base acceleration (:1005-1019) and the rigid-body card
(:1058-1068).  ``to_c()`` packs a deck into the ``sph_scene_desc`` struct of
``include/sph07_amd.h``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


@dataclass
class Material:
    """One material card pair (MSimulationData.cpp:454-459, 504-536)."""
    model: int = 9            # 9 = fluid
    eos: int = 29             # 29 linear (Morris 1997), 28 Tait (Monaghan 1994)
    visc_type: int = 1        # 1 Morris viscosity; 2 adds the Monaghan alpha term
    rho: float = 1000.0       # rho0
    av_l: float = 1.0e-6      # kinematic viscosity nu
    av_q: float = 1.5
    surface_tension: float = 0.0
    mass: float = 0.0         # total material mass (massopt 0)
    h: float = 0.0            # smoothing length
    eosinput: tuple = (0.0, 0.0, 0.0, 0.0)
    strinput0: float = 1.0e7  # pressure cut-off at initialisation
    av_alpha: float = 0.0     # Monaghan alpha (visc_type 2)

    def check_defaults(self) -> "Material":
        """MMaterial::CheckDefaults (MMaterial.cpp:71-81)."""
        if self.visc_type == 0:
            self.visc_type = 1
        if self.av_l == 0.0:
            self.av_l = 0.06
        if self.av_q == 0.0:
            self.av_q = 1.5
        return self


class CMaterial(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32), ("eos", ctypes.c_int32), ("visc_type", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("rho", ctypes.c_double), ("av_l", ctypes.c_double), ("av_q", ctypes.c_double),
        ("surface_tension", ctypes.c_double), ("mass", ctypes.c_double), ("h", ctypes.c_double),
        ("eosinput", ctypes.c_double * 4), ("strinput0", ctypes.c_double),
        ("av_alpha", ctypes.c_double),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class CDeck(ctypes.Structure):
    """Byte layout of ``sph_scene_desc`` (include/sph07_amd.h)."""
    _fields_ = [
        ("ndim", ctypes.c_int32), ("disctype", ctypes.c_int32), ("tcrit_opt", ctypes.c_int32),
        ("maxnbr", ctypes.c_int32),
        ("surface_tension", ctypes.c_int32), ("rigid_body", ctypes.c_int32),
        ("boundary", ctypes.c_int32), ("baseaccel", ctypes.c_int32),
        ("massopt", ctypes.c_int32), ("init_h_opt", ctypes.c_int32),
        ("init_rhoe", ctypes.c_int32), ("nummat", ctypes.c_int32),
        ("np", ctypes.c_int32), ("rb_id", ctypes.c_int32), ("veloc_opt", ctypes.c_int32),
        ("h_opt", ctypes.c_int32),
        ("boundary_code", (ctypes.c_int32 * 3) * 2),
        ("endtime", ctypes.c_double), ("tssf", ctypes.c_double), ("itss", ctypes.c_double),
        ("boundary_x", (ctypes.c_double * 3) * 2),
        ("base_a", ctypes.c_double * 3),
        ("rb_total_mass", ctypes.c_double), ("rb_inertia", ctypes.c_double),
        ("mat", ctypes.POINTER(CMaterial)),
        ("x", _dp), ("v", _dp), ("mat_id", _ip), ("color", _ip), ("dispbc", _ip),
        ("node_mass", _dp), ("node_h", _dp), ("node_rho", _dp),
    ]


@dataclass
class Deck:
    ndim: int
    x: np.ndarray                       # (np, 3) float64
    mat_id: np.ndarray                  # (np,) int32, 1-based
    materials: List[Material]
    v: Optional[np.ndarray] = None      # (np, 3) float64
    color: Optional[np.ndarray] = None  # (np,) int32
    dispbc: Optional[np.ndarray] = None # (np,) int32; 7 fixed wall, 8 moving-lid wall
    disctype: int = 2
    tcrit_opt: int = 3
    maxnbr: int = 40
    surface_tension: int = 0
    rigid_body: int = 0
    boundary: int = 0
    baseaccel: int = 0
    massopt: int = 0
    init_h_opt: int = 0
    init_rhoe: int = 0
    veloc_opt: int = 0
    h_opt: int = 0
    boundary_code: np.ndarray = field(default_factory=lambda: np.zeros((2, 3), np.int32))
    boundary_x: np.ndarray = field(default_factory=lambda: np.zeros((2, 3)))
    base_a: np.ndarray = field(default_factory=lambda: np.zeros(3))
    endtime: float = 1.0e30
    tssf: float = 0.8
    itss: float = 0.0
    rb_id: int = 0
    rb_total_mass: float = 0.0
    rb_inertia: float = 0.0
    node_mass: Optional[np.ndarray] = None
    node_rho: Optional[np.ndarray] = None
    name: str = "scene"
    n_fluid: int = -1                   # fluid particles (the rest are walls / bodies)
    meta: dict = field(default_factory=dict)

    @property
    def np(self) -> int:
        return int(self.x.shape[0])

    def normalized(self) -> "Deck":
        """Contiguous arrays of the dtypes the C structs expect."""
        n = self.np
        self.x = np.ascontiguousarray(self.x, dtype=np.float64).reshape(n, 3)
        self.v = (np.zeros((n, 3)) if self.v is None
                  else np.ascontiguousarray(self.v, dtype=np.float64).reshape(n, 3))
        self.mat_id = np.ascontiguousarray(self.mat_id, dtype=np.int32).reshape(n)
        self.color = (np.ones(n, np.int32) if self.color is None
                      else np.ascontiguousarray(self.color, dtype=np.int32).reshape(n))
        self.dispbc = (np.zeros(n, np.int32) if self.dispbc is None
                       else np.ascontiguousarray(self.dispbc, dtype=np.int32).reshape(n))
        self.boundary_code = np.ascontiguousarray(self.boundary_code, dtype=np.int32).reshape(2, 3)
        self.boundary_x = np.ascontiguousarray(self.boundary_x, dtype=np.float64).reshape(2, 3)
        self.base_a = np.ascontiguousarray(self.base_a, dtype=np.float64).reshape(3)
        if self.node_mass is not None:
            self.node_mass = np.ascontiguousarray(self.node_mass, dtype=np.float64)
        if self.node_rho is not None:
            self.node_rho = np.ascontiguousarray(self.node_rho, dtype=np.float64)
        return self

    def to_c(self):
        """Returns (CDeck, keepalive) -- keep ``keepalive`` referenced while
        the struct is in use."""
        d = self.normalized()
        mats = (CMaterial * len(d.materials))()
        for k, m in enumerate(d.materials):
            mats[k].model, mats[k].eos, mats[k].visc_type = m.model, m.eos, m.visc_type
            mats[k].rho, mats[k].av_l, mats[k].av_q = m.rho, m.av_l, m.av_q
            mats[k].surface_tension, mats[k].mass, mats[k].h = m.surface_tension, m.mass, m.h
            for q in range(4):
                mats[k].eosinput[q] = m.eosinput[q]
            mats[k].strinput0, mats[k].av_alpha = m.strinput0, m.av_alpha
        c = CDeck()
        for name in ("ndim", "disctype", "tcrit_opt", "maxnbr", "surface_tension", "rigid_body",
                     "boundary", "baseaccel", "massopt", "init_h_opt", "init_rhoe", "rb_id",
                     "veloc_opt", "h_opt"):
            setattr(c, name, int(getattr(d, name)))
        c.nummat = len(d.materials)
        c.np = d.np
        for s in range(2):
            for m in range(3):
                c.boundary_code[s][m] = int(d.boundary_code[s, m])
                c.boundary_x[s][m] = float(d.boundary_x[s, m])
        for m in range(3):
            c.base_a[m] = float(d.base_a[m])
        c.endtime, c.tssf, c.itss = d.endtime, d.tssf, d.itss
        c.rb_total_mass, c.rb_inertia = d.rb_total_mass, d.rb_inertia
        c.mat = mats
        c.x = d.x.ctypes.data_as(_dp)
        c.v = d.v.ctypes.data_as(_dp)
        c.mat_id = d.mat_id.ctypes.data_as(_ip)
        c.color = d.color.ctypes.data_as(_ip)
        c.dispbc = d.dispbc.ctypes.data_as(_ip)
        c.node_mass = d.node_mass.ctypes.data_as(_dp) if d.node_mass is not None else None
        c.node_h = None
        c.node_rho = d.node_rho.ctypes.data_as(_dp) if d.node_rho is not None else None
        keep = (mats, d.x, d.v, d.mat_id, d.color, d.dispbc, d.node_mass, d.node_rho)
        return c, keep
