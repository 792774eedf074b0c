"""Writer of sph07 ``.mcm`` decks (the card layouts MSimulationData reads,
MSimulationData.cpp:244-1068), so generated scenes run through the same
deck reader and command-line driver as the reference's example decks.

Reading is done by the library (``native.read_mcm`` -> ``sph_deck_read``).
"""
from __future__ import annotations

import numpy as np

from .deck import Deck


def _f10(v: float) -> str:
    s = f"{v:10.3E}"
    if len(s) > 10:
        s = f"{v:10.2E}"
    return s


def _f20(v: float) -> str:
    return f"{v:20.12E}"


def write_mcm(deck: Deck, path: str, title: str | None = None, status_interval: int = 100) -> None:
    d = deck.normalized()
    n = d.np
    L = []
    L.append(f"{(title or d.name)[:78]:<78s}{2:2d}")
    L.append("* written by sph07_amd.mcm")
    L.append(f"{d.ndim:5d}{d.disctype:5d}{len(d.materials):10d}{n:10d}")                 # card 1
    L.append(f"{d.endtime:10.3E}{d.tssf:10.3E}{d.itss:10.3E}")                          # card 2
    L.append(f"{0.0:10.3E}{0:5d}{0.0:10.3E}{0:10d}{0:5d}{status_interval:10d}{0:5d}{0:5d}")
    init_v = 1 if np.any(d.v != 0.0) else 0
    L.append(f"{0:5d}{init_v:5d}{d.massopt:5d}{d.init_h_opt:5d}{d.init_rhoe:5d}")
    nthp = [1 if d.baseaccel and d.base_a[m] != 0.0 else 0 for m in range(3)]
    L.append(f"{0:5d}{d.tcrit_opt:5d}{d.veloc_opt:5d}{1 if d.boundary else 0:5d}{0:5d}"
             f"{nthp[0]:5d}{nthp[1]:5d}{nthp[2]:5d}{d.surface_tension:5d}{d.rigid_body:5d}")
    L.append(f"{d.h_opt:5d}{1:5d}{d.maxnbr:5d}")                                         # card 6
    L.append("")                                                                         # card 7
    for k, m in enumerate(d.materials, 1):
        L.append(f"{k:5d}{m.model:5d}{_f10(m.rho)}{m.eos:5d}{0:5d}{_f10(0.0)}{m.visc_type:5d}"
                 f"{_f10(m.av_q)}{_f10(m.av_l)}{_f10(m.surface_tension)}")
        L.append(f"{_f10(m.mass)}{_f10(m.h)}")
        L.append("Fluid")
        L.append(f"{_f10(m.strinput0)}")
        L.extend([""] * 5)
        if m.eos:
            L.append("EOS")
            if m.eos == 29:
                L.append(f"{_f10(m.eosinput[0])}")
            elif m.eos == 28:
                L.append(f"{_f10(m.eosinput[0])}{_f10(m.eosinput[1])}")
    for i in range(n):
        xs = "".join(_f20(d.x[i, k]) for k in range(d.ndim))
        L.append(f"{i + 1:8d}{int(d.dispbc[i]):5d}{xs}{int(d.mat_id[i]):7d}{int(d.color[i]):7d}")
    if d.massopt == 1 or d.init_rhoe == 1:
        for i in range(n):
            mm = d.node_mass[i] if d.node_mass is not None else 0.0
            rr = d.node_rho[i] if d.node_rho is not None else 0.0
            L.append(f"{i + 1:8d}{_f20(mm)}{_f20(0.0)}{_f20(rr)}{_f20(0.0)}")
    if init_v:
        for i in range(n):
            L.append(f"{i + 1:8d}" + "".join(f"{d.v[i, k]:10.3E}" for k in range(d.ndim)))
    if d.boundary:
        L.append("".join(f"{int(d.boundary_code[s, m]):10d}" for m in range(d.ndim) for s in range(2)))
        L.append("".join(_f10(d.boundary_x[s, m]) for m in range(d.ndim) for s in range(2)))
    if d.baseaccel:
        L.append("".join(_f20(d.base_a[m]) for m in range(3)))
    L.append(f"{0.0:10.3E}")                                                              # plot card
    if d.rigid_body:
        L.append(f"{d.rb_id:5d}{_f10(d.rb_total_mass)}{_f10(d.rb_inertia)}{_f10(0.0)}{_f10(0.0)}")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
