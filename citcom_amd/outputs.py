"""Readers for the reference's text output files (Output.c:52-160), so that a
run of the reference program (with or without the GPU drop-in) can be
compared field by field.

  <datafile>.coord.<rank>            nno, then x y z per node
  <datafile>.ave.<rank>.<step>       noz timesteps time Nu_top Nu_bot ...,
                                     then per z level T vrms Vi Rho F f C Tadi
  <datafile>.temp.<rank>.<step>      nno timesteps time, then T Vz adv-flux
  <datafile>.velo.<rank>.<step>      nno timesteps time, then V1 V2 V3

Node order is the reference's (node = k + noz*(i + nox*j), 1-based in the
files' line order), the same as this package's 0-based order.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class Ave:
    noz: int
    timesteps: int
    time: float
    nu_top: float
    nu_bot: float
    profiles: np.ndarray   # [noz][8]: T, vrms, Vi, Rho, F, f, C, Tadi


@dataclasses.dataclass
class NodeField:
    nno: int
    timesteps: int
    time: float
    values: np.ndarray     # [nno][ncol]


def read_ave(path: str) -> Ave:
    """Output.c:86-88"""
    with open(path) as fh:
        head = fh.readline().split()
        rows = np.loadtxt(fh, ndmin=2)
    return Ave(int(head[0]), int(head[1]), float(head[2]), float(head[3]), float(head[4]), rows)


def _read_nodes(path: str) -> NodeField:
    with open(path) as fh:
        head = fh.readline().split()
        rows = np.loadtxt(fh, ndmin=2)
    nno = int(head[0])
    if rows.shape[0] != nno:
        raise ValueError(f"{path}: {rows.shape[0]} rows for {nno} nodes")
    return NodeField(nno, int(head[1]), float(head[2]), rows)


def read_temp(path: str) -> NodeField:
    """Output.c:93-120: T, V3, heatflux_adv per node"""
    return _read_nodes(path)


def read_velo(path: str) -> NodeField:
    """Output.c:123-128: V1, V2, V3 per node"""
    return _read_nodes(path)


def read_coord(path: str) -> np.ndarray:
    """Output.c:66-78: x y z per node"""
    with open(path) as fh:
        nno = int(fh.readline().split()[0])
        rows = np.loadtxt(fh, ndmin=2)
    if rows.shape[0] != nno:
        raise ValueError(f"{path}: {rows.shape[0]} rows for {nno} nodes")
    return rows
