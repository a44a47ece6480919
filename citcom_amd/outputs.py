"""Readers for the reference's text output files (Output.c:52-160), so that a
run of the reference program (with or without the GPU drop-in) can be
compared field by field.

  <datafile>.coord.<rank>            nno, then x y z per node
  <datafile>.ave.<rank>.<step>       noz timesteps time Nu_top Nu_bot ...,
                                     then per z level T vrms Vi Rho F f C Tadi
  <datafile>.temp.<rank>.<step>      nno timesteps time, then T Vz adv-flux
  <datafile>.velo.<rank>.<step>      nno timesteps time, then V1 V2 V3

Node order is the reference's (node = k + noz*(i + nox*j), 1-based in the
files' line order), the same as this package's 0-based order; a run on
several MPI ranks writes one file per rank over its local nodes, which
``assemble`` puts back on the global grid.
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


def assemble(fields: list[np.ndarray], dims: tuple[int, int, int], nproc: tuple[int, int, int]) -> np.ndarray:
    """Per-rank node rows (rank order) -> global node rows [nno][ncol];
    shared face nodes must agree between ranks."""
    from .deck import rank_blocks
    nox, noy, noz = dims
    ncol = fields[0].shape[1]
    out = np.full((noy, nox, noz, ncol), np.nan)
    for (me, sy, sx, sz), vals in zip(rank_blocks(dims, nproc), fields):
        blk = out[sy, sx, sz]
        v = vals.reshape(blk.shape)
        if not np.all(np.isnan(blk) | (np.abs(blk - v) <= 1e-6 * np.maximum(np.abs(v), 1e-30))):
            raise ValueError(f"rank {me}: shared nodes disagree")
        out[sy, sx, sz] = v
    if np.isnan(out).any():
        raise ValueError("ranks do not cover the grid")
    return out.reshape(-1, ncol)
