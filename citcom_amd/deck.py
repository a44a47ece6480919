"""CitcomCU input decks: the reference's ``key=value`` format, read and written.

The reference parses decks with ``input_int/float/string`` (Parsing.c:268-803);
keys may share a line, ``#`` starts a comment, strings may be quoted.  This
module reads the same files, supplies the defaults the reference applies to
the keys used by the Stokes/energy hot path (Instructions.c:414-566,
1020-1120; Advection_diffusion.c:60-90; Viscosity_structures.c:57-326) and
writes decks for the workloads this package is measured on.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

_TOKEN = re.compile(r'([A-Za-z_][A-Za-z0-9_]*)\s*=\s*("[^"]*"|\S+)')


def parse_deck(text: str) -> dict[str, str]:
    """Parse deck text into a key -> raw string map (last assignment wins)."""
    out: dict[str, str] = {}
    for line in text.splitlines():
        line = line.split("#", 1)[0]
        for key, val in _TOKEN.findall(line):
            out[key] = val.strip('"')
    return out


def read_deck(path: str) -> dict[str, str]:
    with open(path) as fh:
        return parse_deck(fh.read())


def write_deck(path: str, keys: dict[str, object]) -> None:
    with open(path, "w") as fh:
        for k, v in keys.items():
            if isinstance(v, (list, tuple)):
                v = ",".join(str(x) for x in v)
            elif isinstance(v, str) and k in _QUOTED:
                v = f'"{v}"'
            fh.write(f"{k}={v}\n")


_QUOTED = {"datafile", "oldfile", "use_scratch"}


def _f(d, k, default):
    return float(d.get(k, default))


def _i(d, k, default):
    return int(float(d.get(k, default)))


def _b(d, k, default):
    v = str(d.get(k, default)).lower()
    return 1 if v in ("1", "on", "true", "yes") else 0


def _fv(d, k, n, default):
    raw = d.get(k)
    if raw is None:
        return [default] * n
    vals = [float(x) for x in raw.split(",") if x != ""]
    return (vals + [vals[-1]] * n)[:n]


def axis_coordinates(layer: float, nnodes: int, seg_x=None, seg_n=None) -> np.ndarray:
    """Node coordinates along one axis, as node_locations computes them
    (Nodal_mesh.c:78-160): float32 running sums of a constant step per grid
    segment, with the end node pinned to the box size."""
    x = np.zeros(nnodes + 1, dtype=np.float32)  # 1-based like the reference
    dx = np.float32(np.float32(layer) / np.float32(nnodes - 1))
    x[1] = 0.0
    x[nnodes] = np.float32(layer)
    for i in range(2, nnodes):
        x[i] = np.float32(x[i - 1] + dx)
    if seg_x is not None and len(seg_x) > 1:
        dxx = [None] + [np.float32(np.float32(seg_x[j] - seg_x[j - 1]) / np.float32(seg_n[j] - seg_n[j - 1]))
                        for j in range(1, len(seg_x))]
        j = 1
        for i in range(2, nnodes):
            if i <= seg_n[j]:
                x[i] = np.float32(x[i - 1] + dxx[j])
            if i == seg_n[j]:
                j += 1
    return x[1:].copy()


@dataclass
class Problem:
    """The subset of ``struct All_variables`` control data that the hot path
    reads (global_defs.h:931-1102), filled from a deck with the reference's
    defaults."""
    levels: int
    mgunit: tuple[int, int, int]          # x, y, z
    layer: tuple[float, float, float]     # box size x, y, z
    coords: tuple[np.ndarray, np.ndarray, np.ndarray]
    topvbc: int = 0
    botvbc: int = 0
    toptbc: int = 1
    bottbc: int = 1
    toptbcval: float = 0.0
    bottbcval: float = 1.0
    accuracy: float = 1e-4
    p_iterations: int = 100
    v_steps_low: int = 250
    v_steps_high: int = 1
    down_heavy: int = 1
    up_heavy: int = 1
    mg_cycle: int = 1
    aug_lagr: int = 0
    aug_number: float = 0.0
    precond: int = 0
    rheol: int = 0
    tdepv: int = 0
    visc_update: int = 0
    num_mat: int = 1
    visc0: list = field(default_factory=lambda: [1.0] * 4)
    viscE: list = field(default_factory=lambda: [0.0] * 4)
    viscT: list = field(default_factory=lambda: [0.0] * 4)
    z_layer: list = field(default_factory=lambda: [0.0] * 4)
    vmin_on: int = 0
    vmax_on: int = 0
    visc_min: float = 1e-20
    visc_max: float = 1e20
    rayleigh: float = 1.0
    finetunedt: float = 0.9
    temp_iterations: int = 2
    gamma: float = 0.5
    Q0: float = 0.0
    fixed_timestep: float = 0.0
    maxstep: int = 1000
    restart: int = 0
    num_perturbations: int = 0
    perturbmag: list = field(default_factory=list)
    perturbk: list = field(default_factory=list)
    smooth_cycles: int = 0

    @property
    def levmax(self) -> int:
        return self.levels - 1

    def dims(self, lev: int | None = None) -> tuple[int, int, int]:
        """(nox, noy, noz) node counts at level lev (Instructions.c:635-650)."""
        lev = self.levmax if lev is None else lev
        f = 1 << lev
        return (self.mgunit[0] * f + 1, self.mgunit[1] * f + 1, self.mgunit[2] * f + 1)


def problem_from_deck(d: dict[str, str]) -> Problem:
    levels = _i(d, "levels", 1)
    mg = (_i(d, "mgunitx", 1), _i(d, "mgunity", 1), _i(d, "mgunitz", 1))
    layer = (_f(d, "dimenx", 1.0), _f(d, "dimeny", 1.0), _f(d, "dimenz", 1.0))
    f = 1 << (levels - 1)
    nn = (mg[0] * f + 1, mg[1] * f + 1, mg[2] * f + 1)
    coords = []
    for ax, key, nkey, lkey in ((0, "xx", "nx", "x_grid_layers"), (1, "yy", "ny", "y_grid_layers"),
                                (2, "zz", "nz", "z_grid_layers")):
        nl = _i(d, lkey, 1)
        sx = sn = None
        if nl > 1:
            sx = [float(v) for v in d[key].split(",")][:nl]
            sn = [int(float(v)) for v in d[nkey].split(",")][:nl]
        coords.append(axis_coordinates(layer[ax], nn[ax], sx, sn))
    num_mat = _i(d, "num_mat", 1)
    rheol = _i(d, "rheol", 3)
    if "z_layer" in d:
        z_layer = _fv(d, "z_layer", 4, 0.0)
    else:
        # Viscosity_structures.c:129-141: without z_layer the four layer bases
        # are z_lith, z_410, z_lmantle and the bottom
        if num_mat != 4:
            raise ValueError("either use z_layer for non dim layer depths, or set num_mat to four")
        z_layer = [_f(d, "z_lith", 0.0), _f(d, "z_410", 1.0), _f(d, "z_lmantle", 1.0), 0.0]
    p = Problem(
        levels=levels, mgunit=mg, layer=layer, coords=tuple(coords),
        topvbc=_i(d, "topvbc", 0), botvbc=_i(d, "botvbc", 0),
        toptbc=_i(d, "toptbc", 1), bottbc=_i(d, "bottbc", 1),
        toptbcval=_f(d, "toptbcval", 0.0), bottbcval=_f(d, "bottbcval", 1.0),
        accuracy=_f(d, "accuracy", 1e-4), p_iterations=_i(d, "piterations", 100),
        v_steps_low=_i(d, "vlowstep", 250), v_steps_high=_i(d, "vhighstep", 1),
        down_heavy=_i(d, "down_heavy", 1), up_heavy=_i(d, "up_heavy", 1),
        mg_cycle=_i(d, "mg_cycle", 1), aug_lagr=_b(d, "aug_lagr", "off"),
        aug_number=_f(d, "aug_number", 0.0), precond=_b(d, "precond", "off"),
        rheol=rheol, tdepv=_b(d, "TDEPV", "on"), visc_update=_b(d, "VISC_UPDATE", "on"),
        num_mat=num_mat, visc0=_fv(d, "visc0", 4, 1.0), viscE=_fv(d, "viscE", 4, 0.0),
        viscT=_fv(d, "viscT", 4, 0.0), z_layer=z_layer,
        vmin_on=_b(d, "VMIN", "off"), vmax_on=_b(d, "VMAX", "off"),
        visc_min=_f(d, "visc_min", 1e-20), visc_max=_f(d, "visc_max", 1e20),
        rayleigh=_f(d, "rayleigh", 1.0), finetunedt=_f(d, "finetunedt", 0.9),
        temp_iterations=_i(d, "adv_sub_iterations", 2), gamma=0.5,
        Q0=_f(d, "Q0", 0.0), fixed_timestep=_f(d, "fixed_timestep", 0.0),
        maxstep=_i(d, "maxstep", 1000), restart=_i(d, "restart", 0),
        num_perturbations=_i(d, "num_perturbations", 0),
        perturbmag=_fv(d, "perturbmag", max(_i(d, "num_perturbations", 0), 1), 0.0),
        perturbk=_fv(d, "perturbk", max(_i(d, "num_perturbations", 0), 1), 0.0),
        smooth_cycles=_i(d, "visc_smooth_cycles", 0),
    )
    return p


# --------------------------------------------------------------------------
# the measured workloads

COMMON = {
    "use_scratch": "local",
    "storage_spacing": 100000,
    "Solver": "multigrid", "node_assemble": 1,
    "Geometry": "cart3d", "Problem": "convection",
    "composition": 0, "rayleigh_comp": 0, "Q0": 0, "Q0_enriched": 0,
    "markers_per_ele": 0, "visc_heating": 0, "adi_heating": 0,
    "nprocx": 1, "nprocz": 1, "nprocy": 1,
    "topvbc": 0, "botvbc": 0,
    "toptbc": 1, "bottbc": 1, "toptbcval": 0.0, "bottbcval": 1.0,
    "aug_lagr": "on", "aug_number": 1.0e3, "precond": "on",
    "mg_cycle": 1, "down_heavy": 3, "up_heavy": 3, "vlowstep": 20, "vhighstep": 3,
    "piterations": 375, "accuracy": 1.0e-6,
    "adv_sub_iterations": 2,
    "Viscosity": "system", "visc_smooth_cycles": 1, "SDEPV": "off",
    "num_perturbations": 0,
    "DESCRIBE": "off", "BEGINNER": "off", "VERBOSE": "off", "verbose": "off",
    "COMPRESS": "off", "see_convergence": 1,
}


def _grid(dimen, mg, levels):
    f = 1 << (levels - 1)
    out = {}
    for ax, (key, nkey, lkey, dkey) in enumerate((("xx", "nx", "x_grid_layers", "dimenx"),
                                                 ("yy", "ny", "y_grid_layers", "dimeny"),
                                                 ("zz", "nz", "z_grid_layers", "dimenz"))):
        out[dkey] = dimen[ax]
        out[lkey] = 2
        out[key] = f"0,{dimen[ax]}"
        out[nkey] = f"1,{mg[ax] * f + 1}"
    return out


def _iso(ra, perturb=True):
    d = {"rayleigh": ra, "rheol": 0, "TDEPV": "off", "VISC_UPDATE": "off", "update_every_steps": 1,
         "num_mat": 4, "visc0": "1,1,1,1", "viscE": "0,0,0,0", "viscT": "0,0,0,0", "viscZ": "0,0,0,0",
         "VMIN": "off", "VMAX": "off"}
    if perturb:
        d.update({"num_perturbations": 1, "perturbmag": 0.01, "perturbk": 1.0, "perturbl": 1.0, "perturbm": 0.0})
    return d


def _tdep(ra, contrast, vmin=None, vmax=None, num_mat=1, visc0="1", z_layer=None):
    d = {"rayleigh": ra, "rheol": 3, "TDEPV": "on", "VISC_UPDATE": "on", "update_every_steps": 1,
         "num_mat": num_mat, "visc0": visc0, "viscE": ",".join([repr(math.log(contrast))] * num_mat),
         "viscT": ",".join(["0"] * num_mat), "viscZ": ",".join(["0"] * num_mat),
         "VMIN": "off", "VMAX": "off"}
    if vmin is not None:
        d.update({"VMIN": "on", "visc_min": vmin})
    if vmax is not None:
        d.update({"VMAX": "on", "visc_max": vmax})
    # without z_layer the reference requires num_mat = 4 (Viscosity_structures.c:130-134)
    d["z_layer"] = z_layer if z_layer is not None else ",".join(["0"] * num_mat)
    return d


def workload(name: str) -> dict[str, object]:
    """Deck keys of a named workload.

    tiny / tiny_tdep  8x8x4 elements, 3 levels (test size)
    tiny2  4x4x2 elements, 2 levels: the shortest multigrid hierarchy
    tiny_layer  the same grid with two viscosity layers, contrast 1e5, clips
    iso16  16^3 elements, 3 levels, isoviscous Ra=1e4, 200 steps
    tdep64  64^3 elements, 5 levels, Ra=1e6, eta = exp(-ln(1e4) T), 500 steps
    tdep128  128x128x64 elements in a 2x2x1 box, 6 levels, Ra=1e7,
          eta = exp(-ln(1e4) T) clipped to [1e-4, 1], CFL 0.5
    layer256  256x256x128 elements, 7 levels, Ra=1e7, contrast 1e5, top layer x100
    learn16  16^3 elements, 3 levels, Ra=1e5, eta = exp(-ln(1e3) T): the
          training and test problems of the learned warm start
    busse1a / busse2  the reference's own example decks
          (examples/Busse1993/case1a.input, case2.input): 32x16x32 and 32^3
          elements, 5 levels, rigid top and bottom; case 2 with the
          Arrhenius law (rheol 11) on a z grid refined at the bottom, started
          from the example's case2-tic.temp.0.0 (restart = 1)

    ``name@acc`` is the same workload with ``accuracy=acc`` (e.g. iso16@1e-8).
    """
    if "@" in name:
        base, acc = name.split("@", 1)
        d = workload(base)
        d["accuracy"] = float(acc)
        return d
    d = dict(COMMON)
    if name == "tiny":
        d.update(_grid((1.0, 1.0, 0.5), (2, 2, 1), 3))
        d.update({"mgunitx": 2, "mgunity": 2, "mgunitz": 1, "levels": 3, "maxstep": 5, "finetunedt": 0.75})
        d.update(_iso(1e5))
    elif name == "tiny2":
        d.update(_grid((1.0, 1.0, 0.5), (2, 2, 1), 2))
        d.update({"mgunitx": 2, "mgunity": 2, "mgunitz": 1, "levels": 2, "maxstep": 5, "finetunedt": 0.75})
        d.update(_tdep(1e5, 100.0))
        d.update({"num_perturbations": 1, "perturbmag": 0.01, "perturbk": 1.0, "perturbl": 1.0, "perturbm": 0.0})
    elif name == "tiny_tdep":
        d.update(_grid((1.0, 1.0, 0.5), (2, 2, 1), 3))
        d.update({"mgunitx": 2, "mgunity": 2, "mgunitz": 1, "levels": 3, "maxstep": 5, "finetunedt": 0.75})
        d.update(_tdep(1e5, 100.0))
        d.update({"num_perturbations": 1, "perturbmag": 0.01, "perturbk": 1.0, "perturbl": 1.0, "perturbm": 0.0})
    elif name == "tiny_layer":
        # the viscosity features of the 128x128x64 and 256x256x128 workloads at
        # test size: two material layers (top 0.1 x100), contrast 1e5, both clips
        d.update(_grid((1.0, 1.0, 0.5), (2, 2, 1), 3))
        d.update({"mgunitx": 2, "mgunity": 2, "mgunitz": 1, "levels": 3, "maxstep": 5, "finetunedt": 1.0})
        d.update(_tdep(1e6, 1e5, vmin=1e-3, vmax=10.0, num_mat=2, visc0="100,1", z_layer="0.4,0"))
        d.update({"num_perturbations": 1, "perturbmag": 0.01, "perturbk": 1.0, "perturbl": 1.0, "perturbm": 0.0})
    elif name == "iso16":
        d.update(_grid((1.0, 1.0, 1.0), (4, 4, 4), 3))
        d.update({"mgunitx": 4, "mgunity": 4, "mgunitz": 4, "levels": 3, "maxstep": 200, "finetunedt": 0.75})
        d.update(_iso(1e4))
    elif name == "tdep64":
        d.update(_grid((1.0, 1.0, 1.0), (4, 4, 4), 5))
        d.update({"mgunitx": 4, "mgunity": 4, "mgunitz": 4, "levels": 5, "maxstep": 500, "finetunedt": 0.75})
        d.update(_tdep(1e6, 1e4))
    elif name == "tdep128":
        d.update(_grid((2.0, 2.0, 1.0), (4, 4, 2), 6))
        d.update({"mgunitx": 4, "mgunity": 4, "mgunitz": 2, "levels": 6, "maxstep": 10, "finetunedt": 1.0})
        d.update(_tdep(1e7, 1e4, vmin=1e-4, vmax=1.0))
    elif name == "learn16":
        d.update(_grid((1.0, 1.0, 1.0), (4, 4, 4), 3))
        d.update({"mgunitx": 4, "mgunity": 4, "mgunitz": 4, "levels": 3, "maxstep": 1, "finetunedt": 0.75})
        d.update(_tdep(1e5, 1e3))
    elif name == "layer256":
        d.update(_grid((2.0, 2.0, 1.0), (4, 4, 2), 7))
        d.update({"mgunitx": 4, "mgunity": 4, "mgunitz": 2, "levels": 7, "maxstep": 5, "finetunedt": 1.0})
        d.update(_tdep(1e7, 1e5, num_mat=2, visc0="100,1", z_layer="0.9,0"))
    elif name in ("busse1a", "busse2"):
        d.update(_busse(name))
    else:
        raise KeyError(name)
    return d


def _busse(name: str) -> dict[str, object]:
    """Keys of examples/Busse1993/case1a.input and case2.input (the parts the
    hot path reads; output, sphere and phase-change keys left out)."""
    d = {"rayleigh": 3.0e4 if name == "busse1a" else 2.0e4,
         "maxstep": 20000 if name == "busse1a" else 10000,
         "nprocx": 1, "nprocy": 1, "nprocz": 1, "mgunitx": 2, "mgunitz": 2, "mgunity": 1 if name == "busse1a" else 2,
         "levels": 5, "num_mat": 4, "visc0": "1,1,1,1", "viscZ": "5e-6,5e-6,5e-6,5e-6",
         "z_lmantle": 0.76655052, "z_410": 0.857143, "z_lith": 0.9651568,
         "VMIN": "off", "visc_min": 5.0e-2, "VMAX": "off", "visc_max": 2.0e4, "visc_smooth_cycles": 1,
         "topvbc": 1, "topvbxval": 0.0, "topvbyval": 0.0, "botvbc": 1, "botvbxval": 0.0, "botvbyval": 0.0,
         "toptbc": 1, "bottbc": 1, "toptbcval": 0.0, "bottbcval": 1.0,
         "num_perturbations": 1, "perturbmag": 0.001, "perturbk": 1.0, "perturbl": 6.0, "perturbm": 0.0,
         "aug_lagr": "on", "aug_number": 1.0e3, "precond": "on", "mg_cycle": 1, "down_heavy": 3, "up_heavy": 3,
         "vlowstep": 20, "vhighstep": 3, "piterations": 375, "accuracy": 1.0e-3, "adv_sub_iterations": 2,
         "finetunedt": 0.75}
    if name == "busse1a":
        d.update({"restart": 0, "dimenx": 1.0079, "dimeny": 0.6283, "dimenz": 1.0,
                  "x_grid_layers": 2, "xx": "0,1.0079", "nx": "1,33", "y_grid_layers": 2, "yy": "0,0.6283",
                  "ny": "1,17", "z_grid_layers": 2, "zz": "0.0,1.0", "nz": "1,33",
                  "rheol": 0, "TDEPV": "off", "VISC_UPDATE": "off", "update_every_steps": 1,
                  "viscE": "0,0,0,0", "viscT": "273,273,273,273"})
    else:
        d.update({"restart": 1, "restart_timesteps": 0, "dimenx": 1.0, "dimeny": 1.0, "dimenz": 1.0,
                  "x_grid_layers": 2, "xx": "0,1.0", "nx": "1,33", "y_grid_layers": 2, "yy": "0,1.0", "ny": "1,33",
                  "z_grid_layers": 4, "zz": "0.0,0.04,0.12,1.0", "nz": "1,3,6,33",
                  "rheol": 11, "TDEPV": "on", "VISC_UPDATE": "on", "update_every_steps": 1,
                  "viscE": "74.357912,74.357912,74.357912,74.357912", "viscT": "4.507123,4.507123,4.507123,4.507123"})
    return d


def read_restart_temperature(path: str) -> np.ndarray:
    """T of a restart file as Convection.c:473-490 reads it: the first value
    of each node line (%g into float), clamped to T >= 0."""
    with open(path) as fh:
        head = fh.readline().split()
        vals = [line.split()[0] for line in fh if line.strip()]
    nno = int(head[0])
    if len(vals) != nno:
        raise ValueError(f"{path}: {len(vals)} node lines for {nno} nodes")
    T = np.array(vals, dtype=np.float32)
    return np.maximum(T, np.float32(0.0))


def seeded_temperature(p: Problem, seed: int, amplitude: float = 0.01) -> np.ndarray:
    """T = 1 - z + amplitude * U(-1, 1) on interior nodes, boundary rows exact.

    This is the seeded initial field of the 64^3 and larger workloads; the
    reference reads it through its restart reader (Convection.c:473-490)."""
    nox, noy, noz = p.dims()
    z = p.coords[2]
    rng = np.random.default_rng(seed)
    T = np.empty((noy, nox, noz), dtype=np.float32)
    T[:] = (1.0 - z)[None, None, :].astype(np.float32)
    T[:, :, 1:-1] += (amplitude * rng.uniform(-1.0, 1.0, size=(noy, nox, noz - 2))).astype(np.float32)
    T[:, :, 0] = p.bottbcval
    T[:, :, -1] = p.toptbcval
    return np.maximum(T.reshape(-1), 0.0).astype(np.float32)


def rank_blocks(dims: tuple[int, int, int], nproc: tuple[int, int, int]):
    """(rank, y slice, x slice, z slice) of each MPI rank's local nodes in the
    global (noy, nox, noz) grid, as parallel_domain_decomp0 assigns them
    (Parallel_related.c:108-121): z fastest, then x, then y; neighbouring
    subdomains share their face nodes."""
    px, py, pz = nproc
    nox, noy, noz = dims
    ex, ey, ez = (nox - 1) // px, (noy - 1) // py, (noz - 1) // pz
    if ex * px != nox - 1 or ey * py != noy - 1 or ez * pz != noz - 1:
        raise ValueError(f"grid {dims} does not split over {nproc} ranks")
    for me in range(px * py * pz):
        kz, jy = me % pz, me // (px * pz)
        ix = (me - jy * px * pz) // pz
        yield (me, slice(jy * ey, jy * ey + ey + 1), slice(ix * ex, ix * ex + ex + 1),
               slice(kz * ez, kz * ez + ez + 1))


def write_restart_temperature(path_prefix: str, T: np.ndarray, dims: tuple[int, int, int] | None = None,
                              nproc: tuple[int, int, int] = (1, 1, 1)) -> list[str]:
    """Write ``<prefix>.temp.<rank>.0`` in the reference's restart format
    (read by Convection.c:473-490), one file per MPI rank over that rank's
    local nodes (rank_blocks).  ``dims`` = (nox, noy, noz) of the global grid;
    needed when nproc is not (1, 1, 1)."""
    if tuple(nproc) == (1, 1, 1):
        blocks = [(0, T.reshape(-1))]
    else:
        nox, noy, noz = dims
        Tg = T.reshape(noy, nox, noz)
        blocks = [(me, Tg[sy, sx, sz].reshape(-1)) for me, sy, sx, sz in rank_blocks(dims, nproc)]
    paths = []
    for me, vals in blocks:
        path = f"{path_prefix}.temp.{me}.0"
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        with open(path, "w") as fh:
            fh.write(f"{vals.size} 0 0.0\n")
            fh.write("".join(f"{float(v):.9e}\n" for v in vals))
        paths.append(path)
    return paths


def perturbed_temperature(p: Problem) -> np.ndarray:
    """Initial T of convection_initial_temperature for restart = 0
    (Convection.c:330-351), add_init_t_gradient on, one perturbation."""
    nox, noy, noz = p.dims()
    x, y, z = (c.astype(np.float64) for c in p.coords)
    T = np.empty((noy, nox, noz), dtype=np.float32)
    mag = float(np.float32(p.perturbmag[0])) if p.num_perturbations else 0.0
    kk = float(np.float32(p.perturbk[0])) if p.num_perturbations else 0.0
    for j in range(noy):
        for i in range(nox):
            v = (1.0 - z) + mag * np.sin(math.pi * (1.0 - z)) * math.cos(kk * math.pi * x[i]) * math.cos(kk * math.pi * y[j])
            T[j, i, :] = v.astype(np.float32)
    T = T.reshape(-1)
    # temperatures_conform_bcs (Boundary_conditions.c:949-990)
    Tv = T.reshape(noy, nox, noz)
    if p.bottbc >= 1:
        Tv[:, :, 0] = np.float32(p.bottbcval)
    if p.toptbc >= 1:
        Tv[:, :, -1] = np.float32(p.toptbcval)
    return T
