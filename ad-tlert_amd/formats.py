"""On-disk formats around the forward path: ERT ``.dat`` files, compact ``.npz`` forward data,
ParFlow grid settings and resistivity slices, and the disk cache of precomputed unit potentials.

Each reads and writes what the reference's workflows read and write, so files move between
the two implementations unchanged:

* ``load_terrain_forward_dat`` / ``save_terrain_forward_dat``: the unified (pyGIMLi-style) ERT
  ``.dat`` layout of ``workflows/terrain.py:318-425``: a sensor count, ``x y z`` rows (commented
  header lines allowed), a datum count, a ``# a b m n ...`` header and 1-based rows; sensor
  coordinates are snapped to a 1e-12 grid as GIMLi does;
* ``save_terrain_forward_npz``: the compact forward data and terrain metadata file
  (``workflows/terrain.py:428-452``);
* ``parse_pftcl``, ``parse_resistivity_slice_name``, ``discover_resistivity_slices``,
  ``load_terrain_resistivity_slice``: ParFlow grid settings and the 2D resistivity slices of a
  time-lapse series (``workflows/terrain.py:182-282``);
* ``UnitPotentialCache``: ``(W, E, dofs)`` unit potential stacks stored as
  ``<sha256>.npz`` files in a cache directory (``forward/ert2p5d.py:349-402, 1124-1141``): the
  key hashes a format version, the stack's name, the result dtype, the operator options and the
  arrays that determine the stack; unreadable or mis-shaped files count as misses, and files are
  written to a temporary name and renamed into place.

Everything here is host-side file handling; nothing touches a GPU.
"""

from __future__ import annotations

import hashlib
import os
import re
import zipfile
from dataclasses import dataclass
from pathlib import Path

import numpy as np

GIMLI_COORDINATE_TOLERANCE = 1.0e-12
"""Sensor coordinates read from ``.dat`` files are rounded to this grid (as GIMLi does)."""

TERRAIN_CACHE_VERSION = 1
"""Bumped whenever the cached stacks' meaning changes, so old files stop matching."""


# ----------------------------------------------------------------------------------------------
# ERT .dat files
# ----------------------------------------------------------------------------------------------


@dataclass(frozen=True)
class TerrainForwardData:
    """ERT data parsed from a ``.dat`` file: 0-based ABMN rows, electrode x and z, optional errors."""

    rhoa: np.ndarray
    measurements: np.ndarray
    elec_x: np.ndarray
    elec_z: np.ndarray
    err: np.ndarray | None = None


class _Lines:
    """Non-empty stripped lines of a text file with a read cursor."""

    def __init__(self, path: Path):
        self.path = path
        self.rows = [line.strip() for line in path.read_text(encoding="utf-8").splitlines() if line.strip()]
        self.at = 0

    def comments(self) -> list[str]:
        """Consume consecutive ``#`` lines; return them without the marker."""

        out = []
        while self.at < len(self.rows) and self.rows[self.at].startswith("#"):
            out.append(self.rows[self.at][1:].strip())
            self.at += 1
        return out

    def integer(self, what: str) -> int:
        try:
            return int(self.rows[self.at])
        except (IndexError, ValueError) as exc:
            raise ValueError(f"{self.path}: {what}") from exc
        finally:
            self.at += 1

    def table(self, count: int, what: str, width: int | None = None) -> list[list[float]]:
        """The next ``count`` rows as floats (only their first ``width`` fields if given)."""

        if self.at + count > len(self.rows):
            raise ValueError(f"{self.path}: {what} is truncated")
        block = [[float(v) for v in row.split()[:width]] for row in self.rows[self.at:self.at + count]]
        self.at += count
        return block


def load_terrain_forward_dat(path: str | Path) -> TerrainForwardData:
    """Read an ERT ``.dat`` file (sensors, then ``a b m n ... rhoa`` rows with 1-based indices)."""

    lines = _Lines(Path(path))
    if len(lines.rows) < 4:
        raise ValueError(f"{lines.path} is too short to be an ERT .dat file")
    sensors_n = lines.integer("first line must be sensor count")
    lines.comments()
    sensors = np.asarray(lines.table(sensors_n, "sensor table", width=3), dtype=float)
    sensors = np.rint(sensors / GIMLI_COORDINATE_TOLERANCE) * GIMLI_COORDINATE_TOLERANCE
    lines.comments()
    if lines.at >= len(lines.rows):
        raise ValueError(f"{lines.path}: missing measurement count")
    data_n = lines.integer("measurement count must be an integer")
    headers = lines.comments()
    if not headers:
        raise ValueError(f"{lines.path}: missing measurement header")
    names = headers[-1].split()
    column = {name: i for i, name in enumerate(names)}
    absent = [name for name in ("a", "b", "m", "n", "rhoa") if name not in column]
    if absent:
        raise ValueError(f"{lines.path}: missing required measurement columns {absent}")
    table = lines.table(data_n, "measurement table")
    if any(len(row) < len(names) for row in table):
        raise ValueError(f"{lines.path}: measurement rows have fewer columns than the header")
    values = np.asarray(table, dtype=float)
    abmn = values[:, [column[c] for c in "abmn"]].astype(np.int32) - 1
    return TerrainForwardData(rhoa=values[:, column["rhoa"]], measurements=abmn, elec_x=sensors[:, 0],
                              elec_z=sensors[:, 1], err=values[:, column["err"]] if "err" in column else None)


def _rhoa_for(case, rhoa) -> np.ndarray:
    expected = int(case.survey.measurement_count)
    out = np.asarray(rhoa, dtype=float).reshape(-1)
    if out.shape != (expected,):
        raise ValueError(f"rhoa must have shape ({expected},)")
    return out


def save_terrain_forward_dat(path: str | Path, case, rhoa, *, relative_error: float = 0.03) -> None:
    """Write a pyGIMLi-style ERT ``.dat`` file for ``case`` (``elec_x``, ``elec_z``, ``survey``):
    sensors ``x z 0``, then per datum ``a b m n err i ip iperr k r rhoa u valid`` (1-based
    electrodes, the analytic geometric factor ``k``, zeros for the unused columns)."""

    from .survey import Survey

    rhoa = _rhoa_for(case, rhoa)
    xz = np.column_stack((np.asarray(case.elec_x, dtype=float), np.asarray(case.elec_z, dtype=float)))
    abmn = np.asarray(case.survey.measurements, dtype=np.int32)
    k = Survey.from_arrays(xz, abmn).geometric_factors().tolist()
    zero = f"{0.0:.14e}"
    out = [f"{xz.shape[0]}\n", "# x y z\n"]
    out += [f"{x:.14g}\t{z:.14g}\t0\n" for x, z in xz]
    out += [f"{abmn.shape[0]}\n", "# a b m n err i ip iperr k r rhoa u valid \n"]
    for (a, b, m, n), kk, value in zip(abmn.tolist(), k, rhoa):
        fields = [str(a + 1), str(b + 1), str(m + 1), str(n + 1), f"{float(relative_error):.14e}", zero, zero, zero,
                  f"{kk:.14e}", zero, f"{value:.14e}", zero, "1"]
        out.append("\t".join(fields) + "\n")
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    target.write_text("".join(out), encoding="utf-8")


def save_terrain_forward_npz(path: str | Path, case, rhoa, *, relative_error: float = 0.03) -> None:
    """Compact forward data and terrain metadata (``rhoa``, ``err``, ``a b m n``, electrode and
    surface coordinates, layer thicknesses, the ParFlow ``y_index``)."""

    rhoa = _rhoa_for(case, rhoa)
    abmn = np.asarray(case.survey.measurements, dtype=np.int32)
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    np.savez(target, rhoa=rhoa, err=np.full(rhoa.shape, float(relative_error)),
             a=abmn[:, 0], b=abmn[:, 1], m=abmn[:, 2], n=abmn[:, 3], elec_x=case.elec_x, elec_z=case.elec_z,
             x_nodes=case.x_nodes, z_top=case.z_top, layer_thickness=case.layer_thickness,
             y_index=np.asarray([case.y_index], dtype=np.int32))


# ----------------------------------------------------------------------------------------------
# ParFlow grids and resistivity slices
# ----------------------------------------------------------------------------------------------


@dataclass(frozen=True)
class ParflowGrid:
    """ParFlow grid settings needed to rebuild a 2D terrain slice."""

    dx: float
    dy: float
    dz_base: float
    nx: int
    ny: int
    nz: int
    dz_scales: np.ndarray


_NUMBER = r"[0-9eE+\-.]+"
_GRID_KEYS = {"DX": _NUMBER, "DY": _NUMBER, "DZ": _NUMBER, "NX": r"\d+", "NY": r"\d+", "NZ": r"\d+"}


def parse_pftcl(path: str | Path) -> ParflowGrid:
    """Grid dimensions and per-layer ``dzScale`` values from a ParFlow ``.pftcl`` file."""

    text = Path(path).read_text(encoding="utf-8", errors="ignore")
    settings = {}
    for key, pattern in _GRID_KEYS.items():
        found = re.findall(rf'ComputationalGrid\.{key}\s+"({pattern})"', text)
        if found:
            settings[key.lower()] = float(found[-1])  # a later setting overrides an earlier one
    scales = {int(i): float(v) for i, v in re.findall(rf'Cell\.(\d+)\.dzScale\.Value\s+"({_NUMBER})"', text)}
    absent = [key for key in ("dx", "dy", "dz", "nx", "ny", "nz") if key not in settings]
    if absent:
        raise ValueError(f"failed to parse ComputationalGrid settings from {path}: missing {absent}")
    nz = int(settings["nz"])
    if len(scales) != nz:
        raise ValueError(f"dzScale count {len(scales)} does not match NZ={nz}")
    return ParflowGrid(dx=settings["dx"], dy=settings["dy"], dz_base=settings["dz"], nx=int(settings["nx"]),
                       ny=int(settings["ny"]), nz=nz, dz_scales=np.asarray([scales[i] for i in range(nz)], dtype=float))


_SLICE_PATTERNS = (re.compile(r"resistivity2d_y(\d+)_t(\d+)\.npy$"), re.compile(r"resistivity_t(\d+)\.npy$"))


def parse_resistivity_slice_name(path: str | Path) -> tuple[int, int]:
    """``(y_index, timestep)`` of ``resistivity2d_y{y}_t{step}.npy``; ``resistivity_t{step}.npy``
    (a file holding every y) gives ``y_index = -1``."""

    name = Path(path).name
    per_y, whole = (p.search(name) for p in _SLICE_PATTERNS)
    if per_y:
        return int(per_y.group(1)), int(per_y.group(2))
    if whole:
        return -1, int(whole.group(1))
    raise ValueError(f"cannot parse y-index and timestep from filename: {name}")


def discover_resistivity_slices(input_dir: str | Path, *, y_index: int | None = None, file_stride: int = 1,
                                max_steps: int | None = None) -> list[tuple[int, Path]]:
    """Sorted ``(step, path)`` pairs of the slices in ``input_dir`` for ``y_index`` (files of
    every y always qualify), every ``file_stride``-th one, at most ``max_steps``."""

    if file_stride < 1:
        raise ValueError("file_stride must be >= 1")
    if max_steps is not None and max_steps < 1:
        raise ValueError("max_steps must be >= 1 when set")
    found = {}
    for path in Path(input_dir).iterdir():
        try:
            y, step = parse_resistivity_slice_name(path)
        except ValueError:
            continue
        if y_index is None or y < 0 or y == y_index:
            found[(step, str(path))] = path
    ordered = [(step, found[(step, key)]) for step, key in sorted(found)]
    return ordered[::file_stride][:max_steps]


def load_terrain_resistivity_slice(path: str | Path, grid: ParflowGrid, *, y_index: int) -> np.ndarray:
    """The ``(nz, nx)`` slice at ``y_index`` of a 2D file, or of a 3D file in any of the ParFlow
    axis orders ``(nz, ny, nx)``, ``(ny, nz, nx)``, ``(nx, ny, nz)``."""

    values = np.asarray(np.load(path, allow_pickle=False), dtype=float)
    if values.ndim == 2:
        return values
    if values.ndim != 3:
        raise ValueError(f"{path}: resistivity array must be 2D or 3D, got shape={values.shape}")
    if not 0 <= y_index < grid.ny:
        raise ValueError(f"y_index={y_index} out of range for NY={grid.ny}")
    nz, ny, nx = grid.nz, grid.ny, grid.nx
    layouts = {(nz, ny, nx): lambda v: v[:, y_index, :], (ny, nz, nx): lambda v: v[y_index],
               (nx, ny, nz): lambda v: v[:, y_index, :].T}
    take = layouts.get(values.shape)
    if take is None:
        raise ValueError(f"{path}: cannot infer 3D resistivity axis order from shape={values.shape}; "
                         f"expected ({nz}, {ny}, {nx}), ({ny}, {nz}, {nx}), or ({nx}, {ny}, {nz})")
    return take(values)


# ----------------------------------------------------------------------------------------------
# disk cache of unit potential stacks
# ----------------------------------------------------------------------------------------------


def terrain_cache_dir(cache_dir: str | Path | None) -> Path | None:
    """The cache directory: the argument, else ``$ADTLERT_TERRAIN_CACHE_DIR``, else none."""

    chosen = cache_dir if cache_dir is not None else os.environ.get("ADTLERT_TERRAIN_CACHE_DIR", "").strip()
    return Path(chosen).expanduser() if chosen else None


class UnitPotentialCache:
    """``<sha256>.npz`` files (one array ``value``) in ``directory``; ``None`` disables it."""

    def __init__(self, directory: str | Path | None):
        self.directory = terrain_cache_dir(directory)

    def path(self, name: str, dtype: str, options, arrays) -> Path | None:
        if self.directory is None:
            return None
        digest = hashlib.sha256()
        for label, value in (("version", TERRAIN_CACHE_VERSION), ("name", name), ("dtype", dtype),
                             ("options", tuple(options))):
            digest.update(f"{label}={value!r};".encode())
        for array in arrays:
            a = np.ascontiguousarray(np.asarray(array))
            digest.update(f"{a.dtype}{a.shape}".encode())
            digest.update(a.tobytes())
        return self.directory / f"{digest.hexdigest()}.npz"

    @staticmethod
    def read(path: Path | None, shape) -> np.ndarray | None:
        """The cached array, or None when the file is absent, unreadable or of another shape."""

        if path is None:
            return None
        try:
            with np.load(path, allow_pickle=False) as payload:
                value = payload["value"]
        except (OSError, ValueError, KeyError, EOFError, zipfile.BadZipFile):
            return None
        return value if value.shape == tuple(shape) else None

    @staticmethod
    def write(path: Path | None, value: np.ndarray) -> None:
        """Store ``value`` atomically (temporary file, then rename); failures are ignored."""

        if path is None:
            return
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            tmp = path.with_name(f"{path.name}.{os.getpid()}.tmp.npz")
            np.savez(tmp, value=np.asarray(value))
            tmp.replace(path)
        except OSError:
            pass

    def get(self, name: str, dtype: str, options, arrays, shape, compute) -> np.ndarray:
        """The cached stack of this key, else ``compute()`` (stored for next time)."""

        path = self.path(name, dtype, options, arrays)
        cached = self.read(path, shape)
        if cached is not None:
            return cached
        value = np.asarray(compute())
        self.write(path, value)
        return value
