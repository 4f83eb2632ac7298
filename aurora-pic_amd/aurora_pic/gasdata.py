"""Readers for the reference's gas-data and swarm-benchmark files.

- ``load_gas_dataset``: the ``.gas`` manifest (versions 1 and 2) with its
  collision sections, the two-column cross-section tables and the
  Henyey-Greenstein mean-cosine tables (GasDataset.cpp:151-455,
  CrossSectionTable Collision.cpp:159-219, MeanCosineTable
  Collision.cpp:239-294).  Keys, defaults, validation order and error texts
  follow those functions so a caller sees the same failures.
- ``load_swarm_config``: the ``.swarm`` benchmark file
  (load_swarm_benchmark_config, Swarm.cpp:1252-1480).
- ``swarm_gas`` / ``apply_channel_options`` / ``swarm_requests``: hand a
  loaded dataset and benchmark to the device context (hip.Context).

All values are SI after loading: tables are multiplied by their
``energy_scale`` / ``cross_section_scale`` as the reference does.
"""
from __future__ import annotations

import datetime
import math
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

# process kinds (the device ABI's AURORA_* values)
ELASTIC, EXCITATION, IONIZATION, ATTACHMENT, CHARGE_EXCHANGE = 0, 1, 2, 3, 4
ISOTROPIC, BACKWARD, HENYEY_GREENSTEIN = 0, 1, 2
FRAME_PROJECTILE, FRAME_CENTER_OF_MASS = 0, 1

_TYPES = {"elastic": ELASTIC, "excitation": EXCITATION, "ionization": IONIZATION,
          "attachment": ATTACHMENT, "charge_exchange": CHARGE_EXCHANGE,
          "charge-exchange": CHARGE_EXCHANGE}
_FRAMES = {"projectile": FRAME_PROJECTILE, "center_of_mass": FRAME_CENTER_OF_MASS,
           "center-of-mass": FRAME_CENTER_OF_MASS, "centre_of_mass": FRAME_CENTER_OF_MASS,
           "centre-of-mass": FRAME_CENTER_OF_MASS}
_ANGULAR = {"isotropic": ISOTROPIC, "backward": BACKWARD,
            "henyey_greenstein": HENYEY_GREENSTEIN, "henyey-greenstein": HENYEY_GREENSTEIN}

_GLOBAL_KEYS = {"gas_data_version", "units", "gas", "neutral_mass", "dataset_id",
                "dataset_version", "data_provenance", "citation", "retrieved", "license"}
_CHANNEL_KEYS = {"type", "cross_section_file", "threshold_energy", "energy_scale",
                 "cross_section_scale", "angular_model", "mean_cosine_file",
                 "mean_cosine_energy_scale", "energy_frame"}


class GasDataError(RuntimeError):
    """A malformed gas dataset, table or swarm file (std::runtime_error in
    the reference)."""


@dataclass
class CollisionChannel:
    """CollisionChannelConfig (Collision.hpp) with its tables loaded."""
    name: str
    process: int
    energy: np.ndarray                   # J
    cross_section: np.ndarray            # m^2
    threshold_energy: float = 0.0        # J
    energy_scale: float = 1.0
    cross_section_scale: float = 1.0
    angular: int = ISOTROPIC
    energy_frame: int = FRAME_PROJECTILE
    mean_cosine_energy: Optional[np.ndarray] = None   # J
    mean_cosine: Optional[np.ndarray] = None
    cross_section_file: str = ""
    mean_cosine_file: str = ""


@dataclass
class GasDataset:
    """GasDataset (GasDataset.hpp)."""
    format_version: int
    units: str                           # "si" or "normalized" (version 2); "" for version 1
    gas_name: str
    neutral_mass: float
    dataset_id: str
    dataset_version: str
    data_provenance: str
    citation: str
    retrieved: str
    license: str
    channels: List[CollisionChannel] = field(default_factory=list)
    path: str = ""


def _strip_comment(line: str) -> str:
    cut = [i for i in (line.find("#"), line.find(";")) if i >= 0]
    return line[:min(cut)] if cut else line


def _number(values: Dict[str, str], key: str, fallback, context: str, integer=False):
    if key not in values:
        return fallback
    text = values[key]
    try:
        if integer:
            if not text.isdigit():
                raise ValueError
            return int(text)
        v = float(text)
        if not math.isfinite(v):
            raise ValueError
        return v
    except ValueError:
        raise GasDataError(f"{context} has invalid numeric key '{key}'") from None


def _required(values: Dict[str, str], key: str, context: str) -> str:
    v = values.get(key, "")
    if not v:
        raise GasDataError(f"{context} requires key '{key}'")
    return v


def _valid_identifier(s: str) -> bool:
    return bool(s) and all(ch.isalnum() or ch in "_-.:" for ch in s)


def _valid_iso_date(s: str) -> bool:
    if len(s) != 10 or s[4] != "-" or s[7] != "-":
        return False
    if not all(ch.isdigit() for i, ch in enumerate(s) if i not in (4, 7)):
        return False
    try:
        datetime.date(int(s[:4]), int(s[5:7]), int(s[8:]))
    except ValueError:
        return False
    return True


def _resolved(manifest: str, value: str) -> str:
    p = value if os.path.isabs(value) else os.path.join(os.path.dirname(manifest), value)
    return os.path.normpath(os.path.abspath(p))


def _two_column_rows(path: str, what: str, second: str, label: str):
    try:
        fh = open(path)
    except OSError:
        raise GasDataError(f"cannot open {label} table: {path}") from None
    rows = []
    with fh:
        for number, line in enumerate(fh, 1):
            parts = _strip_comment(line).split()
            if not parts:
                continue
            try:
                a = float(parts[0])
            except ValueError:
                continue            # the reference skips a line whose first token is not a number
            if len(parts) < 2:
                raise GasDataError(f"{what} table {path}:{number} must contain energy and {second}")
            try:
                b = float(parts[1])
            except ValueError:
                raise GasDataError(
                    f"{what} table {path}:{number} must contain energy and {second}") from None
            if len(parts) > 2:
                raise GasDataError(f"{what} table {path}:{number} has trailing columns")
            rows.append((a, b))
    return rows


def read_cross_section_table(path: str, energy_scale: float = 1.0,
                             cross_section_scale: float = 1.0):
    """CrossSectionTable (Collision.cpp:159-219): rows of (energy,
    cross section), scaled, finite, non-negative, energies strictly
    increasing, at least two rows."""
    if not (math.isfinite(energy_scale) and energy_scale > 0.0):
        raise ValueError("collision cross-section energy_scale must be positive and finite")
    if not (math.isfinite(cross_section_scale) and cross_section_scale > 0.0):
        raise ValueError("collision cross_section_scale must be positive and finite")
    e, s = [], []
    for a, b in _two_column_rows(path, "collision", "cross section",
                                 "collision cross-section"):
        a *= energy_scale
        b *= cross_section_scale
        if not (math.isfinite(a) and a >= 0.0 and math.isfinite(b) and b >= 0.0):
            raise GasDataError("collision table values must be finite and non-negative")
        if e and not a > e[-1]:
            raise GasDataError("collision table energies must be strictly increasing")
        e.append(a)
        s.append(b)
    if len(e) < 2:
        raise GasDataError(f"collision cross-section table requires at least two rows: {path}")
    return np.asarray(e), np.asarray(s)


def read_mean_cosine_table(path: str, energy_scale: float = 1.0):
    """MeanCosineTable (Collision.cpp:239-294): rows of (energy, mean
    cosine), |mean cosine| < 1, energies strictly increasing."""
    if not (math.isfinite(energy_scale) and energy_scale > 0.0):
        raise ValueError("mean-cosine energy_scale must be positive and finite")
    e, g = [], []
    for a, b in _two_column_rows(path, "mean-cosine", "mean cosine", "mean-cosine"):
        a *= energy_scale
        if not (math.isfinite(a) and a >= 0.0 and math.isfinite(b) and abs(b) < 1.0):
            raise GasDataError("mean-cosine table requires finite non-negative energies and "
                               "mean cosines strictly between -1 and 1")
        if e and not a > e[-1]:
            raise GasDataError("mean-cosine table energies must be strictly increasing")
        e.append(a)
        g.append(b)
    if len(e) < 2:
        raise GasDataError(f"mean-cosine table requires at least two rows: {path}")
    return np.asarray(e), np.asarray(g)


def load_gas_dataset(path: str) -> GasDataset:
    """load_gas_dataset (GasDataset.cpp:151-455)."""
    try:
        fh = open(path)
    except OSError:
        raise GasDataError(f"cannot open gas dataset: {path}") from None

    def error(number, message):
        raise GasDataError(f"gas dataset '{path}' line {number}: {message}")

    glob: Dict[str, str] = {}
    blocks: List[tuple] = []
    current, allowed, names = glob, _GLOBAL_KEYS, set()
    with fh:
        for number, raw in enumerate(fh, 1):
            line = _strip_comment(raw).strip()
            if not line:
                continue
            if line[0] == "[" and line[-1] == "]":
                section = line[1:-1].strip()
                if not section.lower().startswith("collision."):
                    error(number, f"unknown section '{section}'")
                name = section[len("collision."):].strip()
                if not name or name in names:
                    error(number, "empty or duplicate collision channel")
                names.add(name)
                blocks.append((name, {}))
                current, allowed = blocks[-1][1], _CHANNEL_KEYS
                continue
            if "=" not in line:
                error(number, "expected key = value")
            k, v = line.split("=", 1)
            key, value = k.strip().lower(), v.strip()
            if key not in allowed:
                error(number, f"unknown key '{key}'")
            if not key or not value:
                error(number, "empty key or value")
            if key in current:
                error(number, f"duplicate key '{key}'")
            current[key] = value

    context = f"gas dataset '{path}'"
    version = _number(glob, "gas_data_version", 0, context, integer=True)
    if version not in (1, 2):
        raise GasDataError(f"{context} supports gas_data_version 1 or 2")
    units = ""
    if version == 2:
        units = _required(glob, "units", context).lower()
        if units not in ("si", "normalized"):
            raise GasDataError(f"{context} units must be si or normalized")
    elif "units" in glob:
        raise GasDataError(f"{context} units requires gas_data_version = 2")
    gas_name = _required(glob, "gas", context)
    _required(glob, "neutral_mass", context)
    neutral_mass = _number(glob, "neutral_mass", 0.0, context)
    if not neutral_mass > 0.0:
        raise GasDataError(f"{context} requires positive neutral_mass")
    dataset_id = _required(glob, "dataset_id", context)
    if not _valid_identifier(dataset_id):
        raise GasDataError(f"{context} dataset_id contains unsupported characters")
    ds = GasDataset(version, units, gas_name, neutral_mass, dataset_id,
                    _required(glob, "dataset_version", context),
                    _required(glob, "data_provenance", context),
                    _required(glob, "citation", context),
                    _required(glob, "retrieved", context),
                    _required(glob, "license", context), path=path)
    if not _valid_iso_date(ds.retrieved):
        raise GasDataError(f"{context} retrieved must use a valid YYYY-MM-DD date")
    if not blocks:
        raise GasDataError(f"{context} requires collision channel sections")

    for name, values in blocks:
        cc = f"{context} channel '{name}'"
        if not _valid_identifier(name):
            raise GasDataError(f"{cc} name contains unsupported characters")
        kind = _required(values, "type", cc).lower()
        if kind not in _TYPES:
            raise GasDataError(f"{cc} type must be elastic, excitation, ionization, "
                               "attachment, or charge_exchange")
        process = _TYPES[kind]
        xs_file = _resolved(path, _required(values, "cross_section_file", cc))
        threshold = _number(values, "threshold_energy", 0.0, cc)
        e_scale = _number(values, "energy_scale", 1.0, cc)
        s_scale = _number(values, "cross_section_scale", 1.0, cc)
        frame_name = values.get("energy_frame", "projectile").lower()
        if frame_name not in _FRAMES:
            raise GasDataError(f"{cc} energy_frame must be projectile or center_of_mass")
        frame = _FRAMES[frame_name]
        if frame == FRAME_CENTER_OF_MASS and version != 2:
            raise GasDataError(f"{cc} center-of-mass energy requires gas_data_version = 2")
        has_angular = any(k in values for k in ("angular_model", "mean_cosine_file",
                                                "mean_cosine_energy_scale"))
        if has_angular and version != 2:
            raise GasDataError(f"{cc} angular scattering requires gas_data_version = 2")
        ang_name = values.get("angular_model", "isotropic").lower()
        if ang_name not in _ANGULAR:
            raise GasDataError(f"{cc} angular_model must be isotropic, backward, or "
                               "henyey_greenstein")
        angular = _ANGULAR[ang_name]
        mc_file = _resolved(path, values["mean_cosine_file"]) if "mean_cosine_file" in values \
            else ""
        mc_scale = _number(values, "mean_cosine_energy_scale", 1.0, cc)
        if frame == FRAME_CENTER_OF_MASS and process not in (ELASTIC, CHARGE_EXCHANGE):
            raise GasDataError(f"{cc} center_of_mass energy is supported only for elastic "
                               "and charge-exchange channels")
        if angular != ISOTROPIC and process != ELASTIC:
            raise GasDataError(f"{cc} anisotropic scattering is valid only for elastic channels")
        if angular == HENYEY_GREENSTEIN:
            if not mc_file:
                raise GasDataError(f"{cc} Henyey-Greenstein scattering requires mean_cosine_file")
        elif mc_file or mc_scale != 1.0:
            raise GasDataError(f"{cc} mean-cosine data requires angular_model = "
                               "henyey_greenstein")
        e, s = read_cross_section_table(xs_file, e_scale, s_scale)
        ch = CollisionChannel(name, process, e, s, threshold, e_scale, s_scale, angular, frame,
                              cross_section_file=xs_file, mean_cosine_file=mc_file)
        if angular == HENYEY_GREENSTEIN:
            me, mv = read_mean_cosine_table(mc_file, mc_scale)
            if me[0] > e[0] or me[-1] < e[-1]:
                raise GasDataError(f"{cc} mean-cosine energy range must cover the "
                                   "cross-section table")
            ch.mean_cosine_energy, ch.mean_cosine = me, mv
        ds.channels.append(ch)
    return ds


# ------------------------------------------------------------ swarm files
@dataclass
class SwarmConfig:
    """SwarmBenchmarkConfig (Swarm.hpp:28-57)."""
    gas_data_file: str = ""
    neutral_density: float = 0.0
    neutral_temperature: float = 0.0
    reduced_fields_td: List[float] = field(default_factory=list)
    max_frequency: float = 0.0
    timestep: float = 0.0
    steps: int = 0
    burn_in_steps: int = 0
    particles: int = 0
    population_model: str = "fixed_population_no_avalanche"
    population_limit: int = 0
    uncertainty_blocks: int = 10
    work_item_limit: int = 100_000_000
    initial_mean_energy_ev: float = 0.0
    max_energy_ev: float = 0.0
    seed: int = 12345
    output_file: str = "swarm.csv"
    spatial_histories: int = 0
    spatial_length_m: float = 0.0
    spatial_bins: int = 0
    spatial_fit_begin_bin: int = 0
    spatial_fit_end_bin: int = 0
    spatial_max_steps: int = 0
    spatial_work_item_limit: int = 100_000_000
    spatial_min_r_squared: float = 0.0
    spatial_profile_file: str = "swarm-spatial-profile.csv"


_FLOAT_PREFIX = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")
_INT_PREFIX = re.compile(r"[+-]?\d+")


def _stream_number(text: str, context: str, unsigned: bool):
    """parse_number (Swarm.cpp:36-54): a leading number as a stream
    extraction reads it, nothing after it, no sign on unsigned keys."""
    if unsigned and text.startswith("-"):
        raise GasDataError(f"{context} must not be negative")
    body = text.lstrip()
    m = (_INT_PREFIX if unsigned else _FLOAT_PREFIX).match(body)
    if not m:
        raise GasDataError(f"{context} must be numeric")
    if body[m.end():].strip():
        raise GasDataError(f"{context} has trailing content")
    return int(m.group(0)) if unsigned else float(m.group(0))


_SWARM_KEYS = {
    "swarm_config_version": None, "gas_data_file": None, "neutral_density": float,
    "neutral_temperature": float, "reduced_fields_td": None, "max_frequency": float,
    "timestep": float, "steps": int, "burn_in_steps": int, "particles": int,
    "population_model": None, "population_limit": int, "uncertainty_blocks": int,
    "work_item_limit": int, "initial_mean_energy_ev": float, "max_energy_ev": float,
    "seed": int, "output_file": None, "spatial_histories": int, "spatial_length_m": float,
    "spatial_bins": int, "spatial_fit_begin_bin": int, "spatial_fit_end_bin": int,
    "spatial_max_steps": int, "spatial_work_item_limit": int,
    "spatial_min_r_squared": float, "spatial_profile_file": None,
}
_SWARM_REQUIRED = ("neutral_density", "max_frequency", "timestep", "steps", "burn_in_steps",
                   "particles", "max_energy_ev")


def load_swarm_config(path: str) -> SwarmConfig:
    """load_swarm_benchmark_config (Swarm.cpp:1252-1480): key = value lines,
    '#'/';' comments, unknown and duplicate keys rejected."""
    try:
        fh = open(path)
    except OSError:
        raise GasDataError(f"cannot open swarm config: {path}") from None
    values: Dict[str, str] = {}
    with fh:
        for number, raw in enumerate(fh, 1):
            line = _strip_comment(raw).strip()
            if not line:
                continue
            if "=" not in line:
                raise GasDataError(f"{path}:{number}: expected key = value")
            k, v = line.split("=", 1)
            key, value = k.strip(), v.strip()
            if key not in _SWARM_KEYS:
                raise GasDataError(f"{path}:{number}: unknown key '{key}'")
            if key in values:
                raise GasDataError(f"{path}:{number}: duplicate key '{key}'")
            values[key] = value
    context = f"swarm config '{path}'"

    def parse(key, kind):
        return _stream_number(values[key], f"{context} key '{key}'", kind is int)

    for key in ("swarm_config_version",) + _SWARM_REQUIRED:
        if key not in values or not values[key]:
            raise GasDataError(f"{context} requires '{key}'")
    if parse("swarm_config_version", int) != 1:
        raise GasDataError(f"{context} requires swarm_config_version = 1")
    base = os.path.dirname(path)
    cfg = SwarmConfig()
    if not values.get("gas_data_file"):
        raise GasDataError(f"{context} requires 'gas_data_file'")
    cfg.gas_data_file = os.path.normpath(os.path.join(base, values["gas_data_file"]))
    if not values.get("reduced_fields_td"):
        raise GasDataError(f"{context} requires 'reduced_fields_td'")
    fields = values["reduced_fields_td"]
    if fields[0] == "," or fields[-1] == ",":
        raise GasDataError(f"{context} reduced_fields_td contains an empty value")
    for item in fields.split(","):
        if not item.strip():
            raise GasDataError(f"{context} reduced_fields_td contains an empty value")
        cfg.reduced_fields_td.append(
            _stream_number(item.strip(), f"{context} reduced_fields_td value", False))
    for key, kind in _SWARM_KEYS.items():
        if kind is not None and key in values:
            setattr(cfg, key, parse(key, kind))
    if "population_model" in values:
        model = values["population_model"]
        if model in ("fixed_population_no_avalanche", "fixed"):
            cfg.population_model = "fixed_population_no_avalanche"
        elif model in ("branching_resampled", "branching"):
            cfg.population_model = "branching_resampled"
        else:
            raise GasDataError(f"{context} key 'population_model' must be "
                               "fixed_population_no_avalanche or branching_resampled")
    cfg.output_file = os.path.normpath(os.path.join(base, values.get("output_file",
                                                                     cfg.output_file)))
    if "spatial_profile_file" in values:
        cfg.spatial_profile_file = os.path.normpath(os.path.join(
            base, values["spatial_profile_file"]))
    return cfg


# ------------------------------------------------------- device hand-off
def gas_tables(ds: GasDataset):
    """The (kind, threshold_J, energies_J, sigmas_m2) channel list the
    device loaders take."""
    return [(c.process, c.threshold_energy, c.energy, c.cross_section) for c in ds.channels]


def swarm_gas(ds: GasDataset, cfg: SwarmConfig, max_candidates: int = 64) -> dict:
    """The gas description hip.Context.swarm_gas takes (collision_config,
    Swarm.cpp:561-593).  A positive neutral temperature needs an SI
    dataset (Swarm.cpp:456-460)."""
    if cfg.neutral_temperature > 0.0 and ds.units == "normalized":
        raise GasDataError("positive swarm neutral_temperature requires an SI gas dataset")
    return dict(channels=gas_tables(ds), neutral_density=cfg.neutral_density,
                max_frequency=cfg.max_frequency, max_candidates=max_candidates,
                neutral_mass=ds.neutral_mass, neutral_temperature=cfg.neutral_temperature)


def apply_channel_options(ctx, ds: GasDataset, model: int) -> None:
    """Per-channel angular model and energy frame on a device context
    (model 0: the context's MCC model, 1: the swarm gas)."""
    for k, c in enumerate(ds.channels):
        if c.angular == ISOTROPIC and c.energy_frame == FRAME_PROJECTILE:
            continue
        ctx.mcc_channel_options(k, angular=c.angular, frame=c.energy_frame,
                                mc_energy=c.mean_cosine_energy, mc_value=c.mean_cosine,
                                model=model)


SWARM_ALLOCATION_LIMIT = 10_000_000      # Swarm.cpp:28


def effective_population_limit(cfg: SwarmConfig) -> int:
    """The branching population limit (Swarm.cpp:240-262): the configured
    one, else the larger of twice the particles and particles + 1024, within
    the allocation safety limit."""
    if cfg.population_limit != 0:
        return cfg.population_limit
    if cfg.particles >= SWARM_ALLOCATION_LIMIT:
        return cfg.particles
    capacity = SWARM_ALLOCATION_LIMIT - cfg.particles
    doubled = capacity if cfg.particles > capacity // 2 else 2 * cfg.particles
    buffer = min(capacity, 1024)
    buffered = capacity if cfg.particles > capacity - buffer else cfg.particles + 1024
    return min(capacity, max(doubled, buffered))


_TOWNSEND = 1.0e-21                       # V m^2 per Td
_EV = 1.602176634e-19


def _exceeds_product_limit(first: int, second: int, limit: int) -> bool:
    return first != 0 and second > limit // first


def validate_swarm_config(cfg: SwarmConfig, ds: GasDataset) -> None:
    """validate_config (Swarm.cpp:277-559): the reference's checks of a
    swarm scan against its gas dataset, in its order and with its messages,
    including the work_item_limit product check (fields x particles x steps,
    Swarm.cpp:336-359) and the population_limit rules (Swarm.cpp:317-335).
    Raises GasDataError."""
    def finite_pos(v):
        return math.isfinite(v) and v > 0.0

    if ds.format_version != 2 or ds.units != "si":
        raise GasDataError("swarm benchmarks require a version-2 SI gas dataset")
    if not finite_pos(cfg.neutral_density):
        raise GasDataError("swarm neutral_density must be positive and finite")
    if not finite_pos(cfg.max_frequency):
        raise GasDataError("swarm max_frequency must be positive and finite")
    if not finite_pos(cfg.timestep):
        raise GasDataError("swarm timestep must be positive and finite")
    if cfg.max_frequency * cfg.timestep > 0.1:
        raise GasDataError("swarm max_frequency * timestep must not exceed 0.1")
    if cfg.steps == 0 or cfg.burn_in_steps >= cfg.steps:
        raise GasDataError("swarm steps must exceed burn_in_steps")
    if cfg.particles == 0:
        raise GasDataError("swarm particles must be positive")
    if cfg.particles > SWARM_ALLOCATION_LIMIT:
        raise GasDataError("swarm particles exceeds the safety limit of 10000000")
    limit = effective_population_limit(cfg)
    if cfg.population_model == "fixed_population_no_avalanche":
        if cfg.population_limit != 0:
            raise GasDataError("swarm population_limit is valid only for "
                               "population_model = branching_resampled")
    elif limit <= cfg.particles or limit > SWARM_ALLOCATION_LIMIT - cfg.particles:
        raise GasDataError("branching swarm population_limit must exceed particles and "
                           "population_limit + particles must not exceed the allocation "
                           "safety limit of 10000000")
    if cfg.work_item_limit == 0:
        raise GasDataError("swarm work_item_limit must be positive")
    nf = len(cfg.reduced_fields_td)
    if (_exceeds_product_limit(nf, cfg.particles, cfg.work_item_limit)
            or _exceeds_product_limit(nf * cfg.particles, cfg.steps, cfg.work_item_limit)):
        raise GasDataError("swarm scan exceeds work_item_limit; raise it explicitly only on "
                           "an appropriate compute host")
    sampling = cfg.steps - cfg.burn_in_steps
    if cfg.uncertainty_blocks == 0 or cfg.uncertainty_blocks > sampling:
        raise GasDataError("swarm uncertainty_blocks must be between 1 and the number of "
                           "sampling steps")
    if sampling % cfg.uncertainty_blocks != 0:
        raise GasDataError("swarm sampling steps must be divisible by uncertainty_blocks")
    branching = cfg.population_model == "branching_resampled"
    if branching and sampling // cfg.uncertainty_blocks < 2:
        raise GasDataError("branching swarm requires at least two sampling steps per "
                           "uncertainty block")
    if cfg.spatial_histories == 0:
        if (cfg.spatial_length_m != 0.0 or cfg.spatial_bins != 0
                or cfg.spatial_fit_begin_bin != 0 or cfg.spatial_fit_end_bin != 0
                or cfg.spatial_max_steps != 0 or cfg.spatial_min_r_squared != 0.0):
            raise GasDataError("spatial Townsend controls require positive spatial_histories")
    else:
        # the spatial Townsend experiment is not built (Swarm.cpp:645-938)
        raise GasDataError("spatial Townsend estimation (spatial_histories > 0) is not "
                           "supported by this build")
    if not math.isfinite(cfg.initial_mean_energy_ev) or cfg.initial_mean_energy_ev < 0.0:
        raise GasDataError("swarm initial_mean_energy_ev must be finite and non-negative")
    if not finite_pos(cfg.max_energy_ev):
        raise GasDataError("swarm max_energy_ev must be positive and finite")
    if not math.isfinite(cfg.neutral_temperature) or cfg.neutral_temperature < 0.0:
        raise GasDataError("swarm neutral_temperature must be finite and non-negative")
    if not cfg.reduced_fields_td:
        raise GasDataError("swarm requires at least one reduced field")
    for f in cfg.reduced_fields_td:
        if not math.isfinite(f) or f < 0.0:
            raise GasDataError("swarm reduced fields must be finite and non-negative")
        if not math.isfinite(f * _TOWNSEND * cfg.neutral_density):
            raise GasDataError("swarm reduced field and neutral density overflow the "
                               "electric field")
    if len(set(cfg.reduced_fields_td)) != nf:
        raise GasDataError("swarm reduced fields must not contain duplicates")
    if not ds.channels:
        raise GasDataError("swarm gas dataset has no collision channels")
    kinds = set()
    for ch in ds.channels:
        kinds.add(ch.process)
        if ch.process == CHARGE_EXCHANGE:
            raise GasDataError("electron swarm benchmark does not support charge-exchange "
                               "channels")
        if ch.energy_frame != FRAME_PROJECTILE:
            raise GasDataError("electron swarm benchmark requires projectile-frame "
                               "cross-section energy")
        if cfg.max_energy_ev > ch.energy[-1] / _EV:
            raise GasDataError(f"swarm max_energy_ev exceeds channel '{ch.name}' table coverage")
        if ch.angular == HENYEY_GREENSTEIN and cfg.max_energy_ev > ch.mean_cosine_energy[-1] / _EV:
            raise GasDataError(f"swarm max_energy_ev exceeds channel '{ch.name}' angular "
                               "table coverage")
    if ELASTIC not in kinds:
        raise GasDataError("electron swarm benchmark requires an elastic channel")
    if not branching and ATTACHMENT in kinds:
        raise GasDataError("fixed-population swarm does not support attachment; use "
                           "population_model = branching_resampled")
    if branching and IONIZATION not in kinds and ATTACHMENT not in kinds:
        raise GasDataError("branching swarm requires an ionization or attachment channel")


def swarm_requests(cfg: SwarmConfig):
    """One device request per reduced field, seeded as
    run_swarm_benchmark seeds its fields (config seed + field index,
    Swarm.cpp:1513-1515).  Fixed populations go to swarm_submit, branching
    ones to swarm_branching with effective_population_limit(cfg)."""
    from .hip import SwarmRequest
    out = []
    for index, td in enumerate(cfg.reduced_fields_td):
        out.append(SwarmRequest(reduced_field_td=td, timestep=cfg.timestep,
                                particles=cfg.particles, steps=cfg.steps,
                                burn_in=cfg.burn_in_steps, blocks=cfg.uncertainty_blocks,
                                initial_mean_energy_ev=cfg.initial_mean_energy_ev,
                                max_energy_ev=cfg.max_energy_ev, seed=cfg.seed + index))
    return out
