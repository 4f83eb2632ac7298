"""Swarm benchmark results in the reference's CSV schema.

``write_swarm_benchmark_csv`` writes the columns of the reference's
write_swarm_benchmark_csv (Swarm.cpp:1526-1716) from device results, so a
served run can be compared column for column with ``aurorapic_swarm``
output: the run description (dataset, config, seeds), the collision model
signature (NullCollisionModel's FNV-1a physics fingerprint,
Collision.cpp:365-577), and the statistics -- drift and mean energy with
block standard errors, diffusion, per-channel counts, rates and Poisson
errors.  A fixed population has no growth, attachment or Townsend
estimates and a branching one no diffusion; those columns carry the
reference's values for the mode (zeros, "no" and empty cells).  The
spatial-flux experiment is not built: its columns are always the
reference's "not run" cells.
"""
from __future__ import annotations

import math
import struct
from typing import List, Sequence

from . import gasdata

ELECTRON_MASS = 9.1093837139e-31      # Swarm.cpp:23
ELEMENTARY_CHARGE = 1.602176634e-19
TOWNSEND = 1.0e-21
BOLTZMANN = 1.380649e-23
NEUTRAL_SPEED_LIMIT_SIGMA = 8.0       # Collision.hpp:144

_PROCESS = {gasdata.ELASTIC: "elastic", gasdata.EXCITATION: "excitation",
            gasdata.IONIZATION: "ionization", gasdata.ATTACHMENT: "attachment",
            gasdata.CHARGE_EXCHANGE: "charge_exchange"}
_ANGULAR = {gasdata.ISOTROPIC: "isotropic", gasdata.BACKWARD: "backward",
            gasdata.HENYEY_GREENSTEIN: "henyey_greenstein"}

# the model signature's offset basis (Collision.cpp:365), as the reference has it
_SIGNATURE_OFFSET = 1469598103934665603
_PRIME = 1099511628211
_MASK = (1 << 64) - 1


class _Fnv:
    def __init__(self, offset=_SIGNATURE_OFFSET):
        self.h = offset

    def bytes(self, data: bytes):
        h = self.h
        for b in data:
            h = ((h ^ b) * _PRIME) & _MASK
        self.h = h

    def string(self, s: str):          # bytes, then a 0xff separator
        self.bytes(s.encode())
        self.bytes(b"\xff")

    def u64(self, v: int):
        self.bytes(struct.pack("<Q", v & _MASK))

    def double(self, v: float):
        self.bytes(struct.pack("<d", float(v)))


def neutral_velocity_stddev(ds: gasdata.GasDataset, cfg: gasdata.SwarmConfig) -> float:
    if cfg.neutral_temperature <= 0.0:
        return 0.0
    return math.sqrt(BOLTZMANN * cfg.neutral_temperature / ds.neutral_mass)


def collision_model_signature(ds: gasdata.GasDataset, cfg: gasdata.SwarmConfig,
                              max_candidates: int = 64) -> int:
    """The swarm's NullCollisionModel signature (collision_config,
    Swarm.cpp:561-593, hashed as Collision.cpp:365-577 does)."""
    h = _Fnv()
    h.string("null_collision")
    h.string("electrons")
    h.double(cfg.neutral_density)
    h.double(cfg.max_frequency)
    h.double(ELECTRON_MASS)
    vstd = neutral_velocity_stddev(ds, cfg)
    if ds.gas_name or ds.neutral_mass != 0.0 or cfg.neutral_temperature != 0.0 or \
            ds.data_provenance:
        h.string(ds.gas_name)
        h.double(ds.neutral_mass)
        h.double(cfg.neutral_temperature)
        h.string(ds.data_provenance)
        if vstd > 0.0:
            h.string("bounded_maxwellian_neutral_kinematics_v1")
            h.double(vstd)
            h.double(NEUTRAL_SPEED_LIMIT_SIGMA)
        elif ds.neutral_mass > 0.0:
            h.string("stationary_finite_mass_neutral_kinematics_v2")
    if ds.dataset_id or ds.dataset_version or ds.citation or ds.retrieved or ds.license:
        h.string(ds.dataset_id)
        h.string(ds.dataset_version)
        h.string(ds.citation)
        h.string(ds.retrieved)
        h.string(ds.license)
    if cfg.gas_data_file:
        h.u64(ds.format_version)
        h.string("si" if ds.units in ("si", "") else ds.units)
    h.u64(max_candidates)
    for c in ds.channels:
        h.string(c.name)
        h.string(_PROCESS[c.process])
        h.string(_ANGULAR[c.angular])
        if c.energy_frame != gasdata.FRAME_PROJECTILE:
            h.string("center_of_mass")
        h.double(c.threshold_energy)
        if c.process == gasdata.IONIZATION:
            h.string("fixed_population_electron")
            h.string("diagnostic_ion")
        elif c.process == gasdata.ATTACHMENT:
            h.string("diagnostic_negative_ion")
        for e, s in zip(c.energy, c.cross_section):
            h.double(e)
            h.double(s)
        if c.mean_cosine is not None:
            for e, g in zip(c.mean_cosine_energy, c.mean_cosine):
                h.double(e)
                h.double(g)
    return h.h


def _num(v) -> str:
    if isinstance(v, (int,)) and not isinstance(v, bool):
        return str(v)
    return "%.17g" % v


def _cell(s: str) -> str:
    if not any(ch in s for ch in ',"\r\n'):
        return s
    return '"' + s.replace('"', '""') + '"'


def header(ds: gasdata.GasDataset) -> List[str]:
    cols = ("dataset_id,dataset_version,gas,retrieved,provenance,citation,license,"
            "gas_data_file,population_model,neutral_density_m3,neutral_temperature_k,"
            "neutral_velocity_stddev_m_s,neutral_speed_limit_sigma,timestep_s,steps,"
            "burn_in_steps,particles,population_limit,work_item_limit,seed,"
            "spatial_histories_requested,spatial_length_m,spatial_bins,spatial_fit_begin_bin,"
            "spatial_fit_end_bin,spatial_max_steps,spatial_work_item_limit,"
            "spatial_min_r_squared,spatial_profile_file,reduced_field_td,"
            "collision_model_signature,electric_field_v_m,mean_velocity_x_m_s,"
            "mean_velocity_x_standard_error_m_s,electron_drift_velocity_m_s,"
            "reduced_mobility_1_v_m_s,mean_energy_ev,mean_energy_standard_error_ev,"
            "diffusion_available,longitudinal_diffusion_m2_s,transverse_diffusion_m2_s,"
            "initial_total_electron_weight,final_total_electron_weight,"
            "final_computational_particles,temporal_growth_rate_s,"
            "temporal_growth_rate_standard_error_s,total_ionization_rate_s,"
            "total_ionization_rate_standard_error_s,total_attachment_rate_s,"
            "total_attachment_rate_standard_error_s,net_creation_rate_s,"
            "net_creation_rate_standard_error_s,townsend_available,"
            "growth_over_flux_drift_townsend_1_m,"
            "growth_over_flux_drift_townsend_standard_error_1_m,"
            "rate_balance_effective_townsend_1_m,"
            "rate_balance_effective_townsend_standard_error_1_m,spatial_townsend_available,"
            "spatial_flux_townsend_1_m,spatial_flux_townsend_standard_error_1_m,"
            "spatial_flux_fit_r_squared,spatial_histories_completed,"
            "spatial_maximum_active_particles,spatial_particle_updates,"
            "maximum_observed_energy_ev,collision_candidates,null_collisions").split(",")
    for c in ds.channels:
        cols += [f"{c.name}_collisions", f"{c.name}_rate_per_electron_s",
                 f"{c.name}_rate_standard_error_s"]
    return cols


def rows(cfg: gasdata.SwarmConfig, ds: gasdata.GasDataset, results: Sequence,
         growth: Sequence = None) -> List[List[str]]:
    """One row of cells per reduced field; results[i] is the device result
    (hip.SwarmResult) of cfg.reduced_fields_td[i], seeded seed + i, and for
    population_model = branching_resampled growth[i] its hip.SwarmGrowth."""
    if len(results) == 0:
        raise ValueError("cannot write an empty swarm result set")
    branching = cfg.population_model == "branching_resampled"
    if branching and (growth is None or len(growth) != len(results)):
        raise ValueError("a branching swarm result needs its growth statistics")
    sig = collision_model_signature(ds, cfg)
    vstd = neutral_velocity_stddev(ds, cfg)
    limit = gasdata.effective_population_limit(cfg) if branching else 0
    out = []
    for i, (td, r) in enumerate(zip(cfg.reduced_fields_td, results)):
        efield = td * TOWNSEND * cfg.neutral_density
        mobility = 0.0 if efield == 0.0 else r.drift * cfg.neutral_density / efield
        ion = att = 0.0
        ion_se = att_se = 0.0
        for k, c in enumerate(ds.channels):
            if c.process == gasdata.IONIZATION:
                ion += r.rate[k]
                ion_se = math.hypot(ion_se, r.rate_se[k])
            elif c.process == gasdata.ATTACHMENT:
                att += r.rate[k]
                att_se = math.hypot(att_se, r.rate_se[k])
        if branching:
            g = growth[i]
            diffusion = ["no", "", ""]
            weights = [_num(g.initial_weight), _num(g.final_weight),
                       _num(int(g.final_particles)), _num(g.growth_rate),
                       _num(g.growth_rate_se)]
            townsend = (["yes", _num(g.townsend_growth), _num(g.townsend_growth_se),
                         _num(g.townsend_rate_balance), _num(g.townsend_rate_balance_se)]
                        if g.townsend_available else ["no", "", "", "", ""])
        else:
            diffusion = ["yes", _num(r.longitudinal_diffusion), _num(r.transverse_diffusion)]
            weights = [_num(cfg.particles), _num(float(cfg.particles)), _num(cfg.particles),
                       "0", "0"]
            townsend = ["no", "", "", "", ""]
        cells = [_cell(ds.dataset_id), _cell(ds.dataset_version), _cell(ds.gas_name),
                 _cell(ds.retrieved), _cell(ds.data_provenance), _cell(ds.citation),
                 _cell(ds.license), _cell(cfg.gas_data_file), _cell(cfg.population_model),
                 _num(cfg.neutral_density), _num(cfg.neutral_temperature), _num(vstd),
                 _num(NEUTRAL_SPEED_LIMIT_SIGMA), _num(cfg.timestep), _num(cfg.steps),
                 _num(cfg.burn_in_steps), _num(cfg.particles), _num(limit),
                 _num(cfg.work_item_limit),
                 _num(cfg.seed + i), _num(cfg.spatial_histories), _num(cfg.spatial_length_m),
                 _num(cfg.spatial_bins), _num(cfg.spatial_fit_begin_bin), "0",
                 _num(cfg.spatial_max_steps), _num(cfg.spatial_work_item_limit),
                 _num(cfg.spatial_min_r_squared), _cell(cfg.spatial_profile_file),
                 _num(td), str(sig), _num(efield), _num(-r.drift), _num(r.drift_se),
                 _num(r.drift), _num(mobility), _num(r.mean_energy_ev), _num(r.mean_energy_se),
                 *diffusion, *weights,
                 _num(ion), _num(ion_se), _num(att), _num(att_se), _num(ion - att),
                 _num(math.hypot(ion_se, att_se)), *townsend, "no", "", "", "",
                 "0", "0", "0", _num(r.max_energy_ev), _num(int(r.candidates)),
                 _num(int(r.nulls))]
        for k in range(len(ds.channels)):
            cells += [_num(int(r.collisions[k])), _num(r.rate[k]), _num(r.rate_se[k])]
        out.append(cells)
    return out


def write_swarm_benchmark_csv(path: str, cfg: gasdata.SwarmConfig, ds: gasdata.GasDataset,
                              results: Sequence, growth: Sequence = None) -> None:
    """write_swarm_benchmark_csv (Swarm.cpp:1526-1716) for device results."""
    body = rows(cfg, ds, results, growth)
    with open(path, "w") as f:
        f.write(",".join(header(ds)) + "\n")
        for cells in body:
            f.write(",".join(cells) + "\n")
