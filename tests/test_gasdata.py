"""Gas-dataset and swarm-file readers (CPU).

- The reference's own loader checks on its example files
  (test_core.cpp:5718-5749: manifest fields, the Henyey-Greenstein channel;
  5895-5935: an invalid retrieval date).
- Error texts of the readers against the reference loader itself
  (oracle/_ref, where built) on a set of malformed manifests.
- The oracle's swarm on the reference's examples/synthetic_swarm.swarm
  (300 K neutrals, Henyey-Greenstein elastic scattering) against the
  reference's own results over 16 seeds (tests/golden/swarm)."""
import math
import os
import shutil

import numpy as np
import pytest

import oracle_py
from aurora_pic import gasdata
from conftest import GOLDEN, golden

SWARM_DIR = os.path.join(GOLDEN, "swarm")
GAS_DIR = os.path.join(GOLDEN, "gas")
EV = 1.602176634e-19
KB = 1.380649e-23


def test_reference_example_manifests():
    ds = gasdata.load_gas_dataset(os.path.join(GAS_DIR, "synthetic_ionization.gas"))
    assert ds.format_version == 2 and ds.units == "normalized"
    assert ds.gas_name == "synthetic_validation_gas" and ds.neutral_mass == 40.0
    assert len(ds.channels) == 1 and ds.channels[0].process == gasdata.IONIZATION
    assert ds.channels[0].threshold_energy == 1.0
    ang = gasdata.load_gas_dataset(os.path.join(SWARM_DIR, "synthetic_swarm.gas"))
    assert len(ang.channels) == 3
    el = ang.channels[0]
    assert el.angular == gasdata.HENYEY_GREENSTEIN
    assert os.path.basename(el.mean_cosine_file) == "synthetic_swarm_mean_cosine.dat"
    # tables are scaled to SI as the reference's CrossSectionTable does
    np.testing.assert_array_equal(el.energy, np.array([0, 2, 5, 10, 20]) * EV)
    np.testing.assert_array_equal(el.mean_cosine_energy, np.array([0, 2, 5, 10, 20]) * EV)
    assert el.mean_cosine[0] == 0.1 and el.mean_cosine[-1] == 0.8
    assert ang.channels[1].threshold_energy == 3.204353268e-19
    cx = gasdata.load_gas_dataset(os.path.join(GAS_DIR, "synthetic_charge_exchange.gas"))
    assert cx.channels[0].process == gasdata.CHARGE_EXCHANGE
    at = gasdata.load_gas_dataset(os.path.join(GAS_DIR, "synthetic_attachment.gas"))
    assert at.channels[0].process == gasdata.ATTACHMENT


def _variant(tmp_path, src, edit):
    for name in os.listdir(os.path.dirname(src)):
        shutil.copy(os.path.join(os.path.dirname(src), name), tmp_path)
    path = os.path.join(tmp_path, os.path.basename(src))
    text = open(path).read()
    open(path, "w").write(edit(text))
    return path


BAD_MANIFESTS = [
    ("retrieved = 2026-07-28", "retrieved = 2026-02-30", "valid YYYY-MM-DD"),
    ("gas_data_version = 2", "gas_data_version = 3", "supports gas_data_version 1 or 2"),
    ("units = normalized", "units = cgs", "units must be si or normalized"),
    ("neutral_mass = 40.0", "neutral_mass = -1", "requires positive neutral_mass"),
    ("neutral_mass = 40.0", "neutral_mass = 4x", "has invalid numeric key 'neutral_mass'"),
    ("type = ionization", "type = recombination", "type must be elastic"),
    ("threshold_energy = 1.0", "threshold_energy = 1.0\nthreshold_energy = 2.0",
     "duplicate key 'threshold_energy'"),
    ("threshold_energy = 1.0", "threshold_energy = 1.0\ncolour = red", "unknown key 'colour'"),
    ("[collision.synthetic_ionization]", "[reaction.synthetic_ionization]", "unknown section"),
    ("dataset_id = aurorapic.synthetic.ionization", "dataset_id = a b",
     "dataset_id contains unsupported characters"),
    ("threshold_energy = 1.0", "threshold_energy = 1.0\nangular_model = backward",
     "anisotropic scattering is valid only for elastic channels"),
    ("threshold_energy = 1.0", "threshold_energy = 1.0\nenergy_frame = center_of_mass",
     "center_of_mass energy is supported only for elastic and charge-exchange channels"),
    ("threshold_energy = 1.0", "threshold_energy = 1.0\nmean_cosine_energy_scale = 2",
     "mean-cosine data requires angular_model = henyey_greenstein"),
    ("cross_section_file = mcc_2d3v_ionization.dat", "cross_section_file = missing.dat",
     "cannot open collision cross-section table"),
    ("gas_data_version = 2\nunits = normalized", "gas_data_version = 1\nunits = normalized",
     "units requires gas_data_version = 2"),
    ("license = Project source license, synthetic data only", "", "requires key 'license'"),
]


@pytest.mark.parametrize("old,new,message", BAD_MANIFESTS)
def test_malformed_manifests_raise_reference_messages(tmp_path, old, new, message):
    src = os.path.join(GAS_DIR, "synthetic_ionization.gas")
    path = _variant(str(tmp_path), src, lambda t: t.replace(old, new, 1))
    with pytest.raises((gasdata.GasDataError, ValueError), match=message.replace("(", r"\(")):
        gasdata.load_gas_dataset(path)


def test_malformed_manifests_match_reference_loader(tmp_path, reference):
    """Every malformed manifest above fails in the reference's own loader
    with the same message (the reference loads it through its swarm driver,
    whose first step is load_gas_dataset)."""
    for k, (old, new, _) in enumerate(BAD_MANIFESTS):
        d = tmp_path / f"case{k}"
        d.mkdir()
        gas = _variant(str(d), os.path.join(GAS_DIR, "synthetic_ionization.gas"),
                       lambda t: t.replace(old, new, 1))
        with pytest.raises(gasdata.GasDataError) as mine:
            gasdata.load_gas_dataset(gas)
        swarm = d / "probe.swarm"
        swarm.write_text("swarm_config_version = 1\ngas_data_file = synthetic_ionization.gas\n"
                         "neutral_density = 1e20\nreduced_fields_td = 1\nmax_frequency = 1e7\n"
                         "timestep = 1e-9\nsteps = 10\nburn_in_steps = 5\nparticles = 4\n"
                         "max_energy_ev = 10\n")
        with pytest.raises(RuntimeError) as theirs:
            reference.swarm_file(str(swarm))
        assert str(mine.value) in str(theirs.value), (str(mine.value), str(theirs.value))


def test_tables_reject_reference_cases(tmp_path):
    t = tmp_path / "t.dat"
    t.write_text("# e s\n0 1\n1 2 3\n")
    with pytest.raises(gasdata.GasDataError, match="has trailing columns"):
        gasdata.read_cross_section_table(str(t))
    t.write_text("0 1\n0 2\n")
    with pytest.raises(gasdata.GasDataError, match="strictly increasing"):
        gasdata.read_cross_section_table(str(t))
    t.write_text("0 1\n")
    with pytest.raises(gasdata.GasDataError, match="at least two rows"):
        gasdata.read_cross_section_table(str(t))
    t.write_text("0 0.1\n1 1.0\n")
    with pytest.raises(gasdata.GasDataError, match="strictly between -1 and 1"):
        gasdata.read_mean_cosine_table(str(t))
    t.write_text("0 1e-20 ; comment\n2 2e-20\n")
    e, s = gasdata.read_cross_section_table(str(t), 2.0, 3.0)
    np.testing.assert_array_equal(e, [0.0, 4.0])
    np.testing.assert_array_equal(s, [1e-20 * 3.0, 2e-20 * 3.0])


def test_swarm_config_example():
    cfg = gasdata.load_swarm_config(os.path.join(SWARM_DIR, "synthetic_swarm.swarm"))
    assert cfg.reduced_fields_td == [1.0, 5.0, 10.0]
    assert cfg.neutral_temperature == 300.0 and cfg.neutral_density == 1.0e20
    assert (cfg.steps, cfg.burn_in_steps, cfg.particles, cfg.uncertainty_blocks) == \
        (2000, 1000, 256, 10)
    assert cfg.seed == 271828 and cfg.max_energy_ev == 15.0
    assert os.path.basename(cfg.gas_data_file) == "synthetic_swarm.gas"
    reqs = gasdata.swarm_requests(cfg)
    assert [r.seed for r in reqs] == [271828, 271829, 271830]
    assert [r.reduced_field_td for r in reqs] == [1.0, 5.0, 10.0]


@pytest.mark.parametrize("line,message", [
    ("steps = -5", "key 'steps' must not be negative"),
    ("steps = 2000x", "key 'steps' has trailing content"),
    ("steps = many", "key 'steps' must be numeric"),
    ("reduced_fields_td = 1,,5", "reduced_fields_td contains an empty value"),
    ("colour = red", "unknown key 'colour'"),
    ("population_model = exotic", "must be fixed_population_no_avalanche or"),
])
def test_swarm_config_errors(tmp_path, line, message):
    key = line.split("=")[0].strip()
    src = os.path.join(SWARM_DIR, "synthetic_swarm.swarm")

    def edit(text):
        lines = [ln for ln in text.splitlines() if not ln.startswith(key + " ")]
        return "\n".join(lines + [line]) + "\n"
    path = _variant(str(tmp_path), src, edit)
    with pytest.raises(gasdata.GasDataError, match=message):
        gasdata.load_swarm_config(path)


def example_model():
    cfg = gasdata.load_swarm_config(os.path.join(SWARM_DIR, "synthetic_swarm.swarm"))
    ds = gasdata.load_gas_dataset(cfg.gas_data_file)
    gas = gasdata.swarm_gas(ds, cfg)
    opts = {k: (c.angular, c.energy_frame, c.mean_cosine_energy, c.mean_cosine)
            for k, c in enumerate(ds.channels)}
    vstd = math.sqrt(KB * cfg.neutral_temperature / ds.neutral_mass)
    return cfg, ds, gas, oracle_py.model_from_gas(gas, neutral_vstd=vstd, options=opts)


def check_against_reference(results, tolerance_sigma=4.0):
    """results: per field (drift, mean energy, elastic collisions,
    candidates, longitudinal and transverse diffusion).  The reference's spread over 16 seeds sets the scale; the
    mean of 16 runs carries its own 1/sqrt(16) error."""
    g = golden("swarm/synthetic_swarm_golden.npz")
    cols = list(g["columns"])
    mean, std = g["seed_mean"], g["seed_std"]
    n = int(g["seeds"])
    for f, got in enumerate(results):
        for name, value in zip(("drift", "mean_energy_ev", "collisions0", "candidates",
                                "longitudinal_diffusion", "transverse_diffusion"), got):
            c = cols.index(name)
            scale = std[f, c] * math.sqrt(1.0 + 1.0 / n)
            assert abs(value - mean[f, c]) <= tolerance_sigma * scale, \
                (f, name, value, mean[f, c], std[f, c])


def test_oracle_swarm_matches_reference_example(oracle):
    cfg, ds, gas, model = example_model()
    sigma = math.sqrt(2.0 * cfg.initial_mean_energy_ev * EV / (3.0 * 9.1093837139e-31))
    results = []
    for f, td in enumerate(cfg.reduced_fields_td):
        rng = np.random.default_rng(cfg.seed + f)
        vel0 = rng.normal(0.0, sigma, (cfg.particles, 3))
        c = dict(neutral_density=cfg.neutral_density, max_frequency=cfg.max_frequency,
                 timestep=cfg.timestep, reduced_field_td=td, steps=cfg.steps,
                 burn_in=cfg.burn_in_steps, particles=cfg.particles,
                 blocks=cfg.uncertainty_blocks,
                 initial_mean_energy_ev=cfg.initial_mean_energy_ev,
                 max_energy_ev=cfg.max_energy_ev)
        r = oracle.swarm(model, c, vel0, cfg.seed + f, 0)
        assert r.status == 0
        results.append((r.drift, r.mean_energy_ev, r.collisions[0], r.candidates,
                        r.longitudinal_diffusion, r.transverse_diffusion))
    check_against_reference(results)


def test_neutral_temperature_needs_si_dataset():
    cfg = gasdata.load_swarm_config(os.path.join(SWARM_DIR, "synthetic_swarm.swarm"))
    ds = gasdata.load_gas_dataset(os.path.join(GAS_DIR, "synthetic_ionization.gas"))
    with pytest.raises(gasdata.GasDataError, match="requires an SI gas dataset"):
        gasdata.swarm_gas(ds, cfg)


# ------------------------------------------------------------------ CSV
def _reference_csv():
    import csv
    with open(os.path.join(SWARM_DIR, "synthetic_swarm_reference.csv")) as f:
        rows = list(csv.reader(f))
    return rows[0], rows[1:]


def _result_from_reference_row(head, row, nchan):
    """The statistics of one reference CSV row in the device result's shape."""
    from types import SimpleNamespace
    col = {name: row[i] for i, name in enumerate(head)}
    chans = [n[:-len("_collisions")] for n in head if n.endswith("_collisions")
             and n != "collision_candidates" and n != "null_collisions"]
    return SimpleNamespace(
        drift=float(col["electron_drift_velocity_m_s"]),
        drift_se=float(col["mean_velocity_x_standard_error_m_s"]),
        mean_energy_ev=float(col["mean_energy_ev"]),
        mean_energy_se=float(col["mean_energy_standard_error_ev"]),
        longitudinal_diffusion=float(col["longitudinal_diffusion_m2_s"]),
        transverse_diffusion=float(col["transverse_diffusion_m2_s"]),
        max_energy_ev=float(col["maximum_observed_energy_ev"]),
        candidates=int(col["collision_candidates"]), nulls=int(col["null_collisions"]),
        collisions=[int(col[f"{c}_collisions"]) for c in chans],
        rate=[float(col[f"{c}_rate_per_electron_s"]) for c in chans],
        rate_se=[float(col[f"{c}_rate_standard_error_s"]) for c in chans])


def _strip_paths(cells, head):
    out = list(cells)
    for name in ("gas_data_file", "spatial_profile_file"):
        i = head.index(name)
        out[i] = os.path.basename(out[i])
    return out


def test_csv_writer_reproduces_reference_csv(tmp_path):
    """Given the reference run's own statistics, the writer reproduces the
    reference's CSV cell for cell: header, run description, the collision
    model signature (7215415709351068229), derived fields (E, mobility,
    rates, errors) and the fixed-population columns."""
    from aurora_pic import swarmio
    head, body = _reference_csv()
    cfg = gasdata.load_swarm_config(os.path.join(SWARM_DIR, "synthetic_swarm.swarm"))
    ds = gasdata.load_gas_dataset(cfg.gas_data_file)
    assert swarmio.header(ds) == head
    assert swarmio.collision_model_signature(ds, cfg) == 7215415709351068229
    results = [_result_from_reference_row(head, row, len(ds.channels)) for row in body]
    out = str(tmp_path / "swarm.csv")
    swarmio.write_swarm_benchmark_csv(out, cfg, ds, results)
    import csv
    with open(out) as f:
        mine = list(csv.reader(f))
    assert mine[0] == head
    for got, want in zip(mine[1:], body):
        assert _strip_paths(got, head) == _strip_paths(want, head)


def test_model_signature_tracks_physics():
    from aurora_pic import swarmio
    cfg = gasdata.load_swarm_config(os.path.join(SWARM_DIR, "synthetic_swarm.swarm"))
    ds = gasdata.load_gas_dataset(cfg.gas_data_file)
    base = swarmio.collision_model_signature(ds, cfg)
    cfg.neutral_temperature = 0.0
    assert swarmio.collision_model_signature(ds, cfg) != base
    cfg.neutral_temperature = 300.0
    ds.channels[0].cross_section = ds.channels[0].cross_section * (1 + 1e-15)
    assert swarmio.collision_model_signature(ds, cfg) != base


def _example_config():
    cfg = gasdata.load_swarm_config(os.path.join(SWARM_DIR, "synthetic_swarm.swarm"))
    return cfg, gasdata.load_gas_dataset(cfg.gas_data_file)


def test_swarm_validator_accepts_the_reference_examples():
    cfg, ds = _example_config()
    gasdata.validate_swarm_config(cfg, ds)
    b = gasdata.load_swarm_config(os.path.join(GOLDEN, "swarm_branching", "branching.swarm"))
    gasdata.validate_swarm_config(b, gasdata.load_gas_dataset(b.gas_data_file))


@pytest.mark.parametrize("change,message", [
    # Swarm.cpp:336-359: the work-item gates
    (dict(work_item_limit=0), "swarm work_item_limit must be positive"),
    (dict(work_item_limit=3 * 256 * 2000 - 1), "swarm scan exceeds work_item_limit"),
    (dict(work_item_limit=3 * 256 - 1), "swarm scan exceeds work_item_limit"),
    # Swarm.cpp:317-335: population limits
    (dict(population_limit=1000), "population_limit is valid only for population_model = "
                                  "branching_resampled"),
    (dict(population_model="branching_resampled", population_limit=256),
     "branching swarm population_limit must exceed particles"),
    (dict(population_model="branching_resampled", population_limit=10_000_000 - 255),
     "branching swarm population_limit must exceed particles"),
    (dict(particles=10_000_001), "swarm particles exceeds the safety limit of 10000000"),
    (dict(particles=0), "swarm particles must be positive"),
    (dict(timestep=1.0), "max_frequency \\* timestep must not exceed 0.1"),
    (dict(burn_in_steps=2000), "swarm steps must exceed burn_in_steps"),
    (dict(uncertainty_blocks=0), "uncertainty_blocks must be between 1 and the number"),
    (dict(uncertainty_blocks=7), "sampling steps must be divisible by uncertainty_blocks"),
    (dict(reduced_fields_td=[1.0, 1.0]), "must not contain duplicates"),
    (dict(reduced_fields_td=[]), "at least one reduced field"),
    (dict(max_energy_ev=1e9), "swarm max_energy_ev exceeds channel"),
    (dict(spatial_bins=8), "spatial Townsend controls require positive spatial_histories"),
])
def test_swarm_validator_messages(change, message):
    cfg, ds = _example_config()
    for k, v in change.items():
        setattr(cfg, k, v)
    with pytest.raises(gasdata.GasDataError, match=message):
        gasdata.validate_swarm_config(cfg, ds)


def test_swarm_validator_work_items_boundary():
    # fields x particles x steps equal to the limit passes, one more fails
    cfg, ds = _example_config()
    cfg.work_item_limit = 3 * 256 * 2000
    gasdata.validate_swarm_config(cfg, ds)
    cfg.steps += 10
    with pytest.raises(gasdata.GasDataError, match="work_item_limit"):
        gasdata.validate_swarm_config(cfg, ds)


def test_serving_work_item_gate():
    # the reference's per-scan gate applied to one served field point
    from aurora_pic import hip, serving
    r = hip.SwarmRequest(particles=50_000, steps=2000)
    assert serving.work_items_exceeded(r) == ""                  # 1e8 exactly
    r.steps = 2001
    assert "swarm scan exceeds work_item_limit" in serving.work_items_exceeded(r)
    assert serving.work_items_exceeded(r, 10 ** 12) == ""
    assert "must be positive" in serving.work_items_exceeded(r, 0)
