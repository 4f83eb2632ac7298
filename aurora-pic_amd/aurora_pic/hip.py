"""ctypes binding of libaurora_hip.so (the C ABI in include/aurora_hip.h).

Fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libaurora_hip.so")
HEADER = os.path.join(REPO, "include", "aurora_hip.h")

PERIODIC, DIRICHLET = 0, 1
PB_ABSORB, PB_REFLECT, PB_PERIODIC = 1, 2, 3
VELOCITY_HALF, VELOCITY_CENTRED = 0, 1
ELASTIC, EXCITATION, IONIZATION, ATTACHMENT, CHARGE_EXCHANGE = range(5)
SCATTER_ISOTROPIC, SCATTER_BACKWARD, SCATTER_HENYEY_GREENSTEIN = range(3)
FRAME_PROJECTILE, FRAME_CENTER_OF_MASS = range(2)
SORT_FIXED, SORT_ADAPTIVE = range(2)
SORT_LOCAL, SORT_RADIX = range(2)
(OK, ERR_INVALID_ARGUMENT, ERR_HIP, ERR_RATE_ABOVE_BOUND, ERR_CANDIDATE_LIMIT, ERR_CAPACITY,
 ERR_B_OUT_OF_RANGE, ERR_ENERGY_ABOVE_MAX, ERR_UNSUPPORTED, ERR_NOT_FOUND,
 ERR_POPULATION) = range(11)

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_ip = C.POINTER(C.c_int)


class HipError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class SwarmRequest(C.Structure):
    _fields_ = [
        ("reduced_field_td", C.c_double),
        ("timestep", C.c_double),
        ("particles", C.c_int64),
        ("steps", C.c_int64),
        ("burn_in", C.c_int64),
        ("blocks", C.c_int64),
        ("initial_mean_energy_ev", C.c_double),
        ("max_energy_ev", C.c_double),
        ("seed", C.c_uint64),
    ]


class SwarmResult(C.Structure):
    _fields_ = [
        ("drift", C.c_double),
        ("drift_se", C.c_double),
        ("mean_energy_ev", C.c_double),
        ("mean_energy_se", C.c_double),
        ("ionization_rate", C.c_double),
        ("ionization_rate_se", C.c_double),
        ("rate", C.c_double * 8),
        ("rate_se", C.c_double * 8),
        ("collisions", C.c_uint64 * 8),
        ("candidates", C.c_uint64),
        ("nulls", C.c_uint64),
        ("max_energy_ev", C.c_double),
        ("status", C.c_int),
        ("longitudinal_diffusion", C.c_double),
        ("transverse_diffusion", C.c_double),
    ]


class SwarmGrowth(C.Structure):
    _fields_ = [
        ("initial_weight", C.c_double),
        ("final_weight", C.c_double),
        ("final_particles", C.c_int64),
        ("growth_rate", C.c_double),
        ("growth_rate_se", C.c_double),
        ("attachment_rate", C.c_double),
        ("attachment_rate_se", C.c_double),
        ("net_creation_rate", C.c_double),
        ("net_creation_rate_se", C.c_double),
        ("townsend_available", C.c_int),
        ("townsend_growth", C.c_double),
        ("townsend_growth_se", C.c_double),
        ("townsend_rate_balance", C.c_double),
        ("townsend_rate_balance_se", C.c_double),
    ]


def declared_symbols(header: str = HEADER):
    """Function names declared in include/aurora_hip.h."""
    text = open(header).read()
    return sorted(set(re.findall(r"\b(aurora_hip_\w+)\s*\(", text)))


_LIB = None


def load(path: str = LIB_PATH):
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python aurora-pic_amd/build.py` "
            "(there is no CPU fallback)")
    L = C.CDLL(path)
    ctx = C.c_void_p
    sig = {
        "aurora_hip_create": [C.c_int, C.POINTER(C.c_void_p)],
        "aurora_hip_destroy": [ctx],
        "aurora_hip_synchronize": [ctx],
        "aurora_hip_device_info": [ctx, C.c_char_p, C.c_size_t],
        "aurora_hip_mesh_bind": [ctx, C.c_int, _i64p, _dp, C.c_int, C.c_double, C.c_double],
        "aurora_hip_particle_boundaries": [ctx, _ip],
        "aurora_hip_species_add": [ctx, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_int,
                                   _ip],
        "aurora_hip_species_upload": [ctx, C.c_int, C.c_int64, _dp, _dp, _u8p, C.c_int],
        "aurora_hip_species_init_maxwellian": [ctx, C.c_int, C.c_int64, C.c_double, _dp,
                                               C.c_uint64],
        "aurora_hip_species_count": [ctx, C.c_int, _i64p, _i64p],
        "aurora_hip_species_download": [ctx, C.c_int, C.c_int64, _dp, _dp, _u8p, _u32p, _i64p],
        "aurora_hip_species_download_raw": [ctx, C.c_int, C.c_int64, _i32p, _dp, _dp, _i64p],
        "aurora_hip_species_velocity": [ctx, C.c_int, C.c_double, C.c_int64, _dp],
        "aurora_hip_species_charge_fixed": [ctx, C.c_int, _i64p],
        "aurora_hip_set_uniform_b": [ctx, _dp],
        "aurora_hip_set_b_profile": [ctx, C.c_int, C.c_int64, _dp, _dp],
        "aurora_hip_set_electrodes": [ctx, _dp, _dp],
        "aurora_hip_set_electric": [ctx, _dp],
        "aurora_hip_get_fields": [ctx, _dp, _dp, _dp],
        "aurora_hip_set_rho": [ctx, _dp],
        "aurora_hip_deposit": [ctx],
        "aurora_hip_field_solve": [ctx, C.c_double, C.c_double, C.c_int],
        "aurora_hip_push": [ctx, C.c_double],
        "aurora_hip_initialize": [ctx, C.c_double],
        "aurora_hip_step": [ctx, C.c_double, C.c_int64],
        "aurora_hip_energy": [ctx, C.c_double, _dp, _dp, _dp, _i64p],
        "aurora_hip_boundary_losses": [ctx, C.c_int, _i64p, _dp],
        "aurora_hip_power_transfer": [ctx, C.c_int, _dp],
        "aurora_hip_time": [ctx, _dp, _i64p],
        "aurora_hip_set_sort_interval": [ctx, C.c_int64],
        "aurora_hip_set_sort_threshold": [ctx, C.c_double],
        "aurora_hip_set_sort_policy": [ctx, C.c_int, C.c_int64],
        "aurora_hip_set_sort_method": [ctx, C.c_int],
        "aurora_hip_checkpoint_save": [ctx, C.c_char_p],
        "aurora_hip_checkpoint_load": [ctx, C.c_char_p],
        "aurora_hip_physics_signature": [ctx, _u64p],
        "aurora_hip_sort": [ctx],
        "aurora_hip_set_tiling": [ctx, C.c_int, C.c_int, C.c_int, C.c_int],
        "aurora_hip_set_push_variant": [ctx, C.c_int, C.c_int],
        "aurora_hip_mcc_load": [ctx, C.c_int, C.c_int, _ip, _dp, _i64p, C.POINTER(_dp),
                                C.POINTER(_dp), C.c_double, C.c_double, C.c_int64, C.c_double,
                                C.c_double, _ip, C.c_uint64],
        "aurora_hip_mcc_stats": [ctx, _u64p, _dp],
        "aurora_hip_mcc_add": [ctx, C.c_int, C.c_int, _ip, _dp, _i64p, C.POINTER(_dp),
                               C.POINTER(_dp), C.c_double, C.c_double, C.c_int64, C.c_double,
                               C.c_double, _ip, C.c_uint64, _ip],
        "aurora_hip_mcc_model_options": [ctx, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64,
                                         _dp, _dp],
        "aurora_hip_mcc_model_stats": [ctx, C.c_int, _u64p, _dp],
        "aurora_hip_set_velocity_dimensions": [ctx, C.c_int],
        "aurora_hip_mcc_channel_options": [ctx, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64,
                                           _dp, _dp],
        "aurora_hip_mcc_collide": [ctx, C.c_int64, _dp, C.c_double, C.c_uint64, C.c_uint32,
                                   _u64p, _dp],
        "aurora_hip_mcc_collide_device": [ctx, C.c_int64, C.c_void_p, C.c_double, C.c_uint64,
                                          C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p],
        "aurora_hip_swarm_gas": [ctx, C.c_int, _ip, _dp, _i64p, C.POINTER(_dp), C.POINTER(_dp),
                                 C.c_double, C.c_double, C.c_int64, C.c_double, C.c_double],
        "aurora_hip_swarm_submit": [ctx, C.POINTER(SwarmRequest), C.c_uint64, _i64p],
        "aurora_hip_swarm_advance": [ctx, C.c_int64, _i64p],
        "aurora_hip_swarm_launch": [ctx, C.c_int64, _i64p],
        "aurora_hip_swarm_collect": [ctx, C.c_int, _ip],
        "aurora_hip_swarm_poll": [ctx, C.c_int64, _ip, C.POINTER(SwarmResult)],
        "aurora_hip_swarm_branching": [ctx, C.POINTER(SwarmRequest), C.c_int64,
                                       C.POINTER(SwarmResult), C.POINTER(SwarmGrowth)],
        "aurora_hip_ensemble_create": [ctx, C.c_int64, _i64p, _i64p],
        "aurora_hip_ensemble_load": [ctx, C.c_int64, C.c_int64, C.c_uint64],
        "aurora_hip_ensemble_graph": [ctx, C.c_int64, C.c_int],
        "aurora_hip_ensemble_step": [ctx, C.c_int64, C.c_double, C.c_int64],
        "aurora_hip_ensemble_count": [ctx, C.c_int64, C.c_int64, C.c_int, _i64p, _i64p],
        "aurora_hip_ensemble_download": [ctx, C.c_int64, C.c_int64, C.c_int, C.c_int64, _dp, _dp,
                                         _u8p, _u32p, _i64p],
        "aurora_hip_ensemble_stats": [ctx, C.c_int64, C.c_int64, _u64p, _dp, _u64p, _dp],
        "aurora_hip_ensemble_fields": [ctx, C.c_int64, C.c_int64, _dp, _dp, _dp],
        "aurora_hip_ensemble_time": [ctx, C.c_int64, _dp, _i64p],
        "aurora_hip_ensemble_destroy": [ctx, C.c_int64],
        "aurora_hip_kernel_timing": [ctx, C.c_int],
        "aurora_hip_stream_bandwidth": [ctx, C.c_int, C.c_int64, C.c_int, _dp],
        "aurora_hip_kernel_time": [ctx, C.c_char_p, _dp, _i64p],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.aurora_hip_last_error.argtypes = [ctx]
    L.aurora_hip_last_error.restype = C.c_char_p
    L.aurora_hip_swarm_footprint.argtypes = [C.POINTER(SwarmRequest)]
    L.aurora_hip_swarm_footprint.restype = C.c_uint64
    _LIB = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One device context: a mesh, its species, fields and MCC/swarm state."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.aurora_hip_create(device, C.byref(h))
        if rc:
            raise HipError(rc, "aurora_hip_create failed (no usable HIP device?)")
        self.h = h
        self.dim = 0
        self.nnodes = 0
        self.species = []

    def close(self):
        if self.h:
            self.lib.aurora_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc:
            raise HipError(rc, self.lib.aurora_hip_last_error(self.h).decode())

    def info(self):
        buf = C.create_string_buffer(256)
        self._check(self.lib.aurora_hip_device_info(self.h, buf, 256))
        return buf.value.decode()

    # ---- mesh / species
    def mesh(self, n, L, bc=PERIODIC, eps=1.0, depth=1.0):
        n = np.ascontiguousarray(n, dtype=np.int64)
        L = _f64(L)
        self._check(self.lib.aurora_hip_mesh_bind(self.h, len(n), _p(n, _i64p), _p(L, _dp), bc,
                                                  eps, depth))
        self.dim = len(n)
        self.nnodes = int(np.prod(n))
        self.species = []

    def boundaries(self, policy):
        pol = (C.c_int * 6)(*(list(policy) + [3] * (6 - len(policy))))
        self._check(self.lib.aurora_hip_particle_boundaries(self.h, pol))

    def add_species(self, charge, mass, weight, capacity, precision=32):
        sid = C.c_int()
        self._check(self.lib.aurora_hip_species_add(self.h, charge, mass, weight, int(capacity),
                                                     precision, C.byref(sid)))
        self.species.append(sid.value)
        return sid.value

    def upload(self, sid, pos, vel, alive=None, kind=VELOCITY_HALF):
        pos = _f64(pos).reshape(-1)
        vel = _f64(vel).reshape(-1)
        count = vel.size // 3
        al = None if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
        self._check(self.lib.aurora_hip_species_upload(self.h, sid, count, _p(pos, _dp),
                                                       _p(vel, _dp),
                                                       None if al is None else _p(al, _u8p),
                                                       kind))

    def init_maxwellian(self, sid, count, vth, seed, drift=(0.0, 0.0, 0.0)):
        d = _f64(drift)
        self._check(self.lib.aurora_hip_species_init_maxwellian(self.h, sid, int(count), vth,
                                                                _p(d, _dp), seed))

    def count(self, sid):
        s, l = C.c_int64(), C.c_int64()
        self._check(self.lib.aurora_hip_species_count(self.h, sid, C.byref(s), C.byref(l)))
        return s.value, l.value

    def download(self, sid):
        slots, _ = self.count(sid)
        pos = np.zeros(max(slots, 1) * self.dim)
        vel = np.zeros(max(slots, 1) * 3)
        alive = np.zeros(max(slots, 1), dtype=np.uint8)
        uid = np.zeros(max(slots, 1), dtype=np.uint32)
        n = C.c_int64()
        self._check(self.lib.aurora_hip_species_download(self.h, sid, max(slots, 1),
                                                         _p(pos, _dp), _p(vel, _dp),
                                                         _p(alive, _u8p), _p(uid, _u32p),
                                                         C.byref(n)))
        k = n.value
        return (pos[:k * self.dim].reshape(k, self.dim), vel[:k * 3].reshape(k, 3), alive[:k],
                uid[:k])

    def download_raw(self, sid):
        slots, _ = self.count(sid)
        cell = np.zeros(max(slots, 1), dtype=np.int32)
        off = np.zeros(max(slots, 1) * self.dim)
        vel = np.zeros(max(slots, 1) * 3)
        n = C.c_int64()
        self._check(self.lib.aurora_hip_species_download_raw(self.h, sid, max(slots, 1),
                                                             _p(cell, _i32p), _p(off, _dp),
                                                             _p(vel, _dp), C.byref(n)))
        k = n.value
        return cell[:k], off[:k * self.dim].reshape(k, self.dim), vel[:k * 3].reshape(k, 3)

    def velocity(self, sid, dt):
        slots, _ = self.count(sid)
        v = np.zeros(max(slots, 1) * 3)
        self._check(self.lib.aurora_hip_species_velocity(self.h, sid, dt, max(slots, 1),
                                                         _p(v, _dp)))
        return v[:slots * 3].reshape(slots, 3)

    def charge_fixed(self, sid):
        acc = np.zeros(self.nnodes, dtype=np.int64)
        self._check(self.lib.aurora_hip_species_charge_fixed(self.h, sid, _p(acc, _i64p)))
        return acc

    # ---- fields
    def uniform_b(self, b):
        b = _f64(b)
        self._check(self.lib.aurora_hip_set_uniform_b(self.h, _p(b, _dp)))

    def b_profile(self, axis, coords, values):
        coords, values = _f64(coords), _f64(values).reshape(-1)
        self._check(self.lib.aurora_hip_set_b_profile(self.h, axis, coords.size,
                                                      _p(coords, _dp), _p(values, _dp)))

    def electrodes(self, left4, right4):
        l4, r4 = _f64(left4), _f64(right4)
        self._check(self.lib.aurora_hip_set_electrodes(self.h, _p(l4, _dp), _p(r4, _dp)))

    def set_electric(self, E):
        E = _f64(E).reshape(-1)
        self._check(self.lib.aurora_hip_set_electric(self.h, _p(E, _dp)))

    def set_rho(self, rho):
        rho = _f64(rho)
        self._check(self.lib.aurora_hip_set_rho(self.h, _p(rho, _dp)))

    def fields(self):
        rho = np.zeros(self.nnodes)
        phi = np.zeros(self.nnodes)
        E = np.zeros(self.nnodes * self.dim)
        self._check(self.lib.aurora_hip_get_fields(self.h, _p(rho, _dp), _p(phi, _dp),
                                                   _p(E, _dp)))
        return rho, phi, E.reshape(self.dim, self.nnodes)

    def rho(self):
        """Node charge density alone (no field solve: also on meshes whose
        solve is not built, such as 2D/3D Dirichlet)."""
        rho = np.zeros(self.nnodes)
        self._check(self.lib.aurora_hip_get_fields(self.h, _p(rho, _dp), None, None))
        return rho

    # ---- operations
    def deposit(self):
        self._check(self.lib.aurora_hip_deposit(self.h))

    def solve(self, phi_left=0.0, phi_right=0.0, want_phi=True):
        self._check(self.lib.aurora_hip_field_solve(self.h, phi_left, phi_right, int(want_phi)))

    def push(self, dt):
        self._check(self.lib.aurora_hip_push(self.h, dt))

    def initialize(self, dt):
        self._check(self.lib.aurora_hip_initialize(self.h, dt))

    def step(self, dt, n=1):
        self._check(self.lib.aurora_hip_step(self.h, dt, n))

    def sync(self):
        self._check(self.lib.aurora_hip_synchronize(self.h))

    def energy(self, dt):
        ke = np.zeros(max(len(self.species), 1))
        fe, l1 = C.c_double(), C.c_double()
        live = C.c_int64()
        self._check(self.lib.aurora_hip_energy(self.h, dt, _p(ke, _dp), C.byref(fe), C.byref(l1),
                                               C.byref(live)))
        return ke[:len(self.species)], fe.value, l1.value, live.value

    def losses(self, sid):
        a = np.zeros(6, dtype=np.int64)
        e = np.zeros(2)
        self._check(self.lib.aurora_hip_boundary_losses(self.h, sid, _p(a, _i64p), _p(e, _dp)))
        return a, e

    def power_transfer(self, sid):
        w = C.c_double()
        self._check(self.lib.aurora_hip_power_transfer(self.h, sid, C.byref(w)))
        return w.value

    def time(self):
        t, s = C.c_double(), C.c_int64()
        self._check(self.lib.aurora_hip_time(self.h, C.byref(t), C.byref(s)))
        return t.value, s.value

    def sort_interval(self, k):
        self._check(self.lib.aurora_hip_set_sort_interval(self.h, k))

    def tiling(self, tx, ty=0, tz=0, halo=0):
        self._check(self.lib.aurora_hip_set_tiling(self.h, tx, ty, tz, halo))

    def push_variant(self, block, prefetch):
        self._check(self.lib.aurora_hip_set_push_variant(self.h, block, prefetch))

    def sort_threshold(self, fraction):
        self._check(self.lib.aurora_hip_set_sort_threshold(self.h, fraction))

    def checkpoint_save(self, path):
        self._check(self.lib.aurora_hip_checkpoint_save(self.h, os.fsencode(path)))

    def checkpoint_load(self, path):
        self._check(self.lib.aurora_hip_checkpoint_load(self.h, os.fsencode(path)))

    def physics_signature(self):
        v = C.c_uint64()
        self._check(self.lib.aurora_hip_physics_signature(self.h, C.byref(v)))
        return v.value

    def sort_policy(self, policy, max_interval=0):
        """SORT_FIXED (every sort_interval steps) or SORT_ADAPTIVE (per
        species, when its timed push exceeds its cycle's average cost)."""
        self._check(self.lib.aurora_hip_set_sort_policy(self.h, policy, max_interval))

    def sort_method(self, method):
        """SORT_LOCAL (counting sort of a tile-sorted species, the default)
        or SORT_RADIX (always the full radix sort)."""
        self._check(self.lib.aurora_hip_set_sort_method(self.h, method))

    def sort(self):
        self._check(self.lib.aurora_hip_sort(self.h))

    # ---- timing
    def stream_bandwidth(self, pattern, nbytes=2 << 30, repeats=5):
        """Attainable GB/s of a hand-written streaming kernel (0: copy, 1:
        seven arrays read and rewritten in place)."""
        g = C.c_double()
        self._check(self.lib.aurora_hip_stream_bandwidth(self.h, pattern, nbytes, repeats,
                                                         C.byref(g)))
        return g.value

    def timing(self, enable=True):
        self._check(self.lib.aurora_hip_kernel_timing(self.h, int(enable)))

    def kernel_time(self, name):
        ms, n = C.c_double(), C.c_int64()
        self._check(self.lib.aurora_hip_kernel_time(self.h, name.encode(), C.byref(ms),
                                                    C.byref(n)))
        return ms.value, n.value

    # ---- MCC
    def _tables(self, channels):
        nch = len(channels)
        kinds = (C.c_int * nch)(*[int(c[0]) for c in channels])
        thr = _f64([c[1] for c in channels])
        rows = np.ascontiguousarray([len(c[2]) for c in channels], dtype=np.int64)
        self._keep = [(_f64(c[2]), _f64(c[3])) for c in channels]
        ep = (_dp * nch)(*[_p(e, _dp) for e, _ in self._keep])
        sp = (_dp * nch)(*[_p(s, _dp) for _, s in self._keep])
        return nch, kinds, thr, rows, ep, sp

    def mcc_load(self, target, gas, product_species=(-1, -1), seed=1, neutral_temperature=0.0):
        nch, kinds, thr, rows, ep, sp = self._tables(gas["channels"])
        prod = (C.c_int * 2)(*product_species)
        self._check(self.lib.aurora_hip_mcc_load(
            self.h, target, nch, kinds, _p(thr, _dp), _p(rows, _i64p), ep, sp,
            gas["neutral_density"], gas["max_frequency"], gas.get("max_candidates", 64),
            gas["neutral_mass"], neutral_temperature, prod, seed))
        self._nchan = nch

    def mcc_channel_options(self, channel, angular=0, frame=0, mc_energy=None, mc_value=None,
                            model=0):
        """CollisionChannelConfig angular model / energy frame of one channel of
        the context's MCC model (model=0) or the swarm gas (model=1)."""
        if mc_energy is None:
            self._check(self.lib.aurora_hip_mcc_channel_options(self.h, model, channel, angular,
                                                                frame, 0, None, None))
            return
        e, v = _f64(mc_energy), _f64(mc_value)
        self._check(self.lib.aurora_hip_mcc_channel_options(self.h, model, channel, angular,
                                                            frame, len(e), _p(e, _dp),
                                                            _p(v, _dp)))

    def mcc_add(self, target, gas, product_species=(-1, -1), seed=1, neutral_temperature=0.0):
        """A further named null-collision model (after mcc_load); returns its
        model index."""
        nch, kinds, thr, rows, ep, sp = self._tables(gas["channels"])
        prod = (C.c_int * 2)(*product_species)
        model = C.c_int()
        self._check(self.lib.aurora_hip_mcc_add(
            self.h, target, nch, kinds, _p(thr, _dp), _p(rows, _i64p), ep, sp,
            gas["neutral_density"], gas["max_frequency"], gas.get("max_candidates", 64),
            gas["neutral_mass"], neutral_temperature, prod, seed, C.byref(model)))
        self._model_nchan = getattr(self, "_model_nchan", {})
        self._model_nchan[model.value] = nch
        return model.value

    def mcc_model_stats(self, model):
        nch = self._nchan if model == 0 else self._model_nchan[model]
        counts = np.zeros(3 + nch, dtype=np.uint64)
        ech = np.zeros(nch)
        self._check(self.lib.aurora_hip_mcc_model_stats(self.h, model, _p(counts, _u64p),
                                                        _p(ech, _dp)))
        return counts, ech

    def velocity_dimensions(self, dims):
        self._check(self.lib.aurora_hip_set_velocity_dimensions(self.h, dims))

    def mcc_stats(self):
        counts = np.zeros(3 + self._nchan, dtype=np.uint64)
        ech = np.zeros(self._nchan)
        self._check(self.lib.aurora_hip_mcc_stats(self.h, _p(counts, _u64p), _p(ech, _dp)))
        return counts, ech

    def mcc_collide(self, vel, dt, step, stream_id=0):
        v = _f64(vel).copy().reshape(-1)
        counts = np.zeros(3 + self._nchan, dtype=np.uint64)
        ech = np.zeros(self._nchan)
        self._check(self.lib.aurora_hip_mcc_collide(self.h, v.size // 3, _p(v, _dp), dt, step,
                                                    stream_id, _p(counts, _u64p),
                                                    _p(ech, _dp)))
        return v.reshape(-1, 3), counts, ech

    def mcc_collide_device(self, count, velocity, dt, step, counts, energy_change, status,
                           stream_id=0):
        """aurora_hip_mcc_collide_device: every pointer is a device address
        (int); queued on the context stream, nothing allocated."""
        self._check(self.lib.aurora_hip_mcc_collide_device(self.h, count, velocity, dt, step,
                                                           stream_id, counts, energy_change,
                                                           status))

    # ---- swarms
    def swarm_gas(self, gas):
        nch, kinds, thr, rows, ep, sp = self._tables(gas["channels"])
        self._check(self.lib.aurora_hip_swarm_gas(
            self.h, nch, kinds, _p(thr, _dp), _p(rows, _i64p), ep, sp, gas["neutral_density"],
            gas["max_frequency"], gas.get("max_candidates", 64), gas["neutral_mass"],
            gas.get("neutral_temperature", 0.0)))
        self._nchan = nch

    def swarm_submit(self, request: SwarmRequest, hbm_budget=0):
        h = C.c_int64()
        self._check(self.lib.aurora_hip_swarm_submit(self.h, C.byref(request), hbm_budget,
                                                     C.byref(h)))
        return h.value

    def swarm_advance(self, max_steps):
        a = C.c_int64()
        self._check(self.lib.aurora_hip_swarm_advance(self.h, max_steps, C.byref(a)))
        return a.value

    def swarm_launch(self, max_steps):
        """Queues one swarm launch and returns at once (the number of
        requests still admitted)."""
        a = C.c_int64()
        self._check(self.lib.aurora_hip_swarm_launch(self.h, max_steps, C.byref(a)))
        return a.value

    def swarm_collect(self, wait=True):
        """Completes (wait) or checks the launch in flight; True while it runs."""
        busy = C.c_int()
        self._check(self.lib.aurora_hip_swarm_collect(self.h, 1 if wait else 0, C.byref(busy)))
        return bool(busy.value)

    def swarm_poll(self, handle):
        done = C.c_int()
        r = SwarmResult()
        self._check(self.lib.aurora_hip_swarm_poll(self.h, handle, C.byref(done), C.byref(r)))
        return bool(done.value), r

    def swarm_branching(self, request: SwarmRequest, population_limit):
        """One request with the branching / resampled population model, run
        to completion: (SwarmResult, SwarmGrowth)."""
        r, g = SwarmResult(), SwarmGrowth()
        self._check(self.lib.aurora_hip_swarm_branching(self.h, C.byref(request),
                                                        int(population_limit), C.byref(r),
                                                        C.byref(g)))
        return r, g

    def swarm_footprint(self, request: SwarmRequest):
        return int(self.lib.aurora_hip_swarm_footprint(C.byref(request)))

    # ---- batched 1D ensembles
    def ensemble_create(self, instances, capacity):
        cap = np.ascontiguousarray(capacity, dtype=np.int64)
        h = C.c_int64()
        self._check(self.lib.aurora_hip_ensemble_create(self.h, instances, _p(cap, _i64p),
                                                        C.byref(h)))
        return Ensemble(self, h.value, instances, len(cap))


class Ensemble:
    """Handle of a batched 1D ensemble owned by a Context."""

    def __init__(self, ctx, handle, instances, nspecies):
        self.ctx, self.h, self.instances, self.nspecies = ctx, handle, instances, nspecies

    def _call(self, name, *args):
        self.ctx._check(getattr(self.ctx.lib, name)(self.ctx.h, self.h, *args))

    def load(self, instance, mcc_seed):
        self._call("aurora_hip_ensemble_load", instance, mcc_seed)

    def graph(self, steps_per_graph):
        self._call("aurora_hip_ensemble_graph", steps_per_graph)

    def step(self, dt, n=1):
        self._call("aurora_hip_ensemble_step", dt, n)

    def count(self, instance, sid):
        s, l = C.c_int64(), C.c_int64()
        self._call("aurora_hip_ensemble_count", instance, sid, C.byref(s), C.byref(l))
        return s.value, l.value

    def download(self, instance, sid):
        slots, _ = self.count(instance, sid)
        m = max(slots, 1)
        pos, vel = np.zeros(m), np.zeros(3 * m)
        alive, uid = np.zeros(m, dtype=np.uint8), np.zeros(m, dtype=np.uint32)
        n = C.c_int64()
        self._call("aurora_hip_ensemble_download", instance, sid, m, _p(pos, _dp), _p(vel, _dp),
                   _p(alive, _u8p), _p(uid, _u32p), C.byref(n))
        k = n.value
        return pos[:k].reshape(k, 1), vel[:3 * k].reshape(k, 3), alive[:k], uid[:k]

    def stats(self, instance, nchan):
        counts = np.zeros(nchan + 3, dtype=np.uint64)
        ech = np.zeros(max(nchan, 1))
        ab = np.zeros(2 * self.nspecies, dtype=np.uint64)
        abe = np.zeros(2 * self.nspecies)
        self._call("aurora_hip_ensemble_stats", instance, _p(counts, _u64p), _p(ech, _dp),
                   _p(ab, _u64p), _p(abe, _dp))
        return counts, ech[:nchan], ab.reshape(self.nspecies, 2), abe.reshape(self.nspecies, 2)

    def fields(self, instance):
        n = self.ctx.nnodes
        rho, phi, E = np.zeros(n), np.zeros(n), np.zeros(n)
        self._call("aurora_hip_ensemble_fields", instance, _p(rho, _dp), _p(phi, _dp), _p(E, _dp))
        return rho, phi, E

    def time(self):
        t, s = C.c_double(), C.c_int64()
        self._call("aurora_hip_ensemble_time", C.byref(t), C.byref(s))
        return t.value, s.value

    def destroy(self):
        if self.h:
            self._call("aurora_hip_ensemble_destroy")
            self.h = 0
