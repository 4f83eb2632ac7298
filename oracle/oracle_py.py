"""ctypes front end for the CPU parity checker (test infrastructure only).

Loads ``oracle/build/liboracle.so`` (the C restatement) and, when it was
built, ``oracle/_ref/libaurora_ref.so`` (the reference's own sources compiled
unmodified, plus ``oracle/ref/ref_harness.cpp``).  Only tests, the smoke
check and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libaurora_ref.so")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


def build(with_ref: bool | None = None) -> None:
    """Compile the restatement (and the reference harness when the reference
    tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/src")
    if with_ref:
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


ELASTIC, EXCITATION, IONIZATION, ATTACHMENT, CHARGE_EXCHANGE = range(5)
ISOTROPIC, BACKWARD, HENYEY_GREENSTEIN = range(3)
FRAME_PROJECTILE, FRAME_CENTER_OF_MASS = range(2)


class _MccModel(C.Structure):
    _fields_ = [
        ("nchan", C.c_int),
        ("kind", C.c_int * 8),
        ("threshold", C.c_double * 8),
        ("rows", C.c_int64 * 8),
        ("energy", _dp * 8),
        ("sigma", _dp * 8),
        ("neutral_density", C.c_double),
        ("max_frequency", C.c_double),
        ("max_candidates", C.c_int64),
        ("particle_mass", C.c_double),
        ("neutral_mass", C.c_double),
        ("neutral_vstd", C.c_double),
        ("angular", C.c_int * 8),
        ("frame", C.c_int * 8),
        ("mc_rows", C.c_int64 * 8),
        ("mc_energy", _dp * 8),
        ("mc_value", _dp * 8),
    ]


class _MccStats(C.Structure):
    _fields_ = [
        ("candidates", C.c_uint64),
        ("nulls", C.c_uint64),
        ("chan", C.c_uint64 * 8),
        ("energy_change", C.c_double * 8),
        ("nsec", C.c_int),
        ("sec_v", (C.c_double * 3) * 8),
        ("ion_v", (C.c_double * 3) * 8),
        ("status", C.c_int),
        ("margin", C.c_double),
    ]


class _Stream(C.Structure):
    _fields_ = [
        ("key", C.c_uint32 * 2),
        ("ctr", C.c_uint32 * 4),
        ("buf", C.c_uint32 * 4),
        ("used", C.c_int),
    ]


class _SwarmCfg(C.Structure):
    _fields_ = [
        ("neutral_density", C.c_double),
        ("max_frequency", C.c_double),
        ("timestep", C.c_double),
        ("reduced_field_td", C.c_double),
        ("steps", C.c_int64),
        ("burn_in", C.c_int64),
        ("particles", C.c_int64),
        ("blocks", C.c_int64),
        ("initial_mean_energy_ev", C.c_double),
        ("max_energy_ev", C.c_double),
    ]


class _SwarmGrowth(C.Structure):
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


class _SwarmResult(C.Structure):
    _fields_ = [
        ("drift", C.c_double),
        ("drift_se", C.c_double),
        ("mean_energy_ev", C.c_double),
        ("mean_energy_se", C.c_double),
        ("ion_rate", C.c_double),
        ("ion_rate_se", C.c_double),
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


class MccModel:
    """Owns the table arrays behind an or_mcc_model."""

    def __init__(self, channels, neutral_density, max_frequency,
                 particle_mass, max_candidates=64, neutral_mass=0.0,
                 neutral_vstd=0.0, options=None):
        # channels: list of (kind, threshold_J, energies_J, sigmas_m2)
        # options: {channel: (angular, frame, mc_energy_J or None, mc_value or None)}
        self.channels = [(int(k), float(t), f64(e), f64(s)) for k, t, e, s in channels]
        m = _MccModel()
        m.nchan = len(self.channels)
        for c, (k, t, e, s) in enumerate(self.channels):
            m.kind[c] = k
            m.threshold[c] = t
            m.rows[c] = len(e)
            m.energy[c] = _ptr(e, _dp)
            m.sigma[c] = _ptr(s, _dp)
        self._mc = []
        for c, (ang, frame, me, mv) in (options or {}).items():
            m.angular[c] = int(ang)
            m.frame[c] = int(frame)
            if me is not None:
                me, mv = f64(me), f64(mv)
                self._mc.append((me, mv))
                m.mc_rows[c] = len(me)
                m.mc_energy[c] = _ptr(me, _dp)
                m.mc_value[c] = _ptr(mv, _dp)
        m.neutral_density = neutral_density
        m.max_frequency = max_frequency
        m.max_candidates = max_candidates
        m.particle_mass = particle_mass
        m.neutral_mass = neutral_mass
        m.neutral_vstd = neutral_vstd
        self.struct = m
        self.neutral_density = neutral_density
        self.max_frequency = max_frequency
        self.particle_mass = particle_mass
        self.max_candidates = max_candidates
        self.neutral_mass = neutral_mass
        self.neutral_vstd = neutral_vstd


def model_from_gas(gas: dict, particle_mass: float = None, neutral_vstd: float = 0.0,
                   options=None) -> MccModel:
    """The oracle model for a gas description in the device loaders' form
    (channels, neutral_density, max_frequency, neutral_mass[, particle_mass,
    max_candidates])."""
    m = gas.get("particle_mass", 9.1093837139e-31) if particle_mass is None else particle_mass
    return MccModel(gas["channels"], gas["neutral_density"], gas["max_frequency"], m,
                    gas.get("max_candidates", 64), gas["neutral_mass"], neutral_vstd, options)


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build(with_ref=False)
        L = C.CDLL(path)
        self.lib = L
        L.or_spacing.restype = C.c_double
        L.or_deposit_cic.argtypes = [C.c_int, _i64p, _dp, C.c_int, C.c_int64, _dp, _u8p,
                                     C.c_double, C.c_double, _dp]
        L.or_deposit_fixed.argtypes = [C.c_int, _i64p, C.c_int, C.c_int64, _i32p, _dp, _i64p]
        L.or_fixed_to_rho.argtypes = [C.c_int, _i64p, _dp, C.c_int, C.c_int,
                                      C.POINTER(_i64p), _dp, _dp]
        L.or_gather.argtypes = [C.c_int, _i64p, _dp, C.c_int, C.c_int64, _dp, _dp, _dp]
        L.or_boris_advance.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, _dp]
        L.or_table_eval.argtypes = [C.c_int64, _dp, _dp, C.c_double, _dp]
        L.or_poisson_periodic_1d.argtypes = [C.c_int64, C.c_double, C.c_double, _dp, _dp, _dp]
        L.or_poisson_dirichlet_1d.argtypes = [C.c_int64, C.c_double, C.c_double, _dp,
                                              C.c_double, C.c_double, _dp, _dp]
        L.or_poisson_periodic_2d.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_double,
                                             C.c_double, _dp, _dp, _dp, _dp]
        L.or_poisson_periodic_3d.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_double,
                                             C.c_double, C.c_double, C.c_double, _dp, _dp,
                                             _dp, _dp, _dp]
        push_args = [C.c_int, _i64p, _dp, C.c_int, C.c_int64, _dp, _dp, _u8p, _dp,
                     C.c_double, C.c_double, _dp, C.c_int, C.c_int64, _dp, _dp,
                     C.POINTER(C.c_int), _i64p]
        L.or_push_species.argtypes = push_args
        half_args = [C.c_int, _i64p, _dp, C.c_int, C.c_int64, _dp, _dp, _u8p, _dp,
                     C.c_double, C.c_double, _dp, C.c_int, C.c_int64, _dp, _dp, _dp]
        L.or_sync_species.argtypes = half_args
        L.or_init_half_species.argtypes = half_args
        _pp = C.POINTER(_dp)
        L.or_step3d.argtypes = [_i64p, _dp, C.c_int, _i64p, _pp, _pp, C.POINTER(_u8p), _pp,
                                _dp, _dp, _dp, C.c_double, C.c_double, _dp, C.c_int,
                                C.c_int64, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.or_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.or_u01.argtypes = [C.c_uint32]
        L.or_u01.restype = C.c_double
        L.or_stream_init.argtypes = [C.POINTER(_Stream), C.c_uint64, C.c_uint32, C.c_uint32,
                                     C.c_uint64]
        L.or_stream_u01.argtypes = [C.POINTER(_Stream)]
        L.or_stream_u01.restype = C.c_double
        L.or_xs_eval.argtypes = [C.c_int64, _dp, _dp, C.c_double]
        L.or_xs_eval.restype = C.c_double
        L.or_mcc_collide_3v.argtypes = [C.POINTER(_MccModel), _dp, C.c_double,
                                        C.POINTER(_Stream), C.POINTER(_MccStats)]
        L.or_mcc_collide_1v.argtypes = [C.POINTER(_MccModel), _dp, C.c_double,
                                        C.POINTER(_Stream), C.POINTER(_MccStats)]
        L.or_swarm_run.argtypes = [C.POINTER(_MccModel), C.POINTER(_SwarmCfg), _dp,
                                   C.c_uint64, C.c_uint32, C.POINTER(_SwarmResult)]
        L.or_block_mean_error.argtypes = [C.c_int64, _dp, C.c_int64, _dp, _dp]
        L.or_block_slope_error.argtypes = [C.c_int64, _dp, C.c_double, C.c_int64, _dp, _dp]
        L.or_systematic_source.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_double]
        L.or_systematic_source.restype = C.c_int64
        L.or_swarm_branching_run.argtypes = [C.POINTER(_MccModel), C.POINTER(_SwarmCfg),
                                             C.c_int64, _dp, C.c_uint64, C.c_uint32,
                                             C.POINTER(_SwarmResult), C.POINTER(_SwarmGrowth)]

    # ---- geometry
    @staticmethod
    def _mesh(n, L):
        n = i64(n)
        L = f64(L)
        return n, L, len(n)

    def deposit_cic(self, n, L, bc, pos, charge, weight, alive=None):
        n, L, dim = self._mesh(n, L)
        pos = f64(pos).reshape(-1)
        np_ = pos.size // dim
        rho = np.zeros(int(np.prod(n)))
        al = None if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
        self.lib.or_deposit_cic(dim, _ptr(n, _i64p), _ptr(L, _dp), bc, np_, _ptr(pos, _dp),
                                None if al is None else _ptr(al, _u8p), charge, weight,
                                _ptr(rho, _dp))
        return rho

    def deposit_fixed(self, n, bc, cell, off):
        n = i64(n)
        dim = len(n)
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        off = f64(off).reshape(-1)
        acc = np.zeros(int(np.prod(n)), dtype=np.int64)
        self.lib.or_deposit_fixed(dim, _ptr(n, _i64p), bc, cell.size, _ptr(cell, _i32p),
                                  _ptr(off, _dp), _ptr(acc, _i64p))
        return acc

    def fixed_to_rho(self, n, L, bc, accs, charge_weights):
        n, L, dim = self._mesh(n, L)
        accs = [i64(a) for a in accs]
        arr = (_i64p * len(accs))(*[_ptr(a, _i64p) for a in accs])
        cw = f64(charge_weights)
        rho = np.zeros(int(np.prod(n)))
        self.lib.or_fixed_to_rho(dim, _ptr(n, _i64p), _ptr(L, _dp), bc, len(accs), arr,
                                 _ptr(cw, _dp), _ptr(rho, _dp))
        return rho

    def gather(self, n, L, bc, pos, E):
        n, L, dim = self._mesh(n, L)
        pos = f64(pos).reshape(-1)
        E = f64(E).reshape(-1)
        np_ = pos.size // dim
        out = np.zeros(np_ * dim)
        self.lib.or_gather(dim, _ptr(n, _i64p), _ptr(L, _dp), bc, np_, _ptr(pos, _dp),
                           _ptr(E, _dp), _ptr(out, _dp))
        return out.reshape(np_, dim) if dim > 1 else out

    def boris(self, v, E, B, qm, h):
        v, E, B = f64(v), f64(E), f64(B)
        out = np.zeros(3)
        self.lib.or_boris_advance(_ptr(v, _dp), _ptr(E, _dp), _ptr(B, _dp), qm, h,
                                  _ptr(out, _dp))
        return out

    def table_eval(self, coords, values, q):
        coords, values = f64(coords), f64(values).reshape(-1)
        out = np.zeros(3)
        rc = self.lib.or_table_eval(coords.size, _ptr(coords, _dp), _ptr(values, _dp), q,
                                    _ptr(out, _dp))
        return None if rc else out

    # ---- Poisson
    def poisson(self, n, L, bc, eps, rho, phi_left=0.0, phi_right=0.0):
        n, L, dim = self._mesh(n, L)
        rho = f64(rho).reshape(-1)
        nn = rho.size
        phi = np.zeros(nn)
        E = np.zeros(dim * nn)
        if dim == 1 and bc == 0:
            self.lib.or_poisson_periodic_1d(n[0], L[0], eps, _ptr(rho, _dp), _ptr(phi, _dp),
                                            _ptr(E, _dp))
        elif dim == 1:
            self.lib.or_poisson_dirichlet_1d(n[0], L[0], eps, _ptr(rho, _dp), phi_left,
                                             phi_right, _ptr(phi, _dp), _ptr(E, _dp))
        elif dim == 2:
            Ex, Ey = E[:nn], E[nn:]
            self.lib.or_poisson_periodic_2d(n[0], n[1], L[0], L[1], eps, _ptr(rho, _dp),
                                            _ptr(phi, _dp), _ptr(Ex, _dp), _ptr(Ey, _dp))
        else:
            Ex, Ey, Ez = E[:nn], E[nn:2 * nn], E[2 * nn:]
            self.lib.or_poisson_periodic_3d(n[0], n[1], n[2], L[0], L[1], L[2], eps,
                                            _ptr(rho, _dp), _ptr(phi, _dp), _ptr(Ex, _dp),
                                            _ptr(Ey, _dp), _ptr(Ez, _dp))
        return phi, E.reshape(dim, nn)

    # ---- push
    def _bargs(self, B, btable):
        Buni = f64(B if B is not None else [0.0, 0.0, 0.0])
        if btable is None:
            return Buni, 0, 0, None, None
        axis, coords, values = btable
        coords, values = f64(coords), f64(values).reshape(-1)
        return Buni, axis, coords.size, coords, values

    def push(self, n, L, bc, pos, vh, alive, E, qm, dt, B=None, btable=None, pb=None):
        n, L, dim = self._mesh(n, L)
        pos = f64(pos).copy().reshape(-1)
        vh = f64(vh).copy().reshape(-1)
        alive = np.ascontiguousarray(alive, dtype=np.uint8).copy()
        E = f64(E).reshape(-1)
        Buni, axis, rows, bc_, bv = self._bargs(B, btable)
        if pb is None:
            pb = [3] * 6 if bc == 0 else [1] * 6
        pbv = (C.c_int * 6)(*pb)
        absorbed = np.zeros(6, dtype=np.int64)
        rc = self.lib.or_push_species(dim, _ptr(n, _i64p), _ptr(L, _dp), bc, alive.size,
                                      _ptr(pos, _dp), _ptr(vh, _dp), _ptr(alive, _u8p),
                                      _ptr(E, _dp), qm, dt, _ptr(Buni, _dp), axis, rows,
                                      None if bc_ is None else _ptr(bc_, _dp),
                                      None if bv is None else _ptr(bv, _dp), pbv,
                                      _ptr(absorbed, _i64p))
        if rc:
            raise ValueError("tabulated vector field query lies outside the table")
        return pos.reshape(-1, dim), vh.reshape(-1, 3), alive, absorbed

    def _half(self, fn, n, L, bc, pos, vin, alive, E, qm, dt, B, btable):
        n, L, dim = self._mesh(n, L)
        pos = f64(pos).reshape(-1)
        vin = f64(vin).reshape(-1)
        alive = np.ascontiguousarray(alive, dtype=np.uint8)
        E = f64(E).reshape(-1)
        out = vin.copy()
        Buni, axis, rows, bc_, bv = self._bargs(B, btable)
        rc = fn(dim, _ptr(n, _i64p), _ptr(L, _dp), bc, alive.size, _ptr(pos, _dp),
                _ptr(vin, _dp), _ptr(alive, _u8p), _ptr(E, _dp), qm, dt, _ptr(Buni, _dp),
                axis, rows, None if bc_ is None else _ptr(bc_, _dp),
                None if bv is None else _ptr(bv, _dp), _ptr(out, _dp))
        if rc:
            raise ValueError("tabulated vector field query lies outside the table")
        return out.reshape(-1, 3)

    def sync(self, n, L, bc, pos, vh, alive, E, qm, dt, B=None, btable=None):
        return self._half(self.lib.or_sync_species, n, L, bc, pos, vh, alive, E, qm, dt, B,
                          btable)

    def init_half(self, n, L, bc, pos, v, alive, E, qm, dt, B=None, btable=None):
        return self._half(self.lib.or_init_half_species, n, L, bc, pos, v, alive, E, qm, dt,
                          B, btable)

    # ---- whole 3D steps (the CPU timing baseline)
    def plasma3d_init(self, n, L, eps, dt, species, btable=None, seed=1):
        """A periodic 3D plasma state for step3d: species = [(charge, mass,
        weight, count, thermal_velocity)], positions uniform and velocities
        Maxwellian from numpy (seeded), then Simulation3D::initialize
        (Simulation3D.cpp:288-299): deposit, solve at t = 0 and v_half from
        v with a backward half step."""
        n, L, _ = self._mesh(n, L)
        rng = np.random.default_rng(seed)
        st = dict(n=n, L=L, eps=float(eps), dt=float(dt), btable=btable, species=species,
                  pos=[], vh=[], v=[], alive=[])
        rho = np.zeros(int(np.prod(n)))
        for q, m, w, count, vth in species:
            pos = rng.uniform(0.0, 1.0, size=(count, 3)) * L
            st["pos"].append(np.ascontiguousarray(pos.reshape(-1)))
            st["v"].append(np.ascontiguousarray(rng.normal(0.0, vth, size=3 * count)))
            st["alive"].append(np.ones(count, dtype=np.uint8))
            rho += self.deposit_cic(n, L, 0, pos, q, w)
        _, E = self.poisson(n, L, 0, eps, rho)
        st["E"] = np.ascontiguousarray(E.reshape(-1))
        for k, (q, m, w, count, vth) in enumerate(species):
            vh = self.init_half(n, L, 0, st["pos"][k].reshape(-1, 3), st["v"][k].reshape(-1, 3),
                                st["alive"][k], st["E"].reshape(3, -1), q / m, dt,
                                btable=btable)
            st["vh"].append(np.ascontiguousarray(vh.reshape(-1)))
        st["rho"], st["phi"] = rho, np.zeros_like(rho)
        return st

    def step3d(self, st, threads=1):
        """One Simulation3D::step of the state in place (or_step3d: serial
        push, deposit and FFT solve, the synchronisation loop on `threads`
        OpenMP threads).  Returns the live particle count."""
        sp = st["species"]
        S = len(sp)
        pp = (_dp * S)(*[_ptr(x, _dp) for x in st["pos"]])
        vv = (_dp * S)(*[_ptr(x, _dp) for x in st["vh"]])
        vc = (_dp * S)(*[_ptr(x, _dp) for x in st["v"]])
        al = (_u8p * S)(*[_ptr(x, _u8p) for x in st["alive"]])
        np_ = i64([len(a) for a in st["alive"]])
        qm = f64([q / m for q, m, _, _, _ in sp])
        ch = f64([q for q, _, _, _, _ in sp])
        wt = f64([w for _, _, w, _, _ in sp])
        Buni, axis, rows, bc_, bv = self._bargs(None, st["btable"])
        rc = self.lib.or_step3d(_ptr(st["n"], _i64p), _ptr(st["L"], _dp), S, _ptr(np_, _i64p),
                                pp, vv, al, vc, _ptr(qm, _dp), _ptr(ch, _dp), _ptr(wt, _dp),
                                st["eps"], st["dt"], _ptr(Buni, _dp), axis, rows,
                                None if bc_ is None else _ptr(bc_, _dp),
                                None if bv is None else _ptr(bv, _dp), int(threads),
                                _ptr(st["rho"], _dp), _ptr(st["phi"], _dp), _ptr(st["E"], _dp))
        if rc:
            raise ValueError("tabulated vector field query lies outside the table")
        return int(sum(int(a.sum()) for a in st["alive"]))

    # ---- RNG / MCC / swarm
    def philox(self, ctr, key):
        ctr = np.ascontiguousarray(ctr, dtype=np.uint32)
        key = np.ascontiguousarray(key, dtype=np.uint32)
        out = np.zeros(4, dtype=np.uint32)
        self.lib.or_philox4x32_10(_ptr(ctr, _u32p), _ptr(key, _u32p), _ptr(out, _u32p))
        return out

    def u01(self, u):
        return self.lib.or_u01(int(u))

    def stream(self, seed, stream, uid, step):
        s = _Stream()
        self.lib.or_stream_init(C.byref(s), seed, stream, uid, step)
        return s

    def stream_u01(self, s):
        return self.lib.or_stream_u01(C.byref(s))

    def xs_eval(self, e, s, x):
        e, s = f64(e), f64(s)
        return self.lib.or_xs_eval(e.size, _ptr(e, _dp), _ptr(s, _dp), x)

    def collide(self, model: MccModel, v, dt, seed, stream, uid, step):
        v = f64(v).copy()
        st = _MccStats()
        s = self.stream(seed, stream, uid, step)
        self.lib.or_mcc_collide_3v(C.byref(model.struct), _ptr(v, _dp), dt, C.byref(s),
                                   C.byref(st))
        return v, st

    def collide_1v(self, model: MccModel, v, dt, seed, stream, uid, step):
        """The scalar collide (Collision.cpp:831-905) on v (a float) with the
        stream keyed (seed, stream, uid, step)."""
        x = np.array([float(v)])
        st = _MccStats()
        s = self.stream(seed, stream, uid, step)
        self.lib.or_mcc_collide_1v(C.byref(model.struct), _ptr(x, _dp), dt, C.byref(s),
                                   C.byref(st))
        return float(x[0]), st

    def swarm(self, model: MccModel, cfg: dict, vel0, seed, instance=0):
        c = _SwarmCfg(**cfg)
        vel0 = f64(vel0).reshape(-1)
        r = _SwarmResult()
        self.lib.or_swarm_run(C.byref(model.struct), C.byref(c), _ptr(vel0, _dp), seed,
                              instance, C.byref(r))
        return r

    def swarm_branching(self, model: MccModel, cfg: dict, vel0, seed, population_limit,
                        instance=0):
        c = _SwarmCfg(**cfg)
        vel0 = f64(vel0).reshape(-1)
        r, g = _SwarmResult(), _SwarmGrowth()
        self.lib.or_swarm_branching_run(C.byref(model.struct), C.byref(c), population_limit,
                                        _ptr(vel0, _dp), seed, instance, C.byref(r),
                                        C.byref(g))
        return r, g

    def block_slope_error(self, samples, dt, blocks):
        samples = f64(samples)
        m, se = C.c_double(), C.c_double()
        rc = self.lib.or_block_slope_error(samples.size, _ptr(samples, _dp), dt, blocks,
                                           C.byref(m), C.byref(se))
        if rc:
            raise ValueError("growth statistics need two samples per equal block")
        return m.value, se.value

    def systematic_source(self, j, M, N, u):
        return self.lib.or_systematic_source(j, M, N, u)

    def block_mean_error(self, samples, blocks):
        samples = f64(samples)
        m = C.c_double()
        se = C.c_double()
        self.lib.or_block_mean_error(samples.size, _ptr(samples, _dp), blocks, C.byref(m),
                                     C.byref(se))
        return m.value, se.value


class Reference:
    """The reference's own code (compiled in this container only)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        self.lib = L
        L.ref_deposit.argtypes = [C.c_int, _i64p, _dp, C.c_int, C.c_int64, _dp, _u8p,
                                  C.c_double, C.c_double, _dp]
        L.ref_gather.argtypes = [C.c_int, _i64p, _dp, C.c_int, C.c_int64, _dp, _dp, _dp]
        L.ref_solve.argtypes = [C.c_int, _i64p, _dp, C.c_int, C.c_double, _dp, C.c_double,
                                C.c_double, _dp, _dp]
        L.ref_boris.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, _dp]
        L.ref_table_eval.argtypes = [C.c_int64, _dp, _dp, C.c_double, _dp]
        L.ref_xs_eval.argtypes = [C.c_int64, _dp, _dp, C.c_int64, _dp, _dp]
        L.ref_mcc_collide.argtypes = [C.c_int, C.POINTER(C.c_int), _dp, _i64p,
                                      C.POINTER(_dp), C.POINTER(_dp), C.c_double, C.c_double,
                                      C.c_int64, C.c_double, C.c_double, C.c_double,
                                      C.c_int64, _dp, C.c_double, C.c_uint64, _dp, _u64p,
                                      _dp, C.c_char_p, C.c_int]
        L.ref_swarm.argtypes = [C.c_int, C.POINTER(C.c_int), _dp, _i64p, C.POINTER(_dp),
                                C.POINTER(_dp), C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                C.c_double, C.c_double, C.c_double, C.c_uint64, _dp,
                                C.c_char_p, C.c_int]
        L.ref_swarm_file.argtypes = [C.c_char_p, C.c_int, _dp, C.c_char_p, C.c_char_p, C.c_int]
        L.ref_aps_signature.argtypes = [C.c_char_p, _u64p, C.c_char_p, C.c_int]
        L.ref_sim3d_time.argtypes = [_i64p, _dp, C.c_double, C.c_double, C.c_int, C.c_int,
                                     _dp, C.c_int64, _dp, _dp, C.c_int, _dp, _i64p,
                                     C.c_char_p, C.c_int]
        L.ref_run3d.argtypes = [C.c_char_p, _i64p, _dp, C.c_double, C.c_int64, C.c_int64,
                                C.c_uint64, C.c_int, C.c_int64, _dp, _dp, C.c_int,
                                C.POINTER(C.c_char_p), _dp, C.c_char_p, C.c_int]
        L.ref_run2d.argtypes = [C.c_char_p, _i64p, _dp, C.c_double, C.c_double, C.c_int64,
                                C.c_int64, C.c_uint64, _dp, C.c_int, C.POINTER(C.c_char_p), _dp,
                                C.c_char_p, C.c_int]
        L.ref_run1d.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, C.c_double, C.c_double,
                                C.c_int64, C.c_int64, C.c_int, _dp, C.c_int, C.c_uint64, C.c_int,
                                C.POINTER(C.c_char_p), _dp, C.c_char_p, C.c_int]

    def deposit(self, n, L, bc, pos, charge, weight, alive=None):
        n, L = i64(n), f64(L)
        dim = len(n)
        pos = f64(pos).reshape(-1)
        rho = np.zeros(int(np.prod(n)))
        al = None if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
        rc = self.lib.ref_deposit(dim, _ptr(n, _i64p), _ptr(L, _dp), bc, pos.size // dim,
                                  _ptr(pos, _dp), None if al is None else _ptr(al, _u8p),
                                  charge, weight, _ptr(rho, _dp))
        assert rc == 0
        return rho

    def gather(self, n, L, bc, pos, E):
        n, L = i64(n), f64(L)
        dim = len(n)
        pos = f64(pos).reshape(-1)
        E = f64(E).reshape(-1)
        out = np.zeros(pos.size)
        rc = self.lib.ref_gather(dim, _ptr(n, _i64p), _ptr(L, _dp), bc, pos.size // dim,
                                 _ptr(pos, _dp), _ptr(E, _dp), _ptr(out, _dp))
        assert rc == 0
        return out.reshape(-1, dim) if dim > 1 else out

    def solve(self, n, L, bc, eps, rho, phi_left=0.0, phi_right=0.0):
        n, L = i64(n), f64(L)
        dim = len(n)
        rho = f64(rho).reshape(-1)
        phi = np.zeros(rho.size)
        E = np.zeros(dim * rho.size)
        rc = self.lib.ref_solve(dim, _ptr(n, _i64p), _ptr(L, _dp), bc, eps, _ptr(rho, _dp),
                                phi_left, phi_right, _ptr(phi, _dp), _ptr(E, _dp))
        assert rc == 0
        return phi, E.reshape(dim, rho.size)

    def boris(self, v, E, B, qm, h):
        v, E, B = f64(v), f64(E), f64(B)
        out = np.zeros(3)
        self.lib.ref_boris(_ptr(v, _dp), _ptr(E, _dp), _ptr(B, _dp), qm, h, _ptr(out, _dp))
        return out

    def table_eval(self, coords, values, q):
        coords, values = f64(coords), f64(values).reshape(-1)
        out = np.zeros(3)
        rc = self.lib.ref_table_eval(coords.size, _ptr(coords, _dp), _ptr(values, _dp), q,
                                     _ptr(out, _dp))
        return None if rc else out

    def xs_eval(self, e, s, q):
        e, s, q = f64(e), f64(s), f64(q)
        out = np.zeros(q.size)
        rc = self.lib.ref_xs_eval(e.size, _ptr(e, _dp), _ptr(s, _dp), q.size, _ptr(q, _dp),
                                  _ptr(out, _dp))
        assert rc == 0
        return out

    def _chan_arrays(self, model: MccModel):
        nch = len(model.channels)
        kinds = (C.c_int * nch)(*[c[0] for c in model.channels])
        thr = f64([c[1] for c in model.channels])
        rows = i64([len(c[2]) for c in model.channels])
        ep = (_dp * nch)(*[_ptr(c[2], _dp) for c in model.channels])
        sp = (_dp * nch)(*[_ptr(c[3], _dp) for c in model.channels])
        return nch, kinds, thr, rows, ep, sp

    def mcc_collide(self, model: MccModel, v, dt, seed, neutral_temperature=0.0):
        nch, kinds, thr, rows, ep, sp = self._chan_arrays(model)
        v = f64(v).reshape(-1)
        out = np.zeros_like(v)
        counts = np.zeros(nch + 3, dtype=np.uint64)
        ech = np.zeros(nch)
        err = C.create_string_buffer(512)
        rc = self.lib.ref_mcc_collide(nch, kinds, _ptr(thr, _dp), _ptr(rows, _i64p), ep, sp,
                                      model.neutral_density, model.max_frequency,
                                      model.max_candidates, model.particle_mass,
                                      model.neutral_mass, neutral_temperature, v.size // 3,
                                      _ptr(v, _dp), dt, seed, _ptr(out, _dp),
                                      _ptr(counts, _u64p), _ptr(ech, _dp), err, 512)
        if rc:
            raise RuntimeError(err.value.decode())
        return out.reshape(-1, 3), counts, ech

    def swarm(self, model: MccModel, neutral_mass, cfg: dict, seed):
        nch, kinds, thr, rows, ep, sp = self._chan_arrays(model)
        out = np.zeros(7 + nch)
        err = C.create_string_buffer(512)
        rc = self.lib.ref_swarm(nch, kinds, _ptr(thr, _dp), _ptr(rows, _i64p), ep, sp,
                                neutral_mass, cfg["neutral_density"], cfg["max_frequency"],
                                cfg["timestep"], cfg["steps"], cfg["burn_in"],
                                cfg["particles"], cfg["blocks"],
                                cfg["initial_mean_energy_ev"], cfg["max_energy_ev"],
                                cfg["reduced_field_td"], seed, _ptr(out, _dp), err, 512)
        if rc:
            raise RuntimeError(err.value.decode())
        return out

    SWARM_FILE_COLS = ("reduced_field_td", "drift", "drift_se", "reduced_mobility",
                       "mean_energy_ev", "mean_energy_se", "ionization_rate",
                       "ionization_rate_se", "max_energy_ev", "candidates", "nulls",
                       "neutral_vstd", "collisions0", "collisions1", "collisions2",
                       "collisions3", "longitudinal_diffusion", "transverse_diffusion",
                       "final_weight", "final_particles", "growth_rate", "growth_rate_se",
                       "attachment_rate", "attachment_rate_se", "net_creation_rate",
                       "net_creation_rate_se", "townsend_available", "townsend_growth",
                       "townsend_growth_se", "townsend_rate_balance",
                       "townsend_rate_balance_se")

    def swarm_file(self, path, max_fields=64, csv_path=""):
        """A .swarm file through the reference's own loader and driver; one
        row of SWARM_FILE_COLS per reduced field (csv_path: also write the
        reference's result CSV there)."""
        out = np.zeros((max_fields, len(self.SWARM_FILE_COLS)))
        err = C.create_string_buffer(512)
        n = self.lib.ref_swarm_file(os.fsencode(path), max_fields, _ptr(out, _dp),
                                    os.fsencode(csv_path), err, 512)
        if n < 0:
            raise RuntimeError(err.value.decode())
        return out[:n]

    def aps_signature(self, path):
        """(signature, None) from the reference's .aps reader, or (None, its
        exception text)."""
        sig = C.c_uint64()
        err = C.create_string_buffer(1024)
        if self.lib.ref_aps_signature(os.fsencode(path), C.byref(sig), err, 1024):
            return None, err.value.decode()
        return sig.value, None

    def sim3d_time(self, n, L, dt, eps_rel, si, species, btable, steps):
        n, L = i64(n), f64(L)
        sp = f64(species).reshape(-1)
        if btable is None:
            rows, bcoords, bvalues = 0, None, None
        else:
            bcoords, bvalues = f64(btable[0]), f64(btable[1]).reshape(-1)
            rows = bcoords.size
        sec = C.c_double()
        live = C.c_int64()
        err = C.create_string_buffer(512)
        rc = self.lib.ref_sim3d_time(_ptr(n, _i64p), _ptr(L, _dp), dt, eps_rel, int(si),
                                     sp.size // 5, _ptr(sp, _dp), rows,
                                     None if bcoords is None else _ptr(bcoords, _dp),
                                     None if bvalues is None else _ptr(bvalues, _dp), steps,
                                     C.byref(sec), C.byref(live), err, 512)
        if rc:
            raise RuntimeError(err.value.decode())
        return sec.value, live.value


    RUN1D_SPECIES_KEYS = ("charge", "mass", "weight", "particles", "drift_velocity",
                          "thermal_velocity", "init_x_min", "init_x_max")

    def run1d(self, case: dict, output_dir: str, aps_path: str = ""):
        """The reference's 1D Simulation::run() on `case` (the keys of a
        tests/golden/run1d case file), writing its outputs into output_dir."""
        sp = case["species"]
        names = (C.c_char_p * len(sp))(*[s["name"].encode() for s in sp])
        rows = f64([[s[k] for k in self.RUN1D_SPECIES_KEYS] for s in sp]).reshape(-1)
        el = f64(case["electrodes"])
        err = C.create_string_buffer(512)
        rc = self.lib.ref_run1d(output_dir.encode(), aps_path.encode(), case["nx"],
                                case["length"], case["dt"], case["steps"],
                                case["output_interval"], int(case["dirichlet"]), _ptr(el, _dp),
                                case["velocity_dimensions"], case["seed"], len(sp), names,
                                _ptr(rows, _dp), err, 512)
        if rc:
            raise RuntimeError(err.value.decode())


    def run1d_case(self, case_path: str, output_dir: str, seed: int = -1):
        """The reference's 1D Simulation::run() on a case file (named MCC
        models included), writing its outputs and the final live particles
        (final_<species>.f64) into output_dir; seed >= 0 replaces the case's."""
        err = C.create_string_buffer(512)
        self.lib.ref_run1d_case.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, C.c_char_p,
                                            C.c_int]
        rc = self.lib.ref_run1d_case(case_path.encode(), output_dir.encode(), int(seed), err,
                                     512)
        if rc:
            raise RuntimeError(err.value.decode())

    def run3d(self, case: dict, output_dir: str):
        """The reference's Simulation3D::run() on a tests/golden/run3d case."""
        sp = case["species"]
        names = (C.c_char_p * len(sp))(*[s["name"].encode() for s in sp])
        rows = f64([[s["charge"], s["mass"], s["weight"], s["particles"],
                     s["thermal_velocity"]] for s in sp]).reshape(-1)
        n, L = i64(case["n"]), f64(case["length"])
        bx, bv = f64(case["bcoords"]), f64(case["bvalues"]).reshape(-1)
        err = C.create_string_buffer(512)
        rc = self.lib.ref_run3d(output_dir.encode(), _ptr(n, _i64p), _ptr(L, _dp), case["dt"],
                                case["steps"], case["output_interval"], case["seed"],
                                case["b_axis"], bx.size, _ptr(bx, _dp), _ptr(bv, _dp), len(sp),
                                names, _ptr(rows, _dp), err, 512)
        if rc:
            raise RuntimeError(err.value.decode())


    def run2d(self, case: dict, output_dir: str):
        """The reference's Simulation2D::run() on a tests/golden/run2d case."""
        sp = case["species"]
        names = (C.c_char_p * len(sp))(*[s["name"].encode() for s in sp])
        rows = f64([[s["charge"], s["mass"], s["weight"], s["particles"],
                     s["thermal_velocity"]] for s in sp]).reshape(-1)
        n, L, B = i64(case["n"]), f64(case["length"]), f64(case["B"])
        err = C.create_string_buffer(512)
        rc = self.lib.ref_run2d(output_dir.encode(), _ptr(n, _i64p), _ptr(L, _dp),
                                case["depth"], case["dt"], case["steps"],
                                case["output_interval"], case["seed"], _ptr(B, _dp), len(sp),
                                names, _ptr(rows, _dp), err, 512)
        if rc:
            raise RuntimeError(err.value.decode())


def reference_available() -> bool:
    return os.path.exists(REF_SO)
