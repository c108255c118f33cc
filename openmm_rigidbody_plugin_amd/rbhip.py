"""ctypes binding of the gfx950 rigid-body step library (include/rbhip.h).

Loads the in-tree ``librbhip.so``; there is no fallback: if the library is
missing or no GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "librbhip.so")

MIXED = 0
DOUBLE = 1
PRECISIONS = {"mixed": MIXED, "double": DOUBLE}
KERNELS = ("part1", "forces", "part2", "fused")
BODY_FIELDS = dict(N=0, dof=1, mass=2, I=slice(3, 6), invI=slice(6, 9), rcm=slice(9, 12), pcm=slice(12, 15),
                   q=slice(15, 19), pi=slice(19, 23), force=slice(23, 26), torque=slice(26, 30), twoKt=30,
                   twoKr=31, loc=32)

_D, _I, _P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
_DP, _IP = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
_SIZEP = ctypes.POINTER(ctypes.c_size_t)
_LLP = ctypes.POINTER(ctypes.c_longlong)

_lib = None


class RbhipError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"rbhip error {code}: {message}")
        self.code = code


def lib() -> ctypes.CDLL:
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m openmm_rigidbody_plugin_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "rbhip_last_error": (ctypes.c_char_p, []),
            "rbhip_version": (_I, []),
            "rbhip_device_count": (_I, [_IP]),
            "rbhip_device_synchronize": (_I, []),
            "rbhip_create": (_I, [_I, _DP, _IP, _D, _I, _I, _I, ctypes.POINTER(_P)]),
            "rbhip_create_constrained": (_I, [_I, _DP, _IP, _I, _IP, _DP, _D, _I, _I, _I, ctypes.POINTER(_P)]),
            "rbhip_set_constraint_tolerance": (_I, [_P, _D]),
            "rbhip_set_virtual_sites": (_I, [_P, _I, _IP, _IP, _IP, _DP]),
            "rbhip_constraint_failures": (_I, [_P, _LLP]),
            "rbhip_destroy": (_I, [_P]),
            "rbhip_set_stream": (_I, [_P, _P]),
            "rbhip_device_bytes": (_I, [_P, _SIZEP]),
            "rbhip_get_counts": (_I, [_P, _IP]),
            "rbhip_get_atom_index": (_I, [_P, _IP]),
            "rbhip_get_body_index": (_I, [_P, _IP]),
            "rbhip_set_tether": (_I, [_P, _DP, _DP]),
            "rbhip_add_bonds": (_I, [_P, _I, _IP, _DP, _DP]),
            "rbhip_compute_forces": (_I, [_P, _DP]),
            "rbhip_set_forces": (_I, [_P, _DP]),
            "rbhip_adopt_arrays": (_I, [_P, _P, _P, _P, _P, _I, _IP, _I]),
            "rbhip_set_atom_order": (_I, [_P, _IP, _I]),
            "rbhip_state_changed": (_I, [_P, _I]),
            "rbhip_free_delta": (_I, [_P, _P]),
            "rbhip_part1_constrained": (_I, [_P, _P]),
            "rbhip_set_positions": (_I, [_P, _DP]),
            "rbhip_set_velocities": (_I, [_P, _DP]),
            "rbhip_get_positions": (_I, [_P, _DP]),
            "rbhip_get_velocities": (_I, [_P, _DP]),
            "rbhip_get_forces": (_I, [_P, _DP]),
            "rbhip_set_step_size": (_I, [_P, _D]),
            "rbhip_get_time": (_I, [_P, _DP, _LLP]),
            "rbhip_part1": (_I, [_P]),
            "rbhip_part2": (_I, [_P]),
            "rbhip_execute": (_I, [_P, _I]),
            "rbhip_use_graph": (_I, [_P, _I]),
            "rbhip_use_fusion": (_I, [_P, _I]),
            "rbhip_kinetic_energies": (_I, [_P, _DP]),
            "rbhip_potential_energy": (_I, [_P, _DP]),
            "rbhip_set_refined_energies": (_I, [_P, _I]),
            "rbhip_refined_kinetic_energies": (_I, [_P, _DP]),
            "rbhip_potential_energy_refinement": (_I, [_P, _DP]),
            "rbhip_get_body_states": (_I, [_P, _DP]),
            "rbhip_get_body_fixed_positions": (_I, [_P, _DP]),
            "rbhip_set_launch": (_I, [_P, _I, _I]),
            "rbhip_set_lanes": (_I, [_P, _I, _I]),
            "rbhip_set_shared_frames": (_I, [_P, _I]),
            "rbhip_autotune": (_I, [_P, _I]),
            "rbhip_describe": (_I, [_P, ctypes.c_char_p, _I]),
            "rbhip_time_kernels": (_I, [_P, _I, _DP, _DP]),
            "rbhip_copy_calibration": (_I, [ctypes.c_size_t, _I, _DP]),
            "rbhip_batch_create": (_I, [_I, ctypes.POINTER(_P)]),
            "rbhip_batch_destroy": (_I, [_P]),
            "rbhip_batch_set_members": (_I, [_P, _I, ctypes.POINTER(_P)]),
            "rbhip_batch_set_launch": (_I, [_P, _I, _I]),
            "rbhip_batch_set_stream": (_I, [_P, _P]),
            "rbhip_batch_use_graph": (_I, [_P, _I]),
            "rbhip_batch_use_fusion": (_I, [_P, _I]),
            "rbhip_batch_execute": (_I, [_P, _I]),
            "rbhip_batch_kinetic_energies": (_I, [_P, _DP]),
            "rbhip_batch_time_kernels": (_I, [_P, _I, _DP, _DP]),
            "rbhip_batch_autotune": (_I, [_P, _I]),
            "rbhip_batch_describe": (_I, [_P, ctypes.c_char_p, _I]),
            "rbhip_estimate_bytes": (_I, [_I, _IP, _I, _SIZEP, _IP]),
            "rbhip_body_bookkeeping": (_I, [_I, _P, _P, _P, _P, _P, _IP, _I, _P]),
            "rbhip_body_bookkeeping_noalloc": (_I, [_I, _P, _P, _P, _P, _P, _IP, _I, _P]),
            "rbhip_server_last_error": (ctypes.c_char_p, []),
            "rbhip_server_create": (_I, [_I, ctypes.c_size_t, _I, _I, _I, ctypes.POINTER(_P)]),
            "rbhip_server_destroy": (_I, [_P]),
            "rbhip_server_set_admission_order": (_I, [_P, _I]),
            "rbhip_server_submit": (_I, [_P, _I, _DP, _IP, _DP, _DP, _DP, _DP, _D, _I, ctypes.c_longlong, _LLP]),
            "rbhip_server_run_round": (_I, [_P, _IP]),
            "rbhip_server_status": (_I, [_P, _LLP, _SIZEP, _LLP]),
            "rbhip_server_completed": (_I, [_P, _LLP, _I, _IP]),
            "rbhip_server_profile": (_I, [_P, _LLP, _DP]),
            "rbhip_server_take": (_I, [_P, ctypes.c_longlong, _DP, _DP, _DP, _DP]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names of the C ABI entry points this binding expects."""
    lib()
    return sorted(n for n in dir(_lib) if n.startswith("rbhip_"))


def _check(rc):
    if rc != 0:
        raise RbhipError(rc, lib().rbhip_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


def _vec_in(x, n):
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.shape != (n, 3):
        raise ValueError(f"expected shape ({n}, 3), got {a.shape}")
    return a


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(lib().rbhip_device_count(ctypes.byref(n)))
    return n.value


def synchronize():
    _check(lib().rbhip_device_synchronize())


def copy_calibration(nbytes: int, reps: int = 20) -> float:
    ms = ctypes.c_double()
    _check(lib().rbhip_copy_calibration(nbytes, reps, ctypes.byref(ms)))
    return ms.value


class Context:
    """One rigid-body system on one GPU: RigidBodyIntegrator + its HIP step kernel."""

    def __init__(self, masses, body_indices, dt, rotation_mode=10, precision="mixed", device=0, constraints=None):
        """`constraints`: (atom1, atom2, distance) triples between free atoms
        (System::addConstraint), or None."""
        m = np.ascontiguousarray(masses, dtype=np.float64)
        b = np.ascontiguousarray(body_indices, dtype=np.int32)
        if m.shape != b.shape or m.ndim != 1:
            raise ValueError("masses and body indices must be 1-D arrays of equal length")
        self.n = m.size
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self._h = _P()
        if constraints is not None and len(constraints) > 0:
            pairs = np.ascontiguousarray([[int(a), int(c)] for a, c, _ in constraints], dtype=np.int32)
            dist = np.ascontiguousarray([float(d) for _, _, d in constraints], dtype=np.float64)
            _check(lib().rbhip_create_constrained(self.n, _dp(m), _ip(b), len(dist), _ip(pairs), _dp(dist),
                                                  float(dt), int(rotation_mode), self.precision, int(device),
                                                  ctypes.byref(self._h)))
        else:
            _check(lib().rbhip_create(self.n, _dp(m), _ip(b), float(dt), int(rotation_mode), self.precision,
                                      int(device), ctypes.byref(self._h)))

    @classmethod
    def from_spec(cls, spec, precision="mixed", rotation_mode=None, dt=None, device=0, set_state=True,
                  constraints=None, shared_frames=True):
        c = cls(spec.masses, spec.body_indices, spec.dt if dt is None else dt,
                spec.rotation_mode if rotation_mode is None else rotation_mode, precision, device, constraints)
        if not shared_frames:
            c.set_shared_frames(False)
        c.set_tether(spec.tether_k, spec.tether_r0)
        if set_state:
            c.set_positions(spec.positions)
            c.set_velocities(spec.velocities)
        return c

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().rbhip_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ---- virtual sites ----
    SITE_TWO_AVERAGE, SITE_THREE_AVERAGE, SITE_OUT_OF_PLANE, SITE_LOCAL_COORDINATES = 1, 2, 3, 4

    def set_virtual_sites(self, sites):
        """sites: (kind, site, parents, weights) tuples; kind SITE_TWO_AVERAGE
        (2 parents, weights w1 w2), SITE_THREE_AVERAGE (3, w1 w2 w3) or
        SITE_OUT_OF_PLANE (3, w12 w13 wCross) or SITE_LOCAL_COORDINATES (3,
        origin, x and y weights and the local position: 12 values)."""
        n = len(sites)
        t = np.ascontiguousarray([int(k) for k, _, _, _ in sites], dtype=np.int32)
        a = np.ascontiguousarray([int(x) for _, x, _, _ in sites], dtype=np.int32)
        p = np.zeros((n, 3), dtype=np.int32)
        w = np.zeros((n, 12))
        for i, (_, _, par, wt) in enumerate(sites):
            p[i, :len(par)] = par
            w[i, :len(wt)] = wt
        _check(lib().rbhip_set_virtual_sites(self._h, n, _ip(t), _ip(a), _ip(p), _dp(w)))

    # ---- constraints ----
    def set_constraint_tolerance(self, tol):
        _check(lib().rbhip_set_constraint_tolerance(self._h, float(tol)))

    def constraint_failures(self):
        n = ctypes.c_longlong()
        _check(lib().rbhip_constraint_failures(self._h, ctypes.byref(n)))
        return n.value

    # ---- bookkeeping ----
    def counts(self):
        out = np.empty(5, dtype=np.int32)
        _check(lib().rbhip_get_counts(self._h, _ip(out)))
        return dict(zip(("numFree", "numBodies", "numActualAtoms", "numBodyAtoms", "numDOF"),
                        (int(x) for x in out)))

    def atom_index(self):
        out = np.empty(self.counts()["numActualAtoms"], dtype=np.int32)
        _check(lib().rbhip_get_atom_index(self._h, _ip(out)))
        return out

    def body_index(self):
        out = np.empty(self.n, dtype=np.int32)
        _check(lib().rbhip_get_body_index(self._h, _ip(out)))
        return out

    def device_bytes(self):
        n = ctypes.c_size_t()
        _check(lib().rbhip_device_bytes(self._h, ctypes.byref(n)))
        return n.value

    # ---- force field ----
    def set_tether(self, k, r0):
        kk = np.ascontiguousarray(np.broadcast_to(np.asarray(k, dtype=np.float64), (self.n,)))
        _check(lib().rbhip_set_tether(self._h, _dp(kk), _dp(_vec_in(r0, self.n))))

    def add_bonds(self, pairs, length, k):
        p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        ln = np.ascontiguousarray(np.broadcast_to(np.asarray(length, dtype=np.float64), (p.shape[0],)))
        kk = np.ascontiguousarray(np.broadcast_to(np.asarray(k, dtype=np.float64), (p.shape[0],)))
        _check(lib().rbhip_add_bonds(self._h, p.shape[0], _ip(p), _dp(ln), _dp(kk)))

    def compute_forces(self, energy=False):
        e = ctypes.c_double()
        _check(lib().rbhip_compute_forces(self._h, ctypes.byref(e) if energy else None))
        return e.value if energy else None

    def set_forces(self, f):
        """Forces from outside the library (the force slot): (n, 3) FP64 in
        original atom order, for the next part 2 (see rbhip_set_forces)."""
        _check(lib().rbhip_set_forces(self._h, _dp(_vec_in(f, self.n))))

    # ---- adopted arrays in OpenMM's GPU layout ----
    STATE_POSITIONS, STATE_VELOCITIES = 1, 2

    def adopt_arrays(self, posq, posq_correction, velm, force, padded_num_atoms, atom_order):
        """Device pointers (ints) of posq, posqCorrection (None in double
        precision), velm and the fixed-point force array, their padded
        length, and the atom order (slot -> original atom)."""
        order = np.ascontiguousarray(atom_order, dtype=np.int32)
        _check(lib().rbhip_adopt_arrays(self._h, posq, posq_correction, velm, force, int(padded_num_atoms),
                                        _ip(order), order.size))

    def set_atom_order(self, atom_order):
        order = np.ascontiguousarray(atom_order, dtype=np.int32)
        _check(lib().rbhip_set_atom_order(self._h, _ip(order), order.size))

    def state_changed(self, what):
        _check(lib().rbhip_state_changed(self._h, int(what)))

    # ---- state ----
    def set_positions(self, x):
        _check(lib().rbhip_set_positions(self._h, _dp(_vec_in(x, self.n))))

    def set_velocities(self, v):
        _check(lib().rbhip_set_velocities(self._h, _dp(_vec_in(v, self.n))))

    def _get(self, fn):
        out = np.empty((self.n, 3))
        _check(fn(self._h, _dp(out)))
        return out

    def positions(self):
        return self._get(lib().rbhip_get_positions)

    def velocities(self):
        return self._get(lib().rbhip_get_velocities)

    def forces(self):
        return self._get(lib().rbhip_get_forces)

    def set_step_size(self, dt):
        _check(lib().rbhip_set_step_size(self._h, float(dt)))

    def time(self):
        t, s = ctypes.c_double(), ctypes.c_longlong()
        _check(lib().rbhip_get_time(self._h, ctypes.byref(t), ctypes.byref(s)))
        return t.value, s.value

    # ---- stepping ----
    def part1(self):
        _check(lib().rbhip_part1(self._h))

    def part2(self):
        _check(lib().rbhip_part2(self._h))

    def free_delta(self, pos_delta):
        """adopted arrays: the free atoms' drift displacements into the
        device array pos_delta (double4 per slot) for the caller's
        constraint solver"""
        _check(lib().rbhip_free_delta(self._h, pos_delta))

    def part1_constrained(self, pos_delta):
        """adopted arrays: part 1 with the free atoms moved by the caller's
        constrained displacements (device pointer, double4 per slot)"""
        _check(lib().rbhip_part1_constrained(self._h, pos_delta))

    def step(self, steps=1):
        _check(lib().rbhip_execute(self._h, int(steps)))

    def use_graph(self, enable=True):
        _check(lib().rbhip_use_graph(self._h, int(bool(enable))))

    def use_fusion(self, enable=True):
        _check(lib().rbhip_use_fusion(self._h, int(bool(enable))))

    def kinetic_energies(self):
        out = np.empty(2)
        _check(lib().rbhip_kinetic_energies(self._h, _dp(out)))
        return out

    def potential_energy(self):
        e = ctypes.c_double()
        _check(lib().rbhip_potential_energy(self._h, ctypes.byref(e)))
        return e.value

    def set_refined(self, enable=True):
        """setComputeRefinedEnergies: accumulate the refined-energy terms on the
        last step of every call (the accumulators start empty)."""
        _check(lib().rbhip_set_refined_energies(self._h, int(bool(enable))))

    def refined_kinetic_energies(self):
        """getRefinedKineticEnergies: {translational, rotational}."""
        out = np.empty(2)
        _check(lib().rbhip_refined_kinetic_energies(self._h, _dp(out)))
        return out

    def potential_energy_refinement(self):
        """getPotentialEnergyRefinement: -dt^2/24 sum F.M^-1 F of the current forces."""
        e = ctypes.c_double()
        _check(lib().rbhip_potential_energy_refinement(self._h, ctypes.byref(e)))
        return e.value

    def body_states_array(self):
        nb = self.counts()["numBodies"]
        out = np.empty((nb, 36))
        if nb:
            _check(lib().rbhip_get_body_states(self._h, _dp(out)))
        return out

    def body_states(self):
        arr = self.body_states_array()
        return [{k: row[sl] for k, sl in BODY_FIELDS.items()} for row in arr]

    def body_fixed_positions(self):
        nba = self.counts()["numBodyAtoms"]
        out = np.empty((nba, 3))
        if nba:
            _check(lib().rbhip_get_body_fixed_positions(self._h, _dp(out)))
        return out

    # ---- launch configuration and measurement ----
    def set_launch(self, block=256, max_blocks=0):
        _check(lib().rbhip_set_launch(self._h, int(block), int(max_blocks)))

    def set_lanes(self, atoms_per_body, lanes):
        _check(lib().rbhip_set_lanes(self._h, int(atoms_per_body), int(lanes)))

    def set_shared_frames(self, enable=True):
        """bodies of one shape read shared body-fixed coordinates (mixed
        precision; applies from the next position upload)"""
        _check(lib().rbhip_set_shared_frames(self._h, 1 if enable else 0))

    def autotune(self, reps=5):
        _check(lib().rbhip_autotune(self._h, int(reps)))

    def describe(self):
        buf = ctypes.create_string_buffer(4096)
        _check(lib().rbhip_describe(self._h, buf, len(buf)))
        return buf.value.decode()

    def time_kernels(self, steps):
        """Mean kernel durations (ms) over `steps` two-pass steps and the fused
        path of step(steps + 1); advances the state by 2 * steps + 1 steps."""
        avg = np.empty(len(KERNELS))
        tot = ctypes.c_double()
        _check(lib().rbhip_time_kernels(self._h, int(steps), _dp(avg), ctypes.byref(tot)))
        return dict(zip(KERNELS, avg)), tot.value


class Batch:
    """Several contexts stepped by one set of launches."""

    def __init__(self, device=0):
        self._h = _P()
        self.members = []
        _check(lib().rbhip_batch_create(int(device), ctypes.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().rbhip_batch_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_members(self, contexts):
        arr = (_P * max(len(contexts), 1))(*[c.handle for c in contexts])
        _check(lib().rbhip_batch_set_members(self._h, len(contexts), arr))
        self.members = list(contexts)

    def set_launch(self, block=256, max_blocks=0):
        _check(lib().rbhip_batch_set_launch(self._h, int(block), int(max_blocks)))

    def use_graph(self, enable=True):
        _check(lib().rbhip_batch_use_graph(self._h, int(bool(enable))))

    def use_fusion(self, enable=True):
        _check(lib().rbhip_batch_use_fusion(self._h, int(bool(enable))))

    def step(self, steps=1):
        _check(lib().rbhip_batch_execute(self._h, int(steps)))

    def kinetic_energies(self):
        out = np.empty((max(len(self.members), 1), 2))
        if self.members:
            _check(lib().rbhip_batch_kinetic_energies(self._h, _dp(out)))
        return out[:len(self.members)]

    def autotune(self, reps=3):
        _check(lib().rbhip_batch_autotune(self._h, int(reps)))

    def describe(self):
        buf = ctypes.create_string_buffer(512)
        _check(lib().rbhip_batch_describe(self._h, buf, 512))
        return buf.value.decode()

    def time_kernels(self, steps):
        avg = np.empty(len(KERNELS))
        tot = ctypes.c_double()
        _check(lib().rbhip_batch_time_kernels(self._h, int(steps), _dp(avg), ctypes.byref(tot)))
        return dict(zip(KERNELS, avg)), tot.value


def body_bookkeeping(num_atoms, body_indices_ptr, masses_ptr, body_index_out, atom_index_out, body_start_out,
                     device=0, stream=None):
    """Device bookkeeping (rbhip_body_bookkeeping) on device arrays given by
    their addresses; returns (numFree, numBodies, numMassless, numBodyAtoms)."""
    c = np.zeros(4, dtype=np.int32)
    _check(lib().rbhip_body_bookkeeping(int(num_atoms), body_indices_ptr, masses_ptr, body_index_out, atom_index_out,
                                        body_start_out, _ip(c), int(device), stream))
    return tuple(int(x) for x in c)


def body_bookkeeping_noalloc(num_atoms, body_indices_ptr, masses_ptr, body_index_out, atom_index_out,
                             body_start_out, device=0, stream=None):
    """rbhip_body_bookkeeping_noalloc: the same on the library's static
    scratch, with no device allocation."""
    c = np.zeros(4, dtype=np.int32)
    _check(lib().rbhip_body_bookkeeping_noalloc(int(num_atoms), body_indices_ptr, masses_ptr, body_index_out,
                                                atom_index_out, body_start_out, _ip(c), int(device), stream))
    return tuple(int(x) for x in c)


def estimate_bytes(body_indices, precision="mixed"):
    """Device bytes a context for this system holds (upper bound before bonds)."""
    bi = np.ascontiguousarray(body_indices, dtype=np.int32)
    n = ctypes.c_size_t()
    _check(lib().rbhip_estimate_bytes(bi.size, _ip(bi), PRECISIONS[precision], ctypes.byref(n), None))
    return n.value


class ServerError(RbhipError):
    pass


def _scheck(status):
    if status != 0:
        raise ServerError(status, lib().rbhip_server_last_error().decode())


class Server:
    """Job server: admission by HBM budget, continuous batching (rbhip_server_*)."""

    ADMISSION_ORDERS = {"fifo": 0, "longest_first": 1}

    def __init__(self, hbm_budget, steps_per_round=100, precision="mixed", block=256, device=0,
                 admission="fifo"):
        self._h = _P()
        _scheck(lib().rbhip_server_create(int(device), int(hbm_budget), int(steps_per_round),
                                          PRECISIONS[precision], int(block), ctypes.byref(self._h)))
        self.num_atoms = {}
        if admission != "fifo":
            self.set_admission_order(admission)
        # the server reads a job's arrays when it sets the job up: they are
        # kept alive here until the job is taken or the server closed
        self._inputs = {}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().rbhip_server_destroy(self._h)
            self._h = _P()
        self._inputs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_admission_order(self, order):
        """'fifo' (submission order) or 'longest_first' (most steps first)."""
        if order not in self.ADMISSION_ORDERS:
            raise ValueError(f"unknown admission order {order!r}")
        _scheck(lib().rbhip_server_set_admission_order(self._h, self.ADMISSION_ORDERS[order]))

    def submit(self, spec, steps):
        """Queues a SystemSpec for `steps` steps; returns the job id.  The
        spec's arrays are read when the job is set up, not copied here: do not
        modify them in place until the job has been taken."""
        n = spec.num_atoms
        m = np.ascontiguousarray(spec.masses, dtype=np.float64)
        bi = np.ascontiguousarray(spec.body_indices, dtype=np.int32)
        x = np.ascontiguousarray(spec.positions, dtype=np.float64)
        v = np.ascontiguousarray(spec.velocities, dtype=np.float64)
        k = np.ascontiguousarray(spec.tether_k, dtype=np.float64)
        r0 = np.ascontiguousarray(spec.tether_r0, dtype=np.float64)
        jid = ctypes.c_longlong()
        _scheck(lib().rbhip_server_submit(self._h, n, _dp(m), _ip(bi), _dp(x), _dp(v), _dp(k), _dp(r0),
                                          float(spec.dt), int(spec.rotation_mode), int(steps), ctypes.byref(jid)))
        self.num_atoms[jid.value] = n
        self._inputs[jid.value] = (m, bi, x, v, k, r0)
        return jid.value

    def run_round(self):
        f = ctypes.c_int()
        _scheck(lib().rbhip_server_run_round(self._h, ctypes.byref(f)))
        return f.value

    def status(self):
        c = (ctypes.c_longlong * 3)()
        b, bs = ctypes.c_size_t(), ctypes.c_longlong()
        _scheck(lib().rbhip_server_status(self._h, c, ctypes.byref(b), ctypes.byref(bs)))
        return {"waiting": c[0], "resident": c[1], "completed": c[2], "bytes_in_use": b.value,
                "body_steps": bs.value}

    def profile(self):
        """Rounds queued (and those cut short by a finishing job), host seconds
        per phase and the rounds' summed duration on the device."""
        r = (ctypes.c_longlong * 2)()
        t = np.empty(6)
        _scheck(lib().rbhip_server_profile(self._h, r, _dp(t)))
        return {"rounds": r[0], "short_rounds": r[1], "admit_s": float(t[0]), "collect_s": float(t[1]),
                "wait_s": float(t[2]), "members_s": float(t[3]), "launch_s": float(t[4]), "device_s": float(t[5])}

    def completed(self, max_ids=4096):
        ids = (ctypes.c_longlong * max_ids)()
        n = ctypes.c_int()
        _scheck(lib().rbhip_server_completed(self._h, ids, max_ids, ctypes.byref(n)))
        return [ids[i] for i in range(n.value)]

    def take(self, job_id):
        """Final positions, velocities, {KE_t, KE_r} and timings (wait, resident,
        latency in seconds) of a completed job."""
        n = self.num_atoms.pop(job_id)
        x, v, ke, t = np.empty((n, 3)), np.empty((n, 3)), np.empty(2), np.empty(3)
        _scheck(lib().rbhip_server_take(self._h, int(job_id), _dp(x), _dp(v), _dp(ke), _dp(t)))
        self._inputs.pop(job_id, None)
        return {"positions": x, "velocities": v, "kinetic": ke, "wait": t[0], "resident": t[1], "latency": t[2]}
