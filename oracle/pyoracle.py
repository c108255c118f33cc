"""ctypes binding of the FP64 CPU restatement (librbref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# RBREF_LIB selects another build of the same source (the CPU baseline times
# one built for its host)
LIB_PATH = os.environ.get("RBREF_LIB") or os.path.join(HERE, "librbref.so")
REF_ELLIPTIC_PATH = os.path.join(HERE, "_ref", "libref_elliptic.so")

FIX_PCM_RESET = 1
FIX_LINEAR_FRAME = 2
FIX_LINEAR_EXACT = 4
FIX_STALE_SYSTEM = 8
FIX_ORTHONORMAL = 16
FIX_REFINED_FREE = 32
FIX_ALL = 63
# planted mistake (tests only): body-atom velocities rebuilt as delta x w
PLANT_VELOCITY_SIGN = 1024
SITE_TWO_AVERAGE = 1
SITE_THREE_AVERAGE = 2
SITE_OUT_OF_PLANE = 3
SITE_LOCAL_COORDINATES = 4

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "rbref_last_error": (ctypes.c_char_p, []),
            "rbref_create": (_P, [_I, _DP, _IP, _D, _I, ctypes.c_uint]),
            "rbref_destroy": (None, [_P]),
            "rbref_set_tether": (_I, [_P, _DP, _DP]),
            "rbref_add_bond": (_I, [_P, _I, _I, _D, _D]),
            "rbref_add_constraint": (_I, [_P, _I, _I, _D]),
            "rbref_add_virtual_site": (_I, [_P, _I, _I, _IP, _DP]),
            "rbref_set_constraint_tolerance": (_I, [_P, _D]),
            "rbref_constraint_failures": (ctypes.c_long, [_P]),
            "rbref_set_positions": (_I, [_P, _DP]),
            "rbref_set_velocities": (_I, [_P, _DP]),
            "rbref_set_step_size": (_I, [_P, _D]),
            "rbref_step": (_I, [_P, _I]),
            "rbref_get_positions": (_I, [_P, _DP]),
            "rbref_get_velocities": (_I, [_P, _DP]),
            "rbref_get_forces": (_I, [_P, _DP]),
            "rbref_kinetic_energies": (_I, [_P, _DP]),
            "rbref_potential_energy": (_D, [_P]),
            "rbref_set_refined": (_I, [_P, _I]),
            "rbref_refined_kinetic_energies": (_I, [_P, _DP]),
            "rbref_potential_energy_refinement": (_D, [_P]),
            "rbref_time": (_D, [_P]),
            "rbref_counts": (_I, [_P, _IP]),
            "rbref_atom_index": (_I, [_P, _IP]),
            "rbref_body_index": (_I, [_P, _IP]),
            "rbref_body_state": (_I, [_P, _I, _DP]),
            "rbref_body_fixed_positions": (_I, [_P, _DP]),
            "rbref_carlson_rf": (_D, [_D, _D, _D]),
            "rbref_carlson_rj": (_D, [_D, _D, _D, _D]),
            "rbref_carlson_rc": (_D, [_D, _D]),
            "rbref_omega": (_D, [_D, _D, _D]),
            "rbref_jacobi": (None, [_D, _D, _DP, _DP, _DP]),
            "rbref_eigen": (None, [_DP, _DP, _DP]),
            "rbref_quat_from_matrix": (None, [_DP, _DP]),
            "rbref_quat_rotate": (None, [_DP, _DP, _I, _DP]),
            "rbref_clean_body_indices": (None, [_I, _IP, _IP]),
            "rbref_rotate": (None, [_DP, _DP, _DP, _DP, _I, _D, _I, ctypes.c_uint]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


def _err():
    return lib().rbref_last_error().decode()


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(_err())


class OracleSim:
    """RigidBodyIntegrator on the FP64 reference platform restatement."""

    BODY_FIELDS = dict(N=0, dof=1, mass=2, I=slice(3, 6), invI=slice(6, 9), rcm=slice(9, 12),
                       pcm=slice(12, 15), q=slice(15, 19), pi=slice(19, 23), force=slice(23, 26),
                       torque=slice(26, 30), twoKt=30, twoKr=31, loc=32)

    def __init__(self, masses, body_indices, dt, rotation_mode, flags=FIX_ALL):
        L = lib()
        m = np.ascontiguousarray(masses, dtype=np.float64)
        b = np.ascontiguousarray(body_indices, dtype=np.int32)
        self.n = m.size
        self._h = L.rbref_create(self.n, _dp(m), _ip(b), float(dt), int(rotation_mode), int(flags))
        if not self._h:
            raise OracleError(_err())

    @classmethod
    def from_spec(cls, spec, flags=FIX_ALL, rotation_mode=None, dt=None):
        s = cls(spec.masses, spec.body_indices, spec.dt if dt is None else dt,
                spec.rotation_mode if rotation_mode is None else rotation_mode, flags)
        s.set_tether(spec.tether_k, spec.tether_r0)
        s.set_positions(spec.positions)
        s.set_velocities(spec.velocities)
        return s

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().rbref_destroy(h)
            self._h = None

    def set_tether(self, k, r0):
        k = np.ascontiguousarray(np.broadcast_to(k, (self.n,)), dtype=np.float64)
        r0 = np.ascontiguousarray(r0, dtype=np.float64)
        _check(lib().rbref_set_tether(self._h, _dp(k), _dp(r0)))

    def add_bond(self, a, b, length, k):
        _check(lib().rbref_add_bond(self._h, a, b, length, k))

    def add_constraint(self, a, b, distance):
        _check(lib().rbref_add_constraint(self._h, int(a), int(b), float(distance)))

    def add_virtual_site(self, kind, site, parents, weights):
        """kind: SITE_TWO_AVERAGE, SITE_THREE_AVERAGE, SITE_OUT_OF_PLANE or
        SITE_LOCAL_COORDINATES (12 weights: origin, x, y, local position)."""
        p = np.zeros(3, dtype=np.int32)
        p[:len(parents)] = parents
        w = np.zeros(12)
        w[:len(weights)] = weights
        _check(lib().rbref_add_virtual_site(self._h, int(kind), int(site), _ip(p), _dp(w)))

    def set_constraint_tolerance(self, tol):
        _check(lib().rbref_set_constraint_tolerance(self._h, float(tol)))

    def constraint_failures(self):
        return int(lib().rbref_constraint_failures(self._h))

    def set_positions(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        _check(lib().rbref_set_positions(self._h, _dp(x)))

    def set_velocities(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(lib().rbref_set_velocities(self._h, _dp(v)))

    def set_step_size(self, dt):
        _check(lib().rbref_set_step_size(self._h, float(dt)))

    def step(self, n):
        _check(lib().rbref_step(self._h, int(n)))

    def _vec(self, fn, n=None):
        out = np.empty((self.n if n is None else n, 3))
        _check(fn(self._h, _dp(out)))
        return out

    def positions(self):
        return self._vec(lib().rbref_get_positions)

    def velocities(self):
        return self._vec(lib().rbref_get_velocities)

    def forces(self):
        return self._vec(lib().rbref_get_forces)

    def kinetic_energies(self):
        out = np.empty(2)
        _check(lib().rbref_kinetic_energies(self._h, _dp(out)))
        return out

    def potential_energy(self):
        return lib().rbref_potential_energy(self._h)

    def set_refined(self, enable=True):
        _check(lib().rbref_set_refined(self._h, int(bool(enable))))

    def refined_kinetic_energies(self):
        out = np.empty(2)
        _check(lib().rbref_refined_kinetic_energies(self._h, _dp(out)))
        return out

    def potential_energy_refinement(self):
        return lib().rbref_potential_energy_refinement(self._h)

    def time(self):
        return lib().rbref_time(self._h)

    def counts(self):
        out = np.empty(5, dtype=np.int32)
        _check(lib().rbref_counts(self._h, _ip(out)))
        return dict(zip(("numFree", "numBodies", "numActualAtoms", "numBodyAtoms", "numDOF"),
                        (int(x) for x in out)))

    def atom_index(self):
        c = self.counts()
        out = np.empty(c["numActualAtoms"], dtype=np.int32)
        _check(lib().rbref_atom_index(self._h, _ip(out)))
        return out

    def body_index(self):
        out = np.empty(self.n, dtype=np.int32)
        _check(lib().rbref_body_index(self._h, _ip(out)))
        return out

    def body_state(self, k):
        out = np.empty(36)
        _check(lib().rbref_body_state(self._h, int(k), _dp(out)))
        return {name: out[sl] for name, sl in self.BODY_FIELDS.items()}

    def body_states(self):
        nb = self.counts()["numBodies"]
        return [self.body_state(k) for k in range(nb)]

    def body_fixed_positions(self):
        return self._vec(lib().rbref_body_fixed_positions, self.counts()["numBodyAtoms"])


# ---- unit-level functions ----
def carlson_rf(x, y, z):
    return lib().rbref_carlson_rf(x, y, z)


def carlson_rj(x, y, z, p):
    return lib().rbref_carlson_rj(x, y, z, p)


def carlson_rc(x, y):
    return lib().rbref_carlson_rc(x, y)


def omega(x, n, m):
    return lib().rbref_omega(x, n, m)


def jacobi(u, m):
    sn, cn, dn = _D(), _D(), _D()
    lib().rbref_jacobi(u, m, ctypes.byref(sn), ctypes.byref(cn), ctypes.byref(dn))
    return sn.value, cn.value, dn.value


def eigen(A):
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    w, V = np.empty(3), np.empty(9)
    lib().rbref_eigen(_dp(A), _dp(w), _dp(V))
    return w, V.reshape(3, 3)


def quat_from_matrix(A):
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    q = np.empty(4)
    lib().rbref_quat_from_matrix(_dp(A), _dp(q))
    return q


def quat_rotate(q, x, transpose=False):
    q = np.ascontiguousarray(q, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(3)
    lib().rbref_quat_rotate(_dp(q), _dp(x), int(transpose), _dp(y))
    return y


def clean_body_indices(idx):
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    out = np.empty_like(idx)
    lib().rbref_clean_body_indices(idx.size, _ip(idx), _ip(out))
    return out


def rotate(q, pi, I, invI, dof, dt, mode, flags=FIX_ALL):
    q = np.array(q, dtype=np.float64)
    pi = np.array(pi, dtype=np.float64)
    I = np.ascontiguousarray(I, dtype=np.float64)
    invI = np.ascontiguousarray(invI, dtype=np.float64)
    lib().rbref_rotate(_dp(q), _dp(pi), _dp(I), _dp(invI), int(dof), float(dt), int(mode), int(flags))
    return q, pi


def reference_elliptic():
    """The reference's own elliptic routines (oracle/_ref), or None if not built."""
    if not os.path.exists(REF_ELLIPTIC_PATH):
        return None
    L = ctypes.CDLL(REF_ELLIPTIC_PATH)
    for f, n in (("ref_carlson_rf", 3), ("ref_carlson_rj", 4), ("ref_carlson_rc", 2), ("ref_omega", 3)):
        getattr(L, f).restype = _D
        getattr(L, f).argtypes = [_D] * n
    L.ref_jacobi.restype = None
    L.ref_jacobi.argtypes = [_D, _D, _DP, _DP, _DP]
    return L
