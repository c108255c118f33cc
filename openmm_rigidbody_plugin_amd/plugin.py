"""ctypes binding of the host plugin library (include/rigidbody_hip_plugin.h).

One ``Simulation`` is a System + RigidBodyIntegrator + Context on the HIP
platform of the C++ host layer, which drives the gfx950 library through the
rbhip C ABI.  No fallback: without the built libraries or a GPU the calls
raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import rbhip  # loads librbhip.so first, which the plugin library links

_PKG = os.path.dirname(os.path.abspath(__file__))
PLUGIN_PATH = os.path.join(_PKG, "librigidbody_hip.so")

_D, _I, _P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
_DP, _IP = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)

_SIG = {
    "registerPlatforms": (None, []),
    "registerKernelFactories": (None, []),
    "registerRigidBodyHipKernelFactories": (None, []),
    "rbplugin_last_error": (ctypes.c_char_p, []),
    "rbplugin_create": (_I, [_I, _DP, _IP, _D, _I, _I, _I, _I, _IP, _DP, _DP, _DP, _DP, ctypes.POINTER(_P)]),
    "rbplugin_create_ex": (_I, [_I, _DP, _IP, _D, _I, _I, _I, _I, _I, _IP, _DP, _DP, _DP, _DP,
                                ctypes.POINTER(_P)]),
    "rbplugin_create_with_force": (_I, [_I, _DP, _IP, _D, _I, _I, _I, _I, _I, _IP, _DP, _DP, _DP, _DP, _P, _P,
                                        ctypes.POINTER(_P)]),
    "rbplugin_destroy": (_I, [_P]),
    "rbplugin_set_positions": (_I, [_P, _DP]),
    "rbplugin_set_velocities": (_I, [_P, _DP]),
    "rbplugin_step": (_I, [_P, _I]),
    "rbplugin_set_step_size": (_I, [_P, _D]),
    "rbplugin_get_state": (_I, [_P, _DP, _DP, _DP, _DP]),
    "rbplugin_energies": (_I, [_P, _DP]),
    "rbplugin_counts": (_I, [_P, _IP]),
    "rbplugin_atom_index": (_I, [_P, _IP]),
    "rbplugin_rigid_body": (_I, [_P, _I, _DP]),
}

_lib = None


class PluginError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(PLUGIN_PATH):
            raise RuntimeError(f"{PLUGIN_PATH} is missing: build it with "
                               "`python -m openmm_rigidbody_plugin_amd.build`")
        rbhip.lib()
        L = ctypes.CDLL(PLUGIN_PATH)
        for name, (res, args) in _SIG.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIG)


def _check(status):
    if status != 0:
        raise PluginError(lib().rbplugin_last_error().decode())


# rbplugin_force_fn: double (*)(void* user, int numAtoms, const double* positions, double* forces)
_FORCE_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_int, _DP, _DP)


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


BODY_FIELDS = dict(N=0, dof=1, mass=2, I=slice(3, 6), invI=slice(6, 9), rcm=slice(9, 12), pcm=slice(12, 15),
                   q=slice(15, 19), pi=slice(19, 23), force=slice(23, 26), torque=slice(26, 30), twoKt=30,
                   twoKr=31, loc=32)


class Simulation:
    """System + RigidBodyIntegrator + Context (HIP platform) of the C++ host layer."""

    def __init__(self, masses, body_indices, step_size, rotation_mode=0, precision="mixed", device=0,
                 bonds=None, tether_k=None, tether_r0=None, refined_energies=False, host_force=None):
        """host_force: a force field evaluated on the host, f(positions) ->
        (forces, energy) with (n, 3) arrays in original atom order; it runs
        at every force evaluation (one execute() per step)."""
        self._h = None
        self._callback = None
        self._callback_error = None
        m = np.ascontiguousarray(masses, dtype=np.float64)
        bi = np.ascontiguousarray(body_indices, dtype=np.int32)
        n = m.size
        if bi.size != n:
            raise ValueError("one body index per atom")
        self.num_atoms = n
        prec = {"mixed": 0, "double": 1}[precision]
        if bonds:
            pairs = np.ascontiguousarray([(a, b) for a, b, _, _ in bonds], dtype=np.int32).ravel()
            blen = np.ascontiguousarray([L for _, _, L, _ in bonds], dtype=np.float64)
            bk = np.ascontiguousarray([k for _, _, _, k in bonds], dtype=np.float64)
            nb, pp, lp, kp = len(bonds), _ip(pairs), _dp(blen), _dp(bk)
        else:
            nb, pp, lp, kp = 0, None, None, None
        if tether_k is not None:
            tk = np.ascontiguousarray(tether_k, dtype=np.float64)
            tr = np.ascontiguousarray(tether_r0, dtype=np.float64).reshape(n, 3)
            tkp, trp = _dp(tk), _dp(tr)
        else:
            tkp, trp = None, None
        fn = None
        if host_force is not None:
            def trampoline(_user, count, xp, fp):
                # a Python error cannot cross the C++ layer: it is kept, the
                # forces left at zero, and raised by the call that ran it
                try:
                    x = np.ctypeslib.as_array(xp, shape=(count, 3)).copy()
                    f, e = host_force(x)
                    np.ctypeslib.as_array(fp, shape=(count, 3))[:] += np.asarray(f, dtype=np.float64)
                    return float(e)
                except BaseException as err:   # noqa: BLE001
                    self._callback_error = err
                    return 0.0
            fn = _FORCE_FN(trampoline)
            self._callback = fn
        h = _P()
        self._raise_after(lib().rbplugin_create_with_force(
            n, _dp(m), _ip(bi), float(step_size), int(rotation_mode), int(bool(refined_energies)), prec,
            int(device), nb, pp, lp, kp, tkp, trp, fn, None, ctypes.byref(h)))
        self._h = h

    def _raise_after(self, status):
        err, self._callback_error = self._callback_error, None
        if err is not None:
            raise PluginError(f"host force failed: {err!r}") from err
        _check(status)

    @classmethod
    def from_spec(cls, spec, precision="mixed", device=0, refined_energies=False):
        """A SystemSpec (openmm_rigidbody_plugin_amd.systems) with its tether,
        positions then velocities set, as the reference's tests do."""
        s = cls(spec.masses, spec.body_indices, spec.dt, spec.rotation_mode, precision, device,
                tether_k=spec.tether_k, tether_r0=spec.tether_r0, refined_energies=refined_energies)
        s.set_positions(spec.positions)
        if spec.velocities is not None:
            s.set_velocities(spec.velocities)
        return s

    def close(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.rbplugin_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_positions(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64).reshape(self.num_atoms, 3)
        self._raise_after(lib().rbplugin_set_positions(self._h, _dp(a)))

    def set_velocities(self, v):
        a = np.ascontiguousarray(v, dtype=np.float64).reshape(self.num_atoms, 3)
        self._raise_after(lib().rbplugin_set_velocities(self._h, _dp(a)))

    def step(self, steps):
        self._raise_after(lib().rbplugin_step(self._h, int(steps)))

    def set_step_size(self, dt):
        _check(lib().rbplugin_set_step_size(self._h, float(dt)))

    def state(self, positions=True, velocities=True, forces=False, energy=True):
        """Context::getState: dict with the requested arrays and, with energy,
        kinetic, potential and time."""
        n = self.num_atoms
        x = np.empty((n, 3)) if positions else None
        v = np.empty((n, 3)) if velocities else None
        f = np.empty((n, 3)) if forces else None
        e = np.empty(3) if energy else None
        self._raise_after(lib().rbplugin_get_state(self._h, *(None if a is None else _dp(a) for a in (x, v, f, e))))
        out = {}
        if positions:
            out["positions"] = x
        if velocities:
            out["velocities"] = v
        if forces:
            out["forces"] = f
        if energy:
            out.update(kinetic=e[0], potential=e[1], time=e[2])
        return out

    def energies(self):
        """{kinetic, refined_kinetic, potential_refinement} from the integrator."""
        o = np.empty(5)
        _check(lib().rbplugin_energies(self._h, _dp(o)))
        return {"kinetic": o[0:2].copy(), "refined_kinetic": o[2:4].copy(), "potential_refinement": o[4]}

    def counts(self):
        o = np.empty(5, dtype=np.int32)
        _check(lib().rbplugin_counts(self._h, _ip(o)))
        return dict(zip(("numFree", "numBodies", "numActualAtoms", "numBodyAtoms", "numDOF"), (int(v) for v in o)))

    def atom_index(self):
        o = np.empty(self.counts()["numActualAtoms"], dtype=np.int32)
        _check(lib().rbplugin_atom_index(self._h, _ip(o)))
        return o

    def rigid_body(self, i):
        o = np.empty(33)
        _check(lib().rbplugin_rigid_body(self._h, int(i), _dp(o)))
        return {k: (o[s].copy() if isinstance(s, slice) else o[s]) for k, s in BODY_FIELDS.items()}
