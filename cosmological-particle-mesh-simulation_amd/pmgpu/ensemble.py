"""Batched ensemble of device-resident simulations (one kernel launch per stage for all).

``PMEnsemble`` groups FP32 ``PMSim`` objects, possibly of different grid sizes, and advances
them together: every stage of the PM step is one launch covering the blocks of all members,
and consecutive steps replay one captured graph.  Each member keeps its own scale factor and
step; the reference's loop (pmesh.py:56-63) is applied per member.  Members remain ordinary
sims between calls (state in and out, stepping alone).
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from .sim import PMSim

MAX_MEMBERS = 64
BATCH_SIZES = (32, 48, 64, 96, 128, 192, 256, 384, 512)


class PMEnsemble:
    def __init__(self, sims=()):
        _lib.init()
        h = C.c_void_p()
        _lib.call("pm_ensemble_create", C.byref(h))
        self._h = h
        self.members: list[PMSim] = []
        for s in sims:
            self.add(s)

    def add(self, sim: PMSim):
        _lib.call("pm_ensemble_add", self._h, sim._h)
        self.members.append(sim)

    def remove(self, sim: PMSim):
        _lib.call("pm_ensemble_remove", self._h, sim._h)
        self.members.remove(sim)

    def __len__(self):
        n = C.c_int()
        _lib.call("pm_ensemble_size", self._h, C.byref(n))
        return n.value

    def run(self, a, da, n_steps: int) -> list[float]:
        """n_steps steps of every member; a[i], da[i] per member (in the order added).

        Returns the members' scale factors after the steps."""
        m = len(self.members)
        if len(a) != m or len(da) != m:
            raise ValueError(f"a and da need one value per member ({m})")
        av = (C.c_double * m)(*[float(x) for x in a])
        dv = (C.c_double * m)(*[float(x) for x in da])
        try:
            _lib.call("pm_ensemble_run", self._h, av, dv, int(n_steps))
        except RuntimeError as exc:
            # a member's state became corrupt: the steps done before the check are kept
            for s in self.members:
                s.version += 1
            exc.a_after = list(av)
            raise
        for s in self.members:
            s.version += 1
        return list(av)

    def sync(self):
        _lib.call("pm_ensemble_sync", self._h)

    def set_timing(self, enable: bool):
        _lib.call("pm_ensemble_set_timing", self._h, 1 if enable else 0)

    def timing(self) -> dict:
        k = len(_lib.STAGES)
        ms = (C.c_double * k)()
        calls = (C.c_int64 * k)()
        _lib.call("pm_ensemble_get_timing", self._h, ms, calls, k)
        return {name: (ms[i], calls[i]) for i, name in enumerate(_lib.STAGES)}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().pm_ensemble_destroy(self._h)
            self._h = C.c_void_p()
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
