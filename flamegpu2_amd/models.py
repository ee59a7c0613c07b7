"""Python driver for model libraries (include/fgx_model.h) via ctypes.

    sim = Simulation("boids_spatial3D", params, seed=7)
    sim.set_agents("Boid", [x, y, z, fx, fy, fz])
    sim.step(10)
    arrays, ids = sim.get_agents("Boid")

The model runs entirely in the library (C++ runtime over libfgx.so); this
module only moves host arrays across the C ABI.
"""
import ctypes

import numpy as np

from . import load
from .fgx import FgxError

_DT = {0: np.float32, 1: np.float64, 2: np.int32, 3: np.uint32, 4: np.int64, 5: np.uint64,
       6: np.int8, 7: np.uint8, 8: np.int16, 9: np.uint16}
_libs = {}


def _lib(model):
    if model in _libs:
        return _libs[model]
    L = load(f"libfgxm_{model}.so")
    c = ctypes
    vp, u32p = c.c_void_p, c.POINTER(c.c_uint32)
    sigs = {
        "fgxm_create": [c.c_char_p, c.POINTER(c.c_float), c.c_uint32, c.c_uint64, c.c_int, c.POINTER(vp)],
        "fgxm_destroy": [vp],
        "fgxm_var_count": [vp, c.c_char_p, u32p],
        "fgxm_var_info": [vp, c.c_char_p, c.c_uint32, c.POINTER(c.c_char_p), c.POINTER(c.c_int32), u32p, u32p],
        "fgxm_set_agents": [vp, c.c_char_p, c.c_char_p, c.c_uint32, c.POINTER(vp), u32p],
        "fgxm_count": [vp, c.c_char_p, c.c_char_p, u32p],
        "fgxm_get_agents": [vp, c.c_char_p, c.c_char_p, c.POINTER(vp), u32p],
        "fgxm_step": [vp, c.c_uint32],
        "fgxm_set_option": [vp, c.c_char_p, c.c_int],
        "fgxm_sync": [vp],
        "fgxm_get_pbm": [vp, c.c_char_p, u32p, c.c_uint32, u32p],
        "fgxm_kernel_times": [vp, c.c_char_p, c.c_uint32, c.c_int],
        "fgxm_timed_steps": [vp, c.c_uint32, c.POINTER(c.c_double)],
        "fgxm_step_counter": [vp, u32p],
        "fgxm_stat": [vp, c.c_char_p, c.POINTER(c.c_uint64)],
        "fgxm_ensemble_simulate": [c.c_char_p, c.POINTER(c.c_float), c.c_uint32, c.c_uint32, c.POINTER(c.c_uint64),
                                   u32p, c.c_char_p, c.c_char_p, c.c_char_p, c.POINTER(c.c_int32), u32p,
                                   c.POINTER(c.c_double), c.POINTER(c.c_int32)],
        "fgxm_simulate": [vp, c.c_uint32],
        "fgxm_log_config": [vp, c.c_int, c.c_char_p, c.c_char_p, c.c_char_p, c.c_char_p, c.c_uint32],
        "fgxm_export_log": [vp, c.c_char_p, c.c_int, c.c_int, c.c_int, c.c_int],
        "fgxm_export_state": [vp, c.c_char_p, c.c_int],
        "fgxm_load_state": [vp, c.c_char_p],
        "fgxm_state_summary": [c.c_char_p, c.POINTER(c.c_float), c.c_uint32, c.c_char_p, c.c_char_p, c.c_uint32],
        "fgxm_batch_create": [c.c_char_p, c.POINTER(c.c_float), c.c_uint32, c.c_int, c.POINTER(vp)],
        "fgxm_batch_destroy": [vp],
        "fgxm_batch_var_count": [vp, c.c_char_p, u32p],
        "fgxm_batch_var_info": [vp, c.c_char_p, c.c_uint32, c.POINTER(c.c_char_p), c.POINTER(c.c_int32), u32p, u32p],
        "fgxm_batch_add": [vp, c.POINTER(c.c_float), c.c_uint32, c.c_uint64, c.c_uint32, c.POINTER(c.c_int32)],
        "fgxm_batch_set_agents": [vp, c.c_int32, c.c_char_p, c.c_uint32, c.POINTER(vp), u32p],
        "fgxm_batch_estimate_bytes": [vp, c.c_int32, c.POINTER(c.c_uint64)],
        "fgxm_batch_step": [vp, c.c_uint32],
        "fgxm_batch_finished": [vp, c.POINTER(c.c_int32), c.c_uint32, u32p],
        "fgxm_batch_steps_done": [vp, c.c_int32, u32p],
        "fgxm_batch_count": [vp, c.c_int32, c.c_char_p, u32p],
        "fgxm_batch_get_agents": [vp, c.c_int32, c.c_char_p, c.POINTER(vp), u32p],
        "fgxm_batch_remove": [vp, c.c_int32],
        "fgxm_batch_set_tuning": [vp, c.c_uint32, c.c_uint32, c.c_uint32],
        "fgxm_batch_set_option": [vp, c.c_char_p, c.c_int],
        "fgxm_batch_info": [vp, u32p, u32p, c.POINTER(c.c_uint64), c.POINTER(c.c_uint64)],
        "fgxm_batch_sync": [vp],
        "fgxm_batch_timed_steps": [vp, c.c_uint32, c.POINTER(c.c_double)],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = c.c_int
    L.fgxm_last_error.restype = c.c_char_p
    L.fgxm_last_error.argtypes = []
    _libs[model] = L
    return L


def _b(s):
    return None if s is None else s.encode()


class Simulation:
    def __init__(self, model, params=(), seed=0, variant=None, device=0):
        self.L = _lib(model)
        self.model = model
        p = np.ascontiguousarray(params, dtype=np.float32)
        h = ctypes.c_void_p()
        self._chk(self.L.fgxm_create(_b(variant), p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(p), seed,
                                     device, ctypes.byref(h)))
        self.h = h

    def _chk(self, rc):
        if rc != 0:
            raise FgxError(rc, self.L.fgxm_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.fgxm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def var_info(self, agent):
        n = ctypes.c_uint32()
        self._chk(self.L.fgxm_var_count(self.h, _b(agent), ctypes.byref(n)))
        out = []
        for v in range(n.value):
            name = ctypes.c_char_p()
            t, e, b = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_uint32()
            self._chk(self.L.fgxm_var_info(self.h, _b(agent), v, ctypes.byref(name), ctypes.byref(t), ctypes.byref(e),
                                           ctypes.byref(b)))
            out.append((name.value.decode(), _DT.get(t.value, np.uint8), e.value, b.value))
        return out

    def set_agents(self, agent, arrays, ids=None, state=None, n=None):
        """n: the population size, needed only for an agent without variables."""
        info = self.var_info(agent)
        assert len(arrays) == len(info), f"{agent} has {len(info)} variables"
        n = len(arrays[0]) if arrays else (len(ids) if ids is not None else n)
        assert n is not None, "set_agents: give n for an agent without variables"
        keep = []
        for a, (name, dt, e, b) in zip(arrays, info):
            a = np.ascontiguousarray(a, dtype=dt)
            assert a.nbytes == n * b, f"variable {name}: expected {n * b} bytes"
            keep.append(a)
        vp = (ctypes.c_void_p * max(len(keep), 1))(*[a.ctypes.data for a in keep])
        idp = None
        if ids is not None:
            ida = np.ascontiguousarray(ids, dtype=np.uint32)
            keep.append(ida)
            idp = ida.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self._chk(self.L.fgxm_set_agents(self.h, _b(agent), _b(state), n, vp, idp))

    def count(self, agent, state=None):
        n = ctypes.c_uint32()
        self._chk(self.L.fgxm_count(self.h, _b(agent), _b(state), ctypes.byref(n)))
        return n.value

    def get_agents(self, agent, state=None):
        info = self.var_info(agent)
        n = self.count(agent, state)
        outs = [np.zeros(n * e, dtype=dt) for (_, dt, e, _) in info]
        ids = np.zeros(n, dtype=np.uint32)
        vp = (ctypes.c_void_p * max(len(outs), 1))(*[o.ctypes.data for o in outs])
        self._chk(self.L.fgxm_get_agents(self.h, _b(agent), _b(state), vp,
                                         ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return outs, ids

    def step(self, n=1):
        self._chk(self.L.fgxm_step(self.h, n))

    def set_option(self, key, value):
        self._chk(self.L.fgxm_set_option(self.h, _b(key), int(value)))

    def sync(self):
        self._chk(self.L.fgxm_sync(self.h))

    def pbm(self, msg, max_bins=1 << 26):
        bins = ctypes.c_uint32()
        # first call with capacity 0 only reports the bin count
        rc = self.L.fgxm_get_pbm(self.h, _b(msg), None, 0, ctypes.byref(bins))
        out = np.zeros(bins.value + 1, dtype=np.uint32)
        self._chk(self.L.fgxm_get_pbm(self.h, _b(msg), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                      len(out), ctypes.byref(bins)))
        del rc
        return out

    def kernel_times(self, reset=False):
        buf = ctypes.create_string_buffer(1 << 16)
        self._chk(self.L.fgxm_kernel_times(self.h, buf, len(buf), 1 if reset else 0))
        out = {}
        for line in buf.value.decode().splitlines():
            label, ms, cnt = line.rsplit(" ", 2)
            out[label] = (float(ms), int(cnt))
        return out

    def timed_steps(self, n):
        ms = ctypes.c_double()
        self._chk(self.L.fgxm_timed_steps(self.h, n, ctypes.byref(ms)))
        return ms.value

    def stat(self, key):
        n = ctypes.c_uint64()
        self._chk(self.L.fgxm_stat(self.h, _b(key), ctypes.byref(n)))
        return n.value

    def step_counter(self):
        n = ctypes.c_uint32()
        self._chk(self.L.fgxm_step_counter(self.h, ctypes.byref(n)))
        return n.value

    def simulate(self, steps):
        """CUDASimulation::simulate(): init functions, step log frame 0, `steps` steps,
        exit functions, exit log."""
        self._chk(self.L.fgxm_simulate(self.h, steps))

    def log(self, what, agent=None, name=None, state=None, value=0, exit=False):
        """Adds one item to the step log (exit=False) or the exit log: what is "count",
        "mean", "standard_deviation", "min", "max", "sum" (agent, name = variable),
        "env" (name = property), "timing" or "frequency" (value)."""
        self._chk(self.L.fgxm_log_config(self.h, 1 if exit else 0, _b(what), _b(agent), _b(state), _b(name), int(value)))

    def export_state(self, path, pretty=True):
        """CUDASimulation::exportData to an .xml or .json state file."""
        self._chk(self.L.fgxm_export_state(self.h, _b(str(path)), int(pretty)))

    def load_state(self, path):
        """Loads an .xml or .json state file (SimulationConfig().input_file + applyConfig)."""
        self._chk(self.L.fgxm_load_state(self.h, _b(str(path))))

    def export_log(self, path, steps=True, exit=True, step_time=False, exit_time=False):
        self._chk(self.L.fgxm_export_log(self.h, _b(str(path)), int(steps), int(exit), int(step_time), int(exit_time)))


def ensemble(model, params, seeds, steps, variant=None, options="", sum_agent=None, sum_var=None, nagents=1):
    """CUDAEnsemble::simulate over len(seeds) runs (fgxm_ensemble_simulate).  Returns a dict:
    status per run (0 done, 1 failed, 2 not started), final counts (runs x agents), the
    exit-log sums (sum_agent.sum_var, or the environment property sum_var when sum_agent is
    None), whether the runs ran as one batch, and the error message ("" when none)."""
    L = _lib(model)
    n = len(seeds)
    p = np.ascontiguousarray(params, dtype=np.float32)
    sd = np.ascontiguousarray(seeds, dtype=np.uint64)
    st = np.ascontiguousarray(steps, dtype=np.uint32)
    status = np.zeros(n, np.int32)
    counts = np.zeros((n, nagents), np.uint32)
    sums = np.zeros(n, np.float64)
    batched = ctypes.c_int32(-1)
    rc = L.fgxm_ensemble_simulate(_b(variant), p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(p), n,
                                  sd.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                  st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _b(options), _b(sum_agent),
                                  _b(sum_var), status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                  counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                  sums.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(batched))
    err = "" if rc == 0 else L.fgxm_last_error().decode()
    return {"status": status, "counts": counts, "sums": sums, "batched": batched.value, "error": err}


def state_summary(model, path, params=(), variant=None):
    """Parses a state file against a model on the host only (no GPU): step count,
    environment values and per (agent, state) counts and per-variable sums."""
    import json
    L = _lib(model)
    p = (ctypes.c_float * max(len(params), 1))(*params)
    buf = ctypes.create_string_buffer(1 << 20)
    rc = L.fgxm_state_summary(_b(variant), p, len(params), _b(str(path)), buf, len(buf))
    if rc:
        raise FgxError(rc, L.fgxm_last_error().decode())
    return json.loads(buf.value.decode())


class Batch:
    """Many simulations of one model library co-scheduled in shared launches
    (include/fgx_model.h, fgxm_batch_*).  `params` of each simulation set its
    environment and spatial message geometry, as for Simulation."""

    def __init__(self, model, params=(), variant=None, device=0):
        self.L = _lib(model)
        self.model = model
        p = np.ascontiguousarray(params, dtype=np.float32)
        h = ctypes.c_void_p()
        self._chk(self.L.fgxm_batch_create(_b(variant), p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(p), device,
                                           ctypes.byref(h)))
        self.h = h
        self._info = {}

    def _chk(self, rc):
        if rc != 0:
            raise FgxError(rc, self.L.fgxm_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.fgxm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def var_info(self, agent):
        if agent in self._info:
            return self._info[agent]
        n = ctypes.c_uint32()
        self._chk(self.L.fgxm_batch_var_count(self.h, _b(agent), ctypes.byref(n)))
        out = []
        for v in range(n.value):
            name = ctypes.c_char_p()
            t, e, b = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_uint32()
            self._chk(self.L.fgxm_batch_var_info(self.h, _b(agent), v, ctypes.byref(name), ctypes.byref(t),
                                                 ctypes.byref(e), ctypes.byref(b)))
            out.append((name.value.decode(), _DT.get(t.value, np.uint8), e.value, b.value))
        self._info[agent] = out
        return out

    def add(self, params, seed, steps, populations):
        """populations: {agent: [arrays in declaration order]}; returns the handle."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        h = ctypes.c_int32()
        self._chk(self.L.fgxm_batch_add(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(p), int(seed),
                                        int(steps), ctypes.byref(h)))
        for agent, arrays in populations.items():
            info = self.var_info(agent)
            assert len(arrays) == len(info), f"{agent} has {len(info)} variables"
            n = len(arrays[0])
            keep = [np.ascontiguousarray(a, dtype=dt) for a, (_, dt, _, _) in zip(arrays, info)]
            vp = (ctypes.c_void_p * max(len(keep), 1))(*[a.ctypes.data for a in keep])
            self._chk(self.L.fgxm_batch_set_agents(self.h, h.value, _b(agent), n, vp, None))
        return h.value

    def estimate_bytes(self, handle):
        b = ctypes.c_uint64()
        self._chk(self.L.fgxm_batch_estimate_bytes(self.h, handle, ctypes.byref(b)))
        return b.value

    def step(self, n=1):
        self._chk(self.L.fgxm_batch_step(self.h, n))

    def finished(self):
        cnt = ctypes.c_uint32()
        self._chk(self.L.fgxm_batch_finished(self.h, None, 0, ctypes.byref(cnt)))
        buf = (ctypes.c_int32 * max(cnt.value, 1))()
        self._chk(self.L.fgxm_batch_finished(self.h, buf, cnt.value, ctypes.byref(cnt)))
        return [buf[i] for i in range(cnt.value)]

    def steps_done(self, handle):
        n = ctypes.c_uint32()
        self._chk(self.L.fgxm_batch_steps_done(self.h, handle, ctypes.byref(n)))
        return n.value

    def get_agents(self, handle, agent):
        info = self.var_info(agent)
        n = ctypes.c_uint32()
        self._chk(self.L.fgxm_batch_count(self.h, handle, _b(agent), ctypes.byref(n)))
        outs = [np.zeros(n.value * e, dtype=dt) for (_, dt, e, _) in info]
        ids = np.zeros(n.value, dtype=np.uint32)
        vp = (ctypes.c_void_p * max(len(outs), 1))(*[o.ctypes.data for o in outs])
        self._chk(self.L.fgxm_batch_get_agents(self.h, handle, _b(agent), vp,
                                               ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return outs, ids

    def remove(self, handle):
        self._chk(self.L.fgxm_batch_remove(self.h, handle))

    def set_tuning(self, block=256, agents_per_thread=1, blocks_per_cu=0):
        self._chk(self.L.fgxm_batch_set_tuning(self.h, block, agents_per_thread, blocks_per_cu))

    def set_option(self, key, value):
        self._chk(self.L.fgxm_batch_set_option(self.h, _b(key), int(value)))

    def info(self):
        r, s = ctypes.c_uint32(), ctypes.c_uint32()
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        self._chk(self.L.fgxm_batch_info(self.h, ctypes.byref(r), ctypes.byref(s), ctypes.byref(a), ctypes.byref(b)))
        return {"resident": r.value, "staged": s.value, "resident_agents": a.value, "resident_bytes": b.value}

    def sync(self):
        self._chk(self.L.fgxm_batch_sync(self.h))

    def timed_steps(self, n):
        ms = ctypes.c_double()
        self._chk(self.L.fgxm_batch_timed_steps(self.h, n, ctypes.byref(ms)))
        return ms.value
