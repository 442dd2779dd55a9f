"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py
as the checker; the product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

OR_AGENT = np.dtype([("availableEggs", "<u4"), ("cycleLength", "<u4"), ("humanBloodmealCount", "<u4"),
                     ("delay", "<f4"), ("numEggBatches", "<u4"), ("aux", "<u4")])
OR_AGE = np.dtype([("ageInHours", "<f4"), ("hoursInState", "<f4"), ("isFemale", "<u4"), ("csd", "<f4")])
OR_STATE = np.dtype([("dead", "<u4"), ("state", "<u4")])
ROW = np.dtype([("count", "<u4", (8,)), ("l1de", "<u4"), ("females", "<u4"), ("potInfective", "<u4"),
                ("numCycles", "<u4"), ("sumCycles", "<u4"), ("bites", "<u4"), ("infectBites", "<u4")])
# host record shared with the engine's export (include/sampo_engine.h sampo_agent)
AGENT = np.dtype([("ageInHours", "<f4"), ("hoursInState", "<f4"), ("delay", "<f4"), ("csd", "<f4"),
                  ("aux", "<u4"), ("availableEggs", "<u4"), ("cycleLength", "<u4"),
                  ("humanBloodmealCount", "<u4"), ("numEggBatches", "<u4"), ("isFemale", "<u4")])
ENV = np.dtype([("bloodmealSuccess", "<f4"), ("ITNValue", "<f4"), ("IRSValue", "<f4"),
                ("carryingCapacity", "<f4"), ("larvacideValue", "<f4"), ("oviTrapValue", "<f4")])


# or_species (same layout as the product's species table)
SPECIES = np.dtype([
    ("egg_incubation", "<f4", (2,)), ("hatch_threshold", "<f4", (4,)), ("hatch_slope", "<f4", (5,)),
    ("hatch_offset", "<f4", (5,)), ("larva_delay", "<f4", (2,)), ("dev_temp_range", "<f4", (2,)),
    ("dev_rates", "<f4", (32,)), ("pupa_delay", "<f8", (2,)), ("immature_delay", "<f4", (2,)),
    ("mate_seeking_delay", "<f4"), ("bms_delay", "<f4"), ("bmd_delay", "<f4", (2,)), ("egg_batch", "<f4", (3,)),
    ("lay_probability", "<f4"), ("bite_human", "<f4"), ("bite_indoor", "<f4"), ("rest_indoor", "<f4"),
    ("window", "<u4", (8,)), ("reserved", "<u4")])


class Config(C.Structure):
    _fields_ = [("mode", C.c_int), ("scan_path", C.c_int), ("threads", C.c_int),
                ("capacity", C.c_uint32), ("steps", C.c_uint32), ("first_step", C.c_uint32),
                ("dmr", C.c_float * 4), ("temp", C.c_void_p), ("env", C.c_void_p), ("species", C.c_void_p)]


class Pop(C.Structure):
    _fields_ = [("r", C.c_uint32 * 16), ("numLarvae1DayEquiv", C.c_uint32), ("numFemale", C.c_uint32),
                ("numPotentiallyInfective", C.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from oracle import build as _b
            _b.build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.or_sim_create.restype = vp
        L.or_sim_destroy.argtypes = [vp]
        L.or_sim_init_reference.argtypes = [vp, C.POINTER(Config), C.c_uint32, C.c_uint32]
        L.or_sim_init_population.argtypes = [vp, C.POINTER(Config), C.c_uint32, C.POINTER(Pop), vp, vp, vp]
        L.or_sim_run.argtypes = [vp, C.c_uint32, vp]
        L.or_sim_population.argtypes = [vp, C.POINTER(Pop)]
        L.or_sim_population.restype = None
        L.or_sim_arrays.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
        L.or_sim_arrays.restype = None
        L.or_sim_histogram.argtypes = [vp, vp]
        L.or_sim_histogram.restype = None
        L.or_sim_step.argtypes = [vp]
        L.or_sim_step.restype = C.c_uint32
        L.or_sim_status.argtypes = [vp]
        L.or_seed_stream.argtypes = [C.c_uint32, C.c_uint32, vp]
        L.or_seed_stream.restype = None
        L.or_bolt_scan.argtypes = [vp, vp, C.c_uint32, C.c_uint32]
        L.or_bolt_scan.restype = None
        for f in ("or_expf", "or_logf", "or_sinf", "or_sqrtf", "or_eggs_hatching", "or_larva_development",
                  "or_pupa_delay"):
            getattr(L, f).argtypes = [C.c_float]
            getattr(L, f).restype = C.c_float
        L.or_powf.argtypes = [C.c_float, C.c_float]
        L.or_powf.restype = C.c_float
        L.or_f2u.argtypes = [C.c_float]
        L.or_f2u.restype = C.c_uint32
        L.or_box_muller.argtypes = [C.c_float] * 4
        L.or_box_muller.restype = C.c_float
        L.or_hybrid_taus.argtypes = [vp, C.c_uint32]
        L.or_hybrid_taus.restype = C.c_float
        L.or_hmr_reference.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_float]
        L.or_hmr_reference.restype = C.c_float
        L.or_species_default.argtypes = [vp]
        L.or_species_default.restype = None
        L.or_sp_eggs_hatching.argtypes = [vp, C.c_float]
        L.or_sp_eggs_hatching.restype = C.c_float
        L.or_sp_larva_development.argtypes = [vp, C.c_float]
        L.or_sp_larva_development.restype = C.c_float
        L.or_sp_egg_delay.argtypes = [vp, C.c_float, C.c_float]
        L.or_sp_egg_delay.restype = C.c_float
        _lib = L
    return _lib


def hybrid_taus(seeds4, gid):
    s = np.ascontiguousarray(seeds4, dtype=np.uint32)
    return lib().or_hybrid_taus(s.ctypes.data, gid)


def seed_stream(seed, n):
    out = np.empty(n, dtype=np.uint32)
    lib().or_seed_stream(seed, n, out.ctypes.data)
    return out


def default_species():
    """include/properties.h as the oracle restates it (SPECIES record)."""
    sp = np.zeros(1, SPECIES)
    lib().or_species_default(sp.ctypes.data)
    return sp[0]


def as_species(sp):
    """A one-element SPECIES array holding the bytes of any same-layout record."""
    return np.frombuffer(np.asarray(sp).tobytes(), dtype=SPECIES).copy()


def bolt_scan(keys, vals, wx=128):
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    v = np.array(vals, dtype=np.int32, copy=True)
    lib().or_bolt_scan(k.ctypes.data, v.ctypes.data, len(k), wx)
    return v


class Sim:
    """One oracle run (reference restatement, array-of-structs layout)."""

    def __init__(self, *, temp, env, capacity, seed=1, n_eggs=None, layout=None, agents=None,
                 mode=0, dmr=(0, 0, 0, 0), scan_path=0, threads=1, species=None):
        self.temp = np.ascontiguousarray(temp, dtype=np.float32)
        env = np.ascontiguousarray(env, dtype=ENV)
        if env.shape[0] == 1:
            env = np.repeat(env, self.temp.shape[0])
        self.env = env
        cfg = Config()
        cfg.mode, cfg.scan_path, cfg.threads = mode, scan_path, threads
        cfg.capacity, cfg.steps, cfg.first_step = capacity, self.temp.shape[0], 0
        cfg.dmr = (C.c_float * 4)(*dmr)
        cfg.temp = self.temp.ctypes.data
        cfg.env = self.env.ctypes.data
        if species is not None:
            self.species = as_species(species)
            cfg.species = self.species.ctypes.data
        self.cfg = cfg
        self.h = lib().or_sim_create()
        if layout is None:
            rc = lib().or_sim_init_reference(self.h, C.byref(cfg), seed, n_eggs)
        else:
            pop = Pop()
            span = int(layout[7][1])
            ag = np.zeros(span, dtype=OR_AGENT)
            ge = np.zeros(span, dtype=OR_AGE)
            st = np.zeros(span, dtype=OR_STATE)
            k = 0
            for s in range(8):
                a, b = int(layout[s][0]), int(layout[s][1])
                pop.r[2 * s], pop.r[2 * s + 1] = a, b
                d = agents[k:k + (b - a)]
                k += b - a
                for f in ("availableEggs", "cycleLength", "humanBloodmealCount", "delay", "numEggBatches",
                          "aux"):
                    ag[f][a:b] = d[f]
                for f in ("ageInHours", "hoursInState", "isFemale", "csd"):
                    ge[f][a:b] = d[f]
                st["state"][a:b] = 1 << s
            rc = lib().or_sim_init_population(self.h, C.byref(cfg), seed, C.byref(pop), ag.ctypes.data,
                                              ge.ctypes.data, st.ctypes.data)
        if rc != 0:
            raise RuntimeError("oracle init failed")

    def close(self):
        if self.h:
            lib().or_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, n):
        rows = np.zeros(n, dtype=ROW)
        before = self.step
        lib().or_sim_run(self.h, n, rows.ctypes.data)
        return rows[:self.step - before]

    @property
    def step(self):
        return lib().or_sim_step(self.h)

    @property
    def status(self):
        return lib().or_sim_status(self.h)

    def layout(self):
        p = Pop()
        lib().or_sim_population(self.h, C.byref(p))
        return np.array([[p.r[2 * s], p.r[2 * s + 1]] for s in range(8)], dtype=np.uint32)

    def export(self):
        """(layout, dense agents in reference index order, states of those agents)."""
        lay = self.layout()
        span = int(lay[7][1])
        a, g, s = C.c_void_p(), C.c_void_p(), C.c_void_p()
        lib().or_sim_arrays(self.h, C.byref(a), C.byref(g), C.byref(s))
        ag = np.frombuffer((C.c_char * (span * OR_AGENT.itemsize)).from_address(a.value), dtype=OR_AGENT)
        ge = np.frombuffer((C.c_char * (span * OR_AGE.itemsize)).from_address(g.value), dtype=OR_AGE)
        st = np.frombuffer((C.c_char * (span * OR_STATE.itemsize)).from_address(s.value), dtype=OR_STATE)
        idx = np.concatenate([np.arange(lay[r][0], lay[r][1], dtype=np.int64) for r in range(8)])
        out = np.zeros(idx.shape[0], dtype=AGENT)
        for f in ("availableEggs", "cycleLength", "humanBloodmealCount", "delay", "numEggBatches", "aux"):
            out[f] = ag[f][idx]
        for f in ("ageInHours", "hoursInState", "isFemale", "csd"):
            out[f] = ge[f][idx]
        return lay, out, st["state"][idx].copy()

    def histogram(self):
        h = np.zeros(100, dtype=np.uint32)
        lib().or_sim_histogram(self.h, h.ctypes.data)
        return h
