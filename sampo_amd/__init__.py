"""sampo_amd -- MI355X-native agent-step engine for SAMPO's mosquito model.

Python host binding over the C ABI of libsampo_engine.so (include/
sampo_engine.h).  The compute runs only in the HIP library: importing this
package never falls back to a CPU implementation, and every call fails
loudly when the library or a GPU is missing.

The engine advances scenario instances by reference iterations (one
simulated hour, src/abms.c:15-16); a day is 24 iterations.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ENGINE_LIB = os.path.join(HERE, "libsampo_engine.so")

RUNNING, DONE, CAPACITY, EXTINCT, PHYS_FULL, GROW, HELD, FAULT, EMPTY = 0, 1, 2, 3, 4, 5, 6, 7, 0xFF
MODE_REFERENCE, MODE_CONFIG = 0, 1
STATES = ("egg", "larva", "pupa", "immature", "mateseeking", "bms", "bmd", "gravid")


class SampoError(RuntimeError):
    pass


class Env(C.Structure):
    """struct Environment (include/agent.h:22-35)."""
    _fields_ = [("bloodmealSuccess", C.c_float), ("ITNValue", C.c_float), ("IRSValue", C.c_float),
                ("carryingCapacity", C.c_float), ("larvacideValue", C.c_float),
                ("oviTrapValue", C.c_float)]


class Agent(C.Structure):
    _fields_ = [("ageInHours", C.c_float), ("hoursInState", C.c_float), ("delay", C.c_float),
                ("csd", C.c_float), ("aux", C.c_uint32), ("availableEggs", C.c_uint32),
                ("cycleLength", C.c_uint32), ("humanBloodmealCount", C.c_uint32),
                ("numEggBatches", C.c_uint32), ("isFemale", C.c_uint32)]


class Layout(C.Structure):
    _fields_ = [("start", C.c_uint32 * 8), ("end", C.c_uint32 * 8)]


class Row(C.Structure):
    _fields_ = [("count", C.c_uint32 * 8), ("l1de", C.c_uint32), ("females", C.c_uint32),
                ("potInfective", C.c_uint32), ("numCycles", C.c_uint32), ("sumCycles", C.c_uint32),
                ("bites", C.c_uint32), ("infectBites", C.c_uint32), ("valid", C.c_uint32)]


class Scenario(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("seed", C.c_uint32), ("horizon", C.c_uint32),
                ("n_eggs", C.c_uint32), ("max_agents", C.c_uint64), ("phys_slots", C.c_uint64),
                ("dmr", C.c_float * 4), ("temp", C.POINTER(C.c_float)), ("env", C.POINTER(Env)),
                ("env_count", C.c_uint32), ("init_layout", C.POINTER(Layout)),
                ("init_agents", C.POINTER(Agent)), ("species", C.c_void_p)]


# sampo_species: the species plug-in (the -DSPECIES header, include/properties.h) as data
SPECIES_DTYPE = np.dtype([
    ("egg_incubation", "<f4", (2,)), ("hatch_threshold", "<f4", (4,)), ("hatch_slope", "<f4", (5,)),
    ("hatch_offset", "<f4", (5,)), ("larva_delay", "<f4", (2,)), ("dev_temp_range", "<f4", (2,)),
    ("dev_rates", "<f4", (32,)), ("pupa_delay", "<f8", (2,)), ("immature_delay", "<f4", (2,)),
    ("mate_seeking_delay", "<f4"), ("bms_delay", "<f4"), ("bmd_delay", "<f4", (2,)), ("egg_batch", "<f4", (3,)),
    ("lay_probability", "<f4"), ("bite_human", "<f4"), ("bite_indoor", "<f4"), ("rest_indoor", "<f4"),
    ("window", "<u4", (8,)), ("reserved", "<u4")])
WINDOWS = ("egg", "larva", "pupa", "immature", "mate_seeking", "bms", "bmd", "lay")


AGENT_DTYPE = np.dtype([("ageInHours", "<f4"), ("hoursInState", "<f4"), ("delay", "<f4"),
                        ("csd", "<f4"), ("aux", "<u4"), ("availableEggs", "<u4"),
                        ("cycleLength", "<u4"), ("humanBloodmealCount", "<u4"),
                        ("numEggBatches", "<u4"), ("isFemale", "<u4")])
ROW_DTYPE = np.dtype([("count", "<u4", (8,)), ("l1de", "<u4"), ("females", "<u4"),
                      ("potInfective", "<u4"), ("numCycles", "<u4"), ("sumCycles", "<u4"),
                      ("bites", "<u4"), ("infectBites", "<u4"), ("valid", "<u4")])
assert AGENT_DTYPE.itemsize == C.sizeof(Agent) and ROW_DTYPE.itemsize == C.sizeof(Row)

_lib = None


def lib():
    """Load libsampo_engine.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(ENGINE_LIB):
            raise SampoError(f"{ENGINE_LIB} is not built; run python -m sampo_amd.build")
        L = C.CDLL(ENGINE_LIB)
        vp = C.c_void_p
        L.sampo_last_error.restype = C.c_char_p
        L.sampo_engine_create.argtypes = [C.c_int, C.c_uint32, C.POINTER(vp)]
        L.sampo_engine_destroy.argtypes = [vp]
        L.sampo_engine_destroy.restype = None
        L.sampo_engine_set_tile.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32]
        L.sampo_engine_set_chunk.argtypes = [vp, C.c_uint32, C.c_uint32]
        L.sampo_engine_set_grid.argtypes = [vp, C.c_uint32]
        L.sampo_engine_admit.argtypes = [vp, C.POINTER(Scenario), C.POINTER(C.c_uint32)]
        L.sampo_engine_release.argtypes = [vp, C.c_uint32]
        L.sampo_engine_hold.argtypes = [vp, C.c_uint32, vp]
        L.sampo_engine_footprint.argtypes = [C.POINTER(Scenario), C.c_uint32]
        L.sampo_engine_footprint.restype = C.c_uint64
        L.sampo_engine_poll.argtypes = [vp, C.c_uint32, vp, vp, vp]
        L.sampo_engine_service.argtypes = [vp, C.c_float, C.POINTER(C.c_uint32)]
        L.sampo_engine_set_day.argtypes = [vp, C.c_uint32]
        L.sampo_engine_step.argtypes = [vp, C.c_uint32]
        L.sampo_engine_step_days.argtypes = [vp, C.c_uint32]
        L.sampo_engine_sync.argtypes = [vp]
        L.sampo_engine_time.argtypes = [vp, C.c_uint32, C.c_int, C.POINTER(C.c_float)]
        L.sampo_engine_status.argtypes = [vp, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_uint64)]
        L.sampo_engine_rows.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, vp]
        L.sampo_engine_export.argtypes = [vp, C.c_uint32, C.POINTER(Layout), vp, C.c_uint64]
        L.sampo_engine_histogram.argtypes = [vp, C.c_uint32, vp]
        L.sampo_engine_agent_steps.argtypes = [vp, C.POINTER(C.c_uint64), C.c_int]
        L.sampo_engine_time_pass.argtypes = [vp, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.sampo_engine_time_days.argtypes = [vp, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.sampo_engine_capture.argtypes = [vp, C.c_uint32]
        L.sampo_engine_replay.argtypes = [vp]
        L.sampo_engine_replay_times.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.sampo_stream_copy.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.POINTER(C.c_float)]
        L.sampo_device_memory.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.sampo_seed_stream.argtypes = [C.c_uint32, C.c_uint32, vp]
        L.sampo_seed_stream.restype = None
        L.sampo_gen_temperature.argtypes = [C.c_uint64, C.c_uint32, C.c_float, C.c_float, C.c_float, vp]
        L.sampo_gen_population.argtypes = [C.c_uint64, C.c_uint32, C.c_float, C.POINTER(Layout), vp]
        L.sampo_compact_workspace_bytes.argtypes = [C.c_uint64]
        L.sampo_compact_workspace_bytes.restype = C.c_uint64
        L.sampo_compact_index.argtypes = [vp, C.c_uint64, vp, vp, vp, C.c_uint64, vp]
        L.sampo_compact_index_host.argtypes = [C.c_int, vp, C.c_uint64, vp, vp]
        L.sampo_compact_time.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint64, C.POINTER(C.c_float)]
        L.sampo_species_default.argtypes = [vp]
        L.sampo_species_default.restype = None
        L.sampo_species_check.argtypes = [vp]
        L.sampo_species_read.argtypes = [C.c_char_p, vp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise SampoError(f"sampo error {rc}: {lib().sampo_last_error().decode()}")


def seed_stream(seed, n):
    """The first n glibc rand() values after srand(seed) (src/abms.c:140-158)."""
    out = np.empty(n, dtype=np.uint32)
    lib().sampo_seed_stream(seed, n, out.ctypes.data)
    return out


def gen_temperature(seed, days, mean=25.0, amp=5.0, noise=1.5):
    out = np.empty(days * 24, dtype=np.float32)
    _check(lib().sampo_gen_temperature(seed, days, mean, amp, noise, out.ctypes.data))
    return out


def gen_population(seed, n0, t0):
    """Configuration-mode initial population: (layout (8,2) array, agents)."""
    lay = Layout()
    agents = np.zeros(n0, dtype=AGENT_DTYPE)
    _check(lib().sampo_gen_population(seed, n0, t0, C.byref(lay), agents.ctypes.data))
    return np.array([[lay.start[s], lay.end[s]] for s in range(8)], dtype=np.uint32), agents


def stream_copy_ms(nbytes, iters=20, device=0):
    """Average ms of a plain device-to-device copy of nbytes (read + write)."""
    ms = C.c_float()
    _check(lib().sampo_stream_copy(device, nbytes, iters, C.byref(ms)))
    return ms.value


def device_memory(device=0):
    """(free, total) bytes of device memory."""
    f, t = C.c_uint64(), C.c_uint64()
    _check(lib().sampo_device_memory(device, C.byref(f), C.byref(t)))
    return f.value, t.value


def make_env(bloodmeal=0.25, itn=0.0, irs=0.0, cc=160000.0, larvicide=0.0, ovitrap=0.0):
    return np.array([(bloodmeal, itn, irs, cc, larvicide, ovitrap)],
                    dtype=[(n, "<f4") for n, _ in Env._fields_])


COMPACT_NONE = 0xFFFFFFFF


def compact_index(keys, device=0):
    """Per-key ranks of the compaction step (sampo_compact_index) on host
    data: dest[i] = number of earlier keys equal to keys[i] for keys 0..7,
    COMPACT_NONE otherwise; counts[k] = number of keys equal to k."""
    keys = np.ascontiguousarray(keys, dtype=np.uint8)
    dest = np.empty(keys.shape[0], np.uint32)
    counts = np.zeros(8, np.uint32)
    _check(lib().sampo_compact_index_host(device, keys.ctypes.data, keys.shape[0], dest.ctypes.data,
                                          counts.ctypes.data))
    return dest, counts


def compact_index_device(keys_ptr, n, dest_ptr, counts_ptr, workspace_ptr, workspace_bytes, stream=None):
    """sampo_compact_index on device pointers (integers), asynchronous."""
    _check(lib().sampo_compact_index(keys_ptr, n, dest_ptr, counts_ptr, workspace_ptr, workspace_bytes, stream))


def compact_workspace_bytes(n):
    return int(lib().sampo_compact_workspace_bytes(n))


def compact_time(n, iters=20, seed=1, device=0):
    """Milliseconds per sampo_compact_index call on n device-generated keys."""
    ms = C.c_float()
    _check(lib().sampo_compact_time(device, n, iters, seed, C.byref(ms)))
    return ms.value


def default_species():
    """include/properties.h as a SPECIES_DTYPE record."""
    sp = np.zeros(1, SPECIES_DTYPE)
    lib().sampo_species_default(sp.ctypes.data)
    return sp[0]


def check_species(sp):
    """True if the library accepts the table (sampo_species_check)."""
    a = np.array([sp], dtype=SPECIES_DTYPE)
    return lib().sampo_species_check(a.ctypes.data) == 0


def read_species(path):
    """A species file (sampo_species_read): properties.h with the lines of the file applied."""
    sp = np.zeros(1, SPECIES_DTYPE)
    _check(lib().sampo_species_read(str(path).encode(), sp.ctypes.data))
    return sp[0]


def _hours(mask):
    hs = [h for h in range(24) if (int(mask) >> h) & 1]
    return " ".join(str(h) for h in hs) if hs else "none"


def write_species(path, sp):
    """Write every member of a species record in the species file format."""
    lines = ["# species table (sampo_species_read format)"]
    for name in SPECIES_DTYPE.names:
        if name == "reserved":
            continue
        if name == "window":
            lines += [f"window_{w} {_hours(sp['window'][i])}" for i, w in enumerate(WINDOWS)]
            continue
        v = np.atleast_1d(sp[name])  # shortest text that reads back to the same binary32 / binary64
        lines.append(name + " " + " ".join(str(x) for x in v))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def scenario(*, mode=MODE_REFERENCE, seed=1, horizon, temp, env, n_eggs=0, max_agents, phys_slots=0,
             dmr=(0.0, 0.0, 0.0, 0.0), layout=None, agents=None, species=None):
    """A sampo_scenario and the arrays it points into (keep them alive).
    species: a SPECIES_DTYPE record, or None for include/properties.h."""
    temp = np.ascontiguousarray(temp, dtype=np.float32)
    env = np.ascontiguousarray(env)
    if temp.shape[0] < horizon:
        raise ValueError("temp shorter than horizon")
    sc = Scenario()
    sc.mode, sc.seed, sc.horizon, sc.n_eggs = mode, seed, horizon, n_eggs
    sc.max_agents, sc.phys_slots = max_agents, phys_slots
    sc.dmr = (C.c_float * 4)(*dmr)
    sc.temp = temp.ctypes.data_as(C.POINTER(C.c_float))
    sc.env = C.cast(env.ctypes.data, C.POINTER(Env))
    sc.env_count = env.shape[0]
    keep = [temp, env]
    if layout is not None:
        lay = Layout()
        for s in range(8):
            lay.start[s], lay.end[s] = int(layout[s][0]), int(layout[s][1])
        agents = np.ascontiguousarray(agents, dtype=AGENT_DTYPE)
        sc.init_layout = C.pointer(lay)
        sc.init_agents = C.cast(agents.ctypes.data, C.POINTER(Agent))
        keep += [lay, agents]
    if species is not None:
        sp = np.array([species], dtype=SPECIES_DTYPE)
        sc.species = sp.ctypes.data
        keep.append(sp)
    return sc, keep


def footprint(chunk=0, **kw):
    """Device bytes admitting this scenario would allocate."""
    sc, keep = scenario(**kw)
    return int(lib().sampo_engine_footprint(C.byref(sc), chunk))


class Engine:
    """A set of scenario instances advanced together on one GPU."""

    def __init__(self, device=0, max_slots=16):
        h = C.c_void_p()
        _check(lib().sampo_engine_create(device, max_slots, C.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if self.h:
            lib().sampo_engine_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile(self, wg, apt, waves=0):
        """Work-group size, agents per thread, register cap (min waves/SIMD)."""
        _check(lib().sampo_engine_set_tile(self.h, wg, apt, waves))

    def set_grid(self, per_cu=0):
        """Persistent work-groups of the pass per CU (0: all that fit)."""
        _check(lib().sampo_engine_set_grid(self.h, per_cu))

    def set_day(self, work_groups):
        """Day launches for later admissions: such an instance is cut into at
        most work_groups equal chunks and, while it is the engine's only
        instance, advances by cooperative launches of up to 96 iterations
        (four days) each.  0: off."""
        _check(lib().sampo_engine_set_day(self.h, work_groups))

    def set_chunk(self, agents, target_chunks=0):
        """Minimum agents per chunk (work unit / output segment) and the most
        chunks per instance (0: two per resident work-group) for later
        admissions."""
        _check(lib().sampo_engine_set_chunk(self.h, agents, target_chunks))

    def admit(self, **kw):
        """Admit a scenario (keywords of scenario()); returns its slot."""
        sc, keep = scenario(**kw)
        slot = C.c_uint32()
        _check(lib().sampo_engine_admit(self.h, C.byref(sc), C.byref(slot)))
        del keep
        return slot.value

    def release(self, slot):
        _check(lib().sampo_engine_release(self.h, slot))

    def hold(self, mask):
        """Hold (mask[i] true) or resume the instances in slots 0..len(mask)-1
        from the next step on; only running instances are held and only
        held ones resumed."""
        m = np.ascontiguousarray(np.asarray(mask, dtype=bool).astype(np.uint8))
        _check(lib().sampo_engine_hold(self.h, m.shape[0], m.ctypes.data))

    def poll(self, n):
        """(status, step, agents) arrays for slots 0..n-1 in one read."""
        st = np.zeros(n, np.uint32)
        step = np.zeros(n, np.uint32)
        na = np.zeros(n, np.uint64)
        _check(lib().sampo_engine_poll(self.h, n, st.ctypes.data, step.ctypes.data, na.ctypes.data))
        return st, step, na

    def service(self, headroom=1.25):
        """Grow instances paused for slots or short of headroom; returns how many."""
        g = C.c_uint32()
        _check(lib().sampo_engine_service(self.h, headroom, C.byref(g)))
        return g.value

    def step(self, n):
        _check(lib().sampo_engine_step(self.h, n))

    def step_days(self, days):
        _check(lib().sampo_engine_step_days(self.h, days))

    def sync(self):
        _check(lib().sampo_engine_sync(self.h))

    def time(self, n, graph=False):
        ms = C.c_float()
        _check(lib().sampo_engine_time(self.h, n, 1 if graph else 0, C.byref(ms)))
        return ms.value

    def capture(self, n):
        """Capture the next n iterations as one HIP graph with events around every pass."""
        _check(lib().sampo_engine_capture(self.h, n))

    def replay(self):
        """Launch the captured iterations (asynchronous)."""
        _check(lib().sampo_engine_replay(self.h))

    def replay_times(self):
        """(first pass start to last pass end ms, summed pass ms) of the replay."""
        w, p = C.c_float(), C.c_float()
        _check(lib().sampo_engine_replay_times(self.h, C.byref(w), C.byref(p)))
        return w.value, p.value

    def time_days(self, days):
        """(first day start to last day end ms, summed per-day launch ms) over
        `days` days, an event pair around each day's launch."""
        w, p = C.c_float(), C.c_float()
        _check(lib().sampo_engine_time_days(self.h, days, C.byref(w), C.byref(p)))
        return w.value, p.value

    def time_pass(self, n):
        """(wall ms, summed fused-pass ms) over n iterations."""
        w, p = C.c_float(), C.c_float()
        _check(lib().sampo_engine_time_pass(self.h, n, C.byref(w), C.byref(p)))
        return w.value, p.value

    def status(self, slot):
        st, step, n = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(lib().sampo_engine_status(self.h, slot, C.byref(st), C.byref(step), C.byref(n)))
        return st.value, step.value, n.value

    def rows(self, slot, first, n):
        out = np.zeros(n, dtype=ROW_DTYPE)
        _check(lib().sampo_engine_rows(self.h, slot, first, n, out.ctypes.data))
        return out

    def export(self, slot):
        """(layout (8,2) uint32 array, dense agents in reference index order)."""
        lay = Layout()
        _check(lib().sampo_engine_export(self.h, slot, C.byref(lay), None, 0))
        layout = np.array([[lay.start[s], lay.end[s]] for s in range(8)], dtype=np.uint32)
        n = int((layout[:, 1] - layout[:, 0]).sum())
        agents = np.zeros(n, dtype=AGENT_DTYPE)
        if n:
            _check(lib().sampo_engine_export(self.h, slot, C.byref(lay), agents.ctypes.data, n))
        return layout, agents

    def histogram(self, slot):
        out = np.zeros(100, dtype=np.uint32)
        _check(lib().sampo_engine_histogram(self.h, slot, out.ctypes.data))
        return out

    def agent_steps(self, reset=False):
        v = C.c_uint64()
        _check(lib().sampo_engine_agent_steps(self.h, C.byref(v), 1 if reset else 0))
        return v.value


# ---------------------------------------------------------------- icl path

ICL_LIB = os.path.join(HERE, "libsampo_icl.so")


class AbmsParams(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("n_eggs", C.c_uint32), ("max_agents", C.c_uint32), ("steps", C.c_uint32),
                ("serial_scan", C.c_uint32), ("temp", C.POINTER(C.c_float)), ("env", C.POINTER(Env)),
                ("env_count", C.c_uint32), ("mode", C.c_uint32), ("dmr", C.c_float * 4),
                ("init_layout", C.POINTER(Layout)), ("init_agents", C.POINTER(Agent)), ("species", C.c_char_p)]


class AbmsResult(C.Structure):
    _fields_ = [("rows", C.c_void_p), ("status", C.c_uint32), ("steps_done", C.c_uint32),
                ("agent_steps", C.c_uint64), ("layout", Layout), ("histogram", C.c_uint32 * 100),
                ("loop_ms", C.c_double), ("total_ms", C.c_double), ("agents", C.c_void_p), ("ages", C.c_void_p),
                ("states", C.c_void_p), ("span_capacity", C.c_uint64)]


ICL_AGENT = np.dtype([("availableEggs", "<u4"), ("cycleLength", "<u4"), ("humanBloodmealCount", "<u4"),
                      ("delay", "<f4"), ("numEggBatches", "<u4"), ("aux", "<u4")])
ICL_AGE = np.dtype([("ageInHours", "<f4"), ("hoursInState", "<f4"), ("isFemale", "<u4"), ("csd", "<f4")])
ICL_STATE = np.dtype([("dead", "<u4"), ("state", "<u4")])

_icl = None


def icl_lib():
    """Load libsampo_icl.so (the icl_* boundary and the reference driver loop)."""
    global _icl
    if _icl is None:
        if not os.path.exists(ICL_LIB):
            raise SampoError(f"{ICL_LIB} is not built; run python -m sampo_amd.build")
        L = C.CDLL(ICL_LIB)
        L.sampo_abms_run.argtypes = [C.POINTER(AbmsParams), C.POINTER(AbmsResult)]
        L.sampo_format_stats.argtypes = [C.c_char_p, C.c_size_t, C.c_uint32, C.c_float, C.c_float, C.POINTER(Row)]
        L.sampo_format_population.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(Layout)]
        L.sampo_format_histogram.argtypes = [C.c_char_p, C.c_size_t, C.c_void_p]
        L.sampo_read_single_val_file.argtypes = [C.c_char_p, C.c_uint32, C.c_void_p]
        L.sampo_read_environment.argtypes = [C.c_char_p, C.c_uint32, C.c_float, C.c_void_p]
        _icl = L
    return _icl


def run_abms(*, seed, n_eggs=0, max_agents, steps, temp, env, serial_scan=False, export=True, mode=MODE_REFERENCE,
             dmr=(0.0, 0.0, 0.0, 0.0), layout=None, agents=None, species=None):
    """Run the reference driver loop on the GPU through the icl boundary.
    species: the speciesHeader argument (None: properties.h, else a species file).

    Returns a dict with rows, status, steps_done, agent_steps, layout,
    histogram, loop_ms and (export=True) the final AoS agent arrays."""
    temp = np.ascontiguousarray(temp, dtype=np.float32)
    env = np.ascontiguousarray(env)
    p = AbmsParams(seed, n_eggs, max_agents, steps, 1 if serial_scan else 0,
                   temp.ctypes.data_as(C.POINTER(C.c_float)), C.cast(env.ctypes.data, C.POINTER(Env)), env.shape[0])
    p.mode = mode
    p.dmr = (C.c_float * 4)(*dmr)
    if species is not None:
        p.species = str(species).encode()
    keep = None
    if layout is not None:
        lay = Layout()
        for s in range(8):
            lay.start[s], lay.end[s] = int(layout[s][0]), int(layout[s][1])
        agents = np.ascontiguousarray(agents, dtype=AGENT_DTYPE)
        p.init_layout = C.pointer(lay)
        p.init_agents = C.cast(agents.ctypes.data, C.POINTER(Agent))
        keep = (lay, agents)
    rows = np.zeros(steps, dtype=ROW_DTYPE)
    r = AbmsResult()
    r.rows = rows.ctypes.data
    if export:
        ag = np.zeros(max_agents, dtype=ICL_AGENT)
        ge = np.zeros(max_agents, dtype=ICL_AGE)
        st = np.zeros(max_agents, dtype=ICL_STATE)
        r.agents, r.ages, r.states, r.span_capacity = ag.ctypes.data, ge.ctypes.data, st.ctypes.data, max_agents
    rc = icl_lib().sampo_abms_run(C.byref(p), C.byref(r))
    del keep
    if rc != 0:
        raise SampoError(f"sampo_abms_run failed with {rc}")
    layout = np.array([[r.layout.start[s], r.layout.end[s]] for s in range(8)], dtype=np.uint32)
    out = dict(rows=rows[:r.steps_done], status=r.status, steps_done=r.steps_done, agent_steps=r.agent_steps,
               layout=layout, histogram=np.array(r.histogram[:], dtype=np.uint32), loop_ms=r.loop_ms,
               total_ms=r.total_ms)
    if export:
        span = int(layout[7][1])
        out.update(agents=ag[:span], ages=ge[:span], states=st[:span])
    return out


# ---------------------------------------------------------------- text input/output

def format_stats(step, temp, row, hours_per_step=1.0):
    """printStats line (src/abms.c:67-81) of one stats row (ROW_DTYPE record)."""
    r = Row()
    for s_ in range(8):
        r.count[s_] = int(row["count"][s_])
    for f in ("l1de", "females", "potInfective", "numCycles", "sumCycles", "bites", "infectBites", "valid"):
        setattr(r, f, int(row[f]))
    buf = C.create_string_buffer(512)
    n = icl_lib().sampo_format_stats(buf, 512, step, hours_per_step, temp, C.byref(r))
    if n < 0:
        raise SampoError("format_stats failed")
    return buf.raw[:n].decode()


def format_population(layout):
    """printPopulation text (src/abms.c:83-92) of an (8, 2) range array."""
    lay = Layout()
    for s_ in range(8):
        lay.start[s_], lay.end[s_] = int(layout[s_][0]), int(layout[s_][1])
    buf = C.create_string_buffer(1024)
    n = icl_lib().sampo_format_population(buf, 1024, C.byref(lay))
    if n < 0:
        raise SampoError("format_population failed")
    return buf.raw[:n].decode()


def format_histogram(hist):
    """histogram.txt content (src/abms.c:176-180)."""
    h = np.ascontiguousarray(hist, dtype=np.uint32)
    buf = C.create_string_buffer(2048)
    n = icl_lib().sampo_format_histogram(buf, 2048, h.ctypes.data)
    if n < 0:
        raise SampoError("format_histogram failed")
    return buf.raw[:n].decode()


def read_single_val_file(path, n):
    """readSingleValFile (src/abms.c:94-108): up to n floats."""
    out = np.zeros(n, np.float32)
    k = icl_lib().sampo_read_single_val_file(str(path).encode(), n, out.ctypes.data)
    if k < 0:
        raise SampoError(f"cannot read {path}")
    return out[:k]


def read_environment(path, n, carrying_capacity):
    """readEnvironment (src/abms.c:110-137): up to n environments."""
    out = np.zeros(n, dtype=[(nm, "<f4") for nm, _ in Env._fields_])
    k = icl_lib().sampo_read_environment(str(path).encode(), n, carrying_capacity, out.ctypes.data)
    if k < 0:
        raise SampoError(f"cannot read {path}")
    return out[:k]
