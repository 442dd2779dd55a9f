"""Synthetic scenario sets for the batched and serving workloads.

Every scenario is a configuration-mode run of the agent model: an explicit
initial population (sampo_gen_population), a daily temperature series
T(d) = mean + amp * sin(2 pi d / 365) + noise * N(0, 1) held for the 24
iterations of each day (sampo_gen_temperature), constant per-stage daily
mortality, and the reference's environment with carrying capacity cc.  The
parameters of a set are drawn from a numpy generator seeded per set, so a
set is a pure function of its arguments; each scenario's own random streams
(population, temperature, the reference seed stream) are keyed by its seed.

Sets:
  ensemble  1024 replicates of the single-run parameters with distinct seeds
            1000 + i, initial populations uniform in [1e5, 1e6];
  sweep     4096 scenarios, initial population log-uniform in [1e3, 4e6],
            mean temperature uniform in [18, 32] degC, amplitude in [2, 8],
            per-stage mortality multipliers uniform in [0.5, 1.5], seeds
            100000 + j;
  serving   Poisson arrivals, initial population log-uniform in [1e4, 1e7],
            horizon uniform in 30..365 days, parameters as the sweep, seeds
            1000000 + r.
Carrying capacity is a fixed multiple of the initial population: 0.025 for
the single-run parameters (populations swing between ~0.5x and ~3.4x the
initial size over a year, mean ~1.85x), 0.0125 for the sweep and serving
sets (mean ~1.1x, peak ~1.7x), so that thousands of scenarios fit in HBM.
Capacity (the 95 % check of src/abms.c:305) is 4x the initial population.
"""
from dataclasses import dataclass, field
from concurrent.futures import ThreadPoolExecutor
import math

import numpy as np

DMR = (0.05, 0.03, 0.02, 0.1)  # daily mortality egg, larva, pupa, adult

# The single-run workloads: one scenario with the parameters above, an
# initial population drawn by sampo_gen_population and the temperature
# series of seed 1.
SINGLE_RUN = {
    # one large run: the fused step's bandwidth at scale
    "large": dict(
        desc="plain single run, large: 32M initial agents, buffer capacity 128M, 365-day horizon, "
             "config-mode mortality 0.05/0.03/0.02/0.1 per day, T(d)=25+5sin(2pi d/365)+N(0,1.5)",
        n0=32_000_000, max_agents=128_000_000, cc_per_agent=0.025, seed=1, days=365),
    # one small run, the size a CPU runs in minutes
    "small": dict(
        desc="plain single run, small: 100k initial agents, 365-day horizon, config-mode mortality",
        n0=100_000, max_agents=16_000_000, cc_per_agent=0.025, seed=1, days=365),
}


def single_run_inputs(w):
    """(hourly temperature, environment, layout, agents) of a SINGLE_RUN entry."""
    import sampo_amd as S
    temp = S.gen_temperature(w["seed"], w["days"])
    layout, agents = S.gen_population(w["seed"], w["n0"], float(temp[0]))
    env = S.make_env(cc=w["cc_per_agent"] * w["n0"])
    return temp, env, layout, agents


@dataclass
class Spec:
    index: int
    seed: int
    n0: int
    days: int
    t_mean: float = 25.0
    t_amp: float = 5.0
    t_noise: float = 1.5
    dmr: tuple = DMR
    cc_factor: float = 0.025
    arrival: float = 0.0  # seconds after the start of a serving run
    cap_factor: float = 4.0
    extra: dict = field(default_factory=dict)

    @property
    def horizon(self):
        return 24 * self.days

    @property
    def max_agents(self):
        return int(self.cap_factor * self.n0) + 4096

    @property
    def cc(self):
        return self.cc_factor * self.n0


def ensemble(count=1024, days=365, seed=2):
    rng = np.random.default_rng(seed)
    n0 = rng.integers(100_000, 1_000_000, size=count, endpoint=True)
    return [Spec(index=i, seed=1000 + i, n0=int(n0[i]), days=days) for i in range(count)]


def _swept(rng, count):
    n0 = np.rint(np.exp(rng.uniform(math.log(1e3), math.log(4e6), size=count))).astype(np.int64)
    mean = rng.uniform(18.0, 32.0, size=count)
    amp = rng.uniform(2.0, 8.0, size=count)
    mult = rng.uniform(0.5, 1.5, size=(count, 4))
    return n0, mean, amp, mult


def sweep(count=4096, days=365, seed=3):
    rng = np.random.default_rng(seed)
    n0, mean, amp, mult = _swept(rng, count)
    return [Spec(index=j, seed=100000 + j, n0=int(n0[j]), days=days, t_mean=float(mean[j]), t_amp=float(amp[j]),
                 dmr=tuple(min(1.0, DMR[c] * float(mult[j, c])) for c in range(4)), cc_factor=0.0125)
            for j in range(count)]


def serving(rate=50.0, duration=600.0, seed=4, n0_range=(1e4, 1e7), days_range=(30, 365)):
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        r = len(out)
        n0 = int(round(math.exp(rng.uniform(math.log(n0_range[0]), math.log(n0_range[1])))))
        days = int(rng.integers(days_range[0], days_range[1], endpoint=True))
        mult = rng.uniform(0.5, 1.5, size=4)
        out.append(Spec(index=r, seed=1_000_000 + r, n0=n0, days=days, t_mean=float(rng.uniform(18.0, 32.0)),
                        t_amp=float(rng.uniform(2.0, 8.0)),
                        dmr=tuple(min(1.0, DMR[c] * float(mult[c])) for c in range(4)), cc_factor=0.0125,
                        arrival=t))
    return out


def materialise(spec, phys_factor=0.0):
    """Keyword arguments of Engine.admit / scenario() for a spec.  With
    phys_factor > 0 the instance starts with phys_factor x n0 physical slots
    and grows on demand."""
    import sampo_amd as S
    temp = S.gen_temperature(spec.seed, spec.days, mean=spec.t_mean, amp=spec.t_amp, noise=spec.t_noise)
    layout, agents = S.gen_population(spec.seed, spec.n0, float(temp[0]))
    kw = dict(mode=S.MODE_CONFIG, seed=spec.seed, horizon=spec.horizon, temp=temp, env=S.make_env(cc=spec.cc),
              max_agents=spec.max_agents, dmr=spec.dmr, layout=layout, agents=agents)
    if phys_factor > 0:
        kw["phys_slots"] = int(phys_factor * spec.n0) + 256
    return kw


def materialise_all(specs, phys_factor=0.0, threads=16, ahead=64):
    """Yield (spec, admit kwargs) in order; populations are generated on a
    thread pool (the generator releases the GIL) at most `ahead` ahead."""
    with ThreadPoolExecutor(max_workers=threads) as pool:
        pending = []
        it = iter(specs)
        for sp in it:
            pending.append((sp, pool.submit(materialise, sp, phys_factor)))
            if len(pending) >= ahead:
                break
        while pending:
            sp, fut = pending.pop(0)
            nxt = next(it, None)
            if nxt is not None:
                pending.append((nxt, pool.submit(materialise, nxt, phys_factor)))
            yield sp, fut.result()
