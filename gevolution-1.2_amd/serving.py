"""serving.py -- whole simulations served concurrently on one GPU.

A request is one run of the reference's program: generateIC_basic
(main.cpp:300; ic_basic.hpp:1626-2259) from a BBKS table for its cosmology
and seed, then the main loop (main.cpp:372-879) with the delta power
spectrum written at z_end (output.hpp:1796-1853, 1983-2028) the way the
reference writes it with EXACT_OUTPUT_REDSHIFTS: interpolated in redshift
between the two cycles around z_end, the loop ending after the field stage
of the crossing cycle (gev_run_to_outputs).  gevolution itself runs one
simulation per process (main.cpp:115-149); this loop has no counterpart
there and calls only the C ABI (include/gev_pm.h) through gevpm.

Admission.  Requests are taken in arrival order.  The head of the queue is
admitted when a slot is free and its box's device footprint
(gev_ctx_footprint_bytes) fits in the memory budget next to the running
boxes and the idle pooled ones; idle boxes are evicted, least recently used
first, to make room.  A request larger than the whole budget is rejected
at once rather than left to block the queue.

Residency and slot reuse.  A finished box goes back to a pool keyed by
(grid, form), and the next request of that shape reuses it: the initial
conditions overwrite the whole state, so nothing of the previous run
survives, and the allocation is saved.

Two servers share that admission policy:

- Server runs each admitted request in a host thread on a box of its own
  (its own HIP stream).  With one slot it is the back-to-back plain run.
- BatchServer is continuous batching: requests of the same grid and gravity
  share a batch (include/gev_pm.h gev_batch_*), whose every cycle kernel is
  one launch over all the boxes it holds.  A request that arrives while
  others run takes a free slot of a batch of its shape (or a new batch, if
  the memory budget allows) and joins the next cycle; a request that writes
  its z = 0 spectrum leaves its slot at once.  All running requests advance
  one cycle per round, the batches on their own streams.

A request's results do not depend on what else runs: its kernels touch only
its own box (its slab of a batch), in a fixed order, and every atomic in
them either only counts (the filing, whose order a sort then fixes) or is
order-independent (integer maxima, fixed-point sums); a batch box equals the
box run alone bit for bit.
"""
import collections
import math
import os
import time
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field

import numpy as np

import gevpm

C_LIGHT = 2997.92458  # 100 km/s units (metadata.hpp:98)


@dataclass
class Request:
    """One simulation: grid N with N^3 particles, box in Mpc/h, to z = 0."""
    rid: int
    ngrid: int
    gr: bool = True
    Omega_m: float = 0.31
    A_s: float = 2.1e-9
    seed: int = 0
    boxsize: float = 0.0          # Mpc/h; 0 means 5 Mpc/h per cell
    arrival: float = 0.0          # seconds after the start of the trace
    form: int = gevpm.FORM_FP64
    Omega_b: float = 0.05
    h: float = 0.67
    n_s: float = 0.965
    z_in: float = 100.0
    z_end: float = 0.0

    @property
    def box_mpc(self):
        return self.boxsize if self.boxsize > 0 else 5.0 * self.ngrid

    @property
    def npart(self):
        return self.ngrid ** 3


@dataclass
class Outcome:
    rid: int
    status: str                   # "ok", "rejected" or "failed"
    arrival: float
    admitted: float = math.nan
    finished: float = math.nan
    reused: bool = False
    error: str = ""
    result: dict = field(default_factory=dict)

    @property
    def latency(self):
        """Arrival to completion, seconds."""
        return self.finished - self.arrival

    @property
    def service(self):
        return self.finished - self.admitted


def percentile(values, p):
    """Nearest-rank percentile (p in (0, 100])."""
    v = sorted(values)
    if not v:
        return math.nan
    return v[max(0, math.ceil(p / 100.0 * len(v)) - 1)]


def poisson_trace(n, rate, seed, grids=(64, 96, 128, 192, 256), weights=(0.3, 0.25, 0.2, 0.15, 0.1),
                  gr_fraction=0.5, Omega_m=(0.25, 0.35), A_s=(1.8e-9, 2.4e-9)):
    """n requests with exponential inter-arrival times (rate per second),
    grids drawn with the given weights, GR or Newtonian with probability
    gr_fraction, Omega_m and A_s uniform in the given ranges; request i
    has seed i."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    g = rng.choice(np.asarray(grids), size=n, p=np.asarray(weights) / np.sum(weights))
    gr = rng.random(n) < gr_fraction
    om = rng.uniform(*Omega_m, n)
    a_s = rng.uniform(*A_s, n)
    return [Request(i, int(g[i]), bool(gr[i]), float(om[i]), float(a_s[i]), i, arrival=float(t[i]))
            for i in range(n)]


class Server:
    """Admission, slot reuse and overlap over a backend with
    footprint(req), key(req), create(req), run(box, req) -> dict and
    destroy(box)."""

    def __init__(self, backend, memory_budget, slots=4, reuse=True, clock=time.perf_counter):
        if slots < 1:
            raise ValueError("slots must be at least 1")
        self.backend = backend
        self.budget = int(memory_budget)
        self.slots = int(slots)
        self.reuse = bool(reuse)
        self.clock = clock
        self.in_use = 0                              # bytes of running and pooled boxes
        self._pool = collections.OrderedDict()       # (key, serial) -> (box, bytes), LRU first
        self._serial = 0
        self.peak_in_use = 0
        self.peak_running = 0
        self.created = 0
        self.evicted = 0

    # -- memory and pool (dispatcher thread only)
    def _take(self, req, need):
        """A box for req, or None while the memory is held by running boxes."""
        key = self.backend.key(req)
        if self.reuse:
            for pk in self._pool:
                if pk[0] == key:
                    box, _ = self._pool.pop(pk)
                    return box, True
        pooled = sum(nb for _, nb in self._pool.values())
        if self.in_use - pooled + need > self.budget:
            return None  # running boxes hold the memory: keep the pool for later requests
        while self.in_use + need > self.budget:
            _, (box, nb) = self._pool.popitem(last=False)
            self.backend.destroy(box)
            self.in_use -= nb
            self.evicted += 1
        box = self.backend.create(req)
        self.created += 1
        self.in_use += need
        self.peak_in_use = max(self.peak_in_use, self.in_use)
        return box, False

    def _give_back(self, req, box, need, healthy):
        if self.reuse and healthy:
            self._serial += 1
            self._pool[(self.backend.key(req), self._serial)] = (box, need)
        else:
            self.backend.destroy(box)
            self.in_use -= need

    def close(self):
        """Frees the pooled boxes."""
        while self._pool:
            _, (box, nb) = self._pool.popitem(last=False)
            self.backend.destroy(box)
            self.in_use -= nb

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- the loop
    def serve(self, requests):
        """Runs the requests as they arrive; returns their outcomes in
        request order.  Arrival times are seconds after the call."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        outcomes = {}
        waiting = collections.deque()
        running = {}
        t0 = self.clock()
        i = 0
        with ThreadPoolExecutor(max_workers=self.slots) as ex:
            while i < len(reqs) or waiting or running:
                now = self.clock() - t0
                while i < len(reqs) and reqs[i].arrival <= now:
                    waiting.append(reqs[i])
                    i += 1
                while waiting and len(running) < self.slots:
                    r = waiting[0]
                    need = int(self.backend.footprint(r))
                    if need < 0 or need > self.budget:
                        waiting.popleft()
                        outcomes[r.rid] = Outcome(r.rid, "rejected", r.arrival,
                                                  error="footprint %d B exceeds the budget %d B" % (need, self.budget))
                        continue
                    try:
                        got = self._take(r, need)
                    except Exception as e:  # allocation failed: report it, keep serving
                        waiting.popleft()
                        outcomes[r.rid] = Outcome(r.rid, "failed", r.arrival, error="create: %s" % e)
                        continue
                    if got is None:
                        if not running:  # nothing will free memory: never wait forever
                            waiting.popleft()
                            outcomes[r.rid] = Outcome(r.rid, "rejected", r.arrival,
                                                      error="footprint %d B does not fit next to %d B in use"
                                                      % (need, self.in_use))
                            continue
                        break  # the head waits for memory; arrival order is kept
                    waiting.popleft()
                    box, reused = got
                    adm = self.clock() - t0
                    running[ex.submit(self.backend.run, box, r)] = (r, box, need, adm, reused)
                    self.peak_running = max(self.peak_running, len(running))
                timeout = None
                if i < len(reqs):
                    timeout = max(0.0, reqs[i].arrival - (self.clock() - t0))
                if running:
                    done, _ = wait(list(running), timeout=timeout, return_when=FIRST_COMPLETED)
                else:
                    if timeout:
                        time.sleep(timeout)
                    done = ()
                for f in done:
                    r, box, need, adm, reused = running.pop(f)
                    fin = self.clock() - t0
                    try:
                        res = f.result()
                        outcomes[r.rid] = Outcome(r.rid, "ok", r.arrival, adm, fin, reused, result=res)
                        self._give_back(r, box, need, True)
                    except Exception as e:
                        outcomes[r.rid] = Outcome(r.rid, "failed", r.arrival, adm, fin, reused, error=str(e))
                        self._give_back(r, box, need, False)
        return [outcomes[r.rid] for r in sorted(requests, key=lambda r: r.rid)]


def summarize(outcomes):
    """p50 / p99 time to completion (arrival to z = 0), makespan and the
    aggregate particle-steps/s over the trace."""
    ok = [o for o in outcomes if o.status == "ok"]
    lat = [o.latency for o in ok]
    steps = sum(o.result.get("particle_steps", 0) for o in ok)
    t_first = min((o.arrival for o in outcomes), default=0.0)
    t_last = max((o.finished for o in ok), default=0.0)
    span = t_last - t_first
    return {"completed": len(ok), "rejected": sum(o.status == "rejected" for o in outcomes),
            "failed": sum(o.status == "failed" for o in outcomes),
            "p50_s": round(percentile(lat, 50), 4), "p99_s": round(percentile(lat, 99), 4),
            "makespan_s": round(span, 4), "particle_steps": int(steps),
            "particle_steps_per_s": round(steps / span, 1) if span > 0 else None}


class GevBackend:
    """Runs requests on the current GPU through the C ABI: one box per slot,
    each on its own stream (gev_ctx_create with a NULL stream)."""

    def __init__(self, template=None, numbins=None, kgrid=None):
        import bbks
        self._bbks = bbks
        if template is None:
            # the sc1 crystal's particles, a 4^3 lattice at (i + 1/2) / 4, x
            # fastest (the template file lists them in another order: pass it,
            # read with gevpm.load_homogeneous_template, for the reference's IDs)
            g = (np.arange(4) + 0.5) / 4.0
            z, y, x = np.meshgrid(g, g, g, indexing="ij")
            template = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
        self.template = np.ascontiguousarray(template, dtype=np.float32).reshape(-1, 3)
        self.tile = round(len(self.template) ** (1.0 / 3.0))
        if self.tile ** 3 != len(self.template):
            raise ValueError("the template must hold a cube number of particles")
        self.numbins = numbins
        self.kgrid = np.logspace(-5.0, np.log10(300.0), 400) if kgrid is None else np.asarray(kgrid)

    def key(self, req):
        return (req.ngrid, req.form)

    def footprint(self, req):
        if req.ngrid % self.tile:
            return -1
        return gevpm.footprint_bytes(req.ngrid, req.npart, req.form)

    def create(self, req):
        return gevpm.Box(req.ngrid, req.npart, form=req.form)

    def destroy(self, box):
        box.close()

    @staticmethod
    def setup(req):
        """(sim, state) of main.cpp:278-295 for the request."""
        cosmo = gevpm.make_cosmology(req.Omega_m - req.Omega_b, req.Omega_b, req.h)
        fourpiG = 1.5 * req.box_mpc ** 2 / C_LIGHT ** 2
        sim = gevpm.SimParams(1 if req.gr else 0, 0, 48.0, 0.04, 100.0, fourpiG, cosmo)
        a0 = 1.0 / (1.0 + req.z_in)
        st = gevpm.CycleState(a0, 0.0, gevpm.next_dtau(sim, req.ngrid, a0), 0.0, 0.0, 0.0, 0.0, 0, 0)
        return sim, st

    def run(self, box, req):
        N = req.ngrid
        sim, st = self.setup(req)
        L, h = req.box_mpc, req.h
        k, d, t = self._bbks.bbks_table(self.kgrid, req.z_in, req.Omega_m, h)
        ic = gevpm.ICParams(int(req.seed), 1, 1, N // self.tile, req.z_in, req.A_s, req.n_s, 0.05, L)
        mv, _ = box.generateIC_basic(sim, ic, self.template, k * L, d, t * L / h, d, t * L / h)
        st.maxvel = mv
        out = box.run_to_outputs(sim, st, [req.z_end], self.numbins or N // 2, L)
        phi_sum = box.field_sum(gevpm.PHI)
        # st.cycle particle updates ran: the last cycle stops after its field stage
        return {"a": st.a, "cycles": st.cycle, "particle_steps": req.npart * st.cycle, "z": float(out["z"][0]),
                "k": out["k"][0], "pk": out["pk"][0], "occupation": out["occupation"][0], "phi_sum": phi_sum,
                "maxvel": st.maxvel}


class BatchBackend:
    """Batches on the current GPU through the C ABI, for BatchServer."""

    def __init__(self, template=None, numbins=None, form=gevpm.FORM_FP64, graph=True):
        self.gev = GevBackend(template, numbins)
        self.form = form
        self.graph = graph

    def key(self, req):
        return (req.ngrid, 1 if req.gr else 0, self.form)

    def slot_footprint(self, req):
        if req.ngrid % self.gev.tile:
            return -1
        return gevpm.batch_footprint_bytes(req.ngrid, req.npart, 1, self.form, 1 if req.gr else 0)

    def create(self, key, slots):
        N, gr, form = key
        B = gevpm.Batch(N, N ** 3, slots, gr=gr, form=form)
        B.set_outputs([0.0], self.gev.numbins or N // 2)
        return B

    def destroy(self, batch):
        batch.close()

    def start(self, batch, slot, req):
        """The request's initial conditions into the slot; the slot joins the
        batch's next cycle."""
        N = req.ngrid
        sim, st = GevBackend.setup(req)
        L, h = req.box_mpc, req.h
        k, d, t = self.gev._bbks.bbks_table(self.gev.kgrid, req.z_in, req.Omega_m, h)
        ic = gevpm.ICParams(int(req.seed), 1, 1, N // self.gev.tile, req.z_in, req.A_s, req.n_s, 0.05, L)
        view = batch.box(slot)
        if self.form == gevpm.FORM_FP64:
            view.generateIC_basic(sim, ic, self.gev.template, k * L, d, t * L / h, d, t * L / h)
        else:
            with gevpm.Box(N, N ** 3) as tmp:
                tmp.generateIC_basic(sim, ic, self.gev.template, k * L, d, t * L / h, d, t * L / h)
                view.copy_state_from(tmp)
        batch.start_box(slot, sim, st, L)

    def step(self, batch):
        return batch.step(self.graph)

    def result(self, batch, slot, req):
        out = batch.outputs(slot)
        st = batch.state(slot)
        return {"a": st.a, "cycles": st.cycle, "particle_steps": req.npart * st.cycle, "z": float(out["z"][0]),
                "k": out["k"][0], "pk": out["pk"][0], "occupation": out["occupation"][0],
                "phi_sum": batch.box(slot).field_sum(gevpm.PHI), "maxvel": st.maxvel}

    def release(self, batch, slot):
        batch.stop_box(slot)


class BatchServer:
    """Continuous batching over a backend with key(req), slot_footprint(req),
    create(key, slots), destroy(batch), start(batch, slot, req),
    step(batch) -> finished slots, result(batch, slot, req) and
    release(batch, slot).  slots(key) gives a new batch's slot count."""

    def __init__(self, backend, memory_budget, slots=None, clock=time.perf_counter):
        self.backend = backend
        self.budget = int(memory_budget)
        self.slots = slots or (lambda key: 8 if key[0] <= 128 else (4 if key[0] <= 192 else 2))
        self.clock = clock
        self.batches = []          # [key, batch, slots, {slot: (req, admitted)}, bytes]
        self.in_use = 0
        self.peak_in_use = 0
        self.peak_running = 0
        self.created = 0
        self.evicted = 0
        self.rounds = 0

    def close(self):
        for e in self.batches:
            self.backend.destroy(e[1])
            self.in_use -= e[4]
        self.batches = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _slot_for(self, req):
        """(entry, slot) for req, a new batch if needed and affordable, or None."""
        key = self.backend.key(req)
        for e in self.batches:
            if e[0] == key and len(e[3]) < e[2]:
                return e, min(set(range(e[2])) - set(e[3]))
        n = self.slots(key)
        need = self.backend.slot_footprint(req) * n
        while self.in_use + need > self.budget:
            idle = [e for e in self.batches if not e[3]]
            if not idle:
                return None
            e = idle[0]
            self.backend.destroy(e[1])
            self.batches.remove(e)
            self.in_use -= e[4]
            self.evicted += 1
        e = [key, self.backend.create(key, n), n, {}, need]
        self.batches.append(e)
        self.created += 1
        self.in_use += need
        self.peak_in_use = max(self.peak_in_use, self.in_use)
        return e, 0

    def serve(self, requests):
        """Runs the requests as they arrive; outcomes in request order."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        outcomes, waiting = {}, collections.deque()
        t0 = self.clock()
        i = 0
        while i < len(reqs) or waiting or any(e[3] for e in self.batches):
            now = self.clock() - t0
            while i < len(reqs) and reqs[i].arrival <= now:
                waiting.append(reqs[i])
                i += 1
            while waiting:
                r = waiting[0]
                one = self.backend.slot_footprint(r)
                if one < 0 or one > self.budget:
                    waiting.popleft()
                    outcomes[r.rid] = Outcome(r.rid, "rejected", r.arrival,
                                              error="footprint %d B exceeds the budget %d B" % (one, self.budget))
                    continue
                got = self._slot_for(r)
                if got is None:
                    if not any(e[3] for e in self.batches):
                        waiting.popleft()
                        outcomes[r.rid] = Outcome(r.rid, "rejected", r.arrival,
                                                  error="a batch for it does not fit next to %d B in use" % self.in_use)
                        continue
                    break  # waits for a slot or memory; arrival order is kept
                waiting.popleft()
                e, slot = got
                try:
                    self.backend.start(e[1], slot, r)
                except Exception as ex:
                    outcomes[r.rid] = Outcome(r.rid, "failed", r.arrival, error="start: %s" % ex)
                    continue
                e[3][slot] = (r, self.clock() - t0)
            running = sum(len(e[3]) for e in self.batches)
            self.peak_running = max(self.peak_running, running)
            if not running:
                if i < len(reqs):
                    time.sleep(max(0.0, reqs[i].arrival - (self.clock() - t0)))
                continue
            self.rounds += 1
            broken = []
            for e in self.batches:
                if not e[3]:
                    continue
                try:
                    fin = self.backend.step(e[1])
                except Exception as ex:  # the batch is lost: fail its requests, free it
                    for slot, (r, adm) in e[3].items():
                        outcomes[r.rid] = Outcome(r.rid, "failed", r.arrival, adm, self.clock() - t0, error=str(ex))
                    e[3].clear()
                    broken.append(e)
                    continue
                for slot in fin:
                    r, adm = e[3].pop(slot)
                    res = self.backend.result(e[1], slot, r)
                    self.backend.release(e[1], slot)
                    outcomes[r.rid] = Outcome(r.rid, "ok", r.arrival, adm, self.clock() - t0, result=res)
            for e in broken:
                self.backend.destroy(e[1])
                self.batches.remove(e)
                self.in_use -= e[4]
        return [outcomes[r.rid] for r in sorted(requests, key=lambda r: r.rid)]


def default_budget():
    """Device bytes the server may hold: GEV_SERVING_BUDGET if set, else
    288 GB of HBM less 8 GB for the runtime and other processes."""
    return int(os.environ.get("GEV_SERVING_BUDGET", 280 * 10 ** 9))
