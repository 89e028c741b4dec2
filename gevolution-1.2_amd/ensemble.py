"""ensemble.py -- many whole simulations on one GPU as batches.

An ensemble is a list of boxes (grid N with N^3 particles, box 5 Mpc/h per
cell, cosmology, seed), each one run of the reference's program from its
initial conditions (generateIC_basic, ic_basic.hpp:1626-2259) to its delta
power spectrum at z_end, written as the reference writes it
(main.cpp:605-693, EXACT_OUTPUT_REDSHIFTS).  The reference runs one
simulation per process (main.cpp:115-149); here the boxes of one grid size
form one batch (include/gev_pm.h gev_batch_*): every cycle kernel is one
launch over all of them, and the two stages of a cycle replay HIP graphs.
Each batch has its own stream, and the batches are stepped round-robin, so
the grids' launches overlap on the device.  Every box's result equals the
same box run alone (gev_run_to_outputs) bit for bit.
"""
import time
from dataclasses import dataclass

import numpy as np

import gevpm

C_LIGHT = 2997.92458  # 100 km/s units (metadata.hpp:98)


@dataclass
class BoxSpec:
    """One simulation of the ensemble."""
    index: int
    ngrid: int
    Omega_m: float
    A_s: float
    seed: int
    gr: int = 1
    Omega_b: float = 0.05
    h: float = 0.67
    n_s: float = 0.965
    z_in: float = 100.0

    @property
    def boxsize(self):
        return 5.0 * self.ngrid


def draw(nbox=256, grids=(32, 48, 64, 96, 128), Omega_m=(0.25, 0.35), A_s=(1.8e-9, 2.4e-9), seed=2025, gr=1):
    """nbox boxes: grid uniform over `grids`, Omega_m and A_s uniform in
    their ranges, seeds 0 .. nbox-1."""
    rng = np.random.default_rng(seed)
    g = rng.choice(np.asarray(grids), size=nbox)
    om = rng.uniform(*Omega_m, nbox)
    a_s = rng.uniform(*A_s, nbox)
    return [BoxSpec(i, int(g[i]), float(om[i]), float(a_s[i]), i, gr) for i in range(nbox)]


def setup(spec):
    """(sim, state, ICParams, transfer tables) of a box: main.cpp:278-295 and
    the BBKS table of its cosmology in loadTransferFunctions' units."""
    import bbks
    L, h = spec.boxsize, spec.h
    k, d, t = bbks.bbks_table(np.logspace(-5.0, np.log10(300.0), 400), spec.z_in, spec.Omega_m, h)
    cosmo = gevpm.make_cosmology(spec.Omega_m - spec.Omega_b, spec.Omega_b, h)
    sim = gevpm.SimParams(spec.gr, 0, 48.0, 0.04, spec.z_in, 1.5 * L * L / C_LIGHT ** 2, cosmo)
    a0 = 1.0 / (1.0 + spec.z_in)
    st = gevpm.CycleState(a0, 0.0, gevpm.next_dtau(sim, spec.ngrid, a0), 0.0, 0.0, 0.0, 0.0, 0, 0)
    ic = gevpm.ICParams(spec.seed, 1, 1, spec.ngrid // 4, spec.z_in, spec.A_s, spec.n_s, 0.05, L)
    return sim, st, ic, (k * L, d, t * L / h, d, t * L / h)


class Ensemble:
    """The boxes of `specs` as one batch per grid size (all GR or all
    Newtonian), in the FP64 or the FP32 form."""

    def __init__(self, specs, template, form=gevpm.FORM_FP32, z_end=0.0, numbins=None):
        self.specs = list(specs)
        self.template = template
        self.form = form
        self.z_end = z_end
        self.numbins = numbins
        self.batches = {}  # N -> (Batch, [spec indices])
        for s in self.specs:
            self.batches.setdefault(s.ngrid, [None, []])[1].append(s.index)
        gr = {s.gr for s in self.specs}
        if len(gr) != 1:
            raise ValueError("an ensemble is all GR or all Newtonian")
        self.gr = gr.pop()
        for N, entry in self.batches.items():
            entry[0] = gevpm.Batch(N, N ** 3, len(entry[1]), gr=self.gr, form=form)
        self.slot = {}
        for N, (B, idx) in self.batches.items():
            for j, i in enumerate(idx):
                self.slot[i] = (B, j)

    def close(self):
        for B, _ in self.batches.values():
            B.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def footprint_bytes(self):
        return sum(gevpm.batch_footprint_bytes(N, N ** 3, len(idx), self.form, self.gr)
                   for N, (_, idx) in self.batches.items())

    def generate(self):
        """Initial conditions of every box (an FP64 box per grid size makes
        them; the FP32 form takes them over with gev_copy_state)."""
        for N, (B, idx) in self.batches.items():
            B.set_outputs([self.z_end], self.numbins or N // 2)
            tmp = gevpm.Box(N, N ** 3) if self.form != gevpm.FORM_FP64 else None
            try:
                for j, i in enumerate(idx):
                    s = self.specs[i]
                    sim, st, ic, tabs = setup(s)
                    view = B.box(j)
                    if tmp is None:
                        view.generateIC_basic(sim, ic, self.template, *tabs)
                    else:
                        tmp.generateIC_basic(sim, ic, self.template, *tabs)
                        view.copy_state_from(tmp)
                    B.start_box(j, sim, st, s.boxsize)
            finally:
                if tmp is not None:
                    tmp.close()

    def autotune(self):
        """Each batch's tile autotuner (outside the run): {N: picks}."""
        return {N: B.autotune() for N, (B, _) in sorted(self.batches.items())}

    def run(self, graph=True):
        """Every box to z_end; returns the wall time in seconds (from the
        first launch to the last box's output)."""
        live = {N: len(idx) for N, (_, idx) in self.batches.items()}
        t0 = time.perf_counter()
        while any(live.values()):
            for N, (B, idx) in self.batches.items():
                if live[N]:
                    live[N] -= len(B.step(graph))
        for B, _ in self.batches.values():
            B.synchronize()
        return time.perf_counter() - t0

    def result(self, i):
        """Box i's z_end spectrum table and final state."""
        B, j = self.slot[i]
        out = B.outputs(j)
        st = B.state(j)
        return {"k": out["k"][0], "pk": out["pk"][0], "occupation": out["occupation"][0], "z": float(out["z"][0]),
                "cycles": st.cycle, "a": st.a, "particle_steps": self.specs[i].ngrid ** 3 * st.cycle}
