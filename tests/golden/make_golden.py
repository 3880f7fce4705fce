#!/usr/bin/env python3
"""Write the golden per-history fixtures of tests/golden/ with the oracle.

The oracle is the reference library compiled from its own source
(oracle/Makefile); it is run here, in the build container, on small seeded
subsets of each workload. Each fixture stores the initial states, the final
states and the per-history event, start / end medium, step and random draw
counts, as produced by pumas_context_transport with the Philox stream.

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle as orc  # noqa: E402
from pumas_amd import Event, Mode  # noqa: E402
from pumas_amd import workloads as W  # noqa: E402

SEED = 0


def cases():
    """(name, physics, context, geometry description, initial states)."""
    yield ("csda_rock", "standard_c005", W.csda_rock_context(),
           W.csda_rock_geometry(), W.csda_rock_states(2048, seed=SEED))
    yield ("slabs_straggled", "standard_c005", W.slabs_forward_context(),
           W.slabs_geometry(), W.slabs_forward_states(256, seed=SEED))
    yield ("slabs_mixed", "standard_c005",
           W.slabs_forward_context(energy_loss=Mode.MIXED),
           W.slabs_geometry(), W.slabs_forward_states(256, seed=SEED))
    yield ("slabs_csda_check", "standard_c005",
           W.slabs_forward_context(energy_loss=Mode.CSDA),
           W.slabs_geometry(step_type=0), W.slabs_forward_states(256, seed=SEED))
    # Backward straggled transport from a detector under the slabs.
    ctx = W.slabs_backward_context()
    yield ("slabs_backward", "standard_c005", ctx, W.slabs_geometry(),
           W.slabs_backward_states(256, seed=SEED))
    # Limits: energy, grammage and time limits, forward and backward.
    yield ("limits_forward", "standard_c005",
           W.slabs_forward_context().replace(
               event=Event.LIMIT_ENERGY | Event.LIMIT_GRAMMAGE,
               limit_energy=50.0, limit_grammage=4e5),
           W.slabs_geometry(), W.slabs_forward_states(128, seed=SEED))
    yield ("limits_time_backward", "standard_c005",
           W.slabs_backward_context().replace(
               event=Event.LIMIT_TIME | Event.LIMIT_ENERGY,
               limit_time=300.0, limit_energy=2e3),
           W.slabs_geometry(), W.slabs_backward_states(128, seed=SEED))


def main():
    physics = {}
    for name, pname, ctx, desc, states in cases():
        if pname not in physics:
            physics[pname] = orc.RefPhysics(pname)
        geom = orc.RefGeometry(desc)
        r = orc.transport(physics[pname], ctx, geom, states, seed=SEED,
                          threads=8)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(
            path, initial=states, final=r.states, events=r.events,
            media=r.media, steps=r.steps, draws=r.draws, rc=r.rc,
            seed=np.uint64(SEED))
        print(f"{name}: {len(states)} histories, {r.total_steps} steps, "
              f"events {np.unique(r.events).tolist()} -> {path}")
    for p in physics.values():
        p.close()


if __name__ == "__main__":
    main()
