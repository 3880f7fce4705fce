"""Transport cases through non uniform media (exponential density profiles
with a proposed locals step, src/pumas.c:5528-5552), shared by the host
engine and GPU parity tests. Each case: (name, physics, context, geometry
description, initial states)."""
import pumas_amd as pa
from pumas_amd import workloads as W


def cases(n=512):
    fwd = W.atmosphere_states(n, seed=11)
    bwd = W.atmosphere_states(n, seed=12, forward=False, kmin=1e-1, kmax=1e2)
    straggled_fwd = W.slabs_forward_context()
    # Backward histories stop at 10 TeV: a backward muon scattered to a
    # grazing direction in the horizontally infinite rock layer would gain
    # energy without bound, up to where the reference's elastic hard
    # scattering angle loop no longer ends (beyond ~1e10 GeV).
    cap = dict(event=pa.Event.LIMIT_ENERGY, limit_energy=1e4)
    straggled_bwd = W.slabs_backward_context().replace(**cap)
    csda = pa.Context(energy_loss=pa.Mode.CSDA, decay=pa.Mode.WEIGHTED,
                      direction=pa.Mode.FORWARD,
                      scattering=pa.Mode.DISABLED)
    mixed_bwd = pa.Context(energy_loss=pa.Mode.MIXED, decay=pa.Mode.WEIGHTED,
                           direction=pa.Mode.BACKWARD,
                           scattering=pa.Mode.MIXED, **cap)
    timed = pa.Context(energy_loss=pa.Mode.CSDA, decay=pa.Mode.WEIGHTED,
                       direction=pa.Mode.BACKWARD,
                       scattering=pa.Mode.DISABLED,
                       event=pa.Event.LIMIT_DISTANCE | pa.Event.LIMIT_TIME,
                       limit_distance=300.0, limit_time=1.5)
    water = W.csda_rock_states(n, seed=13, kmin=1e-1, kmax=1e2)
    water["direction"][:, 0] = 0.6
    water["direction"][:, 2] = -0.8
    return [
        ("atmosphere_forward_straggled_raw", "standard_c005", straggled_fwd,
         W.atmosphere_geometry(pa.Step.RAW), fwd),
        ("atmosphere_forward_straggled_check", "standard_c005",
         straggled_fwd, W.atmosphere_geometry(pa.Step.CHECK), fwd),
        ("atmosphere_backward_straggled_check", "standard_c005",
         straggled_bwd, W.atmosphere_geometry(pa.Step.CHECK, scale=200.0),
         bwd),
        ("atmosphere_backward_mixed_raw", "standard_c005", mixed_bwd,
         W.atmosphere_geometry(pa.Step.RAW, cos_min=1.0), bwd),
        ("atmosphere_forward_csda_check", "standard_c005", csda,
         W.atmosphere_geometry(pa.Step.CHECK, scale=300.0), fwd),
        ("exponential_water_csda_limits", "standard_c005", timed,
         W.exponential_infinite_geometry(), water),
    ]
