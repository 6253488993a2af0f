"""Multiple-time-step MD (hdnnp_mts_*): r-RESPA with a surrogate as the fast
force and the full model as the reference, with a force-RMSE gate that
hands whole outer steps to the full model.

- A surrogate identical to the full model makes the slow force exactly
  zero: the run equals plain velocity Verlet at the inner step (positions
  within 1e-9 A after 80 substeps, energies within 1e-10 relative).
- A gate below zero sends every outer step to the full model: again plain
  velocity Verlet at the inner step.
"""
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd import workloads as W
from tests import cases

pytestmark = pytest.mark.gpu

DT_OUTER, N_INNER = 1.0, 4


def say(*a):
    print("[gpu-mts]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def water(tmp_path_factory):
    d, r, e, box = cases.water_case(tmp_path_factory.mktemp("mts"))
    masses = np.array(W.WATER_MASSES)
    v = W.maxwell_boltzmann(masses[e], 300.0, seed=9)
    return d, r, e, box, masses, v


def _plain(d, r, e, box, masses, v, nsteps):
    c = engine.Context(engine.Model.load_dir(d))
    c.set_structure(r, e, box)
    c.md_init(masses, DT_OUTER / N_INNER, skin=1.0, v=v)
    ep, ek = c.md_run(nsteps)
    x, u = c.md_state()
    return ep, ek, x, u


def _mts(d_full, d_sur, r, e, box, masses, v, nouter, gate):
    full = engine.Context(engine.Model.load_dir(d_full))
    sur = engine.Context(engine.Model.load_dir(d_sur))
    for c in (full, sur):
        c.set_structure(r, e, box)
    sur.mts_init(full, masses, DT_OUTER, N_INNER, skin=1.0, v=v,
                 rmse_gate=gate)
    ep, ek, rm, md = sur.mts_run(nouter)
    x, u = sur.md_state()
    return ep, ek, rm, md, x, u


def _same_mod_lattice(x, y, box):
    d = W.remap(x, box) - W.remap(y, box)
    L = np.diag(box)
    d -= L * np.round(d / L)
    return float(np.abs(d).max())


@pytest.mark.parametrize("gate", [1.0, -1.0])
def test_mts_reduces_to_plain_verlet(water, gate):
    d, r, e, box, masses, v = water
    nouter = 20
    ep, ek, rm, md, x, u = _mts(d, d, r, e, box, masses, v, nouter, gate)
    pep, pek, px, pu = _plain(d, r, e, box, masses, v, nouter * N_INNER)
    dx = _same_mod_lattice(x, px, box)
    du = float(np.abs(u - pu).max())
    de = float(np.max(np.abs(ep - pep[N_INNER - 1::N_INNER])
                      / np.abs(pep[N_INNER - 1::N_INNER])))
    dk = float(np.max(np.abs(ek - pek[N_INNER - 1::N_INNER])
                      / np.abs(pek[N_INNER - 1::N_INNER])))
    say("gate %g: modes %s, max rmse %.3g, dx %.2g A, dv %.2g, de %.2g, dk %.2g"
        % (gate, np.bincount(md, minlength=2).tolist(), rm.max(), dx, du, de,
           dk))
    if gate > 0:
        assert np.all(md == 0) and rm.max() == 0.0
    else:
        assert np.all(md == 1)
    assert dx <= 1e-9 and du <= 1e-9 and de <= 1e-10 and dk <= 1e-10


def test_trained_surrogate_gains(water, tmp_path):
    """What training gains, measured where the integrator uses it: r-RESPA
    without the gate (the surrogate drives every substep) along the same
    trajectory start, with the distilled surrogate trained on the full
    model's own MD frames (the engine's gradient and Adam kernels) and with
    the same surrogate untrained (its initial weights).  The trained one
    tracks the full model's forces more closely at every outer step, so the
    slow force F_full - F_sur it leaves to the outer kicks is smaller: mean
    force RMSE below 0.7 of the untrained one's, and energies stay finite."""
    from n2p2_amd import surrogate as S
    d, r, e, box, masses, v = water
    full_model = engine.Model.load_dir(d)
    sd = str(tmp_path / "sur")
    untrained = S.write_model_gpu_scaled(sd, S.surrogate_settings(),
                                         W.WATER_ELEMENTS, (1, 8),
                                         S.SURROGATE_HIDDEN, r, e, box,
                                         seed=21)
    R, E, F = S.generate_frames(full_model, r, e, box, masses, nframes=32,
                                stride=8)
    ctx = engine.Context(untrained)
    hist = S.train(ctx, e, box, R, E, F, epochs=20, lr=2e-3)
    td = str(tmp_path / "trained")
    import os
    os.makedirs(td)
    for f in ("input.nn", "scaling.data"):
        with open(os.path.join(sd, f)) as a, open(os.path.join(td, f), "w") as b:
            b.write(a.read())
    S.write_weights(td, ctx.get_weights(),
                    [untrained.num_sf(k) for k in range(2)], (1, 8))
    ctx.close()
    res = {}
    for name, dd in (("trained", td), ("untrained", sd)):
        ep, ek, rm, md, _, _ = _mts(d, dd, r, e, box, masses, v, 40, 1e30)
        assert np.all(md == 0)
        assert np.all(np.isfinite(ep + ek))
        res[name] = float(np.mean(rm))
    say("force loss %.3g -> %.3g; MTS mean force RMSE trained %.3g, "
        "untrained %.3g eV/A" % (hist[0][1], hist[-1][1], res["trained"],
                                 res["untrained"]))
    assert res["trained"] < 0.7 * res["untrained"]
