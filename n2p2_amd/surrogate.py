"""Distilled surrogate for multiple-time-step MD: a radial-only HDNNP
(8 G2 per element: 4 eta x 2 neighbour elements, network 8-16-16-1 softplus)
trained on the device on energies and forces that the full HDNNP produces
along its own MD trajectory.

The surrogate is an ordinary n2p2 model (settings, scaling, weights files),
so the same engine evaluates it; training uses the engine's gradient and
Adam kernels (hdnnp_train_*), and the multiple-time-step integrator
(hdnnp_mts_*) runs it against the full model.
"""
import os

import numpy as np

from n2p2_amd import engine
from n2p2_amd import workloads as W

SURROGATE_HIDDEN = (16, 16)


def surrogate_settings(elements=W.WATER_ELEMENTS, rc=6.35, cutoff_type=1,
                       etas=(0.001, 0.0078, 0.0606, 0.5)):
    """Settings of the radial-only surrogate: per element one G2 (rs = 0) per
    (neighbour element, eta); softplus hidden layers, linear output."""
    lines = ["number_of_elements %d" % len(elements),
             "elements " + " ".join(elements),
             "cutoff_type %d" % cutoff_type,
             "scale_symmetry_functions", "scale_min_short 0.0",
             "scale_max_short 1.0",
             "global_hidden_layers_short %d" % len(SURROGATE_HIDDEN),
             "global_nodes_short " + " ".join(str(h) for h in SURROGATE_HIDDEN),
             "global_activation_short p p l"]
    for e in elements:
        for e1 in elements:
            for eta in etas:
                lines.append("symfunction_short %s 2 %s %.10e 0.0 %.4f"
                             % (e, e1, eta, rc))
    return "\n".join(lines) + "\n"


def write_model_gpu_scaled(d, settings, elements, z_numbers, hidden, r, elem,
                           box, seed=7):
    """Model files with Xavier weights and min/max scaling from the GPU SF
    statistics of the structure (r, elem, box); returns the loaded model."""
    W.write_model(d, settings, elements, z_numbers, hidden, seed=seed)
    m0 = engine.Model.load_dir(d, scaling=False)
    c0 = engine.Context(m0)
    c0.set_structure(r, elem, box)
    c0.calculate_neighbor_list(m0.max_cutoff)
    stats = []
    for e in range(m0.num_elements):
        _, st = c0.sf_statistics(e)
        st = st.copy()
        flat = st[:, 1] - st[:, 0] <= 0
        st[flat, 1] = st[flat, 0] + 1.0
        stats.append(st)
    c0.close()
    W.write_scaling(os.path.join(d, "scaling.data"), stats)
    return engine.Model.load_dir(d)


def layer_sizes(nsf, hidden=SURROGATE_HIDDEN):
    return [nsf] + list(hidden) + [1]


def flat_to_file_order(w, sizes):
    """One element's weights from the engine's flat layout (per layer
    W[k][j] row-major, then b) into the weight-file order (per layer j
    outer, k inner, then b; NeuralNetwork.cpp:147-169)."""
    out, p = [], 0
    for a, b in zip(sizes[:-1], sizes[1:]):
        Wm = np.asarray(w[p:p + a * b]).reshape(b, a)
        p += a * b
        out.extend(Wm.T.reshape(-1))
        out.extend(w[p:p + b])
        p += b
    return np.array(out)


def write_weights(d, w, nsf_per_element, z_numbers, hidden=SURROGATE_HIDDEN):
    """Writes the engine's flat weights as weights.%03d.data files."""
    p = 0
    for nsf, z in zip(nsf_per_element, z_numbers):
        sizes = layer_sizes(nsf, hidden)
        nw = sum(a * b + b for a, b in zip(sizes[:-1], sizes[1:]))
        fw = flat_to_file_order(w[p:p + nw], sizes)
        p += nw
        with open(os.path.join(d, "weights.%03d.data" % z), "w") as f:
            f.write("".join("%.17e\n" % v for v in fw))
    assert p == len(w)


def temperature(v, atom_masses):
    """Kinetic temperature (K) of velocities v (A/fs), masses in amu."""
    m = np.asarray(atom_masses, dtype=np.float64)
    return float((m[:, None] * v * v).sum()) / (3 * len(v) * W.KB_EV * W.FTOA)


def generate_frames(model, r, elem, box, masses, nframes, stride=4,
                    dt=0.25, temp=300.0, seed=1, precision=None,
                    final_state=False, equil=0, rescale_every=0, log=None):
    """Frames of the full model's own trajectory (velocity Verlet,
    Verlet-skin list): every `stride` steps the positions (wrapped into the
    cell), the energy and the forces.  With rescale_every > 0 the
    velocities are scaled to `temp` every rescale_every steps (a simple
    velocity-rescaling thermostat), through `equil` steps before the first
    frame and on between the frames.  Returns (R, E, F) arrays, and with
    final_state the trajectory's last positions (wrapped) and velocities
    as a fourth item."""
    ctx = engine.Context(model)
    if precision is not None:
        ctx.set_precision(precision)
    ctx.set_structure(r, elem, box)
    m_atom = np.asarray(masses)[elem]
    v = W.maxwell_boltzmann(m_atom, temp, seed=seed)
    ctx.md_init(masses, dt, skin=1.0, v=v)

    def rescale():
        _, u = ctx.md_state()
        t = temperature(u, m_atom)
        ctx.md_init(masses, dt, skin=1.0, v=u * np.sqrt(temp / t))
        return t

    if rescale_every > 0:
        for k in range(equil // rescale_every):
            ctx.md_run(rescale_every)
            t = rescale()
            if log and (k + 1) % 20 == 0:
                log("equilibration step %d: %.0f K before rescaling"
                    % ((k + 1) * rescale_every, t))
    R = np.zeros((nframes, len(r), 3))
    E = np.zeros(nframes)
    F = np.zeros((nframes, len(r), 3))
    for k in range(nframes):
        ep, _ = ctx.md_run(stride)
        x, _ = ctx.md_state()
        R[k] = W.remap(x, box)
        E[k] = ep[-1]
        F[k] = ctx.forces()
        if rescale_every > 0 and ((k + 1) * stride) % rescale_every == 0:
            rescale()
    last = None
    if final_state:
        x, v = ctx.md_state()
        last = (W.remap(x, box), v.copy())
    ctx.close()
    return (R, E, F, last) if final_state else (R, E, F)


def train(ctx, elem, box, R, E, F, epochs=20, lr=1e-3, w_e=1.0, w_f=1.0,
          seed=0, log=None):
    """Adam on the frames (one frame per step, shuffled each epoch).
    Returns the mean (energy, force) loss terms of every epoch, measured on
    the frames as they are visited."""
    rng = np.random.default_rng(seed)
    ctx.set_structure(R[0], elem, box)
    ctx.train_reset()
    hist = []
    for ep in range(epochs):
        le = lf = 0.0
        for k in rng.permutation(len(R)):
            ctx.set_positions(R[k])
            _, (a, b) = ctx.train_gradient(F[k], E[k], w_e, w_f,
                                           want_grad=False)
            ctx.train_adam(lr)
            le += a
            lf += b
        hist.append((le / len(R), lf / len(R)))
        if log:
            log("epoch %d: loss_E %.4g loss_F %.4g" % (ep, hist[-1][0],
                                                       hist[-1][1]))
    return hist
