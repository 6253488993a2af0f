"""Training of a 2G-HDNNP on the GPU with the engine's Jacobian, gradient,
Adam and Kalman-filter kernels: the two updaters of nnp-train on energies
and forces (nnp-train.cpp, Training.cpp, GradientDescent.cpp,
KalmanFilter.cpp).

What it does, in the reference's terms:
- settings from input.nn: epochs, random_seed, test_fraction,
  use_short_forces, force_weight, short_force_fraction, weights_min /
  weights_max, updater_type (0 gradient descent with Adam,
  gradient_adam_eta / beta1 / beta2 / epsilon; 1 the standard Kalman
  filter, kalman_epsilon / q0 / qtau / qmin / eta / etatau / etamax);
- data from input.data (energies and forces), converted to the model's
  internal units as Dataset::toNormalizedUnits does (energy offsets of
  atom_energy removed, mean_energy per atom subtracted, conv_energy and
  conv_length applied);
- scaling.data from the working directory (nnp-scaling or
  n2p2_amd.scaling writes it);
- initial weights uniform in [weights_min, weights_max] (own generator,
  seeded with random_seed) unless weights.%03zu.data files are present;
- Adam: per epoch one step per training structure on
  ((E - Eref)/N)^2 + force_weight / (3N) sum |F - Fref|^2, in random
  order;
- Kalman filter: per epoch one energy update per training structure
  (observation E/N, innovation (Eref - E)/N, Jacobian dE/dw / N) and
  short_force_fraction of all force components as force updates
  (innovation Fref - F, Jacobian dF/dw), in random order; the reference's
  update-candidate selection modes and force weighting are not
  reproduced;
- learning-curve.out: per epoch the energy RMSE per atom and the force
  RMSE of the training and test sets in physical units, and the weights of
  every epoch as weights.%03zu.%06zu.out (the reference's names).

Usage, in a directory with input.nn, scaling.data and input.data:

    python -m n2p2_amd.train
"""
import os
import sys

import numpy as np

from n2p2_amd import engine
from n2p2_amd import io as nio


def _settings(path):
    kw = {}
    with open(path) as f:
        for line in f:
            w = line.split("#", 1)[0].split()
            if w:
                kw[w[0]] = w[1:]
    return kw


class Training:
    def __init__(self, directory=".", device=0):
        self.dir = directory
        self.device = device

    def _p(self, f):
        return os.path.join(self.dir, f)

    def setup(self):
        kw = _settings(self._p("input.nn"))
        g = lambda k, d: type(d)(kw[k][0]) if k in kw else d
        self.epochs = g("epochs", 10)
        self.seed = g("random_seed", 0)
        self.test_fraction = g("test_fraction", 0.1)
        self.use_forces = "use_short_forces" in kw
        self.force_weight = g("force_weight", 1.0)
        self.updater = g("updater_type", 0)
        self.force_fraction = g("short_force_fraction", 0.1)
        self.kalman = dict(epsilon=g("kalman_epsilon", 1e-2),
                           q0=g("kalman_q0", 0.01), qtau=g("kalman_qtau", 2.302),
                           qmin=g("kalman_qmin", 1e-6),
                           eta0=g("kalman_eta", 0.01),
                           etatau=g("kalman_etatau", 2.302),
                           etamax=g("kalman_etamax", 1.0))
        self.eta = g("gradient_adam_eta", 1e-3)
        self.beta1 = g("gradient_adam_beta1", 0.9)
        self.beta2 = g("gradient_adam_beta2", 0.999)
        self.eps = g("gradient_adam_epsilon", 1e-8)
        wmin, wmax = g("weights_min", -1.0), g("weights_max", 1.0)
        self.rng = np.random.default_rng(self.seed)
        fmt = self._p("weights.%03zu.data")
        probe = engine.Model.load(self._p("input.nn"), None, "")
        zs = [probe.atomic_number(e) for e in range(probe.num_elements)]
        have = all(os.path.exists(fmt.replace("%03zu", "%03d" % z))
                   for z in zs)
        self.model = engine.Model.load(self._p("input.nn"),
                                       self._p("scaling.data"),
                                       fmt if have else "")
        self.ctx = engine.Context(self.model, self.device)
        if not have:
            self.ctx.set_weights(self.rng.uniform(wmin, wmax,
                                                  self.ctx.num_weights()))
        self.zs = zs
        (self.normalize, self.mean_energy, self.conv_energy,
         self.conv_length) = self.model.normalization()
        if not self.normalize:
            self.mean_energy, self.conv_energy, self.conv_length = 0.0, 1.0, 1.0

    def read_data(self, name="input.data"):
        raw = nio.read_structures(self._p(name), self.model.element_index,
                                  with_forces=True)
        self.data = []
        for r, elem, box, energy, force in raw:
            off = sum(np.count_nonzero(elem == k) * self.model.energy_offset(k)
                      for k in range(self.model.num_elements))
            n = len(elem)
            if self.normalize:
                r, box = nio.to_normalized(r, box, self.conv_length)
            e_int = self.conv_energy * (energy - off - n * self.mean_energy)
            f_int = force * (self.conv_energy / self.conv_length)
            self.data.append((r, elem, box, e_int, f_int))
        idx = self.rng.permutation(len(self.data))
        ntest = int(round(self.test_fraction * len(self.data)))
        self.test = sorted(idx[:ntest].tolist())
        self.train_set = sorted(idx[ntest:].tolist())

    def errors(self, which):
        """(energy RMSE per atom, force RMSE) of a set, physical units."""
        se = sf = 0.0
        nf = 0
        for k in which:
            r, elem, box, e_int, f_int = self.data[k]
            self.ctx.set_structure(r, elem, box)
            e = self.ctx.evaluate(engine.FORCES_CONSISTENT)
            f = self.ctx.forces()
            n = len(elem)
            se += ((e - e_int) / self.conv_energy / n) ** 2
            sf += np.sum((f - f_int) ** 2) / (self.conv_energy
                                              / self.conv_length) ** 2
            nf += 3 * n
        return (np.sqrt(se / max(len(which), 1)),
                np.sqrt(sf / max(nf, 1)))

    def write_weights(self, epoch):
        w = self.ctx.get_weights()
        p = 0
        for e, z in enumerate(self.zs):
            sizes = self._layer_sizes(e)
            nw = sum(a * b + b for a, b in zip(sizes[:-1], sizes[1:]))
            out, q = [], p
            for a, b in zip(sizes[:-1], sizes[1:]):
                Wm = w[q:q + a * b].reshape(b, a)
                q += a * b
                out.extend(Wm.T.reshape(-1))
                out.extend(w[q:q + b])
                q += b
            p += nw
            with open(self._p("weights.%03d.%06d.out" % (z, epoch)), "w") as f:
                f.write("".join("%.16E\n" % v for v in out))

    def _layer_sizes(self, e):
        kw = _settings(self._p("input.nn"))
        hidden = [int(x) for x in kw["global_nodes_short"]]
        return [self.model.num_sf(e)] + hidden + [1]

    def _kalman_epoch(self):
        """One epoch of Kalman updates (energies, then a fraction of the
        force components), in random order."""
        ups = [("e", k) for k in self.train_set]
        if self.use_forces:
            comps = [(k, i, c) for k in self.train_set
                     for i in range(len(self.data[k][1])) for c in range(3)]
            nf = int(round(self.force_fraction * len(comps)))
            for q in self.rng.choice(len(comps), nf, replace=False):
                ups.append(("f",) + comps[q])
        for q in self.rng.permutation(len(ups)):
            u = ups[q]
            r, elem, box, e_int, f_int = self.data[u[1]]
            n = len(elem)
            self.ctx.set_structure(r, elem, box)
            if u[0] == "e":
                self.ctx.calculate_neighbor_list(self.model.max_cutoff)
                h = self.ctx.energy_weight_jacobian(file_order=False)[0] / n
                e = self.ctx.evaluate(-1)
                self.ctx.kalman_update(h, (e_int - e) / n)
            else:
                _, _, i, c = u
                self.ctx.evaluate(engine.FORCES_CONSISTENT)
                f = self.ctx.forces()
                h = _force_jacobian(self.ctx, f, i, c)
                self.ctx.kalman_update(h, f_int[i, c] - f[i, c])

    def run(self, epochs=None, log=print):
        epochs = self.epochs if epochs is None else epochs
        curve = []
        rows = []

        def record(ep):
            etr, ftr = self.errors(self.train_set)
            ete, fte = self.errors(self.test)
            curve.append((ep, etr, ete, ftr, fte))
            rows.append("%10d %16.8E %16.8E %16.8E %16.8E\n"
                        % (ep, etr, ete, ftr, fte))
            self.write_weights(ep)
            if log:
                log("epoch %d: RMSE E/atom train %.4e test %.4e, F train "
                    "%.4e test %.4e" % (ep, etr, ete, ftr, fte))

        record(0)
        self.ctx.train_reset()
        if self.updater == 1:
            k = self.kalman
            # Per-update annealing from per-epoch settings, as the
            # reference's tau parameters count epochs.
            nupd = max(len(self.train_set), 1)
            self.ctx.kalman_init(k["epsilon"], k["q0"], k["qtau"] / nupd,
                                 k["qmin"], k["eta0"], k["etatau"] / nupd,
                                 k["etamax"])
        wf = self.force_weight if self.use_forces else 0.0
        for ep in range(1, epochs + 1):
            if self.updater == 1:
                self._kalman_epoch()
            else:
                for k in self.rng.permutation(self.train_set):
                    r, elem, box, e_int, f_int = self.data[k]
                    self.ctx.set_structure(r, elem, box)
                    self.ctx.train_gradient(f_int, e_int, 1.0, wf,
                                            want_grad=False)
                    self.ctx.train_adam(self.eta, self.beta1, self.beta2,
                                        self.eps)
            record(ep)
        with open(self._p("learning-curve.out"), "w") as f:
            f.write("# Learning curve (GPU Adam training): epoch, RMSE of "
                    "the energy per atom (train, test) and of the forces "
                    "(train, test), physical units\n")
            f.writelines(rows)
        return curve


def _force_jacobian(ctx, f, i, c):
    """dF_ic/dw from the loss gradient: with Fref = F - e_ic (3N / 2) the
    gradient of 1/(3N) sum |F - Fref|^2 is exactly dF_ic/dw."""
    n = len(f)
    fref = f.copy()
    fref[i, c] -= 1.5 * n
    g, _ = ctx.train_gradient(fref, 0.0, 0.0, 1.0)
    return g




def main():
    t = Training(".")
    t.setup()
    t.read_data()
    t.run()
    return 0


if __name__ == "__main__":
    sys.exit(main())
