"""Prediction: the nnp-predict flow on the HIP engine.

Mirrors nnp::Prediction (Prediction.cpp:34-76): setup() loads input.nn,
scaling.data and weights.%03zu.data; read_structure_from_file() reads the
first configuration of input.data (with remap) and converts it to the
model's normalised units; predict() runs Mode::evaluateNNP and converts
energy and forces back to physical units and adds the atomic energy offsets
(Mode::addEnergyOffset, Mode.cpp:2018-2035).
"""

import os

import numpy as np

from n2p2_amd import io as nio
from n2p2_amd.engine import Context, Model, FORCES_REFERENCE


class Prediction:
    def __init__(self, directory=".", device=0,
                 settings="input.nn", scaling="scaling.data",
                 weights="weights.%03zu.data"):
        self.dir = directory
        self.device = device
        self.file_settings = settings
        self.file_scaling = scaling
        self.format_weights = weights
        self.model = None
        self.ctx = None

    def setup(self):
        p = lambda f: os.path.join(self.dir, f)
        self.model = Model.load(p(self.file_settings), p(self.file_scaling),
                                p(self.format_weights))
        self.ctx = Context(self.model, self.device)
        (self.normalize, self.mean_energy, self.conv_energy,
         self.conv_length) = self.model.normalization()

    def read_structure_from_file(self, name="input.data"):
        r, elem, box, _ = nio.read_structure(os.path.join(self.dir, name),
                                             self.model.element_index)
        self.r_phys, self.elem, self.box_phys = r, elem, box
        if self.normalize:
            r, box = nio.to_normalized(r, box, self.conv_length)
        self.ctx.set_structure(r, elem, box)

    def predict(self, force_mode=FORCES_REFERENCE):
        e = self.ctx.evaluate(force_mode)
        f = self.ctx.forces() if force_mode >= 0 else None
        n = len(self.elem)
        if self.normalize:
            e = e / self.conv_energy + n * self.mean_energy
            if f is not None:
                f = f / (self.conv_energy / self.conv_length)
        for k in range(self.model.num_elements):
            e += np.count_nonzero(self.elem == k) * self.model.energy_offset(k)
        self.energy = e
        self.forces = f
        return e, f
