"""MI355X HDNNP energy-and-force engine (n2p2 Mode API drop-in).

The compute path is the HIP library ``libhdnnp.so`` built from
``n2p2_amd/csrc`` (``make -C n2p2_amd``); its C ABI is ``include/hdnnp.h``.
This package binds that ABI with ctypes and mirrors the reference's
``nnp::Mode`` call sequence (``calculateNeighborList`` ->
``calculateSymmetryFunctionGroups`` -> ``calculateAtomicNeuralNetworks`` ->
``calculateEnergy`` -> ``calculateForces``).  There is no CPU fallback:
importing ``n2p2_amd.engine`` fails loudly if the library is missing.
"""

__all__ = ["engine", "workloads"]
