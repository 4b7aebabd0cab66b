"""jaxhip: the per-time-step hot path of jaxley-pytorch's `jx.integrate` on MI355X.

Drop-in surface (`import jaxhip as jx`): `Compartment`, `Branch`, `Cell`, `Network`,
`connect`, `fully_connect`, `read_swc`, `integrate`, `build_init_and_step_fn`, `step_current`, `datapoint_to_step_currents`; channels `HH`,
`Leak`; synapses `KineticSynapse`, `SigmoidRateSynapse`, `IonotropicSynapse`, `TestSynapse`; the GPU tree-solve operator
`jaxhip.kernels.tree_solve.make_tree_solve_kernel`; the synapse operator
`jaxhip.kernels.synapse_step`. Kernels are hand-written HIP for gfx950 in
`libjxhip.so` (C ABI: include/jxhip.h).
"""
from .channels import HH, K, Km, Leak, Na
from .integrate import build_init_and_step_fn, integrate
from .modules import Branch, Cell, Compartment, Network, connect, fully_connect
from .stimulus import datapoint_to_step_currents, step_current
from .swc import read_swc
from .synapses import IonotropicSynapse, KineticSynapse, SigmoidRateSynapse, TestSynapse

__version__ = "0.1.0"

__all__ = [
    "Branch", "Cell", "Compartment", "Network", "connect", "fully_connect", "read_swc", "integrate", "build_init_and_step_fn",
    "step_current",
    "datapoint_to_step_currents", "HH", "Leak", "Na", "K", "Km", "KineticSynapse", "SigmoidRateSynapse", "IonotropicSynapse",
    "TestSynapse",
]
