"""openusct_amd: MI355X-native acoustic FD solver and exact discrete-adjoint FWI
gradient, a drop-in for the acoustic path of OpenUSCT.

Modules
-------
fwi        drop-in ``forward_fmc`` / ``misfit`` / ``misfit_and_gradient`` /
           ``simulate_wavefield`` / ``invert`` / ``fwi_from_dataset``
plan       reusable device plans over torch tensors (C ABI: include/usct.h)
autograd   ``torch.autograd.Function`` sound speed -> misfit
serve      continuous-batching request loop
plugins    name -> algorithm registry with the ``"fwi"`` plugin
geometry, sources, phantom, solver   host-side inputs (arrays, wavelet, models, sponge)
"""

__version__ = "0.1.0"

from . import geometry, phantom, plugins, solver, sources  # noqa: F401
