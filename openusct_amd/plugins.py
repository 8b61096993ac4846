"""Algorithm registry (simulation/ringfwi/plugins.py:28-66).

Same contract as the reference: ``register(name, description, output)`` is a
decorator, ``run(name, dataset, **params)`` calls the algorithm, unknown names
raise ``KeyError``.  The built-in ``"fwi"`` runs
:func:`openusct_amd.fwi.fwi_from_dataset` with every misfit and gradient on the
device; a ``backend=`` parameter is forwarded to ``invert`` as in the reference.
"""

from __future__ import annotations

_REGISTRY = {}


def register(name, description="", output="image"):
    def deco(fn):
        _REGISTRY[name] = {"fn": fn, "description": description, "output": output}
        return fn
    return deco


def available():
    return {k: v["description"] for k, v in _REGISTRY.items()}


def info(name):
    return _REGISTRY[name]


def run(name, dataset, **params):
    if name not in _REGISTRY:
        raise KeyError(f"unknown algorithm {name!r}; available: {sorted(_REGISTRY)}")
    return _REGISTRY[name]["fn"](dataset, **params)


@register("fwi", description="Adjoint-state full waveform inversion on the MI355X solver",
          output="model")
def _fwi(dataset, **params):
    from .fwi import fwi_from_dataset
    return fwi_from_dataset(dataset, **params)
