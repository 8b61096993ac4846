"""Figure helpers shared by the demonstrations (matplotlib optional)."""

from __future__ import annotations

import os


def pyplot():
    """matplotlib.pyplot with the Agg backend, or None when it is absent."""
    try:
        import matplotlib
    except ImportError:
        return None
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    return plt


def save(fig, out_dir, name):
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, name)
    fig.savefig(path, dpi=130)
    return path


def velocity_panels(plt, panels, extent, vmin, vmax, history, marks=None, title_hist="Misfit convergence"):
    """One row of sound-speed images and the normalised misfit curve."""
    import numpy as np
    fig, ax = plt.subplots(1, len(panels) + 1, figsize=(4.2 * (len(panels) + 1), 4.2))
    for a, (field, title) in zip(ax, panels):
        im = a.imshow(field, origin="lower", extent=extent, cmap="viridis", vmin=vmin, vmax=vmax)
        if marks is not None:
            a.plot(marks[:, 0] * 1e3, marks[:, 1] * 1e3, "w.", ms=5)
        a.set_title(title)
        a.set_xlabel("x (mm)")
        a.set_ylabel("y (mm)")
        fig.colorbar(im, ax=a, fraction=0.046, label="c (m/s)")
    h = np.asarray(history, dtype=float)
    ax[-1].semilogy(h / h[0], "o-")
    ax[-1].set_title(title_hist)
    ax[-1].set_xlabel("iteration")
    ax[-1].set_ylabel("normalised misfit")
    ax[-1].grid(True, which="both", alpha=0.3)
    fig.tight_layout()
    return fig


def snapshots(plt, frames, extent, marks, dt):
    """Pressure frames side by side on one shared, symmetric colour scale
    (each frame divided by the largest |p| over all frames), the elements as
    open circles and one colour bar for the row.  frames: {step: 2D array}."""
    import numpy as np
    steps = sorted(frames)
    peak = max(float(np.abs(frames[s]).max()) for s in steps) or 1.0
    fig, ax = plt.subplots(1, len(steps), figsize=(4.0 * len(steps) + 1.0, 4.0), squeeze=False)
    im = None
    for a, s in zip(ax[0], steps):
        im = a.imshow(frames[s] / peak, origin="lower", extent=extent, cmap="RdBu_r", vmin=-1.0,
                      vmax=1.0)
        a.scatter(marks[:, 0] * 1e3, marks[:, 1] * 1e3, s=14, facecolors="none", edgecolors="k",
                  linewidths=0.6)
        a.set_title(f"step {s}  ({s * dt * 1e6:.2f} us)")
        a.set_xlabel("x (mm)")
    ax[0, 0].set_ylabel("y (mm)")
    fig.colorbar(im, ax=list(ax[0]), shrink=0.85, label="p / max |p|")
    return fig
