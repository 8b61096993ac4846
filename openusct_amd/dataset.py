"""Acquisition dataset and its HDF5 file (simulation/ringfwi/dataset.py:1-174;
layout software/schema/README.md:13-41).

:class:`Dataset` and :class:`ArrayGeometry` carry the same fields, defaults
and derived properties as the reference's.  ``save`` / ``load`` go through the
native reader/writer ``libusct_h5.so`` instead of h5py and produce files h5py
reads unchanged (and read the files h5py writes).  The file layout is the
schema's: root attributes ``format_version``, ``description``,
``created_by``; ``/geometry`` (``n_elements``, ``radius_m``, ``centre_freq_hz``,
``array_type``, dataset ``element_pos``); ``/acquisition``
(``sample_rate_hz``, ``tx_centre_freq_hz``, ``nominal_speed_m_s``, datasets
``data`` float32 gzip-compressed, ``tx_wavelet``, ``tx_elements``); optional
``/ground_truth`` (``h_m``, dataset ``c`` gzip-compressed).

The channel data is chunked by transmit rows, so :func:`read_fmc` can stream
transmits into pinned host buffers and on to the device, inflating the next
block while the previous one is copied.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import h5file

FORMAT_VERSION = "1.0"

# Target stored chunk size before compression.
_CHUNK_BYTES = 1 << 20
_GZIP_LEVEL = 4  # h5py's level for compression="gzip"


@dataclass
class ArrayGeometry:
    """Element positions in metres (``(n_elements, dim)``, origin at the array
    centre), array radius, centre frequency, and "ring" (2D) or "cylinder" (3D)."""

    element_pos: np.ndarray
    radius_m: float
    centre_freq_hz: float
    array_type: str = "ring"

    @property
    def n_elements(self):
        return self.element_pos.shape[0]

    @property
    def dim(self):
        return self.element_pos.shape[1]

    @classmethod
    def ring(cls, n_elements, radius_m, centre_freq_hz):
        """Planar ring of equally spaced elements, the first on the +x axis."""
        theta = np.linspace(0.0, 2.0 * np.pi, n_elements, endpoint=False)
        pos = np.stack([radius_m * np.cos(theta), radius_m * np.sin(theta)], axis=1)
        return cls(pos, radius_m, centre_freq_hz, "ring")

    @classmethod
    def cylinder(cls, n_rings, per_ring, radius_m, height_m, centre_freq_hz):
        """``n_rings`` rings of ``per_ring`` elements stacked along z over
        ``height_m`` (one ring: at z = 0); ring-major element order."""
        z = np.linspace(-height_m / 2, height_m / 2, n_rings) if n_rings > 1 else np.zeros(1)
        theta = np.linspace(0.0, 2.0 * np.pi, per_ring, endpoint=False)
        zz = np.repeat(z, per_ring)
        tt = np.tile(theta, n_rings)
        pos = np.stack([radius_m * np.cos(tt), radius_m * np.sin(tt), zz], axis=1)
        return cls(pos, radius_m, centre_freq_hz, "cylinder")


@dataclass
class Dataset:
    """Full-matrix capture ``data[i, :, j]`` (transmit ``i``, receiver ``j``)
    with its geometry, timing and optional simulation ground truth
    ``{"c": sound speed, "h_m": grid spacing}``."""

    geometry: ArrayGeometry
    data: np.ndarray
    sample_rate_hz: float
    tx_wavelet: np.ndarray
    tx_centre_freq_hz: float
    nominal_speed_m_s: float
    tx_elements: np.ndarray | None = None
    ground_truth: dict | None = None
    metadata: dict = field(default_factory=dict)

    def __post_init__(self):
        n_tx = self.data.shape[0]
        self.tx_elements = np.arange(n_tx) if self.tx_elements is None else np.asarray(self.tx_elements)

    @property
    def n_tx(self):
        return self.data.shape[0]

    @property
    def n_samples(self):
        return self.data.shape[1]

    @property
    def n_rx(self):
        return self.data.shape[2]

    @property
    def dt(self):
        return 1.0 / self.sample_rate_hz

    def save(self, path, threads=0):
        """Write the HDF5 file (float32 channel data, gzip-compressed in
        chunks of whole transmits; ``threads`` compress chunks in parallel)."""
        meta = self.metadata
        geo = self.geometry
        data = np.asarray(self.data, dtype=np.float32)
        with h5file.Writer(path) as w:
            for name, value in (("format_version", FORMAT_VERSION),
                                ("description", meta.get("description", "")),
                                ("created_by", meta.get("created_by", "OpenUSCT"))):
                w.attr("/", name, str(value))
            w.group("/geometry")
            w.attr("/geometry", "n_elements", int(geo.n_elements))
            w.attr("/geometry", "radius_m", _scalar(geo.radius_m))
            w.attr("/geometry", "centre_freq_hz", _scalar(geo.centre_freq_hz))
            w.attr("/geometry", "array_type", str(geo.array_type))
            w.dataset("/geometry/element_pos", np.asarray(geo.element_pos))
            w.group("/acquisition")
            for name in ("sample_rate_hz", "tx_centre_freq_hz", "nominal_speed_m_s"):
                w.attr("/acquisition", name, _scalar(getattr(self, name)))
            w.dataset("/acquisition/data", data, chunks=_row_chunks(data.shape, 4), gzip=_GZIP_LEVEL,
                      threads=threads)
            w.dataset("/acquisition/tx_wavelet", np.asarray(self.tx_wavelet))
            w.dataset("/acquisition/tx_elements", np.asarray(self.tx_elements))
            if self.ground_truth is not None:
                c = np.asarray(self.ground_truth["c"])
                w.group("/ground_truth")
                w.attr("/ground_truth", "h_m", _scalar(self.ground_truth["h_m"]))
                w.dataset("/ground_truth/c", c, chunks=_row_chunks(c.shape, c.itemsize), gzip=_GZIP_LEVEL,
                          threads=threads)

    @classmethod
    def load(cls, path, threads=0):
        """Read a file written by :meth:`save` or by the reference's
        ``Dataset.save``.  The channel data keeps its on-disk float32; a file
        whose ``format_version`` has a major version other than 1 is refused."""
        with h5file.Reader(path) as r:
            version = str(r.attr("/", "format_version")) if "format_version" in r.attr_names("/") else ""
            if version and version.split(".")[0] != FORMAT_VERSION.split(".")[0]:
                raise ValueError(f"{path}: dataset format version {version} is not readable "
                                 f"(this reader knows {FORMAT_VERSION})")
            geom = ArrayGeometry(element_pos=r.read("/geometry/element_pos", threads=threads),
                                 radius_m=float(r.attr("/geometry", "radius_m")),
                                 centre_freq_hz=float(r.attr("/geometry", "centre_freq_hz")),
                                 array_type=str(r.attr("/geometry", "array_type")))
            truth = None
            if r.kind("/ground_truth") == "group":
                truth = {"c": r.read("/ground_truth/c", threads=threads),
                         "h_m": float(r.attr("/ground_truth", "h_m"))}
            tx = r.read("/acquisition/tx_elements") if r.kind("/acquisition/tx_elements") else None
            root = r.attr_names("/")
            meta = {k: (str(r.attr("/", k)) if k in root else "")
                    for k in ("description", "created_by", "format_version")}
            return cls(geometry=geom,
                       data=r.read("/acquisition/data", threads=threads),
                       sample_rate_hz=float(r.attr("/acquisition", "sample_rate_hz")),
                       tx_wavelet=r.read("/acquisition/tx_wavelet"),
                       tx_centre_freq_hz=float(r.attr("/acquisition", "tx_centre_freq_hz")),
                       nominal_speed_m_s=float(r.attr("/acquisition", "nominal_speed_m_s")),
                       tx_elements=tx, ground_truth=truth, metadata=meta)


def _scalar(v):
    """Python numbers as h5py stores them: int -> int64, float -> float64;
    NumPy scalars keep their type."""
    if isinstance(v, np.generic):
        return v
    if isinstance(v, (bool, int)):
        return int(v)
    return float(v)


def _row_chunks(shape, itemsize):
    """Chunks of whole leading-axis rows, about 1 MiB each."""
    if not shape or 0 in shape:
        return None if not shape else tuple(max(1, s) for s in shape)
    row = int(np.prod(shape[1:], dtype=np.int64)) * itemsize
    rows = max(1, min(shape[0], -(-_CHUNK_BYTES // max(row, 1))))
    return (rows,) + tuple(shape[1:])


def read_fmc(path, device="cuda", dtype=None, rows=None, threads=0, block_bytes=64 << 20):
    """The channel data of a dataset file as a device tensor.

    Transmit rows are inflated by the native reader into one of two pinned
    host buffers while the other is copied to the device on a side stream, so
    decompression and the host-to-device copy overlap.  ``rows`` = (first,
    stop) selects transmits; ``dtype`` (default torch.float64, the solver's
    trace type) is applied on the device.
    """
    import torch

    dtype = torch.float64 if dtype is None else dtype
    dev = torch.device(device)
    with h5file.Reader(path) as r:
        shape, fdt, _ = r.info("/acquisition/data")
        if len(shape) != 3:
            raise ValueError(f"{path}: /acquisition/data has shape {shape}, expected (n_tx, n_samples, n_rx)")
        r0, r1 = (0, shape[0]) if rows is None else rows
        if not (0 <= r0 <= r1 <= shape[0]):
            raise ValueError(f"rows {rows} outside 0..{shape[0]}")
        per_row = shape[1] * shape[2]
        staged = torch.float32 if fdt == np.float32 else torch.float64
        np_staged = np.float32 if staged == torch.float32 else np.float64
        out = torch.empty((r1 - r0, shape[1], shape[2]), dtype=staged, device=dev)
        if r1 == r0 or per_row == 0:
            return out.to(dtype)
        blk = max(1, min(r1 - r0, block_bytes // max(per_row * out.element_size(), 1)))
        bufs = [torch.empty((blk, shape[1], shape[2]), dtype=staged, pin_memory=True) for _ in range(2)]
        done = [None, None]
        stream = torch.cuda.Stream(device=dev)
        for k, a in enumerate(range(r0, r1, blk)):
            b = min(r1, a + blk)
            buf = bufs[k % 2]
            if done[k % 2] is not None:
                done[k % 2].synchronize()  # the copy out of this buffer has finished
            r.read_into("/acquisition/data", buf.data_ptr(), (b - a) * per_row, np_staged, a, b - a, threads)
            with torch.cuda.stream(stream):
                out[a - r0:b - r0].copy_(buf[:b - a], non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(stream)
            done[k % 2] = ev
        torch.cuda.current_stream(dev).wait_stream(stream)
        out.record_stream(stream)
        return out if out.dtype == dtype else out.to(dtype)
