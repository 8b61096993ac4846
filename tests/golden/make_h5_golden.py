"""Write the HDF5 fixtures of tests/golden/ (dataset files written by libhdf5).

Run with a Python that has h5py, in a container that holds the reference
checkout (never on a GPU host):

    /opt/conda/bin/python3.9 tests/golden/make_h5_golden.py [/path/to/reference]

* ``ds_ring.h5``, ``ds_cyl.h5``, ``ds_notruth.h5``: written by the reference's
  own ``Dataset.save`` (simulation/ringfwi/dataset.py:112-139), loaded from its
  file alone (it imports only numpy and h5py), on synthetic acquisitions: a 2D
  ring with ground truth and a transmit subset, a 3D cylinder with empty
  metadata, and a 2D ring without ground truth, default transmits and a
  non-ASCII description.  ``ArrayGeometry.ring`` / ``.cylinder``
  (dataset.py:47-61) give the element positions, so the fixtures pin those too.
* ``h5_features.h5``: plain h5py (libhdf5 1.10, default format) exercising what
  a reader of such files meets: shuffle + Fletcher-32 + gzip, big-endian and
  narrow integer / half-precision types, compact layout, unwritten chunks read
  as the fill value, partial edge chunks, a chunk index of several B-tree
  levels, empty and scalar datasets, nested groups, fixed-length and
  variable-length string attributes, array attributes.
* ``h5_latest.h5``: libver="latest" (superblock 3, version-2 object headers,
  link messages); contiguous datasets read, its chunked dataset uses an index
  the native reader does not implement.

``h5_golden.npz`` holds what each file must read back as (data only, no pickles).
"""

from __future__ import annotations

import importlib.util
import os
import sys

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _ref_dataset(ref_root):
    path = os.path.join(ref_root, "simulation", "ringfwi", "dataset.py")
    spec = importlib.util.spec_from_file_location("ref_ringfwi_dataset", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod
    spec.loader.exec_module(mod)
    return mod


def reference_files(ref_root, out):
    D = _ref_dataset(ref_root)
    rng = np.random.default_rng(21)

    # 2D ring, ground truth, transmit subset
    g = D.ArrayGeometry.ring(8, 0.024, 4.0e5)
    data = rng.standard_normal((3, 300, 8)) * 1e-3
    wav = np.sin(np.arange(300) * 0.05) * np.exp(-((np.arange(300) - 60.0) / 20.0) ** 2)
    c = 1480.0 + 40.0 * rng.standard_normal((61, 61))
    ds = D.Dataset(geometry=g, data=data, sample_rate_hz=1.0 / 1.2e-7, tx_wavelet=wav,
                   tx_centre_freq_hz=4.0e5, nominal_speed_m_s=1480.0, tx_elements=[0, 2, 4],
                   ground_truth={"c": c, "h_m": 1.0e-3},
                   metadata={"description": "roundtrip test", "created_by": "OpenUSCT simulator"})
    ds.save(os.path.join(HERE, "ds_ring.h5"))
    out.update(ring_pos=g.element_pos, ring_data=data.astype(np.float32), ring_wav=wav, ring_c=c)

    # 3D cylinder, empty metadata
    g = D.ArrayGeometry.cylinder(2, 6, 0.012, 0.010, 3.0e5)
    data = rng.standard_normal((2, 120, 12))
    wav = rng.standard_normal(120)
    c = 1500.0 + rng.standard_normal((16, 16, 16))
    ds = D.Dataset(geometry=g, data=data, sample_rate_hz=5.0e6, tx_wavelet=wav, tx_centre_freq_hz=3.0e5,
                   nominal_speed_m_s=1480.0, tx_elements=np.array([0, 6]),
                   ground_truth={"c": c, "h_m": 2.0e-3})
    ds.save(os.path.join(HERE, "ds_cyl.h5"))
    out.update(cyl_pos=g.element_pos, cyl_data=data.astype(np.float32), cyl_wav=wav, cyl_c=c)
    out.update(cyl1_pos=D.ArrayGeometry.cylinder(1, 5, 0.01, 0.02, 1e6).element_pos)

    # 2D ring, no ground truth, default transmits, several chunks per axis
    g = D.ArrayGeometry.ring(16, 0.03, 5.0e5)
    data = np.zeros((16, 400, 16))
    data[:, 50:150, :] = rng.standard_normal((16, 100, 16))
    wav = rng.standard_normal(400)
    ds = D.Dataset(geometry=g, data=data, sample_rate_hz=2.5e7, tx_wavelet=wav, tx_centre_freq_hz=5.0e5,
                   nominal_speed_m_s=1500.0,
                   metadata={"description": "phantom Ω, 25 µm", "created_by": "bench rig"})
    ds.save(os.path.join(HERE, "ds_notruth.h5"))
    out.update(nt_pos=g.element_pos, nt_data=data.astype(np.float32), nt_wav=wav)
    with h5py.File(os.path.join(HERE, "ds_notruth.h5"), "r") as f:
        out.update(nt_chunks=np.array(f["acquisition/data"].chunks))


def feature_file(out):
    rng = np.random.default_rng(22)
    path = os.path.join(HERE, "h5_features.h5")
    with h5py.File(path, "w") as f:
        a = rng.standard_normal((20, 33))
        f.create_dataset("shuffle_fletcher", data=a, dtype=">f8", chunks=(7, 10), shuffle=True,
                         fletcher32=True, compression="gzip", compression_opts=9)
        out["shuffle_fletcher"] = a
        i16 = rng.integers(-30000, 30000, (5, 4)).astype(">i2")
        f.create_dataset("i16_be", data=i16)
        out["i16_be"] = i16.astype(np.int16)
        u32 = np.array([0, 1, 4000000000], dtype=np.uint32)
        f.create_dataset("u32", data=u32)
        out["u32"] = u32
        i8 = np.array([-128, -1, 0, 127], dtype=np.int8)
        f.create_dataset("i8", data=i8, chunks=(3,), compression="gzip")
        out["i8"] = i8
        f16 = np.array([0.0, -0.0, 1.5, 65504.0, 6.1e-5, 5.96e-8, np.inf, -2.0], dtype=np.float16)
        f.create_dataset("f16", data=f16)
        out["f16"] = f16
        # compact layout through the low-level API
        comp = rng.standard_normal((4, 3)).astype(np.float32)
        dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
        dcpl.set_layout(h5py.h5d.COMPACT)
        sid = h5py.h5s.create_simple(comp.shape)
        did = h5py.h5d.create(f.id, b"compact", h5py.h5t.IEEE_F32LE, sid, dcpl=dcpl)
        did.write(h5py.h5s.ALL, h5py.h5s.ALL, comp)
        out["compact"] = comp
        # unwritten chunks read as the fill value
        sp = f.create_dataset("sparse", shape=(10, 10), dtype=np.float32, chunks=(3, 3), fillvalue=-2.5)
        blk = rng.standard_normal((3, 3)).astype(np.float32)
        sp[0:3, 0:3] = blk
        sp[6:9, 3:6] = blk * 2
        e = np.full((10, 10), -2.5, np.float32)
        e[0:3, 0:3] = blk
        e[6:9, 3:6] = blk * 2
        out["sparse"] = e
        # partial edge chunks and a chunk index of several B-tree levels
        many = np.zeros((17, 403, 66), np.float32)
        many[::4, ::7, ::5] = rng.standard_normal(many[::4, ::7, ::5].shape).astype(np.float32)
        f.create_dataset("many", data=many, chunks=(2, 47, 10), compression="gzip", compression_opts=1)
        out["many"] = many
        f.create_dataset("empty", data=np.zeros((0, 4)))
        f.create_dataset("scalar", data=3.25)
        out["scalar"] = np.float64(3.25)
        nest = np.arange(4, dtype=np.int64).reshape(2, 2)
        f.create_dataset("nested/a/b/deep", data=nest)
        out["deep"] = nest
        f.attrs["fixed"] = np.bytes_(b"fixed-len")
        f.attrs["vlen"] = "variable é"
        f.attrs["empty"] = ""
        f.attrs["ints"] = np.array([1, 2, 3], dtype=np.int32)
        f.attrs["f32"] = np.float32(0.1)
        f.attrs["one"] = np.array([7.5])
        f["nested"].attrs["level"] = 1
        # a group with more members than one symbol node holds
        g = f.create_group("wide")
        for k in range(40):
            g.create_dataset(f"m{k:02d}", data=np.array([k], dtype=np.int64))


def latest_file(out):
    path = os.path.join(HERE, "h5_latest.h5")
    with h5py.File(path, "w", libver="latest") as f:
        x = np.arange(12.0).reshape(3, 4)
        f.create_dataset("grp/x", data=x)
        f["grp"].attrs["note"] = "v2 headers"
        f.attrs["n"] = 5
        f.create_dataset("chunked", data=np.ones((8, 8), np.float32), chunks=(4, 4))
        out["latest_x"] = x


def main(ref_root):
    out = {}
    reference_files(ref_root, out)
    feature_file(out)
    latest_file(out)
    out["hdf5_version"] = np.array(h5py.version.hdf5_version)
    np.savez_compressed(os.path.join(HERE, "h5_golden.npz"), **out)
    for name in ("ds_ring.h5", "ds_cyl.h5", "ds_notruth.h5", "h5_features.h5", "h5_latest.h5"):
        print(name, os.path.getsize(os.path.join(HERE, name)), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
