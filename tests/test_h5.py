"""Acquisition-dataset files (CPU only): the native HDF5 reader/writer
(libusct_h5.so) against files written by libhdf5.

Pinning: ds_*.h5 were written by the reference's own ``Dataset.save``
(simulation/ringfwi/dataset.py:112-139, through h5py / libhdf5 1.10.6) and
h5_features.h5 / h5_latest.h5 by plain h5py (tests/golden/make_h5_golden.py);
h5_golden.npz holds what they must read back as.  The reverse direction --
files this library writes, read by libhdf5 -- runs where an h5py is installed
(the image's /opt/conda Python 3.9), in a subprocess.  Reads are bit-exact
(the channel data is float32 on disk, as the reference stores it).  The
damaged-file test runs the reader under AddressSanitizer / UBSan.
"""

from __future__ import annotations

import ctypes
import glob
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from helpers import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CONDA_PY = "/opt/conda/bin/python3.9"


def _g(name):
    return os.path.join(GOLD, name)


@pytest.fixture(scope="module")
def expect():
    return golden("h5_golden.npz")


def _have_h5py():
    if not os.path.exists(CONDA_PY):
        return False
    r = subprocess.run([CONDA_PY, "-c", "import h5py"], capture_output=True)
    return r.returncode == 0


def test_library_exports_every_declared_symbol():
    from openusct_amd import h5file
    hdr = open(os.path.join(ROOT, "include", "usct_h5.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(usct_h5_\w+)\(", hdr, re.M))
    assert declared == set(h5file.EXPORTS)
    lib = ctypes.CDLL(h5file.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


@pytest.mark.parametrize("name,key,dim,atype", [("ds_ring.h5", "ring", 2, "ring"),
                                                ("ds_cyl.h5", "cyl", 3, "cylinder")])
def test_load_reference_saved_dataset(expect, name, key, dim, atype):
    """Dataset.load of files the reference's Dataset.save wrote: every field
    as the reference's load returns it (test_dataset.py:43-52 checks a subset)."""
    from openusct_amd.dataset import Dataset
    ds = Dataset.load(_g(name))
    assert ds.data.dtype == np.float32 and np.array_equal(ds.data, expect[f"{key}_data"])
    assert np.array_equal(ds.geometry.element_pos, expect[f"{key}_pos"])
    assert ds.geometry.dim == dim and ds.geometry.array_type == atype
    assert np.array_equal(ds.tx_wavelet, expect[f"{key}_wav"])
    assert np.array_equal(ds.ground_truth["c"], expect[f"{key}_c"])
    if key == "ring":
        assert ds.sample_rate_hz == 1.0 / 1.2e-7 and ds.dt == 1.0 / (1.0 / 1.2e-7)
        assert ds.tx_centre_freq_hz == 4.0e5 and ds.nominal_speed_m_s == 1480.0
        assert list(ds.tx_elements) == [0, 2, 4] and ds.ground_truth["h_m"] == 1.0e-3
        assert ds.geometry.radius_m == 0.024 and ds.geometry.centre_freq_hz == 4.0e5
        assert ds.metadata == {"description": "roundtrip test", "created_by": "OpenUSCT simulator",
                               "format_version": "1.0"}
    else:
        assert list(ds.tx_elements) == [0, 6] and ds.ground_truth["h_m"] == 2.0e-3
        assert ds.metadata == {"description": "", "created_by": "OpenUSCT", "format_version": "1.0"}
    assert (ds.n_tx, ds.n_samples, ds.n_rx) == ds.data.shape


def test_load_reference_saved_dataset_without_truth(expect):
    from openusct_amd.dataset import Dataset
    ds = Dataset.load(_g("ds_notruth.h5"))
    assert ds.ground_truth is None
    assert np.array_equal(ds.data, expect["nt_data"]) and np.array_equal(ds.tx_wavelet, expect["nt_wav"])
    assert np.array_equal(ds.tx_elements, np.arange(16))
    assert ds.metadata["description"] == "phantom Ω, 25 µm" and ds.metadata["created_by"] == "bench rig"
    assert tuple(expect["nt_chunks"]) != tuple(ds.data.shape)  # h5py's own chunking: several per axis


def test_array_geometry_constructors_match_reference(expect):
    from openusct_amd.dataset import ArrayGeometry
    assert np.array_equal(ArrayGeometry.ring(8, 0.024, 4.0e5).element_pos, expect["ring_pos"])
    assert np.array_equal(ArrayGeometry.ring(16, 0.03, 5.0e5).element_pos, expect["nt_pos"])
    assert np.array_equal(ArrayGeometry.cylinder(2, 6, 0.012, 0.010, 3.0e5).element_pos, expect["cyl_pos"])
    assert np.array_equal(ArrayGeometry.cylinder(1, 5, 0.01, 0.02, 1e6).element_pos, expect["cyl1_pos"])


def test_reads_libhdf5_features(expect):
    from openusct_amd.h5file import Reader
    with Reader(_g("h5_features.h5")) as r:
        for k in ("shuffle_fletcher", "i16_be", "u32", "i8", "f16", "compact", "sparse", "many"):
            v = r.read("/" + k)
            assert v.dtype == expect[k].dtype and np.array_equal(v, expect[k]), k
        assert r.info("/many")[2] and not r.info("/i16_be")[2]
        assert r.read("/scalar") == 3.25 and np.ndim(r.read("/scalar")) == 0
        assert r.read("/empty").shape == (0, 4)
        assert np.array_equal(r.read("/nested/a/b/deep"), expect["deep"])
        assert r.kind("/nested/a") == "group" and r.kind("/nested/a/b/deep") == "dataset"
        assert r.kind("/nope") is None and "/nested/a" in r
        wide = r.members("/wide")
        assert wide == [f"m{k:02d}" for k in range(40)]
        assert all(r.read(f"/wide/{m}")[0] == int(m[1:]) for m in wide)
        a = r.attrs("/")
        assert a["fixed"] == "fixed-len" and a["vlen"] == "variable é" and a["empty"] == ""
        assert a["ints"].dtype == np.int32 and list(a["ints"]) == [1, 2, 3]
        assert a["f32"].dtype == np.float32 and a["f32"] == np.float32(0.1) and np.ndim(a["f32"]) == 0
        assert a["one"].shape == (1,) and a["one"][0] == 7.5
        assert r.attr("/nested", "level") == 1
        with pytest.raises(KeyError):
            r.attr("/", "missing")
        with pytest.raises(KeyError):
            r.read("/missing")


def test_row_selection_and_conversion(expect):
    """Rows of chunked data (partial edge chunks, chunks straddling the
    selection) and float32 -> float64 / int conversions are exact."""
    from openusct_amd.h5file import Reader
    many = expect["many"]
    with Reader(_g("h5_features.h5")) as r:
        for a, b in ((0, 17), (0, 1), (3, 8), (16, 17), (5, 5)):
            v = r.read("/many", rows=(a, b))
            assert np.array_equal(v, many[a:b]), (a, b)
        assert np.array_equal(r.read("/many", rows=(1, 4), dtype=np.float64), many[1:4].astype(np.float64))
        assert np.array_equal(r.read("/i16_be", dtype=np.int64), expect["i16_be"].astype(np.int64))
        assert np.array_equal(r.read("/f16", dtype=np.float32), expect["f16"].astype(np.float32))
        assert np.array_equal(r.read("/sparse", rows=(2, 7)), expect["sparse"][2:7])
        with pytest.raises(ValueError):
            r.read("/many", rows=(10, 20))
    with Reader(_g("ds_notruth.h5")) as r:
        for a, b in ((0, 16), (3, 11), (15, 16)):
            assert np.array_equal(r.read("/acquisition/data", rows=(a, b)), expect["nt_data"][a:b])
        for th in (1, 3, 8):
            assert np.array_equal(r.read("/acquisition/data", threads=th), expect["nt_data"])


def test_latest_format_file(expect):
    """Superblock 3, version-2 object headers and link messages read; a chunk
    index the reader does not implement is reported, not misread."""
    from openusct_amd.h5file import Reader
    with Reader(_g("h5_latest.h5")) as r:
        assert sorted(r.members("/")) == ["chunked", "grp"]
        assert np.array_equal(r.read("/grp/x"), expect["latest_x"])
        assert r.attr("/grp", "note") == "v2 headers" and r.attr("/", "n") == 5
        with pytest.raises(NotImplementedError):
            r.read("/chunked")


def _dataset(rng, dim=2, truth=True, n_tx=5, nt=211, n_rx=12, tx=None, meta=None):
    from openusct_amd.dataset import ArrayGeometry, Dataset
    g = ArrayGeometry.ring(n_rx, 0.02, 5e5) if dim == 2 else ArrayGeometry.cylinder(2, n_rx // 2, 0.01,
                                                                                    0.008, 3e5)
    c = 1500.0 + rng.standard_normal((9,) * dim) if truth else None
    return Dataset(geometry=g, data=rng.standard_normal((n_tx, nt, n_rx)), sample_rate_hz=2e7,
                   tx_wavelet=rng.standard_normal(nt), tx_centre_freq_hz=5e5, nominal_speed_m_s=1480.0,
                   tx_elements=tx, ground_truth=None if c is None else {"c": c, "h_m": 2.5e-4},
                   metadata=meta or {})


def _same(a, b):
    assert np.array_equal(a.data, np.asarray(b.data, np.float32))
    assert np.array_equal(a.geometry.element_pos, b.geometry.element_pos)
    assert a.geometry.array_type == b.geometry.array_type and a.geometry.radius_m == b.geometry.radius_m
    assert np.array_equal(a.tx_wavelet, b.tx_wavelet) and np.array_equal(a.tx_elements, b.tx_elements)
    for k in ("sample_rate_hz", "tx_centre_freq_hz", "nominal_speed_m_s"):
        assert getattr(a, k) == getattr(b, k), k
    assert (a.ground_truth is None) == (b.ground_truth is None)
    if a.ground_truth is not None:
        assert np.array_equal(a.ground_truth["c"], b.ground_truth["c"])
        assert a.ground_truth["h_m"] == b.ground_truth["h_m"]


def test_save_load_roundtrip(tmp_path):
    """test_dataset.py:22-86 with the native writer and reader, here bit-exact."""
    from openusct_amd.dataset import Dataset
    rng = np.random.default_rng(3)
    cases = [_dataset(rng, 2, True, tx=[0, 3, 4, 7, 9], meta={"description": "x", "created_by": "y"}),
             _dataset(rng, 3, True, n_tx=2, nt=97),
             _dataset(rng, 2, False, n_tx=1, nt=5, n_rx=4),
             _dataset(rng, 2, True, n_tx=0, nt=50)]
    for i, ds in enumerate(cases):
        p = str(tmp_path / f"d{i}.h5")
        ds.save(p)
        back = Dataset.load(p)
        _same(back, ds)
        assert back.metadata["format_version"] == "1.0"
        assert back.metadata["description"] == ds.metadata.get("description", "")
        assert back.metadata["created_by"] == ds.metadata.get("created_by", "OpenUSCT")


def test_writer_chunk_trees_and_edges(tmp_path):
    """Chunk indexes of one, two and three B-tree levels (64 chunks a node),
    edge chunks, every writable type, scalars and empty arrays."""
    from openusct_amd.h5file import Reader, Writer
    rng = np.random.default_rng(4)
    arrs = {
        "/one": (rng.standard_normal((3, 5)).astype(np.float32), (3, 5)),
        "/two": (rng.standard_normal((130, 7)), (2, 7)),
        "/three": (rng.integers(-9, 9, (4100, 3)).astype(np.int64), (1, 2)),
        "/edges": (rng.standard_normal((11, 13, 5)).astype(np.float32), (4, 6, 3)),
        "/i32": (rng.integers(-2**31, 2**31, 77).astype(np.int32), (10,)),
        "/u8": (rng.integers(0, 256, (6, 6)).astype(np.uint8), None),
        "/plain": (rng.standard_normal((2, 3, 4)), None),
        "/scalar": (np.float64(1.25), None),
        "/empty": (np.zeros((0, 5), np.float32), (1, 5)),
    }
    p = str(tmp_path / "w.h5")
    with Writer(p) as w:
        w.group("/g")
        w.group("/g/h")
        for k, (a, ch) in arrs.items():
            w.dataset(k, a, chunks=ch, gzip=None if ch is None else 3, threads=3)
        w.dataset("/g/h/raw", arrs["/edges"][0], chunks=(5, 5, 5))  # chunked, no filter
        w.attr("/g", "s", "naïve")
        w.attr("/g/h/raw", "arr", np.array([1.5, 2.5]))
        w.attr("/", "i", 3)
    with Reader(p) as r:
        for k, (a, ch) in arrs.items():
            v = r.read(k)
            assert v.dtype == np.asarray(a).dtype and np.array_equal(v, a), k
            assert r.info(k)[2] == (ch is not None and np.asarray(a).size > 0), k
        assert np.array_equal(r.read("/g/h/raw"), arrs["/edges"][0])
        assert r.attr("/g", "s") == "naïve" and list(r.attr("/g/h/raw", "arr")) == [1.5, 2.5]
        assert r.attr("/", "i") == 3 and r.attr("/", "i").dtype == np.int64
        assert np.array_equal(r.read("/three", rows=(4000, 4100)), arrs["/three"][0][4000:])


def test_writer_errors_and_discard(tmp_path):
    from openusct_amd.h5file import Writer
    p = str(tmp_path / "bad.h5")
    with pytest.raises(KeyError):
        with Writer(p) as w:
            w.dataset("/nogroup/x", np.zeros(3))
    assert not os.path.exists(p)  # a failed write leaves no partial file
    with Writer(p) as w:
        w.dataset("/x", np.zeros(3))
        with pytest.raises(ValueError):
            w.dataset("/x", np.zeros(3))
        with pytest.raises(TypeError):
            w.dataset("/c", np.zeros(3, np.complex128))
        with pytest.raises(ValueError):
            w.dataset("/y", np.zeros((4, 4)), gzip=4)  # compression needs chunks
    assert os.path.exists(p)


def test_reader_errors(tmp_path):
    from openusct_amd.dataset import Dataset
    from openusct_amd.h5file import Reader, Writer
    with pytest.raises(FileNotFoundError):
        Reader(str(tmp_path / "none.h5"))
    junk = tmp_path / "junk.h5"
    junk.write_bytes(b"not an hdf5 file" * 100)
    with pytest.raises(OSError):
        Reader(str(junk))
    short = tmp_path / "short.h5"
    short.write_bytes(open(_g("ds_ring.h5"), "rb").read()[:5000])
    with pytest.raises(OSError):
        Dataset.load(str(short))
    # the schema asks readers to refuse unknown major versions
    v2 = str(tmp_path / "v2.h5")
    ds = _dataset(np.random.default_rng(5), n_tx=1, nt=8, n_rx=4)
    ds.save(v2)
    with Reader(v2) as r:
        names = r.attr_names("/")
    assert "format_version" in names
    with Writer(v2) as w:
        w.attr("/", "format_version", "2.0")
        w.group("/geometry")
    with pytest.raises(ValueError):
        Dataset.load(v2)


@pytest.mark.skipif(not _have_h5py(), reason="no h5py-enabled Python in this image")
def test_libhdf5_reads_what_this_library_writes(tmp_path):
    """h5py (libhdf5) opens files written here and reads every value the
    reference's Dataset.load reads (dataset.py:147-174), with its types."""
    from openusct_amd.h5file import Writer
    rng = np.random.default_rng(6)
    ds = _dataset(rng, 2, True, n_tx=70, nt=33, n_rx=10, tx=list(range(0, 140, 2)),
                  meta={"description": "é ω", "created_by": "w"})
    p2, p3, pw = str(tmp_path / "a.h5"), str(tmp_path / "b.h5"), str(tmp_path / "w.h5")
    ds.save(p2)
    ds3 = _dataset(rng, 3, True, n_tx=3, nt=40)
    ds3.save(p3)
    big = np.zeros((4100, 3), np.float32)
    big[::7] = 1.0
    with Writer(pw) as w:
        w.dataset("/three", big, chunks=(1, 3), gzip=1)  # a three-level chunk index
        w.dataset("/edges", np.arange(11 * 13, dtype=np.float64).reshape(11, 13), chunks=(4, 6))
        w.attr("/", "s", "")
    np.savez(str(tmp_path / "e.npz"), d2=np.float32(ds.data), c2=ds.ground_truth["c"], pos2=ds.geometry.element_pos,
             w2=ds.tx_wavelet, tx2=ds.tx_elements, d3=np.float32(ds3.data), c3=ds3.ground_truth["c"], big=big)
    script = r'''
import sys, h5py, numpy as np
p2, p3, pw, e = sys.argv[1:5]
e = np.load(e)
with h5py.File(p2, "r") as f:
    g, a, t = f["geometry"], f["acquisition"], f["ground_truth"]
    assert np.array_equal(a["data"][()], e["d2"]) and a["data"].dtype == np.float32
    assert a["data"].compression == "gzip"
    assert np.array_equal(g["element_pos"][()], e["pos2"]) and np.array_equal(a["tx_wavelet"][()], e["w2"])
    assert np.array_equal(a["tx_elements"][()], e["tx2"]) and np.array_equal(t["c"][()], e["c2"])
    assert isinstance(f.attrs["description"], str) and f.attrs["description"] == "é ω"
    assert f.attrs["created_by"] == "w" and f.attrs["format_version"] == "1.0"
    assert str(g.attrs["array_type"]) == "ring" and int(g.attrs["n_elements"]) == 10
    assert g.attrs["n_elements"].dtype == np.int64 and g.attrs["radius_m"].shape == ()
    assert float(a.attrs["sample_rate_hz"]) == 2e7 and float(t.attrs["h_m"]) == 2.5e-4
    assert a["data"][5, 3:9, 2:4].tolist() == e["d2"][5, 3:9, 2:4].tolist()
with h5py.File(p3, "r") as f:
    assert np.array_equal(f["acquisition/data"][()], e["d3"]) and np.array_equal(f["ground_truth/c"][()], e["c3"])
    assert str(f["geometry"].attrs["array_type"]) == "cylinder"
with h5py.File(pw, "r") as f:
    assert np.array_equal(f["three"][()], e["big"]) and f["three"][4099, 1] == e["big"][4099, 1]
    assert np.array_equal(f["edges"][()], np.arange(11 * 13, dtype=np.float64).reshape(11, 13))
    assert f.attrs["s"] == ""
print("LIBHDF5 OK")
'''
    r = subprocess.run([CONDA_PY, "-c", script, p2, p3, pw, str(tmp_path / "e.npz")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "LIBHDF5 OK" in r.stdout, r.stdout + r.stderr


@pytest.fixture(scope="module")
def walker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    out = str(tmp_path_factory.mktemp("h5walk") / "h5_walk")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fopenmp", os.path.join(ROOT, "tests", "h5_native", "h5_walk.cpp"),
           os.path.join(ROOT, "openusct_amd", "csrc", "usct_h5.cpp"), "-lz", "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_damaged_files_fail_cleanly(walker, tmp_path):
    """Truncated and bit-flipped copies of every golden file: each either reads
    or fails with a status code; AddressSanitizer / UBSan see no bad access."""
    rng = np.random.default_rng(11)
    files = []
    for src in sorted(glob.glob(os.path.join(GOLD, "*.h5"))):
        b = open(src, "rb").read()
        n = len(b)
        cuts = sorted(set(rng.integers(0, n, 24).tolist() + [0, 8, 95, 96, n - 1]))
        for cut in cuts:
            files.append(b[:cut])
        for t in range(160):
            a = bytearray(b)
            for _ in range(int(rng.integers(1, 5))):
                pos = int(rng.integers(0, min(n, 8192) if t % 4 else n))
                a[pos] ^= 1 << int(rng.integers(0, 8))
            files.append(bytes(a))
    paths = []
    for i, data in enumerate(files):
        p = tmp_path / f"x{i}.h5"
        p.write_bytes(data)
        paths.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([walker, *paths], capture_output=True, text=True, timeout=600, env=env)
    tail = "\n".join((r.stdout + r.stderr).splitlines()[-15:])
    assert r.returncode == 0, tail
    lines = r.stdout.splitlines()
    assert len(lines) == len(paths), tail
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
    ok = sum(" ok " in ln for ln in lines)
    assert 0 < ok < len(paths)  # some damage is harmless, most is detected
    r = subprocess.run([walker, *sorted(glob.glob(os.path.join(GOLD, "*.h5")))], capture_output=True,
                       text=True, timeout=120, env=env)
    assert r.returncode == 0 and r.stdout.count(" ok ") == 4, r.stdout + r.stderr
