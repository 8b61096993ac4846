"""Run every layer of the test-suite this machine can run, with a reason for
each layer it cannot (the role of the reference's run_tests.py:53-100).

    python tests/run_all.py

Layers: the CPU suite (oracle, host logic, HDF5, lane emulator, the C example
compiled and linked); the reference's own C++ core against the oracle when it
is built (oracle/_ref, `make -C oracle ref`); the device suite when a GPU and
the built library are present.  Exit status 0 when nothing that ran failed.
"""

from __future__ import annotations

import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _tail(text, n=12):
    lines = [ln for ln in text.splitlines() if ln.strip()]
    return "\n    ".join(lines[-n:])


def _gpu_reason():
    """None when a device run is possible, else why not."""
    if not os.path.exists(os.path.join(ROOT, "openusct_amd", "libusct.so")):
        return "libusct.so not built (python -c \"import __graft_entry__ as g; g.build()\")"
    probe = subprocess.run([sys.executable, "-c", "import torch; print(torch.cuda.device_count())"],
                           capture_output=True, text=True, timeout=600)
    if probe.returncode != 0 or probe.stdout.strip() in ("", "0"):
        return "no HIP device visible"
    return None


def main():
    results = []

    def layer(name, args, skip=None):
        if skip:
            print(f"--- {name}: SKIP ({skip})", flush=True)
            results.append((name, "SKIP", skip))
            return
        print(f"--- {name}: running", flush=True)
        r = subprocess.run([sys.executable, "-m", "pytest", *args], cwd=ROOT, capture_output=True,
                           text=True, timeout=3600)
        ok = r.returncode == 0
        results.append((name, "PASS" if ok else "FAIL", "" if ok else _tail(r.stdout + r.stderr)))

    layer("CPU suite (oracle, host, HDF5, emulator, C example build)",
          ["tests", "-q", "-m", "not gpu"])
    ref_built = glob.glob(os.path.join(ROOT, "oracle", "_ref", "_uap*.so"))
    layer("reference C++ core vs oracle", ["tests/test_oracle.py", "-q", "-k", "reference_cpp"],
          skip=None if ref_built else "oracle/_ref not built (make -C oracle ref)")
    layer("device suite", ["tests", "-q", "-m", "gpu", "--timeout", "300", "--timeout-method",
                           "thread"], skip=_gpu_reason())

    print("\n=== summary ===")
    for name, status, detail in results:
        print(f"{status:5s} {name}" + (f"\n    {detail}" if detail else ""))
    return 1 if any(s == "FAIL" for _, s, _ in results) else 0


if __name__ == "__main__":
    sys.exit(main())
