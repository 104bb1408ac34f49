"""CPU-side checks of the product: the C ABI library loads and exports
every entry point include/mcdfci.h declares, the synthetic generator is
deterministic and well-formed, and the driver's input validation follows
src/run_wrapper.h:49-122 / src/main.cpp:19-36 (no device needed)."""

import ctypes as C
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import mcdfci_amd
from mcdfci_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mcdfci.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mcdfci_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(mcdfci_amd.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    # the Python mirror binds exactly the declared entry points
    assert sorted(mcdfci_amd.SIGNATURES) == syms


def test_no_device_fails_loudly_or_runs():
    """Creating a context either works (GPU present) or raises: never a CPU path."""
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a device is present")
    with pytest.raises(mcdfci_amd.McdfciError):
        mcdfci_amd.Cdfci(os.path.join(ROOT, "tests", "golden", "h2o_sto3g.FCIDUMP"), max_memory=0.01)


def test_table_above_2_32_slots_is_refused_before_any_allocation():
    """Slot indices are 32-bit: max_memory = 250 GB on a one-word FCIDUMP
    asks for 2^33 slots (src/run_wrapper.h:203-213 sizing) and is refused
    with MCDFCI_E_UNSUPPORTED before the device is touched (so this runs
    without one); 2^32 slots (C5's 200 GB at two words) are accepted by the
    same check."""
    with pytest.raises(mcdfci_amd.McdfciError) as ei:
        mcdfci_amd.Cdfci(os.path.join(ROOT, "tests", "golden", "h2o_sto3g.FCIDUMP"), max_memory=250.0)
    assert ei.value.code == 7
    assert "2^32" in str(ei.value)


def test_synthetic_generator_is_deterministic(tmp_path):
    a = synth.generate("c1", str(tmp_path / "a"))
    b = synth.generate("c1", str(tmp_path / "b"))
    assert a["sha256"] == b["sha256"]
    h, eri = a["h"], a["eri"]
    assert np.array_equal(h, h.T)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.array_equal(eri, eri.transpose(perm))
    n = h.shape[0]
    w = np.linalg.eigvalsh(eri.reshape(n * n, n * n))
    assert w.min() > -1e-10  # positive semidefinite pair matrix
    assert a["core"] == 9.0


def test_generator_core_shift_keeps_spectrum_negative(tmp_path):
    spec = synth.SyntheticSpec("hub", norb=4, nelec=4, kind="hubbard")
    info = synth.generate(spec, str(tmp_path))
    assert info["core"] < 0 and info["hf_energy"] < 0


def _driver(tmp_path, doc, fcidump=True):
    inp = tmp_path / "input.json"
    inp.write_text(json.dumps(doc) if isinstance(doc, dict) else doc)
    return subprocess.run([mcdfci_amd.DRIVER_PATH, str(inp)], capture_output=True, text=True, timeout=60)


def test_driver_argument_and_input_errors(tmp_path):
    r = subprocess.run([mcdfci_amd.DRIVER_PATH], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "No input file specified" in r.stderr
    r = _driver(tmp_path, "{not json")
    assert r.returncode == 1 and "Can not read the input file" in r.stderr
    r = _driver(tmp_path, {"hamiltonian": {"fcidump_path": "x"}, "solver": {"cdfci": {}}})
    assert r.returncode == 1 and "max_memory is invalid" in r.stderr
    r = _driver(tmp_path, {"hamiltonian": {"fcidump_path": "x", "type": "hubbard"},
                           "solver": {"cdfci": {"max_memory": 1}}})
    assert r.returncode == 1 and "hamiltonian:type is invalid" in r.stderr
    r = _driver(tmp_path, {"hamiltonian": {"fcidump_path": "x"},
                           "solver": {"cdfci": {"max_memory": 1, "z_threshold_search": True,
                                                "report_interval": 10}}})
    assert r.returncode == 1 and "report_interval >= 1000" in r.stderr
    assert "Warning: z cut-off threshold auto search" in r.stdout


def test_bench_reads_the_committed_traffic_summary():
    """bench.py's roofline.traffic comes from the counter summary under
    profiles/: the line must parse, and the measured bytes per k_scatter
    launch must exceed the algorithmic bytes (32 B per H-element)."""
    sys.path.insert(0, ROOT)
    import bench

    t = bench.measured_traffic("c4", 64, "mixed")
    assert isinstance(t, int)
    assert t > 1903168 * 32
    t = bench.measured_traffic("c4", 256, "mixed")  # the bench's default batch size
    assert isinstance(t, int)
    assert t > 7612672 * 32
    assert bench.measured_traffic("c4", 63, "mixed") is None


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr
