"""Checkpoint and resume: a run saved after some iterations and resumed in a
fresh context continues bit for bit (reports and the whole table), for one-
and two-word determinants; mismatched contexts are refused."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from mcdfci_amd import Cdfci, McdfciError, UPDATE_LS


def _same_state(a, b):
    ra, rb = a.report(), b.report()
    for key in ("energy", "x_size", "z_size", "cols_size", "updates", "scale", "xx", "xz", "dx_max",
                "iterations", "candidates", "max_residual", "sum_diff_inv_iter"):
        assert ra[key] == rb[key], (key, ra[key], rb[key])
    da, xa, za = a.table()
    db, xb, zb = b.table()
    assert np.array_equal(da, db)
    assert np.array_equal(xa, xb)
    assert np.array_equal(za, zb)


@pytest.mark.parametrize("name,k,M,z_thr,mem,n1,n2", [
    ("c3_synth", 8, 8, 0.0, 0.05, 40, 30),
    ("c4", 64, 64, 1e-6, 8.0, 6, 5),
    ("c5", 16, 16, 1e-5, 8.0, 5, 4),     # two-word determinants
])
def test_resume_continues_bit_exact(name, k, M, z_thr, mem, n1, n2, workloads, tmp_path):
    print("case", name, k, M, flush=True)
    path = workloads(name)["fcidump"]
    kw = dict(num_of_coordinate=k, estimate=M, z_threshold=z_thr, max_memory=mem, threshold=0.0)
    a = Cdfci(path, **kw)
    a.init()
    a.step(n1)
    ck = str(tmp_path / "run.ck")
    a.save(ck)
    a.step(n2)
    b = Cdfci(path, **kw)
    b.load(ck)
    b.set_graph(2)
    b.step(n2)
    _same_state(a, b)
    a.close()
    b.close()


def test_resume_refuses_other_options(workloads, tmp_path):
    path = workloads("c3_synth")["fcidump"]
    a = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.05, threshold=0.0)
    a.init()
    a.step(5)
    ck = str(tmp_path / "run.ck")
    a.save(ck)
    b = Cdfci(path, num_of_coordinate=4, estimate=8, max_memory=0.05, threshold=0.0)
    with pytest.raises(McdfciError):
        b.load(ck)
    c = Cdfci(workloads("c1")["fcidump"], num_of_coordinate=8, estimate=8, max_memory=0.05, threshold=0.0)
    with pytest.raises(McdfciError):
        c.load(ck)
    with pytest.raises(McdfciError):
        b.load(str(tmp_path / "missing.ck"))
    for x in (a, b, c):
        x.close()


def test_resume_refuses_another_coordinate_update(workloads, tmp_path):
    """A line-search checkpoint (k = 1) does not load into an eigen-update
    context of the same k; it does load into a line-search context and
    continues the run."""
    path = workloads("c3_synth")["fcidump"]
    kw = dict(num_of_coordinate=1, estimate=8, max_memory=0.05, threshold=0.0)
    a = Cdfci(path, coordinate_update=UPDATE_LS, **kw)
    a.init()
    a.step(5)
    ck = str(tmp_path / "ls.ck")
    a.save(ck)
    b = Cdfci(path, **kw)
    with pytest.raises(McdfciError) as ei:
        b.load(ck)
    assert ei.value.code == 1 and "coordinate_update" in str(ei.value)
    c = Cdfci(path, coordinate_update=UPDATE_LS, **kw)
    c.load(ck)
    a.step(3)
    c.step(3)
    assert a.report()["energy"] == c.report()["energy"]
    for x in (a, b, c):
        x.close()
