"""Autotuner and run-to-run reproducibility on the GPU.

* Default options time nothing: two processes that solve the same system
  (forward and adjoint, FP64 and mixed precision) return bitwise-equal heads
  and gradients.
* DPF_FLAG_AUTOTUNE times the coarse levels' lanes per row and the smoother
  degrees on the topology's first solve; a second process that pins the
  tuned configuration (dpf_topology_set_smoother, dpf_topology_set_lanes)
  reproduces the tuned process's results bit for bit.
* Every lanes-per-row layout gives the same heads to the solve tolerance.

Each child process prints one JSON line naming its case before its GPU work
and the digests of its outputs at the end."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = {"ns": [64, 60, 24], "maxs": [32.0, 30.0, 12.0]}

CHILD = r"""
import hashlib, json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
from conftest import load_package
P = load_package()
from dpfehm_jl_amd import api
cfg = json.loads(sys.argv[2])
print(json.dumps({"case": cfg["case"]}), flush=True)
w = P.workloads.steady_2m(ns=tuple(cfg["ns"]), maxs=tuple(cfg["maxs"]))
K = api.logKs2Ks_harmonic(w.logK, w.neighbors, w.dirichletnodes, w.areasoverlengths)
topo = api.get_topology(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes)
for prec, (f, c, d, r) in cfg.get("pin", {}).items():
    topo.set_smoother(prec, f, c, d, r)
for level, lg in cfg.get("lanes", []):
    topo.set_lanes(level, lg)
out = {"case": cfg["case"]}
for prec in cfg["precisions"]:
    h, pb = api.rrule_groundwater_steadystate(K, w.neighbors, w.areasoverlengths,
                                              w.dirichletnodes, w.dirichleths, w.Qs,
                                              reltol=1e-10, precision=prec,
                                              autotune=cfg.get("autotune", False))
    delta = np.zeros(w.n)
    delta[w.obs - 1] = 2.0 * h[w.obs - 1]
    dK = pb(delta)[1]
    pb.state.close()
    out[prec] = [hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
                 for a in (h, dK)]
lg, trials = topo.lanes()
out["lanes"] = [[l, g] for l, g in enumerate(lg) if g > 0]
out["lane_trials"] = len(trials)
out["degrees"] = topo.tuned_degrees()
print(json.dumps(out), flush=True)
"""


def _child(cfg):
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT, json.dumps(cfg)],
                       capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    sys.stdout.flush()
    assert r.returncode == 0, r.stderr[-3000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_default_bitwise_across_processes():
    """Two processes with default options: the same bits, FP64 and mixed."""
    cfg = dict(CFG, case="default", precisions=["fp64", "mixed"])
    a = _child(cfg)
    b = _child(cfg)
    assert a["lane_trials"] == 0 and b["lane_trials"] == 0
    assert a["degrees"]["fp64"] == [1, 1, 2, 15] and a["degrees"]["mixed"] == [1, 1, 2, 15]
    for prec in ("fp64", "mixed"):
        assert a[prec] == b[prec], prec


def test_autotuned_configuration_pinned_in_another_process():
    """Process 1 autotunes (lanes per row, then degrees and interval); process
    2 pins what it chose and reproduces its results bit for bit."""
    tuned = _child(dict(CFG, case="autotune", precisions=["mixed"], autotune=True))
    assert tuned["lane_trials"] > 0
    f, c, d, r = tuned["degrees"]["mixed"]
    pinned = _child(dict(CFG, case="pinned", precisions=["mixed"],
                         pin={"mixed": [f, c, d, float(r)]}, lanes=tuned["lanes"]))
    assert pinned["lanes"] == tuned["lanes"]
    assert pinned["mixed"] == tuned["mixed"]


def test_every_lanes_layout_same_heads(pkg):
    """Each level's split copy at 2..32 lanes a row: the heads agree with the
    default layout's to the solve tolerance, and the layout reads back."""
    from dpfehm_jl_amd import api
    print("lanes sweep", flush=True)
    w = pkg.workloads.steady_2m(ns=tuple(CFG["ns"]), maxs=tuple(CFG["maxs"]))
    K = api.logKs2Ks_harmonic(w.logK, w.neighbors, w.dirichletnodes, w.areasoverlengths)
    topo = api.Topology(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes)
    lg0, _ = topo.lanes()
    split = [l for l, g in enumerate(lg0) if g > 0]
    assert split, lg0

    def heads(precision):
        import ctypes
        from dpfehm_jl_amd import _lib
        h = np.empty(w.n)
        o = api._opts(dict(reltol=1e-12, precision=precision))
        _lib.check(_lib.lib().dpf_steady_forward(
            topo.handle, api._ptr(api._f64(K)), api._ptr(api._f64(w.dirichleths)),
            api._ptr(api._f64(w.Qs)), api._ptr(h), ctypes.byref(o), None, None))
        return h

    ref = {p: heads(p) for p in ("fp64", "mixed")}
    for lg in range(1, 6):
        for l in split:
            topo.set_lanes(l, lg)
        got, _ = topo.lanes()
        assert all(got[l] == lg for l in split), got
        for p in ("fp64", "mixed"):
            h = heads(p)
            e = np.linalg.norm(h - ref[p]) / np.linalg.norm(ref[p])
            print("lanes", lg, p, e, flush=True)
            assert e < 1e-10
    with pytest.raises(Exception):
        topo.set_lanes(0, 2)
    with pytest.raises(Exception):
        topo.set_lanes(split[0], 7)
    topo.close()
