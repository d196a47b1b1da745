"""The C ABI library loads on a machine without a GPU and exports every entry
point that include/slx.h declares; host-side argument checks fail cleanly."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "slx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(slx_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    abi = __import__("importlib").import_module("starlab-physics_amd._abi")
    if not os.path.exists(abi.LIB_PATH):
        pytest.fail("libslx.so is not built (run __graft_entry__.build())")
    return abi


def test_every_declared_symbol_is_exported_and_bound(lib):
    L = C.CDLL(lib.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/slx.h but not exported"
        assert n in lib.SIGNATURES, f"{n} has no ctypes signature"
    assert set(lib.SIGNATURES) <= set(names)


def test_version_and_host_argument_checks(lib):
    L = lib.lib()
    assert L.slx_abi_version() == 1
    # invalid arguments are rejected before any device work
    assert L.slx_compute_y(-1, None, None, None, 0.0, 0.0, 0.0, 0.0, None) == -1
    assert "nv < 0" in lib.last_error()
    assert L.slx_spring_pass(4, None, None, None, None, None, None, None, 0.0, 0.0, None) == -1
    assert "null pointer" in lib.last_error()
    assert L.slx_traverse(2, 1, None, 0.0, 0, None, None, None, None, None, None, None, None, 1, None, None,
                          None, None) == -1
    with pytest.raises(lib.SlxError):
        lib.call("slx_update_vx", 3, None, None, None, None, 0.0, 0.0, None)
    assert L.slx_velocity_pass(-1, 0, None, None, None, None, None, None, None, 0.8, 40, None, None, None, 0.0, 1,
                               None, None, 0, None) == -1
    assert "bad sizes" in lib.last_error()
    assert L.slx_velocity_pass(4, 0, None, None, None, None, None, None, None, 0.8, 40, None, None, None, 0.0, 1,
                               None, None, 0, None) == -1
    assert "null pointer" in lib.last_error()
    # zero-sized calls are no-ops
    assert L.slx_compute_y(0, None, None, None, 0.0, 0.0, 0.0, 0.0, None) == 0
    assert L.slx_velocity_pass(0, 0, None, None, None, None, None, None, None, 0.8, 40, None, None, None, 0.0, 1,
                               None, None, 0, None) == 0


def test_workspace_queries_are_host_only(lib):
    L = lib.lib()
    assert L.slx_radix_sort_ws_bytes(1 << 20) > 8 << 20
    assert L.slx_scan_ws_bytes(1000) > 0
    assert L.slx_collision_ws_bytes(1000, 10, 5000) > 5000 * 24
    assert L.slx_cell_build_ws_bytes(100) > 0 and L.slx_tri_build_ws_bytes(100) > 0


def test_solver_fails_loudly_without_gpu(slx):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    g = slx.scene.grid_cloth(4)
    m = slx.Mesh(x=g.x, faces=g.faces, device=None)
    with pytest.raises(slx.SlxError):
        slx.Solver(m, None, 1e-4, 1e6, 1e3, (0, -9.8, 0), 1 / 60)


def test_bench_refuses_more_than_one_gpu():
    """The benchmark runs on one GPU: --gpus 2 exits with status 2 and says why,
    before any GPU work."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2, r.stderr[-2000:]
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
