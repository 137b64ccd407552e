"""The C-ABI library loads and exports every entry point include/*.h declares;
host-only calls work without a GPU and device calls fail loudly."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

import cassandra_amd.capi as capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(cad_[a-z0-9_]+)\s*\(", txt))
    return names


def test_every_declared_symbol_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {ln.split()[-1] for ln in out.stdout.splitlines() if " T " in ln}
    want = declared()
    assert len(want) >= 40
    missing = want - exported
    assert not missing, missing
    # and the ctypes table binds exactly the declared set
    assert set(capi.SIGNATURES) == want


def test_library_loads_and_reports_version():
    assert capi.lib().cad_version().startswith(b"cassandra_amd")


def test_device_calls_fail_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        return
    h = C.c_void_p()
    rc = capi.lib().cad_context_create(0, C.byref(h))
    assert rc == -4
    assert capi.lib().cad_last_error()


def test_host_only_group_rejects_device_work():
    arr = np.array([0x42FF], dtype=np.uint64)
    h = C.c_void_p()
    capi.call("cad_patterns_create", None, arr.ctypes.data_as(C.c_void_p), 1, C.byref(h))
    out = C.c_void_p()
    rc = capi.lib().cad_matrix_from_positions(h, C.c_void_p(16), 1, C.byref(out))
    assert rc == -1 and b"host-only" in capi.lib().cad_last_error()
    rc = capi.lib().cad_index_extract(h, C.c_void_p(16), 1, C.c_void_p(16), 0)
    assert rc == -1
    rc = capi.lib().cad_eval_backward_f32_noalloc(h, C.c_void_p(16), 1, C.c_void_p(16), C.c_void_p(16))
    assert rc == -1 and b"host-only" in capi.lib().cad_last_error()
    capi.call("cad_patterns_destroy", h)


def test_argument_errors():
    out = np.zeros(4, np.uint64)
    assert capi.lib().cad_random_play_host(1, 2, 10, 70, out.ctypes.data_as(C.c_void_p)) == -1
    assert capi.lib().cad_random_play_host(1, 2, 30, 20, out.ctypes.data_as(C.c_void_p)) == -1
    o = capi.default_options()
    assert (o.solver, o.precision, o.tol) == (capi.SOLVER_LSQR, capi.PREC_MIXED, 1e-10)
