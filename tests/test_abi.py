"""The C-ABI library loads and exports every symbol include/foilcfd.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from foil_cfd_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "foilcfd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) > 40
    for must in ("fc_create", "fc_init", "fc_step", "fc_forces", "fc_download_f",
                 "fc_upload_f", "fc_compute_scaling", "fc_build_clean_foil_flags"):
        assert must in syms


def test_library_exports_every_declared_symbol(fclib):
    missing = [s for s in declared_symbols() if not hasattr(fclib, s)]
    assert not missing, f"symbols declared in foilcfd.h but not exported: {missing}"


def test_binding_matches_header():
    assert sorted(_abi.SIGNATURES) == declared_symbols()


def test_library_exports_nothing_undeclared():
    # Every exported fc_* symbol (dynamic table) is declared in the header.
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = sorted({ln.split()[-1] for ln in out.splitlines()
                       if re.search(r" T fc_", ln)})
    assert exported == declared_symbols()


def test_device_count_without_gpu_is_safe(fclib):
    assert fclib.fc_device_count() >= 0


def test_init_without_device_fails_cleanly(fclib):
    if fclib.fc_device_count() > 0:
        return
    import numpy as np
    import foil_cfd_amd as fc
    s = fc.compute_scaling()
    h = fclib.fc_create()
    flags = np.zeros(8 * 8 * 8, dtype=np.uint8)
    opt = _abi.FcInitOptions(0, 1, 1)
    err = ctypes.create_string_buffer(256)
    rc = fclib.fc_init(h, _abi.FcDims(8, 8, 8), ctypes.byref(s),
                       flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(opt),
                       err, 256)
    fclib.fc_destroy(h)
    assert rc == _abi.FC_ERR_DEVICE
    assert b"device" in err.value


def test_probe_rejects_bad_arguments_before_touching_the_device(fclib):
    # Argument checks come first, so these run without a GPU.
    ms = ctypes.c_double(0.0)
    d = _abi.FcDims(64, 8, 8)
    for eb, pattern, work, reps in ((3, 0, 0, 1), (2, 3, 0, 1), (2, -1, 0, 1), (2, 0, -1, 1), (2, 0, 0, 0)):
        assert fclib.fc_probe_aa_traffic(d, eb, pattern, work, reps, ctypes.byref(ms)) == _abi.FC_ERR_INVALID
    assert fclib.fc_probe_aa_traffic(_abi.FcDims(2, 8, 8), 2, 0, 0, 1, ctypes.byref(ms)) == _abi.FC_ERR_INVALID
    assert fclib.fc_probe_copy(8, 1, ctypes.byref(ms)) == _abi.FC_ERR_INVALID
