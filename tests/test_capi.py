"""The C-ABI library loads and exports every entry point include/*.h declares.
No device calls: runs without a GPU."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^[A-Za-z_][\w\s\*]*?\b(rmsd_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("rmsd_ctx_create", "rmsd_pairwise", "rmsd_fit_to_ref", "rmsd_cluster_cache"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import cpptraj_amd
    path = cpptraj_amd.library_path()
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (python -m cpptraj_amd.build)")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_gfx950_only():
    import re as _re
    import cpptraj_amd
    blob = open(cpptraj_amd.library_path(), "rb").read()
    targets = set(_re.findall(rb"amdgcn-amd-amdhsa--(gfx\w+)", blob))
    assert targets == {b"gfx950"}, targets


def test_no_cpu_fallback_without_gpu():
    import torch
    import cpptraj_amd
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(cpptraj_amd.RmsdError):
        cpptraj_amd.Context(0)


def test_host_library_exports_its_c_entry_points():
    """libcpptraj_rmsd_host.so (the C++ host layer) exports every chost_*
    function of host/host_capi.h and links the kernels library."""
    from cpptraj_amd.host import _HOST, load_host
    if not os.path.exists(_HOST):
        pytest.fail(f"{_HOST} not built (python -m cpptraj_amd.build)")
    src = open(os.path.join(ROOT, "cpptraj_amd", "host", "host_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(chost_\w+)\s*\(", src)))
    assert "chost_run" in names
    lib = load_host()
    assert not [n for n in names if not hasattr(lib, n)]


def test_host_session_needs_a_gpu():
    import torch
    from cpptraj_amd import RmsdError
    from cpptraj_amd.host import HostSession
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RmsdError):
        HostSession(0)
