"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares; host-side argument checks return errors, not exits."""
import ctypes
import glob
import os
import re

import pytest

from nuc3d_amd import LIB_PATH, PKG_DIR

ROOT = os.path.dirname(PKG_DIR)


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(nuc3d_[a-z_0-9]+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("nuc3d_case_create", "nuc3d_rhs", "nuc3d_stage", "nuc3d_steps", "nuc3d_upload_q",
                     "nuc3d_download_rhs", "nuc3d_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bad_arguments_return_status():
    from nuc3d_amd import solver
    L = solver.lib()
    assert L.nuc3d_case_create(None, 1, None, None) == -1
    assert b"null" in L.nuc3d_last_error()
    assert L.nuc3d_rhs(None, None) == -1
    assert L.nuc3d_steps(None, 1, 0.5, None) == -1
