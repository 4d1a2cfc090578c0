"""The C-ABI library loads without a GPU and exports every function include/*.h declares.
No compute calls are made here (they need a device)."""
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(tg_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_eight_kernels():
    names = declared_functions()
    for k in ("PP_direct", "PC_direct", "CC_direct", "CP_direct", "PP_ewald", "PC_ewald", "CC_ewald", "CP_ewald"):
        assert f"tg_gravity_{k}" in names


def test_library_exports_every_declared_symbol():
    from tigergrav_amd._lib import lib

    L = lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_version_without_device():
    from tigergrav_amd._lib import lib

    assert lib().tg_version() >= 1
