"""The C-ABI library loads without a GPU and exports every function include/*.h declares, and the
Python binding types every one of them (no compute calls: this runs on the CPU)."""
import glob
import os
import re
import subprocess

import dqmc_amd as gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(dqmc_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_sweep_path():
    names = declared_functions()
    for must in ("dqmc_create", "dqmc_init", "dqmc_sweep_up", "dqmc_sweep_down", "dqmc_sweep_dynamic",
                 "dqmc_get_gf", "dqmc_get_stats", "dqmc_apply_bop", "dqmc_get_health"):
        assert must in names


def test_library_exports_every_declared_function():
    lib = gpu.lib()                                    # loads libdqmc_amd.so (HIP runtime, no device)
    out = subprocess.check_output(["nm", "-D", "--defined-only", gpu.LIB_PATH]).decode()
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = sorted(declared_functions() - exported)
    assert not missing, "declared but not exported: %s" % missing
    for name in declared_functions():
        f = getattr(lib, name)                         # typed by the binding's signature table
        assert f.argtypes is not None, name
