"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares; the ctypes structs match the header's layout."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:int|void\s*\*|void|uint64_t|uint32_t)\s+\**\s*(xkrt_hip_\w+)\s*\(", re.M)


def declared():
    names = []
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        names += DECL.findall(open(h).read())
    return names


def hip_lib():
    import torch  # noqa: F401  -- one HIP runtime per process: torch's libamdhip64 is loaded first
    from xkaapi_amd import _loader
    return _loader.load_host_lib("libxkrt_hip.so")


def test_header_declares_entry_points():
    names = declared()
    assert "xkrt_hip_dep_domain_create" in names and "xkrt_hip_dep_submit" in names
    assert len(names) == len(set(names))


@pytest.mark.parametrize("name", declared())
def test_library_exports(name):
    lib = hip_lib()
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_ctypes_layouts_match_header():
    from xkaapi_amd import resolver as R
    assert C.sizeof(R.ConfigT) == 4 * 4 + 5 * 8
    assert C.sizeof(R.StatsT) == 11 * 8
    assert R.RECT_DTYPE.itemsize == 32
    assert C.sizeof(R.PoolConfigT) == 2 * 4 + 3 * 8 + 4 * 4
    assert C.sizeof(R.PoolStatsT) == 8 * 8
    assert C.sizeof(R.ServerConfigT) == 2 * 4 + 2 * 8 + 2 * 4
    assert C.sizeof(R.ServerStatsT) == 9 * 8
    assert R.ACCESS_DTYPE.itemsize == C.sizeof(R.AccessT) == 32


def test_product_fails_loudly_without_library(monkeypatch, tmp_path):
    from xkaapi_amd import _loader
    monkeypatch.setattr(_loader, "LIB_DIR", str(tmp_path))
    monkeypatch.setattr(_loader, "_cache", {})
    with pytest.raises(_loader.LibraryMissing):
        _loader.load_host_lib("libxkrt_hip.so")


def test_compiled_abi_driver_links_against_the_library():
    # the C++ caller (tests/abi/xkrt_abi_driver.cpp) is built and every xkrt_* symbol it imports is
    # declared in include/ and exported by the library (it runs in tests/test_gpu_abi_driver.py)
    import subprocess
    exe = os.path.join(ROOT, "tests", "abi", "xkrt_abi_driver")
    assert os.path.exists(exe), "run __graft_entry__.build()"
    und = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True, check=True).stdout
    used = {l.split()[-1] for l in und.splitlines() if l.split() and l.split()[-1].startswith("xkrt_")}
    assert {"xkrt_hip_dep_submit", "xkrt_hip_pool_stage", "xkrt_hip_mcc_plan"} <= used
    assert used <= set(declared())
    lib = hip_lib()
    assert all(hasattr(lib, n) for n in used)


def test_ctypes_layouts_match_compiled_header(tmp_path):
    # every ctypes struct against the C compiler's sizeof and offsetof of the header's struct
    import subprocess
    from xkaapi_amd import coherency as K
    from xkaapi_amd import resolver as R
    structs = {
        "xkrt_hip_access_t": R.AccessT, "xkrt_hip_dep_config_t": R.ConfigT, "xkrt_hip_dep_stats_t": R.StatsT,
        "xkrt_hip_pool_config_t": R.PoolConfigT, "xkrt_hip_pool_stats_t": R.PoolStatsT,
        "xkrt_hip_server_config_t": R.ServerConfigT, "xkrt_hip_server_stats_t": R.ServerStatsT,
        "xkrt_hip_task_access_t": R.TaskAccessT, "xkrt_hip_taskdom_config_t": R.TaskDomConfigT,
        "xkrt_hip_taskdom_stats_t": R.TaskDomStatsT, "xkrt_hip_handle_access_t": R.HandleAccessT,
        "xkrt_hip_handle_config_t": R.HandleConfigT, "xkrt_hip_mcc_config_t": K.MccConfigT,
        "xkrt_hip_mcc_access_t": K.MccAccessT, "xkrt_hip_mcc_stats_t": K.MccStatsT,
    }
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "xkrt_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            if f[0] != "pad":
                lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {(a, b): int(c) for a, b, c in (l.split() for l in out if l)}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f in py._fields_:
            if f[0] != "pad":
                assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
