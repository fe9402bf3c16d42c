"""The C-ABI library loads and exports every entry point include/ssf_gpu.h declares
(no compute calls: this runs without a GPU)."""
import os
import re
import subprocess

from ssf_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "ssf_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ssf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", gpu.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared()) <= set(gpu.SIGNATURES)


def test_library_loads_and_reports_errors():
    l = gpu.lib()
    assert l.ssf_last_error() is not None
    # argument validation needs no device
    import ctypes
    h = ctypes.c_void_p()
    rc = l.ssf_matrix_create_csr(None, 11, 0, None, None, ctypes.byref(h))
    assert rc == -1 and b"NULL" in l.ssf_last_error()


def _driver():
    from ssf_amd import build
    if not os.path.exists(build.DRIVER):
        build.build_driver()
    return build.DRIVER


def test_struct_layouts_agree_with_the_compiled_caller():
    """The C++ caller (compiled against include/ssf_gpu.h, whose header also holds
    static_asserts of the layouts) and the ctypes structs of ssf_amd/gpu.py see
    the same sizes and field offsets."""
    import json
    out = subprocess.run([_driver(), "--layout"], capture_output=True, text=True, check=True)
    lay = json.loads(out.stdout)
    for name, cls in (("ssf_solve_params", gpu.SolveParams), ("ssf_solve_info", gpu.SolveInfo)):
        import ctypes
        assert lay[name]["size"] == ctypes.sizeof(cls), name
        for field, _ in cls._fields_:
            assert lay[name][field] == getattr(cls, field).offset, (name, field)
