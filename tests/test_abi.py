"""The C-ABI libraries load without a GPU and export every function their
headers declare (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bc250-vaapi_amd")


def declared_functions(path):
    """Names of the functions a C header declares (prototypes ending in ';')."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r"^\s*#[^\n]*", " ", text, flags=re.M)
    text = text.replace('extern "C" {', " ").replace("extern \"C\"", " ")
    names = []
    depth = 0
    stmt = []
    for ch in text:
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
            stmt = []
            continue
        if depth:
            continue
        if ch == ";":
            s = " ".join("".join(stmt).split())
            stmt = []
            if "(" in s and not s.startswith("typedef"):
                m = re.search(r"(\w+)\s*\(", s)
                if m:
                    names.append(m.group(1))
        else:
            stmt.append(ch)
    return names


HEADERS = {
    "libbc250_hip.so": [os.path.join(ROOT, "include", h) for h in ("gpu_compute.h", "bc250_stages.h",
                                                                     "bc250_calib.h", "bc250_hevc.h",
                                                                     "bc250_cpu_me.h", "bc250_bitstream.h")],
    "libbc250_host.so": [os.path.join(PKG, "host", "bc250_host.h")],
}


def test_header_parser_finds_known_entry_points():
    names = declared_functions(os.path.join(ROOT, "include", "gpu_compute.h"))
    for n in ("bc250_gpu_init", "gpu_compute_dispatch_encode_ext", "gpu_compute_get_nz_staging_data_slot",
              "gpu_compute_sync_slot", "gpu_compute_map_surface"):
        assert n in names
    assert len(names) >= 45
    assert "gpu_image_t" not in names


@pytest.mark.parametrize("lib", sorted(HEADERS))
def test_library_exports_every_declared_function(lib):
    path = os.path.join(PKG, lib)
    if not os.path.exists(path):
        pytest.fail(f"{lib} is not built (run __graft_entry__.build())")
    so = C.CDLL(path)
    missing = []
    count = 0
    for h in HEADERS[lib]:
        for name in declared_functions(h):
            count += 1
            try:
                getattr(so, name)
            except AttributeError:
                missing.append(f"{os.path.basename(h)}: {name}")
    assert count > 0
    assert not missing, f"declared but not exported: {missing}"


def test_context_struct_layout_matches_python_mirror():
    """The caller allocates bc250_gpu_context_t; the ctypes mirror used by the
    tests and bench must have the C layout (pointer, ints, two embedded
    structs)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bc250_pkg", os.path.join(PKG, "__init__.py"))
    pkg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(pkg)
    # 8 (pointer) + 4*4 ints, then gpu_image_t (2 pointers + 8 x 4 B = 48 B),
    # then gpu_memory_t (3 x 8 B), then two bools padded to 8
    assert C.sizeof(pkg.GpuImage) == 48
    assert C.sizeof(pkg.GpuMemory) == 24
    assert pkg.GpuContext.recon_image.offset == 24
    assert pkg.GpuContext.has_recon_frame.offset == 8 + 16 + 48 + 24
    assert C.sizeof(pkg.GpuContext) == 8 + 16 + 48 + 24 + 8
