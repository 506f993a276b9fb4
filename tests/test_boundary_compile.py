"""The drop-in boundary at source level: the reference's own callers of
gpu_compute.h -- the H.264 and HEVC encoders, the CPU ME route and both
decoders -- compile against include/gpu_compute.h in place of the Vulkan
header, with implicit declarations as errors (so every function they call
is declared here; tests/test_abi.py checks that the library exports every
declared one).  The reference files are linked into a temporary directory
with every header of theirs except gpu_compute.h; nothing is copied into the
repository.  Skipped when the reference tree is absent."""
import os
import shutil
import subprocess

import pytest

SRC = "/root/reference/approach1-compute-encoder/src"
INC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
CALLERS = ("encoder_h264.c", "encoder_h265.c", "cpu_simd_me.c", "decoder_h264.c", "decoder_h265.c")

pytestmark = pytest.mark.skipif(not os.path.isdir(SRC) or not shutil.which("gcc"),
                                reason="reference tree or gcc absent")


@pytest.fixture(scope="module")
def tree(tmp_path_factory):
    d = tmp_path_factory.mktemp("callers")
    for f in os.listdir(SRC):
        if f.endswith((".h", ".c")) and f != "gpu_compute.h":
            os.symlink(os.path.join(SRC, f), d / f)
    return d


@pytest.mark.parametrize("caller", CALLERS)
def test_reference_caller_compiles_against_the_header(tree, caller):
    r = subprocess.run(["gcc", "-fsyntax-only", "-std=gnu11", "-fopenmp", "-Werror=implicit-function-declaration",
                        "-Werror=incompatible-pointer-types", "-I" + INC, "-I" + str(tree), str(tree / caller)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_the_vulkan_header_is_not_what_they_found(tree):
    """The temporary tree has no gpu_compute.h of its own, so the header the
    callers compiled against is this repository's."""
    assert not (tree / "gpu_compute.h").exists()
    r = subprocess.run(["gcc", "-E", "-std=gnu11", "-I" + INC, "-I" + str(tree), str(tree / "encoder_h264.c")],
                       capture_output=True, text=True)
    assert r.returncode == 0
    assert os.path.join(INC, "gpu_compute.h") in r.stdout
    assert "vulkan/vulkan.h" not in r.stdout
