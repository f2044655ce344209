"""The reference-side binding (integration/photon_mi355x.h) compiles against
the reference's own headers.

integration/enroll_example.cpp is a reference problem generator that enrols
photon_mi355x::base_t with mesh_t::enroll_module (mesh/mesh_enroll.h:77-109)
and overrides its medium.  hipcc parses it for gfx950 with the reference's
src/ on the include path, so every reference member the binding touches
(geo_t, block_t, neighbor_t, mod_base_t virtuals, device::base_t streams and
allocation, input, binary_io) must exist with the types it uses.  Runs on the
CPU only where the reference tree is present (never on a GPU box)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

needs_ref = pytest.mark.skipif(not os.path.isdir(REF_SRC) or not shutil.which(HIPCC),
                               reason="needs the reference tree and hipcc")


def _syntax(tu, extra_inc):
    cmd = [HIPCC, "-std=c++17", "-fsyntax-only", "--offload-arch=gfx950",
           "-include", "hip/hip_runtime.h", "-I", REF_SRC, "-I", os.path.join(ROOT, "include"),
           "-I", extra_inc, tu]
    return subprocess.run(cmd, capture_output=True, text=True)


@needs_ref
def test_binding_compiles_against_the_reference_headers():
    r = _syntax(os.path.join(ROOT, "integration", "enroll_example.cpp"),
                os.path.join(ROOT, "integration"))
    assert r.returncode == 0, r.stderr[-4000:]


@needs_ref
def test_a_member_the_reference_lacks_fails_to_compile(tmp_path):
    """The check has teeth: the binding with a device member the reference
    base_t does not have (stream_ptr) is rejected."""
    src = open(os.path.join(ROOT, "integration", "photon_mi355x.h")).read()
    bad = src.replace("mesh.p_dev->yield_stream(  )", "mesh.p_dev->stream_ptr(  )")
    assert bad != src
    (tmp_path / "photon_mi355x.h").write_text(bad)
    # a copy of the example beside the altered header includes that header
    tu = tmp_path / "enroll_example.cpp"
    tu.write_text(open(os.path.join(ROOT, "integration", "enroll_example.cpp")).read())
    r = _syntax(str(tu), str(tmp_path))
    assert r.returncode != 0 and "stream_ptr" in r.stderr
