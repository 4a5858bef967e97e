"""CPU checks of the C ABI: the library loads without a GPU and exports every function that
include/nscale_hip.h declares (no compute calls)."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "nscale_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-zA-Z_][\w\s\*]*?\b(nsc_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_operator_surface():
    names = declared_functions()
    for required in ("nsc_imreconstruct_u8", "nsc_imfill_holes_u8", "nsc_bwareaopen2_u8", "nsc_bwlabel2_u8",
                     "nsc_dist_transform_u8", "nsc_watershed2", "nsc_segment_nuclei"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from nscale_amd import lib
    L = lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_synthetic_generators_are_deterministic():
    from nscale_amd.synth import synth_he, synth_mask, synth_marker
    a, b = synth_mask(300, 200, 5), synth_mask(300, 200, 5)
    assert np.array_equal(a, b) and a.shape == (300, 200)
    assert not np.array_equal(a, synth_mask(300, 200, 6))
    assert np.array_equal(synth_marker(a), np.where(a > 10, a - 10, 0).astype(np.uint8))
    h = synth_he(128, 96, 1)
    assert h.shape == (128, 96, 3) and np.array_equal(h, synth_he(128, 96, 1))
    bg = (h[..., 0] > 220) & (h[..., 1] > 210)
    assert 0.3 < bg.mean() < 0.99


def test_errors_are_reported_not_thrown():
    from nscale_amd import lib
    L = lib()
    rc = L.nsc_imreconstruct_u8(None, None, None, 0, 0, 8, None)
    assert rc == -10
    assert b"invalid argument" in L.nsc_last_error()


def test_cpp_operator_surface_compiles_and_links(tmp_path):
    """include/nscale_amd.hpp (the nscale:: C++ names over the C ABI) compiles with the host
    compiler and links against the library; the program makes no GPU call."""
    import shutil
    import subprocess
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        import pytest
        pytest.skip("host compiler or ROCm headers absent")
    src = tmp_path / "shim.cpp"
    src.write_text(
        '#include "nscale_amd.hpp"\n#include <cstdio>\n'
        "int main() {\n"
        "  nsc_seg_params p = nscale::hip::HistologicalEntities::defaults();\n"
        "  nscale::hip::Image<const uint8_t> a(nullptr, 0, 0);\n"
        "  nscale::hip::Image<uint8_t> o(nullptr, 0, 0);\n"
        "  try { nscale::hip::imreconstruct<uint8_t>(a, a, o, 8, nullptr); return 2; }\n"
        "  catch (const nscale::hip::Error& e) { std::printf(\"%s|%d|%d\\n\", nsc_version(), e.code, p.minSize); }\n"
        "  if (p.minSize < 0) {  // never true: instantiates and links the whole surface, runs none of it\n"
        "    nscale::hip::Image<const float> f(nullptr, 0, 0, 64);\n"
        "    nscale::hip::Image<float> fo(nullptr, 0, 0);\n"
        "    nscale::hip::Image<int32_t> l(nullptr, 0, 0, 256);\n"
        "    int n = 0; int* bb = nullptr;\n"
        "    nscale::hip::imreconstruct<float>(f, f, fo, 8, nullptr);\n"
        "    nscale::hip::imfillHoles(a, o, 4, nullptr); nscale::hip::bwlabel2(a, l, 8, nullptr);\n"
        "    nscale::hip::bwareaopen2(a, o, 1, 2, 8, n, nullptr); nscale::hip::distanceTransform(a, fo, nullptr);\n"
        "    nscale::hip::watershed2(a, f, l, 8, nullptr); nscale::hip::imhmin(f, fo, 1.f, 8, nullptr);\n"
        "    nscale::hip::boundingBox(l, 0, n, nullptr); nscale::hip::freeData(nscale::hip::boundingBox2(l, n, nullptr));\n"
        "    nscale::hip::HistologicalEntities::segmentNuclei(a, l, n, bb, NSC_BBOX_ALL, nullptr, nullptr);\n"
        "    nscale::hip::HistologicalEntities::plFindNucleusCandidates(a, o, nullptr, nullptr);\n"
        "    nscale::hip::HistologicalEntities::plSeparateNuclei(a, a, o, nullptr, nullptr);\n"
        "  }\n"
        "  return 0;\n}\n")
    exe = tmp_path / "shim"
    libdir = os.path.join(ROOT, "nscale_amd")
    subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
                    "-I/opt/rocm/include", str(src), "-o", str(exe), "-L", libdir, "-lnscale_hip",
                    "-Wl,-rpath," + libdir, "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.strip().split("|")
    assert out[1] == "-10" and int(out[2]) > 0
