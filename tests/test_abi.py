"""The C-ABI boundary, checked without a GPU.

libcvl_hip.so must load (no kernel is launched) and export every function
that include/cvl.h and include/cvl_hip_ext.h declare.  Where the reference
header is present (nesl/include/cvl.h, read as text), every one of its
prototypes must appear in include/cvl.h with the same return type and
argument types, so the VCODE interpreter and fused libraries link unchanged.
"""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_HEADER = "/root/reference/nesl/include/cvl.h"


def _lib_path():
    from neslgpu_amd import cvl
    return cvl.LIB_PATH


def _prototypes(path, strip_p=False):
    """{name: (return type, (arg types...))} from a C header, types normalised."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    if strip_p:                                   # reference: f P_((args));
        text = re.sub(r"\bP_\(\((.*?)\)\)", r"(\1)", text, flags=re.S)
    out = {}
    for ret, name, args in re.findall(r"^\s*([A-Za-z_][A-Za-z_ ]*?[\s*]+)([A-Za-z_]\w*)\s*\(([^()]*)\)\s*;",
                                      text, flags=re.M):
        if ret.strip() in ("typedef", "return"):
            continue
        types = []
        a = args.strip()
        if a and a != "void":
            for arg in a.split(","):
                arg = " ".join(arg.replace("*", " * ").split())
                parts = arg.split(" ")
                # drop the parameter name when there is one
                if len(parts) > 1 and re.match(r"^[A-Za-z_]\w*$", parts[-1]) and parts[-1] not in (
                        "int", "double", "float", "unsigned", "vec_p", "cvl_bool", "void"):
                    parts = parts[:-1]
                types.append(" ".join(parts).replace("unsigned int", "unsigned"))
        r = " ".join(ret.replace("*", " * ").split()).replace("unsigned int", "unsigned")
        r = re.sub(r"^extern ", "", r)
        out[name] = (r, tuple(types))
    return out


def test_headers_parse():
    cvl_h = _prototypes(os.path.join(ROOT, "include", "cvl.h"))
    ext_h = _prototypes(os.path.join(ROOT, "include", "cvl_hip_ext.h"))
    assert len(cvl_h) > 600
    assert "add_suz" in cvl_h and cvl_h["add_suz"] == ("void", ("vec_p", "vec_p", "int", "vec_p"))
    assert "cvlhip_set_stream" in ext_h


def test_library_exports_every_declared_function():
    path = _lib_path()
    if not os.path.exists(path):
        pytest.skip("libcvl_hip.so not built (python -m neslgpu_amd.build)")
    lib = C.CDLL(path)
    missing = []
    for h in ("cvl.h", "cvl_hip_ext.h"):
        for name in _prototypes(os.path.join(ROOT, "include", h)):
            try:
                getattr(lib, name)
            except AttributeError:
                missing.append(name)
    assert not missing, f"{len(missing)} declared but not exported: {missing[:20]}"


def test_python_binding_covers_the_headers():
    from neslgpu_amd import cvl
    protos = cvl.parse_prototypes()
    ours = _prototypes(os.path.join(ROOT, "include", "cvl.h"))
    assert set(ours) <= set(protos)
    for name, (ret, args) in ours.items():
        assert len(protos[name][1]) == len(args), name


@pytest.mark.skipif(not os.path.exists(REF_HEADER), reason="reference header not present")
def test_header_matches_reference_prototypes():
    ref = _prototypes(REF_HEADER, strip_p=True)
    ours = _prototypes(os.path.join(ROOT, "include", "cvl.h"))
    assert len(ref) > 600
    missing = sorted(set(ref) - set(ours))
    assert not missing, f"reference functions missing from include/cvl.h: {missing[:20]}"
    diff = [(k, ref[k], ours[k]) for k in ref if ref[k] != ours[k]]
    assert not diff, f"{len(diff)} signatures differ, first: {diff[:5]}"
