"""CPU checks of the C ABI: the ctypes mirror matches include/lbflow_hip.h,
the built library loads and exports every function the header declares,
and descriptor validation errors come back as codes (no compute calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from lbflow_amd import abi, configs
from lbflow_amd.engine import library_path, load_library
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(lbfx?_\w+)\s*\(", text))
    return names


def test_struct_layout_matches_header():
    oracle.build()
    lib = oracle.load()
    assert lib.lbo_sizeof(0) == C.sizeof(abi.CaseDesc)
    assert lib.lbo_sizeof(1) == C.sizeof(abi.ScalarDesc)
    assert lib.lbo_sizeof(2) == C.sizeof(abi.FaceBC)
    assert lib.lbo_sizeof(3) == C.sizeof(abi.Collision)
    assert lib.lbo_sizeof(4) == C.sizeof(abi.ScalarFaceBC)
    assert lib.lbo_sizeof(5) == abi.CaseDesc.scalars.offset
    assert lib.lbo_sizeof(6) == abi.CaseDesc.shift_scale.offset
    assert lib.lbo_sizeof(7) == abi.ScalarDesc.initial_q_ads.offset
    assert lib.lbo_sizeof(8) == abi.CaseDesc.force_field.offset
    assert lib.lbo_sizeof(9) == abi.CaseDesc.wall_density.offset


def test_entry_point_list_equals_header():
    assert header_functions() == {n for n, _, _ in abi.ENTRY_POINTS + abi.SERVER_ENTRY_POINTS +
                                  abi.TUNING_ENTRY_POINTS}


def test_library_exports_every_declared_function():
    assert os.path.exists(library_path()), "build the engine first (python -m lbflow_amd.build)"
    lib = load_library()
    for name in header_functions():
        assert hasattr(lib, name), name


def test_validation_errors_without_gpu():
    lib = load_library()
    bad = configs.cavity(n=8)
    desc, keep = bad.to_c()
    desc.abi_version = 99
    out = C.c_uint64()
    assert lib.lbf_device_bytes(C.byref(desc), C.byref(out)) == abi.EINVAL
    assert b"ABI" in lib.lbf_last_error()
    zh = configs.cavity(n=8)
    zh.faces[0] = zh.faces[0].zou_he(1.0, (0.1, 0.0))  # density and ux both given: reference panics
    desc, keep = zh.to_c()
    assert lib.lbf_device_bytes(C.byref(desc), C.byref(out)) == abi.EINVAL
    d3 = configs.porous_channel(n=8)
    d3.faces[0] = d3.faces[0].zou_he(None, (0.1, 0.0, 0.0))
    desc, keep = d3.to_c()
    assert lib.lbf_device_bytes(C.byref(desc), C.byref(out)) == abi.EUNSUPPORTED
    ok = configs.porous_channel(n=16, precision=abi.F32)
    desc, keep = ok.to_c()
    assert lib.lbf_device_bytes(C.byref(desc), C.byref(out)) == 0
    # the in-place AA scheme holds one population array
    assert 19 * 16 ** 3 * 4 <= out.value < 2 * 19 * 16 ** 3 * 4
    two = configs.porous_channel(n=16, precision=abi.F32, streaming=abi.STREAM_TWO_ARRAY)
    desc, keep = two.to_c()
    assert lib.lbf_device_bytes(C.byref(desc), C.byref(out)) == 0
    assert out.value >= 2 * 19 * 16 ** 3 * 4


def test_automatic_scheme_choice_without_gpu():
    # lbf_device_bytes sizes the scheme lbf_create would pick: one population
    # array in place, two arrays otherwise
    lib = load_library()
    out = C.c_uint64()

    def arrays(sp):
        desc, keep = sp.to_c()
        assert lib.lbf_device_bytes(C.byref(desc), C.byref(out)) == 0
        one = sp.num_nodes * {abi.D2Q9: 9, abi.D3Q15: 15, abi.D3Q19: 19, abi.D3Q27: 27}[
            sp.velocity_set] * {abi.F64: 8, abi.F32: 4}[sp.precision]
        return 2 if out.value >= 2 * one else 1

    # FP64 runs in place where that gives the reference's bits ...
    assert arrays(configs.porous_channel(n=16, precision=abi.F64)) == 1
    assert arrays(configs.sphere_channel(nx=24, ny=12, nz=12, radius=3)) == 1  # Zou-He x, periodic y/z
    # ... and on two arrays where a Zou-He face meets another non-periodic face
    from lbflow_amd.case import FaceBC, Spec
    channel = Spec(velocity_set=abi.D3Q19, n=(20, 10, 8), collision=configs.Collision.trt(1.2, 0.9),
                   faces=[FaceBC.zou_he(None, (0.03, 0.0, 0.0)), FaceBC.zou_he(1.0, (None, 0.0, 0.0))]
                   + [FaceBC.no_slip()] * 4, initial_density=1.0, initial_velocity=(0.0, 0.0, 0.0),
                   zou_he_3d=True)
    assert arrays(channel) == 2
    assert arrays(channel.replace(precision=abi.F32)) == 1  # FP32 is not held to the bits
    # an in-place array of 4 GiB or more falls back to two arrays
    assert arrays(configs.porous_channel(n=600, precision=abi.F32, p_solid=0.0)) == 2
    assert arrays(configs.porous_channel(n=350, precision=abi.F32, p_solid=0.0)) == 1  # 3.3 GB
