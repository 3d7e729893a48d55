"""CPU checks of the C ABI: the library loads and exports every function the
header declares, and its host-only entry points (model compilation,
algorithm defaults and validation) agree with the oracle and the reference."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import refobjects as ro
from fpmhip import AlternatingProjection, Epry, Fpie, Mpie, abi, runtime
from fpmhip import synthetic as syn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "fpmhip.h")).read()
    return sorted(set(re.findall(r"\b(fpmhip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    runtime.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", runtime.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (fpmhip_[a-z0-9_]+)", out))
    declared = header_functions()
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    assert sorted(runtime.EXPORTS) == sorted(declared)


def test_version_and_error_state():
    L = runtime.lib()
    assert b"gfx950" in L.fpmhip_version()
    with pytest.raises(runtime.FpmError) as e:
        runtime.validate(Mpie(object_step=0.0).to_c())
    assert e.value.kind == "InvalidParameter" and e.value.name == "object_step"


@pytest.mark.parametrize("k,n,policy", [(3, 32, "smooth"), (15, 128, "smooth"), (13, 64, "minimum"),
                                        (9, 128, "power_of_two"), (3, 32, (64, 64)), (15, 256, "smooth")])
def test_host_compile_matches_oracle(k, n, policy):
    opt, arr = syn.base_optics(), syn.planar_array(k)
    a = runtime.compile_model(opt, arr, (n, n), policy)
    b = oracle.compile_model(opt, arr, (n, n), policy)
    assert a.reconstruction_shape == b.reconstruction_shape
    assert np.array_equal(a.crop_start, b.crop_start)
    assert np.array_equal(a.offset, b.offset)
    assert np.array_equal(a.pupil, b.pupil) and np.array_equal(a.support, b.support)
    assert np.array_equal(a.k_vectors, b.k_vectors)


def test_host_compile_rejects_bad_optics_and_shapes():
    opt = syn.base_optics()
    bad = ro.Optics(532e-9, 0.10, 4.0, 60e-6, 1.0, 1.0)  # pixel above the coherent limit
    with pytest.raises(runtime.FpmError) as e:
        runtime.compile_model(bad, syn.planar_array(3), (32, 32))
    assert e.value.name == "camera_pixel_size"
    with pytest.raises(runtime.FpmError) as e:
        runtime.compile_model(opt, syn.planar_array(15), (128, 128), (200, 200))
    assert e.value.kind == "InvalidShape"


@pytest.mark.parametrize("cls,kind", [(AlternatingProjection, abi.AP), (Fpie, abi.FPIE), (Mpie, abi.MPIE),
                                      (Epry, abi.EPRY)])
def test_defaults_agree(cls, kind):
    c = runtime.algorithm_default(kind)
    p = cls().to_c()
    for name, _ in abi.Algorithm._fields_:
        if name in ("iterations",):
            continue
        assert getattr(c, name) == getattr(p, name), name


@pytest.mark.parametrize("alg,name", [
    (Mpie(object_step=0.0), "object_step"), (Mpie(stability=-0.1), "stability"),
    (Mpie(momentum_interval=0), "momentum_interval"), (Mpie(momentum_friction=1.0), "momentum_friction"),
    (Mpie(momentum_feedback=1.1), "momentum_feedback"), (Mpie(epsilon=0.0), "epsilon"),
    (Epry(pupil_step=-1.0), "pupil_step"), (Epry(gain_step=2.0), "gain_step"),
    (Epry(minimum_gain=2.0, maximum_gain=1.0), "gain_bounds"), (Fpie(stability=2.0), "stability"),
    (AlternatingProjection(object_step=float("nan")), "object_step"),
])
def test_validation_uses_stable_parameter_names(alg, name):
    """tests/reconstruction_tests.rs:907 / :2853."""
    with pytest.raises(runtime.FpmError) as e:
        runtime.validate(alg.to_c())
    assert e.value.kind == "InvalidParameter" and e.value.name == name


def test_mpie_batch_size_zero_is_rejected():
    a = Mpie().to_c()
    a.batch_size = 0
    with pytest.raises(runtime.FpmError) as e:
        runtime.validate(a)
    assert e.value.name == "batch_size"
