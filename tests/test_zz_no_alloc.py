"""The allocation-free forward stack (fpmhip_forward_stack_noalloc): inputs
and output in the caller's device memory, the per-source fields in the
library's fixed device array. Two calls in a row, the second on a smaller
problem, each equal to fpmhip_forward_stack bit for bit, and the device's
free memory unchanged by the calls."""
import sys

import numpy as np
import pytest

from fpmhip import runtime
from fpmhip import synthetic as syn

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cases():
    # 225 LEDs, 128^2 -> 375^2 (four chunks of the scratch array), then 9 LEDs,
    # 64^2 with per-frame gains and background and an aberrated pupil
    big = runtime.compile_model(syn.base_optics(), syn.planar_array(15), (128, 128))
    small = runtime.compile_model(syn.base_optics(), syn.planar_array(3), (64, 64))
    rng = np.random.default_rng(4)
    small.with_frame_gains(rng.uniform(0.5, 2.0, small.frame_count))
    small.with_background(rng.uniform(0.0, 3.0, small.frame_count * small.image_len))
    pupil = syn.zernike_pupil(small, 0.4, 9)
    return [(big, syn.object_spectrum(syn.random_field_object(big.reconstruction_shape, 7)), None),
            (small, syn.object_spectrum(syn.random_field_object(small.reconstruction_shape, 8)), pupil)]


def device_free_bytes():
    """hipMemGetInfo of the HIP runtime this process already runs on."""
    import ctypes as ct

    hip = ct.CDLL("libamdhip64.so", mode=ct.RTLD_GLOBAL)
    free, total = ct.c_size_t(), ct.c_size_t()
    assert hip.hipMemGetInfo(ct.byref(free), ct.byref(total)) == 0
    return free.value


def test_noalloc_forward_stack_equals_forward_stack_on_two_calls(ctx):
    for k, (model, spec, pupil) in enumerate(cases()):
        say("no-alloc forward stack case", k, model.image_shape, model.reconstruction_shape, model.frame_count)
        dm = runtime.DeviceModel(ctx, model)
        ref = runtime.forward_stack(model, spec, pupil, context=ctx, dmodel=dm)
        d_spec = runtime.ResidentBuffer("complex", spec.size, ctx)
        d_spec.upload(spec)
        d_pup = None
        if pupil is not None:
            d_pup = runtime.ResidentBuffer("complex", pupil.size, ctx)
            d_pup.upload(pupil)
        d_out = runtime.ResidentBuffer("real", model.frame_count * model.image_len, ctx)
        d_out.upload(np.full(model.frame_count * model.image_len, -1.0))
        ctx.synchronize()
        free0 = device_free_bytes()
        runtime.forward_stack_noalloc(ctx, dm, d_spec.device_address, 0 if d_pup is None else d_pup.device_address,
                                      d_out.device_address)
        ctx.synchronize()
        free1 = device_free_bytes()
        got = d_out.download().reshape(ref.shape)
        say("max abs difference", float(np.abs(got - ref).max()), "free memory change", free0 - free1)
        assert np.array_equal(got, ref)
        assert free1 == free0
        for b in (d_spec, d_pup, d_out):
            if b is not None:
                b.close()
        dm.close()


def test_noalloc_forward_stack_rejects_what_it_cannot_take(ctx):
    """Multiplexed models are Unsupported; a null output is InvalidParameter."""
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(3), (32, 32))
    model.with_multiplexing([[(0, 0.5), (1, 0.5)]] + [[(s, 1.0)] for s in range(2, model.source_count)])
    dm = runtime.DeviceModel(ctx, model)
    with pytest.raises(runtime.FpmError) as e:
        runtime.forward_stack_noalloc(ctx, dm, 1, 0, 1)
    assert e.value.kind == "Unsupported"
    with pytest.raises(runtime.FpmError) as e:
        runtime.forward_stack_noalloc(ctx, dm, 1, 0, 0)
    assert e.value.kind == "InvalidParameter"
    dm.close()
