"""CPU checks of bench.py's contract: it refuses more than one GPU before
touching a device, and its roofline byte count for the headline workload
(225 LEDs x 128^2 tiles, 375^2 object, momentum interval 30) is the per-LED
and per-event sum written out by hand."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


@pytest.mark.parametrize("gpus", [2, 8])
def test_more_than_one_gpu_exits_with_message(gpus):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(gpus)],
                       capture_output=True, text=True, timeout=300, cwd=ROOT,
                       env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and f"--gpus {gpus}" in r.stderr
    assert r.stdout == ""


def test_headline_sweep_bytes_by_hand():
    import bench
    from fpmhip import runtime
    from fpmhip import synthetic as syn

    model = runtime.compile_model(syn.base_optics(), syn.planar_array(15), (128, 128))
    assert model.frame_count == 225 and tuple(model.reconstruction_shape) == (375, 375)
    # per LED: uint16 frame, c128 object window read and written, c64 pupil
    per_led = 128 * 128 * (2 + 2 * 16 + 8)
    # 50 sweeps x 225 LEDs / 30 = 375 momentum events, 7.5 per sweep; each
    # reads and writes the c128 object and the c64 velocity and anchor
    per_event = 375 * 375 * (2 * 16 + 4 * 8)
    expect = 225 * per_led + 7.5 * per_event
    assert expect == 222_328_800
    got = bench.sweep_bytes(model, 50, 30, 2, 16, 8, 8)
    assert got == pytest.approx(expect, rel=1e-12)
    # whole tile: 50 sweeps plus the init read of every frame and four
    # object-sized init/result passes
    whole = bench.tile_bytes(model, 50, 30, 2, 16, 8, 8)
    assert whole == 50 * 225 * per_led + 375 * per_event + 225 * 128 * 128 * 2 + 4 * 375 * 375 * 16


def test_traffic_line_matches_counter_table():
    """The profile's machine-readable traffic line is FETCH_SIZE x 2 +
    WRITE_SIZE x 1 (KiB per 10-sweep dispatch, from its own table) per sweep."""
    import bench

    text = open(bench.TRAFFIC_PROFILE).read()
    fetch = int(re.search(r"FETCH_SIZE ([\d,]+) KiB x 2\.000", text).group(1).replace(",", ""))
    write = int(re.search(r"WRITE_SIZE ([\d,]+) KiB x 1\.000", text).group(1).replace(",", ""))
    line = re.search(r"^traffic_bytes_per_sweep: (\d+) lib_sha256: ([0-9a-f]{64})$", text, re.M)
    assert line, "traffic line missing"
    per_sweep = (2 * fetch + write) * 1024 / 10
    assert float(line.group(1)) == pytest.approx(per_sweep, rel=1e-6)
    # bench.py reports it only for the library build it was measured on
    got = bench.measured_traffic()
    assert got is None or got == float(line.group(1))
