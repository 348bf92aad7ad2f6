"""The reference's compute-pipeline sample as an end-to-end test on the GPU,
through the C ABI (samples/core/compute-pipeline/main.cpp:66-217).

A synthetic sensor log (increasing timestamps, sensor ids, temperatures in
tenths of a degree with rare +60.0 C spikes, the sample's formulas) goes
through the sample's Compute steps: calibration offsets per sensor (the
sample looks up id -> id % 7 in a device hash map, a component outside this
library; here the same offsets come from a transform of the ids), calibrate
in place (transform2 "x + y"), hottest reading (reduce with
ops::maximum), histogram_even into 12 bins over [0, 1200), alarms
(copy_if "x >= 700u"), Fahrenheit (transform "x * 0.18f + 32.0f"), and the
readings ranked hottest first (sequence + stable sort_by_key, descending).
Every result is verified on the host as the sample does (:171-183), and
more strictly: the full alarm list, the Fahrenheit array, and the whole
ranking against a stable descending argsort."""
import numpy as np
import pytest

import clc_hip
from gpu_helpers import dev, host, say

pytestmark = pytest.mark.gpu


def sensor_log(n, seed=42):
    """The sample's generator (main.cpp:91-100) with numpy's MT19937 in
    place of std::mt19937 (the same engine; the stream of draws differs)."""
    rng = np.random.RandomState(seed)
    raw = rng.randint(0, 2**32, size=(n, 5), dtype=np.uint64)
    timestamps = (1_700_000_000 + np.cumsum(raw[:, 0] % 20)).astype(np.uint32)
    sensors = (raw[:, 1] % 1000).astype(np.uint32)
    readings = (150 + raw[:, 2] % 300 + np.where(raw[:, 3] % 1000 == 0, 600, 0)).astype(np.uint32)
    return timestamps, sensors, readings


@pytest.mark.parametrize("n", [1, 1000, 1_000_000])
def test_compute_pipeline(n):
    import torch
    say("compute-pipeline", n)
    ctx = clc_hip.Context()
    try:
        _, sensors, readings = sensor_log(n)
        d_readings = dev(readings)
        d_sensors = dev(sensors)

        # 3. Calibration offsets per sensor id (id % 7), on the device.
        d_offsets = torch.empty(n, dtype=torch.int32, device="cuda")
        ctx.transform(d_sensors, d_offsets, "x % 7u", in_type=clc_hip.U32,
                      out_type=clc_hip.U32)
        # 4. Calibrate in place, then analyse.
        ctx.transform2(d_readings, d_offsets, d_readings, "x + y", type1=clc_hip.U32,
                       type2=clc_hip.U32, out_type=clc_hip.U32)
        calibrated = readings + sensors % 7

        hottest = ctx.reduce(d_readings, "((a < b) ? b : a)", "0", 0, type_code=clc_hip.U32)

        hist = torch.zeros(12, dtype=torch.int32, device="cuda")
        ctx.histogram_even(d_readings, hist, 0.0, 1200.0, type_code=clc_hip.U32)

        d_alarms = torch.empty(n, dtype=torch.int32, device="cuda")
        alarms = ctx.copy_if(d_readings, d_alarms, "x >= 700u", type_code=clc_hip.U32)

        d_fahrenheit = torch.empty(n, dtype=torch.float32, device="cuda")
        ctx.transform(d_readings, d_fahrenheit, "x * 0.18f + 32.0f", in_type=clc_hip.U32,
                      out_type=clc_hip.F32)

        d_order = torch.empty(n, dtype=torch.int32, device="cuda")
        ctx.sequence(d_order, 0, 1, type_code=clc_hip.U32)
        d_keys = d_readings.clone()
        ctx.sort(d_keys, d_order, descending=True, key_type=clc_hip.U32)
        top = host(d_order, np.uint32)

        # Host reference, as the sample's (:153-183).
        ref_hottest = int(calibrated.max())
        ref_alarms = int((calibrated >= 700).sum())
        ref_hist = np.zeros(12, dtype=np.int64)
        np.add.at(ref_hist, calibrated[calibrated < 1200] // 100, 1)
        assert int(hottest) == ref_hottest
        assert alarms == ref_alarms
        np.testing.assert_array_equal(host(hist, np.uint32), ref_hist)
        assert calibrated[top[0]] == ref_hottest

        # Stricter than the sample.
        np.testing.assert_array_equal(host(d_readings, np.uint32), calibrated)
        np.testing.assert_array_equal(host(d_alarms, np.uint32)[:alarms],
                                      calibrated[calibrated >= 700])
        np.testing.assert_array_equal(
            host(d_fahrenheit, np.float32),
            calibrated.astype(np.float32) * np.float32(0.18) + np.float32(32.0))
        order = np.argsort(-calibrated.astype(np.int64), kind="stable").astype(np.uint32)
        np.testing.assert_array_equal(top, order)
        np.testing.assert_array_equal(host(d_keys, np.uint32), calibrated[order])

        stats = ctx.program_stats()
        say("programs built", stats.builds, "memory hits", stats.memory_hits)
        assert stats.builds == 4  # x % 7u, x + y, copy_if, Fahrenheit
    finally:
        ctx.close()
