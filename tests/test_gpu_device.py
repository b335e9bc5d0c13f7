"""GPU tests of the device measurement helpers (include/rqcopt/rqc_device.h; no reference
counterpart): the plain and the tuned streaming copy report a bandwidth between zero and the
MI355X's 8 TB/s HBM peak, and reject invalid arguments with the reference's -3.
"""
import pytest

import rqcopt_hpc_amd as R

pytestmark = pytest.mark.gpu

HBM_PEAK_GBS = 8000.0


@pytest.fixture(autouse=True)
def gpu():
    R.require_gpu()
    yield


def test_copy_and_stream_bandwidth_within_peak():
    copy = R.measure_copy_bandwidth(1 << 28, 3)
    stream = R.measure_stream_bandwidth(1 << 28, 3)
    print(f"copy {copy:.0f} GB/s, tuned stream {stream:.0f} GB/s", flush=True)
    assert 0 < copy < 1.05 * HBM_PEAK_GBS
    assert 0 < stream < 1.05 * HBM_PEAK_GBS
    # sizes that are not a whole number of the tuned kernel's four-element chunks
    assert R.measure_stream_bandwidth((1 << 20) + 48, 1) > 0


def test_bandwidth_rejects_invalid_arguments():
    for fn in (R.measure_copy_bandwidth, R.measure_stream_bandwidth):
        with pytest.raises(R.RqcError):
            fn(1 << 20, 0)
        with pytest.raises(R.RqcError):
            fn(8, 1)
