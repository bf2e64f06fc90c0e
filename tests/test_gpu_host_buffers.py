"""Host-side buffers of the Python view of the C ABI: outputs(into=...) fills
arrays from output_buffers() with exactly what outputs() returns, from
page-locked (pinned_empty) or ordinary memory, and the signal can be loaded
from page-locked memory."""
import numpy as np
import pytest

from fdt_amd import synth

pytestmark = pytest.mark.gpu


def test_outputs_into_pinned_buffers_equal_outputs(fdt):
    v = synth.voxels("c1", 80)
    x = fdt.Xfibres(nfibres=3, model=2, nburn=50, njumps=100, sampleevery=25)
    try:
        x.set_acquisition(v["bvals"], v["bvecs"])
        host = fdt.pinned_empty(v["data"].shape)
        host[...] = v["data"]
        x.load_data(host)
        x.run(0)
        want = x.outputs()
        pinned = x.output_buffers(alloc=fdt.pinned_empty)
        plain = x.output_buffers()
        assert set(pinned) == set(want)
        for bufs in (pinned, plain):
            got = x.outputs(into=bufs)
            assert got is bufs
            for k in want:
                np.testing.assert_array_equal(got[k], want[k], err_msg=k)
        with pytest.raises(ValueError):
            x.outputs(into={k: a[..., :1] for k, a in plain.items()})
        # the same run from the ordinary array gives the same outputs
        x.load_data(v["data"])
        x.run(0)
        again = x.outputs()
        for k in want:
            np.testing.assert_array_equal(again[k], want[k], err_msg=k)
    finally:
        x.close()
