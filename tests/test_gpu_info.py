"""GPU: eup_get_info's device_bytes against the device memory that eup_create actually takes.

The drop in free device memory (hipMemGetInfo through torch) across one eup_create is an
independent measurement of what the handle allocated; device_bytes must account for it up to
the allocator's rounding.
"""
import sys

import pytest

pytestmark = pytest.mark.gpu

MiB = 1 << 20


@pytest.mark.parametrize("mode,storage,shape", [("physics", "f32", (128, 64, 64)),
                                                ("physics", "f16", (128, 64, 64)),
                                                ("parity", "f64", (96, 48, 48))])
def test_device_bytes_matches_the_free_memory_drop(mode, storage, shape):
    import torch
    from euplectella_hpc_amd.lib import Lattice
    print(f"[gpu] device_bytes {mode} {storage} {shape}", flush=True)
    sys.stdout.flush()
    torch.cuda.init()
    kw = dict(svisc=0.02, u_inflow=0.05, mode=mode, storage=storage)
    with Lattice(*shape, **kw):      # first handle of the process: stream and queue set-up
        pass
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info(0)
    with Lattice(*shape, **kw) as lat:
        free1, _ = torch.cuda.mem_get_info(0)
        held = lat.info()["device_bytes"]
    drop = free0 - free1
    assert held > 0
    assert 0.95 * held - 8 * MiB <= drop <= held + 64 * MiB, (held, drop)
