"""ymrt_render_frame_noalloc against ymrt_render_frame, through the C ABI.

The entry point that allocates no device memory runs the same kernels over a
path state carved out of the library's fixed device array, in chunks of at
most 65536 paths; frames do not depend on the chunking (every path writes its
own output slot and the per-pixel resolve sums samples in a fixed order), so
its images must equal ymrt_render_frame's bit for bit.  Two calls in a row,
the second on a smaller frame, check that state left in the shared array by
one call does not leak into the next.
"""
import numpy as np
import pytest
import torch

import ymrt_amd
from ymrt_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _frames(sc, rs, frame):
    ref, st_ref = sc.render_frame(rs, frame)
    got, st_got = sc.render_frame_noalloc(rs, frame)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy(), st_ref, st_got


def test_noalloc_equals_render_frame(gpu_ctx):
    # media, dielectrics and boundaries: every pass kernel runs
    sa = scenes.cornell_media(96, 80, grid=True)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    # 96x80 at 16 spp = 122880 path slots: two chunks of the fixed capacity
    rs = abi.render_settings(96, 80, 16, 6, sampler=abi.SAMPLER_HALTON, volume=1)
    ref, got, st_ref, st_got = _frames(sc, rs, 1)
    assert np.isfinite(ref).all()
    assert np.array_equal(ref.view(np.uint32), got.view(np.uint32))
    assert st_got.chunks >= 2
    assert st_ref.extension_rays == st_got.extension_rays
    assert st_ref.shadow_rays == st_got.shadow_rays
    # second call, smaller frame, other frame index: nothing of the first call
    # may show through
    rs2 = abi.render_settings(40, 24, 4, 6, sampler=abi.SAMPLER_HALTON, volume=1)
    ref2, got2, st_ref2, st_got2 = _frames(sc, rs2, 2)
    assert np.array_equal(ref2.view(np.uint32), got2.view(np.uint32))
    assert st_ref2.extension_rays == st_got2.extension_rays
    sc.close()


def test_noalloc_cornell_and_limits(gpu_ctx):
    sa = scenes.cornell_box(64, 64)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    rs = abi.render_settings(64, 64, 8, 8, sampler=abi.SAMPLER_PCG)
    ref, got, _, _ = _frames(sc, rs, 3)
    assert np.array_equal(ref.view(np.uint32), got.view(np.uint32))
    # sizes beyond the fixed capacity are refused with a message
    bad = abi.render_settings(64, 64, 8, 8)
    bad.path_capacity = 1 << 17
    with pytest.raises(ymrt_amd.YmrtError, match="fixed capacity"):
        sc.render_frame_noalloc(bad, 1)
    bad = abi.render_settings(64, 64, 8, 8)
    bad.path_capacity = 4
    with pytest.raises(ymrt_amd.YmrtError, match="one pixel's samples"):
        sc.render_frame_noalloc(bad, 1)
    bad = abi.render_settings(64, 64, 8, 8)
    bad.persistent_waves = 2048
    with pytest.raises(ymrt_amd.YmrtError, match="persistent_waves"):
        sc.render_frame_noalloc(bad, 1)
    sc.close()
