"""ash_render_noalloc (the allocation-free render) against ash_render, through
the C ABI, on two calls in a row: the second with a smaller film and its own
frame count.

Its output buffer is a second context's film (device memory the library
allocated at set_film), so the call under test allocates nothing.  Bar:
every one of the 9 output planes bit-identical to ash_render's film for the
same camera, push constants and frames (integer-exact float comparison).

Call 1: 640x480 (more pixels than the 2^18-path scratch queue, so each frame
runs as two pixel runs), frames 0..3.  Call 2: 100x37 (3700 pixels, not a
multiple of 64), frames 0..40 in one batch of 41 frames.
"""
import numpy as np
import pytest

import asuna_amd as A

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def _reference(r, cam, pc, w, h, n):
    r.set_film(w, h)
    r.render(cam, pc, 0, n)
    return np.stack([r.read(k) for k in range(A.ASH_NUM_OUTPUT_IMAGES)])


def test_noalloc_render_equals_render_on_two_calls(box_scene):
    s = A.HostScene(box_scene)
    _, pc = s.shot(0)
    r = A.Renderer(0)
    out = A.Renderer(0)
    try:
        r.upload(s)
        r.build()
        for (w, h, n) in ((640, 480, 4), (100, 37, 41)):
            log(f"noalloc call: film {w}x{h}, frames 0..{n - 1}")
            cam = A.camera_lookat(w, h, 45.0, (0.0, 0.0, 3.4), (0.0, 0.0, 0.0))
            ref = _reference(r, cam, pc, w, h, n)
            out.set_film(w, h)
            st = r.render_noalloc(cam, pc, 0, n, w, h, out.output_device_ptr(0))
            got = np.stack([out.read(k) for k in range(A.ASH_NUM_OUTPUT_IMAGES)])
            log(f"  batches {st['batches']}, paths {st['paths']}, extension rays {st['extension_rays']}")
            assert st["paths"] == w * h * n
            assert ref[0, ..., :3].mean() > 0
            for k in range(A.ASH_NUM_OUTPUT_IMAGES):
                assert np.array_equal(got[k], ref[k]), f"image {k} differs on the {w}x{h} call"
    finally:
        out.close()
        r.close()


def test_noalloc_rejects_a_film_above_its_limit(box_scene):
    """A film of more than 2^30 pixels is refused before any device access
    (the buffer address is never touched)."""
    s = A.HostScene(box_scene)
    cam, pc = s.shot(0)
    r = A.Renderer(0)
    try:
        r.upload(s)
        r.build()
        with pytest.raises(A.AsunaError, match="2\\^30"):
            r.render_noalloc(cam, pc, 0, 1, 40000, 40000, 0x1000)
    finally:
        r.close()
