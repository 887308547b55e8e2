"""Batching, HIP-graph replay and the autotuner change how the work is
launched, never the estimator: every form must reproduce the plain eager
render bit for bit (the RNG is keyed by pixel and frame only, and every pixel
folds its frames in order)."""
import numpy as np
import pytest

import asuna_amd as A

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def renderer():
    r = A.Renderer(0)
    yield r
    r.close()


def _load(r, path):
    s = A.HostScene(path)
    r.upload(s)
    r.build()
    r.set_film(*s.film)
    return s


def _render(r, s, frames, **kw):
    cam, pc = s.shot(0)
    r.set_film(*s.film)
    st = r.render(cam, pc, 0, frames, **kw)
    return [r.read(i) for i in (0, 1, 8)], st


def test_graph_replays_equal_eager(renderer, spheres_small):
    """Each replay of the captured bounce loop, with eager work and
    synchronisations between replays, equals the eager render."""
    s = _load(renderer, spheres_small)
    ref, st0 = _render(renderer, s, 6, graphs=False, queue_paths=1 << 16)
    assert st0["graph_launches"] == 0 and st0["batches"] > 1
    for rep in range(3):
        got, st = _render(renderer, s, 6, queue_paths=1 << 16)
        log(f"replay {rep}: batches {st['batches']}, graph launches {st['graph_launches']}")
        assert st["graph_launches"] == st["batches"]
        for a, b in zip(ref, got):
            np.testing.assert_array_equal(a, b)
        _render(renderer, s, 1, graphs=False)  # eager work in between
    assert st["extension_rays"] == st0["extension_rays"] and st["shadow_rays"] == st0["shadow_rays"]


def test_fused_trace_launch_counts(renderer, spheres_small):
    """The default scheme traces each bounce's shadow rays with the next
    bounce's extension rays in one launch: same images and ray counts as the
    separate launches, and the fused launches' own counters add up (every
    path's first extension ray is traced before the first fused launch)."""
    s = _load(renderer, spheres_small)
    cam, pc = s.shot(0)
    pc.maxPathDepth = 4
    kw = dict(queue_paths=1 << 16, count_traversal=True)
    renderer.set_film(*s.film)
    fused = renderer.render(cam, pc, 0, 2, traversal=0, **kw)
    img_f = [renderer.read(i) for i in (0, 1, 8)]
    renderer.set_film(*s.film)
    sep = renderer.render(cam, pc, 0, 2, traversal=1, **kw)
    img_s = [renderer.read(i) for i in (0, 1, 8)]
    log(f"fused: {fused}")
    for a, b in zip(img_f, img_s):
        np.testing.assert_array_equal(a, b)
    assert fused["extension_rays"] == sep["extension_rays"] and fused["shadow_rays"] == sep["shadow_rays"]
    assert sep["trace_launches"] == 0
    assert fused["batches"] > 1 and fused["extend_launches"] == fused["batches"]
    assert fused["trace_launches"] == fused["batches"] * (pc.maxPathDepth - 1)
    assert fused["trace_ext_rays"] == fused["extension_rays"] - fused["paths"]
    assert 0 < fused["trace_shadow_rays"] <= fused["shadow_rays"]
    assert 0 < fused["trace_ext_node_visits"] < fused["ext_node_visits"]
    assert 0 < fused["trace_shadow_node_visits"] <= fused["shadow_node_visits"]
    renderer.set_film(*s.film)
    tm = renderer.render(cam, pc, 0, 2, traversal=0, queue_paths=1 << 16, timing=True)
    assert tm["trace_ms"] > 0 and tm["trace_launches"] == fused["trace_launches"]


def test_batch_plans_are_equivalent(renderer, multishot_scene):
    """64 shots x 2 frames split three ways: several frames per batch, groups
    of whole shots per batch, and pixel runs of one shot per batch."""
    s = _load(renderer, multishot_scene)
    n = s.state.num_shots
    cams, pcs = zip(*[s.shot(i) for i in range(n)])
    w, h = s.film
    outs = []
    for q in (1 << 26, 1 << 22, 1 << 18, 1 << 16):
        st = renderer.render_shots(list(cams), list(pcs), 2, queue_paths=q)
        log(f"queue {q}: {st['batches']} batches, {st['render_ms']:.1f} ms")
        outs.append([renderer.read_shot(i, img) for i in (0, 31, 63) for img in (0, 8)])
        assert st["paths"] == n * w * h * 2
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("branching", [2, 4])
def test_traversal_fetch_schemes_render_identically(renderer, spheres_small, branching):
    """Per-XCD queue slices (dynamic per-lane fetch, speculative
    while-while rounds), wave fetch in queue order and per-lane refill
    change only which lane traces which ray and when leaves are tested:
    every image is bit-identical and the ray counts are the same."""
    s = A.HostScene(spheres_small)
    renderer.upload(s)
    renderer.build(branching=branching)
    cam, pc = s.shot(0)
    w, h = s.film
    imgs, stats = [], []
    for trav in (0, 1, 2):
        for lds in (8, 12):
            renderer.set_film(w, h)
            st = renderer.render(cam, pc, 0, 3, traversal=trav, lds_stack=lds, count_traversal=True,
                                 blocks_per_cu=4)
            imgs.append([renderer.read(i) for i in (0, 1, 2, 8)])
            stats.append(st)
            log(f"branching {branching} traversal {trav} lds {lds}: nodes {st['ext_node_visits']} "
                f"tris {st['ext_tri_tests']}")
    for other in imgs[1:]:
        for a, b in zip(imgs[0], other):
            np.testing.assert_array_equal(a, b)
    for st in stats[1:]:
        assert st["extension_rays"] == stats[0]["extension_rays"] and st["shadow_rays"] == stats[0]["shadow_rays"]
    # graph replays with each scheme equal the eager render
    for trav in (0, 1, 2):
        renderer.set_film(w, h)
        renderer.render(cam, pc, 0, 3, traversal=trav, graphs=True)
        np.testing.assert_array_equal(renderer.read(0), imgs[0][0])


def test_autotune_picks_a_candidate_and_keeps_results(renderer, multishot_scene):
    s = _load(renderer, multishot_scene)
    cams, pcs = zip(*[s.shot(i) for i in range(8)])
    opts, rep = renderer.autotune(list(cams), list(pcs), probe_frames=1)
    log(f"autotune: {rep}")
    # queue sizes 2^18, 2^19, ... up to the first that holds the whole probe in one batch
    w, h = s.film
    top = next(k for k in range(18, 27) if (1 << k) >= w * h * 8 or k == 26)
    assert rep["candidates"] == 6 * (top - 17) + 2
    assert opts.queue_paths in [1 << k for k in range(18, top + 1)] and opts.blocks_per_cu in (4, 8, 16)
    assert opts.lds_stack in (8, 12) and rep["lds_stack"] == opts.lds_stack
    assert opts.traversal in (0, 1, 2) and rep["traversal"] == opts.traversal
    assert 0 < rep["best_ms"] <= rep["worst_ms"]
    renderer.render_shots(list(cams), list(pcs), 2, queue_paths=opts.queue_paths, blocks_per_cu=opts.blocks_per_cu,
                          lds_stack=opts.lds_stack, traversal=opts.traversal)
    tuned = [renderer.read_shot(i, 0) for i in range(8)]
    renderer.render_shots(list(cams), list(pcs), 2, graphs=False)
    for i in range(8):
        np.testing.assert_array_equal(tuned[i], renderer.read_shot(i, 0))
