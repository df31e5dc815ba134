"""GPU training path: shapeInfo_producer's warp (bit-exact with the oracle's fixed-point warpAffine)
and addTemplate over its infos, one by one and batched, against the oracle and the reference's
trained fixture banks."""
import numpy as np
import pytest

import oracle
import shape_based_matching_amd as sbm
from helpers import golden, load_fixture_detector

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def png_bgr(name):
    from PIL import Image
    rgb = np.asarray(Image.open(golden(name)).convert("RGB"))
    return np.ascontiguousarray(rgb[:, :, ::-1])


def levels_equal(got, ref):
    for g, r in zip(got, ref):
        if (g.width, g.height, g.tl_x, g.tl_y) != (r["width"], r["height"], r["tl_x"], r["tl_y"]):
            return False
        if not np.array_equal(g.features, r["features"]):
            return False
        if not np.array_equal(g.theta.view(np.uint32), np.asarray(r["theta"], np.float32).view(np.uint32)):
            return False
    return True


@pytest.mark.parametrize("shape", [(148, 147, 3), (33, 50), (470, 470, 3), (64, 1)])
def test_warp_equals_oracle(shape):
    rng = np.random.default_rng(sum(shape))
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    if len(shape) == 3 or shape[1] > 1:
        img[: shape[0] // 2] = 200  # a flat half and an edge as well as noise
    say("warp", shape)
    for angle, scale in [(0, 1), (90, 1), (180, 1), (37.25, 1), (-73.5, 0.6), (12, 1.7), (300, 0.1), (1, 2.0)]:
        got = sbm.shapeInfo_producer.transform(img, angle, scale)
        ref = oracle.shape_transform(img, angle, scale)
        assert np.array_equal(got, ref), (angle, scale)
        m = sbm.shapeInfo_producer.transform(img[..., 0] if img.ndim == 3 else img, angle, scale, binary_mask=True)
        assert np.array_equal(m, oracle.mask_of(img[..., 0] if img.ndim == 3 else img, angle, scale)), (angle, scale)


def test_case2_flow_one_by_one_and_batched_equal_oracle():
    """noise_test's training loop (test.cpp:385-405) on the GPU: 361 rotations of case2/train.png."""
    bgr = png_bgr("case2_train.png")
    shapes = sbm.shapeInfo_producer(bgr, np.full(bgr.shape[:2], 255, np.uint8))
    shapes.angle_range = [0, 360]
    shapes.angle_step = 1
    shapes.produce_infos()
    assert len(shapes.infos) == 361
    det = sbm.Detector(30, [4, 8])
    one = [det.addTemplate(shapes.src_of(i), "test", shapes.mask_of(i)) for i in shapes.infos]
    det2 = sbm.Detector(30, [4, 8])
    ids = det2.add_templates_transformed(bgr, "test", shapes.infos, mask=shapes.mask)
    say("case2: one-by-one ids", one[:3], "... batched", ids[:3])
    assert one == ids == list(range(361))
    for k, info in enumerate(shapes.infos):
        ref = oracle.add_template(oracle.shape_transform(bgr, info.angle, info.scale), 30, 2,
                                  mask=oracle.mask_of(shapes.mask, info.angle, info.scale))
        assert levels_equal(det.getTemplates("test", k), ref), k
        assert levels_equal(det2.getTemplates("test", k), ref), k


def test_case0_scale_flow_with_refusals_equals_oracle():
    """scale_test's training loop (test.cpp:128-161): num_features = int(150 * scale), and templates
    the reference refuses come back as -1 in both paths."""
    bgr = png_bgr("case0_circle.png")
    shapes = sbm.shapeInfo_producer(bgr)
    shapes.scale_range = [0.1, 1.0]
    shapes.scale_step = 0.01
    shapes.produce_infos()
    nf = [int(np.float32(150) * np.float32(i.scale)) for i in shapes.infos]
    det = sbm.Detector(150, [4, 8])
    ids = det.add_templates_transformed(bgr, "circle", shapes.infos, num_features=nf)
    expect, k = [], 0
    for info, n in zip(shapes.infos, nf):
        ref = oracle.add_template(oracle.shape_transform(bgr, info.angle, info.scale), n, 2,
                                  mask=oracle.mask_of(shapes.mask, info.angle, info.scale))
        if ref is None:
            expect.append(-1)
            continue
        expect.append(k)
        assert levels_equal(det.getTemplates("circle", k), ref), info
        k += 1
    say("case0:", len(shapes.infos), "infos,", k, "templates")
    assert ids == expect
    # the same flow one template at a time
    det1 = sbm.Detector(150, [4, 8])
    one = [det1.addTemplate(shapes.src_of(i), "circle", shapes.mask_of(i), n) for i, n in zip(shapes.infos, nf)]
    assert one == ids


def test_case1_angle_flow_reproduces_fixture_bank():
    """angle_test's training (test.cpp:230-292): addTemplate on the padded crop at 0 degrees through
    shapeInfo_producer, then addTemplate_rotate -- every template equals the reference's file."""
    bgr = png_bgr("case1_train.png")[110:380, 130:400]
    pad = np.zeros((470, 470, 3), np.uint8)
    pad[100:370, 100:370] = bgr
    mask = np.zeros((470, 470), np.uint8)
    mask[100:370, 100:370] = 255
    shapes = sbm.shapeInfo_producer(pad, mask)
    shapes.angle_range = [0, 360]
    shapes.angle_step = 1
    shapes.scale_range = [1]
    shapes.produce_infos()
    det = sbm.Detector(128, [4, 8])
    first = shapes.infos[0]
    tid = det.addTemplate(shapes.src_of(first), "test", shapes.mask_of(first))
    assert tid == 0
    for info in shapes.infos[1:]:
        det.addTemplate_rotate("test", 0, info.angle - first.angle, (shapes.src.shape[1] / 2.0, shapes.src.shape[0] / 2.0))
    fx = load_fixture_detector(1)
    assert det.numTemplates("test") == fx.numTemplates("test") == 361
    for k in range(361):
        for a, b in zip(det.getTemplates("test", k), fx.getTemplates("test", k)):
            assert np.array_equal(a.features, b.features), k
            assert (a.width, a.height, a.tl_x, a.tl_y) == (b.width, b.height, b.tl_x, b.tl_y), k
