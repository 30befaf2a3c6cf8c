"""The oracle's paint functions (calculateColor after the mask) against the
float64 models of paint_cases: colour transform order, GL_LINEAR/GL_NEAREST
with REPEAT and CLAMP_TO_EDGE, the radial gradient's roots, the Gaussian
blur, the colour matrix, SrcIn/DestIn and the fifteen composite ops. Each
planted mistake (ORC_OPT_PLANT_PAINT) must fail the pin that names it, and
the metadata encoding's binary16 rounding must match numpy's."""
import numpy as np
import pytest

import oracle
import paint_cases as pc

CASES = pc.cases()


def _report(case, fb):
    valid = case["valid"]
    exp = pc.to_u8(np.where(valid[..., None], case["expect"], 0.0))
    d = np.abs(fb - exp)[valid]
    return int(d.max()) if d.size else 0, int(valid.sum())


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_paint_matches_model(name):
    case = CASES[name]
    w, h = case["size"]
    fb = pc.oracle_render(case["cmds"], w, h)
    worst, n = _report(case, fb)
    assert n >= case.get("min_valid", 0.9) * w * h, (name, n)
    assert worst <= case["tol"], (name, worst)


PLANTED = sorted((name, p) for name, c in CASES.items() for p in c["plants"])


@pytest.mark.parametrize("name,plant", PLANTED)
def test_planted_paint_mistake_fails_its_pin(name, plant):
    case = CASES[name]
    w, h = case["size"]
    fb = pc.oracle_render(case["cmds"], w, h, plant=plant)
    worst, _ = _report(case, fb)
    assert worst > case["tol"] + 1, (name, plant, worst)


def test_every_plant_is_caught():
    assert {p for _, p in PLANTED} == set(range(1, 9))


def test_f16_encoding_matches_numpy():
    """orc_f32_to_f16 (the metadata texels' RGBA16F) rounds to nearest even,
    with subnormals, overflow to infinity and NaN as numpy's float16."""
    lib = oracle.load()
    rng = np.random.default_rng(3)
    vals = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(1000),
        rng.uniform(-1, 1, 20000).astype(np.float32) * np.float32(1e-4),
        np.float32([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, -1e9, np.inf, -np.inf,
                    5.96e-8, 2.98e-8, 2.99e-8, 6.1e-5, 1.0 / 255.0, 0.3989422804014327]),
        # halfway cases of the 10-bit mantissa
        (np.arange(1024, 2048, dtype=np.float32) + np.float32(0.5)) * np.float32(2.0 ** -10),
    ])
    got = np.asarray([lib.orc_f32_to_f16(float(v)) for v in vals], np.uint16)
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    assert (lib.orc_f32_to_f16(float("nan")) & 0x7c00) == 0x7c00
    assert lib.orc_f32_to_f16(float("nan")) & 0x3ff
