"""The area LUT: the only reference-held fixture on this path.

pathfinder_amd/data/area_lut.rgba holds the pixels of the reference's
resources/textures/area-lut.png (256x256 RGBA8). It is pinned three ways:
its hash, the PNG itself when the reference tree is present, and the
closed-form generator of utils/area-lut/src/main.rs:29-96 restated in binary32.
"""
import hashlib
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LUT = os.path.join(ROOT, "pathfinder_amd", "data", "area_lut.rgba")
PNG = "/root/reference/resources/textures/area-lut.png"
SHA256 = "2352b0ec6b5ba601bb46ac5a3b1ae6106903670ae7a80add38f7ab2ecb9856fc"


def lut():
    return np.fromfile(LUT, np.uint8).reshape(256, 256, 4)


def test_lut_hash():
    assert hashlib.sha256(open(LUT, "rb").read()).hexdigest() == SHA256


@pytest.mark.skipif(not os.path.exists(PNG), reason="reference tree not mounted")
def test_lut_matches_reference_png():
    from PIL import Image
    png = np.asarray(Image.open(PNG).convert("RGBA"))
    assert np.array_equal(png, lut())


def area_closed_form(y, dydx):
    """utils/area-lut/src/main.rs:29-65 area(), vectorised in binary32."""
    f = np.float32
    y = y.astype(f)
    dydx = dydx.astype(f)
    xl, xr = f(-0.5), f(0.5)
    yl = dydx * xl + y
    yr = dydx * xr + y
    with np.errstate(divide="ignore", invalid="ignore"):
        m = (yr - yl) / (xr - xl)

        def solve(yy):
            return xl - (yl - f(yy)) / m

        p2x = solve(-0.5)
        p4x = solve(0.5)

    def tri(ax, ay, bx, by):
        return f(0.5) * (bx - ax) * (ay - by)

    def rect(ax, ay, bx, by):
        return (bx - ax) * (ay - by)

    with np.errstate(invalid="ignore"):
        t01 = tri(xl, yl, xr, yr)
        t21 = tri(p2x, f(-0.5), xr, yr)
        r07 = rect(xl, yl, xr, f(0.5))
        t04 = tri(xl, yl, p4x, f(0.5))
        r03 = rect(xl, yl, xr, f(-0.5))
        case0 = ((t01 - t21) - r07) + t04
        case6 = (t01 - r07) + t04
        case1 = (t01 - t21) - r07
        case4 = t01 - r07
        case2 = -r07 + r03
    out = np.where(yl > 0.5,
                   np.where(yr < -0.5, case0, np.where(yr < 0.5, case6, f(0.0))),
                   np.where(yl > -0.5, np.where(yr < -0.5, case1, case4), case2))
    return out.astype(f)


def generated_lut():
    """utils/area-lut/src/main.rs:77-96: u = column, v = row."""
    u, v = np.meshgrid(np.arange(256), np.arange(256))
    y = (u.astype(np.float32) - np.float32(128)) / np.float32(16)
    dydx = -(v.astype(np.float32)) / np.float32(16)
    out = np.zeros((256, 256, 4), np.uint8)
    for k in range(4):
        a = area_closed_form(y - np.float32(k), dydx) * np.float32(255)
        r = np.sign(a) * np.floor(np.abs(a) + np.float32(0.5))  # f32::round: half away from 0
        out[..., k] = np.clip(np.nan_to_num(r), 0, 255).astype(np.uint8)  # `as u8` saturates
    out[:, 0] = 255
    out[:, 255] = 0
    return out


def test_lut_matches_closed_form():
    # every texel of the fixture is the generator's binary32 result
    assert np.array_equal(lut(), generated_lut())


def test_lut_is_monotone_in_offset():
    # more of the pixel lies below the line as the line moves up (u grows -> y grows)
    ref = lut().astype(int)
    assert np.all(np.diff(ref[:, 1:255, :], axis=1) <= 1)
