"""Seeded synthetic scenes for the five benchmark workloads.

All generators use numpy's PCG64 (``np.random.default_rng(seed)``), so a seed
fixes every byte on every host.  The rules are stated in each docstring.

  workload_match_640      parent 640x480 RGBA i.i.d. bytes, 48x48 crop
  workload_gray_ui_1080p  1920x1080 gray flat-rectangle UI scene + sigma 3 noise,
                          96x32 crop and two +-8 brightness copies
  workload_map_4k         3840x2160 RGBA i.i.d. bytes through a 5x5 box blur,
                          128x128 crop + sigma 2 noise on RGB
  workload_direct_1440p   2560x1440 gray i.i.d., 16/24/32/48 square crops
  workload_fleet          frames of 2560x1440 RGBA from a flat-rectangle canvas
                          with cumulative shifts, templates U[16,128]^2
"""
from __future__ import annotations

import numpy as np


def box_blur5(img: np.ndarray) -> np.ndarray:
    """5x5 box blur per channel: integer window sum over edge-clamped
    neighbours, divided by 25 with round-half-up ((sum + 12) // 25)."""
    a = img.astype(np.int32)
    pad = np.pad(a, ((2, 2), (2, 2)) + ((0, 0),) * (a.ndim - 2), mode="edge")
    c = np.cumsum(np.cumsum(pad, axis=0, dtype=np.int64), axis=1, dtype=np.int64)
    c = np.pad(c, ((1, 0), (1, 0)) + ((0, 0),) * (a.ndim - 2))
    h, w = a.shape[:2]
    s = c[5:5 + h, 5:5 + w] - c[0:h, 5:5 + w] - c[5:5 + h, 0:w] + c[0:h, 0:w]
    return ((s + 12) // 25).astype(np.uint8)


def add_noise(img: np.ndarray, sigma: float, rng, channels=None) -> np.ndarray:
    """Additive N(0, sigma^2), rounded half-to-even and clamped to [0, 255]."""
    out = img.astype(np.float64)
    if channels is None:
        out = out + rng.normal(0.0, sigma, size=out.shape)
    else:
        noise = rng.normal(0.0, sigma, size=out.shape[:2] + (len(channels),))
        out[..., channels] += noise
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


def flat_rect_scene(w: int, h: int, rng, n_rects=None, channels=1) -> np.ndarray:
    """UI-like canvas: a flat background value U{0..255} per channel, then
    n_rects flat rectangles (sizes U[20,400] each side, top-left uniform over
    the canvas so they clip at the edges, value U{0..255} per channel)."""
    if n_rects is None:
        n_rects = max(8, (w * h) // 6000)
    shape = (h, w) if channels == 1 else (h, w, channels)
    img = np.empty(shape, dtype=np.uint8)
    img[...] = rng.integers(0, 256, size=(channels,) if channels > 1 else None, dtype=np.uint8)
    for _ in range(n_rects):
        rw, rh = rng.integers(20, 401, size=2)
        x0 = int(rng.integers(-rw + 1, w))
        y0 = int(rng.integers(-rh + 1, h))
        v = rng.integers(0, 256, size=(channels,) if channels > 1 else None, dtype=np.uint8)
        img[max(0, y0):max(0, min(h, y0 + rh)), max(0, x0):max(0, min(w, x0 + rw))] = v
    return img


def workload_match_640(seed: int = 1, alpha: str = "random"):
    """Parent 640x480 RGBA, every byte i.i.d. U{0..255} (alpha="random", so
    cn = 4) or alpha fixed at 255 (alpha="opaque", cn = 3); 48x48 crop at
    (x, y) ~ U[0,592] x U[0,432]."""
    rng = np.random.default_rng(seed)
    parent = rng.integers(0, 256, size=(480, 640, 4), dtype=np.uint8)
    if alpha == "opaque":
        parent[..., 3] = 255
    x = int(rng.integers(0, 640 - 48 + 1))
    y = int(rng.integers(0, 480 - 48 + 1))
    sub = parent[y:y + 48, x:x + 48].copy()
    return parent, sub, (x, y)


def workload_gray_ui_1080p(seed: int = 2):
    """1920x1080 gray flat-rectangle scene plus N(0, 3^2) noise; a 96x32 crop
    at a seeded spot; two copies of the crop pasted at seeded non-overlapping
    spots with +8 and -8 brightness (clamped).  Returns (parent, sub,
    crop_xy, copies_xy)."""
    rng = np.random.default_rng(seed)
    w, h, tw, th = 1920, 1080, 96, 32
    img = add_noise(flat_rect_scene(w, h, rng), 3.0, rng)
    spots = []
    while len(spots) < 3:
        x, y = int(rng.integers(0, w - tw + 1)), int(rng.integers(0, h - th + 1))
        if all(abs(x - a) >= tw or abs(y - b) >= th for a, b in spots):
            spots.append((x, y))
    (cx, cy), (ax, ay), (bx, by) = spots
    sub = img[cy:cy + th, cx:cx + tw].copy()
    s16 = sub.astype(np.int16)
    img[ay:ay + th, ax:ax + tw] = np.clip(s16 + 8, 0, 255).astype(np.uint8)
    img[by:by + th, bx:bx + tw] = np.clip(s16 - 8, 0, 255).astype(np.uint8)
    return img, sub, (cx, cy), [(ax, ay), (bx, by)]


def workload_map_4k(seed: int = 3, alpha: str = "random"):
    """3840x2160 RGBA, bytes i.i.d. U{0..255} then box_blur5 per channel
    (alpha included for alpha="random", so cn = 4; alpha = 255 for
    "opaque"); 128x128 crop at a seeded spot with N(0, 2^2) on RGB."""
    rng = np.random.default_rng(seed)
    w, h, tw, th = 3840, 2160, 128, 128
    parent = box_blur5(rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8))
    if alpha == "opaque":
        parent[..., 3] = 255
    x, y = int(rng.integers(0, w - tw + 1)), int(rng.integers(0, h - th + 1))
    sub = add_noise(parent[y:y + th, x:x + tw], 2.0, rng, channels=[0, 1, 2])
    return parent, sub, (x, y)


def workload_direct_1440p(seed: int = 4, sizes=(16, 24, 32, 48)):
    """2560x1440 gray i.i.d. U{0..255}; square crops of the given sizes at
    seeded spots."""
    rng = np.random.default_rng(seed)
    w, h = 2560, 1440
    img = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
    subs = []
    for s in sizes:
        x, y = int(rng.integers(0, w - s + 1)), int(rng.integers(0, h - s + 1))
        subs.append((img[y:y + s, x:x + s].copy(), (x, y)))
    return img, subs


def workload_fleet(seed: int = 5, frames: int = 64, members: int = 256, w: int = 2560, h: int = 1440,
                   tmin: int = 16, tmax: int = 128):
    """A flat-rectangle RGBA canvas (alpha 255) 256 px larger than the frame on
    each axis; frame f is the crop at a cumulative offset whose steps are
    U[-16,16]^2 (clamped to the canvas).  Templates: sizes U[tmin,tmax]^2,
    each cropped from a seeded frame; every odd-numbered template gets
    N(0, 2^2) noise on RGB.  Returns (frames list, templates list)."""
    rng = np.random.default_rng(seed)
    margin = 256
    cw, ch = w + 2 * margin, h + 2 * margin
    canvas = flat_rect_scene(cw, ch, rng, n_rects=(cw * ch) // 5000, channels=4)
    canvas[..., 3] = 255
    ox, oy = margin, margin
    offs = []
    for _ in range(frames):
        offs.append((ox, oy))
        dx, dy = rng.integers(-16, 17, size=2)
        ox = int(np.clip(ox + dx, 0, cw - w))
        oy = int(np.clip(oy + dy, 0, ch - h))
    frame_list = [canvas[y:y + h, x:x + w] for (x, y) in offs]
    tpls = []
    for i in range(members):
        tw, th = (int(v) for v in rng.integers(tmin, tmax + 1, size=2))
        f = int(rng.integers(0, frames))
        x, y = int(rng.integers(0, w - tw + 1)), int(rng.integers(0, h - th + 1))
        t = np.ascontiguousarray(frame_list[f][y:y + th, x:x + tw])
        if i % 2 == 1:
            t = add_noise(t, 2.0, rng, channels=[0, 1, 2])
        tpls.append(t)
    return frame_list, tpls


# ---------------------------------------------------------- window scene
# The synthetic application window of the reference's windowrules example
# (examples/windowrules/main.go:32-105): a 1280x800 gradient with bordered
# widgets whose interiors are textured by a 32-bit LCG, so the scene is the
# same on every run.

WINDOW_RECTS = {  # (x0, y0, x1, y1), window coordinates
    "ok": (860, 700, 980, 736),
    "ok_moved": (600, 700, 720, 736),
    "wifi": (1180, 12, 1204, 36),
    "mute": (1216, 12, 1240, 36),
    "panel": (200, 200, 500, 400),
}


def draw_widget(img: np.ndarray, rect, seed: int) -> None:
    """A dark one-pixel border and an interior textured by the LCG
    s = 1664525 s + 1013904223 (mod 2^32), advanced once per pixel in raster
    order (border pixels included)."""
    x0, y0, x1, y1 = rect
    w, h = x1 - x0, y1 - y0
    n = w * h
    s = np.empty(n, dtype=np.uint64)
    v = seed & 0xFFFFFFFF
    for i in range(n):
        v = (v * 1664525 + 1013904223) & 0xFFFFFFFF
        s[i] = v
    s = s.reshape(h, w)
    blk = np.empty((h, w, 4), dtype=np.uint8)
    blk[..., 0] = 140 + ((s >> 8) & 63)
    blk[..., 1] = 120 + ((s >> 16) & 63)
    blk[..., 2] = 90 + ((s >> 24) & 63)
    blk[..., 3] = 255
    blk[0, :] = blk[-1, :] = (20, 20, 24, 255)
    blk[:, 0] = blk[:, -1] = (20, 20, 24, 255)
    img[y0:y1, x0:x1] = blk


def window_frame(move_button: bool = False) -> np.ndarray:
    """The 1280x800 RGBA window: R = 30 + 50x/1280, G = 40 + 50y/800 (integer
    division), B = 64, opaque; wifi, mute, panel and the confirm button
    (moved left when move_button)."""
    h, w = 800, 1280
    img = np.empty((h, w, 4), dtype=np.uint8)
    img[..., 0] = (30 + np.arange(w) * 50 // 1280)[None, :]
    img[..., 1] = (40 + np.arange(h) * 50 // 800)[:, None]
    img[..., 2] = 64
    img[..., 3] = 255
    draw_widget(img, WINDOW_RECTS["wifi"], 7)
    draw_widget(img, WINDOW_RECTS["mute"], 9)
    draw_widget(img, WINDOW_RECTS["panel"], 33)
    draw_widget(img, WINDOW_RECTS["ok_moved" if move_button else "ok"], 21)
    return img


def crop(img: np.ndarray, rect) -> np.ndarray:
    x0, y0, x1, y1 = rect
    return np.ascontiguousarray(img[y0:y1, x0:x1])


def decoy_widget(w: int, h: int, seed: int) -> np.ndarray:
    """A widget drawn nowhere on the window."""
    img = np.zeros((h, w, 4), dtype=np.uint8)
    draw_widget(img, (0, 0, w, h), seed)
    return img
