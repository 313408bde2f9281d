"""Output side of the tracker: overlay frames of the event and rendered
brightness changes, GIF/video of a run, and the run log.

Mirrors utils/visualizer.py:7-73 and utils/auxiliary.py:54-89 with numpy and
PIL (the reference also uses OpenCV, which is absent here):

* ``delta_Ie_image`` (visualizer.py:7-26): ΔI_e scaled by 255 / max|ΔI_e|,
  positive part in red, negative part in blue, uint8 in BGR channel order
  (the reference builds RGB and converts with cv2.cvtColor RGB2BGR);
* ``delta_Ir_image`` (visualizer.py:29-39): ΔI_r scaled the same way and
  mapped to grey (x + 255) / 2, three equal channels;
* ``overlay`` (visualizer.py:42-46): cv2.addWeighted(img2, 0.6, img1, 0.4, 0),
  restated as round-to-nearest of the weighted sum, saturated to [0, 255]
  (OpenCV's own rounding is not pinned: cv2 is absent);
* ``write_png`` stores a BGR frame as cv2.imwrite does (RGB in the file);
* ``save_gif`` (visualizer.py:65-73) and ``save_video`` (visualizer.py:49-62;
  needs OpenCV, which it imports when called);
* ``TrackingLogger`` (auxiliary.py:54-89): file and console handlers with the
  reference's format.
Float-to-uint8 conversions truncate, as numpy's ``astype`` and array
assignment do in the reference.
"""
from __future__ import annotations

import logging
import os
import re

import numpy as np


def _hw(x):
    """[H, W] float32 array from a [1, H, W] tensor or an [H, W] array."""
    if isinstance(x, np.ndarray):
        a = x
    else:
        a = x.detach().cpu().numpy()
    a = np.asarray(a, dtype=np.float32)
    if a.ndim == 3:
        a = a[0]
    return a


def _scaled(a):
    m = max(float(a.max()), abs(float(a.min()))) if a.size else 0.0
    if m == 0.0:   # all zero: the reference would divide by zero; the image is blank
        return np.zeros_like(a)
    return a * np.float32(255.0 / m)


def delta_Ie_image(delta_Ie):
    """BGR uint8 image of an event frame: positive changes red, negative blue."""
    a = _scaled(_hw(delta_Ie))
    img = np.zeros(a.shape + (3,), np.uint8)
    img[:, :, 2] = np.where(a > 0, a, 0).astype(np.uint8)      # red (RGB channel 0)
    img[:, :, 0] = (-np.where(a < 0, a, 0)).astype(np.uint8)   # blue (RGB channel 2)
    return img


def delta_Ir_image(delta_Ir):
    """BGR uint8 grey image of a rendered brightness change."""
    a = _scaled(_hw(delta_Ir))
    g = ((a + np.float32(255.0)) / np.float32(2.0)).astype(np.uint8)
    return np.repeat(g[:, :, None], 3, axis=2)


def overlay(img1, img2, alpha=0.6):
    """alpha * img2 + (1 - alpha) * img1, rounded and saturated (uint8)."""
    s = img2.astype(np.float64) * alpha + img1.astype(np.float64) * (1.0 - alpha)
    return np.clip(np.rint(s), 0, 255).astype(np.uint8)


def write_png(path, bgr):
    from PIL import Image
    Image.fromarray(np.ascontiguousarray(bgr[:, :, ::-1])).save(path)


def _natural_key(name):
    return [int(t) if t.isdigit() else t for t in re.split(r"(\d+)", name)]


def frame_paths(img_dir):
    """PNG files of a directory in natural order (frame_2 before frame_10)."""
    names = sorted((f for f in os.listdir(img_dir) if f.endswith(".png")), key=_natural_key)
    return [os.path.join(img_dir, f) for f in names]


def save_gif(img_dir, save_path, duration=10):
    from PIL import Image
    paths = frame_paths(img_dir)
    if not paths:
        raise FileNotFoundError(f"no .png frames in {img_dir}")
    images = []
    for p in paths:              # decoded and closed one by one: a long run has more frames than open files
        with Image.open(p) as im:
            images.append(im.copy())
    images[0].save(save_path, save_all=True, append_images=images[1:], duration=duration, loop=0)


def save_video(img_dir, save_path, fps=120):
    """mp4v video of the frames (OpenCV required, as in the reference)."""
    try:
        import cv2
    except ImportError as e:
        raise RuntimeError("save_video needs OpenCV (cv2), which is not installed") from e
    paths = frame_paths(img_dir)
    if not paths:
        raise FileNotFoundError(f"no .png frames in {img_dir}")
    first = cv2.imread(paths[0])
    h, w = first.shape[:2]
    video = cv2.VideoWriter(save_path, cv2.VideoWriter_fourcc(*"mp4v"), fps, (w, h))
    for p in paths:
        video.write(cv2.imread(p))
    video.release()


class TrackingLogger:
    """Logger writing to a file and the console with the reference's format
    ("%(asctime)s - %(name)s - %(levelname)s - %(message)s")."""

    def __init__(self, name, log_file, level=logging.DEBUG):
        self.logger = logging.getLogger(name)
        self.logger.setLevel(level)
        fmt = logging.Formatter("%(asctime)s - %(name)s - %(levelname)s - %(message)s")
        self._handlers = [logging.FileHandler(log_file), logging.StreamHandler()]
        for h in self._handlers:
            h.setLevel(level)
            h.setFormatter(fmt)
            self.logger.addHandler(h)

    def close(self):
        for h in self._handlers:
            self.logger.removeHandler(h)
            h.close()

    def debug(self, msg):
        self.logger.debug(msg)

    def info(self, msg):
        self.logger.info(msg)

    def warning(self, msg):
        self.logger.warning(msg)

    def error(self, msg):
        self.logger.error(msg)

    def critical(self, msg):
        self.logger.critical(msg)
