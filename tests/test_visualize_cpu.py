"""Output side (gsevt/visualize.py) against closed forms of
utils/visualizer.py:7-73 and utils/auxiliary.py:54-89.  OpenCV is absent, so
cv2.addWeighted's own rounding and the mp4 writer are not pinned; the numpy
parts are exact."""
import logging
import os

import numpy as np
import pytest

from gsevt import visualize as vz


def test_delta_Ie_image_colours_and_scale():
    x = np.array([[2.0, -1.0, 0.0], [1.0, -2.0, 0.5]], np.float32)
    img = vz.delta_Ie_image(x)
    assert img.dtype == np.uint8 and img.shape == (2, 3, 3)
    # max |x| = 2 -> scale 127.5; red (BGR index 2) holds positives, blue (0) negatives
    np.testing.assert_array_equal(img[:, :, 2], [[255, 0, 0], [127, 0, 63]])
    np.testing.assert_array_equal(img[:, :, 0], [[0, 127, 0], [0, 255, 0]])
    assert not img[:, :, 1].any()
    # a [1, H, W] tensor gives the same image
    torch = pytest.importorskip("torch")
    np.testing.assert_array_equal(vz.delta_Ie_image(torch.from_numpy(x)[None]), img)


def test_delta_Ir_image_grey_mapping():
    x = np.array([[-4.0, 0.0, 4.0, 1.0]], np.float32)
    img = vz.delta_Ir_image(x)
    # (x * 255 / 4 + 255) / 2, truncated: -4 -> 0, 0 -> 127, 4 -> 255, 1 -> 159
    np.testing.assert_array_equal(img[:, :, 0], [[0, 127, 255, 159]])
    assert np.array_equal(img[:, :, 0], img[:, :, 1]) and np.array_equal(img[:, :, 1], img[:, :, 2])


def test_blank_frames_do_not_divide_by_zero():
    z = np.zeros((3, 4), np.float32)
    assert not vz.delta_Ie_image(z).any()
    np.testing.assert_array_equal(vz.delta_Ir_image(z), np.full((3, 4, 3), 127, np.uint8))


def test_overlay_weights_and_saturation():
    a = np.array([[[100, 0, 255]]], np.uint8)
    b = np.array([[[200, 255, 255]]], np.uint8)
    # 0.6 * b + 0.4 * a
    np.testing.assert_array_equal(vz.overlay(a, b), [[[160, 153, 255]]])


def test_png_gif_and_natural_order(tmp_path):
    from PIL import Image
    d = tmp_path / "frames"
    d.mkdir()
    frames = []
    for i in (10, 2, 1):
        f = np.zeros((4, 5, 3), np.uint8)
        f[:, :, 2] = 20 * i     # red in BGR
        f[0, 0, 0] = 200        # one blue pixel
        vz.write_png(str(d / f"frame_{i}.png"), f)
        frames.append((i, f))
    assert [os.path.basename(p) for p in vz.frame_paths(str(d))] == ["frame_1.png", "frame_2.png", "frame_10.png"]
    # the PNG holds RGB, as cv2.imwrite writes a BGR array
    with Image.open(d / "frame_2.png") as im:
        rgb = np.asarray(im)
    np.testing.assert_array_equal(rgb, dict(frames)[2][:, :, ::-1])
    gif = tmp_path / "run.gif"
    vz.save_gif(str(d), str(gif), duration=2)
    with Image.open(gif) as im:
        assert im.n_frames == 3
        for k, i in enumerate((1, 2, 10)):
            im.seek(k)
            np.testing.assert_array_equal(np.asarray(im.convert("RGB")), dict(frames)[i][:, :, ::-1])
    empty = tmp_path / "empty"
    empty.mkdir()
    with pytest.raises(FileNotFoundError):
        vz.save_gif(str(empty), str(gif))


def test_save_video_needs_opencv(tmp_path):
    try:
        import cv2  # noqa: F401
        pytest.skip("OpenCV present")
    except ImportError:
        pass
    with pytest.raises(RuntimeError, match="OpenCV"):
        vz.save_video(str(tmp_path), str(tmp_path / "v.mp4"))


def test_tracking_logger_format(tmp_path):
    path = tmp_path / "tracking_log.log"
    lg = vz.TrackingLogger("TrackingLoggerTest", str(path), level=logging.INFO)
    lg.info("frame_idx:\t3 / 10")
    lg.debug("hidden at INFO")
    lg.error("fine stage optimization iter exceeded the max_optim_iter!")
    lg.close()
    lines = path.read_text().splitlines()
    assert len(lines) == 2
    assert lines[0].endswith(" - TrackingLoggerTest - INFO - frame_idx:\t3 / 10")
    assert " - TrackingLoggerTest - ERROR - fine stage" in lines[1]


def test_gif_of_more_frames_than_open_files(tmp_path):
    """A run's frame count can exceed the open-file limit (1024 is common):
    the GIF writer holds no frame file open."""
    import resource
    soft, hard = resource.getrlimit(resource.RLIMIT_NOFILE)
    d = tmp_path / "many"
    d.mkdir()
    img = np.zeros((4, 4, 3), np.uint8)
    for i in range(300):
        img[0, 0, 0] = i % 256
        vz.write_png(str(d / f"frame_{i}.png"), img)
    resource.setrlimit(resource.RLIMIT_NOFILE, (min(256, soft), hard))
    try:
        vz.save_gif(str(d), str(tmp_path / "many.gif"), duration=2)
    finally:
        resource.setrlimit(resource.RLIMIT_NOFILE, (soft, hard))
    from PIL import Image
    with Image.open(tmp_path / "many.gif") as im:
        assert im.n_frames == 300
