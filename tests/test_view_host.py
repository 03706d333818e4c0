"""The interactive host pm_view without a GPU or a display: its argument
checks and the missing-display path end with a message and a status, before
any GPU call (the GPU runs are in test_gpu_view.py)."""
import os
import subprocess

import pmamd

VIEW = os.path.join(os.path.dirname(pmamd.LIB_PATH), "pm_view")


def run(args, env=None):
    return subprocess.run([VIEW] + args, capture_output=True, text=True, timeout=60, env=env)


def test_view_is_built():
    assert os.access(VIEW, os.X_OK)


def test_view_rejects_bad_arguments():
    p = run(["--bogus"])
    assert p.returncode == 2 and "usage" in p.stderr
    p = run(["--size", "64"])
    assert p.returncode == 2 and "WxH" in p.stderr


def test_view_without_display_says_so():
    env = dict(os.environ)
    env.pop("DISPLAY", None)
    p = run([], env=env)
    assert p.returncode == 1 and "no X display" in p.stderr
