"""Synthetic scene directories in the reference's on-disk layout
(data/SceneBase.cpp:740-869): dataset.ini, camera.ini, poses.txt,
images.txt, masks.txt, camera_indices.txt, a split folder with train.txt /
eval.txt, and the projections as .npy / .npz files under image_dir."""
from __future__ import annotations

import math
import os

import numpy as np


def tilt_quaternion(theta):
    """Unit quaternion (x, y, z, w) of a rotation by theta about +y."""
    return np.array([0.0, math.sin(theta / 2), 0.0, math.cos(theta / 2)])


def write_scene(root, *, camera_model="orthographic", w=24, h=16, num_images=6, raw=None, masks=None,
                fmt=("npy", "npz"), scale_poses=1.0, xray_min=1000.0, xray_max=60000.0, log10=True,
                log_space_input=False, projection_factor=2.0, scene_scale=1.0, split="exp_uniform_50",
                K=None, relative_image_dir=True, seed=0):
    """Write a scene and return (path, raw images [n, h, w], poses [n, 7])."""
    rng = np.random.default_rng(seed)
    os.makedirs(root, exist_ok=True)
    img_dir = os.path.join(root, "projections")
    os.makedirs(img_dir, exist_ok=True)
    if raw is None:
        raw = rng.uniform(xray_min * 0.5, xray_max * 1.05, size=(num_images, h, w)).astype(np.float32)
        raw[:, 0, 0] = 0.0  # clamped to 1e-5 by the log conversion
    if K is None:
        K = ((w - 1) / 2, (h - 1) / 2, (w - 1) / 2, (h - 1) / 2, 0.0)
    with open(os.path.join(root, "dataset.ini"), "w") as f:
        f.write("[DatasetParams]\n")
        f.write(f"camera_model = {camera_model}\n")
        f.write(f"image_dir = {'projections' if relative_image_dir else img_dir}\n")
        f.write(f"mask_dir = {'masks' if masks is not None else ''}\n")
        f.write(f"projection_factor = {projection_factor!r}\n")
        f.write(f"scene_scale = {scene_scale!r}\n")
        f.write(f"xray_min = {xray_min!r}\nxray_max = {xray_max!r}\n")
        f.write(f"log_space_input = {'true' if log_space_input else 'false'}\n")
        f.write(f"use_log10_conversion = {'true' if log10 else 'false'}\n")
        f.write("roi_min = -1 -1 -0.5\nroi_max = 1 1 0.5\n")
        f.write("; a comment\n[OtherSection]\ncamera_model = pinhole\n")
    with open(os.path.join(root, "camera.ini"), "w") as f:
        f.write(f"[CameraBase]\nw = {w}\nh = {h}\n# fx fy cx cy s\nK = {' '.join(repr(float(k)) for k in K)}\n")
    poses = []
    for i in range(num_images):
        th = -math.pi / 3 + (2 * math.pi / 3) * i / max(1, num_images - 1)
        q = tilt_quaternion(th) * scale_poses  # the loader normalises
        t = rng.uniform(-2, 2, size=3)
        poses.append(np.concatenate([q, t]))
    poses = np.array(poses)
    with open(os.path.join(root, "poses.txt"), "w") as f:
        for p in poses:
            f.write(" ".join(repr(float(v)) for v in p) + "\n")
    names = []
    for i in range(num_images):
        ext = fmt[i % len(fmt)]
        name = f"proj_{i:03d}.{ext}"
        if ext == "npy":
            np.save(os.path.join(img_dir, name), raw[i])
        else:
            np.savez_compressed(os.path.join(img_dir, name), proj=raw[i], other=np.zeros(3))
        names.append(name)
    with open(os.path.join(root, "images.txt"), "w") as f:
        f.write("\n".join(names) + "\n")
    if masks is not None:
        mdir = os.path.join(root, "masks")
        os.makedirs(mdir, exist_ok=True)
        with open(os.path.join(root, "masks.txt"), "w") as f:
            for i in range(num_images):
                np.save(os.path.join(mdir, f"mask_{i:03d}.npy"), masks[i])
                f.write(f"mask_{i:03d}.npy\n")
    with open(os.path.join(root, "camera_indices.txt"), "w") as f:
        f.write("\n".join("0" for _ in range(num_images)) + "\n")
    os.makedirs(os.path.join(root, split), exist_ok=True)
    train = [i for i in range(num_images) if i % 3 != 1]
    test = [i for i in range(num_images) if i % 3 == 1]
    with open(os.path.join(root, split, "train.txt"), "w") as f:
        f.write("\n".join(str(i) for i in reversed(train)) + "\n\n")
    with open(os.path.join(root, split, "eval.txt"), "w") as f:
        f.write("\n".join(str(i) for i in test) + "\n")
    return root, raw, poses
