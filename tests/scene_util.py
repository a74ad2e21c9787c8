"""Scene construction shared by the GPU tests and smoke(): a tree built by
the libtorch layer with a chosen active set, matching oracle arrays, rays,
and field parameters."""
import numpy as np
import torch

from adaptivetensordf_amd import native, phantom


def uniform_active(tree_arrays, depth):
    return (tree_arrays["depth"] == depth).astype(np.int32)


def adaptive_active(tree_arrays, num_ellipsoids=30, seed=0, budget=400, start_depth=2, max_depth=None):
    ell = phantom.ellipsoids(num_ellipsoids, seed)
    md = int(tree_arrays["depth"].max()) if max_depth is None else max_depth
    return phantom.adaptive_active_mask(tree_arrays["node_min"], tree_arrays["node_max"], tree_arrays["children"],
                                        tree_arrays["depth"], ell, budget, start_depth, md)


def device_tree(max_depth, active, device="cuda", quad=None):
    """libtorch HyperTreeBase with the given active mask, moved to the GPU;
    quad = (y_min, y_max) builds the quad-tree representation."""
    t = native().HyperTreeBase(3, max_depth)
    if quad is None:
        t.settree()
    else:
        t.settree(True, quad[0], quad[1])
    t.node_active = torch.from_numpy(np.ascontiguousarray(active, np.int32))
    t.UpdateActive()
    t.to(device)
    return t


def rays_to_device(o, d, device="cuda"):
    return (torch.from_numpy(np.ascontiguousarray(o)).to(device),
            torch.from_numpy(np.ascontiguousarray(d)).to(device))


def tilt_rays(num, seed, W=64, H=64, tilts=61):
    ang = phantom.tilt_angles(tilts)
    tl, px, py = phantom.sample_pixels(num, tilts, W, H, seed)
    return phantom.rays_for_pixels(tl, px, py, W, H, ang)
