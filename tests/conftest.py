import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O  # noqa: E402  (tests are the checker's only users)

    O.lib()
    return O


@pytest.fixture(scope="session")
def native():
    import adaptivetensordf_amd

    return adaptivetensordf_amd.native()


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda:0")


def random_rays(n, seed, spread=1.2):
    """Rays through random points of [-spread, spread]^3 with random unit
    directions, origins 3 units back along the direction (f32)."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(-spread, spread, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = p - 3.0 * d
    return o.astype(np.float32), d.astype(np.float32)


def edge_case_rays():
    """Axis-aligned rays (zero direction components), rays on node faces,
    along edges and corners, and grazing rays."""
    o, d = [], []
    for axis in range(3):
        for off in (0.0, 0.25, -0.5, 0.3, 1.0, -1.0, 0.999999):
            oo = [0.1, 0.2, 0.3]
            oo[(axis + 1) % 3] = off
            oo[axis] = -3.0
            dd = [0.0, 0.0, 0.0]
            dd[axis] = 1.0
            o.append(oo)
            d.append(dd)
            o.append([-x if i == axis else x for i, x in enumerate(oo)])
            d.append([-x for x in dd])
    s = 1 / np.sqrt(3.0)
    for sign in ((1, 1, 1), (-1, 1, 1), (1, -1, -1), (-1, -1, -1)):
        dd = [sign[0] * s, sign[1] * s, sign[2] * s]
        o.append([-3 * x for x in dd])
        d.append(dd)
    s2 = 1 / np.sqrt(2.0)
    o += [[-3 * s2, 0.5, -3 * s2], [0.0, 0.0, -3.0], [0.5, 0.5, -3.0], [-3.0, 1.0, 0.0], [1.0 + 1e-7, 0.0, -3.0]]
    d += [[s2, 0.0, s2], [0.0, 0.0, 1.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]]
    return np.array(o, np.float32), np.array(d, np.float32)
