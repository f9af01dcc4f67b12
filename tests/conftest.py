import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.build()
    return oracle


def star_polygons(rng, n, extent, rmin, rmax, kmin=3, kmax=12, holes=0.0):
    """Random star-shaped polygons (closed rings), optionally with one hole."""
    part = [0]
    ring = [0]
    xs, ys = [], []
    for _ in range(n):
        cx, cy = rng.uniform(0, extent, 2)
        r = rng.uniform(rmin, rmax)
        k = int(rng.integers(kmin, kmax + 1))
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        rr = r * rng.uniform(0.7, 1.3, k)
        vx = cx + rr * np.cos(ang)
        vy = cy + rr * np.sin(ang)
        xs += list(vx) + [vx[0]]
        ys += list(vy) + [vy[0]]
        ring.append(ring[-1] + k + 1)
        nr = 1
        if rng.uniform() < holes:
            hr = 0.25 * r
            hang = np.sort(rng.uniform(0, 2 * np.pi, 5))
            hx = cx + hr * np.cos(hang)
            hy = cy + hr * np.sin(hang)
            xs += list(hx) + [hx[0]]
            ys += list(hy) + [hy[0]]
            ring.append(ring[-1] + 6)
            nr += 1
        part.append(part[-1] + nr)
    return (np.array(part, np.int32), np.array(ring, np.int32), np.array(xs), np.array(ys))
