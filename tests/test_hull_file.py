"""Binary hull files (b3amd_hull_file_*): a pool round-trips bit for bit, and files that are
truncated, corrupted or not hull files are refused."""
import os

import numpy as np
import pytest

from bounce_amd.workloads import gen_hull_pool, load_hull_pool, save_hull_pool


def test_hull_file_round_trip(lib, tmp_path):
    pool = gen_hull_pool(lib, 99, 300, 8, 64)
    path = str(tmp_path / "pool.b3h")
    save_hull_pool(lib, pool, path)
    back = load_hull_pool(lib, path)
    for f in ("desc", "centroids", "vertices", "edges", "faces", "planes"):
        assert getattr(back, f).tobytes() == getattr(pool, f).tobytes(), f


def test_hull_file_rejects_damage(lib, tmp_path):
    pool = gen_hull_pool(lib, 98, 40, 8, 32)
    path = str(tmp_path / "pool.b3h")
    save_hull_pool(lib, pool, path)
    raw = bytearray(open(path, "rb").read())
    bad = tmp_path / "bad.b3h"
    for damage in ("flip", "truncate", "magic"):
        b = bytearray(raw)
        if damage == "flip":
            b[len(b) // 2] ^= 0x40
        elif damage == "truncate":
            b = b[:-20]
        else:
            b[0:8] = b"NOTAHULL"
        bad.write_bytes(bytes(b))
        with pytest.raises(RuntimeError):
            load_hull_pool(lib, str(bad))
    with pytest.raises(RuntimeError):
        load_hull_pool(lib, str(tmp_path / "missing.b3h"))
