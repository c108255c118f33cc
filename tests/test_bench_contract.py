"""bench.py's contract and byte accounting, on the CPU.

The algorithmic bytes are the canonical drop-in step's: per 3-site body
204 + 60 N = 384 B over part 1 + part 2 (236 B for the fused boundary of
step(K)), 104 B per free atom over the two passes, 36 B per atom for the
benchmark force field, and 4 B more per atom and pass when the atoms are
not stored body by body."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("N", [2, 3, 6, 20])
def test_two_pass_bytes_are_the_canonical_step(N):
    per_body = bench.algorithmic_bytes("part1", N, 1, 0) + bench.algorithmic_bytes("part2", N, 1, 0)
    assert per_body == 204 + 60 * N
    if N == 3:
        assert per_body == 384
        assert bench.algorithmic_bytes("fused", 3, 1, 0) == 236


def test_free_atom_and_force_bytes():
    assert bench.algorithmic_bytes("part1", 3, 0, 1) + bench.algorithmic_bytes("part2", 3, 0, 1) == 104
    assert bench.algorithmic_bytes("forces", 3, 10, 5) == 36 * (3 * 10 + 5)
    # non-contiguous atoms: an int32 index per atom and pass
    plain = bench.algorithmic_bytes("part1", 3, 10, 5) + bench.algorithmic_bytes("part2", 3, 10, 5)
    indexed = (bench.algorithmic_bytes("part1", 3, 10, 5, indexed=True) +
               bench.algorithmic_bytes("part2", 3, 10, 5, indexed=True))
    assert indexed - plain == 8 * (3 * 10 + 5)


def test_algorithmic_counts_scale_linearly():
    for k in ("part1", "part2", "fused", "forces"):
        a = bench.algorithmic_bytes(k, 3, 1000, 500)
        assert bench.algorithmic_bytes(k, 3, 2000, 1000) == 2 * a
        assert bench.algorithmic_bytes(k, 3, 1000, 0) + bench.algorithmic_bytes(k, 3, 0, 500) == a


def test_shared_frames_implementation_bytes():
    # a 4-byte body index replaces the 12N bytes of coordinates and 16 of inertia
    for k in ("part1", "part2", "fused"):
        full = bench.implementation_bytes(k, 3, 1, 0)
        assert full - bench.implementation_bytes(k, 3, 1, 0, shared_frames=True) == 12 * 3 + 16 - 4
    assert bench.implementation_bytes("fused", 3, 1, 0, shared_frames=True) == 204
