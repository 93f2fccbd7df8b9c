"""bench.py on the CPU: the one-GPU contract, and the per-launch byte model
its roofline figures divide by, checked against the sizes of the arrays each
sweep reads and writes (real N^3 arrays and the half-complex N x N x (N/2+1)
transform, counted from their shapes, not from the model's formula)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_algorithmic_bytes_match_the_array_sizes():
    sys.path.insert(0, ROOT)
    import bench

    nf = 2
    for n in (256, 1024):
        for w in (4, 8):
            real = n ** 3 * w                    # one real N^3 array
            half = n * n * (n // 2 + 1) * 2 * w  # one half-complex transform
            # x R2C (one field): read phi and pi, write phi and the transform
            assert bench.algorithmic_bytes("x_r2c", n, nf, w) == 3 * real + half
            # column passes (all fields): read and write the transforms
            for stage in ("y_fwd", "z_lap", "y_inv"):
                assert bench.algorithmic_bytes(stage, n, nf, w) == nf * 2 * half
            # x C2R + kick (all fields): read the transform and phi, read and write pi
            assert bench.algorithmic_bytes("x_c2r", n, nf, w) == nf * (half + 3 * real)
    # the figure quoted for the mixed 1024^3 x R2C launch
    assert bench.algorithmic_bytes("x_r2c", 1024, nf, 4) == 17188257792
