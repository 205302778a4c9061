"""The ``dtfe`` command line (src/main.c:35-195): arguments, the host steps
before the tessellation (on the CPU), and an end-to-end run on the GPU whose
raw map equals the oracle's on the same particles and sample stream and whose
TIFF follows from that map."""
import ctypes as C
import io
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from sdtfe_amd import cli, pydtfe as mirror

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ARGS = ["0.5", "0.5", "0.5", "1.0", "1.0", "1.0", "0.0", "4", "0.5"]


def planar_file(tmp_path, n, seed=0, name="parts.f32"):
    pos = np.random.default_rng(seed).random((n, 3)).astype(np.float32)
    path = tmp_path / name
    pos.T.ravel().tofile(path)  # all x, then all y, then all z
    return str(path), pos


def test_parse_positionals_and_prompts():
    a = cli.parse(["f.bin", "100", "64"] + ARGS + ["--seed", "7"])
    assert a["n_particles"] == 100 and a["grid_dim"] == 64 and a["n_mc_samples"] == 4
    assert a["sample_factor"] == 0.5 and a["seed"] == 7 and a["precision"] == "fp64"
    # missing arguments are read from stdin in the reference's order
    b = cli.parse(["f.bin", "100", "--seed", "7"],
                  stdin=io.StringIO("64 " + " ".join(ARGS) + "\n"))
    assert b == a
    with pytest.raises(SystemExit):
        cli.parse(["f.bin", "100"], stdin=io.StringIO("64 0.5"))


def test_prepare_follows_main(tmp_path):
    """Shuffle and rotation draw from the C library's rand() after
    srand(seed), in main.c's order: the shuffle's n_shuffle - 1 draws, then
    three angles."""
    n = 200
    path, pos = planar_file(tmp_path, n)
    args = cli.parse([path, str(n), "32"] + ARGS + ["--seed", "11"])
    pd, n_shuffle, mass, theta = cli.prepare(args)
    assert n_shuffle == 100 and mass == 2.0  # 1.0f / 0.5f
    libc = C.CDLL(None)
    libc.rand.restype = C.c_int
    libc.srand(11)
    draws = [libc.rand() for _ in range(n_shuffle - 1 + 3)]
    want = np.zeros((n, 4))
    want[:, :3] = pos.astype(np.float64)
    for i in range(1, n_shuffle):
        r = draws[i - 1] % (n - i) + i
        want[[i - 1, r]] = want[[r, i - 1]]
    th = [d / 2147483647 * 2.0 * math.pi for d in draws[n_shuffle - 1:]]
    assert theta == th
    mirror.rotate3d(want, n_shuffle, th, [0.5, 0.5, 0.5])  # pinned to the reference by G5
    np.testing.assert_array_equal(pd, want)


def test_prepare_rejects_short_file(tmp_path):
    path, _ = planar_file(tmp_path, 10)
    args = cli.parse([path, "11", "32"] + ARGS + ["--seed", "1"])
    with pytest.raises(ValueError, match="fewer than"):
        cli.prepare(args)


@pytest.mark.gpu
def test_cli_end_to_end(tmp_path, oracle):
    """The CLI as a user runs it: the raw map equals the oracle's map of the
    same shuffled, rotated particles (1e-10 per pixel), and the TIFF's
    indices are those of the raw map."""
    print("[gpu] dtfe CLI end to end", flush=True)
    from tests.test_host_exports import read_tiff
    from sdtfe_amd.tessellate import triangulate
    n, G, seed = 3000, 48, 5
    path, _ = planar_file(tmp_path, n, seed=3)
    argv = [path, str(n), str(G), "0.5", "0.5", "0.5", "0.8", "0.6", "1.5", "0.03", "3", "0.7",
            "--seed", str(seed), "--out-dir", str(tmp_path)]
    r = subprocess.run([sys.executable, "-m", "sdtfe_amd.cli", *argv], cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(tmp_path / "parts.f32.rho.bin", dtype=np.float64)
    args = cli.parse(argv)
    pd, n_shuffle, mass, _ = cli.prepare(args)
    pd = np.ascontiguousarray(pd[:n_shuffle])
    tetra = triangulate(pd)
    ref, _ = oracle.compute_density(pd, tetra, G, 0.8, 0.6, mass, (0.5, 0.5, 0.5), 3,
                                    float(np.float32(0.03)), seed)
    rel = np.max(np.abs(raw - ref) / np.maximum(np.abs(ref), 1e-300))
    assert rel <= 1e-10, rel
    tags, pix = read_tiff(tmp_path / "parts.f32.rho.tiff")
    assert tags[256] == G
    pos = raw[raw > 0]
    lo, hi = math.log(pos.min() + 1.0), math.log(raw.max() + 1.0)
    want = np.array([int(128.0 * (math.log(v + 1.0) - lo) / (hi - lo)) if v > 0 else 0
                     for v in raw], dtype=np.uint8)
    np.testing.assert_array_equal(pix, want)
