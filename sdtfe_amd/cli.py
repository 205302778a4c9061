"""``dtfe`` command line: the reference's example driver (src/main.c:35-195)
over libsdtfe_amd.so.

    python -m sdtfe_amd.cli path_to_file n_particles grid_dim center_x center_y
        center_z field_width field_depth particle_mass mc_box_width n_mc_samples
        sample_factor [--seed S] [--precision fp64|mixed] [--out-dir DIR]

Positional arguments and their order are the reference's; any that are
missing are asked for on standard input in the same order (src/main.c:38-118).
The steps are the reference's, through the same C entry points, now exported
by libsdtfe_amd.so:

1. the C library's generator is seeded (``srand``) with ``--seed``, by default
   the current time, and the seed is printed (src/main.c:120-123);
2. particle_mass /= sample_factor in single precision (src/main.c:126);
3. ``n_particles`` x 3 float32 are read from the file, all x, then all y, then
   all z (src/main.c:132-137; the reference's comment calls this row-major, the
   code reads it planar), and widened by ``convert_to_qhdata``;
4. ``n_shuffle = (size_t)(n_particles * sample_factor)`` in single precision;
   ``shuffle_buff`` draws from ``rand()`` (src/main.c:143-145);
5. three angles ``rand() / RAND_MAX * 2 pi``, then ``rotate3d`` about the
   centre (src/main.c:148-152);
6. Delaunay tetrahedralisation of the first n_shuffle particles ("Qz Qt Qbb";
   Qhull stays on the host, the tets are assembled on the GPU);
7. ``compute_density`` into a zeroed map on the GPU; the Monte-Carlo samples
   come from the counter-based stream keyed by the same seed;
8. ``write_image`` to ``<out-dir>/<basename>.rho.tiff`` and ``write_rho`` to
   ``<out-dir>/<basename>.rho.bin`` (src/main.c:182-193).

Errors end the program with status 1 and a message (the reference reads past
a short file and does not check its calls).
"""
from __future__ import annotations

import argparse
import ctypes as C
import math
import os
import sys
import time

import numpy as np

from . import capi

POSITIONAL = [
    ("path_to_file", str, "Enter the file location):"),
    ("n_particles", int, "Enter the number of particles in the file:"),
    ("grid_dim", int, "Enter the grid cells per dimension (square assumed):"),
    ("center_x", float, "Enter the center of the 3D field volume (x):"),
    ("center_y", float, "Enter the center of the 3D field volume (y):"),
    ("center_z", float, "Enter the center of the 3D field volume (z):"),
    ("field_width", float, "Enter the width of the 2D field (square assumed):"),
    ("field_depth", float, "Enter the depth of the 3D field (set limits of integration):"),
    ("particle_mass", float, "Enter the particle mass"),
    ("mc_box_width", float, "Enter the width for the Monte Carlo sample volume of each pixel:"),
    ("n_mc_samples", int, "Enter number of Monte Carlo samples for each pixel:"),
    ("sample_factor", float, "Enter the subsample factor:"),
]
RAND_MAX = 2147483647  # glibc


def parse(argv, stdin=None):
    """Arguments as main.c reads them: positionals from argv, the missing ones
    from stdin, one whitespace-separated token each."""
    ap = argparse.ArgumentParser(prog="dtfe", description=__doc__.split("\n\n")[0])
    ap.add_argument("values", nargs="*")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--out-dir", default=".")
    a = ap.parse_args(argv)
    if len(a.values) > len(POSITIONAL):
        ap.error(f"at most {len(POSITIONAL)} positional arguments")
    vals = list(a.values)
    tokens = None
    for name, typ, prompt in POSITIONAL[len(vals):]:
        print(prompt, flush=True)
        if tokens is None:
            tokens = iter((stdin or sys.stdin).read().split())
        tok = next(tokens, None)
        if tok is None:
            ap.error(f"missing {name}")
        vals.append(tok)
    out = {}
    for (name, typ, _), v in zip(POSITIONAL, vals):
        try:
            out[name] = typ(v)
        except ValueError:
            ap.error(f"{name}: cannot read {v!r}")
    out["seed"] = int(time.time()) if a.seed is None else a.seed
    out["precision"] = a.precision
    out["out_dir"] = a.out_dir
    return out


def _libc():
    libc = C.CDLL(None)
    libc.rand.restype = C.c_int
    libc.srand.argtypes = [C.c_uint]
    return libc


def prepare(args):
    """Steps 1-5: returns (particle_data with the first n_shuffle rows
    shuffled and rotated, n_shuffle, particle mass as float32, theta)."""
    n = args["n_particles"]
    if n <= 0:
        raise ValueError("n_particles must be positive")
    path = args["path_to_file"]
    need = 12 * n
    if os.path.getsize(path) < need:
        raise ValueError(f"{path} holds fewer than {n} x 3 float32 values")
    libc = _libc()
    libc.srand(args["seed"] & 0xFFFFFFFF)
    sf = np.float32(args["sample_factor"])
    mass = np.float32(np.float32(args["particle_mass"]) / sf)
    buf = np.fromfile(path, dtype=np.float32, count=3 * n)
    pd = capi.convert_to_qhdata(buf)
    n_shuffle = int(np.float32(np.float32(n) * sf))
    if n_shuffle < 5 or n_shuffle > n:
        raise ValueError(f"n_particles * sample_factor = {n_shuffle}: need 5 .. {n}")
    capi.shuffle_buff(pd, n, n_shuffle)
    theta = [libc.rand() / RAND_MAX * 2.0 * math.pi for _ in range(3)]
    center = [args["center_x"], args["center_y"], args["center_z"]]
    capi.rotate3d(pd, n_shuffle, theta, center)
    return pd, n_shuffle, float(mass), theta


def run(args) -> tuple[str, str]:
    from .tessellate import triangulate
    print(f"Using prng seed: {args['seed']}")
    pd, n_shuffle, mass, theta = prepare(args)
    print("Rotating with theta=[{:f} {:f} {:f}]...".format(*theta))
    print("Creating Delaunay triangulation...")
    pd = np.ascontiguousarray(pd[:n_shuffle])
    tetra = triangulate(pd, gpu=True)
    print(f"Done creating {tetra.shape[0]} tetrahedra.")
    G = args["grid_dim"]
    center = [args["center_x"], args["center_y"], args["center_z"]]
    print(f"Creating {G}x{G} surface density field...")
    rho = np.zeros(G * G)
    capi.set_seed(args["seed"])
    old = os.environ.get("SDTFE_PRECISION")
    os.environ["SDTFE_PRECISION"] = args["precision"]
    try:
        capi.compute_density(pd, tetra, G, args["field_width"], args["field_depth"], rho, mass,
                             center, args["n_mc_samples"], float(np.float32(args["mc_box_width"])))
    finally:
        if old is None:
            os.environ.pop("SDTFE_PRECISION", None)
        else:
            os.environ["SDTFE_PRECISION"] = old
    base = os.path.join(args["out_dir"], os.path.basename(args["path_to_file"]))
    tiff, raw = base + ".rho.tiff", base + ".rho.bin"
    capi.write_image(tiff, rho)
    if capi.write_rho(raw, rho) != rho.size:
        raise OSError(f"cannot write {raw}")
    print(f"Wrote {tiff} and {raw}")
    return tiff, raw


def main(argv=None) -> int:
    args = parse(sys.argv[1:] if argv is None else argv)
    try:
        run(args)
    except (ValueError, OSError, capi.SdtfeError) as e:
        print(f"dtfe: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
