"""The device walk code (walk_fp64.hip / walk_fp64_core.h) compiled for the
host and run on the golden rays: pins the kernel's logic without a GPU.
(The library never runs this host build; the GPU tests check the device
build.)"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from sdtfe_amd import build as sb

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def host_walk(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("hw") / "host_walk")
    cmd = [hipcc, "-x", "hip", "-std=c++17", "-O2", "-ffp-contract=off",
           f"--offload-arch={sb.ARCH}", "-I" + sb.CSRC, "-I" + os.path.join(sb.ROOT, "include"),
           "-w", os.path.join(HERE, "host_walk.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def records(pos, tetra, rho, nabla):
    t = tetra.shape[0]
    rec = np.zeros(t, dtype=np.dtype([("nb", "<i4", 4), ("v", "<f8", 12), ("f", "<f8", 4)]))
    rec["nb"] = tetra[:, 4:]
    rec["v"] = pos[tetra[:, :4]].reshape(t, 12)
    rec["f"][:, 0] = rho[tetra[:, 0]]
    rec["f"][:, 1:] = nabla
    return rec


def records32(pos, tetra, rho):
    """The mixed walk's 96-byte records: neighbours, A (FP64), B-A, C-A, D-A
    and the vertex densities (FP32)."""
    t = tetra.shape[0]
    rec = np.zeros(t, dtype=np.dtype([("nb", "<i4", 4), ("a", "<f8", 3), ("e", "<f4", 9),
                                      ("rho", "<f4", 4), ("pad", "<f4")]))
    assert rec.dtype.itemsize == 96
    p = pos[tetra[:, :4]]
    rec["nb"] = tetra[:, 4:]
    rec["a"] = p[:, 0]
    rec["e"] = (p[:, 1:] - p[:, :1]).astype(np.float32).reshape(t, 9)
    rec["rho"] = rho[tetra[:, :4]].astype(np.float32)
    return rec


def face_records(pos, tetra, rho):
    """The mixed walk's 32-byte face records (4 per tet): the new vertex
    (x, y as FP32 hi/lo pairs), the tet across face f, a code (one byte per
    label m of this tet, value at bits 3-4: 8 x the label there of vertex m,
    or for m = f 8 x the face's index there; bit 31 when the walk stops: hull
    or missing back link), and the new vertex's z relative to the mean z of
    face f's vertices and its density (FP32)."""
    t = tetra.shape[0]
    rec = np.zeros((t, 4), dtype=np.dtype([("qxh", "<f4"), ("qxl", "<f4"), ("qyh", "<f4"),
                                           ("qyl", "<f4"), ("nb", "<i4"), ("code", "<u4"),
                                           ("dqz", "<f4"), ("rho", "<f4")]))
    assert rec.dtype.itemsize == 32
    stop = np.int64(1) << 31
    verts, nbrs = tetra[:, :4], tetra[:, 4:]
    for f in range(4):
        n = nbrs[:, f]
        rec["nb"][:, f] = n
        ok = n >= 0
        nn = np.where(ok, n, 0)
        back = np.full(t, -1)
        for k in range(3, -1, -1):
            back = np.where(nbrs[nn, k] == np.arange(t), k, back)
        code = np.where(back < 0, stop, (back & 3) << (8 * f + 3)).astype(np.int64)
        zs = np.zeros(t)
        for k in (1, 2, 3):
            v = verts[:, f ^ k]
            zs = zs + pos[v, 2]
            lab = np.full(t, -1)
            for m in range(3, -1, -1):
                lab = np.where(verts[nn, m] == v, m, lab)
            code |= np.where((lab < 0) | (lab == back), stop, 0)
            code |= (lab & 3) << (8 * (f ^ k) + 3)
        q = verts[nn, back & 3]
        rec["code"][:, f] = np.where(ok, code, stop)
        for c, k in ((0, "qx"), (1, "qy")):
            hi = pos[q, c].astype(np.float32)
            lo = (pos[q, c] - hi.astype(np.float64)).astype(np.float32)
            rec[k + "h"][:, f] = np.where(ok, hi, 0)
            rec[k + "l"][:, f] = np.where(ok, lo, 0)
        rec["dqz"][:, f] = np.where(ok, (pos[q, 2] - zs / 3.0).astype(np.float32), 0)
        rec["rho"][:, f] = np.where(ok, rho[q].astype(np.float32), 0)
    return rec.reshape(-1)


def run_walk(host_walk, tmp_path, g, mixed=False, wide=False, face=False):
    rec = records(g["pos"], g["tetra"], g["rho"], g["nabla"])
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([rec.size, g["x"].size], dtype=np.int64).tofile(f)
        np.array([float(g["lz"]), float(g["dxy"]), float(g["off_z"])]).tofile(f)
        np.array([int(g["seed"])], dtype=np.uint64).tofile(f)
        rec.tofile(f)
        for k in ("x", "y"):
            g[k].astype(np.float64).tofile(f)
        g["pixel"].astype(np.uint64).tofile(f)
        g["sample"].astype(np.uint64).tofile(f)
        g["entry"].astype(np.int64).tofile(f)
        if mixed:
            lo, hi = g["pos"].min(0), g["pos"].max(0)
            org = 0.5 * (lo + hi)
            records32(g["pos"], g["tetra"], g["rho"]).tofile(f)
            face_records(g["pos"], g["tetra"], g["rho"]).tofile(f)
            zr = (lo[2], hi[2])
            np.array([*org, float((0.5 * (hi - lo)).max()), *zr]).tofile(f)
        if face:
            pos4 = np.zeros((g["pos"].shape[0], 4))
            pos4[:, :3] = g["pos"]
            pos4[:, 3] = g["rho"]
            np.array([pos4.shape[0]], dtype=np.int64).tofile(f)
            pos4.tofile(f)
            np.ascontiguousarray(g["tetra"], dtype=np.int32).tofile(f)
    mode = (["mixedwide" if wide else "mixed"] if mixed else ["face"] if face else [])
    r = subprocess.run([host_walk, str(inp), str(out)] + mode,
                       check=True, capture_output=True, text=True)
    res = np.fromfile(out).reshape(-1, 2)
    return res, r.stderr


def test_device_walk_code_matches_reference(golden, host_walk, tmp_path):
    g = golden("g2_rays.npz")
    res, _ = run_walk(host_walk, tmp_path, g)
    np.testing.assert_array_equal(res[:, 1].astype(np.int32), g["crossed"])
    np.testing.assert_array_equal(res[:, 0], g["value"])


@pytest.mark.parametrize("slab", ["hull", "cut"])
def test_face_record_walk_matches_reference(golden, host_walk, tmp_path, slab):
    """The FP64 walk over face records (the map kernels' FP64 form) gives
    the reference walk's values and tets crossed bit for bit: on the golden
    rays (16,384 stream rays and 600 rays exactly through projected vertices
    and along projected edges, which take the hand-over to the full face
    test), and through a slab that cuts the hull."""
    g = dict(golden("g2_rays.npz"))
    if slab == "cut":
        g["lz"], g["off_z"] = 0.5, 0.25
        want, _ = run_walk(host_walk, tmp_path, g)
        ref, crossed = want[:, 0], want[:, 1].astype(np.int32)
    else:
        ref, crossed = g["value"], g["crossed"]
    res, err = run_walk(host_walk, tmp_path, g, face=True)
    assert "face walk" in err
    np.testing.assert_array_equal(res[:, 1].astype(np.int32), crossed)
    np.testing.assert_array_equal(res[:, 0], ref)


def test_face_table(host_walk):
    """The mixed walk's sign-mask face table equals the general face choice
    (the FP64 walk's, pinned above) on all 64 sign patterns."""
    r = subprocess.run([host_walk, "faces"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("wide", [False, True], ids=["narrow", "wide"])
@pytest.mark.parametrize("slab", ["hull", "cut"])
def test_device_mixed_walk_code(golden, host_walk, tmp_path, slab, wide):
    """Mixed walk: same tets as the reference walk, values within 1e-5.
    'hull': the golden rays, whose slab holds the hull's z range (midpoints
    are not tested); 'cut': the same rays through the slab 0.25 < z < 0.75,
    against the FP64 walk code, so that chords are cut at both ends."""
    g = dict(golden("g2_rays.npz"))
    if slab == "cut":
        g["lz"], g["off_z"] = 0.5, 0.25
        want, _ = run_walk(host_walk, tmp_path, g)
        ref, crossed = want[:, 0], want[:, 1].astype(np.int32)
    else:
        ref, crossed = g["value"], g["crossed"]
    res, err = run_walk(host_walk, tmp_path, g, mixed=True, wide=wide)
    assert f"full_slab {int(slab == 'hull')}" in err, err
    np.testing.assert_array_equal(res[:, 1].astype(np.int32), crossed)
    v = res[:, 0]
    m = ref != 0
    assert np.max(np.abs(v[m] - ref[m]) / np.abs(ref[m])) <= 1e-5
    assert (v[~m] == 0).all()


@pytest.mark.parametrize("n,seed", [(1195, 5), (5175, 14), (3598, 50), (1092, 1)])
def test_mixed_walk_few_large_tets(host_walk, tmp_path, oracle, n, seed):
    """Halo cutouts of a few thousand particles: rays near the hull cross a
    few large, nearly vertical tets and have small values, where the FP32
    crossings' weight errors are largest.  Every ray is within 1e-5 of the
    FP64 walk (those whose error bound exceeds the tolerance are handed to
    the FP64 walk), on random rays and on the rays of pixels that once
    exceeded it."""
    from sdtfe_amd import workloads
    from sdtfe_amd.tessellate import to_particle_data, triangulate
    pos = workloads.cutout(n, seed)
    tetra = triangulate(pos)
    pd = to_particle_data(pos)
    oracle.density(pd, tetra, 1.0)
    nabla = oracle.gradients(tetra, pd)
    tri = oracle.triangles(tetra, pd)
    rng = np.random.default_rng(seed)
    x, y = rng.random(3000), rng.random(3000)
    ent = np.array([oracle.pt_loc_2d(tri, pd, a, b) for a, b in zip(x, y)])
    m = ent >= 0
    g = dict(pos=pos, tetra=tetra, rho=pd[:, 3].copy(), nabla=nabla, x=x[m], y=y[m],
             pixel=np.arange(m.sum()), sample=np.zeros(m.sum(), int), entry=ent[m], lz=1.0,
             dxy=1.0 / 512, off_z=0.0, seed=seed)
    want, _ = run_walk(host_walk, tmp_path, g)
    got, err = run_walk(host_walk, tmp_path, g, mixed=True)
    np.testing.assert_array_equal(got[:, 1], want[:, 1])
    v, r = got[:, 0], want[:, 0]
    k = r != 0
    assert np.max(np.abs(v[k] - r[k]) / np.abs(r[k])) <= 1e-5
    assert "fp64_redo" in err
