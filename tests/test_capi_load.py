"""The C-ABI library builds, loads and exports every function its header
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os

from sdtfe_amd import capi


def test_library_exports_header_functions():
    names = capi.header_functions()
    assert "compute_density" in names and "sdtfe_map" in names
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every header function has a ctypes signature in the binding
    assert not [n for n in names if n not in capi._SIGS]


def test_reference_prototype_kept():
    """compute_density keeps the prototype of src/dtfe.h:13-15."""
    txt = open(capi.HEADER).read()
    assert ("void compute_density(double *particle_data, int n_particles, int *tetra_data, "
            "int n_tetra,") in " ".join(txt.split())
    assert "double *gradients(int *tetra_data, int n_tetra, double *particle_data);" in txt


def test_binding_loads_without_gpu():
    capi.lib()
    assert os.path.exists(capi.LIB_PATH)
    assert capi.lib().sdtfe_last_error() is not None


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """Without the built library the binding (and pydtfe above it) raises; there is no CPU
    fallback."""
    import numpy as np
    import pytest

    from sdtfe_amd import pydtfe
    monkeypatch.setattr(capi, "LIB_PATH", str(tmp_path / "libsdtfe_amd.so"))
    monkeypatch.setattr(capi, "_lib", None)
    with pytest.raises(capi.SdtfeError, match="not built"):
        capi.lib()
    pos = np.random.default_rng(0).random((64, 3))
    with pytest.raises(capi.SdtfeError, match="not built"):
        pydtfe.compute_surface_density(pos, [0.5, 0.5, 0.5], 8, 1.0, 1.0, 1.0, 1.0, 0.0, 1)


def test_c_caller_links_against_library_alone(tmp_path):
    """A C program declaring the reference's prototypes (compute_density,
    pre_compute_vol, gradients, find_forward_hull, copy_fh_triangles, the
    bitmap and write_rho) links against libsdtfe_amd.so with no other object:
    every symbol it needs is exported."""
    import subprocess
    libdir = os.path.dirname(capi.LIB_PATH)
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c_caller.c")
    exe = str(tmp_path / "c_caller")
    r = subprocess.run(["gcc", "-O2", "-std=c99", "-Wall", "-Werror", src, "-o", exe,
                        "-L" + libdir, "-lsdtfe_amd", "-Wl,-rpath," + libdir,
                        "-Wl,--no-undefined"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_dropin_failure_raises_without_device():
    """With no HIP device the void drop-in fails and the binding raises
    (nothing falls back to a CPU path)."""
    import numpy as np
    import pytest
    if capi.device_count() > 0:
        pytest.skip("a HIP device is present (the GPU tests cover failures there)")
    pd = np.random.default_rng(0).random((8, 4))
    tet = np.zeros((1, 8), dtype=np.int32)
    tet[0, :4] = [0, 1, 2, 3]
    tet[0, 4:] = -1
    with pytest.raises(capi.SdtfeError, match="no HIP device"):
        capi.compute_density(pd, tet, 8, 1.0, 1.0, np.zeros(64), 1.0, (0.5, 0.5, 0.5), 1, 0.0)


def test_link_recipe_keeps_gpu_field_functions(tmp_path):
    """The link recipe for reference callers: an object that, like the
    reference's triangulate.o, defines triangulate beside pre_compute_vol and
    gradients, with those two made local by objcopy, leaves them undefined in
    the program, so they resolve to libsdtfe_amd.so; without the objcopy step
    the program's own copies would be used."""
    import subprocess
    libdir = os.path.dirname(capi.LIB_PATH)
    (tmp_path / "tri.c").write_text(
        "#include <stddef.h>\n"
        "int *triangulate(double *p, int n, size_t *nf, char *o) { *nf = 0; return 0; }\n"
        "void pre_compute_vol(double **p, int n, int *t, int nt, float m) {}\n"
        "double *gradients(int *t, int nt, double *p) { return 0; }\n")
    (tmp_path / "main.c").write_text(
        "#include <stddef.h>\n"
        "int *triangulate(double *, int, size_t *, char *);\n"
        "void pre_compute_vol(double **, int, int *, int, float);\n"
        "double *gradients(int *, int, double *);\n"
        "int main(void) { size_t n; triangulate(0, 0, &n, 0); pre_compute_vol(0, 0, 0, 0, 1.f);"
        " gradients(0, 0, 0); return 0; }\n")

    def run(*cmd):
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp_path)
        assert r.returncode == 0, r.stderr
        return r.stdout

    run("gcc", "-c", "tri.c", "-o", "tri.o")
    run("gcc", "-c", "main.c", "-o", "main.o")
    run("objcopy", "--localize-symbol=pre_compute_vol", "--localize-symbol=gradients",
        "tri.o", "tri_q.o")
    for obj, local in (("tri_q.o", True), ("tri.o", False)):
        run("gcc", "main.o", obj, "-L" + libdir, "-lsdtfe_amd", "-Wl,-rpath," + libdir,
            "-o", "prog")
        syms = run("nm", "prog")
        defined = {ln.split()[-1] for ln in syms.splitlines() if " T " in ln}
        undefined = {ln.split()[-1] for ln in syms.splitlines() if " U " in ln}
        assert "triangulate" in defined
        for name in ("pre_compute_vol", "gradients"):
            if local:
                assert name in undefined and name not in defined, syms
            else:
                assert name in defined, syms
