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
