"""The C-ABI library builds, loads and exports every entry point that
include/*.h declares (no device calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tioga_amd", "libtioga_amd.so")


def declared_c_symbols():
    names = set()
    src = open(os.path.join(ROOT, "include", "tioga_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    for m in re.finditer(r"^\s*(?:void|int|long|double)\s+\**\s*(tioga\w*)\s*\(", src, flags=re.M):
        names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from tioga_amd import build

        build.build()
    return ctypes.CDLL(LIB)


def test_header_declares_the_reference_abi():
    names = set(declared_c_symbols())
    # tiogaInterface.C:36-328
    for n in ("tioga_init_f90_", "tioga_init_", "tioga_registergrid_data_", "tioga_registergrid_data_mb_",
              "tioga_preprocess_grids_", "tioga_performconnectivity_", "tioga_registersolution_",
              "tioga_dataupdate_mb_", "tioga_dataupdate_", "tioga_getdonorcount_", "tioga_getdonorinfo_",
              "tioga_setsymmetry_", "tioga_setresolutions_", "tioga_setresolutions_multi_",
              "tioga_setcelliblank_", "tioga_reduce_fringes_", "tioga_setnfringe_", "tioga_setmexclude_",
              "tioga_writeoutputfiles_", "tioga_delete_", "tioga_register_amr_global_data_",
              "tioga_register_amr_patch_count_", "tioga_register_amr_local_data_",
              "tioga_performconnectivity_highorder_", "tioga_performconnectivity_amr_",
              "tioga_set_highorder_callback_", "tioga_set_p4est_", "tioga_set_amr_callback_",
              "tioga_set_p4est_search_callback_"):
        assert n in names, n


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_c_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_cpp_class_symbols_exported():
    import subprocess

    if not os.path.exists(LIB):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True).stdout
    for m in ("TIOGA::tioga::tioga()", "TIOGA::tioga::registerGridData", "TIOGA::tioga::profile()",
              "TIOGA::tioga::performConnectivity()", "TIOGA::tioga::dataUpdate(int, int, int)",
              "TIOGA::tioga::getDonorInfo", "TIOGA::tioga::getReceptorInfo"):
        assert m in out, m


def test_status_api_without_device(lib):
    lib.tioga_amd_status.restype = ctypes.c_int
    buf = ctypes.create_string_buffer(256)
    lib.tioga_amd_clear_status()
    assert lib.tioga_amd_status(buf, 256) == 0
    # a call before init is an error, reported through the status
    lib.tioga_preprocess_grids_()
    assert lib.tioga_amd_status(buf, 256) == 1
    assert b"tioga_init_" in buf.value
    lib.tioga_amd_clear_status()


@pytest.mark.parametrize("prelude", [
    "",  # no MPI header: an int handle (MPICH's MPI_Comm)
    "#define MPI_VERSION 3\ntypedef int MPI_Comm;\n",  # MPICH-like
    "#define MPI_VERSION 3\ntypedef struct ompi_communicator_t *MPI_Comm;\n",  # Open MPI-like
])
def test_header_binds_any_mpi_comm(tmp_path, prelude):
    """tioga_init_ takes the caller's MPI_Comm when <mpi.h> comes first
    (tiogaInterface.C:61); the C header compiles with each MPI's handle type."""
    import subprocess

    src = tmp_path / "use.c"
    src.write_text(prelude + '#include "tioga_amd.h"\n'
                   "void f(void) { tioga_comm_t c = 0; tioga_init_(c); }\n")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
