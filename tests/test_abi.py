"""The C ABI library loads and exports every symbol include/cac.h declares
(no compute calls: this runs without a GPU)."""
import ctypes

import cac


def test_library_exports_every_declared_symbol():
    lib = cac.load_library()
    declared = cac.declared_symbols()
    assert "cac_cc" in declared and "cac_ctx_create" in declared
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"
    # every declared function has a ctypes signature in the binding
    assert set(declared) <= set(cac.SIGNATURES), set(declared) - set(cac.SIGNATURES)


def test_strerror_names_every_code():
    lib = cac.load_library()
    for code in (cac.CAC_OK, cac.CAC_EINVAL, cac.CAC_ENOMEM, cac.CAC_EDEVICE, cac.CAC_ENOTCONN,
                 cac.CAC_ERANGE, cac.CAC_EIO):
        assert lib.cac_strerror(code) and lib.cac_strerror(code) != b"unknown error"


def test_ctx_create_rejects_null_out():
    lib = cac.load_library()
    assert lib.cac_ctx_create(0, None) == cac.CAC_EINVAL


def test_record_layouts():
    assert cac.UEDGE.itemsize == 8 and cac.EDGE.itemsize == 16
    assert cac.EDGE.fields["w"][1] == 8
