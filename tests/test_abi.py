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


def test_tune_rejects_null_ctx():
    lib = cac.load_library()
    assert lib.cac_tune_set(None, cac.TUNE_KS_CONTRACT_BLOCK, 512, 128) == cac.CAC_EINVAL
    v = ctypes.c_int32()
    assert lib.cac_tune_get(None, cac.TUNE_KS_CONTRACT_BLOCK, 512, ctypes.byref(v)) == cac.CAC_EINVAL


def test_tuning_file_shape():
    # a committed tuning table names only known classes and legal block sizes
    import json

    if not cac.TUNING_PATH.exists():
        return
    t = json.loads(cac.TUNING_PATH.read_text())["ks_contract_block"]
    for key, b in t.items():
        assert int(key) in cac.TUNE_KS_CONTRACT_KEYS
        assert b == 0 or (max(64, int(key) // 8) <= b <= 1024 and b & (b - 1) == 0)
    occ = json.loads(cac.TUNING_PATH.read_text()).get("sample_occupancy", {})
    for key, k in occ.items():
        assert int(key) in cac.TUNE_SAMPLE_OCCUPANCY_KEYS and k in (0, 1, 2)
    wide = json.loads(cac.TUNING_PATH.read_text()).get("wide_sampling_log2", 0)
    assert 0 <= int(wide) <= 63
    comp = json.loads(cac.TUNING_PATH.read_text()).get("ks_compact_block", {})
    for key, b in comp.items():
        assert int(key) in cac.TUNE_KS_COMPACT_KEYS and b in (0, 256, 512, 1024)


def test_record_layouts():
    assert cac.UEDGE.itemsize == 8 and cac.EDGE.itemsize == 16
    assert cac.EDGE.fields["w"][1] == 8
