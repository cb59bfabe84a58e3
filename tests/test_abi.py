"""The C ABI: every built library loads and exports exactly the functions
include/cumayo.h declares, and the Python bindings cover all of them; the
launch configuration calls and the argument checks of the batched entry
points behave as declared. Runs without a GPU: loading a library registers
its code objects lazily, and the entry points called here return before
enqueuing any work."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cumayo.h")

# api.cuh:9-47 names (plus aes128_tables_init, aes128_ctr.cuh:16)
REFERENCE_API = [
    "crypto_sign_keypair", "crypto_sign_keypair_DtH",
    "crypto_sign_expandsk_HtD", "crypto_sign_expandsk", "crypto_sign_expandsk_DtH",
    "crypto_sign_expandpk_HtD", "crypto_sign_expandpk", "crypto_sign_expandpk_DtH",
    "crypto_sign_sign_HtD", "crypto_sign_sign", "crypto_sign_sign_DtH",
    "crypto_sign_signWithesk_HtD", "crypto_sign_signWithesk", "crypto_sign_signWithesk_DtH",
    "crypto_sign_verify_HtD", "crypto_sign_verify", "crypto_sign_verify_DtH",
    "crypto_sign_verifyWithepk_HtD", "crypto_sign_verifyWithepk", "crypto_sign_verifyWithepk_DtH",
    "aes128_tables_init",
]


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = "\n".join(l for l in src.splitlines() if not l.lstrip().startswith("#"))
    src = re.sub(r"typedef\s+struct\s*\{.*?\}\s*\w+\s*;", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_]\w*)\s*\(", src)) - {"CUMAYO_DEFAULT", "void"})


def test_header_declares_reference_api():
    decl = declared_functions()
    for name in REFERENCE_API:
        assert name in decl, name
    assert len(decl) >= 34


def test_bindings_cover_header():
    from cumayo_amd import engine, serve
    bound = set(engine.EXPORTS) | set(serve.EXPORTS)
    assert bound == set(declared_functions())


@pytest.mark.parametrize("version", [1, 2, 3, 5])
def test_library_exports(version):
    from cumayo_amd import build
    path = build.lib_path(version)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")
    lib = ctypes.CDLL(path)
    for name in declared_functions():
        assert hasattr(lib, name), f"{name} missing from {os.path.basename(path)}"
    # the parameter set is reported without touching a GPU
    from cumayo_amd.engine import _Sizes
    s = _Sizes()
    lib.cumayo_get_sizes(ctypes.byref(s))
    assert s.version == version


@pytest.mark.parametrize("version", [1, 2, 3, 5])
def test_launch_variants_and_config(version):
    """The launch autotuner's variant lists and configuration, without a
    GPU: entry 0 of every list is the configuration in force at load, a
    compiled variant is accepted and an unknown one refused unchanged."""
    from cumayo_amd import build
    from cumayo_amd.engine import EXPORTS, LAUNCH_FIELDS, LaunchConfig, load_library
    if not os.path.exists(build.lib_path(version)):
        pytest.skip("library not built")
    lib = load_library(version)
    variants = {}
    for name, f in LAUNCH_FIELDS.items():
        buf = (ctypes.c_int32 * 16)()
        cnt = lib.cumayo_launch_variants(f, buf, 16)
        assert 2 <= cnt <= 16, name
        variants[name] = [buf[i] for i in range(cnt)]
        assert len(set(variants[name])) == cnt, (name, variants[name])
    assert lib.cumayo_launch_variants(99, None, 0) == -1
    # sign_chunks: three kernels over the batch (1, the default), that in two
    # halves (2), one persistent fused kernel (0)
    assert variants["order"] == [0, 1] and variants["sign_chunks"] == [1, 2, 0]
    cfg = LaunchConfig()
    lib.cumayo_get_launch_config(ctypes.byref(cfg))
    assert cfg.as_dict() == {"sign_prep": variants["sign_prep"][0], "sign_solve": variants["sign_solve"][0],
                             "sign_order": 0, "verify": variants["verify"][0], "verify_order": 0,
                             "sign_chunks": 1}
    # quadratic-form variants are coded 10 x rows per lane + waves per SIMD,
    # shared-key variants 1000 + 10 x items per workgroup + table columns
    for code in variants["sign_prep"] + variants["verify"]:
        if code >= 1000:
            assert (code % 1000) // 10 in (4, 8) and code % 10 in (1, 2), code
        else:
            assert code // 10 in (1, 2, 3) and 1 <= code % 10 <= 8, code
    assert [c for c in variants["sign_prep"] if c >= 1000] == [c for c in variants["verify"] if c >= 1000] != []
    alt = LaunchConfig(variants["sign_prep"][1], variants["sign_solve"][2], 1, variants["verify"][1], 1, 2)
    assert lib.cumayo_set_launch_config(ctypes.byref(alt)) == 0
    got = LaunchConfig()
    lib.cumayo_get_launch_config(ctypes.byref(got))
    assert got.as_dict() == alt.as_dict()
    bad = LaunchConfig(variants["sign_prep"][0] + 50, variants["sign_solve"][0], 0, variants["verify"][0], 0, 1)
    assert lib.cumayo_set_launch_config(ctypes.byref(bad)) != 0
    for chunks in (3, -1):
        bad = LaunchConfig(variants["sign_prep"][0], variants["sign_solve"][0], 0, variants["verify"][0], 0, chunks)
        assert lib.cumayo_set_launch_config(ctypes.byref(bad)) != 0
    lib.cumayo_get_launch_config(ctypes.byref(got))
    assert got.as_dict() == alt.as_dict()
    assert lib.cumayo_set_launch_config(ctypes.byref(cfg)) == 0
    assert "cumayo_autotune" in EXPORTS


def _max_batch():
    m = re.search(r"#define\s+CUMAYO_MAX_BATCH\s+(\d+)u", open(HEADER).read())
    assert m, "CUMAYO_MAX_BATCH missing from the header"
    return int(m.group(1))


@pytest.mark.parametrize("version", [1, 2, 3, 5])
def test_batch_entry_points_refuse_bad_arguments(version):
    """Argument checks that return before any work is enqueued (so they run
    without a GPU): an empty batch is a no-op (0), a batch above
    CUMAYO_MAX_BATCH and a null pointer are refused with hipErrorInvalidValue
    (1), also by the void calls of the reference API, which record it for
    cumayo_last_error. The pointer 4096 is a stand-in that is never read."""
    from cumayo_amd import build
    from cumayo_amd.engine import load_library
    if not os.path.exists(build.lib_path(version)):
        pytest.skip("library not built")
    lib = load_library(version)
    big = _max_batch() + 1
    assert big - 1 == 2 ** 24 - 1
    p = 4096
    lib.cumayo_last_error()  # clear
    sign_prep = lambda n, key=p: lib.cumayo_sign_prepared(p, key, p, p, p, p, n, p, None, 3)
    verify_prep = lambda n, key=p: lib.cumayo_verify_prepared(p, key, p, p, p, p, p, n, p, None, 3)
    sign_tab = lambda n, key=p: lib.cumayo_sign_table_stages(p, 64, key, p, p, p, p, n, p, None, 3)
    verify_tab = lambda n, key=p: lib.cumayo_verify_table_stages(p, 64, key, p, p, p, p, p, n, p, None, 3)
    wire_sign = lambda n, key=p: lib.cumayo_crypto_sign(p, key, p, p, p, p, p, n, p, None)
    wire_open = lambda n, key=p: lib.cumayo_crypto_sign_open(p, key, p, p, p, p, p, p, p, n, p, None)
    for name, call in (("sign_prepared", sign_prep), ("verify_prepared", verify_prep),
                       ("sign_table_stages", sign_tab), ("verify_table_stages", verify_tab),
                       ("crypto_sign", wire_sign), ("crypto_sign_open", wire_open)):
        assert call(0) == 0, name
        assert call(big) == 1, name
        assert call(big - 1, key=None) == 1, name  # null key_idx
    # a prepared table must be 16-byte aligned
    assert lib.cumayo_sign_prepared(p + 4, p, p, p, p, p, 8, p, None, 3) == 1
    # the comparison form (stage bit 4) is not available over prepared tables
    assert lib.cumayo_sign_prepared(p, p, p, p, p, p, 8, p, None, 7) == 1
    # key preparation: at most 65,535 keys per call, aligned destination
    assert lib.cumayo_prepare_esk(p, 64, 0, p, None) == 0
    assert lib.cumayo_prepare_esk(p, 64, 65536, p, None) == 1
    assert lib.cumayo_prepare_epk(p, 64, 4, p + 8, None) == 1
    # the reference API's void calls record the refusal
    lib.cumayo_last_error()  # clear (the refusals above are recorded too)
    assert lib.cumayo_last_error() == 0
    lib.crypto_sign_expandsk(p, 1 << 20, big, None)
    assert lib.cumayo_last_error() == 1
    assert lib.cumayo_last_error() == 0  # reading clears it
    lib.crypto_sign_verify(p, 32, 1 << 20, big, None)
    assert lib.cumayo_last_error() == 1
