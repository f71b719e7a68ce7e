"""The C ABI library loads and exports every entry point include/falx_gpu.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from falx_amd import build, gpu, spec as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "falx_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:falx_status|void\*?|double|uint64_t)\s+(falx_\w+)\s*\(", text, re.M)))


def test_header_declares_the_mirror_surface():
    syms = declared_symbols()
    assert set(syms) == set(gpu.EXPORTED), set(syms) ^ set(gpu.EXPORTED)


def test_library_exports_every_declared_symbol():
    path = build.gpu_lib_path()
    if not os.path.exists(path):
        pytest.skip("libfalx_gpu.so not built")
    lib = ctypes.CDLL(path)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_struct_layouts_match_header():
    # sizes the C compiler gives these structs (x86-64 / gfx950 share them)
    assert ctypes.sizeof(S.CSpec) == 16 + 4 * 5 + 8 * 32
    assert ctypes.sizeof(S.CColumnBuf) == 40
    assert ctypes.sizeof(S.CColumns) == 16 + 40 * 32
    assert ctypes.sizeof(S.CColumnStats) == 48
    assert ctypes.sizeof(S.CNdjsonIdScore) == 48
    assert ctypes.sizeof(S.CRequest) == 56
    assert ctypes.sizeof(S.CResult) == 16 + (8 + 48 * 32) + 16 + 32 + 48
    assert ctypes.sizeof(S.CBatchTuning) == 4 * 11


def test_spec_toml_front_end():
    """docs/spec.md examples through the spec.rs validation rules."""
    csv_hash = S.parse_spec_toml('name = "csv_hash"\nstructural = [",", "\\n"]\nquote = \'"\'\ncomment = "#"\n'
                                 '[[columns]]\nindex = 0\nname = "key"\ntype = "string"\n'
                                 '[[columns]]\nindex = 1\nname = "amount"\ntype = "i64"\n')
    assert csv_hash.structural == b",\n" and csv_hash.quote == 34 and csv_hash.comment == 35
    assert [(c.index, c.type) for c in csv_hash.columns] == [(0, S.STR), (1, S.I64)]
    nd = S.parse_spec_toml('name = "ndjson"\nstructural = [",", ":", "\\n"]\nquote = \'"\'\n'
                           'escape = "backslash"\nnesting = ["{}", "[]"]\n')
    assert nd.escape == 92 and set(b"{}[]") <= set(nd.structural)
    for bad in ['name = "x"\nstructural = [","]\ncomment = "\\n"\n',
                'name = "x"\nstructural = [","]\nquote = "#"\ncomment = "#"\n',
                'name = "x"\nstructural = [",,"]\n', 'structural = [","]\n',
                'name = "x"\nstructural = [","]\nescape = "weird"\n']:
        with pytest.raises(ValueError):
            S.parse_spec_toml(bad)


def test_spec_validate_rejections_through_the_c_abi():
    """falx_spec_validate (no device needed) applies the spec.rs / formats.rs
    rules: every reference dialect passes; each broken rule is rejected."""
    path = build.gpu_lib_path()
    if not os.path.exists(path):
        pytest.skip("libfalx_gpu.so not built")
    L = gpu.load()
    for spec in list(S.DIALECTS.values()) + [S.CSV_HASH, S.CSV_TYPED, S.GEO_ALL, S.HASH_TYPED, S.TSV_TYPED]:
        assert L.falx_spec_validate(ctypes.byref(spec.to_c())) == S.OK, spec.name

    def broken(**kw):
        c = S.CSV_HASH.to_c()
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    too_many_cols = S.CSV_HASH.to_c()
    too_many_cols.n_columns = 33
    bad_type = S.CSV_HASH.to_c()
    bad_type.columns[1].type = 4
    for c in [broken(n_structural=0), broken(n_structural=17), broken(quote=256), broken(quote=-2),
              broken(escape=ord('"')), broken(comment=ord("\n")), broken(comment=ord("\r")),
              broken(comment=ord('"')), broken(escape=300), too_many_cols, bad_type]:
        assert L.falx_spec_validate(ctypes.byref(c)) == S.E_INVALID
    assert L.falx_spec_validate(None) == S.E_INVALID
