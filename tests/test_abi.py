"""The C-ABI library loads and exports every entry point declared in include/vpbs.h."""
import ctypes
import pathlib
import re
import subprocess

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "vpbs.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int64_t|int|void|const char \*)\s*\**\s*(vpbs_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "vpbs_create" in syms and "vpbs_compute" in syms and len(syms) >= 10


def test_library_exports_all_declared_symbols(vp):
    from importlib import import_module  # noqa: F401

    lib = ctypes.CDLL(str(ROOT / "vortexpasta.jl_amd" / "libvpbs.so"))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(vp._lib.EXPORTED) == set(declared_symbols())


def _c_layout(struct, fields, tmp_path):
    src = tmp_path / "layout.c"
    body = "".join(f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields)
    src.write_text(
        "#include <stddef.h>\n#include <stdio.h>\n#include \"vpbs.h\"\n"
        f"int main(void){{{body}printf(\"%zu\\n\", sizeof({struct}));return 0;}}\n"
    )
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return [int(x) for x in out]


def test_struct_layouts_match_header(vp, tmp_path):
    for cstruct, pystruct in (("vpbs_params", vp._lib.VpbsParams), ("vpbs_timings", vp._lib.VpbsTimings),
                              ("vpbs_learned_report", vp._lib.VpbsLearnedReport)):
        names = [n for n, _ in pystruct._fields_]
        got = _c_layout(cstruct, names, tmp_path)
        want = [getattr(pystruct, n).offset for n in names] + [ctypes.sizeof(pystruct)]
        assert got == want, (cstruct, got, want)


def test_report_structs_convert_to_dicts(vp):
    """The ctypes mirrors of vpbs_timings and vpbs_learned_report turn into plain dicts."""
    t = vp._lib.VpbsTimings()
    t.pair = 1.5
    t.grid[:] = (768, 768, 640)
    d = t.as_dict()
    assert d["pair"] == 1.5 and d["grid"] == (768, 768, 640) and d["cells"] == (0, 0, 0)
    assert set(d) == {n for n, _ in vp._lib.VpbsTimings._fields_}
    r = vp._lib.VpbsLearnedReport(err_corrected=1e-5, fallback=1)
    assert r.as_dict() == {"err_corrected": 1e-5, "err_uncorrected": 0.0, "nsample": 0, "fallback": 1,
                           "total_ms": 0.0}
