"""The C-ABI library loads and exports every entry point include/dpx.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import dpx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "dpx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dpx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "dpx_zuker" in names and "dpx_swat" in names and "dpx_zuker_batch" in names


def test_library_exports_all_declared():
    lib = ctypes.CDLL(dpx.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_exports():
    assert set(dpx.EXPORTS) <= set(_declared())


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the header's structs have the C sizes and offsets."""
    import subprocess
    structs = {"dpx_trace_t": dpx.Trace, "dpx_swat_scores": dpx.SwatScores, "dpx_server_config": dpx.ServerConfig,
               "dpx_request": dpx.Request, "dpx_result": dpx.Result, "dpx_server_stats": dpx.ServerStats,
               "dpx_tuning": dpx.Tuning, "dpx_tune_report": dpx.TuneReport, "dpx_grammar_info": dpx.GrammarInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dpx.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append('printf("%%zu\\n", sizeof(%s));' % cname)
        for f in py._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (cname, f[0]))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = []
    for py in structs.values():
        want.append(ctypes.sizeof(py))
        want += [getattr(py, f[0]).offset for f in py._fields_]
    assert got == want


def test_tuning_table_file():
    """The committed MI355X tuning table: every kind, every bucket, shapes the
    library accepts (dpx_set_tuning's rules)."""
    import json
    with open(dpx.TUNING_MI355X) as f:
        d = json.load(f)
    t = d["table"]
    assert set(t) == set(dpx.KIND_NAMES.values())
    for kind, rows in t.items():
        assert len(rows) == dpx.TUNE_BUCKETS
        for r in rows:
            assert r["wg_per_cu"] >= 0 and r["chunk"] in (0, 16, 32, 64)
    for kind, rep in d["report"].items():
        for b in range(dpx.TUNE_BUCKETS):
            rows = [x for x in rep if x["bucket"] == b]
            chosen = [x for x in rows if x["chosen"]]
            assert len(chosen) == 1 and chosen[0]["ms"] == min(x["ms"] for x in rows)
            assert t[kind][b] == {"wg_per_cu": chosen[0]["wg_per_cu"], "chunk": chosen[0]["chunk"]}
