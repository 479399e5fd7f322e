"""The C ABI library loads and exports every symbol the public headers
declare; host-side pieces (data generators) behave. CPU only: no kernels."""
import ctypes as C
import glob
import os
import re

import numpy as np

import rasterjoin as rj

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(rj_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(rj.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_create_reports_status_without_crashing():
    ctx = C.c_void_p()
    rc = rj.lib().rj_create(0, 1, C.byref(ctx))
    assert rc in (0, -4)
    if rc == 0:
        assert rj.lib().rj_destroy(ctx) == 0
    assert rj.lib().rj_destroy(None) == -1


def test_point_generator_is_deterministic_and_sliceable():
    a = rj.gen_points(10_000)
    b = rj.gen_points(10_000)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    c = rj.gen_points(4_000, first=3_000)
    assert np.array_equal(c[0], a[0][6_000:14_000])
    assert np.array_equal(c[2], a[2][3_000:7_000])
    xy, t, fare, pas = a
    assert xy.min() >= 0 and xy.max() < 50000
    assert t.min() >= 1230768000 and t.max() < 1230768000 + 31536000
    assert pas.min() >= 1 and pas.max() <= 6
    assert 10 < np.median(fare) < 14          # lognormal(2.5, 0.6): median e^2.5
    u = (xy[0::2] > 10) & (xy[0::2] < 49990)
    assert u.mean() > 0.99


def test_voronoi_cells_share_exact_vertex_chains():
    xy, off = rj.gen_voronoi(60, seed=2, spacing=(60.0, 60.0))
    assert len(off) == 61 and np.all(np.diff(off) >= 3)
    v = xy.reshape(-1, 2)
    assert v[:, 0].min() == 0 and v[:, 0].max() == 50000
    edges = {}
    for p in range(60):
        ring = [tuple(v[k]) for k in range(off[p], off[p + 1])]
        for k in range(len(ring)):
            a, b = ring[k - 1], ring[k]
            assert a != b
            key = (min(a, b), max(a, b))
            edges.setdefault(key, []).append(p)
    on_box = 0
    for (a, b), owners in edges.items():
        border = (a[0] == b[0] and a[0] in (0.0, 50000.0)) or (a[1] == b[1] and a[1] in (0.0, 50000.0))
        if border:
            on_box += 1
            assert len(owners) == 1
        else:
            assert len(owners) == 2 and owners[0] != owners[1]
    assert on_box > 0
    counts = np.diff(off)
    assert counts.min() >= 40


def test_constraint_record_is_16_bytes():
    assert C.sizeof(rj.Constraint) == 16


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes records of the Python binding have the sizes and field
    offsets of the header's structs, as a C compiler lays them out."""
    records = {"rj_constraint": rj.Constraint, "rj_table_options": rj.TableOptions, "rj_query": rj.Query,
               "rj_result": rj.Result, "rj_server_config": rj.ServerConfig, "rj_server_stats": rj.ServerStats}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "rasterjoin.h"', "int main(void) {"]
    for cname, py in records.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            fname = f[0]
            if fname == "val":  # the union of rj_constraint
                continue
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.split() for l in out if l)
    for cname, py in records.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f in py._fields_:
            if f[0] == "val":
                continue
            assert int(got[f"{cname}.{f[0]}"]) == getattr(py, f[0]).offset, (cname, f[0])
