"""Host side of libpumatch.so: ABI exports, graph formats, generators,
pattern parsing and planning.  No GPU needed."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import pumatch_amd as pm
from tests.helpers import C1_EDGES, edges_to_adj, g_text


def test_library_exports_every_declared_symbol(root):
    header = open(os.path.join(root, "include", "pumatch.h")).read()
    declared = set(re.findall(r"\b(pm_[a-z_]+)\s*\(", header))
    assert declared, "no declarations found"
    L = pm.lib()
    for name in sorted(declared):
        assert hasattr(L, name), f"{name} declared but not exported"
    assert declared == set(pm._SIGS), "binding and header disagree"


def test_rmat_is_deterministic_symmetric_sorted():
    a = pm.Graph.rmat(10, 8, seed=3, device=False)
    b = pm.Graph.rmat(10, 8, seed=3, device=False)
    c = pm.Graph.rmat(10, 8, seed=4, device=False)
    ra, ca, _ = a.csr()
    rb, cb, _ = b.csr()
    rc, cc, _ = c.csr()
    assert np.array_equal(ra, rb) and np.array_equal(ca, cb)
    assert not (ra.size == rc.size and ca.size == cc.size and np.array_equal(ca, cc))
    n = ra.size - 1
    src = np.repeat(np.arange(n), np.diff(ra).astype(np.int64))
    assert np.all(src != ca), "self loop"
    key = src.astype(np.int64) * n + ca
    assert np.all(np.diff(key) > 0), "rows not strictly sorted"
    fwd = set(zip(src.tolist(), ca.tolist()))
    assert all((y, x) in fwd for x, y in fwd), "not symmetric"
    assert a.info().max_degree == int(np.diff(ra).max())


def test_chung_lu_shape():
    g = pm.Graph.chung_lu(20000, 2.2, 500, 20, seed=4, device=False)
    r, c, _ = g.csr()
    deg = np.diff(r)
    assert 10 < deg.mean() < 21
    assert deg.max() <= 700 and deg[:10].mean() > 10 * deg.mean()


def test_uniform_labels_range_and_seed():
    g = pm.Graph.rmat(9, 8, seed=1, device=False)
    g.set_uniform_labels(4, 2)
    _, _, lab = g.csr()
    assert lab.min() == 0 and lab.max() == 3
    counts = np.bincount(lab, minlength=4)
    assert counts.min() > 0.15 * lab.size


def test_file_round_trip(tmp_path):
    g = pm.Graph.rmat(9, 8, seed=2, device=False)
    g.set_uniform_labels(3, 7)
    prefix = str(tmp_path / "g")
    g.write(prefix)
    meta = open(prefix + ".meta.txt").read().split()
    r, c, l = g.csr()
    assert meta == [str(r.size - 1), str(c.size), "4", str(int(np.diff(r).max()))]
    h = pm.Graph.open(prefix, labels=True, device=False)
    r2, c2, l2 = h.csr()
    assert np.array_equal(r, r2) and np.array_equal(c, c2) and np.array_equal(l, l2)
    h2 = pm.Graph.open(prefix, labels=False, device=False)
    assert h2.csr()[2] is None


def test_open_missing_and_malformed(tmp_path):
    with pytest.raises(pm.PumatchError):
        pm.Graph.open(str(tmp_path / "nothing"), device=False)
    (tmp_path / "bad.meta.txt").write_text("10 20 8 3\n")
    with pytest.raises(pm.PumatchError):
        pm.Graph.open(str(tmp_path / "bad"), labels=False, device=False)


def test_from_csr_validation():
    with pytest.raises(pm.PumatchError):  # unsorted adjacency
        pm.Graph.from_csr([0, 2, 3, 4], [2, 1, 0, 0], device=False)
    with pytest.raises(pm.PumatchError):  # label above 31
        pm.Graph.from_csr([0, 1, 2], [1, 0], labels=[0, 40], device=False)
    g = pm.Graph.from_csr([0, 1, 2], [1, 0], device=False)
    assert (g.n, g.m) == (2, 2)


def test_reference_patterns_plan_like_the_oracle(patterns_dir):
    for i in range(1, 21):
        path = os.path.join(patterns_dir, f"{i}.g")
        text = open(path).read()
        for labeled in (True, False):
            info = pm.Pattern.load(path, labeled).info()
            ref = oracle.plan_info(oracle.parse_g(text, labeled), labeled)
            k = info.k
            assert k == len(ref["deg"])
            assert info.automorphisms == ref["aut"]
            assert list(info.degree)[:k] == ref["deg"]
            assert list(info.nbr_mask)[1:k] == ref["A"][1:]
            if not labeled:
                assert list(info.sym_mask)[:k] == ref["B"]
                assert info.rc_order == ref["rc_order"]


def test_pattern_errors():
    with pytest.raises(pm.PumatchError):
        pm.Pattern.parse("# t 0\n")
    with pytest.raises(pm.PumatchError):  # 8 vertices
        pm.Pattern.parse("\n".join(f"v {i} 1" for i in range(8)) + "\ne 0 1 1\n")
    with pytest.raises(pm.PumatchError):  # endpoint out of range
        pm.Pattern.parse("v 0 1\nv 1 1\nv 2 1\ne 0 5 1\n")
    with pytest.raises(pm.PumatchError):  # vertex 2 has no earlier neighbour
        pm.Pattern.parse(g_text(4, edges_to_adj(4, [(0, 1), (2, 3), (1, 3)])))
    with pytest.raises(pm.PumatchError):  # label outside 0..31
        pm.Pattern.parse("v 0 40\nv 1 1\nv 2 1\ne 0 1 1\ne 1 2 1\n")


def test_vertex_induced_mode_flag():
    text = g_text(7, edges_to_adj(7, C1_EDGES))
    p = pm.Pattern.parse(text, True, induced=True)
    assert p.info().vertex_induced == 1 and p.info().labeled == 1
    q = pm.Pattern.parse(text, True)
    assert q.info().vertex_induced == 0
    # vertex-induced plans fold nothing: the pendant path stays in the core
    # and needs slots of its own
    assert p.info().nslots > q.info().nslots
    assert pm.Pattern.parse(text, False, induced=True).info().vertex_induced == 1
    h = pm.C.c_void_p()
    assert pm.lib().pm_pattern_parse(text.encode(), 4, pm.C.byref(h)) == pm.PM_ERR_ARG


def test_folding_of_pendant_trees():
    # labeled K4 + path 3-4-5-6: the path folds onto vertex 3, the core is K4
    p = pm.Pattern.parse(g_text(7, edges_to_adj(7, C1_EDGES)), True)
    assert p.info().nslots >= 1
    # a labeled path folds down to the job edge: still plans
    pm.Pattern.parse(g_text(4, edges_to_adj(4, [(0, 1), (1, 2), (2, 3)])), True)


def test_cli_usage(root):
    cli = pm.CLI_PATH
    assert os.path.exists(cli)
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def test_snap_converter(tmp_path):
    """bin/convert: ids renumbered by first appearance, both directions,
    rows sorted and deduplicated, self-loops dropped, labels rand() % 10
    after srand(0) as the reference's labelize step."""
    import ctypes
    import subprocess

    snap = tmp_path / "snap.txt"
    snap.write_text("# comment\n10 20\n20 30\n\n30 10\n10 10\n20 10\n40 40\n50 20\n")
    prefix = str(tmp_path / "g")
    exe = os.path.join(os.path.dirname(pm.CLI_PATH), "convert")
    r = subprocess.run([exe, str(snap), prefix], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert open(prefix + ".meta.txt").read().split() == ["5", "8", "4", "3"]
    g = pm.Graph.open(prefix, labels=True, device=False)
    rp, ci, lb = g.csr()
    # 10->0, 20->1, 30->2, 40->3, 50->4
    rows = [sorted(ci[rp[v]:rp[v + 1]].tolist()) for v in range(5)]
    assert rows == [[1, 2], [0, 2, 4], [0, 1], [], [1]]
    libc = ctypes.CDLL(None)
    libc.srand(0)
    assert lb.tolist() == [libc.rand() % 10 for _ in range(5)]
    r = subprocess.run([exe, str(snap), prefix, "--keep-self-loops", "--labels", "3", "--seed", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    g = pm.Graph.open(prefix, labels=True, device=False)
    rp, ci, lb = g.csr()
    assert sorted(ci[rp[0]:rp[1]].tolist()) == [0, 1, 2] and ci[rp[3]:rp[4]].tolist() == [3]
    assert lb.max() < 3


def test_model_parameters_round_trip_and_init():
    """pm_model_create / get / set on the host: He-uniform weights within
    their bounds, zero biases, exact round trip, layout size."""
    from oracle import mlp

    m = pm.Model(5)
    w = m.params
    assert w.size == pm.N_PARAMS == mlp.NPARAMS
    W1 = w[mlp.OFF["W1"][0]:mlp.OFF["W1"][0] + mlp.OFF["W1"][1]]
    W2 = w[mlp.OFF["W2"][0]:mlp.OFF["W2"][0] + mlp.OFF["W2"][1]]
    assert np.abs(W1).max() <= np.sqrt(6 / 16) and np.abs(W2).max() <= np.sqrt(6 / 64)
    assert np.abs(W1).std() > 0.1
    for name in ("b1", "b2", "b3"):
        o, n = mlp.OFF[name]
        assert not np.any(w[o:o + n])
    assert np.array_equal(pm.Model(5).params, w)  # seeded
    assert not np.array_equal(pm.Model(6).params, w)
    v = np.random.default_rng(0).normal(size=pm.N_PARAMS).astype(np.float32)
    m.params = v
    assert np.array_equal(m.params, v)


def test_batch_and_tune_argument_errors():
    """Argument checks of the batched, tuning and model entry points run
    before any device work."""
    L = pm.lib()
    g = pm.Graph.rmat(6, 4, seed=1, device=False)
    p = pm.Pattern.parse(g_text(3, edges_to_adj(3, [(0, 1), (1, 2)])), False)
    rs = (pm.Result * 2)()
    assert L.pm_match_batch(None, None, 1, None, rs) == pm.PM_ERR_ARG
    assert L.pm_match_batch(g._h, (pm.C.c_void_p * 1)(p._h.value), -1, None, rs) == pm.PM_ERR_ARG
    assert L.pm_match_batch(g._h, (pm.C.c_void_p * 2)(p._h.value, None), 2, None, rs) == pm.PM_ERR_ARG
    assert L.pm_match_batch(g._h, (pm.C.c_void_p * 1)(p._h.value), 0, None, rs) == pm.PM_OK
    best, rep = pm.Opts(), pm.TuneReport()
    hs = (pm.C.c_void_p * 1)(p._h.value)
    assert L.pm_autotune(g._h, hs, 0, None, None, pm.C.byref(best), pm.C.byref(rep)) == pm.PM_ERR_ARG
    assert L.pm_autotune(g._h, hs, 1, None, None, None, pm.C.byref(rep)) == pm.PM_ERR_ARG
    assert L.pm_model_get(None, None) == pm.PM_ERR_ARG
    assert L.pm_model_predict(None, None, 0, None) == pm.PM_ERR_ARG
    assert L.pm_model_train(None, None, None, None, 1, None, None) == pm.PM_ERR_ARG


def _header_fields(struct_name):
    """(name, C type, array length) of each field of `typedef struct {...}
    struct_name;` in include/pumatch.h, in order."""
    import re
    text = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                             "include", "pumatch.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + struct_name + r";", text)
    assert m, struct_name
    out = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        dm = re.match(r"(const )?(struct )?([A-Za-z_0-9]+)\s*(\*?)\s*(.*)", decl)
        ctype, ptr, names = dm.group(3), dm.group(4), dm.group(5)
        for nm in names.split(","):
            nm = nm.strip()
            star = ptr or nm.startswith("*")
            nm = nm.lstrip("*").strip()
            am = re.match(r"([A-Za-z_0-9]+)(\[([A-Z_0-9]+)\])?", nm)
            out.append((am.group(1), "ptr" if star else ctype, am.group(3)))
    return out


@pytest.mark.parametrize("struct_name,cls", [("pm_opts", "Opts"), ("pm_result", "Result"),
                                             ("pm_pattern_info", "PatternInfo"),
                                             ("pm_graph_info", "GraphInfo"),
                                             ("pm_train_opts", "TrainOpts"),
                                             ("pm_tune_opts", "TuneOpts"),
                                             ("pm_tune_report", "TuneReport")])
def test_ctypes_mirrors_match_the_header(struct_name, cls):
    import ctypes as C
    widths = {"int32_t": 4, "uint32_t": 4, "int64_t": 8, "uint64_t": 8, "double": 8,
              "float": 4, "ptr": 8}
    fields = _header_fields(struct_name)
    mirror = getattr(pm, cls)._fields_
    assert [f[0] for f in fields] == [f[0] for f in mirror], struct_name
    for (name, ctype, arr), (_, mtype) in zip(fields, mirror):
        n = 1
        if arr:
            n = {"PM_MAX_PATTERN": 7}.get(arr, None) or int(arr)
        assert C.sizeof(mtype) == widths[ctype] * n, (struct_name, name)
