"""The drop-in command line: parameter file, .mod / FASTA / BED loaders,
element selection and the reference's output files."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth

CLI = os.path.join(os.path.dirname(E.LIBPATH), "phyloacc-st-gpu")


def _write_inputs(tmp, w, extra="", shuffle_rows=True, ids=None):
    tmp = str(tmp)
    with open(os.path.join(tmp, "tree.mod"), "w") as f:
        f.write(w.mod_text)
    order = np.random.default_rng(0).permutation(w.tree.S) if shuffle_rows else np.arange(w.tree.S)
    with open(os.path.join(tmp, "aln.fa"), "w") as f:
        for s in order:
            row = w.align[s].tobytes().decode().upper()
            f.write(f">{w.tree.names[s]}\n")
            for k in range(0, len(row), 60):      # wrapped FASTA lines
                f.write(row[k:k + 60] + "\n")
    with open(os.path.join(tmp, "elem.bed"), "w") as f:
        pos = 1000
        for c in range(w.C):
            f.write(f"elem{c}\t{pos}\t{pos + int(w.sizes[c])}\n")
            pos += int(w.sizes[c]) + 17            # BED gaps do not matter: slices are contiguous
    out = os.path.join(tmp, "out")
    os.makedirs(out, exist_ok=True)
    names = w.tree.names
    lines = [f"PHYTREE_FILE {tmp}/tree.mod", f"ALIGN_FILE {tmp}/aln.fa", f"SEG_FILE {tmp}/elem.bed",
             f"RESULT_FOLDER {out}", "PREFIX run", "CHAIN 1",
             "OUTGROUP " + ";".join(names[i] for i in w.outgroup),
             "CONSERVE " + ";".join(names[i] for i in w.conserve),
             "TARGETSPECIES " + ";".join(names[i] for i in w.target),
             "# a comment line", ""]
    if ids is not None:
        with open(os.path.join(tmp, "ids.txt"), "w") as f:
            f.write("\n".join(str(i) for i in ids) + "\n")
        lines.append(f"ID_FILE {tmp}/ids.txt")
    p = os.path.join(tmp, "params.txt")
    with open(p, "w") as f:
        f.write("\n".join(lines) + "\n" + extra)
    return p, out


def _run(*args, check=True):
    # CPU-only tests: the program runs as its own process
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=600)
    if check:
        assert r.returncode == 0, r.stdout + r.stderr
    return r


def _run_in_process(*args):
    """The same program through pa_cli_main in this process: a process that
    has used the GPU must not start other programs."""
    import ctypes as C
    argv = [CLI.encode(), *[a.encode() for a in args]]
    arr = (C.c_char_p * len(argv))(*argv)
    fn = E.lib().pa_cli_main
    fn.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
    fn.restype = C.c_int
    return fn(len(argv), arr)


def test_dry_run_matches_engine_encoding(tmp_path):
    w = synth.make_workload(1, seed=21, n_elements=12)
    p, _ = _write_inputs(tmp_path, w, extra="MIN_LEN 120\nGAP_PROP 0.5\n")
    r = _run(p, "--dry-run")
    rows = [l.split("\t") for l in r.stdout.splitlines() if l[:1].isdigit() and "\t" in l]
    assert len(rows) == w.C
    eng = E.Engine.from_workload(w, E.default_options(min_len=120, gap_prop=0.5))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    for c, row in enumerate(rows):
        assert int(row[0]) == c and row[1] == f"elem{c}"
        assert (row[2] == "filtered") == (st[c] == E.PA_ELEM_FILTERED)
        if st[c] == E.PA_ELEM_OK:
            d = eng.element(int(ids[c]))
            assert row[3] == f"GG={d['GG']}" and row[4] == f"P={d['P']}" and row[5] == f"root={d['root']}"


def test_id_file_and_errors(tmp_path):
    w = synth.make_workload(1, seed=22, n_elements=9)
    p, _ = _write_inputs(tmp_path, w, ids=[7, 2, 5])
    r = _run(p, "--dry-run")
    rows = [l.split("\t")[0] for l in r.stdout.splitlines() if l[:1].isdigit() and "\t" in l]
    assert rows == ["7", "2", "5"]
    bad = os.path.join(str(tmp_path), "bad.txt")
    with open(bad, "w") as f:
        f.write("BURNIN 10 20\n")
    r = _run(bad, check=False)
    assert r.returncode == 1 and "not formatted correctly" in r.stderr
    with open(bad, "w") as f:
        f.write(f"RESULT_FOLDER {tmp_path}/missing_dir\n")
    r = _run(bad, check=False)
    assert r.returncode == 1 and "output path" in r.stdout


def _table(path):
    with open(path) as f:
        head = f.readline().rstrip("\n").split("\t")
        rows = [l.rstrip("\n").split("\t") for l in f if l.strip()]
    return head, rows


@pytest.mark.gpu
def test_full_run_outputs_match_oracle(require_gpu, tmp_path):
    w = synth.make_workload(1, seed=23, n_elements=5)
    p, out = _write_inputs(tmp_path, w, extra="BURNIN 20\nMCMC 60\nADAPT_FREQ 25\nPRECISION fp64\n")
    assert _run_in_process(p) == 0
    m = O.model_from_workload(w)
    oo = O.default_options(burnin=20, mcmc=60, adapt_freq=25)
    ref = {c: O.Element(m, oo, w.align, w.starts[c], w.sizes[c]).run(oo) for c in range(w.C)}
    head, rows = _table(os.path.join(out, "run_elem_lik.txt"))
    assert head[:3] == ["No.", "ID", "loglik_Null"]
    for row in rows:
        c = int(row[0])
        r = ref[c]
        vals = [float(x) for x in row[2:]]
        exp = [r[0]["vb"], r[1]["vb"], r[2]["vb"], r[1]["vb"] - r[0]["vb"], r[1]["vb"] - r[2]["vb"],
               r[0]["max_loglik"], r[1]["max_loglik"], r[2]["max_loglik"]]
        np.testing.assert_allclose(vals, exp, rtol=2e-7, atol=1e-6)
    for k in range(3):
        head, rows = _table(os.path.join(out, f"run_M{k}_elem_Z.txt"))
        assert head[1:] == list(w.tree.names)
        for row in rows:
            np.testing.assert_array_equal(np.array(row[1:], int), ref[int(row[0])][k]["max_Z"])
        head, rows = _table(os.path.join(out, f"run_rate_postZ_M{k}.txt"))
        assert len(head) == 6 + 4 * w.tree.N
        for row in rows:
            r = ref[int(row[0])][k]
            np.testing.assert_allclose([float(x) for x in row[1:6]], r["med"], rtol=2e-6)
            np.testing.assert_allclose(np.array(row[6:], float), r["postZ"].reshape(-1), atol=2e-6)
    with open(os.path.join(out, "run_species_names.txt")) as f:
        assert f.read().split() == list(w.tree.names)
