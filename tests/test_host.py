"""Host side of the product library (no GPU needed): the C ABI loads and
exports every declared symbol, fails loudly without a device, and its host
helpers (trace reader, bit-parallel evaluator, generators) agree with the
oracle."""
import ctypes
import glob
import os
import random
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(ltl_[a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(ltl):
    L = ltl.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(ltl.EXPORTS) <= set(names)


def test_version(ltl):
    assert b"gfx950" in ltl.lib().ltl_version()


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK),
                    reason="a GPU is visible")
def test_no_device_fails_loudly(ltl):
    with pytest.raises(ltl.LtlError):
        ltl.Learner(budget=1 << 20)


def test_bad_options_rejected(ltl):
    o = ltl.make_options(neg_type=3)
    st = ctypes.c_int32(0)
    assert not ltl.lib().ltl_create(ctypes.byref(o), ctypes.byref(st))
    assert st.value == ltl.E_ARG
    o = ltl.make_options(cost=(0, 1, 1, 1, 1, 1, 1, 1))
    assert not ltl.lib().ltl_create(ctypes.byref(o), ctypes.byref(st))
    assert st.value == ltl.E_ARG


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "traces", "*.trace"))),
                         ids=os.path.basename)
def test_trace_reader_matches_oracle(ltl, oracle, path):
    a = ltl.read_trace_file(path)
    b = oracle.read_trace_file(path)
    assert (a.num_pos, a.num_neg, a.names) == (b.num_pos, b.num_neg, b.names)
    assert np.array_equal(a.lens, b.lens)
    assert np.array_equal(a.atoms, b.atoms)


def random_formula(rng, size, names):
    if size <= 1:
        return rng.choice(names)
    if size == 2 or rng.random() < 0.4:
        return f"{rng.choice('~XFG')}({random_formula(rng, size - 1, names)})"
    k = rng.randint(1, size - 2)
    return f"({random_formula(rng, k, names)}){rng.choice('&|U')}({random_formula(rng, size - 1 - k, names)})"


def test_product_evaluator_matches_textbook(ltl, oracle):
    rng = random.Random(4)
    for trial in range(30):
        W = rng.randint(1, 70)
        A = rng.randint(1, 4)
        lens = np.array([rng.randint(1, 64) for _ in range(W)], dtype=np.uint8)
        atoms = np.array([[rng.getrandbits(int(n)) for n in lens] for _ in range(A)], dtype=np.uint64)
        P = rng.randint(0, W)
        pi = ltl.Instance(P, W - P, lens, atoms)
        oi = oracle.Instance(P, W - P, lens, atoms)
        for _ in range(10):
            f = random_formula(rng, rng.randint(1, 12), pi.names)
            assert np.array_equal(ltl.eval_formula(pi, f), oracle.eval_formula(oi, f)), f
            cs = ltl.eval_formula(pi, f)
            assert ltl.mismatches(pi, cs) == oracle.mismatches(oi, cs)


def test_evaluator_rejects_bad_syntax(ltl):
    inst = ltl.Instance(1, 1, [3, 3], [[1, 2]])
    for bad in ("", "(a)", "a&b", "~a", "(a)&(b", "X(c)", "(a)V(a)"):
        with pytest.raises(ltl.LtlError):
            ltl.eval_formula(inst, bad)


@pytest.mark.parametrize("config,P,N,A,lmin,lmax", [
    (1, 10, 10, 3, 4, 8), (2, 500, 500, 5, 10, 32), (3, 1000, 1000, 8, 64, 64), (5, 1000, 1000, 6, 16, 64)])
def test_generator_shapes_and_labels(ltl, oracle, config, P, N, A, lmin, lmax):
    inst = ltl.generate(config, seed=1)
    assert (inst.num_pos, inst.num_neg, inst.atoms.shape[0]) == (P, N, A)
    assert lmin <= int(inst.lens.min()) and int(inst.lens.max()) <= lmax
    oi = oracle.Instance(inst.num_pos, inst.num_neg, inst.lens, inst.atoms, inst.names)
    cs = oracle.eval_formula(oi, inst.planted)
    # labels come from the planted formula; config 3 swaps 50 + 50 labels
    assert oracle.mismatches(oi, cs) == (100 if config == 3 else 0)
    for a in range(A):  # bits above each trace's length are zero
        assert not np.any(inst.atoms[a] >> inst.lens.astype(np.uint64))
    again = ltl.generate(config, seed=1)
    assert np.array_equal(again.atoms, inst.atoms)  # seeded


def test_generator_config4_batch(ltl, oracle):
    sizes = set()
    for i in range(24):
        inst = ltl.generate(4, seed=3, index=i)
        W = inst.W
        assert 20 <= W <= 400 and 2 <= inst.atoms.shape[0] <= 6
        assert 0.25 <= inst.num_pos / W <= 0.75
        assert 5 <= int(inst.lens.min()) and int(inst.lens.max()) <= 40
        assert 5 <= inst.max_cost <= 9
        oi = oracle.Instance(inst.num_pos, inst.num_neg, inst.lens, inst.atoms, inst.names)
        assert oracle.mismatches(oi, oracle.eval_formula(oi, inst.planted)) == 0
        assert oracle.formula_cost(inst.planted) == inst.max_cost
        sizes.add(W)
    assert len(sizes) > 10


def test_token_encoding_matches_reader(ltl, oracle):
    """LTLI()'s token traces encode like the reference (code/ltli6463.cu:1100-1133)."""
    path = os.path.join(GOLDEN, "traces", "E0.trace")
    ref = oracle.read_trace_file(path)
    lines = open(path).read().split("\n")
    alpha = [c for c in lines[-1] if "a" <= c <= "z"] if lines[-1] else [c for c in lines[-2] if "a" <= c <= "z"]

    def tokens(line):
        return ["".join(alpha[j] for j, v in enumerate(tok.split(",")) if v == "1") for tok in line.split(";")]
    i = lines.index("---")
    j = lines.index("---", i + 1)
    pos = [tokens(l) for l in lines[:i]]
    neg = [tokens(l) for l in lines[i + 1:j]]
    inst = ltl._instance_from_tokens(pos, neg, alpha)
    assert inst.names == ref.names
    assert np.array_equal(inst.atoms, ref.atoms) and np.array_equal(inst.lens, ref.lens)


@pytest.mark.parametrize("variant", ["crlf", "trailing_blank_lines", "no_final_newline", "trailing_spaces"])
def test_trace_reader_tolerates_editor_artifacts(ltl, tmp_path, variant):
    """The same traces with CR-LF line ends, blank lines after the alphabet,
    no final newline or blanks at the end of lines parse identically."""
    src = os.path.join(GOLDEN, "traces", "Ab1.trace")
    text = open(src).read()
    lines = text.rstrip("\n").split("\n")
    if variant == "crlf":
        out = "\r\n".join(lines) + "\r\n"
    elif variant == "trailing_blank_lines":
        out = "\n".join(lines) + "\n\n\n"
    elif variant == "no_final_newline":
        out = "\n".join(lines)
    else:
        out = "\n".join(l + "  " for l in lines) + "\n"
    p = tmp_path / "t.trace"
    p.write_bytes(out.encode())
    a, b = ltl.read_trace_file(src), ltl.read_trace_file(str(p))
    assert (a.num_pos, a.num_neg, a.names) == (b.num_pos, b.num_neg, b.names)
    assert np.array_equal(a.lens, b.lens)
    assert np.array_equal(a.atoms, b.atoms)


def write_named_traces(path, n_props=30, n=24, seed=3):
    """A trace file whose alphabet line names 30 propositions p0..p29 (more
    than the reference's 26 letters); labels by F(p27 & X p3), evaluated
    position by position here."""
    rng = np.random.default_rng(seed)
    pos, neg = [], []
    while len(pos) < n // 2 or len(neg) < n // 2:
        L = int(rng.integers(3, 9))
        bits = rng.integers(0, 2, size=(L, n_props))
        holds = any(bits[k, 27] and bits[k + 1, 3] for k in range(L - 1))
        line = ";".join(",".join(str(int(b)) for b in row) for row in bits)
        side = pos if holds else neg
        if len(side) < n // 2:
            side.append(line)
    names = [f"p{i}" for i in range(n_props)]
    with open(path, "w") as f:
        f.write("\n".join(pos) + "\n---\n" + "\n".join(neg) + "\n---\nF,G,X,!,&,|\n---\n" + ",".join(names))
    return pos, neg, names


def test_reader_more_than_26_propositions(ltl, tmp_path):
    path = str(tmp_path / "named.trace")
    pos, neg, names = write_named_traces(path)
    inst = ltl.read_trace_file(path)
    assert (inst.num_pos, inst.num_neg) == (len(pos), len(neg))
    assert inst.names == ",".join(names) and inst.atoms.shape[0] == 30
    for t, line in enumerate(pos + neg):
        rows = [list(map(int, tok.split(","))) for tok in line.split(";")]
        assert inst.lens[t] == len(rows)
        for a in range(30):
            assert int(inst.atoms[a, t]) == sum(r[a] << k for k, r in enumerate(rows))
    # the planted formula, spelled with the names, separates the traces
    cs = ltl.eval_formula(inst, "F((p27)&(X(p3)))")
    assert ltl.mismatches(inst, cs) == 0
    # a name that is not in the list, or malformed, is a parse error
    for bad in ("F((p30)&(X(p3)))", "F((p27)&(X(p)))", "(p2)&"):
        with pytest.raises(ltl.LtlError):
            ltl.eval_formula(inst, bad)


def test_letters_only_files_read_as_before(ltl):
    # reference files keep the letter rule: sorted letters, plain names
    inst = ltl.read_trace_file(os.path.join(GOLDEN, "traces", "Ab0.trace"))
    assert inst.names == "pq"
