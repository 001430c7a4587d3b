"""The pinned parts of the CPU restatement of the reference compressor (oracle/przm_refcpu.c),
each mutated, must fail a check of tests/test_refcpu.py.

Each case plants one plausible slip in a copy of przm_refcpu.c (a constant of the cost model, a
sublength step, a stride candidate dropped, a hash word offset, a dictionary transform id, the
upper-case probe switched off), builds it with the oracle into a temporary library and runs the
known-answer checks.  The trellis's choices among valid parses have no such check: their frames
decode whatever they are, and no Go output exists to compare with (parity unpinned).  One
written rule cannot change an output and has no mutant: the distance-code cost floor of 6
(compress.go:255), since no code of the static tables costs less.
"""
import ctypes
import os
import shutil
import subprocess

import pytest

from oracle_lib import RefCPU, have_reftables

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(ROOT, "oracle", "przm_refcpu.c")
ORC = os.path.join(ROOT, "oracle", "przm_oracle.c")

MUTANTS = [
    ("sublength_step_13", "steps[4] = {32, 24, 16, 12}", "steps[4] = {32, 24, 16, 13}"),
    ("sublength_boundary_15", "const int cb = 14 + bmin;", "const int cb = 15 + bmin;"),
    ("sublength_delta_3", "for (int delta = 1; delta <= 4; delta++)", "for (int delta = 1; delta <= 3; delta++)"),
    ("lit_cost_floor_2", "return cost < 3 ? 3 : cost;", "return cost < 2 ? 2 : cost;"),
    ("lit_cost_slope", "(uint32_t)(bits * 2.6 + 1.2)", "(uint32_t)(bits * 2.5 + 1.2)"),
    ("lit_cost_absent_35", "if (f == 0) return 36;", "if (f == 0) return 35;"),
    ("dist_cost_rounding", "(uint32_t)(bits * 2.0 + 0.5)", "(uint32_t)(bits * 2.0)"),
    ("repeat_bonus", "const uint32_t bonus = (uint32_t)(4 - k);", "const uint32_t bonus = (uint32_t)(3 - k);"),
    ("repeat_short_nlit_8", "if (L <= 3 + bmin && nlit <= 7) base = 14;", "if (L <= 3 + bmin && nlit <= 8) base = 14;"),
    ("explicit_far_512", "if (L == bmin) { if (dist > 512) dc += 8; }", "if (L == bmin) { if (dist > 1024) dc += 8; }"),
    ("explicit_long_penalty", "base += L <= 14 + bmin + 127 ? 3 : 7;", "base += L <= 14 + bmin + 127 ? 4 : 7;"),
    ("explicit_extra_bits", "rc_dist_cost[code] + 2u * nb", "rc_dist_cost[code] + nb"),
    ("stride_without_7", "{2, 3, 4, 5, 6, 7, 8, 12,", "{2, 3, 4, 5, 6, 8, 12,"),
    ("stride_min_len_256", "if (n < 128) return 1;", "if (n < 256) return 1;"),
    ("hash12_offset", "rc_rd64(s + p + 4) * RC_C1", "rc_rd64(s + p + 8) * RC_C1"),
    ("hash24_constant", "(rc_rd64(s + p + 16) * RC_C2);\n    return", "(rc_rd64(s + p + 16) * RC_C1);\n    return"),
    ("dict_suffix_space_tau", "case ' ': fn(ctx, wid, 1, l + 1); break;", "case ' ': fn(ctx, wid, 2, l + 1); break;"),
    ("dict_upper_off", "if (s[0] >= 'A' && s[0] <= 'Z') {", "if (0) {"),
    ("dict_dot_prefix_tau", "fn(ctx, wid, 32, 1 + l);", "fn(ctx, wid, 33, 1 + l);"),
    ("dict_capitalize_all", "if (op == 10) break; /* capitalize: the first byte only */", ""),
]


def refcpu_failures(R):
    """names of the checks of tests/test_refcpu.py that R fails"""
    import test_refcpu as T
    import oracle_lib
    bad = []
    saved = oracle_lib._refcpu
    oracle_lib._refcpu = R
    try:
        checks = [("sublengths", T.test_pareto_sublengths_by_hand), ("lit_cost", T.test_literal_costs_by_hand),
                  ("dist_cost", T.test_distance_code_costs), ("costs", T.test_repeat_and_explicit_costs_by_hand),
                  ("stride", T.test_stride_detection), ("hashes", T.test_hash_tiers)]
        if have_reftables():
            checks.append(("dict_size", T.test_dictionary_index_size))
            checks.append(("dict_transforms", T.test_dictionary_transformed_matches))
            checks.append(("dict_compact", T.test_dictionary_compact_pairs_indexed))
            for p in (b"return value from function", b"constant variable declaration"):
                checks.append((f"dict_{p[:6]}", lambda p=p: T.test_dictionary_matches_are_the_words(p)))
        for name, fn in checks:
            try:
                fn()
            except AssertionError:
                bad.append(name)
    finally:
        oracle_lib._refcpu = saved
    return bad


def test_real_restatement_passes_every_check():
    from oracle_lib import refcpu
    assert refcpu_failures(refcpu()) == []


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    return tmp_path_factory.mktemp("refcpu_mutants")


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_fails_a_check(mutant_dir, name, old, new):
    if name.startswith("dict_") and not have_reftables():
        pytest.skip("reference dictionary tables not built")
    text = open(SRC).read()
    assert text.count(old) == 1, f"{name}: the mutated text occurs {text.count(old)} times"
    d = mutant_dir / name
    d.mkdir()
    csrc = d / "przm_refcpu.c"
    csrc.write_text(text.replace(old, new))
    so = d / "liboracle.so"
    inc = os.path.join(ROOT, "oracle")
    r = subprocess.run(["gcc", "-O1", "-std=gnu11", "-fPIC", "-shared", "-w", "-I" + inc, ORC, str(csrc), "-o", str(so),
                        "-lm", "-ldl", "-pthread"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    bad = refcpu_failures(RefCPU(ctypes.CDLL(str(so))))
    assert bad, f"mutant {name} passes every known-answer check"
