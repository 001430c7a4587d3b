"""The oracle's decision rules, each mutated, must fail a known-answer check.

Every case below plants one plausible slip in a copy of oracle/przm_oracle.c
(a threshold off by one, a cap moved, a comparison made non-strict, two
candidates tried in the other order, a register update dropped), compiles it
with gcc into a temporary directory and runs the known-answer checks of
tests/test_oracle_kats.py, tests/offsets_cases.py and tests/sheaf_cases.py against it.  A mutant that passes every check means
the checks do not pin that rule.  CPU only; the real oracle is not touched.

One written rule cannot change an output and has no mutant here: the run
test's "12 bytes to the tile end" (przm_tokenizer.cuh:222) is implied by its
"run of at least 12" (:245).
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle_lib import Oracle, have_refdec, refdec
from offsets_cases import offsets_failures
from sheaf_cases import sheaf_failures
from tokenizer_kats import KATS, run_kat

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(ROOT, "oracle", "przm_oracle.c")
TIES = json.load(open(os.path.join(HERE, "golden", "choice_ties.json")))

# (name, text in przm_oracle.c, replacement): each text occurs exactly once
MUTANTS = [
    ("lookahead_margin_2", "(desc[mp + 1] >> 16) > mlen + 1", "(desc[mp + 1] >> 16) > mlen + 2"),
    ("lookahead_non_strict", "(desc[mp + 1] >> 16) > mlen + 1", "(desc[mp + 1] >> 16) >= mlen + 1"),
    ("rle_min_13", "if (r >= 12 && r >= mlen)", "if (r >= 13 && r >= mlen)"),
    ("rle_min_11", "if (r >= 12 && r >= mlen)", "if (r >= 11 && r >= mlen)"),
    ("rle_vs_len_strict", "if (r >= 12 && r >= mlen)", "if (r >= 12 && r > mlen)"),
    ("rle_cap_65536", "#define ORC_RLE_CAP 65535u", "#define ORC_RLE_CAP 65536u"),
    ("ext_cap_65531", "#define ORC_EXT_CAP 65530u", "#define ORC_EXT_CAP 65531u"),
    ("ext_cap_65529", "#define ORC_EXT_CAP 65530u", "#define ORC_EXT_CAP 65529u"),
    ("validate_3_bytes", "rd32(src + mp) == rd32(src + mp - d)",
     "(rd32(src + mp) & 0xFFFFFFu) == (rd32(src + mp - d) & 0xFFFFFFu)"),
    ("validate_skipped", "if (!(d > 0 && d <= mp && mlen >= 4 && rd32(src + mp) == rd32(src + mp - d)))",
     "if (!(d > 0 && d <= mp && mlen >= 4))"),
    ("split_short_8", "if (lit_len > 7) { emit_lits(s, src, lit_start, lit_len - 7); lit_start = mp - 7; lit_len = 7; }",
     "if (lit_len > 6) { emit_lits(s, src, lit_start, lit_len - 6); lit_start = mp - 6; lit_len = 6; }"),
    ("split_long_14", "if (lit_len > 15) { emit_lits(s, src, lit_start, lit_len - 15); lit_start = mp - 15; lit_len = 15; }",
     "if (lit_len > 14) { emit_lits(s, src, lit_start, lit_len - 14); lit_start = mp - 14; lit_len = 14; }"),
    ("split_explicit_2", "if (lit_len > 3) { emit_lits(s, src, lit_start, lit_len - 3); lit_start = mp - 3; lit_len = 3; }",
     "if (lit_len > 2) { emit_lits(s, src, lit_start, lit_len - 2); lit_start = mp - 2; lit_len = 2; }"),
    ("short_repeat_up_to_8", "if (mlen <= 7) {", "if (mlen <= 6) {"),
    ("ring_k1_no_swap", "if (k == 1) { rep1 = rep0; rep0 = (int)d; }", "if (k == 1) { rep0 = (int)d; }"),
    ("ring_k2_rotate_wrong", "else if (k == 2) { rep2 = rep1; rep1 = rep0; rep0 = (int)d; }",
     "else if (k == 2) { rep2 = rep0; rep1 = rep1; rep0 = (int)d; }"),
    ("ring_explicit_no_insert", "put_leb(s->off, &s->noff, d);\n            rep2 = rep1; rep1 = rep0; rep0 = (int)d;",
     "put_leb(s->off, &s->noff, d);"),
    ("leb_literal_threshold", "if (count <= 8) {", "if (count <= 7) {"),
    ("choice_order_swapped",
     "    if (n >= 24) {\n        uint32_t l = orc_meta8_encode(sub, n, tmp);\n        if (l < best) { best = l; ch = 1; memcpy(out, tmp, l); }\n    }\n"
     "    if (n >= 16) {\n        uint32_t l = orc_vans4x_adaptive(sub, n, tmp);\n        if (l < best) { best = l; ch = 2; memcpy(out, tmp, l); }\n    }\n",
     "    if (n >= 16) {\n        uint32_t l = orc_vans4x_adaptive(sub, n, tmp);\n        if (l < best) { best = l; ch = 2; memcpy(out, tmp, l); }\n    }\n"
     "    if (n >= 24) {\n        uint32_t l = orc_meta8_encode(sub, n, tmp);\n        if (l < best) { best = l; ch = 1; memcpy(out, tmp, l); }\n    }\n"),
    ("choice_vans_non_strict", "if (l < best) { best = l; ch = 2;", "if (l <= best) { best = l; ch = 2;"),
    ("choice_meta8_non_strict", "if (l < best) { best = l; ch = 1;", "if (l <= best) { best = l; ch = 1;"),
    ("mode8_keep_non_strict", "if (pos < s->ncmd) {", "if (pos <= s->ncmd) {"),
    ("hazard_dropped", "int hazard = !vans && l4 >= 8 &&", "int hazard = 0 && l4 >= 8 &&"),
    # Section 3 as a distance-prefix container
    ("dist_code_half_bit", "const uint32_t half = (u >> b) & 1u;", "const uint32_t half = (u >> (b + 1)) & 1u;"),
    ("m12_context_class_6", "return prev < 6 ? 0 : prev < 14 ? 1 : prev < 26 ? 2 : 3;",
     "return prev <= 6 ? 0 : prev < 14 ? 1 : prev < 26 ? 2 : 3;"),
    ("m12_first_context", "const int ctx = dist_ctx(i ? codes[i - 1] : 0);", "const int ctx = dist_ctx(i ? codes[i - 1] : 26);"),
    ("m12_renorm_early", "while (x > 4096 * f - 1) { stream[ns++] = (uint8_t)x; x >>= 8; }",
     "while (x >= 4096 * f - 1) { stream[ns++] = (uint8_t)x; x >>= 8; }"),
    ("m12_initial_state", "uint32_t ns = 0, x = 65536;", "uint32_t ns = 0, x = 65537;"),
    ("m12_bytes_not_reversed", "out[5 + k] = stream[ns - 1 - k];", "out[5 + k] = stream[k];"),
    ("bit_flush_dropped", "if (cnt) extra[ne++] = (uint8_t)(acc << (8 - cnt));", ""),
    ("vans_codes_from_65", "if (n >= 64) {\n        const uint32_t al", "if (n > 64) {\n        const uint32_t al"),
    ("vans_codes_non_strict", "if (al < cl) {", "if (al <= cl) {"),
    ("byte_header_varints_254", "if (orig <= 255 && n <= 255 && cl <= 255) {", "if (orig < 255 && n <= 255 && cl <= 255) {"),
    ("byte_header_count_254", "if (orig <= 255 && n <= 255 && cl <= 255) {", "if (orig <= 255 && n < 255 && cl <= 255) {"),
    ("compact_header_varints", "} else if (orig <= 65535 && n <= 65535 && cl <= 65535) {",
     "} else if (orig < 65535 && n <= 65535 && cl <= 65535) {"),
    ("prefix_non_strict", "if (pl < raw && info[1] != n) {", "if (pl <= raw && info[1] != n) {"),
    ("code_length_rule_dropped", "if (pl < raw && info[1] != n) {", "if (pl < raw) {"),
    # Section 4 with the sheaf trials
    ("sheaf_syntax_without_hash", 'strchr("{}[]()<>:;,.\\"\'`=+-*/%&|^!~?\\\\#", b)',
     'strchr("{}[]()<>:;,.\\"\'`=+-*/%&|^!~?\\\\", b)'),
    ("sheaf_ident_without_underscore", "static int sheaf_ident(uint8_t b) { return (b >= 'a' && b <= 'z') || b == '_'; }",
     "static int sheaf_ident(uint8_t b) { return (b >= 'a' && b <= 'z'); }"),
    ("sheaf_mode1_without_colon", "case 1: return (p1 == '\"' || p1 == ':')", "case 1: return (p1 == '\"')"),
    ("sheaf_order2_without_dot", "(sheaf_digit(p2) || p2 == '.' || p2 == '-')", "(sheaf_digit(p2) || p2 == '-')"),
    ("sheaf_first_context", "uint32_t l[4] = {0, 0, 0, 0}, start[4], e[4];\n    uint8_t p1 = 0, p2 = 0;",
     "uint32_t l[4] = {0, 0, 0, 0}, start[4], e[4];\n    uint8_t p1 = ' ', p2 = 0;"),
    ("sheaf_flags_without_4x", "out[3] = (uint8_t)(0x01 | 0x04 | ((mode & 0x0F) << 4));",
     "out[3] = (uint8_t)(0x04 | ((mode & 0x0F) << 4));"),
    ("sheaf_fiber_non_strict", "if (cl < l[k]) { memcpy(out + pos, enc + 8, cl); e[k] = cl; }",
     "if (cl <= l[k]) { memcpy(out + pos, enc + 8, cl); e[k] = cl; }"),
    ("sheaf_lengths_swapped", "wr32(out + 8 + 4 * k, l[k]); wr32(out + 24 + 4 * k, e[k]);",
     "wr32(out + 8 + 4 * k, e[k]); wr32(out + 24 + 4 * k, l[k]);"),
    ("literals_vans_non_strict", "if (lv < best) { best = lv; *kind = 1; memcpy(out, tmp, lv); }",
     "if (lv <= best) { best = lv; *kind = 1; memcpy(out, tmp, lv); }"),
    ("literals_sheaf_non_strict", "if (ls < best) { best = ls; *kind = 2 + m; memcpy(out, tmp, ls); }",
     "if (ls <= best) { best = ls; *kind = 2 + m; memcpy(out, tmp, ls); }"),
    ("literals_modes_reversed", "for (int m = 0; m < 4; m++) {\n        const uint32_t ls",
     "for (int m = 3; m >= 0; m--) {\n        const uint32_t ls"),
]


def _sub(key):
    return np.frombuffer(bytes.fromhex(TIES[key]["sub"]), np.uint8).copy()


def decision_failures(O):
    """Names of the known-answer checks O fails."""
    bad = [k["name"] for k in KATS if run_kat(O, k)]
    p, ch = O.pack_substream(_sub("m8_eq_vans"))
    if ch != 1:
        bad.append("m8_eq_vans")
    for key in ("m8_eq_flat", "vans_eq_flat"):
        s = _sub(key)
        p, ch = O.pack_substream(s)
        if ch != 0 or p != b"\x09" + s.tobytes():
            bad.append(key)
    cmd = bytes.fromhex(TIES["mode8_eq_cmd"]["cmd"])
    if O.package_cmds(cmd)[0] != b"\x09" + cmd:
        bad.append("mode8_eq_cmd")
    cmd = bytes.fromhex(TIES["mode8_lt_cmd"]["cmd"])
    if O.package_cmds(cmd)[0][:1] != b"\x08":
        bad.append("mode8_lt_cmd")
    rng = np.random.default_rng(12)
    lit = np.concatenate([np.frombuffer(b"VA\x01", np.uint8), rng.integers(0, 200, 600).astype(np.uint8)])
    a = np.concatenate([lit, np.frombuffer(b"q" * 5000, np.uint8)])
    if O.encode_block(a)[1].kind != 3:
        bad.append("hazard")
    R = refdec() if have_refdec() else None
    return bad + offsets_failures(O, R) + sheaf_failures(O, R)


def test_real_oracle_passes_every_check():
    from oracle_lib import oracle
    assert decision_failures(oracle()) == []


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    return tmp_path_factory.mktemp("oracle_mutants")


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_fails_a_check(mutant_dir, name, old, new):
    text = open(SRC).read()
    assert text.count(old) == 1, f"{name}: the mutated text occurs {text.count(old)} times"
    d = mutant_dir / name
    d.mkdir()
    csrc = d / "przm_oracle.c"
    csrc.write_text(text.replace(old, new))
    so = d / "liboracle.so"
    inc = os.path.join(ROOT, "oracle")  # the oracle includes ../przm_amd/csrc/przm_meta8_freq.h
    r = subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-shared", "-w", "-I" + inc, str(csrc), "-o", str(so)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    bad = decision_failures(Oracle(str(so)))
    assert bad, f"mutant {name} passes every known-answer check"
