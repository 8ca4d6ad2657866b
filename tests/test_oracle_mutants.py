"""The closed-form truncation and schedule cases of test_oracle.py must catch
plausible mistakes in the oracle.  Each case below plants one mistake in a
temporary copy of oracle/ffo_oracle.cpp, builds it with g++, and runs the
closed-form test that should catch it in a subprocess against that build: the
test must fail.  The unmodified source passes the same tests (test_oracle.py).
CPU only.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# (name, [(old, new), ...], test in test_oracle.py that must fail)
MUTANTS = [
    # truncate_top_k keeps the k smallest |c| (src/pauli/truncate.h:21-38 keeps the largest)
    ("topk_smallest",
     [("std::greater<double>()", "std::less<double>()"),
      ("return std::abs(c) < cutoff; });", "return std::abs(c) > cutoff; });"),
      ("std::abs(obs.vals[it]) > cutoff) ++it;", "std::abs(obs.vals[it]) < cutoff) ++it;")],
     "test_topk_keeps_exactly_the_k_largest"),
    # no trim of boundary terms: k + 1 terms survive
    ("topk_no_trim", [("while ((int)obs.size() > topk) {", "while (false) {")],
     "test_topk_keeps_exactly_the_k_largest"),
    # top-k only at the last truncation point of the run
    ("topk_last_only", [("if (topk > 0 && (int)obs.size() > topk) {",
                         "if (topk > 0 && (int)obs.size() > topk && rot_count == 1) {")],
     "test_topk_runs_at_every_truncation_point"),
    # commuting batches formed from the forward gate list (propagate.h:272-291 batches the reversed one)
    ("batch_forward_list",
     [("for (const auto& batch : batch_commuting_gates<W>(rot_buffer)) {",
       "std::vector<Rot<W>> fwd(rot_buffer.rbegin(), rot_buffer.rend());\n"
       "        auto bs = batch_commuting_gates<W>(fwd);\n"
       "        std::reverse(bs.begin(), bs.end());\n"
       "        for (auto& bb : bs) std::reverse(bb.begin(), bb.end());\n"
       "        for (const auto& batch : bs) {")],
     "test_batch_schedule_closed_form"),
    # batched=True truncating after every rotation (every batch a single gate)
    ("batch_per_gate", [("        if (ok)\n            current.push_back(g);", "        if (ok && current.empty())\n            current.push_back(g);")],
     "test_batch_schedule_closed_form"),
    # X/Y cap tested with the index of the batch's first rotation (propagate.h:343 uses the last)
    ("xcap_first_index", [("do_trunc(rot_count - 1);", "do_trunc(rot_count - static_cast<int>(batch.size()));")],
     "test_xweight_cap_uses_last_rotation_of_a_batch"),
    # Majorana threshold only after the last gate (majorana_propagation.h:24-31 erases after every gate)
    ("majorana_threshold_at_end",
     [("        if (mincoeff > 0)\n            ret.erase_if([mincoeff](const MKey<W>&",
       "        if (mincoeff > 0 && i == 0)\n            ret.erase_if([mincoeff](const MKey<W>&")],
     "test_majorana_threshold_after_every_gate_closed_form"),
    # maxdegree < 0 read as "no filter" (the reference compares literally)
    ("maxdegree_negative_unlimited", [("const int maxdegree = o.maxdegree;", "const int maxdegree = o.maxdegree < 0 ? 1 << 30 : o.maxdegree;")],
     "test_maxdegree_negative_drops_every_partner"),
]


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    return tmp_path_factory.mktemp("oracle_mutants")


def build_mutant(workdir, name, edits):
    src = open(os.path.join(ORACLE, "ffo_oracle.cpp")).read()
    for old, new in edits:
        assert src.count(old) == 1, (name, old)
        src = src.replace(old, new)
    d = os.path.join(str(workdir), name)
    os.makedirs(d, exist_ok=True)
    shutil.copy(os.path.join(ORACLE, "ffo.h"), d)
    with open(os.path.join(d, "ffo_oracle.cpp"), "w") as f:
        f.write("#include <algorithm>\n" + src)
    so = os.path.join(d, "libffo_oracle.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-o", so,
                    os.path.join(d, "ffo_oracle.cpp")], check=True)
    return so


RUNNER = """
import sys
sys.path[:0] = {paths!r}
import pyffo
pyffo.lib({so!r})
import test_oracle
try:
    getattr(test_oracle, {test!r})()
except AssertionError:
    print("CAUGHT")
    sys.exit(0)
print("MISSED")
sys.exit(1)
"""


@pytest.mark.parametrize("name,edits,test", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(workdir, name, edits, test):
    so = build_mutant(workdir, name, edits)
    paths = [ROOT, ORACLE, os.path.join(ROOT, "tests")]
    r = subprocess.run([sys.executable, "-c", RUNNER.format(paths=paths, so=so, test=test)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "CAUGHT" in r.stdout, (name, r.stdout, r.stderr[-2000:])
