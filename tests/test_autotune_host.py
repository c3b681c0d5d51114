"""The tile autotuner's decision rule and its compiled-in table (no GPU)."""
import os
import re

from rcpgenerator_amd import autotune as A


def _r(**ms):
    return {int(k[1:]): {"kernel_ms": v} for k, v in ms.items()}


def test_default_kept_within_margin():
    assert A.choose(_r(t64=1.0, t96=0.98, t128=1.0)) == 128          # 2 % faster: noise
    assert A.choose(_r(t64=0.80, t96=0.95, t128=1.0)) == 64           # 20 % faster
    assert A.choose(_r(t64=1.2, t128=1.1, t256=1.3)) == 128


def test_header_table_matches_its_recorded_sweep():
    p = os.path.join(os.path.dirname(A.__file__), "csrc", "rcp_tile_table.h")
    text = open(p).read()
    table = {int(a): int(b) for a, b in re.findall(r"\{(\d), (\d+)\},", text)}
    for nd in (2, 3):
        line = re.search(rf"//   {nd}-D: (.*)", text).group(1)
        r = {int(t): {"kernel_ms": float(v)} for t, v in re.findall(r"(\d+): ([\d.]+)", line)}
        assert table[nd] == A.choose(r)
