"""Programs compiled by this package's fuser (neslgpu_amd/fusion/fuser.py)
run on the GPU: the .fcode the build wrote for each reference program
(neslgpu_amd/fused/gen/) with the HIP library generated from its super-op
definitions must print what the unfused program printed on the reference
serial CVL (the goldens of tests/golden/vcode)."""
import json
import os

import pytest

from neslgpu_amd import vinterp as V
from neslgpu_amd.build import generated

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "vcode")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))


def stable(text, volatile):
    return [l for i, l in enumerate(text.splitlines()) if i not in volatile]


@pytest.mark.parametrize("name", sorted(k for k in MANIFEST if k.endswith(".vcode")))
def test_fuser_output_runs_on_the_gpu(name):
    base = name[: -len(".vcode")]
    fcode, lib = generated(base, "fcode"), generated(base, "so")
    assert os.path.exists(fcode) and os.path.exists(lib), f"{fcode} not built (python -m neslgpu_amd.build)"
    print("fuser output", base, flush=True)
    rec = MANIFEST[name]
    r = V.run(open(fcode).read(), fused=lib, timeout=240)
    want = open(os.path.join(GOLD, name + ".out")).read()
    assert stable(r.stdout, rec["volatile_stdout_lines"]) == stable(want, rec["volatile_stdout_lines"])
    assert r.stderr == open(os.path.join(GOLD, name + ".err")).read()


def test_fused_quicksort_at_the_benchmark_size():
    """qs.nesl at 2^20 keys in [0, 2^31 - 1), fused by this package's fuser
    at build time (the benchmark's quicksort_fused, printing its keys): the
    sorted keys must hash to what the serial CVL printed for the unfused
    program (the qs.1048576.full digest)."""
    import hashlib
    rec = MANIFEST["qs.1048576.full"]
    fcode, lib = generated("qs_bench_full", "fcode"), generated("qs_bench_full", "so")
    assert os.path.exists(fcode) and os.path.exists(lib), f"{fcode} not built (python -m neslgpu_amd.build)"
    print("fused quicksort 2^20", flush=True)
    r = V.run(open(fcode).read(), fused=lib, timeout=240)
    assert len(r.stdout.splitlines()) == rec["stdout_lines"]
    got = stable(r.stdout, rec["volatile_stdout_lines"])
    assert hashlib.sha256("\n".join(got).encode()).hexdigest() == rec["sha256"]
