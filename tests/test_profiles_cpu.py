"""The committed profile summaries are consistent with the code they describe.

Each summary under profiles/ names the commands that produced it; those
commands must run committed code only (the profiling driver with one of its
workloads), and the HBM traffic file that bench.py quotes as `roofline.traffic`
must hold, for every workload, the dominant kernel the driver profiles, with
bytes = reads + writes.  The split-BF16 entry must be the instantiation the
planner picks for the benchmark's 8192^3 product, so the quoted traffic
belongs to the kernel the benchmark times."""
import json
import os
import re

import pytest

import eigen_amd as E
from eigen_amd import profile_kernels as PK

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILES = os.path.join(ROOT, "profiles")


def _summaries():
    return sorted(f for f in os.listdir(PROFILES) if f.endswith(".md"))


def test_every_summary_names_runnable_commands():
    names = _summaries()
    assert names, "no profile summaries committed"
    for name in names:
        text = open(os.path.join(PROFILES, name)).read()
        assert "rocprofv3" in text, name
        runs = re.findall(r"profile_kernels run (\$?\w+)", text)
        assert runs, name
        for w in runs:
            assert w == "$w" or w in PK.DOMINANT, (name, w)
        # no command reads files outside the repository's committed tree
        assert "_work" not in text, name


def test_traffic_file_entries():
    traffic = json.load(open(os.path.join(PROFILES, "traffic.json")))
    assert "split3" in traffic
    for w, e in traffic.items():
        assert w in PK.DOMINANT, w
        assert PK.DOMINANT[w] in e["kernel"], (w, e["kernel"])
        assert e["bytes"] == e["fetch_bytes"] + e["write_bytes"], w
        assert e["fetch_bytes"] > 0 and e["write_bytes"] > 0, w


def test_split_traffic_is_the_benchmarked_instantiation():
    """The rule's configuration for the 8192^3 split-BF16 product is the
    256x256 tile without split K; the committed traffic entry must be that
    kernel (its template arguments start with the tile shape)."""
    traffic = json.load(open(os.path.join(PROFILES, "traffic.json")))
    prev = E.lib().eh_get_tune_mode()
    E.set_tune_mode(E.TUNE_HEURISTIC)
    try:
        c = E.plan_packed(E.PACKED_SPLIT3, 0, 0, 8192, 8192, 8192)
    finally:
        E.set_tune_mode(prev)
    import ctypes as C
    tm, tn, sk = C.c_int(), C.c_int(), C.c_int()
    E.check(E.lib().eh_packed_config_info(E.PACKED_SPLIT3, c, C.byref(tm), C.byref(tn), C.byref(sk)),
            "eh_packed_config_info")
    assert sk.value == 1
    assert f"gemm_split_packed_kernel<{tm.value}, {tn.value}, " in traffic["split3"]["kernel"]


@pytest.mark.parametrize("name", _summaries())
def test_summary_quotes_its_dominant_kernel(name):
    """A summary's kernel table lists the dominant kernel of its workload."""
    text = open(os.path.join(PROFILES, name)).read()
    runs = [w for w in re.findall(r"profile_kernels run (\w+)", text) if w != "copy"]
    assert runs, name
    assert PK.DOMINANT[runs[0]] in text, name
