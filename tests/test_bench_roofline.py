"""The bench's roofline arithmetic on the host: algorithmic bytes per launch of each front end's
dominant kernel at n = 2^27, k = 16000 (closed forms of the sizes), the 64-byte and 128-byte
gather pricing, and the extraction of a kernel's source text that ties the counter traffic in
profiles/traffic.json to the code it was measured on."""
import importlib.util
import json
import os

import pytest

import sfft_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def prm():
    return S.derive_params(27, 16000, 0.65, 2, 2, 7, 4096, 2)


def test_headline_sizes(prm):
    # the reference README's run (README.md:76-92): B 131072, W 4194304, w 2230607, 9 loops
    assert (prm["n"], prm["B"], prm["W_RLFC"], prm["w"], prm["loops"]) == (
        1 << 27, 131072, 4194304, 2230607, 9)


@pytest.mark.parametrize("prec,e", [("fp64", 16), ("fp32", 8)])
def test_column_fold_bytes(bench, prm, prec, e):
    # one launch: the signal read once, the filter once, the bins and the comb samples
    # written once
    r = bench.kernel_roofline(prm, prec, "columns", {"stream": 0.5}, "n27_k16000")
    n, w, L, B, W, Rl = (1 << 27), 2230607, 9, 131072, 1 << 22, 2
    assert r["bytes_per_launch"] == n * e + w * e + L * B * e + Rl * W * e
    assert r["kernel"] == "fold_slab_kernel"
    assert r["achieved"] == pytest.approx(r["bytes_per_launch"] / 0.5e-3 / 1e9, rel=1e-3)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0, rel=1e-3)
    if prec == "fp64":
        assert r["bytes_per_launch"] == 2336265456


def test_gather_fold_granules(bench, prm):
    r = bench.kernel_roofline(prm, "fp64", "gather", {"fold": 0.4}, "n27_k16000")
    L, w, B = 9, 2230607, 131072
    assert r["bytes_per_launch"] == L * w * 64 + w * 16 + L * B * 16
    g128 = L * w * 128 + w * 16 + L * B * 16
    assert r["granule_bytes"] == 64
    assert r["achieved_granule128"] == pytest.approx(g128 / 0.4e-3 / 1e9, rel=1e-3)


def test_kernel_source_text_bounds(bench):
    """The digest covers each listed function from its template line to its closing brace."""
    path = os.path.join(ROOT, "gpu-sfft_amd", "csrc", "sfft_fft.hip")
    lines = open(path).read().split("\n")
    for kname, fns in bench.KERNEL_SOURCE.items():
        d = bench.kernel_source_digest(kname)
        assert len(d) == 16 and int(d, 16) >= 0
        for fn in fns:
            at = next(i for i, ln in enumerate(lines) if f" {fn}(" in ln and "void" in ln)
            assert lines[at - 1].startswith("template") or lines[at].startswith(
                ("template", "__global__", "__device__")), fn
    # different kernels, different texts
    assert len({bench.kernel_source_digest(k) for k in bench.KERNEL_SOURCE}) == len(
        bench.KERNEL_SOURCE)


def test_traffic_entries_name_known_kernels(bench):
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as fh:
        t = json.load(fh)
    for key, entry in t.items():
        if key.startswith("_"):
            continue
        kname = key.split("_n27_")[0]
        assert kname in bench.KERNEL_SOURCE, key
        assert entry["bytes"] > 0 and os.path.exists(os.path.join(ROOT, entry["source"]))
        assert len(entry["kernel_source_sha256"]) == 16
