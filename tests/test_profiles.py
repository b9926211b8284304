"""The committed rocprofv3 summaries that bench.py reads: well-formed, for the workload the bench measures, and
each naming the command that made it."""
import glob
import json
import os

import bench as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NNZB_BATCH0 = 159256   # blocks of the batch-0 plan of the contour workload


def test_traffic_profile_covers_every_fp32_shape():
    for shape in B.FP32_SHAPES:
        t = B.traffic_of(shape, NNZB_BATCH0)
        assert t is not None and t > 0, shape
        # at least x read once and y written once (the algorithmic bytes), at most a few times that
        alg = NNZB_BATCH0 * 2 * B.BLOCK_BYTES[32]
        assert alg <= t <= 8 * alg, (shape, t / alg)
    assert B.traffic_of((1, 32, 0), NNZB_BATCH0 + 1) is None   # another workload: no figure


def test_mfma_profile_utilisation_is_consistent():
    m = B.mfma_of(NNZB_BATCH0)
    assert m is not None and len(m) == 4
    for k, e in m.items():
        assert 0 < e["mfma_util"] <= 1, k
        # the flop rate over the dense peak and the busy-cycle fraction are two measures of one thing
        assert abs(e["TFLOPs"] / e["peak_TFLOPs"] - e["mfma_util"]) < 0.05, k
    assert B.mfma_of(NNZB_BATCH0 + 1) is None


def test_every_profile_summary_names_its_command():
    files = sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md")))
    assert files
    for f in files:
        text = open(f).read()
        assert "rocprofv3" in text and "bench.py" in text, f
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.json"))):
        t = json.load(open(f))
        assert any("rocprofv3" in c for c in t["commands"]), f


def test_cpu_baseline_single_and_all_cores():
    """The bench's CPU baseline (the oracle's TFQMR on the host, one thread and several at once) runs on the CPU and
    reports a positive rate for both, the all-cores one with its thread count."""
    import types
    from angstromcube_amd import gfa as G
    from angstromcube_amd import synthetic as S
    sysm = S.small_periodic()
    inst = [{"E": complex(-0.5, 0.05), "k": (0.0, 0.0, 0.0), "radius": 6.0}]
    save = B.SOURCE
    B.SOURCE = (4, 4, 4)   # the centre block of the 8^3-block cell
    try:
        out = B.cpu_baseline(G, sysm, inst, [20], types.SimpleNamespace(cpu_halfsteps=2, cpu_cores=2))
    finally:
        B.SOURCE = save
    assert out["value"] > 0 and out["cores"] == 1 and out["kind"] == "port"
    assert out["all_cores"]["value"] > 0 and out["all_cores"]["cores"] == 2
