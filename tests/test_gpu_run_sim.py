"""The run_sim harness (script/run_sim.py:60-207 restated without ranks) in
its three modes on GHZ-20: result files named and shaped as the reference's
(times, sv_MiB, fidelity), the reference's block-size default, and the
per-stage timer keys of get_total_times."""
import json
import os

import pytest

from qompsim import run_sim as R

pytestmark = pytest.mark.gpu

TIMER_KEYS = {"decomp", "gate", "comp", "fused_swap", "allgather", "hmalloc", "dmalloc", "h2d", "d2h", "dfree"}


def _lines(path):
    with open(path) as f:
        return [json.loads(line) for line in f if line.strip()]


def test_run_sim_fidelity_mode_ghz20(tmp_path, capsys):
    n = 20
    res = R.run_sim("fidelity", "ghz", n, outdir=str(tmp_path))
    B = n - 1 - 4 - 1  # run_sim.py:82-86 with one rank, inner_size 4
    q = _lines(tmp_path / f"qulacs_ghz_{n}qubits_1ranks.json")
    c = _lines(tmp_path / f"qompsim_ghz_{n}qubits_1ranks_bs{B}_inner4.json")
    assert len(q) == 1 and len(c) == 1
    assert set(q[0]) == {"times", "sv_MiB", "fidelity"} == set(c[0])
    assert set(q[0]["times"]) == {"init", "overall", "fused_swap"}
    assert q[0]["times"]["fused_swap"] == 0.0
    assert q[0]["sv_MiB"] == 2 ** (n + 4) / 2 ** 20
    assert q[0]["fidelity"] is None
    assert set(c[0]["times"]) == {"init", "overall"} | TIMER_KEYS
    assert c[0]["times"]["gate"] > 0.0  # the stage kernels
    assert 0.0 < c[0]["sv_MiB"] < q[0]["sv_MiB"]  # GHZ compresses
    assert c[0]["fidelity"] >= 0.9999
    assert res["qompsim"].fidelity == c[0]["fidelity"]
    out = capsys.readouterr().out
    assert "Fidelity:" in out and "[Stage 0/" in out and "Total SV bytes:" in out
    # an existing result file is not overwritten without --append
    assert R.run_sim("qompsim", "ghz", n, outdir=str(tmp_path)) is None
    R.run_sim("qompsim", "ghz", n, outdir=str(tmp_path), append=True)
    assert len(_lines(tmp_path / f"qompsim_ghz_{n}qubits_1ranks_bs{B}_inner4.json")) == 2


@pytest.mark.parametrize("mode", ["qulacs", "qompsim"])
def test_run_sim_cli_modes(mode, tmp_path):
    R.main(["-s", mode, "-c", "qft", "-n", "16", "-o", str(tmp_path), "-b", "10", "-i", "3"])
    files = os.listdir(tmp_path)
    assert len(files) == 1
    rec = _lines(tmp_path / files[0])[0]
    assert rec["times"]["overall"] > 0.0
    if mode == "qompsim":
        assert files[0] == "qompsim_qft_16qubits_1ranks_bs10_inner3.json"
        # a multi-pass-free QFT at B=10: the fused launches count under gate
        assert rec["times"]["gate"] > 0.0
