"""The bench command line on the host: the one-GPU guard returns before anything touches a device."""

from __future__ import annotations

import bench


def test_more_than_one_gpu_is_refused_with_a_message(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    err = capsys.readouterr().err
    assert "one GPU only" in err and "--gpus 2" in err


def test_large_workload_is_the_measured_shape():
    assert bench.WORKLOADS["large"] == {"seed": 1, "nao": 1500, "naux": 5000, "nocc": 250}
