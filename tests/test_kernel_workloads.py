"""The per-kernel workload driver whose commands the profile summaries give:
argument checks on the CPU, and its small workloads on the GPU."""
import json

import pytest

from chemsim_amd import kernel_workloads as K


def test_unknown_workload_is_an_error():
    with pytest.raises(SystemExit):
        K.main(["bogus"])


@pytest.mark.gpu
def test_small_workloads_run(capsys):
    assert K.main(["cavity128", "--steps", "3"]) == 0
    recs = [json.loads(line) for line in capsys.readouterr().out.splitlines() if line.startswith("{")]
    assert [r["workload"] for r in recs] == ["cavity128/temporal", "cavity128/one_step"]
    assert recs[0]["steps_per_launch"] == 8
    assert all(r["cells"] == 128 * 128 and r["event_us_per_step"] > 0 for r in recs)
