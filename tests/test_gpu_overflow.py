"""The active-list growth path: a step whose interface outgrows the active capacity is
repeated with larger arrays.  AMCAFE_ACTIVE_CAP_INIT starts every domain with a capacity of
64 active cells, so the first steps of the small deck overflow it several times.  Both the
driver level (temperature fused into the step) and the seam level (the caller's temperature
field, which a repeated step must keep) are checked against the oracle every step."""
import pytest

from test_gpu_parity import _lockstep, seam_run

pytestmark = pytest.mark.gpu


def test_driver_level_across_active_overflow(monkeypatch):
    monkeypatch.setenv("AMCAFE_ACTIVE_CAP_INIT", "64")
    assert _lockstep("t_small.in", 60) == 60


def test_seam_level_across_active_overflow(monkeypatch):
    monkeypatch.setenv("AMCAFE_ACTIVE_CAP_INIT", "64")
    assert seam_run("t_small.in", 60) > 20


def test_seam_level_upload_between_temperature_and_update():
    """A vState upload after amcafe_analytic_temp_curr keeps the device's melt flags in step
    with its temperature field, so the following UpdateVoxels applies SetLiquid3 as the
    reference does."""
    assert seam_run("t_small.in", 120, upload_after_temperature=True) > 20
