"""Errors of the Fortran host come back to the caller as a status and a
message (volcanor_run returns 1, volcanor_last_error), where the reference
stops the process (`error stop`, libCommon.f90:75-77).  These cases fail
while the inputs are read, before any device work, so they run without a
GPU; test_host_gpu.py covers an error after the device context exists."""
import os
import shutil
import tempfile

import pytest

from volcanor_amd import host, vlm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture
def case_copy():
    d = tempfile.mkdtemp(prefix="vlmerr_")
    src = os.path.join(ROOT, "cases", "cfg1_wing.case")
    for f in os.listdir(src):
        if os.path.isfile(os.path.join(src, f)):
            shutil.copy(os.path.join(src, f), d)
    yield d
    shutil.rmtree(d, ignore_errors=True)


def _edit(path, old, new):
    text = open(path).read()
    assert old in text
    with open(path, "w") as fh:
        fh.write(text.replace(old, new))


def test_wrong_namelist_version_returns_an_error(case_copy):
    _edit(os.path.join(case_copy, "config.nml"), "fileFormatVersion = 0.5",
          "fileFormatVersion = 0.4")
    with pytest.raises(vlm.VLMError, match="config.nml template version does not match"):
        host.run_case(case_copy)


def test_missing_case_and_bad_values_return_errors_repeatedly(case_copy):
    with pytest.raises(vlm.VLMError, match="No such file"):
        host.run_case(os.path.join(case_copy, "absent"))
    _edit(os.path.join(case_copy, "geom01.nml"), "nc = 4", "nc = four")
    # twice: the stopped run closed the files it had open
    for _ in range(2):
        with pytest.raises(vlm.VLMError, match=r"geom01.nml &PANELS"):
            host.run_case(case_copy)


def test_unsupported_option_returns_an_error(case_copy):
    _edit(os.path.join(case_copy, "geom01.nml"), "ductSwitch = 0", "ductSwitch = 1")
    with pytest.raises(vlm.VLMError, match="ducts are not supported"):
        host.run_case(case_copy)
