"""The upstream binding of the MI355X backend compiles against the reference's
class declarations: every `override` of the backend classes names a virtual
of the reference class it derives from, no destructor is marked `override`
where the reference's is not virtual, and the backend reads no Config member
the reference lacks.  The reference's names come from
tests/golden/reference_interface.json (written by make_reference_interface.py
from src/grid.h, src/hardware.h, src/data_Acquisition.h and src/config.h)."""
import json
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "hpcscan_amd", "host")
REF = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_interface.json")))

BACKEND = {   # file -> base class of the backend class it declares
    "grid_Mi355x.h": "Grid",
    "hardware_Mi355x.h": "Hardware",
    "data_Acquisition_Mi355x.cpp": "DataAcquisition",
}
SOURCES = ["grid_Mi355x.h", "grid_Mi355x.cpp", "hardware_Mi355x.h", "hardware_Mi355x.cpp",
           "data_Acquisition_Mi355x.cpp"]


def _code(name):
    s = open(os.path.join(HOST, name)).read()
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    return re.sub(r"//[^\n]*", "", s)


def _overrides(code):
    """names of the member functions declared with `override`"""
    out = []
    for m in re.finditer(r"(~?\w+)\s*\(([^()]|\([^()]*\))*\)\s*(const\s*)?override\b", code):
        out.append(m.group(1))
    return out


@pytest.mark.parametrize("fname", sorted(BACKEND))
def test_overrides_name_reference_virtuals(fname):
    base = BACKEND[fname]
    names = _overrides(_code(fname))
    assert names, f"no override found in {fname}"
    allowed = set(REF[base]["virtuals"])
    bad = [n for n in names if n not in allowed]
    assert not bad, f"{fname}: override of {bad}, not virtual in the reference's {base}"
    if not REF[base]["destructor_virtual"]:
        assert not any(n.startswith("~") for n in names), f"{fname}: destructor marked override"


def test_non_virtual_reference_members_are_not_overridden():
    """appliSourceTerm is not virtual in the reference: the backend declares it
    without override and checks the dispatch at initialize()"""
    code = _code("data_Acquisition_Mi355x.cpp")
    assert "appliSourceTerm" not in REF["DataAcquisition"]["virtuals"]
    assert re.search(r"Rtn_code\s+appliSourceTerm\s*\([^)]*\)\s*\{", code)
    assert "PROBE_STEP" in code and "probed" in code


def test_backend_reads_only_reference_config_members():
    members = set(REF["Config"]["members"])
    used = set()
    for name in SOURCES:
        code = _code(name)
        used |= set(re.findall(r"Config::Instance\(\)\s*->\s*(\w+)", code))
        for var in re.findall(r"Config\s*\*\s*(\w+)\s*=\s*Config::Instance\(\)", code):
            used |= set(re.findall(r"\b" + var + r"\s*->\s*(\w+)", code))
    calls = {"Instance"}
    extra = sorted(u for u in used if u not in members and u not in calls)
    assert not extra, f"the backend reads Config members the reference lacks: {extra}"


def test_factories_delete_backend_objects_as_their_own_type():
    """~Grid / ~Hardware / ~DataAcquisition are not virtual in the reference:
    the backend's creators make their shared pointers from the derived type"""
    code = _code("grid_Mi355x.cpp") + _code("data_Acquisition_Mi355x.cpp")
    assert "make_shared<Grid_Mi355x>" in code
    assert "make_shared<Hardware_Mi355x>" in code
    assert "make_shared<DataAcquisition_Mi355x>" in code
    assert not re.search(r"shared_ptr<\w+>\s*\(\s*new\b", code)
