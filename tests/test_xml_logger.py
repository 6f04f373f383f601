"""The driver's run log (gample_amd/host/XmlLogger.cpp) against the format of
the reference's XmlLogger (src/XmlLogger.h:8-93), read back the way
tools/parseXML.py reads it: <Cycle ID> elements of named <Section>s, a
time line in the stream's default format, the timeseries CSV row, nothing
written while the switch is off, and fmtSci of a float in %.6e."""
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "gample_amd", "host")

PROG = r"""
#include <cstdio>
#include "XmlLogger.h"
using namespace gample;
int main()
{
    {
        XmlLogger log("run", "cycle,time,phi1,Ex1,Ey1");
        log.logComment("parsed condition");
        log.switchLog(true);
        log.logCycleStart(5, 1.25e-10);
        log.logSection("solvePoisson", {{"iteration", "7"}, {"error", fmtSci(3.2e-9f, 6)}});
        log.logSection("flowout_electron", {{"Xmin", "-12"}, {"particleNumber", "4096"}});
        log.csvOutput("1.0,2.0,3.0");
        log.logCycleEnd();
        log.switchLog(false);
        log.logCycleStart(6, 1.5e-10);
        log.logSection("solvePoisson", {{"iteration", "9"}});
        log.csvOutput("4.0,5.0,6.0");
        log.logCycleEnd();
        log.switchLog(true);
        log.logCycleStart(10, 2.5e-10);
        log.logSection("elapsedTime", {});
        log.logCycleEnd();
    }
    const double v[] = {0.1, 1.0 / 3.0, -2.5e-12, 6.02214076e23, 0.0};
    for (double x : v) std::printf("%s\n", fmtSci((float)x, 6).c_str());
    return 0;
}
"""


@pytest.fixture(scope="module")
def run_log(tmp_path_factory):
    d = tmp_path_factory.mktemp("xml")
    src = d / "t.cpp"
    src.write_text(PROG)
    exe = d / "t"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", HOST, str(src), os.path.join(HOST, "XmlLogger.cpp"),
                           "-o", str(exe)])
    out = subprocess.run([str(exe)], cwd=str(d), capture_output=True, text=True, timeout=30, check=True)
    return d, out.stdout.split()


def test_xml_structure(run_log):
    d, _ = run_log
    text = (d / "run_log.xml").read_text()
    assert text.startswith('<?xml version="1.0" encoding="UTF-8"?>\n<SimulationLog>\n')
    assert text.endswith("</SimulationLog>\n")
    assert "<![CDATA[\nparsed condition]]>" in text
    root = ET.parse(str(d / "run_log.xml")).getroot()
    cycles = root.findall("Cycle")
    assert [c.get("ID") for c in cycles] == ["5", "10"]             # cycle 6 was off
    c5 = cycles[0]
    assert c5.text.strip() == "time = 1.25e-10"
    secs = {s.get("Name"): {e.tag: e.text for e in s} for s in c5.findall("Section")}
    assert secs == {"solvePoisson": {"iteration": "7", "error": "3.200000e-09"},
                    "flowout_electron": {"Xmin": "-12", "particleNumber": "4096"}}
    # tags in key order inside a section, sections in call order
    assert [s.get("Name") for s in c5.findall("Section")] == ["solvePoisson", "flowout_electron"]
    assert [e.tag for e in c5.find("Section")] == ["error", "iteration"]
    assert cycles[1].find("Section").get("Name") == "elapsedTime" and len(cycles[1].find("Section")) == 0


def test_timeseries_csv(run_log):
    d, _ = run_log
    lines = (d / "run_timeseries.csv").read_text().splitlines()
    assert lines == ["cycle,time,phi1,Ex1,Ey1", "5,1.25e-10,1.0,2.0,3.0", "10,2.5e-10"]


def test_fmtsci_rounds_to_float_first(run_log):
    _, printed = run_log
    vals = [0.1, 1.0 / 3.0, -2.5e-12, 6.02214076e23, 0.0]
    assert printed == ["%.6e" % float(np.float32(v)) for v in vals]
    assert printed[1] == "3.333333e-01"
