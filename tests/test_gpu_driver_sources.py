"""The driver's particle sources on the reference's Hall-thruster condition
(tests/golden/conditions/org.json, narrowed to 40 rows and 10 particles per
cell, 30 cycles): x-sinusoidal artificial ionization in 0.25-1.0 cm
(Collision.mm:464-566), the hollow-cathode electron line at 2.4 cm fed by the
current lost at Xmin (:568-678), the cathode potential ramp
(EMField.mm:1063-1077), and a 1D magnetic field read from files in the
format of tools/makeBfld.py.

From the per-cycle XML log, read as tools/parseXML.py reads it: each
species' count moves by exactly its wall losses plus the sources; the ionization count is src / w * dt rounded down
or up; the cathode injects -(sum of the species' charge lost at Xmin) when
that is positive and reports it as neglected otherwise; every solve
converges.  The fused step and the reference-order pieces both keep the
balance."""
import json
import math
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gample_amd", "bin", "gample")
COND = os.path.join(ROOT, "tests", "golden", "conditions", "org.json")
C = 2.99792458e10
EC = 4.803204197e-10
JTOSJ = 1 / (10 / C * 1e4)
NGY, PPC, END = 40, 10, 30


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(EXE):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "gample_amd", "host")])


def write_1d(path, values):
    """tools/makeBfld.py write_1d_field: int32 n, float32 d, n + 1 float32"""
    v = np.asarray(values, np.float32)
    with open(path, "wb") as f:
        f.write(np.int32(len(v) - 1).tobytes())
        f.write(np.float32(0.005).tobytes())
        f.write(v.tobytes())


def condition(tmp, project):
    raw = json.load(open(COND))
    raw["ParamForTimeIntegration"].update(End=END, LogOutput=1, FieldOutput=1000, ProgressOutput=100000,
                                          ProjectName=project)
    raw["ParamForComputing"].update(MaximumParticleNumber=500000, ToleranceForPoisson=1e-8, MaxiterForPoisson=400)
    F = raw["ParamForField"]
    F["NumberOfGridY"] = NGY
    nx = F["NumberOfGridX"]                          # order 1: no ghost cells
    x = np.arange(nx + 2) * F["GridSizeOfX"]
    bz = 0.015 * np.exp(-((x - 1.0) / 0.5) ** 2)    # T
    for k, v in (("Bx", np.zeros_like(x)), ("By", np.zeros_like(x)), ("Bz", bz)):
        write_1d(os.path.join(tmp, k + ".bin"), v)
        F["FilePathOf" + k] = os.path.join(tmp, k + ".bin")
    for p in raw["ParamForParticle"]:
        p["pNumValue"] = PPC
    path = os.path.join(tmp, project + ".json")
    json.dump(raw, open(path, "w"))
    return path, raw


def read_log(path):
    """the run log as tools/parseXML.py walks it: Cycle elements by ID, their
    Sections by Name, each a dict of tag -> text"""
    out = []
    for cyc in ET.parse(path).getroot().findall("Cycle"):
        r = {"cycle": int(cyc.get("ID"))}
        for sec in cyc.findall("Section"):
            r[sec.get("Name")] = {e.tag: e.text for e in sec}
        out.append(r)
    return out


def run(tmp, cfg, extra=()):
    print(f"[driver sources] {cfg} {' '.join(extra)}", flush=True)
    out = subprocess.run([EXE, "-i", cfg, "-o", "t", *extra], cwd=tmp, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]


@pytest.mark.parametrize("mode", ["fused", "unfused"])
def test_hall_thruster_sources(tmp_path, mode):
    tmp = str(tmp_path)
    cfg, raw = condition(tmp, "src")
    run(tmp, cfg, () if mode == "fused" else ("--unfused",))
    rows = read_log(os.path.join(tmp, "src_t_log.xml"))
    assert [r["cycle"] for r in rows] == list(range(1, END + 1))
    F = raw["ParamForField"]
    n0 = F["NumberOfGridX"] * NGY * PPC
    # addn = float(src) / float(w_e) * dt (Collision.mm:482), src from
    # CurrentDensity[A/m2] (Init.mm:240-241), w from InitialNumDens (Init.mm:192-200)
    ai = raw["SourceForParticle"][0]
    src = np.float32(ai["SrcValue"] * JTOSJ * NGY * F["GridSizeOfY"] / EC)
    w = np.float32(raw["ParamForParticle"][0]["WeightValue"] * F["GridSizeOfX"] * F["GridSizeOfY"] / PPC)
    addn_d = float(src / w) * raw["ParamForTimeIntegration"]["TimeStep"]
    ne, ni = n0, n0
    ai_counts, hc_total = [], 0
    for r in rows:
        walls = ("Xmin", "Xmax", "Ymin", "Ymax")
        fe, fi = r["flowout_electron"], r["flowout_ion_Xe1"]
        pe = [int(fe[k]) for k in walls]                # charge units: minus the electrons lost
        pi = [int(fi[k]) for k in walls]
        assert int(fe["pulledPtclNum"]) == -sum(pe) and int(fi["pulledPtclNum"]) == sum(pi)
        a = int(r["artificialIonization"]["addn_injected"])
        h = int(r["hollowCathode"]["addn_injected"])
        neg = int(r["hollowCathode"]["addn_neglected"])
        assert a in (math.floor(addn_d), math.floor(addn_d) + 1)
        net = -(pe[0] + pi[0])
        assert (h, neg) == ((net, 0) if net > 0 else (0, -net))
        # flowout's particleNumber is the count after the reduce, before the sources
        assert int(fe["particleNumber"]) == ne + sum(pe) and int(fi["particleNumber"]) == ni - sum(pi)
        ne += sum(pe) + a + h
        ni += -sum(pi) + a
        sp = r["solvePoisson"]
        assert float(sp["error"]) <= 1e-8 and int(sp["iteration"]) >= 0
        assert math.isfinite(float(sp["meanCathode"])) and math.isfinite(float(sp["totalEEincrement"]))
        assert set(r["memoryUsage"]) == {"physicalFootprint", "residentSetSize"}
        et = r["elapsedTime"]
        assert all(int(v) >= 0 for v in et.values())
        # the phases tools/elapsedTime.py plots (main.mm:176-228): timed on
        # the host in the reference order, by device events under the fused
        # step, where the deposit is part of update_<s>
        for sp_name in ("electron", "ion_Xe1"):
            assert {"integCDens_" + sp_name, "update_" + sp_name, "reduce_" + sp_name} <= set(et)
            assert int(et["update_" + sp_name]) > 0
        assert int(et["solvePoisson"]) > 0
        if mode == "fused":
            assert int(et["integCDens_electron"]) == 0 and int(et["reduce_electron"]) > 0
        ai_counts.append(a)
        hc_total += h
    assert abs(np.mean(ai_counts) - addn_d) < 0.5
    assert hc_total > 0
