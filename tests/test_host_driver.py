"""The C++ driver's condition-file parsing (gample_amd/host/Init.cpp) against
the semantics of src/Init.mm:27-394 on the reference's own condition files
(tests/golden/conditions/, copied from the reference's data/ directory):
merged defaults, pNum and weight derivations, eV->K, and the failure
paths.  Expected values are derived here from the raw JSON by the
reference's formulas, so the test reads like a restatement check."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gample_amd", "bin", "gample")
COND = os.path.join(ROOT, "tests", "golden", "conditions")
EVTOK = 11604.0


def dump(path):
    out = subprocess.run([EXE, "-i", path, "--dump-config"], capture_output=True, text=True, timeout=30)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(EXE):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "gample_amd", "host")])


@pytest.mark.parametrize("name", sorted(os.listdir(COND)))
def test_condition_files(name):
    raw = json.load(open(os.path.join(COND, name)))
    ok_src = {"Source[1/cm/s]", "CurrentDensity[A/m2]", "BoundaryCurrentAtXmin", "BoundaryCurrentAtXmax",
              "BoundaryCurrentAtYmin", "BoundaryCurrentAtYmax"}
    if any(s.get("SrcSetMethod") not in ok_src for s in raw["SourceForParticle"]):
        # the reference rejects such a file too (Init.mm:249-252)
        out = subprocess.run([EXE, "-i", os.path.join(COND, name), "--dump-config"], capture_output=True,
                             text=True, timeout=30)
        assert out.returncode != 0 and "SrcSet failed" in out.stderr
        return
    d = dump(os.path.join(COND, name))
    fld = raw["ParamForField"]
    ngx, ngy = fld["NumberOfGridX"], fld["NumberOfGridY"]
    order = fld.get("WeightingOrder", 5)
    assert d["ngx"] == ngx and d["ngy"] == ngy and d["weightOrder"] == order and d["ngb"] == order // 2
    assert d["BorisOrder"] == fld.get("BorisOrder", 2)
    comp = raw["ParamForComputing"]
    assert d["maxiter"] == comp.get("MaxiterForPoisson", 200)
    assert d["tolerance"] == comp.get("ToleranceForPoisson", 1e-7)
    assert d["pNumMax"] == int(comp.get("MaximumParticleNumber", 0.08))
    assert d["Particle"] == raw["FlagForEquation"]["Particle"]
    for sp, p in zip(d["species"], raw["ParamForParticle"]):
        n = p["pNumValue"] * (ngx * ngy if p["pNumSetMethod"] == "PtclNumPerCell" else 1)
        assert sp["pNum"] == n
        if p["WeightSetMethod"] == "InitialNumDens":
            w = p["WeightValue"] * fld["GridSizeOfX"] * fld["GridSizeOfY"] / (n / (ngx * ngy))
        else:
            w = p["WeightValue"]
        assert sp["w"] == pytest.approx(w, rel=1e-15)
        assert sp["genT"] == pytest.approx(p.get("InitialTemp", 0) * EVTOK, rel=1e-15)
        assert sp["q"] == p["Charge"] and sp["m"] == p["Mass"]
    assert [tuple(b) for b in d["pbc"]] == [(b["RegionName"], b["BCType"]) for b in raw["BoundaryConditionForParticle"]]


def test_missing_section_is_fatal(tmp_path):
    raw = json.load(open(os.path.join(COND, "periodic.json")))
    del raw["ParamForField"]
    f = tmp_path / "bad.json"
    f.write_text(json.dumps(raw))
    out = subprocess.run([EXE, "-i", str(f), "--dump-config"], capture_output=True, text=True, timeout=30)
    assert out.returncode != 0 and "ParamForField" in out.stderr


def test_bad_pnum_method_and_boris(tmp_path):
    raw = json.load(open(os.path.join(COND, "periodic.json")))
    raw["ParamForParticle"][0]["pNumSetMethod"] = "nope"
    f = tmp_path / "bad.json"
    f.write_text(json.dumps(raw))
    out = subprocess.run([EXE, "-i", str(f), "--dump-config"], capture_output=True, text=True, timeout=30)
    assert out.returncode != 0 and "pNumSet failed" in out.stderr
    raw = json.load(open(os.path.join(COND, "periodic.json")))
    raw["ParamForField"]["BorisOrder"] = 3
    f.write_text(json.dumps(raw))
    out = subprocess.run([EXE, "-i", str(f), "--dump-config"], capture_output=True, text=True, timeout=30)
    assert out.returncode != 0


def test_defaults_merge(tmp_path):
    raw = json.load(open(os.path.join(COND, "periodic.json")))
    del raw["ParamForField"]["WeightingOrder"]
    raw["ParamForComputing"] = {}
    f = tmp_path / "d.json"
    f.write_text(json.dumps(raw))
    d = dump(str(f))
    assert d["weightOrder"] == 5 and d["ngb"] == 2 and d["maxiter"] == 200 and d["tolerance"] == 1e-7
    assert d["pNumMax"] == 0          # default @0.08 -> integerValue 0 (Init.mm:326)


def test_poisson_solver_key(tmp_path):
    """the build's PoissonSolver choices; a network mode needs NetworkFile"""
    raw = json.load(open(os.path.join(COND, "periodic.json")))
    f = tmp_path / "s.json"
    for name, ok in (("MG-CG", True), ("NN-FCG", False), ("bogus", False)):
        raw.setdefault("ParamForComputing", {})["PoissonSolver"] = name
        f.write_text(json.dumps(raw))
        out = subprocess.run([EXE, "-i", str(f), "--dump-config"], capture_output=True, text=True, timeout=30)
        assert (out.returncode == 0) == ok, (name, out.stderr)
        if not ok:
            assert "PoissonSolver" in out.stderr


def test_sort_keys(tmp_path):
    """SortInterval and SortDriftFraction (build keys) parse with their
    defaults (20, off) and reject a fraction outside [0, 1]"""
    raw = json.load(open(os.path.join(COND, "periodic.json")))
    raw.setdefault("ParamForComputing", {}).pop("SortDriftFraction", None)
    f = tmp_path / "s.json"
    f.write_text(json.dumps(raw))
    out = subprocess.run([EXE, "-i", str(f), "--dump-config"], capture_output=True, text=True, timeout=30)
    assert out.returncode == 0, out.stderr
    d = json.loads(out.stdout.strip().splitlines()[-1])
    assert d["SortInterval"] == raw["ParamForComputing"].get("SortInterval", 20)
    assert d["SortDriftFraction"] == 0.0
    for frac, ok in ((3e-3, True), (1.5, False), (-1.0, False)):
        raw["ParamForComputing"]["SortDriftFraction"] = frac
        f.write_text(json.dumps(raw))
        out = subprocess.run([EXE, "-i", str(f), "--dump-config"], capture_output=True, text=True, timeout=30)
        assert (out.returncode == 0) == ok, (frac, out.stderr)
        if ok:
            assert json.loads(out.stdout.strip().splitlines()[-1])["SortDriftFraction"] == frac
