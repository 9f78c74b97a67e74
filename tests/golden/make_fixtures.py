"""Extract golden data from the TeNeS test suite into JSON fixtures.

Reads (once, in the build container) the reference's own test inputs and
outputs under /root/reference/test/data and writes small JSON fixtures next to
this script.  Only data is copied: operator elements, lattice/parameter
values and the expected output numbers.  Run:  python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

REF = Path("/root/reference/test/data")
OUT = Path(__file__).resolve().parent


def parse_toml_subset(text: str) -> dict:
    """Parser for the TOML subset the TeNeS test inputs use: tables,
    arrays of tables, ints/floats/bools/strings, (nested) arrays and
    triple-quoted multi-line strings."""
    root: dict = {}
    cur = root
    lines = text.splitlines()
    i = 0

    def value(s: str):
        s = s.strip()
        if s.startswith("'") or s.startswith('"'):
            return s[1:-1]
        if s in ("true", "false"):
            return s == "true"
        if s.startswith("["):
            # multi-line arrays may end rows with a trailing comma
            return json.loads(re.sub(r",\s*\]", "]", s.replace("'", '"')))
        try:
            return int(s)
        except ValueError:
            return float(s)

    while i < len(lines):
        raw = lines[i]
        line = raw.split("#", 1)[0].strip() if '"""' not in raw else raw.strip()
        i += 1
        if not line:
            continue
        if line.startswith("[["):
            keys = line.strip("[]").split(".")
            d = root
            for k in keys[:-1]:
                d = d.setdefault(k, {})
            d.setdefault(keys[-1], []).append({})
            cur = d[keys[-1]][-1]
            continue
        if line.startswith("["):
            keys = line.strip("[]").split(".")
            d = root
            for k in keys:
                d = d.setdefault(k, {})
            cur = d
            continue
        k, v = line.split("=", 1)
        k = k.strip()
        v = v.strip()
        if v.startswith('"""'):
            body = []
            rest = v[3:]
            if rest.endswith('"""'):
                cur[k] = rest[:-3]
                continue
            if rest:
                body.append(rest)
            while '"""' not in lines[i]:
                body.append(lines[i])
                i += 1
            i += 1
            cur[k] = "\n".join(body)
        else:
            while v.count("[") > v.count("]"):
                v += lines[i].split("#", 1)[0].strip()
                i += 1
            cur[k] = value(v)
    return root


def parse_elements(s: str, rank: int):
    out = []
    for line in s.splitlines():
        f = line.split()
        if not f:
            continue
        out.append([int(x) for x in f[:rank]] + [float(f[rank]), float(f[rank + 1])])
    return out


def parse_bonds(s: str):
    return [[int(x) for x in l.split()] for l in s.splitlines() if l.split()]


def read_dat(path: Path):
    rows = []
    for line in path.read_text().splitlines():
        if not line.strip() or line.lstrip().startswith("#"):
            continue
        rows.append(line)
    return rows


def input_fixture(name: str) -> dict:
    t = parse_toml_subset((REF / f"{name}.toml").read_text())
    fx = {"parameter": t.get("parameter", {}), "tensor": t.get("tensor", {})}
    obs = t.get("observable", {})
    fx["onesite"] = [
        {"name": o["name"], "group": o["group"], "sites": o.get("sites", []), "dim": o["dim"],
         "elements": parse_elements(o["elements"], 2), "coeff": o.get("coeff", 1.0)}
        for o in obs.get("onesite", [])]
    fx["twosite"] = []
    for o in obs.get("twosite", []):
        e = {"name": o["name"], "group": o["group"], "bonds": parse_bonds(o["bonds"]),
             "coeff": o.get("coeff", 1.0)}
        if "elements" in o:
            e["dim"] = o["dim"]
            e["elements"] = parse_elements(o["elements"], 4)
        else:
            e["ops"] = o["ops"]
        fx["twosite"].append(e)
    fx["multisite"] = []
    for o in obs.get("multisite", []):
        rows = [[int(x) for x in l.split()] for l in o["multisites"].splitlines() if l.strip()]
        fx["multisite"].append({"name": o["name"], "group": o["group"], "ops": o["ops"],
                                "coeff": o.get("coeff", 1.0),
                                "sites": [{"source": r[0], "dx": r[1::2], "dy": r[2::2]} for r in rows]})
    fx["evolution"] = [
        {"group": o.get("group", 0), "source_site": o["source_site"], "source_leg": o["source_leg"],
         "dimensions": o["dimensions"], "elements": parse_elements(o["elements"], 4)}
        for o in t.get("evolution", {}).get("simple", [])]
    fx["evolution_full"] = [
        {"group": o.get("group", 0), "source_site": o["source_site"], "source_leg": o["source_leg"],
         "dimensions": o["dimensions"], "elements": parse_elements(o["elements"], 4)}
        for o in t.get("evolution", {}).get("full", [])]
    fx["correlation_length"] = t.get("correlation_length", {})
    fx["correlation"] = t.get("correlation", {})
    return fx


def outputs(dirname: str, prefix: str = "") -> dict:
    d = REF / dirname
    res = {}
    for f in ("density", "onesite_obs", "twosite_obs", "correlation_length", "correlation",
              "multisite_obs_4"):
        p = d / f"{prefix}{f}.dat"
        if p.exists():
            res[f] = read_dat(p)
    return res


def main():
    fixtures = {}
    for name, prefix in (("TE_TFI", "TE_"), ("AntiferroHeisenberg_real", ""), ("J1J2_AFH", ""),
                         ("Honeycomb", ""), ("Honeycomb_skew", ""),
                         ("FT_TFI_square", "FT_"), ("FT_Kitaev", "FT_"), ("RSVD", "")):
        fx = input_fixture(name)
        fx["outputs"] = outputs(f"output_{name}", prefix)
        fixtures[name] = fx
    for name, fx in fixtures.items():
        (OUT / f"{name}.json").write_text(json.dumps(fx, indent=1))
    # bond-update known answers (test/simple_update.cpp, test/full_update.cpp)
    for f in ("simple_update.dat", "full_update.dat"):
        vals = [l.split() for l in (REF / f).read_text().splitlines() if l.strip()]
        (OUT / f.replace(".dat", ".json")).write_text(json.dumps(vals))
    print("wrote", sorted(p.name for p in OUT.glob("*.json")))


if __name__ == "__main__":
    sys.exit(main())
