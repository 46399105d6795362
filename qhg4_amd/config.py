"""Species configuration and initial agents in the reference's file formats.

read_species_xml   a species <class> block (QHG4/tutorial_data/xmldat/*.xml):
                   population class, modules with their <param>s (nested
                   modules named "SingleEvaluator[id]"), <priorities>
params_from_xml    the qhg_params of such a block, through
                   qhg_params_from_modules (the reference's attribute rules)
read_agent_dat     an agent data file (QHG4/tutorial_data/xmldat/*.dat):
                   "Longitude;Latitude;LifeState;AgentID;BirthTime;Gender;..."
                   rows, each placed in the grid cell nearest its position
"""
import ctypes as C
import xml.etree.ElementTree as ET

import numpy as np

from . import abi
from ._lib import check, lib


def _module_name(el):
    """Action name as the Prioritizer knows it: name, or name[id]
    (QHG4/actions/Action.cpp:11-19)."""
    n = el.get("name", "")
    i = el.get("id")
    return "%s[%s]" % (n, i) if i else n


def read_species_xml(src):
    """Parse a species <class> block from a path or an XML string.  Returns
    {"class", "species_name", "species_id", "modules": [(name, {param: value})],
    "priorities": [(name, prio)]} with nested modules listed after their
    parent."""
    text = src if src.lstrip().startswith("<") else open(src).read()
    root = ET.fromstring(text)
    if root.tag != "class":
        root = root.find(".//class")
        if root is None:
            raise ValueError("no <class> element")
    mods = []

    def walk(el):
        for m in el.findall("module"):
            params = {p.get("name"): p.get("value") for p in m.findall("param")}
            mods.append((_module_name(m), params))
            walk(m)
    walk(root)
    prios = []
    pr = root.find("priorities")
    if pr is not None:
        for p in pr.findall("prio"):
            prios.append((p.get("name"), int(p.get("value"))))
    return {"class": root.get("name"), "species_name": root.get("species_name"),
            "species_id": root.get("species_id"), "modules": mods, "priorities": prios}


def params_from_species(spec, seed):
    """qhg_params of a parsed species block (raises QhgError with the
    reference-style message when an attribute is missing or unknown)."""
    keep = []

    def cstrs(items):
        arr = (C.c_char_p * max(1, len(items)))(*[s.encode() for s in items])
        keep.append(arr)
        return arr
    mods = (abi.ModuleAttrs * max(1, len(spec["modules"])))()
    for k, (name, params) in enumerate(spec["modules"]):
        keys = list(params.keys())
        mods[k].name = name.encode()
        mods[k].num_params = len(keys)
        mods[k].keys = cstrs(keys)
        mods[k].values = cstrs([params[x] for x in keys])
    names = cstrs([n for n, _ in spec["priorities"]])
    vals = (C.c_int32 * max(1, len(spec["priorities"])))(*[v for _, v in spec["priorities"]])
    out = abi.Params()
    check(lib().qhg_params_from_modules(spec["class"].encode(), mods, len(spec["modules"]), names, vals,
                                        len(spec["priorities"]), int(seed) & 0xFFFFFFFFFFFFFFFF, C.byref(out)),
          "qhg_params_from_modules")
    return out


def params_from_xml(src, seed):
    return params_from_species(read_species_xml(src), seed)


def get_attribute(params, name):
    v = C.c_double()
    check(lib().qhg_params_get_attribute(C.byref(params), name.encode(), C.byref(v)), "qhg_params_get_attribute")
    return v.value


def _unit(lon, lat):
    lo, la = np.radians(lon), np.radians(lat)
    return np.stack([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)], axis=-1)


def read_agent_dat(src, grid):
    """Agents of a .dat file.  The header line names the columns
    (Longitude, Latitude, LifeState, AgentID, BirthTime, Gender, then model
    fields such as Age, LastBirth).  Each agent goes to the cell nearest its
    longitude/latitude (grid["lon"], grid["lat"] in degrees)."""
    lines = open(src).read().splitlines() if "\n" not in src else src.splitlines()
    header = None
    rows = []
    for ln in lines:
        s = ln.strip()
        if not s:
            continue
        if s.startswith("#"):
            header = [h.strip() for h in s[1:].split(";")]
            continue
        rows.append([x.strip() for x in s.split(";")])
    if header is None:
        raise ValueError("agent file without a header line")
    col = {h: i for i, h in enumerate(header)}
    for need in ("Longitude", "Latitude", "LifeState", "AgentID", "BirthTime", "Gender"):
        if need not in col:
            raise ValueError("agent file lacks column %s" % need)
    data = np.array(rows, dtype=object)
    n = len(rows)

    def f(name, dtype, default=None):
        if name in col:
            return np.array(data[:, col[name]], dtype=np.float64).astype(dtype)
        return np.full(n, default, dtype=dtype)
    pos = _unit(f("Longitude", np.float64), f("Latitude", np.float64))
    cells = _unit(np.asarray(grid["lon"]), np.asarray(grid["lat"]))
    cell = np.empty(n, dtype=np.int32)
    for a in range(0, n, 4096):
        cell[a:a + 4096] = np.argmax(pos[a:a + 4096] @ cells.T, axis=1)
    birth = f("BirthTime", np.float64)
    return {
        "id": f("AgentID", np.int64),
        "cell": cell,
        "life": f("LifeState", np.uint8),
        "gender": f("Gender", np.uint8),
        "birth_time": birth,
        "age": f("Age", np.float64, 0.0) if "Age" in col else -birth,
        "last_birth": f("LastBirth", np.float64, -1.0),
        "mate_id": np.full(n, -3, dtype=np.int64),
        "num_babies": np.zeros(n, dtype=np.int32),
    }
