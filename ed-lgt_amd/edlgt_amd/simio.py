"""Configuration and result layout of the batch runner the reference's
scripts use (the third-party package simsio, github.com/rgbmrc/simsio, a git
submodule the reference does not vendor).  Restated from what the reference
holds: the I/O handlers of `.simsiorc` (results/$uid/res and results/$uid/par
written by an NPZ and a YAML serializer, results/$uid/info the log), the
configuration format of `config_example.yaml` (a `===` header holding a
`template` and named blocks, entries that merge a block with `<<<: name`),
the 32-character hexadecimal uids of `sims.sh`, and the calls of
`scripts/*.py` and `analyze_*.py` (`with run_sim() as sim: sim.par / sim.res`,
`gen_configs(template, params, group)`, `SimsQuery(group_glob=...).uids`,
`uids_grid(uids, keys)`, `get_sim(uid).par / .res`).

    with run_sim(uid) as sim:             # uid defaults to sys.argv[1]
        sim.res.update(run_QED_spectrum(sim.par))
    get_sim(uid).res["energy"]
"""
import copy
import glob
import itertools
import logging
import os
import string
import sys
import uuid
from contextlib import contextmanager
from pathlib import Path

import numpy as np
import yaml

__all__ = ["load_group", "resolve", "gen_configs", "find_config", "Sim", "run_sim", "get_sim", "SimsQuery",
           "uids_grid"]

HEADER, MERGE = "===", "<<<"


def _root(path, default):
    return Path(path if path is not None else os.environ.get(f"EDLGT_SIMS_{default.upper()}", default))


def load_group(path):
    """The mapping of one configuration file (uid -> entry, plus the header)."""
    with open(path) as f:
        data = yaml.safe_load(f) or {}
    if not isinstance(data, dict):
        raise ValueError(f"{path}: a configuration file is a mapping")
    return data


def _merge(base, over):
    out = copy.deepcopy(base)
    for k, v in over.items():
        if isinstance(v, dict) and isinstance(out.get(k), dict):
            out[k] = _merge(out[k], v)
        else:
            out[k] = copy.deepcopy(v)
    return out


def resolve(entry, header, _depth=0):
    """Entry with its `<<<` reference replaced by the named header block
    (itself resolved), the entry's own keys taking precedence, recursively
    in nested mappings."""
    if _depth > 32:
        raise ValueError("'<<<' references form a cycle")
    if not isinstance(entry, dict):
        return entry
    out = {}
    if MERGE in entry:
        name = entry[MERGE]
        if name not in header:
            raise KeyError(f"'<<<: {name}' names no block of the header")
        out = resolve(header[name], header, _depth + 1)
    for k, v in entry.items():
        if k == MERGE:
            continue
        v = resolve(v, header, _depth + 1)
        out[k] = _merge(out[k], v) if isinstance(v, dict) and isinstance(out.get(k), dict) else v
    return out


def _yaml_scalar(v):
    v = v.item() if isinstance(v, np.generic) else v
    return yaml.safe_dump(v, default_flow_style=True).strip().removesuffix("...").strip()


def gen_configs(template, params, group, configs=None):
    """Appends one entry per combination of the parameter values (product in
    the order of params) to configs/<group>.yaml, from the header's template
    string: `$enum` is the running number of the entry in the file, `$uid` a
    fresh 32-character hexadecimal uid, `$<name>` a parameter value.  Returns
    the keys of the new entries."""
    path = _root(configs, "configs") / f"{group}.yaml"
    data = load_group(path)
    tmpl = data.get(HEADER, {}).get(template)
    if not isinstance(tmpl, str):
        raise KeyError(f"{path}: the header has no template {template!r}")
    names = list(params)
    start = sum(1 for k in data if k != HEADER)
    chunks, keys = [], []
    for i, combo in enumerate(itertools.product(*(list(np.atleast_1d(params[n])) for n in names))):
        sub = {n: _yaml_scalar(v) for n, v in zip(names, combo)}
        text = string.Template(tmpl).substitute(enum=start + i, uid=uuid.uuid4().hex, **sub)
        entry = yaml.safe_load(text)
        if not isinstance(entry, dict) or len(entry) != 1:
            raise ValueError("a template must define exactly one entry")
        (key,) = entry
        if key in data or key in keys:
            raise ValueError(f"{path}: entry {key!r} exists")
        keys.append(key)
        chunks.append(yaml.safe_dump(entry, sort_keys=False))
    with open(path, "a") as f:
        for c in chunks:
            f.write("\n" + c)
    return keys


def find_config(uid, configs=None):
    """(path of the configuration file holding uid, resolved parameters)."""
    root = _root(configs, "configs")
    for p in sorted(glob.glob(str(root / "**" / "*.yaml"), recursive=True)):
        data = load_group(p)
        if uid in data and uid != HEADER:
            return Path(p), resolve(data[uid], data.get(HEADER, {}))
    raise KeyError(f"uid {uid!r} is in no configuration under {root}")


class Sim:
    """One simulation: parameters, results, and their files
    results/<uid>/res.npz, results/<uid>/par.yaml, results/<uid>/info."""

    def __init__(self, uid, par, results=None):
        self.uid = uid
        self.par = par
        self.res = {}
        self.dir = _root(results, "results") / uid

    def dump(self):
        self.dir.mkdir(parents=True, exist_ok=True)
        arrays = {}
        for k, v in self.res.items():
            try:
                a = np.asarray(v)
            except ValueError:
                a = np.empty(0, dtype=object)
            if a.dtype == object:
                raise TypeError(f"result {k!r} is not an array of numbers or strings")
            arrays[k] = a
        np.savez(self.dir / "res.npz", **arrays)
        with open(self.dir / "par.yaml", "w") as f:
            yaml.safe_dump(_plain(self.par), f, sort_keys=False)


def _plain(x):
    if isinstance(x, dict):
        return {k: _plain(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_plain(v) for v in x]
    if isinstance(x, np.generic):
        return x.item()
    if isinstance(x, np.ndarray):
        return _plain(x.tolist())
    return x


@contextmanager
def run_sim(uid=None, configs=None, results=None):
    """Context of one run: yields a Sim with the resolved parameters of uid
    (default: the first command-line argument); the log of the run goes to
    results/<uid>/info and, on normal exit, the results and parameters are
    written."""
    uid = uid if uid is not None else sys.argv[1]
    _, par = find_config(uid, configs)
    sim = Sim(uid, par, results)
    sim.dir.mkdir(parents=True, exist_ok=True)
    handler = logging.FileHandler(sim.dir / "info")
    handler.setFormatter(logging.Formatter("%(asctime)s %(levelname)-8s | %(message)s", "%Y-%m-%d %H:%M"))
    root = logging.getLogger()
    root.addHandler(handler)
    try:
        yield sim
        sim.dump()
    finally:
        root.removeHandler(handler)
        handler.close()


def get_sim(uid, configs=None, results=None):
    """Sim with the stored results (res.npz, no pickles) and parameters
    (par.yaml, or the configuration when absent)."""
    d = _root(results, "results") / uid
    par_file = d / "par.yaml"
    if par_file.exists():
        with open(par_file) as f:
            par = yaml.safe_load(f)
    else:
        _, par = find_config(uid, configs)
    sim = Sim(uid, par, results)
    with np.load(d / "res.npz", allow_pickle=False) as z:
        sim.res = {k: (z[k][()] if z[k].ndim == 0 else z[k]) for k in z.files}
    return sim


class SimsQuery:
    """Entries of the configuration files matching group_glob
    (configs/<group_glob>.yaml)."""

    def __init__(self, group_glob, configs=None):
        root = _root(configs, "configs")
        self.files = sorted(glob.glob(str(root / f"{group_glob}.yaml"), recursive=True))
        self.uids, self.pars = [], {}
        for p in self.files:
            data = load_group(p)
            header = data.get(HEADER, {})
            for k, v in data.items():
                if k != HEADER:
                    self.uids.append(k)
                    self.pars[k] = resolve(v, header)


def _lookup(par, key):
    if key in par:
        return par[key]
    for v in par.values():
        if isinstance(v, dict):
            try:
                return _lookup(v, key)
            except KeyError:
                pass
    raise KeyError(key)


def uids_grid(uids, keys, configs=None):
    """(ugrid, vals): vals[k] the sorted distinct values of parameter k over
    the uids (found at any depth of their parameters), ugrid an object array
    of shape (len(vals[k0]), len(vals[k1]), ...) holding the uid with those
    values (None where no entry has them)."""
    pars = {}
    for u in uids:
        _, pars[u] = find_config(u, configs)
    vals = {}
    for k in keys:
        vs = [_lookup(pars[u], k) for u in uids]
        hashable = [tuple(v) if isinstance(v, list) else v for v in vs]
        uniq = sorted(set(hashable))
        vals[k] = np.array([list(v) if isinstance(v, tuple) else v for v in uniq], dtype=object
                           if any(isinstance(v, tuple) for v in uniq) else None)
    ugrid = np.full(tuple(len(vals[k]) for k in keys), None, dtype=object)
    for u in uids:
        idx = []
        for k in keys:
            v = _lookup(pars[u], k)
            v = tuple(v) if isinstance(v, list) else v
            idx.append([tuple(x) if isinstance(x, list) else x for x in vals[k]].index(v))
        ugrid[tuple(idx)] = u
    return ugrid, vals
