"""Writer and reader of the namelist &PARAMETERS_MGBETA (mgbeta.nml,
src.mgbf/mg_parameter.f90:190-212) for the Config record."""
from __future__ import annotations

import re

from .mgbf import Config, _DOUBLE_KEYS, _INT_KEYS

_LOGICAL = {"mgbf_line", "l_loc", "l_loc_vertical", "l_filt_g1", "l_lin_vertical",
            "l_lin_horizontal", "l_quad_horizontal", "l_new_map", "l_vertical_filter",
            "ldelta", "lquart", "lhelm"}


def write_namelist(cfg: Config, path: str) -> None:
    lines = ["&PARAMETERS_MGBETA"]
    for k in _DOUBLE_KEYS:
        lines.append(f"  {k}={float(getattr(cfg, k))!r},")
    for k in _INT_KEYS:
        v = getattr(cfg, k)
        lines.append(f"  {k}={'.true.' if v else '.false.'}," if k in _LOGICAL else f"  {k}={int(v)},")
    lines.append("/")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_namelist(path: str) -> Config:
    """Reads key=value pairs of &PARAMETERS_MGBETA ('!' comments allowed)."""
    txt = open(path).read()
    m = re.search(r"&parameters_mgbeta(.*?)/", txt, re.I | re.S)
    if not m:
        raise ValueError("no &PARAMETERS_MGBETA group")
    body = "\n".join(line.split("!")[0] for line in m.group(1).splitlines())
    cfg = Config()
    names = {k.lower(): k for k in _DOUBLE_KEYS + _INT_KEYS}
    for key, val in re.findall(r"([A-Za-z_0-9]+)\s*=\s*([^,\s]+)", body):
        k = names.get(key.lower())
        if k is None:
            raise ValueError(f"unknown namelist key {key}")
        v = val.strip().lower()
        if k in _LOGICAL:
            setattr(cfg, k, v.startswith(".t") or v.startswith("t"))
        elif k in _DOUBLE_KEYS:
            setattr(cfg, k, float(v.replace("d", "e")))
        else:
            if not re.fullmatch(r"[+-]?\d+", v):
                # the reference declares these integer (e.g. p, mg_parameter.f90:110)
                raise ValueError(f"{k}={val} is not an integer")
            setattr(cfg, k, int(v))
    return cfg
