"""Input files of the gradient-descent driver.

The reference reads its run parameters from a MATLAB/Octave text file of ``name = value;`` lines
(``src/cuda/infile_grad``) with ``read_check_2`` (``src/cuda/read_check_2.m:1-46``): every listed name
must appear, other lines are reported and skipped.  The reference evaluates each line as MATLAB
code; here the values are parsed as literals only (numbers, ``[a, b, c]`` vectors, quoted strings and
``@name`` function handles), so an input file cannot execute anything.
"""
from __future__ import annotations

import re
from dataclasses import asdict, dataclass

# variables of src/cuda/simple_grad_desc.m:16-29, in the reference's order
GRAD_DESC_VARS = ("alpha", "ds", "ns", "final_time", "gamma", "nmod", "Et_seed_file", "icondfile", "grid_file",
                  "S_file", "mex_function", "in_path", "out_path", "jobid")

_ASSIGN = re.compile(r"^\s*(\w+)\s*=(.*)$")


def float_integer(r: float) -> int:
    """r as an int when it is one to within 5e-16*r, 0 for 0, else ValueError
    (src/cuda/float_integer.m:1-13)."""
    if r == 0:
        return 0
    rn = round(r)
    if abs(r - rn) < 5e-16 * abs(r):
        return int(rn)
    raise ValueError(f"{r!r} is not a float integer")


def _strip_comment(line: str) -> str:
    """Drops a % comment that is not inside a quoted string."""
    quote = None
    for i, ch in enumerate(line):
        if quote:
            if ch == quote:
                quote = None
        elif ch in "'\"":
            quote = ch
        elif ch == "%":
            return line[:i]
    return line


def parse_value(text: str):
    """A MATLAB literal: number -> float, [a, b, c] -> list of floats, 'text' -> str, @name -> 'name'."""
    t = text.strip()
    while t.endswith(";"):
        t = t[:-1].rstrip()
    if len(t) >= 2 and t[0] == t[-1] and t[0] in "'\"":
        return t[1:-1].replace(t[0] * 2, t[0])
    if t.startswith("@"):
        name = t[1:].strip()
        if not re.fullmatch(r"\w+", name):
            raise ValueError(f"not a function handle: {text!r}")
        return name
    if t.startswith("[") and t.endswith("]"):
        body = t[1:-1].replace(",", " ").replace(";", " ").split()
        return [float(x) for x in body]
    return float(t)


def read_infile(path_or_text: str, names=GRAD_DESC_VARS, *, text: bool = False, log=None) -> dict:
    """Values of `names` from a ``name = value;`` file (or the text itself with text=True).  A name
    that is missing raises KeyError listing all missing names; unknown or unparsable lines are
    reported through `log` and skipped, as read_check_2 does."""
    if text:
        lines = path_or_text.splitlines()
    else:
        with open(path_or_text) as f:
            lines = f.read().splitlines()
    wanted = set(names)
    out = {}
    for raw in lines:
        line = _strip_comment(raw).strip()
        if not line:
            continue
        m = _ASSIGN.match(line)
        if not m or m.group(1) not in wanted or m.group(1) in out:
            if log:
                log(f"Expression '{raw.strip()}' not evaluated")
            continue
        out[m.group(1)] = parse_value(m.group(2))
    missing = [n for n in names if n not in out]
    if missing:
        raise KeyError("The following variables were not found: " + " ".join(missing))
    return out


@dataclass
class GradDescentConfig:
    """Run parameters of the gradient-descent driver (src/cuda/simple_grad_desc.m:16-52)."""
    alpha: float            # weight of the field-energy term
    ds: float               # pseudo-time step of the descent
    ns: int                 # descent steps
    final_time: float       # T (ms)
    gamma: list             # weights of the u, v, w terms of the objective
    nmod: int               # checkpoint file every nmod steps
    Et_seed_file: str
    icondfile: str
    grid_file: str
    S_file: str
    mex_function: str
    in_path: str
    out_path: str
    jobid: str
    dt: float = 0.1         # fixed in the reference driver (simple_grad_desc.m:45)

    @classmethod
    def from_values(cls, v: dict) -> "GradDescentConfig":
        return cls(alpha=float(v["alpha"]), ds=float(v["ds"]), ns=int(v["ns"]), final_time=float(v["final_time"]),
                   gamma=[float(x) for x in v["gamma"]], nmod=int(v["nmod"]), Et_seed_file=str(v["Et_seed_file"]),
                   icondfile=str(v["icondfile"]), grid_file=str(v["grid_file"]), S_file=str(v["S_file"]),
                   mex_function=str(v["mex_function"]), in_path=str(v["in_path"]), out_path=str(v["out_path"]),
                   jobid=str(v["jobid"]))

    @classmethod
    def from_infile(cls, path: str, log=None) -> "GradDescentConfig":
        return cls.from_values(read_infile(path, log=log))

    @property
    def nt(self) -> int:
        """Steps of one solve; T/dt must be an integer (float_integer, simple_grad_desc.m:53)."""
        return float_integer(self.final_time / self.dt)

    def as_dict(self) -> dict:
        return asdict(self)
