"""FINFLO's input file: the Fortran namelists INPUTS and BLOCKS that main.f reads
(NAMELIST declarations main.f:5172-5204, defaults main.f:5240-5360), and the parameters of
this path derived from them as main.f does: the free stream of BCINI1 (state.f:1936-2085,
perfect-gas branch), the turbulence limits (main.f:3772-3826) and the temperature limits
(main.f:3827-3828).

The reader takes Fortran namelist input as the standard defines it and as a Fortran
runtime writes it: `&GROUP` (or `$GROUP`) ... `/` (or `&END`, `$END`), `NAME = value, ...`,
array elements `NAME(i) = ...`, repeat counts `r*value`, null values, quoted strings with
doubled quotes, logicals (`.TRUE.`, `T`, `.f.` ...), D or E exponents and `!` comments.
Names are upper-cased."""
import re

import numpy as np

# The INPUTS defaults of main.f:5240-5360 that this path reads
INPUTS_DEFAULTS = {
    "FULLNC": "NO", "TURCOC": "YES", "STRESC": "NO", "LUSGS": 0, "IDRXX": 6,
    "CFL": 2.0, "CFLL": 2.0, "CDIFF": 2.0, "RMACH": 0.0, "ALPHA": 0.0, "BETA": 0.0,
    "RE": 0.0, "PR": 0.72, "PRT": 0.9, "RGAS": 287.05287, "GAMMA": 1.4, "VISU0": 1.458e-6,
    "EXPSU": 1.5, "TSU0": 110.4, "T0REF": 0.0, "E0REF": 0.0, "FRSTEM": 0.0, "FRSDEN": 0.0,
    "FRSPRE": 0.0, "FRSVEL": 0.0, "RMULTV": 1.0, "FRSTUR": 0.0, "FRSMUT": 0.01,
    "TURLIM": 5000.0, "CMGK": 0.1, "CMGEPS": 0.2, "CSIMPS": 5.0, "CHLREF": 1.0,
    "REFVEL": 0.0, "ITURB": 0, "ISTRES": 0, "FLUXTY": "ROE",
}

_TOKEN = re.compile(r"""
    (?P<str>'(?:[^']|'')*'|"(?:[^"]|"")*")
  | (?P<grp>[&$][A-Za-z_][A-Za-z0-9_]*)
  | (?P<end>/)
  | (?P<name>[A-Za-z_][A-Za-z0-9_%]*\s*(?:\(\s*[-0-9:,\s]+\))?\s*=)
  | (?P<val>[^\s,'"/=!]+)
  | (?P<comma>,)
  | (?P<comment>![^\n]*)
  | (?P<ws>\s+)
""", re.VERBOSE)


def _value(tok):
    """A namelist constant: integer, real, logical or bare word."""
    t = tok.strip()
    u = t.upper()
    if re.fullmatch(r"[-+]?\d+", t):
        return int(t)
    if re.fullmatch(r"[-+]?(\d+\.?\d*|\.\d+)([EDQ][-+]?\d+)?", u) or \
            re.fullmatch(r"[-+]?(\d+\.?\d*|\.\d+)[-+]\d+", u):
        u = re.sub(r"[DQ]", "E", u)
        if re.fullmatch(r"[-+]?(\d+\.?\d*|\.\d+)[-+]\d+", u):
            u = re.sub(r"(\d|\.)([-+]\d+)$", r"\1E\2", u)
        return float(u)
    m = re.fullmatch(r"\.?([TF])[A-Z]*\.?", u)
    if m:
        return m.group(1) == "T"
    return t


def _unquote(s):
    q = s[0]
    return s[1:-1].replace(q + q, q)


def read_namelists(text):
    """All namelist groups of `text` as {GROUP: {NAME: value}}; a name given several values,
    or an array element, holds a list (1-based elements, None where none was given)."""
    groups = {}
    cur, name, vals, index = None, None, None, None

    def flush():
        if cur is None or name is None:
            return
        d = groups[cur]
        if all(v is _NULL for v in vals):
            return                      # a null value leaves the variable as it was
        if index is None and len(vals) == 1:
            d[name] = vals[0]
            return
        start = index if index is not None else 1
        old = d.get(name)
        arr = list(old) if isinstance(old, list) else ([old] if old is not None and index is not None else [])
        need = start - 1 + len(vals)
        arr += [None] * (need - len(arr))
        for k, v in enumerate(vals):
            if v is not _NULL:
                arr[start - 1 + k] = v
        d[name] = arr

    pos = 0
    pending_null = False
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ValueError(f"namelist: cannot read {text[pos:pos + 20]!r}")
        pos = m.end()
        kind = m.lastgroup
        tok = m.group(kind)
        if kind in ("ws", "comment"):
            continue
        if kind == "grp":
            g = tok[1:].upper()
            if g == "END":
                flush()
                cur, name = None, None
                continue
            flush()
            cur, name = g, None
            groups.setdefault(cur, {})
            continue
        if cur is None:
            continue
        if kind == "end":
            flush()
            cur, name = None, None
            continue
        if kind == "name":
            flush()
            raw = tok[:-1].strip()
            mm = re.fullmatch(r"([A-Za-z_][A-Za-z0-9_%]*)\s*(?:\(\s*([-0-9]+)[^)]*\))?", raw)
            name = mm.group(1).upper()
            index = int(mm.group(2)) if mm.group(2) else None
            vals = []
            pending_null = False
            continue
        if name is None:
            raise ValueError(f"namelist: value {tok!r} before any name")
        if kind == "comma":
            if pending_null:
                vals.append(_NULL)
            pending_null = True
            continue
        pending_null = False
        if kind == "str":
            vals.append(_unquote(tok))
            continue
        rep = re.fullmatch(r"(\d+)\*(.*)", tok)
        if rep:
            n, v = int(rep.group(1)), rep.group(2)
            if v == "" and pos < len(text) and text[pos] in "'\"":
                ms = _TOKEN.match(text, pos)
                pos = ms.end()
                vals += [_unquote(ms.group("str"))] * n
            else:
                vals += [_NULL if v == "" else _value(v)] * n
            continue
        vals.append(_value(tok))
    flush()
    return groups


class _Null:
    pass


_NULL = _Null()


def _fmt(v):
    if isinstance(v, bool):
        return ".TRUE." if v else ".FALSE."
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return repr(float(v)).replace("e", "E") if "e" in repr(float(v)) else repr(float(v))
    if isinstance(v, str):
        return "'" + v.replace("'", "''") + "'"
    raise TypeError(type(v))


def write_namelist(group, values):
    """One namelist group as text a Fortran runtime reads: NAME = value, lists as value
    sequences (None entries as nulls)."""
    lines = [f" &{group.upper()}"]
    for k, v in values.items():
        if isinstance(v, (list, tuple, np.ndarray)):
            body = ", ".join("" if x is None else _fmt(x) for x in v)
        else:
            body = _fmt(v)
        lines.append(f"  {k.upper()} = {body},")
    lines.append(" /")
    return "\n".join(lines) + "\n"


def bcini1(inp):
    """The free stream of BCINI1's perfect-gas branch (state.f:2001-2083, ALTITUDE <
    GROUND): from FRSTEM and FRSPRE or FRSDEN, RMACH or FRSVEL (or REFVEL), and two of
    FRSDEN, RE, CHLREF; returns FRSVEL, FRSPRE, FRSDEN, FRSSSP, FRSVIS, RE, RMACH, CHLREF
    and T0.  FRSVIS = RMUFRS, the Sutherland viscosity times RMULTV."""
    g = dict(INPUTS_DEFAULTS)
    g.update(inp)
    frsvel, frspre, frsden, frstem = g["FRSVEL"], g["FRSPRE"], g["FRSDEN"], g["FRSTEM"]
    rmach, re_, chlref, refvel = g["RMACH"], g["RE"], g["CHLREF"], g["REFVEL"]
    rgas, gamma, t0ref = g["RGAS"], g["GAMMA"], g["T0REF"]
    visu0, expsu, tsu0, rmult = g["VISU0"], g["EXPSU"], g["TSU0"], g["RMULTV"]
    gm1 = gamma - 1.
    if frsden == 0. and frspre != 0.:
        frsden = frspre / (rgas * (frstem - t0ref))
    frsssp = np.sqrt(gamma * rgas * frstem)
    rmufrs = visu0 * frstem ** expsu / (frstem + tsu0) * rmult
    if frsvel == 0. and rmach != 0.:
        frsvel = rmach * frsssp
    elif frsvel != 0. and rmach == 0.:
        rmach = frsvel / frsssp
    elif frsvel == 0. and rmach == 0. and refvel != 0.:
        rmach = refvel / frsssp
    else:
        raise ValueError("BCINI1: only one velocity scale may be given (RMACH, FRSVEL)")
    vel = frsvel if frsvel != 0. else refvel
    if frsden == 0. and re_ != 0. and chlref != 0.:
        frsden = re_ * rmufrs / (vel * chlref)
    elif frsden != 0. and re_ == 0. and chlref != 0.:
        re_ = frsden * vel * chlref / rmufrs
    elif frsden != 0. and re_ != 0. and chlref == 0.:
        chlref = re_ * rmufrs / (frsden * vel)
    elif frsden == 0. and re_ == 0. and chlref != 0. and frspre != 0.:
        frsden = frspre / (rgas * (frstem - t0ref))
        re_ = frsden * vel * chlref / rmufrs
    else:
        raise ValueError("BCINI1: the problem is overdefined (FRSDEN, RE, CHLREF)")
    frspre = rgas * (frstem - t0ref) * frsden
    t0 = frstem * (1. + .5 * gm1 * rmach ** 2)
    frsvis = rmufrs
    return dict(FRSVEL=frsvel, FRSPRE=frspre, FRSDEN=frsden, FRSSSP=frsssp, FRSVIS=frsvis,
                RE=re_, RMACH=rmach, CHLREF=chlref, T0=t0)


def params_from_inputs(groups, block=1):
    """ff_params (finflo_amd.layout.Params) of block `block` from the INPUTS and BLOCKS
    namelists: the scalars this path reads, the free stream of BCINI1, the two-equation
    limits of main.f:3772-3826 (FRSTUR is the turbulence intensity, FRSMUT the
    eddy-to-molecular viscosity ratio; IDIS = 2 for the SST), the temperature limits,
    MUSCL switches and the implicit method (LUSGS = 0: DDADI)."""
    from finflo_amd.layout import default_params
    inp = dict(INPUTS_DEFAULTS)
    inp.update(groups.get("INPUTS", {}))
    blk = groups.get("BLOCKS", {})
    fs = bcini1(inp)
    p = default_params(iturb=int(inp["ITURB"]))
    p.gamma, p.rgas, p.pr, p.prt = inp["GAMMA"], inp["RGAS"], inp["PR"], inp["PRT"]
    p.visu0, p.expsu, p.tsu0 = inp["VISU0"], inp["EXPSU"], inp["TSU0"]
    p.frstem = inp["FRSTEM"]
    p.frspre, p.frsden, p.frsvel = fs["FRSPRE"], fs["FRSDEN"], fs["FRSVEL"]
    # the molecular viscosity the solver uses carries RMULTV (FRSVIS = RMUFRS)
    p.rmult = inp["RMULTV"]
    p.frsvis = fs["FRSVIS"]
    p.re = fs["RE"]
    p.cfl = inp["CFL"]
    p.cfll = inp["CFLL"] if inp["CFLL"] > 0. else inp["CFL"]           # main.f:5538
    p.cdiff = 1.0 / inp["CDIFF"]                                        # main.f:4886
    p.tlolim, p.tuplim = 0.05 * p.frstem, 6. * p.frstem                 # main.f:3827-3828
    p.turlim = inp["TURLIM"]
    p.csimps, p.cmgk, p.cmgeps = inp["CSIMPS"], inp["CMGK"], inp["CMGEPS"]
    p.istres = int(inp["ISTRES"])
    p.fullns = 1 if str(inp["FULLNC"]).upper().startswith("Y") else 0
    p.turcor = 1 if str(inp["TURCOC"]).upper().startswith("Y") else 0
    p.ddadi = 1 if int(inp["LUSGS"]) == 0 else 0
    p.iflux = 0 if str(inp["FLUXTY"]).upper().startswith("VAN") else 1
    if p.iturb >= 3 and p.iturb != 9:
        # main.f:3772-3790: the input RKLIM (FRSTUR) is Tu, EPSLIM (FRSMUT) mu_t/mu
        refvel = inp["REFVEL"] if inp["REFVEL"] != 0. else p.frsvel
        rklim, epslim = inp["FRSTUR"], max(inp["FRSMUT"], 1.E-6)
        rkii = 1.5 * p.frsden * (rklim * refvel) ** 2
        frsmut = epslim * p.frsvis
        if rklim < 1.E-15:
            frsmut = 1.0E-3 * p.frsvis
            rkii = frsmut * 1.0 * refvel / fs["CHLREF"]
        repsi = p.frsden * rkii / frsmut                               # IDIS = 2
        p.rklim, p.epslim, p.frsmut = rkii, repsi, frsmut
    for n, key in ((0, "INTERI"), (1, "INTERJ"), (2, "INTERK")):
        v = blk.get(key)
        if isinstance(v, list):
            v = v[block - 1] if len(v) >= block and v[block - 1] is not None else None
        if v is not None:
            p.inter[n] = int(v)
    return p


def mgrid_levels(groups, block=1):
    """MGRID of block `block` from the BLOCKS namelist (1 when not given)."""
    v = groups.get("BLOCKS", {}).get("MGRID")
    if isinstance(v, list):
        v = v[block - 1] if len(v) >= block else None
    return int(v) if v is not None else 1
