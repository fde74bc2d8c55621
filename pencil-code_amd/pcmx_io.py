"""Host-side I/O in the reference's on-disk formats, next to the engine.

* Fortran unformatted sequential records as the gfortran runtime writes them
  (4-byte byte counts around each record; records over 2147483639 bytes split
  into subrecords, the leading count negated when another subrecord follows,
  the trailing count negated when one precedes).  ``read_records`` parses
  them; ``pcmx.write_record`` / ``pcmx.write_snapshot`` write them.
* ``read_var``: a var.dat snapshot of one process (output_snap,
  src/io_dist.f90:108-199): record 1 f(mx,my,mz,1:mvar), record 2 t, x, y, z,
  dx, dy, dz.
* ``TimeSeries``: time_series.dat lines as the reference's ``prints`` writes
  them (src/diagnostics.f90:314-461): a legend line of dash-padded names
  (``noform``, src/sub.f90:4610-4693), then one line per diagnostic output
  with the print.in edit descriptors (e.g. ``urms(1pe23.16)``) normalised as
  fparse_name does, written with the Fortran rules for those descriptors
  restated here (F, E, D, G, kP; I as F with the point blanked).
"""
from __future__ import annotations

import math
import re
import struct

import numpy as np

NGHOST = 3


# --------------------------------------------------------------- records
def read_records(path):
    """All records of an unformatted sequential file, as bytes objects."""
    out = []
    with open(path, "rb") as fh:
        data = fh.read()
    pos, n = 0, len(data)
    while pos < n:
        parts, more = [], True
        while more:
            if pos + 4 > n:
                raise ValueError(f"{path}: truncated record marker at byte {pos}")
            lead = struct.unpack_from("<i", data, pos)[0]
            ln = abs(lead)
            if pos + 8 + ln > n:
                raise ValueError(f"{path}: record of {ln} bytes at byte {pos} runs past the end")
            parts.append(data[pos + 4:pos + 4 + ln])
            trail = struct.unpack_from("<i", data, pos + 4 + ln)[0]
            if abs(trail) != ln:
                raise ValueError(f"{path}: markers {lead} / {trail} do not match at byte {pos}")
            if (trail < 0) != (len(parts) > 1):
                raise ValueError(f"{path}: trailing marker sign wrong at byte {pos}")
            more = lead < 0
            pos += 8 + ln
        out.append(b"".join(parts))
    return out


def read_var(path, mvar, n):
    """(f, t, x, y, z, dx, dy, dz) of a one-process var.dat: f has shape
    (mvar, mz, my, mx) (the Fortran array f(mx,my,mz,mvar) in C order)."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    mx, my, mz = nx + 2 * NGHOST, ny + 2 * NGHOST, nz + 2 * NGHOST
    recs = read_records(path)
    if len(recs) < 2:
        raise ValueError(f"{path}: expected 2 records, found {len(recs)}")
    f = np.frombuffer(recs[0], dtype="<f8")
    if f.size != mvar * mx * my * mz:
        raise ValueError(f"{path}: record 1 holds {f.size} values, expected {mvar * mx * my * mz}")
    g = np.frombuffer(recs[1], dtype="<f8")
    if g.size != 4 + mx + my + mz:
        raise ValueError(f"{path}: record 2 holds {g.size} values, expected {4 + mx + my + mz}")
    t = g[0]
    x, y, z = g[1:1 + mx], g[1 + mx:1 + mx + my], g[1 + mx + my:1 + mx + my + mz]
    dx, dy, dz = g[-3:]
    return f.reshape(mvar, mz, my, mx), t, x, y, z, dx, dy, dz


# --------------------------------------------------------------- Fortran edit descriptors
_DESC = re.compile(r"^\s*(?:([-+]?\d+)[pP])?\s*([iIfFeEgG])(\d+)(?:\.(\d+))?\s*$")


def parse_descriptor(form):
    """'1pe23.16' -> ('E', 23, 16, 1); 'i8' -> ('I', 8, 0, 0)."""
    m = _DESC.match(form)
    if not m:
        raise ValueError(f"unsupported edit descriptor {form!r}")
    p = int(m.group(1)) if m.group(1) else 0
    return m.group(2).upper(), int(m.group(3)), int(m.group(4) or 0), p


def _stars(w):
    return "*" * w


def _e_field(v, w, d, p):
    """Ew.d with scale factor p (Fortran 2008, 13.7.2.3.3): p digits before
    the point, d - p + 1 after (p > 0), exponent field E+dd or +ddd."""
    if math.isnan(v):
        return "NaN".rjust(w)
    if math.isinf(v):
        return ("-Inf" if v < 0 else "Inf").rjust(w)
    neg = v < 0 or (v == 0 and math.copysign(1.0, v) < 0)
    a = abs(v)
    if p > 0:
        nsig = d + 1  # significant digits: p before the point, d - p + 1 after
    else:
        nsig = d + p  # p <= 0: leading 0., |p| zeros, d + p significant digits
    if a == 0.0:
        e10 = 0
        digits = "0" * max(nsig, 1)
    else:
        s = f"{a:.{max(nsig, 1) - 1}e}"
        mant, ex = s.split("e")
        digits = mant.replace(".", "")
        e10 = int(ex) + 1  # value = 0.digits * 10**e10
        digits = digits[:max(nsig, 1)]
    exp = e10 - p if a != 0.0 else 0
    if p > 0:
        body = digits[:p] + "." + digits[p:p + d - p + 1]
    else:
        body = "0." + "0" * (-p) + digits[:d + p]
    if abs(exp) <= 99:
        ef = "E" + ("-" if exp < 0 else "+") + f"{abs(exp):02d}"
    elif abs(exp) <= 999:
        ef = ("-" if exp < 0 else "+") + f"{abs(exp):03d}"
    else:
        return _stars(w)
    s = ("-" if neg else "") + body + ef
    if len(s) > w:
        if s.startswith("0.") and len(s) - 1 == w:       # optional leading zero dropped
            s = s[1:]
        elif s.startswith("-0.") and len(s) - 1 == w:
            s = "-" + s[2:]
        else:
            return _stars(w)
    return s.rjust(w)


def _f_field(v, w, d):
    if math.isnan(v):
        return "NaN".rjust(w)
    s = f"{v:.{d}f}"
    if d == 0:
        s += "."  # F editing always writes the decimal point
    if len(s) > w:
        if s.startswith("0.") and len(s) - 1 == w:
            s = s[1:]
        elif s.startswith("-0.") and len(s) - 1 == w:
            s = "-" + s[2:]
        else:
            return _stars(w)
    return s.rjust(w)


def format_value(v, form):
    """One value with a Fortran edit descriptor (F, E, D, G, optional kP
    scale factor; I as written by print.in is handled by the caller)."""
    kind, w, d, p = parse_descriptor(form)
    v = float(v)
    if kind == "I":
        s = str(int(round(v)))
        return s.rjust(w) if len(s) <= w else _stars(w)
    if kind == "F":
        return _f_field(v, w, d)
    if kind == "E":
        return _e_field(v, w, d, p)
    if kind == "D":
        return _e_field(v, w, d, p).replace("E", "D")
    # G: F editing when 0.1 <= |v| < 10**d (with 4 trailing blanks), else E
    a = abs(v)
    if a == 0.0 or (0.1 - 0.5 * 10.0 ** (-d - 1) <= a < 10.0 ** d - 0.5):
        if a == 0.0:
            n = 1
        else:
            n = int(math.floor(math.log10(a))) + 1
            n = min(max(n, 0), d)
        return _f_field(v, w - 4, d - n) + "    "
    return _e_field(v, w, d, p)


def _name_form(cname):
    m = re.match(r"^\s*([^(\s]+)\s*(?:\((.*)\))?", cname)
    return m.group(1), m.group(2)


def print_in_column(cname):
    """(name, legend format, output format, integer) of one print.in entry,
    normalised as fparse_name does (src/diagnostics.f90:1538-1612): no format
    means 1pE10.2; E/D formats are widened to w >= d + 7 (the legend then
    shows the widened format) and written with scale factor 1P; an integer
    format Iw becomes Fw.0 whose trailing point is blanked."""
    name, form = _name_form(cname)
    if not form:
        form = "1pE10.2"
    legend_form = form
    out = form
    if form[0] in "eEdDgG":
        if form[0] in "eEdD" and "." in form:
            w, d = int(form[1:form.index(".")]), int(form[form.index(".") + 1:])
            idiff = w - d - 7
            if idiff < 0:
                form = f"{form[0]}{w - idiff}.{d}"
                legend_form = form
        out = "1p" + form
    integer = False
    kind, w, d, p = parse_descriptor(out)
    if kind == "I":
        out = f"f{w}.0"
        integer = True
    return name, legend_form, out, integer


def clean_line(line, comment_char="#"):
    """src/diagnostics.f90:634-660: a point followed by a blank, ')' or ','
    is blanked out; a line holding *** or ??? is commented out."""
    for pat in (". ", ".)", ".,"):
        while pat in line:
            i = line.index(pat)
            line = line[:i] + " " + line[i + 1:]
    if "***" in line or "???" in line:
        line = comment_char + line[1:]
    return line


def noform(cname):
    """Legend text of one diagnostic: its name centred in dashes over the
    width of its format (src/sub.f90:4610-4693; default format e10.2)."""
    name, form = _name_form(cname)
    form = form or "e10.2"
    _, width, _, _ = parse_descriptor(form)
    n1 = max(0, (width - len(name)) // 2)
    n2 = max(1, width - len(name) - n1)
    return "-" * n1 + name + "-" * n2


class TimeSeries:
    """time_series.dat of the diagnostics named in print.in lines, e.g.
    ``["it(i8)", "t(e12.4)", "dt(e12.4)", "urms(1pe23.16)"]``, as the
    reference's ``prints`` writes it (src/diagnostics.f90:314-461): the first
    call of ``append(values)`` writes the '#'-prefixed legend, each call one
    line (values in the order of the names).  Fixed-point columns are widened
    when a value needs more digits, as gen_form_legend does
    (diagnostics.f90:601-620)."""

    def __init__(self, path, print_in, comment_char="#"):
        self.path = path
        self.cols = []
        for ln in print_in:
            ln = re.split(r"[!#]", ln)[0].strip()
            if ln:
                self.cols.append(list(print_in_column(ln)))
        self.comment = comment_char
        self.first = True

    @property
    def names(self):
        return [c[0] for c in self.cols]

    @property
    def legend(self):
        return "".join(noform(f"{c[0]}({c[1]})") for c in self.cols)

    def _widen(self, col, v):
        # a fixed-point format too narrow for the value's integer digits
        # gains the missing digits (diagnostics.f90:601-620)
        kind, w, d, p = parse_descriptor(col[2])
        if kind != "F" or v == 0 or col[3]:
            return
        dlength = math.floor(math.log10(abs(v))) + 2 - float(f"{w}.{d}")
        if dlength > 0:
            col[2] = f"f{w + math.ceil(dlength)}.{d}"

    def line(self, values):
        if len(values) != len(self.cols):
            raise ValueError(f"{len(values)} values for {len(self.cols)} diagnostics")
        out = []
        for col, v in zip(self.cols, values):
            self._widen(col, float(v))
            s = format_value(v, col[2])
            if col[3]:
                s = s[:-1] + " "   # TL1," " over the point of Fw.0
            out.append(s)
        return clean_line("".join(out), self.comment).rstrip()

    def append(self, values):
        with open(self.path, "w" if self.first else "a") as fh:
            if self.first:
                fh.write(f"{self.comment}{self.legend}\n")
                self.first = False
            fh.write(self.line(values) + "\n")
