"""Writes timeseries_reference_samples.json: for the reference's sample runs
(samples/*/print.in with the time series their reference.out holds, the
output the reference's own auto-test compares against), the print.in
entries in legend order, the legend line and the first data lines, each
with the values read from its fields.  Run from the repository root with
the reference checkout at /root/reference; the fixture is data only.

A sample is kept when every legend column names a print.in entry with a
plain (real, non-complex, no-location) format and its lines split into
fields of the formats' widths."""
import glob
import json
import os
import re
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "..", "pencil-code_amd"))
import pcmx_io as I  # noqa: E402

REF = "/root/reference/samples"
OUT = os.path.join(os.path.dirname(__file__), "timeseries_reference_samples.json")
MAX_LINES = 4
# reference outputs that their own print.in does not describe (left out):
EXCLUDE = {
    # the Rmesh column is printed with two decimals, print.in asks for E10.3
    "continuous-forcing-from_file_h5",
    # the legend's last column carries two more dashes than noform gives for
    # e13.3 (a legend of another version of the code)
    "debris-photoelectric-fluid",
}


def entries(print_in):
    out = {}
    for ln in open(print_in, errors="replace"):
        ln = re.split(r"[!#]", ln)[0].strip()
        if not ln:
            continue
        name, form = I._name_form(ln)
        out.setdefault(name, ln)
    return out


def main():
    fixtures = []
    for d in sorted(glob.glob(f"{REF}/*/")):
        pin, ref = os.path.join(d, "print.in"), os.path.join(d, "reference.out")
        if os.path.basename(d.rstrip("/")) in EXCLUDE:
            continue
        if not (os.path.exists(pin) and os.path.exists(ref)):
            continue
        lines = open(ref, errors="replace").read().splitlines()
        if not lines or not lines[0].startswith("#"):
            continue
        legend = lines[0][1:]
        names = re.findall(r"[^-]+", legend)
        ent = entries(pin)
        if not names or any(n not in ent for n in names):
            continue
        cols = [ent[n] for n in names]
        try:
            ts = I.TimeSeries(os.devnull, cols)
        except ValueError:
            continue
        if ts.legend.rstrip("-") != legend.rstrip("-"):
            continue
        widths = [I.parse_descriptor(c[2])[1] for c in ts.cols]
        rows = []
        for ln in lines[1:]:
            if ln.startswith("#") or not ln.strip():
                continue
            if len(rows) >= MAX_LINES:
                break
            vals, pos, ok = [], 0, True
            for w in widths:
                field = ln[pos:pos + w]
                pos += w
                try:
                    vals.append(float(field.replace("D", "E")))
                except ValueError:
                    ok = False
                    break
            if ok and pos >= len(ln.rstrip()):
                rows.append({"values": vals, "line": ln.rstrip()})
        if rows:
            fixtures.append({"sample": os.path.basename(d.rstrip("/")), "columns": cols,
                             "legend": legend, "rows": rows})
    with open(OUT, "w") as fh:
        json.dump(fixtures, fh, indent=1)
    print(f"{len(fixtures)} samples -> {OUT}")


if __name__ == "__main__":
    main()
