"""On-disk formats beside the hot path: Fortran unformatted sequential
records (the var.dat container) and time_series.dat lines.

* Records: the bytes pcmx_write_record produces are compared with bytes
  assembled here from the gfortran runtime's convention (4-byte counts
  around each subrecord, leading count negated when another subrecord
  follows, trailing count negated when one precedes), including subrecord
  splits forced with a small subrecord limit; pcmx_io.read_records reads
  them back and rejects damaged files.  No GPU is needed.
* Time series: pcmx_io.TimeSeries reproduces, byte for byte, the legends and
  the first lines of the reference's own sample outputs
  (samples/*/reference.out with their print.in; fixture written by
  tests/golden/make_timeseries_fixtures.py)."""
import json
import os
import struct

import numpy as np
import pytest

import pcmx
import pcmx_io as I

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "timeseries_reference_samples.json")


def framed(payload, max_sub):
    """The convention, assembled independently of the library."""
    if len(payload) == 0:
        return struct.pack("<ii", 0, 0)
    subs = [payload[i:i + max_sub] for i in range(0, len(payload), max_sub)]
    out = b""
    for k, s in enumerate(subs):
        lead = -len(s) if k < len(subs) - 1 else len(s)
        trail = -len(s) if k > 0 else len(s)
        out += struct.pack("<i", lead) + s + struct.pack("<i", trail)
    return out


@pytest.mark.parametrize("max_sub", [0, 8, 24, 40, 1000])
def test_record_bytes_follow_the_convention(tmp_path, max_sub):
    a = np.arange(7, dtype=np.float64) * 0.5
    b = np.array([np.pi, -1e300, 2.5e-310])
    c = np.zeros(0)
    path = tmp_path / "r.dat"
    pcmx.write_record(str(path), [a, c, b], max_subrecord=max_sub)
    payload = a.tobytes() + b.tobytes()
    assert path.read_bytes() == framed(payload, max_sub or 2147483639)
    assert I.read_records(str(path)) == [payload]


def test_subrecords_split_inside_values_and_at_exact_multiples(tmp_path):
    path = tmp_path / "r.dat"
    x = np.arange(10, dtype=np.float64)            # 80 bytes
    pcmx.write_record(str(path), [x], max_subrecord=20)     # 20/20/20/20: exact multiple, splits mid-value
    raw = path.read_bytes()
    assert raw == framed(x.tobytes(), 20)
    leads = [struct.unpack_from("<i", raw, 28 * k)[0] for k in range(4)]
    assert leads == [-20, -20, -20, 20]
    pcmx.write_record(str(path), [x[:3]], append=True, max_subrecord=16)   # 24 bytes: 16 + 8
    recs = I.read_records(str(path))
    assert recs == [x.tobytes(), x[:3].tobytes()]


def test_empty_record_and_several_records(tmp_path):
    path = tmp_path / "r.dat"
    pcmx.write_record(str(path), [])
    pcmx.write_record(str(path), [np.ones(2)], append=True)
    assert path.read_bytes() == struct.pack("<ii", 0, 0) + framed(np.ones(2).tobytes(), 1 << 30)
    assert I.read_records(str(path)) == [b"", np.ones(2).tobytes()]


def test_reader_rejects_damaged_files(tmp_path):
    path = tmp_path / "r.dat"
    good = framed(np.arange(6, dtype=np.float64).tobytes(), 16)
    for bad in (good[:-1],                                         # truncated
                good[:4] + good[4:-4] + struct.pack("<i", 999),    # trailing count differs
                struct.pack("<i", -16) + good[4:20] + struct.pack("<i", 16)):   # continuation never comes
        path.write_bytes(bad)
        with pytest.raises(ValueError):
            I.read_records(str(path))


def test_write_record_errors():
    with pytest.raises(pcmx.PcmxError, match="cannot open"):
        pcmx.write_record("/nonexistent-dir/x.dat", [np.ones(1)])
    with pytest.raises(pcmx.PcmxError, match="bad arguments"):
        pcmx.write_record("/tmp/x.dat", [np.ones(1)], max_subrecord=-1)


def test_snapshot_needs_an_engine(tmp_path):
    with pytest.raises(pcmx.PcmxError, match="pcmx_initialize"):
        pcmx.write_snapshot(str(tmp_path / "var.dat"), 0.0)


def test_read_var_layout(tmp_path):
    """A file assembled with the record writer in the var.dat layout reads
    back into f(mx,my,mz,mvar) order and the grid record."""
    mvar, n = 2, (4, 5, 6)
    mx, my, mz = n[0] + 6, n[1] + 6, n[2] + 6
    f = np.random.default_rng(1).standard_normal((mvar, mz, my, mx))
    x, y, z = np.arange(mx) * 0.1, np.arange(my) * 0.2, np.arange(mz) * 0.3
    path = tmp_path / "var.dat"
    pcmx.write_record(str(path), [f])
    pcmx.write_record(str(path), [np.array([1.5]), x, y, z, np.array([0.1, 0.2, 0.3])], append=True)
    g, t, xr, yr, zr, dx, dy, dz = I.read_var(str(path), mvar, n)
    np.testing.assert_array_equal(g, f)
    assert t == 1.5 and (dx, dy, dz) == (0.1, 0.2, 0.3)
    np.testing.assert_array_equal(xr, x)
    np.testing.assert_array_equal(zr, z)


def golden():
    with open(GOLDEN) as fh:
        return json.load(fh)


def test_golden_fixture_is_broad():
    fx = golden()
    forms = {c.split("(")[1].rstrip(")").lower() if "(" in c else "" for f in fx for c in f["columns"]}
    assert len(fx) >= 80
    assert any(fm.startswith("i") for fm in forms) and any(fm.startswith("f") for fm in forms)
    assert any(fm.startswith("e") for fm in forms) and any(fm.startswith("1p") for fm in forms)


@pytest.mark.parametrize("case", golden(), ids=lambda c: c["sample"])
def test_time_series_matches_reference_output(case, tmp_path):
    ts = I.TimeSeries(str(tmp_path / "time_series.dat"), case["columns"])
    assert ts.legend == case["legend"]
    for row in case["rows"]:
        assert ts.line(row["values"]) == row["line"]
    # the file: '#' + legend, then the lines
    ts2 = I.TimeSeries(str(tmp_path / "ts2.dat"), case["columns"])
    for row in case["rows"]:
        ts2.append(row["values"])
    text = (tmp_path / "ts2.dat").read_text().splitlines()
    assert text[0] == "#" + case["legend"]
    assert text[1:] == [r["line"] for r in case["rows"]]


def test_descriptor_rules():
    """Closed cases of the restated edit descriptors."""
    assert I.format_value(0.0, "F6.0") == "    0."
    assert I.format_value(-0.446, "F8.4") == " -0.4460"
    assert I.format_value(6.793e-3, "1pE10.3") == " 6.793E-03"
    assert I.format_value(6.793e-3, "E10.3") == " 0.679E-02"
    assert I.format_value(1.0e123, "1pE10.3") == " 1.000+123"
    assert I.format_value(123456.0, "F5.1") == "*****"
    name, legend_form, out, integer = I.print_in_column("dt(e10.4)")   # w < d + 7: widened to e11.4
    assert (legend_form, out, integer) == ("e11.4", "1pe11.4", False)
    assert I.print_in_column("it(i6)")[2:] == ("f6.0", True)
    assert I.noform("urms(1pe23.16)") == "---------urms----------"
