"""fdtd_readConfig, the config.txt reader of the reference's driver (main.c:319-366,
through parser_nextLine, parser.c:3-20), on the library's C host (no GPU call).

Checked against the file format of the reference's configSample.txt (comment
lines, blank lines, trailing "# notes"), all 11 values in main.c's order, and,
where the reference's sources are here, against the reference's own parser
(oracle/_ref/ref_harness mode=config: parser_nextLine plus main.c's atoi /
strtod conversions) on the same files, including the reference's own sample."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
REF_SAMPLE = "/root/reference/configSample.txt"

ORDER = ("width_nm", "height_nm", "h_u_nm", "pml", "lambda_nm", "stepNum", "startAngle", "endAngle", "deltaAngle",
         "model", "solver")

SAMPLE = """# simulation parameters
# "#" starts a comment
# units [nm]

4130  #width
2065  #height
5    #h_u
32    #pml
550.9   #lambda
20000  #step

#incidence angle deg
15     #start
75    #end
10     #delta

#Model -> models.h
3

#Solver -> simulator.h
2
"""


def read(path):
    from mpifdtd_amd import _native as N
    cfg = N.RunConfig()
    rc = N.lib().fdtd_readConfig(str(path).encode(), ctypes.byref(cfg))
    if rc != 0:
        return rc, None
    fi = cfg.field_info
    vals = [fi.width_nm, fi.height_nm, fi.h_u_nm, fi.pml, fi.lambda_nm, fi.stepNum, cfg.startAngle, cfg.endAngle,
            cfg.deltaAngle, cfg.model, cfg.solver]
    assert fi.angle_deg == cfg.startAngle  # the sweep starts at startAngle
    return rc, vals


def reference_parse(path):
    if not os.access(HARNESS, os.X_OK):
        return None
    out = subprocess.run([HARNESS, "mode=config", f"path={path}"], capture_output=True, text=True, timeout=30)
    import json
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_sample_format_all_fields_in_order(tmp_path):
    p = tmp_path / "config.txt"
    p.write_text(SAMPLE)
    rc, vals = read(p)
    assert rc == 0
    assert dict(zip(ORDER, vals)) == dict(width_nm=4130, height_nm=2065, h_u_nm=5, pml=32, lambda_nm=550,
                                          stepNum=20000, startAngle=15, endAngle=75, deltaAngle=10, model=3, solver=2)
    ref = reference_parse(p)
    if ref is not None:
        assert vals == ref


@pytest.mark.skipif(not os.path.exists(REF_SAMPLE), reason="the reference's sources are not here")
def test_reference_sample_file():
    rc, vals = read(REF_SAMPLE)
    assert rc == 0
    assert vals == [2560, 2560, 10, 10, 500, 2000, 0, 90, 5, 3, 2]
    ref = reference_parse(REF_SAMPLE)
    if ref is not None:
        assert vals == ref


def test_reference_parser_quirks(tmp_path):
    """A line of blanks is a value line (0), "2.5e3" is 2 for the atoi fields and the
    lambda field truncates strtod's value; a CRLF file's "\\r\\n" lines are value lines
    too.  Same values as the reference's parser."""
    p = tmp_path / "quirks.txt"
    p.write_text("# c\n2.5e3\n   \n7 # x\n8\n640.99\n100\n0\n0\n1\n1\n4\n")
    rc, vals = read(p)
    assert rc == 0
    assert vals == [2, 0, 7, 8, 640, 100, 0, 0, 1, 1, 4]
    ref = reference_parse(p)
    if ref is not None:
        assert vals == ref
    q = tmp_path / "crlf.txt"
    q.write_bytes(b"#a\r\n2560\r\n\r\n2560\r\n10\r\n10\r\n500\r\n2000\r\n0\r\n90\r\n5\r\n3\r\n2\r\n")
    rc, vals = read(q)
    assert vals == [2560, 0, 2560, 10, 10, 500, 2000, 0, 90, 5, 3]
    ref = reference_parse(q)
    if ref is not None:
        assert vals == ref


def test_missing_file_and_short_file(tmp_path):
    assert read(tmp_path / "none.txt")[0] == -1
    p = tmp_path / "short.txt"
    p.write_text("1\n2\n3\n")
    assert read(p)[0] == -2
