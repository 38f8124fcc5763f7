"""Independent checker of the `-d` snapshot files (main.c:269-330).

Expected values are stated from what main.c computes, not from either
writer's text: every tstep % 10 == 0 line writes `%08ld.raw` named by the
step count (four n^3 float64 fields U, V, W, P, each c2r(.)/n^3) and an
XDMF sidecar `a.%08ld.xdmf2` named by the dump counter idump = tstep / 10.
The sidecar carries the reference's quirks on purpose:
  * its DataItems point at `<idump>.raw`, not at the `<tstep>.raw` on disk
    (main.c:281,317 use idump; main.c:270 uses tstep);
  * the geometry type is spelt "ORIGIn_DXDYDZ" (main.c:294);
  * Seek of variable v is v * n^3 * sizeof(double) (main.c:306,321);
  * Time is t printed with "%+.16e", spacings 2 pi / n with "%.16e".
"""
import math
import os
import re
import xml.etree.ElementTree as ET

import numpy as np

TIME_RE = re.compile(r"^[+-]\d\.\d{16}e[+-]\d{2,3}$")
FLOAT_RE = re.compile(r"^-?\d\.\d{16}e[+-]\d{2,3}$")


def dump_times(dt, nsteps_printed):
    """t of every printed line: t accumulates t += dt from 0.0 (main.c:254,421)."""
    t, out = 0.0, {}
    for step in range(nsteps_printed + 1):
        if step % 10 == 0:
            out[step] = t
        t += dt
    return out


def check_xdmf(path, n, idump, t):
    text = open(path).read()
    root = ET.fromstring(text)
    assert root.tag == "Xdmf" and root.attrib == {"Version": "2"}
    (domain,) = list(root)
    assert domain.tag == "Domain"
    (grid,) = list(domain)
    assert grid.tag == "Grid" and grid.attrib == {}
    kids = list(grid)
    assert [k.tag for k in kids] == ["Time", "Topology", "Geometry"] + ["Attribute"] * 4
    tv = kids[0].attrib["Value"]
    assert TIME_RE.match(tv), tv
    assert float(tv) == t
    assert kids[1].attrib == {"TopologyType": "3DCoRectMesh", "Dimensions": f"{n} {n} {n}"}
    geo = kids[2]
    assert geo.attrib == {"GeometryType": "ORIGIn_DXDYDZ"}
    origin, spacing = list(geo)
    assert origin.tag == spacing.tag == "DataItem"
    assert origin.attrib == spacing.attrib == {"Dimensions": "3"}
    assert origin.text.split() == ["0", "0", "0"]
    sp = spacing.text.split()
    assert len(sp) == 3 and all(FLOAT_RE.match(s) for s in sp), sp
    assert all(float(s) == 2 * math.pi / n for s in sp)
    for v, (att, name) in enumerate(zip(kids[3:], "UVWP")):
        assert att.attrib == {"name": name}
        (item,) = list(att)
        assert item.attrib == {"Format": "Binary", "Seek": str(v * n ** 3 * 8), "Precision": "8",
                               "Dimensions": f"{n} {n} {n}"}
        assert item.text.strip() == f"{idump:08d}.raw"
    # one element or attribute per line, as main.c:283-327 prints them
    assert len(text.splitlines()) == 27 + 4 * 10
    assert text.endswith("</Xdmf>\n")


def check_dump_dir(d, n, dt, last_step, u0, tol=1e-15):
    """Files of a run whose last printed line is `last_step`."""
    times = dump_times(dt, last_step)
    want = set()
    for idump, (step, t) in enumerate(sorted(times.items())):
        raw = f"{step:08d}.raw"
        xdmf = f"a.{idump:08d}.xdmf2"
        want |= {raw, xdmf}
        assert os.path.getsize(os.path.join(d, raw)) == 4 * n ** 3 * 8
        check_xdmf(os.path.join(d, xdmf), n, idump, t)
    assert set(f for f in os.listdir(d) if f.endswith((".raw", ".xdmf2")) and f != "in.raw") == want
    a = np.fromfile(os.path.join(d, "00000000.raw"), dtype="<f8").reshape(4, n, n, n)
    # step 0: the input velocity through r2c and c2r/n^3; P is zero before
    # the first step (the reference writes an uninitialised array there)
    assert np.max(np.abs(a[:3] - u0[:3])) < tol * max(1.0, np.max(np.abs(u0)))
    assert np.all(a[3] == 0)
