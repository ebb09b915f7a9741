"""Writes an rt_amd Scheme as a scheme file in the reference's XML format.

The format is the one common-lib/scheme.cpp reads with mxml: attributes are
child elements whose text is the value (scheme.cpp:130-212), under
<general> (name, result_file, log_file; :367-402), one <surface> per surface
in order (class, type, shape, params, size, center, distance, angles,
comment, and for gratings grating_type, order, const, n1, n2; for lenses n1,
n2: numbers or the names of "x y" table files; :406-664) and <beam> (type,
shape, profile, range_distr, params, center, range, nlambda; :668-856).
Efficiency attributes (alpha, gamma, blaze_angle, rule_params, qe_curve;
:1157-1201) are written when given.
"""
from __future__ import annotations

import os

from rt_amd import scheme as S

CLASS = {S.RT_PLANE: "PLANE", S.RT_CONIC: "CONIC", S.RT_TORIC: "TORIC"}
TYPE = {S.RT_MIRROR: "MIRROR", S.RT_LENS: "LENS", S.RT_GRATING: "GRATING", S.RT_STOP: "STOP",
        S.RT_SHADE: "SHADE", S.RT_DETECTOR: "DETECTOR", S.RT_AUXTYPE: "AUX"}
SHAPE = {S.RT_CIRCLE: "CIRCLE", S.RT_RECTANGLE: "RECTANGLE"}
BEAM = {S.RT_BEAM_PARALLEL: "PARALLEL", S.RT_BEAM_CONIC: "CONIC", S.RT_BEAM_GAUSS: "GAUSS"}
DISTR = {S.RT_RANGE_RANDOM: "RANDOM", S.RT_RANGE_UNIFORM: "UNIFORM", S.RT_RANGE_MONO: "MONOCHROMATIC"}


def num(v) -> str:
    return "%.17g" % float(v)


def nums(vs) -> str:
    return " ".join(num(v) for v in vs)


def write_table(path: str, spline: "S.Spline"):
    with open(path, "w") as f:
        f.write("# x y\n")
        for x, y in zip(spline.x, spline.y):
            f.write("%s %s\n" % (num(x), num(y)))


def write_scheme(sc: "S.Scheme", dirpath: str, result_root: str = None, ranges=None, qe=None,
                 extra_surface=None, drop=None) -> str:
    """scheme.xml in dirpath (tables beside it).  ranges: a list of range rows
    (w1 w2 [orders...]) written to a range file, or None for the beam's
    range.  qe: {surface index: dict(alpha=, gamma=, blaze_angle=,
    rule_params=, qe_curve=)}.  extra_surface / drop: {index: {attr: text}} /
    {index: [attr, ...]} to plant attributes or leave them out."""
    os.makedirs(dirpath, exist_ok=True)
    root = result_root or os.path.join(dirpath, "result")
    tables = {}

    def fn(v, tag):
        if isinstance(v, S.Spline):
            key = id(v)
            if key not in tables:
                path = os.path.join(dirpath, "%s.txt" % tag)
                write_table(path, v)
                tables[key] = path
            return tables[key]
        return num(v)

    lines = ['<?xml version="1.0"?>', "<scheme>",
             "  <general>", "    <name>%s</name>" % (sc.name or "scheme"),
             "    <result_file>%s</result_file>" % root, "  </general>"]
    for i, s in enumerate(sc.surfaces):
        a = {}
        if s.type == S.RT_AUXTYPE:
            a["class"] = "AUX"
        else:
            a["class"] = CLASS[s.cls]
            a["type"] = TYPE[s.type]
            a["shape"] = SHAPE[s.shape]
            a["params"] = nums(s.params)
            a["size"] = nums(s.size if s.shape == S.RT_RECTANGLE else s.size[:1])
            a["center"] = nums(s.center)
        a["distance"] = nums(s.distance)
        a["angles"] = nums(s.angles)
        a["comment"] = s.comment or "surface %d" % i
        if s.type == S.RT_GRATING:
            if s.order_mode != S.RT_ORDER_FIXED:
                raise ValueError("the reference's gratings take fixed orders")
            a["grating_type"] = "REFLECTIVE" if s.grating_type == S.RT_REFLECTIVE else "TRANSPARENT"
            a["order"] = str(int(s.order))
            a["const"] = num(s.gr_const)
        if s.type in (S.RT_LENS, S.RT_GRATING):
            a["n1"] = fn(s.n1, "n1_%d" % i)
            a["n2"] = fn(s.n2, "n2_%d" % i)
        for k, v in ((qe or {}).get(i) or {}).items():
            a[k] = fn(v, "qe_%d" % i) if k == "qe_curve" else (nums(v) if isinstance(v, (list, tuple)) else str(v))
        for k, v in ((extra_surface or {}).get(i) or {}).items():
            a[k] = v
        for k in ((drop or {}).get(i) or []):
            a.pop(k, None)
        lines.append("  <surface>")
        lines += ["    <%s>%s</%s>" % (k, v, k) for k, v in a.items()]
        lines.append("  </surface>")
    b = sc.beam
    lines += ["  <beam>", "    <type>%s</type>" % BEAM[b.type], "    <shape>%s</shape>" % SHAPE[b.shape],
              "    <profile>RANDOM</profile>", "    <range_distr>%s</range_distr>" % DISTR[b.range_distr],
              "    <params>%s</params>" % nums(b.params), "    <center>%s</center>" % nums(b.center)]
    if ranges is None:
        lines.append("    <range>%s</range>" % nums(b.range))
    else:
        path = os.path.join(dirpath, "ranges.txt")
        with open(path, "w") as f:
            f.write("# w1 w2 orders\n")
            for r in ranges:
                f.write(nums(r) + "\n")
        lines.append("    <range>%s</range>" % path)
    if b.composite:
        lines.append("    <nlambda>%d</nlambda>" % int(b.nlambda))
    lines += ["  </beam>", "</scheme>"]
    path = os.path.join(dirpath, "scheme.xml")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path
