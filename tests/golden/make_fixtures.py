#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden from the reference's own files.

Runs only where the reference checkout exists (default /root/reference); the
fixtures it writes are committed, so the tests never read the reference.

  ref_tables.json   basis_coeff rows/cols 0..20 (src/2d/basis.cu:2-47) and the
                    2D / 1D quadrature rules (src/2d/quadrature.cu:7-169), the
                    1D sqrt() expressions evaluated in IEEE double
  t0/               the reference's projection golden vectors
                    src/2d/tests/t0/alpha{a}_{mesh}{p}.out for the meshes
                    canonical, box and supersimple, plus those meshes
                    (src/2d/mesh/*.pmsh, advection-prototype 6-field dialect)
  sv1.pmsh          conservation-dialect meshes with boundary sides (the
  sv1refined.pmsh   supersonic vortex's annulus sector and its two
  sv1refined1.pmsh  refinements; src/2d/conservation/euler/supersonic/mesh/)
  euler_U0.msh      the reference's gmsh output of density on sv1refined1
                    (src/2d/conservation/euler/output/U0.msh)
  gmsh/             .msh / .pmsh pairs the reference's genmesh.py made:
                    sv1, sv1refined (supersonic/mesh), cylinder, cylinder1
                    (cylinder/mesh)
  cpu_output/       the gmsh views the reference's CPU Euler driver wrote on
                    sv1 (src/2d/conservation/euler/cpu/output/uniform_rho.out,
                    uniform_u.out, uniform_E.out, p.out; main.c:201-263)
  t1/ t10/ t100/    the reference's time-step golden vectors
                    src/2d/tests/t{1,10,100}/alpha{a}_{mesh}{p}.out (central
                    flux, exact boundary values, RK4; test_all.cu:275-498,
                    614-705) for the meshes canonical and box.  The uniform
                    mesh's dumps are left out: at p >= 1 they grow without
                    bound (dt = 0.01 exceeds that mesh's stable step, values
                    up to 1e244), so they pin nothing.
"""
import json
import math
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def c_values(body):
    vals, depth, cur = [], 0, ""
    for ch in body:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            vals.append(cur)
            cur = ""
        else:
            cur += ch
    vals.append(cur)
    out = []
    for v in vals:
        v = v.strip()
        if v:
            out.append(float(eval(re.sub(r"\bsqrt\b", "math.sqrt", v), {"__builtins__": {}}, {"math": math})))
    return out


def main():
    src2d = os.path.join(REF, "src", "2d")
    basis = open(os.path.join(src2d, "basis.cu")).read()
    body = re.search(r"basis_coeff\[size\]\[size\]\s*=\s*\{(.*?)\n\};", basis, re.S).group(1)
    rows = [[float(x) for x in r.split(",")] for r in re.findall(r"\{([^{}]*)\}", body)]
    quad = open(os.path.join(src2d, "quadrature.cu")).read()
    q2 = {}
    for d in range(1, 11):
        b = re.search(r"double\s+quad_2d_degree%d\[\]\s*=\s*\{(.*?)\};" % d, quad, re.S).group(1)
        q2[d] = c_values(b)
    q1 = {}
    for n in range(1, 9):
        b = re.search(r"double\s+quad_1d_degree%d\[\]\s*=\s*\{(.*?)\};" % n, quad, re.S).group(1)
        q1[n] = c_values(b)
    data = {
        "source": "src/2d/basis.cu:2-47, src/2d/quadrature.cu:7-169",
        "basis_coeff": [r[:21] for r in rows[:21]],
        "quad2d": {str(d): q2[d] for d in q2},
        "quad1d": {str(n): q1[n] for n in q1},
    }
    with open(os.path.join(OUT, "ref_tables.json"), "w") as f:
        json.dump(data, f, indent=0)
    t0 = os.path.join(OUT, "t0")
    os.makedirs(t0, exist_ok=True)
    for mesh in ("canonical", "box", "supersimple"):
        shutil.copyfile(os.path.join(src2d, "mesh", mesh + ".pmsh"), os.path.join(t0, mesh + ".pmsh"))
        for a in range(4):
            for p in range(a, 6):
                name = "alpha%d_%s%d.out" % (a, mesh, p)
                shutil.copyfile(os.path.join(src2d, "tests", "t0", name), os.path.join(t0, name))
    for name in ("sv1.pmsh", "sv1refined.pmsh", "sv1refined1.pmsh"):
        shutil.copyfile(os.path.join(src2d, "conservation", "euler", "supersonic", "mesh", name), os.path.join(OUT, name))
    shutil.copyfile(os.path.join(src2d, "conservation", "euler", "output", "U0.msh"), os.path.join(OUT, "euler_U0.msh"))
    gm = os.path.join(OUT, "gmsh")
    os.makedirs(gm, exist_ok=True)
    for sub, name in (("supersonic", "sv1"), ("supersonic", "sv1refined"), ("cylinder", "cylinder"),
                      ("cylinder", "cylinder1")):
        for ext in (".msh", ".pmsh"):
            shutil.copyfile(os.path.join(src2d, "conservation", "euler", sub, "mesh", name + ext),
                            os.path.join(gm, name + ext))
    co = os.path.join(OUT, "cpu_output")
    os.makedirs(co, exist_ok=True)
    for name in ("uniform_rho.out", "uniform_u.out", "uniform_E.out", "p.out"):
        shutil.copyfile(os.path.join(src2d, "conservation", "euler", "cpu", "output", name), os.path.join(co, name))
    for steps in (1, 10, 100):
        d = os.path.join(OUT, "t%d" % steps)
        os.makedirs(d, exist_ok=True)
        src = os.path.join(src2d, "tests", "t%d" % steps)
        for name in sorted(os.listdir(src)):
            if re.fullmatch(r"alpha\d_(canonical|box)\d\.out", name):
                shutil.copyfile(os.path.join(src, name), os.path.join(d, name))
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
