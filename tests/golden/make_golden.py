"""Writes the golden vectors of tests/golden/*.npz from the reference's own
Fortran modules.

The modules whose dependency closure needs no GSL (Constants, ErrorCheck,
Conversion, MathUtilities, SortAndSearch, VelocityGrid, Remapping, Scaling,
SpeciesAndReferenceData, DistFunc) are compiled from the reference sources by
`make -C oracle ref` into oracle/_ref/ and driven by oracle/ref_harness.f90.
Inputs are generated here from fixed seeds; outputs are the reference's exact
doubles (bit patterns). Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import struct
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT = os.path.join(ROOT, "tests", "golden")


def bits(x):
    return struct.unpack("<q", struct.pack("<d", float(x)))[0]


def unbits(i):
    return struct.unpack("<d", struct.pack("<q", int(i)))[0]


def run(cmds):
    r = subprocess.run([HARNESS], input=cmds, capture_output=True, text=True, check=True)
    return r.stdout.splitlines()


def grid_cmd(kind, n, arg):
    """kind 1: uniform with bounds arg=(lo3, hi3); kind 0: file-read axes arg=(x,y,z)."""
    s = "grid\n%d %d %d %d\n" % (kind, n[0], n[1], n[2])
    if kind == 1:
        lo, hi = arg
        s += " ".join(str(bits(v)) for v in list(lo) + list(hi)) + "\n"
    else:
        for ax in arg:
            s += "".join("%d\n" % bits(v) for v in ax)
    return s


def parse_grid(lines):
    gi = [l for l in lines if l.startswith("GRIDI")][0].split()[1:]
    gk = [l for l in lines if l.startswith("GRIDK")][0].split()[1:]
    b3 = unbits([l for l in lines if l.startswith("B3MIN")][0].split()[1])
    axes, betas = [], []
    for tag in ("AX", "AY", "AZ"):
        rows = [l.split()[1:] for l in lines if l.startswith(tag + " ")]
        axes.append(np.array([unbits(r[0]) for r in rows]))
        betas.append(np.array([unbits(r[1]) for r in rows]))
    ints = np.array([int(v) for v in gi] + [int(v) for v in gk], dtype=np.int64)
    return axes, betas, ints, b3


def remap_inputs(axes, rng, n_random):
    lo = np.array([a[0] for a in axes])
    hi = np.array([a[-1] for a in axes])
    span = hi - lo
    pts = [rng.uniform(lo - 0.3 * span, hi + 0.3 * span, size=(n_random, 3)),
           rng.uniform(lo, hi, size=(n_random, 3))]
    # exact nodes, midpoints (nearest-node ties), extreme corners, signed zeros
    nodes = np.stack([rng.choice(a, 200) for a in axes], axis=1)
    pts.append(nodes)
    mids = np.stack([(lambda a: 0.5 * (a[:-1] + a[1:]))(a)[rng.integers(0, len(a) - 1, 200)] for a in axes], axis=1)
    pts.append(mids)
    corners = np.array([[x, y, z] for x in (lo[0], hi[0], lo[0] - 0.1, hi[0] + 0.4)
                        for y in (lo[1], hi[1], hi[1] + 0.2) for z in (lo[2], hi[2], lo[2] - 0.3)])
    pts.append(corners)
    pts.append(np.array([[-0.0, 0.0, -0.0], [0.0, -0.0, 0.0], [-0.0, -0.0, -0.0]]))
    return np.concatenate(pts, axis=0)


def make_remap_golden():
    rng = np.random.default_rng(20261203)
    cases = {}
    nonuni = np.cumsum(np.r_[0.0, rng.uniform(0.2, 1.0, 10)])
    nonuni = nonuni - 0.5 * nonuni[-1]
    specs = {
        "u25": (1, (25, 25, 25), ((-5.0, -5.0, -5.0), (5.0, 5.0, 5.0))),
        "u18": (1, (18, 18, 18), ((-3.5, -3.5, -3.5), (3.5, 3.5, 3.5))),
        "u357": (1, (3, 5, 7), ((-1.0, -2.0, -3.0), (1.0, 2.0, 3.0))),
        "nonuni": (0, (11, 11, 11), (nonuni, nonuni * 1.3, nonuni * 0.7)),
        "shift": (1, (9, 9, 9), ((-2.0, -1.5, -4.0), (3.0, 2.5, 4.0))),
    }
    for name, (kind, n, arg) in specs.items():
        lines = run(grid_cmd(kind, n, arg) + "end\n")
        axes, betas, ints, b3 = parse_grid(lines)
        uvw = remap_inputs(axes, rng, 1500)
        foff = 1.0
        cmd = grid_cmd(kind, n, arg) + "remap\n%d %d\n" % (len(uvw), bits(foff))
        cmd += "".join("%d %d %d\n" % tuple(bits(v) for v in p) for p in uvw) + "end\n"
        out = [l.split()[1:] for l in run(cmd) if l.startswith("RM ")]
        st = np.array([[int(v) for v in r[:7]] for r in out], dtype=np.int32)
        ms = np.array([[unbits(v) for v in r[7:]] for r in out])
        # global2local_map on a few node indices
        N = n[0] * n[1] * n[2]
        ls = sorted(set([0, 1, n[0] - 1, n[0], N - 1] + list(rng.integers(0, N, 20))))
        cmd = grid_cmd(kind, n, arg) + "".join("g2l\n%d\n" % l for l in ls) + "end\n"
        g2l = np.array([[int(v) for v in l.split()[1:]] for l in run(cmd) if l.startswith("GL ")], dtype=np.int32)
        cases[name] = dict(kind=np.int32(kind), n=np.array(n, dtype=np.int32), ax=axes[0], ay=axes[1], az=axes[2],
                           bx=betas[0], by=betas[1], bz=betas[2], grid_ints=ints, beta3_min=np.float64(b3),
                           uvw=uvw, f_off=np.float64(foff), stencil=st, mass=ms,
                           g2l_l=np.array(ls, dtype=np.int32), g2l_ijk=g2l)
    flat = {}
    for name, d in cases.items():
        for k, v in d.items():
            flat[f"{name}__{k}"] = v
    np.savez_compressed(os.path.join(OUT, "remap_grid.npz"), **flat)
    return sorted(cases)


def buckets(cumul):
    """Bucket table input for the reference's referenced_search; built with the
    bracket rule of variance_reduction.f90:468-493."""
    n = len(cumul) - 1
    left = np.zeros(1000, dtype=np.int64)
    right = np.zeros(1000, dtype=np.int64)
    l = 1
    mx = 0.001 * cumul[n]
    for ref in range(1, 1001):
        left[ref - 1] = l - 1
        while l <= n and ref * mx > cumul[l]:
            l += 1
        right[ref - 1] = l
    right[999] = n
    return left, right


def make_search_golden():
    rng = np.random.default_rng(77)
    out = {}
    for name, n in (("n27", 27), ("n343", 343), ("n4096", 4096)):
        w = rng.exponential(size=n)
        w[rng.random(n) < 0.2] = 0.0  # empty nodes
        cumul = np.r_[0.0, np.cumsum(w)]
        left, right = buckets(cumul)
        u = np.r_[rng.random(3000), [0.0, 5e-14, 1e-13, 0.5, 1.0 - 2 ** -53],
                  np.arange(1, 1000) / 1000.0, (np.arange(1, 1000) / 1000.0) + 1e-12]
        cmd = "search\n%d %d\n" % (n, len(u))
        cmd += "".join("%d\n" % bits(v) for v in cumul)
        cmd += "".join("%d %d\n" % (a, b) for a, b in zip(left, right))
        cmd += "".join("%d\n" % bits(v) for v in u) + "end\n"
        res = np.array([[int(v) for v in l.split()[1:]] for l in run(cmd) if l.startswith("SR ")], dtype=np.int64)
        out[f"{name}__cumul"] = cumul
        out[f"{name}__left"] = left
        out[f"{name}__right"] = right
        out[f"{name}__u"] = u
        out[f"{name}__referenced"] = res[:, 0]
        out[f"{name}__binary"] = res[:, 1]
    np.savez_compressed(os.path.join(OUT, "search.npz"), **out)


def make_scalar_golden():
    rng = np.random.default_rng(5)
    pi_ref = 3.1459265358979323
    rows = []
    for _ in range(300):
        x, y, z = rng.uniform(-6, 6, 3)
        u, v, w = rng.uniform(-2, 2, 3)
        temp = rng.uniform(0.3, 4.0)
        mass = rng.choice([1.0, 0.5, 2.0])
        dens = rng.uniform(0.1, 3.0)
        coeff = dens * np.sqrt(mass * mass * mass / (pi_ref * pi_ref * pi_ref * temp * temp * temp))
        rows.append([x, y, z, coeff, mass, u, v, w, temp])
    rows = np.array(rows)
    cmd = "maxw\n%d\n" % len(rows) + "".join(" ".join(str(bits(v)) for v in r) + "\n" for r in rows) + "end\n"
    mx = np.array([unbits(l.split()[1]) for l in run(cmd) if l.startswith("MX ")])
    cs_in, cs_out = [], []
    for model_coeff in (1, 2, 3):
        for m_red, omega, d1, visc, tv in ((0.5, 0.5, 1.0, 1.0, 1.0), (0.5, 0.81, 1.0, 2.1e-5, 273.0),
                                           (0.25, 0.74, 0.8, 1.7e-5, 300.0), (1.0, 1.0, 1.2, 1.0, 1.0)):
            args = [m_red, omega, d1, d1, visc, visc, tv, tv]
            cmd = "csec\n3 %d %s\nend\n" % (model_coeff, " ".join(str(bits(v)) for v in args))
            cs_out.append(unbits([l for l in run(cmd) if l.startswith("CS ")][0].split()[1]))
            cs_in.append([model_coeff] + args)
    np.savez_compressed(os.path.join(OUT, "scalars.npz"), maxw_in=rows, maxw_out=mx,
                        csec_in=np.array(cs_in), csec_out=np.array(cs_out))


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    print("remap grids:", make_remap_golden())
    make_search_golden()
    make_scalar_golden()
    print("golden vectors written to", OUT)
