"""Runs the reference's Qfep6 (built from its sources into oracle/_ref/ by
`make -C oracle/ref`) on energy files and parses its printed tables (test
helper; Qfep6 prints Part 1 with 3 decimals and Parts 2/3 with 2)."""
import os
import subprocess

from conftest import ROOT

QFEP = os.path.join(ROOT, "oracle", "_ref", "Qfep6")


def available():
    return os.path.exists(QFEP)


def run(workdir, files, nstates, kT, nskip, nbins, nmin, alphas, coeff, noffd=1, scale_hij=1.0, couplings=None):
    """couplings: list of (i, j, A, mu, eta, r0); Qfep6 then computes Hij from
    the r_kl of the files' first off-diagonal record instead of taking the
    files' Hij."""
    if couplings:
        noffd = 0
    lines = [str(len(files)), f"{nstates} {noffd}", f"{kT} {nskip}", str(nbins), str(nmin)]
    lines += [str(a) for a in alphas[1:]]
    if noffd:
        lines.append(str(scale_hij))
    else:
        lines.append(str(len(couplings or [])))
        lines += [" ".join(str(v) for v in c) for c in couplings or []]
    lines.append(" ".join(str(c) for c in coeff))
    lines += list(files)
    r = subprocess.run([QFEP], input="\n".join(lines) + "\n", cwd=workdir, capture_output=True, text=True,
                       timeout=120)
    if r.returncode != 0:
        raise RuntimeError("Qfep6 failed: " + r.stdout[-2000:] + r.stderr[-2000:])
    return parse(r.stdout)


def parse(out):
    part, p1, p2, p3 = None, [], [], []
    for ln in out.splitlines():
        if ln.startswith("# Part 1"):
            part = 1
        elif ln.startswith("# Part 2"):
            part = 2
        elif ln.startswith("# Part 3"):
            part = 3
        elif ln.startswith("#") or not ln.strip():
            continue
        elif part == 1:
            p1.append([float(t) for t in ln.split()])
        elif part == 2:
            t = ln.split()
            p2.append(dict(lam=float(t[0]), bin=int(t[1]) - 1, gap=float(t[2]), dGa=float(t[3]), dGb=float(t[4]),
                           dGg=float(t[5]), pts=int(t[6]), c1=float(t[7]), c2=float(t[8])))
        elif part == 3:
            t = ln.split()
            p3.append(dict(bin=int(t[0]) - 1, gap=float(t[1]), dGg=float(t[2]), norm=float(t[3]), pts=int(t[4]),
                           c1=float(t[5]), c2=float(t[6]), r=float(t[7])))
    return dict(part1=p1, part2=p2, part3=p3)
