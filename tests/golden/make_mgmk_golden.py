"""Writes the fixtures of the reference's own solver (mgMK) under ``tests/golden/``.

Runs in a container that holds the reference tree: builds mgMK from its
Fortran sources with ``oracle/Makefile.ref`` (flang + MPICH, ``-DPARALLEL``)
and drives it with ``oracle/mgmk_harness.F90``.  Run from the repository
root: ``python tests/golden/make_mgmk_golden.py [all|vectors|cycles]``.

``mgmk_uniform_{16,32}.npz`` -- the uniform-grid, config-1-style problem
(2*pi x 2*pi x 1, periodic x/y, Neumann z, RHS = Psource of seeded N(0,1)
face velocities, volume mean removed) at 16^3 and 32^3, solved by the
reference's ``mg_solver`` (RBPGS V(2,2), coarsened to 4^3 at 16^3 / 32^3,
tolmax 1e-14) on one rank and on four (2x2x1); both runs must give the
same bits.  Stored: the interior RHS, the interior solution, V-cycles, the
final relmax and relative residual, and the relative residual after each
of the first cycles.

``mgmk_cycles.json`` -- V-cycles the reference needs to reach a relative
residual of 1e-10 on the 64^3 (config 1), 128^3 and 256^3 versions of that
problem (the level counts and MPI splits of RUNS), and its relative
residual every 50 cycles.  The
benchmark's CPU baseline times the reference's V-cycle at 256^3 on the
GPU box's cores and scales it by this cycle count.  Also the 64^3 solve
to 1e-10 timed here on 1 and 8 ranks, with this container's CPU count.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

REF = "/root/reference/ANELASTIC_SL"
HARNESS = os.path.join(ROOT, "oracle", "_ref", "mgmk_harness")
MPIEXEC = "/opt/conda/bin/mpiexec"
L = (2 * np.pi, 2 * np.pi, 1.0)
# (n, levels, ranks, cycles traced).  The reference's MPI build splits x
# and y only here, as its production run does (16x8x1): a z split returns
# NaN with Neumann z walls, four ranks in y give other bits than one rank,
# and splits that leave one coarsest-level point per rank fault.  64^3 and
# 128^3 coarsen to 4^3 (config 1's hierarchy) on 2x2 ranks; 256^3 stops
# at 8^3 so that 4x2 ranks keep two coarsest points each.
RUNS = ((64, 5, (2, 2, 1), 800), (128, 6, (2, 2, 1), 1500), (256, 6, (4, 2, 1), 3500))


def ini_text(levels, tolmax=1e-14, maxcy=3000):
    lv = ",".join(["2"] * levels)
    return f"""[MGsolver]
Smoother=RBPGS
Strict=yes
Verbose=no
nxlevels={levels}
nylevels={levels}
nzlevels={levels}
maxcy={maxcy}
tolmax={tolmax:.1e}
PreRelax={lv}
PstRelax={lv}
[BCs]
X1=periodic
X2=periodic
X3=read
[X3min]
PTYPE=Neumann
PVALUE=0.d0
[X3max]
PTYPE=Neumann
PVALUE=0.d0
""".replace("e-", "d-")


def rhs(n):
    """Ghosted RHS (interior = compatible Psource of the seeded field)."""
    from anelastic_solve_amd import workloads
    from conftest import oracle_levels
    from oracle import ops
    g, U, V, W = workloads.make(workloads.UNIFORM_64, n=(n, n, n))
    b = workloads.psource_host(U, V, W, g)
    levs, _ = oracle_levels(g, (0, 1, 0, 1, 0, 0))
    full = np.zeros_like(b)
    full[1:-1, 1:-1, 1:-1] = ops.compat_project(b[1:-1, 1:-1, 1:-1], levs[0])
    return full


def run(n, levels, procs, mode, count, tmp, out="-", **kw):
    ini = os.path.join(tmp, f"mg{n}.ini")
    with open(ini, "w") as fh:
        fh.write(ini_text(levels, **kw))
    rf = os.path.join(tmp, f"rhs{n}.bin")
    if not os.path.exists(rf):
        rhs(n).astype("<f8").tofile(rf)
    cmd = [MPIEXEC, "-n", str(procs[0] * procs[1] * procs[2]), HARNESS, ini, str(n), str(n), str(n),
           *map(str, procs), *map(repr, L), rf, out, mode, str(count), tmp]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp, check=True)
    return [ln.split() for ln in r.stdout.splitlines() if ln.strip()]


def main(what="all"):
    subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "Makefile.ref"),
                    f"REF={REF}"], check=True, cwd=ROOT)
    with tempfile.TemporaryDirectory() as tmp:
        for n, levels in ((16, 3), (32, 4)) if what in ("all", "vectors") else ():
            sols = []
            for procs in ((1, 1, 1), (2, 2, 1)):
                out = os.path.join(tmp, f"phi{n}_{procs[0]}.bin")
                rows = run(n, levels, procs, "solve", 0, tmp, out=out)
                row = [r for r in rows if r[0] == "SOLVE"][0]
                phi = np.fromfile(out, dtype="<f8").reshape(n + 2, n + 2, n + 2)
                sols.append((phi, int(row[1]), float(row[2]), float(row[3])))
            assert np.array_equal(sols[0][0], sols[1][0]), "1 rank and 4 ranks differ"
            trace = run(n, levels, (1, 1, 1), "trace", 30, tmp)
            phi, it, relmax, relres = sols[0]
            np.savez(os.path.join(HERE, f"mgmk_uniform_{n}.npz"),
                     rhs=rhs(n)[1:-1, 1:-1, 1:-1], phi=phi[1:-1, 1:-1, 1:-1], iters=it,
                     relmax=relmax, relres=relres,
                     trace_relres=np.array([float(r[3]) for r in trace if r[0] == "CYCLE"]))
            print(f"{n}^3: {it} cycles, relmax {relmax:.2e}, relres {relres:.2e}")
        if what not in ("all", "cycles"):
            return
        cyc = {}
        for n, levels, procs, count in RUNS:
            # the split must not change the reference's answer: its first
            # cycles on one rank and on the split give the same bits
            a = run(n, levels, (1, 1, 1), "trace", 2, tmp)
            b = run(n, levels, procs, "trace", 2, tmp)
            assert a == b, f"{n}^3: {procs} ranks differ from one rank"
            rows = [r for r in run(n, levels, procs, "trace", count, tmp) if r[0] == "CYCLE"]
            res = [float(r[3]) for r in rows]
            hit = next((q + 1 for q, v in enumerate(res) if v <= 1e-10), None)
            cyc[str(n)] = {"levels": levels, "ranks": list(procs), "cycles_to_1e-10": hit,
                           "relres_every_50": res[49::50], "cycles_run": len(res)}
            print(f"{n}^3: 1e-10 after {hit} cycles")
        # config 1 (64^3) solved to 1e-10 by the reference in this container,
        # timed as mgMK/test/test.F90:443-457 times it, on 1 and on 8 ranks
        need = cyc["64"]["cycles_to_1e-10"]
        here = {"cpus": os.cpu_count(), "cycles": need}
        for procs in ((1, 1, 1), (2, 2, 1)):
            rows = run(64, 5, procs, "time", 3, tmp, maxcy=need)
            row = [r for r in rows if r[0] == "TIMED"][0]
            here[f"ranks_{procs[0] * procs[1]}"] = {"s_per_solve_best_of_3": float(row[3]),
                                                    "relres": float(row[5])}
            print(f"64^3 on {procs}: {float(row[3]):.2f} s per solve")
        with open(os.path.join(HERE, "mgmk_cycles.json"), "w") as fh:
            json.dump({"problem": "uniform n^3 on 2pi x 2pi x 1, periodic x/y, Neumann z, "
                                  "RHS = Psource of N(0,1) face velocities (seed 0), "
                                  "RBPGS V(2,2); levels and MPI ranks per size",
                       "stop": "max|b - A p| / max|b| <= 1e-10",
                       "sizes": cyc,
                       "config1_timed_in_build_container": here}, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "all")
