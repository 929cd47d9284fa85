#!/usr/bin/env python3
"""Benchmark of the dense 2-D Laplace BEM path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--ne NE] [--no-cpu] [--no-fp64]
                    [--no-adjoint] [--phase all|assemble]

Workload: random star r(t) = 1 + sum_{m=2..8} a_m cos(m t + phi_m) (seed 1),
quadratic elements with +-10% jittered angular widths, N = 32768 requested ->
N_eff = 32769 collocation nodes (10923 elements x 3 Gauss nodes), npg = 16,
T = e^x cos y, Dirichlet/Neumann arcs of 64 elements.

A step = one pass of the hot path with the geometry resident in HBM:
assemble H, G (far field FP32, near/singular/diagonal FP64) -> mixed-BC
column swap and FP64 right-hand side -> FP32 LU (trailing updates on the bf16
matrix cores, each FP32 operand split exactly into three bf16 parts) ->
FP64 iterative refinement until ||dx|| <= 1e-10 ||x|| (or until the error left,
estimated from the observed contraction, is ten times below that). The all-FP64 pipeline
(FP64 assembly + FP64 LU) runs once after the timed steps for the speed-up and
the accuracy check (mixed vs FP64, per field).

value = H,G entries per second of the assembly phase (2 nt n entries per
assembly / assembly seconds, HIP events on the library stream, averaged over
the timed steps). ms_per_step = wall time of a whole step.

After the timed steps (outside them): 2-D elasticity (Kelvin) on a star of
N = 32766 dofs (5461 quadratic elements, plane strain E = 1, nu = 0.3, uniform
strain field, traction on the Neumann arcs): the mixed pipeline (FP32 storage,
FP32 LU, FP64 refinement) and the FP64 pipeline, the Kelvin far kernel's
roofline and an oracle CPU sample; the
adjoint shape gradient on the
N = 4095 star (1365 elements, npg = 12, target fluxes from the FP64 solve on
the shape with a_m shifted by U(-0.01, 0.01)) after an FP64 and after a mixed
solve; the 64-domain sweep (stars with N ~ U[2048, 16384], npg = 16): forward
solve and adjoint gradient per domain with the factors resident, mixed and
FP64, as entries/s, solves/s and gradients/s; and the CPU baseline (the FP64
restatement of the reference algorithm on a bounded sample, in a child process
whose OpenMP and BLAS thread counts are set before anything loads them).
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import bem_gmsh_amd as B  # noqa: E402
from bem_gmsh_amd import _lib  # noqa: E402
from bem_gmsh_amd._lib import dptr, i64ptr  # noqa: E402

PEAK_HBM = 8.0e12          # B/s
PEAK_MFMA_F32 = 157.3e12   # FLOP/s dense
PEAK_MFMA_F64 = 78.6e12
PEAK_MFMA_BF16 = 2.5e15    # FLOP/s dense
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_star_N32769.json")
KELVIN_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_kelvin_star_N32766.json")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_flops_star_N32769.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_problem(ne: int, seed: int = 1):
    mesh = B.star_mesh(ne, seed=seed)
    dad = B.format2d(mesh, B.Laplace(1.0))
    B.apply_bc_from_physical(dad, B.ana_exp_cos())
    return dad


class Ctx(B.HipContext):
    """A library context driven through the raw C ABI (closed at exit like every
    HipContext)."""

    def __init__(self):
        super().__init__(0)
        self.L = _lib.lib()

    def chk(self, st):
        _lib.check(st, self.h)

    def set_geometry(self, dad):
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (
            dad.collocation, dad.Normal, dad.node_J(), dad.elem_weight,
            dad.elem_geo.reshape(-1, 2), dad.elem_L)]
        a = self._keep
        self.chk(self.L.bemhip_set_geometry(self.h, dad.p, dad.ne, dad.n, dad.ni, dptr(a[0]),
                                            dptr(a[1]), dptr(a[2]), dptr(a[3]), dptr(a[4]),
                                            dptr(a[5]), float(getattr(dad.properties, "k", 1.0))))

    def pipeline(self, dad, npg, precision, mode, tol=1e-10, maxit=30):
        """assemble -> applyBC -> solve; the mixed solve refines until
        ||dx|| <= tol ||x||, tol = the metric's 1e-10 relative accuracy."""
        BC = np.ascontiguousarray(dad.BC, dtype=np.int64)
        BV = np.ascontiguousarray(dad.BV, dtype=np.float64)
        self.chk(self.L.bemhip_assemble(self.h, npg, 1.5, precision))
        self.chk(self.L.bemhip_apply_bc(self.h, i64ptr(BC), dptr(BV), None, None))
        x = np.empty(dad.nt)
        it = C.c_int(0)
        rel = C.c_double(0)
        self.chk(self.L.bemhip_solve(self.h, mode, tol, maxit, dptr(x), C.byref(it),
                                     C.byref(rel)))
        self.chk(self.L.bemhip_synchronize(self.h))
        return x, it.value, rel.value

    def interior(self, pts, T, q, npg, near_factor):
        P = np.ascontiguousarray(pts, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        u = np.empty(P.shape[0])
        self.chk(self.L.bemhip_interior_laplace(self.h, P.shape[0], dptr(P), dptr(T), dptr(q),
                                                npg, near_factor, dptr(u)))
        return u


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_threads():
    """Threads for the CPU leg: the cores this process may run on, capped by
    OMP_NUM_THREADS when the environment sets it (the job's CPU share)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(avail, env) if env > 0 else avail


def cpu_leg(ne, npg, rows, lu_n):
    """Child process: reference-algorithm CPU restatement (the FP64 oracle,
    OpenMP over rows) on a bounded sample, i.e. `rows` assembly rows of this
    workload and an LAPACK LU (scipy/OpenBLAS) of size lu_n, extrapolated to the
    workload as n^3. OMP_NUM_THREADS / OPENBLAS_NUM_THREADS are set by the
    parent before this process imports scipy; threadpoolctl reports what the
    libraries actually use."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import scipy.linalg as sl
    from threadpoolctl import threadpool_info
    dad = build_problem(ne)
    threads = int(os.environ["OMP_NUM_THREADS"])
    t = time.perf_counter()
    O.assemble(dad, npg, nthreads=threads, rows=(0, rows))
    t_asm = time.perf_counter() - t
    rng = np.random.default_rng(0)
    A = rng.standard_normal((lu_n, lu_n)) + lu_n * np.eye(lu_n)
    sl.lu_factor(A[:256, :256].copy())           # load the BLAS before timing
    t = time.perf_counter()
    sl.lu_factor(A, overwrite_a=True, check_finite=False)
    t_lu = time.perf_counter() - t
    pools = [{"api": p.get("internal_api"), "threads": p.get("num_threads"),
              "version": p.get("version")} for p in threadpool_info()]
    print(json.dumps({"entries_per_s": 2.0 * rows * dad.n / t_asm, "assembly_sample_s": t_asm,
                      "rows": rows, "lu_sample_n": lu_n, "lu_sample_s": t_lu,
                      "lu_extrapolated_s": t_lu * (dad.nt / lu_n) ** 3,
                      "threadpools": pools}))


def cpu_baseline(ne, npg, rows, lu_n, threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OPENBLAS_NUM_THREADS=str(threads),
               MKL_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", "--ne", str(ne),
                        "--npg", str(npg), "--cpu-rows", str(rows), "--cpu-lu-n", str(lu_n)],
                       env=env, capture_output=True, text=True, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def load_traffic(nt, path=TRAFFIC_FILE):
    """HBM bytes per launch measured by rocprofv3 PMC (FETCH_SIZE + WRITE_SIZE,
    scaled by a plain-copy calibration), from the committed summary of the
    same workload; None when absent or for another size."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if int(t.get("nt", -1)) != nt:
        return None
    return {k: v["bytes_per_launch"] for k, v in t.get("kernels", {}).items()}


def adjoint_leg(ne=1365, seed=1, npg=12):
    """Adjoint shape gradient on the N = 4095 star: times (ms, HIP events) of the
    transposed solve and of the gradient pass, after an FP64 and a mixed solve."""
    mesh = B.star_mesh(ne, seed=seed)
    dad = B.format2d(mesh, B.Laplace(1.0))
    B.apply_bc_from_physical(dad, B.ana_exp_cos())
    mp = B.star_mesh(ne, seed=seed, params=B.star_params(seed, perturb=True))
    dp = B.format2d(mp, B.Laplace(1.0))
    B.apply_bc_from_physical(dp, B.ana_exp_cos())
    B.assemble(dp, npg, fetch=False)
    B.solve(dp, mode="fp64")
    qt = dp.q.copy()
    out = {"workload": f"star N={dad.n} ({ne} quadratic elements), npg={npg}"}
    grads = {}
    for prec, mode in (("fp64", "fp64"), ("mixed", "mixed")):
        B.assemble(dad, npg, precision=prec, fetch=False)
        B.solve(dad, mode=mode)
        B.adjoint(dad, qt)                       # warm-up (kernel loading)
        t = time.perf_counter()
        J, g = B.adjoint(dad, qt)
        wall = time.perf_counter() - t
        ms = dad.hip.counters()["ms"]
        grads[mode] = g
        out[mode] = {"J": J, "adjoint_ms": ms["adjoint"],
                     "transposed_solve_ms": ms["adjoint_solve"],
                     "gradient_pass_ms": ms["adjoint_grad"], "wall_s": wall,
                     "solve_ms": ms.get("solve64" if mode == "fp64" else "solve_mixed")}
    out["grad_mixed_vs_fp64"] = float(np.abs(grads["mixed"] - grads["fp64"]).max()
                                      / np.abs(grads["fp64"]).max())
    log(f"[bench] adjoint N={dad.n}: fp64 {out['fp64']['adjoint_ms']:.2f} ms, mixed "
        f"{out['mixed']['adjoint_ms']:.2f} ms, mixed vs fp64 {out['grad_mixed_vs_fp64']:.1e}")
    return out


def elastic_leg(ne=5461, npg=16, seed=1, cpu_rows=1024, threads=1, reps=2):
    """Kelvin elasticity on the star, the headline's two pipelines: mixed
    (FP32 H, G with FP64 near / diagonal values -> per-dof BCs with the FP64
    right-hand side from the operator pass -> FP32 LU + FP64 refinement with the
    recomputed FP64 operator) and FP64 (FP64 H, G -> FP64 LU). HIP-event times
    from the library; accuracy against the exact uniform-strain field.
    Roofline of kelvin_far_kernel: 2x2 blocks of H and G written per far (row
    node, column node) pair, 32 B (FP32) or 64 B (FP64)."""
    mesh = B.star_mesh(ne, seed=seed)
    d = B.format2d(mesh, B.Elasticity(1.0, 0.3, 1.0))
    neu = np.nonzero(d.BC.reshape(-1, 2)[:, 0] == 1)[0]
    ana = B.ana_elasticity_patch(1.0, 0.3, 0.01, -0.004, 0.003)
    B.apply_analytical_bc(d, ana, neu)
    nt, n = d.nt, d.n
    ctx = B.HipContext(0)
    d.hip = ctx

    def phases(c):
        m = c["ms"]
        return {"assemble_ms": m["assemble"], "far_ms": m["assemble_far"],
                "classify_ms": m["assemble_classify"], "scan_ms": m["assemble_scan"],
                "near_chain_ms": m["assemble_near"], "finalize_ms": m["assemble_finalize"],
                "rhs_ms": m["apply_bc_rhs"], "swap_ms": m["apply_bc_swap"]}

    res = []
    for r in range(reps):
        B.assemble(d, npg, precision="mixed", fetch=False)
        B.solve(d, mode="mixed", tol=1e-10)
        c = ctx.counters()
        res.append(dict(phases(c), lu32_ms=c["ms"]["lu32"], refine_ms=c["ms"]["refine"],
                        refine_iters=d.solve_info["iters"], relres=d.solve_info["relres"],
                        near_pairs=c["stats"]["near_pairs"]))
        log(f"[bench] elastic mixed rep {r + 1}/{reps}: assemble {res[-1]['assemble_ms']:.2f} ms "
            f"(far {res[-1]['far_ms']:.2f}), lu32 {res[-1]['lu32_ms']:.1f} ms, refine "
            f"{res[-1]['refine_ms']:.1f} ms ({res[-1]['refine_iters']} it)")
    last = res[-1]
    xm = d.solve_info["x"].copy()
    err_u_mixed = B.rel_error(d)
    err_t_mixed = B.rel_error_flux(d)
    for r in range(2):  # the first FP64 assembly allocates its matrices
        B.assemble(d, npg, precision="fp64", fetch=False)
        B.solve(d, mode="fp64")
    c = ctx.counters()
    p64 = dict(phases(c), lu64_ms=c["ms"]["lu64"])
    x64 = d.solve_info["x"]

    traffic = load_traffic(nt, KELVIN_TRAFFIC_FILE) or {}

    def roof(nbytes, ms, name, key):
        a = nbytes / (ms * 1e-3)
        return {"kernel": name, "bound": "hbm", "achieved": a / 1e9, "peak": PEAK_HBM / 1e9,
                "unit": "GB/s", "frac": a / PEAK_HBM, "traffic": traffic.get(key),
                "algorithmic_bytes": nbytes}

    out = {
        "workload": f"star N={2 * nt} dofs ({d.ne} quadratic elements, {nt} nodes), npg={npg}, "
                    "Kelvin plane strain E=1 nu=0.3, uniform strain, traction on the "
                    "Neumann arcs",
        "mixed_reps": res, "fp64": p64,
        "assemble_ms": last["assemble_ms"],
        "entries_per_s": 8.0 * nt * n / (last["assemble_ms"] * 1e-3),
        "fp64_assemble_ms": p64["assemble_ms"],
        "fp64_entries_per_s": 8.0 * nt * n / (p64["assemble_ms"] * 1e-3),
        "lu32_ms": last["lu32_ms"], "refine_ms": last["refine_ms"],
        "solve_mixed_s": (last["lu32_ms"] + last["refine_ms"]) * 1e-3,
        "lu64_ms": p64["lu64_ms"], "solve_fp64_s": p64["lu64_ms"] * 1e-3,
        "mixed_vs_fp64": float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)),
        "rel_error_u_mixed": err_u_mixed, "rel_error_traction_mixed": err_t_mixed,
        "rel_error_u_fp64": B.rel_error(d), "rel_error_traction_fp64": B.rel_error_flux(d),
        "roofline_far": roof(32.0 * nt * n, last["far_ms"],
                             "kelvin_far_kernel<float, 3> (Kelvin H,G far field, FP32 storage)",
                             "kelvin_far_kernel<float, 3, true>"),
        "roofline_far_fp64": roof(64.0 * nt * n, p64["far_ms"],
                                  "kelvin_far_kernel<double, 3> (FP64 storage)",
                                  "kelvin_far_kernel<double, 3, true>"),
    }
    ctx.close()
    if cpu_rows:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        rows = min(cpu_rows, nt)
        t = time.perf_counter()
        O.assemble_kelvin(d, npg, nthreads=threads, rows=(0, rows))
        ts = time.perf_counter() - t
        out["cpu_baseline"] = {"value": 8.0 * rows * n / ts, "unit": "entries/s",
                               "cores": threads, "kind": "port",
                               "sample": f"oracle Kelvin assembly of rows 0..{rows} "
                                         f"({ts:.2f} s)"}
    log(f"[bench] elastic N={2 * nt}: mixed assemble {out['assemble_ms']:.2f} ms "
        f"({out['entries_per_s']:.3g} entries/s, far {out['roofline_far']['achieved']:.0f} GB/s), "
        f"fp64 assemble {out['fp64_assemble_ms']:.2f} ms; solve mixed {out['solve_mixed_s']:.3f} s "
        f"vs fp64 {out['solve_fp64_s']:.3f} s, mixed vs fp64 {out['mixed_vs_fp64']:.1e}, "
        f"u err {out['rel_error_u_mixed']:.2e}")
    return out


def sweep_sizes(nseeds=64):
    """The 64-domain sweep: N ~ U[2048, 16384] (SplitMix64 seed 1000 + s), rounded
    to the nearest multiple of 3; domain s is the seed-(s + 1) star."""
    out = []
    for s in range(nseeds):
        rng = B.SplitMix64(1000 + s)
        out.append((s, int(round(rng.uniform(2048.0, 16384.0) / 3.0)) * 3))
    return out


def sweep_leg(npg=16, nseeds=64):
    """Per domain, in one context (buffers reused, factors resident from the
    solve to the adjoint): mixed pipeline (FP32 far field, FP32 LU, FP64
    refinement to 1e-10) + adjoint gradient, then the FP64 pipeline + adjoint.
    Objective J = 1/2 sum_j w_j (q_j - q_exact_j)^2, w_j = J_j w_k. Rates from the
    library's HIP-event times: entries/s = sum 2 nt n / sum assembly time,
    solves/s = domains / sum solve time (LU + solves / refinement), gradients/s =
    domains / sum adjoint time (transposed solve + gradient pass)."""
    ctx = B.HipContext(0)
    acc = {m: {"entries": 0.0, "assemble_ms": 0.0, "solve_ms": 0.0, "adjoint_ms": 0.0}
           for m in ("mixed", "fp64")}
    worst = {"T": 0.0, "q": 0.0, "grad": 0.0}
    sizes = sweep_sizes(nseeds)
    t0 = time.perf_counter()
    for s, N in sizes:
        d = build_problem(N // 3, seed=s + 1)
        d.hip = ctx
        qt = B.analytical_flux(d)[:d.n]
        res = {}
        for mode in ("mixed", "fp64"):
            B.assemble(d, npg, precision=mode, fetch=False)
            B.solve(d, mode=mode, tol=1e-10)
            J, g = B.adjoint(d, qt)
            ms = ctx.counters()["ms"]
            a = acc[mode]
            a["entries"] += 2.0 * d.nt * d.n
            a["assemble_ms"] += ms["assemble"]
            a["solve_ms"] += ms["solve_mixed" if mode == "mixed" else "solve64"]
            a["adjoint_ms"] += ms["adjoint"]
            res[mode] = (d.T.copy(), d.q.copy(), g)
        (Tm, qm, gm), (T6, q6, g6) = res["mixed"], res["fp64"]
        worst["T"] = max(worst["T"], float(np.linalg.norm(Tm - T6) / np.linalg.norm(T6)))
        worst["q"] = max(worst["q"], float(np.linalg.norm(qm - q6) / np.linalg.norm(q6)))
        worst["grad"] = max(worst["grad"], float(np.abs(gm - g6).max() / np.abs(g6).max()))
    ctx.close()
    out = {"workload": f"{nseeds} seeded stars, N ~ U[2048, 16384] "
                       f"({min(n for _, n in sizes)}..{max(n for _, n in sizes)}), npg={npg}, "
                       "forward solve + adjoint gradient per domain",
           "wall_s": time.perf_counter() - t0,
           "mixed_vs_fp64_max": worst}
    for mode, a in acc.items():
        out[mode] = {"entries_per_s": a["entries"] / (a["assemble_ms"] * 1e-3),
                     "solves_per_s": nseeds / (a["solve_ms"] * 1e-3),
                     "gradients_per_s": nseeds / (a["adjoint_ms"] * 1e-3),
                     "assemble_s": a["assemble_ms"] * 1e-3, "solve_s": a["solve_ms"] * 1e-3,
                     "adjoint_s": a["adjoint_ms"] * 1e-3}
    m, f = out["mixed"], out["fp64"]
    out["speedup_mixed_vs_fp64"] = {"assembly": m["entries_per_s"] / f["entries_per_s"],
                                    "solve": m["solves_per_s"] / f["solves_per_s"],
                                    "gradient": m["gradients_per_s"] / f["gradients_per_s"]}
    log(f"[bench] sweep {nseeds} domains: mixed {m['solves_per_s']:.1f} solves/s, "
        f"{m['gradients_per_s']:.1f} gradients/s; fp64 {f['solves_per_s']:.1f} solves/s, "
        f"{f['gradients_per_s']:.1f} gradients/s; worst mixed vs fp64 {worst}")
    return out


def interior_leg(ctx, dad, npg, T, q, m_far=65536, m_near=2048, seed=3):
    """Potential at random points inside the star from the boundary solution:
    m_far points with the assembly's rules (far lumping beyond 1.5 L) and m_near
    points with every (point, element) pair through the near rule. Times from
    HIP events around the library call; accuracy against the exact field."""
    rng = np.random.default_rng(seed)
    ana = B.ana_exp_cos()
    out = {}
    for name, m, nf in (("assembly_rules", m_far, 1.5), ("all_near", m_near, 1e300)):
        th = rng.uniform(0, 2 * np.pi, m)
        r = 0.85 * np.sqrt(rng.uniform(0, 1, m))
        pts = np.stack([r * np.cos(th), r * np.sin(th)], axis=1)
        ctx.interior(pts[:64], T, q, npg, nf)                    # warm-up
        u = ctx.interior(pts, T, q, npg, nf)
        c = ctx.counters()
        ms = c["ms"]["interior"]
        pairs = float(m) * dad.ne
        out[name] = {"points": m, "ms": ms, "points_per_s": m / (ms * 1e-3),
                     "pair_integrals_per_s": pairs / (ms * 1e-3),
                     "near_pairs": c["stats"]["interior_pairs"],
                     "max_abs_err_vs_exact": float(np.abs(u - ana.u(pts)).max())}
        log(f"[bench] interior {name}: {m} points in {ms:.2f} ms, "
            f"err vs exact {out[name]['max_abs_err_vs_exact']:.1e}")
    return out


def gmres_large(spec, npg, far="dense", rtol=1e-10, restart=100, itmax=4000):
    """Matrix-free assembly (near pairs + FP64 diagonal, no stored H, G) and
    GMRES, the far field of every product recomputed in FP64 (far "dense") or
    evaluated by the FMM (far "fmm"), for boundaries beyond the dense path's
    memory; accuracy against the exact field."""
    kind, size = spec.split(":")
    if kind == "square":
        dad = B.format2d(B.quadrado(ndiv=int(size), cells=False), B.Laplace(1.0),
                         pontointerno=False)
        B.attach_analytical(dad, B.ana_laplace_linear())
        workload = f"square N={dad.n} ({dad.ne} linear elements), T = x, npg={npg}"
    elif kind == "star":
        dad = build_problem(int(size))
        workload = f"star N={dad.n} ({dad.ne} quadratic elements), npg={npg}"
    elif kind == "esquare":
        if far == "fmm":
            raise SystemExit("--gmres-large esquare: elasticity has no FMM far field")
        dad = B.format2d(B.quadrado_elasticity(ndiv=int(size), cells=False),
                         B.Elasticity(1.0, 0.3, 1.0), pontointerno=False)
        B.apply_bc_from_physical(dad, B.ana_elasticity_patch(1.0, 0.3, 0.01, -0.004, 0.003))
        workload = (f"elastic square N={2 * dad.n} dofs ({dad.ne} linear elements), uniform "
                    f"strain, left edge clamped, traction elsewhere, npg={npg}")
    else:
        raise SystemExit(f"--gmres-large: unknown geometry {kind!r}")
    elastic = isinstance(dad.properties, B.Elasticity)
    ns = (2 if elastic else 1) * dad.nt
    ctx = Ctx()
    ctx.set_geometry(dad)
    if elastic:
        ctx.chk(ctx.L.bemhip_set_elasticity(ctx.h, float(dad.properties.mu),
                                            float(dad.properties.effective_nu())))
    BC = np.ascontiguousarray(dad.BC, dtype=np.int64)
    BV = np.ascontiguousarray(dad.BV, dtype=np.float64)
    t = time.perf_counter()
    ctx.chk(ctx.L.bemhip_assemble(ctx.h, npg, 1.5, 3 if far == "fmm" else 2))
    ctx.chk(ctx.L.bemhip_apply_bc(ctx.h, i64ptr(BC), dptr(BV), None, None))
    ms = ctx.counters()["ms"]
    log(f"[bench] matrix-free ({far} far field) N={ns}: assembly {ms['assemble']:.1f} ms, "
        f"rhs {ms['apply_bc_rhs']:.1f} ms")
    x = np.empty(ns)
    it = C.c_int(0)
    rel = C.c_double(0)
    status = ctx.L.bemhip_solve_gmres(ctx.h, 0.0, rtol, restart, itmax, dptr(x), C.byref(it),
                                      C.byref(rel))
    wall = time.perf_counter() - t
    c = ctx.counters()
    if elastic:
        u, tr = B.split_sol(dad, x)
        dad.set_cache(u=u, traction=tr, T=u)
    else:
        T, q = B.split_sol(dad, x)
        dad.set_cache(T=T, q=q)
    out = {"gmres_large": {"workload": workload, "far_field": far, "converged": status == 0,
                           "assemble_ms": c["ms"]["assemble"], "rhs_ms": c["ms"]["apply_bc_rhs"],
                           "gmres_ms": c["ms"]["gmres"], "products": it.value,
                           "relres": rel.value,
                           "ms_per_product": c["ms"]["gmres"] / max(it.value, 1), "wall_s": wall,
                           "rel_error": B.rel_error(dad), "restart": restart, "rtol": rtol,
                           "dense_fp32_HG_GB": 2 * 4.0 * ns * (ns // dad.nt) * dad.n / 1e9}}
    log(f"[bench] gmres: {it.value} products, {c['ms']['gmres']:.0f} ms, relres {rel.value:.1e}, "
        f"rel_error {out['gmres_large']['rel_error']:.2e}")
    ctx.close()
    print(json.dumps(out))


def getrf_yardstick(n: int):
    """Same-box yardstick for the LU, outside the library: rocSOLVER's
    rocsolver_sgetrf / rocsolver_dgetrf (blocked LU with partial pivoting) on a
    random n x n matrix already on the GPU (torch allocates it), timed with
    events on the stream the handle uses, beside this library's FP32 / FP64 LU
    of the same matrix (device time of the factorization alone)."""
    import torch
    roc = C.CDLL("/opt/rocm/lib/librocsolver.so.0")
    rb = C.CDLL("/opt/rocm/lib/librocblas.so.5")
    handle = C.c_void_p()
    assert rb.rocblas_create_handle(C.byref(handle)) == 0
    stream = torch.cuda.current_stream()
    rb.rocblas_set_stream(handle, C.c_void_p(stream.cuda_stream))
    out = {"yardstick": "rocsolver getrf vs this library's LU, random normal matrix",
           "n": n}
    g = torch.Generator(device="cuda").manual_seed(0)
    for dt, fn, key in ((torch.float32, roc.rocsolver_sgetrf, "s"),
                        (torch.float64, roc.rocsolver_dgetrf, "d")):
        A0 = torch.randn((n, n), dtype=dt, device="cuda", generator=g)
        ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        times = []
        for rep in range(2):  # the first call includes rocSOLVER's set-up
            A = A0.clone()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            st = fn(handle, n, n, C.c_void_p(A.data_ptr()), n, C.c_void_p(ipiv.data_ptr()),
                    C.c_void_p(info.data_ptr()))
            e1.record(stream)
            torch.cuda.synchronize()
            assert st == 0 and int(info.item()) == 0, (st, int(info.item()))
            times.append(e0.elapsed_time(e1))
            del A
        out[f"rocsolver_{key}getrf_ms"] = times[-1]
        # this library's LU of the same matrix (host round trip; device time read back)
        Ah = np.asfortranarray(A0.cpu().numpy())
        del A0
        torch.cuda.empty_cache()
        ctx = B.HipContext(0)
        B.lu_factor(Ah, ctx=ctx)
        ms = ctx.counters()["ms"]
        ctx.close()
        lk = "lu32" if key == "s" else "lu64"
        out[f"bemhip_{lk}_ms"] = ms[lk]
        flops = 2.0 / 3.0 * float(n) ** 3
        out[f"tflops_rocsolver_{key}getrf"] = flops / (times[-1] * 1e-3) / 1e12
        out[f"tflops_bemhip_{lk}"] = flops / (ms[lk] * 1e-3) / 1e12
        log(f"[bench] {key}getrf {times[-1]:.1f} ms vs {lk} {ms[lk]:.1f} ms")
    rb.rocblas_destroy_handle(handle)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ne", type=int, default=10923, help="elements (N_eff = 3 ne)")
    ap.add_argument("--npg", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-interior", action="store_true")
    ap.add_argument("--no-elastic", action="store_true")
    ap.add_argument("--elastic-only", action="store_true",
                    help="only the elasticity leg (for profiling)")
    ap.add_argument("--gmres-large", default="", metavar="SPEC",
                    help="only: matrix-free assembly + GMRES; SPEC square:NDIV (linear elements, "
                         "T = x), star:NE (quadratic elements) or esquare:NDIV (elasticity, "
                         "uniform strain)")
    ap.add_argument("--far", choices=["dense", "fmm"], default="dense",
                    help="far field of the --gmres-large products")
    ap.add_argument("--phase", choices=["all", "assemble"], default="all",
                    help="'assemble': only assembly + BC passes (for counter runs)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--getrf-yardstick", action="store_true",
                    help="only: rocSOLVER sgetrf / dgetrf vs this library's LU at n = N_eff")
    ap.add_argument("--cpu-leg", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-rows", type=int, default=16384, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-lu-n", type=int, default=10240, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the path is single-GPU; "
              "see the sharding notes); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.cpu_leg:
        cpu_leg(args.ne, args.npg, args.cpu_rows, args.cpu_lu_n)
        return
    if args.gmres_large:
        gmres_large(args.gmres_large, args.npg, far=args.far)
        return
    if args.getrf_yardstick:
        print(json.dumps(getrf_yardstick(3 * args.ne)))
        return
    if args.elastic_only:
        print(json.dumps({"elastic": elastic_leg(npg=args.npg, cpu_rows=0)}))
        return
    t0 = time.perf_counter()
    dad = build_problem(args.ne)
    nt, n = dad.nt, dad.n
    log(f"[bench] star ne={dad.ne} nt={nt} npg={args.npg} built in "
        f"{time.perf_counter() - t0:.1f}s")
    ctx = Ctx()
    ctx.set_geometry(dad)
    if args.phase == "assemble":
        BC = np.ascontiguousarray(dad.BC, dtype=np.int64)
        BV = np.ascontiguousarray(dad.BV, dtype=np.float64)
        res = {}
        for prec, name in ((1, "mixed"), (0, "fp64")):
            for s in range(args.warmup + args.steps):
                ctx.chk(ctx.L.bemhip_assemble(ctx.h, args.npg, 1.5, prec))
                ctx.chk(ctx.L.bemhip_apply_bc(ctx.h, i64ptr(BC), dptr(BV), None, None))
            res[name] = ctx.counters()["ms"]
            log(f"[bench] {name} assembly: {res[name]}")
        ctx.close()
        print(json.dumps({"phase": "assemble", "nt": nt, "ms": res}))
        return
    for w in range(args.warmup):
        ctx.pipeline(dad, args.npg, 1, 1)
        log(f"[bench] warmup {w + 1}/{args.warmup} done")
    asm_ms, lu_ms, ref_ms, solve_ms, iters, far_ms = [], [], [], [], [], []
    big_ms_l, big_n_l, big_flops_l, big_bytes_l = [], [], [], []
    t_start = time.perf_counter()
    x_mixed = None
    for s in range(args.steps):
        x_mixed, it, rel = ctx.pipeline(dad, args.npg, 1, 1)
        cnt = ctx.counters()["ms"]
        asm_ms.append(cnt["assemble"])
        far_ms.append(cnt["assemble_far"])
        lu_ms.append(cnt["lu32"])
        st = ctx.counters()["stats"]
        big_ms_l.append(cnt["lu32_big_gemm"])
        big_n_l.append(st["lu32_big_gemm_launches"])
        big_flops_l.append(st["lu32_big_gemm_flops"])
        big_bytes_l.append(st["lu32_big_gemm_bytes"])
        ref_ms.append(cnt.get("refine", 0.0))
        solve_ms.append(cnt["solve_mixed"])
        iters.append(it)
        log(f"[bench] step {s + 1}/{args.steps}: assemble {cnt['assemble']:.2f} ms, "
            f"lu32 {cnt['lu32']:.1f} ms, refine {cnt.get('refine', 0):.1f} ms ({it} it), "
            f"relres {rel:.2e}")
    t_total = time.perf_counter() - t_start
    counters = ctx.counters()
    # the LU's phase breakdown from one more, untimed, pipeline with the phase
    # timers on (their event records lengthen the LU, so the timed steps run
    # without them)
    ctx.chk(ctx.L.bemhip_set_lu_timers(ctx.h, 1))
    ctx.pipeline(dad, args.npg, 1, 1)
    ph = ctx.counters()["ms"]
    ctx.chk(ctx.L.bemhip_set_lu_timers(ctx.h, 0))
    lu_phases = {k: ph[k] for k in ("lu32", "lu32_panel", "lu32_swap", "lu32_trsm", "lu32_gemm",
                                    "lu32_big_gemm") if k in ph}
    ms_step = 1e3 * t_total / args.steps
    asm_avg = float(np.mean(asm_ms))
    entries = 2.0 * nt * n
    value = entries / (asm_avg * 1e-3)
    solve_mixed_s = float(np.mean(solve_ms)) * 1e-3    # FP32 LU + FP64 refinement

    detail = {"valu": (json.load(open(VALU_FILE)) if os.path.exists(VALU_FILE) and nt == 32769
                       else None),
              "assemble_ms": asm_avg, "far_kernel_ms": float(np.mean(far_ms)),
              "lu32_ms": float(np.mean(lu_ms)), "refine_ms": float(np.mean(ref_ms)),
              "refine_iters": iters, "solve_mixed_s": solve_mixed_s,
              "counters": counters, "lu32_phases_ms_timers_on": lu_phases}

    # all-FP64 pipeline once: speed-up and accuracy reference
    if not args.no_fp64:
        t = time.perf_counter()
        x64, _, rel64 = ctx.pipeline(dad, args.npg, 0, 0)
        c64 = ctx.counters()["ms"]
        detail.update({
            "fp64_pipeline_s": time.perf_counter() - t,
            "fp64_assemble_ms": c64["assemble"], "fp64_far_kernel_ms": c64["assemble_far"],
            "lu64_ms": c64["lu64"], "solve_fp64_s": c64["solve64"] * 1e-3,  # LU + solves
            "fp64_relres": rel64,
            "fp64_entries_per_s": entries / (c64["assemble"] * 1e-3),
        })
        st64 = ctx.counters()["stats"]
        n64 = int(st64.get("lu64_big_gemm_launches", 0))
        if n64 and c64.get("lu64_big_gemm"):
            ach64 = st64["lu64_big_gemm_flops"] / (c64["lu64_big_gemm"] * 1e-3)
            detail["roofline_lu64"] = {
                "kernel": "gemm_tile_kernel<double, 256, 128, 4, 2, false> (FP64 LU trailing update)",
                "bound": "mfma", "achieved": ach64 / 1e12, "peak": PEAK_MFMA_F64 / 1e12,
                "unit": "TFLOP/s", "frac": ach64 / PEAK_MFMA_F64, "launches": n64,
                "avg_launch_us": 1e3 * c64["lu64_big_gemm"] / n64,
                "flops_per_launch": st64["lu64_big_gemm_flops"] / n64}
        lu_flops = 2.0 / 3.0 * float(nt) ** 3
        detail["lu_rates"] = {
            "lu32_tflops": lu_flops / (float(np.mean(lu_ms)) * 1e-3) / 1e12,
            "lu64_tflops": lu_flops / (c64["lu64"] * 1e-3) / 1e12,
            "lu32_frac_of_fp32_peak": lu_flops / (float(np.mean(lu_ms)) * 1e-3) / PEAK_MFMA_F32,
            "lu64_frac_of_fp64_peak": lu_flops / (c64["lu64"] * 1e-3) / PEAK_MFMA_F64}
        T64, q64 = B.split_sol(dad, x64)
        Tm, qm = B.split_sol(dad, x_mixed)
        eT = float(np.linalg.norm(Tm - T64) / np.linalg.norm(T64))
        eq = float(np.linalg.norm(qm - q64) / np.linalg.norm(q64))
        dad.set_cache(T=T64, q=q64)
        detail.update({"err_T_mixed_vs_fp64": eT, "err_q_mixed_vs_fp64": eq,
                       "rel_error_fp64": B.rel_error(dad),
                       "rel_error_flux_fp64": B.rel_error_flux(dad),
                       "solve_speedup_mixed_vs_fp64": (c64["solve64"] * 1e-3) / solve_mixed_s,
                       "assembly_speedup_mixed_vs_fp64": c64["assemble"] / asm_avg,
                       "lu64_gemm_ms": c64.get("lu64_gemm")})
        log(f"[bench] fp64 pipeline: assemble {c64['assemble']:.2f} ms, lu64 {c64['lu64']:.1f} ms;"
            f" mixed vs fp64: T {eT:.2e} q {eq:.2e}")
    if not args.no_interior:
        Tm, qm = B.split_sol(dad, x_mixed)
        detail["interior"] = interior_leg(ctx, dad, args.npg, Tm[:n], qm)
    ctx.close()

    if not args.no_adjoint:
        detail["adjoint"] = adjoint_leg()
    if not args.no_sweep:
        detail["sweep"] = sweep_leg(npg=args.npg)
    if not args.no_elastic:
        detail["elastic"] = elastic_leg(npg=args.npg, cpu_rows=0 if args.no_cpu else 1024,
                                        threads=host_threads())

    # roofline of the dominant arithmetic kernel: the LU's FP32 trailing
    # updates on 256x256 tiles (split-bf16: every operand split into three bf16
    # parts, six partial products, so the matrix cores execute 6 bf16 flops per
    # FP32 flop of the product), each launch timed with HIP events on its stream
    big_ms = float(np.mean(big_ms_l))
    big_n = int(big_n_l[-1])
    big_flops = float(np.mean(big_flops_l))
    ach = big_flops / (big_ms * 1e-3)
    traffic = load_traffic(nt)
    big_name = "gemm_split_kernel"
    roofline = {"kernel": big_name + " (LU trailing update, split-bf16 MFMA)", "bound": "mfma",
                "achieved": 6.0 * ach / 1e12, "peak": PEAK_MFMA_BF16 / 1e12, "unit": "TFLOP/s",
                "frac": 6.0 * ach / PEAK_MFMA_BF16,
                "traffic": traffic.get(big_name) if traffic else None,
                "fp32_equivalent": {"achieved": ach / 1e12, "peak": PEAK_MFMA_F32 / 1e12,
                                    "frac": ach / PEAK_MFMA_F32},
                "launches": big_n, "flops_per_launch": big_flops / max(big_n, 1),
                "mfma_bf16_flops_per_launch": 6.0 * big_flops / max(big_n, 1),
                "algorithmic_bytes_per_launch": float(np.mean(big_bytes_l)) / max(big_n, 1),
                "avg_launch_us": 1e3 * big_ms / max(big_n, 1)}
    far_bytes = entries * 4.0
    far_s = float(np.mean(far_ms)) * 1e-3
    roofline_far = {"kernel": "far_kernel<float,3> (H,G far field)", "bound": "hbm",
                    "achieved": far_bytes / far_s / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                    "frac": far_bytes / far_s / PEAK_HBM,
                    "traffic": traffic.get("far_kernel<float, 3>") if traffic else None,
                    "algorithmic_bytes": far_bytes}

    cpu = None
    if not args.no_cpu:
        threads = host_threads()
        rows, lu_n = min(nt, 16384), min(nt, 10240)
        log(f"[bench] cpu baseline on {threads} threads ...")
        cb = cpu_baseline(dad.ne, args.npg, rows, lu_n, threads)
        asm_cpu_s = entries / cb["entries_per_s"]
        cpu = {"value": cb["entries_per_s"], "unit": "entries/s", "cores": threads,
               "kind": "port",
               "sample": f"oracle assembly of rows 0..{rows} of this workload "
                         f"({cb['assembly_sample_s']:.2f} s); LAPACK LU n={lu_n} "
                         f"({cb['lu_sample_s']:.2f} s) -> {cb['lu_extrapolated_s']:.1f} s at "
                         f"n={nt} (n^3)",
               "host": {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
                        "OMP_NUM_THREADS_env": os.environ.get("OMP_NUM_THREADS"),
                        "cpu_model": cpu_model(), "threadpools": cb["threadpools"]},
               "assembly_extrapolated_s": asm_cpu_s,
               "lu_extrapolated_s": cb["lu_extrapolated_s"],
               "gpu_vs_cpu": {"assembly": asm_cpu_s / (asm_avg * 1e-3),
                              "solve_mixed": cb["lu_extrapolated_s"] / solve_mixed_s,
                              "target": 50.0}}
    out = {
        "metric": "H,G entries/s; solve s to 1e-10 rel, mixed vs FP64, N=32768, 1 GPU",
        "value": value, "unit": "entries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded star boundary, analytic harmonic BCs)",
        "config": {"workload": f"star N={nt} ({dad.ne} quadratic elements), npg={args.npg}, "
                               "mixed FP32 far field + FP32 LU + FP64 refinement",
                   "N_eff": nt, "elements": dad.ne, "npg": args.npg},
        "roofline": roofline, "roofline_far": roofline_far, "cpu_baseline": cpu,
        "detail": detail,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
