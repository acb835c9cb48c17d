"""Autotuner of the FFT tile schedules (elements per thread E, lines per block P)
of the Wigner column passes (V) and row passes (K), of the density-matrix row
and column passes, and of the batched propagator (E per grid size).

Each candidate is compiled, run for a few steps on the device, and timed with
HIP events on the propagator's stream (fw_wigner_kernel_times) in the default
two-pass steps (the fused VV column pass and the normalising K row pass). The
column and row schedules are independent kernels and are chosen
independently. The winners are written in the format of
``fastwigner_amd/tuning.json``.

Usage (on the GPU):
    python -m fastwigner_amd.autotune --grids 1024 2048 --precision fp64 fp32 --out tuning.json
    python -m fastwigner_amd.autotune --kind rho --grids 1024 --out tuning.json
    python -m fastwigner_amd.autotune --kind batch --precision fp32 --out tuning.json
"""
import argparse
import ctypes
import json
import sys

import numpy as np

from . import _lib, tuning

PHYS = dict(X_amplitude=8., P_amplitude=8., dt=0.01, t=0., V="0.5*X*X+0.1*X*X*X*X", K="0.5*P*P",
            diff_V="X+0.4*X*X*X", diff_K="P")


def twiddle_entries(n, fp32):
    """LDS twiddle entries (16 bytes each): the full padded table in FP32; none
    in FP64, whose twiddles are computed in the kernel."""
    return n * 17 // 16 + 1 if fp32 else 1


def candidates(n_line, n_lines_pairs, esz_c, fp32, max_threads=1024, max_lds=160 * 1024):
    out = []
    for e in (4, 8, 16, 32):
        if e > n_line:
            continue
        tpl = n_line // e
        for p in (1, 2, 4, 8, 16):
            nt = p * tpl
            if nt < 64 or nt > max_threads or n_lines_pairs % p:
                continue
            raw = p * n_line
            if (raw + raw // 16 + 1) * esz_c + twiddle_entries(n_line, fp32) * 16 + 1152 > max_lds:
                continue
            out.append((e, p))
    return out


def measure(n, prec, tiles, steps=64):
    from .wigner_moyal import WignerMoyalHIP1D
    s = WignerMoyalHIP1D(X_gridDIM=n, P_gridDIM=n, precision="fp32" if prec else "fp64", tiles=tiles, **PHYS)
    s.set_wignerfunction(np.exp(-(s.X - 1.) ** 2 - (s.P - 0.5) ** 2))
    s.propagate(8)
    _lib.call("fw_wigner_set_timing", s._handle, 1)
    s.propagate(steps)
    t = (ctypes.c_double * 3)()
    _lib.call("fw_wigner_kernel_times", s._handle, t)
    return list(t)


def tune(n, prec, log=sys.stderr, sweeps=2):
    """Every candidate is timed in `sweeps` passes over the list (the minimum
    counts), so that a drift of the device's clocks during the sweep does not
    pick the winner."""
    esz_c = 8 if prec else 16
    col = candidates(n, n // 2, esz_c, bool(prec))
    row = candidates(n, n // 2, esz_c, bool(prec))
    m = max(len(col), len(row))
    best = {}
    for sweep in range(sweeps):
        for i in range(m):
            c = col[i % len(col)]
            r = row[i % len(row)]
            tiles = (c[0], c[1], r[0], r[1])
            try:
                t = measure(n, prec, tiles)
            except _lib.FastWignerError as ex:
                if sweep == 0:
                    print("  tiles %s failed: %s" % (tiles, str(ex).splitlines()[0]), file=log, flush=True)
                continue
            print("  %dx%d %s tiles %s: vv %.2f k %.2f us" % (
                n, n, "fp32" if prec else "fp64", tiles, t[0], t[1]), file=log, flush=True)
            old = best.get(i, (1e30, 1e30))
            best[i] = (min(old[0], t[0]), min(old[1], t[1]))
    best_col, best_row = (None, 1e30), (None, 1e30)
    for i, (vt, kt) in best.items():   # vt: the fused column pass (vv), kt: the row pass
        if i < len(col) and vt < best_col[1]:
            best_col = (col[i], vt)
        if i < len(row) and kt < best_row[1]:
            best_row = (row[i], kt)
    tiles = (best_col[0][0], best_col[0][1], best_row[0][0], best_row[0][1])
    return tiles, {"vv_pass_us": best_col[1], "k_pass_us": best_row[1]}


# --------------------------------------------------------------------------- density matrix
RHO_PHYS = dict(X_amplitude=10., dt=0.005, t=0., V="-0.5*X*X + X*X*X*X/20.", K="0.5*P*P", diff_V="-X + X*X*X/5.",
                diff_K="P")


def rho_candidates(n, fp32, rows):
    """(E, P) tiles of the density-matrix row (A, C) or column (B) passes that
    fit 1024 threads and the 160 KiB of LDS (the library's own check)."""
    esz = 8 if fp32 else 16
    tws = (n * 17 // 16 + 1) * 16 if fp32 else 16
    out = []
    for e in ((4, 8, 16) if rows else (8, 16, 32)):
        if e > n:
            continue
        for p in (1, 2, 4, 8, 16):
            nt = p * (n // e)
            lines = p * n
            if nt < 64 or nt > 1024 or n % p or (lines + lines // 16 + 1) * esz + tws + 20 * 32 * 8 > 160 * 1024:
                continue
            out.append((e, p))
    return out


def measure_rho(n, prec, tiles, steps=64):
    """Mean device time (us) of the passes A, B, C of the von Neumann step with
    the Ehrenfest sums (HIP events around each pass, fw_rho_kernel_times)."""
    from .rho_vneumann import RhoVNeumannHIP1D
    s = RhoVNeumannHIP1D(X_gridDIM=n, precision="fp32" if prec else "fp64", tiles=tiles, **RHO_PHYS)
    x = s.X
    psi = np.exp(-(x - 1.) ** 2 / 2. + 0.5j * x)
    s.set_rho(np.outer(psi.conj(), psi))
    s.propagate(8)
    _lib.call("fw_rho_set_timing", s._handle, 1)
    s.propagate(steps)
    t = (ctypes.c_double * 3)()
    _lib.call("fw_rho_kernel_times", s._handle, t)
    return list(t)


def tune_rho(n, prec, log=sys.stderr, sweeps=2):
    """Rows (passes A and C) and columns (pass B) are independent kernels and
    are chosen independently, the minimum over `sweeps` passes counting."""
    row = rho_candidates(n, bool(prec), True)
    col = rho_candidates(n, bool(prec), False)
    m = max(len(row), len(col))
    best = {}
    for sweep in range(sweeps):
        for i in range(m):
            r, c = row[i % len(row)], col[i % len(col)]
            tiles = (r[0], r[1], c[0], c[1])
            try:
                t = measure_rho(n, prec, tiles)
            except _lib.FastWignerError as ex:
                if sweep == 0:
                    print("  rho tiles %s failed: %s" % (tiles, str(ex).splitlines()[0]), file=log, flush=True)
                continue
            print("  rho %d^2 %s tiles %s: A %.2f B %.2f C %.2f us" % (n, "fp32" if prec else "fp64", tiles, *t),
                  file=log, flush=True)
            old = best.get(i, (1e30, 1e30))
            best[i] = (min(old[0], t[0] + t[2]), min(old[1], t[1]))
    best_row, best_col = (None, 1e30), (None, 1e30)
    for i, (rt, ct) in best.items():
        if i < len(row) and rt < best_row[1]:
            best_row = (row[i], rt)
        if i < len(col) and ct < best_col[1]:
            best_col = (col[i], ct)
    tiles = (best_row[0][0], best_row[0][1], best_col[0][0], best_col[0][1])
    return tiles, {"row_passes_us": round(best_row[1], 3), "col_pass_us": round(best_col[1], 3)}


# --------------------------------------------------------------------------- batched propagator
BATCH_SIZES = (64, 128, 256, 512, 1024)


def batch_e_ok(n, e, bt=256):
    """The library's rule for E at grid size N in a 256-thread tile."""
    if e & (e - 1) or e < 2 or e > 16 or e > n:
        return False
    tpl = n // e
    return tpl <= bt and (n // 2) % (bt // tpl) == 0


def measure_batch(prec, e_for_n, inst, steps=640):
    """Device time (us) of one step of the batched random-collision workload
    (Gibbs sources prepared, graph replays between HIP events)."""
    from .batch import QUARTIC_DV, QUARTIC_V, WignerBatchHIP1D, gaussian_state
    b = WignerBatchHIP1D(inst, V=QUARTIC_V, K="0.5*P*P", dt=0.01, param_names=("a2", "a3", "a4"), diff_V=QUARTIC_DV,
                         diff_K="P", precision="fp32" if prec else "fp64", e_for_n=list(e_for_n))
    b.prepare_random_collision(dbeta=0.01)
    for i, d in enumerate(inst):
        b.set_wignerfunction(i, gaussian_state(d["X_gridDIM"], 8., d["x0"], d["p0"]))
    b.propagate(64, record=False)
    return b.step_time_us(steps)


def tune_batch(prec, log=sys.stderr, count=256, seed=4):
    """Coordinate search over E per grid size on the config-4 workload (256
    instances drawn from 128..1024^2): each size in turn takes the E that
    gives the shortest step with the others fixed."""
    from .batch import random_quartic_instances
    inst, _ = random_quartic_instances(count, seed=seed)
    present = sorted({d["X_gridDIM"] for d in inst})
    cur = [0] * len(BATCH_SIZES)          # 0: the library default
    best = measure_batch(prec, cur, inst)
    print("  batch %s default E: %.1f us/step" % ("fp32" if prec else "fp64", best), file=log, flush=True)
    for n in present:
        k = BATCH_SIZES.index(n)
        for e in (2, 4, 8, 16):
            if not batch_e_ok(n, e):
                continue
            trial = list(cur)
            trial[k] = e
            try:
                t = measure_batch(prec, trial, inst)
            except _lib.FastWignerError as ex:
                print("  batch E %s failed: %s" % (trial, str(ex).splitlines()[0]), file=log, flush=True)
                continue
            print("  batch E %s: %.1f us/step" % (trial, t), file=log, flush=True)
            if t < best:
                best, cur = t, trial
    return cur, {"step_us": round(best, 3), "instances": count, "seed": seed}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kind", choices=("wigner", "rho", "batch"), default="wigner")
    ap.add_argument("--grids", type=int, nargs="+", default=[256, 512, 1024, 2048])
    ap.add_argument("--precision", nargs="+", default=["fp64", "fp32"])
    ap.add_argument("--out", default=None, help="write the winners to this tuning file")
    args = ap.parse_args()
    result = {}
    if args.kind == "batch":
        for p in args.precision:
            prec = 1 if p == "fp32" else 0
            e_for_n, info = tune_batch(prec)
            print("[autotune] batch %s -> E %s" % (p, e_for_n), file=sys.stderr, flush=True)
            result["batch:%s" % p] = {"e_for_n": e_for_n, **info}
            if args.out:
                tuning.store_entry("batch:%s" % p, {"e_for_n": e_for_n, **info}, path=args.out)
        print(json.dumps(result))
        return
    if args.kind == "rho":
        for n in args.grids:
            for p in args.precision:
                prec = 1 if p == "fp32" else 0
                tiles, info = tune_rho(n, prec)
                print("[autotune] rho %dx%d %s -> %s" % (n, n, p, tiles), file=sys.stderr, flush=True)
                result[tuning.key("rho", n, n, prec)] = {"tiles": list(tiles), **info}
                if args.out:
                    tuning.store("rho", n, n, prec, tiles, info, path=args.out)
        print(json.dumps(result))
        return
    for n in args.grids:
        for p in args.precision:
            prec = 1 if p == "fp32" else 0
            tiles, info = tune(n, prec)
            print("[autotune] wigner %dx%d %s -> %s" % (n, n, p, tiles), file=sys.stderr, flush=True)
            result[tuning.key("wigner", n, n, prec)] = {"tiles": list(tiles), "vv_pass_us": info["vv_pass_us"],
                                                         "k_pass_us": info["k_pass_us"]}
            if args.out:
                tuning.store("wigner", n, n, prec, tiles,
                             {"vv_pass_us": round(info["vv_pass_us"], 3), "k_pass_us": round(info["k_pass_us"], 3)},
                             path=args.out)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
