"""Autotuner of the FFT tile schedules (elements per thread E, lines per block P)
of the Wigner column passes (V) and row passes (K).

Each candidate is compiled, run for a few steps on the device, and timed with
HIP events on the propagator's stream (fw_wigner_kernel_times) in the default
two-pass steps (the fused VV column pass and the normalising K row pass). The
column and row schedules are independent kernels and are chosen
independently. The winners are written in the format of
``fastwigner_amd/tuning.json``.

Usage (on the GPU):
    python -m fastwigner_amd.autotune --grids 1024 2048 --precision fp64 fp32 --out tuning.json
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grids", type=int, nargs="+", default=[256, 512, 1024, 2048])
    ap.add_argument("--precision", nargs="+", default=["fp64", "fp32"])
    ap.add_argument("--out", default=None, help="write the winners to this tuning file")
    args = ap.parse_args()
    result = {}
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
